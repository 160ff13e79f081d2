"""Fit the erfc approximation used by the nonbonded kernel.

erfc(x) = exp(-x^2) * g(x),  g(x) = erfc(x) exp(x^2) ~ t * P(t),  t = 1 / (1 + p x),
P a polynomial of degree K-1, coefficients chosen to minimise the maximum RELATIVE error of g
on [0, XMAX] (least squares at Chebyshev nodes, then iteratively reweighted towards the
minimax solution, Lawson's algorithm). The form is that of Abramowitz & Stegun 7.1.26; this
refit trades two more multiply-adds for ~1e-9 relative error instead of A&S's 1.5e-7 absolute,
which matters near the cutoff where erfc itself is ~1e-5. With the Ewald tolerance 1e-5 rule
beta = sqrt(-ln(2 tol))/rc, beta*rc = 3.29, so the fit range [0, 3.6] covers every pair.

    python -m mdpy_amd.tools.fit_erfc
"""
import numpy as np
from scipy.special import erfcx

XMAX = 3.6


def fit(p, K, iters=60):
    n = 4000
    x = 0.5 * XMAX * (1 - np.cos(np.pi * (np.arange(n) + 0.5) / n))
    g = erfcx(x)
    t = 1.0 / (1.0 + p * x)
    A = np.stack([t ** (k + 1) for k in range(K)], 1) / g[:, None]
    b = np.ones(n)
    w = np.ones(n)
    for _ in range(iters):
        c, *_ = np.linalg.lstsq(A * w[:, None], b * w, rcond=None)
        err = np.abs(A @ c - 1)
        w = w * (err + 1e-30) ** 0.5
        w /= w.max()
    return c, np.abs(A @ c - 1).max()


def best(K=7):
    res = []
    for p in np.linspace(0.2, 0.6, 81):
        c, e = fit(p, K)
        res.append((e, p, c))
    e, p, c = min(res, key=lambda r: r[0])
    return p, c, e


if __name__ == "__main__":
    for K in (5, 6, 7):
        p, c, e = best(K)
        print("K=%d p=%.6f max rel err %.3e" % (K, p, e))
        print("  coeffs:", ", ".join("%.10ef" % v for v in c))
