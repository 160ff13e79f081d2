"""CPU test: the erfc approximation the nonbonded kernel's energy launches use.

erfc(x) = exp(-x^2) t P(t), t = 1/(1 + p x) (the form of Abramowitz & Stegun 7.1.26) with a
7-term polynomial whose coefficients were fitted for minimum maximum relative error of
erfc(x) exp(x^2) on [0, 3.6] (beta * rc = 3.29 at the 1e-5 Ewald tolerance). The coefficients
are read from the kernel header, rounded to float32 as the kernel holds them, and evaluated in
float64 against scipy's erfcx."""
import os
import re

import numpy as np
from scipy.special import erfcx

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdpy_amd", "csrc",
                   "mdx_nonbonded_body.h")


def coefficients():
    text = open(HDR).read()
    vals = dict(re.findall(r"#define (ERFC_\w+) ([-+0-9.eE]+)f", text))
    p = float(np.float32(vals["ERFC_P"]))
    c = [float(np.float32(vals["ERFC_C%d" % k])) for k in range(1, 8)]
    return p, c


def test_erfc_polynomial_relative_error_on_fit_range():
    p, c = coefficients()
    x = np.linspace(0.0, 3.6, 200001)
    t = 1.0 / (1.0 + p * x)
    poly = np.zeros_like(t)
    for ck in reversed(c):
        poly = poly * t + ck
    rel = np.abs(t * poly / erfcx(x) - 1.0)
    # 4.53e-8 with the decimal coefficients as fitted, 5.53e-8 once rounded to float32
    assert rel.max() < 6e-8, rel.max()
    # the textbook A&S 7.1.26 (5 terms, p = 0.3275911) is ~100x worse in relative terms near the
    # cutoff; the refit is what keeps float32 accuracy where erfc ~ 1e-5
    a = [0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429]
    ta = 1.0 / (1.0 + 0.3275911 * x)
    pa = np.zeros_like(ta)
    for ak in reversed(a):
        pa = pa * ta + ak
    assert np.abs(ta * pa / erfcx(x) - 1.0).max() > 100 * rel.max()
