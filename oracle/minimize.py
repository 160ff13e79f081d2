"""Float64 restatement of mdpy's energy minimisers -- TEST INFRASTRUCTURE ONLY (see
oracle/__init__.py). `force_energy(x)` returns (forces (N,3), potential energy); `wrap(x)` is
the reference's wrap_positions (mdref.wrap_positions with the box)."""
import numpy as np


def steepest_descent(x, force_energy, wrap, alpha, energy_tolerance, max_iterations):
    """mdpy/minimizer/steepest_descent_minimizer.py:30-53: x <- wrap(x + alpha F/|F|) per atom
    until |(E - E_prev)/E_prev| < energy_tolerance. Returns (x, iterations, energies)."""
    F, E = force_energy(x)
    pre, energies, it = E, [E], 0
    while it < max_iterations:
        nrm = np.linalg.norm(F, axis=1).reshape(-1, 1)
        x = wrap(x + alpha * F / nrm)
        F, E = force_energy(x)
        energies.append(E)
        it += 1
        if abs((E - pre) / pre) < energy_tolerance:
            break
        pre = E
    return x, it, energies


def conjugate_gradient(x, force_energy, wrap, theta, force_tol, max_sub_iterations, energy_tolerance,
                       max_iterations):
    """mdpy/minimizer/conjugate_gradient_minimizer.py:33-77 as written: each sub-iteration probes
    from the current (already probed) positions along t = F, takes alpha = |F|^2 / (omega . t)
    with omega = -(F(x + theta t) - F)/theta, and only the last sub-iteration's d = alpha t is
    applied to the outer iteration's starting positions."""
    F, E = force_energy(x)
    energies, it = [E], 0
    while it < max_iterations:
        pre = E
        x0 = x.copy()
        f0 = F.reshape(-1, 1)
        d = None
        for _ in range(max_sub_iterations):
            cur_f = F.reshape(-1, 1)
            t = cur_f.copy()
            x = wrap(x + theta * t.reshape(-1, 3))
            F, E = force_energy(x)
            omega = -(F.reshape(-1, 1) - cur_f) / theta
            alpha = (cur_f.T @ cur_f).item() / (omega.T @ t).item()
            d = alpha * t
            cur_f = cur_f + alpha * omega
            if (cur_f ** 2).sum() / (f0 ** 2).sum() < force_tol:
                break
        x = wrap(x0 + d.reshape(-1, 3))
        F, E = force_energy(x)
        energies.append(E)
        it += 1
        if abs((E - pre) / pre) < energy_tolerance:
            break
    return x, it, energies
