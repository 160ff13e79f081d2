"""Energy minimisers (mdpy/minimizer/*.py) on the device force path.

The iterations run in libmdx (mdx_minimize_sd / mdx_minimize_cg): position updates and the
conjugate-gradient sums are kernels, every force evaluation is the fused one of Ensemble.update,
and only the energy and the stopping sums come back to the host per iteration. The update
rules are the reference's, as written (including the conjugate-gradient minimiser's inner loop,
which probes from already-probed positions and applies only its last step)."""
import ctypes

from .. import _lib as L
from ..unit import KJ_PER_MOL


class Minimizer:
    """mdpy/minimizer/minimizer.py:14-34."""

    def __init__(self, output_unit=KJ_PER_MOL, output_unit_label="kj/mol", is_verbose=False, log_freq=5) -> None:
        self._output_unit = output_unit
        self._output_unit_label = output_unit_label
        self._is_verbose = is_verbose
        self._log_freq = log_freq
        self.iterations = 0

    def minimize(self, ensemble, energy_tolerance=0.001, max_iterations: int = 1000):
        raise NotImplementedError("%s must implement minimize()" % type(self).__name__)

    def _energy2str(self, energy):
        return "%.5f %s" % (energy / self._output_unit, self._output_unit_label)

    def _finish(self, ensemble, start_energy, iterations, energy):
        ensemble.state.invalidate()
        ensemble.update()
        self.iterations = iterations
        if self._is_verbose:
            print("Initial potential energy: %s" % self._energy2str(start_energy))
            print("Final potential energy: %s after %d iterations" % (self._energy2str(energy), iterations))


class SteepestDescentMinimizer(Minimizer):
    """mdpy/minimizer/steepest_descent_minimizer.py:15-53: x <- x + alpha F_i/|F_i| until the
    relative energy change drops below the tolerance."""

    def __init__(self, alpha=0.01, output_unit=KJ_PER_MOL, output_unit_label="kj/mol", is_verbose=False,
                 log_freq=5) -> None:
        super().__init__(output_unit, output_unit_label, is_verbose, log_freq)
        self._alpha = float(alpha)

    def minimize(self, ensemble, energy_tolerance=0.001, max_iterations: int = 1000):
        ensemble.update()
        e0 = ensemble.potential_energy
        it, en = ctypes.c_int(), ctypes.c_double()
        L.call("mdx_minimize_sd", ensemble.state.engine.ctx, int(ensemble.term_bits), self._alpha,
               float(energy_tolerance), int(max_iterations), ctypes.byref(it), ctypes.byref(en))
        self._finish(ensemble, e0, it.value, en.value)

    alpha = property(lambda self: self._alpha)


class ConjugateGradientMinimizer(Minimizer):
    """mdpy/minimizer/conjugate_gradient_minimizer.py:16-77."""

    def __init__(self, theta=0.001, force_tol=0.01, max_sub_iterations=10, output_unit=KJ_PER_MOL,
                 output_unit_label="kj/mol", is_verbose=False, log_freq=5) -> None:
        super().__init__(output_unit, output_unit_label, is_verbose, log_freq)
        self._theta = float(theta)
        self._force_tol = float(force_tol)
        self._max_sub_iterations = int(max_sub_iterations)

    def minimize(self, ensemble, energy_tolerance=0.001, max_iterations: int = 1000):
        ensemble.update()
        e0 = ensemble.potential_energy
        it, en = ctypes.c_int(), ctypes.c_double()
        L.call("mdx_minimize_cg", ensemble.state.engine.ctx, int(ensemble.term_bits), self._theta, self._force_tol,
               self._max_sub_iterations, float(energy_tolerance), int(max_iterations), ctypes.byref(it),
               ctypes.byref(en))
        self._finish(ensemble, e0, it.value, en.value)


__all__ = ["Minimizer", "SteepestDescentMinimizer", "ConjugateGradientMinimizer"]
