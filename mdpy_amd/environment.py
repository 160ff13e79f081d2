"""Global environment flags (mdpy/environment.py:13-82).

Precision selects the dtype of the arrays handed back to callers (positions, velocities,
forces), exactly as the reference's SINGLE/DOUBLE switch does. The device arithmetic is fixed
by design: float32 pair/PME kernels, float64 positions/velocities/integrator, int64 fixed-point
force accumulation. The platform flag is kept for API compatibility; every constraint and
integrator of this package runs on the HIP device whatever it says.
"""
import numpy as np

from .error import EnvironmentVariableError


class MDPYEnvironment:
    def __init__(self) -> None:
        self._supported_precisions = ["SINGLE", "DOUBLE"]
        self._default_precisions = "SINGLE"
        self._supported_platforms = ["CPU", "CUDA", "HIP"]
        self._default_platforms = "HIP"
        self.set_precision(self._default_precisions)
        self.set_platform(self._default_platforms)

    def set_precision(self, precision: str):
        precision = precision.upper()
        if precision not in self._supported_precisions:
            raise EnvironmentVariableError(
                "unknown precision %r; the choices are listed in mdpy_amd.env.supported_precisions" % precision)
        self._precision = precision
        if precision == "SINGLE":
            self.NUMPY_FLOAT, self.NUMPY_INT = np.float32, np.int32
        else:
            self.NUMPY_FLOAT, self.NUMPY_INT = np.float64, np.int64

    def set_platform(self, platform: str):
        platform = platform.upper()
        if platform not in self._supported_platforms:
            raise EnvironmentVariableError(
                "unknown platform %r; the choices are listed in mdpy_amd.env.supported_platforms" % platform)
        self._platform = platform

    def set_default(self):
        self.set_precision(self._default_precisions)
        self.set_platform(self._default_platforms)

    @property
    def supported_precisions(self):
        return self._supported_precisions

    @property
    def precision(self):
        return self._precision

    @property
    def supported_platforms(self):
        return self._supported_platforms

    @property
    def platform(self):
        return self._platform


env = MDPYEnvironment()
