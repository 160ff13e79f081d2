"""Force-field base class (mdpy/forcefield/forcefield.py:12-33)."""
from ..core import Topology


class Forcefield:
    def __init__(self, topology: Topology) -> None:
        self._topology = topology
        self._parameters = None

    def set_param_files(self, *file_paths):
        raise NotImplementedError("%s must implement set_param_files()" % type(self).__name__)

    def check_parameters(self):
        raise NotImplementedError("%s must implement check_parameters()" % type(self).__name__)

    def create_ensemble(self):
        raise NotImplementedError("%s must implement create_ensemble()" % type(self).__name__)

    @property
    def parameters(self):
        return self._parameters
