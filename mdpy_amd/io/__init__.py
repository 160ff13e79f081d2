"""File readers and writers (mdpy/io/__init__.py), without MDAnalysis or h5py."""
from .charmm_toppar_parser import CharmmTopparParser
from .dcd_parser import DCDParser
from .psf_parser import PSFParser
from .pdb_parser import PDBParser
from .pdb_writer import PDBWriter

__all__ = ["CharmmTopparParser", "DCDParser", "PSFParser", "PDBParser", "PDBWriter"]
