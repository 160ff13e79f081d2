"""File readers and writers (mdpy/io/__init__.py), without MDAnalysis or h5py."""
from .charmm_toppar_parser import CharmmTopparParser
from .dcd_parser import DCDParser
from .hdf5 import HDF5Parser, HDF5Writer
from .psf_parser import PSFParser
from .pdb_parser import PDBParser
from .pdb_writer import PDBWriter

HDF5_FILE_HIERARCHY = """Hierarchy of the HDF5 trajectory files (mdpy's layout)
/ topology / particles / particle_id, particle_type, particle_name, matrix_id, molecule_id,
                         molecule_type, chain_id, mass, charge
  topology / num_particles, bonds, num_bonds, angles, num_angles, dihedrals, num_dihedrals,
             impropers, num_impropers
/ positions / frame-0, frame-1, ..., num_frames
/ pbc_matrix
"""

__all__ = ["CharmmTopparParser", "DCDParser", "HDF5Parser", "HDF5Writer", "HDF5_FILE_HIERARCHY", "PSFParser",
           "PDBParser", "PDBWriter"]
