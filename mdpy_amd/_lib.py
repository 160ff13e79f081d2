"""ctypes binding of libmdx.so (the C-ABI declared in include/mdx.h).

The library is loaded from this package directory only; there is no CPU fallback. Importing
works without a GPU (symbols can be inspected); calling into the library needs a HIP device.
"""
import ctypes
import os

from . import error

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdx.so")

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_OVERFLOW = -3
ERR_PARTICLE_LOSS = -4
ERR_NONFINITE = -5
ERR_CELL_CUTOFF = -6

TERM_LJ = 1 << 0
TERM_COUL_DIRECT = 1 << 1
TERM_COUL_RECIP = 1 << 2
TERM_BOND = 1 << 3
TERM_ANGLE = 1 << 4
TERM_DIHEDRAL = 1 << 5
TERM_IMPROPER = 1 << 6
TERM_COUL_PLAIN = 1 << 7
TERM_MLP = 1 << 8
NTERMS = 9
E_KINETIC = NTERMS
TERM_PME = TERM_COUL_DIRECT | TERM_COUL_RECIP
TERM_BONDED = TERM_BOND | TERM_ANGLE | TERM_DIHEDRAL | TERM_IMPROPER
TERM_NAMES = ["lj", "coulomb_direct", "coulomb_reciprocal", "bond", "angle", "dihedral", "improper",
              "coulomb_plain", "mlp", "kinetic"]

PAIRS_EXACT = 0
PAIRS_REFERENCE = 1

INT_VELOCITY_VERLET = 0
INT_LANGEVIN = 1
INT_VERLET_REF = 2
INT_LANGEVIN_REF = 3

BATCH_ONE_GRAPH = 0
BATCH_STREAMS = 1
BATCH_SEQUENTIAL = 2
BATCH_FUSED = 3
BATCH_MODES = {"one_graph": BATCH_ONE_GRAPH, "streams": BATCH_STREAMS, "sequential": BATCH_SEQUENTIAL,
               "fused": BATCH_FUSED}

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "mdx_last_error": (ctypes.c_char_p, []),
    "mdx_version": (ctypes.c_int, []),
    "mdx_coulomb_constant": (ctypes.c_double, []),
    "mdx_ctx_create": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "mdx_ctx_destroy": (ctypes.c_int, [_vp]),
    "mdx_ctx_set_stream": (ctypes.c_int, [_vp, _vp]),
    "mdx_system_set": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _c_dbl_p]),
    "mdx_set_box": (ctypes.c_int, [_vp, _c_dbl_p]),
    "mdx_set_masses": (ctypes.c_int, [_vp, _c_dbl_p]),
    "mdx_set_charges": (ctypes.c_int, [_vp, _c_dbl_p]),
    "mdx_set_lj": (ctypes.c_int, [_vp, _c_dbl_p]),
    "mdx_set_exclusions": (ctypes.c_int, [_vp, _c_int_p, _c_int_p]),
    "mdx_set_scaling": (ctypes.c_int, [_vp, _c_int_p, _c_int_p]),
    "mdx_set_bonds": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_dbl_p]),
    "mdx_set_angles": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_dbl_p]),
    "mdx_set_dihedrals": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_dbl_p, ctypes.c_int]),
    "mdx_set_impropers": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_dbl_p]),
    "mdx_wrap_positions": (ctypes.c_int, [_vp]),
    "mdx_positions_changed": (ctypes.c_int, [_vp]),
    "mdx_energy_history_copy_async": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _c_int_p]),
    "mdx_minimize_sd": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       _c_int_p, _c_dbl_p]),
    "mdx_minimize_cg": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, _c_int_p, _c_dbl_p]),
    "mdx_pair_histogram": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _c_dbl_p, _vp, ctypes.c_int, _vp,
                                          ctypes.c_int, _vp, ctypes.c_int, _vp]),
    "mdx_celllist_reference_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _c_dbl_p, ctypes.c_double, _vp,
                                                      _c_int_p, _c_int_p]),
    "mdx_celllist_reference": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _c_int_p, _c_int_p]),
    "mdx_celllist_table": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "mdx_nonbonded_config": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]),
    "mdx_neighbor_pairs": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _c_i64_p]),
    "mdx_pme_config": (ctypes.c_int, [_vp, _c_int_p, ctypes.c_int, ctypes.c_double]),
    "mdx_pme_set_fused": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mdx_force_phases_config": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "mdx_compute_forces": (ctypes.c_int, [_vp, ctypes.c_uint, _vp]),
    "mdx_compute_forces_only": (ctypes.c_int, [_vp, ctypes.c_uint, _vp]),
    "mdx_energies_fetch": (ctypes.c_int, [_vp, _c_dbl_p]),
    "mdx_integrator_config": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_uint64]),
    "mdx_integrator_reset": (ctypes.c_int, [_vp]),
    "mdx_set_external_noise": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "mdx_integrate": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_int, ctypes.c_int]),
    "mdx_set_energy_interval": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mdx_energy_history_fetch": (ctypes.c_int, [_vp, _c_dbl_p, ctypes.c_int, _c_int_p]),
    "mdx_status_fetch": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint)]),
    "mdx_stats_fetch": (ctypes.c_int, [_vp, _c_i64_p, _c_i64_p]),
    "mdx_pair_counts": (ctypes.c_int, [_vp, _c_i64_p, _c_i64_p]),
    "mdx_time_kernel": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint, ctypes.c_int, _c_dbl_p]),
    "mdx_nonbonded_tune": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_int, ctypes.c_int, _c_int_p]),
    "mdx_nonbonded_persist": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mdx_nonbonded_rebuild_every": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mdx_nonbonded_cells": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mdx_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "mdx_batch_integrate": (ctypes.c_int, [_vp, ctypes.c_uint, ctypes.c_int, ctypes.c_int]),
    "mdx_batch_time_kernel": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "mdx_batch_destroy": (ctypes.c_int, [_vp]),
    "mdx_mlp_create": (ctypes.c_int, [_vp, _c_int_p, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_vp)]),
    "mdx_mlp_destroy": (ctypes.c_int, [_vp]),
    "mdx_mlp_param_count": (ctypes.c_int, []),
    "mdx_mlp_set_params": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "mdx_mlp_get_params": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "mdx_mlp_set_norm": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "mdx_mlp_descriptors": (ctypes.c_int, [_vp, _vp]),
    "mdx_mlp_eval": (ctypes.c_int, [_vp, _vp, _c_dbl_p]),
    "mdx_ctx_set_mlp": (ctypes.c_int, [_vp, _vp]),
    "mdx_mlp_train_step": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, _c_dbl_p]),
    "mdx_mlp_adam_config": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "mdx_mlp_get_grads": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "mdx_mlp_gemm_check": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load libmdx.so from the package directory (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmdx.so not built: run `python -m mdpy_amd.build` (HIP path required, "
                          "no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


_EXC = {
    ERR_INVALID: ValueError,
    ERR_HIP: error.DeviceError,
    ERR_OVERFLOW: error.CellListPoorDefinedError,
    ERR_PARTICLE_LOSS: error.ParticleLossError,
    ERR_NONFINITE: error.NumericalInstabilityError,
    ERR_CELL_CUTOFF: error.CellListPoorDefinedError,
}


def check(rc):
    if rc == OK:
        return
    msg = load().mdx_last_error().decode(errors="replace")
    raise _EXC.get(rc, RuntimeError)(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def dptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array's data."""
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
