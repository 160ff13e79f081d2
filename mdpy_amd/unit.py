"""Internal unit system and physical constants of the force-and-integrate path.

The reference works in Angstrom, Dalton, femtosecond, elementary charge and Kelvin
(mdpy/unit/__init__.py:31-36); energy is Da*A^2/fs^2 (:41). Every number that crosses
the C-ABI is in these units. The SI relative values of the base units are those of
mdpy/unit/unit_definition.py (:45 angstrom, :54 dalton, :68 femtosecond, :90 e,
:112/:117 kJ/mol and kcal/mol with N_A = 6.0221e23); KB and EPSILON0 are the SI values
of mdpy/unit/__init__.py:54-58.

The reference builds KB and EPSILON0 while the global precision is still SINGLE, so
its constants are float32-rounded (relative offset <= 6.5e-9). This module keeps the
exact float64 values.
"""
import math

SPATIAL_DIM = 3

# SI value of one internal base unit (mdpy/unit/unit_definition.py)
ANGSTROM_SI = 1e-10
DALTON_SI = 1.66053904e-27
FEMTOSECOND_SI = 1e-15
E_SI = 1.602176634e-19
AVOGADRO = 6.0221e23

# Derived internal units in SI
ENERGY_SI = DALTON_SI * ANGSTROM_SI ** 2 / FEMTOSECOND_SI ** 2  # J
FORCE_SI = ENERGY_SI / ANGSTROM_SI                              # N

KCAL_PER_MOL = 4.184e3 / AVOGADRO / ENERGY_SI   # 1 kcal/mol in internal energy units
KJ_PER_MOL = 1e3 / AVOGADRO / ENERGY_SI         # 1 kJ/mol in internal energy units
KJ_PER_MOL_PER_NM = KJ_PER_MOL / 10.0           # 1 kJ/mol/nm in internal force units

KB = 1.38064852e-23 / ENERGY_SI                  # internal energy per Kelvin
EPSILON0 = 8.85418e-12 / (FEMTOSECOND_SI ** 2 * E_SI ** 2 / (ANGSTROM_SI ** 3 * DALTON_SI))
COULOMB_K = 4.0 * math.pi * EPSILON0             # the reference divides q_i q_j by this (electrostatic_constraint.py:60)
ONE_OVER_4PI_EPS0 = 1.0 / COULOMB_K

RMIN_TO_SIGMA_FACTOR = 2.0 ** (-1.0 / 6.0)      # mdpy/io/charmm_toppar_parser.py:16

NS_PER_DAY_FROM_FS_PER_S = 86400.0 * 1e-6       # ns/day = fs per wall second * 0.0864


def ns_per_day(time_step_fs: float, seconds_per_step: float) -> float:
    """Simulated ns per wall-clock day (mdpy/dumper/log_dumper.py:199-208)."""
    return time_step_fs * 1e-6 / (seconds_per_step / 86400.0)


def kinetic_to_temperature(kinetic_energy: float, num_particles: int) -> float:
    """T = 2 KE / (3 N kB), the definition used by the reference's tests (test_state.py:77)."""
    return 2.0 * kinetic_energy / (3.0 * num_particles * KB)


def check_quantity_value(val, *_):
    """Coerce an argument to a float in internal units.

    The reference accepts Quantity objects or raw numbers already in internal units
    (mdpy/utils/check_quantity.py:20-26). Quantity objects are not part of this path;
    raw numbers pass through unchanged.
    """
    if val is None:
        return None
    if hasattr(val, "value"):
        v = val.value
        try:
            return float(v[0]) if len(v) == 1 else v
        except TypeError:
            return float(v)
    return val
