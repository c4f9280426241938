"""Unit constants of fennol_md (Angstrom, ps, kcal/mol) and element data.

Values restate FeNNol's utils/atomic_units.py:16-134 (atomic-unit constants),
UnitSystem :142-260 (derived mass unit, us.DA, us.K_B) and md/utils.py:9
(the MD unit system), plus utils/periodic_table.py:18-46 (symbols, masses).
"""

from __future__ import annotations

import numpy as np

ANGSTROM = 0.52917721          # bohr -> Angstrom
FS = 2.4188843e-2              # au time -> fs
PS = FS / 1000.0
KCALPERMOL = 627.5096080305927 # Hartree -> kcal/mol
EV = 27.211386024367243        # Hartree -> eV
MPROT = 1836.1526734252586
K_B_AU = 3.166811563e-6        # Hartree / K

# us_to_au for L, T, E and the derived mass unit M = E T^2 L^-2
_L = 1.0 / ANGSTROM
_T = 1.0 / PS
_E = 1.0 / KCALPERMOL
_M = _E * _T**2 * _L**-2

DA = _M * (1.0 / MPROT)        # us.DA: mass_internal = mass_Da / DA
K_B = K_B_AU * KCALPERMOL      # us.K_B in kcal/mol/K
FS_PER_PS = _T * FS            # us.FS = 1000: dt_internal = dt_fs / FS_PER_PS
THZ = 1.0                      # us.THZ: gamma_internal = gamma_THz / THZ (ps^-1)

ENERGY_UNITS = {"HA": 1.0, "HARTREE": 1.0, "AU": 1.0, "EV": EV, "MEV": 1.0e3 * EV, "KCALPERMOL": KCALPERMOL,
                "KCAL/MOL": KCALPERMOL, "KJPERMOL": 4.184 * KCALPERMOL, "KJ/MOL": 4.184 * KCALPERMOL}


def energy_multiplier(unit: str) -> float:
    """us.get_multiplier(unit): kcal/mol -> unit (md/integrate.py:21)."""
    return ENERGY_UNITS[unit.upper()] / KCALPERMOL


PERIODIC_TABLE = [
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
]
SYMBOL_TO_Z = {s: z for z, s in enumerate(PERIODIC_TABLE)}
ATOMIC_MASSES = np.array([
    0.0, 1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999,
    18.99840316, 20.1797, 22.98976928, 24.305, 26.9815385, 28.085,
    30.973762, 32.06, 35.45, 39.948,
])
