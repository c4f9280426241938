"""Unit constants of fennol_md (Angstrom, ps, kcal/mol) and element data.

Values restate FeNNol's utils/atomic_units.py:16-134 (atomic-unit constants
and unit tables), UnitSystem :142-310 (derived base unit, multipliers, unit
strings with products, quotients, powers and parentheses: utils/
unit_parser.py), md/utils.py:9 (the MD unit system L=ANGSTROM, T=PS,
E=KCALPERMOL) and utils/periodic_table.py:18-46 (symbols, masses).

A multiplier follows FeNNol's convention: ``us.get_multiplier(u)`` is the
value of one unit of the unit system expressed in ``u`` (a value given in
``u`` is divided by it, a value in system units is multiplied by it).
"""

from __future__ import annotations

import math
import re

import numpy as np

# ---------------------------------------------------------------- atomic units
ANGSTROM = 0.52917721          # bohr -> Angstrom
FS = 2.4188843e-2              # au time -> fs
PS = FS / 1000.0
KCALPERMOL = 627.5096080305927 # Hartree -> kcal/mol
EV = 27.211386024367243        # Hartree -> eV
KJPERMOL = 2625.5002           # Hartree -> kJ/mol
MPROT = 1836.1526734252586
K_B_AU = 3.166811563e-6        # Hartree / K
AVOGADRO = 6.02214129e23
_MOL = 1.0 / AVOGADRO
KBAR_AU = 294210.2648438959    # Hartree/bohr^3 -> kbar

# value of one atomic unit in the named unit, and the unit's dimension as powers of the base
# quantities length L, time T, energy E and mass M (atomic_units.py:50-134)
_AU_UNITS: dict[str, tuple[float, dict]] = {}


def _add(names, value, dims):
    for n in names:
        _AU_UNITS[n] = (value, dims)


_add(("A0", "BOHR"), 1.0, {"L": 1})
_add(("ANGSTROM", "ANG", "A"), ANGSTROM, {"L": 1})
_add(("NM",), ANGSTROM * 1e-1, {"L": 1})
_add(("CM",), ANGSTROM * 1e-8, {"L": 1})
_add(("AU_T",), 1.0, {"T": 1})
_add(("FS", "FEMTOSECONDS"), FS, {"T": 1})
_add(("PS", "PICOSECONDS"), PS, {"T": 1})
_add(("NS", "NANOSECONDS"), PS / 1000.0, {"T": 1})
_add(("HARTREE", "HA"), 1.0, {"E": 1})
_add(("EV",), EV, {"E": 1})
_add(("MEV",), 1.0e3 * EV, {"E": 1})
_add(("KCALPERMOL",), KCALPERMOL, {"E": 1})
_add(("KJPERMOL",), KJPERMOL, {"E": 1})
_add(("RY",), 2.0, {"E": 1})
_add(("KELVIN",), 1.0 / K_B_AU, {"E": 1})
_add(("KCAL",), KCALPERMOL * _MOL, {"E": 1})
_add(("KJ",), KJPERMOL * _MOL, {"E": 1})
_add(("AU_M", "ME"), 1.0, {"M": 1})
_add(("DA",), 1.0 / MPROT, {"M": 1})
_add(("GRAM",), _MOL / MPROT, {"M": 1})
_add(("KG", "KILOGRAM"), _MOL / MPROT * 1e-3, {"M": 1})
_add(("MOL",), _MOL, {})
_add(("DEBYE",), 2.541746, {"L": 1})
_add(("THZ",), 1000.0 / FS, {"T": -1})
_add(("CM-1", "CM1"), 219471.52, {"T": -1})
_add(("KBAR",), KBAR_AU, {"L": -3, "E": 1})
_add(("ATM",), KBAR_AU * 1000.0 / 1.01325, {"L": -3, "E": 1})
_add(("GPA",), 0.1 * KBAR_AU, {"L": -3, "E": 1})
_add(("NNEWTON",), 82.387, {"L": -1, "E": 1})

_BASE_TABLE = {"L": ("A0", "BOHR", "ANGSTROM", "ANG", "A", "NM", "CM"),
               "T": ("AU_T", "FS", "FEMTOSECONDS", "PS", "PICOSECONDS", "NS", "NANOSECONDS"),
               "E": ("HARTREE", "HA", "EV", "MEV", "KCALPERMOL", "KJPERMOL", "RY", "KELVIN", "KCAL", "KJ"),
               "M": ("AU_M", "ME", "DA", "GRAM", "KG", "KILOGRAM")}

_TOKEN = re.compile(r"\s*(?:(?P<name>[A-Z0-9_-]+)|(?P<op>[*/()^]))")
_POWER = re.compile(r"\s*(?:\{(?P<b>[^}]*)\}|\((?P<p>[^)]*)\)|(?P<n>[+-]?\d+(?:\.\d+)?))")


class UnitSystem:
    """A unit system fixed by three of the base units L, T, E, M (the fourth derived from
    E = M L^2 T^-2), as FeNNol's UnitSystem (atomic_units.py:142-260)."""

    def __init__(self, L: str | None = None, T: str | None = None, E: str | None = None,
                 M: str | None = None):
        given = {k: v.strip().upper() for k, v in (("L", L), ("T", T), ("E", E), ("M", M)) if v is not None}
        if len(given) != 3:
            raise ValueError(f"a unit system needs exactly three base units, got {given}")
        to_au = {}
        for dim, name in given.items():
            if name not in _BASE_TABLE[dim]:
                raise ValueError(f"{name} is not a unit of {dim}")
            to_au[dim] = 1.0 / _AU_UNITS[name][0]
        if "L" not in to_au:
            to_au["L"] = to_au["E"] ** 0.5 * to_au["T"] * to_au["M"] ** -0.5
        elif "T" not in to_au:
            to_au["T"] = to_au["E"] ** -0.5 * to_au["L"] * to_au["M"] ** 0.5
        elif "E" not in to_au:
            to_au["E"] = to_au["L"] ** 2 * to_au["T"] ** -2 * to_au["M"]
        else:
            to_au["M"] = to_au["E"] * to_au["T"] ** 2 * to_au["L"] ** -2
        self.base = given
        self.us_to_au = to_au
        self.us_to_unit = {name: value * self._conv(dims) for name, (value, dims) in _AU_UNITS.items()}
        self.us_to_unit["1"] = 1.0
        self.us_to_unit["2PI"] = 2.0 * math.pi
        # constants in system units (atomic_units.py:31-45): the Boltzmann constant and the
        # proton mass are the two the MD path uses
        self.K_B = K_B_AU / self._conv({"E": 1})
        self.MPROT = MPROT / self._conv({"M": 1})

    def _conv(self, dims: dict) -> float:
        """Multiplier from system units to atomic units of a quantity of dimension ``dims``."""
        return math.prod(self.us_to_au[d] ** p for d, p in dims.items())

    def __getattr__(self, name):
        table = self.__dict__.get("us_to_unit")
        if table is not None and name.upper() in table:
            return table[name.upper()]
        raise AttributeError(name)

    def get_multiplier(self, unit: str) -> float:
        """Multiplier of a unit string: names, ``*``, ``/`` (applies to the next factor only),
        powers ``^2``, ``^{-1}``, ``^(2.5)`` and parenthesised groups (unit_parser.py)."""
        s = unit.strip().upper()
        if not s:
            raise ValueError("empty unit string")
        toks = []
        pos = 0
        while pos < len(s):
            m = _TOKEN.match(s, pos)
            if m is None or m.end() == pos:
                if s[pos:].strip() == "":
                    break
                raise ValueError(f"invalid character {s[pos]!r} in unit {unit!r}")
            pos = m.end()
            if m.group("op") == "^":
                p = _POWER.match(s, pos)
                if p is None:
                    raise ValueError(f"'^' not followed by a power in unit {unit!r}")
                toks.append(("^", float(next(g for g in p.groups() if g is not None))))
                pos = p.end()
            elif m.group("op"):
                toks.append((m.group("op"), None))
            else:
                toks.append(("name", m.group("name")))
        value, k = self._product(toks, 0, unit)
        if k != len(toks):
            raise ValueError(f"unbalanced parentheses in unit {unit!r}")
        return value

    def _product(self, toks, k, unit):
        out = 1.0
        sign = 1.0
        expect_factor = True
        while k < len(toks):
            kind, val = toks[k]
            if kind == ")":
                break
            if kind in ("*", "/"):
                if expect_factor:
                    raise ValueError(f"misplaced operator {kind!r} in unit {unit!r}")
                sign = 1.0 if kind == "*" else -1.0
                expect_factor = True
                k += 1
                continue
            if kind == "^" or not expect_factor:
                raise ValueError(f"invalid unit {unit!r}")
            if kind == "(":
                factor, k = self._product(toks, k + 1, unit)
                if k >= len(toks) or toks[k][0] != ")":
                    raise ValueError(f"unbalanced parentheses in unit {unit!r}")
                k += 1
            else:
                if val not in self.us_to_unit:
                    raise ValueError(f"unknown unit {val!r} in {unit!r}")
                factor = self.us_to_unit[val]
                k += 1
            power = 1.0
            if k < len(toks) and toks[k][0] == "^":
                power = toks[k][1]
                k += 1
            out *= factor ** (sign * power)
            sign = 1.0
            expect_factor = False
        if expect_factor:
            raise ValueError(f"incomplete unit {unit!r}")
        return out, k


# the MD unit system (md/utils.py:9)
us = UnitSystem(L="ANGSTROM", T="PS", E="KCALPERMOL")

DA = us.DA                     # mass_internal = mass_Da / DA (md/initial.py:79-108)
K_B = us.K_B                   # kcal/mol/K
FS_PER_PS = us.FS              # dt_internal = dt_fs / FS_PER_PS
THZ = us.THZ                   # gamma_internal = gamma_THz / THZ (ps^-1)

ENERGY_UNITS = {"HA": 1.0, "HARTREE": 1.0, "AU": 1.0, "EV": EV, "MEV": 1.0e3 * EV, "KCALPERMOL": KCALPERMOL,
                "KCAL/MOL": KCALPERMOL, "KJPERMOL": KJPERMOL, "KJ/MOL": KJPERMOL}


def energy_multiplier(unit: str) -> float:
    """us.get_multiplier(unit) for an energy unit: kcal/mol -> unit (md/integrate.py:21)."""
    u = unit.strip().upper()
    if u in ENERGY_UNITS:
        return ENERGY_UNITS[u] / KCALPERMOL
    return us.get_multiplier(u)


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
