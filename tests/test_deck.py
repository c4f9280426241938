"""fennol_md-style decks: the unit system, the parser, the run plan and the xyz reader (CPU)."""

from __future__ import annotations

import glob
import math
import os

import numpy as np
import pytest

from fennol_amd import units
from fennol_amd.deck import parse_deck, plan_deck, read_last_frame, unit_multiplier, wrap_into_box

REFERENCE_DECKS = sorted(glob.glob("/root/reference/examples/md/*/input.fnl"))


DECK = """\
device cuda:0   # ignored here
model_file model.fnx
xyz_input{
  file water.xyz
  indexed yes
  has_comment_line no
}
cell = 6.2 0. 0. 0. 6.2 0. 0. 0. 6.2
minimum_image no
nsteps = 20
dt[fs] =  .5
tdump[fs] = 2.5
gamma[THz] = 10.
temperature = 300
nblist_skin 1.
thermostat LGV
traj_format xyz
print_timings
! a comment line
nprint = 5
&qtb
  omegacut[cm1] = 15000.
  tseg{ps} = 0.25
&end
empty{}
"""


def test_parse_deck_units_sections_and_types(tmp_path):
    f = tmp_path / "input.fnl"
    f.write_text(DECK)
    d = parse_deck(f)
    assert d.get("model_file") == "model.fnx"
    assert d.get("xyz_input/file") == "water.xyz"
    assert d.get("xyz_input/indexed") is True and d.get("xyz_input/has_comment_line") is False
    assert d.get("cell") == [6.2, 0.0, 0.0, 0.0, 6.2, 0.0, 0.0, 0.0, 6.2]
    assert d.get("minimum_image") is False
    assert d.get("nsteps") == 20 and isinstance(d.get("nsteps"), int)
    assert d.get("dt") == pytest.approx(5e-4)        # ps
    assert d.get("tdump") == pytest.approx(2.5e-3)   # ps
    assert d.get("gamma") == pytest.approx(10.0)     # 1/ps
    assert d.get("print_timings") is True            # keyword alone
    assert d.get("thermostat") == "LGV"
    assert d.get("NPRINT") == 5                      # case-insensitive keys
    # &name ... &end section, braces unit, cm-1 (an angular frequency in 1/ps)
    assert d.get("qtb/omegacut") == pytest.approx(15000.0 / unit_multiplier("cm1"))
    assert d.get("qtb/tseg") == pytest.approx(0.25)
    assert d.get("empty") == {}


def test_unit_system_against_the_reference_docstring_values():
    # utils/atomic_units.py:164-176 (UnitSystem docstring) and :284-289 (get_multiplier)
    au = units.UnitSystem(L="BOHR", T="AU_T", E="HARTREE")
    assert au.get_multiplier("ANGSTROM") == pytest.approx(0.52917721)
    assert au.get_multiplier("EV") == pytest.approx(27.211386024367243)
    assert au.FS == pytest.approx(2.4188843e-2)
    u2 = units.UnitSystem(L="ANGSTROM", E="EV", M="DA")
    assert u2.BOHR == pytest.approx(1.889725989, rel=1e-7)  # the docstring rounds 1/0.52917721
    assert u2.fs == pytest.approx(10.217477, rel=1e-6)


def test_unit_strings_follow_dimensional_analysis():
    us = units.us
    # products, quotients (the '/' applies to the next factor only), powers, parentheses
    ev, ang, fs = us.get_multiplier("EV"), us.get_multiplier("ANGSTROM"), us.get_multiplier("FS")
    assert ang == 1.0 and us.get_multiplier("ps") == 1.0 and fs == 1000.0
    assert us.get_multiplier("EV/ANGSTROM") == pytest.approx(ev)
    assert us.get_multiplier("EV*ANGSTROM^{2}/FS^{3}") == pytest.approx(ev / fs ** 3)
    assert us.get_multiplier("(EV*FS)^2") == pytest.approx((ev * fs) ** 2)
    assert us.get_multiplier("EV/(FS*FS)") == pytest.approx(ev / fs ** 2)
    assert us.get_multiplier("FS^-1") == pytest.approx(1e-3)
    assert us.get_multiplier("NM^(2.5)") == pytest.approx(0.1 ** 2.5)
    # a wavenumber is an angular frequency: 1 cm-1 = 2 pi c x 100 /m; 1/ps = 5.3088 cm-1
    c = 299792458.0
    # (the reference's constant 219471.52 cm-1 per atomic unit is 1.4e-5 below CODATA's)
    assert us.get_multiplier("CM1") == pytest.approx(1.0e12 / (2 * math.pi * c * 100.0), rel=3e-5)
    # kcal/mol/A^3 in atm: 4184 J/mol / N_A / 1e-30 m^3 / 101325 Pa
    assert us.get_multiplier("ATM") == pytest.approx(4184.0 / 6.02214129e23 / 1e-30 / 101325.0, rel=2e-5)
    # the mass unit kcal/mol ps^2 / A^2 is 4184e-4 kg/mol = 418.4 g/mol; the reference's Dalton is
    # 1/MPROT electron masses (the proton's, atomic_units.py:97), so us.DA = 418.4 m_u / m_p
    assert units.DA == pytest.approx(418.4 * 1822.888486 / units.MPROT, rel=1e-6)
    with pytest.raises(ValueError):
        unit_multiplier("furlong")
    with pytest.raises(ValueError):
        unit_multiplier("EV*")
    with pytest.raises(ValueError):
        unit_multiplier("(EV")


def test_unit_multipliers():
    assert unit_multiplier("fs") == 1000.0 and unit_multiplier("ps") == 1.0
    assert unit_multiplier("eV") == pytest.approx(27.211386024367243 / 627.5096080305927)
    assert unit_multiplier("bohr") == pytest.approx(1.0 / 0.52917721)
    assert unit_multiplier("THz") == pytest.approx(1.0)


def test_plan_refuses_only_active_features(tmp_path):
    f = tmp_path / "input.fnl"
    f.write_text(DECK)
    plan = plan_deck(parse_deck(f), tmp_path)
    assert plan.thermostat == "LGV" and plan.dt_fs == pytest.approx(0.5) and plan.gamma_thz == pytest.approx(10.0)
    assert plan.minimum_image is False and plan.nblist_skin == 1.0 and plan.tdump_ps == pytest.approx(2.5e-3)
    for extra, word in (("thermostat ADQTB\n", "thermostat"), ("barostat LGV\n", "barostat"), ("nbeads 8\n", "nbeads"),
                        ("ir_spectrum yes\n", "ir_spectrum"), ("hmr 2\n", "hmr")):
        g = tmp_path / "bad.fnl"
        g.write_text(DECK + extra)
        with pytest.raises(NotImplementedError, match=word):
            plan_deck(parse_deck(g), tmp_path)
    g = tmp_path / "ok.fnl"
    g.write_text(DECK + "barostat NONE\nwrap_box yes\nestimate_pressure yes\npressure_unit bar\n"
                 "double_precision\nnsummary 100\n".replace("pressure_unit bar", "pressure_unit kbar"))
    plan = plan_deck(parse_deck(g), tmp_path)
    assert plan.wrap_box and plan.estimate_pressure and plan.pressure_unit == "kbar"


@pytest.mark.skipif(not REFERENCE_DECKS, reason="the reference's example decks are not present")
@pytest.mark.parametrize("path", REFERENCE_DECKS, ids=lambda p: os.path.basename(os.path.dirname(p)))
def test_reference_example_decks_parse_and_plan(path):
    """Every examples/md/*/input.fnl of the reference parses and plans: their inactive qtb{}
    blocks (omegacut[cm1]) no longer stop them."""
    base = os.path.dirname(path)
    d = parse_deck(path)
    plan = plan_deck(d, __import__("pathlib").Path(base))
    assert plan.thermostat == "LGV"
    assert plan.dt_fs == pytest.approx(0.5)
    assert plan.model_file.name == "ani2x.fnx"
    assert plan.xyz_file.is_file()
    if d.get("qtb/omegacut") is not None:
        assert d.get("qtb/omegacut") == pytest.approx(15000.0 / unit_multiplier("cm1"))
    name = os.path.basename(base)
    if name == "aspirin":
        assert plan.cell is None and plan.energy_unit == "Ha" and plan.per_atom_energy is False
    else:
        assert plan.cell is not None and plan.nblist_skin == 2.0
        assert plan.minimum_image is (name != "watertiny")
    if name == "watersmall":
        assert plan.cell[0, 0] == pytest.approx(18.643) and plan.gamma_thz == pytest.approx(10.0)
        sym, x = read_last_frame(plan.xyz_file, plan.indexed, plan.has_comment_line)
        assert len(sym) == 648 and sym[:3] == ["O", "H", "H"]


def test_read_last_frame_plain_and_indexed(tmp_path):
    plain = tmp_path / "a.xyz"
    plain.write_text("2\nfirst\nO 0 0 0\nH 1 0 0\n2\nsecond\nO 0 0 1\nH 1 0 1\n")
    sym, x = read_last_frame(plain)
    assert sym == ["O", "H"]
    np.testing.assert_array_equal(x, [[0, 0, 1], [1, 0, 1]])
    arc = tmp_path / "b.arc"
    arc.write_text("     2  water\n  6.2 6.2 6.2 90. 90. 90.\n     1  O  0.0 0.0 0.0  1  2\n"
                   "     2  H  0.9 0.0 0.0  2  1\n")
    sym, x = read_last_frame(arc, indexed=True, has_comment_line=False)
    assert sym == ["O", "H"]
    np.testing.assert_allclose(x, [[0, 0, 0], [0.9, 0, 0]])


def test_wrap_into_box():
    cell = np.array([[10.0, 0, 0], [2.0, 9.0, 0], [0, 1.0, 8.0]])
    rng = np.random.default_rng(0)
    x = rng.normal(scale=30.0, size=(50, 3))
    w = wrap_into_box(x, cell)
    q = w @ np.linalg.inv(cell)
    assert np.all(q >= -1e-12) and np.all(q < 1 + 1e-12)
    np.testing.assert_allclose(np.round((x - w) @ np.linalg.inv(cell)), (x - w) @ np.linalg.inv(cell), atol=1e-9)
