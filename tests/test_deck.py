"""fennol_md-style decks: the parser and the xyz reader (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

from fennol_amd.deck import parse_deck, read_last_frame, unit_multiplier


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


def test_unit_multipliers():
    assert unit_multiplier("fs") == 1000.0 and unit_multiplier("ps") == 1.0
    assert unit_multiplier("eV") == pytest.approx(27.211386024367243 / 627.5096080305927)
    assert unit_multiplier("bohr") == pytest.approx(1.0 / 0.52917721)
    with pytest.raises(ValueError):
        unit_multiplier("furlong")


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
