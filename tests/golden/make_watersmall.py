"""Write tests/golden/watersmall.npz: the 216-water box of the reference's watersmall example.

Input data only (species, coordinates, cell) read from the reference's
examples/md/watersmall/watersmall.xyz (Tinker format: indexed, no comment
line) and the cell of examples/md/watersmall/input.fnl (18.643 A cube).
This is the box of the reference's published NVE conservation runs
(examples/md/watersmall/conservation_tests).  Run from the repository root:

    python tests/golden/make_watersmall.py
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from fennol_amd.deck import parse_deck, read_last_frame  # noqa: E402
from fennol_amd.units import SYMBOL_TO_Z  # noqa: E402

SRC = Path(os.environ.get("FENNOL_REFERENCE", "/root/reference")) / "examples" / "md" / "watersmall"


def main() -> None:
    deck = parse_deck(SRC / "input.fnl")
    sym, x = read_last_frame(SRC / deck.get("xyz_input/file"), indexed=bool(deck.get("xyz_input/indexed")),
                             has_comment_line=bool(deck.get("xyz_input/has_comment_line")))
    cell = np.asarray(deck.get("cell"), np.float64).reshape(3, 3)
    species = np.array([SYMBOL_TO_Z[s.capitalize()] for s in sym], np.int32)
    out = Path(__file__).resolve().parent / "watersmall.npz"
    np.savez_compressed(out, species=species, coordinates=np.asarray(x, np.float64), cell=cell)
    print(f"wrote {out}: {len(species)} atoms, cell {cell[0, 0]} A")


if __name__ == "__main__":
    main()
