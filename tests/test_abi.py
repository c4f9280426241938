"""CPU checks of the C ABI: the library loads and exports every declared entry point."""

from __future__ import annotations

import re
from pathlib import Path

import pytest

from fennol_amd import _abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "fnl_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fnl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_functions():
    assert set(declared_functions()) == set(_abi.EXPORTED)


def test_library_loads_and_exports_all_symbols():
    if not _abi.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.fnl_version() >= 10000
    assert lib.fnl_strerror(-1).decode() == "invalid argument"


def test_host_side_argument_checks():
    """Calls rejected before any GPU work (null pointers, bad sizes)."""
    if not _abi.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    lib = _abi.load()
    assert lib.fnl_nblist_workspace(10, 0, 1) == 0
    assert lib.fnl_nblist_workspace(100, 1, 8) > 0
    assert lib.fnl_mlp_tile_rows(_abi.FNL_F32) == 32
    assert lib.fnl_mlp_tile_rows(_abi.FNL_BF16) == 96
    assert lib.fnl_nblist_build(None, None, None, None, 10, 1, 0, None, None, None, 1, 5.2, 3.5, 8, 8,
                                None, 0, None, None, None, None, None, None, None) == -1
    # general PBC without a shift output, skin update with an unknown mode
    assert lib.fnl_nblist_build(None, None, None, None, 10, 1, _abi.FNL_PBC_GENERAL, None, None, None, 1, 5.2,
                                3.5, 8, 8, None, 0, None, None, None, None, None, None, None) == -1
    assert lib.fnl_nblist_skin_update(None, None, 10, 7, None, None, None, None, None, 8, 5.2, 3.5, 8, 8, None,
                                      None, None, None, None, None, None) == -1
    p = _abi.AevParams()
    p.n_species = 9  # above the maximum
    assert lib.fnl_aev_forward(None, None, None, 0, 0, None, None, None, None, 8, None, None, None, 8, p, None, 0,
                               0, 0xFFFFFFFF, None) == -1
    assert lib.fnl_fill_i32(None, 0, 0, None) == 0
    assert lib.fnl_fill_i32(None, 0, 5, None) == -1
    assert lib.fnl_species_mlp(None, 0, None, None, None, None, None, 0, None, None, 0, 0, 0, None) == -1
