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
    assert lib.fnl_species_mlp(None, 0, None, None, None, None, None, 0, None, None, 0, 0, 0, None, None) == -1
    # the fused step + sums: more systems than its last workgroup sums, no segments (the pointers
    # are never dereferenced: the host checks return first)
    dummy = 64
    args = [dummy] * 5 + [1.0, dummy, 0.5, 1.0, dummy, None, 1, 0, dummy, 10, 1, dummy, dummy]
    tail = [512, dummy, dummy, dummy, dummy, dummy, None]
    assert lib.fnl_baoab_step_sums(*args, 17, 17, *tail) == _abi.FNL_ERR_UNSUPPORTED
    assert lib.fnl_baoab_step_sums(*args, 1, 0, *tail) == -1
    assert lib.fnl_baoab_step_sums(*args[:-1], None, 1, 1, *tail) == -1  # no seg_offsets


def test_header_constants_match_the_bindings():
    text = HEADER.read_text()
    for name in ("FNL_REDUCE_VALUES", "FNL_REDUCE_FUSED_MAX_SYSTEMS", "FNL_POOL_HEADER_WORDS"):
        m = re.search(rf"#define {name} (\d+)", text)
        assert m and int(m.group(1)) == getattr(_abi, name), name


def _c_layout(struct: str, fields: list[str], tmp_path) -> dict:
    """sizeof and offsetof of a header struct, from the host C++ compiler."""
    import shutil
    import subprocess

    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    lines = [f'#include "{HEADER}"', "#include <cstddef>", "#include <cstdio>", "int main() {",
             f'  std::printf("size %zu\\n", sizeof({struct}));']
    lines += [f'  std::printf("{f} %zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cxx, "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


@pytest.mark.parametrize("name", ["Pool", "PoolRequest", "MlpWeights", "MlpNet", "AevParams", "TrainNet"])
def test_ctypes_structs_match_the_header(name, tmp_path):
    """Every ctypes structure has the size and field offsets of its C declaration."""
    import ctypes as C

    cls = getattr(_abi, name)
    cname = {"Pool": "fnl_pool", "PoolRequest": "fnl_pool_request", "MlpWeights": "fnl_mlp_weights",
             "MlpNet": "fnl_mlp_net", "AevParams": "fnl_aev_params", "TrainNet": "fnl_train_net"}[name]
    fields = [f for f, _ in cls._fields_]
    lay = _c_layout(cname, fields, tmp_path)
    assert lay["size"] == C.sizeof(cls)
    for f in fields:
        assert lay[f] == getattr(cls, f).offset, f


def test_serving_record_size_matches_the_library():
    if not _abi.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    from fennol_amd.serving import REQ_DTYPE, out_bytes

    lib = _abi.load()
    for args in [(0, 0, 0), (1, 1, 9), (7, 3, 100), (2048, 2048, 65536)]:
        assert lib.fnl_pool_out_bytes(*args) == out_bytes(*args)
    assert REQ_DTYPE.itemsize == __import__("ctypes").sizeof(_abi.PoolRequest)
    assert [REQ_DTYPE.fields[f][1] for f, _ in _abi.PoolRequest._fields_] == \
        [getattr(_abi.PoolRequest, f).offset for f, _ in _abi.PoolRequest._fields_]
