"""Compile the gfx950 HIP sources into ``fennol_amd/_lib/libfnl_hip.so``.

Usage: ``python -m fennol_amd.build`` (also called by ``__graft_entry__.build``).
Every call recompiles.  hipcc cross-compiles for gfx950 without a GPU.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIBNAME = "libfnl_hip.so"
SOURCES = ["nblist.hip", "aev.hip", "mlp.hip", "md.hip", "serving.hip", "train.hip"]
ARCH = os.environ.get("FNL_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (v_pk_* FP32 forms are avoided on this target)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return exe


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; returns its path."""
    hipcc = _hipcc()
    objdir = LIBDIR / "obj"
    objdir.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        jobs.append((cmd, obj))
    with cf.ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in jobs]
        for f in futs:
            f.result()
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o in jobs]])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
