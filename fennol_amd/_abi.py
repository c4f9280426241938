"""ctypes binding of the C ABI declared in ``include/fnl_hip.h``.

The shared library is built in-tree by :mod:`fennol_amd.build`.  Loading fails
loudly when it is missing: there is no CPU fallback on the product path.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libfnl_hip.so"

FNL_OK = 0
FNL_ERR_INVALID = -1
FNL_ERR_HIP = -2
FNL_ERR_UNSUPPORTED = -3
FNL_REDUCE_FUSED_MAX_SYSTEMS = 16  # systems fnl_baoab_step_sums sums in its own launch
FNL_ERR_SCRATCH_TOO_SMALL = -4
FNL_NBLIST_SCRATCH_BYTES = 64 << 20
FNL_PBC_NONE = 0
FNL_PBC_MINIMUM_IMAGE = 1
FNL_PBC_GENERAL = 2
FNL_F32 = 0
FNL_BF16 = 1
FNL_STATUS_WORDS = 8
FNL_STATUS_OVERFLOW = 0
FNL_STATUS_MAX_NBR = 1
FNL_STATUS_MAX_NBR_ANG = 2
FNL_STATUS_MAX_CELL = 3
FNL_STATUS_STICKY = 4
FNL_STATUS_STICKY_MAX_NBR = 5
FNL_STATUS_STICKY_MAX_NBR_ANG = 6
FNL_MAX_RADIAL = 32
FNL_MAX_ANG_DIST = 16
FNL_MAX_ANG_SECT = 16
FNL_MAX_SPECIES = 8
FNL_MLP_H1, FNL_MLP_H2, FNL_MLP_H3 = 256, 192, 160
FNL_MLP_MAX_NETS = 8
FNL_REDUCE_VALUES = 19

vp = C.c_void_p
i32 = C.c_int32
u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
f64 = C.c_double
sz = C.c_size_t


class AevParams(C.Structure):
    _fields_ = [
        ("n_species", i32),
        ("n_radial", i32),
        ("n_ang_dist", i32),
        ("n_ang_sect", i32),
        ("cutoff", f32),
        ("cutoff_ang", f32),
        ("radial_eta", f32),
        ("angular_eta", f32),
        ("zeta", f32),
        ("radial_shifts", f32 * FNL_MAX_RADIAL),
        ("ang_shifts", f32 * FNL_MAX_ANG_DIST),
        ("ang_cos", f32 * FNL_MAX_ANG_SECT),
        ("ang_sin", f32 * FNL_MAX_ANG_SECT),
    ]


class MlpNet(C.Structure):
    _fields_ = [(n, vp) for n in ("w1t", "w2t", "w3t", "w1", "w2", "w3", "w4", "b1", "b2", "b3", "b4",
                                  "w1t_lo", "w2t_lo", "w3t_lo", "w1_lo", "w2_lo", "w3_lo")]


class MlpWeights(C.Structure):
    _fields_ = [
        ("n_nets", i32),
        ("k_pad", i32),
        ("dtype", i32),
        ("celu_alpha", f32),
        ("tile_rows", i32),
        ("nets", MlpNet * FNL_MLP_MAX_NETS),
    ]


class TrainNet(C.Structure):
    _fields_ = [(n, vp) for n in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4")]


FNL_POOL_N_ATOMS = 0
FNL_POOL_N_SYS = 1
FNL_POOL_N_TILES = 2
FNL_POOL_ERROR = 3
FNL_POOL_COUNTS = 8
FNL_POOL_HEADER_WORDS = 16


class PoolRequest(C.Structure):
    _fields_ = [
        ("n_atoms", i32),
        ("src_row", i32),
        ("stage_row", i32),
        ("md_left", i32),
        ("evaluated", i32),
        ("noise_key", u32),
        ("vel_key", u32),
        ("reserved", u32),
        ("step", u64),
    ]


POOL_TABLE_FIELDS = ("z_species", "z_net", "z_dtm", "z_a2", "z_vscale")
POOL_ROW_FIELDS = ("x", "v", "x32", "z", "grad_self", "grad_nbr", "e_atom", "virial_atom", "dtm", "a2",
                   "old_x", "old_v", "old_grad_self", "old_grad_nbr", "old_z", "batch_index", "species_idx", "perm")
POOL_SYS_FIELDS = ("sys_offsets", "seg_offsets", "grid_cap", "tile_start", "tile_rows_tab", "tile_net",
                   "sys_left", "sys_eval", "sys_step", "sys_key", "sys_fin", "sys_out_row", "counts")


class Pool(C.Structure):
    _fields_ = ([(n, i32) for n in ("max_atoms", "max_sys", "max_tiles", "n_nets", "tile_rows", "seg_atoms",
                                    "n_z", "stage_rows")]
                + [("seed", u64)]
                + [(n, vp) for n in POOL_TABLE_FIELDS + POOL_ROW_FIELDS + POOL_SYS_FIELDS])


# name -> (restype, argtypes)
_SIGNATURES = {
    "fnl_version": (C.c_int, []),
    "fnl_strerror": (C.c_char_p, [C.c_int]),
    "fnl_nblist_workspace": (sz, [i32, i32, i32]),
    "fnl_nblist_build": (C.c_int, [vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, i32, f64, f64, i32, i32,
                                   vp, sz, vp, vp, vp, vp, vp, vp, vp]),
    "fnl_nblist_build_noalloc": (C.c_int, [vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, i32, f64, f64, i32, i32,
                                           vp, vp, vp, vp, vp, vp, vp]),
    "fnl_nblist_skin_update": (C.c_int, [vp, vp, i32, i32, vp, vp, vp, vp, vp, i32, f64, f64, i32, i32, vp, vp,
                                         vp, vp, vp, vp, vp]),
    "fnl_nblist_export": (C.c_int, [vp, vp, i32, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp]),
    "fnl_aev_forward": (C.c_int, [vp, vp, vp, i32, i32, vp, vp, vp, vp, i32, vp, vp, vp, i32,
                                  C.POINTER(AevParams), vp, i32, i32, u32, vp]),
    "fnl_aev_backward": (C.c_int, [vp, vp, vp, i32, i32, vp, vp, vp, vp, i32, vp, vp, vp, i32,
                                   C.POINTER(AevParams), vp, i32, i32, vp, vp, vp, u32, vp]),
    "fnl_aev_tangent": (C.c_int, [vp, vp, vp, vp, i32, i32, vp, vp, vp, vp, i32, vp, vp, vp, i32,
                                  C.POINTER(AevParams), vp, i32, vp]),
    "fnl_mlp_tile_rows": (C.c_int, [i32]),
    "fnl_species_mlp": (C.c_int, [vp, i32, C.POINTER(MlpWeights), vp, vp, vp, vp, i32, vp, vp, i32, i32, u64,
                                  vp, vp]),
    "fnl_system_reduce": (C.c_int, [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "fnl_forces": (C.c_int, [vp, vp, i32, f32, vp, vp]),
    "fnl_baoab_step": (C.c_int, [vp, vp, vp, vp, vp, f32, vp, f64, f64, vp, vp, u64, u64, vp, i32, i32, vp]),
    "fnl_baoab_step_sums": (C.c_int, [vp, vp, vp, vp, vp, f32, vp, f64, f64, vp, vp, u64, u64, vp, i32, i32,
                                      vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp]),
    "fnl_counter_add": (C.c_int, [vp, u64, vp]),
    "fnl_baoab_step_batched": (C.c_int, [vp, vp, vp, vp, vp, f32, vp, f64, f64, vp, vp, vp, vp, vp, u64, i32, vp]),
    "fnl_philox_normals": (C.c_int, [vp, i32, u64, u64, u32, vp]),
    "fnl_train_workspace": (sz, [i32, i32, i32]),
    "fnl_train_grad_stride": (sz, [i32]),
    "fnl_train_mlp_grad": (C.c_int, [vp, i32, vp, i32, i32, C.POINTER(TrainNet), i32, C.POINTER(i32), vp, vp, f32,
                                     f32, u64, vp, sz, vp, vp]),
    "fnl_train_residuals": (C.c_int, [vp, vp, i32, vp, vp, f32, vp, vp, vp, i32, f64, f64, f64, f64, vp, vp, vp, vp,
                                      vp, vp]),
    "fnl_pool_out_bytes": (sz, [i32, i32, i32]),
    "fnl_pool_save_rows": (C.c_int, [C.POINTER(Pool), i32, vp]),
    "fnl_pool_layout": (C.c_int, [C.POINTER(Pool), vp, i32, vp, vp, vp]),
    "fnl_pool_integrate": (C.c_int, [C.POINTER(Pool), f32, f64, f64, i32, vp]),
    "fnl_pool_finish": (C.c_int, [C.POINTER(Pool), vp, vp, vp, i32, i32, vp, f32, i32, vp, sz, vp]),
    "fnl_fill_i32": (C.c_int, [vp, i32, sz, vp]),
    "fnl_fill_f32": (C.c_int, [vp, f32, sz, vp]),
}

EXPORTED = tuple(_SIGNATURES.keys())

_lib = None


class FnlError(RuntimeError):
    pass


def load(path: Path | str | None = None):
    """Load (once) and return the ctypes library handle; raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"fennol_amd HIP library not found at {p}; build it with `python -m fennol_amd.build`"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != FNL_OK:
        msg = load().fnl_strerror(rc).decode()
        raise FnlError(f"{what} failed: {msg} ({rc})")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
