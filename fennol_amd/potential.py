"""Device-resident ANI potential: neighbour lists -> AEV -> species MLP -> forces.

This is the engine behind the FeNNol-shaped modules of :mod:`fennol_amd.fennix`.
It owns the device buffers of one batch of systems (:class:`SystemBuffers`) and
launches the HIP kernels of ``libfnl_hip.so`` on the current torch stream.

Reference path (FeNNol src/fennol): GraphGenerator / GraphFilter /
GraphAngularExtension / SpeciesIndexer (models/preprocessing.py), ANIAEV
(models/embeddings/ani.py:47-142), SpeciesIndexNet (models/misc/nets.py:
301-426), total energy and gradients (models/fennix.py:216-347).
"""

from __future__ import annotations

import contextlib
import functools
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from .units import PERIODIC_TABLE, SYMBOL_TO_Z

CELU_ALPHA = 0.1
SEG_ATOMS = 512  # atoms per partial-sum segment of the per-system reductions


@dataclass(frozen=True)
class ANIConfig:
    """ANI_AEV hyper-parameters (models/embeddings/ani.py:18-41) plus the graph cutoffs."""

    species_order: tuple = ("H", "C", "N", "O")
    cutoff: float = 5.2
    cutoff_angular: float = 3.5
    radial_eta: float = 19.7
    angular_eta: float = 12.5
    radial_dist_divisions: int = 16
    angular_dist_divisions: int = 4
    zeta: float = 14.1
    angle_sections: int = 8
    radial_start: float = 0.8
    angular_start: float = 0.8

    @property
    def n_species(self) -> int:
        return len(self.species_order)

    @property
    def n_pairs(self) -> int:
        return self.n_species * (self.n_species + 1) // 2

    @property
    def aev_length(self) -> int:
        return (self.n_species * self.radial_dist_divisions
                + self.n_pairs * self.angular_dist_divisions * self.angle_sections)

    @functools.lru_cache(maxsize=256)
    def live_column_blocks(self, species_mask: int) -> int:
        """Bit q set if AEV columns [32q, 32q + 32) hold a channel block written for a batch whose
        AEV species indices are the set bits of ``species_mask`` (radial blocks of those species,
        angular blocks of pairs of them); 0 (= every column) for the full mask.  The MLP and the
        training kernels skip the products of the other columns, which are zero in every row of
        such a batch (columns from 2048 on are always treated as live)."""
        if species_mask == 0xFFFFFFFF or species_mask == 0:
            return 0
        R = self.radial_dist_divisions
        A = self.angular_dist_divisions * self.angle_sections
        spans = []
        pr = 0
        for a in range(self.n_species):
            if (species_mask >> a) & 1:
                spans.append((a * R, (a + 1) * R))
        for a in range(self.n_species):
            for b in range(a, self.n_species):  # pair_index order: row-major upper triangle
                if (species_mask >> a) & 1 and (species_mask >> b) & 1:
                    c0 = self.n_species * R + pr * A
                    spans.append((c0, c0 + A))
                pr += 1
        live = 0
        for c0, c1 in spans:
            for q in range(c0 // 32, (c1 - 1) // 32 + 1):
                if q < 64:
                    live |= 1 << q
        return live or 1

    @property
    def k_pad(self) -> int:
        """AEV row length padded to the MLP kernel's K chunk (128)."""
        return ((self.aev_length + 127) // 128) * 128

    def c_params(self) -> _abi.AevParams:
        """Kernel parameters; shifts computed in float64 and rounded to float32 (ani.py:73-110)."""
        p = _abi.AevParams()
        p.n_species = self.n_species
        p.n_radial = self.radial_dist_divisions
        p.n_ang_dist = self.angular_dist_divisions
        p.n_ang_sect = self.angle_sections
        p.cutoff = self.cutoff
        p.cutoff_ang = self.cutoff_angular
        p.radial_eta = self.radial_eta
        p.angular_eta = self.angular_eta
        p.zeta = self.zeta
        rs = np.linspace(self.radial_start, self.cutoff, self.radial_dist_divisions + 1)[:-1]
        asf = np.linspace(self.angular_start, self.cutoff_angular, self.angular_dist_divisions + 1)[:-1]
        tz = (np.linspace(0, np.pi, self.angle_sections + 1) + np.pi / (2 * self.angle_sections))[:-1]
        for k, v in enumerate(rs):
            p.radial_shifts[k] = float(v)
        for k, v in enumerate(asf):
            p.ang_shifts[k] = float(v)
        for k, v in enumerate(tz):
            p.ang_cos[k] = math.cos(v)
            p.ang_sin[k] = math.sin(v)
        return p

    def species_map(self) -> np.ndarray:
        """Atomic number -> AEV species index, 0 for elements not listed (ani.py:52-60)."""
        conv = np.zeros(len(PERIODIC_TABLE) + 1, dtype=np.int32)
        for i, s in enumerate(self.species_order):
            conv[SYMBOL_TO_Z[s]] = i
        return conv


def _status_errors(st) -> None:
    """Raise for the status bits that regrowing the rows cannot fix."""
    if st[_abi.FNL_STATUS_OVERFLOW] & 2:
        raise RuntimeError("cell grid capacities inconsistent with the system")
    if st[_abi.FNL_STATUS_OVERFLOW] & 4:
        raise RuntimeError("an atom lies more than 511 cells from its cell's origin (general PBC shift range)")


def grid_caps(cutoff: float, natoms, cells) -> tuple[np.ndarray, int]:
    """Per-system cell-grid capacities (cap_x, cap_y, cap_z, first_cell) for fnl_nblist_build and
    the total cell count.  Periodic systems get room for cells half a cutoff wide (the build
    searches them two deep); open systems about one cell per atom.  Systems without atoms (the
    padding system of AtomPadding, whose cell is 1000 A wide) get one cell."""
    B = len(natoms)
    g = np.zeros((B, 4), np.int32)
    first = 0
    rbin = cutoff * 1.0001 + 1.0e-4
    for b in range(B):
        if int(natoms[b]) == 0:
            caps = np.array([1, 1, 1])
        elif cells is not None:
            R = np.linalg.inv(np.asarray(cells[b], np.float64))
            spacing = 1.0 / np.sqrt((R ** 2).sum(axis=0))
            caps = np.maximum(np.floor(2.0 * spacing / rbin).astype(int), 1) + 1
        else:
            k = max(1, int(math.ceil(max(int(natoms[b]), 1) ** (1.0 / 3.0))))
            caps = np.array([k, k, k])
        g[b, :3] = caps
        g[b, 3] = first
        first += int(np.prod(caps))
    return g, max(first, 1)


_CUS = []


def _n_cus() -> int:
    """Compute units of the current device (256 on an MI355X)."""
    if not _CUS:
        _CUS.append(torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
                    if torch.cuda.is_available() else 256)
    return _CUS[0]


def _next_mult(x: int, m: int) -> int:
    return ((x + m - 1) // m) * m


def fragment_order(a) -> np.ndarray:
    """[R][C] matrix (R % 16 == 0, C % 32 == 0) in the BF16 MLP kernel's fragment order: blocks of
    16 rows x 32 columns, block (r/16, c/32) at element offset ((r/16)*(C/32) + c/32)*512, inside
    which element (r, c) sits at ((c%32)/8)*128 + (r%16)*8 + c%8 (one MFMA A operand per block,
    each lane's 8 consecutive columns at lane*8)."""
    a = np.asarray(a)
    R, C = a.shape
    if R % 16 or C % 32:
        raise ValueError(f"fragment order needs rows % 16 == 0 and columns % 32 == 0, got {a.shape}")
    return np.ascontiguousarray(a.reshape(R // 16, 16, C // 32, 4, 8).transpose(0, 2, 3, 1, 4)).reshape(R, C)


class SpeciesNetworks:
    """Per-species MLP weights on the device in both GEMM layouts (BF16: in fragment order).

    ``params``: {symbol: [(W [in, out], b [out]) x 4]} with widths
    A -> 256 -> 192 -> 160 -> 1 (FullyConnectedNet layers, nets.py:81-94).

    BF16 with ``compensated``: each hidden-layer matrix is stored as hi = bf16(W) and
    lo = bf16(W - hi), and a product that takes both on the MFMA sees weights of about 16
    significant bits.  Plain BF16 weights (bf16(W) only, half the MFMA work and weight traffic)
    bias every atom of a similar environment the same way, an energy error that grows with the
    system's energy scale: -0.08 meV/atom on a 216-water box with random weights against -0.004
    compensated, and along a 1,000-step trajectory of a 23,328-atom box 0.09 -> 1.7 meV/atom
    against 0.0006 -> 0.007.  ``compensated="forward"`` (the default): hi + lo in the forward
    products, hi only in the backward products that form dE/dAEV -- the same energies as
    ``True`` (the forward is the same arithmetic) and forces within a few per cent of its error
    (force RMS vs FP64 at step 1,000 of that trajectory 7.3e-3 against 7.1e-3 eV/A; plain
    9.7e-3), for a quarter less MFMA work.
    """

    def __init__(self, params: dict, k_in: int, k_pad: int, dtype: str = "fp32", device="cuda",
                 celu_alpha: float = CELU_ALPHA, tile_rows: int = 0, compensated: bool | str = "forward"):
        if dtype not in ("fp32", "bf16"):
            raise ValueError("dtype must be 'fp32' or 'bf16'")
        if len(params) > _abi.FNL_MLP_MAX_NETS:
            raise ValueError("too many networks")
        if compensated not in (True, False, "forward"):
            raise ValueError("compensated must be True, False or 'forward'")
        self.species = list(params.keys())
        self.dtype = dtype
        self.compensated = bool(compensated) and dtype == "bf16"  # forward products
        self.compensated_backward = compensated is True and dtype == "bf16"
        self.k_pad = k_pad
        tdt = torch.float32 if dtype == "fp32" else torch.bfloat16
        self._keep = []
        self.cw = _abi.MlpWeights()
        self.cw.n_nets = len(self.species)
        self.cw.k_pad = k_pad
        self.cw.dtype = _abi.FNL_F32 if dtype == "fp32" else _abi.FNL_BF16
        self.cw.celu_alpha = celu_alpha
        dims = [k_in, _abi.FNL_MLP_H1, _abi.FNL_MLP_H2, _abi.FNL_MLP_H3, 1]
        for n, s in enumerate(self.species):
            layers = params[s]
            if len(layers) != 4:
                raise ValueError("networks must have 3 hidden layers (256, 192, 160) and a scalar output")
            for l, (W, b) in enumerate(layers):
                if tuple(np.shape(W)) != (dims[l], dims[l + 1]) or np.shape(b) != (dims[l + 1],):
                    raise ValueError(f"layer {l + 1} of network {s} has shape {np.shape(W)}, expected "
                                     f"{(dims[l], dims[l + 1])}")
            W1 = np.zeros((k_pad, dims[1]))
            W1[:k_in] = np.asarray(layers[0][0])
            mats = [W1, np.asarray(layers[1][0]), np.asarray(layers[2][0])]
            net = self.cw.nets[n]
            for l, Wm in enumerate(mats):
                if dtype == "bf16":  # the kernel's MFMA operand layout
                    w = torch.tensor(fragment_order(Wm), dtype=tdt, device=device).contiguous()
                    wt = torch.tensor(fragment_order(Wm.T), dtype=tdt, device=device).contiguous()
                    if self.compensated:
                        W64 = torch.tensor(Wm, dtype=torch.float64)
                        lo = (W64 - W64.to(tdt).to(torch.float64)).numpy()  # exact in float64
                        w_lo = torch.tensor(fragment_order(lo), dtype=tdt, device=device).contiguous()
                        wt_lo = torch.tensor(fragment_order(lo.T), dtype=tdt, device=device).contiguous()
                        self._keep += [w_lo, wt_lo]
                        setattr(net, f"w{l + 1}t_lo", wt_lo.data_ptr())  # forward: W^T
                        if self.compensated_backward:
                            setattr(net, f"w{l + 1}_lo", w_lo.data_ptr())
                else:
                    w = torch.tensor(Wm, dtype=tdt, device=device).contiguous()
                    wt = torch.tensor(Wm.T, dtype=tdt, device=device).contiguous()
                self._keep += [w, wt]
                setattr(net, f"w{l + 1}", w.data_ptr())
                setattr(net, f"w{l + 1}t", wt.data_ptr())
            w4 = torch.tensor(np.asarray(layers[3][0])[:, 0], dtype=torch.float32, device=device)
            bs = [torch.tensor(np.asarray(layers[l][1]), dtype=torch.float32, device=device) for l in range(4)]
            self._keep += [w4, *bs]
            net.w4 = w4.data_ptr()
            net.b1, net.b2, net.b3, net.b4 = (x.data_ptr() for x in bs)
        self.tile_rows = tile_rows or _abi.load().fnl_mlp_tile_rows(self.cw.dtype)
        self.cw.tile_rows = self.tile_rows
        # BF16, small systems (fewer tiles than half the CUs): 32-atom tiles, since one tile's
        # latency is then the MLP's time (every tile size gives the same results; measured on 648
        # atoms: 0.104 vs 0.116 ms per MD step; FP32 16-atom tiles were slower than 32); a tile
        # size given by the caller is kept
        self.small_tile_rows = 32 if (dtype == "bf16" and not tile_rows) else 0


@dataclass
class SystemBuffers:
    """Device state of one batch of systems with fixed composition."""

    n_atoms: int
    n_sys: int
    pbc: int
    species: np.ndarray                 # atomic numbers (host copy)
    natoms: np.ndarray
    sys_offsets: torch.Tensor           # int32 [B+1]
    batch_index: torch.Tensor           # int32 [N]
    species_idx: torch.Tensor           # int32 [N] AEV species index, -1 = padding
    cells: torch.Tensor | None          # float32 [B,3,3]
    rcells: torch.Tensor | None
    grid: torch.Tensor                  # int32 [B,4]
    n_cells: int
    max_nbr: int
    max_nbr_ang: int
    nbr: torch.Tensor = None
    n_nbr: torch.Tensor = None
    nbr_ang: torch.Tensor = None
    n_nbr_ang: torch.Tensor = None
    nbr_shift: torch.Tensor = None      # int32 [N, max_nbr] packed cell shifts (general PBC only)
    status: torch.Tensor = None
    workspace: torch.Tensor = None
    perm: torch.Tensor = None
    tile_start: torch.Tensor = None
    tile_rows: torch.Tensor = None
    tile_net: torch.Tensor = None
    n_tiles: int = 0
    aev: torch.Tensor = None
    g_aev: torch.Tensor = None
    e_atom: torch.Tensor = None
    grad_self: torch.Tensor = None
    grad_nbr: torch.Tensor = None       # int64 [N,3] angular neighbour gradients, 2^-32 fixed point
    vir_atom: torch.Tensor = None
    energy: torch.Tensor = None         # float64 [B]
    virial: torch.Tensor = None         # float64 [B,9]
    seg_offsets: torch.Tensor = None    # int32 [B+1] reduction segments per system
    n_seg: int = 0
    partials: torch.Tensor = None       # float64 [1 + n_seg * 19]: completion count, partial sums
    # Verlet skin: rows of the list built with cutoff + skin (radial only), re-filtered to the
    # radial/angular rows above on the steps between builds; skin == 0: no skin list
    skin: float = 0.0
    max_skin: int = 0
    nbr_skin: torch.Tensor = None
    skin_shift: torch.Tensor = None
    n_skin: torch.Tensor = None
    status_skin: torch.Tensor = None
    extra: dict = field(default_factory=dict)


class StageTimer:
    """HIP-event timing of the kernel stages on the current stream (for measurement only)."""

    def __init__(self):
        self.events: dict[str, list] = {}

    @contextlib.contextmanager
    def stage(self, name: str):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        yield
        b.record()
        self.events.setdefault(name, []).append((a, b))

    def times_ms(self) -> dict[str, np.ndarray]:
        torch.cuda.synchronize()
        return {k: np.array([a.elapsed_time(b) for a, b in v]) for k, v in self.events.items()}


class ANIPotential:
    """ANI AEV + species MLPs on the GPU with an analytic backward.

    ``precision``: "fp32" (AEV, GEMMs and activations in FP32) or "bf16" (AEV
    computed in FP32 and stored in BF16, MLP activations in BF16 on MFMA with FP32
    accumulation; hidden-layer weights as BF16 hi + lo pairs, about 16 significant bits, in the
    forward products and BF16 in the backward products with ``bf16_compensated="forward"``, the
    default; pairs in both with True, plain BF16 weights with False).  Per-system energies and virials are
    reduced in FP64.
    """

    def __init__(self, config: ANIConfig, params: dict, precision: str = "fp32", device="cuda",
                 max_nbr: int = 64, max_nbr_ang: int = 32, mlp_tile_rows: int = 0,
                 unhandled_species: str = "raise", bf16_compensated: bool | str = "forward"):
        self.lib = _abi.load()
        self.config = config
        self.precision = precision
        self.device = torch.device(device)
        self.cparams = config.c_params()
        self.nets = SpeciesNetworks(params, config.aev_length, config.k_pad, precision, self.device,
                                    tile_rows=mlp_tile_rows, compensated=bf16_compensated)
        self.net_of_z = {SYMBOL_TO_Z[s]: n for n, s in enumerate(self.nets.species)}
        self._row_dtype = _abi.FNL_F32 if precision == "fp32" else _abi.FNL_BF16  # AEV and dE/dAEV rows
        self.init_max_nbr = max_nbr
        self.init_max_nbr_ang = max_nbr_ang
        # atoms of an element without a network: "raise" (SpeciesIndexNet, nets.py:394-397) or
        # "zero" (ChemicalNet, nets.py:480-505: their output is 0, while their AEV and their
        # place in their neighbours' AEVs use the AEV's species map as for every atom)
        if unhandled_species not in ("raise", "zero"):
            raise ValueError("unhandled_species must be 'raise' or 'zero'")
        self.unhandled_species = unhandled_species
        self.timer: StageTimer | None = None

    def _stage(self, name: str):
        return self.timer.stage(name) if self.timer is not None else contextlib.nullcontext()

    # ------------------------------------------------------------------ setup
    def grid_caps(self, natoms, cells) -> tuple[np.ndarray, int]:
        """Per-system cell-grid capacities (cap_x, cap_y, cap_z, first_cell)."""
        return grid_caps(self.config.cutoff, natoms, cells)

    def setup(self, species, natoms=None, cells=None, skin: float = 0.0,
              minimum_image: bool = True) -> SystemBuffers:
        """Device buffers of one batch.  ``skin`` > 0 (A) adds a Verlet skin list: full builds
        use cutoff + skin and :meth:`update_skin` re-filters it between builds
        (preprocessing.py:640-696).  With cells, ``minimum_image`` selects the minimum-image
        convention (the reference's ``flags={"minimum_image": None}``) or, if False, general
        PBC with every image within the cutoff (preprocessing.py:144-257)."""
        species = np.asarray(species, np.int32)
        N = species.shape[0]
        natoms = np.array([N], np.int32) if natoms is None else np.asarray(natoms, np.int32)
        B = len(natoms)
        if natoms.sum() > N:
            raise ValueError("natoms exceeds the number of atoms")
        dev = self.device
        offs = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int32)
        batch_index = np.full(N, max(B - 1, 0), np.int32)
        batch_index[: offs[-1]] = np.repeat(np.arange(B, dtype=np.int32), natoms)
        conv = self.config.species_map()
        sidx = np.where(species > 0, conv[np.clip(species, 0, conv.shape[0] - 1)], -1).astype(np.int32)
        sidx[offs[-1]:] = -1
        real = species[: offs[-1]]
        real = real[real > 0]
        missing = sorted({PERIODIC_TABLE[z] for z in np.unique(real)} - set(self.nets.species))
        if missing and self.unhandled_species == "raise":
            raise ValueError(f"Species {missing} not found in networks. Handled species are "
                             f"{self.nets.species} (nets.py:394-397)")
        pbc = _abi.FNL_PBC_NONE
        tc = tr = cells64 = None
        if cells is not None:
            cells = np.asarray(cells, np.float64).reshape(-1, 3, 3)
            if cells.shape[0] == 1 and B > 1:
                cells = np.repeat(cells, B, axis=0)
            pbc = _abi.FNL_PBC_MINIMUM_IMAGE if minimum_image else _abi.FNL_PBC_GENERAL
            tc = torch.tensor(cells.astype(np.float32), device=dev).contiguous()
            cells64 = cells
            tr = torch.tensor(np.linalg.inv(cells).astype(np.float32), device=dev).contiguous()
        grid, n_cells = self.grid_caps(natoms, cells)
        sb = SystemBuffers(
            n_atoms=N, n_sys=B, pbc=pbc, species=species, natoms=natoms,
            sys_offsets=torch.tensor(offs, device=dev),
            batch_index=torch.tensor(batch_index, device=dev),
            species_idx=torch.tensor(sidx, device=dev),
            cells=tc, rcells=tr, grid=torch.tensor(grid, device=dev), n_cells=n_cells,
            max_nbr=self.init_max_nbr, max_nbr_ang=self.init_max_nbr_ang,
        )
        if skin < 0:
            raise ValueError("skin must be >= 0")
        if skin > 0:
            sb.skin = float(skin)
            grow = ((self.config.cutoff + skin) / self.config.cutoff) ** 3
            sb.max_skin = _next_mult(int(grow * self.init_max_nbr), 8)
        sb.extra["cells64"] = cells64
        self._alloc_nblist(sb)
        self._setup_tiles(sb)
        kp = self.config.k_pad
        adt = torch.float32 if self.precision == "fp32" else torch.bfloat16
        # AEV rows zeroed once; the forward then writes only the channel blocks of the species
        # present in this batch (the others stay zero for every atom of it)
        sb.aev = torch.zeros((N, kp), dtype=adt, device=dev)
        sb.extra["aev_mask"] = int(np.bitwise_or.reduce(1 << np.unique(sidx[sidx >= 0]).astype(np.int64),
                                                        initial=0)) & 0xFFFFFFFF
        sb.g_aev = torch.zeros((N, kp), dtype=adt, device=dev)  # dE/dAEV rows: BF16 in the BF16 form
        sb.e_atom = torch.zeros(N, dtype=torch.float32, device=dev)
        sb.grad_self = torch.zeros((N, 3), dtype=torch.float32, device=dev)
        sb.grad_nbr = torch.zeros((N, 3), dtype=torch.int64, device=dev)  # 2^-32 fixed point
        sb.vir_atom = torch.zeros((N, 9), dtype=torch.float32, device=dev)
        sb.energy = torch.zeros(B, dtype=torch.float64, device=dev)
        sb.virial = torch.zeros((B, 9), dtype=torch.float64, device=dev)
        nseg = (natoms.astype(np.int64) + SEG_ATOMS - 1) // SEG_ATOMS
        segoff = np.concatenate([[0], np.cumsum(nseg)]).astype(np.int32)
        sb.seg_offsets = torch.tensor(segoff, device=dev)
        sb.n_seg = int(segoff[-1])
        # completion count + per-segment partial sums (fnl_system_reduce)
        sb.partials = torch.zeros(1 + max(sb.n_seg, 1) * _abi.FNL_REDUCE_VALUES, dtype=torch.float64, device=dev)
        return sb

    # ------------------------------------------------------- capacity pools
    def allocate_pool(self, max_atoms: int, max_sys: int, max_nbr: int = 208, max_nbr_ang: int = 72,
                      cells_per_atom: int = 4) -> SystemBuffers:
        """Buffers sized for up to max_atoms atoms in up to max_sys open-boundary systems.

        The layout is set (and changed) by :meth:`set_pool_layout`; the kernels only touch the
        first n_atoms rows, so one pool serves batches of changing composition (serving)."""
        dev = self.device
        n_cells = cells_per_atom * max_atoms + max_sys
        sb = SystemBuffers(
            n_atoms=0, n_sys=0, pbc=_abi.FNL_PBC_NONE, species=np.zeros(0, np.int32),
            natoms=np.zeros(0, np.int32),
            sys_offsets=torch.zeros(max_sys + 1, dtype=torch.int32, device=dev),
            batch_index=torch.zeros(max_atoms, dtype=torch.int32, device=dev),
            species_idx=torch.full((max_atoms,), -1, dtype=torch.int32, device=dev),
            cells=None, rcells=None, grid=torch.zeros((max_sys, 4), dtype=torch.int32, device=dev),
            n_cells=n_cells, max_nbr=max_nbr, max_nbr_ang=max_nbr_ang)
        sb.extra["max_atoms"] = max_atoms
        sb.extra["max_sys"] = max_sys
        sb.extra["max_cells"] = n_cells
        sb.n_atoms = max_atoms  # size the neighbour and workspace buffers for the capacity
        sb.n_sys = max_sys
        self._alloc_nblist(sb)
        M = self.nets.tile_rows
        max_tiles = max_atoms // M + max_sys * len(self.nets.species) + 1
        sb.perm = torch.zeros(max_atoms, dtype=torch.int32, device=dev)
        sb.tile_start = torch.zeros(max_tiles, dtype=torch.int32, device=dev)
        sb.tile_rows = torch.zeros(max_tiles, dtype=torch.int32, device=dev)
        sb.tile_net = torch.zeros(max_tiles, dtype=torch.int32, device=dev)
        kp = self.config.k_pad
        adt = torch.float32 if self.precision == "fp32" else torch.bfloat16
        sb.aev = torch.zeros((max_atoms, kp), dtype=adt, device=dev)
        sb.g_aev = torch.zeros((max_atoms, kp), dtype=adt, device=dev)
        sb.e_atom = torch.zeros(max_atoms, dtype=torch.float32, device=dev)
        sb.grad_self = torch.zeros((max_atoms, 3), dtype=torch.float32, device=dev)
        sb.grad_nbr = torch.zeros((max_atoms, 3), dtype=torch.int64, device=dev)  # 2^-32 fixed point
        sb.vir_atom = torch.zeros((max_atoms, 9), dtype=torch.float32, device=dev)
        sb.energy = torch.zeros(max_sys, dtype=torch.float64, device=dev)
        sb.virial = torch.zeros((max_sys, 9), dtype=torch.float64, device=dev)
        max_seg = max_atoms // SEG_ATOMS + max_sys + 1
        sb.seg_offsets = torch.zeros(max_sys + 1, dtype=torch.int32, device=dev)
        sb.partials = torch.zeros(1 + max_seg * _abi.FNL_REDUCE_VALUES, dtype=torch.float64, device=dev)
        sb.n_atoms = 0
        sb.n_sys = 0
        return sb

    def pool_layout_arrays(self, species: np.ndarray, natoms: np.ndarray) -> dict:
        """Host arrays of a pool layout (systems packed from atom 0, no padding between them)."""
        species = np.asarray(species, np.int32)
        natoms = np.asarray(natoms, np.int32)
        B = natoms.shape[0]
        N = int(natoms.sum())
        offs = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int32)
        batch_index = np.repeat(np.arange(B, dtype=np.int32), natoms)
        conv = self.config.species_map()
        sidx = np.where(species > 0, conv[np.clip(species, 0, conv.shape[0] - 1)], -1).astype(np.int32)
        k = np.maximum(1, np.ceil(np.maximum(natoms, 1) ** (1.0 / 3.0) - 1e-9)).astype(np.int32)
        ncell = k.astype(np.int64) ** 3
        grid = np.stack([k, k, k, np.concatenate([[0], np.cumsum(ncell)[:-1]]).astype(np.int32)], 1)
        n_cells = int(ncell.sum())
        net = np.full(N, -1, np.int64)
        for z, n in self.net_of_z.items():
            net[species == z] = n
        if np.any((net < 0) & (species > 0)) and self.unhandled_species == "raise":
            missing = sorted({PERIODIC_TABLE[z] for z in np.unique(species[(net < 0) & (species > 0)])})
            raise ValueError(f"Species {missing} not found in networks")
        order = np.argsort(net, kind="stable")
        order = order[net[order] >= 0]
        counts = np.bincount(net[net >= 0], minlength=len(self.nets.species))
        M = self.nets.tile_rows
        ts, tr, tn = [], [], []
        base = 0
        for n, c in enumerate(counts):
            for s0 in range(0, int(c), M):
                ts.append(base + s0)
                tr.append(min(M, int(c) - s0))
                tn.append(n)
            base += int(c)
        nseg = (natoms.astype(np.int64) + SEG_ATOMS - 1) // SEG_ATOMS
        segoff = np.concatenate([[0], np.cumsum(nseg)]).astype(np.int32)
        return dict(n_atoms=N, n_sys=B, sys_offsets=offs, batch_index=batch_index, species_idx=sidx, grid=grid,
                    n_cells=max(n_cells, 1), perm=order.astype(np.int32), tile_start=np.array(ts, np.int32),
                    tile_rows=np.array(tr, np.int32), tile_net=np.array(tn, np.int32), seg_offsets=segoff,
                    n_seg=int(segoff[-1]), species=species, natoms=natoms)

    def set_pool_layout(self, sb: SystemBuffers, lay: dict) -> None:
        """Copy a layout (from pool_layout_arrays) into the pool's device tables."""
        if lay["n_atoms"] > sb.extra["max_atoms"] or lay["n_sys"] > sb.extra["max_sys"]:
            raise ValueError("layout exceeds the pool capacity")
        if lay["n_cells"] > sb.extra["max_cells"] or len(lay["tile_start"]) > sb.tile_start.numel():
            raise ValueError("layout exceeds the pool's cell or tile capacity")
        N, B, T = lay["n_atoms"], lay["n_sys"], len(lay["tile_start"])

        def put(dst, src):
            if src.size:
                dst[: src.shape[0]].copy_(torch.from_numpy(np.ascontiguousarray(src)), non_blocking=False)

        put(sb.sys_offsets, lay["sys_offsets"])
        put(sb.batch_index, lay["batch_index"])
        put(sb.species_idx, lay["species_idx"])
        put(sb.grid, lay["grid"])
        put(sb.perm, lay["perm"])
        put(sb.tile_start, lay["tile_start"])
        put(sb.tile_rows, lay["tile_rows"])
        put(sb.tile_net, lay["tile_net"])
        put(sb.seg_offsets, lay["seg_offsets"])
        sb.n_atoms, sb.n_sys, sb.n_tiles, sb.n_seg = N, B, T, lay["n_seg"]
        sb.n_cells = lay["n_cells"]
        sb.species, sb.natoms = lay["species"], lay["natoms"]

    def _alloc_nblist(self, sb: SystemBuffers) -> None:
        dev = self.device
        N = sb.extra.get("max_atoms", sb.n_atoms)
        n_sys = sb.extra.get("max_sys", sb.n_sys)
        n_cells = sb.extra.get("max_cells", sb.n_cells)
        sb.nbr = torch.zeros((max(N, 1), sb.max_nbr), dtype=torch.int32, device=dev)
        sb.n_nbr = torch.zeros(max(N, 1), dtype=torch.int32, device=dev)
        sb.nbr_ang = torch.zeros((max(N, 1), sb.max_nbr_ang), dtype=torch.int32, device=dev)
        sb.n_nbr_ang = torch.zeros(max(N, 1), dtype=torch.int32, device=dev)
        sb.nbr_shift = None
        if sb.pbc == _abi.FNL_PBC_GENERAL:
            sb.nbr_shift = torch.zeros((max(N, 1), sb.max_nbr), dtype=torch.int32, device=dev)
        sb.status = torch.zeros(_abi.FNL_STATUS_WORDS, dtype=torch.int32, device=dev)
        wsb = self.lib.fnl_nblist_workspace(N, max(n_sys, 1), max(n_cells, 1))
        sb.workspace = torch.empty(max(int(wsb), 256), dtype=torch.uint8, device=dev)
        if sb.skin > 0:
            self._alloc_skin(sb)

    def _alloc_skin(self, sb: SystemBuffers) -> None:
        N = max(sb.extra.get("max_atoms", sb.n_atoms), 1)
        sb.nbr_skin = torch.zeros((N, sb.max_skin), dtype=torch.int32, device=self.device)
        sb.skin_shift = None
        if sb.pbc == _abi.FNL_PBC_GENERAL:
            sb.skin_shift = torch.zeros((N, sb.max_skin), dtype=torch.int32, device=self.device)
        sb.n_skin = torch.zeros(N, dtype=torch.int32, device=self.device)
        sb.status_skin = torch.zeros(_abi.FNL_STATUS_WORDS, dtype=torch.int32, device=self.device)

    def _setup_tiles(self, sb: SystemBuffers) -> None:
        """Atoms grouped by network, cut into tiles of at most tile_rows rows (SpeciesIndexer role)."""
        net = np.array([self.net_of_z.get(int(z), -1) if z > 0 else -1 for z in sb.species], np.int64)
        net[int(np.sum(sb.natoms)):] = -1
        M = self.nets.tile_rows
        counts = [int(np.sum(net == n)) for n in range(len(self.nets.species))]
        if self.nets.small_tile_rows and sum(-(-c // M) for c in counts) < _n_cus() // 2:
            M = self.nets.small_tile_rows
        sb.extra["mlp_tile_rows"] = M
        perm, ts, tr, tn = [], [], [], []
        for n in range(len(self.nets.species)):
            idx = np.nonzero(net == n)[0]
            base = len(perm)
            perm.extend(idx.tolist())
            for s in range(0, idx.size, M):
                ts.append(base + s)
                tr.append(min(M, idx.size - s))
                tn.append(n)
        dev = self.device
        sb.perm = torch.tensor(np.array(perm if perm else [0], np.int32), device=dev)
        sb.tile_start = torch.tensor(np.array(ts if ts else [0], np.int32), device=dev)
        sb.tile_rows = torch.tensor(np.array(tr if tr else [0], np.int32), device=dev)
        sb.tile_net = torch.tensor(np.array(tn if tn else [0], np.int32), device=dev)
        sb.n_tiles = len(ts)

    # --------------------------------------------------------------- kernels
    def build_nblist(self, sb: SystemBuffers, coords32: torch.Tensor, stream=None) -> None:
        """Full build (GraphGenerator.process + GraphFilter.process, preprocessing.py:369-638,
        909-956); with a skin, the cutoff + skin list followed by :meth:`update_skin`."""
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        if sb.skin > 0:
            rc = self.lib.fnl_nblist_build(
                P(coords32), P(sb.species_idx), P(sb.batch_index), P(sb.sys_offsets), sb.n_atoms, sb.n_sys,
                sb.pbc, P(sb.cells), P(sb.rcells), P(sb.grid), sb.n_cells, self.config.cutoff + sb.skin,
                self.config.cutoff_angular, sb.max_skin, 1, P(sb.workspace), sb.workspace.numel(),
                P(sb.nbr_skin), P(sb.skin_shift), P(sb.n_skin), None, None, P(sb.status_skin), s)
            _abi.check(rc, "fnl_nblist_build")
            self.update_skin(sb, coords32, stream)
            return
        rc = self.lib.fnl_nblist_build(
            P(coords32), P(sb.species_idx), P(sb.batch_index), P(sb.sys_offsets), sb.n_atoms, sb.n_sys, sb.pbc,
            P(sb.cells), P(sb.rcells), P(sb.grid), sb.n_cells, self.config.cutoff, self.config.cutoff_angular,
            sb.max_nbr, sb.max_nbr_ang, P(sb.workspace), sb.workspace.numel(), P(sb.nbr), P(sb.nbr_shift),
            P(sb.n_nbr), P(sb.nbr_ang), P(sb.n_nbr_ang), P(sb.status), s)
        _abi.check(rc, "fnl_nblist_build")

    def update_skin(self, sb: SystemBuffers, coords32: torch.Tensor, stream=None) -> None:
        """Radial and angular rows at coords32 from the skin list of the last full build
        (GraphGenerator.update_skin preprocessing.py:640-696, GraphFilter.update_skin :958-960,
        GraphAngularExtension.update_skin :1278-1280)."""
        if sb.skin <= 0:
            raise ValueError("these buffers have no skin list (setup(..., skin>0))")
        P = _abi.ptr
        rc = self.lib.fnl_nblist_skin_update(
            P(coords32), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells), P(sb.nbr_skin),
            P(sb.skin_shift), P(sb.n_skin), sb.max_skin, self.config.cutoff, self.config.cutoff_angular,
            sb.max_nbr, sb.max_nbr_ang, P(sb.nbr), P(sb.nbr_shift), P(sb.n_nbr), P(sb.nbr_ang),
            P(sb.n_nbr_ang), P(sb.status), _abi.stream_handle(stream))
        _abi.check(rc, "fnl_nblist_skin_update")

    def check_reallocate(self, sb: SystemBuffers, coords32: torch.Tensor, mult: float = 1.05,
                         add: int = 4) -> bool:
        """Read the status words; on overflow grow the rows and rebuild (preprocessing.py:357-367).

        With a skin list, an overflow of the skin rows grows them and redoes the full build; an
        overflow of the rows re-filtered from it grows those and redoes only the re-filter, so
        the skin list of the last full build is kept.  Returns True if the lists were
        reallocated.  This synchronises the stream.
        """
        grown = False
        if sb.skin > 0:
            st = sb.status_skin.cpu().numpy()
            _status_errors(st)
            if st[_abi.FNL_STATUS_OVERFLOW]:
                sb.max_skin = _next_mult(int(mult * max(st[_abi.FNL_STATUS_MAX_NBR], sb.max_skin)) + add, 8)
                self._alloc_skin(sb)
                self.build_nblist(sb, coords32)
                if sb.status_skin.cpu().numpy()[_abi.FNL_STATUS_OVERFLOW]:
                    raise RuntimeError("skin list still overflows after reallocation")
                grown = True
        st = sb.status.cpu().numpy()
        _status_errors(st)
        if st[_abi.FNL_STATUS_OVERFLOW] == 0:
            return grown
        sb.max_nbr = _next_mult(int(mult * max(st[_abi.FNL_STATUS_MAX_NBR], sb.max_nbr)) + add, 8)
        sb.max_nbr_ang = _next_mult(int(mult * max(st[_abi.FNL_STATUS_MAX_NBR_ANG], sb.max_nbr_ang)) + add, 8)
        if sb.skin > 0:
            kept = sb.nbr_skin, sb.skin_shift, sb.n_skin, sb.status_skin
            self._alloc_nblist(sb)
            sb.nbr_skin, sb.skin_shift, sb.n_skin, sb.status_skin = kept
            self.update_skin(sb, coords32)
        else:
            self._alloc_nblist(sb)
            self.build_nblist(sb, coords32)
        st = sb.status.cpu().numpy()
        if st[_abi.FNL_STATUS_OVERFLOW]:
            raise RuntimeError("neighbour list still overflows after reallocation")
        return True

    def aev_forward(self, sb: SystemBuffers, coords32: torch.Tensor, stream=None) -> None:
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        dt = _abi.FNL_F32 if self.precision == "fp32" else _abi.FNL_BF16
        rc = self.lib.fnl_aev_forward(
            P(coords32), P(sb.species_idx), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells),
            P(sb.nbr), P(sb.n_nbr), sb.max_nbr, P(sb.nbr_shift), P(sb.nbr_ang), P(sb.n_nbr_ang), sb.max_nbr_ang,
            C.byref(self.cparams), P(sb.aev), sb.aev.shape[1], dt, sb.extra.get("aev_mask", 0xFFFFFFFF), s)
        _abi.check(rc, "fnl_aev_forward")

    def mlp(self, sb: SystemBuffers, with_grad: bool = True, stream=None) -> None:
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        self.nets.cw.tile_rows = sb.extra.get("mlp_tile_rows", self.nets.tile_rows)
        rc = self.lib.fnl_species_mlp(
            P(sb.aev), sb.aev.shape[1], C.byref(self.nets.cw), P(sb.perm), P(sb.tile_start), P(sb.tile_rows),
            P(sb.tile_net), sb.n_tiles, P(sb.e_atom), P(sb.g_aev) if with_grad else None, sb.g_aev.shape[1],
            self._row_dtype, self.config.live_column_blocks(sb.extra.get("aev_mask", 0xFFFFFFFF)),
            sb.extra.get("n_tiles_dev"), s)
        _abi.check(rc, "fnl_species_mlp")

    def zero_grad_nbr(self, sb: SystemBuffers, stream=None) -> None:
        """Clear the neighbour-gradient rows the adjoint accumulates into (a fill kernel over the
        int64 words)."""
        _abi.check(self.lib.fnl_fill_i32(_abi.ptr(sb.grad_nbr), 0, 2 * sb.grad_nbr.numel(),
                                         _abi.stream_handle(stream)), "fnl_fill_i32")

    def aev_backward(self, sb: SystemBuffers, coords32: torch.Tensor, stream=None, zero: bool = True) -> None:
        """dE/dx from dE/dAEV as grad_self + grad_nbr (grad_self: written per atom, the centre's
        terms and the radial terms gathered from its neighbours; grad_nbr: angular neighbour
        terms accumulated with int64 atomics in 2^-32 fixed point -- exact sums, the same bits
        every run -- cleared first unless ``zero`` is False), virial per atom."""
        if zero:
            self.zero_grad_nbr(sb, stream)
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        rc = self.lib.fnl_aev_backward(
            P(coords32), P(sb.species_idx), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells),
            P(sb.nbr), P(sb.n_nbr), sb.max_nbr, P(sb.nbr_shift), P(sb.nbr_ang), P(sb.n_nbr_ang), sb.max_nbr_ang,
            C.byref(self.cparams), P(sb.g_aev), sb.g_aev.shape[1], self._row_dtype, P(sb.grad_self), P(sb.grad_nbr),
            P(sb.vir_atom), sb.extra.get("aev_mask", 0xFFFFFFFF), s)
        _abi.check(rc, "fnl_aev_backward")

    def reduce(self, sb: SystemBuffers, with_virial: bool = True, stream=None) -> None:
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        rc = self.lib.fnl_system_reduce(
            P(sb.sys_offsets), P(sb.seg_offsets), sb.n_sys, sb.n_seg, SEG_ATOMS, P(sb.e_atom), P(sb.energy),
            P(sb.vir_atom) if with_virial else None, P(sb.virial), None, None, None, P(sb.partials), s)
        _abi.check(rc, "fnl_system_reduce")

    def compute(self, sb: SystemBuffers, coords32: torch.Tensor, forces: bool = True,
                rebuild: bool | str = True, check: bool = True, stream=None, zero_grad: bool = True) -> None:
        """Energy (and gradients) of the batch at coords32 (float32 [N,3] device tensor).

        ``rebuild``: True = full neighbour-list build, "skin" = re-filter of the skin list
        (:meth:`update_skin`), False = keep the current rows.  ``zero_grad`` False: grad_nbr is
        already zero (the MD integrator clears it when it reads the forces)."""
        if rebuild == "skin":
            with self._stage("skin"):
                self.update_skin(sb, coords32, stream)
            if check:
                self.check_reallocate(sb, coords32)
        elif rebuild:
            with self._stage("nblist"):
                self.build_nblist(sb, coords32, stream)
            if check:
                self.check_reallocate(sb, coords32)
        with self._stage("aev_fwd"):
            self.aev_forward(sb, coords32, stream)
        with self._stage("mlp"):
            self.mlp(sb, with_grad=forces, stream=stream)
        if forces:
            if zero_grad:
                with self._stage("zero_grad"):
                    self.zero_grad_nbr(sb, stream)
            with self._stage("aev_bwd"):
                self.aev_backward(sb, coords32, stream, zero=False)
        with self._stage("reduce"):
            self.reduce(sb, with_virial=forces, stream=stream)

    def clear_sticky(self, sb: SystemBuffers, stream=None) -> None:
        """Zero the sticky overflow word and the sticky row maxima (status words 4-6) with a
        fill kernel (no memset node if the caller's stream is being captured)."""
        s = _abi.stream_handle(stream)
        for st in (sb.status, sb.status_skin if sb.skin > 0 else None):
            if st is not None:
                _abi.check(self.lib.fnl_fill_i32(st.data_ptr() + 4 * _abi.FNL_STATUS_STICKY, 0, 3, s),
                           "fnl_fill_i32")

    def check_sticky(self, sb: SystemBuffers, clear: bool = True) -> None:
        """Raise if any build since the last clear overflowed a capacity (synchronises)."""
        v = int(sb.status[_abi.FNL_STATUS_STICKY].item())
        if sb.skin > 0:
            v |= int(sb.status_skin[_abi.FNL_STATUS_STICKY].item())
        if clear:
            self.clear_sticky(sb)
        if v:
            raise RuntimeError("neighbour-list capacity overflowed during the run")

    def fit_capacity(self, sb: SystemBuffers, coords32: torch.Tensor, margin: float = 1.25, add: int = 8) -> None:
        """Size the rows from one build with headroom for an MD run without per-step host checks."""
        self.build_nblist(sb, coords32)
        self.check_reallocate(sb, coords32)
        n = max(sb.n_atoms, 1)
        want = _next_mult(int(margin * int(sb.n_nbr[:n].max())) + add, 8)
        want_a = _next_mult(int(margin * int(sb.n_nbr_ang[:n].max())) + add, 8)
        want_s = sb.max_skin
        if sb.skin > 0:
            want_s = _next_mult(int(margin * int(sb.n_skin[:n].max())) + add, 8)
        if want > sb.max_nbr or want_a > sb.max_nbr_ang or want_s > sb.max_skin:
            sb.max_nbr = max(sb.max_nbr, want)
            sb.max_nbr_ang = max(sb.max_nbr_ang, want_a)
            sb.max_skin = max(sb.max_skin, want_s)
            self._alloc_nblist(sb)
            self.build_nblist(sb, coords32)
        self.clear_sticky(sb)

    def forces(self, sb: SystemBuffers, scale: float = 1.0, stream=None) -> torch.Tensor:
        out = torch.empty((sb.n_atoms, 3), dtype=torch.float32, device=self.device)
        s = _abi.stream_handle(stream)
        P = _abi.ptr
        _abi.check(self.lib.fnl_forces(P(sb.grad_self), P(sb.grad_nbr), sb.n_atoms, scale, P(out), s),
                   "fnl_forces")
        return out

    def energy_forces_virial(self, species, coords, natoms=None, cells=None, sb: SystemBuffers | None = None):
        """Convenience: (E [B] f64, F [N,3] f32, virial [B,3,3] f64) as CPU NumPy arrays."""
        if sb is None:
            sb = self.setup(species, natoms, cells)
        x = torch.as_tensor(np.asarray(coords, np.float32), device=self.device).contiguous()
        self.compute(sb, x)
        F = self.forces(sb)
        torch.cuda.synchronize()
        return (sb.energy.cpu().numpy(), F.cpu().numpy(), sb.virial.cpu().numpy().reshape(-1, 3, 3))
