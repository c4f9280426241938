"""Continuous batching of energy/force and short-MD requests on one GPU.

FeNNol has no serving loop; its closest analogues are multi-system batches
(``natoms`` / ``batch_index``, models/preprocessing.py:1763-1791) and the
fixed-size batches of analyze.py:108-200.  Here requests of uneven size share
one capacity-static pool of device buffers:

* Admission is by HBM bytes: the pool's buffers are counted per atom row and
  per request slot, and a waiting request (FIFO) is admitted while the bytes
  its atoms and its slot hold fit the pool's budget.
* When the set of requests changes, the host uploads one small record per
  request (size, previous rows, steps left, noise keys) and the atoms of new
  ones; ``fnl_pool_layout`` builds every per-atom and per-system table on the
  device (segment offsets, cell-grid capacities, species index, the
  network-sorted atom list and MLP tiles) and moves continuing requests' rows.
* One iteration -- BAOAB of the MD requests whose forces are known, neighbour
  lists, AEV, MLP, forces, FP64 per-request energies, bookkeeping and the
  packed record of the requests that finish -- is a HIP graph, captured once
  per power-of-two row extent and replayed; the live sizes are device words.
* The host never waits on an iteration it has just launched: it keeps up to
  ``depth`` iterations in flight, each ending with an asynchronous copy of its
  record into pinned memory, and reads records as their events complete.  The
  host knows when each request finishes (steps are deterministic), so it
  re-plans the next layout without reading the device.
* A request whose neighbour rows overflowed in an iteration is flagged in that
  iteration's record: it is restarted from its inputs (its noise is keyed by
  its id and step, so the rerun is the same trajectory) once the rows have
  been regrown, and the graphs are captured again.

Each result does not depend on the other requests in the pool: a request's
neighbour rows, AEV rows, tree reductions and Philox noise counters are its
own, so it is the same bits alone and in any mix.
"""

from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi, units
from .potential import SEG_ATOMS, ANIPotential

REQ_DTYPE = np.dtype([("n_atoms", "<i4"), ("src_row", "<i4"), ("stage_row", "<i4"), ("md_left", "<i4"),
                      ("evaluated", "<i4"), ("noise_key", "<u4"), ("vel_key", "<u4"), ("reserved", "<u4"),
                      ("step", "<u8")])
assert REQ_DTYPE.itemsize == 40


def noise_stream(rid: int) -> int:
    """Philox stream of request rid's O-step noise (counter (local atom, step, stream)): even
    streams, so that no request's noise repeats another's initial velocities.  Request 0 uses
    stream 0, the stream of a single LangevinMD run."""
    return 2 * int(rid)


def velocity_stream(rid: int) -> tuple[int, int]:
    """(step, stream) of request rid's initial velocities: odd streams at step 0 (request 0:
    stream 1, as a single LangevinMD run draws them)."""
    return 0, 2 * int(rid) + 1


def out_bytes(n_sys: int, n_fin: int, n_fin_atoms: int) -> int:
    """Size of an iteration's result record (fnl_pool_out_bytes)."""
    return 4 * _abi.FNL_POOL_HEADER_WORDS + ((4 * n_sys + 7) & ~7) + 8 * n_fin + 36 * n_fin_atoms


@dataclass
class Request:
    species: np.ndarray
    coordinates: np.ndarray
    md_steps: int = 0                 # 0: one energy + forces evaluation
    arrival: float = 0.0              # seconds (or iterations), relative to the start of the run
    rid: int = -1
    # filled by the engine
    energy: float | None = None       # potential energy at the end (model unit)
    forces: np.ndarray | None = None  # forces at the end (model unit / A)
    positions: np.ndarray | None = None
    t_admit: float | None = None
    t_done: float | None = None
    steps_done: int = field(default=0, repr=False)
    evaluated: bool = field(default=False, repr=False)
    restarts: int = field(default=0, repr=False)
    gen: int = field(default=0, repr=False)
    row0: int = field(default=-1, repr=False)

    @property
    def n(self) -> int:
        return len(self.species)

    @property
    def atom_steps(self) -> int:
        return self.n * (self.md_steps + 1)

    @property
    def kind(self) -> str:
        return "md" if self.md_steps > 0 else "ef"

    def _reset(self) -> None:
        self.steps_done = 0
        self.evaluated = False
        self.row0 = -1
        self.gen += 1


@dataclass
class _Flight:
    """One launched iteration whose record has not been read yet."""
    slot: int
    event: torch.cuda.Event
    order: list            # [(request, gen)] in pool order
    finishing: list        # [(b, request, gen, first result row)]
    n_fin_atoms: int
    nbytes: int


class ServingEngine:
    """Continuous-batching server for ANI energy/force and Langevin-MD requests.

    ``max_atoms`` / ``max_systems`` size the pool (rows, request slots); with ``hbm_budget``
    (bytes) the row capacity is instead the largest that fits that budget.  ``depth``: iterations
    in flight before the host waits for the oldest record.  ``clock``: "wall" (arrival times in
    seconds) or "iterations" (arrival times in iterations; deterministic runs).  ``graphs``:
    replay captured HIP graphs (False: the same launches issued one by one).
    """

    def __init__(self, potential: ANIPotential, max_atoms: int = 65536, max_systems: int = 2048,
                 dt_fs: float = 0.5, temperature: float = 300.0, gamma_thz: float = 1.0, seed: int = 0,
                 energy_unit: str = "eV", max_nbr: int = 208, max_nbr_ang: int = 72, hbm_budget: int | None = None,
                 depth: int = 4, graphs: bool = True, clock: str = "wall"):
        if clock not in ("wall", "iterations"):
            raise ValueError("clock must be 'wall' or 'iterations'")
        self.pot = potential
        self.lib = potential.lib
        dev = potential.device
        self.dev = dev
        S = int(max_systems)
        if hbm_budget is not None:
            per_row = self.estimate_row_bytes(potential, max_nbr, max_nbr_ang)
            max_atoms = int((hbm_budget - S * self.estimate_sys_bytes()) // per_row)
            if max_atoms < 64:
                raise ValueError("hbm_budget too small for the pool")
        R = int(max_atoms)
        self.max_atoms, self.max_systems = R, S
        self.depth = max(1, int(depth))
        self.graphs_on = bool(graphs)
        self.clock = clock
        self.sb = sb = potential.allocate_pool(R, S, max_nbr=max_nbr, max_nbr_ang=max_nbr_ang)
        self.dt = dt_fs / units.FS_PER_PS
        self.kT = units.K_B * temperature
        self.a1 = math.exp(-(gamma_thz / units.THZ) * self.dt)
        self.seed = int(seed)
        self.force_scale = 1.0 / units.energy_multiplier(energy_unit)
        z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt, device=dev)
        i32 = torch.int32
        # per-row state and the saved rows of the previous layout
        self.x, self.v, self.x32 = z(R, 3), z(R, 3), z(R, 3, dt=torch.float32)
        self.z, self.dtm, self.a2 = z(R, dt=i32), z(R), z(R)
        self.old = {"x": z(R, 3), "v": z(R, 3), "grad_self": z(R, 3, dt=torch.float32),
                    "grad_nbr": z(R, 3, dt=torch.int64), "z": z(R, dt=i32)}
        # per-request state, device words, uploads and the result record
        self.sys = {k: z(S, dt=i32) for k in ("sys_left", "sys_eval", "sys_key", "sys_fin", "sys_out_row")}
        self.sys["sys_step"] = z(S, dt=torch.int64)
        self.counts = z(_abi.FNL_POOL_COUNTS, dt=i32)
        self.req_dev = z(S * REQ_DTYPE.itemsize, dt=torch.uint8)
        self.stage_z, self.stage_x = z(R, dt=i32), z(R, 3)
        self.out_cap = out_bytes(S, S, R)
        self.out_dev = z(self.out_cap, dt=torch.uint8)
        # per-element tables (atomic number -> AEV species, network, integrator factors)
        conv = potential.config.species_map()
        nz = conv.shape[0]
        z_species = np.where(np.arange(nz) > 0, conv, -1).astype(np.int32)
        z_net = np.full(nz, -1, np.int32)
        for zz, n in potential.net_of_z.items():
            if zz < nz:
                z_net[zz] = n
        masses = np.zeros(nz)
        mm = units.ATOMIC_MASSES[: nz] / units.DA
        masses[: mm.shape[0]] = mm
        ok = masses > 0
        z_dtm, z_a2, z_vs = np.zeros(nz), np.zeros(nz), np.zeros(nz)
        # the factors of the single-run integrator, in the same float64 operations
        z_dtm[ok] = self.dt / masses[ok]
        z_a2[ok] = np.sqrt((1.0 - self.a1 * self.a1) * self.kT / masses[ok])
        z_vs[ok] = np.sqrt(self.kT / masses[ok])
        z_dtm[0] = z_a2[0] = z_vs[0] = 0.0
        self.tables = {"z_species": torch.tensor(z_species, device=dev), "z_net": torch.tensor(z_net, device=dev),
                       "z_dtm": torch.tensor(z_dtm, device=dev), "z_a2": torch.tensor(z_a2, device=dev),
                       "z_vscale": torch.tensor(z_vs, device=dev)}
        self._z_species, self._z_net = z_species, z_net
        self.n_z = nz
        self.n_nets = len(potential.nets.species)
        self.tile_rows = potential.nets.tile_rows
        sb.extra["n_tiles_dev"] = self.counts.data_ptr() + 4 * _abi.FNL_POOL_N_TILES
        self.pool = self._pool_struct()
        # pinned host buffers: one slot per iteration in flight (+1)
        nslot = self.depth + 1
        pin = lambda n, dt: torch.empty(n, dtype=dt, pin_memory=True)
        self.out_host = [pin(self.out_cap, torch.uint8) for _ in range(nslot)]
        self.events = [torch.cuda.Event() for _ in range(nslot)]
        self.up_req = [pin(S * REQ_DTYPE.itemsize, torch.uint8) for _ in range(nslot)]
        self.up_z = [pin(R, torch.int32) for _ in range(nslot)]
        self.up_x = [pin(3 * R, torch.float64) for _ in range(nslot)]
        # host-side scheduler state
        self.pending: deque[Request] = deque()
        self.active: list[Request] = []
        self.inflight: deque[_Flight] = deque()
        self.next_rid = 0
        self.atom_steps = 0
        self.iterations = 0
        self.layouts = 0
        self.restarts = 0
        self.regrowths = 0
        self.n_rows = 0
        self._freed_rows = 0  # rows of finished or restarted requests still in the current layout
        self._dirty = False
        self._grow: tuple[int, int] | None = None
        self._graphs: dict[int, torch.cuda.CUDAGraph] = {}
        self._t0 = time.perf_counter()
        self._clock_iter = 0.0
        # HBM held by the pool: bytes per atom row and per request slot (admission budget)
        self._account()

    # ------------------------------------------------------------ sizes
    @staticmethod
    def estimate_row_bytes(potential: ANIPotential, max_nbr: int, max_nbr_ang: int) -> int:
        """Device bytes per atom row of a pool (before allocation; the exact figure comes from
        the buffers once allocated)."""
        e = 2 if potential.precision == "bf16" else 4
        kp = potential.config.k_pad
        pool = 3 * 24 + 12 + 4 + 16 + 88 + 28 + 36            # x, v, old rows, x32, z, dtm, a2, stage, record
        model = 12 + 4 * (max_nbr + max_nbr_ang + 2) + 2 * kp * e + 4 + 12 + 24 + 36
        # cell-list workspace at the pool's capacity of 4 cells per atom row
        rows = 1 << 16
        ws = int(potential.lib.fnl_nblist_workspace(rows, 1, 4 * rows)) // rows + 1
        return pool + model + ws

    @staticmethod
    def estimate_sys_bytes() -> int:
        return 256

    def _account(self) -> None:
        R, S = self.max_atoms, self.max_systems
        rows = [self.x, self.v, self.x32, self.z, self.dtm, self.a2, self.stage_z, self.stage_x, *self.old.values()]
        sysb = [self.req_dev, *self.sys.values()]
        sb = self.sb
        for name in ("batch_index", "species_idx", "perm", "nbr", "n_nbr", "nbr_ang", "n_nbr_ang", "aev", "g_aev",
                     "e_atom", "grad_self", "grad_nbr", "vir_atom", "workspace"):
            rows.append(getattr(sb, name))
        for name in ("sys_offsets", "seg_offsets", "grid", "energy", "virial", "partials", "tile_start",
                     "tile_rows", "tile_net"):
            sysb.append(getattr(sb, name))
        nb = lambda ts: sum(t.numel() * t.element_size() for t in ts)
        rec = self.out_dev.numel()
        self.row_bytes = (nb(rows) + rec) / R
        self.sys_bytes = nb(sysb) / S
        self.pool_bytes = nb(rows) + rec + nb(sysb)
        self.budget = self.pool_bytes

    def request_bytes(self, r: Request) -> float:
        return r.n * self.row_bytes + self.sys_bytes

    @property
    def used_bytes(self) -> float:
        """HBM bytes the active requests hold (their rows and slots)."""
        return (self.n_rows - self._freed_rows) * self.row_bytes + len(self.active) * self.sys_bytes

    def _pool_struct(self) -> _abi.Pool:
        sb = self.sb
        p = _abi.Pool()
        p.max_atoms, p.max_sys = self.max_atoms, self.max_systems
        p.max_tiles = sb.tile_start.numel()
        p.n_nets, p.tile_rows, p.seg_atoms = self.n_nets, self.tile_rows, SEG_ATOMS
        p.n_z, p.stage_rows, p.seed = self.n_z, self.max_atoms, self.seed
        for k, t in self.tables.items():
            setattr(p, k, t.data_ptr())
        rows = {"x": self.x, "v": self.v, "x32": self.x32, "z": self.z, "grad_self": sb.grad_self,
                "grad_nbr": sb.grad_nbr, "e_atom": sb.e_atom, "virial_atom": sb.vir_atom, "dtm": self.dtm,
                "a2": self.a2, "batch_index": sb.batch_index, "species_idx": sb.species_idx, "perm": sb.perm,
                "sys_offsets": sb.sys_offsets, "seg_offsets": sb.seg_offsets, "grid_cap": sb.grid,
                "tile_start": sb.tile_start, "tile_rows_tab": sb.tile_rows, "tile_net": sb.tile_net,
                "counts": self.counts}
        rows.update({f"old_{k}": t for k, t in self.old.items()})
        rows.update(self.sys)
        for k, t in rows.items():
            setattr(p, k, t.data_ptr())
        return p

    # ------------------------------------------------------------ admission
    def submit(self, req: Request) -> int:
        if req.n > self.max_atoms:
            raise ValueError("request larger than the pool")
        if req.n < 1:
            raise ValueError("empty request")
        sp = np.asarray(req.species)
        if np.any(sp < 0) or np.any(sp >= self.n_z):
            raise ValueError("atomic numbers out of range")
        # every element needs a network here, whatever the potential's unhandled_species rule:
        # pool rows are reused, and only the MLP writes a row's dE/dAEV
        bad = (sp > 0) & (self._z_net[sp] < 0)
        if bad.any():
            from .units import PERIODIC_TABLE
            missing = sorted({PERIODIC_TABLE[z] for z in np.unique(sp[bad])})
            raise ValueError(f"Species {missing} not found in networks")
        if req.rid < 0:
            req.rid = self.next_rid
        self.next_rid = max(self.next_rid, req.rid + 1)
        self.pending.append(req)
        return req.rid

    def _admit(self, now: float) -> list[Request]:
        rows = self.n_rows - self._freed_rows  # rows of the requests that stay
        used = rows * self.row_bytes + len(self.active) * self.sys_bytes
        n_sys = len(self.active)
        new = []
        while self.pending and self.pending[0].arrival <= now:
            r = self.pending[0]
            cost = self.request_bytes(r)
            if used + cost > self.budget or rows + r.n > self.max_atoms or n_sys + 1 > self.max_systems:
                break
            self.pending.popleft()
            if r.t_admit is None:
                r.t_admit = now
            used += cost
            rows += r.n
            n_sys += 1
            new.append(r)
        return new

    # ------------------------------------------------------------ layout
    def _layout(self, new: list[Request]) -> None:
        """Upload the request records (and new requests' atoms) and build the tables on the device."""
        slot = self.iterations % (self.depth + 1)
        order = self.active + new
        B = len(order)
        req = np.zeros(B, REQ_DTYPE)
        n_new = sum(r.n for r in new)
        zs = self.up_z[slot].numpy()
        xs = self.up_x[slot].numpy()
        row = 0
        stage = 0
        for b, r in enumerate(order):
            cont = r.row0 >= 0
            req[b] = (r.n, r.row0 if cont else -1, -1 if cont else stage, r.md_steps - r.steps_done,
                      1 if r.evaluated else 0, noise_stream(r.rid), velocity_stream(r.rid)[1], 0, r.steps_done)
            if not cont:
                zs[stage: stage + r.n] = r.species
                xs[3 * stage: 3 * (stage + r.n)] = np.asarray(r.coordinates, np.float64).reshape(-1)
                stage += r.n
            r.row0 = row
            row += r.n
        up = self.up_req[slot].numpy()
        up[: B * REQ_DTYPE.itemsize] = req.view(np.uint8)
        if B:
            self.req_dev[: B * REQ_DTYPE.itemsize].copy_(self.up_req[slot][: B * REQ_DTYPE.itemsize], non_blocking=True)
        if n_new:
            self.stage_z[:n_new].copy_(self.up_z[slot][:n_new], non_blocking=True)
            self.stage_x.view(-1)[: 3 * n_new].copy_(self.up_x[slot][: 3 * n_new], non_blocking=True)
        s = _abi.stream_handle()
        if self.n_rows:
            _abi.check(self.lib.fnl_pool_save_rows(ctypes.byref(self.pool), self.n_rows, s), "fnl_pool_save_rows")
        _abi.check(self.lib.fnl_pool_layout(ctypes.byref(self.pool), self.req_dev.data_ptr(), B, self.stage_z.data_ptr(),
                                            self.stage_x.data_ptr(), s), "fnl_pool_layout")
        self.active = order
        self.n_rows = row
        self._freed_rows = 0
        self.layouts += 1

    # ------------------------------------------------------------ one iteration
    def _bucket(self, n: int) -> int:
        return min(self.max_atoms, max(256, 1 << max(0, (n - 1).bit_length())))

    def _launch(self, E: int) -> None:
        """The launches of one iteration over the first E rows (capacity-sized grids)."""
        pot, sb, lib = self.pot, self.sb, self.lib
        S = self.max_systems
        s = _abi.stream_handle()
        _abi.check(lib.fnl_pool_integrate(ctypes.byref(self.pool), self.force_scale, 0.5 * self.dt, self.a1, E, s),
                   "fnl_pool_integrate")
        sb.n_atoms, sb.n_sys = E, S
        sb.n_cells = min(sb.extra["max_cells"], 4 * E + S)
        sb.n_seg = E // SEG_ATOMS + S
        sb.n_tiles = min(sb.tile_start.numel(), E // self.tile_rows + self.n_nets + 1)
        pot.build_nblist(sb, self.x32)
        pot.aev_forward(sb, self.x32)
        pot.mlp(sb, with_grad=True)
        _abi.check(lib.fnl_fill_i32(sb.grad_nbr.data_ptr(), 0, 6 * E, s), "fnl_fill_i32")
        pot.aev_backward(sb, self.x32, zero=False)
        pot.reduce(sb, with_virial=False)
        _abi.check(lib.fnl_pool_finish(ctypes.byref(self.pool), sb.energy.data_ptr(), sb.n_nbr.data_ptr(),
                                       sb.n_nbr_ang.data_ptr(), sb.max_nbr, sb.max_nbr_ang, sb.status.data_ptr(),
                                       1.0, E, self.out_dev.data_ptr(), self.out_cap, s), "fnl_pool_finish")

    def _run_iteration(self, E: int) -> None:
        if not self.graphs_on:
            self._launch(E)
            return
        g = self._graphs.get(E)
        if g is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch(E)
            self._graphs[E] = g
        g.replay()

    def _regrow(self) -> None:
        """Grow the neighbour rows after an overflow (buffers move: graphs are captured again)."""
        torch.cuda.synchronize()
        mn, ma = self._grow
        sb = self.sb
        sb.max_nbr = max(sb.max_nbr, mn)
        sb.max_nbr_ang = max(sb.max_nbr_ang, ma)
        self.pot._alloc_nblist(sb)
        self._graphs.clear()
        self._grow = None
        self.regrowths += 1
        self._account()

    def _iterate(self, now: float) -> None:
        if self._grow is not None:
            self._regrow()
        new = self._admit(now)
        if new or self._dirty:
            self._layout(new)
            self._dirty = False
        if not self.active:
            return
        self._run_iteration(self._bucket(self.n_rows))
        # record of this iteration: finished requests (known on the host), copied asynchronously
        order = [(r, r.gen) for r in self.active]
        finishing = []
        row = 0
        for b, r in enumerate(self.active):
            if r.evaluated and r.steps_done < r.md_steps:
                r.steps_done += 1
            r.evaluated = True
            if r.steps_done >= r.md_steps:
                finishing.append((b, r, r.gen, row))
                row += r.n
        nb = out_bytes(len(order), len(finishing), row)
        slot = self.iterations % (self.depth + 1)
        self.out_host[slot][:nb].copy_(self.out_dev[:nb], non_blocking=True)
        ev = self.events[slot]  # its previous flight has been read (at most depth in flight)
        ev.record()
        self.inflight.append(_Flight(slot, ev, order, finishing, row, nb))
        self.iterations += 1
        self.atom_steps += self.n_rows
        if finishing:
            gone = {id(r) for _, r, _, _ in finishing}
            self.active = [r for r in self.active if id(r) not in gone]
            self._dirty = True
            self._freed_rows += row

    # ------------------------------------------------------------ results
    def _process(self, f: _Flight) -> list[Request]:
        buf = self.out_host[f.slot].numpy()[: f.nbytes]
        hdr = buf[: 4 * _abi.FNL_POOL_HEADER_WORDS].view(np.int32)
        B, nfin = len(f.order), len(f.finishing)
        if hdr[8] != 0 or hdr[0] != B or hdr[1] != nfin or hdr[2] != f.n_fin_atoms:
            raise RuntimeError(f"serving record does not match the schedule: header {hdr[:9].tolist()}, expected "
                               f"{B} requests, {nfin} finishing with {f.n_fin_atoms} atoms")
        o = 4 * _abi.FNL_POOL_HEADER_WORDS
        flags = buf[o: o + 4 * B].view(np.int32)
        o += (4 * B + 7) & ~7
        energy = buf[o: o + 8 * nfin].view(np.float64)
        o += 8 * nfin
        xs = buf[o: o + 24 * f.n_fin_atoms].view(np.float64).reshape(-1, 3)
        o += 24 * f.n_fin_atoms
        fs = buf[o: o + 12 * f.n_fin_atoms].view(np.float32).reshape(-1, 3)
        if np.any(flags & 2):
            self._grow = (int(1.25 * max(hdr[5], self.sb.max_nbr)) + 8,
                          int(1.25 * max(hdr[6], self.sb.max_nbr_ang)) + 8)
            for b in np.nonzero(flags & 2)[0]:
                r, gen = f.order[b]
                if r.gen == gen:
                    self._restart(r)
        done = []
        t = time.perf_counter()
        for k, (b, r, gen, row) in enumerate(f.finishing):
            if r.gen != gen:
                continue  # restarted: this result came from truncated neighbour rows
            r.energy = float(energy[k])
            r.positions = xs[row: row + r.n].copy()
            r.forces = fs[row: row + r.n].copy()
            r.t_done = t
            done.append(r)
        return done

    def _restart(self, r: Request) -> None:
        if any(q is r for q in self.active):
            self.active = [q for q in self.active if q is not r]
            self._dirty = True
            self._freed_rows += r.n
        r._reset()
        r.restarts += 1
        self.restarts += 1
        self.pending.appendleft(r)

    def harvest(self, block: bool = False) -> list[Request]:
        """Read the records of completed iterations (waiting for the oldest if ``block``)."""
        done = []
        while self.inflight:
            f = self.inflight[0]
            if block:
                f.event.synchronize()
                block = False
            elif not f.event.query():
                break
            self.inflight.popleft()
            done += self._process(f)
        return done

    def drain(self) -> list[Request]:
        done = []
        while self.inflight:
            done += self.harvest(block=True)
        return done

    # ------------------------------------------------------------ loop
    def now(self) -> float:
        return time.perf_counter() - self._t0 if self.clock == "wall" else float(self.iterations)

    def step(self) -> list[Request]:
        """One scheduling step: launch an iteration if there is work and room in flight, else
        wait for the oldest record.  Returns the requests whose results were read."""
        done = self.harvest(block=False)
        now = self.now()
        ready = bool(self.active) or (self.pending and self.pending[0].arrival <= now)
        if ready and len(self.inflight) < self.depth:
            self._iterate(now)
        elif self.inflight:
            done += self.harvest(block=True)
        elif self.pending:
            if self.clock == "wall":
                time.sleep(min(0.001, max(0.0, self.pending[0].arrival - now)))
            else:
                self.iterations = max(self.iterations, int(math.ceil(self.pending[0].arrival)))
        return done

    def run(self, requests: list[Request], realtime: bool = True) -> dict:
        """Serve a request stream (arrivals against the engine clock when realtime, else all at
        once) until every request has its result."""
        for r in sorted(requests, key=lambda q: q.arrival):
            if not realtime:
                r.arrival = 0.0
            self.submit(r)
        self._t0 = time.perf_counter()
        t0 = self._t0
        finished = []
        total = len(requests)
        while len(finished) < total:
            finished += self.step()
        wall = time.perf_counter() - t0
        for r in finished:
            r.t_done -= t0
        lat = np.array([r.t_done - (r.arrival if self.clock == "wall" else 0.0) for r in finished])
        return {"wall_s": wall, "requests": len(finished), "atom_steps": self.atom_steps,
                "atom_steps_per_s": self.atom_steps / wall, "iterations": self.iterations, "layouts": self.layouts,
                "restarts": self.restarts, "regrowths": self.regrowths, "graphs": len(self._graphs),
                "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_ms": float(lat.mean() * 1e3)}


def make_requests(n_requests: int, seed: int = 3, md_fraction: float = 0.3, md_steps: int = 200,
                  n_min: int = 8, n_max: int = 512) -> list[Request]:
    """Request stream of clusters with log-uniform sizes, a fraction of them short MD runs."""
    from .systems import serving_requests

    out = []
    for k, q in enumerate(serving_requests(n_requests, seed, n_min, n_max, md_fraction, md_steps)):
        s = q["system"]
        out.append(Request(s["species"], s["coordinates"], md_steps=q["md_steps"], rid=k))
    return out


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def serve_stream(potential: ANIPotential, requests: list[Request], arrivals: np.ndarray | None,
                 **engine_kw) -> dict:
    """Serve copies of `requests` (arrival times `arrivals` in seconds, None = all at t=0);
    latency percentiles per request class (energy+forces "ef", Langevin "md")."""
    eng = ServingEngine(potential, **engine_kw)
    reqs = [Request(r.species, r.coordinates, r.md_steps, 0.0 if arrivals is None else float(arrivals[k]), r.rid)
            for k, r in enumerate(requests)]
    res = eng.run(reqs, realtime=arrivals is not None)
    res["requests_per_s"] = len(reqs) / res["wall_s"]
    for kind in ("ef", "md"):
        lat = np.array([r.t_done - r.arrival for r in reqs if r.kind == kind])
        if lat.size:
            res[f"{kind}_requests"] = int(lat.size)
            res[f"{kind}_p50_ms"] = float(np.percentile(lat, 50) * 1e3)
            res[f"{kind}_p99_ms"] = float(np.percentile(lat, 99) * 1e3)
    res["pool_bytes"] = eng.pool_bytes
    res["row_bytes"] = eng.row_bytes
    res["sys_bytes"] = eng.sys_bytes
    res["max_atoms"] = eng.max_atoms
    res["max_systems"] = eng.max_systems
    res["_requests"] = reqs
    return res


def per_request_baseline(potential: ANIPotential, requests: list[Request], dt_fs: float = 0.5,
                         temperature: float = 300.0, gamma_thz: float = 1.0, seed: int = 0) -> dict:
    """Plain per-request launches: each request alone (own buffers and launches), in order."""
    from .md import LangevinMD

    times = []
    atom_steps = 0
    t0 = time.perf_counter()
    for r in requests:
        t1 = time.perf_counter()
        system = {"species": r.species, "coordinates": r.coordinates}
        if r.md_steps == 0:
            sb = potential.setup(r.species)
            x = torch.as_tensor(np.asarray(r.coordinates, np.float32), device=potential.device)
            potential.compute(sb, x)
            potential.forces(sb).cpu()
        else:
            md = LangevinMD(potential, system, dt_fs=dt_fs, temperature=temperature, gamma_thz=gamma_thz,
                            seed=seed + r.rid, check_every=0)
            md.run(r.md_steps)
            md.energies()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
        atom_steps += r.atom_steps
    wall = time.perf_counter() - t0
    t = np.array(times)
    return {"wall_s": wall, "requests": len(requests), "atom_steps": atom_steps,
            "atom_steps_per_s": atom_steps / wall, "requests_per_s": len(requests) / wall,
            "service_p50_ms": float(np.percentile(t, 50) * 1e3), "service_p99_ms": float(np.percentile(t, 99) * 1e3)}
