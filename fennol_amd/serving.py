"""Continuous batching of energy/force and short-MD requests on one GPU.

FeNNol has no serving loop; its closest analogues are multi-system batches
(``natoms`` / ``batch_index``, models/preprocessing.py:1763-1791) and the
fixed-size batches of analyze.py:108-200.  Here requests of uneven size share
one device-resident pool of atoms:

* every iteration the engine admits waiting requests (FIFO) while the pool's
  atom and system capacity -- its HBM budget -- allows, packs the active
  requests as consecutive systems (segment offsets), moves the positions,
  velocities and last gradients of continuing requests to their new rows on
  the device, then runs one batched step: BAOAB for the MD requests whose
  forces are known, the neighbour lists, AEV, MLP, forces and FP64 per-system
  energies for all;
* a single energy+force request finishes after its first evaluation; an MD
  request after ``md_steps`` Langevin steps;
* per-request state (x, v in FP64, step counter, noise key) stays on the
  device; the noise of a request depends on its own key, step and local atom
  index only, so its result does not depend on the other requests in the mix.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi, units
from .potential import ANIPotential


def noise_stream(rid: int) -> int:
    """Philox stream of request rid's O-step noise (counter (local atom, step, stream)): even
    streams, so that no request's noise repeats another's initial velocities.  Request 0 uses
    stream 0, the stream of a single LangevinMD run."""
    return 2 * int(rid)


def velocity_stream(rid: int) -> tuple[int, int]:
    """(step, stream) of request rid's initial velocities: odd streams at step 0 (request 0:
    stream 1, as a single LangevinMD run draws them)."""
    return 0, 2 * int(rid) + 1


@dataclass
class Request:
    species: np.ndarray
    coordinates: np.ndarray
    md_steps: int = 0                 # 0: one energy + forces evaluation
    arrival: float = 0.0              # seconds, relative to the start of the run
    rid: int = -1
    # filled by the engine
    energy: float | None = None       # potential energy at the end (model unit)
    forces: np.ndarray | None = None  # forces at the end (model unit / A)
    positions: np.ndarray | None = None
    t_admit: float | None = None
    t_done: float | None = None
    steps_done: int = field(default=0, repr=False)
    evaluated: bool = field(default=False, repr=False)

    @property
    def n(self) -> int:
        return int(self.species.shape[0])

    @property
    def atom_steps(self) -> int:
        return self.n * (self.md_steps + 1)


class ServingEngine:
    """Continuous-batching server for ANI energy/force and Langevin-MD requests."""

    def __init__(self, potential: ANIPotential, max_atoms: int = 65536, max_systems: int = 2048,
                 dt_fs: float = 0.5, temperature: float = 300.0, gamma_thz: float = 1.0, seed: int = 0,
                 energy_unit: str = "eV", max_nbr: int = 208, max_nbr_ang: int = 72):
        self.pot = potential
        self.lib = potential.lib
        dev = potential.device
        self.dev = dev
        self.max_atoms = max_atoms
        self.max_systems = max_systems
        self.sb = potential.allocate_pool(max_atoms, max_systems, max_nbr=max_nbr, max_nbr_ang=max_nbr_ang)
        self.dt = dt_fs / units.FS_PER_PS
        self.kT = units.K_B * temperature
        self.a1 = math.exp(-(gamma_thz / units.THZ) * self.dt)
        self.seed = int(seed)
        self.force_scale = 1.0 / units.energy_multiplier(energy_unit)
        z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt, device=dev)
        self.x = z(max_atoms, 3)
        self.v = z(max_atoms, 3)
        self.x32 = z(max_atoms, 3, dt=torch.float32)
        self.dtm = z(max_atoms)
        self.a2 = z(max_atoms)
        self.sys_step = torch.zeros(max_systems, dtype=torch.int64, device=dev)
        self.sys_key = torch.zeros(max_systems, dtype=torch.int32, device=dev)
        self.pending: deque[Request] = deque()
        self.active: list[Request] = []
        self.next_rid = 0
        self.atom_steps = 0
        self.iterations = 0
        self.changed = True
        self._row0: dict[int, int] = {}
        # HBM held by the pool (the admission budget is its atom/system capacity)
        self.pool_bytes = sum(t.numel() * t.element_size() for t in (
            self.x, self.v, self.x32, self.dtm, self.a2, self.sb.nbr, self.sb.nbr_ang, self.sb.aev, self.sb.g_aev,
            self.sb.workspace, self.sb.grad_self, self.sb.grad_nbr, self.sb.vir_atom))

    # ------------------------------------------------------------ admission
    def submit(self, req: Request) -> int:
        if req.n > self.max_atoms:
            raise ValueError("request larger than the pool")
        if req.rid < 0:
            req.rid = self.next_rid
        self.next_rid = max(self.next_rid, req.rid + 1)
        self.pending.append(req)
        return req.rid

    def _admit(self, now: float) -> list[Request]:
        n_atoms = sum(r.n for r in self.active)
        n_sys = len(self.active)
        new = []
        while self.pending and self.pending[0].arrival <= now:
            r = self.pending[0]
            if n_atoms + r.n > self.max_atoms or n_sys + 1 > self.max_systems:
                break
            self.pending.popleft()
            r.t_admit = now
            n_atoms += r.n
            n_sys += 1
            new.append(r)
        return new

    def _relayout(self, new: list[Request]) -> None:
        """Pack active requests as consecutive systems; move continuing state on the device."""
        order = self.active + new
        natoms = np.array([r.n for r in order], np.int32)
        species = np.concatenate([r.species for r in order]).astype(np.int32) if order else np.zeros(0, np.int32)
        lay = self.pot.pool_layout_arrays(species, natoms)
        offs = lay["sys_offsets"]
        # gather map for continuing requests (old row -> new row)
        src, dst = [], []
        for b, r in enumerate(self.active):
            o = self._row0[r.rid]
            src.append(np.arange(o, o + r.n))
            dst.append(np.arange(offs[b], offs[b] + r.n))
        if src:
            s_idx = torch.from_numpy(np.concatenate(src)).to(self.dev)
            d_idx = torch.from_numpy(np.concatenate(dst)).to(self.dev)
            # the next BAOAB kick reads the gradient rows of the last evaluation, so they move too
            for t in (self.x, self.v, self.sb.grad_self, self.sb.grad_nbr):
                t[d_idx] = t[s_idx].clone()
        # new requests: positions, velocities ~ N(0, kT/m) for MD requests
        n_cont = len(self.active)
        for k, r in enumerate(new):
            b = n_cont + k
            o = int(offs[b])
            self.x[o:o + r.n] = torch.from_numpy(np.asarray(r.coordinates, np.float64)).to(self.dev)
            if r.md_steps > 0:
                m = units.ATOMIC_MASSES[r.species] / units.DA
                vv = torch.empty((r.n, 3), dtype=torch.float64, device=self.dev)
                step0, stream0 = velocity_stream(r.rid)
                _abi.check(self.lib.fnl_philox_normals(vv.data_ptr(), r.n, self.seed, step0, stream0,
                                                       _abi.stream_handle()), "fnl_philox_normals")
                self.v[o:o + r.n] = vv * torch.from_numpy(np.sqrt(self.kT / m)).to(self.dev)[:, None]
            else:
                self.v[o:o + r.n] = 0.0
        N = lay["n_atoms"]
        self.x32[:N] = self.x[:N].float()
        self.pot.set_pool_layout(self.sb, lay)
        self.active = order
        self._row0 = {r.rid: int(offs[b]) for b, r in enumerate(order)}
        self._masses = np.concatenate([units.ATOMIC_MASSES[r.species] / units.DA for r in order]) if order \
            else np.zeros(0)

    def _integration_tables(self) -> None:
        """Per-atom dt/m and a2 (zero for requests not integrated this iteration) and the
        per-system step numbers and noise keys."""
        N = self.sb.n_atoms
        dtm = np.zeros(N)
        a2 = np.zeros(N)
        steps = np.zeros(len(self.active), np.int64)
        keys = np.zeros(len(self.active), np.int32)
        for b, r in enumerate(self.active):
            o = self._row0[r.rid]
            keys[b] = noise_stream(r.rid)
            steps[b] = r.steps_done
            if r.md_steps > 0 and r.evaluated and r.steps_done < r.md_steps:
                m = self._masses[o:o + r.n]
                dtm[o:o + r.n] = self.dt / m
                a2[o:o + r.n] = np.sqrt((1.0 - self.a1 * self.a1) * self.kT / m)
        if N:
            self.dtm[:N].copy_(torch.from_numpy(dtm))
            self.a2[:N].copy_(torch.from_numpy(a2))
        B = len(self.active)
        if B:
            self.sys_step[:B].copy_(torch.from_numpy(steps))
            self.sys_key[:B].copy_(torch.from_numpy(keys))

    # ------------------------------------------------------------ iteration
    def iterate(self, now: float) -> list[Request]:
        """One batched step; returns the requests completed by it."""
        new = self._admit(now)
        if new or self.changed:
            self._relayout(new)
            self.changed = False
        if not self.active:
            return []
        sb = self.sb
        self._integration_tables()
        s = _abi.stream_handle()
        _abi.check(self.lib.fnl_baoab_step_batched(
            self.x.data_ptr(), self.v.data_ptr(), self.x32.data_ptr(), sb.grad_self.data_ptr(),
            sb.grad_nbr.data_ptr(), self.force_scale, self.dtm.data_ptr(), 0.5 * self.dt, self.a1,
            self.a2.data_ptr(), sb.batch_index.data_ptr(), sb.sys_offsets.data_ptr(), self.sys_step.data_ptr(),
            self.sys_key.data_ptr(), self.seed, sb.n_atoms, s), "fnl_baoab_step_batched")
        self.pot.compute(sb, self.x32, forces=True, rebuild=True, check=False)
        # read back what this iteration decides: capacities, energies of finishing requests
        st = sb.status.cpu().numpy()
        if st[_abi.FNL_STATUS_STICKY]:
            self._grow(st)
            self.pot.compute(sb, self.x32, forces=True, rebuild=True, check=False)
            st = sb.status.cpu().numpy()
            if st[_abi.FNL_STATUS_OVERFLOW]:
                raise RuntimeError("neighbour rows overflow after regrowth")
        self.iterations += 1
        self.atom_steps += sb.n_atoms
        done = []
        for r in self.active:
            if r.evaluated and r.md_steps > 0 and r.steps_done < r.md_steps:
                r.steps_done += 1
            r.evaluated = True
            if r.steps_done >= r.md_steps:
                done.append(r)
        if done:
            energies = sb.energy[: len(self.active)].cpu().numpy()
            F = self.pot.forces(sb)
            t = time.perf_counter()
            idx = {r.rid: b for b, r in enumerate(self.active)}
            for r in done:
                o = self._row0[r.rid]
                r.energy = float(energies[idx[r.rid]])
                r.forces = F[o:o + r.n].cpu().numpy()
                r.positions = self.x[o:o + r.n].cpu().numpy()
                r.t_done = t
            gone = {r.rid for r in done}
            self.active = [r for r in self.active if r.rid not in gone]
            self.changed = True
        return done

    def _grow(self, st) -> None:
        sb = self.sb
        sb.max_nbr = max(sb.max_nbr, int(1.25 * st[_abi.FNL_STATUS_MAX_NBR]) + 8)
        sb.max_nbr_ang = max(sb.max_nbr_ang, int(1.25 * st[_abi.FNL_STATUS_MAX_NBR_ANG]) + 8)
        self.pot._alloc_nblist(sb)

    def run(self, requests: list[Request], realtime: bool = True) -> dict:
        """Serve a request stream; arrivals are honoured against the wall clock when realtime."""
        for r in sorted(requests, key=lambda q: q.arrival):
            self.submit(r)
        t0 = time.perf_counter()
        finished = []
        total = len(requests)
        while len(finished) < total:
            now = time.perf_counter() - t0 if realtime else float("inf")
            if not self.active and self.pending and self.pending[0].arrival > now:
                time.sleep(min(0.001, self.pending[0].arrival - now))
                continue
            done = self.iterate(now)
            for r in done:
                r.t_done -= t0
            finished.extend(done)
        wall = time.perf_counter() - t0
        lat = np.array([r.t_done - r.arrival for r in finished])
        return {"wall_s": wall, "requests": len(finished), "atom_steps": self.atom_steps,
                "atom_steps_per_s": self.atom_steps / wall, "iterations": self.iterations,
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
    """Serve copies of `requests` (arrival times `arrivals`, None = all at t=0)."""
    eng = ServingEngine(potential, **engine_kw)
    reqs = [Request(r.species, r.coordinates, r.md_steps, 0.0 if arrivals is None else float(arrivals[k]), r.rid)
            for k, r in enumerate(requests)]
    res = eng.run(reqs, realtime=arrivals is not None)
    lat_ef = np.array([r.t_done - r.arrival for r in reqs if r.md_steps == 0])
    lat_md = np.array([r.t_done - r.arrival for r in reqs if r.md_steps > 0])
    res["requests_per_s"] = len(reqs) / res["wall_s"]
    if lat_ef.size:
        res["ef_p50_ms"] = float(np.percentile(lat_ef, 50) * 1e3)
        res["ef_p99_ms"] = float(np.percentile(lat_ef, 99) * 1e3)
    if lat_md.size:
        res["md_p50_ms"] = float(np.percentile(lat_md, 50) * 1e3)
        res["md_p99_ms"] = float(np.percentile(lat_md, 99) * 1e3)
    res["pool_bytes"] = eng.pool_bytes
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
