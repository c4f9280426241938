"""Langevin (BAOAB) and NVE molecular dynamics on the device.

Mirrors fennol_md's step (md/integrate.py:570-600): ``integrate`` (BAOAB with
merged half kicks, :286-294, LGV O-step md/thermostats.py:108-120), then the
neighbour-list update (``update_graphs``, :515-566: a full rebuild every
step without a skin; with ``nblist_skin`` a full rebuild every
``nblist_stride`` steps and re-filters of the skin list in between), then
``update_observables`` (energy, forces, kinetic tensor and pressure,
:298-375).  Positions and velocities are FP64 on the device; the model reads
a float32 copy written by the integrator kernel.  Units are fennol_md's
(Angstrom, ps, kcal/mol); the model's energies are converted with its energy
unit (integrate.py:303-305).

Capacity overflows.  ``check_every=1`` reads the neighbour-list overflow word
after every list update and regrows the rows before the forces are used, as
the reference's ``check_reallocate`` does (models/preprocessing.py:1859-1878,
md/integrate.py:524-530).  Otherwise steps run without host reads (eagerly
or as HIP-graph replays) and :meth:`LangevinMD.run` keeps a checkpoint: every
``check_every`` steps (``checkpoint_every`` when ``check_every`` is 0) it
reads the sticky overflow word; if a list update overflowed since the last
checkpoint, the rows are regrown from the sticky maxima, positions,
velocities and the step counter are restored from the checkpoint, and the
steps are run again.  The noise of step n depends on n only, so the redone
steps are those an ample-capacity run takes.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _abi, units
from .potential import SEG_ATOMS, ANIPotential, SystemBuffers, _next_mult, _status_errors


def nblist_schedule(skin: float, dt_fs: float, stride: int = -1) -> int:
    """Steps between full rebuilds (md/integrate.py:488-507): 1 without a skin; else the
    given stride, or by default the time a 2 A skin lasts at 300 K (40 fs) scaled by
    skin / 2 A."""
    if skin <= 0:
        return 1
    if stride <= 0:
        stride = int(math.floor(skin / 2.0 * 40.0 / dt_fs))
    return max(int(stride), 1)


def grown_capacity(current: int, needed: int, mult: float = 1.25, add: int = 8) -> int:
    """Row capacity after an overflow that needed ``needed`` slots (unchanged if it fits)."""
    if needed <= current:
        return current
    return _next_mult(int(mult * needed) + add, 8)


class LangevinMD:
    """Device-resident BAOAB Langevin (``thermostat="LGV"``) or NVE dynamics.

    ``system``: FeNNol input dict (species, coordinates [A], natoms, optional
    cells).  ``seed`` keys the Philox streams: stream 1 draws the initial
    velocities v ~ N(0, kT/m) (thermostats.py:64-67), stream 0 the O-step
    noise of step n.  For NVE the initial velocities are rescaled to exactly
    kT per degree of freedom (thermostats.py:172-189).

    ``nblist_skin`` (A, > 0 enables it), ``nblist_stride`` (steps between full
    rebuilds; <= 0: from the skin, md/integrate.py:498-504) and
    ``nblist_warmup_fs`` (full rebuilds every step before that time) follow
    fennol_md's keywords of the same names (integrate.py:477-513);
    ``minimum_image`` False selects general PBC for periodic systems.

    ``check_every``: 1 = the reference's per-update overflow check; K > 1 =
    checkpointed runs checked every K steps; 0 = capacities sized with headroom
    up front and checkpointed runs checked every ``checkpoint_every`` steps
    (the form HIP-graph replay uses).
    """

    def __init__(self, potential: ANIPotential, system: dict, dt_fs: float = 0.5, temperature: float = 300.0,
                 gamma_thz: float = 1.0, seed: int = 0, thermostat: str = "LGV", energy_unit: str = "eV",
                 check_every: int = 1, sb: SystemBuffers | None = None, nblist_skin: float = -1.0,
                 nblist_stride: int = -1, nblist_warmup_fs: float = -1.0, minimum_image: bool = True,
                 checkpoint_every: int = 200):
        self.pot = potential
        self.lib = potential.lib
        dev = potential.device
        species = np.asarray(system["species"], np.int32)
        natoms = system.get("natoms")
        cells = system.get("cells")
        skin = float(nblist_skin) if nblist_skin > 0 else 0.0
        self.sb = sb if sb is not None else potential.setup(species, natoms, cells, skin=skin,
                                                            minimum_image=minimum_image)
        if self.sb.skin != skin:
            raise ValueError(f"buffers built with skin {self.sb.skin}, MD asks for {skin}")
        self.nblist_stride = nblist_schedule(skin, dt_fs, nblist_stride)
        self.nblist_warmup = int(nblist_warmup_fs / dt_fs) if nblist_warmup_fs > 0 else 0
        self.nblist_countdown = 0
        self.n = species.shape[0]
        mass_da = units.ATOMIC_MASSES[np.clip(species, 0, None)]
        self.mass = np.where(species > 0, mass_da / units.DA, 1.0)       # md/initial.py:79-108
        self.dt = dt_fs / units.FS_PER_PS                                  # ps
        self.gamma = gamma_thz / units.THZ                                 # ps^-1
        self.kT = units.K_B * temperature                                  # kcal/mol
        self.thermostat = thermostat.upper()
        self.seed = int(seed)
        self.istep = 0
        if check_every < 0 or checkpoint_every < 1:
            raise ValueError("check_every must be >= 0 and checkpoint_every >= 1")
        self.check_every = int(check_every)
        self.checkpoint_every = int(check_every) if check_every > 1 else int(checkpoint_every)
        self.rollbacks = 0
        self.model_energy_unit = units.energy_multiplier(energy_unit)
        self.force_scale = 1.0 / self.model_energy_unit                    # model unit -> kcal/mol
        self.t_mass = torch.tensor(self.mass, dtype=torch.float64, device=dev)
        self.t_dtm = torch.tensor(self.dt / self.mass, dtype=torch.float64, device=dev)
        if self.thermostat in ("LGV", "LANGEVIN"):
            self.a1 = math.exp(-self.gamma * self.dt)                      # thermostats.py:62-63
            a2 = np.sqrt((1.0 - self.a1 * self.a1) * self.kT / self.mass)
            self.t_a2 = torch.tensor(np.where(species > 0, a2, 0.0), dtype=torch.float64, device=dev)
        elif self.thermostat in ("NVE", "NONE"):
            self.a1 = 1.0
            self.t_a2 = None
        else:
            raise NotImplementedError(f"thermostat {thermostat}")
        # step counter on the device: the O-step noise of a captured graph advances on replay
        self.step_dev = torch.zeros(2, dtype=torch.int64, device=dev)  # {step, completion count}
        self._sums_stale = False  # a replayed step leaves its per-system sums to the next one
        self.graph = None
        self.graph_skin = None
        self.graph_skin_run = None
        self.steps_per_graph = 1
        self.x = torch.tensor(np.asarray(system["coordinates"], np.float64), device=dev).contiguous()
        self.x32 = self.x.float().contiguous()
        self.v = torch.empty((self.n, 3), dtype=torch.float64, device=dev)
        _abi.check(self.lib.fnl_philox_normals(self.v.data_ptr(), self.n, self.seed, 0, 1, _abi.stream_handle()),
                   "fnl_philox_normals")
        self.v *= torch.sqrt(self.kT / self.t_mass)[:, None]
        self.v[torch.tensor(species <= 0, device=dev)] = 0.0
        if self.t_a2 is None:
            nreal = int((species > 0).sum())
            kt_sys = float((self.t_mass[:, None] * self.v ** 2).sum()) / (3 * nreal)
            self.v *= math.sqrt(self.kT / kt_sys)
        if self.check_every == 0:
            self.pot.fit_capacity(self.sb, self.x32)
        # initial forces (integrate.py:604-616)
        self.pot.compute(self.sb, self.x32)
        self.pot.clear_sticky(self.sb)
        self._ckpt = None

    def verify(self) -> None:
        """Raise if a neighbour-list capacity overflowed since the last verify()."""
        self.pot.check_sticky(self.sb)

    # --------------------------------------------------------------------
    def _integrate(self, stream=None) -> None:
        s = _abi.stream_handle(stream)
        sb = self.sb
        with self.pot._stage("baoab"):
            self._baoab(s, sb)

    def _baoab(self, s, sb) -> None:
        rc = self.lib.fnl_baoab_step(
            self.x.data_ptr(), self.v.data_ptr(), self.x32.data_ptr(), sb.grad_self.data_ptr(),
            sb.grad_nbr.data_ptr(), self.force_scale, self.t_dtm.data_ptr(), 0.5 * self.dt, self.a1,
            _abi.ptr(self.t_a2), None, self.seed, 0, self.step_dev.data_ptr(), self.n, 1, s)
        _abi.check(rc, "fnl_baoab_step")

    def _baoab_sums(self) -> None:
        """The BAOAB step with, in the same launch, the per-system energy and virial sums of the
        atom energies and virials the last force evaluation left (fnl_baoab_step_sums)."""
        sb, P = self.sb, _abi.ptr
        rc = self.lib.fnl_baoab_step_sums(
            self.x.data_ptr(), self.v.data_ptr(), self.x32.data_ptr(), sb.grad_self.data_ptr(),
            sb.grad_nbr.data_ptr(), self.force_scale, self.t_dtm.data_ptr(), 0.5 * self.dt, self.a1,
            P(self.t_a2), None, self.seed, 0, self.step_dev.data_ptr(), self.n, 1, P(sb.sys_offsets),
            P(sb.seg_offsets), sb.n_sys, sb.n_seg, SEG_ATOMS, P(sb.e_atom), P(sb.energy), P(sb.vir_atom),
            P(sb.virial), P(sb.partials), _abi.stream_handle())
        _abi.check(rc, "fnl_baoab_step_sums")

    def _graph_update(self) -> bool | str:
        """The neighbour-list action of the next step (update_graphs, integrate.py:515-566):
        True = full rebuild, "skin" = skin re-filter; advances the countdown."""
        if self.nblist_countdown <= 0 or self.istep < self.nblist_warmup:
            self.nblist_countdown = self.nblist_stride - 1
            return True
        self.nblist_countdown -= 1
        return "skin"

    def step(self, check: bool | None = None, stream=None) -> None:
        """integrate -> update neighbour lists -> energy and forces at the new positions.
        With ``check_every=1`` the overflow word is read after the list update (the reference's
        protocol); otherwise no host read happens here (see :meth:`run`)."""
        rebuild = self._graph_update()
        self._integrate(stream)
        self.istep += 1
        if check is None:
            check = self.check_every == 1
        self.pot.compute(self.sb, self.x32, forces=True, rebuild=rebuild, check=check, stream=stream,
                         zero_grad=False)  # the integrator cleared grad_nbr
        self._sums_stale = False

    # --------------------------------------------------------- checkpoints
    def _checkpoint(self) -> None:
        """Device copies of x, v and the step counter, with the host-side step state."""
        if self._ckpt is None:
            self._ckpt = {"x": torch.empty_like(self.x), "v": torch.empty_like(self.v),
                          "step": torch.empty_like(self.step_dev)}
        self._ckpt["x"].copy_(self.x)
        self._ckpt["v"].copy_(self.v)
        self._ckpt["step"].copy_(self.step_dev)
        self._ckpt["istep"] = self.istep

    def _rollback(self) -> None:
        """Regrow the overflowed rows from the sticky maxima, restore the checkpoint and
        recompute the forces there with a full rebuild (which also restarts the skin schedule)."""
        sb, pot = self.sb, self.pot
        st = sb.status.cpu().numpy()
        _status_errors(st)
        sb.max_nbr = grown_capacity(sb.max_nbr, int(st[_abi.FNL_STATUS_STICKY_MAX_NBR]))
        sb.max_nbr_ang = grown_capacity(sb.max_nbr_ang, int(st[_abi.FNL_STATUS_STICKY_MAX_NBR_ANG]))
        if sb.skin > 0:
            sts = sb.status_skin.cpu().numpy()
            _status_errors(sts)
            sb.max_skin = grown_capacity(sb.max_skin, int(sts[_abi.FNL_STATUS_STICKY_MAX_NBR]))
        pot._alloc_nblist(sb)
        self.x.copy_(self._ckpt["x"])
        self.v.copy_(self._ckpt["v"])
        self.step_dev.copy_(self._ckpt["step"])
        self.x32.copy_(self.x)
        self.istep = self._ckpt["istep"]
        pot.compute(sb, self.x32, forces=True, rebuild=True, check=True)
        self._sums_stale = False
        self.nblist_countdown = self.nblist_stride - 1
        pot.clear_sticky(sb)
        self.rollbacks += 1
        if self.graph is not None:  # the rows moved: record the steps again
            self.capture(self.steps_per_graph, warmup=False)

    def overflowed(self) -> bool:
        """True if a list update overflowed a capacity since the sticky words were last cleared
        (one host read: synchronises)."""
        v = int(self.sb.status[_abi.FNL_STATUS_STICKY].item())
        if self.sb.skin > 0:
            v |= int(self.sb.status_skin[_abi.FNL_STATUS_STICKY].item())
        return v != 0

    def run(self, nsteps: int, graph: bool = False, max_rollbacks: int = 8) -> None:
        """``nsteps`` MD steps.  With ``check_every=1`` (and no graph) each step checks its own
        list update; otherwise the steps run in segments of ``checkpoint_every`` without host
        reads (as HIP-graph replays if ``graph``: :meth:`capture` is called first if needed),
        each segment followed by one read of the sticky overflow word and, on overflow, a
        rollback to the segment's start with regrown rows."""
        if self.check_every == 1 and not graph:
            for _ in range(nsteps):
                self.step()
            return
        if graph and self.graph is None:
            self.capture(1)
        per = self.steps_per_graph if graph else 1
        K = max(per, (self.checkpoint_every // per) * per)
        done = 0
        tries = 0
        self._checkpoint()
        while done < nsteps:
            n = min(K, nsteps - done)
            if graph:
                if self.graph_skin is None and n % per:
                    raise ValueError(f"graphs hold {per} steps: nsteps must be a multiple of it")
                self.replay(n, max_steps=n)
            else:
                for _ in range(n):
                    self.step(check=False)
            if self.overflowed():
                tries += 1
                if tries > max_rollbacks:
                    raise RuntimeError("neighbour-list capacity keeps overflowing after regrowth")
                self._rollback()
                continue
            tries = 0
            done += n
            self._checkpoint()

    # ---------------------------------------------------------- graph replay
    def _record(self, rebuild: bool | str) -> None:
        """One step with a fixed neighbour-list action and no host reads or counters, for
        capture.  With at most 16 systems the per-system sums (energy, virial) of the previous
        step are formed by the integrator's launch (:meth:`_baoab_sums`) rather than by a launch
        of their own at the end of the step; :meth:`energies` and :meth:`virial_kcal` form the
        last replayed step's sums when they are read."""
        pot, sb = self.pot, self.sb
        if sb.n_sys > _abi.FNL_REDUCE_FUSED_MAX_SYSTEMS:
            self._integrate()
            pot.compute(sb, self.x32, forces=True, rebuild=rebuild, check=False, zero_grad=False)
            return
        self._baoab_sums()
        if rebuild == "skin":
            pot.update_skin(sb, self.x32)
        else:
            pot.build_nblist(sb, self.x32)
        pot.aev_forward(sb, self.x32)
        pot.mlp(sb, with_grad=True)
        pot.aev_backward(sb, self.x32, zero=False)  # the integrator cleared grad_nbr

    def capture(self, steps_per_graph: int = 1, warmup: bool = True) -> None:
        """Record MD steps in HIP graphs (no host checks inside; construct with check_every != 1
        and use :meth:`run` with ``graph=True`` for checked runs).  Without a skin list one
        graph holds `steps_per_graph` steps; with one, one-step graphs of a full rebuild and of
        a skin re-filter are recorded, and for steps_per_graph > 1 also a graph of that many
        re-filter steps, which :meth:`replay` uses wherever the rebuild schedule has that many
        re-filters ahead.  ``warmup`` runs one (two with a skin) ordinary steps first, outside
        the capture."""
        if self.check_every == 1:
            raise ValueError("graph capture needs check_every != 1 (no host reads inside a step)")
        if steps_per_graph < 1:
            raise ValueError("steps_per_graph must be at least 1")
        skin = self.nblist_stride > 1
        timer, self.pot.timer = self.pot.timer, None
        if warmup:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):  # warm the allocator and kernels outside the capture
                self.step(check=False)
                if skin:
                    self.step(check=False)
            torch.cuda.current_stream().wait_stream(side)
        # (name, neighbour-list action, steps)
        kinds = [(True, True, 1), ("skin", "skin", 1)] if skin else [(True, True, steps_per_graph)]
        if skin and steps_per_graph > 1:
            kinds.append(("skin_run", "skin", steps_per_graph))
        graphs = {}
        for name, action, steps in kinds:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                for _ in range(steps):
                    self._record(action)
            graphs[name] = g
        self.graph = graphs[True]
        self.graph_skin = graphs.get("skin")
        self.graph_skin_run = graphs.get("skin_run")
        self.steps_per_graph = steps_per_graph
        self.pot.timer = timer

    def replay(self, n_graphs: int = 1, max_steps: int | None = None) -> int:
        """Replay ``n_graphs`` captured graphs, or fewer once ``max_steps`` steps ran; returns the
        number of MD steps advanced.  With a skin list each graph follows the rebuild schedule
        (:meth:`_graph_update`): a full rebuild or a re-filter step, or the graph of
        steps_per_graph re-filters where the schedule has that many ahead and as many steps are
        left."""
        done = 0
        S = self.steps_per_graph
        for _ in range(n_graphs):
            left = None if max_steps is None else max_steps - done
            if left is not None and left <= 0:
                break
            if self.graph_skin is None:
                self.graph.replay()
                k = S
            elif (self.graph_skin_run is not None and self.istep >= self.nblist_warmup
                  and self.nblist_countdown >= S and (left is None or left >= S)):
                self.graph_skin_run.replay()  # S re-filters: what S calls of _graph_update give
                self.nblist_countdown -= S
                k = S
            else:
                (self.graph if self._graph_update() is True else self.graph_skin).replay()
                k = 1
            self.istep += k
            done += k
        if done > 0 and self.sb.n_sys <= _abi.FNL_REDUCE_FUSED_MAX_SYSTEMS:
            self._sums_stale = True
        return done

    def _current_sums(self) -> None:
        """Form the per-system sums of the last replayed step (its graph leaves them to the next
        step's integrator launch)."""
        if self._sums_stale:
            self.pot.reduce(self.sb, with_virial=True)
            self._sums_stale = False

    # -------------------------------------------------------------- observables
    def energies(self) -> np.ndarray:
        """Potential energy per system in the model's energy unit."""
        self._current_sums()
        return self.sb.energy.cpu().numpy()

    def kinetic_tensor(self) -> np.ndarray:
        """0.5 sum m v v^T per system in kcal/mol (integrate.py:320-339)."""
        sb = self.sb
        out = torch.zeros((sb.n_sys, 9), dtype=torch.float64, device=self.pot.device)
        _abi.check(self.lib.fnl_system_reduce(sb.sys_offsets.data_ptr(), sb.seg_offsets.data_ptr(), sb.n_sys,
                                              sb.n_seg, SEG_ATOMS, None, None, None, None, self.v.data_ptr(),
                                              self.t_mass.data_ptr(), out.data_ptr(), sb.partials.data_ptr(),
                                              _abi.stream_handle()), "fnl_system_reduce")
        return out.cpu().numpy().reshape(-1, 3, 3)

    def temperature(self) -> np.ndarray:
        """2 K / (3 N k_B) per system (md/dynamic.py printing)."""
        ek = np.trace(self.kinetic_tensor(), axis1=1, axis2=2)
        return 2.0 * ek / (3.0 * np.asarray(self.sb.natoms) * units.K_B)

    def virial_kcal(self) -> np.ndarray:
        """dE/dstrain per system [B,3,3] in kcal/mol at the current positions."""
        self._current_sums()
        return self.sb.virial.cpu().numpy().reshape(-1, 3, 3) * self.force_scale

    def pressure_tensor(self) -> np.ndarray:
        """P = 2 K / V - virial / V per system [B,3,3] in kcal/mol/A^3 (md/integrate.py:364-371,
        the kinetic tensor of the stored mid-step velocities, pressure_o_weight = 1)."""
        cells = self.sb.extra.get("cells64")
        if cells is None:
            raise ValueError("the pressure needs periodic cells")
        vol = np.abs(np.linalg.det(cells))[:, None, None]
        return self.kinetic_tensor() * (2.0 / vol) - self.virial_kcal() / vol

    def pressure(self) -> np.ndarray:
        """Scalar pressure trace(P)/3 per system in kcal/mol/A^3 (multiply by
        ``units.us.get_multiplier('atm')`` for atm)."""
        return np.trace(self.pressure_tensor(), axis1=1, axis2=2) / 3.0

    def total_energy_kcal(self) -> np.ndarray:
        return self.energies() / self.model_energy_unit + np.trace(self.kinetic_tensor(), axis1=1, axis2=2)
