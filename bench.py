"""Benchmark: ANI energy/force + BAOAB Langevin MD on the 23,328-atom periodic water box.

One step = BAOAB integrate -> cell-list neighbour lists (rebuilt every step) ->
AEV -> species MLP (forward + input gradient) -> AEV adjoint (forces, virial)
-> FP64 per-system reductions, all on one GPU with inputs resident in HBM.
The timed steps replay a HIP graph of one MD step (one graph launch per step);
a separate eager run gives the per-stage breakdown.

Prints ONE JSON line.  ``value`` is the BF16-MLP form's atom-steps/s; the
plain FP32 form, their ratio, the per-step p50/p99 latency, the energy and
force errors of both forms against an FP64 evaluation, the roofline of the
dominant kernel and a CPU baseline (the NumPy oracle on the host cores) are
reported beside it, and under "skin" the same BF16 MD with a Verlet skin
(fennol_md's nblist_skin, 2 A by default: full rebuilds every 80 steps,
re-filters of the skin list in between).  "reference_box" times the setting
the reference's only published MD rates come from (216 waters, NVE, 0.1 fs)
next to that rate.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the stages that are one kernel launch each (candidates for the roofline entry)
KERNEL_OF_STAGE = {"aev_fwd": "k_aev_fwd", "mlp": "k_species_mlp", "aev_bwd": "k_aev_bwd"}
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
ISSUE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "issue.json")
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp32": 157.3}
SPECIES7 = ("H", "C", "N", "O", "S", "F", "Cl")
MODEL_SEED = 11
MD_SEED = 2024


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def workload(name: str) -> dict:
    from fennol_amd import systems

    if name == "water23k":
        return systems.water_config2()
    if name == "water7k":
        return systems.water_config1()
    raise SystemExit(f"unknown workload {name}")


def model_params(cfg):
    """Seeded random-init weights (truncated lecun normal, zero biases) for H, C, N, O networks."""
    from fennol_amd.params import init_species_params

    return init_species_params(MODEL_SEED, cfg.aev_length)


# Issue model of the AEV kernels: the VALU lane-operations one triplet and one radial slot need
# with every lane busy and no bookkeeping (per triplet: cosine, sine, switch product and mean
# distance, NA Gaussians, NZ powers, and the 32 channel products -- twice that with the
# derivatives and the dE/dAEV contraction in the adjoint; per radial slot: distance, switch and
# 16 shifted Gaussians, for the centre's and the neighbour's row in the adjoint)
VALU_MODEL = {"k_aev_fwd": (100, 90), "k_aev_bwd": (205, 180)}


def live_channels(cfg, species_mask: int) -> tuple[int, int]:
    """(radial, angular) AEV channels written for a batch whose species are the bits of the mask."""
    S = sum(1 for s in range(cfg.n_species) if (species_mask >> s) & 1)
    return S * cfg.radial_dist_divisions, S * (S + 1) // 2 * cfg.angular_dist_divisions * cfg.angle_sections


def algorithmic_units(sb, cfg, precision: str, compensated: bool = False, compensated_backward: bool | None = None) -> dict:
    """Per-launch algorithmic bytes / flops of each stage, from the measured neighbour counts and
    the channel blocks the batch's species use (the other blocks are never read or written).
    ``compensated`` (``compensated_backward``, default the same): the BF16 weights of the forward
    (backward) products are hi + lo pairs (two MFMA products each)."""
    nr = sb.n_nbr.cpu().numpy().astype(np.float64)
    na = sb.n_nbr_ang.cpu().numpy().astype(np.float64)
    N = sb.n_atoms
    rad, ang = live_channels(cfg, sb.extra.get("aev_mask", (1 << cfg.n_species) - 1))
    L = rad + ang
    out_b = 2 if precision == "bf16" else 4
    H = 256 * 192 + 192 * 160
    per_atom_mlp = 2.0 * (L * 256 + H + 160) + 2.0 * (H + 256 * L)  # forward and input gradient
    # executed on the MFMA: the 128 (BF16) / 64 (FP32) column chunks holding a live column (hi and
    # lo with compensated weights)
    kc = 128 if precision == "bf16" else 64
    live = cfg.live_column_blocks(sb.extra.get("aev_mask", 0xFFFFFFFF))
    pieces = sum(1 for q in range(min(64, cfg.k_pad // 32)) if (live >> q) & 1) + max(0, cfg.k_pad // 32 - 64)
    k_exec = cfg.k_pad if not live else kc * -(-pieces // (kc // 32))
    comp_b = compensated if compensated_backward is None else compensated_backward
    comp = 2.0 if precision == "bf16" and compensated else 1.0
    comp_b = 2.0 if precision == "bf16" and comp_b else 1.0
    exec_mlp = comp * 2.0 * (k_exec * 256 + H + 160) + comp_b * 2.0 * (H + 256 * k_exec)
    return {
        # xyz + species in, neighbour rows (radial + angular) and counts out
        "nblist": ("hbm", float(N * (12 + 4 + 8) + 4 * nr.sum() + 4 * na.sum())),
        # xyz, species, rows in; the live channel blocks of the AEV row out
        "aev_fwd": ("hbm", float(N * (12 + 4 + 8) + 4 * nr.sum() + 4 * na.sum() + N * L * out_b)),
        "mlp": ("mfma", float(N * per_atom_mlp)),
        # xyz, species, rows and the live dE/dAEV blocks in (plus each atom's radial blocks once,
        # which its neighbours gather); forces (self + int64 neighbour sums) and virial out
        "aev_bwd": ("hbm", float(N * (12 + 4 + 8) + 4 * nr.sum() + 4 * na.sum() + N * (L + rad) * out_b
                                 + N * (12 + 24 + 36))),
        "reduce": ("hbm", float(N * (4 + 36))),
        "baoab": ("hbm", float(N * (48 + 24 + 8 + 8 + 48 + 12))),
        "_mlp_executed": float(N * exec_mlp),
        "_counts": {"n_r": float(nr[:N].mean()), "n_a": float(na[:N].mean()),
                    "triplets": float((na[:N] * (na[:N] - 1) / 2).mean()), "live_channels": L},
    }


def measured_traffic(kernel: str, workload_name: str, precision: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter's reading for a plain copy of
    known size), or None when that kernel/workload has not been profiled."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    ent = table.get(f"{workload_name}/{precision}/{kernel}")
    return None if ent is None else float(ent["bytes_per_launch"])


def measured_issue(kernel: str, workload_name: str, precision: str):
    """VALU busy fraction, active lanes per VALU instruction and MFMA busy fraction of ``kernel``
    from the committed rocprofv3 issue-counter summary, or None when it was not profiled."""
    try:
        with open(ISSUE_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    return table.get(f"{workload_name}/{precision}/{kernel}")


def copy_calibration(mib: int) -> None:
    """One plain streaming kernel of known size (reads and writes ``mib`` MiB each), launched so
    that a counter run of this command can scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {mib} MiB read + {mib} MiB written by one elementwise kernel")
    del a, b


def run_md(system, cfg, params, precision, steps, warmup, tile_rows=0, stage_steps=20, skin=-1.0,
           compensated="forward", steps_per_graph=1):
    """Warm-up, then a per-stage breakdown (eager steps, HIP events around each stage on the
    launch stream), then the timed run: ``steps_per_graph`` MD steps per HIP graph (one with a
    skin list), replayed back to back until ``steps`` steps ran (the host enqueues one graph
    launch per replay).  Returns the potential, the MD state, the wall time of the timed steps,
    the per-step latencies (ms; a replay's time over its steps), the stage times and the eager
    wall time per step."""
    import torch

    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIPotential, StageTimer

    pot = ANIPotential(cfg, params, precision=precision, mlp_tile_rows=tile_rows, bf16_compensated=compensated)
    md = LangevinMD(pot, system, dt_fs=0.5, temperature=300.0, gamma_thz=1.0, seed=MD_SEED, check_every=0,
                    nblist_skin=skin)
    for _ in range(warmup):
        md.step()
    torch.cuda.synchronize()
    md.verify()
    timer = StageTimer()
    pot.timer = timer
    torch.cuda.synchronize()
    te0 = time.perf_counter()
    for _ in range(stage_steps):
        md.step()
    torch.cuda.synchronize()
    eager_ms = 1000.0 * (time.perf_counter() - te0) / stage_steps
    pot.timer = None
    stages = timer.times_ms()
    spg = steps_per_graph if skin > 0 or steps % steps_per_graph == 0 else 1
    md.capture(steps_per_graph=spg)
    md.replay(1)
    torch.cuda.synchronize()
    step_ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    sizes = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ev[0].record()
    done = 0
    while done < steps:  # one graph per replay: spg steps, or one rebuild / re-filter step
        k = md.replay(1, max_steps=steps - done)
        done += k
        sizes.append(k)
        step_ev[len(sizes)].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    md.verify()
    lat = np.array([step_ev[r].elapsed_time(step_ev[r + 1]) / k for r, k in enumerate(sizes)])
    return pot, md, (t1 - t0), lat, stages, eager_ms


def form_precision(form: str) -> tuple[str, bool | str]:
    """(precision, compensated weights) of a benchmarked form: fp32, bf16 (hi + lo BF16 weights in
    the forward products, BF16 in the backward ones), bf16_full (hi + lo in both) or bf16_plain
    (BF16 weights only)."""
    if form == "fp32":
        return "fp32", False
    return "bf16", {"bf16": "forward", "bf16_full": True, "bf16_plain": False}[form]


def gpu_energy_forces(system, x, cfg, params, form):
    """E [B] and F [N,3] of the device potential at coordinates x (model unit, eV)."""
    from fennol_amd.potential import ANIPotential

    precision, comp = form_precision(form)
    pot = ANIPotential(cfg, params, precision=precision, bf16_compensated=comp)
    E, F, _ = pot.energy_forces_virial(system["species"], x, system["natoms"], system.get("cells"))
    return E, F


def oracle_step(system, x, cfg, params):
    """The cpu_baseline leg: the FP64 NumPy oracle runs one MD step at x -- the reference's
    all-pairs neighbour list, AEV, MLP, forces and a BAOAB update -- and is timed; its energy
    and forces are the FP64 reference the GPU results are compared with."""
    import oracle as O

    p = O.AEVParams(species_order=cfg.species_order)
    model = O.Model(p, params)
    x = np.asarray(x, np.float64)
    t0 = time.perf_counter()
    g = O.build_graphs(x, system["natoms"], p, system.get("cells"), system.get("reciprocal_cells"))
    E, F, V, _ = O.energy_forces_virial(model, x, system["species"], system["natoms"], system.get("cells"),
                                        system.get("reciprocal_cells"), graphs=g)
    mass = O.ATOMIC_MASSES[system["species"]] / O.md_units.DA
    a1, a2 = O.langevin_coefficients(5e-4, 1.0, O.md_units.K_B * 300.0, mass)
    noise = O.philox_normals(MD_SEED, 0, x.shape[0])
    O.baoab_step(x, np.zeros_like(x), F / O.md_units.energy_multiplier("eV"), mass, 5e-4, a1, a2, noise)
    return (E, F), time.perf_counter() - t0


def errors(E, F, Er, Fr, n):
    return {"energy_err_mev_per_atom": float(1000.0 * np.abs(np.asarray(E) - np.asarray(Er)).max() / n),
            "force_rms_err_ev_per_a": float(np.sqrt(np.mean((np.asarray(F) - np.asarray(Fr)) ** 2)))}


def trajectory_checkpoints(system, cfg, params, steps: int, points) -> dict:
    """Coordinates of the BF16 Langevin trajectory of the benchmark (graph replay) at the given
    step numbers."""
    import torch

    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIPotential

    pot = ANIPotential(cfg, params, precision="bf16")
    md = LangevinMD(pot, system, dt_fs=0.5, temperature=300.0, gamma_thz=1.0, seed=MD_SEED, check_every=0)
    out = {0: md.x.cpu().numpy().copy()} if 0 in points else {}
    done = 0
    for p_ in sorted(k for k in points if k > 0):
        md.run(p_ - done, graph=True)
        done = p_
        torch.cuda.synchronize()
        out[p_] = md.x.cpu().numpy().copy()
    md.verify()
    del md, pot
    torch.cuda.empty_cache()
    return out


def accuracy_and_cpu_baseline(system, cfg, params, precisions, steps: int = 1000,
                              points=(0, 500, 1000)) -> tuple[dict, dict]:
    """Errors of every form against the FP64 oracle at checkpoints of a BF16 trajectory of
    ``steps`` steps (the oracle evaluated at the GPU trajectory's coordinates), and the CPU
    baseline: the median time of those oracle steps (one per checkpoint)."""
    n = len(system["species"])
    xs = trajectory_checkpoints(system, cfg, params, steps, points)
    rows, times = [], []
    for k in sorted(xs):
        log(f"[bench] FP64 oracle at trajectory step {k} (CPU)")
        (Er, Fr), t = oracle_step(system, xs[k], cfg, params)
        times.append(t)
        row = {"step": k}
        for prec in precisions:
            E, F = gpu_energy_forces(system, xs[k], cfg, params, prec)
            row[prec] = errors(E, F, Er, Fr, n)
        rows.append(row)
        log(f"[bench]   step {k}: " + ", ".join(f"{p} {row[p]['energy_err_mev_per_atom']:.4f} meV/atom "
                                                   f"{row[p]['force_rms_err_ev_per_a']:.2e} eV/A" for p in precisions))
    worst = {p: {key: max(r[p][key] for r in rows) for key in rows[0][p]} for p in precisions}
    acc = {"checkpoints": rows, "worst": worst,
           "trajectory": f"BF16 Langevin MD of the benchmark box, {steps} steps, dt 0.5 fs, 300 K"}
    med = float(np.median(times))
    threads = os.environ.get("OMP_NUM_THREADS")
    cpu = {"value": round(n / med, 1), "unit": "atom-steps/s",
           "cores": int(threads) if threads and threads.isdigit() else len(os.sched_getaffinity(0)),
           "kind": "port",
           "sample": (f"median of {len(times)} single MD steps (all-pairs neighbour list as the reference, AEV, "
                      f"MLP, forces, BAOAB) of the same {n}-atom box at steps {sorted(xs)} of the trajectory, "
                      f"FP64 NumPy oracle, {', '.join(f'{t:.1f}' for t in times)} s; element-wise NumPy on one "
                      f"thread, BLAS products on OMP_NUM_THREADS={threads}; "
                      f"{len(os.sched_getaffinity(0))} host cores available, CPU {cpu_model()}")}
    return acc, cpu


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_nve_drift(cfg, params, steps: int = 20000, every: int = 200) -> dict:
    """NVE energy conservation on the reference's watersmall box (its 216-water geometry, 18.643 A
    cube, dt 0.1 fs, initial velocities at 300 K, list rebuilt every step), FP32 and BF16:
    total energy sampled every ``every`` steps, drift = least-squares slope, reported per ns
    and per atom, for the reference-style total energy (kinetic energy of the stored mid-step
    velocities, as fennol_md prints) and the synchronised one (v + dt/2 F/m)."""
    import torch

    from fennol_amd import units
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIPotential

    ws = np.load(os.path.join(ROOT, "tests", "golden", "watersmall.npz"))
    system = {"species": ws["species"], "coordinates": ws["coordinates"], "cells": ws["cell"][None],
              "natoms": np.array([len(ws["species"])], np.int32)}
    n = len(ws["species"])
    out = {"workload": f"watersmall: 216 waters ({n} atoms) of the reference example, 18.643 A cube, NVE, "
                       f"dt 0.1 fs, {steps} steps ({steps * 1e-4:.1f} ps), seeded random-init weights",
           "reference": "examples/md/watersmall/conservation_tests/*/watersmall.log.bak (trained ANI-2x "
                        "weights, 1 ns): FP64 +6.7 ueV/atom/ns, FP32 highest -16 ueV/atom/ns, "
                        "high +0.18 meV/atom/ns, default +0.39 meV/atom/ns"}
    for prec in ("fp32", "bf16"):
        pot = ANIPotential(cfg, params, precision=prec)
        md = LangevinMD(pot, system, dt_fs=0.1, temperature=300.0, seed=MD_SEED, thermostat="NVE", check_every=0)
        mass = torch.tensor(md.mass, dtype=torch.float64, device="cuda")[:, None]
        t, e_ref, e_sync = [], [], []
        for k in range(steps // every + 1):
            if k:
                md.run(every, graph=True)
            torch.cuda.synchronize()
            ep = float(md.energies()[0]) / md.model_energy_unit
            v = md.v
            F = pot.forces(md.sb).double() * md.force_scale       # kcal/mol/A
            ek = float(0.5 * (mass * v * v).sum())
            vs = v + (0.5 * md.dt) * F / mass
            eks = float(0.5 * (mass * vs * vs).sum())
            t.append(md.istep * md.dt)                             # ps
            e_ref.append(ep + ek)
            e_sync.append(ep + eks)
        md.verify()
        t = np.array(t)
        ev = units.us.get_multiplier("eV")                         # kcal/mol -> eV
        res = {}
        for name, e in (("reference_style", np.array(e_ref)), ("synchronised", np.array(e_sync))):
            slope = np.polyfit(t, e, 1)[0]                          # kcal/mol per ps
            res[name] = {"drift_mev_per_atom_per_ns": round(float(slope * 1000.0 * ev * 1000.0 / n), 4),
                         "first_last_mev_per_atom": round(float((e[-1] - e[0]) * ev * 1000.0 / n), 4),
                         "fluct_rms_mev_per_atom": round(float(np.std(e - np.polyval(np.polyfit(t, e, 1), t))
                                                                * ev * 1000.0 / n), 4)}
        res["temperature_K_end"] = round(float(md.temperature()[0]), 1)
        out[prec] = res
        log(f"[bench]   NVE {prec}: drift {res['reference_style']['drift_mev_per_atom_per_ns']} (reference-style), "
            f"{res['synchronised']['drift_mev_per_atom_per_ns']} (synchronised) meV/atom/ns")
        del md, pot
        torch.cuda.empty_cache()
    return out


def bench_serving(cfg, params, n_requests: int, n_baseline: int, n_check: int = 24) -> dict:
    """Continuous batching of cluster requests (BF16 MLP) vs plain FP32 per-request launches:
    saturated (every request waiting at t=0) and Poisson arrivals at 70 % of the saturated rate;
    latency per request class; FP64 error of a sample of the served results."""
    import torch

    import oracle as O
    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.serving import make_requests, per_request_baseline, poisson_arrivals, serve_stream

    cfg4 = ANIConfig()  # 4-species AEV (clusters of H, C, N, O)
    from fennol_amd.params import init_species_params

    p4 = init_species_params(MODEL_SEED, cfg4.aev_length)
    reqs = make_requests(n_requests, seed=3)
    pot = ANIPotential(cfg4, p4, precision="bf16")
    pool = dict(max_atoms=65536, max_systems=2048)
    log(f"[bench] serving: {n_requests} requests, saturated")
    sat = serve_stream(pot, reqs, None, **pool)
    rate = 0.7 * sat["requests_per_s"]
    log(f"[bench] serving: {sat['requests_per_s']:.0f} req/s, {sat['atom_steps_per_s'] / 1e6:.1f} M atom-steps/s; "
        f"Poisson arrivals at {rate:.1f} req/s")
    poi = serve_stream(pot, reqs, poisson_arrivals(len(reqs), rate, seed=1), **pool)
    # FP64 error of served results: energy + forces requests at their inputs, MD requests at the
    # positions they returned
    served = poi["_requests"]
    model = O.Model(O.AEVParams(species_order=cfg4.species_order), p4)
    rng = np.random.Generator(np.random.PCG64(7))
    pick = rng.choice(len(served), size=min(n_check, len(served)), replace=False)
    e_err, f_err = {"ef": [], "md": []}, {"ef": [], "md": []}
    for k in pick:
        r = served[int(k)]
        x = r.positions if r.md_steps else np.asarray(r.coordinates)
        x64 = np.asarray(x, np.float32).astype(np.float64)
        Er, Fr, _, _ = O.energy_forces_virial(model, x64, r.species, np.array([r.n]))
        e_err[r.kind].append(1000.0 * abs(r.energy - Er[0]) / r.n)
        f_err[r.kind].append(float(np.sqrt(np.mean((r.forces - Fr) ** 2))))
    accuracy = {kind: {"requests": len(e_err[kind]),
                       "max_energy_err_mev_per_atom": round(max(e_err[kind]), 5) if e_err[kind] else None,
                       "max_force_rms_err_ev_per_a": round(max(f_err[kind]), 5) if f_err[kind] else None}
                for kind in ("ef", "md")}
    del pot
    torch.cuda.empty_cache()
    log(f"[bench] serving: FP32 per-request baseline on {n_baseline} requests")
    pot32 = ANIPotential(cfg4, p4, precision="fp32")
    base = per_request_baseline(pot32, reqs[:n_baseline])
    keep = ("wall_s", "requests", "atom_steps", "atom_steps_per_s", "requests_per_s", "iterations", "layouts",
            "restarts", "regrowths", "graphs", "p50_ms", "p99_ms", "ef_requests", "ef_p50_ms", "ef_p99_ms",
            "md_requests", "md_p50_ms", "md_p99_ms", "service_p50_ms", "service_p99_ms")
    rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items() if k in keep}
    return {"workload": f"{n_requests} cluster requests, 8-512 atoms log-uniform, 30% 200-step Langevin, "
                        f"70% single energy+forces",
            "pool": {"max_atoms": sat["max_atoms"], "max_systems": sat["max_systems"],
                     "hbm_bytes": int(sat["pool_bytes"]), "bytes_per_atom_row": round(sat["row_bytes"], 1),
                     "bytes_per_request_slot": round(sat["sys_bytes"], 1),
                     "admission": "FIFO while the request's rows and slot fit the pool's HBM bytes"},
            "bf16_saturated": rnd(sat),
            "bf16_poisson": {**rnd(poi), "arrival_rate_per_s": round(rate, 2)},
            "accuracy_vs_fp64": {**accuracy, "sample": f"{len(pick)} random served requests of the Poisson run"},
            "fp32_per_request": rnd(base),
            "speedup_atom_steps": round(sat["atom_steps_per_s"] / base["atom_steps_per_s"], 2)}


# The reference's own MD throughput figures: FP32 NVE conservation runs of the same ANI-2x
# architecture on a 216-water box (examples/md/watersmall/conservation_tests/*/watersmall.log.bak,
# unstated NVIDIA GPU): 689.6 steps/s median with matmul precision "highest".
REFERENCE_BOX_FP32_ATOM_STEPS = 689.6 * 648


def bench_reference_box(cfg, params, steps: int = 2000, warmup: int = 20) -> dict:
    """The reference's watersmall setting: 216 waters (648 atoms) in an 18.64 A cube, NVE with a
    0.1 fs step and the neighbour list rebuilt every step, one MD step per HIP graph replay."""
    import torch

    from fennol_amd import systems
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIPotential

    ws = np.load(os.path.join(ROOT, "tests", "golden", "watersmall.npz"))
    system = {"species": ws["species"], "coordinates": ws["coordinates"], "cells": ws["cell"][None],
              "reciprocal_cells": np.linalg.inv(ws["cell"])[None], "natoms": np.array([len(ws["species"])], np.int32)}
    n = int(system["natoms"].sum())
    out = {"workload": f"the reference's watersmall box: its 216-water geometry ({n} atoms), "
                       f"{float(system['cells'][0, 0, 0]):.3f} A cube, NVE, dt 0.1 fs, neighbour list rebuilt "
                       "every step (the reference's watersmall conservation runs)",
           "reference_fp32_atom_steps_per_s": round(REFERENCE_BOX_FP32_ATOM_STEPS, 1),
           "reference_source": "examples/md/watersmall/conservation_tests/ani2x_dt0.1_highest_noskin/"
                               "watersmall.log.bak (unstated NVIDIA GPU)"}
    for prec in ("fp32", "bf16"):
        pot = ANIPotential(cfg, params, precision=prec)
        md = LangevinMD(pot, system, dt_fs=0.1, temperature=300.0, seed=MD_SEED, thermostat="NVE", check_every=0)
        for _ in range(warmup):
            md.step()
        md.capture(steps_per_graph=1)
        md.replay(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        md.replay(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        md.verify()
        out[prec] = {"ms_per_step": round(1000.0 * wall / steps, 4), "atom_steps_per_s": round(n * steps / wall, 1)}
        # error against FP64 at the end of the run (the oracle at the same coordinates)
        x_end = md.x.cpu().numpy()
        (Er, Fr), _ = oracle_step(system, x_end, cfg, params)
        E, F = gpu_energy_forces(system, x_end, cfg, params, prec)
        out[prec]["vs_fp64_at_end"] = errors(E, F, Er, Fr, n)
        del pot, md
    out["fp32_vs_reference"] = round(out["fp32"]["atom_steps_per_s"] / REFERENCE_BOX_FP32_ATOM_STEPS, 2)
    return out


def bench_training(n_boxes: int, steps: int) -> dict:
    """cfg 4: energy + force loss and weight gradient through forces on boxes of 64 waters."""
    import torch

    from fennol_amd import systems
    from fennol_amd.params import init_species_params
    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.training import TrainingStep

    cfg = ANIConfig(species_order=SPECIES7)
    params = init_species_params(MODEL_SEED, cfg.aev_length)
    ref = init_species_params(MODEL_SEED + 1, cfg.aev_length)
    batch = systems.training_boxes(n_boxes=n_boxes, seed=4)
    N = len(batch["species"])
    ref_pot = ANIPotential(cfg, ref, precision="fp32")
    e_ref, f_ref, _ = ref_pot.energy_forces_virial(batch["species"], batch["coordinates"], batch["natoms"],
                                                   batch["cells"])
    del ref_pot
    pot = ANIPotential(cfg, params, precision="fp32")
    sb = pot.setup(batch["species"], batch["natoms"], batch["cells"])
    x32 = torch.tensor(np.asarray(batch["coordinates"], np.float32), device="cuda")
    ts = TrainingStep(pot, params)
    er = torch.tensor(e_ref, device="cuda")
    fr = torch.tensor(f_ref, device="cuda")
    ts(sb, x32, er, fr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, grads, _, _ = ts(sb, x32, er, fr)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"workload": f"{n_boxes} periodic boxes x 64 waters ({N} atoms), loss = MSE(E) + 10 MSE(F)",
            "ms_per_step": round(1000 * dt, 3), "atoms_per_s": round(N / dt, 1), "loss": loss}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--steps-per-graph", type=int, default=10,
                    help="MD steps captured in one HIP graph (1 if it does not divide --steps)")
    ap.add_argument("--workload", default="water23k", choices=["water23k", "water7k"])
    ap.add_argument("--skip-cpu", action="store_true", help="skip the FP64 oracle (accuracy and CPU baseline)")
    ap.add_argument("--serving-requests", type=int, default=8192, help="0 skips the serving measurement")
    ap.add_argument("--serving-baseline", type=int, default=64)
    ap.add_argument("--training-boxes", type=int, default=256, help="0 skips the training measurement")
    ap.add_argument("--tile-bf16", type=int, default=0, help="MLP atoms per tile, BF16 form (0: default)")
    ap.add_argument("--tile-fp32", type=int, default=0, help="MLP atoms per tile, FP32 form (0: default)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="launch one plain copy kernel of this size first (for counter runs)")
    ap.add_argument("--precisions", default="fp32,bf16,bf16_plain",
                    help="MD forms to run (bf16 gives value; its weights are hi + lo BF16 pairs in the forward "
                         "products; bf16_full: pairs in the backward products too; bf16_plain: BF16 weights only)")
    ap.add_argument("--no-reference-box", dest="reference_box", action="store_false",
                    help="skip the 216-water NVE run that the reference's published MD rate is quoted on")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="length of the BF16 trajectory whose start, middle and end are checked against FP64")
    ap.add_argument("--nve-steps", type=int, default=20000,
                    help="NVE steps (dt 0.1 fs) on the watersmall box for the energy drift; 0 skips")
    ap.add_argument("--skin", type=float, default=2.0,
                    help="also time BF16 MD with a Verlet skin of this many A (fennol_md's nblist_skin; "
                         "rebuild stride from the skin); 0 skips")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (single-GPU build)",
              file=sys.stderr)
        return 2

    import torch

    from fennol_amd.potential import ANIConfig

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    system = workload(args.workload)
    N = len(system["species"])
    cfg = ANIConfig(species_order=SPECIES7)
    params = model_params(cfg)
    log(f"[bench] workload {args.workload}: {N} atoms, AEV {cfg.aev_length}")

    if args.calibrate_copy_mib > 0:
        copy_calibration(args.calibrate_copy_mib)
    precisions = [p for p in ("fp32", "bf16", "bf16_full", "bf16_plain") if p in args.precisions.split(",")]
    if "bf16" not in precisions:
        print("the bf16 form gives the headline value; --precisions must include it", file=sys.stderr)
        return 2
    results = {}
    for form in precisions:
        log(f"[bench] MD {form}: warmup {args.warmup}, steps {args.steps}")
        precision, comp = form_precision(form)
        tile = args.tile_bf16 if precision == "bf16" else args.tile_fp32
        pot, md, wall, lat, stages, eager_ms = run_md(system, cfg, params, precision, args.steps, args.warmup,
                                                      tile_rows=tile, compensated=comp,
                                                      steps_per_graph=args.steps_per_graph)
        units = algorithmic_units(md.sb, cfg, precision, pot.nets.compensated, pot.nets.compensated_backward)
        results[form] = dict(wall=wall, lat=lat, stages=stages, units=units, eager_ms=eager_ms,
                             max_nbr=md.sb.max_nbr, max_nbr_ang=md.sb.max_nbr_ang)
        log(f"[bench]   {form}: {1000 * wall / args.steps:.3f} ms/step (graph replay), eager "
            f"{eager_ms:.3f} ms/step; stages (ms): " + ", ".join(f"{k} {v.mean():.3f}" for k, v in stages.items()))
        del pot, md
        torch.cuda.empty_cache()

    skin = None
    if args.skin > 0:
        from fennol_amd.md import nblist_schedule

        stride = nblist_schedule(args.skin, 0.5)
        k_skin = max(args.steps, 2 * stride)
        log(f"[bench] MD bf16 with a {args.skin} A skin (rebuild every {stride} steps): {k_skin} steps")
        pot, md, wall, lat, stages, eager_ms = run_md(system, cfg, params, "bf16", k_skin, args.warmup,
                                                      tile_rows=args.tile_bf16, skin=args.skin,
                                                      steps_per_graph=args.steps_per_graph)
        skin = {"nblist_skin_A": args.skin, "rebuild_stride": stride, "steps": k_skin,
                "steps_per_graph": args.steps_per_graph,
                "value": round(N * k_skin / wall, 1), "unit": "atom-steps/s",
                "ms_per_step": round(1000.0 * wall / k_skin, 4),
                "p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
                "refilter_ms": round(float(stages["skin"].mean()), 4) if "skin" in stages else None,
                "max_skin": md.sb.max_skin}
        # the re-filtered lists are the full build's lists (Verlet guarantee): the last step's
        # energy and forces against a fresh full build at the same coordinates, so the errors
        # against FP64 are those of the rebuild-every-step form (accuracy_vs_fp64)
        x_end = md.x32.cpu().numpy().astype(np.float64)
        E_skin, F_skin = md.energies(), pot.forces(md.sb).cpu().numpy()
        E_full, F_full = gpu_energy_forces(system, x_end, cfg, params, "bf16")
        skin["vs_full_rebuild"] = {"max_abs_energy_diff_ev": float(np.abs(E_skin - E_full).max()),
                                   "max_abs_force_diff_ev_per_a": float(np.abs(F_skin - F_full).max())}
        log(f"[bench]   skin: {skin['ms_per_step']:.3f} ms/step, re-filter {skin['refilter_ms']} ms, "
            f"vs full rebuild {skin['vs_full_rebuild']}")
        del pot, md
        torch.cuda.empty_cache()

    bf = results["bf16"]
    fp = results.get("fp32")
    cp = results.get("bf16_plain")
    cf = results.get("bf16_full")
    ms_bf = 1000.0 * bf["wall"] / args.steps
    value = N * args.steps / bf["wall"]

    # roofline of the dominant kernel of the headline (BF16) form: the longest single-kernel
    # stage, its algorithmic bytes (or flops) per launch over its mean launch time (HIP events on
    # the launch stream)
    st = bf["stages"]
    dom = max((k for k in KERNEL_OF_STAGE if k in st), key=lambda k: st[k].mean())
    bound, work = bf["units"][dom]
    t_s = st[dom].mean() / 1000.0
    if bound == "hbm":
        achieved, peak, unit = work / t_s / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, unit = work / t_s / 1e12, MFMA_PEAK_TFLOPS["bf16"], "TFLOP/s"
    traffic = measured_traffic(KERNEL_OF_STAGE[dom], args.workload, "bf16")
    roofline = {"bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": KERNEL_OF_STAGE[dom],
                "launch_us": round(1e6 * t_s, 2), "algorithmic_per_launch": work}
    issue = measured_issue(KERNEL_OF_STAGE[dom], args.workload, "bf16")
    if issue is not None:  # what actually limits the kernel: its instruction issue (counters)
        roofline["issue"] = dict(issue)
        kern = KERNEL_OF_STAGE[dom]
        if kern in VALU_MODEL and "valu_insts" in issue:
            c = bf["units"]["_counts"]
            per_t, per_slot = VALU_MODEL[kern]
            model = (per_t * c["triplets"] + per_slot * c["n_r"]) / 64.0
            measured = issue["valu_insts"] / N
            roofline["issue"].update({
                "valu_insts_per_centre": round(measured, 1), "model_insts_per_centre": round(model, 1),
                "model_fraction": round(model / measured, 3),
                "model": f"{per_t} lane-ops per triplet ({c['triplets']:.1f} per centre) + {per_slot} per radial "
                         f"slot ({c['n_r']:.1f}), 64 lanes per instruction"})
    stage_table = {}
    for prec, r in results.items():
        stage_table[prec] = {}
        for k, v in r["stages"].items():
            ent = {"ms": round(float(v.mean()), 4)}
            if k in r["units"]:
                b, w = r["units"][k]
                if b == "hbm":
                    ent["GB/s"] = round(w / (v.mean() / 1000.0) / 1e9, 1)
                else:
                    ent["TFLOP/s"] = round(w / (v.mean() / 1000.0) / 1e12, 2)
                    if k == "mlp":  # the MFMA work actually run (live chunks; hi and lo in BF16)
                        ent["executed_TFLOP/s"] = round(r["units"]["_mlp_executed"] / (v.mean() / 1000.0) / 1e12, 2)
            stage_table[prec][k] = ent

    acc = {}
    cpu = None
    if not args.skip_cpu:
        log("[bench] FP64 oracle on the host at checkpoints of a 1,000-step BF16 trajectory "
            "(accuracy reference + CPU baseline)")
        acc, cpu = accuracy_and_cpu_baseline(system, cfg, params, precisions, steps=args.accuracy_steps,
                                             points=(0, args.accuracy_steps // 2, args.accuracy_steps))
    nve = bench_nve_drift(cfg, params, steps=args.nve_steps) if args.nve_steps > 0 else None
    serving = bench_serving(cfg, params, args.serving_requests, args.serving_baseline) \
        if args.serving_requests > 0 else None
    training = bench_training(args.training_boxes, 3) if args.training_boxes > 0 else None
    ref_box = None
    if args.reference_box:
        log("[bench] reference box: 216 waters, NVE, dt 0.1 fs")
        ref_box = bench_reference_box(cfg, params)
        log(f"[bench]   fp32 {ref_box['fp32']['ms_per_step']} ms/step ({ref_box['fp32_vs_reference']}x the "
            f"reference's published FP32 rate), bf16 {ref_box['bf16']['ms_per_step']} ms/step")
    out = {
        "metric": "atom-steps/s and p50/p99 latency, batched BF16 vs plain FP32, at <=1 meV/atom err",
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_bf, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded water box, seeded random-init weights)",
        "config": {"workload": f"periodic water box, {N // 3} waters ({N} atoms), min-image, neighbour list "
                               f"rebuilt every step, BAOAB Langevin dt=0.5 fs 300 K gamma=1/ps",
                   "aev_length": cfg.aev_length, "mlp": "256-192-160-1 CELU(0.1), 4 species networks",
                   "precision": "BF16 MLP GEMMs on MFMA (FP32 accumulate; weights as BF16 hi + lo pairs in the "
                                "forward products, BF16 in the backward products), FP32 AEV, FP64 reductions and "
                                "integrator",
                   "steps_per_graph": args.steps_per_graph if args.steps % args.steps_per_graph == 0 else 1},
        "eager_ms_per_step": round(bf["eager_ms"], 4),
        "fp32": None if fp is None else {
            "value": round(N * args.steps / fp["wall"], 1), "ms_per_step": round(1000.0 * fp["wall"] / args.steps, 4),
            "eager_ms_per_step": round(fp["eager_ms"], 4),
            "p50_ms": round(float(np.percentile(fp["lat"], 50)), 4),
            "p99_ms": round(float(np.percentile(fp["lat"], 99)), 4)},
        "bf16_over_fp32": None if fp is None else round(fp["wall"] / bf["wall"], 3),
        "bf16_plain": None if cp is None else {
            "weights": "BF16 weights only (one MFMA product each); its FP64 errors are in accuracy_vs_fp64",
            "value": round(N * args.steps / cp["wall"], 1), "ms_per_step": round(1000.0 * cp["wall"] / args.steps, 4),
            "p50_ms": round(float(np.percentile(cp["lat"], 50)), 4),
            "p99_ms": round(float(np.percentile(cp["lat"], 99)), 4)},
        **({} if cf is None else {"bf16_full": {
            "weights": "hi + lo BF16 pairs in the forward and the backward products; its FP64 errors are in "
                       "accuracy_vs_fp64",
            "value": round(N * args.steps / cf["wall"], 1), "ms_per_step": round(1000.0 * cf["wall"] / args.steps, 4),
            "p50_ms": round(float(np.percentile(cf["lat"], 50)), 4),
            "p99_ms": round(float(np.percentile(cf["lat"], 99)), 4)}}),
        "p50_ms": round(float(np.percentile(bf["lat"], 50)), 4),
        "p99_ms": round(float(np.percentile(bf["lat"], 99)), 4),
        "accuracy_vs_fp64": acc,
        "stages": stage_table,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "skin": skin,
        "serving": serving,
        "training": training,
        "reference_box": ref_box,
        "nve_drift": nve,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
