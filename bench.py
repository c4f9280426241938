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


def algorithmic_units(sb, cfg, precision: str) -> dict:
    """Per-launch algorithmic bytes / flops of each stage, from the measured neighbour counts."""
    nr = sb.n_nbr.cpu().numpy().astype(np.float64)
    na = sb.n_nbr_ang.cpu().numpy().astype(np.float64)
    N = sb.n_atoms
    A = cfg.aev_length
    out_b = 2 if precision == "bf16" else 4
    per_atom_mlp = 2.0 * (A * 256 + 256 * 192 + 192 * 160 + 160)   # forward
    per_atom_mlp += 2.0 * (160 * 192 + 192 * 256 + 256 * A)         # input gradient
    return {
        # xyz + species in, neighbour rows (radial + angular) and counts out
        "nblist": ("hbm", float(N * (12 + 4 + 8) + 4 * nr.sum() + 4 * na.sum())),
        # xyz, rows in; AEV row out
        "aev_fwd": ("hbm", float(N * (12 + 8) + 4 * nr.sum() + 4 * na.sum() + N * A * out_b)),
        "mlp": ("mfma", float(N * per_atom_mlp)),
        # dE/dAEV in (fp32, or bf16 in the BF16 form), rows in; forces (self + neighbour) and virial out
        "aev_bwd": ("hbm", float(N * (12 + 8) + 4 * nr.sum() + 4 * na.sum() + N * A * out_b + N * (24 + 36))),
        "reduce": ("hbm", float(N * (4 + 36))),
        "baoab": ("hbm", float(N * (48 + 24 + 8 + 8 + 48 + 12))),
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


def run_md(system, cfg, params, precision, steps, warmup, tile_rows=0, stage_steps=20, skin=-1.0):
    """Warm-up, then a per-stage breakdown (eager steps, HIP events around each stage on the
    launch stream), then the timed run: one MD step per HIP graph, replayed ``steps`` times back
    to back (the host enqueues one graph launch per step).  Returns the potential, the MD state,
    the wall time of the timed steps, the per-step latencies (ms), the stage times and the eager
    wall time per step."""
    import torch

    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIPotential, StageTimer

    pot = ANIPotential(cfg, params, precision=precision, mlp_tile_rows=tile_rows)
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
    md.capture(steps_per_graph=1)
    md.replay(1)
    torch.cuda.synchronize()
    step_ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ev[0].record()
    for k in range(steps):
        md.replay(1)
        step_ev[k + 1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    md.verify()
    lat = np.array([step_ev[k].elapsed_time(step_ev[k + 1]) for k in range(steps)])
    return pot, md, (t1 - t0), lat, stages, eager_ms


def accuracy(system, cfg, params, precision, fp64):
    """Energy (meV/atom) and force-RMS (eV/A) errors against the FP64 evaluation."""
    from fennol_amd.potential import ANIPotential

    pot = ANIPotential(cfg, params, precision=precision)
    E, F, V = pot.energy_forces_virial(system["species"], system["coordinates"], system["natoms"],
                                       system.get("cells"))
    Er, Fr = fp64
    n = len(system["species"])
    return {"energy_err_mev_per_atom": float(1000.0 * abs(E[0] - Er[0]) / n),
            "force_rms_err_ev_per_a": float(np.sqrt(np.mean((F - Fr) ** 2)))}


def cpu_reference(system, cfg, params):
    """The cpu_baseline leg: the FP64 NumPy oracle runs one MD step (energy, forces, BAOAB) on
    the host cores and is timed; its energy and forces are also the FP64 accuracy reference."""
    import oracle as O

    p = O.AEVParams(species_order=cfg.species_order)
    model = O.Model(p, params)
    x = np.asarray(system["coordinates"], np.float64)
    t0 = time.perf_counter()
    g = O.build_graphs(x, system["natoms"], p, system.get("cells"), system.get("reciprocal_cells"))
    E, F, V, _ = O.energy_forces_virial(model, x, system["species"], system["natoms"], system.get("cells"),
                                        system.get("reciprocal_cells"), graphs=g)
    mass = O.ATOMIC_MASSES[system["species"]] / O.md_units.DA
    a1, a2 = O.langevin_coefficients(5e-4, 1.0, O.md_units.K_B * 300.0, mass)
    noise = O.philox_normals(MD_SEED, 0, x.shape[0])
    O.baoab_step(x, np.zeros_like(x), F / O.md_units.energy_multiplier("eV"), mass, 5e-4, a1, a2, noise)
    t1 = time.perf_counter()
    return (E, F), t1 - t0


def bench_serving(cfg, params, n_requests: int, n_baseline: int) -> dict:
    """cfg 3: continuous batching (BF16 MLP) vs plain FP32 per-request launches."""
    import torch

    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.serving import make_requests, per_request_baseline, poisson_arrivals, serve_stream

    cfg4 = ANIConfig()  # 4-species AEV (clusters of H, C, N, O)
    from fennol_amd.params import init_species_params

    p4 = init_species_params(MODEL_SEED, cfg4.aev_length)
    reqs = make_requests(n_requests, seed=3)
    pot = ANIPotential(cfg4, p4, precision="bf16")
    log(f"[bench] serving: {n_requests} requests, saturated")
    sat = serve_stream(pot, reqs, None, max_atoms=65536, max_systems=2048)
    rate = 0.7 * sat["requests_per_s"]
    log(f"[bench] serving: Poisson arrivals at {rate:.1f} req/s")
    poi = serve_stream(pot, reqs, poisson_arrivals(len(reqs), rate, seed=1), max_atoms=65536, max_systems=2048)
    del pot
    torch.cuda.empty_cache()
    log(f"[bench] serving: FP32 per-request baseline on {n_baseline} requests")
    pot32 = ANIPotential(cfg4, p4, precision="fp32")
    base = per_request_baseline(pot32, reqs[:n_baseline])
    keep = ("wall_s", "requests", "atom_steps", "atom_steps_per_s", "requests_per_s", "iterations", "p50_ms",
            "p99_ms", "ef_p50_ms", "ef_p99_ms", "md_p50_ms", "md_p99_ms", "service_p50_ms", "service_p99_ms")
    rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items() if k in keep}
    return {"workload": f"{n_requests} cluster requests, 8-512 atoms log-uniform, 30% 200-step Langevin, "
                        f"70% single energy+forces", "bf16_saturated": rnd(sat),
            "bf16_poisson": {**rnd(poi), "arrival_rate_per_s": round(rate, 2)},
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

    system = systems.water_box(216, 6, seed=5)
    n = int(system["natoms"].sum())
    out = {"workload": f"216 waters ({n} atoms), {float(system['cells'][0, 0, 0]):.2f} A cube, NVE, dt 0.1 fs, "
                       "neighbour list rebuilt every step (the reference's watersmall conservation runs)",
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
    ap.add_argument("--workload", default="water23k", choices=["water23k", "water7k"])
    ap.add_argument("--skip-cpu", action="store_true", help="skip the FP64 oracle (accuracy and CPU baseline)")
    ap.add_argument("--serving-requests", type=int, default=1024, help="0 skips the serving measurement")
    ap.add_argument("--serving-baseline", type=int, default=64)
    ap.add_argument("--training-boxes", type=int, default=256, help="0 skips the training measurement")
    ap.add_argument("--tile-bf16", type=int, default=0, help="MLP atoms per tile, BF16 form (0: default)")
    ap.add_argument("--tile-fp32", type=int, default=0, help="MLP atoms per tile, FP32 form (0: default)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="launch one plain copy kernel of this size first (for counter runs)")
    ap.add_argument("--precisions", default="fp32,bf16", help="MD forms to run (bf16 gives value)")
    ap.add_argument("--no-reference-box", dest="reference_box", action="store_false",
                    help="skip the 216-water NVE run that the reference's published MD rate is quoted on")
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
    precisions = [p for p in ("fp32", "bf16") if p in args.precisions.split(",")]
    if "bf16" not in precisions:
        print("the bf16 form gives the headline value; --precisions must include it", file=sys.stderr)
        return 2
    results = {}
    for precision in precisions:
        log(f"[bench] MD {precision}: warmup {args.warmup}, steps {args.steps}")
        tile = args.tile_bf16 if precision == "bf16" else args.tile_fp32
        pot, md, wall, lat, stages, eager_ms = run_md(system, cfg, params, precision, args.steps, args.warmup,
                                                      tile_rows=tile)
        units = algorithmic_units(md.sb, cfg, precision)
        results[precision] = dict(wall=wall, lat=lat, stages=stages, units=units, eager_ms=eager_ms,
                                  max_nbr=md.sb.max_nbr, max_nbr_ang=md.sb.max_nbr_ang)
        log(f"[bench]   {precision}: {1000 * wall / args.steps:.3f} ms/step (graph replay), eager "
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
                                                      tile_rows=args.tile_bf16, skin=args.skin)
        skin = {"nblist_skin_A": args.skin, "rebuild_stride": stride, "steps": k_skin,
                "value": round(N * k_skin / wall, 1), "unit": "atom-steps/s",
                "ms_per_step": round(1000.0 * wall / k_skin, 4),
                "p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
                "refilter_ms": round(float(stages["skin"].mean()), 4) if "skin" in stages else None,
                "max_skin": md.sb.max_skin}
        log(f"[bench]   skin: {skin['ms_per_step']:.3f} ms/step, re-filter {skin['refilter_ms']} ms")
        del pot, md
        torch.cuda.empty_cache()

    bf = results["bf16"]
    fp = results.get("fp32")
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
        roofline["issue"] = issue
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
            stage_table[prec][k] = ent

    acc = {}
    cpu = None
    if not args.skip_cpu:
        log("[bench] FP64 oracle on the host (accuracy reference + CPU baseline)")
        fp64, t_cpu = cpu_reference(system, cfg, params)
        for precision in precisions:
            acc[precision] = accuracy(system, cfg, params, precision, fp64)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": round(N / t_cpu, 1), "unit": "atom-steps/s", "cores": cores, "kind": "port",
               "sample": f"1 MD step (all-pairs nblist, AEV, MLP, forces, BAOAB) of the same {N}-atom box, "
                         f"FP64 NumPy oracle, {t_cpu:.1f} s"}
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
                   "precision": "BF16 MLP GEMMs (FP32 accumulate), FP32 AEV, FP64 reductions and integrator"},
        "eager_ms_per_step": round(bf["eager_ms"], 4),
        "fp32": None if fp is None else {
            "value": round(N * args.steps / fp["wall"], 1), "ms_per_step": round(1000.0 * fp["wall"] / args.steps, 4),
            "eager_ms_per_step": round(fp["eager_ms"], 4),
            "p50_ms": round(float(np.percentile(fp["lat"], 50)), 4),
            "p99_ms": round(float(np.percentile(fp["lat"], 99)), 4)},
        "bf16_over_fp32": None if fp is None else round(fp["wall"] / bf["wall"], 3),
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
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
