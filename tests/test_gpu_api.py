"""GPU tests of the FeNNol-shaped API: FENNIX wrapper, preprocessing dictionaries, MD driver."""

from __future__ import annotations

from collections import OrderedDict

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd import systems
from fennol_amd.fennix import FENNIX
from tests.helpers import canonical_pairs_oracle, water_box_small

pytestmark = pytest.mark.gpu


def ani_model(species=("H", "C", "N", "O"), precision="fp32", seed=3):
    return FENNIX(
        cutoff=5.2,
        modules=OrderedDict(
            embedding={"module_name": "ANI_AEV", "species_order": list(species), "graph_angle_key": "graph_angle",
                       "radial_eta": 19.7, "angular_eta": 12.5, "radial_dist_divisions": 16,
                       "angular_dist_divisions": 4, "zeta": 14.1, "angle_sections": 8},
            energy={"module_name": "SPECIES_INDEX_NET", "output_dim": 1, "hidden_neurons": [256, 192, 160],
                    "species_order": ["H", "C", "N", "O"], "activation": "partial(celu,alpha=0.1)",
                    "input_key": "embedding", "squeeze": True}),
        preprocessing=OrderedDict(
            graph_angle={"module_name": "GRAPH_FILTER", "cutoff": 3.5, "parent_graph": "graph",
                         "graph_key": "graph_angle"},
            angle_extension={"module_name": "GRAPH_ANGULAR_EXTENSION", "graph_key": "graph_angle"},
            species_indexer={"module_name": "SPECIES_INDEXER", "output_key": "species_index"}),
        energy_terms=["energy"], energy_unit="eV", seed=seed, precision=precision)


def test_fennix_energy_forces_virial_match_oracle(gpu):
    print("[api] FENNIX cluster", flush=True)
    s = systems.cluster_config0(5)
    model = ani_model()
    e, f, v, out = model.energy_and_forces_and_virial(species=s["species"], coordinates=s["coordinates"])
    p = O.AEVParams()
    m = O.Model(p, model.params)
    x32 = np.asarray(s["coordinates"], np.float32).astype(np.float64)
    Er, Fr, Vr, eat = O.energy_forces_virial(m, x32, s["species"])
    np.testing.assert_allclose(e, Er, rtol=1e-5)
    assert np.sqrt(np.mean((f - Fr) ** 2)) < 1e-4
    np.testing.assert_allclose(out["energy"], eat, rtol=1e-4, atol=1e-5)
    # unit conversion as in FENNIX.energy_and_forces (fennix.py:517-524)
    e2, f2, _ = model.energy_and_forces(unit="kcalpermol", species=s["species"], coordinates=s["coordinates"])
    np.testing.assert_allclose(e2, e * 627.5096080305927 / 27.211386024367243, rtol=1e-12)
    # gradient function (MD's energy_and_gradient)
    gfun = model.get_gradient_function("coordinates", "strain")
    eg, de, _ = gfun(model.variables, {"species": s["species"], "coordinates": s["coordinates"]})
    # forces are summed over neighbours with float atomics: equal to float32 rounding only
    np.testing.assert_allclose(de["coordinates"], -f, rtol=1e-5, atol=1e-6)
    out2 = model(species=s["species"], coordinates=s["coordinates"])
    g = O.build_graphs(x32, np.array([64]), p)
    np.testing.assert_allclose(out2["embedding"], O.ani_aev(x32, s["species"], g, p), rtol=2e-5, atol=2e-5)


def test_fennix_general_pbc_without_minimum_image_flag(gpu):
    """Periodic inputs without flags={'minimum_image': None} use every image within the
    cutoff (the reference's general branch); the 6.2 A water box needs it (rc = 5.2)."""
    print("[api] FENNIX general PBC", flush=True)
    s = water_box_small(8, 4)
    model = ani_model(("H", "C", "N", "O", "S", "F", "Cl"))
    inp = {"species": s["species"], "coordinates": s["coordinates"], "cells": s["cells"]}
    e, f, v, _ = model.energy_and_forces_and_virial(**inp)
    p = O.AEVParams(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    x32 = np.asarray(s["coordinates"], np.float32).astype(np.float64)
    Er, Fr, Vr, _ = O.energy_forces_virial(O.Model(p, model.params, minimum_image=False), x32, s["species"],
                                           None, s["cells"], np.linalg.inv(s["cells"]))
    np.testing.assert_allclose(e, Er, rtol=1e-5)
    assert np.sqrt(np.mean((f - Fr) ** 2)) < 1e-4
    np.testing.assert_allclose(v, Vr, rtol=1e-4, atol=1e-4 * np.abs(Vr).max())
    # the minimum-image result differs on this box (images beyond the nearest are missed)
    e_mi, _, _ = model.energy_and_forces(**inp, flags={"minimum_image": None})
    assert abs(e_mi[0] - e[0]) > 1e-4 * abs(e[0])
    # host-path export: pbc_shifts of the general branch (several images per pair)
    out = model.preprocess(use_gpu=False, **inp)
    g = out["graph"]
    P = g["edge_src"].shape[0] // 2
    valid = g["edge_src"][:P] < len(s["species"])
    pairs = np.stack([g["edge_src"][:P][valid], g["edge_dst"][:P][valid]], 1)
    assert len({tuple(r) for r in pairs.tolist()}) < len(pairs)


def test_md_general_pbc_skin(gpu):
    """Langevin MD in general PBC with a skin list stays finite and follows the
    rebuild-every-step run."""
    print("[api] MD general PBC", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(8, 4)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, cfg.aev_length)
    xs = []
    for skin in (-1.0, 1.0):
        pot = ANIPotential(cfg, params, precision="fp32")
        md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=4, nblist_skin=skin, nblist_stride=8,
                        minimum_image=False)
        md.run(24)
        md.verify()
        xs.append(md.x.cpu().numpy())
        assert np.all(np.isfinite(md.energies()))
    np.testing.assert_allclose(xs[1], xs[0], rtol=0, atol=1e-6)


def test_host_preprocessing_layout_matches_reference(gpu):
    print("[api] host preprocessing", flush=True)
    s = water_box_small(64, 4)
    model = ani_model(("H", "C", "N", "O", "S", "F", "Cl"))
    out = model.preprocess(use_gpu=False, **s)
    g = out["graph"]
    nat = len(s["species"])
    P = g["edge_src"].shape[0] // 2
    src, dst = g["edge_src"][:P], g["edge_dst"][:P]
    valid = src < nat
    ref, ref_d12 = canonical_pairs_oracle(s["coordinates"], s["natoms"], 5.2, s["cells"])
    got = np.stack([src[valid], dst[valid]], 1)
    np.testing.assert_array_equal(got, ref)                      # reference order: row-major i < j
    np.testing.assert_array_equal(g["d12"][:P][valid], ref_d12)
    np.testing.assert_array_equal(g["edge_src"][P:], g["edge_dst"][:P])   # symmetrised
    np.testing.assert_array_equal(g["pbc_shifts"][P:], -g["pbc_shifts"][:P])
    assert np.all(src[~valid] == nat)                            # padding value N
    ga = out["graph_angle"]
    fi = ga["filter_indices"]
    ok = fi < g["edge_src"].shape[0]
    assert np.all(g["d12"][fi[ok]] < np.float32(3.5 ** 2))
    nang = np.count_nonzero(g["d12"] < np.float32(3.5 ** 2))
    assert ok.sum() == nang
    # triplets: count identity T = sum_i C(n_i, 2)
    es = ga["edge_src"][ok]
    n_i = np.bincount(es, minlength=nat)
    T = int((n_i * (n_i - 1) // 2).sum())
    assert np.count_nonzero(ga["angle_src"] < ga["edge_src"].shape[0]) == T
    assert set(out["species_index"].keys()) == {"H", "O"}


def test_md_energy_conservation_nve(gpu):
    """NVE with dt = 0.1 fs on the 64-atom cluster: total energy drift stays small."""
    print("[api] NVE drift", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = systems.cluster_config0(0)
    cfg = ANIConfig()
    pot = ANIPotential(cfg, O.init_mlp_params(11, O.aev_length(O.AEVParams())), precision="fp32")
    md = LangevinMD(pot, s, dt_fs=0.1, temperature=100.0, seed=3, thermostat="NVE")
    e0 = md.total_energy_kcal()[0]
    md.run(200)
    e1 = md.total_energy_kcal()[0]
    ek = np.trace(md.kinetic_tensor()[0])
    assert np.isfinite(e1)
    # mid-step velocities (merged half kicks) make Etot fluctuate at O(dt^2): bound relative to K
    assert abs(e1 - e0) < 0.05 * ek, (e0, e1, ek)


def test_langevin_temperature(gpu):
    print("[api] LGV temperature", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(216, 9)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    # a random-init network is far from equilibrium at the lattice start and keeps releasing
    # potential energy; scaled down 20x the thermostat, not the release, sets the temperature
    params = {s_: [(W, b) for W, b in layers[:-1]] + [(0.05 * layers[-1][0], layers[-1][1])]
              for s_, layers in O.init_mlp_params(11, cfg.aev_length).items()}
    pot = ANIPotential(cfg, params, precision="bf16")
    md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, gamma_thz=10.0, seed=5, check_every=0)
    temps = []
    for _ in range(40):
        md.run(10)
        temps.append(md.temperature()[0])
    md.verify()
    assert 200.0 < np.mean(temps[10:]) < 420.0, temps


def test_md_graph_replay_matches_eager(gpu):
    """A captured MD step replays like the eager step (several replays, eager work in between):
    the same positions, energies and virials bit for bit.  Each replayed step's integrator launch
    forms the previous step's per-system sums, so after the replays the stored energies are one
    step old until energies() forms the last step's."""
    print("[api] graph replay", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(64, 11)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, cfg.aev_length)
    runs = []
    for use_graph in (False, True):
        pot = ANIPotential(cfg, params, precision="fp32")
        md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=7, check_every=0)
        if use_graph:
            md.step(check=False)          # the same first step as the eager run
            md.capture(steps_per_graph=2)  # capture warms up with one more eager step
            traj = []
            for _ in range(3):
                md.replay(1)
                torch.cuda.synchronize()
                _ = md.x.sum().item()      # eager work between replays
                traj.append(md.x.cpu().numpy().copy())
            stale = md.sb.energy.cpu().numpy().copy()  # the sums of step 7
        else:
            traj = []
            md.step(check=False)
            md.step(check=False)
            for k in range(3):
                md.step(check=False)
                if k == 2:
                    e7 = md.energies().copy()
                md.step(check=False)
                torch.cuda.synchronize()
                traj.append(md.x.cpu().numpy().copy())
        md.verify()
        runs.append((traj, md.istep, md.energies().copy(), md.virial_kcal().copy()))
    (te, ne, ee, ve), (tg, ng, eg, vg) = runs
    assert ne == ng == 8
    # fixed-order sums throughout (int64 neighbour-force sums, ordered reductions): the same bits
    for a, b in zip(te, tg):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ee, eg)
    np.testing.assert_array_equal(ve, vg)
    np.testing.assert_array_equal(stale, e7)
    assert not np.array_equal(stale, eg)


def test_md_skin_matches_rebuild_every_step(gpu):
    """Langevin MD with a 1 A skin, full rebuilds every 10 steps and re-filters in between,
    follows the rebuild-every-step trajectory (same seed) to rounding: no atom moves 0.5 A
    in 10 steps of 0.5 fs, so every re-filtered list is the fresh list (only the order of
    the rows, hence of the AEV sums, differs)."""
    print("[api] skin MD", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(64, 13)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, cfg.aev_length)
    out = []
    for skin in (-1.0, 1.0):
        pot = ANIPotential(cfg, params, precision="fp32", max_nbr=96, max_nbr_ang=48)  # no regrowth
        md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=9, nblist_skin=skin, nblist_stride=10)
        assert md.nblist_stride == (10 if skin > 0 else 1)
        kinds = []
        build, upd = pot.build_nblist, pot.update_skin
        pot.build_nblist = lambda *a, **k: (kinds.append("full"), build(*a, **k))[1]
        pot.update_skin = lambda *a, **k: (kinds.append("skin"), upd(*a, **k))[1]
        energies, x_prev, moved = [], md.x.cpu().numpy(), 0.0
        for k in range(60):
            md.step()
            energies.append(md.energies()[0])
            if k % 10 == 9:
                x = md.x.cpu().numpy()
                moved = max(moved, float(np.linalg.norm(x - x_prev, axis=1).max()))
                x_prev = x
        md.verify()
        out.append((md.x.cpu().numpy(), np.array(energies), kinds.count("full"), kinds.count("skin"), moved))
    (x_a, e_a, f_a, s_a, _), (x_b, e_b, f_b, s_b, moved) = out
    assert moved < 0.5, moved
    assert (f_a, s_a) == (60, 0)
    assert (f_b, s_b) == (6, 60)  # full builds at steps 1, 11, ..., 51, each with its re-filter
    np.testing.assert_allclose(x_b, x_a, rtol=0, atol=1e-6)
    np.testing.assert_allclose(e_b, e_a, rtol=1e-6, atol=1e-6)


def test_md_skin_graph_replay_matches_eager(gpu):
    """Two captured one-step graphs (full rebuild, skin re-filter) replayed on the rebuild
    schedule give the eager skin trajectory."""
    print("[api] skin graph replay", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(64, 11)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, cfg.aev_length)
    runs = []
    for use_graph in (False, True):
        pot = ANIPotential(cfg, params, precision="fp32")
        md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=7, check_every=0, nblist_skin=1.0,
                        nblist_stride=5)
        traj = []
        if use_graph:
            md.capture()  # two eager warm-up steps
            for _ in range(12):
                md.replay(1)
                torch.cuda.synchronize()
                _ = md.x.sum().item()
                traj.append(md.x.cpu().numpy().copy())
        else:
            md.step(check=False)
            md.step(check=False)
            for _ in range(12):
                md.step(check=False)
                torch.cuda.synchronize()
                traj.append(md.x.cpu().numpy().copy())
        md.verify()
        runs.append((traj, md.istep, md.nblist_countdown))
    (te, ne, ce), (tg, ng, cg) = runs
    assert ne == ng == 14 and ce == cg
    for a, b in zip(te, tg):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)


def test_md_skin_multi_step_graphs_match_eager(gpu):
    """With a skin list and steps_per_graph=3, replay() uses the 3-re-filter graph where the
    rebuild schedule (stride 5) has three re-filters ahead and one-step graphs elsewhere: the
    positions after every replay and the final energies and virials equal the eager steps bit
    for bit, and the step count and schedule end where the eager run's do."""
    print("[api] skin multi-step graphs", flush=True)
    from fennol_amd.md import LangevinMD
    from fennol_amd.potential import ANIConfig, ANIPotential

    s = water_box_small(64, 11)
    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, cfg.aev_length)

    def make():
        pot = ANIPotential(cfg, params, precision="fp32")
        return LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=7, check_every=0, nblist_skin=1.0,
                          nblist_stride=5)

    md = make()
    eager = {}
    for k in range(2 + 23):
        md.step(check=False)
        eager[k + 1] = md.x.cpu().numpy().copy()
    e_ref, v_ref, c_ref = md.energies().copy(), md.virial_kcal().copy(), md.nblist_countdown

    md = make()
    md.capture(steps_per_graph=3)  # two eager warm-up steps
    done, sizes = 0, []
    while done < 23:
        k = md.replay(1, max_steps=23 - done)
        done += k
        sizes.append(k)
        np.testing.assert_array_equal(md.x.cpu().numpy(), eager[2 + done])
    md.verify()
    assert md.istep == 25 and md.nblist_countdown == c_ref
    assert 3 in sizes and 1 in sizes
    np.testing.assert_array_equal(md.energies(), e_ref)
    np.testing.assert_array_equal(md.virial_kcal(), v_ref)


def test_fnx_save_load_roundtrip(gpu, tmp_path):
    print("[api] fnx roundtrip", flush=True)
    s = systems.cluster_config0(6)
    model = ani_model(seed=9)
    e, f, _ = model.energy_and_forces(species=s["species"], coordinates=s["coordinates"])
    path = tmp_path / "ani.fnx"
    model.save(path)
    back = FENNIX.load(path)
    e2, f2, _ = back.energy_and_forces(species=s["species"], coordinates=s["coordinates"])
    np.testing.assert_array_equal(e, e2)
    np.testing.assert_allclose(f, f2, rtol=1e-6, atol=1e-6)


def test_deck_runs_like_the_md_driver(gpu, tmp_path):
    """A fennol_md-style deck (general PBC, Langevin, skin) runs through the device MD driver:
    the trajectory file has one frame per tdump and the final positions equal a direct
    LangevinMD run with the same parameters and seed."""
    print("[api] deck", flush=True)
    from fennol_amd.deck import run_deck
    from fennol_amd.md import LangevinMD

    s = water_box_small(8, 4)
    model = ani_model(("H", "C", "N", "O", "S", "F", "Cl"))
    model.save(tmp_path / "model.fnx")
    sym = {1: "H", 8: "O"}
    lines = [f"{len(s['species'])}  water8"]
    lines += [f"{k + 1:6d}  {sym[int(z)]}  {x[0]:.8f} {x[1]:.8f} {x[2]:.8f}" for k, (z, x) in
              enumerate(zip(s["species"], s["coordinates"]))]
    (tmp_path / "water8.xyz").write_text("\n".join(lines) + "\n")
    L = float(s["cells"][0][0, 0])
    (tmp_path / "input.fnl").write_text(
        "model_file model.fnx\nxyz_input{\n file water8.xyz\n indexed yes\n has_comment_line no\n}\n"
        f"cell = {L} 0 0 0 {L} 0 0 0 {L}\nminimum_image no\nnsteps = 20\ndt[fs] = 0.5\ntdump[fs] = 2.5\n"
        "temperature = 300\ngamma[THz] = 1.\nrandom_seed 12\nnblist_skin 1.\nnblist_stride 6\n"
        "nprint 5\ntraj_format xyz\nenergy_unit eV\n")
    res = run_deck(tmp_path / "input.fnl")
    assert len(res["rows"]) == 4 and all(np.isfinite(r[1]) for r in res["rows"])
    text = (tmp_path / "water8.traj.xyz").read_text().splitlines()
    assert len(text) == 4 * (2 + len(s["species"]))
    x_deck = res["md"].x.cpu().numpy()
    xyz = np.array([[float(v) for v in ln.split()[2:5]] for ln in lines[1:]])
    md = LangevinMD(model.potential, {"species": s["species"], "coordinates": xyz, "cells": s["cells"]},
                    dt_fs=0.5, temperature=300.0, gamma_thz=1.0, seed=12, energy_unit=model.energy_unit,
                    nblist_skin=1.0, nblist_stride=6, minimum_image=False)
    md.run(20)
    np.testing.assert_allclose(x_deck, md.x.cpu().numpy(), rtol=0, atol=1e-6)
