"""GPU tests of the MD driver: recovery from capacity overflows between checks, the pressure
estimator, and the reference's watersmall deck run through the deck runner."""

from __future__ import annotations

from collections import OrderedDict
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd import units
from fennol_amd.fennix import FENNIX
from tests.helpers import water_box_small

pytestmark = pytest.mark.gpu

SPECIES7 = ("H", "C", "N", "O", "S", "F", "Cl")
GOLDEN = Path(__file__).resolve().parent / "golden"


def _flush(msg):
    print(msg, flush=True)


def _potential(precision="fp32", max_nbr=64, max_nbr_ang=32):
    from fennol_amd.potential import ANIConfig, ANIPotential

    cfg = ANIConfig(species_order=SPECIES7)
    params = O.init_mlp_params(11, cfg.aev_length)
    return cfg, params, ANIPotential(cfg, params, precision=precision, max_nbr=max_nbr, max_nbr_ang=max_nbr_ang)


@pytest.mark.parametrize("graph", [False, True])
def test_overflow_between_checks_rolls_back(gpu, graph):
    """Rows forced too small in the middle of a checkpointed run (no host reads between
    checks): the overflow is found at the next check, the rows are regrown, the run goes back
    to the checkpoint and redoes the steps; the trajectory is the one an ample-capacity run
    takes with the same seed."""
    _flush(f"[md] overflow rollback graph={graph}")
    from fennol_amd.md import LangevinMD

    s = water_box_small(64, 21)
    out = []
    for force in (False, True):
        _, _, pot = _potential()
        md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=4, check_every=0, checkpoint_every=6)
        md.run(12, graph=graph)
        if force:
            sb = md.sb
            sb.max_nbr, sb.max_nbr_ang = 16, 8
            pot._alloc_nblist(sb)
            if graph:
                md.capture(1, warmup=False)
        md.run(18, graph=graph)
        assert md.istep == (30 if not graph else 31)  # capture(warmup) took one eager step
        assert md.rollbacks == (1 if force else 0)
        assert md.sb.max_nbr > 16 and md.sb.max_nbr_ang > 8
        assert not md.overflowed()
        out.append(md.x.cpu().numpy())
    np.testing.assert_allclose(out[1], out[0], rtol=0, atol=1e-6)


def test_pressure_matches_virial_and_kinetic_tensor(gpu):
    """P = 2K/V - virial/V (md/integrate.py:364-371): the virial against the oracle's
    (itself the finite-difference strain derivative), K from the device velocities."""
    _flush("[md] pressure")
    from fennol_amd.md import LangevinMD

    s = water_box_small(64, 23)
    cfg, params, pot = _potential()
    md = LangevinMD(pot, s, dt_fs=0.5, temperature=300.0, seed=6, energy_unit="eV")
    md.run(5)
    P = md.pressure_tensor()[0]
    x32 = md.x32.cpu().numpy().astype(np.float64)
    p = O.AEVParams(species_order=SPECIES7)
    _, _, V, _ = O.energy_forces_virial(O.Model(p, params), x32, s["species"], s["natoms"], s["cells"],
                                        s["reciprocal_cells"])
    v = md.v.cpu().numpy()
    m = units.ATOMIC_MASSES[s["species"]] / units.DA
    K = 0.5 * np.einsum("i,ia,ib->ab", m, v, v)
    vol = abs(np.linalg.det(s["cells"][0]))
    kin, vir = 2.0 * K / vol, V[0] / units.energy_multiplier("eV") / vol
    P_ref = kin - vir
    scale = max(np.abs(kin).max(), np.abs(vir).max())
    np.testing.assert_allclose(P, P_ref, rtol=0, atol=2e-4 * scale)
    assert md.pressure()[0] == pytest.approx(np.trace(P_ref) / 3.0, abs=2e-4 * scale)


def test_watersmall_deck_runs(gpu, tmp_path):
    """The reference's watersmall deck (examples/md/watersmall/input.fnl: Tinker xyz, 18.643 A
    cube, minimum image, 2 A skin, LGV at 10/ps, arc trajectory, an inactive qtb{} block with
    omegacut[cm1]) on its 216-water geometry, with a model file written by FENNIX.save: the
    runner's final positions equal a direct LangevinMD run with the same parameters."""
    _flush("[md] watersmall deck")
    from fennol_amd.deck import run_deck
    from fennol_amd.md import LangevinMD

    ws = np.load(GOLDEN / "watersmall.npz")
    species, x0, cell = ws["species"], ws["coordinates"], ws["cell"]
    model = FENNIX(
        cutoff=5.2,
        modules=OrderedDict(
            embedding={"module_name": "ANI_AEV", "species_order": list(SPECIES7), "graph_angle_key": "graph_angle",
                       "radial_eta": 19.7, "angular_eta": 12.5, "radial_dist_divisions": 16,
                       "angular_dist_divisions": 4, "zeta": 14.1, "angle_sections": 8},
            energy={"module_name": "SPECIES_INDEX_NET", "output_dim": 1, "hidden_neurons": [256, 192, 160],
                    "species_order": ["H", "C", "N", "O"], "activation": "partial(celu,alpha=0.1)",
                    "input_key": "embedding", "squeeze": True}),
        preprocessing=OrderedDict(
            graph_angle={"module_name": "GRAPH_FILTER", "cutoff": 3.5, "parent_graph": "graph",
                         "graph_key": "graph_angle"},
            angle_extension={"module_name": "GRAPH_ANGULAR_EXTENSION", "graph_key": "graph_angle"}),
        energy_terms=["energy"], energy_unit="eV", seed=5)
    (tmp_path / "models").mkdir()
    model.save(tmp_path / "models" / "ani2x.fnx")
    work = tmp_path / "watersmall"
    work.mkdir()
    sym = {1: "H", 8: "O"}
    lines = [f"{len(species):6d}  Water Cubic Box (18.643 Ang, 216 AMOEBA)"]
    lines += [f"{k + 1:6d}  {sym[int(z)]}  {x[0]:11.6f} {x[1]:11.6f} {x[2]:11.6f}     {1 if z == 8 else 2}"
              for k, (z, x) in enumerate(zip(species, x0))]
    (work / "watersmall.xyz").write_text("\n".join(lines) + "\n")
    deck = f"""device cuda:0
matmul_prec highest
print_timings no
model_file ../models/ani2x.fnx

xyz_input{{
  file watersmall.xyz
  indexed yes
  has_comment_line no
}}
cell = {cell[0, 0]} 0. 0. 0. {cell[1, 1]} 0. 0. 0. {cell[2, 2]}
minimum_image yes
wrap_box no
estimate_pressure yes
nsteps = 40
dt[fs] =  .5
nblist_skin 2.
traj_format arc
tdump[fs] = 10.
nprint = 10
nsummary = 10000
thermostat LGV
temperature = 300.
gamma[THz] = 10.
random_seed 31
qtb{{
  tseg[ps]=0.25
  omegacut[cm1]=15000.
  skipseg = 5
  startsave = 50
  agamma  = 1.
}}
"""
    (work / "input.fnl").write_text(deck)
    res = run_deck(work / "input.fnl", out=open(tmp_path / "log.txt", "w"))
    assert len(res["rows"]) == 4 and all(np.isfinite(r[1]) and np.isfinite(r[5]) for r in res["rows"])
    assert res["md"].nblist_stride == 80
    traj = (work / "watersmall.traj.arc").read_text().splitlines()
    assert len(traj) == 2 * (2 + len(species))  # 2 frames, title + cell line + atoms
    xyz = np.array([[float(v) for v in ln.split()[2:5]] for ln in lines[1:]])
    md = LangevinMD(FENNIX.load(tmp_path / "models" / "ani2x.fnx").potential,
                    {"species": species, "coordinates": xyz, "cells": cell[None]}, dt_fs=0.5, temperature=300.0,
                    gamma_thz=10.0, seed=31, energy_unit="eV", nblist_skin=2.0)
    md.run(40)
    np.testing.assert_allclose(res["md"].x.cpu().numpy(), md.x.cpu().numpy(), rtol=0, atol=1e-6)
