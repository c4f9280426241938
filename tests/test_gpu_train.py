"""GPU training step: loss and weight gradient through forces vs FP64 finite differences."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd import systems

pytestmark = pytest.mark.gpu

SPECIES7 = ("H", "C", "N", "O", "S", "F", "Cl")


def _oracle_loss(model, batch, e_ref, f_ref, graphs, w_e=1.0, w_f=10.0):
    x = np.asarray(batch["coordinates"], np.float32).astype(np.float64)
    E, F, _, _ = O.energy_forces_virial(model, x, batch["species"], batch["natoms"], batch["cells"],
                                        batch["reciprocal_cells"], graphs=graphs)
    n_e = len(batch["natoms"])
    n_f = 3 * len(batch["species"])
    return w_e * np.sum((E - e_ref) ** 2) / n_e + w_f * np.sum((F - f_ref) ** 2) / n_f


def test_aev_tangent_matches_finite_differences(gpu):
    print("[train] AEV tangent", flush=True)
    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.training import TrainingStep

    s = systems.training_boxes(n_boxes=2, seed=3)
    cfg = ANIConfig(species_order=SPECIES7)
    p = O.AEVParams(species_order=SPECIES7)
    params = O.init_mlp_params(11, cfg.aev_length)
    pot = ANIPotential(cfg, params, precision="fp32")
    sb = pot.setup(s["species"], s["natoms"], s["cells"])
    x32 = torch.tensor(np.asarray(s["coordinates"], np.float32), device="cuda")
    pot.compute(sb, x32)
    rng = np.random.Generator(np.random.PCG64(2))
    r = rng.normal(size=x32.shape).astype(np.float32)
    ts = TrainingStep(pot, params)
    dA = ts.tangent(sb, x32, torch.tensor(r, device="cuda")).cpu().numpy()[:, : cfg.aev_length]
    x = np.asarray(s["coordinates"], np.float32).astype(np.float64)
    g = O.build_graphs(x, s["natoms"], p, s["cells"], s["reciprocal_cells"])
    bi = np.repeat(np.arange(len(s["natoms"])), s["natoms"])
    h = 1e-5
    ap = O.ani_aev(x + h * r, s["species"], g, p, s["cells"], bi)
    am = O.ani_aev(x - h * r, s["species"], g, p, s["cells"], bi)
    num = (ap - am) / (2 * h)
    err = np.abs(dA - num)
    assert np.all(err <= 2e-3 * np.maximum(1.0, np.abs(num))), err.max()


def test_training_gradient_finite_differences(gpu):
    print("[train] weight gradient", flush=True)
    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.training import TrainingStep

    batch = systems.training_boxes(n_boxes=4, seed=5)
    cfg = ANIConfig(species_order=SPECIES7)
    p = O.AEVParams(species_order=SPECIES7)
    params = O.init_mlp_params(11, cfg.aev_length)
    ref_params = O.init_mlp_params(12, cfg.aev_length)
    x = np.asarray(batch["coordinates"], np.float32).astype(np.float64)
    g = O.build_graphs(x, batch["natoms"], p, batch["cells"], batch["reciprocal_cells"])
    e_ref, f_ref, _, _ = O.energy_forces_virial(O.Model(p, ref_params), x, batch["species"], batch["natoms"],
                                                batch["cells"], batch["reciprocal_cells"], graphs=g)
    pot = ANIPotential(cfg, params, precision="fp32")
    sb = pot.setup(batch["species"], batch["natoms"], batch["cells"])
    x32 = torch.tensor(np.asarray(batch["coordinates"], np.float32), device="cuda")
    ts = TrainingStep(pot, params)
    loss, grads, E, F = ts(sb, x32, torch.tensor(e_ref, device="cuda"),
                           torch.tensor(f_ref, dtype=torch.float32, device="cuda"))
    model = O.Model(p, params)
    loss_ref = _oracle_loss(model, batch, e_ref, f_ref, g)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    rng = np.random.Generator(np.random.PCG64(7))
    h = 1e-4
    checked = 0
    for s in ("H", "O"):
        for layer in range(4):
            W = params[s][layer][0]
            for _ in range(2):
                i = int(rng.integers(0, W.shape[0] if layer else 128))  # layer 1: channels that are non-zero
                j = int(rng.integers(0, W.shape[1]))
                old = W[i, j]
                W[i, j] = old + h
                lp = _oracle_loss(model, batch, e_ref, f_ref, g)
                W[i, j] = old - h
                lm = _oracle_loss(model, batch, e_ref, f_ref, g)
                W[i, j] = old
                num = (lp - lm) / (2 * h)
                got = float(grads[s][layer][0][i, j])
                assert got == pytest.approx(num, rel=5e-3, abs=1e-6 * max(1.0, abs(loss_ref))), (s, layer, i, j)
                checked += 1
            b = params[s][layer][1]
            k = int(rng.integers(0, b.shape[0]))
            old = b[k]
            b[k] = old + h
            lp = _oracle_loss(model, batch, e_ref, f_ref, g)
            b[k] = old - h
            lm = _oracle_loss(model, batch, e_ref, f_ref, g)
            b[k] = old
            num = (lp - lm) / (2 * h)
            got = float(grads[s][layer][1][k])
            assert got == pytest.approx(num, rel=5e-3, abs=1e-6 * max(1.0, abs(loss_ref))), (s, "bias", layer, k)
            checked += 1
    assert checked == 24
