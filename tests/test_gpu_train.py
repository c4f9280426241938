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


def _celu_parts(z, alpha):
    e = torch.exp(torch.clamp(z, max=0.0) / alpha)
    pos = z > 0
    return (torch.where(pos, z, alpha * (e - 1.0)), torch.where(pos, torch.ones_like(z), e),
            torch.where(pos, torch.zeros_like(z), e / alpha))


def _torch_species_grad(weights, A, dA, cE, cD, a=0.1):
    """Plain PyTorch FP32 restatement of the MLP part of the gradient (forward primal and
    tangent chains, reverse of both), the reference the HIP kernels are compared with."""
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = weights
    z1 = A @ W1 + b1
    h1, s1, t1 = _celu_parts(z1, a)
    dz1 = dA @ W1
    dh1 = s1 * dz1
    z2 = h1 @ W2 + b2
    h2, s2, t2 = _celu_parts(z2, a)
    dz2 = dh1 @ W2
    dh2 = s2 * dz2
    z3 = h2 @ W3 + b3
    h3, s3, t3 = _celu_parts(z3, a)
    dz3 = dh2 @ W3
    dh3 = s3 * dz3
    w4 = W4[:, 0]
    gW4 = (h3 * cE[:, None]).sum(0) + cD * dh3.sum(0)
    bh3 = cE[:, None] * w4[None, :]
    bdh3 = torch.full_like(dh3, cD) * w4[None, :]
    bdz3 = s3 * bdh3
    bz3 = t3 * dz3 * bdh3 + s3 * bh3
    bdh2, bh2 = bdz3 @ W3.T, bz3 @ W3.T
    bdz2 = s2 * bdh2
    bz2 = t2 * dz2 * bdh2 + s2 * bh2
    bdh1, bh1 = bdz2 @ W2.T, bz2 @ W2.T
    bdz1 = s1 * bdh1
    bz1 = t1 * dz1 * bdh1 + s1 * bh1
    return [(dA.T @ bdz1 + A.T @ bz1, bz1.sum(0)), (dh1.T @ bdz2 + h1.T @ bz2, bz2.sum(0)),
            (dh2.T @ bdz3 + h2.T @ bz3, bz3.sum(0)), (gW4[:, None], cE.sum().reshape(1))]


def _near_kink_rows(weights, A, tol=3e-5):
    """Rows whose pre-activation at some layer lies within tol of 0, where CELU's second
    derivative jumps from 1/alpha to 0: two FP32 summation orders may put such a z on different
    sides of the kink."""
    (W1, b1), (W2, b2), (W3, b3), _ = weights
    near = torch.zeros(A.shape[0], dtype=torch.bool, device=A.device)
    h = A
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        z = h @ W + b
        near |= (z.abs() < tol).any(dim=1)
        h, _, _ = _celu_parts(z, 0.1)
    return near


def _training_setup(n_boxes=4, seed=5):
    from fennol_amd.potential import ANIConfig, ANIPotential
    from fennol_amd.training import TrainingStep

    batch = systems.training_boxes(n_boxes=n_boxes, seed=seed)
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
    return batch, p, params, g, e_ref, f_ref, pot, sb, x32, ts


def test_training_kernels_match_a_torch_fp32_reference(gpu):
    """fnl_train_mlp_grad against the same chains written with torch FP32 products, on every
    weight and bias of both networks present (H and O); dead layer-1 rows are exact zeros; a
    second call gives the same bits (fixed-order reductions).  Atoms with a pre-activation
    within 3e-5 of CELU's kink (about the FP32 rounding of a 256-term sum) get a zero tangent
    row, which removes their second-derivative terms: the two FP32 summation orders may put such
    a z on different sides of the kink, where CELU'' jumps by 1/alpha."""
    print("[train] kernels vs torch fp32", flush=True)
    batch, p, params, g, e_ref, f_ref, pot, sb, x32, ts = _training_setup(n_boxes=8, seed=6)
    pot.compute(sb, x32)
    rng = np.random.Generator(np.random.PCG64(3))
    r = torch.tensor(rng.normal(size=x32.shape).astype(np.float32), device="cuda")
    dA = ts.tangent(sb, x32, r)
    N = sb.n_atoms
    cE = torch.tensor(rng.normal(size=N).astype(np.float32), device="cuda")
    cD = -0.37
    from fennol_amd.units import SYMBOL_TO_Z
    z_all = torch.as_tensor(batch["species"], device="cuda")
    for s in ts.species:
        idx = torch.nonzero(z_all == SYMBOL_TO_Z[s]).flatten()
        if idx.numel():
            near = _near_kink_rows(ts.weights[s], sb.aev[idx, : ts.k_in].float())
            dA[idx[near]] = 0.0
    grads = ts.mlp_grad(sb, dA, cE, cD)
    again = ts.mlp_grad(sb, dA, cE, cD)
    z = z_all
    K = ts.k_in
    live = pot.config.live_column_blocks(sb.extra["aev_mask"])
    dead = torch.tensor([((live >> (c // 64 * 2)) & 3) == 0 for c in range(K)], device="cuda")
    for s in ts.species:
        idx = torch.nonzero(z == SYMBOL_TO_Z[s]).flatten()
        ref = _torch_species_grad(ts.weights[s], sb.aev[idx, :K].float(), dA[idx, :K], cE[idx], cD)
        for l in range(4):
            for k in range(2):
                got, want = grads[s][l][k], ref[l][k]
                assert torch.equal(got, again[s][l][k])
                if idx.numel() == 0:
                    assert not got.any()
                    continue
                scale = float(want.abs().max()) + 1e-30
                err = float((got - want).abs().max())
                assert err <= 2e-4 * scale, (s, l, k, err, scale)
        if idx.numel():
            assert not grads[s][0][0][dead].any()


def test_training_gradient_finite_differences(gpu):
    """Loss and 64 sampled weights (plus 8 biases) against central differences of the FP64
    oracle loss (h = 1e-4).  Layer-1 samples come from AEV channels the batch's atoms use (H and
    O blocks), so zeroing a layer-1 gradient fails the test."""
    print("[train] weight gradient", flush=True)
    batch, p, params, g, e_ref, f_ref, pot, sb, x32, ts = _training_setup()
    loss, grads, E, F = ts(sb, x32, torch.tensor(e_ref, device="cuda"),
                           torch.tensor(f_ref, dtype=torch.float32, device="cuda"))
    model = O.Model(p, params)
    loss_ref = _oracle_loss(model, batch, e_ref, f_ref, g)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    live = pot.config.live_column_blocks(sb.extra["aev_mask"])
    aev = sb.aev.float().cpu().numpy()
    rng = np.random.Generator(np.random.PCG64(7))
    h = 1e-4
    checked = 0
    for s, zz in (("H", 1), ("O", 8)):
        used = np.nonzero(np.abs(aev[batch["species"] == zz]).max(0) > 1e-3)[0]
        used = used[[((live >> (c // 32)) & 1) == 1 for c in used]]
        for layer in range(4):
            W = params[s][layer][0]
            for _ in range(8):
                i = int(rng.choice(used)) if layer == 0 else int(rng.integers(0, W.shape[0]))
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
                if layer == 0:
                    assert abs(num) > 1e-8 * max(1.0, abs(loss_ref)), (s, i, j)  # a live channel
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
    assert checked == 64
