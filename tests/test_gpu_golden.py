"""The HIP path (FP32) against the committed golden fixtures.

Bars: neighbour pair sets and their float32 d12 bit-exact; AEV within 2e-5
relative; energies 1e-5 relative; forces RMS 1e-4 eV/A; virial 1e-4
relative; MLP energies and dE/dAEV 1e-5; ten BAOAB steps of the device MD
driver (same seed, same Philox counters) to 1e-5 A and 1e-3 A/ps.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd.md import LangevinMD
from fennol_amd.potential import ANIPotential
from tests.golden import make_golden as G
from tests.helpers import canonical_half, configs, device_half_list

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


def _flush(msg):
    print(msg, flush=True)


def _potential():
    p, c = configs(G.SPECIES4)
    params = O.init_mlp_params(G.WEIGHT_SEED, O.aev_length(p))
    assert G.weights_digest(params) == str(_load("water2")["weights_sha256"])
    return c, ANIPotential(c, params, precision="fp32")


@pytest.mark.parametrize("name", ["cluster64", "water2", "triclinic5"])
def test_golden_system(gpu, name):
    _flush(f"[golden] {name}")
    f = _load(name)
    c, pot = _potential()
    cells = f["cells"] if "cells" in f else None
    # triclinic5: cell edges below twice the cutoff, general PBC (several images per pair)
    sb = pot.setup(f["species"], None, cells, minimum_image=name != "triclinic5")
    x = torch.tensor(f["coordinates"].astype(np.float32), device="cuda")
    pot.compute(sb, x, forces=True)
    F = pot.forces(sb)
    torch.cuda.synchronize()
    src, dst, d12, sh = device_half_list(pot, sb, x, with_shifts=True)
    if "pair_shifts" in f:
        rs, rd, r12, rsh = canonical_half(f["pair_src"], f["pair_dst"], f["pair_d12"], f["pair_shifts"])
        np.testing.assert_array_equal(sh, rsh)
    else:
        rs, rd, r12 = f["pair_src"], f["pair_dst"], f["pair_d12"]
    np.testing.assert_array_equal(src, rs)
    np.testing.assert_array_equal(dst, rd)
    np.testing.assert_array_equal(d12.view(np.uint32), r12.view(np.uint32))
    aev = sb.aev.float().cpu().numpy()[:, : c.aev_length]
    ref = f["aev"].astype(np.float64)
    err = np.abs(aev - ref)
    assert np.all(err <= 2e-5 * np.maximum(1.0, np.abs(ref))), err.max()
    E = sb.energy.cpu().numpy()
    V = sb.virial.cpu().numpy().reshape(3, 3)
    Er = f["energy"]
    np.testing.assert_allclose(E, Er, rtol=1e-5, atol=1e-5 * np.abs(Er).max())
    np.testing.assert_allclose(sb.e_atom.cpu().numpy(), f["atomic_energies"], rtol=1e-5, atol=1e-6)
    rms = np.sqrt(np.mean((F.cpu().numpy() - f["forces"]) ** 2))
    assert rms < 1e-4, rms
    Vr = f["virial"].reshape(3, 3)
    np.testing.assert_allclose(V, Vr, rtol=1e-4, atol=1e-4 * np.abs(Vr).max())


def test_golden_small_aevs(gpu):
    _flush("[golden] small aevs")
    f = _load("small_aevs")
    c, pot = _potential()
    for name in ("dimer", "trimer"):
        sb = pot.setup(f[f"{name}_species"])
        x = torch.tensor(f[f"{name}_coordinates"].astype(np.float32), device="cuda")
        pot.build_nblist(sb, x)
        pot.check_reallocate(sb, x)
        pot.aev_forward(sb, x)
        aev = sb.aev.float().cpu().numpy()
        ref = f[f"{name}_aev"]
        np.testing.assert_allclose(aev[:, : c.aev_length], ref, rtol=2e-5, atol=2e-6, err_msg=name)
        assert np.all(aev[:, c.aev_length:] == 0)


def test_golden_mlp8(gpu):
    _flush("[golden] mlp8")
    f = _load("mlp8")
    c, pot = _potential()
    sb = pot.setup(f["species"])
    sb.aev[:, : c.aev_length] = torch.tensor(f["aev"].astype(np.float32), device="cuda")
    pot.mlp(sb, with_grad=True)
    torch.cuda.synchronize()
    e = sb.e_atom.cpu().numpy()
    g = sb.g_aev.float().cpu().numpy()
    np.testing.assert_allclose(e, f["energy"], rtol=1e-5, atol=1e-6)
    gr = f["grad"]
    np.testing.assert_allclose(g[:, : c.aev_length], gr, rtol=1e-4, atol=1e-5 * np.abs(gr).max())
    assert np.all(g[:, c.aev_length:] == 0)


def test_golden_baoab10(gpu):
    """The device MD driver from the fixture's seed reproduces the FP64 trajectory: same
    initial velocities, same O-step noise, forces within FP32 accuracy."""
    _flush("[golden] baoab10")
    f = _load("baoab10")
    _, pot = _potential()
    system = {"species": f["species"], "coordinates": f["x0"], "cells": f["cells"]}
    md = LangevinMD(pot, system, dt_fs=float(f["dt"]) * 1000.0, temperature=300.0, gamma_thz=float(f["gamma"]),
                    seed=77, thermostat="LGV", energy_unit="eV")
    assert abs(md.kT - float(f["kT"])) < 1e-12
    np.testing.assert_allclose(md.v.cpu().numpy(), f["v0"], rtol=1e-12, atol=1e-12)
    energies = []
    for _ in range(10):
        md.step()
        energies.append(float(md.sb.energy.cpu().numpy()[0]))
    torch.cuda.synchronize()
    np.testing.assert_allclose(md.x.cpu().numpy(), f["x"], rtol=0, atol=1e-5)
    np.testing.assert_allclose(md.v.cpu().numpy(), f["v"], rtol=0, atol=1e-3)
    np.testing.assert_allclose(energies, f["energies"], rtol=1e-5, atol=1e-6)
