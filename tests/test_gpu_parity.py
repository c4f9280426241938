"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bars: neighbour pair sets and their float32 d12 bit-exact; FP32 AEV, energies
(1e-5 relative) and forces (RMS 1e-4 eV/A) against the FP64 oracle; BF16 MLP
within 1 meV/atom and 0.02 eV/A force RMS of FP64.
"""

from __future__ import annotations

import inspect

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd import _abi, systems
from fennol_amd.potential import ANIPotential
from tests.helpers import (canonical_half, canonical_pairs_oracle, configs, device_half_list, rows_to_pairs,
                           triclinic_box, water_box_small)

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _cases():
    return {
        "cluster64": systems.cluster_config0(0),
        "water64": water_box_small(64, 1),
        "water216": water_box_small(216, 2),
        "triclinic": triclinic_box(150, 3),
    }


def _potential(cfg, precision="fp32", seed=11, max_nbr=64, max_nbr_ang=32):
    p, c = cfg
    params = O.init_mlp_params(seed, O.aev_length(p))
    return params, ANIPotential(c, params, precision=precision, max_nbr=max_nbr, max_nbr_ang=max_nbr_ang)


def _nblist(pot, s):
    sb = pot.setup(s["species"], s["natoms"], s.get("cells"))
    x = torch.tensor(np.asarray(s["coordinates"], np.float32), device="cuda")
    pot.build_nblist(sb, x)
    pot.check_reallocate(sb, x)
    torch.cuda.synchronize()
    return sb, x


@pytest.mark.parametrize("name", ["cluster64", "water64", "water216", "triclinic"])
def test_nblist_pair_set(gpu, name):
    _flush(f"[nblist] {name}")
    s = _cases()[name]
    cfg = configs()
    _, pot = _potential(cfg, max_nbr=16, max_nbr_ang=8)  # small rows: exercises regrowth
    sb, x = _nblist(pot, s)
    nbr = sb.nbr.cpu().numpy()
    n_nbr = sb.n_nbr.cpu().numpy()
    i, j = rows_to_pairs(nbr, n_nbr, sb.max_nbr)
    # symmetric directed list
    fwd = set(zip(i.tolist(), j.tolist()))
    assert fwd == set(zip(j.tolist(), i.tolist()))
    assert len(fwd) == len(i)  # no duplicate edge
    half = np.array(sorted((a, b) for a, b in fwd if a < b), dtype=np.int64).reshape(-1, 2)
    ref, ref_d12 = canonical_pairs_oracle(s["coordinates"], s["natoms"], cfg[1].cutoff, s.get("cells"))
    np.testing.assert_array_equal(half, ref)
    # d12 bit-exact via the reference-layout export
    rs = np.concatenate([[0], np.cumsum(np.minimum(n_nbr, sb.max_nbr))]).astype(np.int32)
    E = int(rs[-1])
    dev = torch.device("cuda")
    es = torch.empty(E, dtype=torch.int32, device=dev)
    ed = torch.empty(E, dtype=torch.int32, device=dev)
    d12 = torch.empty(E, dtype=torch.float32, device=dev)
    sh = torch.empty((E, 3), dtype=torch.float32, device=dev)
    rst = torch.tensor(rs[:-1], device=dev)
    P = _abi.ptr
    _abi.check(pot.lib.fnl_nblist_export(P(x), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells),
                                         P(sb.nbr), P(sb.nbr_shift), P(sb.n_nbr), sb.max_nbr, P(rst), P(es), P(ed),
                                         P(d12), P(sh), _abi.stream_handle()), "export")
    es, ed, d12 = es.cpu().numpy(), ed.cpu().numpy(), d12.cpu().numpy()
    m = es < ed
    order = np.lexsort((ed[m], es[m]))
    np.testing.assert_array_equal(d12[m][order].view(np.uint32), ref_d12.view(np.uint32))
    # angular rows: the radial slots with d12 < rca^2
    na = sb.n_nbr_ang.cpu().numpy()
    nba = sb.nbr_ang.cpu().numpy()
    rca2 = np.float32(cfg[1].cutoff_angular ** 2)
    for a in range(sb.n_atoms):
        row = nbr[a, : n_nbr[a]]
        slots = nba[a, : na[a]]
        dd = d12[rs[a]: rs[a + 1]]
        expect = np.nonzero(dd < rca2)[0]
        np.testing.assert_array_equal(np.sort(slots), expect)
        assert np.all(row[slots] == ed[rs[a] + slots])


@pytest.mark.parametrize("name", ["water216", "triclinic", "cluster64"])
def test_nblist_depth_one_and_two_cells_agree(gpu, name):
    """The build searches half-cutoff cells two deep where the grid capacity allows it and
    cutoff-wide cells one deep where it does not: capacities that force the cutoff-wide grid give
    every row the same neighbours (radial and angular) as the default grid, in another order."""
    _flush(f"[nblist] depth 1 vs 2: {name}")
    s = _cases()[name]
    cfg = configs()
    _, pot = _potential(cfg, max_nbr=96, max_nbr_ang=48)
    sb, x = _nblist(pot, s)

    def rows():
        n, na = sb.n_nbr.cpu().numpy(), sb.n_nbr_ang.cpu().numpy()
        nbr, nbr_ang = sb.nbr.cpu().numpy(), sb.nbr_ang.cpu().numpy()
        rad = [sorted(nbr[i, : n[i]].tolist()) for i in range(sb.n_atoms)]
        ang = [sorted(nbr[i, nbr_ang[i, : na[i]]].tolist()) for i in range(sb.n_atoms)]
        return rad, ang, nbr[:, : n.max()].copy()

    rad2, ang2, raw2 = rows()
    rbin = cfg[1].cutoff * 1.0001 + 1.0e-4
    caps = sb.grid.cpu().numpy().copy()
    if s.get("cells") is not None:
        R = np.linalg.inv(np.asarray(s["cells"][0], np.float64))
        width = 1.0 / np.sqrt((R ** 2).sum(axis=0))
    else:
        xs = np.asarray(s["coordinates"], np.float64)
        width = xs.max(axis=0) - xs.min(axis=0)
    caps[0, :3] = np.maximum(np.floor(width / rbin).astype(int), 1)  # cells at least a cutoff wide
    assert (caps[0, :3] <= sb.grid.cpu().numpy()[0, :3]).all()
    sb.grid.copy_(torch.tensor(caps, device=sb.grid.device))
    pot.build_nblist(sb, x)
    torch.cuda.synchronize()
    assert int(sb.status[_abi.FNL_STATUS_OVERFLOW].item()) == 0
    rad1, ang1, raw1 = rows()
    assert rad1 == rad2 and ang1 == ang2
    assert sum(len(r) for r in rad1) > 0 and sum(len(a) for a in ang1) > 0
    if name == "water216":  # 7 half-cutoff cells per axis against 3: another candidate order
        assert not np.array_equal(raw1, raw2)


def test_nblist_batched_clusters(gpu):
    _flush("[nblist] batched clusters")
    rng = np.random.Generator(np.random.PCG64(5))
    sysl = [systems.lattice_cluster(int(n), rng) for n in (8, 33, 120, 300, 512)]
    s = systems.batch_systems(sysl)
    cfg = configs()
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    i, j = rows_to_pairs(sb.nbr.cpu().numpy(), sb.n_nbr.cpu().numpy(), sb.max_nbr)
    half = np.array(sorted((a, b) for a, b in zip(i.tolist(), j.tolist()) if a < b), dtype=np.int64)
    ref, _ = canonical_pairs_oracle(s["coordinates"], s["natoms"], cfg[1].cutoff)
    np.testing.assert_array_equal(half, ref)


def test_nblist_crowded_cells(gpu):
    """Cells with more atoms than the binning kernel stages (32): the cell sort gathers them from
    the system's range instead; a collapsed 1,100-atom cluster (0.28 A lattice, one cell) also has
    more candidates per centre than the pair kernel maps through LDS (1,024), so they are found by
    the binary search over the range offsets.  Same pair set as the oracle."""
    _flush("[nblist] crowded cells")
    rng = np.random.Generator(np.random.PCG64(9))
    sysl = [systems.lattice_cluster(int(n), rng) for n in (40, 1000, 64)]
    sysl.insert(2, systems.lattice_cluster(1100, rng, spacing=0.28, jitter=0.01, min_dist=0.25))
    s = systems.batch_systems(sysl)
    cfg = configs()
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    assert int(sb.status[_abi.FNL_STATUS_MAX_CELL].item()) > 64
    i, j = rows_to_pairs(sb.nbr.cpu().numpy(), sb.n_nbr.cpu().numpy(), sb.max_nbr)
    half = np.array(sorted((a, b) for a, b in zip(i.tolist(), j.tolist()) if a < b), dtype=np.int64)
    ref, _ = canonical_pairs_oracle(s["coordinates"], s["natoms"], cfg[1].cutoff)
    np.testing.assert_array_equal(half, ref)


def test_padding_and_isolated_atoms(gpu):
    """Ragged batch with padding atoms (species 0), an atom beyond every cutoff and a one-atom
    system: energies, forces and virial equal the oracle's on the real atoms; padding atoms get
    zero force and no neighbours, isolated atoms the energy of a zero AEV and zero force."""
    _flush("[edge] padding and isolated atoms")
    water = np.array([[0.0, 0.0, 0.0], [0.9572, 0.0, 0.0], [-0.24, 0.927, 0.0]])
    xa = np.concatenate([water, [[20.0, 0.0, 0.0]], [[3.0, 3.0, 3.0], [1.0, 1.0, 1.0]]])
    sa = np.array([8, 1, 1, 8, 0, 0])
    xb = np.array([[50.0, 50.0, 50.0]])
    sb_ = np.array([1])
    cfg = configs()
    params, pot = _potential(cfg)
    species = np.concatenate([sa, sb_])
    coords = np.concatenate([xa, xb])
    E, F, V = pot.energy_forces_virial(species, coords, np.array([6, 1]))
    real = species > 0
    model = O.Model(cfg[0], params)
    Er, Fr, Vr, _ = O.energy_forces_virial(model, coords[real].astype(np.float32).astype(np.float64),
                                           species[real], np.array([4, 1]))
    np.testing.assert_allclose(E, Er, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(F[real], Fr, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(V, Vr, rtol=1e-4, atol=1e-6)
    assert np.all(F[~real] == 0.0)
    assert np.all(F[3] == 0.0) and np.all(F[6] == 0.0)  # isolated atoms


def test_nblist_near_cutoff(gpu):
    """Pairs placed within a few ulp of the cutoff keep the reference's decision."""
    _flush("[nblist] near cutoff")
    cfg = configs()
    rc = np.float32(cfg[1].cutoff)
    xs = []
    for k in range(40):
        base = np.array([10.0 * k, 0.0, 0.0], np.float32)
        d = np.nextafter(rc, np.float32(np.inf) if k % 2 else np.float32(0), dtype=np.float32)
        for _ in range(k // 2):
            d = np.nextafter(d, np.float32(np.inf) if k % 2 else np.float32(0), dtype=np.float32)
        xs += [base, base + np.array([d, 0, 0], np.float32)]
    coords = np.array(xs, np.float32)
    s = {"species": np.ones(len(xs), np.int32), "coordinates": coords, "natoms": np.array([len(xs)], np.int32)}
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    i, j = rows_to_pairs(sb.nbr.cpu().numpy(), sb.n_nbr.cpu().numpy(), sb.max_nbr)
    half = np.array(sorted((a, b) for a, b in zip(i.tolist(), j.tolist()) if a < b), dtype=np.int64).reshape(-1, 2)
    ref, _ = canonical_pairs_oracle(coords, s["natoms"], cfg[1].cutoff)
    np.testing.assert_array_equal(half, ref.reshape(-1, 2))


def _window_pairs(n_pairs, seed=3, rc=5.2):
    """Isolated atom pairs whose float32 d12 lies in [float32(float32(rc)^2), float32(rc^2)):
    inside the reference's cutoff, outside a threshold squared in float32."""
    lo = np.float32(np.float64(np.float32(rc)) ** 2)
    hi = np.float32(rc ** 2)
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = []
    while len(xs) < 2 * n_pairs:
        base = np.array([30.0 * (len(xs) // 2), 0.0, 0.0], np.float32) + rng.uniform(-1, 1, 3).astype(np.float32)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        xj = (base.astype(np.float64) + u * rc * (1 + rng.uniform(-3e-7, 3e-7))).astype(np.float32)
        v = (xj - base).astype(np.float32)
        d12 = np.float32(np.float32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
        if lo <= d12 < hi:
            xs += [base, xj]
    return np.array(xs, np.float32)


def test_nblist_cutoff_rounding_window(gpu):
    """Pairs whose d12 falls between the float32 squares of float32(rc) and of rc are kept,
    as the reference keeps them (its threshold is float32(rc**2))."""
    _flush("[nblist] cutoff rounding window")
    coords = _window_pairs(24)
    s = {"species": np.ones(len(coords), np.int32), "coordinates": coords,
         "natoms": np.array([len(coords)], np.int32)}
    cfg = configs()
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    src, dst, d12 = device_half_list(pot, sb, x)
    ref, ref_d12 = canonical_pairs_oracle(coords, s["natoms"], cfg[1].cutoff)
    assert ref.shape[0] == 24
    np.testing.assert_array_equal(np.stack([src, dst], 1), ref)
    np.testing.assert_array_equal(d12.view(np.uint32), ref_d12.view(np.uint32))


def _skin_cases():
    return {
        "cluster64": systems.cluster_config0(0),
        "water216": water_box_small(216, 2),
        "triclinic": triclinic_box(150, 3),
        "batched": systems.batch_systems([systems.lattice_cluster(int(n), np.random.Generator(
            np.random.PCG64(n))) for n in (8, 40, 200)]),
    }


@pytest.mark.parametrize("name", ["cluster64", "water216", "triclinic", "batched"])
def test_skin_update_matches_oracle(gpu, name):
    """Build with a 1 A skin, move every atom by less than 0.5 A, re-filter on the device:
    pair set and d12 bits equal the oracle's skin_update and a fresh cutoff build."""
    _flush(f"[skin] {name}")
    s = _skin_cases()[name]
    cfg = configs()
    rc, skin = cfg[1].cutoff, 1.0
    _, pot = _potential(cfg, max_nbr=16, max_nbr_ang=8)  # small rows: every capacity regrows
    cells = s.get("cells")
    rec = None if cells is None else np.linalg.inv(cells)
    sb = pot.setup(s["species"], s["natoms"], cells, skin=skin)
    x0 = np.asarray(s["coordinates"], np.float32)
    t0 = torch.tensor(x0, device="cuda")
    pot.build_nblist(sb, t0)
    pot.check_reallocate(sb, t0)
    src, dst, d12 = device_half_list(pot, sb, t0)
    ref, ref_d12 = canonical_pairs_oracle(x0, s["natoms"], rc, cells)
    np.testing.assert_array_equal(np.stack([src, dst], 1), ref)
    np.testing.assert_array_equal(d12.view(np.uint32), ref_d12.view(np.uint32))
    # the skin rows hold the cutoff + skin pair set
    i, j = rows_to_pairs(sb.nbr_skin.cpu().numpy(), sb.n_skin.cpu().numpy(), sb.max_skin)
    half = np.array(sorted((a, b) for a, b in zip(i.tolist(), j.tolist()) if a < b), np.int64).reshape(-1, 2)
    ss, sd, _, ssh = O.half_pairs(x0, s["natoms"], rc, cells, rec, skin=skin)
    np.testing.assert_array_equal(half, np.stack(canonical_half(ss, sd, np.zeros(len(ss)))[:2], 1))
    # move, re-filter
    rng = np.random.Generator(np.random.PCG64(23))
    u = rng.normal(size=x0.shape)
    x1 = (x0 + 0.49 * skin * u / np.linalg.norm(u, axis=1, keepdims=True) * rng.random((len(x0), 1))).astype(
        np.float32)
    t1 = torch.tensor(x1, device="cuda")
    sb.max_nbr, sb.max_nbr_ang = 16, 8  # regrow again from the re-filter's own status
    pot._alloc_nblist(sb)
    pot.build_nblist(sb, t0)
    pot.check_reallocate(sb, t0)
    nbr_skin_ptr = sb.nbr_skin.data_ptr()
    pot.update_skin(sb, t1)
    pot.check_reallocate(sb, t1)
    assert sb.nbr_skin.data_ptr() == nbr_skin_ptr  # the skin list of the build is kept
    src, dst, d12 = device_half_list(pot, sb, t1)
    bidx = np.repeat(np.arange(len(s["natoms"])), s["natoms"])
    us, ud, u12, _ = O.skin_update(ss, sd, ssh, x1, rc, cells, bidx)
    cs, cd, c12 = canonical_half(us, ud, u12)
    np.testing.assert_array_equal(src, cs)
    np.testing.assert_array_equal(dst, cd)
    np.testing.assert_array_equal(d12.view(np.uint32), c12.view(np.uint32))
    fresh, _ = canonical_pairs_oracle(x1, s["natoms"], rc, cells)
    np.testing.assert_array_equal(np.stack([src, dst], 1), fresh)


def _small_triclinic(n_atoms, seed, cell):
    rng = np.random.Generator(np.random.PCG64(seed))
    cell = np.asarray(cell, np.float64)
    # atoms kept >= 0.9 A apart (a smooth energy for the force checks)
    pts = []
    while len(pts) < n_atoms:
        x = rng.random(3) @ cell
        ok = True
        for y in pts:
            d = x - y
            for k in np.array(np.meshgrid(*[np.arange(-2, 3)] * 3)).T.reshape(-1, 3):
                if np.linalg.norm(d + k @ cell) < 0.9:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            pts.append(x)
    species = rng.choice(np.array([1, 6, 7, 8], np.int32), size=n_atoms)
    return {"species": species, "coordinates": np.array(pts), "natoms": np.array([n_atoms], np.int32),
            "cells": cell[None], "reciprocal_cells": np.linalg.inv(cell)[None]}


def _general_cases():
    tri = _small_triclinic(14, 21, [[6.4, 0.0, 0.0], [1.3, 6.1, 0.0], [-0.8, 1.1, 6.6]])
    tri2 = _small_triclinic(9, 22, [[5.2, 0.0, 0.0], [0.0, 5.6, 0.0], [1.5, -0.7, 5.9]])
    w8 = water_box_small(8, 4)
    both = systems.batch_systems([{k: w8[k] for k in ("species", "coordinates", "cells")},
                                  {k: tri[k] for k in ("species", "coordinates", "cells")}])
    # atoms moved out of the cell by different lattice vectors (unwrapped MD positions): the
    # stored shifts carry the wrap
    n_far = np.random.Generator(np.random.PCG64(23)).integers(-3, 4, size=(9, 3)).astype(np.float64)
    far = {**tri2, "coordinates": tri2["coordinates"] + n_far @ tri2["cells"][0]}
    return {"water8": w8, "triclinic14": tri, "batched": both, "unwrapped9": far}


def _oracle_general(s, rc, skin=0.0):
    cells = np.asarray(s["cells"])
    return O.half_pairs(s["coordinates"], s["natoms"], rc, cells, np.linalg.inv(cells), minimum_image=False,
                        skin=skin)


@pytest.mark.parametrize("name", ["water8", "triclinic14", "batched", "unwrapped9"])
def test_general_pbc_pair_set(gpu, name):
    """Cells narrower than twice the cutoff: every image within the cutoff, with its cell shift,
    pair set / shifts / d12 bit-exact against the oracle (general branch of the reference)."""
    _flush(f"[nblist general] {name}")
    s = _general_cases()[name]
    cfg = configs()
    _, pot = _potential(cfg, max_nbr=16, max_nbr_ang=8)
    sb = pot.setup(s["species"], s["natoms"], s["cells"], minimum_image=False)
    assert sb.pbc == _abi.FNL_PBC_GENERAL
    x = torch.tensor(np.asarray(s["coordinates"], np.float32), device="cuda")
    pot.build_nblist(sb, x)
    pot.check_reallocate(sb, x)
    src, dst, d12, sh = device_half_list(pot, sb, x, with_shifts=True)
    os_, od, o12, osh = _oracle_general(s, cfg[1].cutoff)
    cs, cd, c12, csh = canonical_half(os_, od, o12, osh)
    assert np.any(np.bincount(cs * 10000 + cd) > 1)  # several images of one pair
    np.testing.assert_array_equal(src, cs)
    np.testing.assert_array_equal(dst, cd)
    np.testing.assert_array_equal(sh, csh)
    np.testing.assert_array_equal(d12.view(np.uint32), c12.view(np.uint32))


@pytest.mark.parametrize("name", ["water8", "triclinic14", "unwrapped9"])
def test_general_pbc_energy_forces_virial(gpu, name):
    _flush(f"[efv general] {name}")
    s = _general_cases()[name]
    cfg = configs()
    params, pot = _potential(cfg)
    sb = pot.setup(s["species"], s["natoms"], s["cells"], minimum_image=False)
    E, F, V = pot.energy_forces_virial(s["species"], s["coordinates"], sb=sb)
    model = O.Model(cfg[0], params, minimum_image=False)
    x32 = np.asarray(s["coordinates"], np.float32).astype(np.float64)
    Er, Fr, Vr, _ = O.energy_forces_virial(model, x32, s["species"], s["natoms"], s["cells"],
                                           np.linalg.inv(s["cells"]))
    np.testing.assert_allclose(E, Er, rtol=1e-5, atol=1e-5 * np.abs(Er).max())
    rms = np.sqrt(np.mean((F - Fr) ** 2))
    assert rms < 1e-4, rms
    np.testing.assert_allclose(V, Vr, rtol=1e-4, atol=1e-4 * np.abs(Vr).max())


def test_general_pbc_skin_update(gpu):
    """Skin list in general PBC: the stored shifts are applied to the moved positions."""
    _flush("[skin general] triclinic14")
    s = _general_cases()["triclinic14"]
    cfg = configs()
    rc, skin = cfg[1].cutoff, 1.0
    _, pot = _potential(cfg)
    sb = pot.setup(s["species"], s["natoms"], s["cells"], skin=skin, minimum_image=False)
    x0 = np.asarray(s["coordinates"], np.float32)
    pot.build_nblist(sb, torch.tensor(x0, device="cuda"))
    pot.check_reallocate(sb, torch.tensor(x0, device="cuda"))
    rng = np.random.Generator(np.random.PCG64(31))
    u = rng.normal(size=x0.shape)
    x1 = (x0 + 0.45 * skin * u / np.linalg.norm(u, axis=1, keepdims=True)).astype(np.float32)
    t1 = torch.tensor(x1, device="cuda")
    pot.update_skin(sb, t1)
    pot.check_reallocate(sb, t1)
    ss, sd, _, ssh = _oracle_general(s, rc, skin=skin)
    us, ud, u12, ush = O.skin_update(ss, sd, ssh, x1, rc, s["cells"])
    cs, cd, _, csh = canonical_half(us, ud, u12, ush)
    # the device export recomputes d12 from wrapped positions (the build's formula): compare the
    # pair set and shifts, and d12 against a fresh general build at x1
    src, dst, d12, sh = device_half_list(pot, sb, t1, with_shifts=True)
    np.testing.assert_array_equal(src, cs)
    np.testing.assert_array_equal(dst, cd)
    np.testing.assert_array_equal(sh, csh)
    fs, fd, f12, fsh = canonical_half(*_oracle_general({**s, "coordinates": x1}, rc))
    np.testing.assert_array_equal(src, fs)
    np.testing.assert_array_equal(sh, fsh)
    np.testing.assert_array_equal(d12.view(np.uint32), f12.view(np.uint32))


@pytest.mark.parametrize("name", ["cluster64", "water64", "water216"])
def test_aev_forward(gpu, name):
    _flush(f"[aev] {name}")
    s = _cases()[name]
    order = ("H", "C", "N", "O", "S", "F", "Cl") if name.startswith("water") else ("H", "C", "N", "O")
    cfg = configs(order)
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    pot.aev_forward(sb, x)
    aev = sb.aev.float().cpu().numpy()[:, : cfg[1].aev_length]
    g = O.build_graphs(s["coordinates"], s["natoms"], cfg[0], s.get("cells"), s.get("reciprocal_cells"))
    ref = O.ani_aev(np.asarray(s["coordinates"], np.float32).astype(np.float64), s["species"], g, cfg[0],
                    s.get("cells"))
    err = np.abs(aev - ref)
    tol = 2e-5 * np.maximum(1.0, np.abs(ref))
    assert np.all(err <= tol), f"max err {err.max()} at {np.unravel_index(err.argmax(), err.shape)}"
    assert np.all(sb.aev.float().cpu().numpy()[:, cfg[1].aev_length:] == 0)


def test_aev_species_mask_writes_every_live_block(gpu):
    """The forward with the batch's species mask (blocks of absent species left as the zeros of
    the buffer) gives the rows a full write gives on a buffer filled with NaN."""
    _flush("[aev] species mask")
    s = water_box_small(64, 1)
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    _, pot = _potential(cfg)
    sb, x = _nblist(pot, s)
    assert sb.extra["aev_mask"] == (1 << 0) | (1 << 3)  # H and O
    pot.aev_forward(sb, x)
    masked = sb.aev.clone()
    sb.aev.fill_(float("nan"))
    sb.extra["aev_mask"] = 0xFFFFFFFF
    pot.aev_forward(sb, x)
    full = sb.aev.clone()
    torch.cuda.synchronize()
    assert not torch.isnan(full).any()
    assert torch.equal(full, masked)


@pytest.mark.parametrize("name", ["cluster64", "water64", "water216", "triclinic"])
def test_energy_forces_virial_fp32(gpu, name):
    _flush(f"[efv fp32] {name}")
    s = _cases()[name]
    if name == "triclinic":  # keep atoms apart for a smooth energy
        s = water_box_small(27, 7)
        cell = np.array([[12.0, 0.0, 0.0], [1.5, 11.0, 0.0], [-1.0, 2.0, 11.5]])
        frac = np.asarray(s["coordinates"]) @ np.linalg.inv(s["cells"][0])
        s = {**s, "coordinates": frac @ cell, "cells": cell[None], "reciprocal_cells": np.linalg.inv(cell)[None]}
    order = ("H", "C", "N", "O", "S", "F", "Cl") if name.startswith("water") else ("H", "C", "N", "O")
    cfg = configs(order)
    params, pot = _potential(cfg)
    E, F, V = pot.energy_forces_virial(s["species"], s["coordinates"], s["natoms"], s.get("cells"))
    model = O.Model(cfg[0], params)
    x32 = np.asarray(s["coordinates"], np.float32).astype(np.float64)
    Er, Fr, Vr, _ = O.energy_forces_virial(model, x32, s["species"], s["natoms"], s.get("cells"),
                                           s.get("reciprocal_cells"))
    np.testing.assert_allclose(E, Er, rtol=1e-5, atol=1e-5 * np.abs(Er).max())
    rms = np.sqrt(np.mean((F - Fr) ** 2))
    assert rms < 1e-4, rms
    np.testing.assert_allclose(V, Vr, rtol=1e-4, atol=1e-4 * np.abs(Vr).max())


@pytest.mark.parametrize("name", ["cluster64", "water216"])
def test_energy_forces_bf16(gpu, name):
    _flush(f"[efv bf16] {name}")
    s = _cases()[name]
    order = ("H", "C", "N", "O", "S", "F", "Cl") if name.startswith("water") else ("H", "C", "N", "O")
    cfg = configs(order)
    params, pot = _potential(cfg, precision="bf16")
    E, F, V = pot.energy_forces_virial(s["species"], s["coordinates"], s["natoms"], s.get("cells"))
    model = O.Model(cfg[0], params)
    Er, Fr, Vr, _ = O.energy_forces_virial(model, np.asarray(s["coordinates"], np.float64), s["species"],
                                           s["natoms"], s.get("cells"), s.get("reciprocal_cells"))
    n = len(s["species"])
    e_err_mev = 1000.0 * np.abs(E - Er).max() / n
    f_rms = np.sqrt(np.mean((F - Fr) ** 2))
    _flush(f"  bf16 energy err {e_err_mev:.4f} meV/atom, force rms {f_rms:.5f} eV/A")
    assert e_err_mev <= 1.0
    assert f_rms <= 0.02


def test_bf16_compensated_weights_remove_the_rounding_bias(gpu):
    """BF16 weights stored as hi + lo pairs: the energy of the 7-species water216 box is within
    0.02 meV/atom of the FP64 oracle and at least 5 times closer than with plain BF16 weights
    (whose rounding shifts every similar atom the same way: measured -0.081 vs -0.004
    meV/atom); forces keep the BF16 bar.  The default form takes the pairs in the forward
    products only: its energies are those of the pairs in both (the same forward arithmetic,
    bit for bit) and its force error within 1.5 times theirs."""
    _flush("[bf16 compensated] water216")
    s = _cases()["water216"]
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(11, O.aev_length(cfg[0]))
    model = O.Model(cfg[0], params)
    Er, Fr, _, _ = O.energy_forces_virial(model, np.asarray(s["coordinates"], np.float32).astype(np.float64),
                                          s["species"], s["natoms"], s.get("cells"), s.get("reciprocal_cells"))
    n = len(s["species"])
    err, ferr, energies = {}, {}, {}
    for comp in (True, "forward", False):
        pot = ANIPotential(cfg[1], params, precision="bf16", max_nbr=64, max_nbr_ang=32, bf16_compensated=comp)
        assert pot.nets.compensated == bool(comp)
        assert pot.nets.compensated_backward == (comp is True)
        E, F, _ = pot.energy_forces_virial(s["species"], s["coordinates"], s["natoms"], s.get("cells"))
        energies[comp] = E
        err[comp] = 1000.0 * abs(E[0] - Er[0]) / n
        ferr[comp] = float(np.sqrt(np.mean((F - Fr) ** 2)))
        assert ferr[comp] <= 0.02
    _flush(f"  energy error: compensated {err[True]:.5f}, forward only {err['forward']:.5f}, plain "
           f"{err[False]:.5f} meV/atom; force RMS {ferr[True]:.2e}, {ferr['forward']:.2e}, {ferr[False]:.2e} eV/A")
    assert err[True] <= 0.02
    assert 5 * err[True] <= err[False]
    np.testing.assert_array_equal(energies["forward"], energies[True])
    assert ferr["forward"] <= 1.5 * ferr[True]
    assert inspect.signature(ANIPotential).parameters["bf16_compensated"].default == "forward"


def test_bf16_aev_rows_are_rounded_fp32_rows(gpu):
    """The BF16 form's AEV rows are the FP32 rows rounded to nearest even (the same float
    arithmetic, one conversion at the store): bit-exact against torch's rounding for normal
    values, within the smallest normal for subnormal ones."""
    _flush("[aev bf16 rounding] water216")
    s = _cases()["water216"]
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    rows = {}
    for prec in ("fp32", "bf16"):
        _, pot = _potential(cfg, precision=prec)
        sb, x = _nblist(pot, s)
        pot.aev_forward(sb, x)
        torch.cuda.synchronize()
        rows[prec] = sb.aev[:, : cfg[1].aev_length].clone()
    ref = rows["fp32"].to(torch.bfloat16)
    got = rows["bf16"]
    normal = rows["fp32"].abs() >= 1.1754944e-38
    assert torch.equal(got[normal].view(torch.int16), ref[normal].view(torch.int16))
    assert torch.all((got[~normal].float() - rows["fp32"][~normal]).abs() <= 1.1754944e-38)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_aev_backward_species_mask(gpu, precision):
    """The adjoint reading only the dE/dAEV blocks of the species present (water in a 7-species
    model: 5 of 35 blocks) gives the result of reading every block."""
    _flush(f"[aev bwd mask] water216 {precision}")
    s = _cases()["water216"]
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    _, pot = _potential(cfg, precision=precision)
    sb, x = _nblist(pot, s)
    pot.aev_forward(sb, x)
    pot.mlp(sb, with_grad=True)
    out = {}
    for mask in (sb.extra["aev_mask"], 0xFFFFFFFF):
        sb.extra["aev_mask"] = mask
        pot.aev_backward(sb, x)
        torch.cuda.synchronize()
        out[mask] = [t.clone() for t in (sb.grad_self, sb.grad_nbr, sb.vir_atom)]
    part, full = out.values()
    assert bin(list(out)[0]).count("1") == 2
    assert torch.equal(part[0], full[0]) and torch.equal(part[2], full[2])
    assert torch.equal(part[1], full[1])  # fixed-point sums: exact in any atomic order


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_mlp_skips_dead_columns_exactly(gpu, precision):
    """Layer 1 over only the 32-column pieces that hold the batch's species (water in a
    7-species model: 8 of 32 pieces) gives, bit for bit, the atom energies and the live
    dE/dAEV columns of the product over all 1024 columns (the skipped products are exact
    zeros); dE/dAEV is written only in the live pieces (the others keep the buffer's zeros)."""
    _flush(f"[mlp dead columns] water216 {precision}")
    s = _cases()["water216"]
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    _, pot = _potential(cfg, precision=precision)
    sb, x = _nblist(pot, s)
    pot.aev_forward(sb, x)
    live = cfg[1].live_column_blocks(sb.extra["aev_mask"])
    # 32-column pieces: H, O radial (0, 1); HH (3, 4), HO (6, 7), OO (21, 22) angular
    assert live == sum(1 << q for q in (0, 1, 3, 4, 6, 7, 21, 22))
    out = []
    for mask in (sb.extra["aev_mask"], 0xFFFFFFFF):
        sb.extra["aev_mask"] = mask
        sb.g_aev.zero_()
        pot.mlp(sb, with_grad=True)
        torch.cuda.synchronize()
        out.append((sb.e_atom.clone(), sb.g_aev.float().clone()))
    (e_p, g_p), (e_f, g_f) = out
    assert torch.equal(e_p, e_f)
    on = torch.tensor([(live >> (c // 32)) & 1 == 1 for c in range(g_p.shape[1])], device=g_p.device)
    assert torch.equal(g_p[:, on], g_f[:, on])
    assert int(on.sum()) == 8 * 32
    assert not g_p[:, ~on].any() and g_f[:, ~on].any()


@pytest.mark.parametrize("precision,tiles", [("bf16", (32, 64, 96)), ("fp32", (16, 32))])
def test_mlp_tile_sizes_agree(gpu, precision, tiles):
    """Every atom's MLP result depends only on its own row: all tile sizes give the same atom
    energies and dE/dAEV rows (same k order per atom)."""
    _flush(f"[mlp tiles] {precision} {tiles}")
    s = water_box_small(216, 4)
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(13, O.aev_length(cfg[0]))
    outs = []
    for m in tiles:
        pot = ANIPotential(cfg[1], params, precision=precision, mlp_tile_rows=m)
        sb = pot.setup(s["species"], s["natoms"], s["cells"])
        x = torch.tensor(np.asarray(s["coordinates"], np.float32), device="cuda")
        pot.compute(sb, x, forces=True)
        torch.cuda.synchronize()
        outs.append((sb.e_atom[: sb.n_atoms].cpu().numpy().copy(), sb.g_aev[: sb.n_atoms].float().cpu().numpy().copy()))
    for e, g in outs[1:]:
        np.testing.assert_allclose(e, outs[0][0], rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(g, outs[0][1], rtol=1e-5, atol=1e-6 * np.abs(outs[0][1]).max())


def test_batched_systems_match_single(gpu):
    _flush("[efv] batched vs single")
    rng = np.random.Generator(np.random.PCG64(9))
    sysl = [systems.lattice_cluster(int(n), rng) for n in (12, 64, 150)]
    cfg = configs()
    params, pot = _potential(cfg)
    s = systems.batch_systems(sysl)
    E, F, V = pot.energy_forces_virial(s["species"], s["coordinates"], s["natoms"])
    off = 0
    for b, sysd in enumerate(sysl):
        e1, f1, v1 = pot.energy_forces_virial(sysd["species"], sysd["coordinates"])
        n = len(sysd["species"])
        np.testing.assert_allclose(E[b], e1[0], rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(F[off:off + n], f1, rtol=1e-5, atol=1e-5)
        off += n


def test_philox_and_baoab_injected_noise(gpu):
    _flush("[md] philox + baoab")
    n = 257
    lib = _abi.load()
    out = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    _abi.check(lib.fnl_philox_normals(out.data_ptr(), n, 1234567, 42, 0, _abi.stream_handle()), "normals")
    ref = O.philox_normals(1234567, 42, n, 0)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)
    rng = np.random.Generator(np.random.PCG64(1))
    x = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    gs = rng.normal(size=(n, 3)).astype(np.float32)
    # the angular neighbour gradients are int64 fixed point in units of 2^-32
    gn_fixed = np.rint(rng.normal(size=(n, 3)) * 2.0 ** 32).astype(np.int64)
    gn = gn_fixed.astype(np.float64) * 2.0 ** -32
    mass = rng.uniform(1, 16, size=n)
    noise = rng.normal(size=(n, 3))
    dt = 5e-4
    a1, a2 = O.langevin_coefficients(dt, 1.0, 0.596, mass)
    scale = 1.0 / 0.04336409463078789
    F = -(gs.astype(np.float64) + gn) * np.float64(np.float32(scale))
    xr, vr = O.baoab_step(x, v, F, mass, dt, a1, a2, noise)
    dev = "cuda"
    tx = torch.tensor(x, device=dev)
    tv = torch.tensor(v, device=dev)
    x32 = torch.empty((n, 3), dtype=torch.float32, device=dev)
    tdtm = torch.tensor(dt / mass, device=dev)
    ta2 = torch.tensor(a2, device=dev)
    tno = torch.tensor(noise, device=dev)
    tgs = torch.tensor(gs, device=dev)
    tgn = torch.tensor(gn_fixed, device=dev)
    _abi.check(lib.fnl_baoab_step(tx.data_ptr(), tv.data_ptr(), x32.data_ptr(), tgs.data_ptr(), tgn.data_ptr(),
                                  scale, tdtm.data_ptr(), 0.5 * dt, a1, ta2.data_ptr(), tno.data_ptr(), 0, 0, None, n,
                                  0, _abi.stream_handle()), "baoab")
    np.testing.assert_allclose(tx.cpu().numpy(), xr, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(tv.cpu().numpy(), vr, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(x32.cpu().numpy(), xr.astype(np.float32))
    np.testing.assert_array_equal(tgn.cpu().numpy(), gn_fixed)  # not cleared without the flag
    # device step counter {step, done}: two calls draw the noise of steps 7 and 8 (Philox), the
    # last workgroup advances the step; the neighbour-gradient rows are cleared once read
    step_dev = torch.tensor([7, 0], dtype=torch.int64, device=dev)
    xs, vs = x.copy(), v.copy()
    for k in (7, 8):
        Fk = -(gs.astype(np.float64) + (gn if k == 7 else 0.0)) * np.float64(np.float32(scale))
        xs, vs = O.baoab_step(xs, vs, Fk, mass, dt, a1, a2, O.philox_normals(99, k, n))
    tx, tv, tgn = torch.tensor(x, device=dev), torch.tensor(v, device=dev), torch.tensor(gn_fixed, device=dev)
    for _ in range(2):
        _abi.check(lib.fnl_baoab_step(tx.data_ptr(), tv.data_ptr(), x32.data_ptr(), tgs.data_ptr(),
                                      tgn.data_ptr(), scale, tdtm.data_ptr(), 0.5 * dt, a1, ta2.data_ptr(), None, 99,
                                      0, step_dev.data_ptr(), n, 1, _abi.stream_handle()), "baoab counter")
    assert step_dev.cpu().tolist() == [9, 0]
    assert np.all(tgn.cpu().numpy() == 0)
    np.testing.assert_allclose(tx.cpu().numpy(), xs, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(tv.cpu().numpy(), vs, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_forces_are_bitwise_reproducible(gpu, precision):
    """The angular force scatter sums int64 fixed point (exact in any order): repeated
    evaluations give the same bits."""
    _flush(f"[determinism] {precision}")
    s = water_box_small(512, 41)
    cfg = configs(("H", "C", "N", "O", "S", "F", "Cl"))
    _, pot = _potential(cfg, precision=precision, max_nbr=96, max_nbr_ang=40)
    sb = pot.setup(s["species"], s["natoms"], s["cells"])
    x = torch.tensor(np.asarray(s["coordinates"], np.float32), device="cuda")
    outs = []
    for _ in range(3):
        pot.compute(sb, x)
        outs.append((pot.forces(sb).cpu().numpy(), sb.vir_atom.cpu().numpy().copy(), sb.energy.cpu().numpy()))
    for f, v, e in outs[1:]:
        np.testing.assert_array_equal(f.view(np.uint32), outs[0][0].view(np.uint32))
        np.testing.assert_array_equal(v.view(np.uint32), outs[0][1].view(np.uint32))
        np.testing.assert_array_equal(e, outs[0][2])
