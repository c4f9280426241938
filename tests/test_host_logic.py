"""CPU tests of the host-side logic (no kernel launches)."""

from __future__ import annotations

from collections import OrderedDict

import numpy as np
import pytest

import oracle as O
from fennol_amd import _abi, systems

needs_lib = pytest.mark.skipif(not _abi.LIB_PATH.exists(), reason="HIP library not built")


@needs_lib
def test_pool_layout_tables():
    from fennol_amd.potential import ANIConfig, ANIPotential

    cfg = ANIConfig()
    pot = ANIPotential(cfg, O.init_mlp_params(1, cfg.aev_length), device="cpu")
    rng = np.random.Generator(np.random.PCG64(0))
    sysl = [systems.lattice_cluster(int(n), rng) for n in (8, 100, 513 - 1, 37)]
    b = systems.batch_systems(sysl)
    lay = pot.pool_layout_arrays(b["species"], b["natoms"])
    N = len(b["species"])
    assert lay["n_atoms"] == N and lay["n_sys"] == 4
    np.testing.assert_array_equal(lay["sys_offsets"], np.concatenate([[0], np.cumsum(b["natoms"])]))
    # perm: every atom once, grouped by network in species order
    assert sorted(lay["perm"].tolist()) == list(range(N))
    z_of_net = [1, 6, 7, 8]
    covered = []
    for s, r, n in zip(lay["tile_start"], lay["tile_rows"], lay["tile_net"]):
        assert 0 < r <= pot.nets.tile_rows
        rows = lay["perm"][s:s + r]
        assert np.all(b["species"][rows] == z_of_net[n])
        covered.extend(rows.tolist())
    assert sorted(covered) == list(range(N))
    # cell grids: caps k^3 with k = ceil(n^(1/3)), consecutive first cells
    k = lay["grid"][:, 0]
    assert np.all(k == np.ceil(b["natoms"] ** (1 / 3) - 1e-9).astype(int))
    np.testing.assert_array_equal(lay["grid"][1:, 3], np.cumsum(k.astype(np.int64) ** 3)[:-1])
    assert lay["n_cells"] == int((k.astype(np.int64) ** 3).sum())
    np.testing.assert_array_equal(np.diff(lay["seg_offsets"]), np.ceil(b["natoms"] / 512).astype(int))


@needs_lib
def test_unhandled_species_is_rejected():
    from fennol_amd.potential import ANIConfig, ANIPotential

    cfg = ANIConfig(species_order=("H", "C", "N", "O", "S", "F", "Cl"))
    params = O.init_mlp_params(1, cfg.aev_length, species=("H", "O"))
    pot = ANIPotential(cfg, params, device="cpu")
    with pytest.raises(ValueError):
        pot.pool_layout_arrays(np.array([1, 6, 8], np.int32), np.array([3], np.int32))


def test_fennix_rejects_other_modules():
    from fennol_amd.fennix import FENNIX

    with pytest.raises(NotImplementedError):
        FENNIX(cutoff=5.0, modules=OrderedDict(emb={"module_name": "SCHNET_EMBEDDING"}), device="cpu")


def test_serving_request_stream_properties():
    from fennol_amd.serving import make_requests, poisson_arrivals

    reqs = make_requests(400, seed=3)
    n = np.array([r.n for r in reqs])
    assert n.min() >= 8 and n.max() <= 512
    md = np.array([r.md_steps > 0 for r in reqs])
    assert 0.2 < md.mean() < 0.4
    assert np.all(np.array([r.md_steps for r in reqs])[md] == 200)
    t = poisson_arrivals(20000, 50.0, seed=1)
    assert np.all(np.diff(t) > 0)
    assert abs(20000 / t[-1] - 50.0) < 2.0


def test_fnx_roundtrip_and_layout(tmp_path):
    """msgpack .fnx encoding: arrays as ext type 1 (shape, dtype, bytes), scalars as ext type 3."""
    import msgpack

    from fennol_amd import fnx
    from fennol_amd.params import init_species_params, to_variables

    variables = to_variables(init_species_params(2, 40, species=("H", "O")))
    state = {"cutoff": 5.2, "modules": [["energy", {"module_name": "SPECIES_INDEX_NET"}]],
             "preprocessing": [], "energy_unit": "Ha", "energy_terms": ["energy"], "variables": variables,
             "scalar": np.float32(1.5)}
    data = fnx.dumps(state)
    back = fnx.loads(data)
    W = back["variables"]["params"]["energy"]["H"]["Layer_1"]["kernel"]
    np.testing.assert_array_equal(W, variables["params"]["energy"]["H"]["Layer_1"]["kernel"])
    assert back["scalar"] == np.float32(1.5) and back["cutoff"] == 5.2
    # raw structure: ext code 1 with a msgpack (shape, dtype name, bytes) payload
    raw = msgpack.unpackb(data, raw=False, strict_map_key=False)
    ext = raw["variables"]["params"]["energy"]["O"]["Layer_2"]["bias"]
    assert isinstance(ext, msgpack.ExtType) and ext.code == 1
    shape, dt, buf = msgpack.unpackb(ext.data, raw=False)
    assert shape == [192] and dt == "float64" and len(buf) == 192 * 8


def test_nblist_schedule_matches_fennol_md():
    """Rebuild stride from the skin (md/integrate.py:488-507): a 2 A skin lasts 40 fs."""
    from fennol_amd.md import nblist_schedule

    assert nblist_schedule(-1.0, 0.5) == 1
    assert nblist_schedule(0.0, 0.5) == 1
    assert nblist_schedule(2.0, 0.5) == 80
    assert nblist_schedule(2.0, 1.0) == 40
    assert nblist_schedule(1.0, 0.5) == 40
    assert nblist_schedule(1.5, 0.7) == 42  # floor(0.75 * 40 / 0.7)
    assert nblist_schedule(2.0, 0.5, stride=10) == 10


def test_bf16_fragment_order_layout():
    """BF16 weights as the MLP kernel reads them: element (r, c) of an [R][C] matrix at
    ((r/16)*(C/32) + c/32)*512 + ((c%32)/8)*128 + (r%16)*8 + c%8, element by element."""
    from fennol_amd.potential import fragment_order

    rng = np.random.default_rng(5)
    R, C = 48, 96
    a = rng.standard_normal((R, C))
    flat = fragment_order(a).reshape(-1)
    assert flat.shape == (R * C,)
    for r in range(R):
        for c in range(C):
            off = ((r // 16) * (C // 32) + c // 32) * 512 + ((c % 32) // 8) * 128 + (r % 16) * 8 + c % 8
            assert flat[off] == a[r, c]
    with pytest.raises(ValueError):
        fragment_order(np.zeros((20, 32)))


def test_atom_padding_and_unpadding():
    """AtomPadding (preprocessing.py:1654-1729): atoms to int(1.2 N) + 1 when N grows, one padding
    system (natoms 0, cell 1000 I), masks; atom_unpadding (:1732-1760) restores the arrays."""
    from fennol_amd.preprocessing import AtomPadding, atom_unpadding, check_input

    pad = AtomPadding()
    rng = np.random.default_rng(0)
    x = rng.normal(size=(10, 3))
    inp = check_input({"species": np.full(10, 8, np.int32), "coordinates": x, "cells": 9.0 * np.eye(3)})
    st, out = pad(pad.init(), inp)
    assert st == {"prev_nat": 13, "prev_nsys": 1}
    assert out["species"].shape == (13,) and np.all(out["species"][10:] == -1)
    np.testing.assert_array_equal(out["natoms"], [10, 0])
    np.testing.assert_array_equal(out["batch_index"][10:], [1, 1, 1])
    np.testing.assert_array_equal(out["cells"][1], 1000.0 * np.eye(3))
    assert out["true_atoms"].sum() == 10 and list(out["true_sys"]) == [True, False]
    back = atom_unpadding({**out, "forces": np.ones((13, 3)), "total_energy": np.array([1.0, 0.0])})
    np.testing.assert_array_equal(back["coordinates"], x)
    assert back["forces"].shape == (10, 3) and back["total_energy"].shape == (1,)
    # a smaller system keeps the padded size
    inp2 = check_input({"species": np.full(7, 1, np.int32), "coordinates": x[:7], "cells": 9.0 * np.eye(3)})
    st2, out2 = pad(st, inp2)
    assert st2 == st and out2["species"].shape == (13,)


def test_serving_noise_and_velocity_streams_are_disjoint():
    """The O-step noise of request r (counter (atom, step, 2r)) never reuses the counters of any
    request's initial velocities (atom, 0, 2r' + 1); request 0 matches a single LangevinMD run
    (noise stream 0, velocities stream 1)."""
    from fennol_amd.serving import noise_stream, velocity_stream

    noise = {noise_stream(r) for r in range(5000)}
    vel = {velocity_stream(r)[1] for r in range(5000)}
    assert not noise & vel
    assert noise_stream(0) == 0 and velocity_stream(0) == (0, 1)
    assert len(vel) == 5000 and len(noise) == 5000


def test_grown_capacity():
    from fennol_amd.md import grown_capacity

    assert grown_capacity(64, 60) == 64
    assert grown_capacity(64, 70) == 96  # int(1.25 * 70) + 8 = 95 -> multiple of 8
    assert grown_capacity(16, 69) % 8 == 0 and grown_capacity(16, 69) >= 69


def test_grid_caps_leave_room_for_half_cutoff_cells():
    """Periodic grid capacities hold floor(2 w / rc') + 1 cells per lattice direction (w the
    lattice-plane spacing, rc' the cutoff with the kernel's margin), so the build can use cells
    half a cutoff wide; open systems get about one cell per atom; empty systems one cell; the
    first-cell offsets are the running sums of the capacities."""
    from fennol_amd.potential import grid_caps

    rc = 5.2
    rbin = rc * 1.0001 + 1.0e-4
    box = systems.water_config1()
    cells = np.asarray(box["cells"], np.float64)
    natoms = np.array([len(box["species"]), 0, 64], np.int32)
    g, n_cells = grid_caps(rc, natoms, np.stack([cells[0], np.eye(3) * 1000.0, cells[0]]))
    w = 1.0 / np.sqrt((np.linalg.inv(cells[0]) ** 2).sum(axis=0))
    np.testing.assert_array_equal(g[0, :3], np.floor(2.0 * w / rbin).astype(int) + 1)
    assert (g[0, :3] >= 2 * np.floor(w / rbin)).all()
    np.testing.assert_array_equal(g[1, :3], [1, 1, 1])
    np.testing.assert_array_equal(g[:, 3], [0, int(np.prod(g[0, :3])), int(np.prod(g[0, :3])) + 1])
    assert n_cells == int(np.prod(g[0, :3])) + 1 + int(np.prod(g[2, :3]))
    g_open, _ = grid_caps(rc, np.array([512], np.int32), None)
    np.testing.assert_array_equal(g_open[0, :3], [8, 8, 8])
