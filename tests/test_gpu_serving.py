"""GPU tests of the continuous-batching server (device-built layout, graph-replayed iterations,
pipelined results): each request's result is the same bits alone and in any mix, also when it is
admitted while others are mid-trajectory or restarted after a neighbour-row overflow; the device
tables equal the host layout; single evaluations match the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from fennol_amd import _abi, systems
from fennol_amd.potential import ANIConfig, ANIPotential
from fennol_amd.serving import Request, ServingEngine

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _params():
    return O.init_mlp_params(11, O.aev_length(O.AEVParams()))


def _requests(sizes, md_every=3, md_steps=12, seed=21, arrivals=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k, n in enumerate(sizes):
        c = systems.lattice_cluster(n, rng)
        out.append(Request(c["species"], c["coordinates"], md_steps=(md_steps if k % md_every == 0 else 0),
                           arrival=0.0 if arrivals is None else float(arrivals[k]), rid=k))
    return out


def _copy(r: Request, arrival=None) -> Request:
    return Request(r.species, r.coordinates, r.md_steps, r.arrival if arrival is None else arrival, r.rid)


def _serve(reqs, precision="fp32", realtime=False, **kw):
    pot = ANIPotential(ANIConfig(), _params(), precision=precision)
    kw.setdefault("max_systems", 16)
    kw.setdefault("seed", 4)
    eng = ServingEngine(pot, **kw)
    batch = [_copy(r) for r in reqs]
    eng.run(batch, realtime=realtime)
    return {r.rid: r for r in batch}, eng


def _alone(reqs, precision="fp32"):
    out = {}
    for r in reqs:
        res, _ = _serve([_copy(r, 0.0)], precision=precision, max_atoms=512)
        out[r.rid] = res[r.rid]
    return out


def _assert_same(a: Request, b: Request):
    assert a.steps_done == b.steps_done
    # the same bits: nothing of a request depends on its rows or on the other requests
    assert a.energy == b.energy
    np.testing.assert_array_equal(a.forces, b.forces)
    np.testing.assert_array_equal(a.positions, b.positions)


SIZES = (9, 40, 77, 130, 256, 17, 64, 300)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_serving_results_same_alone_and_in_mix(gpu, precision):
    """Continuous batching with a pool too small for all requests at once (staggered
    admission): every result equals the result of the request served alone."""
    _flush(f"[serving] mix vs alone {precision}")
    reqs = _requests(SIZES)
    mix, eng = _serve(reqs, precision=precision, max_atoms=700)
    assert eng.layouts > 2 and eng.iterations > 12
    alone = _alone(reqs, precision)
    for r in reqs:
        assert mix[r.rid].steps_done == r.md_steps
        _assert_same(mix[r.rid], alone[r.rid])


def test_serving_admits_while_md_requests_are_mid_trajectory(gpu):
    """Arrivals at iterations 0..10 while 12-step MD requests run: the pool changes under them
    (rows move when earlier requests leave), and every result is still its alone result."""
    _flush("[serving] admission mid-trajectory")
    arrivals = [0, 0, 1, 3, 4, 6, 8, 10]
    reqs = _requests(SIZES, md_every=2, arrivals=arrivals)
    mix, eng = _serve(reqs, max_atoms=1024, clock="iterations", realtime=True)
    assert eng.layouts >= 5
    admitted = {r.rid: r.t_admit for r in mix.values()}
    assert admitted[7] >= 10 and admitted[0] == 0
    alone = _alone(reqs)
    for r in reqs:
        _assert_same(mix[r.rid], alone[r.rid])


def test_serving_restarts_requests_whose_rows_overflowed(gpu):
    """Rows far too short (16 radial / 8 angular slots): the first iteration flags every request,
    the rows are regrown, graphs captured again and the requests rerun from their inputs --
    results equal to a pool with ample rows."""
    _flush("[serving] overflow restart")
    reqs = _requests(SIZES[:5])
    small, eng = _serve(reqs, max_atoms=1024, max_nbr=16, max_nbr_ang=8)
    assert eng.regrowths >= 1 and eng.sb.max_nbr > 16 and eng.sb.max_nbr_ang > 8
    ample, _ = _serve(reqs, max_atoms=1024)
    for r in reqs:
        # the 9-atom cluster fits 16 slots; every larger one overflowed and was rerun
        assert (small[r.rid].restarts >= 1) == (r.n > 17)
        _assert_same(small[r.rid], ample[r.rid])


def test_serving_graph_replay_equals_eager_launches(gpu):
    _flush("[serving] graphs vs eager")
    reqs = _requests(SIZES[:6])
    g, eng_g = _serve(reqs, max_atoms=800, graphs=True)
    e, _ = _serve(reqs, max_atoms=800, graphs=False)
    assert eng_g._graphs  # at least one row extent captured
    for r in reqs:
        _assert_same(g[r.rid], e[r.rid])


def test_serving_device_layout_equals_host_layout(gpu):
    """fnl_pool_layout's tables equal the host layout of the same requests (segment offsets,
    cell-grid capacities, species index, network-sorted atoms and tiles)."""
    _flush("[serving] device layout")
    pot = ANIPotential(ANIConfig(), _params(), precision="bf16")
    eng = ServingEngine(pot, max_atoms=2048, max_systems=32, clock="iterations")
    reqs = _requests((9, 40, 77, 130, 256, 17, 64, 300, 1, 2, 27, 28), md_every=4)
    for r in reqs:
        eng.submit(r)
    eng._layout(eng._admit(0.0))
    torch.cuda.synchronize()
    species = np.concatenate([r.species for r in eng.active])
    natoms = np.array([r.n for r in eng.active])
    lay = pot.pool_layout_arrays(species, natoms)
    sb = eng.sb
    N, B, T = lay["n_atoms"], lay["n_sys"], len(lay["tile_start"])
    cnt = eng.counts.cpu().numpy()
    assert (cnt[_abi.FNL_POOL_N_ATOMS], cnt[_abi.FNL_POOL_N_SYS], cnt[_abi.FNL_POOL_N_TILES],
            cnt[_abi.FNL_POOL_ERROR]) == (N, B, T, 0)
    np.testing.assert_array_equal(sb.sys_offsets[: B + 1].cpu().numpy(), lay["sys_offsets"])
    np.testing.assert_array_equal(sb.seg_offsets[: B + 1].cpu().numpy(), lay["seg_offsets"])
    np.testing.assert_array_equal(sb.grid[:B].cpu().numpy(), lay["grid"])
    np.testing.assert_array_equal(sb.batch_index[:N].cpu().numpy(), lay["batch_index"])
    np.testing.assert_array_equal(sb.species_idx[:N].cpu().numpy(), lay["species_idx"])
    assert np.all(sb.species_idx[N:].cpu().numpy() == -1)
    np.testing.assert_array_equal(sb.perm[: len(lay["perm"])].cpu().numpy(), lay["perm"])
    for name in ("tile_start", "tile_rows", "tile_net"):
        np.testing.assert_array_equal(getattr(sb, name)[:T].cpu().numpy(), lay[name])
    np.testing.assert_array_equal(eng.x[:N].cpu().numpy(), np.concatenate([r.coordinates for r in eng.active]))
    # a layout beyond the row capacity is refused on the device: empty pool, error word set
    req = np.zeros(1, dtype=[("n", "<i4"), ("rest", "<i4", 9)])
    req["n"] = eng.max_atoms + 1
    eng.req_dev[:40].copy_(torch.from_numpy(req.view(np.uint8)))
    _abi.check(eng.lib.fnl_pool_layout(__import__("ctypes").byref(eng.pool), eng.req_dev.data_ptr(), 1,
                                       eng.stage_z.data_ptr(), eng.stage_x.data_ptr(), _abi.stream_handle()),
               "fnl_pool_layout")
    cnt = eng.counts.cpu().numpy()
    assert cnt[_abi.FNL_POOL_ERROR] == 1 and cnt[_abi.FNL_POOL_N_ATOMS] == 0 and cnt[_abi.FNL_POOL_N_SYS] == 0


def test_serving_single_evaluations_match_the_fp64_oracle(gpu):
    """Energy + forces requests served in a mix: FP32 within 1e-5 relative (energy) and 1e-4 eV/A
    RMS (forces) of the FP64 oracle; BF16 within 1 meV/atom and 0.02 eV/A."""
    _flush("[serving] FP64 oracle")
    reqs = _requests((9, 40, 77, 130, 64), md_steps=0)
    model = O.Model(O.AEVParams(), _params())
    for precision in ("fp32", "bf16"):
        res, _ = _serve(reqs, precision=precision, max_atoms=512)
        for r in reqs:
            x64 = np.asarray(r.coordinates, np.float32).astype(np.float64)
            Er, Fr, _, _ = O.energy_forces_virial(model, x64, r.species, np.array([r.n]))
            m = res[r.rid]
            f_rms = np.sqrt(np.mean((m.forces - Fr) ** 2))
            if precision == "fp32":
                np.testing.assert_allclose(m.energy, Er[0], rtol=1e-5, atol=1e-6)
                assert f_rms < 1e-4
            else:
                assert 1000 * abs(m.energy - Er[0]) / r.n <= 1.0 and f_rms <= 0.02
