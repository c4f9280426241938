"""The committed golden fixtures (tests/golden/*.npz) against the oracle.

The fixtures were written by tests/golden/make_golden.py from the FP64 oracle;
recomputing every case here pins the oracle against silent drift (a changed
hyper-parameter, order of terms, unit or Philox counter shows up as a
mismatch).  The oracle itself is pinned by the closed forms, finite
differences and invariances of test_oracle.py.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
from tests.golden import make_golden as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


@pytest.fixture(scope="module")
def recomputed():
    return {}


def _case(recomputed, name):
    if name not in recomputed:
        recomputed[name] = G.CASES[name]()
    return recomputed[name]


def test_weights_digest():
    p = O.AEVParams(species_order=G.SPECIES4)
    digest = G.weights_digest(O.init_mlp_params(G.WEIGHT_SEED, O.aev_length(p)))
    for name in G.CASES:
        f = _load(name)
        assert int(f["weight_seed"]) == G.WEIGHT_SEED
        assert str(f["weights_sha256"]) == digest, name


@pytest.mark.parametrize("name", ["cluster64", "water2", "triclinic5"])
def test_systems(recomputed, name):
    f = _load(name)
    r = _case(recomputed, name)
    np.testing.assert_array_equal(f["species"], r["species"])
    np.testing.assert_array_equal(f["coordinates"], r["coordinates"])
    # float32 coordinates exactly: the device path starts from the same numbers
    assert np.all(f["coordinates"].astype(np.float32).astype(np.float64) == f["coordinates"])
    np.testing.assert_array_equal(f["pair_src"], r["pair_src"])
    np.testing.assert_array_equal(f["pair_dst"], r["pair_dst"])
    np.testing.assert_array_equal(f["pair_d12"].view(np.uint32), r["pair_d12"].view(np.uint32))
    if "pair_shifts" in f:
        np.testing.assert_array_equal(f["pair_shifts"], r["pair_shifts"])
    np.testing.assert_allclose(f["aev"], r["aev"], rtol=1e-6, atol=1e-7)
    for k in ("energy", "forces", "virial", "atomic_energies"):
        np.testing.assert_allclose(f[k], r[k], rtol=1e-10, atol=1e-12, err_msg=k)


def test_system_invariants():
    """Properties the fixtures must hold whatever the oracle: translation invariance
    (sum of forces 0 for the isolated cluster), symmetric virial, E = sum of atomic energies."""
    for name in ("cluster64", "water2", "triclinic5"):
        f = _load(name)
        np.testing.assert_allclose(f["atomic_energies"].sum(), f["energy"].sum(), rtol=1e-12)
        np.testing.assert_allclose(f["forces"].sum(0), 0.0, atol=1e-9)
        V = f["virial"].reshape(3, 3)
        np.testing.assert_allclose(V, V.T, atol=1e-9)
    # general PBC with cell edges below 2 rc: images of the same atom pair appear
    f = _load("triclinic5")
    keys = set()
    dup = False
    for s, d in zip(f["pair_src"].tolist(), f["pair_dst"].tolist()):
        dup |= (s, d) in keys
        keys.add((s, d))
    assert dup
    # cluster: no periodic shifts, every pair within the cutoff
    f = _load("cluster64")
    assert np.all(f["pair_d12"] < np.float32(5.2 ** 2))


def test_small_aevs(recomputed):
    f = _load("small_aevs")
    r = _case(recomputed, "small_aevs")
    for k in r:
        np.testing.assert_allclose(f[k], r[k], rtol=1e-12, atol=1e-14, err_msg=k)
    # the dimer has no angular terms and exactly one radial block per atom
    p = O.AEVParams(species_order=G.SPECIES4)
    nr = p.radial_length
    aev = f["dimer_aev"]
    assert np.all(aev[:, nr:] == 0.0)
    assert np.count_nonzero(aev[0, :nr]) == np.count_nonzero(aev[1, :nr]) > 0


def test_mlp8(recomputed):
    f = _load("mlp8")
    r = _case(recomputed, "mlp8")
    np.testing.assert_allclose(f["energy"], r["energy"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f["grad"], r["grad"], rtol=1e-12, atol=1e-14)


def test_baoab10(recomputed):
    f = _load("baoab10")
    r = _case(recomputed, "baoab10")
    for k in ("x0", "v0", "noise", "x", "v", "energies"):
        np.testing.assert_allclose(f[k], r[k], rtol=1e-12, atol=1e-12, err_msg=k)
    # the noise of step k is Philox stream 0 at counter k
    np.testing.assert_array_equal(f["noise"][3], O.philox_normals(77, 3, 6))
