"""Seeded synthetic systems for the workloads this package is measured on.

All generators use NumPy ``PCG64`` with an explicit seed and return plain
NumPy arrays in FeNNol's input-dict layout (preprocessing.py:1763-1791):
``species`` int32[N] (atomic numbers), ``coordinates`` float64[N,3] in
Angstrom, ``natoms`` int32[B] and, for periodic boxes, ``cells`` float64[B,3,3]
(rows are lattice vectors) with ``flags = {"minimum_image": None}``.
"""

from __future__ import annotations

import math

import numpy as np

# species drawn for clusters: H, C, N, O with weights 1 : 1 : 0.6 : 0.4
CLUSTER_SPECIES = np.array([1, 6, 7, 8], dtype=np.int32)
CLUSTER_WEIGHTS = np.array([1.0, 1.0, 0.6, 0.4])

WATER_MOLAR_MASS = 15.999 + 2 * 1.008  # g/mol, masses of utils/periodic_table.py:37-46
AVOGADRO = 6.02214129e23               # utils/atomic_units.py:23
OH_BOND = 0.9572                       # Angstrom
HOH_ANGLE = 104.52                     # degrees


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _min_pair_distance(x: np.ndarray) -> float:
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    d[np.diag_indices_from(d)] = np.inf
    return float(d.min())


def lattice_cluster(n_atoms: int, rng: np.random.Generator, spacing: float = 1.45,
                    jitter: float = 0.15, min_dist: float = 0.9) -> dict:
    """One molecule-like cluster: a jittered cubic lattice, species H/C/N/O.

    The n sites closest to the lattice centre are used (a full 4x4x4 block for
    n = 64).  Jitter is uniform in [-jitter, jitter] per coordinate and is
    redrawn until every pair is at least ``min_dist`` apart.
    """
    k = max(1, math.ceil(n_atoms ** (1.0 / 3.0) - 1e-9))
    grid = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    centre = (k - 1) / 2.0
    order = np.lexsort((grid[:, 2], grid[:, 1], grid[:, 0], ((grid - centre) ** 2).sum(-1)))
    sites = grid[order[:n_atoms]] * spacing
    for _ in range(1000):
        x = sites + rng.uniform(-jitter, jitter, size=sites.shape)
        if n_atoms < 2 or _min_pair_distance(x) >= min_dist:
            break
    else:  # pragma: no cover - the jitter is far below spacing - min_dist
        raise RuntimeError("could not place atoms above the minimum distance")
    p = CLUSTER_WEIGHTS / CLUSTER_WEIGHTS.sum()
    species = rng.choice(CLUSTER_SPECIES, size=n_atoms, p=p).astype(np.int32)
    return {"species": species, "coordinates": x, "natoms": np.array([n_atoms], np.int32)}


def cluster_config0(seed: int = 0) -> dict:
    """64-atom cluster on a jittered 1.45 A lattice, no PBC."""
    return lattice_cluster(64, _rng(seed))


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random((3, n))
    q = np.stack([
        np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
        np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
        np.sqrt(u1) * np.sin(2 * np.pi * u3),
        np.sqrt(u1) * np.cos(2 * np.pi * u3),
    ], -1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box_length(n_waters: int, density_g_cm3: float = 1.0) -> float:
    """Cubic box edge (A) holding n waters at the given density."""
    vol_cm3 = n_waters * WATER_MOLAR_MASS / AVOGADRO / density_g_cm3
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def water_box(n_waters: int, sites_per_dim: int, seed: int, jitter: float = 0.15) -> dict:
    """Rigid waters randomly rotated on a jittered simple-cubic lattice in a periodic cube.

    ``n_waters`` of the ``sites_per_dim**3`` sites are chosen at random; the
    oxygen sits on the jittered site, the hydrogens at 0.9572 A with a
    104.52 degree H-O-H angle.  Atoms are ordered O, H, H per molecule.
    """
    rng = _rng(seed)
    L = water_box_length(n_waters)
    a = L / sites_per_dim
    grid = np.stack(np.meshgrid(*[np.arange(sites_per_dim)] * 3, indexing="ij"), -1).reshape(-1, 3)
    chosen = np.sort(rng.choice(grid.shape[0], size=n_waters, replace=False))
    o = (grid[chosen] + 0.5) * a + rng.uniform(-jitter, jitter, size=(n_waters, 3))
    half = math.radians(HOH_ANGLE) / 2.0
    local = np.array([
        [0.0, 0.0, 0.0],
        [OH_BOND * math.sin(half), OH_BOND * math.cos(half), 0.0],
        [-OH_BOND * math.sin(half), OH_BOND * math.cos(half), 0.0],
    ])
    R = _random_rotations(rng, n_waters)
    mol = np.einsum("nij,kj->nki", R, local) + o[:, None, :]
    coords = mol.reshape(-1, 3)
    species = np.tile(np.array([8, 1, 1], np.int32), n_waters)
    cells = (np.eye(3) * L)[None]
    return {
        "species": species,
        "coordinates": coords,
        "natoms": np.array([coords.shape[0]], np.int32),
        "cells": cells,
        "reciprocal_cells": np.linalg.inv(cells),
        "flags": {"minimum_image": None},
    }


def water_config1(seed: int = 1) -> dict:
    """2,592 waters (7,776 atoms) at 1 g/cm^3, 14^3 lattice sites."""
    return water_box(2592, 14, seed)


def water_config2(seed: int = 2) -> dict:
    """7,776 waters (23,328 atoms) at 1 g/cm^3, 20^3 lattice sites."""
    return water_box(7776, 20, seed)


def batch_systems(systems: list[dict]) -> dict:
    """Concatenate systems into one batch (natoms / batch_index, preprocessing.py:1773-1777)."""
    species = np.concatenate([s["species"] for s in systems]).astype(np.int32)
    coords = np.concatenate([s["coordinates"] for s in systems])
    natoms = np.array([s["species"].shape[0] for s in systems], np.int32)
    out = {"species": species, "coordinates": coords, "natoms": natoms,
           "batch_index": np.repeat(np.arange(len(systems), dtype=np.int32), natoms)}
    if all("cells" in s for s in systems):
        out["cells"] = np.concatenate([s["cells"] for s in systems])
        out["reciprocal_cells"] = np.linalg.inv(out["cells"])
        out["flags"] = {"minimum_image": None}
    return out


def training_boxes(n_boxes: int = 256, n_waters: int = 64, seed: int = 4) -> dict:
    """Batch of small periodic water boxes (64 waters on a 4^3 lattice each)."""
    k = round(n_waters ** (1.0 / 3.0))
    boxes = [water_box(n_waters, k, seed * 100003 + i) for i in range(n_boxes)]
    return batch_systems(boxes)


def serving_requests(n_requests: int = 8192, seed: int = 3, n_min: int = 8, n_max: int = 512,
                     md_fraction: float = 0.3, md_steps: int = 200):
    """Request stream: cluster sizes log-uniform in [n_min, n_max], 30% short MD runs."""
    rng = _rng(seed)
    sizes = np.floor(np.exp(rng.uniform(math.log(n_min), math.log(n_max + 1), size=n_requests))).astype(int)
    sizes = np.clip(sizes, n_min, n_max)
    kinds = rng.random(n_requests) < md_fraction
    reqs = []
    for i in range(n_requests):
        sysd = lattice_cluster(int(sizes[i]), rng)
        reqs.append({"system": sysd, "md_steps": md_steps if kinds[i] else 0})
    return reqs
