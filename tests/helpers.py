"""Shared test helpers: matching oracle / product configurations and canonical pair sets."""

from __future__ import annotations

import numpy as np

import oracle as O
from fennol_amd import systems


def configs(species_order=("H", "C", "N", "O"), **kw):
    """(oracle AEVParams, product ANIConfig) with identical hyper-parameters."""
    from fennol_amd.potential import ANIConfig

    p = O.AEVParams(species_order=tuple(species_order), **kw)
    c = ANIConfig(species_order=tuple(species_order), **kw)
    return p, c


def canonical_pairs_oracle(coords, natoms, cutoff, cells=None, minimum_image=True):
    cells_ = None if cells is None else np.asarray(cells)
    rec = None if cells is None else np.linalg.inv(cells_)
    s, d, d12, sh = O.half_pairs(coords, natoms, cutoff, cells_, rec, minimum_image)
    order = np.lexsort((d, s))
    return np.stack([s[order], d[order]], 1), d12[order]


def rows_to_pairs(nbr, n_nbr, max_nbr):
    """All directed (i, j) of the ELL rows (host arrays)."""
    n = np.minimum(n_nbr, max_nbr)
    i = np.repeat(np.arange(nbr.shape[0]), n)
    mask = np.arange(max_nbr)[None, :] < n[:, None]
    j = nbr[mask]
    return i, j


def water_box_small(n_waters=64, seed=0):
    k = int(round(n_waters ** (1 / 3)))
    if k ** 3 < n_waters:
        k += 1
    return systems.water_box(n_waters, k, seed)


def triclinic_box(n_atoms=120, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    cell = np.array([[13.0, 0.0, 0.0], [2.5, 12.0, 0.0], [-1.5, 3.0, 12.5]])
    frac = rng.random((n_atoms, 3))
    x = frac @ cell
    species = rng.choice(np.array([1, 6, 7, 8], np.int32), size=n_atoms)
    return {"species": species, "coordinates": x, "natoms": np.array([n_atoms], np.int32),
            "cells": cell[None], "reciprocal_cells": np.linalg.inv(cell)[None]}
