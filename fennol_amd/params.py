"""Seeded random initialisation of species-network weights.

FullyConnectedNet layers use Flax's ``lecun_normal`` (models/misc/nets.py:
59-63, utils/initializers.py:33-47): a normal truncated to [-2, 2] scaled by
sqrt(1/fan_in) / 0.87962566103423978, and zero biases.  Values are drawn with
NumPy's PCG64 from an explicit seed (Flax's threefry streams are not
reproduced); the layout is {symbol: [(kernel [in, out], bias [out]), ...]},
matching the parameter tree {"Layer_1": {"kernel", "bias"}, ...} per species.
"""

from __future__ import annotations

import math

import numpy as np

TRUNC_STD = 0.87962566103423978  # std of a standard normal truncated to [-2, 2]


def lecun_truncated(rng: np.random.Generator, fan_in: int, fan_out: int) -> np.ndarray:
    w = rng.standard_normal((fan_in, fan_out))
    bad = np.abs(w) > 2.0
    while np.any(bad):
        w[bad] = rng.standard_normal(int(bad.sum()))
        bad = np.abs(w) > 2.0
    return w * (math.sqrt(1.0 / fan_in) / TRUNC_STD)


def init_species_params(seed: int, n_in: int, hidden=(256, 192, 160), n_out: int = 1,
                        species=("H", "C", "N", "O"), zero_bias: bool = True) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = [n_in, *hidden, n_out]
    out = {}
    for s in species:
        layers = []
        for a, b in zip(dims[:-1], dims[1:]):
            w = lecun_truncated(rng, a, b)
            bias = np.zeros(b) if zero_bias else 0.1 * rng.standard_normal(b)
            layers.append((w, bias))
        out[s] = layers
    return out


def to_variables(params: dict, module_name: str = "energy") -> dict:
    """Flax-style variables tree of a SpeciesIndexNet (nets.py:339-369)."""
    tree = {}
    for s, layers in params.items():
        tree[s] = {f"Layer_{i + 1}": {"kernel": np.asarray(w), "bias": np.asarray(b)}
                   for i, (w, b) in enumerate(layers)}
    return {"params": {module_name: tree}}


def to_chemical_variables(params: dict, module_name: str = "energy") -> dict:
    """Flax-style variables tree of a ChemicalNet (nets.py:429-512): the species networks
    vmapped over a leading species axis, {"VmapFullyConnectedNet_0": {"Layer_i": {"kernel"
    [S, in, out], "bias" [S, out]}}}, species in the module's species_order."""
    species = list(params.keys())
    n = len(params[species[0]])
    tree = {f"Layer_{i + 1}": {"kernel": np.stack([np.asarray(params[s][i][0]) for s in species]),
                               "bias": np.stack([np.asarray(params[s][i][1]) for s in species])}
            for i in range(n)}
    return {"params": {module_name: {"VmapFullyConnectedNet_0": tree}}}


def from_chemical_variables(variables: dict, species_order, module_name: str = "energy") -> dict:
    """Per-species layers of a ChemicalNet variables tree (the inverse of to_chemical_variables;
    the vmapped sub-tree is the one child holding Layer_* entries)."""
    tree = variables["params"][module_name]
    if "Layer_1" not in tree:
        subs = [v for v in tree.values() if isinstance(v, dict) and "Layer_1" in v]
        if len(subs) != 1:
            raise ValueError("ChemicalNet variables: expected one vmapped FullyConnectedNet sub-tree")
        tree = subs[0]
    n = len([k for k in tree if k.startswith("Layer_")])
    out = {}
    for k, s in enumerate(species_order):
        out[s] = [(np.asarray(tree[f"Layer_{i + 1}"]["kernel"])[k], np.asarray(tree[f"Layer_{i + 1}"]["bias"])[k])
                  for i in range(n)]
    return out


def from_variables(variables: dict, module_name: str = "energy") -> dict:
    tree = variables["params"][module_name]
    out = {}
    for s, layers in tree.items():
        n = len(layers)
        out[s] = [(np.asarray(layers[f"Layer_{i + 1}"]["kernel"]), np.asarray(layers[f"Layer_{i + 1}"]["bias"]))
                  for i in range(n)]
    return out
