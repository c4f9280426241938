"""FeNNol preprocessing modules for the ANI path, backed by the HIP neighbour lists.

Same names, ``FPID`` identifiers, state keys and dictionary contracts as
FeNNol's models/preprocessing.py:
  GraphGenerator        GRAPH                    (:21-696)
  GraphFilter           GRAPH_FILTER             (:780-960)
  GraphAngularExtension GRAPH_ANGULAR_EXTENSION  (:1046-1280)
  SpeciesIndexer        SPECIES_INDEXER          (:1320-1490)
  PreprocessingChain                             (:1824-1925)
  check_input                                    (:1763-1791)

``process(state, inputs)`` is the device path: graphs are kept in the cell-list
row layout of ``fnl_nblist_build`` (keys ``nbr``, ``n_nbr`` and, for an angular
graph, ``nbr_ang``/``n_nbr_ang`` holding radial slots) with an ``overflow``
flag on the device; ``check_reallocate`` reads it and grows the rows as the
reference does (:357-367).  ``__call__(state, inputs)`` is the host path: it
runs the device path and returns the reference's flat arrays (``edge_src``,
``edge_dst``, ``d12``, ``pbc_shifts``, ``filter_indices``, ``angle_src``,
``angle_dst``, ``central_atom``) as NumPy arrays, padded to capacities that
grow by ``mult_size`` like the reference's (``int(mult*max(n, prev)) + 1``).
A GraphFilter whose parent is the radial graph is fused into the generator's
kernel (one candidate scan fills both rows).
"""

from __future__ import annotations

import dataclasses
from typing import ClassVar

import numpy as np
import torch

from . import _abi
from .units import PERIODIC_TABLE, SYMBOL_TO_Z


def check_input(inputs: dict) -> dict:
    """Defaults for natoms / batch_index, cell shapes (preprocessing.py:1763-1791)."""
    assert "species" in inputs, "species must be provided"
    assert "coordinates" in inputs, "coordinates must be provided"
    species = np.asarray(inputs["species"]).astype(np.int32)
    ifake = int(np.argmax(species <= 0)) if np.any(species <= 0) else 0
    if ifake > 0:
        assert np.all(species[:ifake] > 0), "species must be positive"
    nat = species.shape[0]
    natoms = np.asarray(inputs.get("natoms", np.array([nat], np.int32))).astype(np.int32)
    batch_index = np.asarray(inputs.get("batch_index", np.repeat(np.arange(len(natoms), dtype=np.int32), natoms)))
    out = {**inputs, "species": species, "natoms": natoms, "batch_index": batch_index.astype(np.int32)}
    if "cells" in inputs:
        cells = np.asarray(inputs["cells"], np.float64)
        rec = np.asarray(inputs["reciprocal_cells"]) if "reciprocal_cells" in inputs else np.linalg.inv(cells)
        out["cells"] = cells[None] if cells.ndim == 2 else cells
        out["reciprocal_cells"] = rec[None] if rec.ndim == 2 else rec
    return out


def _grow(n: int, prev: int, mult: float) -> int:
    return int(mult * max(n, prev)) + 1


@dataclasses.dataclass(frozen=True)
class GraphGenerator:
    """Radial graph within ``cutoff`` (cell list on the GPU).  FPID: GRAPH"""

    cutoff: float
    graph_key: str = "graph"
    switch_params: dict = dataclasses.field(default_factory=dict, hash=False)
    mult_size: float = 1.05
    FPID: ClassVar[str] = "GRAPH"

    def init(self):
        return {"max_nat": 1, "npairs": 1, "nblist_mult_size": self.mult_size}

    def get_graph_properties(self):
        return {self.graph_key: {"cutoff": self.cutoff, "directed": True}}


@dataclasses.dataclass(frozen=True)
class GraphFilter:
    """Sub-graph with d12 < cutoff^2 of a parent graph.  FPID: GRAPH_FILTER"""

    cutoff: float
    parent_graph: str
    graph_key: str
    remove_hydrogens: int = False
    switch_params: dict = dataclasses.field(default_factory=dict, hash=False)
    mult_size: float = 1.05
    FPID: ClassVar[str] = "GRAPH_FILTER"

    def __post_init__(self):
        if self.remove_hydrogens:
            raise NotImplementedError("remove_hydrogens is not part of the ANI path")

    def init(self):
        return {"npairs": 1, "nblist_mult_size": self.mult_size}

    def get_graph_properties(self):
        return {self.graph_key: {"cutoff": self.cutoff, "directed": True, "parent_graph": self.parent_graph}}


@dataclasses.dataclass(frozen=True)
class GraphAngularExtension:
    """All neighbour pairs of each centre of a graph (angle triplets).  FPID: GRAPH_ANGULAR_EXTENSION"""

    mult_size: float = 1.05
    add_neigh: int = 5
    graph_key: str = "graph"
    FPID: ClassVar[str] = "GRAPH_ANGULAR_EXTENSION"

    def init(self):
        return {"nangles": 0, "nblist_mult_size": self.mult_size, "max_neigh": self.add_neigh,
                "add_neigh": self.add_neigh}

    def get_graph_properties(self):
        return {self.graph_key: {"has_angles": True}}


@dataclasses.dataclass(frozen=True)
class SpeciesIndexer:
    """Per-species index arrays padded with N.  FPID: SPECIES_INDEXER"""

    output_key: str = "species_index"
    species_order: str | None = None
    add_atoms: int = 0
    add_atoms_margin: int = 10
    FPID: ClassVar[str] = "SPECIES_INDEXER"

    def init(self):
        return {"sizes": {}}

    def __call__(self, state, inputs, return_state_update=False, add_margin=False):
        """Host path (preprocessing.py:1345-1408)."""
        species = np.asarray(inputs["species"], np.int32)
        nat = species.shape[0]
        set_species, counts = np.unique(species, return_counts=True)
        sizes = dict(state.get("sizes", {}))
        new_state = {**state}
        state_up = {}
        up = False
        for s, c in zip(set_species, counts):
            if s <= 0:
                continue
            if c > sizes.get(int(s), 0):
                add = state.get("add_atoms", self.add_atoms)
                if add_margin:
                    add += state.get("add_atoms_margin", self.add_atoms_margin)
                sizes[int(s)] = int(c) + add
                up = True
        if up:
            state_up["sizes"] = (sizes, state.get("sizes", {}))
            new_state["sizes"] = sizes
        if self.species_order is not None:
            order = [el.strip() for el in self.species_order.split(",")]
            max_size = max(sizes.values()) if sizes else 0
            idx = np.full((len(order), max_size), nat, np.int32)
            for i, el in enumerate(order):
                z = SYMBOL_TO_Z[el]
                w = np.nonzero(species == z)[0]
                idx[i, : w.size] = w
            species_index = idx
        else:
            species_index = {}
            for z, c in sizes.items():
                w = np.nonzero(species == z)[0]
                a = np.full(c, nat, np.int32)
                a[: w.size] = w
                species_index[PERIODIC_TABLE[z]] = a
        out = {**inputs, self.output_key: species_index, self.output_key + "_overflow": False}
        if return_state_update:
            return new_state, out, state_up
        return new_state, out


PREPROCESSING = {c.FPID: c for c in (GraphGenerator, GraphFilter, GraphAngularExtension, SpeciesIndexer)}


def register_fennix_preprocessing(module, FPID: str | None = None):
    """Register a preprocessing class; clashes with a different class raise (modules.py:46-65)."""
    names = [FPID.upper()] if FPID is not None else [module.FPID.upper()]
    for name in names:
        if name in PREPROCESSING and PREPROCESSING[name] is not module:
            raise ValueError(f"A different module identified as '{name}' is already registered !")
        PREPROCESSING[name] = module


class PreprocessingChain:
    """Ordered preprocessing layers (preprocessing.py:1824-1925) driving the HIP neighbour lists.

    Supported layer set (the ANI path): one GraphGenerator, an optional
    GraphFilter of it (the angular graph), an optional GraphAngularExtension
    on the filtered graph, an optional SpeciesIndexer.
    """

    def __init__(self, layers, potential=None):
        self.layers = tuple(layers)
        gens = [l for l in self.layers if isinstance(l, GraphGenerator)]
        if len(gens) != 1:
            raise NotImplementedError("exactly one GraphGenerator is supported")
        self.generator = gens[0]
        filters = [l for l in self.layers if isinstance(l, GraphFilter)]
        if len(filters) > 1 or (filters and filters[0].parent_graph != self.generator.graph_key):
            raise NotImplementedError("one GraphFilter of the radial graph is supported")
        self.filter = filters[0] if filters else None
        exts = [l for l in self.layers if isinstance(l, GraphAngularExtension)]
        self.angular = exts[0] if exts else None
        if self.angular is not None and (self.filter is None or self.angular.graph_key != self.filter.graph_key):
            raise NotImplementedError("the angular extension must apply to the filtered graph")
        idx = [l for l in self.layers if isinstance(l, SpeciesIndexer)]
        self.indexer = idx[0] if idx else None
        self.potential = potential

    def get_graphs_properties(self):
        props = {}
        for l in self.layers:
            if hasattr(l, "get_graph_properties"):
                for k, v in l.get_graph_properties().items():
                    props[k] = {**props.get(k, {}), **v}
        return props

    def init(self):
        return {i: l.init() for i, l in enumerate(self.layers)}

    @property
    def cutoff_angular(self) -> float:
        return self.filter.cutoff if self.filter is not None else self.generator.cutoff

    # -------------------------------------------------------------- device
    def process(self, state, inputs, sb):
        """Device path: rebuild the rows of ``sb`` at inputs['coordinates'] (no host sync)."""
        x32 = inputs["coordinates"]
        self.potential.build_nblist(sb, x32)
        graph = {"nbr": sb.nbr, "n_nbr": sb.n_nbr, "max_nbr": sb.max_nbr, "overflow": sb.status[0:1],
                 "cutoff": self.generator.cutoff}
        out = {**inputs, self.generator.graph_key: graph}
        if self.filter is not None:
            out[self.filter.graph_key] = {"nbr_ang": sb.nbr_ang, "n_nbr_ang": sb.n_nbr_ang,
                                          "max_nbr_ang": sb.max_nbr_ang, "parent_graph": self.generator.graph_key,
                                          "overflow": sb.status[0:1], "cutoff": self.filter.cutoff}
        return out

    def check_reallocate(self, state, inputs, sb):
        """Read the overflow flag; on overflow grow the rows and rebuild (preprocessing.py:1859-1878)."""
        grown = self.potential.check_reallocate(sb, inputs["coordinates"])
        return state, grown

    # ---------------------------------------------------------------- host
    def export(self, state, inputs, sb, x32) -> dict:
        """The reference's host-path arrays (NumPy), from the device rows."""
        lib = self.potential.lib
        n_nbr = np.minimum(sb.n_nbr.cpu().numpy(), sb.max_nbr)
        rs = np.concatenate([[0], np.cumsum(n_nbr)]).astype(np.int32)
        E = int(rs[-1])
        dev = self.potential.device
        es = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
        ed = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
        d12 = torch.empty(max(E, 1), dtype=torch.float32, device=dev)
        sh = torch.empty((max(E, 1), 3), dtype=torch.float32, device=dev)
        rst = torch.tensor(rs[:-1], device=dev)
        P = _abi.ptr
        _abi.check(lib.fnl_nblist_export(P(x32), P(sb.batch_index), sb.n_atoms, sb.pbc, P(sb.cells), P(sb.rcells),
                                         P(sb.nbr), P(sb.nbr_shift), P(sb.n_nbr), sb.max_nbr, P(rst), P(es), P(ed),
                                         P(d12), P(sh),
                                         _abi.stream_handle()), "fnl_nblist_export")
        es, ed, d12, sh = (t.cpu().numpy()[:E] for t in (es, ed, d12, sh))
        nat = sb.n_atoms
        rc2 = np.float32(self.generator.cutoff ** 2)
        # half list (src < dst) in the reference's row-major pair order, then symmetrised
        m = es < ed
        order = np.lexsort((ed[m], es[m]))
        hs, hd, h12, hsh = es[m][order], ed[m][order], d12[m][order], sh[m][order]
        gstate = state.setdefault("graph", {"npairs": 1})
        npairs = hs.shape[0]
        if npairs > gstate["npairs"]:
            gstate["npairs"] = _grow(npairs, gstate["npairs"], self.generator.mult_size)
        P_ = gstate["npairs"]
        src = np.full(P_, nat, np.int32)
        dst = np.full(P_, nat, np.int32)
        dd = np.full(P_, rc2, np.float32)
        shp = np.zeros((P_, 3), np.float32)
        src[:npairs], dst[:npairs], dd[:npairs], shp[:npairs] = hs, hd, h12, hsh
        graph = {"edge_src": np.concatenate([src, dst]), "edge_dst": np.concatenate([dst, src]),
                 "d12": np.concatenate([dd, dd]), "overflow": False,
                 "pbc_shifts": np.concatenate([shp, -shp]) if sb.pbc else None}
        out = {**inputs, self.generator.graph_key: graph}
        if self.filter is not None:
            rca2 = np.float32(self.filter.cutoff ** 2)
            keep = np.nonzero(graph["d12"] < rca2)[0].astype(np.int32)
            fstate = state.setdefault("filter", {"npairs": 1})
            nf = keep.shape[0]
            if nf > fstate["npairs"]:
                fstate["npairs"] = _grow(nf, fstate["npairs"], self.filter.mult_size)
            Pf = fstate["npairs"]
            fi = np.full(Pf, graph["edge_src"].shape[0], np.int32)
            fs = np.full(Pf, nat, np.int32)
            fd = np.full(Pf, nat, np.int32)
            f12 = np.full(Pf, rca2, np.float32)
            fi[:nf], fs[:nf], fd[:nf], f12[:nf] = keep, graph["edge_src"][keep], graph["edge_dst"][keep], \
                graph["d12"][keep]
            g2 = {"edge_src": fs, "edge_dst": fd, "filter_indices": fi, "d12": f12, "overflow": False}
            if self.angular is not None:
                g2.update(self._triplets(state, fs, nat))
            out[self.filter.graph_key] = g2
        return out

    def _triplets(self, state, edge_src, nat):
        """angle_src / angle_dst / central_atom of GraphAngularExtension.__call__ (:1085-1181)."""
        nedge = edge_src.shape[0]
        valid = edge_src < nat
        order = np.argsort(np.where(valid, edge_src, nat), kind="stable")
        counts = np.bincount(edge_src[valid], minlength=nat)
        starts = np.concatenate([[0], np.cumsum(counts)])
        a_s, a_d = [], []
        for i in range(nat):
            e = order[starts[i]:starts[i + 1]]
            if e.size < 2:
                continue
            l1, l2 = np.triu_indices(e.size, 1)
            a_s.append(e[l1])
            a_d.append(e[l2])
        a_s = np.concatenate(a_s).astype(np.int32) if a_s else np.zeros(0, np.int32)
        a_d = np.concatenate(a_d).astype(np.int32) if a_d else np.zeros(0, np.int32)
        astate = state.setdefault("angles", {"nangles": 0, "max_neigh": self.angular.add_neigh})
        n = a_s.shape[0]
        if n > astate["nangles"]:
            astate["nangles"] = _grow(n, astate["nangles"], self.angular.mult_size)
        mc = int(counts.max()) if counts.size else 0
        if mc > astate["max_neigh"]:
            astate["max_neigh"] = mc + self.angular.add_neigh
        T = astate["nangles"]
        s_ = np.full(T, nedge, np.int32)
        d_ = np.full(T, nedge, np.int32)
        c_ = np.full(T, nat, np.int32)
        s_[:n], d_[:n] = a_s, a_d
        c_[:n] = edge_src[a_s]
        return {"angle_src": s_, "angle_dst": d_, "central_atom": c_, "angle_overflow": False,
                "max_neigh": astate["max_neigh"]}
