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


class _LayerBuffers:
    """Device arrays of one input layout (species, natoms, cells) for the per-layer device path:
    batch tables, the cell-list workspace and rows of a given capacity, reused while the layout
    and capacities stay the same (the role of the reference's static shapes)."""

    _cache: dict = {}

    def __init__(self, inputs: dict, cutoff: float, minimum_image: bool):
        from .potential import grid_caps

        dev = torch.device("cuda")
        self.lib = _abi.load()
        species = np.asarray(inputs["species"], np.int32)
        natoms = np.asarray(inputs["natoms"], np.int32)
        N, B = species.shape[0], natoms.shape[0]
        offs = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int32)
        batch_index = np.full(N, max(B - 1, 0), np.int32)
        batch_index[: offs[-1]] = np.repeat(np.arange(B, dtype=np.int32), natoms)
        # the neighbour-list kernels only tell real atoms (index >= 0) from padding
        sidx = np.where(species > 0, 0, -1).astype(np.int32)
        sidx[offs[-1]:] = -1
        self.n_atoms, self.n_sys = N, B
        self.sys_offsets = torch.tensor(offs, device=dev)
        self.batch_index = torch.tensor(batch_index, device=dev)
        self.species_idx = torch.tensor(sidx, device=dev)
        self.pbc = _abi.FNL_PBC_NONE
        self.cells = self.rcells = None
        cells = inputs.get("cells")
        if cells is not None:
            cells = np.asarray(cells, np.float64).reshape(-1, 3, 3)
            if cells.shape[0] == 1 and B > 1:
                cells = np.repeat(cells, B, axis=0)
            self.pbc = _abi.FNL_PBC_MINIMUM_IMAGE if minimum_image else _abi.FNL_PBC_GENERAL
            self.cells = torch.tensor(cells.astype(np.float32), device=dev)
            self.rcells = torch.tensor(np.linalg.inv(cells).astype(np.float32), device=dev)
        g, self.n_cells = grid_caps(cutoff, natoms, cells)
        self.grid = torch.tensor(g, device=dev)
        wsb = self.lib.fnl_nblist_workspace(N, max(B, 1), self.n_cells)
        self.workspace = torch.empty(max(int(wsb), 256), dtype=torch.uint8, device=dev)
        self.rows: dict = {}

    @classmethod
    def get(cls, inputs: dict, cutoff: float) -> "_LayerBuffers":
        minimum_image = "minimum_image" in inputs.get("flags", {}) or inputs.get("cells") is None
        key = (np.asarray(inputs["species"], np.int32).tobytes(), np.asarray(inputs["natoms"], np.int32).tobytes(),
               None if inputs.get("cells") is None else np.asarray(inputs["cells"], np.float64).tobytes(),
               minimum_image, float(cutoff))
        lb = cls._cache.get(key)
        if lb is None:
            if len(cls._cache) > 16:
                cls._cache.clear()
            lb = cls(inputs, cutoff, minimum_image)
            cls._cache[key] = lb
        return lb

    def row_buffers(self, name: str, width: int, shifts: bool = False) -> dict:
        """Rows [N, width] (+ counts, status, shifts for general PBC), kept per name and width."""
        key = (name, width)
        if key not in self.rows:
            dev = self.species_idx.device
            n = max(self.n_atoms, 1)
            self.rows[key] = {
                "nbr": torch.zeros((n, width), dtype=torch.int32, device=dev),
                "n": torch.zeros(n, dtype=torch.int32, device=dev),
                "shift": (torch.zeros((n, width), dtype=torch.int32, device=dev)
                          if shifts and self.pbc == _abi.FNL_PBC_GENERAL else None),
                "status": torch.zeros(_abi.FNL_STATUS_WORDS, dtype=torch.int32, device=dev)}
        return self.rows[key]


def _coords32(inputs) -> torch.Tensor:
    x = inputs["coordinates"]
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=torch.float32).contiguous()
    return torch.as_tensor(np.asarray(x, np.float32), device="cuda").contiguous()


def _overflow_of(graph) -> bool:
    """Host value of a graph's overflow flag (reads the device status word: synchronises)."""
    ov = graph.get("overflow", False)
    if isinstance(ov, torch.Tensor):
        return bool(int(ov.reshape(-1)[0].item()) != 0)
    return bool(ov)


def _grown(n: int, prev: int, mult: float, add_margin: bool) -> int:
    """Capacity after an overflow: int(mult * max(n, prev)) + 1 (preprocessing.py:269-282), with
    one more growth step when add_margin."""
    c = _grow(n, prev, mult)
    return _grow(c, c, mult) if add_margin else c


@dataclasses.dataclass(frozen=True)
class GraphGenerator:
    """Radial graph within ``cutoff`` (cell list on the GPU).  FPID: GRAPH

    ``process`` (device path, models/preprocessing.py:369-638) builds the rows on the GPU with the
    capacity held in the state (``max_nbr``); with ``nblist_skin`` in the state it builds the
    list with cutoff + skin and re-filters it (``update_skin``, :640-696).  ``__call__`` (host
    path, :74-355) returns the reference's flat arrays.  ``check_reallocate`` (:357-367) grows
    the capacity and redoes the build when the overflow flag is set."""

    cutoff: float
    graph_key: str = "graph"
    switch_params: dict = dataclasses.field(default_factory=dict, hash=False)
    mult_size: float = 1.05
    FPID: ClassVar[str] = "GRAPH"

    def init(self):
        return {"max_nat": 1, "npairs": 1, "nblist_mult_size": self.mult_size}

    def get_graph_properties(self):
        return {self.graph_key: {"cutoff": self.cutoff, "directed": True}}

    def get_processor(self):
        """The differentiable part of the graph (GraphProcessor, :699-777: edge vectors, safe
        distances, cosine switch) runs inside the AEV kernels; this names it and its settings."""
        return FusedProcessor, {"kind": "GraphProcessor", "cutoff": self.cutoff, "graph_key": self.graph_key,
                                "switch_params": self.switch_params, "name": f"{self.graph_key}_Processor"}

    # ------------------------------------------------------------ device
    def process(self, state, inputs):
        """Rows of the radial graph at inputs['coordinates'] on the GPU (no host sync)."""
        skin = float(state.get("nblist_skin", 0.0) or 0.0)
        lb = _LayerBuffers.get(inputs, self.cutoff + max(skin, 0.0))
        x = _coords32(inputs)
        cap = int(state.get("max_nbr", 64))
        P = _abi.ptr
        s = _abi.stream_handle()
        graph = {"cutoff": self.cutoff, "pbc": lb.pbc, "_layout": lb, "coordinates": x}
        if skin > 0:
            cap_s = int(state.get("max_nbr_skin", _next8(int(cap * ((self.cutoff + skin) / self.cutoff) ** 3))))
            sk = lb.row_buffers(f"{self.graph_key}/skin", cap_s, shifts=True)
            _abi.check(lb.lib.fnl_nblist_build(
                P(x), P(lb.species_idx), P(lb.batch_index), P(lb.sys_offsets), lb.n_atoms, lb.n_sys, lb.pbc,
                P(lb.cells), P(lb.rcells), P(lb.grid), lb.n_cells, self.cutoff + skin, self.cutoff, cap_s, 1,
                P(lb.workspace), lb.workspace.numel(), P(sk["nbr"]), P(sk["shift"]), P(sk["n"]), None, None,
                P(sk["status"]), s), "fnl_nblist_build")
            graph.update(nbr_skin=sk["nbr"], n_skin=sk["n"], skin_shift=sk["shift"], max_skin=cap_s,
                         status_skin=sk["status"], skin=skin)
            out = {**inputs, self.graph_key: graph}
            return self.update_skin(out, state)
        rows = lb.row_buffers(self.graph_key, cap, shifts=True)
        _abi.check(lb.lib.fnl_nblist_build(
            P(x), P(lb.species_idx), P(lb.batch_index), P(lb.sys_offsets), lb.n_atoms, lb.n_sys, lb.pbc,
            P(lb.cells), P(lb.rcells), P(lb.grid), lb.n_cells, self.cutoff, self.cutoff, cap, 1,
            P(lb.workspace), lb.workspace.numel(), P(rows["nbr"]), P(rows["shift"]), P(rows["n"]), None, None,
            P(rows["status"]), s), "fnl_nblist_build")
        graph.update(nbr=rows["nbr"], n_nbr=rows["n"], nbr_shift=rows["shift"], max_nbr=cap,
                     status=rows["status"], overflow=rows["status"][0:1])
        return {**inputs, self.graph_key: graph}

    def update_skin(self, inputs, state=None):
        """Rows at the current coordinates re-filtered from the skin list of the last build
        (preprocessing.py:640-696); rows keep the skin list's order."""
        graph = dict(inputs[self.graph_key])
        if "nbr_skin" not in graph:
            raise ValueError("this graph was built without nblist_skin")
        lb = graph["_layout"]
        x = _coords32(inputs)
        cap = int((state or {}).get("max_nbr", graph.get("max_nbr", 64)))
        rows = lb.row_buffers(self.graph_key, cap, shifts=True)
        tmp_a = lb.row_buffers(self.graph_key + "/unused_angular", 1)
        P = _abi.ptr
        _abi.check(lb.lib.fnl_nblist_skin_update(
            P(x), P(lb.batch_index), lb.n_atoms, lb.pbc, P(lb.cells), P(lb.rcells), P(graph["nbr_skin"]),
            P(graph["skin_shift"]), P(graph["n_skin"]), graph["max_skin"], self.cutoff, 0.0, cap, 1,
            P(rows["nbr"]), P(rows["shift"]), P(rows["n"]), P(tmp_a["nbr"]), P(tmp_a["n"]), P(rows["status"]),
            _abi.stream_handle()), "fnl_nblist_skin_update")
        graph.update(nbr=rows["nbr"], n_nbr=rows["n"], nbr_shift=rows["shift"], max_nbr=cap, status=rows["status"],
                     overflow=rows["status"][0:1], coordinates=x)
        return {**inputs, self.graph_key: graph}

    def check_reallocate(self, state, inputs, parent_overflow=False):
        """(state, state_up, inputs, overflow): on overflow, capacities grow from the status maxima
        and the rows are rebuilt."""
        graph = inputs[self.graph_key]
        overflow = parent_overflow or _overflow_of(graph)
        if not overflow:
            return state, {}, inputs, False
        new = dict(state)
        up = {}
        mult = state.get("nblist_mult_size", self.mult_size)
        st = graph["status"].cpu().numpy()
        if st[_abi.FNL_STATUS_OVERFLOW]:
            prev = int(state.get("max_nbr", graph["max_nbr"]))
            new["max_nbr"] = _next8(_grown(int(st[_abi.FNL_STATUS_MAX_NBR]), prev, mult, True))
            up["max_nbr"] = (new["max_nbr"], prev)
        if "status_skin" in graph:
            sts = graph["status_skin"].cpu().numpy()
            if sts[_abi.FNL_STATUS_OVERFLOW]:
                prev = int(state.get("max_nbr_skin", graph["max_skin"]))
                new["max_nbr_skin"] = _next8(_grown(int(sts[_abi.FNL_STATUS_MAX_NBR]), prev, mult, True))
                up["max_nbr_skin"] = (new["max_nbr_skin"], prev)
        out = self.process(new, inputs)
        if _overflow_of(out[self.graph_key]):
            raise RuntimeError("neighbour list still overflows after reallocation")
        return new, up, out, True

    # -------------------------------------------------------------- host
    def __call__(self, state, inputs, return_state_update=False, add_margin=False):
        """Host path: the flat, symmetrised, capacity-padded arrays of the reference
        (edge_src, edge_dst, d12, pbc_shifts; preprocessing.py:259-341), built from the device
        rows; ``npairs`` grows to int(mult * max(npairs, prev)) + 1."""
        inputs = check_input(inputs)
        st = {**state}
        dev_in = self.process(st, inputs)
        st, _, dev_in, _ = self.check_reallocate(st, dev_in)
        g = dev_in[self.graph_key]
        lb = g["_layout"]
        es, ed, d12, sh = _export_rows(lb, g["coordinates"], g["nbr"], g["nbr_shift"], g["n_nbr"], g["max_nbr"])
        nat = lb.n_atoms
        m = es < ed
        order = np.lexsort((sh[m][:, 2], sh[m][:, 1], sh[m][:, 0], ed[m], es[m]))
        hs, hd, h12, hsh = es[m][order], ed[m][order], d12[m][order], sh[m][order]
        npairs = hs.shape[0]
        mult = st.get("nblist_mult_size", self.mult_size)
        prev = int(st.get("npairs", 1))
        state_up = {}
        if npairs > prev or add_margin:
            st["npairs"] = _grown(npairs, prev, mult, add_margin)
            state_up["npairs"] = (st["npairs"], prev)
        P_ = int(st["npairs"])
        rc2 = np.float32(self.cutoff ** 2)
        src = np.full(P_, nat, np.int32)
        dst = np.full(P_, nat, np.int32)
        dd = np.full(P_, rc2, np.float32)
        shp = np.zeros((P_, 3), np.float32)
        src[:npairs], dst[:npairs], dd[:npairs], shp[:npairs] = hs, hd, h12, hsh
        graph = {"edge_src": np.concatenate([src, dst]), "edge_dst": np.concatenate([dst, src]),
                 "d12": np.concatenate([dd, dd]), "overflow": False}
        if "cells" in inputs:
            graph["pbc_shifts"] = np.concatenate([shp, -shp])
        out = {**inputs, self.graph_key: graph}
        if return_state_update:
            return st, out, state_up
        return st, out


def _next8(n: int) -> int:
    return ((int(n) + 7) // 8) * 8


def _export_rows(lb, x, nbr, nbr_shift, n_nbr, max_nbr):
    """Flat directed edges of device rows (fnl_nblist_export) as NumPy arrays."""
    counts = np.minimum(n_nbr.cpu().numpy(), max_nbr)
    rs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    E = int(rs[-1])
    dev = x.device
    es = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
    ed = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
    d12 = torch.empty(max(E, 1), dtype=torch.float32, device=dev)
    sh = torch.zeros((max(E, 1), 3), dtype=torch.float32, device=dev)
    rst = torch.tensor(rs[:-1], device=dev)
    P = _abi.ptr
    _abi.check(lb.lib.fnl_nblist_export(P(x), P(lb.batch_index), lb.n_atoms, lb.pbc, P(lb.cells), P(lb.rcells),
                                        P(nbr), P(nbr_shift), P(n_nbr), max_nbr, P(rst), P(es), P(ed), P(d12),
                                        P(sh), _abi.stream_handle()), "fnl_nblist_export")
    return tuple(t.cpu().numpy()[:E] for t in (es, ed, d12, sh))


@dataclasses.dataclass(frozen=True)
class GraphFilter:
    """Sub-graph with d12 < cutoff^2 of a parent graph.  FPID: GRAPH_FILTER

    Device path (:909-956): the angular rows are the slots of the parent's radial rows with
    d12 < cutoff^2, recomputed with the build's float32 arithmetic (fnl_nblist_skin_update over
    the parent rows), so they hold radial slot numbers of the parent graph.  Host path
    (:834-907): ``filter_indices`` and the filtered edges from the parent's flat arrays."""

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

    def get_processor(self):
        """GraphFilterProcessor (:963-1043) runs inside the AEV kernels."""
        return FusedProcessor, {"kind": "GraphFilterProcessor", "cutoff": self.cutoff, "graph_key": self.graph_key,
                                "parent_graph": self.parent_graph, "switch_params": self.switch_params,
                                "name": f"{self.graph_key}_Filter_{self.parent_graph}"}

    def process(self, state, inputs):
        parent = inputs[self.parent_graph]
        lb = parent["_layout"]
        x = _coords32(inputs)
        cap = int(state.get("max_nbr_ang", 32))
        pcap = int(parent["max_nbr"])
        radial = lb.row_buffers(self.graph_key + "/radial", pcap, shifts=True)
        rows = lb.row_buffers(self.graph_key, cap)
        P = _abi.ptr
        # the parent's rows re-filtered at the parent's own cutoff (every slot kept, same order)
        # and at this cutoff (the angular rows); d12 with the build's arithmetic
        _abi.check(lb.lib.fnl_nblist_skin_update(
            P(x), P(lb.batch_index), lb.n_atoms, lb.pbc, P(lb.cells), P(lb.rcells), P(parent["nbr"]),
            P(parent["nbr_shift"]), P(parent["n_nbr"]), pcap, parent["cutoff"], self.cutoff, pcap, cap,
            P(radial["nbr"]), P(radial["shift"]), P(radial["n"]), P(rows["nbr"]), P(rows["n"]), P(rows["status"]),
            _abi.stream_handle()), "fnl_nblist_skin_update")
        graph = {"nbr_ang": rows["nbr"], "n_nbr_ang": rows["n"], "max_nbr_ang": cap, "status": rows["status"],
                 "overflow": rows["status"][0:1], "parent_graph": self.parent_graph, "cutoff": self.cutoff,
                 "_layout": lb}
        return {**inputs, self.graph_key: graph}

    def update_skin(self, inputs, state=None):
        """The angular rows of the parent's re-filtered rows (GraphFilter.update_skin :958-960)."""
        return self.process(state or {"max_nbr_ang": inputs[self.graph_key]["max_nbr_ang"]}, inputs)

    def check_reallocate(self, state, inputs, parent_overflow=False):
        """After a parent overflow the rows are first re-derived from the regrown parent (the
        counts seen on truncated parent rows cannot size this graph); then this graph's own
        overflow grows its capacity from the status maxima."""
        out = self.process(state, inputs) if parent_overflow else inputs
        new = dict(state)
        up = {}
        for _ in range(4):
            graph = out[self.graph_key]
            if not _overflow_of(graph):
                return new, up, out, bool(parent_overflow or up)
            st = graph["status"].cpu().numpy()
            prev = int(new.get("max_nbr_ang", graph["max_nbr_ang"]))
            mult = new.get("nblist_mult_size", self.mult_size)
            new["max_nbr_ang"] = _next8(_grown(int(st[_abi.FNL_STATUS_MAX_NBR_ANG]), prev, mult, True))
            up["max_nbr_ang"] = (new["max_nbr_ang"], up.get("max_nbr_ang", (None, prev))[1])
            out = self.process(new, inputs)
        raise RuntimeError("filtered graph still overflows after reallocation")

    def __call__(self, state, inputs, return_state_update=False, add_margin=False):
        """Host path from the parent's flat arrays (preprocessing.py:834-907)."""
        parent = inputs[self.parent_graph]
        d12 = np.asarray(parent["d12"], np.float32)
        nat = np.asarray(inputs["species"]).shape[0]
        rca2 = np.float32(self.cutoff ** 2)
        keep = np.nonzero(d12 < rca2)[0].astype(np.int32)
        st = {**state}
        nf = keep.shape[0]
        prev = int(st.get("npairs", 1))
        state_up = {}
        if nf > prev or add_margin:
            st["npairs"] = _grown(nf, prev, st.get("nblist_mult_size", self.mult_size), add_margin)
            state_up["npairs"] = (st["npairs"], prev)
        Pf = int(st["npairs"])
        fi = np.full(Pf, d12.shape[0], np.int32)
        fs = np.full(Pf, nat, np.int32)
        fd = np.full(Pf, nat, np.int32)
        f12 = np.full(Pf, rca2, np.float32)
        fi[:nf], fs[:nf], fd[:nf], f12[:nf] = keep, parent["edge_src"][keep], parent["edge_dst"][keep], d12[keep]
        g = {"edge_src": fs, "edge_dst": fd, "filter_indices": fi, "d12": f12, "overflow": False}
        if "pbc_shifts" in parent:
            sh = np.zeros((Pf, 3), np.float32)
            sh[:nf] = np.asarray(parent["pbc_shifts"])[keep]
            g["pbc_shifts"] = sh
        out = {**inputs, self.graph_key: g}
        if return_state_update:
            return st, out, state_up
        return st, out


@dataclasses.dataclass(frozen=True)
class GraphAngularExtension:
    """All neighbour pairs of each centre of a graph (angle triplets).  FPID: GRAPH_ANGULAR_EXTENSION

    Device path: the triplets of centre i are all slot pairs a < b of its angular row
    (:1237-1245), enumerated inside the AEV kernels, so ``process`` records the triplet layout
    and passes the angular row capacity check through (``max_neigh`` is the row capacity).
    Host path (:1085-1181): ``angle_src``, ``angle_dst``, ``central_atom`` padded to
    ``nangles``."""

    mult_size: float = 1.05
    add_neigh: int = 5
    graph_key: str = "graph"
    FPID: ClassVar[str] = "GRAPH_ANGULAR_EXTENSION"

    def init(self):
        return {"nangles": 0, "nblist_mult_size": self.mult_size, "max_neigh": self.add_neigh,
                "add_neigh": self.add_neigh}

    def get_graph_properties(self):
        return {self.graph_key: {"has_angles": True}}

    def get_processor(self):
        """GraphAngleProcessor (:1283-1317) runs inside the AEV kernels (cos(theta - theta_z)
        without an arccos)."""
        return FusedProcessor, {"kind": "GraphAngleProcessor", "graph_key": self.graph_key,
                                "name": f"{self.graph_key}_AngleProcessor"}

    def process(self, state, inputs):
        graph = {**inputs[self.graph_key], "angles": "slot pairs a < b of each angular row",
                 "max_neigh": inputs[self.graph_key].get("max_nbr_ang")}
        graph["angle_overflow"] = graph.get("overflow", False)
        return {**inputs, self.graph_key: graph}

    def update_skin(self, inputs, state=None):
        return self.process(state or {}, inputs)

    def check_reallocate(self, state, inputs, parent_overflow=False):
        """The angular row capacity belongs to the filter that fills the rows: an overflow there is
        regrown by the filter's check (the parent_overflow passed in); this layer redoes its
        annotation."""
        if not parent_overflow:
            return state, {}, inputs, False
        return state, {}, self.process(state, inputs), True

    def __call__(self, state, inputs, return_state_update=False, add_margin=False):
        g = inputs[self.graph_key]
        nat = np.asarray(inputs["species"]).shape[0]
        st = {**state}
        trip = _triplets_host(st, np.asarray(g["edge_src"], np.int32), nat, self.mult_size, self.add_neigh,
                              add_margin)
        state_up = trip.pop("_state_up")
        out = {**inputs, self.graph_key: {**g, **trip}}
        if return_state_update:
            return st, out, state_up
        return st, out


def _triplets_host(astate, edge_src, nat, mult_size, add_neigh, add_margin=False):
    """angle_src / angle_dst / central_atom of GraphAngularExtension.__call__ (:1085-1181),
    padded to ``nangles`` (grown by mult_size) with ``max_neigh`` tracked in ``astate``."""
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
    state_up = {}
    n = a_s.shape[0]
    prev = int(astate.get("nangles", 0))
    if n > prev or add_margin:
        astate["nangles"] = _grown(n, prev, mult_size, add_margin)
        state_up["nangles"] = (astate["nangles"], prev)
    mc = int(counts.max()) if counts.size else 0
    prev_mn = int(astate.get("max_neigh", add_neigh))
    if mc > prev_mn or add_margin:
        astate["max_neigh"] = max(mc, prev_mn) + int(astate.get("add_neigh", add_neigh))
        state_up["max_neigh"] = (astate["max_neigh"], prev_mn)
    T = int(astate["nangles"])
    s_ = np.full(T, nedge, np.int32)
    d_ = np.full(T, nedge, np.int32)
    c_ = np.full(T, nat, np.int32)
    s_[:n], d_[:n] = a_s, a_d
    c_[:n] = edge_src[a_s]
    return {"angle_src": s_, "angle_dst": d_, "central_atom": c_, "angle_overflow": False,
            "max_neigh": astate["max_neigh"], "_state_up": state_up}


@dataclasses.dataclass(frozen=True)
class FusedProcessor:
    """What ``get_processor`` returns for the graph layers of this path: the differentiable graph
    processors (edge vectors, distances, switches, angles) are computed inside the HIP AEV
    kernels, so the model holds no separate processor module.  ``kind`` names the FeNNol
    processor it stands for; the keyword arguments are the ones that processor takes."""

    kind: str
    graph_key: str
    name: str
    cutoff: float | None = None
    parent_graph: str | None = None
    switch_params: dict = dataclasses.field(default_factory=dict, hash=False)


@dataclasses.dataclass(frozen=True)
class AtomPadding:
    """Pad atomic arrays to a fixed size (preprocessing.py:1654-1729): atoms to
    int(mult_size * N) + 1 when N grows (species -1, zero arrays, batch index of the padding
    system), systems to the largest count seen + add_sys plus one padding system (natoms 0,
    cell 1000 I), and ``true_atoms`` / ``true_sys`` masks, so that the device buffers keep their
    shapes across calls of varying size."""

    mult_size: float = 1.2
    add_sys: int = 0

    def init(self):
        return {"prev_nat": 0, "prev_nsys": 0}

    def __call__(self, state, inputs):
        species = np.asarray(inputs["species"])
        nat = species.shape[0]
        nsys = len(inputs["natoms"])
        prev_nat = int(state.get("prev_nat", 0))
        if nat > prev_nat:
            prev_nat = int(self.mult_size * nat) + 1
        prev_nsys = int(state.get("prev_nsys", 0))
        if nsys > prev_nsys:
            prev_nsys = nsys + self.add_sys
        add_atoms = prev_nat - nat
        add_sys = prev_nsys - nsys + 1
        out = {**inputs}
        if add_atoms > 0:
            for k, v in inputs.items():
                if not isinstance(v, np.ndarray) or v.ndim == 0:
                    continue
                if v.shape[0] == nat:
                    out[k] = np.concatenate([v, np.zeros((add_atoms, *v.shape[1:]), v.dtype)])
                elif v.shape[0] == nsys:
                    pad = (1000.0 * np.eye(3, dtype=v.dtype)[None].repeat(add_sys, 0) if k == "cells"
                           else np.zeros((add_sys, *v.shape[1:]), v.dtype))
                    out[k] = np.concatenate([v, pad])
            out["natoms"] = np.concatenate([np.asarray(inputs["natoms"], np.int32), np.zeros(add_sys, np.int32)])
            out["species"] = np.concatenate([species, -np.ones(add_atoms, species.dtype)])
            out["batch_index"] = np.concatenate([np.asarray(inputs["batch_index"]),
                                                 np.full(add_atoms, out["natoms"].shape[0] - 1,
                                                         np.asarray(inputs["batch_index"]).dtype)])
            if "reciprocal_cells" in out and "cells" in out:
                out["reciprocal_cells"] = np.linalg.inv(out["cells"])
        out["true_atoms"] = np.asarray(out["species"]) > 0
        out["true_sys"] = np.arange(len(out["natoms"])) < nsys
        return {**state, "prev_nat": prev_nat, "prev_nsys": prev_nsys}, out


def atom_unpadding(inputs: dict) -> dict:
    """Remove the padding atoms and systems of AtomPadding (preprocessing.py:1732-1760)."""
    if "true_atoms" not in inputs:
        return inputs
    species = np.asarray(inputs["species"])
    true_atoms = np.asarray(inputs["true_atoms"])
    true_sys = np.asarray(inputs["true_sys"])
    natall = species.shape[0]
    if int(np.argmax(species <= 0)) == 0:
        return inputs
    nsysall = len(inputs["natoms"])
    out = {**inputs}
    for k, v in inputs.items():
        if isinstance(v, torch.Tensor):
            v = v.detach().cpu().numpy()
        if isinstance(v, np.ndarray):
            if v.ndim == 0:
                out[k] = v
            elif v.shape[0] == natall:
                out[k] = v[true_atoms]
            elif v.shape[0] == nsysall:
                out[k] = v[true_sys]
    del out["true_sys"]
    del out["true_atoms"]
    return out


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


    def process(self, state, inputs):
        """Device path (preprocessing.py:1410-1466): per-species index arrays of the sizes in the
        state, padded with N, as device tensors, and the overflow flag (a count above its size)."""
        species = torch.as_tensor(np.asarray(inputs["species"], np.int32), device="cuda")
        nat = species.shape[0]
        sizes = state.get("sizes", {})
        out = {}
        overflow = False
        for z, size in sizes.items():
            idx = torch.nonzero(species == int(z)).flatten().to(torch.int32)
            overflow |= idx.numel() > size
            a = torch.full((size,), nat, dtype=torch.int32, device="cuda")
            a[: min(size, idx.numel())] = idx[:size]
            out[PERIODIC_TABLE[int(z)]] = a
        known = set(int(z) for z in sizes)
        present = set(int(z) for z in np.unique(np.asarray(inputs["species"])) if z > 0)
        overflow |= bool(present - known)
        return {**inputs, self.output_key: out, self.output_key + "_overflow": overflow}

    def check_reallocate(self, state, inputs, parent_overflow=False):
        """Grow the sizes from the current counts on overflow (preprocessing.py:1468-1490)."""
        if not inputs.get(self.output_key + "_overflow", False):
            return state, {}, inputs, False
        new, out, up = self(state, inputs, return_state_update=True, add_margin=True)
        return new, up, self.process(new, inputs), True


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

    def __init__(self, layers, potential=None, use_atom_padding: bool = False):
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
        self.padder = AtomPadding() if use_atom_padding else None

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
        """angle_src / angle_dst / central_atom (GraphAngularExtension.__call__ :1085-1181)."""
        astate = state.setdefault("angles", {"nangles": 0, "max_neigh": self.angular.add_neigh})
        out = _triplets_host(astate, edge_src, nat, self.angular.mult_size, self.angular.add_neigh)
        out.pop("_state_up")
        return out

    # ------------------------------------------------ per-layer (FeNNol's chain protocol)
    def atom_padding(self, state, inputs):
        """AtomPadding of the inputs when the chain pads atoms (preprocessing.py:1840-1846)."""
        if self.padder is None:
            return state, inputs
        pst, out = self.padder(state.get("atom_padding", self.padder.init()), inputs)
        return {**state, "atom_padding": pst}, out

    def process_layers(self, state, inputs):
        """Every layer's device ``process`` in order (preprocessing.py:1848-1857)."""
        for i, layer in enumerate(self.layers):
            if hasattr(layer, "process"):
                inputs = layer.process(state[i], inputs)
        return inputs

    def check_reallocate_layers(self, state, inputs):
        """Every layer's ``check_reallocate`` in order, each told whether its parent overflowed
        (preprocessing.py:1859-1878); returns (state, state_up, inputs, overflow)."""
        state = dict(state)
        state_up = {}
        parent = False
        for i, layer in enumerate(self.layers):
            if hasattr(layer, "check_reallocate"):
                st, up, inputs, ov = layer.check_reallocate(state[i], inputs, parent)
                state[i] = st
                if up:
                    state_up[i] = up
                parent = parent or ov
        return state, state_up, inputs, parent

    def update_skin_layers(self, inputs):
        """Every graph layer's ``update_skin`` in order (preprocessing.py:1880-1887)."""
        for layer in self.layers:
            if hasattr(layer, "update_skin"):
                inputs = layer.update_skin(inputs)
        return inputs

    def call_layers(self, state, inputs):
        """The host path: every layer's ``__call__`` in order (preprocessing.py:1824-1838)."""
        inputs = check_input(inputs)
        state = dict(state)
        for i, layer in enumerate(self.layers):
            state[i], inputs = layer(state[i], inputs)
        return state, inputs

    def get_processors(self):
        return [l.get_processor() for l in self.layers if hasattr(l, "get_processor")]
