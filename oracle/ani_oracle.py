"""NumPy restatement of FeNNol's ANI energy/force/MD path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product package.

Reference files are cited relative to the FeNNol source tree ``src/fennol``.

What pins each part (checked in ``tests/test_oracle.py``):

* ``half_pairs`` (preprocessing.py:74-355): brute force over explicit periodic
  images (min-image result == nearest of the 27 images when rc < L/2),
  shell counts 6/12/8/6/24 of a perfect simple-cubic lattice, symmetry of the
  directed list.
* ``skin_update`` (preprocessing.py:640-696): at the build positions it
  equals ``half_pairs`` at the cutoff (pairs, d12 bits, shifts); after moves
  below skin/2 per atom it equals ``half_pairs`` at the new positions (the
  Verlet-list guarantee); after larger moves it is a subset of it.
* ``angular_triplets`` (preprocessing.py:1085-1181): T = sum_i C(n_i, 2) and a
  brute-force enumeration of neighbour pairs.
* ``cosine_switch`` (misc.py:674-678): fc(0)=1, fc(rc)=0, fc'(rc)=0.
* ``ani_aev`` (ani.py:47-142): dimer closed form (radial), isolated trimer
  closed form (angular), rotation / translation / permutation invariance.
* ``species_mlp`` (nets.py:13-102, 301-426): explicit per-layer product and
  the CELU closed form.
* ``energy_forces_virial``: central finite differences of the energy (FP64),
  sum of forces = 0, virial == finite-difference strain derivative, symmetric.
* ``baoab_step`` (integrate.py:286-294, thermostats.py:108-120): gamma=0 is
  velocity Verlet (free particle exact, harmonic oscillator bounded and
  time-reversible); injected noise makes the O-step deterministic.
* ``philox4x32`` (Salmon et al., SC'11, Random123 known-answer vectors).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "PERIODIC_TABLE",
    "ATOMIC_MASSES",
    "MDUnits",
    "md_units",
    "half_pairs",
    "skin_update",
    "symmetrize",
    "filter_graph",
    "angular_triplets",
    "cosine_switch",
    "AEVParams",
    "aev_length",
    "Graphs",
    "build_graphs",
    "edge_vectors",
    "safe_distance",
    "ani_aev",
    "ani_aev_backward",
    "langevin_coefficients",
    "init_mlp_params",
    "truncated_lecun_normal",
    "celu_grad",
    "cosine_switch_grad",
    "celu",
    "species_mlp",
    "Model",
    "model_energy",
    "energy_forces_virial",
    "philox4x32",
    "philox_normals",
    "baoab_step",
]

# ---------------------------------------------------------------------------
# Periodic table and masses (utils/periodic_table.py:18-46)
# ---------------------------------------------------------------------------
PERIODIC_TABLE = [
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
]
ATOMIC_MASSES = np.array([
    0.0, 1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999,
    18.99840316, 20.1797, 22.98976928, 24.305, 26.9815385, 28.085,
    30.973762, 32.06, 35.45, 39.948,
])
SYMBOL_TO_Z = {s: z for z, s in enumerate(PERIODIC_TABLE)}


# ---------------------------------------------------------------------------
# Units (utils/atomic_units.py:16-134, UnitSystem :142-260, md/utils.py:9)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class MDUnits:
    """fennol_md's unit system: Angstrom, ps, kcal/mol (md/utils.py:9)."""

    ANGSTROM: float = 0.52917721
    FS: float = 2.4188843e-2
    KCALPERMOL: float = 627.5096080305927
    EV: float = 27.211386024367243
    MPROT: float = 1836.1526734252586
    K_B_AU: float = 3.166811563e-6

    @property
    def PS(self) -> float:
        return self.FS / 1000.0

    @property
    def mass_unit(self) -> float:
        # M derived from E, T, L: us_to_au[M] = E^1 T^2 L^-2 (atomic_units.py:205-206)
        e = 1.0 / self.KCALPERMOL
        t = 1.0 / self.PS
        l = 1.0 / self.ANGSTROM
        return e * t**2 * l**-2

    @property
    def DA(self) -> float:
        """Internal-mass units per Dalton multiplier ``us.DA`` (atomic_units.py:101, :224-225)."""
        return self.mass_unit * (1.0 / self.MPROT)

    @property
    def K_B(self) -> float:
        """Boltzmann constant in kcal/mol/K (CONSTANTS, atomic_units.py:234-238)."""
        return self.K_B_AU * self.KCALPERMOL

    def energy_multiplier(self, unit: str) -> float:
        """``us.get_multiplier(unit)`` for energy units (atomic_units.py:77-89, :219-221)."""
        table = {"HA": 1.0, "HARTREE": 1.0, "EV": self.EV, "KCALPERMOL": self.KCALPERMOL}
        return table[unit.upper()] / self.KCALPERMOL

    @property
    def fs(self) -> float:
        """Multiplier ps -> fs (``us.FS``): a value in fs is divided by it."""
        return (1.0 / self.PS) * self.FS


md_units = MDUnits()


# ---------------------------------------------------------------------------
# Neighbour list (GraphGenerator.__call__, preprocessing.py:74-355)
# ---------------------------------------------------------------------------
def _f32_dot3(v: np.ndarray, m: np.ndarray) -> np.ndarray:
    """Row-vector times 3x3 in float32 with a fixed order ((a0*m0j + a1*m1j) + a2*m2j).

    The reference calls ``np.dot(vec, R)`` (preprocessing.py:131-133); for
    axis-aligned cells every order gives the same bits, and for skewed cells
    the order only matters at exact rounding ties of the fractional image.
    """
    v = v.astype(np.float32)
    m = m.astype(np.float32)
    out = np.empty_like(v)
    for j in range(3):
        out[:, j] = (v[:, 0] * m[0, j] + v[:, 1] * m[1, j]) + v[:, 2] * m[2, j]
    return out


def _f32_d2(vec: np.ndarray) -> np.ndarray:
    """d12 = (vec**2).sum(-1) in float32, sequential order (preprocessing.py:265)."""
    sq = (vec * vec).astype(np.float32)
    return (sq[:, 0] + sq[:, 1]) + sq[:, 2]


def half_pairs(coords, natoms, cutoff, cells=None, reciprocal_cells=None,
               minimum_image=True, skin=0.0):
    """All pairs i<j per system with d12 < (cutoff+skin)^2, in the reference's order.

    Restates GraphGenerator.__call__ (preprocessing.py:74-355): triu pairs per
    system (:105-118), minimum image (:127-143) or general PBC (:144-257)
    in float32, ``d12 = sum(vec**2)`` and the strict test ``d12 < rc^2``
    (:265-267).  Returns (edge_src, edge_dst, d12, pbc_shifts) of the *half*
    list, unpadded; pbc_shifts is None without cells.  For general PBC the
    shifts include the wrapping shifts of both atoms (:288-293).
    """
    coords = np.asarray(coords, dtype=np.float32)
    natoms = np.asarray(natoms, dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(natoms)])
    rc2 = np.float32((cutoff + skin) ** 2)
    src_all, dst_all, d12_all, sh_all = [], [], [], []
    for b in range(len(natoms)):
        n = int(natoms[b])
        off = int(offsets[b])
        if n < 2:
            continue
        x = coords[off:off + n]
        if (cells is None or minimum_image) and n > 3000:
            # same pairs and order as triu_indices(n, 1), produced in row blocks
            Cb = Rb = None
            if cells is not None:
                C_ = np.asarray(cells, dtype=np.float32)
                R_ = np.asarray(reciprocal_cells, dtype=np.float32)
                Cb = C_[b] if C_.shape[0] > 1 else C_[0]
                Rb = R_[b] if R_.shape[0] > 1 else R_[0]
            rows = max(1, 20_000_000 // n)
            for i0 in range(0, n - 1, rows):
                i1 = min(n - 1, i0 + rows)
                cnt = n - 1 - np.arange(i0, i1)
                p1 = np.repeat(np.arange(i0, i1), cnt)
                starts = np.repeat(np.cumsum(cnt) - cnt, cnt)
                p2 = np.arange(p1.shape[0]) - starts + p1 + 1
                vec = x[p2] - x[p1]
                shifts = None
                if Cb is not None:
                    shifts = -np.round(_f32_dot3(vec, Rb))
                    vec = vec + _f32_dot3(shifts, Cb)
                d12 = _f32_d2(vec.astype(np.float32))
                keep = d12 < rc2
                src_all.append(p1[keep] + off)
                dst_all.append(p2[keep] + off)
                d12_all.append(d12[keep])
                if shifts is not None:
                    sh_all.append(shifts.astype(np.float32)[keep])
            continue
        p1, p2 = np.triu_indices(n, 1)
        if cells is None:
            vec = x[p2] - x[p1]
            shifts = None
            keep_idx = None
        else:
            C = np.asarray(cells, dtype=np.float32)
            R = np.asarray(reciprocal_cells, dtype=np.float32)
            Cb = C[b] if C.shape[0] > 1 else C[0]
            Rb = R[b] if R.shape[0] > 1 else R[0]
            if minimum_image:
                vec = x[p2] - x[p1]
                shifts = -np.round(_f32_dot3(vec, Rb))
                vec = vec + _f32_dot3(shifts, Cb)
            else:
                frac = _f32_dot3(x, Rb)
                at_shifts = -np.floor(frac)
                xw = x + _f32_dot3(at_shifts, Cb)
                vec0 = xw[p2] - xw[p1]
                # float32 as the reference: (sum(R**2, axis=1))**0.5 and cutoff_skin * it
                # (preprocessing.py:152-154)
                Rf = Rb.astype(np.float32)
                inv_d = np.sqrt((Rf[0] * Rf[0] + Rf[1] * Rf[1]) + Rf[2] * Rf[2]).astype(np.float32)
                nrep = np.ceil(np.float32(cutoff + skin) * inv_d).astype(int)
                grid = np.array(
                    np.meshgrid(*[np.arange(-k, k + 1) for k in nrep]), dtype=np.float32
                ).T.reshape(-1, 3)
                dvec = _f32_dot3(grid, Cb)
                vec = (vec0[:, None, :] + dvec[None, :, :]).reshape(-1, 3)
                shifts = np.broadcast_to(grid[None], (p1.shape[0], grid.shape[0], 3)).reshape(-1, 3)
                p1 = np.repeat(p1, grid.shape[0])
                p2 = np.repeat(p2, grid.shape[0])
                shifts = shifts + at_shifts[p2] - at_shifts[p1]
        d12 = _f32_d2(vec.astype(np.float32))
        keep = d12 < rc2
        src_all.append(p1[keep] + off)
        dst_all.append(p2[keep] + off)
        d12_all.append(d12[keep])
        if shifts is not None:
            sh_all.append(np.asarray(shifts, dtype=np.float32)[keep])
    if not src_all:
        empty = np.zeros(0, np.int32)
        return empty, empty, np.zeros(0, np.float32), (None if cells is None else np.zeros((0, 3), np.float32))
    src = np.concatenate(src_all).astype(np.int32)
    dst = np.concatenate(dst_all).astype(np.int32)
    d12 = np.concatenate(d12_all).astype(np.float32)
    sh = None if cells is None else np.concatenate(sh_all).astype(np.float32)
    return src, dst, d12, sh


def skin_update(src_skin, dst_skin, shifts_skin, coords, cutoff, cells=None, batch_index=None):
    """Pairs of a skin list still within ``cutoff`` at new coordinates.

    Restates GraphGenerator.update_skin (preprocessing.py:640-696) on the half
    list: ``vec = x[dst] - x[src] + shift . C`` with the shifts stored at the
    build (the cell of each pair's system via ``batch_index[src]`` when there
    are several, :655-659), ``d12 = sum(vec**2)`` in float32 and the strict test
    ``d12 < rc^2`` (:662-663), keeping the skin list's order (mask_filter_1d).
    Returns (src, dst, d12, shifts); shifts is None without cells.
    """
    src = np.asarray(src_skin, np.int64)
    dst = np.asarray(dst_skin, np.int64)
    x = np.asarray(coords, np.float32)
    vec = (x[dst] - x[src]).astype(np.float32)
    sh = None
    if cells is not None:
        C = np.asarray(cells, np.float32).reshape(-1, 3, 3)
        sh = np.asarray(shifts_skin, np.float32)
        if C.shape[0] == 1:
            vec = vec + _f32_dot3(sh, C[0])
        else:
            bidx = np.asarray(batch_index)[src]
            for b in range(C.shape[0]):
                m = bidx == b
                if np.any(m):
                    vec[m] = vec[m] + _f32_dot3(sh[m], C[b])
    d12 = _f32_d2(vec.astype(np.float32))
    keep = d12 < np.float32(cutoff ** 2)
    return (src[keep].astype(np.int32), dst[keep].astype(np.int32), d12[keep],
            None if sh is None else sh[keep])


def symmetrize(src, dst, d12, shifts):
    """Directed list: (src,dst) then (dst,src) with negated shifts (preprocessing.py:320-326)."""
    s = np.concatenate([src, dst])
    d = np.concatenate([dst, src])
    dd = np.concatenate([d12, d12])
    sh = None if shifts is None else np.concatenate([shifts, -shifts])
    return s, d, dd, sh


def filter_graph(d12, cutoff):
    """Indices of directed parent edges with d12 < cutoff^2 (GraphFilter, preprocessing.py:852-870)."""
    return np.nonzero(np.asarray(d12) < np.float32(cutoff) ** 2)[0].astype(np.int32)


def angular_triplets(edge_src, nat):
    """All slot pairs a<b of each centre's angular edges (preprocessing.py:1085-1181).

    Returns (angle_src, angle_dst, central_atom) as indices into the angular
    edge list; the pair set does not depend on the sort order of equal keys.
    """
    edge_src = np.asarray(edge_src)
    order = np.argsort(edge_src, kind="stable")
    counts = np.bincount(edge_src, minlength=nat)[:nat]
    starts = np.concatenate([[0], np.cumsum(counts)])
    a_src, a_dst, cen = [], [], []
    for i in range(nat):
        e = order[starts[i]:starts[i + 1]]
        if e.shape[0] < 2:
            continue
        l1, l2 = np.triu_indices(e.shape[0], 1)
        a_src.append(e[l1])
        a_dst.append(e[l2])
        cen.append(np.full(l1.shape[0], i))
    if not a_src:
        z = np.zeros(0, np.int32)
        return z, z, z
    return (np.concatenate(a_src).astype(np.int32), np.concatenate(a_dst).astype(np.int32),
            np.concatenate(cen).astype(np.int32))


def _scatter_add_rows(n_rows, idx, vals):
    """out[idx[e]] += vals[e] for 2-D vals, via one bincount per column (= np.add.at)."""
    vals = np.asarray(vals)
    out = np.empty((n_rows, vals.shape[1]))
    for c in range(vals.shape[1]):
        out[:, c] = np.bincount(idx, weights=vals[:, c], minlength=n_rows)[:n_rows]
    return out


# ---------------------------------------------------------------------------
# Graph processing and switch (GraphProcessor :699-777, SwitchFunction misc.py:617-716)
# ---------------------------------------------------------------------------
def edge_vectors(coords, src, dst, shifts=None, cells=None, batch_index=None):
    """vec = x[dst] - x[src] + shifts @ cell (preprocessing.py:722-735), FP64."""
    x = np.asarray(coords, dtype=np.float64)
    vec = x[dst] - x[src]
    if shifts is not None:
        C = np.asarray(cells, dtype=np.float64)
        if C.shape[0] == 1:
            vec = vec + np.asarray(shifts, np.float64) @ C[0]
        else:
            vec = vec + np.einsum("ej,eji->ei", np.asarray(shifts, np.float64), C[batch_index[src]])
    return vec


def safe_distance(vec):
    """safe_sqrt(sum vec^2) = sqrt(max(d2, 1e-5)) (activations.py:173-174)."""
    return np.sqrt(np.maximum((vec * vec).sum(-1), 1.0e-5))


def cosine_switch(d, cutoff):
    """(0.5 cos(pi d / rc) + 0.5) for d < rc else 0 (misc.py:674-678, :708)."""
    d = np.asarray(d, dtype=np.float64)
    return np.where(d < cutoff, 0.5 * np.cos(d * (math.pi / cutoff)) + 0.5, 0.0)


def cosine_switch_grad(d, cutoff):
    d = np.asarray(d, dtype=np.float64)
    return np.where(d < cutoff, -0.5 * (math.pi / cutoff) * np.sin(d * (math.pi / cutoff)), 0.0)


# ---------------------------------------------------------------------------
# ANI AEV (ani.py:47-142)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class AEVParams:
    """ANI_AEV hyper-parameters (ani.py:18-41).

    Defaults are the values this build uses for its workloads: ANI-2x radial
    and angular widths (examples/training/ani.yaml:59-68) with 4 angular
    distance divisions x 8 angle sections.
    """

    species_order: tuple = ("H", "C", "N", "O")
    cutoff: float = 5.2
    cutoff_angular: float = 3.5
    radial_eta: float = 19.7
    angular_eta: float = 12.5
    radial_dist_divisions: int = 16
    angular_dist_divisions: int = 4
    zeta: float = 14.1
    angle_sections: int = 8
    radial_start: float = 0.8
    angular_start: float = 0.8

    @property
    def n_species(self) -> int:
        return len(self.species_order)

    @property
    def n_pairs(self) -> int:
        s = self.n_species
        return s * (s + 1) // 2

    @property
    def radial_length(self) -> int:
        return self.n_species * self.radial_dist_divisions

    @property
    def angular_channels(self) -> int:
        return self.angular_dist_divisions * self.angle_sections

    def radial_shifts(self) -> np.ndarray:
        """R_k = linspace(radial_start, rc, n_R + 1)[:-1] (ani.py:73-78)."""
        return np.linspace(self.radial_start, self.cutoff, self.radial_dist_divisions + 1)[:-1]

    def angular_shifts(self) -> np.ndarray:
        """A_d = linspace(angular_start, rc_a, n_A + 1)[:-1] (ani.py:105-110)."""
        return np.linspace(self.angular_start, self.cutoff_angular, self.angular_dist_divisions + 1)[:-1]

    def angle_centres(self) -> np.ndarray:
        """theta_z = linspace(0, pi, n_Z + 1)[:-1] + pi / (2 n_Z) (ani.py:100-104)."""
        return np.linspace(0, np.pi, self.angle_sections + 1)[:-1] + np.pi / (2 * self.angle_sections)

    def species_map(self) -> np.ndarray:
        """conv[Z] = position in species_order, 0 for others (ani.py:52-60)."""
        conv = np.zeros(len(PERIODIC_TABLE) + 1, dtype=np.int32)
        for i, s in enumerate(self.species_order):
            conv[SYMBOL_TO_Z[s]] = i
        return conv

    def pair_index(self) -> np.ndarray:
        """Symmetric S x S table of row-major upper-triangle positions (ani.py:124-130)."""
        s = self.n_species
        t = np.zeros((s, s), dtype=np.int32)
        a, b = np.triu_indices(s, 0)
        t[a, b] = np.arange(a.shape[0])
        t[b, a] = np.arange(a.shape[0])
        return t


def aev_length(p: AEVParams) -> int:
    return p.radial_length + p.n_pairs * p.angular_channels


@dataclass
class Graphs:
    """Directed radial edges plus angular edges and triplets, as the reference builds them."""

    src: np.ndarray
    dst: np.ndarray
    shifts: np.ndarray | None
    ang_edges: np.ndarray          # indices into the radial directed list
    angle_src: np.ndarray          # indices into ang_edges
    angle_dst: np.ndarray
    central: np.ndarray


def build_graphs(coords, natoms, params: AEVParams, cells=None, reciprocal_cells=None,
                 minimum_image=True) -> Graphs:
    """GraphGenerator -> GraphFilter -> GraphAngularExtension on the host (preprocessing.py)."""
    s, d, d12, sh = half_pairs(coords, natoms, params.cutoff, cells, reciprocal_cells, minimum_image)
    s, d, d12, sh = symmetrize(s, d, d12, sh)
    ang = filter_graph(d12, params.cutoff_angular)
    nat = int(np.sum(natoms))
    a_src, a_dst, cen = angular_triplets(s[ang], nat)
    return Graphs(s, d, sh, ang, a_src, a_dst, cen)


def _aev_geometry(coords, g: Graphs, params: AEVParams, cells, batch_index):
    vec = edge_vectors(coords, g.src, g.dst, g.shifts, cells, batch_index)
    return vec


def ani_aev(coords, species, g: Graphs, params: AEVParams, cells=None, batch_index=None,
            return_parts=False):
    """ANI AEV in FP64 (ani.py:47-142, GraphProcessor :722-743, GraphAngleProcessor :1296-1307).

    Radial channel ``s(dst) * n_R + k``; angular channel
    ``n_S*n_R + pair(s_j, s_k) * n_A*n_Z + d * n_Z + z``.
    """
    nat = species.shape[0]
    conv = params.species_map()
    sidx = conv[np.asarray(species)]
    S, nR = params.n_species, params.radial_dist_divisions
    nA, nZ = params.angular_dist_divisions, params.angle_sections
    vec = _aev_geometry(coords, g, params, cells, batch_index)
    dist = safe_distance(vec)
    fc = cosine_switch(dist, params.cutoff)
    Rk = params.radial_shifts()
    rad_terms = np.exp(-params.radial_eta * (dist[:, None] - Rk[None, :]) ** 2) * (0.25 * fc)[:, None]
    radial = _scatter_add_rows(nat * S, g.src * S + sidx[g.dst], rad_terms).reshape(nat, S * nR)

    # angular graph: gathered vec / distances, switch with the angular cutoff
    avec = vec[g.ang_edges]
    adist = dist[g.ang_edges]
    afc = cosine_switch(adist, params.cutoff_angular)
    dirs = avec / np.maximum(adist, 1.0e-5)[:, None]
    cos_t = (dirs[g.angle_src] * dirs[g.angle_dst]).sum(-1)
    theta = np.arccos(0.95 * cos_t)
    tz = params.angle_centres()
    Ad = params.angular_shifts()
    pref = 2.0 * 0.5 ** params.zeta
    f1 = pref * afc[g.angle_src] * afc[g.angle_dst]
    f1 = f1[:, None] * (1.0 + np.cos(theta[:, None] - tz[None, :])) ** params.zeta
    dm = 0.5 * (adist[g.angle_src] + adist[g.angle_dst])
    f2 = np.exp(-params.angular_eta * (dm[:, None] - Ad[None, :]) ** 2)
    terms = (f1[:, None, :] * f2[:, :, None]).reshape(-1, nA * nZ)
    P = params.n_pairs
    ptab = params.pair_index()
    adst = sidx[g.dst[g.ang_edges]]
    aidx = g.central * P + ptab[adst[g.angle_src], adst[g.angle_dst]]
    angular = _scatter_add_rows(nat * P, aidx, terms).reshape(nat, P * nA * nZ)
    aev = np.concatenate([radial, angular], axis=1)
    if return_parts:
        return aev, dict(vec=vec, dist=dist)
    return aev


def ani_aev_backward(coords, species, g: Graphs, params: AEVParams, dE_daev,
                     cells=None, batch_index=None):
    """Adjoint of :func:`ani_aev`: dE/dvec per directed radial edge (FP64).

    Hand-derived: radial d/dd of 0.25 fc(d) exp(-eta (d-R)^2); angular through
    d_j, d_k and cos(theta) with theta = arccos(0.95 cos), cos = u_j . u_k.
    """
    nat = species.shape[0]
    conv = params.species_map()
    sidx = conv[np.asarray(species)]
    S, nR = params.n_species, params.radial_dist_divisions
    nA, nZ = params.angular_dist_divisions, params.angle_sections
    G = np.asarray(dE_daev, dtype=np.float64)
    Grad = G[:, : S * nR].reshape(nat, S, nR)
    Gang = G[:, S * nR:].reshape(nat, params.n_pairs, nA, nZ)

    vec = edge_vectors(coords, g.src, g.dst, g.shifts, cells, batch_index)
    d2 = (vec * vec).sum(-1)
    dist = np.sqrt(np.maximum(d2, 1.0e-5))
    ddist_dvec = np.where((d2 > 1.0e-5)[:, None], vec / dist[:, None], 0.0)
    fc = cosine_switch(dist, params.cutoff)
    dfc = cosine_switch_grad(dist, params.cutoff)
    Rk = params.radial_shifts()
    e = np.exp(-params.radial_eta * (dist[:, None] - Rk[None, :]) ** 2)
    dterm = 0.25 * (dfc[:, None] * e + fc[:, None] * e * (-2.0 * params.radial_eta) * (dist[:, None] - Rk[None, :]))
    gsel = Grad[g.src, sidx[g.dst], :]
    dE_dd = (gsel * dterm).sum(-1)
    grad_vec = dE_dd[:, None] * ddist_dvec

    # angular
    ae = g.ang_edges
    avec = vec[ae]
    adist = dist[ae]
    add_dvec = ddist_dvec[ae]
    afc = cosine_switch(adist, params.cutoff_angular)
    adfc = cosine_switch_grad(adist, params.cutoff_angular)
    clip = np.maximum(adist, 1.0e-5)
    u = avec / clip[:, None]
    ja, ka = g.angle_src, g.angle_dst
    c = (u[ja] * u[ka]).sum(-1)
    x = 0.95 * c
    theta = np.arccos(x)
    dtheta_dc = -0.95 / np.sqrt(np.maximum(1.0 - x * x, 1e-300))
    tz = params.angle_centres()
    Ad = params.angular_shifts()
    z = params.zeta
    pref = 2.0 * 0.5 ** z
    base = 1.0 + np.cos(theta[:, None] - tz[None, :])
    F = base ** z
    dF_dtheta = z * base ** (z - 1.0) * (-np.sin(theta[:, None] - tz[None, :]))
    dm = 0.5 * (adist[ja] + adist[ka])
    Rd = np.exp(-params.angular_eta * (dm[:, None] - Ad[None, :]) ** 2)
    dRd_dm = Rd * (-2.0 * params.angular_eta) * (dm[:, None] - Ad[None, :])
    P = params.n_pairs
    ptab = params.pair_index()
    adst = sidx[g.dst[ae]]
    Gt = Gang[g.central, ptab[adst[ja], adst[ka]]]        # [T, nA, nZ]
    sw = pref * afc[ja] * afc[ka]
    # E_t = sum_dz G * sw * F_z * R_d
    GF = np.einsum("tdz,tz->td", Gt, F)                   # sum over z
    GFR = (GF * Rd).sum(-1)                               # sum G F R
    GFdR = (GF * dRd_dm).sum(-1)
    GdFR = np.einsum("tdz,tz,td->t", Gt, dF_dtheta, Rd)
    dE_dswj = pref * adfc[ja] * afc[ka] * GFR
    dE_dswk = pref * afc[ja] * adfc[ka] * GFR
    dE_ddj = dE_dswj + sw * GFdR * 0.5
    dE_ddk = dE_dswk + sw * GFdR * 0.5
    dE_dc = sw * GdFR * dtheta_dc
    # d c / d vec_j = (u_k - c u_j)/clip_j (when d > 1e-5)
    dc_dvj = (u[ka] - c[:, None] * u[ja]) / clip[ja][:, None]
    dc_dvk = (u[ja] - c[:, None] * u[ka]) / clip[ka][:, None]
    gj = dE_ddj[:, None] * add_dvec[ja] + dE_dc[:, None] * dc_dvj
    gk = dE_ddk[:, None] * add_dvec[ka] + dE_dc[:, None] * dc_dvk
    gang = _scatter_add_rows(ae.shape[0], ja, gj) + _scatter_add_rows(ae.shape[0], ka, gk)
    grad_vec = grad_vec + _scatter_add_rows(grad_vec.shape[0], ae, gang)
    return grad_vec, vec


# ---------------------------------------------------------------------------
# Species MLP (FullyConnectedNet nets.py:13-102, SpeciesIndexNet :301-426)
# ---------------------------------------------------------------------------
CELU_ALPHA = 0.1


def celu(x, alpha=CELU_ALPHA):
    """jax.nn.celu: max(0,x) + alpha * (exp(min(0,x)/alpha) - 1) (activations.py:306-331)."""
    return np.where(x > 0, x, alpha * np.expm1(np.minimum(x, 0.0) / alpha))


def celu_grad(x, alpha=CELU_ALPHA):
    return np.where(x > 0, 1.0, np.exp(np.minimum(x, 0.0) / alpha))


def truncated_lecun_normal(rng: np.random.Generator, fan_in: int, fan_out: int) -> np.ndarray:
    """Flax lecun_normal: truncated normal on [-2, 2], std sqrt(1/fan_in)/0.87962566 (initializers.py:33-47)."""
    out = rng.standard_normal((fan_in, fan_out))
    bad = np.abs(out) > 2.0
    while np.any(bad):
        out[bad] = rng.standard_normal(int(bad.sum()))
        bad = np.abs(out) > 2.0
    return out * (math.sqrt(1.0 / fan_in) / 0.87962566103423978)


def init_mlp_params(seed: int, n_in: int, hidden=(256, 192, 160), n_out: int = 1,
                    species=("H", "C", "N", "O"), zero_bias: bool = True):
    """Seeded per-species weights: {sym: [(W[in,out], b[out]), ...]} in FP64."""
    rng = np.random.Generator(np.random.PCG64(seed))
    params = {}
    dims = [n_in, *hidden, n_out]
    for s in species:
        layers = []
        for a, b in zip(dims[:-1], dims[1:]):
            W = truncated_lecun_normal(rng, a, b)
            bias = np.zeros(b) if zero_bias else 0.1 * rng.standard_normal(b)
            layers.append((W, bias))
        params[s] = layers
    return params


def species_mlp(aev, species, params: dict, return_grad: bool = False):
    """Atomic energies by species-specific MLPs, optional dE/dAEV (FP64)."""
    aev = np.asarray(aev, dtype=np.float64)
    nat = aev.shape[0]
    out = np.zeros(nat)
    grad = np.zeros_like(aev) if return_grad else None
    sym = np.array([PERIODIC_TABLE[z] if z > 0 else "" for z in np.asarray(species)])
    present = {s for s in sym if s}
    missing = present - set(params.keys())
    if missing:
        raise ValueError(f"Species {sorted(missing)} not found in networks (nets.py:394-397)")
    for s, layers in params.items():
        idx = np.nonzero(sym == s)[0]
        if idx.size == 0:
            continue
        h = aev[idx]
        zs = []
        for l, (W, b) in enumerate(layers):
            zz = h @ W + b
            if l < len(layers) - 1:
                zs.append(zz)
                h = celu(zz)
            else:
                h = zz
        out[idx] = h[:, 0]
        if return_grad:
            gcur = np.broadcast_to(layers[-1][0][:, 0], (idx.size, layers[-1][0].shape[0])).copy()
            for l in range(len(layers) - 2, -1, -1):
                gcur = gcur * celu_grad(zs[l])
                gcur = gcur @ layers[l][0].T
            grad[idx] = gcur
    return (out, grad) if return_grad else out


# ---------------------------------------------------------------------------
# Model, energy, forces, virial (fennix.py:216-347)
# ---------------------------------------------------------------------------
@dataclass
class Model:
    aev: AEVParams
    params: dict
    energy_unit: str = "eV"
    minimum_image: bool = True


def model_energy(model: Model, coords, species, natoms=None, cells=None, reciprocal_cells=None,
                 graphs: Graphs | None = None):
    """Per-system total energy in FP64 (fennix.py:224-258)."""
    species = np.asarray(species)
    nat = species.shape[0]
    natoms = np.array([nat]) if natoms is None else np.asarray(natoms)
    batch_index = np.repeat(np.arange(len(natoms)), natoms)
    if graphs is None:
        graphs = build_graphs(coords, natoms, model.aev, cells, reciprocal_cells, model.minimum_image)
    aev = ani_aev(coords, species, graphs, model.aev, cells, batch_index)
    eat = species_mlp(aev, species, model.params)
    eat = np.where(species > 0, eat, 0.0)
    return np.bincount(batch_index, weights=eat, minlength=len(natoms)), eat


def energy_forces_virial(model: Model, coords, species, natoms=None, cells=None,
                         reciprocal_cells=None, graphs: Graphs | None = None):
    """E [B], F = -dE/dx [N,3], virial dE/dstrain [B,3,3] (fennix.py:260-347), FP64.

    Virial[a][b] = sum over directed edges of vec_a * dE/dvec_b, equal to the
    derivative with respect to the strain S in x -> x S, cell -> cell S.
    """
    species = np.asarray(species)
    nat = species.shape[0]
    natoms = np.array([nat]) if natoms is None else np.asarray(natoms)
    batch_index = np.repeat(np.arange(len(natoms)), natoms)
    if graphs is None:
        graphs = build_graphs(coords, natoms, model.aev, cells, reciprocal_cells, model.minimum_image)
    aev = ani_aev(coords, species, graphs, model.aev, cells, batch_index)
    eat, gaev = species_mlp(aev, species, model.params, return_grad=True)
    eat = np.where(species > 0, eat, 0.0)
    energy = np.bincount(batch_index, weights=eat, minlength=len(natoms))
    gvec, vec = ani_aev_backward(coords, species, graphs, model.aev, gaev, cells, batch_index)
    grad = _scatter_add_rows(nat, graphs.dst, gvec) - _scatter_add_rows(nat, graphs.src, gvec)
    vir = _scatter_add_rows(len(natoms), batch_index[graphs.src],
                            (vec[:, :, None] * gvec[:, None, :]).reshape(-1, 9)).reshape(-1, 3, 3)
    return energy, -grad, vir, eat


# ---------------------------------------------------------------------------
# Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
# ---------------------------------------------------------------------------
_PHILOX_M0 = 0xD2511F53
_PHILOX_M1 = 0xCD9E8D57
_PHILOX_W0 = 0x9E3779B9
_PHILOX_W1 = 0xBB67AE85
_MASK32 = 0xFFFFFFFF


def philox4x32(ctr, key, rounds: int = 10):
    """Vectorised Philox4x32: ctr uint32[..., 4], key uint32[..., 2] -> uint32[..., 4]."""
    c = np.asarray(ctr, dtype=np.uint64) & _MASK32
    k = np.asarray(key, dtype=np.uint64) & _MASK32
    c0, c1, c2, c3 = (c[..., i].copy() for i in range(4))
    k0, k1 = k[..., 0].copy(), k[..., 1].copy()
    for r in range(rounds):
        p0 = np.uint64(_PHILOX_M0) * c0
        p1 = np.uint64(_PHILOX_M1) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(_MASK32)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(_MASK32)
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
        if r < rounds - 1:
            k0 = (k0 + np.uint64(_PHILOX_W0)) & np.uint64(_MASK32)
            k1 = (k1 + np.uint64(_PHILOX_W1)) & np.uint64(_MASK32)
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def philox_normals(seed: int, step: int, n_atoms: int, stream: int = 0):
    """Standard normals [n_atoms, 3] from one Philox block per atom.

    counter = (atom, step_lo, step_hi, stream), key = (seed_lo, seed_hi);
    uniforms u = (r + 0.5) * 2^-32; Box-Muller on (r0, r1) and (r2, r3),
    keeping three of the four normals.
    """
    atoms = np.arange(n_atoms, dtype=np.uint64)
    ctr = np.stack([atoms,
                    np.full(n_atoms, step & _MASK32, np.uint64),
                    np.full(n_atoms, (step >> 32) & _MASK32, np.uint64),
                    np.full(n_atoms, stream, np.uint64)], axis=-1)
    key = np.broadcast_to(np.array([seed & _MASK32, (seed >> 32) & _MASK32], np.uint64), (n_atoms, 2))
    r = philox4x32(ctr, key).astype(np.float64)
    u = (r + 0.5) * (2.0 ** -32)
    rad0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    rad1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    ang0 = 2.0 * math.pi * u[:, 1]
    ang1 = 2.0 * math.pi * u[:, 3]
    return np.stack([rad0 * np.cos(ang0), rad0 * np.sin(ang0), rad1 * np.cos(ang1)], axis=-1)


# ---------------------------------------------------------------------------
# BAOAB Langevin (md/integrate.py:286-294, md/thermostats.py:56-120)
# ---------------------------------------------------------------------------
def langevin_coefficients(dt, gamma, kT, mass):
    """a1 = exp(-gamma dt), a2 = sqrt((1 - a1^2) kT / m) (thermostats.py:62-63)."""
    a1 = math.exp(-gamma * dt)
    a2 = np.sqrt((1.0 - a1 * a1) * kT / np.asarray(mass, np.float64))
    return a1, a2


def baoab_step(x, v, forces, mass, dt, a1, a2, noise):
    """One step: v += dt/m F; x += dt/2 v; v = a1 v + a2 xi; x += dt/2 v (integrate.py:286-294)."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    m = np.asarray(mass, np.float64)[:, None]
    v = v + (dt / m) * forces
    x = x + (0.5 * dt) * v
    v = a1 * v + np.asarray(a2, np.float64)[:, None] * noise
    x = x + (0.5 * dt) * v
    return x, v
