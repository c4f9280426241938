// Cell-list neighbour lists for gfx950.
//
// Replaces FeNNol's all-pairs GraphGenerator (models/preprocessing.py:74-638)
// and the angular GraphFilter (:780-960) with an O(N) cell list that keeps the
// reference's pair set bit for bit: every candidate pair is tested with the
// reference's float32 arithmetic (fnl_min_image / fnl_d12), and the candidate
// set is a superset of the pairs within the cutoff: cells at least half a
// cutoff wide searched two cells deep along each axis (5x5x5 neighbourhood),
// or, where the caller's grid capacity leaves them at least one cutoff wide,
// one cell deep (3x3x3); deduplicated for grids narrower than the
// neighbourhood.  Half-cutoff cells halve the candidates per centre (about
// 240 instead of 470 in liquid water at a 5.2 A cutoff).
//
// Pipeline (all on the caller's stream):
//   k_grid      one workgroup per system: grid dims (+ bounding box if open);
//               clears the cell counters
//   k_bin       one thread per atom: cell id, slot in the cell (population
//               count), atom staged at its cell's slot; the last workgroup
//               scans the populations into cell offsets
//   k_cell_sort one half-wave per cell: the staged atoms ordered by index
//               (deterministic rows) into the cell's segment, with
//               cell-sorted float4 positions, and the cell's table of
//               neighbouring-cell candidate ranges
//   k_pairs     one half-wave per centre, centres in cell order: candidates
//               of the neighbouring cells (contiguous z-runs of the
//               cell-sorted arrays) tested 32 at a time, kept pairs compacted
//               with a ballot + prefix popcount into the centre's row
//   (row maxima go to the status words only from rows that overflowed: they
//   size the regrown rows; a same-address atomic per centre would serialise)
//
// Skin update (a list built with cutoff + skin, re-filtered at the current
// positions on the steps between builds):
//   k_status_clear, k_skin_filter (one half-wave per centre)

#include <vector>

#include "fnl_common.h"

namespace {

// per system: nx ny nz first | ox oy oz (open: box origin) | iwx iwy iwz (cells per unit of
// fractional or Cartesian extent, float bits) | pbc mode | search depth per axis (1 or 2:
// bits 0-3 x, 4-7 y, 8-11 z)
constexpr int kGridWords = 12;
constexpr int kCellCap = 32;    // atoms per cell staged by k_bin (more: k_cell_sort gathers them)
constexpr int kMaxRanges = 50;  // 5 x 5 columns x at most 2 z-runs

struct NblistWs {
  int32_t* grid;       // [n_sys][kGridWords] (floats stored as bits)
  int32_t* atom_cell;  // [N]
  int32_t* cell_count; // [n_cells]
  int32_t* cell_start; // [n_cells + 1]
  int32_t* stage;      // [n_cells][kCellCap] atoms in binning order
  unsigned* done;      // completion count of k_bin (left at zero)
  int32_t* slot_cell;  // [N] cell of each position of the cell-sorted arrays
  float4* pos;         // [N] cell-sorted x, y, z, atom index bits
  int2* ranges;        // [n_cells][kMaxRanges] candidate ranges of each cell's neighbourhood
};

size_t ws_layout(int n_atoms, int n_sys, int n_cells, char* base, NblistWs* ws) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = fnl_align256(off + bytes);
    return o;
  };
  size_t o_grid = take((size_t)n_sys * kGridWords * 4);
  size_t o_ac = take((size_t)n_atoms * 4);
  size_t o_cc = take((size_t)n_cells * 4);
  size_t o_cs = take((size_t)(n_cells + 1) * 4);
  size_t o_st = take((size_t)n_cells * kCellCap * 4);
  size_t o_dn = take(4);
  size_t o_sc = take((size_t)n_atoms * 4);
  size_t o_p = take((size_t)n_atoms * 16);
  size_t o_r = take((size_t)n_cells * kMaxRanges * 8);
  if (base && ws) {
    ws->grid = (int32_t*)(base + o_grid);
    ws->atom_cell = (int32_t*)(base + o_ac);
    ws->cell_count = (int32_t*)(base + o_cc);
    ws->cell_start = (int32_t*)(base + o_cs);
    ws->stage = (int32_t*)(base + o_st);
    ws->done = (unsigned*)(base + o_dn);
    ws->slot_cell = (int32_t*)(base + o_sc);
    ws->pos = (float4*)(base + o_p);
    ws->ranges = (int2*)(base + o_r);
  }
  return off;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f32(float* __restrict__ p, float v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Cells along one axis and the search depth: cells half a cutoff wide (two
// deep) where the capacity allows, else at least one cutoff wide (one deep).
// Depth x width exceeds the cutoff with a margin so that rounding in the bin
// coordinates cannot push two atoms within the cutoff more than depth cells
// apart.
__device__ __forceinline__ int cells_along(float width, float rc, int cap, int* depth) {
  const float rbin = rc * 1.0001f + 1.0e-4f;
  const int n1 = max((int)floorf(width / rbin), 1);         // widths >= rbin
  const int n = min(max((int)floorf(2.0f * width / rbin), 1), max(cap, 1));  // widths >= rbin / 2
  *depth = n > n1 ? 2 : 1;
  return n;
}

__global__ void k_grid(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
                       const int32_t* __restrict__ sys_offsets, int pbc_mode,
                       const float* __restrict__ rec, const int32_t* __restrict__ grid_cap,
                       float cutoff, int32_t* __restrict__ grid, int32_t* __restrict__ status,
                       int32_t* __restrict__ cell_count, unsigned* __restrict__ done, int n_cells, int n_sys) {
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  if (b == 0 && t < FNL_STATUS_STICKY) status[t] = 0;  // the sticky word persists
  if (b == 0 && t == 0) *done = 0u;
  // every workgroup clears a share of the cell counters; the first n_sys also set up a system
  for (int c = b * blockDim.x + t; c < n_cells; c += gridDim.x * blockDim.x) cell_count[c] = 0;
  if (b >= n_sys) return;
  const int capx = grid_cap[b * 4 + 0], capy = grid_cap[b * 4 + 1], capz = grid_cap[b * 4 + 2];
  const int first = grid_cap[b * 4 + 3];
  int32_t* g = grid + b * kGridWords;
  if (pbc_mode == FNL_PBC_GENERAL) {
    // images per direction: ceil(cutoff * |column a of R|) in float32 (preprocessing.py:152-154)
    if (t == 0) {
      const float* R = rec + b * 9;
      for (int a = 0; a < 3; ++a) {
        const float nrm = sqrtf((R[a] * R[a] + R[3 + a] * R[3 + a]) + R[6 + a] * R[6 + a]);
        g[a] = (int)ceilf(cutoff * nrm);
      }
      g[3] = first;
      g[4] = g[5] = g[6] = g[7] = g[8] = g[9] = 0;
      g[10] = 2;
      g[11] = 0;
    }
    return;
  }
  if (pbc_mode == FNL_PBC_MINIMUM_IMAGE) {
    if (t == 0) {
      const float* R = rec + b * 9;
      // lattice-plane spacing along a = 1 / |column a of R|
      int n[3], dp[3];
      for (int a = 0; a < 3; ++a) {
        const float nrm = sqrtf(R[a] * R[a] + R[3 + a] * R[3 + a] + R[6 + a] * R[6 + a]);
        const int cap = a == 0 ? capx : (a == 1 ? capy : capz);
        n[a] = cells_along(1.0f / nrm, cutoff, cap, &dp[a]);
      }
      g[0] = n[0]; g[1] = n[1]; g[2] = n[2]; g[3] = first;
      g[4] = g[5] = g[6] = 0;
      g[7] = __float_as_int((float)n[0]);
      g[8] = __float_as_int((float)n[1]);
      g[9] = __float_as_int((float)n[2]);
      g[10] = 1;
      g[11] = dp[0] | (dp[1] << 4) | (dp[2] << 8);
    }
    return;
  }
  // open boundaries: bounding box of the system's real atoms
  __shared__ float smin[3][256];
  __shared__ float smax[3][256];
  const int a0 = sys_offsets[b], a1 = sys_offsets[b + 1];
  float mn[3] = {3.0e38f, 3.0e38f, 3.0e38f}, mx[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
  for (int i = a0 + t; i < a1; i += blockDim.x) {
    if (species_idx[i] < 0) continue;
    for (int a = 0; a < 3; ++a) {
      const float x = coords[3 * i + a];
      mn[a] = fminf(mn[a], x);
      mx[a] = fmaxf(mx[a], x);
    }
  }
  for (int a = 0; a < 3; ++a) {
    smin[a][t] = mn[a];
    smax[a][t] = mx[a];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      for (int a = 0; a < 3; ++a) {
        smin[a][t] = fminf(smin[a][t], smin[a][t + s]);
        smax[a][t] = fmaxf(smax[a][t], smax[a][t + s]);
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    int n[3], dp[3];
    float inv_w[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = smax[a][0] - smin[a][0];
      const int cap = a == 0 ? capx : (a == 1 ? capy : capz);
      dp[a] = 1;
      n[a] = (ext > 0.0f) ? cells_along(ext, cutoff, cap, &dp[a]) : 1;
      inv_w[a] = (n[a] > 1) ? (float)n[a] / ext : 0.0f;
    }
    g[0] = n[0]; g[1] = n[1]; g[2] = n[2]; g[3] = first;
    g[4] = __float_as_int(smin[0][0] < 1.0e38f ? smin[0][0] : 0.0f);
    g[5] = __float_as_int(smin[1][0] < 1.0e38f ? smin[1][0] : 0.0f);
    g[6] = __float_as_int(smin[2][0] < 1.0e38f ? smin[2][0] : 0.0f);
    g[7] = __float_as_int(inv_w[0]);
    g[8] = __float_as_int(inv_w[1]);
    g[9] = __float_as_int(inv_w[2]);
    g[10] = 0;
    g[11] = dp[0] | (dp[1] << 4) | (dp[2] << 8);
  }
}

__device__ __forceinline__ int3 atom_cell3(const float* __restrict__ coords, int i,
                                           const int32_t* __restrict__ g, const float* __restrict__ rec,
                                           int b) {
  const float x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const int nx = g[0], ny = g[1], nz = g[2];
  float f[3];
  if (g[10]) {
    const float* R = rec + b * 9;
    f[0] = (x * R[0] + y * R[3]) + z * R[6];
    f[1] = (x * R[1] + y * R[4]) + z * R[7];
    f[2] = (x * R[2] + y * R[5]) + z * R[8];
    for (int a = 0; a < 3; ++a) f[a] = (f[a] - floorf(f[a])) * __int_as_float(g[7 + a]);
  } else {
    f[0] = (x - __int_as_float(g[4])) * __int_as_float(g[7]);
    f[1] = (y - __int_as_float(g[5])) * __int_as_float(g[8]);
    f[2] = (z - __int_as_float(g[6])) * __int_as_float(g[9]);
  }
  int3 c;
  c.x = min(max((int)floorf(f[0]), 0), nx - 1);
  c.y = min(max((int)floorf(f[1]), 0), ny - 1);
  c.z = min(max((int)floorf(f[2]), 0), nz - 1);
  return c;
}

// General PBC: every atom wrapped into its cell (x + at . C, at = -floor(x . R)), the
// integer wrap shift packed into w.
__global__ void k_wrap(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
                       const int32_t* __restrict__ batch_index, int n_atoms, const float* __restrict__ cells,
                       const float* __restrict__ rec, float4* __restrict__ pos, int32_t* __restrict__ n_nbr,
                       int32_t* __restrict__ n_nbr_ang, int32_t* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  if (species_idx[i] < 0) {
    n_nbr[i] = 0;
    if (n_nbr_ang) n_nbr_ang[i] = 0;
    return;
  }
  FnlCell cl;
  fnl_load_cell(cells, rec, batch_index[i], cl);
  float xw[3];
  int at[3];
  fnl_wrap(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], cl, xw, at);
  if (!fnl_shift_fits(at[0], at[1], at[2])) {  // an atom more than 511 cells away
    at[0] = at[1] = at[2] = 0;
    atomicOr(&status[FNL_STATUS_OVERFLOW], 4);
    atomicOr(&status[FNL_STATUS_STICKY], 4);
  }
  pos[i] = make_float4(xw[0], xw[1], xw[2], __int_as_float(fnl_shift_pack(at[0], at[1], at[2])));
}

// General PBC pairs (GraphGenerator, preprocessing.py:144-257, 288-293): one wave per centre
// i; the candidates are every atom j of i's system times every cell shift s with
// |s_a| <= n_a, j-major, tested 64 at a time with the reference's float32 arithmetic
// vec = (xw_j - xw_i) + s . C.  A kept slot stores j and the total shift
// s + at_j - at_i (packed), which maps the unwrapped positions onto the same image.
// O(n_sys_atoms * n_shifts) per centre: the convention for cells narrower than twice the
// cutoff, where systems are small.
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_pairs_general(const int32_t* __restrict__ species_idx, const int32_t* __restrict__ batch_index,
                const int32_t* __restrict__ sys_offsets, int n_atoms, const float* __restrict__ cells,
                const float* __restrict__ rec, const int32_t* __restrict__ grid, const float4* __restrict__ pos,
                float rc2, float rca2, int max_nbr, int max_nbr_ang, int32_t* __restrict__ nbr,
                int32_t* __restrict__ nbr_shift, int32_t* __restrict__ n_nbr, int32_t* __restrict__ nbr_ang,
                int32_t* __restrict__ n_nbr_ang, int32_t* __restrict__ status) {
  const int lane = fnl_lane();
  const int i = blockIdx.x * WAVES + (threadIdx.x >> 6);
  if (i >= n_atoms || species_idx[i] < 0) return;
  const int b = batch_index[i];
  const int32_t* g = grid + b * kGridWords;
  const int n0 = g[0], n1 = g[1], n2 = g[2];
  const int w1 = 2 * n1 + 1, w2 = 2 * n2 + 1;
  const int n_sh = (2 * n0 + 1) * w1 * w2;
  const int a0 = sys_offsets[b];
  const int total = (sys_offsets[b + 1] - a0) * n_sh;
  FnlCell cl;
  fnl_load_cell(cells, rec, b, cl);
  const float4 pi = pos[i];
  const int32_t ati = __float_as_int(pi.w);
  const int ai0 = (ati & 1023) - 512, ai1 = ((ati >> 10) & 1023) - 512, ai2 = ((ati >> 20) & 1023) - 512;
  int32_t* row = nbr + (size_t)i * max_nbr;
  int32_t* row_s = nbr_shift + (size_t)i * max_nbr;
  int32_t* row_a = nbr_ang + (size_t)i * max_nbr_ang;
  int count = 0, count_a = 0;
  bool range_ok = true;
  for (int base = 0; base < total; base += 64) {
    const int c = base + lane;
    const bool valid = c < total;
    const int jl = valid ? c / n_sh : 0;
    const int sidx = valid ? c - jl * n_sh : 0;
    const int j = a0 + jl;
    const int s0 = sidx / (w1 * w2) - n0;
    const int r12 = sidx % (w1 * w2);
    const int s1 = r12 / w2 - n1, s2 = r12 % w2 - n2;
    bool keep = false, keep_a = false;
    int32_t code = 0;
    if (valid && j != i && species_idx[j] >= 0) {  // no self images (the reference pairs i < j)
      const float4 pj = pos[j];
      const float fs0 = (float)s0, fs1 = (float)s1, fs2 = (float)s2;
      float vx, vy, vz;
      {
#pragma clang fp contract(off)
        vx = (pj.x - pi.x) + ((fs0 * cl.c[0] + fs1 * cl.c[3]) + fs2 * cl.c[6]);
        vy = (pj.y - pi.y) + ((fs0 * cl.c[1] + fs1 * cl.c[4]) + fs2 * cl.c[7]);
        vz = (pj.z - pi.z) + ((fs0 * cl.c[2] + fs1 * cl.c[5]) + fs2 * cl.c[8]);
      }
      const float d12 = fnl_d12(vx, vy, vz);
      keep = d12 < rc2;
      keep_a = d12 < rca2;
      if (keep) {
        const int32_t atj = __float_as_int(pj.w);
        const int t0 = s0 + ((atj & 1023) - 512) - ai0;
        const int t1 = s1 + (((atj >> 10) & 1023) - 512) - ai1;
        const int t2 = s2 + (((atj >> 20) & 1023) - 512) - ai2;
        range_ok = range_ok && fnl_shift_fits(t0, t1, t2);
        code = fnl_shift_pack(t0, t1, t2);
      }
    }
    const unsigned long long m = __ballot(keep);
    const unsigned long long ma = __ballot(keep_a);
    if (keep) {
      const int e = count + fnl_rank_below(m);
      if (e < max_nbr) {
        row[e] = j;
        row_s[e] = code;
      }
      if (keep_a) {
        const int ea = count_a + fnl_rank_below(ma);
        if (ea < max_nbr_ang && e < max_nbr) row_a[ea] = e;
      }
    }
    count += __popcll(m);
    count_a += __popcll(ma);
  }
  const bool bad_range = __ballot(!range_ok) != 0ull;
  if (lane == 0) {
    n_nbr[i] = count;
    if (n_nbr_ang) n_nbr_ang[i] = count_a;
    if (count > max_nbr || count_a > max_nbr_ang) {
      atomicOr(&status[FNL_STATUS_OVERFLOW], 1);
      atomicOr(&status[FNL_STATUS_STICKY], 1);
      atomicMax(&status[FNL_STATUS_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_MAX_NBR_ANG], count_a);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR_ANG], count_a);
    }
    if (bad_range) {
      atomicOr(&status[FNL_STATUS_OVERFLOW], 4);
      atomicOr(&status[FNL_STATUS_STICKY], 4);
    }
  }
}

// One thread per atom: cell id and slot (population count), the atom index
// staged at stage[cell * kCellCap + slot] while slot < kCellCap.  The last
// workgroup to finish (completion count `done`, left at zero) turns the
// populations into the exclusive prefix sum cell_start and records the
// largest population.
__global__ void __launch_bounds__(256) k_bin(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
                      const int32_t* __restrict__ batch_index, int n_atoms,
                      const float* __restrict__ rec, const int32_t* __restrict__ grid,
                      int32_t* __restrict__ atom_cell, int32_t* __restrict__ cell_count,
                      int32_t* __restrict__ n_nbr, int32_t* __restrict__ n_nbr_ang, int n_cells,
                      int32_t* __restrict__ status, int32_t* __restrict__ stage, unsigned* __restrict__ done,
                      int32_t* __restrict__ cell_start) {
  const int t = threadIdx.x;
  const int i = blockIdx.x * blockDim.x + t;
  if (i < n_atoms) {
    int cell = -1;
    if (species_idx[i] >= 0) {
      const int b = batch_index[i];
      const int32_t* g = grid + b * kGridWords;
      const int3 c = atom_cell3(coords, i, g, rec, b);
      cell = g[3] + (c.x * g[1] + c.y) * g[2] + c.z;
      if (cell < 0 || cell >= n_cells) {  // inconsistent grid capacities: drop and flag
        cell = -1;
        atomicOr(&status[FNL_STATUS_OVERFLOW], 2);
        atomicOr(&status[FNL_STATUS_STICKY], 2);
      }
    }
    atom_cell[i] = cell;
    if (cell < 0) {
      n_nbr[i] = 0;
      if (n_nbr_ang) n_nbr_ang[i] = 0;
    } else {
      const int slot = atomicAdd(&cell_count[cell], 1);
      if (slot < kCellCap) stage[(size_t)cell * kCellCap + slot] = i;
    }
  }
  // completion count: the last workgroup sees every population.  The populations are only ever
  // changed by device-scope atomics whose results this workgroup's threads waited for before the
  // barrier below, so they are final at the device's coherence point when the count is taken;
  // the last workgroup reads them there (agent-scope atomic loads), without cache-wide fences.
  __shared__ bool last;
  __shared__ int wsum[4];
  __shared__ int carry;
  __syncthreads();
  if (t == 0) {
    last = atomicAdd(done, 1u) == gridDim.x - 1;
    carry = 0;
  }
  __syncthreads();
  if (!last) return;
  // exclusive scan in chunks of 256 * kScanPer cells: each chunk's populations are loaded
  // coalesced (thread t holds cells base + 256 k + t) and requested before the previous chunk is
  // scanned, then transposed through LDS so that thread t scans kScanPer consecutive cells, and
  // the starts go back through LDS to coalesced stores
  constexpr int kScanPer = 16;
  constexpr int kChunk = 256 * kScanPer;
  __shared__ int chunk[kChunk];
  const int lane = t & 63, w = t >> 6;
  auto ld = [&](int q) {
    return q < n_cells ? __hip_atomic_load(cell_count + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  };
  int cmax = 0;
  int nxt[kScanPer];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) nxt[k] = ld(256 * k + t);
  for (int base = 0; base < n_cells; base += kChunk) {
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      chunk[256 * k + t] = nxt[k];
      cmax = max(cmax, nxt[k]);
    }
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) nxt[k] = ld(base + kChunk + 256 * k + t);
    __syncthreads();
    int mine[kScanPer];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k += 4) {
      const int4 q = *reinterpret_cast<const int4*>(&chunk[kScanPer * t + k]);
      mine[k] = q.x; mine[k + 1] = q.y; mine[k + 2] = q.z; mine[k + 3] = q.w;
      tot += (q.x + q.y) + (q.z + q.w);
    }
    const int incl = fnl_wave_incl_scan(tot);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = carry;
    for (int q = 0; q < w; ++q) before += wsum[q];
    int e = before + incl - tot;  // exclusive start of this thread's cells
#pragma unroll
    for (int k = 0; k < kScanPer; k += 4) {
      int4 q;
      q.x = e; e += mine[k];
      q.y = e; e += mine[k + 1];
      q.z = e; e += mine[k + 2];
      q.w = e; e += mine[k + 3];
      *reinterpret_cast<int4*>(&chunk[kScanPer * t + k]) = q;
    }
    __syncthreads();
    if (t == 255) carry = before + incl;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      const int idx = base + 256 * k + t;
      if (idx < n_cells) cell_start[idx] = chunk[256 * k + t];
    }
    __syncthreads();
  }
  if (t == 0) {
    cell_start[n_cells] = carry;
    *done = 0u;
  }
  cmax = fnl_wave_max_i(cmax);
  if (lane == 0 && cmax > 0) atomicMax(&status[FNL_STATUS_MAX_CELL], cmax);
}

// Cells along one axis of the neighbourhood c-depth..c+depth, deduplicated: a periodic axis of
// at most 2 depth + 1 cells contributes each of its cells once.  axis_count gives how many,
// axis_cell the q-th of them (q < axis_count), in a fixed order.
__device__ __forceinline__ int axis_count(int c, int n, int depth, bool pbc) {
  if (pbc) return min(n, 2 * depth + 1);
  return min(c + depth, n - 1) - max(c - depth, 0) + 1;
}

__device__ __forceinline__ int axis_cell(int c, int n, int depth, bool pbc, int q) {
  if (pbc) return n <= 2 * depth + 1 ? q : (c - depth + q + n) % n;
  return max(c - depth, 0) + q;
}

// z-runs of the neighbourhood along one (x, y) column of cells: cells cz-depth..cz+depth are
// consecutive in memory, so their atoms form one contiguous range of the cell-sorted arrays
// (two ranges when the periodic wrap splits it).  Returns the number of runs; run r (< 2) is
// the cells lo..hi of the column.
__device__ __forceinline__ int z_run(int cz, int nz, int depth, bool pbc, int r, int* lo, int* hi) {
  if (!pbc) {
    *lo = max(cz - depth, 0);
    *hi = min(cz + depth, nz - 1);
    return 1;
  }
  if (nz <= 2 * depth + 1) {  // every cell of the column
    *lo = 0;
    *hi = nz - 1;
    return 1;
  }
  if (cz - depth < 0) {
    *lo = r == 0 ? 0 : cz - depth + nz;
    *hi = r == 0 ? cz + depth : nz - 1;
    return 2;
  }
  if (cz + depth > nz - 1) {
    *lo = r == 0 ? cz - depth : 0;
    *hi = r == 0 ? nz - 1 : cz + depth - nz;
    return 2;
  }
  *lo = cz - depth;
  *hi = cz + depth;
  return 1;
}

// One wave per two cells (32 lanes each): the staged atoms of the cell ranked
// by index (rows come out in a fixed order), written to the cell's segment of
// the cell-sorted arrays (float4 position with the atom index in w, and the
// cell).  A cell holding more than kCellCap atoms is gathered instead by the
// whole wave, scanning its system's atoms in index order.  Each half also
// writes its cell's candidate ranges for k_pairs: range r < kMaxRanges =
// (first position in the cell-sorted arrays, length) of one z-run of one of
// the (2 depth_x + 1) x (2 depth_y + 1) neighbouring columns; unused ranges
// are (0, 0).
__global__ void k_cell_sort(int n_cells, const int32_t* __restrict__ cell_start,
                            const int32_t* __restrict__ stage, const float* __restrict__ coords,
                            const int32_t* __restrict__ atom_cell, const int32_t* __restrict__ grid,
                            const int32_t* __restrict__ sys_offsets, int n_sys,
                            int32_t* __restrict__ slot_cell,
                            float4* __restrict__ pos, int2* __restrict__ ranges) {
  static_assert(kCellCap <= 32, "one half-wave ranks a staged cell");
  const int lane = fnl_lane();
  const int h = lane >> 5, hl = lane & 31;
  const int c = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 2 + h;
  const bool live = c < n_cells;
  // the cell's population and staged atoms, requested before the range table is built
  int s = 0, cnt = 0, v = 0x7fffffff;
  if (live) {
    s = cell_start[c];
    cnt = cell_start[c + 1] - s;
    if (hl < min(cnt, kCellCap)) v = stage[(size_t)c * kCellCap + hl];
  }
  // the system whose cell range holds c: largest b with first cell <= c
  int sys = 0;
  if (live) {
    int hi = n_sys - 1;
    while (sys < hi) {
      const int mid = (sys + hi + 1) >> 1;
      if (grid[mid * kGridWords + 3] <= c) sys = mid; else hi = mid - 1;
    }
    const int32_t* g = grid + sys * kGridWords;
    const int nx = g[0], ny = g[1], nz = g[2], first = g[3];
    const bool pbc = g[10] != 0;
    const int dx = g[11] & 15, dy = (g[11] >> 4) & 15, dz = (g[11] >> 8) & 15;
    const int cell = c - first;
    const int cz = cell % nz, cy = (cell / nz) % ny, cx = cell / (ny * nz);
    int nox = 0, noy = 0, nzr = 1, zlo, zhi;
    if (cell < nx * ny * nz) {  // cells past the system's grid (capacity slack) have no atoms
      nox = axis_count(cx, nx, dx, pbc);
      noy = axis_count(cy, ny, dy, pbc);
      nzr = z_run(cz, nz, dz, pbc, 0, &zlo, &zhi);
    }
    for (int e = hl; e < kMaxRanges; e += 32) {
      int2 rr = make_int2(0, 0);
      if (e < nox * noy * nzr) {
        const int zr = e % nzr, xy = e / nzr;
        const int qx = xy / noy, qy = xy % noy;
        z_run(cz, nz, dz, pbc, zr, &zlo, &zhi);
        const int col = first + (axis_cell(cx, nx, dx, pbc, qx) * ny + axis_cell(cy, ny, dy, pbc, qy)) * nz;
        rr.x = cell_start[col + zlo];
        rr.y = cell_start[col + zhi + 1] - rr.x;
      }
      ranges[(size_t)c * kMaxRanges + e] = rr;
    }
  }
  // staged cells: rank among the half's atoms (lanes past the population hold INT_MAX)
  const int qn = cnt <= kCellCap ? cnt : 0;
  const int qmax = max(__shfl(qn, 0, 64), __shfl(qn, 32, 64));
  int rank = 0;
  for (int q = 0; q < qmax; ++q) rank += __shfl(v, (h << 5) + q, 64) < v ? 1 : 0;
  if (hl < qn) {
    slot_cell[s + rank] = c;
    pos[s + rank] = make_float4(coords[3 * v], coords[3 * v + 1], coords[3 * v + 2], __int_as_float(v));
  }
  // crowded cells: the whole wave gathers the cell's atoms from its system's range
  for (int hh = 0; hh < 2; ++hh) {
    const int cc = __shfl(c, hh * 32, 64), ccnt = __shfl(cnt, hh * 32, 64);
    if (cc >= n_cells || ccnt <= kCellCap) continue;  // wave-uniform
    const int csys = __shfl(sys, hh * 32, 64), cs = __shfl(s, hh * 32, 64);
    const int a_end = sys_offsets[csys + 1];
    int k = 0;
    for (int a0 = sys_offsets[csys]; a0 < a_end; a0 += 64) {
      const int a = a0 + lane;
      const bool mine = a < a_end && atom_cell[a] == cc;
      const unsigned long long m = __ballot(mine);
      if (mine) {
        const int r = k + fnl_rank_below(m);
        slot_cell[cs + r] = cc;
        pos[cs + r] = make_float4(coords[3 * a], coords[3 * a + 1], coords[3 * a + 2], __int_as_float(a));
      }
      k += __popcll(m);
    }
  }
}

// inclusive prefix sum and maximum over each half-wave (32 lanes)
__device__ __forceinline__ int half_incl_scan(int v) {
  const int hl = fnl_lane() & 31;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    const int t = __shfl_up(v, o, 32);
    if (hl >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int half_max(int v) {
#pragma unroll
  for (int o = 16; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 32));
  return v;
}

constexpr int kCmapHalf = 512;  // candidates per centre mapped through LDS (more: the search below)

// One half-wave (32 lanes) per centre, two centres per wave (centres in cell order), so that a
// wave has two centres' load chains in flight.  The candidates are the atoms of up to 50
// contiguous ranges (columns x z-runs) of the cell-sorted arrays, read from the centre cell's
// range table (ranges hl and 32 + hl in lane hl), concatenated into one list (a half-wave
// prefix sum over the range lengths; candidate v is mapped through LDS, or beyond kCmapHalf
// candidates found by a binary search over the range offsets), so every pass tests 32
// candidates per centre, and the next pass's positions are loaded before the current pass is
// compacted.  Kept pairs are compacted with a ballot and a prefix popcount over the half's 32
// bits into the radial row and, for d12 < rca^2, the angular row, in candidate order.  Only
// rows that overflow update the maxima in the status words (a same-address atomic per centre
// would serialise at the device's coherence point).
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_pairs(const int32_t* __restrict__ batch_index, int n_binned_cap, const float* __restrict__ cells,
        const float* __restrict__ rec, const int32_t* __restrict__ grid,
        const int32_t* __restrict__ cell_start, const float4* __restrict__ pos,
        const int32_t* __restrict__ slot_cell, const int2* __restrict__ ranges, float rc2, float rca2,
        int max_nbr, int max_nbr_ang, int32_t* __restrict__ nbr, int32_t* __restrict__ n_nbr,
        int32_t* __restrict__ nbr_ang, int32_t* __restrict__ n_nbr_ang, int32_t* __restrict__ status,
        int n_cells) {
  static_assert(kMaxRanges <= 64, "two ranges per lane of a half-wave");
  const int w = threadIdx.x >> 6;
  const int lane = fnl_lane();
  const int h = lane >> 5, hl = lane & 31;
  const int slot = (blockIdx.x * WAVES + w) * 2 + h;
  const int n_binned = min(cell_start[n_cells], n_binned_cap);
  if (slot - h >= n_binned) return;  // both centres of the wave past the binned atoms
  const bool live = slot < n_binned;
  // the centre (its cell-sorted position and atom index) and its cell, side by side
  float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int cell = 0;
  if (live) {
    pi = pos[slot];
    cell = slot_cell[slot];
  }
  const int i = __float_as_int(pi.w);
  // ranges hl (a) and 32 + hl (b) of the centre's cell: start and length (k_cell_sort)
  int lo_a = 0, len_a = 0, lo_b = 0, len_b = 0;
  if (live) {
    const int2 ra = ranges[(size_t)cell * kMaxRanges + hl];
    lo_a = ra.x;
    len_a = ra.y;
    if (32 + hl < kMaxRanges) {
      const int2 rb = ranges[(size_t)cell * kMaxRanges + 32 + hl];
      lo_b = rb.x;
      len_b = rb.y;
    }
  }
  int b = 0;
  bool pbc = false;
  if (live) {
    b = batch_index[i];
    pbc = grid[b * kGridWords + 10] != 0;
  }
  const int inc_a = half_incl_scan(len_a);
  const int tot_a = __shfl(inc_a, (h << 5) + 31, 64);
  const int inc_b = half_incl_scan(len_b) + tot_a;
  const int off_a = inc_a - len_a, off_b = inc_b - len_b;
  const int total = __shfl(inc_b, (h << 5) + 31, 64);  // candidates of this half's centre
  // position in the cell-sorted arrays of candidate v < total: a per-centre LDS map (one LDS
  // write per candidate instead of a search per candidate and pass), filled by the range lanes
  // in parallel (as many steps as the longest range); beyond kCmapHalf candidates, v + shift of
  // the last range that starts at or before v (empty ranges start where the next one does),
  // found by a binary search over the range offsets in the same LDS
  __shared__ int cmap_all[WAVES][2][kCmapHalf];
  int* cmap = cmap_all[w][h];
  int2* rtab = (int2*)cmap;  // [64] when the map is not used
  const bool use_map = total <= kCmapHalf;
  if (use_map) {
    const int longest = half_max(max(len_a, len_b));
    for (int k = 0; k < longest; ++k) {
      if (k < len_a) cmap[off_a + k] = lo_a + k;
      if (k < len_b) cmap[off_b + k] = lo_b + k;
    }
  } else {
    rtab[hl] = make_int2(off_a, lo_a - off_a);  // lanes past the ranges: (total, 0)
    rtab[32 + hl] = make_int2(off_b, lo_b - off_b);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  auto candidate = [&](int v) {
    if (use_map) return cmap[min(v, kCmapHalf - 1)];
    int lo = 0, hi = kMaxRanges - 1;  // rtab[0].x = 0 <= v
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rtab[mid].x <= v) lo = mid; else hi = mid - 1;
    }
    return v + rtab[lo].y;
  };

  FnlCell cl;
  if (pbc) fnl_load_cell(cells, rec, b, cl);
  const bool diag = pbc && fnl_cell_diagonal(cl);  // orthorhombic box: 3 products per candidate
  const float xi = pi.x, yi = pi.y, zi = pi.z;  // the coordinates as given (k_cell_sort copies them)
  int count = 0, count_a = 0;
  int32_t* row = nbr + (size_t)i * max_nbr;
  int32_t* row_a = nbr_ang + (size_t)i * max_nbr_ang;
  const unsigned below = (1u << hl) - 1u;
  const int passes = max(__shfl(total, 0, 64), __shfl(total, 32, 64));  // wave-uniform
  float4 next = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(i));
  if (hl < total) next = pos[candidate(hl)];
  for (int base = 0; base < passes; base += 32) {
    const float4 pj = next;
    const bool valid = base + hl < total;
    if (base + 32 + hl < total) next = pos[candidate(base + 32 + hl)];
    bool keep = false, keep_a = false;
    const int j = __float_as_int(pj.w);
    if (valid && j != i) {
      float vx = pj.x - xi, vy = pj.y - yi, vz = pj.z - zi;
      if (diag)
        fnl_min_image_diag(vx, vy, vz, cl);
      else if (pbc)
        fnl_min_image(vx, vy, vz, cl);
      const float d12 = fnl_d12(vx, vy, vz);
      keep = d12 < rc2;
      keep_a = d12 < rca2;
    }
    const unsigned m = (unsigned)(__ballot(keep) >> (32 * h));
    const unsigned ma = (unsigned)(__ballot(keep_a) >> (32 * h));
    if (keep) {
      const int e = count + __popc(m & below);
      if (e < max_nbr) row[e] = j;
      if (keep_a) {
        const int ea = count_a + __popc(ma & below);
        if (ea < max_nbr_ang && e < max_nbr) row_a[ea] = e;
      }
    }
    count += __popc(m);
    count_a += __popc(ma);
  }
  if (live && hl == 0) {
    n_nbr[i] = count;
    if (n_nbr_ang) n_nbr_ang[i] = count_a;
    if (count > max_nbr || count_a > max_nbr_ang) {  // rare: the maxima only matter on overflow
      atomicOr(&status[FNL_STATUS_OVERFLOW], 1);
      atomicOr(&status[FNL_STATUS_STICKY], 1);
      atomicMax(&status[FNL_STATUS_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_MAX_NBR_ANG], count_a);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR_ANG], count_a);
    }
  }
}

// Status words of one list update: the overflow flag and the row maxima start at zero (the
// sticky word persists).
__global__ void k_status_clear(int32_t* __restrict__ status) {
  const int t = threadIdx.x;
  if (t < FNL_STATUS_STICKY) status[t] = 0;
}

// Skin update (GraphGenerator.update_skin, preprocessing.py:640-696, with the
// GraphFilter re-filter :958-960): one half-wave per centre (two centres' load
// chains per wave) reads the centre's row of the list built with cutoff + skin,
// 32 slots a pass, recomputes d12 at the current positions with the same
// float32 arithmetic as the full build, and compacts the slots with
// d12 < rc^2 (radial row) and d12 < rca^2 (angular row) in row order, as
// k_pairs does.  For the minimum-image convention the image chosen at the
// build is the image the current positions select for every pair still
// inside the cutoff (rc below half the lattice-plane spacing), so the
// recomputed image reproduces the reference's stored pbc_shifts; for general
// PBC the stored shift of each slot is applied to the unwrapped positions
// (the reference's update_skin formula) and carried to the output rows.  The
// slot indices are loaded two passes ahead and the neighbours' positions one
// pass ahead of the test.
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES)
k_skin_filter(const float* __restrict__ coords, const int32_t* __restrict__ batch_index, int n_atoms,
              int pbc_mode, const float* __restrict__ cells, const float* __restrict__ rec,
              const int32_t* __restrict__ nbr_skin, const int32_t* __restrict__ skin_shift,
              const int32_t* __restrict__ n_skin, int max_skin, float rc2, float rca2, int max_nbr,
              int max_nbr_ang, int32_t* __restrict__ nbr, int32_t* __restrict__ nbr_shift,
              int32_t* __restrict__ n_nbr, int32_t* __restrict__ nbr_ang, int32_t* __restrict__ n_nbr_ang,
              int32_t* __restrict__ status) {
  const int lane = fnl_lane();
  const int h = lane >> 5, hl = lane & 31;
  const int i = (blockIdx.x * WAVES + (threadIdx.x >> 6)) * 2 + h;
  if (i - h >= n_atoms) return;  // both centres of the wave past the atoms
  const bool live = i < n_atoms;
  const int ic = live ? i : 0;  // the centre, or atom 0 for a half without one (never written)
  const int ns = live ? n_skin[i] : 0;
  const int n = min(ns, max_skin);
  const int32_t* srow = nbr_skin + (size_t)ic * max_skin;
  FnlCell cl;
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  const bool gen = pbc_mode == FNL_PBC_GENERAL;
  if (pbc || gen) fnl_load_cell(cells, rec, batch_index[ic], cl);
  const float xi = coords[3 * ic], yi = coords[3 * ic + 1], zi = coords[3 * ic + 2];
  int32_t* row = nbr + (size_t)ic * max_nbr;
  int32_t* row_a = nbr_ang + (size_t)ic * max_nbr_ang;
  const unsigned below = (1u << hl) - 1u;
  int count = 0, count_a = 0;
  // slot v's neighbour (the centre itself past the row end: a harmless load, never kept)
  auto slot_atom = [&](int v) { return v < n ? srow[v] : ic; };
  int j0 = slot_atom(hl);
  float x0 = coords[3 * j0], y0 = coords[3 * j0 + 1], z0 = coords[3 * j0 + 2];
  int j1 = slot_atom(32 + hl);
  const int passes = max(__shfl(n, 0, 64), __shfl(n, 32, 64));  // wave-uniform
  for (int base = 0; base < passes; base += 32) {
    const bool valid = base + hl < n;
    const int j = j0;
    float vx = x0 - xi, vy = y0 - yi, vz = z0 - zi;
    j0 = j1;
    x0 = coords[3 * j1];
    y0 = coords[3 * j1 + 1];
    z0 = coords[3 * j1 + 2];
    j1 = slot_atom(base + 64 + hl);
    int32_t code = 0;
    if (gen) {
      code = valid ? skin_shift[(size_t)ic * max_skin + base + hl] : 0;
      fnl_apply_shift(vx, vy, vz, cl, code);
    } else if (pbc) {
      fnl_min_image(vx, vy, vz, cl);
    }
    const float d12 = fnl_d12(vx, vy, vz);
    const bool keep = valid && d12 < rc2;
    const bool keep_a = valid && d12 < rca2;
    const unsigned m = (unsigned)(__ballot(keep) >> (32 * h));
    const unsigned ma = (unsigned)(__ballot(keep_a) >> (32 * h));
    if (keep) {
      const int e = count + __popc(m & below);
      if (e < max_nbr) {
        row[e] = j;
        if (gen) nbr_shift[(size_t)ic * max_nbr + e] = code;
      }
      if (keep_a) {
        const int ea = count_a + __popc(ma & below);
        if (ea < max_nbr_ang && e < max_nbr) row_a[ea] = e;
      }
    }
    count += __popc(m);
    count_a += __popc(ma);
  }
  if (live && hl == 0) {
    n_nbr[i] = count;
    n_nbr_ang[i] = count_a;
    if (count > max_nbr || count_a > max_nbr_ang || ns > max_skin) {
      atomicOr(&status[FNL_STATUS_OVERFLOW], 1);
      atomicOr(&status[FNL_STATUS_STICKY], 1);
      atomicMax(&status[FNL_STATUS_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_MAX_NBR_ANG], count_a);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR], count);
      atomicMax(&status[FNL_STATUS_STICKY_MAX_NBR_ANG], count_a);
    }
  }
}

// Flat edge list from the rows.  d12 is recomputed as the build computed it: minimum image
// from the positions, or, for general PBC, from the wrapped positions and the cell shift
// (total shift - at_j + at_i), so that it equals the reference's exported d12 bit for bit.
__global__ void k_export(const float* __restrict__ coords, const int32_t* __restrict__ batch_index,
                         int n_atoms, int pbc_mode, const float* __restrict__ cells,
                         const float* __restrict__ rec, const int32_t* __restrict__ nbr,
                         const int32_t* __restrict__ nbr_shift, const int32_t* __restrict__ n_nbr, int max_nbr,
                         const int32_t* __restrict__ row_start, int32_t* __restrict__ edge_src,
                         int32_t* __restrict__ edge_dst, float* __restrict__ d12,
                         float* __restrict__ shifts) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = fnl_lane();
  if (wave >= n_atoms) return;
  const int i = wave;
  const int n = min(n_nbr[i], max_nbr);
  const int b = batch_index[i];
  FnlCell cl;
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  const bool gen = pbc_mode == FNL_PBC_GENERAL;
  if (pbc || gen) fnl_load_cell(cells, rec, b, cl);
  const int o = row_start[i];
  for (int e = lane; e < n; e += 64) {
    const int j = nbr[(size_t)i * max_nbr + e];
    float vx, vy, vz;
    float s[3] = {0.f, 0.f, 0.f};
    if (gen) {
      float xwi[3], xwj[3];
      int ati[3], atj[3];
      fnl_wrap(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2], cl, xwi, ati);
      fnl_wrap(coords[3 * j], coords[3 * j + 1], coords[3 * j + 2], cl, xwj, atj);
      fnl_shift_unpack(nbr_shift[(size_t)i * max_nbr + e], s[0], s[1], s[2]);
      vx = xwj[0] - xwi[0];
      vy = xwj[1] - xwi[1];
      vz = xwj[2] - xwi[2];
      fnl_apply_shift(vx, vy, vz, cl,
                      fnl_shift_pack((int)s[0] - atj[0] + ati[0], (int)s[1] - atj[1] + ati[1],
                                     (int)s[2] - atj[2] + ati[2]));
    } else {
      vx = coords[3 * j] - coords[3 * i];
      vy = coords[3 * j + 1] - coords[3 * i + 1];
      vz = coords[3 * j + 2] - coords[3 * i + 2];
      if (pbc) fnl_min_image(vx, vy, vz, cl, s);
    }
    edge_src[o + e] = i;
    edge_dst[o + e] = j;
    d12[o + e] = fnl_d12(vx, vy, vz);
    if (shifts) {
      shifts[3 * (o + e)] = s[0];
      shifts[3 * (o + e) + 1] = s[1];
      shifts[3 * (o + e) + 2] = s[2];
    }
  }
}

}  // namespace

// Scratch of the allocation-free build (fnl_nblist_build_noalloc): a file-scope device array, one
// copy per GPU at its own address.  Host code never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address): the address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept in scratch_ptrs.
static __device__ __align__(256) unsigned char scratch[FNL_NBLIST_SCRATCH_BYTES];

namespace {

int scratch_address(void** out) {
  static std::mutex mu;
  static std::vector<void*> scratch_ptrs;  // per device; nullptr until looked up
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return FNL_ERR_HIP;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev >= scratch_ptrs.size()) scratch_ptrs.resize((size_t)dev + 1, nullptr);
  if (!scratch_ptrs[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FNL_ERR_HIP;
    scratch_ptrs[dev] = p;
  }
  *out = scratch_ptrs[dev];
  return FNL_OK;
}

}  // namespace

extern "C" int fnl_version(void) { return 10000; }

extern "C" const char* fnl_strerror(int code) {
  switch (code) {
    case FNL_OK: return "ok";
    case FNL_ERR_INVALID: return "invalid argument";
    case FNL_ERR_HIP: return "HIP launch error";
    case FNL_ERR_UNSUPPORTED: return "unsupported configuration";
    case FNL_ERR_SCRATCH_TOO_SMALL:
      return "the cell-list temporaries exceed the library's static scratch (FNL_NBLIST_SCRATCH_BYTES)";
    default: return "unknown error";
  }
}

extern "C" int fnl_fill_i32(int32_t* p, int32_t value, size_t count, fnl_stream_t stream) {
  if (count == 0) return FNL_OK;
  if (!p) return FNL_ERR_INVALID;
  const int blocks = (int)std::min<size_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, (hipStream_t)stream, p, value, count);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_fill_f32(float* p, float value, size_t count, fnl_stream_t stream) {
  if (count == 0) return FNL_OK;
  if (!p) return FNL_ERR_INVALID;
  const int blocks = (int)std::min<size_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, (hipStream_t)stream, p, value, count);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" size_t fnl_nblist_workspace(int32_t n_atoms, int32_t n_sys, int32_t n_cells) {
  if (n_atoms < 0 || n_sys < 1 || n_cells < 1) return 0;
  return ws_layout(n_atoms, n_sys, n_cells, nullptr, nullptr);
}

extern "C" int fnl_nblist_build(const float* coords, const int32_t* species_idx,
                                const int32_t* batch_index, const int32_t* sys_offsets,
                                int32_t n_atoms, int32_t n_sys, int32_t pbc_mode,
                                const float* cells, const float* reciprocal_cells,
                                const int32_t* grid, int32_t n_cells, double cutoff,
                                double cutoff_ang, int32_t max_nbr, int32_t max_nbr_ang,
                                void* workspace, size_t workspace_bytes, int32_t* nbr, int32_t* nbr_shift,
                                int32_t* n_nbr, int32_t* nbr_ang, int32_t* n_nbr_ang,
                                int32_t* status, fnl_stream_t stream) {
  if (n_atoms < 0 || n_sys < 1 || n_cells < 1 || max_nbr < 1 || max_nbr_ang < 1) return FNL_ERR_INVALID;
  if (!(cutoff > 0.0) || cutoff_ang > cutoff) return FNL_ERR_INVALID;
  if (pbc_mode != FNL_PBC_NONE && pbc_mode != FNL_PBC_MINIMUM_IMAGE && pbc_mode != FNL_PBC_GENERAL)
    return FNL_ERR_UNSUPPORTED;
  if (pbc_mode != FNL_PBC_NONE && (!cells || !reciprocal_cells)) return FNL_ERR_INVALID;
  if (pbc_mode == FNL_PBC_GENERAL && n_atoms > 0 && !nbr_shift) return FNL_ERR_INVALID;
  if (!status || !grid || !sys_offsets) return FNL_ERR_INVALID;
  if (n_atoms > 0 && (!coords || !species_idx || !batch_index || !nbr || !n_nbr)) return FNL_ERR_INVALID;
  if ((nbr_ang == nullptr) != (n_nbr_ang == nullptr)) return FNL_ERR_INVALID;
  if (!workspace || workspace_bytes < fnl_nblist_workspace(n_atoms, n_sys, n_cells)) return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  NblistWs ws;
  ws_layout(n_atoms, n_sys, n_cells, (char*)workspace, &ws);
  // thresholds: the float32 rounding of the squared double cutoff, as the reference's
  // float32 d12 < python-float rc**2 comparison (preprocessing.py:266, :557)
  const float rc2 = (float)(cutoff * cutoff);
  // without angular rows no slot passes the angular test
  const float rca2 = nbr_ang ? (float)(cutoff_ang * cutoff_ang) : -1.0f;

  // one workgroup per system, at least one per 2,048 cell counters to clear (up to 64)
  const int grid_blocks = std::max(n_sys, std::min((n_cells + 2047) / 2048, 64));
  hipLaunchKernelGGL(k_grid, dim3(grid_blocks), dim3(256), 0, s, coords, species_idx, sys_offsets, pbc_mode,
                     reciprocal_cells, grid, (float)cutoff, ws.grid, status, ws.cell_count, ws.done, n_cells, n_sys);
  FNL_LAUNCH_CHECK();
  if (n_atoms == 0) return FNL_OK;
  constexpr int W = 4;
  if (pbc_mode == FNL_PBC_GENERAL) {
    hipLaunchKernelGGL(k_wrap, dim3(fnl_blocks(n_atoms, 256)), dim3(256), 0, s, coords, species_idx, batch_index,
                       n_atoms, cells, reciprocal_cells, ws.pos, n_nbr, n_nbr_ang, status);
    FNL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pairs_general<W>, dim3(fnl_blocks(n_atoms, W)), dim3(64 * W), 0, s, species_idx,
                       batch_index, sys_offsets, n_atoms, cells, reciprocal_cells, ws.grid, ws.pos, rc2, rca2,
                       max_nbr, max_nbr_ang, nbr, nbr_shift, n_nbr, nbr_ang, n_nbr_ang, status);
    FNL_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(k_bin, dim3(fnl_blocks(n_atoms, 256)), dim3(256), 0, s, coords, species_idx,
                       batch_index, n_atoms, reciprocal_cells, ws.grid, ws.atom_cell, ws.cell_count,
                       n_nbr, n_nbr_ang, n_cells, status, ws.stage, ws.done, ws.cell_start);
    FNL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cell_sort, dim3(fnl_blocks((long long)(n_cells + 1) / 2 * 64, 256)), dim3(256), 0, s, n_cells,
                       ws.cell_start, ws.stage, coords, ws.atom_cell, ws.grid, sys_offsets, n_sys,
                       ws.slot_cell, ws.pos, ws.ranges);
    FNL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pairs<W>, dim3(fnl_blocks(n_atoms, 2 * W)), dim3(64 * W), 0, s, batch_index, n_atoms,
                       cells, reciprocal_cells, ws.grid, ws.cell_start, ws.pos, ws.slot_cell, ws.ranges, rc2,
                       rca2, max_nbr, max_nbr_ang, nbr, n_nbr, nbr_ang, n_nbr_ang, status, n_cells);
    FNL_LAUNCH_CHECK();
  }
  return FNL_OK;
}

extern "C" int fnl_nblist_build_noalloc(const float* coords, const int32_t* species_idx,
                                        const int32_t* batch_index, const int32_t* sys_offsets,
                                        int32_t n_atoms, int32_t n_sys, int32_t pbc_mode,
                                        const float* cells, const float* reciprocal_cells,
                                        const int32_t* grid, int32_t n_cells, double cutoff,
                                        double cutoff_ang, int32_t max_nbr, int32_t max_nbr_ang,
                                        int32_t* nbr, int32_t* nbr_shift, int32_t* n_nbr,
                                        int32_t* nbr_ang, int32_t* n_nbr_ang, int32_t* status,
                                        fnl_stream_t stream) {
  if (n_atoms < 0 || n_sys < 1 || n_cells < 1) return FNL_ERR_INVALID;
  const size_t need = fnl_nblist_workspace(n_atoms, n_sys, n_cells);
  if (need == 0) return FNL_ERR_INVALID;
  if (need > FNL_NBLIST_SCRATCH_BYTES) return FNL_ERR_SCRATCH_TOO_SMALL;
  void* ws = nullptr;
  if (const int rc = scratch_address(&ws)) return rc;
  // the build's first kernel (k_grid) clears the completion count and the cell counters in the
  // workspace; everything else in it is written before it is read
  return fnl_nblist_build(coords, species_idx, batch_index, sys_offsets, n_atoms, n_sys, pbc_mode, cells,
                          reciprocal_cells, grid, n_cells, cutoff, cutoff_ang, max_nbr, max_nbr_ang, ws,
                          FNL_NBLIST_SCRATCH_BYTES, nbr, nbr_shift, n_nbr, nbr_ang, n_nbr_ang, status, stream);
}

extern "C" int fnl_nblist_skin_update(const float* coords, const int32_t* batch_index, int32_t n_atoms,
                                      int32_t pbc_mode, const float* cells, const float* reciprocal_cells,
                                      const int32_t* nbr_skin, const int32_t* skin_shift, const int32_t* n_skin,
                                      int32_t max_skin, double cutoff, double cutoff_ang, int32_t max_nbr,
                                      int32_t max_nbr_ang, int32_t* nbr, int32_t* nbr_shift, int32_t* n_nbr,
                                      int32_t* nbr_ang, int32_t* n_nbr_ang, int32_t* status, fnl_stream_t stream) {
  if (n_atoms < 0 || max_skin < 1 || max_nbr < 1 || max_nbr_ang < 1) return FNL_ERR_INVALID;
  if (!(cutoff > 0.0) || cutoff_ang > cutoff || !status) return FNL_ERR_INVALID;
  if (pbc_mode != FNL_PBC_NONE && pbc_mode != FNL_PBC_MINIMUM_IMAGE && pbc_mode != FNL_PBC_GENERAL)
    return FNL_ERR_UNSUPPORTED;
  if (pbc_mode != FNL_PBC_NONE && (!cells || !reciprocal_cells)) return FNL_ERR_INVALID;
  if (pbc_mode == FNL_PBC_GENERAL && n_atoms > 0 && (!skin_shift || !nbr_shift)) return FNL_ERR_INVALID;
  if (n_atoms > 0 && (!coords || !batch_index || !nbr_skin || !n_skin || !nbr || !n_nbr || !nbr_ang || !n_nbr_ang))
    return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const float rc2 = (float)(cutoff * cutoff);
  const float rca2 = (float)(cutoff_ang * cutoff_ang);
  hipLaunchKernelGGL(k_status_clear, dim3(1), dim3(64), 0, s, status);
  FNL_LAUNCH_CHECK();
  if (n_atoms == 0) return FNL_OK;
  constexpr int W = 4;
  hipLaunchKernelGGL(k_skin_filter<W>, dim3(fnl_blocks(n_atoms, 2 * W)), dim3(64 * W), 0, s, coords, batch_index,
                     n_atoms, pbc_mode, cells, reciprocal_cells, nbr_skin, skin_shift, n_skin, max_skin, rc2,
                     rca2, max_nbr, max_nbr_ang, nbr, nbr_shift, n_nbr, nbr_ang, n_nbr_ang, status);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_nblist_export(const float* coords, const int32_t* batch_index, int32_t n_atoms,
                                 int32_t pbc_mode, const float* cells, const float* reciprocal_cells,
                                 const int32_t* nbr, const int32_t* nbr_shift, const int32_t* n_nbr,
                                 int32_t max_nbr, const int32_t* row_start, int32_t* edge_src,
                                 int32_t* edge_dst, float* d12, float* pbc_shifts, fnl_stream_t stream) {
  if (n_atoms < 0 || max_nbr < 1) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!coords || !batch_index || !nbr || !n_nbr || !row_start || !edge_src || !edge_dst || !d12)
    return FNL_ERR_INVALID;
  if (pbc_mode != FNL_PBC_NONE && pbc_mode != FNL_PBC_MINIMUM_IMAGE && pbc_mode != FNL_PBC_GENERAL)
    return FNL_ERR_UNSUPPORTED;
  if (pbc_mode != FNL_PBC_NONE && (!cells || !reciprocal_cells)) return FNL_ERR_INVALID;
  if (pbc_mode == FNL_PBC_GENERAL && !nbr_shift) return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_export, dim3(fnl_blocks((long long)n_atoms * 64, 256)), dim3(256), 0,
                     (hipStream_t)stream, coords, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                     nbr, nbr_shift, n_nbr, max_nbr, row_start, edge_src, edge_dst, d12, pbc_shifts);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
