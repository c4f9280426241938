// Serving pool: the per-atom and per-system tables of a continuous batch built on the device from
// one row per request, and the two launches that open and close a serving iteration (BAOAB of the
// MD requests; bookkeeping and the packed result record).  Every grid is sized by the pool's
// capacity and reads the live sizes from device words, so the launches of an iteration can be
// replayed from a HIP graph while requests come and go.
//
// References: the batched-systems convention (natoms / batch_index, models/preprocessing.py:
// 1763-1791), the per-species atom index (SpeciesIndexer, preprocessing.py:1583-1651), BAOAB with
// the LGV O-step (md/integrate.py:286-294, md/thermostats.py:108-120).

#include "fnl_common.h"

namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanWaves = kScanThreads / 64;

// Exclusive prefix sum of one int per thread over a workgroup of kScanThreads threads; *total is
// the workgroup's sum.  Every thread must call it (it synchronises the workgroup).
__device__ int block_excl_scan(int v, int* lds, int* total) {
  const int lane = fnl_lane(), w = threadIdx.x >> 6;
  const int inc = fnl_wave_incl_scan(v);
  if (lane == 63) lds[w] = inc;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int s = threadIdx.x < kScanWaves ? lds[threadIdx.x] : 0;
    const int si = fnl_wave_incl_scan(s);
    if (threadIdx.x < kScanWaves) lds[threadIdx.x] = si - s;
    if (threadIdx.x == 63) lds[kScanWaves] = si;
  }
  __syncthreads();
  const int r = lds[w] + inc - v;
  *total = lds[kScanWaves];
  __syncthreads();
  return r;
}

// side of the open-boundary cell-grid capacity: the smallest k with k^3 >= n (n >= 1)
__device__ __forceinline__ int cube_side(int n) {
  int k = 1;
  while (k * k * k < n) ++k;
  return k;
}

// Per-system tables: offsets (rows, reduction segments, cells), grid capacities and the
// per-request serving state.  One workgroup; each thread takes a contiguous run of systems.
__global__ void __launch_bounds__(kScanThreads)
k_pool_systems(const fnl_pool p, const fnl_pool_request* __restrict__ req, int n_req) {
  __shared__ int lds[kScanWaves + 1];
  const int S = p.max_sys;
  const int per = (S + kScanThreads - 1) / kScanThreads;
  const int b0 = min((int)threadIdx.x * per, S), b1 = min(b0 + per, S);
  long long sn = 0;
  int sbad = 0;
  for (int b = b0; b < min(b1, n_req); ++b) {
    const int n = req[b].n_atoms;
    sn += n;
    sbad += (n < 1 || n > p.max_atoms) ? 1 : 0;
  }
  int tn, tbad;
  int on = block_excl_scan((int)min(sn, (long long)p.max_atoms + 1), lds, &tn);
  block_excl_scan(sbad, lds, &tbad);
  // a refused layout leaves an empty pool (every system without atoms) and the error word set
  const bool bad = tbad > 0 || tn > p.max_atoms || n_req > S;
  int sseg = 0, scell = 0;
  for (int b = b0; b < b1; ++b) {
    const int n = (!bad && b < n_req) ? req[b].n_atoms : 0;
    const int k = cube_side(max(n, 1));
    sseg += (n + p.seg_atoms - 1) / p.seg_atoms;
    scell += k * k * k;
  }
  int tseg, tcell;
  int oseg = block_excl_scan(sseg, lds, &tseg);
  int ocell = block_excl_scan(scell, lds, &tcell);
  if (bad) on = 0;
  for (int b = b0; b < b1; ++b) {
    const bool live = !bad && b < n_req;
    const int n = live ? req[b].n_atoms : 0;
    const int k = cube_side(max(n, 1));
    p.sys_offsets[b] = on;
    p.seg_offsets[b] = oseg;
    p.grid_cap[4 * b + 0] = k;
    p.grid_cap[4 * b + 1] = k;
    p.grid_cap[4 * b + 2] = k;
    p.grid_cap[4 * b + 3] = ocell;
    on += n;
    oseg += (n + p.seg_atoms - 1) / p.seg_atoms;
    ocell += k * k * k;
    p.sys_left[b] = live ? max(req[b].md_left, 0) : 0;
    p.sys_eval[b] = live ? (req[b].evaluated != 0) : 0;
    p.sys_step[b] = live ? req[b].step : 0ull;
    p.sys_key[b] = live ? req[b].noise_key : 0u;
    p.sys_fin[b] = -1;
    p.sys_out_row[b] = 0;
  }
  if (threadIdx.x == 0) {
    p.sys_offsets[S] = bad ? 0 : tn;
    p.seg_offsets[S] = tseg;
    p.counts[FNL_POOL_N_ATOMS] = bad ? 0 : tn;
    p.counts[FNL_POOL_N_SYS] = bad ? 0 : n_req;
    p.counts[FNL_POOL_N_TILES] = 0;
    p.counts[FNL_POOL_ERROR] = bad ? 1 : 0;
  }
}

// Rows: continuing requests from the saved previous layout, new ones from the staging arrays;
// species index, system index, integrator factors; rows beyond the requests become padding.
__global__ void __launch_bounds__(256)
k_pool_rows(const fnl_pool p, const fnl_pool_request* __restrict__ req, const int32_t* __restrict__ stage_z,
            const double* __restrict__ stage_x) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= p.max_atoms) return;
  const int N = p.counts[FNL_POOL_N_ATOMS], B = p.counts[FNL_POOL_N_SYS];
  p.e_atom[r] = 0.0f;  // atoms without a network keep a zero energy (the MLP skips them)
#pragma unroll
  for (int k = 0; k < 9; ++k) p.virial_atom[9 * r + k] = 0.0f;
  int z = 0, b = 0;
  double x3[3] = {0.0, 0.0, 0.0}, v3[3] = {0.0, 0.0, 0.0};
  float g3[3] = {0.0f, 0.0f, 0.0f};
  long long n3[3] = {0, 0, 0};
  if (r < N && B > 0) {
    // system of row r: the largest b < B with sys_offsets[b] <= r
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (p.sys_offsets[mid] <= r) lo = mid; else hi = mid - 1;
    }
    b = lo;
    const int local = r - p.sys_offsets[b];
    const fnl_pool_request q = req[b];
    if (q.src_row >= 0) {
      const long long s = (long long)q.src_row + local;
      if (s < p.max_atoms) {
        z = p.old_z[s];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          x3[k] = p.old_x[3 * s + k];
          v3[k] = p.old_v[3 * s + k];
          g3[k] = p.old_grad_self[3 * s + k];
          n3[k] = (long long)p.old_grad_nbr[3 * s + k];
        }
      }
    } else {
      const long long s = (long long)q.stage_row + local;
      if (q.stage_row >= 0 && s < p.stage_rows) {
        z = stage_z[s];
#pragma unroll
        for (int k = 0; k < 3; ++k) x3[k] = stage_x[3 * s + k];
        if (q.md_left > 0 && z > 0 && z < p.n_z) {
          double xi[3];
          fnl_normals3(p.seed, 0ull, q.vel_key, (uint32_t)local, xi);
          const double vs = p.z_vscale[z];
#pragma unroll
          for (int k = 0; k < 3; ++k) v3[k] = xi[k] * vs;
        }
      }
    }
  }
  const bool zok = r < N && z >= 0 && z < p.n_z;
  p.z[r] = zok ? z : 0;
  p.species_idx[r] = zok ? p.z_species[z] : -1;
  p.batch_index[r] = b;
  p.dtm[r] = zok ? p.z_dtm[z] : 0.0;
  p.a2[r] = zok ? p.z_a2[z] : 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p.x[3 * r + k] = x3[k];
    p.v[3 * r + k] = v3[k];
    p.x32[3 * r + k] = (float)x3[k];
    p.grad_self[3 * r + k] = g3[k];
    p.grad_nbr[3 * r + k] = (int64_t)n3[k];
  }
}

// Network-sorted atom list and MLP tile table (the order of a stable sort by network, padding
// and atoms without a network left out; each network's atoms cut into tiles of tile_rows).
// One workgroup; each thread counts, then places, a contiguous run of rows.
__global__ void __launch_bounds__(kScanThreads) k_pool_tiles(const fnl_pool p) {
  __shared__ int lds[kScanWaves + 1];
  __shared__ int net_tot[FNL_MLP_MAX_NETS], net_base[FNL_MLP_MAX_NETS], tile_base[FNL_MLP_MAX_NETS + 1];
  const int N = p.counts[FNL_POOL_N_ATOMS];
  const int NN = p.n_nets, M = p.tile_rows;
  const int per = (N + kScanThreads - 1) / kScanThreads;
  const int r0 = min((int)threadIdx.x * per, N), r1 = min(r0 + per, N);
  auto net_of = [&](int r) -> int {
    if (p.species_idx[r] < 0) return -1;
    return p.z_net[p.z[r]];
  };
  int cnt[FNL_MLP_MAX_NETS];
  int off[FNL_MLP_MAX_NETS];
#pragma unroll
  for (int n = 0; n < FNL_MLP_MAX_NETS; ++n) cnt[n] = 0;
  for (int r = r0; r < r1; ++r) {
    const int n = net_of(r);
#pragma unroll
    for (int m = 0; m < FNL_MLP_MAX_NETS; ++m) cnt[m] += (n == m) ? 1 : 0;
  }
#pragma unroll
  for (int n = 0; n < FNL_MLP_MAX_NETS; ++n) {
    int tot = 0;
    off[n] = n < NN ? block_excl_scan(cnt[n], lds, &tot) : 0;
    if (threadIdx.x == 0 && n < NN) net_tot[n] = tot;
  }
  if (threadIdx.x == 0) {
    int base = 0, tb = 0;
    for (int n = 0; n < NN; ++n) {
      net_base[n] = base;
      tile_base[n] = tb;
      base += net_tot[n];
      tb += (net_tot[n] + M - 1) / M;
    }
    tile_base[NN] = tb;
    p.counts[FNL_POOL_N_TILES] = min(tb, p.max_tiles);
    if (tb > p.max_tiles) p.counts[FNL_POOL_ERROR] |= 2;
  }
  __syncthreads();
  for (int r = r0; r < r1; ++r) {
    const int n = net_of(r);
#pragma unroll
    for (int m = 0; m < FNL_MLP_MAX_NETS; ++m)
      if (n == m) p.perm[net_base[m] + off[m]++] = r;
  }
  const int T = min(tile_base[NN], p.max_tiles);
  for (int t = threadIdx.x; t < T; t += kScanThreads) {
    int n = 0;
    while (n + 1 < NN && tile_base[n + 1] <= t) ++n;
    const int j = t - tile_base[n];
    p.tile_start[t] = net_base[n] + j * M;
    p.tile_rows_tab[t] = min(M, net_tot[n] - j * M);
    p.tile_net[t] = n;
  }
}

__global__ void __launch_bounds__(256) k_pool_save(const fnl_pool p, int n_rows) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  p.old_z[r] = p.z[r];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p.old_x[3 * r + k] = p.x[3 * r + k];
    p.old_v[3 * r + k] = p.v[3 * r + k];
    p.old_grad_self[3 * r + k] = p.grad_self[3 * r + k];
    p.old_grad_nbr[3 * r + k] = p.grad_nbr[3 * r + k];
  }
}

// BAOAB of the rows of requests that integrate this iteration (evaluated, steps left)
__global__ void __launch_bounds__(256)
k_pool_integrate(const fnl_pool p, float force_scale, double half_dt, double a1, int n_rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rows || i >= p.counts[FNL_POOL_N_ATOMS]) return;
  if (p.species_idx[i] < 0) return;
  const int b = p.batch_index[i];
  if (!(p.sys_left[b] > 0 && p.sys_eval[b])) return;
  const double m = p.dtm[i];
  if (m == 0.0) return;
  double xi[3];
  fnl_normals3(p.seed, p.sys_step[b], p.sys_key[b], (uint32_t)(i - p.sys_offsets[b]), xi);
  fnl_baoab_row(p.x, p.v, p.x32, p.grad_self, (const long long*)p.grad_nbr, force_scale, m, half_dt, a1, p.a2[i],
                xi, i);
}

__host__ __device__ __forceinline__ size_t out_flags_bytes(int n_sys) { return ((size_t)n_sys * 4 + 7) & ~(size_t)7; }

// Iteration bookkeeping and the record's header, flags and energies.  One workgroup.
__global__ void __launch_bounds__(kScanThreads)
k_pool_end(const fnl_pool p, const double* __restrict__ energy_sys, const int32_t* __restrict__ status,
           unsigned char* __restrict__ out, size_t out_bytes) {
  __shared__ int lds[kScanWaves + 1];
  const int B = p.counts[FNL_POOL_N_SYS], N = p.counts[FNL_POOL_N_ATOMS];
  const int per = (B + kScanThreads - 1) / kScanThreads;
  const int b0 = min((int)threadIdx.x * per, B), b1 = min(b0 + per, B);
  int nf = 0, na = 0;
  for (int b = b0; b < b1; ++b) {
    int left = p.sys_left[b];
    if (left > 0 && p.sys_eval[b]) {
      --left;
      p.sys_left[b] = left;
      p.sys_step[b] = p.sys_step[b] + 1;
    }
    p.sys_eval[b] = 1;
    if (left == 0) {
      ++nf;
      na += p.sys_offsets[b + 1] - p.sys_offsets[b];
    }
  }
  int tf, ta;
  int of = block_excl_scan(nf, lds, &tf);
  int oa = block_excl_scan(na, lds, &ta);
  const size_t off_e = FNL_POOL_HEADER_WORDS * 4 + out_flags_bytes(B);
  const size_t total = off_e + 8 * (size_t)tf + 36 * (size_t)ta;
  const bool fits = total <= out_bytes;
  int32_t* hdr = (int32_t*)out;
  int32_t* flags = hdr + FNL_POOL_HEADER_WORDS;
  double* e_out = (double*)(out + off_e);
  for (int b = b0; b < b1; ++b) {
    const bool fin = p.sys_left[b] == 0 && fits;
    flags[b] = fin ? 1 : 0;
    p.sys_fin[b] = fin ? of : -1;
    p.sys_out_row[b] = oa;
    if (fin) {
      e_out[of] = energy_sys[b];
      ++of;
      oa += p.sys_offsets[b + 1] - p.sys_offsets[b];
    }
  }
  if (threadIdx.x < FNL_POOL_HEADER_WORDS) {
    int v = 0;
    switch (threadIdx.x) {
      case 0: v = B; break;
      case 1: v = fits ? tf : 0; break;
      case 2: v = fits ? ta : 0; break;
      case 3: v = N; break;
      case 4: case 5: case 6: case 7: v = status[threadIdx.x - 4]; break;
      case 8: v = p.counts[FNL_POOL_ERROR] | (fits ? 0 : 4); break;
      default: v = 0;
    }
    hdr[threadIdx.x] = v;
  }
}

// Positions and forces of the finished requests into the record; row-overflow flags of every
// request (a truncated neighbour row makes its result invalid).
__global__ void __launch_bounds__(256)
k_pool_gather(const fnl_pool p, const int32_t* __restrict__ n_nbr, const int32_t* __restrict__ n_nbr_ang,
              int max_nbr, int max_nbr_ang, float force_scale, int n_rows, unsigned char* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rows || i >= p.counts[FNL_POOL_N_ATOMS]) return;
  int32_t* hdr = (int32_t*)out;
  int32_t* flags = hdr + FNL_POOL_HEADER_WORDS;
  const int B = hdr[0], tf = hdr[1], ta = hdr[2];
  const int b = p.batch_index[i];
  if (b >= B) return;
  if (p.species_idx[i] >= 0 && (n_nbr[i] > max_nbr || n_nbr_ang[i] > max_nbr_ang)) atomicOr(flags + b, 2);
  const int f = p.sys_fin[b];
  if (f < 0) return;
  const size_t off_x = FNL_POOL_HEADER_WORDS * 4 + out_flags_bytes(B) + 8 * (size_t)tf;
  double* x_out = (double*)(out + off_x);
  float* f_out = (float*)(out + off_x + 24 * (size_t)ta);
  const int dst = p.sys_out_row[b] + (i - p.sys_offsets[b]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x_out[3 * dst + k] = p.x[3 * i + k];
    f_out[3 * dst + k] =
        (float)(-((double)p.grad_self[3 * i + k] + fnl_from_fixed((long long)p.grad_nbr[3 * i + k])) *
                (double)force_scale);
  }
}

bool pool_ok(const fnl_pool* p) {
  if (!p || p->max_atoms < 1 || p->max_sys < 1 || p->max_tiles < 1 || p->n_nets < 1 ||
      p->n_nets > FNL_MLP_MAX_NETS || p->tile_rows < 1 || p->seg_atoms < 1 || p->n_z < 1 || p->stage_rows < 0)
    return false;
  const void* ptrs[] = {p->z_species, p->z_net, p->z_dtm, p->z_a2, p->z_vscale, p->x, p->v, p->x32, p->z,
                        p->grad_self, p->grad_nbr, p->e_atom, p->virial_atom, p->dtm, p->a2, p->old_x, p->old_v,
                        p->old_grad_self, p->old_grad_nbr, p->old_z, p->batch_index, p->species_idx, p->perm,
                        p->sys_offsets, p->seg_offsets, p->grid_cap, p->tile_start, p->tile_rows_tab, p->tile_net,
                        p->sys_left, p->sys_eval, p->sys_step, p->sys_key, p->sys_fin, p->sys_out_row, p->counts};
  for (const void* q : ptrs)
    if (!q) return false;
  return true;
}

}  // namespace

extern "C" size_t fnl_pool_out_bytes(int32_t n_sys, int32_t n_fin, int32_t n_fin_atoms) {
  if (n_sys < 0 || n_fin < 0 || n_fin_atoms < 0) return 0;
  return FNL_POOL_HEADER_WORDS * 4 + out_flags_bytes(n_sys) + 8 * (size_t)n_fin + 36 * (size_t)n_fin_atoms;
}

extern "C" int fnl_pool_save_rows(const fnl_pool* pool, int32_t n_rows, fnl_stream_t stream) {
  if (!pool_ok(pool) || n_rows < 0 || n_rows > pool->max_atoms) return FNL_ERR_INVALID;
  if (n_rows == 0) return FNL_OK;
  hipLaunchKernelGGL(k_pool_save, dim3(fnl_blocks(n_rows, 256)), dim3(256), 0, (hipStream_t)stream, *pool, n_rows);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_pool_layout(const fnl_pool* pool, const fnl_pool_request* req, int32_t n_req,
                               const int32_t* stage_z, const double* stage_x, fnl_stream_t stream) {
  if (!pool_ok(pool) || n_req < 0 || n_req > pool->max_sys || (n_req > 0 && !req)) return FNL_ERR_INVALID;
  if (pool->stage_rows > 0 && (!stage_z || !stage_x)) return FNL_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pool_systems, dim3(1), dim3(kScanThreads), 0, st, *pool, req, n_req);
  hipLaunchKernelGGL(k_pool_rows, dim3(fnl_blocks(pool->max_atoms, 256)), dim3(256), 0, st, *pool, req, stage_z,
                     stage_x);
  hipLaunchKernelGGL(k_pool_tiles, dim3(1), dim3(kScanThreads), 0, st, *pool);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_pool_integrate(const fnl_pool* pool, float force_scale, double half_dt, double a1,
                                  int32_t n_rows, fnl_stream_t stream) {
  if (!pool_ok(pool) || n_rows < 0 || n_rows > pool->max_atoms) return FNL_ERR_INVALID;
  if (n_rows == 0) return FNL_OK;
  hipLaunchKernelGGL(k_pool_integrate, dim3(fnl_blocks(n_rows, 256)), dim3(256), 0, (hipStream_t)stream, *pool,
                     force_scale, half_dt, a1, n_rows);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_pool_finish(const fnl_pool* pool, const double* energy_sys, const int32_t* n_nbr,
                               const int32_t* n_nbr_ang, int32_t max_nbr, int32_t max_nbr_ang, const int32_t* status,
                               float force_scale, int32_t n_rows, void* out, size_t out_bytes, fnl_stream_t stream) {
  if (!pool_ok(pool) || !energy_sys || !n_nbr || !n_nbr_ang || !status || !out || n_rows < 0 ||
      n_rows > pool->max_atoms || max_nbr < 1 || max_nbr_ang < 1)
    return FNL_ERR_INVALID;
  // the header and the flags of every request always fit
  if (out_bytes < fnl_pool_out_bytes(pool->max_sys, 0, 0)) return FNL_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pool_end, dim3(1), dim3(kScanThreads), 0, st, *pool, energy_sys, status,
                     (unsigned char*)out, out_bytes);
  if (n_rows > 0)
    hipLaunchKernelGGL(k_pool_gather, dim3(fnl_blocks(n_rows, 256)), dim3(256), 0, st, *pool, n_nbr, n_nbr_ang,
                       max_nbr, max_nbr_ang, force_scale, n_rows, (unsigned char*)out);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
