// ANI atomic environment vectors, forward and adjoint, for gfx950.
//
// One wave per centre atom, four independent waves per workgroup, each with
// its own slice of dynamic LDS.  Reference: models/embeddings/ani.py:47-142
// (radial and angular terms, channel layout), GraphProcessor
// models/preprocessing.py:699-777 (edge vectors, safe distance), the cosine
// SwitchFunction models/misc/misc.py:617-716 and GraphAngleProcessor
// preprocessing.py:1283-1317 (theta = arccos(0.95 cos)).
//
// The angular factor uses cos(theta - theta_z) = x cos(theta_z) + sqrt(1-x^2)
// sin(theta_z) with x = 0.95 cos(angle), equal to the reference's
// cos(arccos(x) - theta_z) without the arccos.
//
// Forward (per centre i):
//   1. radial rows: neighbour j, distance, 0.25*fc, species, vector -> LDS
//   2. radial channels: lanes = 4 neighbour groups x 16 radial shifts, each
//      lane owns one LDS accumulator column (no atomics, fixed order)
//   3. angular rows: the radial slots listed in the angular row
//   4. triplets (a<b of the angular row) in chunks of 64: one lane per
//      triplet evaluates the n_Z angle and n_A distance factors into LDS; the
//      wave then re-maps to (half, channel) lanes that accumulate the 32
//      channels of their half's 32 triplets into species-pair columns
//   5. the row (radial | angular | zero pad) is written coalesced
// Adjoint: the same terms are recomputed; dE/dvec per radial slot is
// accumulated in LDS (radial part by owner lanes, angular part with LDS
// float atomics), then scattered: grad_self[i] = -sum_e g_e (plain store),
// grad_nbr[j] += g_e (global atomics), virial_atom[i] = sum_e vec_e (x) g_e.

#include <hip/hip_bfloat16.h>

#include "fnl_common.h"

namespace {

constexpr int kWaves = 4;
constexpr float kPi = 3.14159265358979323846f;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ float cos_switch(float d, float rc) {
  return d < rc ? 0.5f * cosf(d * (kPi / rc)) + 0.5f : 0.0f;
}

__device__ __forceinline__ float cos_switch_grad(float d, float rc) {
  return d < rc ? -0.5f * (kPi / rc) * sinf(d * (kPi / rc)) : 0.0f;
}

// map a linear index t < n(n-1)/2 to the pair (a, b), a < b, row-major
__device__ __forceinline__ void tri_pair(int t, int n, int& a, int& b) {
  const float nn = (float)(2 * n - 1);
  int r = (int)floorf(0.5f * (nn - sqrtf(nn * nn - 8.0f * (float)t)));
  r = max(0, min(r, n - 2));
  // start(r) = r*n - r(r+1)/2 ; fix rounding
  while (r > 0 && r * n - (r * (r + 1)) / 2 > t) --r;
  while (r < n - 2 && (r + 1) * n - ((r + 1) * (r + 2)) / 2 <= t) ++r;
  a = r;
  b = t - (r * n - (r * (r + 1)) / 2) + r + 1;
}

struct RadialRow {
  float* d;    // distance
  float* fc;   // radial switch (0.25 folded in for the forward)
  int* s;      // neighbour species index
  float* vx;
  float* vy;
  float* vz;
};

__device__ __forceinline__ void load_radial_rows(int i, int nr, const float* __restrict__ coords,
                                                 const int32_t* __restrict__ species_idx,
                                                 const int32_t* __restrict__ row, bool pbc,
                                                 const FnlCell& cl, float rc, float fc_scale,
                                                 RadialRow rr) {
  const int lane = fnl_lane();
  const float xi = coords[3 * i], yi = coords[3 * i + 1], zi = coords[3 * i + 2];
  for (int e = lane; e < nr; e += 64) {
    const int j = row[e];
    float vx = coords[3 * j] - xi, vy = coords[3 * j + 1] - yi, vz = coords[3 * j + 2] - zi;
    if (pbc) fnl_min_image(vx, vy, vz, cl);
    const float d2 = vx * vx + vy * vy + vz * vz;
    const float d = sqrtf(fmaxf(d2, 1.0e-5f));
    rr.d[e] = d;
    rr.fc[e] = fc_scale * cos_switch(d, rc);
    rr.s[e] = species_idx[j];
    rr.vx[e] = vx;
    rr.vy[e] = vy;
    rr.vz[e] = vz;
  }
}

template <typename T>
__device__ __forceinline__ T to_out(float v);
template <>
__device__ __forceinline__ float to_out<float>(float v) { return v; }
template <>
__device__ __forceinline__ hip_bfloat16 to_out<hip_bfloat16>(float v) { return hip_bfloat16(v); }

struct AevDims {
  int S, P, nR, rad_len, aev_len;
};

__device__ __forceinline__ int pair_index(int a, int b, int S) {
  // row-major upper triangle incl. diagonal of an S x S table (ani.py:124-130)
  const int lo = min(a, b), hi = max(a, b);
  return lo * S - (lo * (lo - 1)) / 2 + (hi - lo);
}

// per-wave LDS footprint (in 4-byte words) of the forward kernel
__host__ __device__ inline int fwd_lds_words(int max_nbr, int max_nbr_ang, int S, int nR, int P, int NA,
                                             int NZ) {
  return 6 * max_nbr + 6 * max_nbr_ang + 4 * S * 16 + 2 * P * 32 + 64 * (NA + NZ + 1);
}

template <int NA, int NZ, typename TOut>
__global__ void __launch_bounds__(64 * kWaves)
k_aev_fwd(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
          const int32_t* __restrict__ batch_index, int n_atoms, int pbc_mode,
          const float* __restrict__ cells, const float* __restrict__ rec,
          const int32_t* __restrict__ nbr, const int32_t* __restrict__ n_nbr, int max_nbr,
          const int32_t* __restrict__ nbr_ang, const int32_t* __restrict__ n_nbr_ang, int max_nbr_ang,
          fnl_aev_params p, AevDims dm, TOut* __restrict__ aev, int stride) {
  extern __shared__ float lds[];
  const int w = threadIdx.x >> 6;
  const int lane = fnl_lane();
  const int i = blockIdx.x * kWaves + w;
  if (i >= n_atoms) return;
  const int words = fwd_lds_words(max_nbr, max_nbr_ang, dm.S, dm.nR, dm.P, NA, NZ);
  float* base = lds + (size_t)w * words;
  RadialRow rr;
  rr.d = base;
  rr.fc = rr.d + max_nbr;
  rr.s = (int*)(rr.fc + max_nbr);
  rr.vx = (float*)(rr.s + max_nbr);
  rr.vy = rr.vx + max_nbr;
  rr.vz = rr.vy + max_nbr;
  float* ax = rr.vz + max_nbr;  // angular rows: ux uy uz d fc s
  float* ay = ax + max_nbr_ang;
  float* az = ay + max_nbr_ang;
  float* ad = az + max_nbr_ang;
  float* afc = ad + max_nbr_ang;
  int* as = (int*)(afc + max_nbr_ang);
  float* racc = (float*)(as + max_nbr_ang);  // [4][S*16]
  float* aacc = racc + 4 * dm.S * 16;        // [2][P][32]
  float* trip = aacc + 2 * dm.P * 32;        // [64][NZ + NA + 1]
  constexpr int TW = NA + NZ + 1;

  TOut* out = aev + (size_t)i * stride;
  const int si = species_idx[i];
  if (si < 0) {
    for (int c = lane; c < stride; c += 64) out[c] = to_out<TOut>(0.0f);
    return;
  }
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  FnlCell cl;
  if (pbc) fnl_load_cell(cells, rec, batch_index[i], cl);
  const int nr = min(n_nbr[i], max_nbr);
  const int na = min(n_nbr_ang[i], max_nbr_ang);

  for (int c = lane; c < 4 * dm.S * 16; c += 64) racc[c] = 0.0f;
  for (int c = lane; c < 2 * dm.P * 32; c += 64) aacc[c] = 0.0f;
  load_radial_rows(i, nr, coords, species_idx, nbr + (size_t)i * max_nbr, pbc, cl, p.cutoff, 0.25f, rr);
  wave_sync();

  // radial channels
  {
    const int q = lane >> 4, k = lane & 15;
    const bool kv = k < dm.nR;
    const float shift = kv ? p.radial_shifts[k] : 0.0f;
    float* col = racc + q * dm.S * 16 + k;
    for (int e = q; e < nr; e += 4) {
      const float x = rr.d[e] - shift;
      const float term = rr.fc[e] * __expf(-p.radial_eta * x * x);
      if (kv) col[rr.s[e] * 16] += term;
    }
  }

  // angular rows
  const int32_t* arow = nbr_ang + (size_t)i * max_nbr_ang;
  for (int a = lane; a < na; a += 64) {
    const int e = arow[a];
    const float d = rr.d[e];
    const float inv = 1.0f / fmaxf(d, 1.0e-5f);
    ax[a] = rr.vx[e] * inv;
    ay[a] = rr.vy[e] * inv;
    az[a] = rr.vz[e] * inv;
    ad[a] = d;
    afc[a] = cos_switch(d, p.cutoff_ang);
    as[a] = rr.s[e];
  }
  wave_sync();

  const float pref = 2.0f * exp2f(-p.zeta);  // 2 * 0.5^zeta (ani.py:112)
  const int T = na * (na - 1) / 2;
  const int h = lane >> 5, ch = lane & 31;
  const int cd = ch / NZ, cz = ch % NZ;
  for (int base_t = 0; base_t < T; base_t += 64) {
    const int t = base_t + lane;
    float* tr = trip + lane * TW;
    if (t < T) {
      int a, b;
      tri_pair(t, na, a, b);
      const float c = ax[a] * ax[b] + ay[a] * ay[b] + az[a] * az[b];
      const float x = 0.95f * c;
      const float st = sqrtf(fmaxf(1.0f - x * x, 0.0f));
      const float sw = pref * afc[a] * afc[b];
#pragma unroll
      for (int z = 0; z < NZ; ++z) {
        const float bz = 1.0f + x * p.ang_cos[z] + st * p.ang_sin[z];
        tr[z] = bz > 0.0f ? sw * exp2f(p.zeta * __log2f(bz)) : 0.0f;
      }
      const float m = 0.5f * (ad[a] + ad[b]);
#pragma unroll
      for (int d = 0; d < NA; ++d) {
        const float y = m - p.ang_shifts[d];
        tr[NZ + d] = __expf(-p.angular_eta * y * y);
      }
      tr[NZ + NA] = __int_as_float(pair_index(as[a], as[b], dm.S));
    } else {
      tr[NZ + NA] = __int_as_float(-1);
    }
    wave_sync();
    float* col = aacc + h * dm.P * 32 + ch;
#pragma unroll 4
    for (int tt = 0; tt < 32; ++tt) {
      const float* q = trip + (h * 32 + tt) * TW;
      const int pi = __float_as_int(q[NZ + NA]);
      if (pi >= 0) col[pi * 32] += q[cz] * q[NZ + cd];
    }
    wave_sync();
  }
  wave_sync();

  // write the row: radial | angular | zero padding
  const int ang_len = dm.P * NA * NZ;
  for (int c = lane; c < stride; c += 64) {
    float v = 0.0f;
    if (c < dm.rad_len) {
      const int s = c / dm.nR, k = c % dm.nR;
      const int o = s * 16 + k;
      v = racc[o] + racc[dm.S * 16 + o] + racc[2 * dm.S * 16 + o] + racc[3 * dm.S * 16 + o];
    } else if (c < dm.rad_len + ang_len) {
      const int o = c - dm.rad_len;
      const int pi = o / (NA * NZ), cc = o % (NA * NZ);
      v = aacc[pi * 32 + cc] + aacc[dm.P * 32 + pi * 32 + cc];
    }
    out[c] = to_out<TOut>(v);
  }
}

__host__ __device__ inline int bwd_lds_words(int max_nbr, int max_nbr_ang, int aev_len) {
  return aev_len + 9 * max_nbr + 9 * max_nbr_ang;
}

template <int NA, int NZ>
__global__ void __launch_bounds__(64 * kWaves)
k_aev_bwd(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
          const int32_t* __restrict__ batch_index, int n_atoms, int pbc_mode,
          const float* __restrict__ cells, const float* __restrict__ rec,
          const int32_t* __restrict__ nbr, const int32_t* __restrict__ n_nbr, int max_nbr,
          const int32_t* __restrict__ nbr_ang, const int32_t* __restrict__ n_nbr_ang, int max_nbr_ang,
          fnl_aev_params p, AevDims dm, const float* __restrict__ g, int g_stride,
          float* __restrict__ grad_self, float* __restrict__ grad_nbr, float* __restrict__ virial_atom) {
  extern __shared__ float lds[];
  const int w = threadIdx.x >> 6;
  const int lane = fnl_lane();
  const int i = blockIdx.x * kWaves + w;
  if (i >= n_atoms) return;
  const int words = bwd_lds_words(max_nbr, max_nbr_ang, dm.aev_len);
  float* G = lds + (size_t)w * words;
  RadialRow rr;
  rr.d = G + dm.aev_len;
  rr.fc = rr.d + max_nbr;  // unscaled switch
  rr.s = (int*)(rr.fc + max_nbr);
  rr.vx = (float*)(rr.s + max_nbr);
  rr.vy = rr.vx + max_nbr;
  rr.vz = rr.vy + max_nbr;
  float* gx = rr.vz + max_nbr;  // per radial slot dE/dvec
  float* gy = gx + max_nbr;
  float* gz = gy + max_nbr;
  float* ax = gz + max_nbr;  // angular rows
  float* ay = ax + max_nbr_ang;
  float* az = ay + max_nbr_ang;
  float* ad = az + max_nbr_ang;
  float* afc = ad + max_nbr_ang;
  float* adfc = afc + max_nbr_ang;
  float* aok = adfc + max_nbr_ang;  // 1 if d2 > 1e-5 (distance differentiable)
  int* as = (int*)(aok + max_nbr_ang);
  int* ae = as + max_nbr_ang;

  const int si = species_idx[i];
  if (si < 0) {
    if (lane < 3) grad_self[3 * i + lane] = 0.0f;
    if (lane < 9) virial_atom[9 * i + lane] = 0.0f;
    return;
  }
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  FnlCell cl;
  if (pbc) fnl_load_cell(cells, rec, batch_index[i], cl);
  const int nr = min(n_nbr[i], max_nbr);
  const int na = min(n_nbr_ang[i], max_nbr_ang);
  const float* grow = g + (size_t)i * g_stride;
  for (int c = lane; c < dm.aev_len; c += 64) G[c] = grow[c];
  for (int e = lane; e < nr; e += 64) gx[e] = gy[e] = gz[e] = 0.0f;
  load_radial_rows(i, nr, coords, species_idx, nbr + (size_t)i * max_nbr, pbc, cl, p.cutoff, 1.0f, rr);
  wave_sync();

  // radial: dE/dd_e = sum_k G[s_e, k] * d/dd (0.25 fc exp(-eta (d - R_k)^2))
  {
    const int q = lane >> 4, k = lane & 15;
    const bool kv = k < dm.nR;
    const float shift = kv ? p.radial_shifts[k] : 0.0f;
    const int nrr = (nr + 3) & ~3;
    for (int e = q; e < nrr; e += 4) {
      float v = 0.0f;
      if (e < nr && kv) {
        const float d = rr.d[e];
        const float x = d - shift;
        const float ex = __expf(-p.radial_eta * x * x);
        const float fc = rr.fc[e];
        const float dfc = cos_switch_grad(d, p.cutoff);
        v = G[rr.s[e] * dm.nR + k] * 0.25f * ex * (dfc - 2.0f * p.radial_eta * x * fc);
      }
      v += __shfl_xor(v, 8, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 1, 64);
      if (k == 0 && e < nr) {
        const float vx = rr.vx[e], vy = rr.vy[e], vz = rr.vz[e];
        const float d2 = vx * vx + vy * vy + vz * vz;
        if (d2 > 1.0e-5f) {
          const float s = v / rr.d[e];
          gx[e] += s * vx;
          gy[e] += s * vy;
          gz[e] += s * vz;
        }
      }
    }
  }

  // angular rows
  const int32_t* arow = nbr_ang + (size_t)i * max_nbr_ang;
  for (int a = lane; a < na; a += 64) {
    const int e = arow[a];
    const float d = rr.d[e];
    const float inv = 1.0f / fmaxf(d, 1.0e-5f);
    ax[a] = rr.vx[e] * inv;
    ay[a] = rr.vy[e] * inv;
    az[a] = rr.vz[e] * inv;
    ad[a] = d;
    afc[a] = cos_switch(d, p.cutoff_ang);
    adfc[a] = cos_switch_grad(d, p.cutoff_ang);
    const float d2 = rr.vx[e] * rr.vx[e] + rr.vy[e] * rr.vy[e] + rr.vz[e] * rr.vz[e];
    aok[a] = d2 > 1.0e-5f ? 1.0f : 0.0f;
    as[a] = rr.s[e];
    ae[a] = e;
  }
  wave_sync();

  const float pref = 2.0f * exp2f(-p.zeta);
  const int T = na * (na - 1) / 2;
  const float* Gang = G + dm.rad_len;
  for (int t = lane; t < T; t += 64) {
    int a, b;
    tri_pair(t, na, a, b);
    const float c = ax[a] * ax[b] + ay[a] * ay[b] + az[a] * az[b];
    const float x = 0.95f * c;
    const float st = sqrtf(fmaxf(1.0f - x * x, 1.0e-12f));
    float F[NZ], dF[NZ], R[NA], dR[NA];
#pragma unroll
    for (int z = 0; z < NZ; ++z) {
      const float bz = 1.0f + x * p.ang_cos[z] + st * p.ang_sin[z];
      if (bz > 0.0f) {
        const float fz = exp2f(p.zeta * __log2f(bz));
        F[z] = fz;
        // d/dc (1 + cos(theta - theta_z))^zeta with theta = arccos(0.95 c)
        dF[z] = 0.95f * p.zeta * (fz / bz) * (p.ang_cos[z] - x * p.ang_sin[z] / st);
      } else {
        F[z] = 0.0f;
        dF[z] = 0.0f;
      }
    }
    const float m = 0.5f * (ad[a] + ad[b]);
#pragma unroll
    for (int d = 0; d < NA; ++d) {
      const float y = m - p.ang_shifts[d];
      R[d] = __expf(-p.angular_eta * y * y);
      dR[d] = -2.0f * p.angular_eta * y * R[d];
    }
    const float* gp = Gang + pair_index(as[a], as[b], dm.S) * (NA * NZ);
    float S0 = 0.0f, S1 = 0.0f, S2 = 0.0f;
#pragma unroll
    for (int d = 0; d < NA; ++d) {
      float gf = 0.0f, gdf = 0.0f;
#pragma unroll
      for (int z = 0; z < NZ; ++z) {
        const float gv = gp[d * NZ + z];
        gf += gv * F[z];
        gdf += gv * dF[z];
      }
      S0 += gf * R[d];
      S1 += gf * dR[d];
      S2 += gdf * R[d];
    }
    const float sw = pref * afc[a] * afc[b];
    const float dEda = pref * adfc[a] * afc[b] * S0 + 0.5f * sw * S1;
    const float dEdb = pref * afc[a] * adfc[b] * S0 + 0.5f * sw * S1;
    const float dEdc = sw * S2;
    const float ia = 1.0f / fmaxf(ad[a], 1.0e-5f), ib = 1.0f / fmaxf(ad[b], 1.0e-5f);
    const float ra = dEda * aok[a], rb = dEdb * aok[b];
    const float gax = ra * ax[a] + dEdc * (ax[b] - c * ax[a]) * ia;
    const float gay = ra * ay[a] + dEdc * (ay[b] - c * ay[a]) * ia;
    const float gaz = ra * az[a] + dEdc * (az[b] - c * az[a]) * ia;
    const float gbx = rb * ax[b] + dEdc * (ax[a] - c * ax[b]) * ib;
    const float gby = rb * ay[b] + dEdc * (ay[a] - c * ay[b]) * ib;
    const float gbz = rb * az[b] + dEdc * (az[a] - c * az[b]) * ib;
    const int ea = ae[a], eb = ae[b];
    atomicAdd(&gx[ea], gax);
    atomicAdd(&gy[ea], gay);
    atomicAdd(&gz[ea], gaz);
    atomicAdd(&gx[eb], gbx);
    atomicAdd(&gy[eb], gby);
    atomicAdd(&gz[eb], gbz);
  }
  wave_sync();

  // scatter: neighbours by atomics, centre by a wave reduction; virial
  float sx = 0.f, sy = 0.f, sz = 0.f;
  float v[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int32_t* row = nbr + (size_t)i * max_nbr;
  for (int e = lane; e < nr; e += 64) {
    const float x = gx[e], y = gy[e], z = gz[e];
    const int j = row[e];
    atomicAdd(&grad_nbr[3 * j], x);
    atomicAdd(&grad_nbr[3 * j + 1], y);
    atomicAdd(&grad_nbr[3 * j + 2], z);
    sx += x;
    sy += y;
    sz += z;
    const float vx = rr.vx[e], vy = rr.vy[e], vz = rr.vz[e];
    v[0] += vx * x; v[1] += vx * y; v[2] += vx * z;
    v[3] += vy * x; v[4] += vy * y; v[5] += vy * z;
    v[6] += vz * x; v[7] += vz * y; v[8] += vz * z;
  }
  sx = fnl_wave_sum(sx);
  sy = fnl_wave_sum(sy);
  sz = fnl_wave_sum(sz);
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = fnl_wave_sum(v[k]);
  if (lane == 0) {
    grad_self[3 * i] = -sx;
    grad_self[3 * i + 1] = -sy;
    grad_self[3 * i + 2] = -sz;
#pragma unroll
    for (int k = 0; k < 9; ++k) virial_atom[9 * i + k] = v[k];
  }
}

bool check_params(const fnl_aev_params* p, AevDims* dm) {
  if (!p) return false;
  if (p->n_species < 1 || p->n_species > FNL_MAX_SPECIES) return false;
  if (p->n_radial < 1 || p->n_radial > 16) return false;
  if (p->n_ang_dist < 1 || p->n_ang_sect < 1 || p->n_ang_dist * p->n_ang_sect > 32) return false;
  if (!(p->cutoff > 0.0f) || !(p->cutoff_ang > 0.0f) || p->cutoff_ang > p->cutoff) return false;
  dm->S = p->n_species;
  dm->P = p->n_species * (p->n_species + 1) / 2;
  dm->nR = p->n_radial;
  dm->rad_len = dm->S * dm->nR;
  dm->aev_len = dm->rad_len + dm->P * p->n_ang_dist * p->n_ang_sect;
  return true;
}

template <int NA, int NZ>
int launch_fwd(const float* coords, const int32_t* species_idx, const int32_t* batch_index, int n_atoms,
               int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
               const int32_t* n_nbr, int max_nbr, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
               int max_nbr_ang, const fnl_aev_params& p, const AevDims& dm, void* aev, int stride,
               int out_dtype, hipStream_t s) {
  const size_t lds = (size_t)kWaves * 4 * fwd_lds_words(max_nbr, max_nbr_ang, dm.S, dm.nR, dm.P, NA, NZ);
  if (lds > 160 * 1024) return FNL_ERR_UNSUPPORTED;
  dim3 grid(fnl_blocks(n_atoms, kWaves)), block(64 * kWaves);
  const void* kf = out_dtype == FNL_F32 ? reinterpret_cast<const void*>(&k_aev_fwd<NA, NZ, float>)
                                        : reinterpret_cast<const void*>(&k_aev_fwd<NA, NZ, hip_bfloat16>);
  if (hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return FNL_ERR_HIP;
  if (out_dtype == FNL_F32) {
    hipLaunchKernelGGL((k_aev_fwd<NA, NZ, float>), grid, block, lds, s, coords, species_idx, batch_index,
                       n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang,
                       p, dm, (float*)aev, stride);
  } else {
    hipLaunchKernelGGL((k_aev_fwd<NA, NZ, hip_bfloat16>), grid, block, lds, s, coords, species_idx,
                       batch_index, n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang,
                       max_nbr_ang, p, dm, (hip_bfloat16*)aev, stride);
  }
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

template <int NA, int NZ>
int launch_bwd(const float* coords, const int32_t* species_idx, const int32_t* batch_index, int n_atoms,
               int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
               const int32_t* n_nbr, int max_nbr, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
               int max_nbr_ang, const fnl_aev_params& p, const AevDims& dm, const float* g, int g_stride,
               float* grad_self, float* grad_nbr, float* virial_atom, hipStream_t s) {
  const size_t lds = (size_t)kWaves * 4 * bwd_lds_words(max_nbr, max_nbr_ang, dm.aev_len);
  if (lds > 160 * 1024) return FNL_ERR_UNSUPPORTED;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_aev_bwd<NA, NZ>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return FNL_ERR_HIP;
  hipLaunchKernelGGL((k_aev_bwd<NA, NZ>), dim3(fnl_blocks(n_atoms, kWaves)), dim3(64 * kWaves), lds, s,
                     coords, species_idx, batch_index, n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr,
                     nbr_ang, n_nbr_ang, max_nbr_ang, p, dm, g, g_stride, grad_self, grad_nbr, virial_atom);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

bool common_args_ok(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                    int n_atoms, int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
                    const int32_t* n_nbr, int max_nbr, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
                    int max_nbr_ang) {
  if (n_atoms < 0 || max_nbr < 1 || max_nbr_ang < 1) return false;
  if (n_atoms == 0) return true;
  if (!coords || !species_idx || !batch_index || !nbr || !n_nbr || !nbr_ang || !n_nbr_ang) return false;
  if (pbc_mode != FNL_PBC_NONE && pbc_mode != FNL_PBC_MINIMUM_IMAGE) return false;
  if (pbc_mode == FNL_PBC_MINIMUM_IMAGE && (!cells || !rec)) return false;
  return true;
}

}  // namespace

extern "C" int fnl_aev_forward(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                               int32_t n_atoms, int32_t pbc_mode, const float* cells,
                               const float* reciprocal_cells, const int32_t* nbr, const int32_t* n_nbr,
                               int32_t max_nbr, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
                               int32_t max_nbr_ang, const fnl_aev_params* params, void* aev,
                               int32_t aev_stride, int32_t out_dtype, fnl_stream_t stream) {
  AevDims dm;
  if (!check_params(params, &dm)) return FNL_ERR_INVALID;
  if (!common_args_ok(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, nbr,
                      n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang))
    return FNL_ERR_INVALID;
  if (aev_stride < dm.aev_len || (out_dtype != FNL_F32 && out_dtype != FNL_BF16)) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!aev) return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int na = params->n_ang_dist, nz = params->n_ang_sect;
  if (na == 4 && nz == 8)
    return launch_fwd<4, 8>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, aev,
                            aev_stride, out_dtype, s);
  if (na == 8 && nz == 4)
    return launch_fwd<8, 4>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, aev,
                            aev_stride, out_dtype, s);
  return FNL_ERR_UNSUPPORTED;
}

extern "C" int fnl_aev_backward(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                                int32_t n_atoms, int32_t pbc_mode, const float* cells,
                                const float* reciprocal_cells, const int32_t* nbr, const int32_t* n_nbr,
                                int32_t max_nbr, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
                                int32_t max_nbr_ang, const fnl_aev_params* params, const float* g,
                                int32_t g_stride, float* grad_self, float* grad_nbr, float* virial_atom,
                                fnl_stream_t stream) {
  AevDims dm;
  if (!check_params(params, &dm)) return FNL_ERR_INVALID;
  if (!common_args_ok(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, nbr,
                      n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang))
    return FNL_ERR_INVALID;
  if (g_stride < dm.aev_len) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!g || !grad_self || !grad_nbr || !virial_atom) return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int na = params->n_ang_dist, nz = params->n_ang_sect;
  if (na == 4 && nz == 8)
    return launch_bwd<4, 8>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, g, g_stride,
                            grad_self, grad_nbr, virial_atom, s);
  if (na == 8 && nz == 4)
    return launch_bwd<8, 4>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, g, g_stride,
                            grad_self, grad_nbr, virial_atom, s);
  return FNL_ERR_UNSUPPORTED;
}
