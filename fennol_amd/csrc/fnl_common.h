// Shared device helpers for the gfx950 kernels (wave64 CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <set>
#include <utility>

#include "../../include/fnl_hip.h"

#define FNL_WAVE 64

#define FNL_LAUNCH_CHECK()                                  \
  do {                                                      \
    if (hipGetLastError() != hipSuccess) return FNL_ERR_HIP; \
  } while (0)

static inline int fnl_blocks(long long n, int per_block) {
  return (int)((n + per_block - 1) / per_block);
}

static inline size_t fnl_align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Raise a kernel's dynamic-LDS limit to the whole 160 KB of a CU, once per kernel and device.
// The only state the library keeps: a mutex-protected set of the (kernel, device) pairs already
// raised, so concurrent callers on several host threads neither race nor lower a limit another
// thread relies on.  hipFuncSetAttribute is not a stream operation, so a call made while the
// caller's stream is being captured into a graph records nothing.
inline int fnl_allow_full_lds(const void* kernel) {
  static std::mutex mu;
  static std::set<std::pair<const void*, int>> raised;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FNL_ERR_HIP;
  std::lock_guard<std::mutex> lock(mu);
  if (raised.count({kernel, dev})) return FNL_OK;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
    return FNL_ERR_HIP;
  raised.insert({kernel, dev});
  return FNL_OK;
}

// -------------------------------------------------------------------------
// wave-level helpers
// -------------------------------------------------------------------------
__device__ __forceinline__ int fnl_lane() { return (int)__lane_id(); }

// number of set bits of m below this lane
__device__ __forceinline__ int fnl_rank_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ float fnl_wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, FNL_WAVE);
  return v;
}

__device__ __forceinline__ int fnl_wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, FNL_WAVE));
  return v;
}

// inclusive prefix sum over the wave
__device__ __forceinline__ int fnl_wave_incl_scan(int v) {
  const int lane = fnl_lane();
#pragma unroll
  for (int o = 1; o < FNL_WAVE; o <<= 1) {
    int t = __shfl_up(v, o, FNL_WAVE);
    if (lane >= o) v += t;
  }
  return v;
}

// -------------------------------------------------------------------------
// Minimum-image pair vector, float32, the reference's operation order
// (models/preprocessing.py:127-143): vec = xj - xi; s = -rint(vec . R);
// vec += s . C; no FMA contraction so that d12 matches bit for bit.
// -------------------------------------------------------------------------
struct FnlCell {
  float c[9];  // rows = lattice vectors
  float r[9];  // reciprocal (inverse)
};

__device__ __forceinline__ void fnl_load_cell(const float* __restrict__ cells,
                                              const float* __restrict__ rec, int b,
                                              FnlCell& cl) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    cl.c[k] = cells[b * 9 + k];
    cl.r[k] = rec[b * 9 + k];
  }
}

__device__ __forceinline__ void fnl_min_image(float& vx, float& vy, float& vz, const FnlCell& cl,
                                              float* shift_out = nullptr) {
#pragma clang fp contract(off)
  const float f0 = (vx * cl.r[0] + vy * cl.r[3]) + vz * cl.r[6];
  const float f1 = (vx * cl.r[1] + vy * cl.r[4]) + vz * cl.r[7];
  const float f2 = (vx * cl.r[2] + vy * cl.r[5]) + vz * cl.r[8];
  const float s0 = -rintf(f0), s1 = -rintf(f1), s2 = -rintf(f2);
  vx = vx + ((s0 * cl.c[0] + s1 * cl.c[3]) + s2 * cl.c[6]);
  vy = vy + ((s0 * cl.c[1] + s1 * cl.c[4]) + s2 * cl.c[7]);
  vz = vz + ((s0 * cl.c[2] + s1 * cl.c[5]) + s2 * cl.c[8]);
  if (shift_out) {
    shift_out[0] = s0;
    shift_out[1] = s1;
    shift_out[2] = s2;
  }
}

// fnl_min_image for a cell whose lattice vectors lie along the axes (c and r diagonal): the
// off-diagonal products are exact zeros and adding a zero leaves a sum unchanged (up to the sign
// of a zero, which neither rint nor d12 sees), so this gives the same vec bit for bit
__device__ __forceinline__ bool fnl_cell_diagonal(const FnlCell& cl) {
  return cl.c[1] == 0.0f && cl.c[2] == 0.0f && cl.c[3] == 0.0f && cl.c[5] == 0.0f && cl.c[6] == 0.0f &&
         cl.c[7] == 0.0f && cl.r[1] == 0.0f && cl.r[2] == 0.0f && cl.r[3] == 0.0f && cl.r[5] == 0.0f &&
         cl.r[6] == 0.0f && cl.r[7] == 0.0f;
}

__device__ __forceinline__ void fnl_min_image_diag(float& vx, float& vy, float& vz, const FnlCell& cl) {
#pragma clang fp contract(off)
  vx = vx + (-rintf(vx * cl.r[0])) * cl.c[0];
  vy = vy + (-rintf(vy * cl.r[4])) * cl.c[4];
  vz = vz + (-rintf(vz * cl.r[8])) * cl.c[8];
}

// Integer cell shifts (s0, s1, s2) packed into one int32 (general PBC rows):
// 10 bits per component, offset 512, so |s| <= 511 lattice vectors.
__device__ __host__ __forceinline__ int32_t fnl_shift_pack(int s0, int s1, int s2) {
  return (int32_t)((s0 + 512) | ((s1 + 512) << 10) | ((s2 + 512) << 20));
}

__device__ __host__ __forceinline__ bool fnl_shift_fits(int s0, int s1, int s2) {
  return s0 >= -511 && s0 <= 511 && s1 >= -511 && s1 <= 511 && s2 >= -511 && s2 <= 511;
}

__device__ __forceinline__ void fnl_shift_unpack(int32_t p, float& s0, float& s1, float& s2) {
  s0 = (float)((p & 1023) - 512);
  s1 = (float)(((p >> 10) & 1023) - 512);
  s2 = (float)(((p >> 20) & 1023) - 512);
}

// vec += s . C for a stored shift (general PBC), the operation order of fnl_min_image
__device__ __forceinline__ void fnl_apply_shift(float& vx, float& vy, float& vz, const FnlCell& cl,
                                                int32_t packed) {
#pragma clang fp contract(off)
  float s0, s1, s2;
  fnl_shift_unpack(packed, s0, s1, s2);
  vx = vx + ((s0 * cl.c[0] + s1 * cl.c[3]) + s2 * cl.c[6]);
  vy = vy + ((s0 * cl.c[1] + s1 * cl.c[4]) + s2 * cl.c[7]);
  vz = vz + ((s0 * cl.c[2] + s1 * cl.c[5]) + s2 * cl.c[8]);
}

// Atom wrapped into the cell for the general-PBC build (preprocessing.py:145-149):
// at = -floor(x . R), xw = x + at . C, float32 in a fixed order.
__device__ __forceinline__ void fnl_wrap(float x, float y, float z, const FnlCell& cl, float* xw, int* at) {
#pragma clang fp contract(off)
  const float f0 = (x * cl.r[0] + y * cl.r[3]) + z * cl.r[6];
  const float f1 = (x * cl.r[1] + y * cl.r[4]) + z * cl.r[7];
  const float f2 = (x * cl.r[2] + y * cl.r[5]) + z * cl.r[8];
  const float a0 = -floorf(f0), a1 = -floorf(f1), a2 = -floorf(f2);
  xw[0] = x + ((a0 * cl.c[0] + a1 * cl.c[3]) + a2 * cl.c[6]);
  xw[1] = y + ((a0 * cl.c[1] + a1 * cl.c[4]) + a2 * cl.c[7]);
  xw[2] = z + ((a0 * cl.c[2] + a1 * cl.c[5]) + a2 * cl.c[8]);
  at[0] = (int)a0;
  at[1] = (int)a1;
  at[2] = (int)a2;
}

__device__ __forceinline__ float fnl_d12(float vx, float vy, float vz) {
#pragma clang fp contract(off)
  return (vx * vx + vy * vy) + vz * vz;
}

// -------------------------------------------------------------------------
// Fixed-point accumulation of the angular neighbour gradients: each float term
// is rounded to a multiple of 2^-32 in an int64, and the int64 sums are exact,
// so the result does not depend on the order the atomics land in (forces are
// the same bits every run).
// -------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long fnl_to_fixed(float v) {
  return (unsigned long long)__double2ll_rn((double)v * (double)FNL_GRAD_FIXED_ONE);
}

__device__ __forceinline__ double fnl_from_fixed(long long v) {
  return (double)v * (1.0 / (double)FNL_GRAD_FIXED_ONE);
}

// -------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) and Box-Muller normals.
// -------------------------------------------------------------------------
struct FnlU4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ FnlU4 fnl_philox4x32_10(FnlU4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    FnlU4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// three N(0,1) doubles for (atom, step, stream) under key seed
__device__ __forceinline__ void fnl_normals3(uint64_t seed, uint64_t step, uint32_t stream,
                                             uint32_t atom, double out[3]) {
  FnlU4 c{atom, (uint32_t)step, (uint32_t)(step >> 32), stream};
  FnlU4 r = fnl_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = ((double)r.x + 0.5) * s, u1 = ((double)r.y + 0.5) * s;
  const double u2 = ((double)r.z + 0.5) * s, u3 = ((double)r.w + 0.5) * s;
  const double two_pi = 6.283185307179586;
  const double rad0 = sqrt(-2.0 * log(u0));
  const double rad1 = sqrt(-2.0 * log(u2));
  double s0, c0, s1, c1;
  sincos(two_pi * u1, &s0, &c0);
  sincos(two_pi * u3, &s1, &c1);
  out[0] = rad0 * c0;
  out[1] = rad0 * s0;
  out[2] = rad1 * c1;
}

// One atom's BAOAB step with merged half kicks (md/integrate.py:286-294, LGV O-step of
// md/thermostats.py:108-120): F = -(grad_self + grad_nbr 2^-32) * force_scale in float64,
// v += dtm F; x += dt/2 v; v = a1 v + a2 xi; x += dt/2 v.  Shared by the batched kernels so
// that a request integrates to the same bits wherever it is served.
__device__ __forceinline__ void fnl_baoab_row(double* __restrict__ x, double* __restrict__ v,
                                              float* __restrict__ x32, const float* __restrict__ gs,
                                              const long long* __restrict__ gn, float force_scale, double dtm,
                                              double half_dt, double a1, double a2, const double xi[3], int i) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double f = -((double)gs[3 * i + k] + fnl_from_fixed(gn[3 * i + k])) * (double)force_scale;
    double vv = v[3 * i + k] + dtm * f;
    double xx = x[3 * i + k] + half_dt * vv;
    vv = a1 * vv + a2 * xi[k];
    xx = xx + half_dt * vv;
    v[3 * i + k] = vv;
    x[3 * i + k] = xx;
    x32[3 * i + k] = (float)xx;
  }
}
