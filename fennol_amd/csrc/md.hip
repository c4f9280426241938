// Per-system FP64 reductions, force assembly and the BAOAB Langevin step.
//
// References: total_energy segment_sum (models/fennix.py:240-246), virial
// (fennix.py:320-347), kinetic tensor (md/integrate.py:320-339), BAOAB
// integrate (md/integrate.py:286-294) with the LGV thermostat O-step
// (md/thermostats.py:108-120).

#include "fnl_common.h"

namespace {

constexpr int kRedVals = 1 + 9 + 9;  // energy, virial[9], kinetic[9]
constexpr int kFuseMaxSystems = FNL_REDUCE_FUSED_MAX_SYSTEMS;  // summed by the last segment workgroup

// Per-system sums of the partials of its segments: one wave per system, lane k
// sums segments k, k+64, ...; then a fixed xor butterfly (deterministic order).
__device__ __forceinline__ void reduce_system(int b, int lane, const int32_t* __restrict__ seg_offsets,
                                              const double* partials, double* __restrict__ energy_sys,
                                              double* __restrict__ virial_sys, double* __restrict__ kinetic_sys) {
  double acc[kRedVals];
#pragma unroll
  for (int k = 0; k < kRedVals; ++k) acc[k] = 0.0;
  // partials read at the device's coherence point (agent-scope atomic loads: see k_reduce_segments)
  for (int s = seg_offsets[b] + lane; s < seg_offsets[b + 1]; s += 64)
#pragma unroll
    for (int k = 0; k < kRedVals; ++k)
      acc[k] += __longlong_as_double((long long)__hip_atomic_load(
          reinterpret_cast<const unsigned long long*>(partials + (size_t)s * kRedVals + k), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT));
#pragma unroll
  for (int k = 0; k < kRedVals; ++k)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  if (lane == 0 && energy_sys) energy_sys[b] = acc[0];
  if (lane < 9) {
    // lane k writes component k (the values are wave-uniform after the butterfly)
    double vk = 0.0, kk = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (lane == k) {
        vk = acc[1 + k];
        kk = acc[10 + k];
      }
    if (virial_sys) virial_sys[9 * b + lane] = vk;
    if (kinetic_sys) kinetic_sys[9 * b + lane] = kk;
  }
}

// One workgroup (256 threads) per segment (a run of at most seg_atoms atoms of
// one system) -> FP64 partial sums in a fixed tree order.  The last of the
// n_seg segment workgroups to finish (a completion count in the word before the
// partials) then forms the per-system sums, so the result is bitwise
// reproducible and one launch does both stages.
__device__ __forceinline__ void reduce_segment_block(
    int sgi, int n_seg, const int32_t* __restrict__ sys_offsets, const int32_t* __restrict__ seg_offsets,
    int n_sys, int seg_atoms, const float* __restrict__ energy_atom, const float* __restrict__ virial_atom,
    const double* __restrict__ vel, const double* __restrict__ mass, double* __restrict__ partials,
    unsigned* __restrict__ done, double* __restrict__ energy_sys, double* __restrict__ virial_sys,
    double* __restrict__ kinetic_sys, int fuse_final) {
  __shared__ double wsum[4][kRedVals];
  const int t = threadIdx.x;
  // system of this segment: largest b with seg_offsets[b] <= sgi
  int lo = 0, hi = n_sys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_offsets[mid] <= sgi) lo = mid; else hi = mid - 1;
  }
  const int b = lo;
  const int a0 = sys_offsets[b] + (sgi - seg_offsets[b]) * seg_atoms;
  const int a1 = min(a0 + seg_atoms, sys_offsets[b + 1]);
  double acc[kRedVals];
#pragma unroll
  for (int k = 0; k < kRedVals; ++k) acc[k] = 0.0;
  for (int i = a0 + t; i < a1; i += 256) {
    if (energy_atom) acc[0] += (double)energy_atom[i];
    if (virial_atom) {
#pragma unroll
      for (int k = 0; k < 9; ++k) acc[1 + k] += (double)virial_atom[9 * i + k];
    }
    if (vel) {
      const double m = 0.5 * mass[i];
      const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
      acc[10] += m * vx * vx; acc[11] += m * vx * vy; acc[12] += m * vx * vz;
      acc[13] += m * vy * vx; acc[14] += m * vy * vy; acc[15] += m * vy * vz;
      acc[16] += m * vz * vx; acc[17] += m * vz * vy; acc[18] += m * vz * vz;
    }
  }
  // wave sums by a fixed xor butterfly, then the four wave sums in order (the same tree every
  // launch: bitwise reproducible)
  const bool need_v = virial_atom != nullptr, need_k = vel != nullptr;
#pragma unroll
  for (int k = 0; k < kRedVals; ++k) {
    if ((k >= 1 && k < 10 && !need_v) || (k >= 10 && !need_k)) continue;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
  }
  if ((t & 63) == 0) {
#pragma unroll
    for (int k = 0; k < kRedVals; ++k) wsum[t >> 6][k] = acc[k];
  }
  __syncthreads();
  double seg = 0.0;
  if (t < kRedVals) seg = ((wsum[0][t] + wsum[1][t]) + wsum[2][t]) + wsum[3][t];
  // Each partial is written by a device-scope exchange whose result the thread waits for, so it
  // sits at the device's coherence point before the completion count is taken; the last
  // workgroup reads the partials there (agent-scope atomic loads), without cache-wide fences.
  if (!fuse_final) {  // k_reduce_systems forms the per-system sums in the next launch
    if (t < kRedVals) partials[(size_t)sgi * kRedVals + t] = seg;
    return;
  }
  if (t < kRedVals) {
    const unsigned long long old = atomicExch(
        reinterpret_cast<unsigned long long*>(partials + (size_t)sgi * kRedVals + t),
        (unsigned long long)__double_as_longlong(seg));
    __asm__ volatile("" : : "v"(old));  // keep the returning form (wait for completion)
  }
  __shared__ bool last;
  __syncthreads();
  if (t == 0) last = atomicAdd(done, 1u) == (unsigned)(n_seg - 1);
  __syncthreads();
  if (!last) return;
  for (int b = t >> 6; b < n_sys; b += blockDim.x >> 6)
    reduce_system(b, t & 63, seg_offsets, partials, energy_sys, virial_sys, kinetic_sys);
  if (t == 0) atomicExch(done, 0u);
}

__global__ void __launch_bounds__(256)
k_reduce_segments(const int32_t* __restrict__ sys_offsets, const int32_t* __restrict__ seg_offsets, int n_sys,
                  int seg_atoms, const float* __restrict__ energy_atom, const float* __restrict__ virial_atom,
                  const double* __restrict__ vel, const double* __restrict__ mass, double* __restrict__ partials,
                  unsigned* __restrict__ done, double* __restrict__ energy_sys, double* __restrict__ virial_sys,
                  double* __restrict__ kinetic_sys, int fuse_final) {
  reduce_segment_block(blockIdx.x, gridDim.x, sys_offsets, seg_offsets, n_sys, seg_atoms, energy_atom, virial_atom,
                       vel, mass, partials, done, energy_sys, virial_sys, kinetic_sys, fuse_final);
}

// Systems without segments (no atoms at all): one wave per system writes zeros.
__global__ void __launch_bounds__(64) k_reduce_systems(const int32_t* __restrict__ seg_offsets, int n_sys,
                                                       const double* __restrict__ partials,
                                                       double* __restrict__ energy_sys, double* __restrict__ virial_sys,
                                                       double* __restrict__ kinetic_sys) {
  const int b = blockIdx.x;
  if (b >= n_sys) return;
  reduce_system(b, threadIdx.x, seg_offsets, partials, energy_sys, virial_sys, kinetic_sys);
}

__global__ void k_forces(const float* __restrict__ gs, const long long* __restrict__ gn, int n3, float scale,
                         float* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) f[i] = (float)(-((double)gs[i] + fnl_from_fixed(gn[i])) * (double)scale);
}

// One workgroup (256 threads) of the BAOAB step: block blk of n_blocks covers
// atoms [256 blk, 256 blk + 256).  step_dev (optional): {step, done}.  Every
// workgroup reads the step before it counts itself in `done`; the last one to
// finish advances the step and clears `done`, so a captured MD step replays
// with the next noise without a counter kernel.  clear_gn: the
// neighbour-gradient rows are zeroed once read (the adjoint of the same MD step
// accumulates into them).
__device__ __forceinline__ void baoab_block(int blk, int n_blocks, double* __restrict__ x, double* __restrict__ v,
                                            float* __restrict__ x32, const float* __restrict__ gs,
                                            long long* __restrict__ gn, float force_scale,
                                            const double* __restrict__ dtm, double half_dt, double a1,
                                            const double* __restrict__ a2, const double* __restrict__ noise,
                                            uint64_t seed, uint64_t step_arg, uint64_t* __restrict__ step_dev,
                                            int n_atoms, int clear_gn) {
  const int i = blk * 256 + (int)threadIdx.x;
  const uint64_t step = step_dev ? step_dev[0] : step_arg;
  if (i < n_atoms) {
    double xi[3];
    if (a2 && !noise) fnl_normals3(seed, step, 0u, (uint32_t)i, xi);
    const double m = dtm[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double f = -((double)gs[3 * i + k] + fnl_from_fixed(gn[3 * i + k])) * (double)force_scale;
      double vv = v[3 * i + k] + m * f;           // B (full dt: merged half kicks)
      double xx = x[3 * i + k] + half_dt * vv;    // A
      if (a2) {                                   // O
        const double nz = noise ? noise[3 * i + k] : xi[k];
        vv = a1 * vv + a2[i] * nz;
      }
      xx = xx + half_dt * vv;                     // A
      v[3 * i + k] = vv;
      x[3 * i + k] = xx;
      x32[3 * i + k] = (float)xx;
      if (clear_gn) gn[3 * i + k] = 0;
    }
  }
  if (step_dev) {
    __syncthreads();  // every thread of this workgroup has read the step (and used it)
    if (threadIdx.x == 0) {
      const unsigned prev = atomicAdd(reinterpret_cast<unsigned*>(step_dev + 1), 1u);
      if (prev == (unsigned)(n_blocks - 1)) {
        atomicExch(reinterpret_cast<unsigned long long*>(step_dev), (unsigned long long)(step + 1));
        atomicExch(reinterpret_cast<unsigned*>(step_dev + 1), 0u);
      }
    }
  }
}

__global__ void __launch_bounds__(256)
k_baoab(double* __restrict__ x, double* __restrict__ v, float* __restrict__ x32,
        const float* __restrict__ gs, long long* __restrict__ gn, float force_scale,
        const double* __restrict__ dtm, double half_dt, double a1,
        const double* __restrict__ a2, const double* __restrict__ noise, uint64_t seed,
        uint64_t step_arg, uint64_t* __restrict__ step_dev, int n_atoms, int clear_gn) {
  baoab_block(blockIdx.x, gridDim.x, x, v, x32, gs, gn, force_scale, dtm, half_dt, a1, a2, noise, seed, step_arg,
              step_dev, n_atoms, clear_gn);
}

// The BAOAB step and the per-system energy and virial sums of the last force
// evaluation in one launch: workgroups [0, n_seg) reduce the segments (the
// last of them forms the per-system sums), workgroups [n_seg, n_seg +
// n_blocks) integrate.  The two parts touch disjoint memory (the sums read
// the atom energies and virials, which the step does not write), so they run
// side by side and the launch takes the longer of the two.
struct BaoabArgs {
  double* x;
  double* v;
  float* x32;
  const float* gs;
  long long* gn;
  float force_scale;
  const double* dtm;
  double half_dt, a1;
  const double* a2;
  const double* noise;
  uint64_t seed, step;
  uint64_t* step_dev;
  int n_atoms, clear_gn;
};

struct SumArgs {
  const int32_t* sys_offsets;
  const int32_t* seg_offsets;
  int n_sys, n_seg, seg_atoms;
  const float* energy_atom;
  const float* virial_atom;
  double* rows;
  unsigned* done;
  double* energy_sys;
  double* virial_sys;
};

__global__ void __launch_bounds__(256) k_baoab_sums(BaoabArgs b, SumArgs r) {
  if ((int)blockIdx.x < r.n_seg) {
    reduce_segment_block(blockIdx.x, r.n_seg, r.sys_offsets, r.seg_offsets, r.n_sys, r.seg_atoms, r.energy_atom,
                         r.virial_atom, nullptr, nullptr, r.rows, r.done, r.energy_sys, r.virial_sys, nullptr, 1);
  } else {
    baoab_block(blockIdx.x - r.n_seg, gridDim.x - r.n_seg, b.x, b.v, b.x32, b.gs, b.gn, b.force_scale, b.dtm,
                b.half_dt, b.a1, b.a2, b.noise, b.seed, b.step, b.step_dev, b.n_atoms, b.clear_gn);
  }
}

// Batched form: every system b carries its own step number and stream key, so
// the noise of an atom depends only on (key_b, step_b, local atom index) and
// not on where the system sits in the batch.
__global__ void k_baoab_batched(double* __restrict__ x, double* __restrict__ v, float* __restrict__ x32,
                                const float* __restrict__ gs, const long long* __restrict__ gn, float force_scale,
                                const double* __restrict__ dtm, double half_dt, double a1,
                                const double* __restrict__ a2, const int32_t* __restrict__ batch_index,
                                const int32_t* __restrict__ sys_offsets, const uint64_t* __restrict__ sys_step,
                                const uint32_t* __restrict__ sys_key, uint64_t seed, int n_atoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  const double m = dtm[i];
  if (m == 0.0) return;  // atoms of systems that are not integrated this step
  const int b = batch_index[i];
  double xi[3];
  fnl_normals3(seed, sys_step[b], sys_key[b], (uint32_t)(i - sys_offsets[b]), xi);
  fnl_baoab_row(x, v, x32, gs, gn, force_scale, m, half_dt, a1, a2[i], xi, i);
}

__global__ void k_counter_add(uint64_t* __restrict__ p, uint64_t v) {
  if (threadIdx.x == 0) *p += v;
}

__global__ void k_normals(double* __restrict__ out, int n_atoms, uint64_t seed, uint64_t step, uint32_t sid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  double z[3];
  fnl_normals3(seed, step, sid, (uint32_t)i, z);
  out[3 * i] = z[0];
  out[3 * i + 1] = z[1];
  out[3 * i + 2] = z[2];
}

}  // namespace

extern "C" int fnl_system_reduce(const int32_t* sys_offsets, const int32_t* seg_offsets, int32_t n_sys,
                                 int32_t n_seg, int32_t seg_atoms, const float* energy_atom, double* energy_sys,
                                 const float* virial_atom, double* virial_sys, const double* vel,
                                 const double* mass, double* kinetic_sys, double* partials,
                                 fnl_stream_t stream) {
  if (n_sys < 0 || n_seg < 0 || seg_atoms < 1) return FNL_ERR_INVALID;
  if (n_sys == 0) return FNL_OK;
  if (!sys_offsets || !seg_offsets || !partials) return FNL_ERR_INVALID;
  if ((energy_atom && !energy_sys) || (virial_atom && !virial_sys) || (vel && (!mass || !kinetic_sys)))
    return FNL_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  // partials: [0] completion count (zero between calls), then n_seg rows of kRedVals
  unsigned* done = reinterpret_cast<unsigned*>(partials);
  double* rows = partials + 1;
  double* e_out = energy_atom ? energy_sys : nullptr;
  double* v_out = virial_atom ? virial_sys : nullptr;
  double* k_out = vel ? kinetic_sys : nullptr;
  // few systems: the last segment workgroup forms the per-system sums (one launch); many: one
  // wave per system in a second launch (the last workgroup would loop over all of them)
  const bool fuse = n_sys <= kFuseMaxSystems;
  if (n_seg > 0) {
    hipLaunchKernelGGL(k_reduce_segments, dim3(n_seg), dim3(256), 0, st, sys_offsets, seg_offsets, n_sys,
                       seg_atoms, energy_atom, virial_atom, vel, mass, rows, done, e_out, v_out, k_out,
                       (int)fuse);
  }
  if (n_seg == 0 || !fuse) {
    hipLaunchKernelGGL(k_reduce_systems, dim3(n_sys), dim3(64), 0, st, seg_offsets, n_sys, rows, e_out, v_out,
                       k_out);
  }
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_forces(const float* grad_self, const int64_t* grad_nbr, int32_t n_atoms, float scale,
                          float* forces, fnl_stream_t stream) {
  if (n_atoms < 0) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!grad_self || !grad_nbr || !forces) return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_forces, dim3(fnl_blocks(3LL * n_atoms, 256)), dim3(256), 0, (hipStream_t)stream,
                     grad_self, (const long long*)grad_nbr, 3 * n_atoms, scale, forces);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_baoab_step(double* x, double* v, float* x32, const float* grad_self, int64_t* grad_nbr,
                              float force_scale, const double* dtm, double half_dt, double a1, const double* a2,
                              const double* noise, uint64_t seed, uint64_t step, uint64_t* step_dev,
                              int32_t n_atoms, int32_t clear_grad_nbr, fnl_stream_t stream) {
  if (n_atoms < 0) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!x || !v || !x32 || !grad_self || !grad_nbr || !dtm) return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_baoab, dim3(fnl_blocks(n_atoms, 256)), dim3(256), 0, (hipStream_t)stream, x, v, x32,
                     grad_self, (long long*)grad_nbr, force_scale, dtm, half_dt, a1, a2, noise, seed, step, step_dev, n_atoms,
                     (int)(clear_grad_nbr != 0));
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_baoab_step_sums(double* x, double* v, float* x32, const float* grad_self, int64_t* grad_nbr,
                                   float force_scale, const double* dtm, double half_dt, double a1, const double* a2,
                                   const double* noise, uint64_t seed, uint64_t step, uint64_t* step_dev,
                                   int32_t n_atoms, int32_t clear_grad_nbr, const int32_t* sys_offsets,
                                   const int32_t* seg_offsets, int32_t n_sys, int32_t n_seg, int32_t seg_atoms,
                                   const float* energy_atom, double* energy_sys, const float* virial_atom,
                                   double* virial_sys, double* partials, fnl_stream_t stream) {
  if (n_atoms < 0 || n_sys < 0 || n_seg < 0 || seg_atoms < 1) return FNL_ERR_INVALID;
  if (!x || !v || !x32 || !grad_self || !grad_nbr || !dtm) return FNL_ERR_INVALID;
  if (!sys_offsets || !seg_offsets || !partials || !energy_atom || !energy_sys) return FNL_ERR_INVALID;
  if (virial_atom && !virial_sys) return FNL_ERR_INVALID;
  // the last segment workgroup forms the per-system sums itself only for a few systems
  if (n_sys > kFuseMaxSystems) return FNL_ERR_UNSUPPORTED;
  if (n_atoms == 0 || n_sys == 0 || n_seg == 0) return FNL_ERR_INVALID;
  BaoabArgs b{x, v, x32, grad_self, (long long*)grad_nbr, force_scale, dtm, half_dt, a1, a2, noise, seed, step,
              step_dev, n_atoms, (int)(clear_grad_nbr != 0)};
  SumArgs r{sys_offsets, seg_offsets, n_sys, n_seg, seg_atoms, energy_atom, virial_atom, partials + 1,
            reinterpret_cast<unsigned*>(partials), energy_sys, virial_atom ? virial_sys : nullptr};
  const int n_blocks = fnl_blocks(n_atoms, 256);
  hipLaunchKernelGGL(k_baoab_sums, dim3(n_seg + n_blocks), dim3(256), 0, (hipStream_t)stream, b, r);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_baoab_step_batched(double* x, double* v, float* x32, const float* grad_self,
                                      const int64_t* grad_nbr, float force_scale, const double* dtm, double half_dt,
                                      double a1, const double* a2, const int32_t* batch_index,
                                      const int32_t* sys_offsets, const uint64_t* sys_step, const uint32_t* sys_key,
                                      uint64_t seed, int32_t n_atoms, fnl_stream_t stream) {
  if (n_atoms < 0) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!x || !v || !x32 || !grad_self || !grad_nbr || !dtm || !a2 || !batch_index || !sys_offsets || !sys_step ||
      !sys_key)
    return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_baoab_batched, dim3(fnl_blocks(n_atoms, 256)), dim3(256), 0, (hipStream_t)stream, x, v, x32,
                     grad_self, (const long long*)grad_nbr, force_scale, dtm, half_dt, a1, a2, batch_index, sys_offsets, sys_step,
                     sys_key, seed, n_atoms);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_counter_add(uint64_t* counter, uint64_t value, fnl_stream_t stream) {
  if (!counter) return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, (hipStream_t)stream, counter, value);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_philox_normals(double* out, int32_t n_atoms, uint64_t seed, uint64_t step, uint32_t stream_id,
                                  fnl_stream_t stream) {
  if (n_atoms < 0) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!out) return FNL_ERR_INVALID;
  hipLaunchKernelGGL(k_normals, dim3(fnl_blocks(n_atoms, 256)), dim3(256), 0, (hipStream_t)stream, out, n_atoms,
                     seed, step, stream_id);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
