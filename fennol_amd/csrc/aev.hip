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
// Forward (per centre i), one wave:
//   1. radial rows (vector, distance, species) and the angular row -> LDS
//   2. radial slots sorted by species (ballot counting sort); each species
//      block summed over its neighbours in registers, lanes = (neighbour
//      group, radial shift), one store per channel
//   3. angular slots sorted by species; the triplets of the present species
//      pairs enumerated pair-major in chunks of 64: one lane per triplet
//      evaluates the R_d and F_z factors, then lanes = (half, channel) sum
//      R_d F_z over each pair's triplets in registers, one store per channel
//   4. empty blocks and the row padding written as zeros
// Adjoint: the same terms are recomputed.  Radial: the lane of slot e sums
// dE_i/dvec_ie from the centre's dE/dAEV row and dE_j/dvec_ji from the
// neighbour's row, so radial forces are gathered, not scattered.  Angular:
// triplets in the rounds of the circle method (every slot in one pair of each
// round), per-round terms stored to LDS and summed per slot in round order;
// the angular slot gradients are then scattered to grad_nbr with global int64
// atomics on 2^-32 fixed point (exact, order-independent sums: the same bits
// every run).  grad_self[i] is written once, virial_atom[i] = sum vec (x) dE/dvec.

#include <hip/hip_bfloat16.h>

#include "fnl_common.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kLog2e = 1.44269504088896340736f;

// Hardware transcendentals (v_exp_f32, v_log_f32, v_sqrt_f32, v_rcp_f32, v_cos_f32, v_sin_f32):
// about one ulp and no special-case code.  The arguments here stay in the ranges where they are
// exact to that level: distances and angles of atoms within the cutoff, bases of the angular
// power clamped away from zero.
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// cos(pi t) and sin(pi t): the hardware takes the angle in turns
__device__ __forceinline__ float hw_cospi(float t) { return __builtin_amdgcn_cosf(0.5f * t); }
__device__ __forceinline__ float hw_sinpi(float t) { return __builtin_amdgcn_sinf(0.5f * t); }
// exp(-eta x^2)
__device__ __forceinline__ float gauss(float x, float eta) { return hw_exp2(x * x * (-eta * kLog2e)); }
// (1 + cos(theta - theta_z))^zeta for a base b = 1 + cos(..) >= 0 up to rounding and zeta > 0: |b| keeps a
// rounded negative base away from log2's NaN, and b = 0 gives exp2(-inf) = 0
__device__ __forceinline__ float ang_pow(float b, float zeta) { return hw_exp2(zeta * hw_log2(fabsf(b))); }

// Each wave owns its LDS slice, so the lanes of one wave only exchange data among
// themselves: the LDS executes a wave's DS instructions in issue order, and wavefront-scope
// fences keep the compiler from moving LDS accesses across this point without waiting for
// outstanding global loads (a workgroup-scope fence would drain them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float cos_switch(float d, float rc) {
  return d < rc ? 0.5f * hw_cospi(d * hw_rcp(rc)) + 0.5f : 0.0f;
}

__device__ __forceinline__ float cos_switch_grad(float d, float rc) {
  return d < rc ? (-0.5f * kPi) * hw_rcp(rc) * hw_sinpi(d * hw_rcp(rc)) : 0.0f;
}

// map a linear index t < n(n-1)/2 to the pair (a, b), a < b, row-major.  Row r starts at
// start(r) = r n - r(r+1)/2; r is the root of start(r) = t rounded down, from a float square
// root whose operand (2n-1)^2 - 8t is exact (n < 2^11) and whose result is within a few ulp, so
// the estimate is off by at most one row and one correction each way is exact.
__device__ __forceinline__ void tri_pair(int t, int n, int& a, int& b) {
  const float nn = (float)(2 * n - 1);
  int r = (int)floorf(0.5f * (nn - hw_sqrt(nn * nn - 8.0f * (float)t)));
  r = max(0, min(r, n - 2));
  if (r > 0 && r * n - (r * (r + 1)) / 2 > t) --r;
  if (r < n - 2 && (r + 1) * n - ((r + 1) * (r + 2)) / 2 <= t) ++r;
  a = r;
  b = t - (r * n - (r * (r + 1)) / 2) + r + 1;
}

// map a linear index k < m(m+1)/2 to the block (A, B), A <= B, of the upper triangle of an
// m x m table with its diagonal, row-major: row A starts at A m - A(A-1)/2 (same rounding
// argument as tri_pair, m < 2^11).
__device__ __forceinline__ void tri_block(int k, int m, int& A, int& B) {
  const float mm = (float)(2 * m + 1);
  int r = (int)floorf(0.5f * (mm - hw_sqrt(fmaxf(mm * mm - 8.0f * (float)k, 0.0f))));
  r = max(0, min(r, m - 1));
  if (r > 0 && r * m - (r * (r - 1)) / 2 > k) --r;
  if (r < m - 1 && (r + 1) * m - ((r + 1) * r) / 2 <= k) ++r;
  A = r;
  B = r + (k - (r * m - (r * (r - 1)) / 2));
}

template <typename T>
__device__ __forceinline__ T to_out(float v);
template <>
__device__ __forceinline__ float to_out<float>(float v) { return v; }
template <>
__device__ __forceinline__ hip_bfloat16 to_out<hip_bfloat16>(float v) {
  // round to nearest even in one hardware conversion (v_cvt_pk_bf16_f32), the rounding of
  // hip_bfloat16(float) for every finite value
  hip_bfloat16 h;
  h.data = __builtin_bit_cast(unsigned short, static_cast<__bf16>(v));
  return h;
}

// dE/dAEV rows: float32, or bfloat16 bits (the BF16 form)
__device__ __forceinline__ float g2f(float v) { return v; }
__device__ __forceinline__ float g2f(unsigned short v) { return __uint_as_float((uint32_t)v << 16); }

constexpr int kMaxGroups = 192;  // 8-channel groups of the longest AEV (8 species: 176)

struct AevDims {
  int S, P, nR, rad_len, aev_len;
  uint64_t live;  // adjoint: channel blocks (radial s at bit s, angular pair p at bit S+p) read from g
  int n_groups;   // adjoint: the live blocks as 8-channel groups (first channel / 8), or 0
  uint16_t groups[kMaxGroups];
  // Gaussians as exp2(-y^2) of scaled distances, y = k (d - shift) with k = sqrt(eta log2 e);
  // rs / as are the radial / angular shifts times k, ka_half is k_a / 2 (the scaled mean distance of
  // a triplet is ka_half (d_a + d_b)); for the derivatives 2 eta x = br y (radial, br = 2 eta_r / k_r)
  // and eta_a x = ang_b y (angular, ang_b = eta_a / k_a)
  float kr, ka_half, br, ang_b;
  float rs[16], as[16];
};

// blocks of the species in species_mask and of the pairs of two such species (all blocks when
// the mask is ~0u or a 4-channel chunk can straddle two radial blocks)
inline uint64_t live_blocks(uint32_t species_mask, const AevDims& dm) {
  if (species_mask == 0xFFFFFFFFu || dm.nR % 4 != 0) return ~0ull;
  uint64_t live = 0;
  int pr = 0;
  for (int a = 0; a < dm.S; ++a) {
    if ((species_mask >> a) & 1u) live |= 1ull << a;
    for (int b = a; b < dm.S; ++b, ++pr)  // pair_index order: row-major upper triangle
      if (((species_mask >> a) & 1u) && ((species_mask >> b) & 1u)) live |= 1ull << (dm.S + pr);
  }
  return live;
}

// 4-channel chunks of the blocks in dm.live (the others are never read: no neighbour or
// neighbour pair of their species exists in the batch)
template <typename GT>
__device__ __forceinline__ void load_g_row(float* G, const GT* __restrict__ row, const AevDims& dm, int lane) {
  if (dm.n_groups > 0) {  // one 16-byte (BF16) or two (FP32) loads per live 8-channel group
    for (int q = lane; q < dm.n_groups; q += 64) {
      const int c = 8 * (int)dm.groups[q];
      if constexpr (sizeof(GT) == 4) {
        ((float4*)(G + c))[0] = ((const float4*)(row + c))[0];
        ((float4*)(G + c))[1] = ((const float4*)(row + c))[1];
      } else {
        const uint4 u = *(const uint4*)(row + c);
        ((float4*)(G + c))[0] = make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                                            __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
        ((float4*)(G + c))[1] = make_float4(__uint_as_float(u.z << 16), __uint_as_float(u.z & 0xffff0000u),
                                            __uint_as_float(u.w << 16), __uint_as_float(u.w & 0xffff0000u));
      }
    }
    return;
  }
  const int n = dm.aev_len;
  auto live = [&](int c) {
    const int ch = 4 * c;
    const int blk = ch < dm.rad_len ? ch / dm.nR : dm.S + ((ch - dm.rad_len) >> 5);
    return ((dm.live >> blk) & 1ull) != 0;
  };
  if constexpr (sizeof(GT) == 4) {
    const float4* r4 = (const float4*)row;
    for (int c = lane; c < (n + 3) / 4; c += 64)
      if (live(c)) ((float4*)G)[c] = r4[c];
  } else {
    const ushort4* r4 = (const ushort4*)row;
    for (int c = lane; c < (n + 3) / 4; c += 64) {
      if (!live(c)) continue;
      const ushort4 u = r4[c];
      ((float4*)G)[c] = make_float4(g2f(u.x), g2f(u.y), g2f(u.z), g2f(u.w));
    }
  }
}

__device__ __forceinline__ int pair_index(int a, int b, int S) {
  // row-major upper triangle incl. diagonal of an S x S table (ani.py:124-130)
  const int lo = min(a, b), hi = max(a, b);
  return lo * S - (lo * (lo - 1)) / 2 + (hi - lo);
}

constexpr int kFacLd = 65;     // row stride of the forward triplet-factor table (odd: the 4 R rows and
                               // 8 F rows read at one column fall in different banks)
constexpr int kTabWords = 16;  // per-wave table: species counts and offsets

__host__ __device__ inline int round4(int n) { return (n + 3) & ~3; }

// per-wave LDS layout of the forward kernel (offsets in 4-byte words, 16-byte aligned)
struct FwdLayout {
  int nbt, rdf, av, at, af, ag, fac, tab, words;
};
// per-wave LDS slice of the forward (4-byte words): the triplet-factor table is only used after
// the slot vectors (nbv) and the sorted radial slots (rdf) are dead, so it shares their space;
// 8 waves per SIMD at the water box's row capacities (104 radial, 40 angular slots).  The
// tangent (tan) adds the slot displacements (nbt), the angular slots' unit-vector and distance
// tangents (at) and switch derivatives (ag), float4 radial entries and a table twice as tall.
__host__ __device__ inline FwdLayout fwd_layout(int max_nbr, int max_nbr_ang, int na_nz, bool tan = false) {
  FwdLayout L;
  L.fac = 0;            // nbv at 0 as well
  L.nbt = 4 * max_nbr;  // tangent only
  L.rdf = (tan ? 8 : 4) * max_nbr;
  const int lo = L.rdf + (tan ? 4 : 2) * max_nbr, fw = (tan ? 2 : 1) * na_nz * kFacLd;
  L.av = round4(lo > fw ? lo : fw);
  L.at = L.av + 4 * max_nbr_ang;                   // tangent only
  L.af = L.at + (tan ? 4 * max_nbr_ang : 0);
  L.ag = L.af + 2 * max_nbr_ang;                   // tangent only
  L.tab = round4(L.ag + (tan ? 2 * max_nbr_ang : 0));
  L.words = round4(L.tab + kTabWords);
  return L;
}

__device__ __forceinline__ void lds_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// DPP lane moves (VALU, no LDS traffic): row_ror:n = 0x120 + n, quad_perm = 0x00..0xFF
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// sum over the wave, returned in every lane: rotations within each row of 16 lanes, then the
// four row sums read as scalars
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += dpp<0x128>(v);
  v += dpp<0x124>(v);
  v += dpp<0x122>(v);
  v += dpp<0x121>(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return (r0 + r1) + (r2 + r3);
}

// Stable counting sort of a row of n entries by species (values < S, S <= FNL_MAX_SPECIES),
// with ballots: on return cnt[s] / off[s] hold the per-species counts and start positions
// (wave-uniform), and pos_of(e) is called by the lane that owns entry e with its sorted position.
template <typename SpecFn, typename PlaceFn>
__device__ __forceinline__ void species_sort(int n, int S, SpecFn spec, PlaceFn place,
                                             int (&cnt)[FNL_MAX_SPECIES], int (&off)[FNL_MAX_SPECIES],
                                             uint32_t smask) {
  const int lane = fnl_lane();
  const uint64_t below = (1ull << lane) - 1ull;
  if (n <= 64) {  // one ballot per species serves both the counts and the ranks
    const int sa = lane < n ? spec(lane) : S;
    int pos = 0, acc = 0;
#pragma unroll
    for (int s = 0; s < FNL_MAX_SPECIES; ++s) {
      cnt[s] = 0;
      off[s] = acc;
      if (s < S && ((smask >> s) & 1u)) {  // species outside the batch have no slot
        const uint64_t m = __ballot(sa == s);
        cnt[s] = __popcll(m);
        if (sa == s) pos = acc + __popcll(m & below);
        acc += cnt[s];
      }
    }
    if (lane < n) place(lane, pos, sa);
    return;
  }
#pragma unroll
  for (int s = 0; s < FNL_MAX_SPECIES; ++s) cnt[s] = 0;
  for (int c = 0; c < n; c += 64) {
    const int sb = c + lane < n ? spec(c + lane) : S;
#pragma unroll
    for (int s = 0; s < FNL_MAX_SPECIES; ++s)
      if (s < S && ((smask >> s) & 1u)) cnt[s] += __popcll(__ballot(sb == s));
  }
  int acc = 0;
#pragma unroll
  for (int s = 0; s < FNL_MAX_SPECIES; ++s) {
    off[s] = acc;
    acc += cnt[s];
  }
  int run[FNL_MAX_SPECIES];
#pragma unroll
  for (int s = 0; s < FNL_MAX_SPECIES; ++s) run[s] = off[s];
  for (int c = 0; c < n; c += 64) {
    const int e = c + lane;
    const int sa = e < n ? spec(e) : S;
    int pos = 0;
#pragma unroll
    for (int s = 0; s < FNL_MAX_SPECIES; ++s) {
      if (s < S && ((smask >> s) & 1u)) {  // species outside the batch have no slot
        const uint64_t m = __ballot(sa == s);
        if (sa == s) pos = run[s] + __popcll(m & below);
        run[s] += __popcll(m);
      }
    }
    if (e < n) place(e, pos, sa);
  }
}

// Forward, one wave per centre.  LDS slice of the wave (4-byte words):
//   nbv [max_nbr]      float4 (vx, vy, vz, species bits) per radial slot (slot order);
//                      d is recomputed from the vector where it is needed
//   rdf [max_nbr]      float2 (d, 0.25 fc(d)), radial slots sorted by species
//   av  [max_nbr_ang]  float4 (ux, uy, uz, d), angular slots sorted by species
//   af  [max_nbr_ang]  float2 (fc_ang(d), species), same order
//   fac [NA+NZ][65]    per-triplet factors R_d and F_z of one chunk of 64
//                      triplets, component-major (lane t owns column t)
//   tab [16]           species counts and offsets in the sorted angular row
// Radial: lanes (q = lane/16, k = lane%16); for each species s the lanes sum
// 0.25 fc exp(-eta (d - R_k)^2) over neighbours q, q+4, ... of the species-s
// segment in a register, reduce over q with two shuffles and store channel
// s*nR + k.  Angular: triplets are enumerated pair-major -- all triplets of
// species pair p (a from the s1 segment, b from the s2 segment, a < b) are
// consecutive -- so the sum over a pair's triplets is a branch-free loop.
// Per chunk of 64 triplets, lane t evaluates R_d (NA) and F_z (NZ) of
// triplet t into fac; then lane (h = lane/32, c = d*NZ + z = lane%32) adds
// R_d F_z of the triplets h, h+2, ... of each pair segment in the chunk to a
// register; when a segment ends, the halves combine (one shuffle) and the 32
// channels of the pair are stored.  Every channel is written exactly once.
//
// TAN: the tangent (forward-mode JVP) of the AEV along a displacement field r instead,
//   dA_i = sum_e dA_i/dvec_e . (r_j - r_i)
// (the same edge vectors and terms; used for weight gradients through forces, where
// F . r = -sum_i dE/dA_i . dA_i).  Radial: d/dd (0.25 fc exp(-eta x^2)) dd.  Angular, per triplet
// the tangent of sw F_z(c) R_d(m) is F_z (dsw R_d + sw R_d' dm) + F_z' (sw R_d dc) =
// F_z P_d + F_z' Q_d: the table holds P_d, Q_d, F_z, F_z' and each channel adds two products.
template <int NA, int NZ, typename TOut, bool TAN = false>
__global__ void __launch_bounds__(64 * kWaves)
k_aev_fwd(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
          const int32_t* __restrict__ batch_index, int n_atoms, int pbc_mode,
          const float* __restrict__ cells, const float* __restrict__ rec,
          const int32_t* __restrict__ nbr, const int32_t* __restrict__ n_nbr, int max_nbr,
          const int32_t* __restrict__ nbr_shift,
          const int32_t* __restrict__ nbr_ang, const int32_t* __restrict__ n_nbr_ang, int max_nbr_ang,
          fnl_aev_params p, AevDims dm, TOut* __restrict__ aev, int stride, uint32_t smask,
          const float* __restrict__ disp) {
  static_assert(NA * NZ == 32, "angular channel layout");
  extern __shared__ float4 lds4[];
  const int w = threadIdx.x >> 6;
  const int lane = fnl_lane();
  const int i = blockIdx.x * kWaves + w;
  if (i >= n_atoms) return;
  const FwdLayout L = fwd_layout(max_nbr, max_nbr_ang, NA + NZ, TAN);
  float* base = (float*)lds4 + (size_t)w * L.words;
  float4* nbv = (float4*)base;
  float4* nbt = (float4*)(base + L.nbt);
  float2* rdf = (float2*)(base + L.rdf);
  float4* rdf4 = (float4*)(base + L.rdf);
  float4* av = (float4*)(base + L.av);
  float4* at = (float4*)(base + L.at);
  float2* af = (float2*)(base + L.af);
  float2* ag = (float2*)(base + L.ag);
  float* fac = base + L.fac;
  int* tab = (int*)(base + L.tab);

  TOut* out = aev + (size_t)i * stride;
  const int si = species_idx[i];
  if (si < 0) {
    for (int c = lane; c < stride; c += 64) out[c] = to_out<TOut>(0.0f);
    return;
  }
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  FnlCell cl;
  if (pbc || nbr_shift) fnl_load_cell(cells, rec, batch_index[i], cl);
  const int nr = min(n_nbr[i], max_nbr);
  const int na = min(n_nbr_ang[i], max_nbr_ang);
  const int S = dm.S;
  {
    const float xi = coords[3 * i], yi = coords[3 * i + 1], zi = coords[3 * i + 2];
    const int32_t* row = nbr + (size_t)i * max_nbr;
    float ri0 = 0.0f, ri1 = 0.0f, ri2 = 0.0f;
    if constexpr (TAN) {
      ri0 = disp[3 * i];
      ri1 = disp[3 * i + 1];
      ri2 = disp[3 * i + 2];
    }
    for (int e = lane; e < nr; e += 64) {
      const int j = row[e];
      float vx = coords[3 * j] - xi, vy = coords[3 * j + 1] - yi, vz = coords[3 * j + 2] - zi;
      if (nbr_shift)  // general PBC: the stored image (GraphProcessor, preprocessing.py:727-735)
        fnl_apply_shift(vx, vy, vz, cl, nbr_shift[(size_t)i * max_nbr + e]);
      else if (pbc)
        fnl_min_image(vx, vy, vz, cl);
      nbv[e] = make_float4(vx, vy, vz, __int_as_float(species_idx[j]));
      if constexpr (TAN) nbt[e] = make_float4(disp[3 * j] - ri0, disp[3 * j + 1] - ri1, disp[3 * j + 2] - ri2, 0.0f);
    }
  }
  // the angular row (radial slot numbers): entries < 64 in a register, fetched with the radial row
  const int32_t* arow = nbr_ang + (size_t)i * max_nbr_ang;
  const int ae0 = lane < na ? arow[lane] : 0;
  auto ae_of = [&](int a) { return a < 64 ? ae0 : arow[a]; };  // a is the lane's own entry
  auto dist = [](const float4& v) { return hw_sqrt(fmaxf(v.x * v.x + v.y * v.y + v.z * v.z, 1.0e-5f)); };
  // tangent of the distance along the slot displacement t (0 for coincident atoms)
  auto dist_dot = [](const float4& v, const float4& t, float d) {
    const float d2 = v.x * v.x + v.y * v.y + v.z * v.z;
    return d2 > 1.0e-5f ? (v.x * t.x + v.y * t.y + v.z * t.z) * hw_rcp(d) : 0.0f;
  };
  wave_sync();

  int cnt[FNL_MAX_SPECIES], off[FNL_MAX_SPECIES];
  // radial slots sorted by species
  species_sort(
      nr, S, [&](int e) { return __float_as_int(nbv[e].w); },
      [&](int e, int pos, int) {
        const float d = dist(nbv[e]);
        if constexpr (TAN)
          rdf4[pos] = make_float4(d, 0.25f * cos_switch(d, p.cutoff), 0.25f * cos_switch_grad(d, p.cutoff),
                                  dist_dot(nbv[e], nbt[e], d));
        else
          rdf[pos] = make_float2(d * dm.kr, 0.25f * cos_switch(d, p.cutoff));
      },
      cnt, off, smask);
  wave_sync();

  const int q = lane >> 4, k = lane & 15;
  uint32_t rad_present = 0;  // species with at least one radial neighbour
  {
    const float shift = k < dm.nR ? (TAN ? p.radial_shifts[k] : dm.rs[k]) : 0.0f;
#pragma unroll
    for (int s = 0; s < FNL_MAX_SPECIES; ++s) {
      if (s >= S) break;
      if (cnt[s] == 0) continue;  // written as zeros with the other empty blocks
      rad_present |= 1u << s;
      const int lo = off[s], hi = off[s] + cnt[s];
      float run = 0.0f;
#pragma unroll 4
      for (int e = lo + q; e < hi; e += 4) {
        if constexpr (TAN) {  // (0.25 fc' - 2 eta x 0.25 fc) exp(-eta x^2) dd
          const float4 df = rdf4[e];
          const float x = df.x - shift;
          run += (df.z - 2.0f * p.radial_eta * x * df.y) * gauss(x, p.radial_eta) * df.w;
        } else {
          const float2 df = rdf[e];  // (k_r d, 0.25 fc)
          const float y = df.x - shift;
          run += df.y * hw_exp2(-(y * y));
        }
      }
      run += __shfl_xor(run, 16, 64);
      run += __shfl_xor(run, 32, 64);
      if (q == 0 && k < dm.nR) out[s * dm.nR + k] = to_out<TOut>(run);
    }
  }

  // angular slots sorted by species: unit vector, distance, switch
  {
    species_sort(
        na, S, [&](int a) { return __float_as_int(nbv[ae_of(a)].w); },
        [&](int a, int pos, int sa) {
          const int e = ae_of(a);
          const float4 v = nbv[e];
          const float d = dist(v);
          const float inv = hw_rcp(fmaxf(d, 1.0e-5f));
          const float ux = v.x * inv, uy = v.y * inv, uz = v.z * inv;
          av[pos] = make_float4(ux, uy, uz, TAN ? d : d * dm.ka_half);  // forward: the scaled half distance
          af[pos] = make_float2(cos_switch(d, p.cutoff_ang), __int_as_float(sa));
          if constexpr (TAN) {  // du = (t - u (u . t)) / d, dd
            const float4 t = nbt[e];
            const float ut = ux * t.x + uy * t.y + uz * t.z;
            at[pos] = make_float4((t.x - ux * ut) * inv, (t.y - uy * ut) * inv, (t.z - uz * ut) * inv,
                                  dist_dot(v, t, d));
            ag[pos] = make_float2(cos_switch_grad(d, p.cutoff_ang), 0.0f);
          }
        },
        cnt, off, smask);
  }
  // species pairs: lane p < P holds pair p's species blocks in the sorted angular row (sizes,
  // offsets) and its triplet count; the present pairs form a wave-uniform bit mask
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < FNL_MAX_SPECIES; ++s) {
      tab[s] = cnt[s];
      tab[8 + s] = off[s];
    }
  }
  wave_sync();
  const int P = dm.P;
  int pn1 = 0, pn2 = 0, po1 = 0, po2 = 0, np = 0;
  bool live = false;  // both species of the pair in smask: its block is written
  if (lane < P) {
    int s1 = 0, s2 = 0, r0 = 0;
    for (int a = 0; a < S; ++a) {
      if (lane < r0 + (S - a)) {
        s1 = a;
        s2 = a + (lane - r0);
        break;
      }
      r0 += S - a;
    }
    pn1 = tab[s1];
    pn2 = tab[s2];
    po1 = tab[8 + s1];
    po2 = tab[8 + s2];
    np = s1 == s2 ? pn1 * (pn1 - 1) / 2 : pn1 * pn2;
    live = ((smask >> s1) & (smask >> s2) & 1u) != 0u;
  }
  const uint64_t present = __ballot(np > 0);
  const uint64_t pair_live = __ballot(live);

  const float pref = 2.0f * exp2f(-p.zeta);  // 2 * 0.5^zeta (ani.py:112)
  const int h = lane >> 5, ch = lane & 31;
  const int dd = ch / NZ, zz = ch % NZ;
  // table rows: forward R_d, sw F_z; tangent P_d, Q_d, F_z, F_z'
  const float* fR = fac + dd * kFacLd;
  const float* fF = fac + ((TAN ? 2 * NA : NA) + zz) * kFacLd;
  const float* fQ = fac + (NA + dd) * kFacLd;
  const float* fG = fac + (2 * NA + NZ + zz) * kFacLd;
  TOut* aout = out + dm.rad_len;
  // empty blocks (species without radial neighbours, pairs without triplets) and the padding:
  // zeros in 16-byte stores.  Blocks of species outside smask (and, unless smask holds every
  // species, the padding) are not written: the caller keeps them zero.
  const bool pad_live = smask == 0xFFFFFFFFu;
  const uint64_t pmask = dm.P >= 64 ? ~0ull : ((1ull << dm.P) - 1ull);
  const bool any_zero = ((smask & ~rad_present) & ((1u << S) - 1u)) != 0u ||
                        ((pair_live & ~present) & pmask) != 0ull || (pad_live && stride > dm.aev_len);
  if (any_zero) {  // (most centres have every written block non-empty)
    constexpr int CH = 16 / (int)sizeof(TOut);  // channels per 16-byte chunk
    // radial block of channel c: (c + 0.5) / nR rounded down, exact in float for these sizes
    const float inv_nr = 1.0f / (float)dm.nR;
    const bool aligned = (dm.nR % CH) == 0 && (dm.rad_len % CH) == 0 && (stride % CH) == 0 &&
                         ((uintptr_t)aev & 15) == 0;
    if (aligned) {
      for (int c0 = lane * CH; c0 < stride; c0 += 64 * CH) {
        bool empty, wr;
        if (c0 < dm.rad_len) {
          const int sb = (int)(((float)c0 + 0.5f) * inv_nr);
          empty = !((rad_present >> sb) & 1u);
          wr = (smask >> sb) & 1u;
        } else if (c0 < dm.aev_len) {
          const int pb = (c0 - dm.rad_len) >> 5;
          empty = !((present >> pb) & 1ull);
          wr = (pair_live >> pb) & 1ull;
        } else {
          empty = true;
          wr = pad_live;
        }
        if (empty && wr) *(uint4*)(out + c0) = make_uint4(0u, 0u, 0u, 0u);
      }
    } else {
      for (int c = lane; c < stride; c += 64) {
        bool empty, wr;
        if (c < dm.rad_len) {
          const int sb = (int)(((float)c + 0.5f) * inv_nr);
          empty = !((rad_present >> sb) & 1u);
          wr = (smask >> sb) & 1u;
        } else if (c < dm.aev_len) {
          const int pb = (c - dm.rad_len) >> 5;
          empty = !((present >> pb) & 1ull);
          wr = (pair_live >> pb) & 1ull;
        } else {
          empty = true;
          wr = pad_live;
        }
        if (empty && wr) out[c] = to_out<TOut>(0.0f);
      }
    }
  }
  // Triplets of the present pairs, pair-major and concatenated (pair p's triplets start at
  // tstart of lane p), in chunks of 64: lane t evaluates triplet u0 + t into fac (its pair is
  // the last present pair starting at or before it), then lane (h, c) adds R_d F_z over the
  // chunk's triplets h, h+2, ... of each pair segment in the chunk to a register, and stores
  // the pair's 32 channels when its segment ends.  Pairs share chunks, so no lanes idle at a
  // pair boundary.
  const int t_incl = fnl_wave_incl_scan(np);
  const int tstart = t_incl - np;
  const int T = __builtin_amdgcn_readlane(t_incl, 63);
  float run = 0.0f;
  for (int u0 = 0; u0 < T; u0 += 64) {
    const int u = u0 + lane;
    const int c_end = min(u0 + 64, T);
    int n1 = 1, n2 = 1, o1 = 0, o2 = 0, ts = 0;
    for (uint64_t pm = present; pm;) {
      const int pr = __builtin_ctzll(pm);
      pm &= pm - 1ull;
      const int s0 = __builtin_amdgcn_readlane(tstart, pr);
      if (s0 >= c_end) break;
      const int m1 = __builtin_amdgcn_readlane(pn1, pr), m2 = __builtin_amdgcn_readlane(pn2, pr);
      const int q1 = __builtin_amdgcn_readlane(po1, pr), q2 = __builtin_amdgcn_readlane(po2, pr);
      if (u >= s0) {
        ts = s0;
        n1 = m1;
        n2 = m2;
        o1 = q1;
        o2 = q2;
      }
    }
    if (u < T) {
      const int uu = u - ts;
      int a, b;
      if (o1 == o2) {
        tri_pair(uu, n1, a, b);
        a += o1;
        b += o1;
      } else {
        int qd = (int)((float)uu * hw_rcp((float)n2));
        if (qd * n2 > uu) --qd;
        if ((qd + 1) * n2 <= uu) ++qd;
        a = o1 + qd;
        b = o2 + (uu - qd * n2);
      }
      const float4 ua = av[a], ub = av[b];
      const float c = ua.x * ub.x + ua.y * ub.y + ua.z * ub.z;
      const float x = 0.95f * c;
      const float sw = pref * af[a].x * af[b].x;
      const float m = 0.5f * (ua.w + ub.w);
      if constexpr (TAN) {
        const float4 ta = at[a], tb = at[b];
        const float dc = ta.x * ub.x + ta.y * ub.y + ta.z * ub.z + ua.x * tb.x + ua.y * tb.y + ua.z * tb.z;
        const float st = hw_sqrt(fmaxf(1.0f - x * x, 1.0e-12f));
        const float ist = hw_rcp(st);
        const float dsw = pref * (ag[a].x * af[b].x * ta.w + af[a].x * ag[b].x * tb.w);
        const float dm_ = 0.5f * (ta.w + tb.w);
#pragma unroll
        for (int d = 0; d < NA; ++d) {
          const float y = m - p.ang_shifts[d];
          const float R = gauss(y, p.angular_eta);
          fac[d * kFacLd + lane] = dsw * R + sw * (-2.0f * p.angular_eta * y * R) * dm_;  // P_d
          fac[(NA + d) * kFacLd + lane] = sw * R * dc;                                       // Q_d
        }
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
          const float bz = 1.0f + x * p.ang_cos[z] + st * p.ang_sin[z];
          float F = 0.0f, dF = 0.0f;
          if (bz > 0.0f) {
            const float lb = hw_log2(bz);
            const float tz = p.zeta * lb;
            F = hw_exp2(tz);
            dF = (0.95f * p.zeta) * hw_exp2(tz - lb) * (p.ang_cos[z] - x * p.ang_sin[z] * ist);
          }
          fac[(2 * NA + z) * kFacLd + lane] = F;
          fac[(2 * NA + NZ + z) * kFacLd + lane] = dF;
        }
      } else {
        const float st = hw_sqrt(fmaxf(1.0f - x * x, 0.0f));
#pragma unroll
        for (int d = 0; d < NA; ++d) {
          const float y = (ua.w + ub.w) - dm.as[d];  // k_a (m - R_d)
          fac[d * kFacLd + lane] = hw_exp2(-(y * y));
        }
#pragma unroll
        for (int z = 0; z < NZ; ++z)
          fac[(NA + z) * kFacLd + lane] = sw * ang_pow(1.0f + x * p.ang_cos[z] + st * p.ang_sin[z], p.zeta);
      }
    }
    wave_sync();
    for (uint64_t pm = present; pm;) {
      const int pr = __builtin_ctzll(pm);
      pm &= pm - 1ull;
      const int s0 = __builtin_amdgcn_readlane(tstart, pr);
      if (s0 >= c_end) break;
      const int e0 = s0 + __builtin_amdgcn_readlane(np, pr);
      if (e0 <= u0) continue;
      const int hi = min(e0, c_end) - u0;
#pragma unroll 4
      for (int tt = max(s0, u0) - u0 + h; tt < hi; tt += 2) {
        if constexpr (TAN)
          run += fR[tt] * fF[tt] + fQ[tt] * fG[tt];
        else
          run += fR[tt] * fF[tt];
      }
      if (e0 <= c_end) {  // the pair's last triplet is in this chunk
        const float v = run + __shfl_xor(run, 32, 64);
        if (h == 0) aout[pr * 32 + ch] = to_out<TOut>(v);
        run = 0.0f;
      }
    }
    wave_sync();
  }
}

// per-wave LDS footprint (4-byte words) of the adjoint kernel; gs starts 16-byte aligned
// The region after G holds the radial slot vectors (4 words a slot) and, once the angular slots
// are set up, the triplet terms of one pass of rounds (at most 128 slot entries of 3 words for
// rounds of up to 32 pairs, one round of n2 <= max_nbr_ang + 1 slots beyond).
__host__ __device__ inline int bwd_nbv_words(int max_nbr, int max_nbr_ang) {
  const int t = 3 * (max_nbr_ang + 1) > 384 ? 3 * (max_nbr_ang + 1) : 384;
  return round4(4 * max_nbr > t ? 4 * max_nbr : t);
}
__host__ __device__ inline int bwd_gs_offset(int max_nbr, int max_nbr_ang, int aev_len) {
  return round4(aev_len) + bwd_nbv_words(max_nbr, max_nbr_ang);
}
__host__ __device__ inline int bwd_lds_words(int max_nbr, int max_nbr_ang, int aev_len) {
  return bwd_gs_offset(max_nbr, max_nbr_ang, aev_len) + 12 * max_nbr_ang;
}

// Adjoint, one wave per centre.  LDS slice of the wave (4-byte words):
//   G   [aev_len]      dE/dAEV row of the centre
//   nbv [max_nbr]      float4 (vx, vy, vz, species bits) per radial slot
//   gs  [max_nbr_ang]  float4 dE/dvec of the angular slot, angular terms only (xyz), neighbour atom (w)
//                      (per-pass round sums)
//   av  [max_nbr_ang]  float4 (ux, uy, uz, d) per angular slot
//   af  [max_nbr_ang]  float4 (fc_ang, fc_ang', species, 1/d; negative for a degenerate slot)
// Radial terms: the lane of slot e (neighbour j) evaluates both dE_i/dvec_ij (from the
// centre's row G_i: self term and virial of i) and dE_j/dvec_ji (from j's row G_j, read from
// global memory: the neighbour term that j's own radial block gives i), so the radial
// neighbour terms are gathered per atom and written with the self term -- no atomics.  The
// angular terms of each angular slot are scattered to its atom with global int64 atomics on
// fixed point (exact sums in any order).
// Outputs: grad_self[i] = -sum_e dE_i/dvec_ie - sum_j dE_j/dvec_ji|radial;
// grad_nbr[j] += dE_i/dvec_ij|angular * 2^32 (int64); virial_atom[i] = sum_e vec_ie (x) dE_i/dvec_ie.
template <int NA, int NZ, typename GT>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(5)))
k_aev_bwd(const float* __restrict__ coords, const int32_t* __restrict__ species_idx,
          const int32_t* __restrict__ batch_index, int n_atoms, int pbc_mode,
          const float* __restrict__ cells, const float* __restrict__ rec,
          const int32_t* __restrict__ nbr, const int32_t* __restrict__ n_nbr, int max_nbr,
          const int32_t* __restrict__ nbr_shift,
          const int32_t* __restrict__ nbr_ang, const int32_t* __restrict__ n_nbr_ang, int max_nbr_ang,
          fnl_aev_params p, AevDims dm, const GT* __restrict__ g, int g_stride,
          float* __restrict__ grad_self, long long* __restrict__ grad_nbr, float* __restrict__ virial_atom) {
  static_assert(NA * NZ == 32, "angular channel layout");
  extern __shared__ float4 lds4[];
  const int w = threadIdx.x >> 6;
  const int lane = fnl_lane();
  const int i = blockIdx.x * kWaves + w;
  if (i >= n_atoms) return;
  const int words = bwd_lds_words(max_nbr, max_nbr_ang, dm.aev_len);
  float* G = (float*)lds4 + (size_t)w * words;
  float4* nbv = (float4*)(G + ((dm.aev_len + 3) & ~3));
  float4* gs = (float4*)(G + bwd_gs_offset(max_nbr, max_nbr_ang, dm.aev_len));
  float4* av = gs + max_nbr_ang;
  float4* af = av + max_nbr_ang;

  const int si = species_idx[i];
  if (si < 0) {
    if (lane < 3) grad_self[3 * i + lane] = 0.0f;
    if (lane < 9) virial_atom[9 * i + lane] = 0.0f;
    return;
  }
  const bool pbc = pbc_mode == FNL_PBC_MINIMUM_IMAGE;
  FnlCell cl;
  if (pbc || nbr_shift) fnl_load_cell(cells, rec, batch_index[i], cl);
  const int nr = min(n_nbr[i], max_nbr);
  const int na = min(n_nbr_ang[i], max_nbr_ang);
  // the first 64 entries of the angular row, fetched with the dE/dAEV and radial rows
  const int32_t* arow = nbr_ang + (size_t)i * max_nbr_ang;
  const int ae0 = lane < na ? arow[lane] : 0;
  load_g_row(G, g + (size_t)i * g_stride, dm, lane);
  const int32_t* row = nbr + (size_t)i * max_nbr;
  // per-lane partial sums: dE/dx_i (self and gathered radial neighbour terms) and the virial
  float sx = 0.f, sy = 0.f, sz = 0.f;
  float v[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  {
    const float xi = coords[3 * i], yi = coords[3 * i + 1], zi = coords[3 * i + 2];
    for (int e = lane; e < nr; e += 64) {
      const int j = row[e];
      float vx = coords[3 * j] - xi, vy = coords[3 * j + 1] - yi, vz = coords[3 * j + 2] - zi;
      if (nbr_shift)  // general PBC: the stored image (GraphProcessor, preprocessing.py:727-735)
        fnl_apply_shift(vx, vy, vz, cl, nbr_shift[(size_t)i * max_nbr + e]);
      else if (pbc)
        fnl_min_image(vx, vy, vz, cl);
      const float d2 = vx * vx + vy * vy + vz * vz;
      const float d = hw_sqrt(fmaxf(d2, 1.0e-5f));
      const float fc = cos_switch(d, p.cutoff);
      const int sj = species_idx[j];
      nbv[e] = make_float4(vx, vy, vz, __int_as_float(sj));
      // radial: dE/dd = sum_k G[s, k] * 0.25 exp(-eta x^2) (fc' - 2 eta x fc), x = d - R_k,
      // for the centre's block of species sj (G_i) and j's block of species si (G_j)
      const float dfc = cos_switch_grad(d, p.cutoff);
      const float* gr = G + sj * dm.nR;
      const GT* gj = g + (size_t)j * g_stride + si * dm.nR;
      // with y = k_r (d - R_k): exp(-eta x^2) = exp2(-y^2) and fc' - 2 eta x fc = fc' - br fc y
      const float dk = d * dm.kr, bfc = dm.br * fc;
      float vi = 0.0f, vj = 0.0f;
      if (sizeof(GT) == 2 && dm.nR == 16) {
        // BF16 rows, the usual block: unrolled, rows read as 16-byte vectors, 8 shifts a step
        // (FP32 rows keep the loop below: unrolling them spills at this register budget)
#pragma unroll
        for (int q = 0; q < 16; q += 8) {
          float gi[8], gn[8];
          *(float4*)gi = *(const float4*)(gr + q);
          *(float4*)(gi + 4) = *(const float4*)(gr + q + 4);
          const uint4 u = *(const uint4*)((const unsigned short*)gj + q);
          const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            gn[2 * r] = __uint_as_float(w[r] << 16);
            gn[2 * r + 1] = __uint_as_float(w[r] & 0xffff0000u);
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const float y = dk - dm.rs[q + k];
            const float t = hw_exp2(-(y * y)) * fmaf(-bfc, y, dfc);
            vi += gi[k] * t;
            vj += gn[k] * t;
          }
        }
      } else {
        for (int k = 0; k < dm.nR; ++k) {
          const float y = dk - dm.rs[k];
          const float t = hw_exp2(-(y * y)) * fmaf(-bfc, y, dfc);
          vi += gr[k] * t;
          vj += g2f(gj[k]) * t;
        }
      }
      const float id = d2 > 1.0e-5f ? 0.25f * hw_rcp(d) : 0.0f;
      const float sc = vi * id, sj_ = vj * id;
      const float gx = sc * vx, gy = sc * vy, gz = sc * vz;
      // dE/dx_i: -dE_i/dvec_ij + dE_j/dvec_ji, with vec_ji = -vec_ij
      sx -= gx + sj_ * vx;
      sy -= gy + sj_ * vy;
      sz -= gz + sj_ * vz;
      v[0] += vx * gx; v[1] += vx * gy; v[2] += vx * gz;
      v[3] += vy * gx; v[4] += vy * gy; v[5] += vy * gz;
      v[6] += vz * gx; v[7] += vz * gy; v[8] += vz * gz;
    }
  }
  wave_sync();
  {
    for (int a = lane; a < na; a += 64) {
      const int e = a < 64 ? ae0 : arow[a];
      const float4 vv = nbv[e];
      const float d2 = vv.x * vv.x + vv.y * vv.y + vv.z * vv.z;
      const float d = hw_sqrt(fmaxf(d2, 1.0e-5f));
      const float inv = hw_rcp(fmaxf(d, 1.0e-5f));
      av[a] = make_float4(vv.x * inv, vv.y * inv, vv.z * inv, d);
      const bool ok = d2 > 1.0e-5f;
      // af.w: 1/d, negative for a degenerate slot (coincident atoms: no radial term)
      af[a] = make_float4(cos_switch(d, p.cutoff_ang), ok ? cos_switch_grad(d, p.cutoff_ang) : 0.0f, vv.w,
                          ok ? inv : -inv);
      gs[a] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(row[e]));  // .w: the neighbour atom
    }
  }
  wave_sync();

  const float pref = 2.0f * exp2f(-p.zeta);
  const float zm1 = p.zeta - 1.0f;
  const float dpref = (0.95f * p.zeta) * pref;  // d/dc of (1 + cos(theta - theta_z))^zeta: 0.95 zeta b^(zeta-1) (..)
  const float* Gang = G + dm.rad_len;
  // Triplets in the rounds of the circle method: the slots, padded to an even count n2 (slot na
  // is a dummy when na is odd), are paired n2/2 at a time so that each slot is in exactly one
  // pair of a round, and the n2 - 1 rounds hold every pair once.  Round r pairs slot n2 - 1 with
  // r and slots (r + k) mod (n2 - 1), (r - k) mod (n2 - 1) for k = 1 .. n2/2 - 1.  A pass holds
  // rpc = 64 / (n2/2) whole rounds, so all lanes but 64 mod (n2/2) evaluate a triplet.  Each lane
  // stores the dE/dvec terms of its two slots at tmp[round][slot] (plain stores: a slot occurs
  // once in a round), then lanes (slot, component) add the pass's rounds to gs in round order --
  // no LDS atomics and the same order every run.  With more than 64 pairs in a round (na > 128)
  // a round is split over passes.  tmp reuses the radial slot vectors' space (dead here).
  float* tmp = (float*)nbv;
  const int n2 = (na + 1) & ~1;
  const int P = n2 >> 1;
  const int n_rounds = n2 - 1;
  const int rpc = P > 0 ? max(1, 64 / P) : 1;
  const int lane_round = P > 0 ? lane / P : 0, lane_pair = lane - lane_round * P;
  for (int r0 = 0; r0 < n_rounds; r0 += rpc) {
    for (int k0 = 0; k0 < (rpc > 1 ? 1 : P); k0 += 64) {
      const int rl = rpc > 1 ? lane_round : 0;
      const int kk = rpc > 1 ? lane_pair : k0 + lane;
      const int r = r0 + rl;
      const bool live = rl < rpc && r < n_rounds && kk < P;
      int q1 = 0, q2 = 0;
      if (live) {
        if (kk == 0) {
          q1 = n2 - 1;
          q2 = r;
        } else {
          q1 = r + kk;
          if (q1 >= n_rounds) q1 -= n_rounds;
          q2 = r - kk;
          if (q2 < 0) q2 += n_rounds;
        }
      }
      const int a = min(q1, q2), b = max(q1, q2);
      const bool valid = live && b < na;
      float gax = 0.f, gay = 0.f, gaz = 0.f, gbx = 0.f, gby = 0.f, gbz = 0.f;
      if (valid) {
        const float4 ua = av[a], ub = av[b];
        const float4 fa = af[a], fb = af[b];
        const float c = ua.x * ub.x + ua.y * ub.y + ua.z * ub.z;
        const float x = 0.95f * c;
        const float st = hw_sqrt(fmaxf(1.0f - x * x, 1.0e-12f));
        const float xi = x * hw_rcp(st);
        // b = 1 + cos(theta - theta_z), theta = arccos(0.95 c), clamped at 1e-30 (where every term
        // vanishes): F = b^zeta = b^(zeta-1) b, and d/dc b^zeta = 0.95 zeta b^(zeta-1) (cos_z - x sin_z
        // / st), whose factor 0.95 zeta is applied once per triplet (dpref); one log2 and one exp2 per
        // section
        float F[NZ], dF[NZ];
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
          const float bz = fmaxf(1.0f + x * p.ang_cos[z] + st * p.ang_sin[z], 1.0e-30f);
          const float pm1 = hw_exp2(zm1 * hw_log2(bz));
          F[z] = pm1 * bz;
          dF[z] = pm1 * (p.ang_cos[z] - xi * p.ang_sin[z]);
        }
        const float mk = (ua.w + ub.w) * dm.ka_half;  // k_a m
        const float* gp = Gang + pair_index(__float_as_int(fa.z), __float_as_int(fb.z), dm.S) * (NA * NZ);
        // with y_d = k_a (m - R_d): R_d = exp2(-y_d^2), S0 = sum_d R_d g.F, S1 = sum_d y_d R_d g.F
        // (dR_d/dm = -2 (eta_a / k_a) y_d R_d), S2 = sum_d R_d g.dF
        float S0 = 0.0f, S1 = 0.0f, S2 = 0.0f;
#pragma unroll
        for (int d = 0; d < NA; ++d) {
          const float y = mk - dm.as[d];
          const float R = hw_exp2(-(y * y));
          float gz[NZ];
#pragma unroll
          for (int q = 0; q < NZ; q += 4) *(float4*)(gz + q) = *(const float4*)(gp + d * NZ + q);
          float gf = 0.0f, gdf = 0.0f;
#pragma unroll
          for (int z = 0; z < NZ; ++z) {
            gf += gz[z] * F[z];
            gdf += gz[z] * dF[z];
          }
          const float gR = gf * R;
          S0 += gR;
          S1 += gR * y;
          S2 += gdf * R;
        }
        // dE/dd_a = pref fb (fa' S0 + fa dS0/dm / 2), with dS0/dm = -2 (eta_a / k_a) S1; dE/dc = 0.95 zeta
        // pref fa fb S2
        const float S1e = dm.ang_b * S1;
        const float dEda = pref * fb.x * (fa.y * S0 - fa.x * S1e);
        const float dEdb = pref * fa.x * (fb.y * S0 - fb.x * S1e);
        const float dEdc = dpref * fa.x * fb.x * S2;
        // 1/d of the slots (af.w), negative for a degenerate slot, whose radial term is dropped
        const float ra = fa.w < 0.0f ? 0.0f : dEda, rb = fb.w < 0.0f ? 0.0f : dEdb;
        const float ia = fabsf(fa.w), ib = fabsf(fb.w);
        gax = ra * ua.x + dEdc * (ub.x - c * ua.x) * ia;
        gay = ra * ua.y + dEdc * (ub.y - c * ua.y) * ia;
        gaz = ra * ua.z + dEdc * (ub.z - c * ua.z) * ia;
        gbx = rb * ub.x + dEdc * (ua.x - c * ub.x) * ib;
        gby = rb * ub.y + dEdc * (ua.y - c * ub.y) * ib;
        gbz = rb * ub.z + dEdc * (ua.z - c * ub.z) * ib;
      }
      if (live) {  // every slot of the round, the dummy included (zeros)
        float* ta = tmp + (rl * n2 + a) * 3;
        float* tb = tmp + (rl * n2 + b) * 3;
        ta[0] = gax; ta[1] = gay; ta[2] = gaz;
        tb[0] = gbx; tb[1] = gby; tb[2] = gbz;
      }
    }
    wave_sync();
    const int n_pass = min(rpc, n_rounds - r0);
    for (int q = lane; q < 3 * na; q += 64) {
      const int sl = q / 3, c = q - 3 * sl;
      const float* tq = tmp + q;  // entry (round 0, slot sl, component c); rounds 3 n2 apart
      float acc = 0.0f;
#pragma unroll 4
      for (int rl = 0; rl < n_pass; ++rl, tq += 3 * n2) acc += *tq;
      ((float*)(gs + sl))[c] += acc;
    }
    wave_sync();
  }

  // angular terms: scatter to the neighbour atoms, add to the centre's sums and virial
  for (int a = lane; a < na; a += 64) {
    const float4 gg = gs[a];
    const float4 ua = av[a];
    const int j = __float_as_int(gg.w);
    unsigned long long* gn = reinterpret_cast<unsigned long long*>(grad_nbr + 3 * j);
    __hip_atomic_fetch_add(gn, fnl_to_fixed(gg.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(gn + 1, fnl_to_fixed(gg.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(gn + 2, fnl_to_fixed(gg.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sx -= gg.x;
    sy -= gg.y;
    sz -= gg.z;
    const float vx = ua.x * ua.w, vy = ua.y * ua.w, vz = ua.z * ua.w;
    v[0] += vx * gg.x; v[1] += vx * gg.y; v[2] += vx * gg.z;
    v[3] += vy * gg.x; v[4] += vy * gg.y; v[5] += vy * gg.z;
    v[6] += vz * gg.x; v[7] += vz * gg.y; v[8] += vz * gg.z;
  }
  // the 12 sums (dE/dx_i, virial) as a transposing reduction: halves of the values cross between
  // lane halves (xor 32, then 16), so that the 16 lanes of row g end up with partials of values
  // 3g..3g+2 only, summed over the row with DPP; lane r < 3 of row g writes value 3g + r
  {
    const float x[12] = {sx, sy, sz, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
    const bool hi32 = lane >= 32, hi16 = (lane & 16) != 0;
    float y[6], z[3];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float send = hi32 ? x[k] : x[6 + k];
      y[k] = (hi32 ? x[6 + k] : x[k]) + __shfl_xor(send, 32, 64);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float send = hi16 ? y[k] : y[3 + k];
      z[k] = (hi16 ? y[3 + k] : y[k]) + __shfl_xor(send, 16, 64);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      z[k] += dpp<0x128>(z[k]);
      z[k] += dpp<0x124>(z[k]);
      z[k] += dpp<0x122>(z[k]);
      z[k] += dpp<0x121>(z[k]);
    }
    const int r = lane & 15, idx = 3 * (lane >> 4) + r;
    if (r < 3) {
      const float val = r == 0 ? z[0] : (r == 1 ? z[1] : z[2]);
      if (idx < 3)
        grad_self[3 * i + idx] = val;
      else
        virial_atom[9 * i + idx - 3] = val;
    }
  }
}

bool check_params(const fnl_aev_params* p, AevDims* dm) {
  if (!p) return false;
  if (p->n_species < 1 || p->n_species > FNL_MAX_SPECIES) return false;
  if (p->n_radial < 1 || p->n_radial > 16) return false;
  if (p->n_ang_dist < 1 || p->n_ang_sect < 1 || p->n_ang_dist * p->n_ang_sect > 32) return false;
  if (!(p->cutoff > 0.0f) || !(p->cutoff_ang > 0.0f) || p->cutoff_ang > p->cutoff) return false;
  if ((p->n_species * p->n_radial) % 4 != 0) return false;  // angular block 16-byte aligned
  dm->S = p->n_species;
  dm->P = p->n_species * (p->n_species + 1) / 2;
  dm->nR = p->n_radial;
  dm->rad_len = dm->S * dm->nR;
  dm->aev_len = dm->rad_len + dm->P * p->n_ang_dist * p->n_ang_sect;
  dm->live = ~0ull;
  dm->n_groups = 0;
  if (!(p->radial_eta >= 0.0f) || !(p->angular_eta >= 0.0f) || !(p->zeta > 0.0f)) return false;
  const float l2e = 1.44269504088896340736f;
  dm->kr = sqrtf(p->radial_eta * l2e);
  const float ka = sqrtf(p->angular_eta * l2e);
  dm->ka_half = 0.5f * ka;
  dm->br = dm->kr > 0.0f ? 2.0f * p->radial_eta / dm->kr : 0.0f;
  dm->ang_b = ka > 0.0f ? p->angular_eta / ka : 0.0f;
  for (int k = 0; k < 16; ++k) {
    dm->rs[k] = k < p->n_radial ? p->radial_shifts[k] * dm->kr : 0.0f;
    dm->as[k] = k < p->n_ang_dist ? p->ang_shifts[k] * ka : 0.0f;
  }
  return true;
}

template <int NA, int NZ>
int launch_fwd(const float* coords, const int32_t* species_idx, const int32_t* batch_index, int n_atoms,
               int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
               const int32_t* n_nbr, int max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
               int max_nbr_ang, const fnl_aev_params& p, const AevDims& dm, void* aev, int stride,
               int out_dtype, uint32_t smask, hipStream_t s) {
  const size_t lds = (size_t)kWaves * 4 * fwd_layout(max_nbr, max_nbr_ang, p.n_ang_dist + p.n_ang_sect).words;
  if (lds > 160 * 1024) return FNL_ERR_UNSUPPORTED;
  dim3 grid(fnl_blocks(n_atoms, kWaves)), block(64 * kWaves);
  const void* kf = out_dtype == FNL_F32 ? reinterpret_cast<const void*>(&k_aev_fwd<NA, NZ, float>)
                                        : reinterpret_cast<const void*>(&k_aev_fwd<NA, NZ, hip_bfloat16>);
  if (const int rc = fnl_allow_full_lds(kf)) return rc;
  if (out_dtype == FNL_F32) {
    hipLaunchKernelGGL((k_aev_fwd<NA, NZ, float>), grid, block, lds, s, coords, species_idx, batch_index,
                       n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang,
                       p, dm, (float*)aev, stride, smask, nullptr);
  } else {
    hipLaunchKernelGGL((k_aev_fwd<NA, NZ, hip_bfloat16>), grid, block, lds, s, coords, species_idx,
                       batch_index, n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang,
                       max_nbr_ang, p, dm, (hip_bfloat16*)aev, stride, smask, nullptr);
  }
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

template <int NA, int NZ, typename GT>
int launch_bwd(const float* coords, const int32_t* species_idx, const int32_t* batch_index, int n_atoms,
               int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
               const int32_t* n_nbr, int max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang, const int32_t* n_nbr_ang,
               int max_nbr_ang, const fnl_aev_params& p, const AevDims& dm, const GT* g, int g_stride,
               float* grad_self, long long* grad_nbr, float* virial_atom, hipStream_t s) {
  const size_t lds = (size_t)kWaves * 4 * bwd_lds_words(max_nbr, max_nbr_ang, dm.aev_len);
  if (lds > 160 * 1024) return FNL_ERR_UNSUPPORTED;
  if (const int rc = fnl_allow_full_lds(reinterpret_cast<const void*>(&k_aev_bwd<NA, NZ, GT>))) return rc;
  hipLaunchKernelGGL((k_aev_bwd<NA, NZ, GT>), dim3(fnl_blocks(n_atoms, kWaves)), dim3(64 * kWaves), lds, s,
                     coords, species_idx, batch_index, n_atoms, pbc_mode, cells, rec, nbr, n_nbr, max_nbr,
                     nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang, p, dm, g, g_stride, grad_self, grad_nbr,
                     virial_atom);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

// nbr_shift: required for FNL_PBC_GENERAL (the stored image of every slot), NULL otherwise
bool common_args_ok(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                    int n_atoms, int pbc_mode, const float* cells, const float* rec, const int32_t* nbr,
                    const int32_t* n_nbr, int max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang,
                    const int32_t* n_nbr_ang, int max_nbr_ang) {
  if (n_atoms < 0 || max_nbr < 1 || max_nbr_ang < 1) return false;
  if (n_atoms == 0) return true;
  if (!coords || !species_idx || !batch_index || !nbr || !n_nbr || !nbr_ang || !n_nbr_ang) return false;
  if (pbc_mode != FNL_PBC_NONE && pbc_mode != FNL_PBC_MINIMUM_IMAGE && pbc_mode != FNL_PBC_GENERAL) return false;
  if (pbc_mode != FNL_PBC_NONE && (!cells || !rec)) return false;
  if ((pbc_mode == FNL_PBC_GENERAL) != (nbr_shift != nullptr)) return false;
  return true;
}

}  // namespace

extern "C" int fnl_aev_forward(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                               int32_t n_atoms, int32_t pbc_mode, const float* cells,
                               const float* reciprocal_cells, const int32_t* nbr, const int32_t* n_nbr,
                               int32_t max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang,
                               const int32_t* n_nbr_ang,
                               int32_t max_nbr_ang, const fnl_aev_params* params, void* aev,
                               int32_t aev_stride, int32_t out_dtype, uint32_t species_mask,
                               fnl_stream_t stream) {
  AevDims dm;
  if (!check_params(params, &dm)) return FNL_ERR_INVALID;
  if (!common_args_ok(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, nbr,
                      n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang))
    return FNL_ERR_INVALID;
  if (aev_stride < dm.aev_len || (out_dtype != FNL_F32 && out_dtype != FNL_BF16)) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!aev) return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int na = params->n_ang_dist, nz = params->n_ang_sect;
  if (na == 4 && nz == 8)
    return launch_fwd<4, 8>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, aev,
                            aev_stride, out_dtype, species_mask, s);
  if (na == 8 && nz == 4)
    return launch_fwd<8, 4>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells,
                            nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm, aev,
                            aev_stride, out_dtype, species_mask, s);
  return FNL_ERR_UNSUPPORTED;
}

extern "C" int fnl_aev_backward(const float* coords, const int32_t* species_idx, const int32_t* batch_index,
                                int32_t n_atoms, int32_t pbc_mode, const float* cells,
                                const float* reciprocal_cells, const int32_t* nbr, const int32_t* n_nbr,
                                int32_t max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang,
                                const int32_t* n_nbr_ang,
                                int32_t max_nbr_ang, const fnl_aev_params* params, const void* g,
                                int32_t g_stride, int32_t g_dtype, float* grad_self, int64_t* grad_nbr,
                                float* virial_atom, uint32_t species_mask, fnl_stream_t stream) {
  AevDims dm;
  if (!check_params(params, &dm)) return FNL_ERR_INVALID;
  dm.live = live_blocks(species_mask, dm);
  // the live blocks as 8-channel groups (blocks are whole groups when nR % 8 == 0)
  if (dm.nR % 8 == 0 && g_stride % 8 == 0) {
    int ng = 0;
    for (int blk = 0; blk < dm.S + dm.P && ng >= 0; ++blk) {
      if (!((dm.live >> blk) & 1ull)) continue;
      const int c0 = blk < dm.S ? blk * dm.nR : dm.rad_len + (blk - dm.S) * 32;
      const int c1 = blk < dm.S ? c0 + dm.nR : c0 + 32;
      for (int c = c0; c < c1; c += 8) {
        if (ng >= kMaxGroups) { ng = -1; break; }
        dm.groups[ng++] = (uint16_t)(c / 8);
      }
    }
    dm.n_groups = ng > 0 ? ng : 0;
  }
  if (!common_args_ok(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, nbr,
                      n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang))
    return FNL_ERR_INVALID;
  if (g_stride < ((dm.aev_len + 3) & ~3) || g_stride % 4 != 0) return FNL_ERR_INVALID;
  if (g_dtype != FNL_F32 && g_dtype != FNL_BF16) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!g || !grad_self || !grad_nbr || !virial_atom) return FNL_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const int na = params->n_ang_dist, nz = params->n_ang_sect;
  const float* gf = (const float*)g;
  const unsigned short* gb = (const unsigned short*)g;
  const bool bf = g_dtype == FNL_BF16;
#define FNL_BWD(NA_, NZ_)                                                                                         \
  return bf ? launch_bwd<NA_, NZ_>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, \
                                   nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm,  \
                                   gb, g_stride, grad_self, (long long*)grad_nbr, virial_atom, s)                \
            : launch_bwd<NA_, NZ_>(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, \
                                   nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang, *params, dm,  \
                                   gf, g_stride, grad_self, (long long*)grad_nbr, virial_atom, s)
  if (na == 4 && nz == 8) FNL_BWD(4, 8);
  if (na == 8 && nz == 4) FNL_BWD(8, 4);
#undef FNL_BWD
  return FNL_ERR_UNSUPPORTED;
}

extern "C" int fnl_aev_tangent(const float* coords, const float* disp, const int32_t* species_idx,
                               const int32_t* batch_index, int32_t n_atoms, int32_t pbc_mode, const float* cells,
                               const float* reciprocal_cells, const int32_t* nbr, const int32_t* n_nbr,
                               int32_t max_nbr, const int32_t* nbr_shift, const int32_t* nbr_ang,
                               const int32_t* n_nbr_ang,
                               int32_t max_nbr_ang, const fnl_aev_params* params, float* daev, int32_t daev_stride,
                               fnl_stream_t stream) {
  AevDims dm;
  if (!check_params(params, &dm)) return FNL_ERR_INVALID;
  if (!common_args_ok(coords, species_idx, batch_index, n_atoms, pbc_mode, cells, reciprocal_cells, nbr,
                      n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang, max_nbr_ang))
    return FNL_ERR_INVALID;
  if (daev_stride < dm.aev_len) return FNL_ERR_INVALID;
  if (n_atoms == 0) return FNL_OK;
  if (!disp || !daev) return FNL_ERR_INVALID;
  const int na = params->n_ang_dist, nz = params->n_ang_sect;
  const size_t lds = (size_t)kWaves * 4 * fwd_layout(max_nbr, max_nbr_ang, na + nz, true).words;
  if (lds > 160 * 1024) return FNL_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  const void* kf = nullptr;
  if (na == 4 && nz == 8) kf = reinterpret_cast<const void*>(&k_aev_fwd<4, 8, float, true>);
  else if (na == 8 && nz == 4) kf = reinterpret_cast<const void*>(&k_aev_fwd<8, 4, float, true>);
  else return FNL_ERR_UNSUPPORTED;
  if (const int rc = fnl_allow_full_lds(kf)) return rc;
  dim3 grid(fnl_blocks(n_atoms, kWaves)), block(64 * kWaves);
  // every channel block written (zeros where no neighbour or pair contributes)
  if (na == 4)
    hipLaunchKernelGGL((k_aev_fwd<4, 8, float, true>), grid, block, lds, s, coords, species_idx, batch_index,
                       n_atoms, pbc_mode, cells, reciprocal_cells, nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang,
                       max_nbr_ang, *params, dm, daev, daev_stride, 0xFFFFFFFFu, disp);
  else
    hipLaunchKernelGGL((k_aev_fwd<8, 4, float, true>), grid, block, lds, s, coords, species_idx, batch_index,
                       n_atoms, pbc_mode, cells, reciprocal_cells, nbr, n_nbr, max_nbr, nbr_shift, nbr_ang, n_nbr_ang,
                       max_nbr_ang, *params, dm, daev, daev_stride, 0xFFFFFFFFu, disp);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
