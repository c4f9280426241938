// Species MLPs (SpeciesIndexNet over FullyConnectedNet, models/misc/nets.py:
// 13-102, 301-426) as one fused forward + input-gradient kernel on MFMA.
//
// One workgroup (4 waves) per tile of M atoms that share a network.  The
// tile walks K -> 256 -> 192 -> 160 -> 1 and back:
//   forward   Z_l = H_{l-1} W_l + b_l, H_l = celu(Z_l) -> LDS (compute dtype),
//             D_l = celu'(Z_l) kept in registers in the accumulator layout
//   energy    E = H_3 . w4 + b4 (fp32, from the accumulators)
//   backward  G_3 = w4 * D_3;  G_2 = (G_3 W_3^T) * D_2;  G_1 = (G_2 W_2^T) * D_1;
//             dE/dAEV = G_1 W_1^T -> global (fp32)
// Activations H_l are overwritten by G_l in the same LDS buffers.
//
// MFMA shapes (wave64):
//   FP32: v_mfma_f32_16x16x4_f32    A: lane (row l%16, k l/16)
//   BF16: v_mfma_f32_16x16x32_bf16  A: lane (row l%16, k 8*(l/16) .. +7)
//   result: lane l, reg r -> (row 4*(l/16) + r, col l%16)
// A fragments come from LDS, B fragments straight from global memory (the
// weights stay resident in L2 across tiles) in the layout that makes each
// lane's load contiguous: K-contiguous (transposed) matrices for BF16 (8 K
// values per lane), N-contiguous ones for FP32 (16 lanes read 16 columns).

#include <hip/hip_bfloat16.h>

#include "fnl_common.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int H1 = FNL_MLP_H1, H2 = FNL_MLP_H2, H3 = FNL_MLP_H3;
constexpr int KC = 64;  // layer-1 K chunk staged through LDS

struct NetPtrs {
  const void* fw[3];   // B operand of the forward GEMMs
  const void* bw[3];   // B operand of the backward GEMMs
  const float* w4;
  const float* b[4];
};

struct MlpArgs {
  NetPtrs nets[FNL_MLP_MAX_NETS];
  int k_pad;
  float alpha;
};

template <typename T>
struct Tr;

template <>
struct Tr<float> {
  static constexpr int KS = 4;  // K per MFMA
  static constexpr int M = 32;  // rows per tile
  static constexpr int PAD = 4; // LDS row padding (elements)
  typedef float frag;
  static __device__ __forceinline__ f4 mma(frag a, frag b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // A fragment from an LDS matrix (row-major, ld elements)
  static __device__ __forceinline__ frag a_lds(const float* buf, int ld, int row0, int k0) {
    const int l = fnl_lane();
    return buf[(row0 + (l & 15)) * ld + k0 + (l >> 4)];
  }
  // B fragment: element (k, n) of an N-contiguous matrix p[k*ld + n]
  static __device__ __forceinline__ frag b_glb(const float* p, int ld, int k0, int n0) {
    const int l = fnl_lane();
    return p[(size_t)(k0 + (l >> 4)) * ld + n0 + (l & 15)];
  }
  static __device__ __forceinline__ float to_f(float v) { return v; }
  static __device__ __forceinline__ float from_f(float v) { return v; }
};

__device__ __forceinline__ unsigned short f2bf(float f) {
  // round to nearest even
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16 | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <>
struct Tr<unsigned short> {  // bf16 stored as raw bits
  static constexpr int KS = 32;
  static constexpr int M = 64;
  static constexpr int PAD = 8;
  typedef bf8 frag;
  static __device__ __forceinline__ f4 mma(frag a, frag b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ frag a_lds(const unsigned short* buf, int ld, int row0, int k0) {
    const int l = fnl_lane();
    const us8 v = *(const us8*)(buf + (row0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
    return __builtin_bit_cast(bf8, v);
  }
  // B fragment: element (k, n) of a K-contiguous matrix p[n*ld + k]
  static __device__ __forceinline__ frag b_glb(const unsigned short* p, int ld, int k0, int n0) {
    const int l = fnl_lane();
    const us8 v = *(const us8*)(p + (size_t)(n0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
    return __builtin_bit_cast(bf8, v);
  }
  static __device__ __forceinline__ float to_f(unsigned short v) { return bf2f(v); }
  static __device__ __forceinline__ unsigned short from_f(float v) { return f2bf(v); }
};

__device__ __forceinline__ float celu(float z, float a) { return z > 0.0f ? z : a * expm1f(z / a); }
__device__ __forceinline__ float celu_d(float z, float a) { return z > 0.0f ? 1.0f : __expf(z / a); }

// C[M x N] += A[M x K] (LDS) * B[K x N] (global) for the col-blocks of this wave
// (cb = wave + 4 j).  K is a multiple of KS.
template <typename T, int RB, int CBW>
__device__ __forceinline__ void gemm_lds(f4 (&acc)[RB][CBW], const T* A, int lda, int K, const T* B,
                                         int ldb, int ncb, int wave) {
  typedef Tr<T> tr;
  for (int k0 = 0; k0 < K; k0 += tr::KS) {
    typename tr::frag a[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) a[rb] = tr::a_lds(A, lda, rb * 16, k0);
#pragma unroll
    for (int j = 0; j < CBW; ++j) {
      const int cb = wave + 4 * j;
      if (cb < ncb) {
        const typename tr::frag b = tr::b_glb(B, ldb, k0, cb * 16);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb][j] = tr::mma(a[rb], b, acc[rb][j]);
      }
    }
  }
}

template <int RB, int CBW>
__device__ __forceinline__ void zero_acc(f4 (&acc)[RB][CBW]) {
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int j = 0; j < CBW; ++j) acc[rb][j] = f4{0.f, 0.f, 0.f, 0.f};
}

// forward epilogue: z = acc + bias; H -> LDS (dtype T); D = celu'(z) -> regs
template <typename T, int RB, int CBW>
__device__ __forceinline__ void fwd_epilogue(f4 (&acc)[RB][CBW], f4 (&D)[RB][CBW], const float* bias, T* H,
                                             int ldh, int ncb, int wave, float alpha) {
  const int l = fnl_lane();
#pragma unroll
  for (int j = 0; j < CBW; ++j) {
    const int cb = wave + 4 * j;
    if (cb >= ncb) continue;
    const int col = cb * 16 + (l & 15);
    const float bv = bias[col];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + 4 * (l >> 4) + r;
        const float z = acc[rb][j][r] + bv;
        D[rb][j][r] = celu_d(z, alpha);
        acc[rb][j][r] = celu(z, alpha);
        H[row * ldh + col] = Tr<T>::from_f(acc[rb][j][r]);
      }
    }
  }
}

// backward epilogue: G = acc * D -> LDS
template <typename T, int RB, int CBW>
__device__ __forceinline__ void bwd_epilogue(const f4 (&acc)[RB][CBW], const f4 (&D)[RB][CBW], T* G, int ldg,
                                             int ncb, int wave) {
  const int l = fnl_lane();
#pragma unroll
  for (int j = 0; j < CBW; ++j) {
    const int cb = wave + 4 * j;
    if (cb >= ncb) continue;
    const int col = cb * 16 + (l & 15);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + 4 * (l >> 4) + r;
        G[row * ldg + col] = Tr<T>::from_f(acc[rb][j][r] * D[rb][j][r]);
      }
  }
}

template <typename T>
__global__ void __launch_bounds__(256, 1)
k_species_mlp(const T* __restrict__ aev, int aev_stride, MlpArgs args, const int32_t* __restrict__ perm,
              const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tile_rows,
              const int32_t* __restrict__ tile_net, float* __restrict__ energy, float* __restrict__ grad,
              int grad_stride) {
  typedef Tr<T> tr;
  constexpr int M = tr::M;
  constexpr int RB = M / 16;
  constexpr int LDX = KC + tr::PAD;
  constexpr int LD1 = H1 + tr::PAD, LD2 = H2 + tr::PAD, LD3 = H3 + tr::PAD;
  __shared__ __attribute__((aligned(16))) T Xs[M * LDX];
  __shared__ __attribute__((aligned(16))) T B1[M * LD1];
  __shared__ __attribute__((aligned(16))) T B2[M * LD2];
  __shared__ __attribute__((aligned(16))) T B3[M * LD3];
  __shared__ float Epart[4][M];
  __shared__ int rows_atom[M];

  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int l = fnl_lane();
  const int tile = blockIdx.x;
  const int net = tile_net[tile];
  const int nrows = tile_rows[tile];
  const int r0 = tile_start[tile];
  const NetPtrs& P = args.nets[net];
  const int K = args.k_pad;
  const float alpha = args.alpha;
  if (t < M) rows_atom[t] = t < nrows ? perm[r0 + t] : -1;
  __syncthreads();

  // ---------------- layer 1: X (global, staged) x W1 ----------------
  f4 acc1[RB][4];
  f4 D1[RB][4];
  zero_acc(acc1);
  constexpr int VEC = 16 / sizeof(T);          // elements per 16-byte load
  constexpr int TPR = KC / VEC;                // threads per row
  constexpr int ROWS_PER_PASS = 256 / TPR;
  for (int kc = 0; kc < K; kc += KC) {
#pragma unroll
    for (int rp = 0; rp < M; rp += ROWS_PER_PASS) {
      const int row = rp + t / TPR;
      const int c = (t % TPR) * VEC;
      if (row < M) {
        const int atom = rows_atom[row];
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (atom >= 0) v = *(const uint4*)(aev + (size_t)atom * aev_stride + kc + c);
        *(uint4*)(Xs + row * LDX + c) = v;
      }
    }
    __syncthreads();
    gemm_lds<T, RB, 4>(acc1, Xs, LDX, KC, (const T*)P.fw[0] + (sizeof(T) == 2 ? kc : (size_t)kc * H1),
                       sizeof(T) == 2 ? K : H1, H1 / 16, wave);
    __syncthreads();
  }
  fwd_epilogue<T, RB, 4>(acc1, D1, P.b[0], B1, LD1, H1 / 16, wave, alpha);
  __syncthreads();

  // ---------------- layer 2 ----------------
  f4 acc2[RB][3];
  f4 D2[RB][3];
  zero_acc(acc2);
  gemm_lds<T, RB, 3>(acc2, B1, LD1, H1, (const T*)P.fw[1], sizeof(T) == 2 ? H1 : H2, H2 / 16, wave);
  fwd_epilogue<T, RB, 3>(acc2, D2, P.b[1], B2, LD2, H2 / 16, wave, alpha);
  __syncthreads();

  // ---------------- layer 3 + energy ----------------
  f4 acc3[RB][3];
  f4 D3[RB][3];
  zero_acc(acc3);
  gemm_lds<T, RB, 3>(acc3, B2, LD2, H2, (const T*)P.fw[2], sizeof(T) == 2 ? H2 : H3, H3 / 16, wave);
  fwd_epilogue<T, RB, 3>(acc3, D3, P.b[2], B3, LD3, H3 / 16, wave, alpha);
  {
    // E_row = sum_c H3[row][c] w4[c]: lanes hold 4 rows x 1 column per block
    float part[RB][4];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[rb][r] = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int cb = wave + 4 * j;
      if (cb >= H3 / 16) continue;
      const float wv = P.w4[cb * 16 + (l & 15)];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) part[rb][r] += acc3[rb][j][r] * wv;
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = part[rb][r];
        v += __shfl_xor(v, 8, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 1, 64);
        if ((l & 15) == 0) Epart[wave][rb * 16 + 4 * (l >> 4) + r] = v;
      }
  }
  __syncthreads();
  if (t < M && rows_atom[t] >= 0) {
    energy[rows_atom[t]] = ((Epart[0][t] + Epart[1][t]) + (Epart[2][t] + Epart[3][t])) + P.b[3][0];
  }
  if (grad == nullptr) return;

  // ---------------- backward ----------------
  // G3 = w4 * D3 -> B3
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int cb = wave + 4 * j;
    if (cb >= H3 / 16) continue;
    const int col = cb * 16 + (l & 15);
    const float wv = P.w4[col];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + 4 * (l >> 4) + r;
        B3[row * LD3 + col] = tr::from_f(wv * D3[rb][j][r]);
      }
  }
  __syncthreads();
  // G2 = (G3 W3^T) * D2 -> B2
  zero_acc(acc2);
  gemm_lds<T, RB, 3>(acc2, B3, LD3, H3, (const T*)P.bw[2], sizeof(T) == 2 ? H3 : H2, H2 / 16, wave);
  bwd_epilogue<T, RB, 3>(acc2, D2, B2, LD2, H2 / 16, wave);
  __syncthreads();
  // G1 = (G2 W2^T) * D1 -> B1
  zero_acc(acc1);
  gemm_lds<T, RB, 4>(acc1, B2, LD2, H2, (const T*)P.bw[1], sizeof(T) == 2 ? H2 : H1, H1 / 16, wave);
  bwd_epilogue<T, RB, 4>(acc1, D1, B1, LD1, H1 / 16, wave);
  __syncthreads();
  // dE/dX = G1 W1^T, 64 output columns per wave per pass
  const int ncb_x = K / 16;
  for (int cbase = 0; cbase < ncb_x; cbase += 16) {
    f4 accx[RB][4];
    zero_acc(accx);
    // col-blocks of this pass: cbase + wave + 4 j
    typedef Tr<T> trr;
    for (int k0 = 0; k0 < H1; k0 += trr::KS) {
      typename trr::frag a[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) a[rb] = trr::a_lds(B1, LD1, rb * 16, k0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cb = cbase + wave + 4 * j;
        if (cb < ncb_x) {
          const typename trr::frag b = trr::b_glb((const T*)P.bw[0], sizeof(T) == 2 ? H1 : K, k0, cb * 16);
#pragma unroll
          for (int rb = 0; rb < RB; ++rb) accx[rb][j] = trr::mma(a[rb], b, accx[rb][j]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cb = cbase + wave + 4 * j;
      if (cb >= ncb_x) continue;
      const int col = cb * 16 + (l & 15);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rb * 16 + 4 * (l >> 4) + r;
          const int atom = rows_atom[row];
          if (atom >= 0) grad[(size_t)atom * grad_stride + col] = accx[rb][j][r];
        }
    }
  }
}

}  // namespace

extern "C" int fnl_mlp_tile_rows(int32_t dtype) {
  if (dtype == FNL_F32) return Tr<float>::M;
  if (dtype == FNL_BF16) return Tr<unsigned short>::M;
  return FNL_ERR_INVALID;
}

extern "C" int fnl_species_mlp(const void* aev, int32_t aev_stride, const fnl_mlp_weights* w,
                               const int32_t* perm, const int32_t* tile_start, const int32_t* tile_rows,
                               const int32_t* tile_net, int32_t n_tiles, float* energy, float* grad,
                               int32_t grad_stride, fnl_stream_t stream) {
  if (!w || n_tiles < 0) return FNL_ERR_INVALID;
  if (n_tiles == 0) return FNL_OK;
  if (!aev || !perm || !tile_start || !tile_rows || !tile_net || !energy) return FNL_ERR_INVALID;
  if (w->n_nets < 1 || w->n_nets > FNL_MLP_MAX_NETS) return FNL_ERR_INVALID;
  if (w->k_pad < KC || w->k_pad % KC != 0 || aev_stride < w->k_pad) return FNL_ERR_INVALID;
  if (grad && grad_stride < w->k_pad) return FNL_ERR_INVALID;
  MlpArgs a;
  a.k_pad = w->k_pad;
  a.alpha = w->celu_alpha;
  const bool bf = w->dtype == FNL_BF16;
  if (!bf && w->dtype != FNL_F32) return FNL_ERR_INVALID;
  for (int n = 0; n < FNL_MLP_MAX_NETS; ++n) {
    NetPtrs& p = a.nets[n];
    if (n >= w->n_nets) {
      p = a.nets[0];
      continue;
    }
    const fnl_mlp_net& s = w->nets[n];
    if (!s.w1t || !s.w2t || !s.w3t || !s.w1 || !s.w2 || !s.w3 || !s.w4 || !s.b1 || !s.b2 || !s.b3 || !s.b4)
      return FNL_ERR_INVALID;
    // forward B operands: K-contiguous for BF16, N-contiguous for FP32 (and the
    // opposite for the backward GEMMs)
    p.fw[0] = bf ? s.w1t : s.w1;
    p.fw[1] = bf ? s.w2t : s.w2;
    p.fw[2] = bf ? s.w3t : s.w3;
    p.bw[0] = bf ? s.w1 : s.w1t;
    p.bw[1] = bf ? s.w2 : s.w2t;
    p.bw[2] = bf ? s.w3 : s.w3t;
    p.w4 = s.w4;
    p.b[0] = s.b1;
    p.b[1] = s.b2;
    p.b[2] = s.b3;
    p.b[3] = s.b4;
  }
  hipStream_t st = (hipStream_t)stream;
  if (bf) {
    hipLaunchKernelGGL(k_species_mlp<unsigned short>, dim3(n_tiles), dim3(256), 0, st,
                       (const unsigned short*)aev, aev_stride, a, perm, tile_start, tile_rows, tile_net, energy,
                       grad, grad_stride);
  } else {
    hipLaunchKernelGGL(k_species_mlp<float>, dim3(n_tiles), dim3(256), 0, st, (const float*)aev, aev_stride, a,
                       perm, tile_start, tile_rows, tile_net, energy, grad, grad_stride);
  }
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
