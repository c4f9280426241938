// Species MLPs (SpeciesIndexNet over FullyConnectedNet, models/misc/nets.py:
// 13-102, 301-426) as one fused forward + input-gradient kernel on MFMA.
//
// One workgroup (4 waves; 8 for 96-atom BF16 tiles) per tile of M atoms that share a network.  Every
// GEMM is computed transposed, out^T = W^T act^T: the weights are the MFMA A
// operand (streamed from global/L2, prefetched PF k-steps ahead in registers),
// the activations the B operand (read from LDS), so that each lane's result
// registers hold 4 consecutive output features of one atom and store as one
// 8/16-byte vector into the row-major activation buffers.
//
//   forward   H_l = celu(W_l^T H_{l-1} + b_l)  -> LDS [M][N_l] (compute dtype)
//   energy    E = H_3 . w4 + b4                 (fp32 accumulators)
//   backward  G_3 = w4 * D_3 ; G_2 = (W_3 G_3) * D_2 ; G_1 = (W_2 G_2) * D_1
//             dE/dAEV = W_1 G_1 -> global (fp32, or bf16 rows)
// with D_l = celu'(Z_l) recovered from the stored activation:
//   celu'(z) = 1 (z > 0), exp(z/alpha) = 1 + h/alpha (z <= 0).
// G_l overwrites H_l in place.  Layer-1 inputs (AEV rows) are staged through
// a double-buffered LDS area that aliases the H_2/H_3 buffers, which are free
// until layer 2.  BF16 weights are stored in MFMA fragment order (one 1 KB
// contiguous block per 16 x 32 A operand), so every weight load is whole cache lines.
//
// MFMA (wave64): FP32 v_mfma_f32_16x16x4_f32, BF16 v_mfma_f32_16x16x32_bf16.
// A operand lane l: row (l%16), k-group (l/16); B operand lane l: column
// (l%16), k-group (l/16); result lane l, register r: row 4*(l/16)+r, col l%16.

#include "fnl_common.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

constexpr int H1 = FNL_MLP_H1, H2 = FNL_MLP_H2, H3 = FNL_MLP_H3;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X

struct NetPtrs {
  const void* fw[3];     // A operand of the forward GEMMs
  const void* bw[3];     // A operand of the backward GEMMs
  const void* fw_lo[3];  // compensated BF16: the rounding remainders W - bf16(W), same layout
  const void* bw_lo[3];
  const float* w4;
  const float* b[4];
};

constexpr int FNL_MLP_MAX_CHUNKS = 128;  // layer-1 K chunks per row: 8192 FP32 / 16384 BF16 columns

struct MlpArgs {
  NetPtrs nets[FNL_MLP_MAX_NETS];
  int k_pad;
  float alpha;
  unsigned long long k_live;  // bit q: input columns [32q, 32q + 32) may be non-zero in the batch
  const int32_t* n_tiles_dev;  // tiles in use, read on the device (the grid covers n_tiles); or null
};

__device__ __forceinline__ unsigned short f2bf(float f) {
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(f));
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <typename T>
struct Tr;

template <>
struct Tr<float> {
  static constexpr int KS = 4;    // K per MFMA
  static constexpr int PAD = 4;   // activation row padding (elements)
  static constexpr int KC = 64;   // layer-1 K chunk
  typedef float frag;
  static __device__ __forceinline__ f4 mma(frag a, frag b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // weights: element (n, k) of an N-contiguous matrix p[k*ld + n]
  static __device__ __forceinline__ frag wload(const float* p, int ld, int n0, int k0) {
    const int l = fnl_lane();
    return p[(size_t)(k0 + (l >> 4)) * ld + n0 + (l & 15)];
  }
  // activations: element (k, m) = buf[m*ld + k]
  static __device__ __forceinline__ frag aload(const float* buf, int ld, int m0, int k0) {
    const int l = fnl_lane();
    return buf[(m0 + (l & 15)) * ld + k0 + (l >> 4)];
  }
  static __device__ __forceinline__ void store4(float* p, f4 v) { *(f4*)p = v; }
  static __device__ __forceinline__ f4 load4(const float* p) { return *(const f4*)p; }
};

template <>
struct Tr<unsigned short> {  // bf16 stored as raw bits
  static constexpr int KS = 32;
  static constexpr int PAD = 8;
  static constexpr int KC = 128;
  typedef bf8 frag;
  static __device__ __forceinline__ f4 mma(frag a, frag b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  // weights in fragment order: the (n0/16, k0/32) block of 16 x 32 elements of a matrix with ld
  // columns is one contiguous 1 KB fragment, lane l's 8 elements at l*8 (element (n, k) at
  // ((n/16)*(ld/32) + k/32)*512 + ((k%32)/8)*128 + (n%16)*8 + k%8): each load is 1 KB in full lines
  static __device__ __forceinline__ frag wload(const unsigned short* p, int ld, int n0, int k0) {
    const int l = fnl_lane();
    return __builtin_bit_cast(bf8, *(const us8*)(p + ((size_t)(n0 >> 4) * (ld >> 5) + (k0 >> 5)) * 512 + l * 8));
  }
  static __device__ __forceinline__ frag aload(const unsigned short* buf, int ld, int m0, int k0) {
    const int l = fnl_lane();
    return __builtin_bit_cast(bf8, *(const us8*)(buf + (m0 + (l & 15)) * ld + k0 + 8 * (l >> 4)));
  }
  static __device__ __forceinline__ void store4(unsigned short* p, f4 v) {
    us4 o;
    o.x = f2bf(v.x);
    o.y = f2bf(v.y);
    o.z = f2bf(v.z);
    o.w = f2bf(v.w);
    *(us4*)p = o;
  }
  static __device__ __forceinline__ f4 load4(const unsigned short* p) {
    const us4 v = *(const us4*)p;
    return f4{bf2f(v.x), bf2f(v.y), bf2f(v.z), bf2f(v.w)};
  }
};

template <typename T, int M_, int WV_ = 4>
struct Lds {
  static constexpr int M = M_;
  static constexpr int LD1 = H1 + Tr<T>::PAD, LD2 = H2 + Tr<T>::PAD, LD3 = H3 + Tr<T>::PAD;
  static constexpr int LDX = Tr<T>::KC + Tr<T>::PAD;
  static constexpr int OFF1 = 0;
  static constexpr int OFF2 = OFF1 + M * LD1 * (int)sizeof(T);
  static constexpr int OFF3 = OFF2 + M * LD2 * (int)sizeof(T);
  static constexpr int OFFE = OFF3 + M * LD3 * (int)sizeof(T);
  static constexpr int BYTES = OFFE;
  static_assert((H3 / 16) * M * 4 <= M * LD3 * (int)sizeof(T), "energy partials fit the H3 buffer");
  static constexpr int XBUF = M * LDX * (int)sizeof(T);  // one staging buffer, at OFF2
  static_assert(2 * XBUF <= OFFE - OFF2, "layer-1 staging must fit in the H2/H3 buffers");
  static_assert(BYTES <= 160 * 1024, "fits the LDS of one CU");
  // workgroups per CU allowed by the LDS footprint, at most 3 for BF16 (the
  // prefetch ring then keeps <= 168 VGPRs per lane without spills)
  static constexpr int BLOCKS_LDS = (160 * 1024) / BYTES;
  static constexpr int BLOCKS_PER_CU =
      sizeof(T) == 2 ? (BLOCKS_LDS > 3 ? 3 : BLOCKS_LDS) : (BLOCKS_LDS > 4 ? 4 : BLOCKS_LDS);
};

template <int NB, int MB>
__device__ __forceinline__ void zero_acc(f4 (&acc)[NB][MB]) {
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
}

// The weight fragments of a GEMM's first PF k-steps (A operands of this wave's n-blocks), loaded
// ahead: they do not depend on the activations, so the caller issues them before the previous
// layer's epilogue and barrier, whose time then covers their latency.
template <typename T, int NB, int PF, bool COMP>
struct WRing {
  typename Tr<T>::frag q[PF][NB];
  typename Tr<T>::frag lo[COMP ? PF : 1][NB];
  int ncol[NB];
};

// n-blocks nb0 + WV*nb (< ncb) of this wave.  With pieces, those indices count a list of
// 16-column n-blocks instead: index j is n-block 2 pieces[j / 2] + j % 2 (the two halves of
// 32-column piece pieces[j / 2]).  Every prefetch load is issued unconditionally (k-step and
// n-block clamped into range), so that the compiler can count the loads in flight and wait only
// for the one it needs; a conditional load makes it wait for all of them.
template <typename T, int NB, int PF, int WV, bool COMP>
__device__ __forceinline__ void ring_load(WRing<T, NB, PF, COMP>& r, const T* __restrict__ W, int ldw, int nb0,
                                          int ncb, int K, const T* __restrict__ Wlo, const int* pieces = nullptr) {
  typedef Tr<T> tr;
  const int nsteps = K / tr::KS;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int j = min(nb0 + WV * nb, ncb - 1);
    r.ncol[nb] = (pieces ? 2 * pieces[j >> 1] + (j & 1) : j) * 16;
  }
#pragma unroll
  for (int u = 0; u < PF; ++u)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      r.q[u][nb] = tr::wload(W, ldw, r.ncol[nb], min(u, nsteps - 1) * tr::KS);
      if constexpr (COMP) r.lo[u][nb] = tr::wload(Wlo, ldw, r.ncol[nb], min(u, nsteps - 1) * tr::KS);
    }
}

// acc[nb][mb] += sum_k W(n, k) act(k, m) over k in [0, K) for the n-blocks of ring_load, whose
// first PF k-steps r holds; the later ones are loaded PF steps ahead.
// COMP: the weights are the sum of two BF16 matrices (value and rounding remainder, Wlo), each
// multiplied on the MFMA: the product keeps ~16 significant bits of W instead of 8.
template <typename T, int NB, int MB, int PF, int WV, bool COMP = false>
__device__ __forceinline__ void gemm_r(f4 (&acc)[NB][MB], WRing<T, NB, PF, COMP>& r, const T* __restrict__ W,
                                       int ldw, int nb0, int ncb, const T* act, int lda, int K,
                                       const T* __restrict__ Wlo = nullptr) {
  typedef Tr<T> tr;
  const int nsteps = K / tr::KS;
  typename tr::frag b[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) b[mb] = tr::aload(act, lda, mb * 16, 0);
  for (int s0 = 0; s0 < nsteps; s0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int s = s0 + u;
      if (s < nsteps) {
        // activations of the next k-step, read before this step's MFMAs
        typename tr::frag bn[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) bn[mb] = tr::aload(act, lda, mb * 16, min(s + 1, nsteps - 1) * tr::KS);
        // keep the next step's LDS reads ahead of this step's MFMAs (the scheduler would otherwise
        // sink them to their uses and expose the LDS latency)
        __builtin_amdgcn_sched_barrier(0);
        const int sn = min(s + PF, nsteps - 1);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          if (nb0 + WV * nb < ncb) {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc[nb][mb] = tr::mma(r.q[u][nb], b[mb], acc[nb][mb]);
            if constexpr (COMP) {
#pragma unroll
              for (int mb = 0; mb < MB; ++mb) acc[nb][mb] = tr::mma(r.lo[u][nb], b[mb], acc[nb][mb]);
            }
          }
          r.q[u][nb] = tr::wload(W, ldw, r.ncol[nb], sn * tr::KS);
          if constexpr (COMP) r.lo[u][nb] = tr::wload(Wlo, ldw, r.ncol[nb], sn * tr::KS);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) b[mb] = bn[mb];
      }
    }
  }
}

// forward epilogue: h = celu(acc + b) -> acc and (STORE) LDS [m][n]
template <typename T, int NB, int MB, int WV, bool STORE = true>
__device__ __forceinline__ void fwd_epilogue(f4 (&acc)[NB][MB], const float* __restrict__ bias, T* H, int ldh,
                                             int nb0, int ncb, float alpha) {
  const int l = fnl_lane();
  const float ia = 1.0f / alpha;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int cb = nb0 + WV * nb;
    if (cb >= ncb) continue;
    const int n = cb * 16 + 4 * (l >> 4);
    const f4 bv = *(const f4*)(bias + n);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      f4 v = acc[nb][mb] + bv;
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = v[r] > 0.0f ? v[r] : alpha * (__expf(v[r] * ia) - 1.0f);
      acc[nb][mb] = v;
      if constexpr (STORE) Tr<T>::store4(H + (mb * 16 + (l & 15)) * ldh + n, v);
    }
  }
}

// backward epilogue: G = acc * celu'(z) with celu'(z) = h > 0 ? 1 : 1 + h/alpha,
// h read from the LDS buffer that G then overwrites
template <typename T, int NB, int MB, int WV>
__device__ __forceinline__ void bwd_epilogue(const f4 (&acc)[NB][MB], T* HG, int ld, int nb0, int ncb,
                                             float alpha) {
  const int l = fnl_lane();
  const float ia = 1.0f / alpha;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int cb = nb0 + WV * nb;
    if (cb >= ncb) continue;
    const int n = cb * 16 + 4 * (l >> 4);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      T* p = HG + (mb * 16 + (l & 15)) * ld + n;
      const f4 h = Tr<T>::load4(p);
      f4 g;
#pragma unroll
      for (int r = 0; r < 4; ++r) g[r] = acc[nb][mb][r] * (h[r] > 0.0f ? 1.0f : 1.0f + h[r] * ia);
      Tr<T>::store4(p, g);
    }
  }
}

// COMP / COMPB: compensated BF16 weights (hi + lo products) in the forward / backward GEMMs
template <typename T, int M_, int PF_, int WV_, bool COMP = false, bool COMPB = COMP>
__global__ void __launch_bounds__(64 * WV_, (Lds<T, M_, WV_>::BLOCKS_PER_CU))
k_species_mlp(const T* __restrict__ aev, int aev_stride, MlpArgs args, const int32_t* __restrict__ perm,
              const int32_t* __restrict__ tile_start, const int32_t* __restrict__ tile_rows,
              const int32_t* __restrict__ tile_net, int n_tiles, float* __restrict__ energy,
              void* __restrict__ grad, int grad_stride, int grad_bf16) {
  typedef Tr<T> tr;
  typedef Lds<T, M_, WV_> L;
  constexpr int M = M_;
  constexpr int PF = PF_;
  constexpr int WV = WV_;                    // waves per workgroup
  constexpr int NB1 = 16 / WV;               // n-blocks per wave: H1 = 256 features
  constexpr int NB2 = (H2 / 16 + WV - 1) / WV;
  constexpr int NB3 = (H3 / 16 + WV - 1) / WV;
  constexpr int MB = M / 16;
  constexpr int KC = tr::KC;
  __shared__ __attribute__((aligned(16))) unsigned char smem[L::BYTES];
  T* B1 = (T*)(smem + L::OFF1);
  T* B2 = (T*)(smem + L::OFF2);
  T* B3 = (T*)(smem + L::OFF3);
  float* Epart = (float*)(smem + L::OFF3);  // [H3/16][M], in the H3 buffer (h3 is kept in registers)
  T* Xs0 = B2;
  T* Xs1 = (T*)(smem + L::OFF2 + L::XBUF);

  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int l = fnl_lane();
  // Workgroups go to the XCDs round-robin (workgroup b -> XCD b % 8).  Each XCD takes a contiguous
  // range of the network-sorted tile list, so its L2 holds the weights of one or two networks.
  const int nt = args.n_tiles_dev ? min(*args.n_tiles_dev, n_tiles) : n_tiles;
  const int per_xcd = (nt + kXcds - 1) / kXcds;
  const int tile = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (tile >= nt) return;
  const int net = tile_net[tile];
  const int nrows = tile_rows[tile];
  if (nrows <= 0) return;
  const int r0 = tile_start[tile];
  const NetPtrs& P = args.nets[net];
  const int K = args.k_pad;
  const float alpha = args.alpha;

  // ---------------- layer 1: X (global, staged) -> H1 ----------------
  constexpr int VEC = 16 / sizeof(T);          // elements per 16-byte load
  constexpr int TPR = KC / VEC;                // threads per row of a chunk
  constexpr int RPP = 64 * WV / TPR;           // rows per pass
  constexpr int NPASS = M / RPP;
  static_assert(M % RPP == 0, "staging passes");
  // the tile's atoms, read once: the energy and dE/dAEV stores index them (no global load there)
  __shared__ int atom_sh[M];
  if (t < M) atom_sh[t] = t < nrows ? perm[r0 + t] : -1;
  int srow[NPASS];
  const int row_safe = perm[r0];
  const int scol = (t % TPR) * VEC;
#pragma unroll
  for (int p = 0; p < NPASS; ++p) {
    const int row = p * RPP + t / TPR;
    srow[p] = row < nrows ? perm[r0 + row] : -1;
  }
  // Input pieces of 32 columns that hold a non-zero column for some atom of the batch (k_live,
  // bit q: columns [32q, 32q + 32); columns from 2048 on are always live): only those are
  // multiplied (the others contribute exact zeros).  live_piece[v] is the v-th of them, padded
  // to whole chunks with a dead piece (its columns are zero in every row); chunk c stages pieces
  // c*PPC .. c*PPC + PPC - 1, and k-step s of the loop is column
  // live_piece[s / SPP] * PW + (s % SPP) * KS.
  constexpr int PW = 32;            // piece width (columns)
  constexpr int PPC = KC / PW;      // pieces per chunk
  constexpr int SPP = PW / tr::KS;  // k-steps per piece
  static_assert(KC % PW == 0 && PW % tr::KS == 0 && PW % VEC == 0, "pieces tile the chunks");
  __shared__ int live_piece[FNL_MLP_MAX_CHUNKS * 4];
  __shared__ int n_chunks_sh, n_live_sh;
  if (t < 64) {  // wave 0 compacts the list with a ballot (pieces from 64 on are always live)
    const int nq = K / PW;
    const bool lv = t < nq && ((args.k_live >> t) & 1ull);
    const unsigned long long m = __ballot(lv);
    if (lv) live_piece[fnl_rank_below(m)] = t;
    int n = __popcll(m);
    for (int q = 64 + t; q < nq; q += 64) live_piece[n + q - 64] = q;
    n += max(nq - 64, 0);
    if (t == 0) {
      n_live_sh = n;  // the live pieces: the columns whose dE/dAEV is written
      const unsigned long long dead_m = ~m & (nq >= 64 ? ~0ull : (1ull << nq) - 1ull);
      const int dead = dead_m ? __builtin_ctzll(dead_m) : -1;
      if (n == 0) live_piece[n++] = dead;  // no live column: one zero piece, the bias-only result
      while (n % PPC) live_piece[n++] = dead;  // a dead piece exists unless every piece is live
      n_chunks_sh = n / PPC;
    }
  }
  __syncthreads();
  const int nlive = n_chunks_sh;
  // two register sets: the chunk two ahead is in flight while the next one waits for its LDS buffer
  u4v xa[NPASS], xb[NPASS];
  // chunk c of the live pieces: this thread's 16 bytes lie in piece c * PPC + scol / 32
  auto stage_load = [&](u4v (&xr)[NPASS], int c) {
    const int col = live_piece[c * (KC / 32) + scol / 32] * 32 + scol % 32;
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
      const int rr = srow[p] >= 0 ? srow[p] : row_safe;  // a real row of the tile: the load is always issued
      const u4v v = __builtin_nontemporal_load((const u4v*)(aev + (size_t)rr * aev_stride + col));
      xr[p] = srow[p] >= 0 ? v : u4v{0u, 0u, 0u, 0u};
    }
  };
  auto stage_store = [&](const u4v (&xr)[NPASS], T* buf) {
#pragma unroll
    for (int p = 0; p < NPASS; ++p) *(u4v*)(buf + (p * RPP + t / TPR) * L::LDX + scol) = xr[p];
  };


  f4 acc1[NB1][MB];
  zero_acc(acc1);
  {
    constexpr int SPC = KC / tr::KS;  // k-steps per chunk
    static_assert(SPC % PF == 0, "prefetch ring aligned to chunks");
    const T* W = (const T*)P.fw[0];
    const T* Wl = (const T*)P.fw_lo[0];
    const int ldw = sizeof(T) == 2 ? K : H1;
    const int nsteps = nlive * SPC;
    auto kstep = [&](int s) { return live_piece[s / SPP] * PW + (s % SPP) * tr::KS; };
    typename tr::frag wq[PF][NB1];
    typename tr::frag wl[COMP ? PF : 1][NB1];
#pragma unroll
    for (int u = 0; u < PF; ++u)
#pragma unroll
      for (int nb = 0; nb < NB1; ++nb) {
        wq[u][nb] = tr::wload(W, ldw, (wave + WV * nb) * 16, kstep(min(u, nsteps - 1)));
        if constexpr (COMP) wl[u][nb] = tr::wload(Wl, ldw, (wave + WV * nb) * 16, kstep(min(u, nsteps - 1)));
      }
    const int nchunks = nlive;
    stage_load(xa, 0);
    stage_load(xb, min(1, nchunks - 1));
    stage_store(xa, Xs0);
    __syncthreads();
    // chunk c: computed from LDS buffer c&1; chunk c+1 in flight in one register set (stored at the
    // end of the iteration), chunk c+2 issued into the other.  All loads are issued (clamped chunk
    // index) so that the compiler can wait for the older set only.
    auto chunk = [&](int c, u4v (&xnext)[NPASS], u4v (&xfree)[NPASS]) {
      const T* X = (c & 1) ? Xs1 : Xs0;
      stage_load(xfree, min(c + 2, nchunks - 1));
      typename tr::frag b[MB];
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) b[mb] = tr::aload(X, L::LDX, mb * 16, 0);
      for (int s0 = 0; s0 < SPC; s0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int s = c * SPC + s0 + u;  // global k-step
          // activations of the next k-step of the chunk, read before this step's MFMAs
          typename tr::frag bn[MB];
          const int kn = min(s0 + u + 1, SPC - 1) * tr::KS;
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) bn[mb] = tr::aload(X, L::LDX, mb * 16, kn);
          __builtin_amdgcn_sched_barrier(0);  // next step's LDS reads stay ahead (see gemm_r)
          const int kpf = kstep(min(s + PF, nsteps - 1));
#pragma unroll
          for (int nb = 0; nb < NB1; ++nb) {
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc1[nb][mb] = tr::mma(wq[u][nb], b[mb], acc1[nb][mb]);
            if constexpr (COMP) {
#pragma unroll
              for (int mb = 0; mb < MB; ++mb) acc1[nb][mb] = tr::mma(wl[u][nb], b[mb], acc1[nb][mb]);
            }
            wq[u][nb] = tr::wload(W, ldw, (wave + WV * nb) * 16, kpf);
            if constexpr (COMP) wl[u][nb] = tr::wload(Wl, ldw, (wave + WV * nb) * 16, kpf);
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) b[mb] = bn[mb];
        }
      }
      if (c + 1 < nchunks) stage_store(xnext, (c & 1) ? Xs0 : Xs1);
      __syncthreads();
    };
    for (int c = 0; c < nchunks; c += 2) {
      chunk(c, xb, xa);
      if (c + 1 < nchunks) chunk(c + 1, xa, xb);
    }
  }
  // each GEMM's first weight fragments are requested before the previous layer's epilogue (AHEAD;
  // the 32-atom BF16 tiles, three workgroups per CU, have no registers to spare for it and load
  // them at the GEMM)
  constexpr bool AHEAD = !(sizeof(T) == 2 && M <= 32);
  const int ld_f2 = sizeof(T) == 2 ? H1 : H2, ld_f3 = sizeof(T) == 2 ? H2 : H3;
  const int ld_b3 = sizeof(T) == 2 ? H3 : H2, ld_b2 = sizeof(T) == 2 ? H2 : H1, ld_b1 = sizeof(T) == 2 ? H1 : K;
  WRing<T, NB2, PF, COMP> ring2;
  auto load2 = [&] { ring_load<T, NB2, PF, WV, COMP>(ring2, (const T*)P.fw[1], ld_f2, wave, H2 / 16, H1, (const T*)P.fw_lo[1]); };
  if constexpr (AHEAD) load2();
  fwd_epilogue<T, NB1, MB, WV>(acc1, P.b[0], B1, L::LD1, wave, H1 / 16, alpha);
  __syncthreads();

  // ---------------- layer 2 ----------------
  WRing<T, NB3, PF, COMP> ring3;
  auto load3 = [&] { ring_load<T, NB3, PF, WV, COMP>(ring3, (const T*)P.fw[2], ld_f3, wave, H3 / 16, H2, (const T*)P.fw_lo[2]); };
  {
    f4 acc2[NB2][MB];
    zero_acc(acc2);
    if constexpr (!AHEAD) load2();
    gemm_r<T, NB2, MB, PF, WV, COMP>(acc2, ring2, (const T*)P.fw[1], ld_f2, wave, H2 / 16, B1, L::LD1, H1,
                                     (const T*)P.fw_lo[1]);
    if constexpr (AHEAD) load3();
    fwd_epilogue<T, NB2, MB, WV>(acc2, P.b[1], B2, L::LD2, wave, H2 / 16, alpha);
  }
  __syncthreads();

  // ---------------- layer 3 + energy ----------------
  f4 acc3[NB3][MB];
  zero_acc(acc3);
  if constexpr (!AHEAD) load3();
  gemm_r<T, NB3, MB, PF, WV, COMP>(acc3, ring3, (const T*)P.fw[2], ld_f3, wave, H3 / 16, B2, L::LD2, H2,
                                   (const T*)P.fw_lo[2]);
  WRing<T, NB2, PF, COMPB> ring_b3;
  auto load_b3 = [&] { ring_load<T, NB2, PF, WV, COMPB>(ring_b3, (const T*)P.bw[2], ld_b3, wave, H2 / 16, H3, (const T*)P.bw_lo[2]); };
  if constexpr (AHEAD) load_b3();
  // h3 stays in registers (energy and G3 are formed from acc3); B3 holds the energy partials
  fwd_epilogue<T, NB3, MB, WV, false>(acc3, P.b[2], B3, L::LD3, wave, H3 / 16, alpha);
  {
    // H3 . w4 per 16-feature block into Epart[block][atom], then summed over the blocks in order:
    // the same arithmetic for every tile size and wave count
#pragma unroll
    for (int nb = 0; nb < NB3; ++nb) {
      const int cb = wave + WV * nb;
      if (cb >= H3 / 16) continue;
      const f4 wv = *(const f4*)(P.w4 + cb * 16 + 4 * (l >> 4));
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        float v = acc3[nb][mb][0] * wv[0] + acc3[nb][mb][1] * wv[1] + acc3[nb][mb][2] * wv[2] +
                  acc3[nb][mb][3] * wv[3];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (l < 16) Epart[cb * M + mb * 16 + l] = v;
      }
    }
  }
  __syncthreads();
  if (t < nrows) {
    float e = 0.0f;
#pragma unroll
    for (int cb = 0; cb < H3 / 16; ++cb) e += Epart[cb * M + t];
    energy[atom_sh[t]] = e + P.b[3][0];
  }
  if (grad == nullptr) return;
  __syncthreads();  // energy partials read: B3 takes G3

  // ---------------- backward ----------------
  {
    // G3 = w4 * celu'(z3) -> B3 (h3 is in acc3)
    const float ia = 1.0f / alpha;
#pragma unroll
    for (int nb = 0; nb < NB3; ++nb) {
      const int cb = wave + WV * nb;
      if (cb >= H3 / 16) continue;
      const int n = cb * 16 + 4 * (l >> 4);
      const f4 wv = *(const f4*)(P.w4 + n);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        f4 g;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = acc3[nb][mb][r];
          g[r] = wv[r] * (h > 0.0f ? 1.0f : 1.0f + h * ia);
        }
        tr::store4(B3 + (mb * 16 + (l & 15)) * L::LD3 + n, g);
      }
    }
  }
  __syncthreads();
  // the live 16-column n-blocks of dE/dAEV (two per live piece), in passes of 16
  const int nlb = 2 * n_live_sh;
  WRing<T, NB1, PF, COMPB> ring_b2;
  auto load_b2 = [&] { ring_load<T, NB1, PF, WV, COMPB>(ring_b2, (const T*)P.bw[1], ld_b2, wave, H1 / 16, H2, (const T*)P.bw_lo[1]); };
  {
    f4 acc2[NB2][MB];
    zero_acc(acc2);
    if constexpr (!AHEAD) load_b3();
    gemm_r<T, NB2, MB, PF, WV, COMPB>(acc2, ring_b3, (const T*)P.bw[2], ld_b3, wave, H2 / 16, B3, L::LD3, H3,
                                     (const T*)P.bw_lo[2]);
    if constexpr (AHEAD) load_b2();
    bwd_epilogue<T, NB2, MB, WV>(acc2, B2, L::LD2, wave, H2 / 16, alpha);
  }
  __syncthreads();
  WRing<T, NB1, PF, COMPB> ring_b1;
  {
    f4 accg[NB1][MB];
    zero_acc(accg);
    if constexpr (!AHEAD) load_b2();
    gemm_r<T, NB1, MB, PF, WV, COMPB>(accg, ring_b2, (const T*)P.bw[1], ld_b2, wave, H1 / 16, B2, L::LD2, H2,
                                     (const T*)P.bw_lo[1]);
    if (AHEAD && nlb > 0)
      ring_load<T, NB1, PF, WV, COMPB>(ring_b1, (const T*)P.bw[0], ld_b1, wave, nlb, H1, (const T*)P.bw_lo[0],
                                      live_piece);
    bwd_epilogue<T, NB1, MB, WV>(accg, B1, L::LD1, wave, H1 / 16, alpha);
  }
  __syncthreads();
  // dE/dX = W1 G1: input features in passes of 16 n-blocks (16/WV per wave)
  int atom[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int m = mb * 16 + (l & 15);
    atom[mb] = atom_sh[m];
  }
  // BF16 rows: each pass's 16 n-blocks of the tile go through LDS (the H2/H3 buffers, free now)
  // so that the rows leave as 16-byte chunks
  constexpr int GLD = 256 + 8;
  static_assert(M * GLD * 2 <= L::OFFE - L::OFF2, "BF16 gradient staging fits the H2/H3 buffers");
  unsigned short* gst = (unsigned short*)(smem + L::OFF2);
  for (int base = 0; base < nlb; base += 16) {
    f4 accx[NB1][MB];
    zero_acc(accx);
    if (!AHEAD || base > 0)  // (AHEAD) the first pass's fragments were requested before the last epilogue
      ring_load<T, NB1, PF, WV, COMPB>(ring_b1, (const T*)P.bw[0], ld_b1, base + wave, nlb, H1, (const T*)P.bw_lo[0],
                                      live_piece);
    gemm_r<T, NB1, MB, PF, WV, COMPB>(accx, ring_b1, (const T*)P.bw[0], ld_b1, base + wave, nlb, B1, L::LD1, H1,
                                     (const T*)P.bw_lo[0]);
    if (grad_bf16) {
#pragma unroll
      for (int nb = 0; nb < NB1; ++nb) {
        if (base + wave + WV * nb >= nlb) continue;
        const int nl = (wave + WV * nb) * 16 + 4 * (l >> 4);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) Tr<unsigned short>::store4(gst + (mb * 16 + (l & 15)) * GLD + nl, accx[nb][mb]);
      }
      __syncthreads();
      const int cpr = 2 * min(16, nlb - base);  // 16-byte chunks per row in this pass: two per n-block
      unsigned short* gout = (unsigned short*)grad;
      for (int idx = t; idx < M * cpr; idx += 64 * WV) {
        const int m = idx / cpr, q = idx - m * cpr;
        const int j = base + (q >> 1);
        const int col = (2 * live_piece[j >> 1] + (j & 1)) * 16 + (q & 1) * 8;
        if (m < nrows)
          __builtin_nontemporal_store(*(const u4v*)(gst + m * GLD + q * 8),
                                      (u4v*)(gout + (size_t)atom_sh[m] * grad_stride + col));
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int nb = 0; nb < NB1; ++nb) {
        const int j = base + wave + WV * nb;
        if (j >= nlb) continue;
        const int n = (2 * live_piece[j >> 1] + (j & 1)) * 16 + 4 * (l >> 4);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          if (atom[mb] >= 0)
            __builtin_nontemporal_store(accx[nb][mb], (f4*)((float*)grad + (size_t)atom[mb] * grad_stride + n));
      }
    }
  }
}

}  // namespace

extern "C" int fnl_mlp_tile_rows(int32_t dtype) {
  if (dtype == FNL_F32) return 32;
  if (dtype == FNL_BF16) return 96;
  return FNL_ERR_INVALID;
}

extern "C" int fnl_species_mlp(const void* aev, int32_t aev_stride, const fnl_mlp_weights* w,
                               const int32_t* perm, const int32_t* tile_start, const int32_t* tile_rows,
                               const int32_t* tile_net, int32_t n_tiles, float* energy, void* grad,
                               int32_t grad_stride, int32_t grad_dtype, uint64_t k_live,
                               const int32_t* n_tiles_dev, fnl_stream_t stream) {
  if (!w || n_tiles < 0) return FNL_ERR_INVALID;
  if (n_tiles == 0) return FNL_OK;
  if (!aev || !perm || !tile_start || !tile_rows || !tile_net || !energy) return FNL_ERR_INVALID;
  if (w->n_nets < 1 || w->n_nets > FNL_MLP_MAX_NETS) return FNL_ERR_INVALID;
  const bool bf = w->dtype == FNL_BF16;
  if (!bf && w->dtype != FNL_F32) return FNL_ERR_INVALID;
  const int kc = bf ? Tr<unsigned short>::KC : Tr<float>::KC;
  if (w->k_pad < kc || w->k_pad % kc != 0 || aev_stride < w->k_pad ||
      aev_stride % 8 != 0)
    return FNL_ERR_INVALID;
  if (w->k_pad > FNL_MLP_MAX_CHUNKS * kc) return FNL_ERR_UNSUPPORTED;  // the kernel's live-chunk table
  if (grad && (grad_stride < w->k_pad || grad_stride % 8 != 0)) return FNL_ERR_INVALID;
  if (grad_dtype != FNL_F32 && !(grad_dtype == FNL_BF16 && bf)) return FNL_ERR_INVALID;
  const int gbf = grad_dtype == FNL_BF16 ? 1 : 0;
  MlpArgs a;
  a.k_pad = w->k_pad;
  a.alpha = w->celu_alpha;
  a.k_live = k_live ? (unsigned long long)k_live : ~0ull;
  a.n_tiles_dev = n_tiles_dev;
  // compensated BF16 forward (backward) products when every network carries the three forward
  // (backward) remainder matrices; backward remainders only with the forward ones
  bool comp = bf, comp_b = bf;
  for (int n = 0; n < FNL_MLP_MAX_NETS; ++n) {
    NetPtrs& p = a.nets[n];
    if (n >= w->n_nets) {
      p = a.nets[0];
      continue;
    }
    const fnl_mlp_net& s = w->nets[n];
    if (!s.w1t || !s.w2t || !s.w3t || !s.w1 || !s.w2 || !s.w3 || !s.w4 || !s.b1 || !s.b2 || !s.b3 || !s.b4)
      return FNL_ERR_INVALID;
    const bool f_lo = s.w1t_lo && s.w2t_lo && s.w3t_lo, b_lo = s.w1_lo && s.w2_lo && s.w3_lo;
    const bool any_f = s.w1t_lo || s.w2t_lo || s.w3t_lo, any_b = s.w1_lo || s.w2_lo || s.w3_lo;
    // a set is all three or none, BF16 only, the backward set only with the forward one
    if ((any_f && !f_lo) || (any_b && !b_lo) || ((any_f || any_b) && !bf) || (b_lo && !f_lo))
      return FNL_ERR_INVALID;
    comp = comp && f_lo;
    comp_b = comp_b && b_lo;
    p.fw_lo[0] = s.w1t_lo;
    p.fw_lo[1] = s.w2t_lo;
    p.fw_lo[2] = s.w3t_lo;
    p.bw_lo[0] = s.w1_lo;
    p.bw_lo[1] = s.w2_lo;
    p.bw_lo[2] = s.w3_lo;
    // A operands: K-contiguous for BF16 (8 k per lane), N-contiguous for FP32
    // (16 lanes read 16 consecutive n); forward uses W^T, backward W.
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
  const int M = w->tile_rows > 0 ? w->tile_rows : fnl_mlp_tile_rows(w->dtype);
  const int grid = (n_tiles + kXcds - 1) / kXcds * kXcds;
#define FNL_MLP_BF16(M_, WV_)                                                                                  \
  do {                                                                                                         \
    if (comp && comp_b)                                                                                        \
      hipLaunchKernelGGL((k_species_mlp<unsigned short, M_, 2, WV_, true>), dim3(grid), dim3(64 * WV_), 0, st, \
                         (const unsigned short*)aev, aev_stride, a, perm, tile_start, tile_rows, tile_net,     \
                         n_tiles, energy, grad, grad_stride, gbf);                                             \
    else if (comp)                                                                                             \
      hipLaunchKernelGGL((k_species_mlp<unsigned short, M_, 2, WV_, true, false>), dim3(grid), dim3(64 * WV_), \
                         0, st, (const unsigned short*)aev, aev_stride, a, perm, tile_start, tile_rows,        \
                         tile_net, n_tiles, energy, grad, grad_stride, gbf);                                   \
    else                                                                                                       \
      hipLaunchKernelGGL((k_species_mlp<unsigned short, M_, 2, WV_, false>), dim3(grid), dim3(64 * WV_), 0,   \
                         st, (const unsigned short*)aev, aev_stride, a, perm, tile_start, tile_rows, tile_net, \
                         n_tiles, energy, grad, grad_stride, gbf);                                             \
  } while (0)
  if (bf && M == 64) {
    FNL_MLP_BF16(64, 4);
  } else if (bf && M == 96) {  // eight waves: two per SIMD cover each other's load latency
    FNL_MLP_BF16(96, 8);
  } else if (bf && M == 32) {
    FNL_MLP_BF16(32, 4);
  } else if (!bf && M == 32) {
    hipLaunchKernelGGL((k_species_mlp<float, 32, 4, 4>), dim3(grid), dim3(256), 0, st, (const float*)aev,
                       aev_stride, a, perm, tile_start, tile_rows, tile_net, n_tiles, energy, grad,
                       grad_stride, gbf);
  } else if (!bf && M == 16) {
    hipLaunchKernelGGL((k_species_mlp<float, 16, 2, 4>), dim3(grid), dim3(256), 0, st, (const float*)aev,
                       aev_stride, a, perm, tile_start, tile_rows, tile_net, n_tiles, energy, grad,
                       grad_stride, gbf);
  } else {
    return FNL_ERR_UNSUPPORTED;
  }
#undef FNL_MLP_BF16
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
