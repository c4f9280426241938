// Weight gradient of an energy + force loss through the species MLPs (FP32 on MFMA).
//
// The loss of FeNNol's train_step (training/utils.py:140-525, jax.grad of the force-carrying loss
// at :508) differentiates the forces once more with respect to the weights.  Its MLP part, for
// one network with input rows A (AEV), tangent rows dA = dAEV/dx . r (r = F - F_ref), per-atom
// energy adjoints cE and the force adjoint cD, is
//   forward   z_l = h_{l-1} W_l + b_l, dz_l = dh_{l-1} W_l, h_l = celu(z_l), dh_l = celu'(z_l) dz_l
//   reverse   bh3 = cE w4, bdh3 = cD w4,
//             bdz_l = s_l bdh_l, bz_l = t_l dz_l bdh_l + s_l bh_l   (s = celu', t = celu'')
//             bdh_{l-1} = bdz_l W_l^T, bh_{l-1} = bz_l W_l^T
//   gradient  dL/dW_l = dh_{l-1}^T bdz_l + h_{l-1}^T bz_l, dL/db_l = sum bz_l,
//             dL/dw4 = sum (h3 cE + cD dh3), dL/db4 = sum cE
// (h_0 = A, dh_0 = dA).  Three kernels:
//   k_train_rows   one 64-atom tile x 64 output columns, both chains on the same weight operands,
//                  with the CELU parts, the start of the reverse pass and per-tile column sums
//                  fused into the epilogue (modes FWD1..3, BWD2, BWD1);
//   k_train_wgrad  X^T Y for both chains of one layer, one 64 x 64 block of dL/dW per workgroup,
//                  over a chunk of RC atoms (the reduction runs over atoms);
//   k_train_reduce the chunk and tile partials of each network summed in a fixed order, so the
//                  gradient is the same bits every run.
// Layer-1 work covers only the 32-column AEV pieces in k_live (the other columns are zero).

#include "fnl_common.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int H1 = FNL_MLP_H1, H2 = FNL_MLP_H2, H3 = FNL_MLP_H3;
constexpr int TM = 64;       // atoms per row tile
constexpr int TN = 64;       // output columns per workgroup
constexpr int KC = 32;       // K chunk staged in LDS
constexpr int RK = 32;       // atoms per staged step of the weight-gradient kernel
constexpr int RC = 1024;     // atoms per weight-gradient chunk
constexpr int LDX = KC + 1;  // LDS row strides (floats)
constexpr int LDW = TN + 4;
constexpr int LDG = 64 + 4;
// per-tile partial sums: gb1 [H1], gb2 [H2], gb3 [H3], gw4 [H3], gb4 [1]
constexpr int TP_GB1 = 0, TP_GB2 = H1, TP_GB3 = H1 + H2, TP_GW4 = H1 + H2 + H3, TP_GB4 = H1 + H2 + 2 * H3;
constexpr int TP_STRIDE = TP_GB4 + 4;

enum Mode { FWD1 = 0, FWD2 = 1, FWD3 = 2, BWD2 = 3, BWD1 = 4 };

struct TrainArgs {
  int n_nets;
  int k_in;
  int rows[FNL_MLP_MAX_NETS + 1];    // rows of each network in perm order (offsets)
  int tiles[FNL_MLP_MAX_NETS + 1];   // 64-atom tiles (offsets)
  int chunks[FNL_MLP_MAX_NETS + 1];  // RC-atom chunks (offsets)
  unsigned long long k_live;
  float alpha;
  float c_force;
  fnl_train_net net[FNL_MLP_MAX_NETS];
};

struct TrainBufs {
  const float* aev;
  const float* daev;
  int aev_ld, daev_ld;
  const int32_t* perm;
  const float* c_energy;
  float *z1, *dz1, *h1, *dh1;  // [rows][H1]; z1 / dz1 hold bz1 / bdz1 after BWD1
  float *z2, *dz2, *h2, *dh2;  // [rows][H2]; z2 / dz2 hold bz2 / bdz2 after BWD2
  float *bz3, *bdz3;           // [rows][H3]
  float* tile_part;            // [tiles][TP_STRIDE]
  float* part1;                // [chunks][k_in][H1]
  float* part2;                // [chunks][H1][H2]
  float* part3;                // [chunks][H2][H3]
};

__device__ __forceinline__ f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// CELU(alpha) and its first two derivatives
__device__ __forceinline__ void celu3(float z, float alpha, float& h, float& s, float& t) {
  if (z > 0.0f) {
    h = z;
    s = 1.0f;
    t = 0.0f;
  } else {
    const float e = expf(z / alpha);
    h = alpha * (e - 1.0f);
    s = e;
    t = e / alpha;
  }
}

__device__ __forceinline__ int find_net(const int* off, int n_nets, int idx) {
  int n = 0;
  while (n + 1 < n_nets && off[n + 1] <= idx) ++n;
  return n;
}

// k_live bit q: AEV columns [32q, 32q + 32) may be non-zero (columns from 2048 on: always)
__device__ __forceinline__ bool chunk_live(unsigned long long live, int k0) {
  const int q = k0 >> 5;
  return q >= 64 || ((live >> q) & 1ull);
}

// the 64 columns from k0 (a multiple of 64): live if either of their pieces is
__device__ __forceinline__ bool block64_live(unsigned long long live, int k0) {
  const int q = k0 >> 5;
  return q >= 64 || ((live >> q) & 3ull);
}

// ---------------------------------------------------------------------------------------------
// Row tiles: [Zp | Zt] = [Xp | Xt] W over one 64-atom tile and 64 output columns, then the
// mode's epilogue.
// ---------------------------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(256) k_train_rows(const TrainArgs a, const TrainBufs bf) {
  constexpr int K_ = MODE == FWD1 ? 0 : (MODE == FWD2 ? H1 : (MODE == FWD3 ? H2 : (MODE == BWD2 ? H3 : H2)));
  constexpr int N = MODE == FWD1 ? H1 : (MODE == FWD2 ? H2 : (MODE == FWD3 ? H3 : (MODE == BWD2 ? H2 : H1)));
  constexpr bool TRANS = MODE == BWD2 || MODE == BWD1;  // W^T: element (k, n) = W[n * ldw + k]
  __shared__ float Xs[2][TM][LDX];
  __shared__ float Ws[KC][LDW];
  __shared__ float red[4][TN];
  __shared__ float red4[4];
  const int t = threadIdx.x, wave = t >> 6, l = fnl_lane();
  const int tile = blockIdx.x;
  if (tile >= a.tiles[a.n_nets]) return;
  const int n0 = blockIdx.y * TN;
  if (n0 >= N) return;
  const int net = find_net(a.tiles, a.n_nets, tile);
  const int r0 = a.rows[net] + (tile - a.tiles[net]) * TM;
  const int nrows = min(TM, a.rows[net + 1] - r0);
  const int K = MODE == FWD1 ? a.k_in : K_;
  const fnl_train_net& P = a.net[net];
  const float* W = MODE == FWD1 ? P.w1 : (MODE == FWD2 ? P.w2 : (MODE == FWD3 ? P.w3 : (MODE == BWD2 ? P.w3 : P.w2)));
  const int ldw = MODE == FWD1 ? H1 : (MODE == FWD2 ? H2 : (MODE == FWD3 ? H3 : (MODE == BWD2 ? H3 : H2)));
  const float* xp;
  const float* xt;
  int ldx;
  if (MODE == FWD1) { xp = bf.aev; xt = bf.daev; ldx = 0; }
  else if (MODE == FWD2) { xp = bf.h1; xt = bf.dh1; ldx = H1; }
  else if (MODE == FWD3) { xp = bf.h2; xt = bf.dh2; ldx = H2; }
  else if (MODE == BWD2) { xp = bf.bz3; xt = bf.bdz3; ldx = H3; }
  else { xp = bf.z2; xt = bf.dz2; ldx = H2; }  // bz2 / bdz2

  // staging coordinates: X rows (t >> 2, 8 columns from 8 (t & 3)); W (forward: k row t >> 3,
  // 8 columns from 8 (t & 7); transposed: column t >> 2, 8 k from 8 (t & 3))
  const int xr = t >> 2, xc = (t & 3) * 8;
  const bool xrow_ok = xr < nrows;
  size_t xoff_p = 0, xoff_t = 0;
  if (xrow_ok) {
    if (MODE == FWD1) {
      const int atom = bf.perm[r0 + xr];
      xoff_p = (size_t)atom * bf.aev_ld;
      xoff_t = (size_t)atom * bf.daev_ld;
    } else {
      xoff_p = xoff_t = (size_t)(r0 + xr) * ldx;
    }
  }
  f4 acc[2][4];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[c][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};

  // the next live K chunk's X rows and W block, loaded into registers one chunk ahead so that
  // their latency overlaps the current chunk's MFMAs
  f4 sx[4], sw[2];
  auto next_live = [&](int k) {
    if (MODE == FWD1)
      while (k < K && !chunk_live(a.k_live, k)) k += KC;
    return k;
  };
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) sx[q] = f4{0.f, 0.f, 0.f, 0.f};
    sw[0] = sw[1] = f4{0.f, 0.f, 0.f, 0.f};
    const int k = k0 + xc;
    if (xrow_ok && k < K) {
      sx[0] = *(const f4*)(xp + xoff_p + k);
      sx[1] = *(const f4*)(xp + xoff_p + k + 4);
      sx[2] = *(const f4*)(xt + xoff_t + k);
      sx[3] = *(const f4*)(xt + xoff_t + k + 4);
    }
    if (!TRANS) {
      const int kk = t >> 3, nn = (t & 7) * 8;
      if (k0 + kk < K && n0 + nn < N) {
        sw[0] = *(const f4*)(W + (size_t)(k0 + kk) * ldw + n0 + nn);
        sw[1] = *(const f4*)(W + (size_t)(k0 + kk) * ldw + n0 + nn + 4);
      }
    } else {
      const int nn = t >> 2, kk = (t & 3) * 8;
      if (n0 + nn < N && k0 + kk < K) {
        sw[0] = *(const f4*)(W + (size_t)(n0 + nn) * ldw + k0 + kk);
        sw[1] = *(const f4*)(W + (size_t)(n0 + nn) * ldw + k0 + kk + 4);
      }
    }
  };
  int k0 = next_live(0);
  if (k0 < K) fetch(k0);
  while (k0 < K) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      Xs[0][xr][xc + e] = sx[0][e];
      Xs[0][xr][xc + 4 + e] = sx[1][e];
      Xs[1][xr][xc + e] = sx[2][e];
      Xs[1][xr][xc + 4 + e] = sx[3][e];
    }
    if (!TRANS) {
      const int kk = t >> 3, nn = (t & 7) * 8;
      *(f4*)&Ws[kk][nn] = sw[0];
      *(f4*)&Ws[kk][nn + 4] = sw[1];
    } else {
      const int nn = t >> 2, kk = (t & 3) * 8;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        Ws[kk + e][nn] = sw[0][e];
        Ws[kk + 4 + e][nn] = sw[1][e];
      }
    }
    __syncthreads();
    const int kn = next_live(k0 + KC);
    if (kn < K) fetch(kn);
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const float bp = Xs[0][wave * 16 + (l & 15)][kk + (l >> 4)];
      const float bt = Xs[1][wave * 16 + (l & 15)][kk + (l >> 4)];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const float aw = Ws[kk + (l >> 4)][nb * 16 + (l & 15)];
        acc[0][nb] = mma(aw, bp, acc[0][nb]);
        acc[1][nb] = mma(aw, bt, acc[1][nb]);
      }
    }
    __syncthreads();
    k0 = kn;
  }

  // epilogue: lane holds column n0 + nb*16 + 4*(l >> 4) + v of tile row m = wave*16 + (l & 15)
  const int m = wave * 16 + (l & 15);
  const bool row_ok = m < nrows;
  const size_t row = (size_t)(r0 + m);
  const float alpha = a.alpha;
  float cE = 0.0f;
  if (MODE == FWD3 && row_ok) cE = bf.c_energy[bf.perm[r0 + m]];
  float colsum[4][4];   // the column sums of this mode (gb) over this lane's row
  float colsum2[4][4];  // FWD3: gw4
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const int nc = n0 + nb * 16 + 4 * (l >> 4);
    const bool col_ok = nc < N;  // N is a multiple of 16: the 4 columns are all in or all out
    f4 p = acc[0][nb], q = acc[1][nb];
    f4 o1 = f4{0.f, 0.f, 0.f, 0.f}, o2 = o1, o3 = o1, o4 = o1;
    float cs[4] = {0.f, 0.f, 0.f, 0.f}, cs2[4] = {0.f, 0.f, 0.f, 0.f};
    if (row_ok && col_ok) {
      if (MODE == FWD1 || MODE == FWD2 || MODE == FWD3) {
        const float* bias = MODE == FWD1 ? P.b1 : (MODE == FWD2 ? P.b2 : P.b3);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float z = p[v] + bias[nc + v];
          const float dz = q[v];
          float h, s, tt;
          celu3(z, alpha, h, s, tt);
          if (MODE == FWD3) {
            const float w4 = P.w4[nc + v];
            const float bdh = a.c_force * w4, bh = cE * w4;
            o1[v] = s * bdh;                   // bdz3
            o2[v] = tt * dz * bdh + s * bh;    // bz3
            cs[v] = o2[v];                     // gb3
            cs2[v] = h * cE + a.c_force * (s * dz);  // gw4
          } else {
            o1[v] = z;
            o2[v] = dz;
            o3[v] = h;
            o4[v] = s * dz;
          }
        }
        if (MODE == FWD1) {
          *(f4*)(bf.z1 + row * H1 + nc) = o1;
          *(f4*)(bf.dz1 + row * H1 + nc) = o2;
          *(f4*)(bf.h1 + row * H1 + nc) = o3;
          *(f4*)(bf.dh1 + row * H1 + nc) = o4;
        } else if (MODE == FWD2) {
          *(f4*)(bf.z2 + row * H2 + nc) = o1;
          *(f4*)(bf.dz2 + row * H2 + nc) = o2;
          *(f4*)(bf.h2 + row * H2 + nc) = o3;
          *(f4*)(bf.dh2 + row * H2 + nc) = o4;
        } else {
          *(f4*)(bf.bdz3 + row * H3 + nc) = o1;
          *(f4*)(bf.bz3 + row * H3 + nc) = o2;
        }
      } else {
        // p = bh, q = bdh of this layer; its pre-activations z, dz give s and t
        float* zb = MODE == BWD2 ? bf.z2 : bf.z1;
        float* dzb = MODE == BWD2 ? bf.dz2 : bf.dz1;
        const int ld = MODE == BWD2 ? H2 : H1;
        const f4 z = *(const f4*)(zb + row * ld + nc);
        const f4 dz = *(const f4*)(dzb + row * ld + nc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float h, s, tt;
          celu3(z[v], alpha, h, s, tt);
          o1[v] = s * q[v];                         // bdz
          o2[v] = tt * dz[v] * q[v] + s * p[v];     // bz
          cs[v] = o2[v];
        }
        *(f4*)(zb + row * ld + nc) = o2;   // bz over z
        *(f4*)(dzb + row * ld + nc) = o1;  // bdz over dz
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      colsum[nb][v] = cs[v];
      colsum2[nb][v] = cs2[v];
    }
  }
  // column sums over the tile's rows: 16 lanes of a column group, then the 4 waves (fixed order)
  float* tp = bf.tile_part + (size_t)tile * TP_STRIDE;
  const int gb_off = MODE == FWD3 ? TP_GB3 : (MODE == BWD2 ? TP_GB2 : TP_GB1);
  if (MODE == FWD3 || MODE == BWD2 || MODE == BWD1) {
    for (int pass = 0; pass < (MODE == FWD3 ? 2 : 1); ++pass) {
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float x = pass == 0 ? colsum[nb][v] : colsum2[nb][v];
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) x += __shfl_xor(x, o, 64);
          if ((l & 15) == 0) red[wave][nb * 16 + 4 * (l >> 4) + v] = x;
        }
      __syncthreads();
      if (t < TN && n0 + t < N) {
        const float s = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
        tp[(pass == 0 ? gb_off : TP_GW4) + n0 + t] = s;
      }
      __syncthreads();
    }
  }
  if (MODE == FWD3 && blockIdx.y == 0) {  // gb4 = sum of cE over the tile
    float x = cE;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) x += __shfl_xor(x, o, 64);
    if (l == 0) red4[wave] = x;
    __syncthreads();
    if (t == 0) tp[TP_GB4] = ((red4[0] + red4[1]) + red4[2]) + red4[3];
  }
}

// ---------------------------------------------------------------------------------------------
// Weight gradient of one layer: G = Xt^T Yt + Xp^T Yp over a chunk of atoms, one 64 x 64 block.
// ---------------------------------------------------------------------------------------------
template <int LAYER>
__global__ void __launch_bounds__(256) k_train_wgrad(const TrainArgs a, const TrainBufs bf) {
  constexpr int NI_ = LAYER == 1 ? 0 : (LAYER == 2 ? H1 : H2);
  constexpr int N = LAYER == 1 ? H1 : (LAYER == 2 ? H2 : H3);
  __shared__ float Xs[2][RK][LDG];
  __shared__ float Ys[2][RK][LDG];
  const int t = threadIdx.x, wave = t >> 6, l = fnl_lane();
  const int chunk = blockIdx.x;
  if (chunk >= a.chunks[a.n_nets]) return;
  const int NI = LAYER == 1 ? a.k_in : NI_;
  const int i0 = blockIdx.y * 64, j0 = blockIdx.z * 64;
  if (i0 >= NI || j0 >= N) return;
  if (LAYER == 1 && !block64_live(a.k_live, i0)) return;
  const int net = find_net(a.chunks, a.n_nets, chunk);
  const int rb = a.rows[net] + (chunk - a.chunks[net]) * RC;
  const int re = min(rb + RC, a.rows[net + 1]);
  const float *xp, *xt, *yp, *yt;
  int ldx;
  if (LAYER == 1) { xp = bf.aev; xt = bf.daev; ldx = 0; yp = bf.z1; yt = bf.dz1; }
  else if (LAYER == 2) { xp = bf.h1; xt = bf.dh1; ldx = H1; yp = bf.z2; yt = bf.dz2; }
  else { xp = bf.h2; xt = bf.dh2; ldx = H2; yp = bf.bz3; yt = bf.bdz3; }
  f4 acc[4];
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) acc[jb] = f4{0.f, 0.f, 0.f, 0.f};
  const int sr = t >> 3, sc = (t & 7) * 8;  // staging: row, 8 columns
  // the staged step's rows (X, X tangent, Y, Y tangent: 8 columns each), loaded one step ahead so
  // that their latency overlaps the previous step's MFMAs
  f4 stg[8];
  auto fetch = [&](int r) {
#pragma unroll
    for (int q = 0; q < 8; ++q) stg[q] = f4{0.f, 0.f, 0.f, 0.f};
    const int rr = r + sr;
    if (rr < re) {
      if (i0 + sc < NI) {
        size_t op, ot;
        if (LAYER == 1) {
          const int atom = bf.perm[rr];
          op = (size_t)atom * bf.aev_ld + i0 + sc;
          ot = (size_t)atom * bf.daev_ld + i0 + sc;
        } else {
          op = ot = (size_t)rr * ldx + i0 + sc;
        }
        stg[0] = *(const f4*)(xp + op);
        stg[1] = *(const f4*)(xp + op + 4);
        stg[2] = *(const f4*)(xt + ot);
        stg[3] = *(const f4*)(xt + ot + 4);
      }
      if (j0 + sc < N) {
        const size_t oy = (size_t)rr * N + j0 + sc;
        stg[4] = *(const f4*)(yp + oy);
        stg[5] = *(const f4*)(yp + oy + 4);
        stg[6] = *(const f4*)(yt + oy);
        stg[7] = *(const f4*)(yt + oy + 4);
      }
    }
  };
  fetch(rb);
  for (int r = rb; r < re; r += RK) {
    *(f4*)&Xs[0][sr][sc] = stg[0];
    *(f4*)&Xs[0][sr][sc + 4] = stg[1];
    *(f4*)&Xs[1][sr][sc] = stg[2];
    *(f4*)&Xs[1][sr][sc + 4] = stg[3];
    *(f4*)&Ys[0][sr][sc] = stg[4];
    *(f4*)&Ys[0][sr][sc + 4] = stg[5];
    *(f4*)&Ys[1][sr][sc] = stg[6];
    *(f4*)&Ys[1][sr][sc + 4] = stg[7];
    __syncthreads();
    if (r + RK < re) fetch(r + RK);
#pragma unroll
    for (int kk = 0; kk < RK; kk += 4) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float ax = Xs[c][kk + (l >> 4)][wave * 16 + (l & 15)];
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) acc[jb] = mma(ax, Ys[c][kk + (l >> 4)][jb * 16 + (l & 15)], acc[jb]);
      }
    }
    __syncthreads();
  }
  // lane holds G[i0 + wave*16 + 4*(l >> 4) + v][j0 + jb*16 + (l & 15)]
  float* part = LAYER == 1 ? bf.part1 : (LAYER == 2 ? bf.part2 : bf.part3);
  part += (size_t)chunk * NI * N;
#pragma unroll
  for (int jb = 0; jb < 4; ++jb) {
    const int j = j0 + jb * 16 + (l & 15);
    if (j >= N) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + wave * 16 + 4 * (l >> 4) + v;
      if (i < NI) part[(size_t)i * N + j] = acc[jb][v];
    }
  }
}

// Per network: sums of the chunk partials (weights) and tile partials (biases, w4), fixed order.
__global__ void __launch_bounds__(256) k_train_reduce(const TrainArgs a, const TrainBufs bf, float* __restrict__ grad,
                                                      size_t gstride) {
  const int net = blockIdx.y;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= gstride) return;
  const size_t s1 = (size_t)a.k_in * H1, s2 = s1 + H1, s3 = s2 + (size_t)H1 * H2, s4 = s3 + H2,
               s5 = s4 + (size_t)H2 * H3, s6 = s5 + H3, s7 = s6 + H3;
  const int c0 = a.chunks[net], c1 = a.chunks[net + 1];
  const int t0 = a.tiles[net], t1 = a.tiles[net + 1];
  float s = 0.0f;
  // chunk partials: eight independent loads in flight, summed in chunk order
  auto sum_chunks = [&](const float* part, size_t cstride, size_t off) {
    float acc = 0.0f;
    int c = c0;
    for (; c + 8 <= c1; c += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(c + u) * cstride + off];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; c < c1; ++c) acc += part[(size_t)c * cstride + off];
    return acc;
  };
  (void)t0;
  (void)t1;
  if (e < s1) {
    const int i = (int)(e / H1);
    if (block64_live(a.k_live, i & ~63)) s = sum_chunks(bf.part1, s1, e);
  } else if (e >= s2 && e < s3) {
    s = sum_chunks(bf.part2, (size_t)H1 * H2, e - s2);
  } else if (e >= s4 && e < s5) {
    s = sum_chunks(bf.part3, (size_t)H2 * H3, e - s4);
  } else {
    return;  // biases and w4: k_train_reduce_tiles
  }
  grad[(size_t)net * gstride + e] = s;
}

// Biases and w4 of each network: one workgroup per value sums the network's tile partials
// (thread t: tiles t, t+256, ..., then a fixed tree).
__global__ void __launch_bounds__(256) k_train_reduce_tiles(const TrainArgs a, const TrainBufs bf,
                                                            float* __restrict__ grad, size_t gstride) {
  __shared__ float red[256];
  const int v = blockIdx.x, net = blockIdx.y, t = threadIdx.x;
  const int t0 = a.tiles[net], t1 = a.tiles[net + 1];
  float s = 0.0f;
  for (int k = t0 + t; k < t1; k += 256) s += bf.tile_part[(size_t)k * TP_STRIDE + v];
  red[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t != 0) return;
  const size_t s1 = (size_t)a.k_in * H1, s3 = s1 + H1 + (size_t)H1 * H2, s5 = s3 + H2 + (size_t)H2 * H3;
  size_t e;
  if (v < TP_GB2) e = s1 + (v - TP_GB1);
  else if (v < TP_GB3) e = s3 + (v - TP_GB2);
  else if (v < TP_GW4) e = s5 + (v - TP_GB3);
  else if (v < TP_GB4) e = s5 + H3 + (v - TP_GW4);
  else e = s5 + 2 * H3;
  grad[(size_t)net * gstride + e] = red[0];
}

// Residuals of the loss, per atom: forces, masked force residuals, energy adjoints and the
// workgroup's FP64 sum of squared force residuals (fixed tree).
__global__ void __launch_bounds__(256)
k_train_residuals(const double* __restrict__ energy, const double* __restrict__ e_ref,
                  const float* __restrict__ grad_self, const long long* __restrict__ grad_nbr, float force_scale,
                  const float* __restrict__ f_ref, const int32_t* __restrict__ species_idx,
                  const int32_t* __restrict__ batch_index, int n_atoms, double ce_scale, float* __restrict__ forces,
                  float* __restrict__ r_forces, float* __restrict__ c_energy, double* __restrict__ partials) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double sq = 0.0;
  if (i < n_atoms) {
    const bool real = species_idx[i] >= 0;
    const int b = batch_index[i];
    c_energy[i] = (float)(ce_scale * (energy[b] - e_ref[b]));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float f = (float)(-((double)grad_self[3 * i + k] + fnl_from_fixed(grad_nbr[3 * i + k])) * (double)force_scale);
      const float r = real ? f - f_ref[3 * i + k] : 0.0f;
      forces[3 * i + k] = f;
      r_forces[3 * i + k] = r;
      sq += (double)r * (double)r;
    }
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// loss = w_e sum_b rE_b^2 / n_e + w_f sum rF^2 / n_f (fixed order)
__global__ void __launch_bounds__(256)
k_train_loss(const double* __restrict__ energy, const double* __restrict__ e_ref, int n_sys,
             const double* __restrict__ partials, int n_part, double w_e, double w_f, double n_e, double n_f,
             double* __restrict__ loss) {
  __shared__ double re[256], rf[256];
  const int t = threadIdx.x;
  double se = 0.0, sf = 0.0;
  for (int b = t; b < n_sys; b += 256) {
    const double r = energy[b] - e_ref[b];
    se += r * r;
  }
  for (int k = t; k < n_part; k += 256) sf += partials[k];
  re[t] = se;
  rf[t] = sf;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      re[t] += re[t + o];
      rf[t] += rf[t + o];
    }
    __syncthreads();
  }
  if (t == 0) loss[0] = w_e * re[0] / n_e + w_f * rf[0] / n_f;
}

struct WsPlan {
  size_t rows_f, tiles, chunks;  // capacities
  size_t off[16];
  size_t bytes;
};

WsPlan plan(int n_rows, int n_nets, int k_in) {
  WsPlan p{};
  p.rows_f = (size_t)n_rows;
  p.tiles = (size_t)n_rows / TM + n_nets + 1;
  p.chunks = (size_t)n_rows / RC + n_nets + 1;
  const size_t sizes[] = {
      p.rows_f * H1, p.rows_f * H1, p.rows_f * H1, p.rows_f * H1,  // z1 dz1 h1 dh1
      p.rows_f * H2, p.rows_f * H2, p.rows_f * H2, p.rows_f * H2,  // z2 dz2 h2 dh2
      p.rows_f * H3, p.rows_f * H3,                                // bz3 bdz3
      p.tiles * TP_STRIDE,                                         // tile partials
      p.chunks * (size_t)k_in * H1, p.chunks * (size_t)H1 * H2, p.chunks * (size_t)H2 * H3};
  size_t o = 0;
  for (int i = 0; i < 14; ++i) {
    p.off[i] = o;
    o += fnl_align256(sizes[i] * sizeof(float));
  }
  p.bytes = o;
  return p;
}

}  // namespace

extern "C" size_t fnl_train_workspace(int32_t n_rows, int32_t n_nets, int32_t k_in) {
  if (n_rows < 0 || n_nets < 1 || k_in < 1) return 0;
  return plan(n_rows, n_nets, k_in).bytes;
}

extern "C" size_t fnl_train_grad_stride(int32_t k_in) {
  return (size_t)k_in * H1 + H1 + (size_t)H1 * H2 + H2 + (size_t)H2 * H3 + H3 + H3 + 1;
}

extern "C" int fnl_train_mlp_grad(const float* aev, int32_t aev_stride, const float* daev, int32_t daev_stride,
                                  int32_t k_in, const fnl_train_net* nets, int32_t n_nets, const int32_t* net_rows,
                                  const int32_t* perm, const float* c_energy, float c_force, float celu_alpha,
                                  uint64_t k_live, void* workspace, size_t workspace_bytes, float* grad,
                                  fnl_stream_t stream) {
  if (!aev || !daev || !nets || !net_rows || !perm || !c_energy || !workspace || !grad) return FNL_ERR_INVALID;
  if (n_nets < 1 || n_nets > FNL_MLP_MAX_NETS || k_in < 16 || k_in % 16 != 0 || aev_stride < k_in ||
      daev_stride < k_in || aev_stride % 4 != 0 || daev_stride % 4 != 0 || !(celu_alpha > 0.0f))
    return FNL_ERR_INVALID;
  TrainArgs a{};
  a.n_nets = n_nets;
  a.k_in = k_in;
  a.k_live = k_live ? (unsigned long long)k_live : ~0ull;
  a.alpha = celu_alpha;
  a.c_force = c_force;
  a.rows[0] = a.tiles[0] = a.chunks[0] = 0;
  if (net_rows[0] != 0) return FNL_ERR_INVALID;
  for (int n = 0; n < n_nets; ++n) {
    const int cnt = net_rows[n + 1] - net_rows[n];
    if (cnt < 0) return FNL_ERR_INVALID;
    const fnl_train_net& s = nets[n];
    if (!s.w1 || !s.b1 || !s.w2 || !s.b2 || !s.w3 || !s.b3 || !s.w4 || !s.b4) return FNL_ERR_INVALID;
    a.net[n] = s;
    a.rows[n + 1] = net_rows[n + 1];
    a.tiles[n + 1] = a.tiles[n] + (cnt + TM - 1) / TM;
    a.chunks[n + 1] = a.chunks[n] + (cnt + RC - 1) / RC;
  }
  const int n_rows = a.rows[n_nets];
  const WsPlan p = plan(n_rows, n_nets, k_in);
  if (workspace_bytes < p.bytes) return FNL_ERR_INVALID;
  unsigned char* ws = (unsigned char*)workspace;
  auto fp = [&](int i) { return (float*)(ws + p.off[i]); };
  TrainBufs b{};
  b.aev = aev;
  b.daev = daev;
  b.aev_ld = aev_stride;
  b.daev_ld = daev_stride;
  b.perm = perm;
  b.c_energy = c_energy;
  b.z1 = fp(0); b.dz1 = fp(1); b.h1 = fp(2); b.dh1 = fp(3);
  b.z2 = fp(4); b.dz2 = fp(5); b.h2 = fp(6); b.dh2 = fp(7);
  b.bz3 = fp(8); b.bdz3 = fp(9);
  b.tile_part = fp(10);
  b.part1 = fp(11); b.part2 = fp(12); b.part3 = fp(13);
  hipStream_t st = (hipStream_t)stream;
  const size_t gstride = fnl_train_grad_stride(k_in);
  const int T = a.tiles[n_nets], C = a.chunks[n_nets];
  if (T > 0) {
    hipLaunchKernelGGL(k_train_rows<FWD1>, dim3(T, H1 / TN), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_rows<FWD2>, dim3(T, (H2 + TN - 1) / TN), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_rows<FWD3>, dim3(T, (H3 + TN - 1) / TN), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_rows<BWD2>, dim3(T, (H2 + TN - 1) / TN), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_rows<BWD1>, dim3(T, H1 / TN), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_wgrad<1>, dim3(C, (k_in + 63) / 64, H1 / 64), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_wgrad<2>, dim3(C, H1 / 64, (H2 + 63) / 64), dim3(256), 0, st, a, b);
    hipLaunchKernelGGL(k_train_wgrad<3>, dim3(C, (H2 + 63) / 64, (H3 + 63) / 64), dim3(256), 0, st, a, b);
  }
  hipLaunchKernelGGL(k_train_reduce, dim3((unsigned)((gstride + 255) / 256), n_nets), dim3(256), 0, st, a, b, grad,
                     gstride);
  hipLaunchKernelGGL(k_train_reduce_tiles, dim3(TP_GB4 + 1, n_nets), dim3(256), 0, st, a, b, grad, gstride);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}

extern "C" int fnl_train_residuals(const double* energy, const double* e_ref, int32_t n_sys, const float* grad_self,
                                   const int64_t* grad_nbr, float force_scale, const float* f_ref,
                                   const int32_t* species_idx, const int32_t* batch_index, int32_t n_atoms,
                                   double w_e, double w_f, double n_e, double n_f, float* forces, float* r_forces,
                                   float* c_energy, double* loss, double* partials, fnl_stream_t stream) {
  if (n_sys < 1 || n_atoms < 0 || !(n_e > 0.0) || !(n_f > 0.0)) return FNL_ERR_INVALID;
  if (!energy || !e_ref || !loss || !partials) return FNL_ERR_INVALID;
  if (n_atoms > 0 && (!grad_self || !grad_nbr || !f_ref || !species_idx || !batch_index || !forces || !r_forces ||
                      !c_energy))
    return FNL_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const int nb = fnl_blocks(n_atoms, 256);
  if (nb > 0)
    hipLaunchKernelGGL(k_train_residuals, dim3(nb), dim3(256), 0, st, energy, e_ref, grad_self,
                       (const long long*)grad_nbr, force_scale, f_ref, species_idx, batch_index, n_atoms,
                       2.0 * w_e / n_e, forces, r_forces, c_energy, partials);
  hipLaunchKernelGGL(k_train_loss, dim3(1), dim3(256), 0, st, energy, e_ref, n_sys, partials, nb, w_e, w_f, n_e, n_f,
                     loss);
  FNL_LAUNCH_CHECK();
  return FNL_OK;
}
