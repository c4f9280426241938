/*
 * fnl_hip.h -- C ABI of the MI355X (gfx950) kernels for FeNNol's ANI
 * energy / force / MD path.
 *
 * Conventions
 *   - Every buffer is device memory owned by the caller; the library never
 *     allocates.  Temporaries live in a caller-provided workspace whose size is
 *     returned by the matching *_workspace() call.  The only process-wide state
 *     is a mutex-protected record of the kernels whose dynamic-LDS limit has
 *     been raised on each device, so every entry point may be called from
 *     several host threads at once (on different buffers).
 *   - All work is enqueued on the caller's stream; no call synchronises.
 *   - Return value: 0 on success, <0 on invalid arguments or a HIP launch error
 *     (fnl_strerror() names it).  Capacity overflows are reported on the device
 *     in the caller's status buffer, read by the host only when it chooses to
 *     (the reference reads its overflow flag once per rebuild:
 *     models/preprocessing.py:357-367, :1859-1878).
 *   - Atoms of one system are contiguous; sys_offsets[b] .. sys_offsets[b+1].
 *     Atoms with species index < 0 are padding: they get no edges, zero AEV,
 *     zero energy and zero force.
 *   - Cells are row-major [3][3] with rows = lattice vectors; reciprocal cells
 *     are their inverses, so fractional coordinates are x . R.
 *
 * Reference citations are relative to FeNNol's src/fennol.
 */
#ifndef FNL_HIP_H
#define FNL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* fnl_stream_t; /* == hipStream_t */

#define FNL_OK 0
#define FNL_ERR_INVALID (-1)
#define FNL_ERR_HIP (-2)
#define FNL_ERR_UNSUPPORTED (-3)

/* periodic boundary modes */
#define FNL_PBC_NONE 0
#define FNL_PBC_MINIMUM_IMAGE 1
/* every image within the cutoff (cells narrower than twice the cutoff); rows carry
 * a packed integer cell shift per slot (nbr_shift) */
#define FNL_PBC_GENERAL 2

/* packed cell shift of a general-PBC slot: (s0+512) | (s1+512)<<10 | (s2+512)<<20,
 * |s_a| <= 511; vec_ij = x_j - x_i + s . C */
#define FNL_SHIFT_PACK(s0, s1, s2) \
  ((int32_t)(((s0) + 512) | (((s1) + 512) << 10) | (((s2) + 512) << 20)))

/* fixed-point unit of the angular neighbour gradients (fnl_aev_backward's grad_nbr) */
#define FNL_GRAD_FIXED_ONE 4294967296.0

/* element types */
#define FNL_F32 0
#define FNL_BF16 1

/* status words written by fnl_nblist_build (int32[FNL_STATUS_WORDS]) */
#define FNL_STATUS_OVERFLOW 0       /* bit 0: a row capacity was exceeded; bit 1:
                                       grid capacities inconsistent; bit 2: a
                                       general-PBC shift beyond +-511 cells        */
#define FNL_STATUS_MAX_NBR 1        /* largest radial count among rows that overflowed
                                       (0 if none): sizes the regrown rows          */
#define FNL_STATUS_MAX_NBR_ANG 2    /* the same for the angular rows                */
#define FNL_STATUS_MAX_CELL 3       /* largest cell population seen                 */
#define FNL_STATUS_STICKY 4         /* overflow flags OR-ed over every build since the
                                       caller last cleared this word                */
#define FNL_STATUS_STICKY_MAX_NBR 5     /* largest radial count of an overflowing row
                                           over every build since the caller last
                                           cleared words 4-6 (sizes the rows after a
                                           rollback to a checkpoint)                 */
#define FNL_STATUS_STICKY_MAX_NBR_ANG 6 /* the same for the angular rows              */
#define FNL_STATUS_WORDS 8

int fnl_version(void);
const char* fnl_strerror(int code);

/* ------------------------------------------------------------------------
 * Neighbour lists: cell-list radial graph + angular sub-graph.
 *
 * Replaces GraphGenerator.process (models/preprocessing.py:369-638, same pair
 * set as the host path :74-355), GraphFilter.process (:909-956) and the
 * neighbour-slot table of GraphAngularExtension.process (:1196-1276).
 *
 * Output is one row per atom ("ELL" layout):
 *   nbr[i*max_nbr + e]          neighbour atom j of centre i, e < n_nbr[i]
 *   nbr_ang[i*max_nbr_ang + a]  radial slot e whose d12 < cutoff_ang^2
 * A pair (i, j) is kept iff d12 < cutoff^2 with d12 computed as the reference
 * does in float32: vec = x_j - x_i; s = -rint(vec . R); vec += s . C;
 * d12 = (vx*vx + vy*vy) + vz*vz (no contraction).  The cutoffs are doubles
 * (the model's Python floats): the threshold is cutoff^2 squared in double and
 * rounded to float32, as the reference compares float32 d12 with rc**2.  The triplets of centre i are
 * all slot pairs a < b of its angular row (preprocessing.py:1237-1245).
 *
 * grid: int32[n_sys][4] = (cap_x, cap_y, cap_z, first_cell) -- the caller's
 * upper bound on the cell grid of each system and the offset of its cells in
 * the global cell arrays; n_cells = sum of cap_x*cap_y*cap_z.  The grid
 * actually used has floor(2 width / cutoff) cells per direction, clamped to
 * [1, cap]; width is the lattice-plane spacing (periodic) or the bounding box
 * extent (open).  Cells narrower than the cutoff are searched two deep
 * (5x5x5 neighbourhood, half the candidates of 3x3x3 cutoff-wide cells),
 * cells at least a cutoff wide one deep; any cap gives the same pair set.
 * ---------------------------------------------------------------------- */
size_t fnl_nblist_workspace(int32_t n_atoms, int32_t n_sys, int32_t n_cells);

int fnl_nblist_build(const float* coords, const int32_t* species_idx,
                     const int32_t* batch_index, const int32_t* sys_offsets,
                     int32_t n_atoms, int32_t n_sys, int32_t pbc_mode,
                     const float* cells, const float* reciprocal_cells,
                     const int32_t* grid, int32_t n_cells,
                     double cutoff, double cutoff_ang,
                     int32_t max_nbr, int32_t max_nbr_ang,
                     void* workspace, size_t workspace_bytes,
                     int32_t* nbr, int32_t* nbr_shift, int32_t* n_nbr,
                     int32_t* nbr_ang, int32_t* n_nbr_ang,
                     int32_t* status, fnl_stream_t stream);

/* nbr_ang and n_nbr_ang may both be NULL: radial rows only (the list built
 * with cutoff + skin for fnl_nblist_skin_update).
 *
 * pbc_mode FNL_PBC_GENERAL replaces the general-PBC branch of the reference
 * (preprocessing.py:144-257, pbc_shifts :288-293): atoms are wrapped into the
 * cell (at = -floor(x . R), xw = x + at . C), every atom j != i of the system
 * is tested at every shift s with |s_a| <= ceil(cutoff * |column a of R|)
 * with d12 of vec = (xw_j - xw_i) + s . C, and nbr_shift[i*max_nbr + e] holds
 * the packed total shift s + at_j - at_i.  Self images are not pairs (the
 * reference pairs i < j).  Cost O(atoms * shifts) per centre, meant for the
 * small cells that need it.  nbr_shift is ignored (may be NULL) otherwise. */

/* fnl_nblist_build without a workspace argument and without any allocation:
 * the cell-list temporaries live in a scratch array of the library
 * (FNL_NBLIST_SCRATCH_BYTES per GPU, held until the process ends).  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP-graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Arguments and results as
 * fnl_nblist_build.  Returns FNL_ERR_SCRATCH_TOO_SMALL, before any work, if
 * fnl_nblist_workspace(n_atoms, n_sys, n_cells) exceeds the scratch size.
 * Every call on one GPU shares that GPU's scratch (each call starts by
 * clearing the counters it needs at zero): calls that overlap on one GPU --
 * on different streams, or from several host threads -- race and must not be
 * made. */
#define FNL_NBLIST_SCRATCH_BYTES ((size_t)64 << 20)
#define FNL_ERR_SCRATCH_TOO_SMALL (-4)

int fnl_nblist_build_noalloc(const float* coords, const int32_t* species_idx,
                             const int32_t* batch_index, const int32_t* sys_offsets,
                             int32_t n_atoms, int32_t n_sys, int32_t pbc_mode,
                             const float* cells, const float* reciprocal_cells,
                             const int32_t* grid, int32_t n_cells,
                             double cutoff, double cutoff_ang,
                             int32_t max_nbr, int32_t max_nbr_ang,
                             int32_t* nbr, int32_t* nbr_shift, int32_t* n_nbr,
                             int32_t* nbr_ang, int32_t* n_nbr_ang,
                             int32_t* status, fnl_stream_t stream);

/* Skin update: the radial and angular rows at the current positions selected
 * from the rows of a list built by fnl_nblist_build with cutoff + skin
 * (nbr_skin / n_skin / max_skin, radial rows only).  Replaces
 * GraphGenerator.update_skin (models/preprocessing.py:640-696) with the
 * re-filter of GraphFilter.update_skin (:958-960) and the slot table of
 * GraphAngularExtension.update_skin (:1278-1280).  d12 is recomputed with the
 * full build's float32 arithmetic; rows keep the order of the skin rows.
 * Exact as long as no atom moved more than skin/2 since the build (the
 * reference's approximation); with the minimum-image convention, cutoff +
 * skin must stay below half of every lattice-plane spacing.  With
 * FNL_PBC_GENERAL the skin rows' shifts (skin_shift) are applied to the
 * current positions, vec = x_j - x_i + s . C (update_skin :648-659), and
 * copied to nbr_shift.  status: as for fnl_nblist_build (overflow also if a
 * skin row had overflowed). */
int fnl_nblist_skin_update(const float* coords, const int32_t* batch_index,
                           int32_t n_atoms, int32_t pbc_mode,
                           const float* cells, const float* reciprocal_cells,
                           const int32_t* nbr_skin, const int32_t* skin_shift,
                           const int32_t* n_skin, int32_t max_skin,
                           double cutoff, double cutoff_ang,
                           int32_t max_nbr, int32_t max_nbr_ang,
                           int32_t* nbr, int32_t* nbr_shift, int32_t* n_nbr,
                           int32_t* nbr_ang, int32_t* n_nbr_ang,
                           int32_t* status, fnl_stream_t stream);

/* Flat directed edge list in the reference's dictionary layout
 * (edge_src / edge_dst / d12 / pbc_shifts, preprocessing.py:328-341) from the
 * rows above: edges of centre i occupy [row_start[i], row_start[i] + n_nbr[i]);
 * row_start is an exclusive prefix sum of n_nbr computed by the caller. */
int fnl_nblist_export(const float* coords, const int32_t* batch_index,
                      int32_t n_atoms, int32_t pbc_mode,
                      const float* cells, const float* reciprocal_cells,
                      const int32_t* nbr, const int32_t* nbr_shift,
                      const int32_t* n_nbr, int32_t max_nbr,
                      const int32_t* row_start,
                      int32_t* edge_src, int32_t* edge_dst, float* d12,
                      float* pbc_shifts, fnl_stream_t stream);

/* ------------------------------------------------------------------------
 * ANI AEV (models/embeddings/ani.py:47-142, with GraphProcessor
 * preprocessing.py:699-777, GraphFilterProcessor :963-1043,
 * GraphAngleProcessor :1283-1317 and the cosine SwitchFunction
 * models/misc/misc.py:617-716 folded in).
 * ---------------------------------------------------------------------- */
#define FNL_MAX_RADIAL 32
#define FNL_MAX_ANG_DIST 16
#define FNL_MAX_ANG_SECT 16
#define FNL_MAX_SPECIES 8

typedef struct fnl_aev_params {
  int32_t n_species;          /* S <= 8                                    */
  int32_t n_radial;           /* radial_dist_divisions <= 16 (kernel tile) */
  int32_t n_ang_dist;         /* angular_dist_divisions                    */
  int32_t n_ang_sect;         /* angle_sections; n_ang_dist*n_ang_sect <= 32 */
  float cutoff;               /* radial graph cutoff                       */
  float cutoff_ang;           /* angular graph cutoff                      */
  float radial_eta;            /* >= 0                                      */
  float angular_eta;           /* >= 0                                      */
  float zeta;                  /* > 0                                       */
  float radial_shifts[FNL_MAX_RADIAL];   /* linspace(radial_start, rc, nR+1)[:-1]   */
  float ang_shifts[FNL_MAX_ANG_DIST];    /* linspace(angular_start, rca, nA+1)[:-1] */
  float ang_cos[FNL_MAX_ANG_SECT];       /* cos(theta_z)                            */
  float ang_sin[FNL_MAX_ANG_SECT];       /* sin(theta_z)                            */
} fnl_aev_params;

/* AEV length: S*nR + S(S+1)/2 * nA*nZ.  Row i of aev (stride aev_stride
 * elements, element type out_dtype) gets the channels of centre i in the
 * reference order; columns [length, aev_stride) are written as zero.
 * species_mask: bit s set for every AEV species index s present in the batch.
 * Channel blocks of species (radial) and species pairs (angular) outside the
 * mask are not written, nor is the padding unless the mask is ~0u: they are
 * zero for every atom of such a batch, and the caller keeps them zero (a
 * buffer zeroed once and only ever used for batches with species in the
 * mask).  ~0u writes every column.
 * Pair vectors: x_j - x_i, plus the minimum image (FNL_PBC_MINIMUM_IMAGE) or
 * the slot's stored shift . C (FNL_PBC_GENERAL: nbr_shift from
 * fnl_nblist_build, required; NULL for the other modes).  The same holds for
 * fnl_aev_backward and fnl_aev_tangent. */
int fnl_aev_forward(const float* coords, const int32_t* species_idx,
                    const int32_t* batch_index, int32_t n_atoms, int32_t pbc_mode,
                    const float* cells, const float* reciprocal_cells,
                    const int32_t* nbr, const int32_t* n_nbr, int32_t max_nbr,
                    const int32_t* nbr_shift,
                    const int32_t* nbr_ang, const int32_t* n_nbr_ang, int32_t max_nbr_ang,
                    const fnl_aev_params* params,
                    void* aev, int32_t aev_stride, int32_t out_dtype,
                    uint32_t species_mask, fnl_stream_t stream);

/* Adjoint of fnl_aev_forward (the reverse pass jax.grad runs through
 * ani.py:47-142 and the graph processors).  AEV terms are recomputed rather
 * than stored.  Given g = dE/dAEV (rows of stride g_stride, g_dtype FNL_F32 or
 * FNL_BF16, read as float32), with
 * E = sum_i E_i and vec_ij = x_j - x_i:
 *   grad_self[i] = -sum_j dE_i/dvec_ij - sum_j dE_j/dvec_ji|radial
 *                  (written, float32[N][3]: the centre's own terms plus the
 *                  radial terms its neighbours' rows give it, gathered)
 *   grad_nbr[j] += dE_i/dvec_ij|angular for every angular edge, in fixed
 *                  point: int64[N][3] in units of 2^-32 (FNL_GRAD_FIXED_ONE per
 *                  unit of the energy per length); each term is rounded to
 *                  that grid and the integer sums are exact, so the result is
 *                  the same bits whatever order the device-scope atomics land
 *                  in; the caller zeroes it first
 *   virial_atom[i] = sum_j vec_ij (x) dE_i/dvec_ij   (float32[N][9], row-major a,b)
 * so that dE/dx = grad_self + grad_nbr (only the sum is meaningful) and the
 * per-system virial dE/dstrain (fennix.py:320-347) is the per-system sum of
 * virial_atom.  The neighbour lists must be symmetric (as built).
 * species_mask: as for fnl_aev_forward (every species present in the batch);
 * only the channel blocks of those species and of their pairs are read from
 * g, the others are not used by any atom of the batch.  ~0u reads every
 * block. */
int fnl_aev_backward(const float* coords, const int32_t* species_idx,
                     const int32_t* batch_index, int32_t n_atoms, int32_t pbc_mode,
                     const float* cells, const float* reciprocal_cells,
                     const int32_t* nbr, const int32_t* n_nbr, int32_t max_nbr,
                     const int32_t* nbr_shift,
                     const int32_t* nbr_ang, const int32_t* n_nbr_ang, int32_t max_nbr_ang,
                     const fnl_aev_params* params,
                     const void* g, int32_t g_stride, int32_t g_dtype,
                     float* grad_self, int64_t* grad_nbr, float* virial_atom,
                     uint32_t species_mask, fnl_stream_t stream);

/* Tangent (forward-mode derivative) of the AEV along a displacement field
 * disp (float32[N][3]):  daev_i = sum_e dAEV_i/dvec_e . (disp_j - disp_i).
 * With disp = F - F_ref it gives the input tangent whose contraction with
 * dE/dAEV is F . disp -- the term a force loss differentiates with respect to
 * the weights (the reference obtains it by jax.grad through jax.grad,
 * training/utils.py:508 over fennix.py:260-272).  Rows float32, stride
 * daev_stride, columns past the AEV length written as zero. */
int fnl_aev_tangent(const float* coords, const float* disp, const int32_t* species_idx,
                    const int32_t* batch_index, int32_t n_atoms, int32_t pbc_mode,
                    const float* cells, const float* reciprocal_cells,
                    const int32_t* nbr, const int32_t* n_nbr, int32_t max_nbr,
                    const int32_t* nbr_shift,
                    const int32_t* nbr_ang, const int32_t* n_nbr_ang, int32_t max_nbr_ang,
                    const fnl_aev_params* params, float* daev, int32_t daev_stride,
                    fnl_stream_t stream);

/* ------------------------------------------------------------------------
 * Species MLPs as grouped GEMMs on MFMA (SpeciesIndexNet models/misc/nets.py:
 * 301-426 over FullyConnectedNet :13-102, activation CELU(alpha)).
 *
 * Network shape: K -> H1 -> H2 -> H3 -> 1 with H1 = 256, H2 = 192, H3 = 160
 * (the ANI-2x widths, examples/training/ani.yaml:78).  Weights per network
 * (element type = compute dtype), K padded to k_pad (multiple of 64, zero
 * rows):
 *   w1t [H1][k_pad]  w2t [H2][H1]  w3t [H3][H2]   forward, K-contiguous
 *   w1  [k_pad][H1]  w2  [H1][H2]  w3  [H2][H3]   backward, N-contiguous
 *   w4 [H3], b1 [H1], b2 [H2], b3 [H3], b4 [1]   float32
 * FP32 matrices are row-major as shown.  BF16 matrices ([R][C] as shown) are
 * in MFMA fragment order: blocks of 16 rows x 32 columns, block (r/16, c/32)
 * at element offset ((r/16)*(C/32) + c/32)*512, element (r, c) inside it at
 * ((c%32)/8)*128 + (r%16)*8 + c%8 (k_pad a multiple of 128 for BF16).
 * Atoms are processed in tiles of rows of one network: tile t covers
 * perm[tile_start[t] .. tile_start[t] + tile_rows[t]) with network
 * tile_net[t]; rows per tile <= the kernel's tile size (tile_rows).
 * Outputs: energy[perm[r]] and, if grad != NULL, dE/dAEV rows grad[perm[r]]
 * (grad_dtype FNL_F32, or FNL_BF16 with BF16 weights: halves the bytes the
 * rows cost to write and to read back in fnl_aev_backward; stride
 * grad_stride >= k_pad, a multiple of 8; columns >= k_pad untouched).
 * ---------------------------------------------------------------------- */
#define FNL_MLP_H1 256
#define FNL_MLP_H2 192
#define FNL_MLP_H3 160
#define FNL_MLP_MAX_NETS 8

typedef struct fnl_mlp_net {
  const void* w1t; const void* w2t; const void* w3t;
  const void* w1;  const void* w2;  const void* w3;
  const float* w4; const float* b1; const float* b2; const float* b3; const float* b4;
  /* compensated BF16 (optional, BF16 only): the rounding remainders
   * bf16(W - bf16(W)) of the matrices above, same layout; the forward set
   * (w1t_lo, w2t_lo, w3t_lo) alone, or all six.  A product whose remainder is
   * given uses both parts on the MFMA (twice the matrix work), so its weights
   * keep ~16 significant bits: rounding W to 8 bits shifts every atom of a
   * similar environment the same way, a bias that grows with the energy scale
   * (measured 1.7 meV/atom on the 23,328-atom water box after 1,000 steps with
   * plain BF16 weights, 0.007 with the forward remainders; the backward ones
   * change the force error by a few per cent) */
  const void* w1t_lo; const void* w2t_lo; const void* w3t_lo;
  const void* w1_lo;  const void* w2_lo;  const void* w3_lo;
} fnl_mlp_net;

typedef struct fnl_mlp_weights {
  int32_t n_nets;
  int32_t k_pad;              /* multiple of 64 (FP32) / 128 (BF16), at most 128 such chunks */
  int32_t dtype;              /* FNL_F32 or FNL_BF16 (weights and GEMM inputs) */
  float celu_alpha;
  int32_t tile_rows;          /* atoms per tile: 0 = fnl_mlp_tile_rows(dtype);
                                 FP32: 16 or 32; BF16: 32, 64 or 96 (96:
                                 one 120 KB workgroup per CU, the default) */
  fnl_mlp_net nets[FNL_MLP_MAX_NETS];
} fnl_mlp_weights;

/* default atoms per tile for a dtype */
int fnl_mlp_tile_rows(int32_t dtype);

int fnl_species_mlp(const void* aev, int32_t aev_stride, const fnl_mlp_weights* w,
                    const int32_t* perm, const int32_t* tile_start,
                    const int32_t* tile_rows, const int32_t* tile_net, int32_t n_tiles,
                    float* energy, void* grad, int32_t grad_stride, int32_t grad_dtype,
                    uint64_t k_live, const int32_t* n_tiles_dev, fnl_stream_t stream);
/* n_tiles_dev: NULL, or a device word holding the number of tiles in use
 * (<= n_tiles, which then sizes the grid only): tables built on the device
 * (fnl_pool_layout) and launches a HIP graph replays with changing tiles. */
/* k_live: bit q set if input columns [32q, 32q + 32) may be non-zero for some atom
 * of the batch (0 = all).  The other columns must be zero in every AEV row: the
 * kernel skips their products (exact zeros) and writes dE/dAEV only in the live
 * 32-column pieces, leaving the others as they are (fnl_aev_backward with the
 * matching species_mask never reads them).
 * Columns at and beyond 2048 are always treated as live. */

/* ------------------------------------------------------------------------
 * Per-system reductions in FP64 (total_energy segment_sum, fennix.py:240-246;
 * virial tensor, fennix.py:320-347; kinetic tensor, md/integrate.py:320-339).
 * Deterministic two-stage tree in one launch: each system is cut into
 * segments of at most seg_atoms atoms; seg_offsets[b] = index of the first
 * segment of system b (exclusive prefix sum of ceil(natoms_b / seg_atoms)),
 * seg_offsets[n_sys] = n_seg; partials is caller scratch of 1 + n_seg * 19
 * doubles whose first word is a completion count: zero it once when
 * allocating, every call leaves it at zero (the last segment workgroup forms
 * the per-system sums).  Calls sharing one partials buffer must not overlap.
 * energy_atom may be NULL (energy_sys untouched), likewise virial_atom /
 * virial_sys, and vel (kinetic_sys untouched; mass is then ignored).
 * ---------------------------------------------------------------------- */
#define FNL_REDUCE_VALUES 19

int fnl_system_reduce(const int32_t* sys_offsets, const int32_t* seg_offsets,
                      int32_t n_sys, int32_t n_seg, int32_t seg_atoms,
                      const float* energy_atom, double* energy_sys,
                      const float* virial_atom, double* virial_sys,
                      const double* vel, const double* mass, double* kinetic_sys,
                      double* partials, fnl_stream_t stream);

/* forces = -(grad_self + grad_nbr * 2^-32) * scale  (float32[N][3], formed in
 * float64) */
int fnl_forces(const float* grad_self, const int64_t* grad_nbr, int32_t n_atoms,
               float scale, float* forces, fnl_stream_t stream);

/* ------------------------------------------------------------------------
 * BAOAB Langevin step (md/integrate.py:286-294 with the LGV O-step of
 * md/thermostats.py:108-120), FP64 positions and velocities:
 *   F  = -(grad_self + grad_nbr * 2^-32) * force_scale   (float64)
 *   v += dtm * F;  x += dt/2 v;  v = a1 v + a2 xi;  x += dt/2 v
 * xi ~ N(0,1) from Philox4x32-10 with counter (atom, step_lo, step_hi, 0) and
 * key (seed_lo, seed_hi), Box-Muller on (r0,r1), (r2,r3); or, if noise != NULL,
 * xi = noise (float64[N][3]).  a2 == NULL means no O-step (NVE).  If step_dev
 * is not NULL it points to two zero-initialised device words {step, done}:
 * the step number is read from the first instead of `step`, and the kernel's
 * last workgroup advances it by one (so that a captured graph replays with a
 * fresh counter; `done` is left at zero).  clear_grad_nbr != 0 zeroes the
 * grad_nbr rows once read, for an adjoint of the same MD step that
 * accumulates into them.  Also writes the float32 copy of x the model reads.
 * ---------------------------------------------------------------------- */
int fnl_baoab_step(double* x, double* v, float* x32,
                   const float* grad_self, int64_t* grad_nbr, float force_scale,
                   const double* dtm, double half_dt, double a1, const double* a2,
                   const double* noise, uint64_t seed, uint64_t step,
                   uint64_t* step_dev, int32_t n_atoms, int32_t clear_grad_nbr,
                   fnl_stream_t stream);

/* fnl_baoab_step and the per-system energy (and virial) sums of
 * fnl_system_reduce over the atom energies and virials of the last force
 * evaluation, in one launch (the sums of models/fennix.py:240-246 and
 * :320-347, the step of md/integrate.py:286-294): the segment workgroups and
 * the integrator's workgroups run side by side, as the two touch disjoint
 * memory.  For MD loops that read the sums of step k while integrating step
 * k+1.  Same arguments and results as the two calls; energy_atom and
 * energy_sys are required, virial_atom may be NULL (virial_sys untouched), no
 * kinetic sums.  At most FNL_REDUCE_FUSED_MAX_SYSTEMS systems
 * (FNL_ERR_UNSUPPORTED beyond: use the two calls), at least one atom and one
 * segment. */
#define FNL_REDUCE_FUSED_MAX_SYSTEMS 16

int fnl_baoab_step_sums(double* x, double* v, float* x32,
                        const float* grad_self, int64_t* grad_nbr, float force_scale,
                        const double* dtm, double half_dt, double a1, const double* a2,
                        const double* noise, uint64_t seed, uint64_t step,
                        uint64_t* step_dev, int32_t n_atoms, int32_t clear_grad_nbr,
                        const int32_t* sys_offsets, const int32_t* seg_offsets,
                        int32_t n_sys, int32_t n_seg, int32_t seg_atoms,
                        const float* energy_atom, double* energy_sys,
                        const float* virial_atom, double* virial_sys,
                        double* partials, fnl_stream_t stream);

/* Batched BAOAB for many independent systems (continuous-batching server):
 * atoms with dtm == 0 are left untouched (systems not integrated this step);
 * the O-step noise of atom i in system b uses counter (i - sys_offsets[b],
 * step_lo, step_hi, sys_key[b]) with step = sys_step[b] and key (seed_lo,
 * seed_hi), so it does not depend on the system's place in the batch. */
int fnl_baoab_step_batched(double* x, double* v, float* x32,
                           const float* grad_self, const int64_t* grad_nbr, float force_scale,
                           const double* dtm, double half_dt, double a1, const double* a2,
                           const int32_t* batch_index, const int32_t* sys_offsets,
                           const uint64_t* sys_step, const uint32_t* sys_key, uint64_t seed,
                           int32_t n_atoms, fnl_stream_t stream);

/* *counter += value on the device (single thread). */
int fnl_counter_add(uint64_t* counter, uint64_t value, fnl_stream_t stream);

/* Standard normals float64[N][3] from the same generator (initial velocities
 * use stream = 1; the O-step uses stream = 0). */
int fnl_philox_normals(double* out, int32_t n_atoms, uint64_t seed, uint64_t step,
                       uint32_t stream_id, fnl_stream_t stream);

/* ------------------------------------------------------------------------
 * Weight gradient of the energy + force loss through the species MLPs.
 *
 * Replaces the MLP part of jax.grad of the force-carrying loss in FeNNol's
 * train_step (training/utils.py:140-525, :508): for each network, with input
 * rows A (AEV), tangent rows dA = dAEV/dx . r (fnl_aev_tangent along
 * r = F - F_ref), per-atom energy adjoints c_energy = dL/dE_atom and the force
 * adjoint c_force = dL/dD (D = sum_i g_i . dA_i):
 *   forward   z_l = h_{l-1} W_l + b_l, dz_l = dh_{l-1} W_l,
 *             h_l = celu(z_l), dh_l = celu'(z_l) dz_l            (h_0 = A, dh_0 = dA)
 *   reverse   bh3 = c_energy w4, bdh3 = c_force w4,
 *             bdz_l = s_l bdh_l, bz_l = t_l dz_l bdh_l + s_l bh_l  (s = celu', t = celu'')
 *             bdh_{l-1} = bdz_l W_l^T, bh_{l-1} = bz_l W_l^T
 *   gradient  dW_l = dh_{l-1}^T bdz_l + h_{l-1}^T bz_l, db_l = sum bz_l,
 *             dw4 = sum (h3 c_energy + c_force dh3), db4 = sum c_energy
 * FP32 on MFMA; the reductions over atoms are split in chunks and summed in a
 * fixed order (the gradient is the same bits every run).
 *   weights   fnl_train_net per network, row-major W [in][out] (w1: k_in x 256)
 *   rows      perm lists the atoms grouped by network in network order;
 *             net_rows (host, [n_nets+1]) are the offsets of the groups
 *   aev/daev  rows in atom order (float32, strides >= k_in, multiples of 4)
 *   k_live    bit q: AEV columns [32q, 32q+32) may be non-zero (0 = all);
 *             the gradient rows of the others are zero
 *   grad      float32 [n_nets][fnl_train_grad_stride(k_in)]: dW1 [k_in][256],
 *             db1 [256], dW2 [256][192], db2 [192], dW3 [192][160], db3 [160],
 *             dw4 [160], db4 [1]
 *   workspace fnl_train_workspace(rows, n_nets, k_in) bytes
 * ---------------------------------------------------------------------- */
typedef struct fnl_train_net {
  const float* w1; const float* b1;
  const float* w2; const float* b2;
  const float* w3; const float* b3;
  const float* w4; const float* b4;
} fnl_train_net;

size_t fnl_train_workspace(int32_t n_rows, int32_t n_nets, int32_t k_in);
size_t fnl_train_grad_stride(int32_t k_in);
int fnl_train_mlp_grad(const float* aev, int32_t aev_stride, const float* daev, int32_t daev_stride,
                       int32_t k_in, const fnl_train_net* nets, int32_t n_nets,
                       const int32_t* net_rows, const int32_t* perm, const float* c_energy,
                       float c_force, float celu_alpha, uint64_t k_live, void* workspace,
                       size_t workspace_bytes, float* grad, fnl_stream_t stream);

/* Residuals of the energy + force loss (the mse terms of training/utils.py:
 * 329-343 with weights :496-501):
 *   F_i = -(grad_self + grad_nbr 2^-32) * force_scale       (float32, as fnl_forces)
 *   rF_i = (F_i - F_ref,i) for real atoms (species index >= 0), else 0
 *   c_energy_i = 2 w_e / n_e * (E_b - E_ref,b), b the system of atom i
 *   loss = w_e sum_b (E_b - E_ref,b)^2 / n_e + w_f sum rF^2 / n_f   (float64)
 * Every sum runs in a fixed order.  partials: float64[ceil(n_atoms/256)]. */
int fnl_train_residuals(const double* energy, const double* e_ref, int32_t n_sys,
                        const float* grad_self, const int64_t* grad_nbr, float force_scale,
                        const float* f_ref, const int32_t* species_idx, const int32_t* batch_index,
                        int32_t n_atoms, double w_e, double w_f, double n_e, double n_f,
                        float* forces, float* r_forces, float* c_energy, double* loss,
                        double* partials, fnl_stream_t stream);

/* ------------------------------------------------------------------------
 * Serving pool: continuous batching of independent open-boundary systems
 * (energy + forces requests and short Langevin runs) in one set of
 * capacity-static device buffers.  No FeNNol function corresponds: the
 * reference batches systems on the host (models/preprocessing.py:1763-1791,
 * analyze.py:108-200) and builds its per-species index on the host
 * (SpeciesIndexer, preprocessing.py:1583-1651).  Here the caller uploads one
 * fnl_pool_request per active request (and the atoms of new ones) when the set
 * changes; fnl_pool_layout builds every per-atom and per-system table on the
 * device, and one serving iteration is a fixed sequence of launches over
 * capacity-sized grids whose live sizes are read from device words (a HIP
 * graph can hold it):
 *   fnl_pool_integrate -> fnl_nblist_build -> fnl_aev_forward ->
 *   fnl_species_mlp (n_tiles_dev = counts + FNL_POOL_N_TILES) -> clear of
 *   grad_nbr -> fnl_aev_backward -> fnl_system_reduce -> fnl_pool_finish.
 * A request's results do not depend on the other requests or on its rows:
 * its neighbour rows, AEV rows, tree reductions (segments counted from its
 * first row) and noise (counter = its local atom index, its step, its key)
 * are its own.
 * ---------------------------------------------------------------------- */
typedef struct fnl_pool_request {
  int32_t n_atoms;      /* rows of the request (>= 1)                               */
  int32_t src_row;      /* first row of its atoms in the saved previous layout
                           (fnl_pool_save_rows), or -1 for a new request           */
  int32_t stage_row;    /* new request: first row of its atoms in stage_z/stage_x  */
  int32_t md_left;      /* Langevin steps still to do (0: energy + forces only)     */
  int32_t evaluated;    /* 1 once it has been evaluated (its forces are known)      */
  uint32_t noise_key;   /* Philox stream of its O-step noise                        */
  uint32_t vel_key;     /* new MD request: Philox stream of its initial velocities,
                           drawn at step 0                                          */
  uint32_t reserved;
  uint64_t step;        /* its O-step counter (Langevin steps done)                 */
} fnl_pool_request;

/* device words of fnl_pool.counts */
#define FNL_POOL_N_ATOMS 0   /* rows in use (the requests are packed from row 0)   */
#define FNL_POOL_N_SYS 1     /* active requests                                     */
#define FNL_POOL_N_TILES 2   /* MLP tiles of the layout                             */
#define FNL_POOL_ERROR 3     /* non-zero: the last layout was refused (sizes beyond
                                the capacities; the pool is then empty)             */
#define FNL_POOL_COUNTS 8
#define FNL_POOL_HEADER_WORDS 16  /* int32 words of fnl_pool_finish's record header */

typedef struct fnl_pool {
  int32_t max_atoms;    /* row capacity                                             */
  int32_t max_sys;      /* request capacity                                         */
  int32_t max_tiles;    /* MLP tile capacity: >= max_atoms / tile_rows + n_nets     */
  int32_t n_nets;       /* MLP networks (<= FNL_MLP_MAX_NETS)                       */
  int32_t tile_rows;    /* atoms per MLP tile                                       */
  int32_t seg_atoms;    /* atoms per reduction segment (fnl_system_reduce)          */
  int32_t n_z;          /* entries of the per-element tables (atomic numbers)       */
  int32_t stage_rows;   /* rows of stage_z / stage_x                                */
  uint64_t seed;        /* Philox key of the velocities and the O-step noise        */
  /* per-element tables [n_z], indexed by atomic number (0 = padding) */
  const int32_t* z_species;  /* AEV species index, -1 = padding                     */
  const int32_t* z_net;      /* MLP network, -1 = none (energy 0, ChemicalNet rule) */
  const double* z_dtm;       /* dt / m                                              */
  const double* z_a2;        /* sqrt((1 - a1^2) kT / m)                             */
  const double* z_vscale;    /* sqrt(kT / m)                                        */
  /* per-row state [max_atoms] */
  double* x; double* v; float* x32; int32_t* z;
  float* grad_self; int64_t* grad_nbr; float* e_atom; float* virial_atom;
  double* dtm; double* a2;
  /* the previous layout's rows, written by fnl_pool_save_rows [max_atoms] */
  double* old_x; double* old_v; float* old_grad_self; int64_t* old_grad_nbr; int32_t* old_z;
  /* layout tables */
  int32_t* batch_index; int32_t* species_idx; int32_t* perm;   /* [max_atoms]       */
  int32_t* sys_offsets; int32_t* seg_offsets;                  /* [max_sys + 1]     */
  int32_t* grid_cap;                                           /* [max_sys][4]      */
  int32_t* tile_start; int32_t* tile_rows_tab; int32_t* tile_net; /* [max_tiles]    */
  /* per-request serving state [max_sys] */
  int32_t* sys_left; int32_t* sys_eval; uint64_t* sys_step; uint32_t* sys_key;
  int32_t* sys_fin;     /* slot among this iteration's finished requests, or -1     */
  int32_t* sys_out_row; /* first result row of a finished request                   */
  int32_t* counts;      /* [FNL_POOL_COUNTS] device words                           */
} fnl_pool;

/* Copy rows [0, n_rows) of x, v, grad_self, grad_nbr and z to the old_* arrays
 * (before a layout that moves continuing requests). */
int fnl_pool_save_rows(const fnl_pool* pool, int32_t n_rows, fnl_stream_t stream);

/* Layout of n_req requests (req: device array, in pool order) packed from row
 * 0.  Builds sys_offsets, seg_offsets, grid_cap (open-boundary cell-grid
 * capacities k^3 >= n), batch_index, species_idx, the network-sorted perm and
 * tile tables (the order of pool_layout_arrays on the host), the per-request
 * state, and the rows: continuing requests from the old_* arrays, new ones
 * from stage_z / stage_x (device) with velocities v = N(0,1) * z_vscale from
 * (seed, step 0, vel_key, local atom) for MD requests; rows beyond the last
 * request are cleared to padding.  Rows and requests beyond the capacities
 * are refused on the device (FNL_POOL_ERROR, empty pool). */
int fnl_pool_layout(const fnl_pool* pool, const fnl_pool_request* req, int32_t n_req,
                    const int32_t* stage_z, const double* stage_x, fnl_stream_t stream);

/* BAOAB step (fnl_baoab_step_batched's arithmetic) of the rows of requests
 * with md_left > 0 that have been evaluated; grid over n_rows rows (>= the
 * rows in use; the rest exit). */
int fnl_pool_integrate(const fnl_pool* pool, float force_scale, double half_dt, double a1,
                       int32_t n_rows, fnl_stream_t stream);

/* End of an iteration: every evaluated request is marked evaluated, an
 * integrated one advances its step and md_left; requests whose md_left is 0
 * finish.  Their energies (energy_sys), final positions (x) and forces
 * (-(grad_self + grad_nbr 2^-32) * force_scale) are packed into `out`
 * (device, out_bytes):
 *   int32 header[FNL_POOL_HEADER_WORDS]: n_sys, n_fin, n_fin_atoms, n_atoms,
 *     status[0..3] (FNL_STATUS_OVERFLOW, _MAX_NBR, _MAX_NBR_ANG, _MAX_CELL of
 *     this iteration's build), error (non-zero: the record did not fit
 *     out_bytes, or the layout was refused), then zeros;
 *   int32 flags[n_sys] (bit 0: finished now; bit 1: a row of this request
 *     exceeded max_nbr / max_nbr_ang, its result is not valid), padded to 8 B;
 *   double energy[n_fin] (finished requests in pool order);
 *   double x[n_fin_atoms][3]; float forces[n_fin_atoms][3].
 * fnl_pool_out_bytes gives the size of that record. */
int fnl_pool_finish(const fnl_pool* pool, const double* energy_sys, const int32_t* n_nbr,
                    const int32_t* n_nbr_ang, int32_t max_nbr, int32_t max_nbr_ang,
                    const int32_t* status, float force_scale, int32_t n_rows, void* out,
                    size_t out_bytes, fnl_stream_t stream);
size_t fnl_pool_out_bytes(int32_t n_sys, int32_t n_fin, int32_t n_fin_atoms);

/* Device fill helpers (graph-capture safe replacements for memset). */
int fnl_fill_i32(int32_t* p, int32_t value, size_t count, fnl_stream_t stream);
int fnl_fill_f32(float* p, float value, size_t count, fnl_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* FNL_HIP_H */
