// kernels.hpp -- host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

namespace scl_amd {

// Operator view handed to the launchers: CSR with int32 indices.
// Device scratch of the merge-path SpMM: per wave, its (row, nonzero) start
// on the merge path and the partial sum of the row it leaves unfinished.
struct MergeScratch {
    int32_t* split = nullptr;     // 2 * (waves + 1): (row, nonzero) pairs
    int32_t* carry_row = nullptr; // waves: unfinished row, or -1
    float* carry_val = nullptr;   // waves * F: its partial sum
    int64_t waves = 0;            // capacity in waves
    int64_t floats = 0;           // capacity of carry_val
};

// Row-grouped form of an operator (the "grouped" SpMM shapes): rows are
// taken R at a time in visiting order; a group's entries are the union of its
// rows' column indices in increasing order, each with the mask of the rows
// holding it.  gval lists the values entry by entry (row slots in order), so
// every row still meets its own nonzeros in stored (sorted) order.
// Entries past the last stored one in the transposed pattern's device
// arrays (t_col: row 0, t_val: 0): the adjoint kernels load a chunk of
// indices and weights without clamping it to the row.
constexpr int kTransposePad = 8;

struct GroupedCsr {
    int R = 0;                    // rows per group (4 or 8); 0: not built
    int64_t n_groups = 0;
    int64_t n_entries = 0;
    int32_t* gptr = nullptr;      // n_groups + 1: first entry of each group
    uint32_t* gword = nullptr;    // n_entries + 32 (zero padding): column | mask << 24
    int32_t* gvptr = nullptr;     // n_groups + 1: first value of each group
    float* gval = nullptr;        // nnz + 128 (padded for two whole-wave loads)
    int32_t* gperm = nullptr;     // nnz: gval[i] = val[gperm[i]]
    int32_t* grow = nullptr;      // n_groups * R: the rows, -1 past the end
    uint64_t fingerprint = 0;     // pattern it was built for
    bool schedule = false;        // built over the locality order
};

// LDS-staged form of the forward operator: groups of kStagedRows
// consecutive rows (in the visiting order) with the sorted union of their
// columns.  A workgroup stages the union's X rows through LDS in rounds and
// every row of the group reads its neighbours there; each row keeps its own
// stored-order FMA chain, reading its values from the operator (any values
// with the same pattern work).  Needs strictly increasing columns per row.
constexpr int kStagedRows = 64;
struct StagedCsr {
    bool built = false;
    int64_t n_groups = 0;
    int64_t n_union = 0;
    int32_t* gptr = nullptr;   // n_groups + 1: first union column of each group
    int32_t* ucol = nullptr;   // n_union: each group's distinct columns, ascending
    int32_t* upos = nullptr;   // nnz: every stored entry's place in its group's union
    int32_t* grow = nullptr;   // n_groups * kStagedRows: the rows, -1 past the end
    uint64_t fingerprint = 0;  // pattern it was built for
};

// LDS-tiled form of the forward operator (the "tiled" SpMM shapes).  Rows
// are cut, in visiting order, into tiles whose column union fits the LDS of
// a workgroup; a workgroup stages the union's slab of X rows once per slab
// (lpr * 4 features: 16 or 32) with direct global-to-LDS loads, and each of
// its lane groups (lpr lanes, one row of the tile) reads its neighbours
// there in stored order.  Within a tile the rows run longest first, so the
// rows of a wave have similar lengths.  A tile's entries (one union slot
// and one value per stored nonzero, each row padded to a multiple of four)
// are copied to LDS once per workgroup and serve every slab it computes.
constexpr int kTiledRows = 128;   // SpmmShape::rows_per_group code of the form
constexpr int kTiledThreads = 512;
// LDS slots, union slots and padded entries of one tile: lpr 4 or 8 lanes
// per row, budget 1 or 2 workgroups per CU (160 or 80 KiB of LDS each).  The
// union stops one wave load (64 / lpr slots) short of the capacity so that
// the last slot stays zero (the padding entries read it).
constexpr int tiled_slot_capacity(int lpr, int budget) {
    return (lpr == 8 ? 960 : 1920) / budget;
}
constexpr int tiled_union_max(int lpr, int budget) {
    return tiled_slot_capacity(lpr, budget) - 64 / lpr;
}
constexpr int tiled_entries_max(int budget) { return 6400 / budget; }
struct TiledCsr {
    int lpr = 0;                  // 0: not built
    int budget = 0;
    int64_t n_tiles = 0;
    int64_t n_union = 0;
    int64_t n_slots = 0;          // padded entries
    int32_t* info = nullptr;      // (n_tiles + 1) x 4: union offset, union size,
                                  // entry offset, first row (sentinel last)
    int32_t* ucol = nullptr;      // n_union: each tile's columns, ascending
    int32_t* trow = nullptr;      // n: rows in tile order
    int32_t* trbeg = nullptr;     // n: first entry of the row (tile-local)
    int32_t* trlen = nullptr;     // n: stored entries of the row
    uint16_t* tslot = nullptr;    // n_slots: union slot of every entry, in
                                  // 16-byte LDS units (lpr units per slot)
    float* tval = nullptr;        // n_slots: its value (-0.0 in padding)
    int32_t* tperm = nullptr;     // n_slots: position in (col, val), -1 padding
    uint64_t fingerprint = 0;     // pattern it was built for
};
inline int tiled_index(int lpr, int budget) { return (lpr == 8 ? 2 : 0) + (budget == 2 ? 1 : 0); }

struct CsrView {
    int32_t n;            // rows of the applied operator
    int32_t n_sec;        // columns
    int64_t nnz;
    const int32_t* row_ptr;
    const int32_t* col;
    const float* val;
    const int32_t* sched;  // row visiting order or null
    const MergeScratch* merge = nullptr;  // for merge-path shapes
    const GroupedCsr* grouped = nullptr;  // [0]: R = 4, [1]: R = 8
    const StagedCsr* staged = nullptr;    // LDS-staged groups of kStagedRows rows
    const TiledCsr* tiled = nullptr;      // [tiled_index(lpr, budget)]
};

// Launch shape of the row-split SpMM (see kernels.hip).
struct SpmmShape {
    int vec = 4;        // floats per lane per access (1, 2 or 4)
    int group = 64;     // lanes sharing one row (power of two, 1..64)
    int unroll = 8;     // nonzeros in flight per lane group (4, 8 or 16)
    bool nontemporal = false;  // streaming (non-temporal) stores of Y
    bool schedule = true;      // visit rows in the operator's locality order
    bool xcd_rows = false;     // each XCD owns a row range (else a slab range)
    int xcd_block = 0;         // b > 0: each XCD owns a row range and runs it
                               // in blocks of 16 << b workgroup row blocks,
                               // every slab of a block before the next block
    bool scalar_rows = false;  // group 16/32: rows' indices on the scalar unit
    int wg_per_cu = 0;  // cap on resident workgroups per CU (0: none), set by
                        // reserving LDS the kernel does not use
    int merge = 0;      // 0: row-split; m > 0: merge path, 256 << (m - 1)
                        // path items (row ends + nonzeros) per wave
    int rows_per_group = 0;  // 0: one row per lane group; 4 or 8: grouped form;
                             // kStagedRows: LDS-staged form; kTiledRows:
                             // LDS-tiled form (group = lanes per row, 4 or 8;
                             // wg_per_cu = LDS budget, 1 or 2 workgroups per
                             // CU; unroll = slabs per workgroup)
};

// Flags of the C ABI's shape calls (scl_gpu_spmm_set_shape).
constexpr int kShapeNontemporal = 1;
constexpr int kShapeNoSchedule = 2;
constexpr int kShapeXcdRows = 4;
constexpr int kShapeScalarRows = 8;
// bits 8..11: workgroups-per-CU cap (0 = none, else 1..15)
constexpr int kShapeOccupancyShift = 8;
constexpr int kShapeOccupancyMask = 0xF;
// LDS reserved per workgroup so that at most cap workgroups fit on a CU.
int occupancy_lds_bytes(int wg_per_cu);
// bits 12..15: merge-path code (0 = row-split, 1..8)
constexpr int kShapeMergeShift = 12;
constexpr int kShapeMergeMask = 0xF;
constexpr int kMergeMax = 8;
// bits 16..17: grouped form (0 = none, 1 = 4 rows, 2 = 8 rows per group,
// 3 = LDS-staged groups of kStagedRows rows)
constexpr int kShapeGroupShift = 16;
constexpr int kShapeGroupMask = 0x3;
// bit 18: LDS-tiled form (see TiledCsr)
constexpr int kShapeTiled = 1 << 18;
// bits 20..23: XCD row blocks (0 = off, b = 1..8: 16 << b row blocks)
constexpr int kShapeXcdBlockShift = 20;
constexpr int kShapeXcdBlockMask = 0xF;
// Waves a merge-path launch with code m uses for n rows and nnz nonzeros.
int64_t merge_waves(int64_t n, int64_t nnz, int m);

// True when the shape can run on F features with these X / Y pointers.
bool spmm_shape_valid(const SpmmShape& s, int64_t F, const void* X, const void* Y);

// Picks vec/group for a feature count and the alignment of X and Y.
SpmmShape spmm_default_shape(int64_t F, const void* X, const void* Y);

// Candidate shapes the autotuner times for F features.
std::vector<SpmmShape> spmm_candidate_shapes(int64_t F, const void* X, const void* Y,
                                             bool has_schedule);

// One step Y = P X (Y[i,:] = sum over row i, stored order, FMA in fp32).
// layout bit 0 / bit 1: X / Y are slab-planar: slab s (features s * 64 * vec
// onward, ws = min(64 * vec, F - s * 64 * vec) of them) of all n rows is an
// n x ws row-major block starting at float s * 64 * vec * n; the same n * F
// floats as row-major, with every full slab's rows aligned to its width.
// Only the row-grouped shapes on square operators take a non-zero layout
// (spmm_layout_supported).
void launch_spmm_f32(const CsrView& P, const float* X, float* Y, int64_t F,
                     const SpmmShape& shape, hipStream_t s, int layout = 0);
bool spmm_layout_supported(const SpmmShape& shape, const CsrView& P);

// One fused backward step over the transposed pattern PT (rows = columns j
// of P, entries = source rows i in increasing order, values p_ij):
//   Gp[j, :] = sum_i p_ij Gs[i, :]               (fp32 FMA, stored order)
//   dPT[p]  (+)= <Gs[i, :], Xprev[j, :]>          (when Xprev and dPT are set)
// dPT is indexed by transposed position; accumulate adds onto its contents.
// shape: vec, group and schedule on/off of the launch (null or invalid for F
// and the buffers' alignment: the widest vec, one row per group, the
// operator's schedule); unroll is fixed at 8.
void launch_adjoint_edge(const CsrView& PT, const float* Gs, const float* Xprev,
                         float* Gp, float* dPT, bool accumulate, int64_t F,
                         hipStream_t s, const SpmmShape* shape = nullptr);

// Whether the fused adjoint kernel can run vec x group on F features with
// buffers at a, b, c (null pointers are not checked for alignment).
bool adjoint_shape_valid(const SpmmShape& sh, int64_t F, const void* a, const void* b,
                         const void* c);

// (vec, group, schedule on/off) the autotuner times for the fused adjoint
// kernel.
std::vector<SpmmShape> adjoint_candidate_shapes(int64_t F, const void* a, const void* b,
                                                const void* c, bool has_schedule);

// Edge gradient of one step over the forward pattern (CSR order):
// dP[k] (+)= <Gs[i, :], Xprev[j, :]> for every nonzero k = (i, j).
void launch_edge_grad(const CsrView& P, const float* Gs, const float* Xprev, float* dP,
                      bool accumulate, int64_t F, hipStream_t s);

// c = a - b elementwise (m floats).
void launch_sub_f32(const float* a, const float* b, float* c, int64_t m, hipStream_t s);

// dst[idx[p]] = src[p] and dst[p] = src[idx[p]] for p < m.
void launch_scatter_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                        hipStream_t s);
// F = 1 product through LDS-staged windows of 256 * unroll entries (unroll
// 4, 8 or 16); with out_row set, row r of P is stored at y[out_row[r]].
// Requires P.n_sec < 2^30.
void launch_spmv_staged(const CsrView& P, const int32_t* out_row, const float* x, float* y,
                        int unroll, hipStream_t s);
// dst[i, :] = src[idx[i], :] for the n rows of a row-major n x K matrix, or
// with scatter set dst[idx[i], :] = src[i, :] (idx a permutation).
void launch_rows_f32(const int32_t* idx, const float* src, float* dst, int64_t n, int64_t K,
                     bool scatter, hipStream_t s);
void launch_gather_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                       hipStream_t s);
// dst[p] = idx[p] >= 0 ? src[idx[p]] : -0.0 for p < m.
void launch_gather_pad_f32(const int32_t* idx, const float* src, float* dst, int64_t m,
                           hipStream_t s);

// Input / output storage of one reduced-precision SpMM step.
enum class LowpIo { F32_TO_BF16, BF16_TO_BF16, BF16_TO_F32, F32_TO_F32 };

// One step Y = P X with fp32 accumulation, X and Y stored as the io pair
// says (16-bit side: bf16, or fp16 when fp16 is set), P's values read from
// vals_bf16 when not null (else P.val, fp32).  A 16-bit input is multiplied
// by in_scale after widening, a 16-bit output by out_scale before rounding
// (powers of two, so both are exact).  F must be even.
void launch_spmm_lowp(const CsrView& P, const uint16_t* vals_bf16, LowpIo io,
                      const void* X, void* Y, int64_t F, bool use_sched, bool fp16,
                      float in_scale, float out_scale, hipStream_t s);
// Compensated bf16 storage of m values (lowp.hip): each value rounded once
// to 24 bits (15 mantissa bits), stored as hi = its upper 16 bits (a bf16)
// and lo = the next 8 (the mantissa bits bf16 drops); read back exactly as
// the fp32 word hi << 16 | lo << 8.
struct CompPlanes {
    uint16_t* hi = nullptr;  // m
    uint8_t* lo = nullptr;   // m
};
// dst = compensated(src), m values in two planes (P's values).
void launch_encode_comp(const float* src, const CompPlanes& dst, int64_t m, hipStream_t s);
// Packed compensated iterates: groups of 4 consecutive values in 12 bytes
// (the 4 hi halves, then the 4 lo bytes), 3 m / 4 uint32 words.
void launch_pack_comp(const float* src, uint32_t* dst, int64_t m, hipStream_t s);
// One step Y = P X with fp32 accumulation; X read from the packed Xc when set
// (else fp32 X), Y written to the packed Yc when set (else fp32 Y), P's
// values read from vals when set (else P.val).  n_features % 4 == 0.
void launch_spmm_comp(const CsrView& P, const CompPlanes* vals, const float* X,
                      const uint32_t* Xc, float* Y, uint32_t* Yc, int64_t F,
                      bool use_sched, hipStream_t s);
// partial[b] = max |x| over a grid-stride slice, for b < n_partial blocks.
void launch_absmax(const float* x, int64_t m, float* partial, int n_partial, hipStream_t s);
// dst = round16(src * scale) (bf16 or fp16, nearest even), m floats.
void launch_narrow_f32(const float* src, uint16_t* dst, int64_t m, bool fp16, float scale,
                       hipStream_t s);

void launch_to_bf16(const float* src, uint16_t* dst, int64_t m, hipStream_t s);
void launch_from_bf16(const uint16_t* src, float* dst, int64_t m, hipStream_t s);

// FP64 degree normalisation of values in place: deg[i] = sum_k a_ik (FP64,
// stored order); symmetric == 0: a_ik <- (float)(a_ik / deg[i]) (0 if
// deg <= 1e-15); symmetric != 0: a_ik <- (float)(a_ik * r_i * r_j),
// r = 1/sqrt(deg) (0 if deg <= 1e-15).  deg (n doubles) is written.
void launch_row_sums_f64(const int32_t* row_ptr, const float* val, int32_t n,
                         double* deg, hipStream_t s);
void launch_normalize_rows(const int32_t* row_ptr, const int32_t* col,
                           float* val, int32_t n, const double* deg,
                           int symmetric, hipStream_t s);

// FP64 stationarity residual r = sum_j | sum_{i->j} p_ij pi_i - pi_j |,
// pi = deg / sum(deg), evaluated through the transposed pattern (t_row_ptr,
// t_col = source row, t_val).  partial must hold 2 * 1024 doubles; the
// result is written to out (one double, device memory).
void launch_stationarity_residual(const int32_t* t_row_ptr, const int32_t* t_col,
                                  const float* t_val, int32_t n,
                                  const double* deg, double* partial,
                                  double* out, hipStream_t s);

// ---- vector-side consumers (walks.hip) ----------------------------------
// x = fma(a, x, b * t) elementwise.
void launch_blend(float* x, const float* t, int64_t n, float a, float b, hipStream_t s);
// y = fma(alpha, y, beta * r); *total = sum |y - x| (device scalar), and
// *done = 1 once *total < tol.  With *done set, y = x and total keeps the
// converged step's value.
void launch_rwr_update(float* y, const float* x, const float* r, int64_t n, float alpha,
                       float beta, float* part, float* total, float tol, float* done,
                       hipStream_t s);
// One DPT bookkeeping step; *max_unvisited = max(1 - vis) (device scalar),
// *done = 1 once it is below tol.  With *done set, nothing changes.
void launch_dpt_update(const float* pn, float* vis, float* hit, int64_t n, float it,
                       float* part, float* max_unvisited, float tol, float* done,
                       hipStream_t s);
// hit /= vis where vis > 0.5, pt = vis > 0.5 ? hit / max : 1.
void launch_dpt_finish(const float* vis, float* hit, float* pt, int64_t n, float* part,
                       float* max_time, hipStream_t s);
// x = value * e_at (x = 0 when at is outside [0, n)).
void launch_set_unit(float* x, int64_t n, int64_t at, float value, hipStream_t s);
// x[0, n) = value
void launch_fill(float* x, int64_t n, float value, hipStream_t s);
void launch_add(float* acc, const float* x, int64_t n, hipStream_t s);
void launch_scale(float* x, int64_t n, float a, hipStream_t s);
// Diffusion-map start vectors and modified Gram-Schmidt (one workgroup).
void launch_dmap_init(float* Q, int64_t n, int64_t K, hipStream_t s);
void launch_mgs(float* Q, float* w, int64_t n, int64_t K, hipStream_t s);
// The same over the whole GPU on a column-major K x n matrix (columns
// contiguous); part: gram_schmidt_part_floats(n, K) floats, S: 8 floats of
// device scratch.  reorth: second pass and unit-vector fallback (diffusion
// maps); otherwise one pass and FastRNG(v * rng_mult) directions
// (pseudotime).
void launch_gram_schmidt(float* C, float* w, int64_t n, int64_t K, float* part, float* S,
                         bool reorth, uint64_t rng_mult, hipStream_t s);
int64_t gram_schmidt_part_floats(int64_t n, int64_t K);
// Cholesky QR of a row-major n x K matrix in place (K <= kCholQrMaxK): the
// Gram matrix and the factor in FP64, M <- M R^-1 rounded once.  A pivot
// (squared norm of a column after projection) at or below kCholQrMinPivot
// sets *degenerate and leaves M unchanged; callers then redo the work with
// modified Gram-Schmidt, which keeps the reference's degenerate-column rules.
// work: cholqr_workspace_doubles(n, K) doubles.
constexpr int kCholQrMaxK = 22;
constexpr double kCholQrMinPivot = 1e-12;
void launch_cholqr(float* M, int64_t n, int64_t K, double* work, int* degenerate,
                   hipStream_t s);
int64_t cholqr_workspace_doubles(int64_t n, int64_t K);
// B (cols x rows) = A (rows x cols)^T, both row-major.
void launch_transpose_rect(const float* A, float* B, int64_t rows, int64_t cols,
                           hipStream_t s);
// A = I; B = A^T; pairwise row distances D of an n x n row-major matrix.
void launch_identity(float* A, int64_t n, hipStream_t s);
void launch_transpose_square(const float* A, float* B, int64_t n, hipStream_t s);
void launch_pairwise_rows(const float* A, float* D, int64_t n, hipStream_t s);

// ---- other callers of the SpMM (callers.hip) ----
// flag = 1 when some primary slot of the int64 pattern is not strictly
// increasing; false when there is nothing to check (flag untouched).
bool launch_find_unsorted(const int64_t* ip, const int64_t* ix, int64_t slots, int* flag,
                          hipStream_t s);
// Scatters the stored entries of `slots` primary slots into the zeroed dense
// n_cells x G matrix (cells-major input when is_csr, genes-major otherwise);
// `ordered` walks each slot in stored order on one lane.
void launch_densify(const int64_t* ip, const int64_t* ix, const float* v, int64_t slots,
                    int64_t n_cells, int64_t G, bool is_csr, bool ordered, float* out,
                    hipStream_t s);
// MAGIC weights a_ij / (FP32 row sum); rows summing below 1e-12 get zero
// weights and are appended to copy[] (count in *n_copy, cleared by the caller).
void launch_magic_weights(const int32_t* row_ptr, const float* val, int32_t n, float* w,
                          int32_t* copy, int* n_copy, hipStream_t s);
void launch_copy_rows(const int32_t* rows, int32_t m, const float* in, float* out, int64_t G,
                      hipStream_t s);
// Pseudotime distances to the root row of the n x K components, divided by
// their maximum when it is positive; part: 1024 floats.
void launch_pt_distance(const float* Q, int64_t n, int64_t K, int64_t root, float* pt,
                        float* part, hipStream_t s);
// Y = alpha A X + beta Y with zero weights skipped (one wave per row).
void launch_spmm_axpby(const CsrView& A, const float* X, float* Y, int64_t F, float alpha,
                       float beta, hipStream_t s);

// ---- construction of the operator (construct.hip) ----
// KernelParams of kernel.hpp:108-122 for one bandwidth and kernel kind (the
// C ABI's scl_kernel_type_t value).
struct KernelWeights {
    float inv_h, inv_h2, inv_2h2, h2;
    int kind;
};
KernelWeights kernel_weights(float bandwidth, int kind);
// kernel value of every stored distance
void launch_kernel_values(const float* d, int64_t nnz, const KernelWeights& p, float* out,
                          hipStream_t s);
// per-row sums of the kernel values in stored order (int64 row pointer)
void launch_kernel_row_sums(const int64_t* ip, const float* d, int64_t n,
                            const KernelWeights& p, float* sums, hipStream_t s);
// row_normalize_to_stochastic of n rows (int64 row pointer)
void launch_row_normalize(const int64_t* ip, const float* v, int64_t n, float* out,
                          hipStream_t s);
// dense (T + T^T)/2 into the zeroed n x n S, then row normalisation; serial:
// some row is not strictly increasing (one thread, stored order)
void launch_symmetrize(const int64_t* ip, const int64_t* ix, const float* v, int64_t n,
                       bool serial, float* S, hipStream_t s);
// dense velocity transition into the zeroed n x n S
void launch_from_velocity(const int64_t* ip, const int64_t* ix, const float* v, int64_t n,
                          float* S, hipStream_t s);
// after y = T^T pi: y /= vector_sum(y) (when above 1e-15) and diff[0] =
// max |y - pi|; scale: 1 float, part: 1024 floats
void launch_stationary_update(float* y, const float* pi, int64_t n, float* scale, float* part,
                              float* diff, hipStream_t s);

// ---- k-nearest neighbours (neighbors.hip) ----
// vectorize::sum_squared of each row's stored values (int64 row pointer)
void launch_sum_squared(const int64_t* ip, const float* v, int64_t n, float* norms,
                        hipStream_t s);
// padded shape of the dense sample matrix the scan reads
int64_t knn_rows_padded(int64_t n);
int64_t knn_features_padded(int64_t f);
// k nearest of every sample of the zero-padded dense X (n_pad x f_pad);
// ld / lj: n x k list scratch; out_*: n x k
void launch_knn(const float* X, const float* norms, int64_t n, int64_t f_pad, int64_t k,
                float* ld, int32_t* lj, float* out_dist, int64_t* out_idx, hipStream_t s);

}  // namespace scl_amd
