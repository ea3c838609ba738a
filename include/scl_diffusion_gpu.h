/*
 * scl_diffusion_gpu.h -- device-pointer extensions of the diffusion ABI.
 *
 * These entry points have no counterpart in the reference (which is CPU
 * only).  They exist so that callers whose data already lives in GPU memory
 * (a torch.autograd.Function, a benchmark, a pipeline stage) can run the same
 * kernels as scl_diffusion_diffuse_matrix without the host round trip the
 * drop-in's host-pointer signature forces (scl/binding/c_api/diffusion.h:29-35).
 *
 * Conventions:
 *   - every pointer named *_dev is device memory on the current HIP device;
 *   - `stream` is a hipStream_t (NULL = the legacy default stream); the call
 *     enqueues work on it and returns without waiting;
 *   - dense matrices are row-major, n_nodes x n_features, fp32, contiguous;
 *   - the operator is the handle's primary-slot CSR (a CSC handle applies the
 *     transpose), exactly as in scl_diffusion.h;
 *   - errors are reported through the same codes and thread-local message.
 */
#ifndef SCL_DIFFUSION_GPU_H
#define SCL_DIFFUSION_GPU_H

#include "scl_core.h"
#include "scl_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of usable HIP devices (0 when none; never an error). */
SCL_EXPORT scl_error_t scl_gpu_device_count(int* out);

/* Drops the handle's device copies; the next call re-uploads them.  Needed
 * after the caller edits the arrays of a handle made by scl_sparse_wrap. */
SCL_EXPORT scl_error_t scl_gpu_sparse_invalidate(scl_sparse_t matrix);

/* Uploads the handle's arrays to the current device now (otherwise done on
 * first use), so that later stream-ordered calls do no host work. */
SCL_EXPORT scl_error_t scl_gpu_sparse_prepare(scl_sparse_t matrix);

/* Y_dev = P^t X_dev.  X_dev is not modified; Y_dev must not overlap X_dev.
 * work_dev (n_nodes x n_features floats, overlapping neither) is required
 * when n_steps >= 2 and ignored otherwise; on return it holds an
 * intermediate iterate in an unspecified layout (wide operators keep the
 * iterates between the first and last step slab-planar).  n_steps >= 1. */
SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, void* stream);

/* Y_dev = P^t X_dev like scl_gpu_diffuse_matrix_device, for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * It allocates nothing: the intermediate iterate (n_steps >= 2) lives in a
 * fixed array of SCL_GPU_NOALLOC_SCRATCH_FLOATS floats (64 MiB) that the
 * library holds on every GPU it runs on until the process ends; larger
 * n_nodes * n_features with n_steps >= 2 is SCL_ERROR_INVALID_ARGUMENT.  The
 * operator must already be on the current device and unchanged since
 * (scl_gpu_sparse_prepare, a set_shape or an earlier device call), else
 * SCL_ERROR_INVALID_ARGUMENT.  All calls on one GPU share its array: calls
 * that overlap on one GPU (other streams or threads) race.  Results equal
 * scl_gpu_diffuse_matrix_device's bit for bit. */
#define SCL_GPU_NOALLOC_SCRATCH_FLOATS ((size_t)16 * 1024 * 1024)
SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_device_noalloc(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_index_t n_nodes, scl_index_t n_features, scl_index_t n_steps, void* stream);

/* One step Y_dev = op(P) X_dev with op(P) = P (transpose == 0) or P^T
 * (transpose != 0, applied through a cached transposed pattern). */
SCL_EXPORT scl_error_t scl_gpu_spmm_device(scl_sparse_t transition,
                                           const scl_real_t* X_dev,
                                           scl_real_t* Y_dev,
                                           scl_index_t n_nodes,
                                           scl_index_t n_features,
                                           int transpose, void* stream);

/* Like scl_diffusion_compute_transition_matrix (the values of `adjacency`
 * are normalised in place, FP64 on the GPU), and additionally returns the
 * FP64 degrees d_i = sum_j a_ij (degrees_out: host array of rows doubles, or
 * NULL) and the FP64 stationarity residual r = || P^T pi - pi ||_1 with
 * pi = d / sum(d) evaluated with the normalised fp32 values (residual_out:
 * host double, or NULL; a rectangular matrix with residual_out set is
 * SCL_ERROR_DIMENSION_MISMATCH).  r vanishes, up to rounding, for symmetric A
 * with symmetric == 0. */
SCL_EXPORT scl_error_t scl_gpu_compute_transition_matrix_ex(
    scl_sparse_t adjacency, int symmetric, double* degrees_out,
    double* residual_out);

/* Host-only (needs no GPU): the locality row order the device uses for
 * operators of >= 16384 rows -- greedy: each next row is the unplaced row
 * with the most pattern edges (either direction) to the last 512 placed
 * rows.  order_out receives a permutation of 0..primary-1. */
SCL_EXPORT scl_error_t scl_sparse_locality_order(scl_sparse_t matrix,
                                                 scl_index_t* order_out);

/* ---------------------------------------------------------------- shapes --
 * Launch shape of the row-split SpMM for one operator (transpose == 0) or its
 * transpose (transpose == 1), or of the fused adjoint + edge-gradient step of
 * the reverse pass (transpose == 2: vec, group, SCL_GPU_SHAPE_NO_SCHEDULE and
 * SCL_GPU_SHAPE_OCCUPANCY; unroll is 8 and the other flags are ignored), and
 * one feature count:
 *   vec     floats per lane per access: 1, 2 or 4 (must divide n_features)
 *   group   lanes sharing one row: a power of two in 1..64; one launch covers
 *           group * vec features of every row per workgroup column (a slab)
 *   unroll  nonzeros in flight per lane group: 4, 8 or 16
 *   flags   SCL_GPU_SHAPE_NONTEMPORAL: streaming stores of the output;
 *           SCL_GPU_SHAPE_NO_SCHEDULE: natural row order instead of the
 *           operator's locality order (operators of >= 16384 rows get one);
 *           SCL_GPU_SHAPE_XCD_ROWS: each of the 8 XCDs owns a contiguous range
 *           of rows and walks it slab by slab (default: each XCD owns a
 *           range of slabs);
 *           SCL_GPU_SHAPE_SCALAR_ROWS: with group 16 or 32, the 64/group
 *           rows of a wave keep their indices and values on the scalar unit
 *           (each lane selects its row's) instead of loading them per lane;
 *           SCL_GPU_SHAPE_OCCUPANCY(c), c in 1..15: at most c workgroups of
 *           256 lanes resident per CU (LDS is reserved to enforce it); fewer
 *           rows in flight per XCD keep more of their shared neighbours in L2;
 *           SCL_GPU_SHAPE_MERGE_PATH(m), m in 1..8 (group 64 only): merge-path
 *           SpMM -- the path of row ends and nonzeros is cut into pieces of
 *           256 << (m - 1) items, one per wave, so work is even whatever the
 *           row lengths (rows in natural order; the schedule flag is
 *           ignored).  A row cut into p pieces is summed as p stored-order
 *           chains added in order: p - 1 extra fp32 roundings, so such rows
 *           are not bit-identical to the row-split shapes.  Setting it
 *           allocates the per-wave scratch (calls on one operator with a
 *           merge-path shape must not overlap on two streams).  The
 *           autotuner never picks it;
 *           SCL_GPU_SHAPE_ROW_GROUPS(r), r = 4 or 8 (group 64, transpose 0):
 *           one wave computes r consecutive rows of the visiting order,
 *           walking the sorted union of their column indices and gathering
 *           each X row once for all the rows that hold it; every row keeps
 *           its sorted-order FMA chain (bit-identical); unroll is the
 *           union entries gathered per chunk (4 rows: 8 or 16, 8 rows: 4, 8
 *           or 16; larger values are clamped).  Needs strictly
 *           increasing column indices in every row and < 2^24 columns;
 *           setting it builds the grouped form on the host (one pass over
 *           the pattern), which follows later value changes
 *           SCL_GPU_SHAPE_STAGED_GROUPS (vec 4, group 64, transpose 0, no
 *           occupancy cap): a workgroup computes 64 consecutive rows of the
 *           visiting order for 256 features; the union of their columns is
 *           staged through LDS in rounds of 64 X rows, each gathered once
 *           for the whole group, and every row reads its neighbours there in
 *           sorted order (bit-identical).  Needs strictly increasing column
 *           indices in every row; setting it builds the staged pattern on
 *           the host (the values stay the operator's own);
 *           SCL_GPU_SHAPE_XCD_BLOCKS(b), b in 1..8 (row-split and grouped
 *           shapes): each XCD owns a contiguous range of rows and runs it in
 *           blocks of 16 << b workgroup row blocks, every slab of a block
 *           before the next block, so the X rows a block gathers stay in
 *           that XCD's L2 while all its slabs run;
 *           SCL_GPU_SHAPE_TILED (vec 4, transpose 0): LDS tiles -- rows of
 *           the visiting order are cut into tiles whose column union fits
 *           the LDS of one workgroup; per slab of group * 4 features
 *           (group = 4 or 8 lanes per row) the workgroup copies the union's
 *           X rows into LDS once, with direct global-to-LDS loads, and each
 *           row (one group of lanes) reads its neighbours there in stored
 *           order (bit-identical).  SCL_GPU_SHAPE_OCCUPANCY(b), b = 1 or 2,
 *           is the LDS budget (1: 160 KiB tiles, one workgroup per CU; 2: 80
 *           KiB tiles, two); unroll (4, 8 or 16) is the slabs one
 *           workgroup computes.  Needs strictly increasing column indices in
 *           every row, rows of at most 480-1920 nonzeros (budget and group)
 *           and columns * n_features < 2^32; setting it builds the tiled
 *           form on the host (one pass over the pattern plus a sort per
 *           tile), which follows later value changes
 * Any row-split SpMM shape gives bit-identical results; only the speed
 * differs.  The
 * adjoint shape leaves dX bit-identical; dP's per-nonzero sums over features
 * are added in a shape-dependent order (fp32 rounding differences only). */
#define SCL_GPU_SHAPE_NONTEMPORAL 1
#define SCL_GPU_SHAPE_NO_SCHEDULE 2
#define SCL_GPU_SHAPE_XCD_ROWS 4
#define SCL_GPU_SHAPE_SCALAR_ROWS 8
#define SCL_GPU_SHAPE_OCCUPANCY(c) (((c) & 0xF) << 8)
#define SCL_GPU_SHAPE_MERGE_PATH(m) (((m) & 0xF) << 12)
#define SCL_GPU_SHAPE_ROW_GROUPS(r) ((((r) / 4) & 0x3) << 16)
#define SCL_GPU_SHAPE_STAGED_GROUPS (3 << 16)
#define SCL_GPU_SHAPE_TILED (1 << 18)
#define SCL_GPU_SHAPE_XCD_BLOCKS(b) (((b) & 0xF) << 20)

SCL_EXPORT scl_error_t scl_gpu_spmm_set_shape(scl_sparse_t transition,
                                              scl_index_t n_features,
                                              int transpose, int vec, int group,
                                              int unroll, int flags);

/* The shape the next call would use with 16-byte aligned buffers. */
SCL_EXPORT scl_error_t scl_gpu_spmm_get_shape(scl_sparse_t transition,
                                              scl_index_t n_features,
                                              int transpose, int* vec,
                                              int* group, int* unroll,
                                              int* flags);

/* Times the candidate shapes for this operator and feature count on the
 * caller's buffers (Y_dev is overwritten), keeps the fastest for later calls
 * and returns its time per launch in best_ms (nullable).  Synchronises
 * `stream`.  transpose == 2 tunes the fused adjoint + edge-gradient step
 * (X_dev stands in for both of its inputs). */
SCL_EXPORT scl_error_t scl_gpu_spmm_autotune(scl_sparse_t transition,
                                             const scl_real_t* X_dev,
                                             scl_real_t* Y_dev,
                                             scl_index_t n_nodes,
                                             scl_index_t n_features,
                                             int transpose, void* stream,
                                             float* best_ms);

/* ------------------------------------------------- 16-bit iterate storage --
 * Y_dev = P^t X_dev (both fp32) with the t-1 intermediate iterates stored in
 * 16 bits and fp32 accumulation: the first step gathers the fp32 input, the
 * last writes fp32, every row result in between is rounded once (round to
 * nearest even).  flags:
 *   SCL_GPU_LOWP_FP16         iterates as fp16 scaled by 2^scale_exp (else bf16,
 *                             scale_exp must be 0)
 *   SCL_GPU_LOWP_BF16_VALUES  read P's values rounded to bf16 (else fp32)
 *   SCL_GPU_LOWP_NO_SCHEDULE  natural row order (the default follows the
 *                             fp32 SpMM shape tuned or set for n_features)
 *   SCL_GPU_LOWP_ROUND_INPUT  store X in 16 bits too (one rounding of
 *                             X * 2^scale_exp up front), so that the first
 *                             step also gathers 16-bit rows
 *   SCL_GPU_LOWP_COMPENSATED  compensated bf16 (not with FP16): every stored
 *                             value is rounded once to 24 bits (15 mantissa
 *                             bits, nearest even) and kept as a bf16 plane
 *                             (its upper 16 bits) plus an 8-bit plane with
 *                             the mantissa bits bf16 drops, read back
 *                             exactly (3 bytes per value); applies to the
 *                             iterates, to X with ROUND_INPUT and to P's
 *                             values with BF16_VALUES; n_features % 4 == 0
 * n_features must be even; work_dev (required when n_steps >= 2 or with
 * SCL_GPU_LOWP_ROUND_INPUT) holds 2 planes: scl_gpu_lowp_workspace gives its
 * size (n * F * 4 bytes for 16-bit storage, two 256-byte aligned planes of
 * n * F * 3 bytes compensated). */
#define SCL_GPU_LOWP_BF16_VALUES 1
#define SCL_GPU_LOWP_NO_SCHEDULE 2
#define SCL_GPU_LOWP_FP16 4
#define SCL_GPU_LOWP_ROUND_INPUT 8
#define SCL_GPU_LOWP_COMPENSATED 16
SCL_EXPORT scl_error_t scl_gpu_lowp_workspace(scl_index_t n_nodes, scl_index_t n_features,
                                              int flags, scl_size_t* bytes);
SCL_EXPORT scl_error_t scl_gpu_diffuse_matrix_lowp_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    void* work_dev, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t n_steps, int flags, int scale_exp, void* stream);

/* A safe fp16 scale: the largest k with 2^k * max_s max|X_s| <= 2^15 for the
 * intermediate iterates, bounded through the largest row |sum| of P
 * (|X_s| <= rho^s max|X|).  Reads X_dev on the device; synchronises
 * `stream`. */
SCL_EXPORT scl_error_t scl_gpu_lowp_scale_exp(scl_sparse_t transition,
                                              const scl_real_t* X_dev,
                                              scl_index_t n_nodes,
                                              scl_index_t n_features,
                                              scl_index_t n_steps, int* scale_exp,
                                              void* stream);

/* ------------------------------------------------------------- gradients --
 * Reverse mode of Y = P^t X for a loss L(Y): given dY = dL/dY, returns
 *   dX = (P^T)^t dY   and   dP_k = sum_{s=1..t} <G_s[i,:], X_{s-1}[j,:]>
 * for every stored nonzero k = (i, j) (CSR order), where X_s = P X_{s-1},
 * G_t = dY and G_{s-1} = P^T G_s.  The forward keeps the iterates X_c, X_2c,
 * ... (c = every) as checkpoints; the backward recomputes the others. */

/* Replaces the operator's values with values_dev (nnz floats, device, stored
 * order) in the device copies and in the handle's host arrays; synchronises
 * `stream`. */
SCL_EXPORT scl_error_t scl_gpu_sparse_set_values_device(scl_sparse_t matrix,
                                                        const scl_real_t* values_dev,
                                                        void* stream);

/* Y_dev = P^t X_dev, also writing the checkpoints X_every, X_2every, ...
 * (ceil(t/every) - 1 planes of n x F floats) to ckpt_dev; work_dev holds
 * 2 planes and is required when n_steps >= 2. */
SCL_EXPORT scl_error_t scl_gpu_diffuse_forward_ckpt_device(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* ckpt_dev, scl_real_t* work_dev, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream);

/* The same with P's pattern and the values values_dev (nnz floats, the
 * handle's stored order, device memory) in place of the handle's values,
 * which stay untouched: for callers that differentiate with respect to the
 * values (torch.autograd).  values_dev == NULL means the handle's values.
 * Shapes that read a grouped or tiled copy of the handle's values fall back
 * to the default shape. */
SCL_EXPORT scl_error_t scl_gpu_diffuse_forward_ckpt_values_device(
    scl_sparse_t transition, const scl_real_t* values_dev, const scl_real_t* X_dev,
    scl_real_t* Y_dev, scl_real_t* ckpt_dev, scl_real_t* work_dev, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream);

/* Edge gradient of one step on its own (CSR order):
 * dP_dev[k] (+)= <G_dev[i,:], Xprev_dev[j,:]> for every nonzero k = (i, j);
 * accumulate == 0 overwrites dP_dev.  The reverse pass above computes this
 * fused with the adjoint; this call is the unfused building block. */
SCL_EXPORT scl_error_t scl_gpu_edge_gradient_device(
    scl_sparse_t transition, const scl_real_t* G_dev, const scl_real_t* Xprev_dev,
    scl_real_t* dP_dev, scl_index_t n_nodes, scl_index_t n_features, int accumulate,
    void* stream);

/* Bytes of device workspace scl_gpu_diffuse_backward_device needs
 * ((every + 1) planes, plus nnz floats when want_dp). */
SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_workspace(
    scl_sparse_t transition, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t every, int want_dp, scl_size_t* bytes);

/* dX_dev (and dP_dev, nnz floats, when not NULL) from dY_dev, the forward
 * input X_dev and the checkpoints written by the forward call with the same
 * n_steps and every.  dX_dev must differ from dY_dev. */
SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_device(
    scl_sparse_t transition, const scl_real_t* X_dev, const scl_real_t* ckpt_dev,
    const scl_real_t* dY_dev, scl_real_t* dX_dev, scl_real_t* dP_dev,
    void* work_dev, scl_size_t work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream);

/* The reverse pass with P's pattern and the values values_dev (as in
 * scl_gpu_diffuse_forward_ckpt_values_device); its workspace, from the
 * query below, holds nnz more floats (the values in transposed order). */
SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_values_workspace(
    scl_sparse_t transition, scl_index_t n_nodes, scl_index_t n_features,
    scl_index_t every, int want_dp, scl_size_t* bytes);
SCL_EXPORT scl_error_t scl_gpu_diffuse_backward_values_device(
    scl_sparse_t transition, const scl_real_t* values_dev, const scl_real_t* X_dev,
    const scl_real_t* ckpt_dev, const scl_real_t* dY_dev, scl_real_t* dX_dev,
    scl_real_t* dP_dev, void* work_dev, scl_size_t work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every, void* stream);

/* A counter that changes whenever the handle's values or pattern may have
 * changed (any value update through the library, and every host-pointer call
 * on a handle made by scl_sparse_wrap, whose arrays the caller may edit). */
SCL_EXPORT scl_error_t scl_gpu_sparse_generation(scl_sparse_t matrix, uint64_t* out);

/* ----------------------------------------------------------------- plans --
 * A fixed sequence recorded once into a hipGraph and replayed with one
 * launch: Y = P^t X with checkpoints (scl_gpu_diffuse_forward_ckpt_device);
 * then, when T_dev is set, dY = Y - T (the gradient of 0.5 ||Y - T||^2);
 * then, when dY_dev is set, the reverse pass into dX_dev and dP_dev
 * (scl_gpu_diffuse_backward_device).  All buffers are the caller's and are
 * bound at record time.  The operator's device arrays are pinned until the
 * plan is destroyed: value updates are applied in place, a pattern change is
 * refused; destroy plans before their operator.  The graph holds kernel
 * nodes only. */
typedef struct scl_gpu_plan scl_gpu_plan;
typedef scl_gpu_plan* scl_gpu_plan_t;

SCL_EXPORT scl_error_t scl_gpu_plan_diffusion(
    scl_sparse_t transition, const scl_real_t* X_dev, scl_real_t* Y_dev,
    scl_real_t* ckpt_dev, scl_real_t* fwd_work_dev, const scl_real_t* T_dev,
    scl_real_t* dY_dev, scl_real_t* dX_dev, scl_real_t* dP_dev,
    void* bwd_work_dev, scl_size_t bwd_work_bytes, scl_index_t n_nodes,
    scl_index_t n_features, scl_index_t n_steps, scl_index_t every,
    scl_gpu_plan_t* out);
SCL_EXPORT scl_error_t scl_gpu_plan_launch(scl_gpu_plan_t plan, void* stream);
SCL_EXPORT scl_error_t scl_gpu_plan_destroy(scl_gpu_plan_t* plan);

/* Batched plan (no reference counterpart): the same sequence for every graph
 * of a block-diagonal stack (scl_sparse_block_diag) in ONE persistent kernel
 * launch, recorded into a hipGraph: t forward steps, dY = Y - T and t fused
 * adjoint + edge-gradient steps, each graph's steps ordered by the graph's
 * own progress rather than by launch boundaries; a gather recorded after the
 * kernel puts dP in CSR order.  block_rows[g] are the rows of block g (they add up to the
 * operator's rows, and no stored entry leaves its block: refused otherwise).
 * Each graph runs on one XCD (bin-packed by rows + nonzeros); rows_per_item
 * rows form one work item (0: 16) and the queue of an XCD walks its graphs
 * graphs_per_chunk at a time through all their steps (0: 4).
 * n_features % 32 == 0, 32..512; X, Y, T, dY, dX 128-byte aligned.  dX_dev
 * null: forward only (T, dY unused); dP_dev null: no edge gradient.  Results
 * are bit-identical to scl_gpu_plan_diffusion with every = 1 and the default
 * adjoint group (vec 4, one lane group per row covering F).  The plan
 * allocates its own intermediates ((t - 1) planes of iterates, (t - 1)
 * adjoint planes, t dP partials and the dP gather map).  Launch with scl_gpu_plan_launch. */
SCL_EXPORT scl_error_t scl_gpu_plan_diffusion_batched(
    scl_sparse_t transition, const scl_index_t* block_rows, scl_size_t n_blocks,
    const scl_real_t* X_dev, scl_real_t* Y_dev, const scl_real_t* T_dev, scl_real_t* dY_dev,
    scl_real_t* dX_dev, scl_real_t* dP_dev, scl_index_t n_features, scl_index_t n_steps,
    scl_index_t rows_per_item, scl_index_t graphs_per_chunk, scl_gpu_plan_t* out);
/* Waits for the stream, then reports whether the last launch of a batched
 * plan completed every work item (SCL_ERROR_INTERNAL with the reason if a
 * wait gave up or an XCD's queue was never served); SCL_OK for other plans. */
SCL_EXPORT scl_error_t scl_gpu_plan_check(scl_gpu_plan_t plan, void* stream);

#ifdef __cplusplus
}
#endif
#endif
