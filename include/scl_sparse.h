/*
 * scl_sparse.h -- sparse-matrix handles (the operator P of the diffusion).
 *
 * Replaces the subset of scl/binding/c_api/core/sparse.h that callers of the
 * diffusion path use to create, inspect and release P.  Signatures, argument
 * meaning and error codes are those of the reference; storage is always one
 * contiguous CSR/CSC block, so every handle is "contiguous" and the block
 * strategy argument only selects how the host copy is made.
 *
 * A handle keeps a device copy of its arrays (int32 indices, values as
 * float) on each GPU it is used on.  The copy is refreshed when the handle's
 * contents change through this library, and before every host-pointer
 * diffusion call on handles made by scl_sparse_wrap (whose arrays the caller
 * may edit at any time).  Device-pointer extension calls use the cached copy;
 * call scl_gpu_sparse_invalidate() after editing a wrapped handle's arrays.
 */
#ifndef SCL_SPARSE_H
#define SCL_SPARSE_H

#include "scl_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sparse.h:53-70 */
typedef enum {
    SCL_BLOCK_STRATEGY_CONTIGUOUS = 0,
    SCL_BLOCK_STRATEGY_SMALL = 1,
    SCL_BLOCK_STRATEGY_LARGE = 2,
    SCL_BLOCK_STRATEGY_ADAPTIVE = 3
} scl_block_strategy_t;

/* sparse.h:103-112 / sparse.cpp:51-110: deep copy of a traditional CSR
 * (is_csr != 0) or CSC array triple.  rows, cols > 0; nnz >= 0. */
SCL_EXPORT scl_error_t scl_sparse_create(scl_sparse_t* out, scl_index_t rows,
                                         scl_index_t cols, scl_index_t nnz,
                                         const scl_index_t* indptr,
                                         const scl_index_t* indices,
                                         const scl_real_t* data,
                                         scl_bool_t is_csr);

/* sparse.h:125-135 */
SCL_EXPORT scl_error_t scl_sparse_create_with_strategy(
    scl_sparse_t* out, scl_index_t rows, scl_index_t cols, scl_index_t nnz,
    const scl_index_t* indptr, const scl_index_t* indices,
    const scl_real_t* data, scl_bool_t is_csr, scl_block_strategy_t strategy);

/* sparse.h:149-158 / sparse.cpp:176-233: zero-copy; the caller keeps
 * ownership and must keep the arrays alive until scl_sparse_destroy. */
SCL_EXPORT scl_error_t scl_sparse_wrap(scl_sparse_t* out, scl_index_t rows,
                                       scl_index_t cols, scl_index_t nnz,
                                       scl_index_t* indptr,
                                       scl_index_t* indices, scl_real_t* data,
                                       scl_bool_t is_csr);

/* sparse.h:172-181: zero-copy and takes ownership of indices and data
 * (allocated with new[]); indptr stays the caller's. */
SCL_EXPORT scl_error_t scl_sparse_wrap_and_own(scl_sparse_t* out,
                                               scl_index_t rows,
                                               scl_index_t cols,
                                               scl_index_t nnz,
                                               scl_index_t* indptr,
                                               scl_index_t* indices,
                                               scl_real_t* data,
                                               scl_bool_t is_csr);

/* sparse.h:189 */
SCL_EXPORT scl_error_t scl_sparse_clone(scl_sparse_t src, scl_sparse_t* out);

/* sparse.h:196: releases host and device storage; sets *matrix to NULL;
 * NULL or *NULL is accepted. */
SCL_EXPORT scl_error_t scl_sparse_destroy(scl_sparse_t* matrix);

/* sparse.h:203-221 */
SCL_EXPORT scl_error_t scl_sparse_rows(scl_sparse_t matrix, scl_index_t* out);
SCL_EXPORT scl_error_t scl_sparse_cols(scl_sparse_t matrix, scl_index_t* out);
SCL_EXPORT scl_error_t scl_sparse_nnz(scl_sparse_t matrix, scl_index_t* out);
SCL_EXPORT scl_error_t scl_sparse_is_csr(scl_sparse_t matrix, scl_bool_t* out);
SCL_EXPORT scl_error_t scl_sparse_is_csc(scl_sparse_t matrix, scl_bool_t* out);
SCL_EXPORT scl_error_t scl_sparse_is_valid(scl_sparse_t matrix, scl_bool_t* out);
SCL_EXPORT scl_error_t scl_sparse_is_contiguous(scl_sparse_t matrix,
                                                scl_bool_t* out);

/* sparse.h:241-246: copies to caller arrays of size primary+1, nnz, nnz. */
SCL_EXPORT scl_error_t scl_sparse_export(scl_sparse_t matrix,
                                         scl_index_t* indptr,
                                         scl_index_t* indices,
                                         scl_real_t* data);

/* sparse.h:256-286: internal pointers (valid until the handle changes). */
SCL_EXPORT scl_error_t scl_sparse_get_lengths(scl_sparse_t matrix,
                                              const scl_index_t** lengths,
                                              scl_size_t* lengths_size);
SCL_EXPORT scl_error_t scl_sparse_get_indices(scl_sparse_t matrix,
                                              const scl_index_t** indices,
                                              scl_size_t* indices_size);
SCL_EXPORT scl_error_t scl_sparse_get_data(scl_sparse_t matrix,
                                           const scl_real_t** data,
                                           scl_size_t* data_size);

/* sparse.h:297 / scl/core/sparse.hpp:1141-1197: a handle of the opposite
 * orientation whose primary slots are the source's secondary slots (entries
 * regrouped by the source's column, indices sorted), with rows and cols
 * swapped.  Diffusion through the result applies the source's transpose. */
SCL_EXPORT scl_error_t scl_sparse_transpose(scl_sparse_t src, scl_sparse_t* out);

/* Block-diagonal stack of `count` handles of one orientation (a new deep
 * copy): block g's slots and indices are offset by the sizes of blocks
 * 0..g-1, values and in-slot order are kept.  No counterpart in the
 * reference (which has vstack / hstack, sparse.h:444-460): it lets many
 * independent graphs be diffused by one launch per step, since rows of
 * different blocks never mix. */
SCL_EXPORT scl_error_t scl_sparse_block_diag(const scl_sparse_t* mats,
                                             scl_size_t count, scl_sparse_t* out);

/* sparse.h:304 */
SCL_EXPORT scl_error_t scl_sparse_to_contiguous(scl_sparse_t src,
                                                scl_sparse_t* out);

#ifdef __cplusplus
}
#endif
#endif
