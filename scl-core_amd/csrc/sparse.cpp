// sparse.cpp -- sparse handle lifecycle and queries (counterpart of
// scl/binding/c_api/core/sparse.cpp:51-700).  Storage is one contiguous
// traditional CSR/CSC triple with int64 indices and float values; a handle
// also caches device copies (see device.cpp).
#include "internal.hpp"

#include <algorithm>
#include <cstring>

using scl_amd::fail;

scl_sparse_matrix::~scl_sparse_matrix() {
    dev.clear();
    if (owns_new) {
        delete[] indices;
        delete[] data;
    }
}

namespace {

// Validates a traditional triple the way the device needs it: monotone
// indptr starting at 0 and ending at nnz, indices inside the secondary range.
void validate_triple(int64_t primary, int64_t secondary, int64_t nnz,
                     const int64_t* indptr, const int64_t* indices) {
    if (indptr[0] != 0 || indptr[primary] != nnz)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "indptr must start at 0 and end at nnz");
    for (int64_t i = 0; i < primary; ++i)
        if (indptr[i + 1] < indptr[i])
            fail(SCL_ERROR_INVALID_ARGUMENT, "indptr must be non-decreasing");
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || indices[k] >= secondary)
            fail(SCL_ERROR_INDEX_OUT_OF_BOUNDS, "index out of range");
}

scl_sparse_matrix* new_handle(int64_t rows, int64_t cols, int64_t nnz,
                              bool is_csr) {
    auto* h = new scl_sparse_matrix();
    h->rows = rows;
    h->cols = cols;
    h->nnz = nnz;
    h->is_csr = is_csr;
    return h;
}

// Deep copy of a traditional triple into a new handle.
scl_sparse_matrix* copy_handle(int64_t rows, int64_t cols, int64_t nnz,
                               const int64_t* indptr, const int64_t* indices,
                               const float* data, bool is_csr) {
    auto* h = new_handle(rows, cols, nnz, is_csr);
    try {
        const int64_t p = h->primary();
        validate_triple(p, h->secondary(), nnz, indptr, indices);
        h->indptr_store.assign(indptr, indptr + p + 1);
        h->indices_store.assign(indices, indices + nnz);
        h->data_store.assign(data, data + nnz);
        h->indptr = h->indptr_store.data();
        h->indices = h->indices_store.data();
        h->data = h->data_store.data();
        // sparse.hpp:30-33 asks callers for sorted slots and checks it only
        // in debug builds; release builds keep the given order, and so do we
        // (each slot is accumulated in stored order).
    } catch (...) {
        delete h;
        throw;
    }
    return h;
}

scl_error_t create_impl(scl_sparse_t* out, scl_index_t rows, scl_index_t cols,
                        scl_index_t nnz, const scl_index_t* indptr,
                        const scl_index_t* indices, const scl_real_t* data,
                        scl_bool_t is_csr) {
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_CHECK_NULL(indptr, "Offset array pointer is null");
    SCL_API_CHECK_NULL(indices, "Index array pointer is null");
    SCL_API_CHECK_NULL(data, "Data array pointer is null");
    SCL_API_CHECK(rows > 0 && cols > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Matrix dimensions must be positive");
    SCL_API_CHECK(nnz >= 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Number of non-zeros must be non-negative");
    SCL_API_TRY
        *out = copy_handle(rows, cols, nnz, indptr, indices, data, is_csr != SCL_FALSE);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

scl_error_t wrap_impl(scl_sparse_t* out, scl_index_t rows, scl_index_t cols,
                      scl_index_t nnz, scl_index_t* indptr,
                      scl_index_t* indices, scl_real_t* data,
                      scl_bool_t is_csr, bool own) {
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_CHECK_NULL(indptr, "Offset array pointer is null");
    SCL_API_CHECK_NULL(indices, "Index array pointer is null");
    SCL_API_CHECK_NULL(data, "Data array pointer is null");
    SCL_API_CHECK(rows > 0 && cols > 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Matrix dimensions must be positive");
    SCL_API_CHECK(nnz >= 0, SCL_ERROR_INVALID_ARGUMENT,
                  "Number of non-zeros must be non-negative");
    SCL_API_TRY
        auto* h = new_handle(rows, cols, nnz, is_csr != SCL_FALSE);
        try {
            validate_triple(h->primary(), h->secondary(), nnz, indptr, indices);
        } catch (...) {
            delete h;
            throw;
        }
        // wrap_traditional (sparse.hpp:795-840) reads indptr once to set up
        // per-slot pointers and aliases indices/data: keep our own indptr.
        h->indptr_store.assign(indptr, indptr + h->primary() + 1);
        h->indptr = h->indptr_store.data();
        h->indices = indices;
        h->data = data;
        h->wrapped = !own;
        h->owns_new = own;
        *out = h;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

}  // namespace

extern "C" {

SCL_EXPORT scl_error_t scl_sparse_create(scl_sparse_t* out, scl_index_t rows,
                                         scl_index_t cols, scl_index_t nnz,
                                         const scl_index_t* indptr,
                                         const scl_index_t* indices,
                                         const scl_real_t* data,
                                         scl_bool_t is_csr) {
    return create_impl(out, rows, cols, nnz, indptr, indices, data, is_csr);
}

SCL_EXPORT scl_error_t scl_sparse_create_with_strategy(
    scl_sparse_t* out, scl_index_t rows, scl_index_t cols, scl_index_t nnz,
    const scl_index_t* indptr, const scl_index_t* indices,
    const scl_real_t* data, scl_bool_t is_csr, scl_block_strategy_t strategy) {
    SCL_API_CHECK(strategy >= SCL_BLOCK_STRATEGY_CONTIGUOUS &&
                      strategy <= SCL_BLOCK_STRATEGY_ADAPTIVE,
                  SCL_ERROR_INVALID_ARGUMENT, "Unknown block strategy");
    return create_impl(out, rows, cols, nnz, indptr, indices, data, is_csr);
}

SCL_EXPORT scl_error_t scl_sparse_wrap(scl_sparse_t* out, scl_index_t rows,
                                       scl_index_t cols, scl_index_t nnz,
                                       scl_index_t* indptr,
                                       scl_index_t* indices, scl_real_t* data,
                                       scl_bool_t is_csr) {
    return wrap_impl(out, rows, cols, nnz, indptr, indices, data, is_csr, false);
}

SCL_EXPORT scl_error_t scl_sparse_wrap_and_own(scl_sparse_t* out,
                                               scl_index_t rows,
                                               scl_index_t cols,
                                               scl_index_t nnz,
                                               scl_index_t* indptr,
                                               scl_index_t* indices,
                                               scl_real_t* data,
                                               scl_bool_t is_csr) {
    return wrap_impl(out, rows, cols, nnz, indptr, indices, data, is_csr, true);
}

SCL_EXPORT scl_error_t scl_sparse_clone(scl_sparse_t src, scl_sparse_t* out) {
    SCL_API_CHECK_NULL(src, "Source matrix is null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_TRY
        *out = copy_handle(src->rows, src->cols, src->nnz, src->indptr,
                           src->indices, src->data, src->is_csr);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_sparse_destroy(scl_sparse_t* matrix) {
    if (matrix == nullptr || *matrix == nullptr) SCL_API_RETURN_OK;
    SCL_API_TRY
        delete *matrix;
        *matrix = nullptr;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

#define SCL_QUERY(name, type, expr)                                     \
    SCL_EXPORT scl_error_t name(scl_sparse_t matrix, type* out) {       \
        SCL_API_CHECK_NULL(matrix, "Matrix is null");                   \
        SCL_API_CHECK_NULL(out, "Output pointer is null");              \
        *out = (expr);                                                  \
        SCL_API_RETURN_OK;                                              \
    }

SCL_QUERY(scl_sparse_rows, scl_index_t, matrix->rows)
SCL_QUERY(scl_sparse_cols, scl_index_t, matrix->cols)
SCL_QUERY(scl_sparse_nnz, scl_index_t, matrix->nnz)
SCL_QUERY(scl_sparse_is_csr, scl_bool_t, matrix->is_csr ? SCL_TRUE : SCL_FALSE)
SCL_QUERY(scl_sparse_is_csc, scl_bool_t, matrix->is_csr ? SCL_FALSE : SCL_TRUE)
SCL_QUERY(scl_sparse_is_valid, scl_bool_t,
          (matrix->rows > 0 && matrix->cols > 0) ? SCL_TRUE : SCL_FALSE)
SCL_QUERY(scl_sparse_is_contiguous, scl_bool_t, SCL_TRUE)

#undef SCL_QUERY

SCL_EXPORT scl_error_t scl_sparse_export(scl_sparse_t matrix,
                                         scl_index_t* indptr,
                                         scl_index_t* indices,
                                         scl_real_t* data) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(indptr, "Offset array pointer is null");
    SCL_API_CHECK_NULL(indices, "Index array pointer is null");
    SCL_API_CHECK_NULL(data, "Data array pointer is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(matrix->mu);
        const int64_t p = matrix->primary();
        std::memcpy(indptr, matrix->indptr, sizeof(int64_t) * (size_t)(p + 1));
        std::memcpy(indices, matrix->indices, sizeof(int64_t) * (size_t)matrix->nnz);
        std::memcpy(data, matrix->data, sizeof(float) * (size_t)matrix->nnz);
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_sparse_get_lengths(scl_sparse_t matrix,
                                              const scl_index_t** lengths,
                                              scl_size_t* lengths_size) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(lengths, "Output pointer is null");
    SCL_API_CHECK_NULL(lengths_size, "Size pointer is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(matrix->mu);
        const int64_t p = matrix->primary();
        matrix->lengths_cache.resize((size_t)p);
        for (int64_t i = 0; i < p; ++i)
            matrix->lengths_cache[i] = matrix->indptr[i + 1] - matrix->indptr[i];
        *lengths = matrix->lengths_cache.data();
        *lengths_size = (scl_size_t)p;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_sparse_get_indices(scl_sparse_t matrix,
                                              const scl_index_t** indices,
                                              scl_size_t* indices_size) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(indices, "Output pointer is null");
    SCL_API_CHECK_NULL(indices_size, "Size pointer is null");
    *indices = matrix->indices;
    *indices_size = (scl_size_t)matrix->nnz;
    SCL_API_RETURN_OK;
}

SCL_EXPORT scl_error_t scl_sparse_get_data(scl_sparse_t matrix,
                                           const scl_real_t** data,
                                           scl_size_t* data_size) {
    SCL_API_CHECK_NULL(matrix, "Matrix is null");
    SCL_API_CHECK_NULL(data, "Output pointer is null");
    SCL_API_CHECK_NULL(data_size, "Size pointer is null");
    *data = matrix->data;
    *data_size = (scl_size_t)matrix->nnz;
    SCL_API_RETURN_OK;
}

SCL_EXPORT scl_error_t scl_sparse_transpose(scl_sparse_t src, scl_sparse_t* out) {
    SCL_API_CHECK_NULL(src, "Source matrix is null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_TRY
        std::lock_guard<std::mutex> lk(src->mu);
        // counting sort of the entries by their secondary index; visiting the
        // source slots in order keeps every new slot sorted
        const int64_t p = src->primary(), q = src->secondary(), nnz = src->nnz;
        auto* h = new_handle(src->cols, src->rows, nnz, !src->is_csr);
        try {
            h->indptr_store.assign((size_t)q + 1, 0);
            h->indices_store.resize((size_t)nnz);
            h->data_store.resize((size_t)nnz);
            for (int64_t k = 0; k < nnz; ++k) h->indptr_store[src->indices[k] + 1]++;
            for (int64_t j = 0; j < q; ++j) h->indptr_store[j + 1] += h->indptr_store[j];
            std::vector<int64_t> pos(h->indptr_store.begin(), h->indptr_store.end() - 1);
            for (int64_t i = 0; i < p; ++i)
                for (int64_t k = src->indptr[i]; k < src->indptr[i + 1]; ++k) {
                    const int64_t d = pos[src->indices[k]]++;
                    h->indices_store[d] = i;
                    h->data_store[d] = src->data[k];
                }
            h->indptr = h->indptr_store.data();
            h->indices = h->indices_store.data();
            h->data = h->data_store.data();
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_sparse_block_diag(const scl_sparse_t* mats,
                                             scl_size_t count, scl_sparse_t* out) {
    SCL_API_CHECK_NULL(mats, "Matrix array is null");
    SCL_API_CHECK_NULL(out, "Output pointer is null");
    SCL_API_CHECK(count > 0, SCL_ERROR_INVALID_ARGUMENT, "count must be positive");
    for (scl_size_t g = 0; g < count; ++g)
        SCL_API_CHECK(mats[g] != nullptr, SCL_ERROR_NULL_POINTER, "Matrix is null");
    SCL_API_TRY
        const bool csr = mats[0]->is_csr;
        int64_t rows = 0, cols = 0, nnz = 0;
        for (scl_size_t g = 0; g < count; ++g) {
            if (mats[g]->is_csr != csr)
                fail(SCL_ERROR_TYPE_MISMATCH, "all blocks must have the same orientation");
            rows += mats[g]->rows;
            cols += mats[g]->cols;
            nnz += mats[g]->nnz;
        }
        auto* h = new_handle(rows, cols, nnz, csr);
        try {
            h->indptr_store.reserve((size_t)h->primary() + 1);
            h->indices_store.reserve((size_t)nnz);
            h->data_store.reserve((size_t)nnz);
            h->indptr_store.push_back(0);
            int64_t prim_off = 0, sec_off = 0, nz_off = 0;
            for (scl_size_t g = 0; g < count; ++g) {
                scl_sparse_matrix* m = mats[g];
                std::lock_guard<std::mutex> lk(m->mu);
                const int64_t p = m->primary();
                for (int64_t i = 0; i < p; ++i)
                    h->indptr_store.push_back(nz_off + m->indptr[i + 1]);
                for (int64_t k = 0; k < m->nnz; ++k) {
                    h->indices_store.push_back(sec_off + m->indices[k]);
                    h->data_store.push_back(m->data[k]);
                }
                prim_off += p;
                sec_off += m->secondary();
                nz_off += m->nnz;
            }
            h->indptr = h->indptr_store.data();
            h->indices = h->indices_store.data();
            h->data = h->data_store.data();
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        SCL_API_RETURN_OK;
    SCL_API_CATCH
}

SCL_EXPORT scl_error_t scl_sparse_to_contiguous(scl_sparse_t src,
                                                scl_sparse_t* out) {
    return scl_sparse_clone(src, out);
}

}  // extern "C"
