// internal.hpp -- shared declarations of the scl diffusion library (host side).
//
// Error channel and ABI guard macros mirror scl/binding/c_api/core/internal.hpp
// (307-343) and core.cpp (33-253): argument checks return a code early, bodies
// run inside a try block, any exception becomes a code plus a thread-local
// message, and success clears the last error.
#pragma once

#include "kernels.hpp"
#include "scl_core.h"
#include "scl_sparse.h"

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace scl_amd {

// ----------------------------------------------------------------- errors --

struct Error {
    scl_error_t code;
    std::string msg;
};

[[noreturn]] void fail(scl_error_t code, const std::string& msg);
void set_last_error(scl_error_t code, const char* msg) noexcept;
void clear_last_error() noexcept;
scl_error_t handle_exception() noexcept;

// HIP status -> exception: out-of-memory maps to SCL_ERROR_OUT_OF_MEMORY,
// a missing device to SCL_ERROR_FEATURE_UNAVAILABLE, anything else to
// SCL_ERROR_INTERNAL.
void hip_check(hipError_t e, const char* what);

}  // namespace scl_amd

#define SCL_API_CHECK(cond, code, msg)                   \
    do {                                                 \
        if (!(cond)) {                                   \
            ::scl_amd::set_last_error((code), (msg));    \
            return (code);                               \
        }                                                \
    } while (0)

#define SCL_API_CHECK_NULL(p, msg) \
    SCL_API_CHECK((p) != nullptr, SCL_ERROR_NULL_POINTER, msg)

#define SCL_API_TRY try {
#define SCL_API_CATCH                               \
    }                                               \
    catch (...) {                                   \
        return ::scl_amd::handle_exception();       \
    }

#define SCL_API_RETURN_OK                 \
    do {                                  \
        ::scl_amd::clear_last_error();    \
        return SCL_OK;                    \
    } while (0)

namespace scl_amd {

// ------------------------------------------------------------ device data --

// A handle's arrays on one GPU: the primary-slot CSR with int32 indices, plus
// lazily built companions (transposed pattern, bf16 values, FP64 degrees).
struct DeviceCSR {
    int device = -1;
    uint64_t generation = ~uint64_t(0);
    int32_t n = 0;       // primary dimension (rows of the applied operator)
    int32_t n_sec = 0;   // secondary dimension
    int64_t nnz = 0;
    int32_t* row_ptr = nullptr;  // n + 1
    int32_t* col = nullptr;      // nnz
    float* val = nullptr;        // nnz
    int32_t* sched = nullptr;    // n: row visiting order (locality), may be null
    uint64_t fingerprint = 0;    // pattern the schedule was built for
    // transposed pattern (for adjoints): t_row_ptr (n_sec + 1), t_col, t_val,
    // t_src[p] = position in (col, val) of transposed entry p
    uint64_t t_generation = ~uint64_t(0);
    int32_t* t_row_ptr = nullptr;
    int32_t* t_col = nullptr;
    float* t_val = nullptr;
    int32_t* t_src = nullptr;
    // t_dst[k] = transposed position of entry k (the inverse of t_src), so
    // that putting transposed-order values in CSR order is a gather
    int32_t* t_dst = nullptr;
    int32_t* t_sched = nullptr;
    // number of captured plans that hold raw pointers into this copy: while
    // non-zero, refreshes update the arrays in place and never reallocate
    int pins = 0;
    // values rounded to bf16 (built on first use, for the current values)
    uint16_t* val_bf16 = nullptr;
    uint64_t bf16_generation = ~uint64_t(0);
    // values in compensated bf16 (hi + 8-bit residual), built on first use
    CompPlanes val_comp;
    uint64_t comp_generation = ~uint64_t(0);
    // launch shapes chosen per feature count (autotuned or set by the
    // caller) for the operator and for its transpose
    std::map<int64_t, SpmmShape> shapes, t_shapes;
    std::map<int64_t, SpmmShape> adj_shapes;  // fused adjoint + edge kernel
    // scratch of merge-path launches (both directions share it; grown when a
    // merge-path shape is set, never while plans hold pointers into it)
    MergeScratch merge;
    // row-grouped forms of the forward operator, R = 4 and R = 8 (built when
    // a grouped shape is set or tuned; values refreshed with the operator's)
    GroupedCsr grouped[2];
    // LDS-staged groups of kStagedRows rows (pattern only: no values)
    StagedCsr staged;
    // LDS-tiled forms, [tiled_index(lpr, budget)] (values refreshed with the
    // operator's)
    TiledCsr tiled[4];
    // the operator renumbered in the locality order, for single-vector
    // products: row r' is row sched[r'] with its entries in stored order,
    // column c becomes inv(sched)[c]; src[k'] = the entry's place in (col,
    // val).  Built on first use (square operators with a schedule), values
    // re-gathered with the operator's.
    struct {
        int32_t* row_ptr = nullptr;
        int32_t* col = nullptr;
        float* val = nullptr;
        int32_t* src = nullptr;
        uint64_t fingerprint = 0;
    } ren;

    ~DeviceCSR();
    void release_grouped();
    void release_transpose();
    void release_renumbered();
};

}  // namespace scl_amd

// The opaque handle of the C ABI (scl_sparse_t points at one of these).
struct scl_sparse_matrix {
    int64_t rows = 0, cols = 0, nnz = 0;
    bool is_csr = true;
    bool wrapped = false;      // scl_sparse_wrap: arrays belong to the caller
    bool owns_new = false;     // scl_sparse_wrap_and_own: delete[] on destroy
    std::vector<int64_t> indptr_store, indices_store;
    std::vector<float> data_store;
    int64_t* indptr = nullptr;   // primary + 1 (always our own copy)
    int64_t* indices = nullptr;  // nnz
    float* data = nullptr;       // nnz
    std::vector<int64_t> lengths_cache;
    uint64_t generation = 0;     // bumped whenever values/pattern may change
    std::mutex mu;
    std::map<int, std::unique_ptr<scl_amd::DeviceCSR>> dev;

    int64_t primary() const { return is_csr ? rows : cols; }
    int64_t secondary() const { return is_csr ? cols : rows; }
    ~scl_sparse_matrix();
};

namespace scl_amd {

// ---------------------------------------------------------- device runtime --

// Current HIP device, or throws SCL_ERROR_FEATURE_UNAVAILABLE when none.
int current_device();
// Library stream of a device (default priority, non-blocking).
hipStream_t library_stream(int device);
// Serialises the host-pointer entry points on one device (they share the
// device's staging buffers).
std::mutex& device_mutex(int device);
// Copies between pageable caller memory and device memory, stream-ordered
// on s: copies of 8 MiB or more go through the device's two 64 MiB pinned
// bounce buffers in chunks (host side on several threads, overlapped with
// the copy engine).  upload_host
// returns once the caller's buffer has been read; download_host returns
// once the data is in the caller's buffer.
void upload_host(void* dst, const void* src, size_t bytes, hipStream_t s, int device);
void download_host(void* dst, const void* src, size_t bytes, hipStream_t s, int device);
// Grow-only device buffer `slot` of a device, at least `bytes` long.
void* device_workspace(int device, int slot, size_t bytes);
// Grows d.merge to hold `waves` waves of F features.
void ensure_merge_scratch(DeviceCSR& d, int64_t waves, int64_t F);
// Builds d.grouped[R == 8] for the handle's current pattern when missing or
// stale; false when the pattern cannot be grouped (a row with unsorted or
// repeated column indices, or >= 2^24 columns).
bool ensure_grouped(scl_sparse_matrix* h, DeviceCSR& d, int R);
// Builds d.staged for the handle's current pattern when missing or stale;
// false when the pattern cannot be staged (a row with unsorted or repeated
// column indices).
bool ensure_staged(scl_sparse_matrix* h, DeviceCSR& d);
// Builds d.tiled[tiled_index(lpr, budget)] for the handle's current pattern
// when missing or stale; false when the pattern cannot be tiled (a row with
// unsorted or repeated column indices, or too long for one tile).
bool ensure_tiled(scl_sparse_matrix* h, DeviceCSR& d, int lpr, int budget);
// Re-gathers the grouped forms' values from d.val (after the device values
// changed in place), stream-ordered on s.
void refresh_grouped_values(DeviceCSR& d, hipStream_t s);
// Builds d.ren for the handle's current pattern when missing or stale; false
// when the operator has no schedule or is not square.
bool ensure_renumbered(scl_sparse_matrix* h, DeviceCSR& d);

// Operators with at least this many rows get a locality schedule.
constexpr int64_t kScheduleMinRows = 16384;

// Greedy locality visiting order of a CSR pattern: each next row is the
// unplaced one with the most pattern edges to the last 512 placed rows
// (schedule.cpp).
template <typename I>
std::vector<int32_t> locality_schedule(int64_t n, const I* ptr, const I* idx);

// Launch shape of the forward (or transposed) SpMM for F features: the one
// tuned or set for this operator when it fits X and Y; else, for wide
// forward operators with sorted rows, the row-grouped shape (its grouped
// form built on the first such call when `build` is set); else the default.
// The caller holds the handle's mutex (diffusion.cpp).
SpmmShape choose_shape(scl_sparse_matrix* h, DeviceCSR& d, bool transposed, int64_t F,
                       const void* X, const void* Y, bool build = true);

// Rows of a square pattern in groups of R for the row-grouped SpMM form:
// walks `order`; each row not yet placed opens a group, which is filled
// with the unplaced neighbour of its rows sharing the most columns with the
// group's union so far, among rows at most 512 places after the group's
// first row in `order` (the next unplaced row of `order` when none is left).  Returns every row once, each group's rows consecutive, groups
// in the order of their first row; or `order` itself when its consecutive
// groups have fewer distinct columns in all.
std::vector<int32_t> neighbour_groups(int64_t n, const int64_t* ptr, const int64_t* idx,
                                      const std::vector<int32_t>& order, int R);

// Device copy of a handle, refreshed when its generation changed.
// Caller holds h->mu.
DeviceCSR& device_csr(scl_sparse_matrix* h, int device);
// The device copy when it exists and matches the handle's generation, else
// null; never uploads or allocates.  Caller holds h->mu.
DeviceCSR* device_csr_current(scl_sparse_matrix* h, int device);
// Builds the transposed pattern on the device if missing or stale.
void ensure_transpose(scl_sparse_matrix* h, DeviceCSR& d, hipStream_t s);

}  // namespace scl_amd
