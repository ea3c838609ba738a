// session.hpp -- device buffers and the per-call operator session shared by
// the entry points that keep whole iterations on the GPU (walks.cpp,
// callers.cpp).
#pragma once

#include "internal.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

namespace scl_amd {

// Device buffer of `count` elements of `elem` bytes, freed at scope exit.
class DevBuf {
public:
    explicit DevBuf(size_t count, size_t elem = sizeof(float)) {
        if (hipMalloc(&p_, std::max<size_t>(count, 1) * elem) != hipSuccess) {
            (void)hipGetLastError();
            fail(SCL_ERROR_OUT_OF_MEMORY, "device allocation failed");
        }
    }
    ~DevBuf() { (void)hipFree(p_); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    float* get() const { return static_cast<float*>(p_); }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p_);
    }

private:
    void* p_ = nullptr;
};

// The operator on the current device, its view and the stream, with the
// device and handle locks held for the whole call.  A square session also
// checks that n_nodes covers the operator's rows.
struct Session {
    int dev;
    std::unique_lock<std::mutex> call_lk, lk;
    DeviceCSR* d;
    CsrView v;
    hipStream_t s;
    int64_t n;
    scl_sparse_matrix* h;
    std::unique_ptr<DevBuf> xs;  // x in the locality order (renumbered products)
    int ren_state = -1;          // -1 unknown, 0 no renumbered copy, 1 ready

    bool renumbered() {
        if (ren_state < 0) ren_state = ensure_renumbered(h, *d) ? 1 : 0;
        return ren_state == 1;
    }

    Session(scl_sparse_matrix* P, int64_t n_nodes, bool square = true)
        : dev(current_device()), call_lk(device_mutex(dev)), lk(P->mu), h(P) {
        if (square && P->rows != P->cols)
            fail(SCL_ERROR_DIMENSION_MISMATCH, "the transition matrix must be square");
        n = P->primary();
        if (square && n_nodes < n)
            fail(SCL_ERROR_DIMENSION_MISMATCH,
                 "output buffer smaller than the transition matrix");
        if (P->wrapped) P->generation++;
        d = &device_csr(P, dev);
        v = CsrView{d->n, d->n_sec, d->nnz, d->row_ptr, d->col, d->val, d->sched, &d->merge,
                    d->grouped, &d->staged, d->tiled};
        s = library_stream(dev);
    }

    // The shape tuned or set for F features when it fits X and Y; else, for
    // wide operators, the row-grouped shape the drop-in diffusion runs (its
    // grouped form built on first use); else the default.
    SpmmShape shape(int64_t F, const float* X, const float* Y) const {
        return choose_shape(h, *d, false, F, X, Y);
    }
    // y = P x.  With the default single-vector shape, square operators with
    // a locality schedule run their renumbered copy (built on the first such
    // call): x is gathered into the locality order, and each row's result
    // is stored at its original place; every row keeps its chain.
    void spmv(const float* x, float* y) {
        const SpmmShape sh = shape(1, x, y);
        if (sh.group == 1 && sh.vec == 1 && !sh.merge && !sh.rows_per_group &&
            v.n_sec < (1 << 30) && renumbered()) {
            if (!xs) xs.reset(new DevBuf((size_t)v.n));
            launch_gather_f32(d->sched, x, xs->get(), v.n, s);
            CsrView r = v;
            r.row_ptr = d->ren.row_ptr;
            r.col = d->ren.col;
            r.val = d->ren.val;
            launch_spmv_staged(r, d->sched, xs->get(), y, sh.unroll, s);
            return;
        }
        launch_spmm_f32(v, x, y, 1, sh, s);
    }
    void spmm(const float* X, float* Y, int64_t F) {
        launch_spmm_f32(v, X, Y, F, shape(F, X, Y), s);
    }
    // Y = P X in the locality numbering (X and Y hold row r' for row
    // sched[r'] of the operator); needs renumbered().  Every row keeps its
    // stored-order chain.
    void spmm_renumbered(const float* X, float* Y, int64_t F) {
        CsrView r = v;
        r.row_ptr = d->ren.row_ptr;
        r.col = d->ren.col;
        r.val = d->ren.val;
        r.sched = nullptr;
        r.grouped = nullptr;
        r.staged = nullptr;
        r.tiled = nullptr;
        launch_spmm_f32(r, X, Y, F, spmm_default_shape(F, X, Y), s);
    }
    // rows into (scatter = false) or out of (true) the locality numbering
    void renumber_rows(const float* src, float* dst, int64_t F, bool back) {
        launch_rows_f32(d->sched, src, dst, v.n, F, back, s);
    }
    // Whether a walk keeps its vectors in the locality numbering: the
    // default single-vector shape on an operator with a renumbered copy.
    // Its products are then spmv_renumbered(); the element-wise steps and
    // the maxima of a walk do not depend on the order.
    bool walk_renumbered() {
        const auto it = d->shapes.find(1);
        if (it != d->shapes.end() &&
            (it->second.group != 1 || it->second.vec != 1 || it->second.merge ||
             it->second.rows_per_group))
            return false;
        return v.n_sec < (1 << 30) && renumbered();
    }
    // y = P x with x and y in the locality numbering
    void spmv_renumbered(const float* x, float* y) {
        CsrView r = v;
        r.row_ptr = d->ren.row_ptr;
        r.col = d->ren.col;
        r.val = d->ren.val;
        const auto it = d->shapes.find(1);
        launch_spmv_staged(r, nullptr, x, y, it != d->shapes.end() ? it->second.unroll : 8, s);
    }
    // Y = P' X for the operator's pattern with other values (same order as
    // the stored ones).
    void spmm_values(const float* val, const float* X, float* Y, int64_t F) {
        CsrView w = v;
        w.val = val;
        w.grouped = nullptr;
        w.tiled = nullptr;
        SpmmShape sh = shape(F, X, Y);
        // the staged form reads the view's values; the grouped and tiled
        // forms hold copies of the operator's
        if (sh.rows_per_group && sh.rows_per_group != kStagedRows)
            sh = spmm_default_shape(F, X, Y);
        launch_spmm_f32(w, X, Y, F, sh, s);
    }
    void upload(float* dst, const float* src, int64_t m) {
        upload_host(dst, src, (size_t)m * 4, s, dev);
    }
    void download(float* dst, const float* src, int64_t m) {
        download_host(dst, src, (size_t)m * 4, s, dev);
        hip_check(hipStreamSynchronize(s), "download");
    }
    float scalar(const float* src) {
        float h = 0.0f;
        download(&h, src, 1);
        return h;
    }
};

// The first `slots` primary slots of a handle on the device, as stored (int64
// pattern rebased to start at 0, float values).  `serial` is set when some
// slot is not strictly increasing (a repeated or descending index): such
// slots are walked in stored order by one lane, so that a repeated index
// keeps its last value.  Copied under the handle's lock; take it before any
// other lock.
struct HostPattern {
    int64_t slots;
    DevBuf ip, ix, v;
    bool serial = false;

    HostPattern(scl_sparse_matrix* X, int64_t slots_, hipStream_t s)
        : slots(slots_), ip(slots_ + 1, 8), ix(count(X, slots_), 8), v(count(X, slots_)) {
        std::lock_guard<std::mutex> lk(X->mu);
        const int64_t nnz = X->indptr[slots] - X->indptr[0];
        std::vector<int64_t> rebased((size_t)slots + 1);
        for (int64_t i = 0; i <= slots; ++i) rebased[i] = X->indptr[i] - X->indptr[0];
        hip_check(hipMemcpyAsync(ip.as<int64_t>(), rebased.data(), (size_t)(slots + 1) * 8,
                                 hipMemcpyHostToDevice, s),
                  "upload pattern");
        if (nnz > 0) {
            const int dev = current_device();
            upload_host(ix.as<int64_t>(), X->indices + X->indptr[0], (size_t)nnz * 8, s, dev);
            upload_host(v.get(), X->data + X->indptr[0], (size_t)nnz * 4, s, dev);
        }
        DevBuf flag(1, sizeof(int));
        hip_check(hipMemsetAsync(flag.get(), 0, sizeof(int), s), "clear flag");
        int h = 0;
        if (launch_find_unsorted(ip.as<int64_t>(), ix.as<int64_t>(), slots, flag.as<int>(), s))
            hip_check(hipMemcpyAsync(&h, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s),
                      "download flag");
        hip_check(hipStreamSynchronize(s), "upload pattern");
        serial = h != 0;
    }

    static int64_t count(scl_sparse_matrix* X, int64_t slots) {
        std::lock_guard<std::mutex> lk(X->mu);
        return X->indptr[slots] - X->indptr[0];
    }

    void densify(float* out, int64_t n_cells, int64_t G, bool is_csr, hipStream_t s) const {
        launch_densify(ip.as<int64_t>(), ix.as<int64_t>(), v.get(), slots, n_cells, G, is_csr,
                       serial, out, s);
    }
};

// The row pointer (rebased to 0) and stored values of the first `slots`
// primary slots on the device, without the column indices: for the
// per-row operations that never read them.
struct HostRows {
    int64_t slots;
    DevBuf ip, v;

    HostRows(scl_sparse_matrix* X, int64_t slots_, hipStream_t s)
        : slots(slots_), ip(slots_ + 1, 8), v(HostPattern::count(X, slots_)) {
        std::lock_guard<std::mutex> lk(X->mu);
        const int64_t nnz = X->indptr[slots] - X->indptr[0];
        std::vector<int64_t> rebased((size_t)slots + 1);
        for (int64_t i = 0; i <= slots; ++i) rebased[i] = X->indptr[i] - X->indptr[0];
        hip_check(hipMemcpyAsync(ip.as<int64_t>(), rebased.data(), (size_t)(slots + 1) * 8,
                                 hipMemcpyHostToDevice, s),
                  "upload row pointer");
        if (nnz > 0)
            upload_host(v.get(), X->data + X->indptr[0], (size_t)nnz * 4, s, current_device());
        hip_check(hipStreamSynchronize(s), "upload rows");
    }
};

}  // namespace scl_amd
