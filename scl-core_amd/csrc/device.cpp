// device.cpp -- device runtime of the library: device discovery, the
// library's streams, grow-only staging buffers, and the device copies of
// sparse handles (int32 indices, float values, optional transposed pattern).
#include "internal.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace scl_amd {

namespace {

struct DeviceState {
    hipStream_t stream = nullptr;
    std::mutex call_mu;   // serialises host-pointer entry points
    std::mutex ws_mu;
    std::vector<std::pair<void*, size_t>> ws;
    // pinned bounce buffers of the host copies (two, kBounceBytes each) and
    // the events that say when the copy engine is done with each
    std::mutex bounce_mu;
    void* bounce[2] = {nullptr, nullptr};
    hipEvent_t bounce_done[2] = {nullptr, nullptr};
};

// Host copies at least this large go through the pinned bounce buffers, in
// chunks of a quarter of the copy (at least kBounceChunkMin, at most
// kBounceBytes) so that the host copy of one chunk overlaps the transfer of
// the previous one; the runtime's own pageable path is as fast only for
// small copies.
constexpr size_t kBounceMin = 8u << 20;
constexpr size_t kBounceChunkMin = 4u << 20;
constexpr size_t kBounceBytes = 64u << 20;
constexpr int kCopyThreads = 8;

size_t bounce_chunk(size_t bytes) {
    const size_t quarter = (bytes / 4 + 4095) / 4096 * 4096;
    return std::min(kBounceBytes, std::max(kBounceChunkMin, quarter));
}

// dst = src on up to kCopyThreads host threads (chunks of 1 MiB or more)
void parallel_memcpy(void* dst, const void* src, size_t bytes) {
    const size_t per = std::max<size_t>(1u << 20, (bytes + kCopyThreads - 1) / kCopyThreads);
    const int n = (int)std::min<size_t>(kCopyThreads, (bytes + per - 1) / per);
    if (n <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)n - 1);
    for (int i = 1; i < n; ++i) {
        const size_t off = per * (size_t)i, len = std::min(per, bytes - off);
        th.emplace_back([=] {
            std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len);
        });
    }
    std::memcpy(dst, src, std::min(per, bytes));
    for (auto& t : th) t.join();
}

void ensure_bounce(DeviceState& st) {
    for (int b = 0; b < 2; ++b) {
        if (!st.bounce[b])
            hip_check(hipHostMalloc(&st.bounce[b], kBounceBytes, hipHostMallocDefault),
                      "pinned staging buffer");
        if (!st.bounce_done[b])
            hip_check(hipEventCreateWithFlags(&st.bounce_done[b], hipEventDisableTiming),
                      "staging event");
    }
}

std::mutex g_mu;
// Never destroyed: device memory and streams are reclaimed by the runtime at
// process exit, and destroying them from a static destructor can run after
// the HIP runtime itself has been torn down.
std::map<int, DeviceState*>* g_dev = new std::map<int, DeviceState*>();

DeviceState& state(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_dev->find(device);
    if (it != g_dev->end()) return *it->second;
    auto* st = new DeviceState();
    (*g_dev)[device] = st;
    return *st;
}

template <typename T>
T* dev_alloc(size_t count, const char* what) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hip_check(hipMalloc(&p, count * sizeof(T)), what);
    return static_cast<T*>(p);
}

void dev_free(void* p) {
    // a failed free must not leave its error for the caller's next HIP call
    if (p && hipFree(p) != hipSuccess) (void)hipGetLastError();
}

}  // namespace

int current_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        fail(SCL_ERROR_FEATURE_UNAVAILABLE, "No HIP device available");
    }
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    return d;
}

hipStream_t library_stream(int device) {
    DeviceState& st = state(device);
    std::lock_guard<std::mutex> lk(st.ws_mu);
    if (!st.stream)
        hip_check(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking),
                  "hipStreamCreate");
    return st.stream;
}

std::mutex& device_mutex(int device) { return state(device).call_mu; }

void upload_host(void* dst, const void* src, size_t bytes, hipStream_t s, int device) {
    if (bytes == 0) return;
    if (bytes < kBounceMin) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "upload");
        return;
    }
    DeviceState& st = state(device);
    std::lock_guard<std::mutex> lk(st.bounce_mu);
    ensure_bounce(st);
    // chunk i: wait until buffer i % 2 is free, fill it on the host, queue
    // its copy; the copy engine moves one buffer while the host fills the
    // other
    const size_t chunk = bounce_chunk(bytes);
    for (size_t off = 0, i = 0; off < bytes; off += chunk, ++i) {
        const int b = (int)(i & 1);
        const size_t len = std::min(chunk, bytes - off);
        hip_check(hipEventSynchronize(st.bounce_done[b]), "staging wait");
        parallel_memcpy(st.bounce[b], static_cast<const char*>(src) + off, len);
        hip_check(hipMemcpyAsync(static_cast<char*>(dst) + off, st.bounce[b], len,
                                 hipMemcpyHostToDevice, s),
                  "upload");
        hip_check(hipEventRecord(st.bounce_done[b], s), "staging event");
    }
    // the buffers stay busy until their events: later users wait on them
}

void download_host(void* dst, const void* src, size_t bytes, hipStream_t s, int device) {
    if (bytes == 0) return;
    if (bytes < kBounceMin) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "download");
        hip_check(hipStreamSynchronize(s), "download");
        return;
    }
    DeviceState& st = state(device);
    std::lock_guard<std::mutex> lk(st.bounce_mu);
    ensure_bounce(st);
    // chunk i is copied into buffer i % 2 by the copy engine while the host
    // empties chunk i - 1 from the other buffer
    size_t prev_off = 0, prev_len = 0;
    int prev_b = -1;
    const size_t chunk = bounce_chunk(bytes);
    for (size_t off = 0, i = 0; off < bytes; off += chunk, ++i) {
        const int b = (int)(i & 1);
        const size_t len = std::min(chunk, bytes - off);
        hip_check(hipEventSynchronize(st.bounce_done[b]), "staging wait");
        hip_check(hipMemcpyAsync(st.bounce[b], static_cast<const char*>(src) + off, len,
                                 hipMemcpyDeviceToHost, s),
                  "download");
        hip_check(hipEventRecord(st.bounce_done[b], s), "staging event");
        if (prev_b >= 0) {
            hip_check(hipEventSynchronize(st.bounce_done[prev_b]), "download");
            parallel_memcpy(static_cast<char*>(dst) + prev_off, st.bounce[prev_b], prev_len);
        }
        prev_off = off;
        prev_len = len;
        prev_b = b;
    }
    hip_check(hipEventSynchronize(st.bounce_done[prev_b]), "download");
    parallel_memcpy(static_cast<char*>(dst) + prev_off, st.bounce[prev_b], prev_len);
}

void* device_workspace(int device, int slot, size_t bytes) {
    DeviceState& st = state(device);
    std::lock_guard<std::mutex> lk(st.ws_mu);
    if ((int)st.ws.size() <= slot) st.ws.resize(slot + 1, {nullptr, 0});
    auto& w = st.ws[slot];
    if (w.second < bytes) {
        dev_free(w.first);
        w = {nullptr, 0};
        void* p = nullptr;
        hip_check(hipMalloc(&p, bytes), "workspace allocation");
        w = {p, bytes};
    }
    return w.first;
}

void ensure_merge_scratch(DeviceCSR& d, int64_t waves, int64_t F) {
    const int64_t floats = waves * F;
    if (waves <= d.merge.waves && floats <= d.merge.floats) return;
    if (d.pins > 0)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "a diffusion plan holds this operator: merge-path scratch cannot grow");
    const int64_t w = std::max(waves, d.merge.waves);
    const int64_t f = std::max(floats, d.merge.floats);
    dev_free(d.merge.split);
    dev_free(d.merge.carry_row);
    dev_free(d.merge.carry_val);
    d.merge = MergeScratch{};
    void* a = nullptr;
    void* b = nullptr;
    void* c = nullptr;
    if (hipMalloc(&a, (size_t)(2 * (w + 1)) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b, (size_t)w * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c, (size_t)std::max<int64_t>(f, 1) * sizeof(float)) != hipSuccess) {
        dev_free(a);
        dev_free(b);
        dev_free(c);
        (void)hipGetLastError();
        fail(SCL_ERROR_OUT_OF_MEMORY, "merge-path scratch allocation failed");
    }
    d.merge.split = static_cast<int32_t*>(a);
    d.merge.carry_row = static_cast<int32_t*>(b);
    d.merge.carry_val = static_cast<float*>(c);
    d.merge.waves = w;
    d.merge.floats = f;
}

DeviceCSR::~DeviceCSR() {
    int cur = -1;
    if (device >= 0 && hipGetDevice(&cur) == hipSuccess && cur != device)
        (void)hipSetDevice(device);
    dev_free(row_ptr);
    dev_free(col);
    dev_free(val);
    dev_free(sched);
    dev_free(val_bf16);
    dev_free(val_comp.hi);  // lo points into the same allocation
    dev_free(merge.split);
    dev_free(merge.carry_row);
    dev_free(merge.carry_val);
    release_grouped();
    release_transpose();
    release_renumbered();
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
}

void DeviceCSR::release_renumbered() {
    dev_free(ren.row_ptr);
    dev_free(ren.col);
    dev_free(ren.val);
    dev_free(ren.src);
    ren.row_ptr = ren.col = ren.src = nullptr;
    ren.val = nullptr;
    ren.fingerprint = 0;
}

void DeviceCSR::release_grouped() {
    for (GroupedCsr& g : grouped) {
        dev_free(g.gptr);
        dev_free(g.gword);
        dev_free(g.gvptr);
        dev_free(g.gval);
        dev_free(g.gperm);
        dev_free(g.grow);
        g = GroupedCsr{};
    }
    dev_free(staged.gptr);
    dev_free(staged.ucol);
    dev_free(staged.upos);
    dev_free(staged.grow);
    staged = StagedCsr{};
    for (TiledCsr& t : tiled) {
        dev_free(t.info);
        dev_free(t.ucol);
        dev_free(t.trow);
        dev_free(t.trbeg);
        dev_free(t.trlen);
        dev_free(t.tslot);
        dev_free(t.tval);
        dev_free(t.tperm);
        t = TiledCsr{};
    }
    // shapes that read the grouped form go with it
    for (auto* m : {&shapes, &t_shapes})
        for (auto it = m->begin(); it != m->end();)
            it = it->second.rows_per_group ? m->erase(it) : std::next(it);
}

void DeviceCSR::release_transpose() {
    dev_free(t_row_ptr);
    dev_free(t_col);
    dev_free(t_val);
    dev_free(t_src);
    dev_free(t_sched);
    dev_free(t_dst);
    t_row_ptr = t_col = t_src = t_dst = t_sched = nullptr;
    t_val = nullptr;
    t_generation = ~uint64_t(0);
}

namespace {

// FNV-1a over the pattern: decides whether a refreshed copy can keep its
// schedule (values may change without the pattern changing).
template <typename I>
uint64_t pattern_fingerprint(int64_t n, const I* ptr, const I* idx) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ULL;
    };
    mix((uint64_t)n);
    for (int64_t i = 0; i <= n; ++i) mix((uint64_t)ptr[i]);
    for (int64_t k = 0; k < (int64_t)ptr[n]; ++k) mix((uint64_t)idx[k]);
    return h;
}

int32_t* upload_schedule(const std::vector<int32_t>& order) {
    int32_t* d = dev_alloc<int32_t>(order.size(), "schedule");
    hip_check(hipMemcpy(d, order.data(), order.size() * 4, hipMemcpyHostToDevice),
              "upload schedule");
    return d;
}

}  // namespace

DeviceCSR& device_csr(scl_sparse_matrix* h, int device) {
    auto& slot = h->dev[device];
    if (slot && slot->generation == h->generation) return *slot;
    const int64_t p = h->primary(), q = h->secondary(), nnz = h->nnz;
    const int64_t lim = std::numeric_limits<int32_t>::max();
    if (p >= lim || q >= lim || nnz >= lim)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "operator too large for int32 device indices (dims and nnz must be < 2^31-1)");
    const bool same_shape = slot && slot->n == p && slot->n_sec == q && slot->nnz == nnz;
    if (slot && slot->pins > 0 &&
        (!same_shape || (p >= kScheduleMinRows &&
                         pattern_fingerprint(p, h->indptr, h->indices) != slot->fingerprint)))
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "the operator's pattern changed while a captured plan uses it");
    if (!same_shape) {
        slot.reset(new DeviceCSR());
        slot->device = device;
        slot->n = (int32_t)p;
        slot->n_sec = (int32_t)q;
        slot->nnz = nnz;
        slot->row_ptr = dev_alloc<int32_t>((size_t)p + 1, "operator row_ptr");
        slot->col = dev_alloc<int32_t>((size_t)nnz, "operator col");
        slot->val = dev_alloc<float>((size_t)nnz, "operator val");
    }
    std::vector<int32_t> rp((size_t)p + 1), ci((size_t)nnz);
    for (int64_t i = 0; i <= p; ++i) rp[i] = (int32_t)h->indptr[i];
    for (int64_t k = 0; k < nnz; ++k) ci[k] = (int32_t)h->indices[k];
    hip_check(hipMemcpy(slot->row_ptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice),
              "upload row_ptr");
    if (nnz > 0) {
        hip_check(hipMemcpy(slot->col, ci.data(), ci.size() * 4, hipMemcpyHostToDevice),
                  "upload col");
        hip_check(hipMemcpy(slot->val, h->data, (size_t)nnz * 4, hipMemcpyHostToDevice),
                  "upload val");
    }
    if (p >= kScheduleMinRows) {
        const uint64_t fp = pattern_fingerprint(p, h->indptr, h->indices);
        if (!slot->sched || slot->fingerprint != fp) {
            dev_free(slot->sched);
            slot->sched = nullptr;
            slot->sched = upload_schedule(locality_schedule(p, h->indptr, h->indices));
            slot->fingerprint = fp;
            slot->shapes.clear();
            slot->t_shapes.clear();
            slot->adj_shapes.clear();
        }
    }
    bool any_tiled = false;
    for (const TiledCsr& t : slot->tiled) any_tiled = any_tiled || t.lpr;
    if (slot->grouped[0].R || slot->grouped[1].R || slot->staged.built || any_tiled) {
        // same pattern: re-gather the grouped values (the staged form holds
        // no values); else drop the forms
        const uint64_t fp = pattern_fingerprint(p, h->indptr, h->indices);
        bool keep = same_shape;
        for (const GroupedCsr& g : slot->grouped)
            if (g.R && g.fingerprint != fp) keep = false;
        if (slot->staged.built && slot->staged.fingerprint != fp) keep = false;
        for (const TiledCsr& t : slot->tiled)
            if (t.lpr && t.fingerprint != fp) keep = false;
        if (keep) {
            for (GroupedCsr& g : slot->grouped)
                if (g.R) launch_gather_f32(g.gperm, slot->val, g.gval, nnz, nullptr);
            for (TiledCsr& t : slot->tiled)
                if (t.lpr) launch_gather_pad_f32(t.tperm, slot->val, t.tval, t.n_slots, nullptr);
            hip_check(hipStreamSynchronize(nullptr), "refresh grouped values");
        } else {
            if (slot->pins > 0)
                fail(SCL_ERROR_INVALID_ARGUMENT,
                     "the operator's pattern changed while a captured plan uses it");
            slot->release_grouped();
        }
    }
    if (slot->ren.row_ptr) {
        // the renumbered copy follows the schedule's pattern
        if (same_shape && slot->sched && slot->ren.fingerprint == slot->fingerprint) {
            launch_gather_f32(slot->ren.src, slot->val, slot->ren.val, nnz, nullptr);
            hip_check(hipStreamSynchronize(nullptr), "refresh renumbered values");
        } else {
            slot->release_renumbered();
        }
    }
    if (slot->pins > 0 && slot->t_row_ptr) {
        // a plan holds the transposed arrays: refresh their values in place
        launch_gather_f32(slot->t_src, slot->val, slot->t_val, nnz, nullptr);
        hip_check(hipStreamSynchronize(nullptr), "refresh transposed values");
        slot->t_generation = h->generation;
    } else {
        slot->release_transpose();
    }
    slot->generation = h->generation;
    return *slot;
}

DeviceCSR* device_csr_current(scl_sparse_matrix* h, int device) {
    const auto it = h->dev.find(device);
    if (it == h->dev.end() || !it->second || it->second->generation != h->generation)
        return nullptr;
    return it->second.get();
}

void ensure_transpose(scl_sparse_matrix* h, DeviceCSR& d, hipStream_t s) {
    (void)s;
    if (d.t_row_ptr && d.t_generation == d.generation) return;
    d.release_transpose();
    const int64_t p = h->primary(), q = h->secondary(), nnz = h->nnz;
    // counting sort by column; visiting source rows in order keeps each
    // transposed row sorted by source row
    std::vector<int32_t> trp((size_t)q + 1, 0), tcol((size_t)nnz), tsrc((size_t)nnz),
        tdst((size_t)nnz);
    std::vector<float> tval((size_t)nnz);
    for (int64_t k = 0; k < nnz; ++k) trp[h->indices[k] + 1]++;
    for (int64_t j = 0; j < q; ++j) trp[j + 1] += trp[j];
    std::vector<int32_t> pos(trp.begin(), trp.end() - 1);
    for (int64_t i = 0; i < p; ++i)
        for (int64_t k = h->indptr[i]; k < h->indptr[i + 1]; ++k) {
            const int32_t dst = pos[h->indices[k]]++;
            tcol[dst] = (int32_t)i;
            tval[dst] = h->data[k];
            tsrc[dst] = (int32_t)k;
            tdst[(size_t)k] = dst;
        }
    d.t_row_ptr = dev_alloc<int32_t>((size_t)q + 1, "transpose row_ptr");
    d.t_col = dev_alloc<int32_t>((size_t)nnz + kTransposePad, "transpose col");
    d.t_val = dev_alloc<float>((size_t)nnz + kTransposePad, "transpose val");
    hip_check(hipMemset(d.t_col + nnz, 0, kTransposePad * sizeof(int32_t)), "transpose pad");
    hip_check(hipMemset(d.t_val + nnz, 0, kTransposePad * sizeof(float)), "transpose pad");
    d.t_src = dev_alloc<int32_t>((size_t)nnz, "transpose src");
    d.t_dst = dev_alloc<int32_t>((size_t)nnz, "transpose dst");
    hip_check(hipMemcpy(d.t_row_ptr, trp.data(), trp.size() * 4, hipMemcpyHostToDevice),
              "upload transpose");
    if (nnz > 0) {
        hip_check(hipMemcpy(d.t_col, tcol.data(), tcol.size() * 4, hipMemcpyHostToDevice),
                  "upload transpose");
        hip_check(hipMemcpy(d.t_val, tval.data(), tval.size() * 4, hipMemcpyHostToDevice),
                  "upload transpose");
        hip_check(hipMemcpy(d.t_src, tsrc.data(), tsrc.size() * 4, hipMemcpyHostToDevice),
                  "upload transpose");
        hip_check(hipMemcpy(d.t_dst, tdst.data(), tdst.size() * 4, hipMemcpyHostToDevice),
                  "upload transpose");
    }
    if (q >= kScheduleMinRows) d.t_sched = upload_schedule(locality_schedule(q, trp.data(), tcol.data()));
    d.t_generation = d.generation;
}


bool ensure_staged(scl_sparse_matrix* h, DeviceCSR& d) {
    StagedCsr& g = d.staged;
    const int64_t p = h->primary(), q = h->secondary(), nnz = h->nnz;
    const uint64_t fp = pattern_fingerprint(p, h->indptr, h->indices);
    if (g.built && g.fingerprint == fp) return true;
    for (int64_t i = 0; i < p; ++i)
        for (int64_t k = h->indptr[i] + 1; k < h->indptr[i + 1]; ++k)
            if (h->indices[k] <= h->indices[k - 1]) return false;
    if (d.pins > 0)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "a diffusion plan holds this operator: its staged form cannot be rebuilt");
    // groups follow the device's visiting order (the locality schedule)
    std::vector<int32_t> order((size_t)p);
    if (d.sched) {
        hip_check(hipMemcpy(order.data(), d.sched, (size_t)p * 4, hipMemcpyDeviceToHost),
                  "download schedule");
    } else {
        for (int64_t i = 0; i < p; ++i) order[i] = (int32_t)i;
    }
    const int64_t R = kStagedRows, n_groups = (p + R - 1) / R;
    std::vector<int32_t> gptr((size_t)n_groups + 1, 0), grow((size_t)n_groups * R, -1);
    std::vector<int32_t> ucol, upos((size_t)std::max<int64_t>(nnz, 1), 0);
    std::vector<int32_t> place((size_t)std::max<int64_t>(q, 1), -1);
    std::vector<int32_t> cols;
    ucol.reserve((size_t)nnz / 2 + 16);
    for (int64_t gi = 0; gi < n_groups; ++gi) {
        cols.clear();
        for (int64_t r = 0; r < R && gi * R + r < p; ++r) {
            const int32_t row = order[gi * R + r];
            grow[gi * R + r] = row;
            for (int64_t k = h->indptr[row]; k < h->indptr[row + 1]; ++k)
                cols.push_back((int32_t)h->indices[k]);
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        for (size_t u = 0; u < cols.size(); ++u) place[cols[u]] = (int32_t)u;
        for (int64_t r = 0; r < R && gi * R + r < p; ++r) {
            const int32_t row = order[gi * R + r];
            for (int64_t k = h->indptr[row]; k < h->indptr[row + 1]; ++k)
                upos[k] = place[h->indices[k]];
        }
        for (int32_t c : cols) place[c] = -1;
        ucol.insert(ucol.end(), cols.begin(), cols.end());
        if (ucol.size() > (size_t)INT32_MAX) return false;
        gptr[gi + 1] = (int32_t)ucol.size();
    }
    dev_free(g.gptr);
    dev_free(g.ucol);
    dev_free(g.upos);
    dev_free(g.grow);
    g = StagedCsr{};
    g.gptr = dev_alloc<int32_t>(gptr.size(), "staged gptr");
    g.grow = dev_alloc<int32_t>(grow.size(), "staged rows");
    g.ucol = dev_alloc<int32_t>(std::max<size_t>(ucol.size(), 1), "staged columns");
    g.upos = dev_alloc<int32_t>(upos.size(), "staged places");
    hip_check(hipMemcpy(g.gptr, gptr.data(), gptr.size() * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(g.grow, grow.data(), grow.size() * 4, hipMemcpyHostToDevice), "upload");
    if (!ucol.empty())
        hip_check(hipMemcpy(g.ucol, ucol.data(), ucol.size() * 4, hipMemcpyHostToDevice),
                  "upload");
    hip_check(hipMemcpy(g.upos, upos.data(), upos.size() * 4, hipMemcpyHostToDevice), "upload");
    g.n_groups = n_groups;
    g.n_union = (int64_t)ucol.size() - 1;
    g.fingerprint = fp;
    g.built = true;
    return true;
}

bool ensure_renumbered(scl_sparse_matrix* h, DeviceCSR& d) {
    if (!d.sched || d.n != d.n_sec) return false;
    if (d.ren.row_ptr && d.ren.fingerprint == d.fingerprint) return true;
    d.release_renumbered();
    const int64_t n = d.n, nnz = d.nnz;
    std::vector<int32_t> order((size_t)n), inv((size_t)n);
    hip_check(hipMemcpy(order.data(), d.sched, (size_t)n * 4, hipMemcpyDeviceToHost),
              "download schedule");
    for (int64_t i = 0; i < n; ++i) inv[order[i]] = (int32_t)i;
    std::vector<int32_t> rp((size_t)n + 1), ci((size_t)std::max<int64_t>(nnz, 1)),
        src((size_t)std::max<int64_t>(nnz, 1));
    rp[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t r = order[i], b = h->indptr[r], e = h->indptr[r + 1];
        int32_t k2 = rp[i];
        for (int64_t k = b; k < e; ++k, ++k2) {
            ci[k2] = inv[h->indices[k]];
            src[k2] = (int32_t)k;
        }
        rp[i + 1] = k2;
    }
    d.ren.row_ptr = dev_alloc<int32_t>((size_t)n + 1, "renumbered row_ptr");
    d.ren.col = dev_alloc<int32_t>(ci.size(), "renumbered col");
    d.ren.val = dev_alloc<float>(ci.size(), "renumbered val");
    d.ren.src = dev_alloc<int32_t>(src.size(), "renumbered entry map");
    hip_check(hipMemcpy(d.ren.row_ptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice),
              "upload");
    hip_check(hipMemcpy(d.ren.col, ci.data(), ci.size() * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d.ren.src, src.data(), src.size() * 4, hipMemcpyHostToDevice),
              "upload");
    launch_gather_f32(d.ren.src, d.val, d.ren.val, nnz, nullptr);
    hip_check(hipStreamSynchronize(nullptr), "renumbered values");
    d.ren.fingerprint = d.fingerprint;
    return true;
}

void refresh_grouped_values(DeviceCSR& d, hipStream_t s) {
    if (d.ren.row_ptr) launch_gather_f32(d.ren.src, d.val, d.ren.val, d.nnz, s);
    for (GroupedCsr& g : d.grouped)
        if (g.R) launch_gather_f32(g.gperm, d.val, g.gval, d.nnz, s);
    for (TiledCsr& t : d.tiled)
        if (t.lpr) launch_gather_pad_f32(t.tperm, d.val, t.tval, t.n_slots, s);
}

bool ensure_tiled(scl_sparse_matrix* h, DeviceCSR& d, int lpr, int budget) {
    if (!(lpr == 4 || lpr == 8) || !(budget == 1 || budget == 2)) return false;
    TiledCsr& g = d.tiled[tiled_index(lpr, budget)];
    const int64_t p = h->primary(), q = h->secondary();
    const uint64_t fp = pattern_fingerprint(p, h->indptr, h->indices);
    if (g.lpr == lpr && g.budget == budget && g.fingerprint == fp) return true;
    const int64_t smax = tiled_union_max(lpr, budget);
    // padding entries read the last LDS slot, which holds zeros
    const uint16_t zero_unit = (uint16_t)((tiled_slot_capacity(lpr, budget) - 1) * lpr);
    // a tile's padded entries, rounded up to a multiple of 8 at its end
    const int64_t emax = tiled_entries_max(budget) - 4;
    const int64_t rmax = kTiledThreads / lpr;
    for (int64_t i = 0; i < p; ++i) {
        const int64_t len = h->indptr[i + 1] - h->indptr[i];
        if (len > smax || ((len + 3) & ~int64_t(3)) > emax) return false;
        for (int64_t k = h->indptr[i] + 1; k < h->indptr[i + 1]; ++k)
            if (h->indices[k] <= h->indices[k - 1]) return false;
    }
    if (d.pins > 0)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "a diffusion plan holds this operator: its tiled form cannot be rebuilt");
    // tiles follow the device's visiting order (the locality schedule)
    std::vector<int32_t> order((size_t)p);
    if (d.sched) {
        hip_check(hipMemcpy(order.data(), d.sched, (size_t)p * 4, hipMemcpyDeviceToHost),
                  "download schedule");
    } else {
        for (int64_t i = 0; i < p; ++i) order[i] = (int32_t)i;
    }
    // greedy cut: a row joins the current tile while rows, union slots and
    // padded entries all fit
    std::vector<int32_t> stamp((size_t)std::max<int64_t>(q, 1), -1);
    std::vector<int64_t> cut{0};
    {
        int64_t nu = 0, ne = 0, nr = 0;
        int32_t t = 0;
        for (int64_t i = 0; i < p; ++i) {
            const int32_t row = order[i];
            const int64_t b = h->indptr[row], e = h->indptr[row + 1];
            const int64_t pad = (e - b + 3) & ~int64_t(3);
            int64_t fresh = 0;
            for (int64_t k = b; k < e; ++k) fresh += stamp[h->indices[k]] != t;
            if (nr > 0 && (nr + 1 > rmax || nu + fresh > smax || ne + pad > emax)) {
                cut.push_back(i);
                ++t;
                nu = ne = nr = 0;
                fresh = e - b;
            }
            for (int64_t k = b; k < e; ++k) stamp[h->indices[k]] = t;
            nu += fresh;
            ne += pad;
            ++nr;
        }
        if (p > 0) cut.push_back(p);
    }
    const int64_t n_tiles = (int64_t)cut.size() - 1;
    std::vector<int32_t> info((size_t)(n_tiles + 1) * 4, 0), ucol, trow((size_t)p),
        trbeg((size_t)p), trlen((size_t)p), tperm;
    std::vector<uint16_t> tslot;
    std::vector<int32_t> rows, cols;
    std::vector<int32_t> place((size_t)std::max<int64_t>(q, 1), -1);
    ucol.reserve((size_t)h->nnz / 2 + 16);
    tslot.reserve((size_t)h->nnz + (size_t)p * 3);
    tperm.reserve((size_t)h->nnz + (size_t)p * 3);
    for (int64_t t = 0; t < n_tiles; ++t) {
        info[t * 4 + 0] = (int32_t)ucol.size();
        info[t * 4 + 2] = (int32_t)tslot.size();
        info[t * 4 + 3] = (int32_t)cut[t];
        rows.assign(order.begin() + cut[t], order.begin() + cut[t + 1]);
        // longest rows first, so the rows of one wave have similar lengths
        std::stable_sort(rows.begin(), rows.end(), [&](int32_t a, int32_t b) {
            return h->indptr[a + 1] - h->indptr[a] > h->indptr[b + 1] - h->indptr[b];
        });
        cols.clear();
        for (int32_t r : rows)
            for (int64_t k = h->indptr[r]; k < h->indptr[r + 1]; ++k)
                cols.push_back((int32_t)h->indices[k]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        for (size_t u = 0; u < cols.size(); ++u) place[cols[u]] = (int32_t)u;
        info[t * 4 + 1] = (int32_t)cols.size();
        const size_t e0 = tslot.size();
        for (size_t j = 0; j < rows.size(); ++j) {
            const int32_t r = rows[j];
            const int64_t at = cut[t] + (int64_t)j;
            trow[at] = r;
            trbeg[at] = (int32_t)(tslot.size() - e0);
            trlen[at] = (int32_t)(h->indptr[r + 1] - h->indptr[r]);
            for (int64_t k = h->indptr[r]; k < h->indptr[r + 1]; ++k) {
                tslot.push_back((uint16_t)(place[h->indices[k]] * lpr));
                tperm.push_back((int32_t)k);
            }
            while ((tslot.size() - e0) % 4) {
                tslot.push_back(zero_unit);
                tperm.push_back(-1);
            }
        }
        while ((tslot.size() - e0) % 8) {
            tslot.push_back(zero_unit);
            tperm.push_back(-1);
        }
        for (int32_t c : cols) place[c] = -1;
        ucol.insert(ucol.end(), cols.begin(), cols.end());
        if (ucol.size() > (size_t)INT32_MAX || tslot.size() > (size_t)INT32_MAX) return false;
    }
    info[n_tiles * 4 + 0] = (int32_t)ucol.size();
    info[n_tiles * 4 + 2] = (int32_t)tslot.size();
    info[n_tiles * 4 + 3] = (int32_t)p;
    for (TiledCsr* t : {&g}) {
        dev_free(t->info);
        dev_free(t->ucol);
        dev_free(t->trow);
        dev_free(t->trbeg);
        dev_free(t->trlen);
        dev_free(t->tslot);
        dev_free(t->tval);
        dev_free(t->tperm);
        *t = TiledCsr{};
    }
    const size_t ns = std::max<size_t>(tslot.size(), 8);
    g.info = dev_alloc<int32_t>(info.size(), "tiled info");
    ucol.push_back(0);  // spare entry: the kernel's clamped reads stay inside
    g.ucol = dev_alloc<int32_t>(ucol.size(), "tiled columns");
    g.trow = dev_alloc<int32_t>(std::max<size_t>(trow.size(), 1), "tiled rows");
    g.trbeg = dev_alloc<int32_t>(std::max<size_t>(trbeg.size(), 1), "tiled row starts");
    g.trlen = dev_alloc<int32_t>(std::max<size_t>(trlen.size(), 1), "tiled row lengths");
    g.tslot = dev_alloc<uint16_t>(ns, "tiled slots");
    g.tval = dev_alloc<float>(ns, "tiled values");
    g.tperm = dev_alloc<int32_t>(ns, "tiled permutation");
    auto up = [](void* dst, const void* src, size_t bytes) {
        if (bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "upload tiles");
    };
    up(g.info, info.data(), info.size() * 4);
    up(g.ucol, ucol.data(), ucol.size() * 4);
    up(g.trow, trow.data(), trow.size() * 4);
    up(g.trbeg, trbeg.data(), trbeg.size() * 4);
    up(g.trlen, trlen.data(), trlen.size() * 4);
    up(g.tslot, tslot.data(), tslot.size() * 2);
    up(g.tperm, tperm.data(), tperm.size() * 4);
    g.n_tiles = n_tiles;
    g.n_union = (int64_t)ucol.size() - 1;
    g.n_slots = (int64_t)tslot.size();
    launch_gather_pad_f32(g.tperm, d.val, g.tval, g.n_slots, nullptr);
    hip_check(hipStreamSynchronize(nullptr), "tiled values");
    g.lpr = lpr;
    g.budget = budget;
    g.fingerprint = fp;
    return true;
}

bool ensure_grouped(scl_sparse_matrix* h, DeviceCSR& d, int R) {
    GroupedCsr& g = d.grouped[R == 8 ? 1 : 0];
    const int64_t p = h->primary(), q = h->secondary(), nnz = h->nnz;
    const uint64_t fp = pattern_fingerprint(p, h->indptr, h->indices);
    if (g.R == R && g.fingerprint == fp) return true;
    if (q >= (int64_t)1 << 24) return false;
    for (int64_t i = 0; i < p; ++i)
        for (int64_t k = h->indptr[i] + 1; k < h->indptr[i + 1]; ++k)
            if (h->indices[k] <= h->indices[k - 1]) return false;
    if (d.pins > 0)
        fail(SCL_ERROR_INVALID_ARGUMENT,
             "a diffusion plan holds this operator: its grouped form cannot be rebuilt");
    // groups follow the device's visiting order (the locality schedule);
    // on square operators each group gathers the neighbours that share the
    // most columns (fewer distinct columns per group, so fewer gathers)
    std::vector<int32_t> order((size_t)p);
    if (d.sched) {
        hip_check(hipMemcpy(order.data(), d.sched, (size_t)p * 4, hipMemcpyDeviceToHost),
                  "download schedule");
    } else {
        for (int64_t i = 0; i < p; ++i) order[i] = (int32_t)i;
    }
    if (p == q) order = neighbour_groups(p, h->indptr, h->indices, order, R);
    const int64_t n_groups = (p + R - 1) / R;
    std::vector<int32_t> gptr((size_t)n_groups + 1, 0), gvptr((size_t)n_groups + 1, 0);
    std::vector<int32_t> grow((size_t)n_groups * R, -1), gperm;
    std::vector<uint32_t> gword;
    gperm.reserve((size_t)nnz);
    gword.reserve((size_t)nnz);
    // (column, slot, position) of the group's nonzeros, ordered by column then
    // slot: one entry per distinct column, values in slot order
    std::vector<std::array<int64_t, 3>> items;
    for (int64_t gi = 0; gi < n_groups; ++gi) {
        items.clear();
        for (int r = 0; r < R; ++r) {
            const int64_t rr = gi * R + r;
            if (rr >= p) break;
            const int32_t row = order[rr];
            grow[gi * R + r] = row;
            for (int64_t k = h->indptr[row]; k < h->indptr[row + 1]; ++k)
                items.push_back({(int64_t)h->indices[k], r, k});
        }
        std::sort(items.begin(), items.end());
        for (size_t a = 0; a < items.size();) {
            uint32_t mask = 0;
            size_t b = a;
            for (; b < items.size() && items[b][0] == items[a][0]; ++b) {
                mask |= 1u << items[b][1];
                gperm.push_back((int32_t)items[b][2]);
            }
            gword.push_back((uint32_t)items[a][0] | (mask << 24));
            a = b;
        }
        gptr[gi + 1] = (int32_t)gword.size();
        gvptr[gi + 1] = (int32_t)gperm.size();
    }
    // replace the old form (if any) by the new one
    dev_free(g.gptr);
    dev_free(g.gword);
    dev_free(g.gvptr);
    dev_free(g.gval);
    dev_free(g.gperm);
    dev_free(g.grow);
    g = GroupedCsr{};
    const int64_t ne = (int64_t)gword.size();
    g.gptr = dev_alloc<int32_t>(gptr.size(), "grouped gptr");
    g.gvptr = dev_alloc<int32_t>(gvptr.size(), "grouped gvptr");
    g.grow = dev_alloc<int32_t>(grow.size(), "grouped rows");
    // padded by 32 zero words (column 0, no rows): a chunk of up to 16
    // entries and the next chunk's words load without clamping
    g.gword = dev_alloc<uint32_t>((size_t)ne + 32, "grouped words");
    g.gperm = dev_alloc<int32_t>((size_t)std::max<int64_t>(nnz, 1), "grouped perm");
    g.gval = dev_alloc<float>((size_t)nnz + 128, "grouped values");
    hip_check(hipMemcpy(g.gptr, gptr.data(), gptr.size() * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(g.gvptr, gvptr.data(), gvptr.size() * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(g.grow, grow.data(), grow.size() * 4, hipMemcpyHostToDevice), "upload");
    gword.resize((size_t)ne + 32, 0u);
    hip_check(hipMemcpy(g.gword, gword.data(), gword.size() * 4, hipMemcpyHostToDevice), "upload");
    if (nnz > 0)
        hip_check(hipMemcpy(g.gperm, gperm.data(), (size_t)nnz * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(g.gval, 0, ((size_t)nnz + 128) * sizeof(float)), "clear grouped values");
    launch_gather_f32(g.gperm, d.val, g.gval, nnz, nullptr);
    hip_check(hipStreamSynchronize(nullptr), "grouped values");
    g.R = R;
    g.n_groups = n_groups;
    g.n_entries = ne;
    g.fingerprint = fp;
    g.schedule = d.sched != nullptr;
    return true;
}

}  // namespace scl_amd
