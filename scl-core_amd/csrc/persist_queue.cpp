// persist_queue.cpp -- work queues of the persistent batched plan (see
// persist.hpp).  Plain host code: no HIP calls.
#include "persist.hpp"

#include <algorithm>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace scl_amd {

PersistQueues build_persist_queues(const std::vector<int64_t>& off,
                                   const std::vector<int64_t>& nz, int64_t t, bool backward,
                                   int64_t rows_per_item,
                                   int64_t graphs_per_chunk, int fwd_slabs) {
    const int64_t G = (int64_t)off.size() - 1;
    const int64_t R = rows_per_item > 0 ? rows_per_item : 16;
    const int64_t row_phases = backward ? 2 * t : t;
    const int64_t n_phase = row_phases;
    const int64_t S = std::max(fwd_slabs, 1);
    auto slabs_of = [&](int64_t p) -> int64_t { return p < t ? S : 1; };
    auto items_of = [&](int64_t g, int64_t p) -> int64_t {
        return slabs_of(p) * ((off[g + 1] - off[g] + R - 1) / R);
    };
    auto work = [&](int64_t g) { return (off[g + 1] - off[g]) + (nz[g + 1] - nz[g]); };
    std::vector<int64_t> order(G);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return work(x) > work(y); });
    std::vector<std::vector<int64_t>> bins(kPersistQueues);
    std::vector<int64_t> load(kPersistQueues, 0);
    for (int64_t g : order) {
        const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        bins[q].push_back(g);
        load[q] += work(g);
    }
    PersistQueues Q;
    Q.q_off.assign(kPersistLanes + 1, 0);
    Q.expect.assign(G, 0);
    for (int q = 0; q < kPersistQueues; ++q) {
        std::vector<PersistItem> list;
        auto& bin = bins[q];
        std::sort(bin.begin(), bin.end());
        const size_t K = (size_t)(graphs_per_chunk > 0 ? graphs_per_chunk : 4);
        for (size_t c0 = 0; c0 < bin.size(); c0 += K) {
            const size_t c1 = std::min(bin.size(), c0 + K);
            for (int64_t p = 0; p < n_phase; ++p)
                for (size_t ci = c0; ci < c1; ++ci) {
                    const int64_t g = bin[ci];
                    const int64_t m = items_of(g, p);
                    const int64_t lo = off[g], hi = off[g + 1];
                    const int64_t step = R;
                    const int64_t ns = slabs_of(p);
                    const int64_t per = m / ns;
                    for (int64_t sl = 0; sl < ns; ++sl)
                        for (int64_t i = 0; i < per; ++i) {
                            PersistItem it{};
                            it.g = (int32_t)g;
                            it.phase = (int32_t)p;
                            it.begin = (int32_t)(lo + i * step);
                            it.end = (int32_t)std::min(hi, lo + (i + 1) * step);
                            it.need = Q.expect[g];  // the graph's items of earlier phases
                            it.slab = (int32_t)sl;
                            list.push_back(it);
                        }
                    Q.expect[g] += (int32_t)m;
                }
        }
        // longest run of items of one phase: at most that many can run at once
        int64_t run = 0;
        for (size_t i = 0; i < list.size(); ++i) {
            run = (i > 0 && list[i].phase == list[i - 1].phase) ? run + 1 : 1;
            Q.widest = std::max(Q.widest, run);
        }
        for (int sub = 0; sub < kPersistSubQueues; ++sub) {
            for (size_t i = (size_t)sub; i < list.size(); i += kPersistSubQueues)
                Q.items.push_back(list[i]);
            if (Q.items.size() > (size_t)std::numeric_limits<int32_t>::max())
                throw std::length_error("Too many work items");
            Q.q_off[q * kPersistSubQueues + sub + 1] = (int32_t)Q.items.size();
        }
    }
    Q.dpad.assign(G, 0);
    for (int64_t g = 0; g < G; ++g) {
        Q.dpad[g] = (int32_t)(Q.dstride - nz[g]);
        Q.dstride += (nz[g + 1] - nz[g] + 31) / 32 * 32;
    }
    Q.dstride += 32;
    return Q;
}

}  // namespace scl_amd
