// schedule.cpp -- row visiting order for the SpMM (a locality schedule).
//
// The kernels gather X[j, slab] for every nonzero (i, j).  When rows that
// share neighbours are processed at about the same time on the same XCD,
// those gathers hit that XCD's L2 instead of HBM.  The order is a
// Cuthill-McKee breadth-first traversal (neighbours visited by increasing
// degree, starting each component at a pseudo-peripheral node), then sorted
// by decreasing row length inside windows of 64 consecutive rows so that the
// rows sharing one wavefront have similar trip counts.  Results do not
// depend on the order: every output row is computed independently.
#include "internal.hpp"

#include <algorithm>
#include <numeric>

namespace scl_amd {

namespace {

template <typename I>
int64_t bfs(int64_t start, const I* ptr, const I* idx, const std::vector<int64_t>& deg,
            std::vector<uint8_t>& seen, uint8_t mark, std::vector<int32_t>& out,
            std::vector<int32_t>& scratch) {
    // returns the last node reached; appends the visit order to out
    size_t head = out.size();
    out.push_back((int32_t)start);
    seen[start] = mark;
    int64_t last = start;
    while (head < out.size()) {
        const int32_t u = out[head++];
        last = u;
        scratch.clear();
        for (I k = ptr[u]; k < ptr[u + 1]; ++k) {
            const int64_t v = idx[k];
            // mark 2 = already placed in the order: never walked again, so
            // every node is traversed at most twice (probe + placement) even
            // when the pattern is not symmetric
            if (seen[v] != mark && seen[v] != 2) {
                seen[v] = mark;
                scratch.push_back((int32_t)v);
            }
        }
        std::stable_sort(scratch.begin(), scratch.end(),
                         [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
        out.insert(out.end(), scratch.begin(), scratch.end());
    }
    return last;
}

}  // namespace

template <typename I>
std::vector<int32_t> locality_schedule(int64_t n, const I* ptr, const I* idx) {
    std::vector<int64_t> deg((size_t)n);
    for (int64_t i = 0; i < n; ++i) deg[i] = (int64_t)(ptr[i + 1] - ptr[i]);
    std::vector<uint8_t> seen((size_t)n, 0);
    std::vector<int32_t> order, probe, scratch;
    order.reserve((size_t)n);
    for (int64_t s = 0; s < n; ++s) {
        if (seen[s] == 2) continue;
        // pseudo-peripheral start: the last node of a traversal from s
        probe.clear();
        const int64_t p = bfs(s, ptr, idx, deg, seen, 1, probe, scratch);
        // the probe marked its component 1; the real pass marks it 2
        bfs(p, ptr, idx, deg, seen, 2, order, scratch);
        // nodes of the probe the directed traversal from p did not reach
        for (int32_t v : probe)
            if (seen[v] != 2) bfs(v, ptr, idx, deg, seen, 2, order, scratch);
    }
    for (size_t w = 0; w < order.size(); w += 64) {
        auto b = order.begin() + w;
        auto e = order.begin() + std::min(order.size(), w + 64);
        std::stable_sort(b, e, [&](int32_t a, int32_t c) { return deg[a] > deg[c]; });
    }
    return order;
}

template std::vector<int32_t> locality_schedule<int64_t>(int64_t, const int64_t*, const int64_t*);
template std::vector<int32_t> locality_schedule<int32_t>(int64_t, const int32_t*, const int32_t*);

}  // namespace scl_amd
