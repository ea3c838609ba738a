// schedule.cpp -- row visiting order for the SpMM (a locality schedule).
//
// The kernels gather X[j, slab] for every nonzero (i, j).  When rows that
// share neighbours are processed at about the same time on the same XCD,
// those gathers hit that XCD's L2 instead of going to HBM.  The order is
// built greedily: the next row is the unplaced row with the most pattern
// edges (in either direction) to the last kWindow placed rows, so every
// stretch of consecutive rows -- what one XCD has in flight at once -- is a
// tightly connected piece of the graph.  This is the window score of Gorder
// (Wei et al., SIGMOD 2016) without its common-neighbour term: on kNN graphs
// the edge term alone gives nearly all of the gain at O(nnz) cost.  Scores
// live in a bucket queue (one doubly linked list per score), so each
// increment, decrement and extract-max is O(1).  Results do not depend on
// the order: every output row is computed independently.
#include "internal.hpp"

#include <algorithm>

namespace scl_amd {

namespace {

// Rows placed within this distance of a candidate count towards its score.
// On knn30_200k (F = 2000, row groups of 4) windows of 64, 128, 256, 512,
// 1024 and 2048 rows gave 961, 992, 1012, 1019, 1014 and 974 GB/s: 512 is
// the best by a small, repeatable margin.
constexpr int64_t kWindow = 512;

// Unplaced rows bucketed by score.
class BucketQueue {
public:
    BucketQueue(int64_t n, int64_t max_score)
        : next_((size_t)n, -1), prev_((size_t)n, -1), score_((size_t)n, 0),
          head_((size_t)max_score + 1, -1) {
        // every row starts at score 0; pushed in reverse so that ties at
        // score 0 pop in increasing row order
        for (int64_t v = n - 1; v >= 0; --v) link((int32_t)v);
    }
    void bump(int32_t v, int d) {
        unlink(v);
        score_[v] += d;
        link(v);
    }
    void remove(int32_t v) { unlink(v); }
    int32_t pop_max() {
        while (top_ > 0 && head_[top_] < 0) --top_;
        const int32_t v = head_[top_];
        unlink(v);
        return v;
    }

private:
    void link(int32_t v) {
        const int64_t s = score_[v];
        prev_[v] = -1;
        next_[v] = head_[s];
        if (head_[s] >= 0) prev_[head_[s]] = v;
        head_[s] = v;
        top_ = std::max(top_, s);
    }
    void unlink(int32_t v) {
        const int64_t s = score_[v];
        if (prev_[v] >= 0) next_[prev_[v]] = next_[v];
        else head_[s] = next_[v];
        if (next_[v] >= 0) prev_[next_[v]] = prev_[v];
    }
    std::vector<int32_t> next_, prev_;
    std::vector<int64_t> score_;
    std::vector<int32_t> head_;
    int64_t top_ = 0;
};

}  // namespace

template <typename I>
std::vector<int32_t> locality_schedule(int64_t n, const I* ptr, const I* idx) {
    std::vector<int32_t> order;
    if (n <= 0) return order;
    order.reserve((size_t)n);
    // in-edges by a counting sort (the pattern need not be symmetric)
    const int64_t nnz = (int64_t)ptr[n];
    std::vector<int64_t> tptr((size_t)n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t j = (int64_t)idx[k];
        if (j >= 0 && j < n) ++tptr[(size_t)j + 1];
    }
    for (int64_t i = 0; i < n; ++i) tptr[i + 1] += tptr[i];
    std::vector<int32_t> tidx((size_t)tptr[n]);
    {
        std::vector<int64_t> cur(tptr.begin(), tptr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (I k = ptr[i]; k < ptr[i + 1]; ++k) {
                const int64_t j = (int64_t)idx[k];
                if (j >= 0 && j < n) tidx[(size_t)cur[j]++] = (int32_t)i;
            }
    }
    int64_t max_score = 0;
    for (int64_t i = 0; i < n; ++i)
        max_score = std::max<int64_t>(max_score, (int64_t)(ptr[i + 1] - ptr[i]) +
                                                     (tptr[i + 1] - tptr[i]));
    BucketQueue q(n, max_score);
    std::vector<uint8_t> placed((size_t)n, 0);
    auto touch = [&](int32_t v, int d) {
        auto one = [&](int64_t u) {
            if (u >= 0 && u < n && !placed[u]) q.bump((int32_t)u, d);
        };
        for (I k = ptr[v]; k < ptr[v + 1]; ++k) one((int64_t)idx[k]);
        for (int64_t k = tptr[v]; k < tptr[v + 1]; ++k) one(tidx[k]);
    };
    for (int64_t i = 0; i < n; ++i) {
        const int32_t v = q.pop_max();
        placed[v] = 1;
        order.push_back(v);
        touch(v, +1);
        if (i >= kWindow) touch(order[(size_t)(i - kWindow)], -1);
    }
    return order;
}

template std::vector<int32_t> locality_schedule<int64_t>(int64_t, const int64_t*, const int64_t*);
template std::vector<int32_t> locality_schedule<int32_t>(int64_t, const int32_t*, const int32_t*);

namespace {

// How far past a group's first row (in the visiting order) its other rows
// may come from.  An L2 model of one XCD (LRU over 1 KiB slab rows, 4 MiB,
// 512 groups in flight) on knn30_200k with groups of 4: consecutive rows
// gather 7.34 M union columns with 1.30 M misses; unrestricted neighbour
// groups 6.29 M with 1.73 M; a reach of 256, 512 and 1024 rows 6.65, 6.53
// and 6.51 M with 1.38, 1.40 and 1.45 M.  A miss costs about 1.5 union
// gathers on the device, so 512 is the best of these.
constexpr int64_t kGroupReach = 512;

// Distinct columns summed over the groups of R consecutive rows of `rows`:
// the gathers of the row-grouped form.
int64_t grouped_union(int64_t n, const int64_t* ptr, const int64_t* idx,
                      const std::vector<int32_t>& rows, int R) {
    std::vector<int64_t> mark((size_t)n, -1);
    int64_t total = 0;
    for (size_t k = 0; k < rows.size(); ++k) {
        const int64_t g = (int64_t)(k / (size_t)R);
        const int32_t r = rows[k];
        for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e) {
            const int64_t c = idx[e];
            if (c >= 0 && c < n && mark[(size_t)c] != g) {
                mark[(size_t)c] = g;
                ++total;
            }
        }
    }
    return total;
}

}  // namespace

std::vector<int32_t> neighbour_groups(int64_t n, const int64_t* ptr, const int64_t* idx,
                                      const std::vector<int32_t>& order, int R) {
    std::vector<int32_t> out;
    out.reserve((size_t)n);
    // position of each row in the order: a group takes only rows placed at
    // most kGroupReach after its first row, so the groups keep the order's
    // locality (what an XCD has in flight still shares its L2)
    std::vector<int64_t> pos((size_t)n, -1);
    for (size_t k = 0; k < order.size(); ++k) pos[(size_t)order[k]] = (int64_t)k;
    std::vector<uint8_t> used((size_t)n, 0);
    // stamp[c] == current group: column c is in the group's union
    std::vector<int32_t> stamp((size_t)n, -1);
    int32_t group = 0;
    size_t cursor = 0;  // first possibly unplaced position of the order
    std::vector<int32_t> members;
    members.reserve((size_t)R);
    auto add = [&](int32_t row) {
        used[row] = 1;
        out.push_back(row);
        members.push_back(row);
        for (int64_t k = ptr[row]; k < ptr[row + 1]; ++k) stamp[(size_t)idx[k]] = group;
    };
    for (size_t at = 0; at < order.size(); ++at) {
        const int32_t first = order[at];
        if (used[first]) continue;
        members.clear();
        add(first);
        while ((int)members.size() < R) {
            // the unplaced neighbour of a member sharing most of the union
            int32_t best = -1;
            int64_t best_score = -1;
            for (size_t m = 0; m < members.size(); ++m) {
                const int32_t row = members[m];
                for (int64_t k = ptr[row]; k < ptr[row + 1]; ++k) {
                    const int64_t c = idx[k];
                    if (c < 0 || c >= n || used[c]) continue;
                    if (pos[(size_t)c] < (int64_t)at || pos[(size_t)c] > (int64_t)at + kGroupReach)
                        continue;
                    int64_t score = 0;
                    for (int64_t e = ptr[c]; e < ptr[c + 1]; ++e)
                        score += stamp[(size_t)idx[e]] == group;
                    if (score > best_score) {
                        best_score = score;
                        best = (int32_t)c;
                    }
                }
            }
            if (best < 0) {
                // no unplaced neighbour: the next unplaced row of the order
                while (cursor < order.size() && used[order[cursor]]) ++cursor;
                if (cursor == order.size()) break;
                best = order[cursor];
            }
            add(best);
        }
        ++group;
    }
    // greedy groups can lose to plain consecutive rows (a pattern the
    // locality order already walks well): keep whichever gathers less
    if (grouped_union(n, ptr, idx, out, R) >= grouped_union(n, ptr, idx, order, R)) return order;
    return out;
}

}  // namespace scl_amd
