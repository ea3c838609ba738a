"""Host-side row orders of the SpMM forms, checked on the CPU.

scl-core_amd/csrc/schedule.cpp is compiled with a small C++ driver.  On
random square patterns (clustered neighbourhoods, as kNN graphs have) the
neighbour grouping of the row-grouped form must return every row exactly
once with full groups of R, and its groups' column unions must never exceed
those of consecutive rows of the locality order (the gathers it saves; on
clustered 2-D neighbourhoods it must save some); the locality order itself
must be a permutation."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "scl-core_amd" / "csrc"

DRIVER = r"""
#include "internal.hpp"
#include <cstdio>
#include <random>
#include <vector>
#include <algorithm>
using namespace scl_amd;
#define REQUIRE(c) do { if (!(c)) { std::printf("FAIL %s line %d case %d\n", #c, __LINE__, cs); return 1; } } while (0)

static double union_fraction(const std::vector<int64_t>& ptr, const std::vector<int64_t>& idx,
                             const std::vector<int32_t>& order, int R) {
    const int64_t n = (int64_t)ptr.size() - 1;
    std::vector<int64_t> mark(n, -1);
    int64_t uni = 0;
    for (int64_t g = 0; g * R < n; ++g)
        for (int64_t k = g * R; k < std::min<int64_t>(n, (g + 1) * R); ++k) {
            const int32_t r = order[k];
            for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e)
                if (mark[idx[e]] != g) { mark[idx[e]] = g; ++uni; }
        }
    return ptr[n] ? (double)uni / ptr[n] : 0.0;
}

int main() {
    std::mt19937_64 rng(7);
    int cs = 0;
    for (cs = 0; cs < 40; ++cs) {
        // points on a line in clusters: each row's neighbours are nearby rows
        const int64_t n = 200 + (int64_t)(rng() % 3000);
        const int deg = 4 + (int)(rng() % 20);
        std::vector<double> pos(n);
        for (auto& x : pos) x = (double)(rng() % 100000);
        std::vector<int64_t> ptr(n + 1, 0), idx;
        std::vector<int64_t> by(n);
        for (int64_t i = 0; i < n; ++i) by[i] = i;
        std::sort(by.begin(), by.end(), [&](int64_t a, int64_t b) { return pos[a] < pos[b]; });
        std::vector<int64_t> rank(n);
        for (int64_t i = 0; i < n; ++i) rank[by[i]] = i;
        for (int64_t i = 0; i < n; ++i) {
            std::vector<int64_t> nb;
            for (int d = -deg; d <= deg; ++d) {
                const int64_t r = rank[i] + d;
                if (d != 0 && r >= 0 && r < n && rng() % 3) nb.push_back(by[r]);
            }
            std::sort(nb.begin(), nb.end());
            for (int64_t c : nb) idx.push_back(c);
            ptr[i + 1] = (int64_t)idx.size();
        }
        std::vector<int32_t> order = locality_schedule<int64_t>(n, ptr.data(), idx.data());
        std::vector<int> seen(n, 0);
        REQUIRE((int64_t)order.size() == n);
        for (int32_t r : order) { REQUIRE(r >= 0 && r < n); seen[r]++; }
        for (int64_t i = 0; i < n; ++i) REQUIRE(seen[i] == 1);
        for (int R : {4, 8}) {
            std::vector<int32_t> grp = neighbour_groups(n, ptr.data(), idx.data(), order, R);
            std::fill(seen.begin(), seen.end(), 0);
            REQUIRE((int64_t)grp.size() == n);
            for (int32_t r : grp) { REQUIRE(r >= 0 && r < n); seen[r]++; }
            for (int64_t i = 0; i < n; ++i) REQUIRE(seen[i] == 1);
            // groups open in the order of their first row
            std::vector<int64_t> where(n);
            for (int64_t i = 0; i < n; ++i) where[order[i]] = i;
            for (int64_t g = 1; g * R < n; ++g)
                REQUIRE(where[grp[g * R]] > where[grp[(g - 1) * R]]);
            const double a = union_fraction(ptr, idx, order, R);
            const double b = union_fraction(ptr, idx, grp, R);
            REQUIRE(b <= a + 1e-12);
            if (cs == 0) std::printf("R=%d consecutive %.3f grouped %.3f\n", R, a, b);
        }
    }
    // 2-D points: neighbours are the k nearest by a brute-force scan; the
    // greedy groups must gather strictly less than consecutive rows
    {
        cs = 1000;
        const int64_t n = 3000;
        const int k = 12;
        std::vector<double> x(n), y(n);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = (double)(rng() % 100000) / 1000.0;
            y[i] = (double)(rng() % 100000) / 1000.0;
        }
        std::vector<int64_t> ptr(n + 1, 0), idx;
        std::vector<std::pair<double, int64_t>> d(n);
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t j = 0; j < n; ++j)
                d[j] = {(x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]), j};
            std::partial_sort(d.begin(), d.begin() + k + 1, d.end());
            std::vector<int64_t> nb;
            for (int q = 1; q <= k; ++q) nb.push_back(d[q].second);
            std::sort(nb.begin(), nb.end());
            for (int64_t c : nb) idx.push_back(c);
            ptr[i + 1] = (int64_t)idx.size();
        }
        std::vector<int32_t> order = locality_schedule<int64_t>(n, ptr.data(), idx.data());
        std::vector<int32_t> grp = neighbour_groups(n, ptr.data(), idx.data(), order, 4);
        const double a = union_fraction(ptr, idx, order, 4), b = union_fraction(ptr, idx, grp, 4);
        std::printf("2-D kNN R=4 consecutive %.3f grouped %.3f\n", a, b);
        REQUIRE(b < a);
    }
    std::printf("ok %d cases\n", cs);
    return 0;
}
"""


def test_neighbour_groups_and_locality_order(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{CSRC}", f"-I{ROOT / 'include'}", str(src), str(CSRC / "schedule.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ok" in out.stdout


MODEL = r"""
#include "internal.hpp"
#include <cstdio>
#include <vector>
#include <algorithm>
using namespace scl_amd;

static std::vector<int64_t> rd(const char* f) {
    FILE* p = std::fopen(f, "rb");
    std::fseek(p, 0, SEEK_END);
    const long sz = std::ftell(p);
    std::fseek(p, 0, SEEK_SET);
    std::vector<int64_t> v((size_t)sz / 8);
    if (std::fread(v.data(), 8, v.size(), p) != v.size()) v.clear();
    std::fclose(p);
    return v;
}

// LRU over slab rows: a doubly linked list with an in-cache flag per row
struct Lru {
    int64_t cap, head = -1, tail = -1, size = 0;
    std::vector<int64_t> prev, next;
    std::vector<char> in;
    Lru(int64_t n, int64_t c) : cap(c), prev(n, -1), next(n, -1), in(n, 0) {}
    void unlink(int64_t x) {
        if (prev[x] >= 0) next[prev[x]] = next[x]; else head = next[x];
        if (next[x] >= 0) prev[next[x]] = prev[x]; else tail = prev[x];
    }
    void front(int64_t x) {
        prev[x] = -1; next[x] = head;
        if (head >= 0) prev[head] = x;
        head = x;
        if (tail < 0) tail = x;
    }
    bool access(int64_t x) {
        if (in[x]) { unlink(x); front(x); return true; }
        if (size == cap) { const int64_t t = tail; unlink(t); in[t] = 0; --size; }
        front(x); in[x] = 1; ++size;
        return false;
    }
};

// union entries and misses of groups of R consecutive rows of `rows`
static void model(const std::vector<int64_t>& ip, const std::vector<int64_t>& ix,
                  const std::vector<int32_t>& rows, int R, int64_t cap, int W,
                  int64_t& acc, int64_t& miss) {
    const int64_t n = (int64_t)ip.size() - 1, G = (n + R - 1) / R;
    std::vector<int64_t> uptr(G + 1, 0), ucol, mark(n, -1);
    for (int64_t g = 0; g < G; ++g) {
        std::vector<int64_t> cols;
        for (int64_t k = g * R; k < std::min<int64_t>(n, (g + 1) * R); ++k)
            for (int64_t e = ip[rows[k]]; e < ip[rows[k] + 1]; ++e)
                if (mark[ix[e]] != g) { mark[ix[e]] = g; cols.push_back(ix[e]); }
        std::sort(cols.begin(), cols.end());
        ucol.insert(ucol.end(), cols.begin(), cols.end());
        uptr[g + 1] = (int64_t)ucol.size();
    }
    Lru L(n, cap);
    acc = miss = 0;
    std::vector<int64_t> sg(W, -1), se(W, 0);
    int64_t ng = 0, active = 0;
    for (int s = 0; s < W && ng < G; ++s, ++ng, ++active) { sg[s] = ng; se[s] = uptr[ng]; }
    while (active > 0)
        for (int s = 0; s < W; ++s) {
            if (sg[s] < 0) continue;
            const int64_t g = sg[s];
            for (int u = 0; u < 16 && se[s] < uptr[g + 1]; ++u) {
                ++acc;
                if (!L.access(ucol[se[s]++])) ++miss;
            }
            if (se[s] >= uptr[g + 1]) {
                if (ng < G) { sg[s] = ng; se[s] = uptr[ng]; ++ng; }
                else { sg[s] = -1; --active; }
            }
        }
}

int main(int argc, char** argv) {
    const std::vector<int64_t> ip = rd(argv[1]), ix = rd(argv[2]);
    const int64_t n = (int64_t)ip.size() - 1;
    const int64_t cap = 4096 * n / 200000;
    const int W = (int)(512 * n / 200000);
    const std::vector<int32_t> order = locality_schedule<int64_t>(n, ip.data(), ix.data());
    const std::vector<int32_t> grp = neighbour_groups(n, ip.data(), ix.data(), order, 4);
    int64_t a_acc, a_miss, b_acc, b_miss;
    model(ip, ix, order, 4, cap, W, a_acc, a_miss);
    model(ip, ix, grp, 4, cap, W, b_acc, b_miss);
    const double a = a_acc + 1.5 * a_miss, b = b_acc + 1.5 * b_miss;
    std::printf("consecutive: union %lld misses %lld cost %.0f\n", (long long)a_acc,
                (long long)a_miss, a);
    std::printf("groups:      union %lld misses %lld cost %.0f\n", (long long)b_acc,
                (long long)b_miss, b);
    if (!(b_acc < a_acc && b <= a)) { std::printf("FAIL\n"); return 1; }
    std::printf("ok\n");
    return 0;
}
"""


def test_l2_model_of_grouped_order(tmp_path):
    import numpy as np
    from scl_amd import synth
    g = synth.knn_graph(20_000, 30, seed=2)
    ip_f, ix_f = tmp_path / "indptr.bin", tmp_path / "indices.bin"
    g.indptr.astype(np.int64).tofile(ip_f)
    g.indices.astype(np.int64).tofile(ix_f)
    src = tmp_path / "model.cpp"
    src.write_text(MODEL)
    exe = tmp_path / "model"
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{CSRC}", f"-I{ROOT / 'include'}", str(src), str(CSRC / "schedule.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe), str(ip_f), str(ix_f)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0 and "ok" in out.stdout, out.stdout + out.stderr
