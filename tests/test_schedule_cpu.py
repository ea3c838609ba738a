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
