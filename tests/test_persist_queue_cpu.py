"""Work queues of the persistent batched plan, checked on the host.

The queue builder (scl-core_amd/csrc/persist_queue.cpp) is compiled here
with a small C++ driver.  For random batches and every option it checks that
each graph lives in one XCD's queue (dealt round-robin into sub-queues), that every phase's items tile the graph's rows
exactly once (forward phases per feature slab), that each item's `need` counts the graph's items
of its earlier phases (all listed before it in the same queue), and that draining an XCD's sub-queues in order with any number of
workers on each, under any interleaving, never deadlocks and completes every
item (randomised schedules).  The dP partial offsets are padded to 128 bytes and
disjoint."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "scl-core_amd" / "csrc"

DRIVER = r"""
#include "persist.hpp"
#include <cstdio>
#include <random>
#include <vector>
using namespace scl_amd;
#define REQUIRE(c) do { if (!(c)) { std::printf("FAIL %s line %d case %d\n", #c, __LINE__, cs); return 1; } } while (0)

int main() {
    std::mt19937_64 rng(12345);
    int cs = 0;
    for (cs = 0; cs < 400; ++cs) {
        const int G = 1 + (int)(rng() % 40);
        std::vector<int64_t> off(G + 1, 0), nz(G + 1, 0);
        for (int g = 0; g < G; ++g) {
            off[g + 1] = off[g] + 1 + (int64_t)(rng() % 3000);
            const int64_t rows = off[g + 1] - off[g];
            nz[g + 1] = nz[g] + (rng() % 7 == 0 ? 0 : rows * (int64_t)(rng() % 40));
        }
        const int64_t t = 1 + (int64_t)(rng() % 6);
        const bool backward = rng() % 3 != 0;
        const int64_t R = (int64_t)(rng() % 3 == 0 ? 0 : 1 + rng() % 300);
        const int64_t K = (int64_t)(rng() % 3 == 0 ? 0 : 1 + rng() % 6);
        const int S = 1 + (int)(rng() % 4);
        PersistQueues Q = build_persist_queues(off, nz, t, backward, R, K, S);
        const int64_t row_phases = backward ? 2 * t : t;
        const int64_t n_phase = row_phases;
        REQUIRE((int)Q.q_off.size() == kPersistLanes + 1 && Q.q_off[0] == 0);
        REQUIRE(Q.q_off[kPersistLanes] == (int)Q.items.size());
        for (int l = 0; l < kPersistLanes; ++l) REQUIRE(Q.q_off[l] <= Q.q_off[l + 1]);
        // each XCD's list, rebuilt from its sub-queues dealt round-robin
        std::vector<std::vector<PersistItem>> lists(kPersistQueues);
        for (int q = 0; q < kPersistQueues; ++q) {
            const int l0 = q * kPersistSubQueues;
            size_t total = (size_t)(Q.q_off[l0 + kPersistSubQueues] - Q.q_off[l0]);
            for (size_t i = 0; i < total; ++i) {
                const int l = l0 + (int)(i % kPersistSubQueues);
                const int at = Q.q_off[l] + (int)(i / kPersistSubQueues);
                REQUIRE(at < Q.q_off[l + 1]);
                lists[q].push_back(Q.items[at]);
            }
        }
        // one XCD per graph; need = the graph's items of earlier phases
        std::vector<int> home(G, -1);
        std::vector<int> seen(G, 0), phase_start(G, 0), cur_phase(G, -1);
        // cover[g][p * S + slab]: next expected start of phase p's slab
        // (forward phases have S slabs, the others one), slabs in order
        std::vector<std::vector<int64_t>> cover(G, std::vector<int64_t>(n_phase * S));
        auto lo_of = [&](int g, int64_t p) { return p < row_phases ? off[g] : nz[g]; };
        auto hi_of = [&](int g, int64_t p) { return p < row_phases ? off[g + 1] : nz[g + 1]; };
        auto slabs = [&](int64_t p) { return p < t ? S : 1; };
        for (int g = 0; g < G; ++g)
            for (int64_t p = 0; p < n_phase; ++p)
                for (int sl = 0; sl < S; ++sl) cover[g][p * S + sl] = lo_of(g, p);
        int64_t widest = 0;
        for (int q = 0; q < kPersistQueues; ++q) {
            int64_t run = 0;
            for (size_t i = 0; i < lists[q].size(); ++i) {
                const PersistItem& it = lists[q][i];
                run = (i > 0 && it.phase == lists[q][i - 1].phase) ? run + 1 : 1;
                widest = std::max(widest, run);
                REQUIRE(it.g >= 0 && it.g < G && it.phase >= 0 && it.phase < n_phase);
                REQUIRE(home[it.g] == -1 || home[it.g] == q);
                home[it.g] = q;
                if (it.phase != cur_phase[it.g]) {
                    REQUIRE(it.phase == cur_phase[it.g] + 1);
                    cur_phase[it.g] = it.phase;
                    phase_start[it.g] = seen[it.g];
                }
                // every item of the graph's earlier phases, and nothing else
                REQUIRE(it.need == phase_start[it.g]);
                seen[it.g]++;
                REQUIRE(it.slab >= 0 && it.slab < slabs(it.phase));
                int64_t& c = cover[it.g][it.phase * S + it.slab];
                REQUIRE(it.begin == c && it.end > it.begin);
                c = it.end;
                // phases of a graph, and slabs of a phase, appear in order
                for (int64_t p = 0; p < it.phase; ++p)
                    for (int sl = 0; sl < slabs(p); ++sl)
                        REQUIRE(cover[it.g][p * S + sl] == hi_of(it.g, p));
                for (int sl = 0; sl < it.slab; ++sl)
                    REQUIRE(cover[it.g][it.phase * S + sl] == hi_of(it.g, it.phase));
            }
        }
        REQUIRE(widest == Q.widest && Q.widest >= 1);
        for (int g = 0; g < G; ++g) {
            REQUIRE(home[g] >= 0 && seen[g] == Q.expect[g]);
            for (int64_t p = 0; p < n_phase; ++p)
                for (int sl = 0; sl < slabs(p); ++sl)
                    REQUIRE(cover[g][p * S + sl] == hi_of(g, p));
        }
        // dP partials: 32-float aligned, disjoint, inside the stride
        int64_t prev_end = 0;
        for (int g = 0; g < G; ++g) {
            const int64_t base = Q.dpad[g] + nz[g];
            REQUIRE(base % 32 == 0 && base >= prev_end);
            prev_end = base + (nz[g + 1] - nz[g]);
        }
        REQUIRE(prev_end + 32 <= Q.dstride);
        // drain each XCD's sub-queues together with the kernel's policy:
        // W_l waves on sub-queue l, each holding its current item and the
        // next one, claimed as the current one starts; a wave
        // publishes its completed items lazily, per run of one (graph,
        // phase), and always before it waits and when it leaves; it skips the
        // check when its last satisfied dependency covers the item.  Random
        // interleavings must never leave every wave blocked, and no item may
        // start before the items it needs are published.
        for (int rep = 0; rep < 3; ++rep) {
            std::vector<int> done(G, 0);
            for (int q = 0; q < kPersistQueues; ++q) {
                struct Wave {
                    int lane;
                    std::vector<int> held;
                    int stage;  // 0: start the front item, 1: waiting, 2: running, 3: gone
                    int pend_g, pend_phase, pend_n, ok_g, ok_need;
                };
                std::vector<Wave> ws;
                std::vector<int> head(kPersistSubQueues);
                auto claim = [&](Wave& w) {
                    const int sq = w.lane - q * kPersistSubQueues;
                    if (head[sq] < Q.q_off[w.lane + 1]) w.held.push_back(head[sq]++);
                    else w.held.push_back(-1);  // a claim past the end
                };
                auto publish = [&](Wave& w) {
                    if (w.pend_n > 0) done[w.pend_g] += w.pend_n;
                    w.pend_n = 0;
                };
                for (int sq = 0; sq < kPersistSubQueues; ++sq) {
                    const int l = q * kPersistSubQueues + sq;
                    head[sq] = Q.q_off[l];
                    const int W = 1 + (int)(rng() % 12);
                    for (int k = 0; k < W; ++k) ws.push_back({l, {}, 0, -1, -1, 0, -1, 0});
                }
                for (auto& w : ws) {
                    claim(w);
                    if (w.held[0] < 0) w.stage = 3;
                }
                while (true) {
                    std::vector<int> movable;
                    for (size_t k = 0; k < ws.size(); ++k) {
                        Wave& w = ws[k];
                        if (w.stage == 3) continue;
                        if (w.stage == 1) {
                            const PersistItem& it = Q.items[w.held[0]];
                            if (done[it.g] < it.need) continue;
                        }
                        movable.push_back((int)k);
                    }
                    if (movable.empty()) break;
                    Wave& w = ws[movable[rng() % movable.size()]];
                    const PersistItem& it = Q.items[w.held[0]];
                    if (w.stage == 0) {
                        claim(w);
                        if (it.g != w.pend_g || it.phase != w.pend_phase) publish(w);
                        if (it.need > 0 && !(it.g == w.ok_g && it.need <= w.ok_need)) {
                            publish(w);
                            w.stage = 1;
                        } else {
                            // safety: an item runs only once its dependencies
                            // are published (their stores drained)
                            REQUIRE(done[it.g] >= it.need);
                            w.stage = 2;
                        }
                    } else if (w.stage == 1) {
                        REQUIRE(done[it.g] >= it.need);
                        w.ok_g = it.g;
                        w.ok_need = it.need;
                        w.stage = 2;
                    } else {
                        w.pend_g = it.g;
                        w.pend_phase = it.phase;
                        w.pend_n++;
                        w.held.erase(w.held.begin());
                        if (w.held.empty() || w.held[0] < 0) {
                            publish(w);
                            w.stage = 3;
                        } else {
                            w.stage = 0;
                        }
                    }
                }
                for (const Wave& w : ws) REQUIRE(w.stage == 3);   // no deadlock
                for (int sq = 0; sq < kPersistSubQueues; ++sq)
                    REQUIRE(head[sq] == Q.q_off[q * kPersistSubQueues + sq + 1]);
            }
            for (int g = 0; g < G; ++g) REQUIRE(done[g] == Q.expect[g]);
        }
    }
    std::printf("ok %d cases\n", cs);
    return 0;
}
"""


def test_persistent_queues_tile_and_never_deadlock(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{CSRC}", f"-I{ROOT / 'include'}", str(src), str(CSRC / "persist_queue.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
