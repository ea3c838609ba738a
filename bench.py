#!/usr/bin/env python3
"""Diffusion benchmark: Y = P^t X on one MI355X.

Headline (value): knn30_200k -- n = 200,000 cells, geometric kNN graph with
k = 30 and union-max symmetrisation (nnz ~ 8.9e6, rows 30..~100), rows
normalised in FP64 on the GPU; X = 200,000 x 2000 fp32 log1p(Poisson)
counts; t = 8.  One bench step is one full diffusion (t SpMM launches) over
device-resident P and X.

  value             algorithmic HBM GB/s of the diffusion:
                    t * (nnz*(4+4) + 4*(n+1) + n*F*(4+4)) / time
  roofline.achieved the same bytes per SpMM launch / mean launch time (HIP
                    events on the stream the kernels run on)
  roofline.traffic  HBM bytes per launch from rocprofv3 FETCH_SIZE and
                    WRITE_SIZE (profiles/traffic_knn30_200k.json)
  cpu_baseline      the FP32 CPU restatement of the reference algorithm
                    (OpenMP, all rows, a column sample, scaled to full width)

Further legs (in "legs"; --legs selects, "c2" alone skips them; "cr" runs
only when named: the headline kernel on operators whose columns are random
within the first M rows, M = 2048 .. 200,000, to time L2, Infinity Cache
and HBM-served misses):
  c3  knn30_500k, F=128, t=6: forward with checkpoints every 2 + dL/dY of
      0.5||Y-T||^2 + fused adjoint/edge-gradient reverse pass (one hipGraph),
      vs the same work as plain kernels (6 SpMM + 6 transposed SpMM + 6 edge-
      gradient passes with every iterate kept)
  c4  uniform15_1m, F=256, t=4: bf16 iterates (fp32 or bf16 P values),
      fp16 iterates (2^k scaled; X in fp32 or stored in fp16 too) vs the
      fp32 kernel
  c5  256 graphs, n_i log-uniform in [2000, 50000], k_i in 10..30, F=64,
      t=5, fwd+bwd storing all t iterates: one hipGraph over the
      block-diagonal stack vs the 256
      graphs one after another with plain (non-graph) launches
  mp  a skewed pattern (Pareto row lengths, hub rows of 200k nonzeros),
      n=1M, F=64: merge-path SpMM vs the autotuned row-split SpMM
  wk  the vector-side entry points (lazy walk, heat kernel, RWR, DPT,
      diffusion map, diffusion distance) through the C ABI vs the CPU
      restatement
  cl  the other SpMM callers (diffusion imputation, MAGIC, diffusion-
      component pseudotime, Y = alpha A X + beta Y) through the C ABI vs the
      CPU restatement
  kn  building the operator: exact kNN, kernel weights, row normalisation,
      stationary distribution, through the C ABI vs the CPU restatement
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "scl-core_amd"))

PEAK_HBM_GBPS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "traffic_knn30_200k.json"
METRIC = "diffusion HBM GB/s (% peak) + fwd+bwd steps/s vs plain FP32 at stated err"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def step_bytes(n: int, nnz: int, F: int, vb=4, xb=4, yb=4) -> int:
    return nnz * (vb + 4) + 4 * (n + 1) + n * F * (xb + yb)


def host_median_ms(fn, reps: int) -> float:
    """Median wall-clock milliseconds of fn() over reps calls after one
    untimed call (host-array entry points: their copies are part of it)."""
    import numpy as np
    fn()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(times))


def timed(fn, reps: int, stream) -> float:
    """Mean milliseconds of fn() over reps, HIP events on `stream`."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def host_cpu() -> dict:
    """lscpu model, sockets, physical cores and SIMD flags of this host."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        sockets = int(kv.get("Socket(s)", "0") or 0)
        per = int(kv.get("Core(s) per socket", "0") or 0)
        info = {"model": kv.get("Model name"), "sockets": sockets,
                "physical_cores": sockets * per if sockets and per else None,
                "threads_per_core": kv.get("Thread(s) per core"),
                "logical_cpus": kv.get("CPU(s)")}
        flags = kv.get("Flags", "").split()
        info["simd"] = [f for f in ("avx2", "fma", "avx512f", "avx512bw", "avx512_bf16")
                        if f in flags]
    except Exception as e:  # lscpu missing: report what os says
        import os
        info = {"model": None, "logical_cpus": os.cpu_count(), "error": str(e)}
    return info


def cpu_baseline(graph_arrays, n, F, t, seed, bytes_full_step, budget_s=12.0):
    """Times the CPU restatement on all rows and a column sample, on all the
    OpenMP threads and on one."""
    import numpy as np
    from oracle import oracle as orc
    from scl_amd import synth
    ip, ix, p = graph_arrays
    threads = orc.num_threads()
    cols = np.arange(0, F, max(1, F // 32))[:32]
    Xs = synth.log1p_poisson(n, F, seed=seed, cols=cols)
    orc.diffuse_matrix_f32(ip, ix, p, Xs, 1)  # warm the pages and threads
    t0 = time.perf_counter()
    orc.diffuse_matrix_f32(ip, ix, p, Xs, t)
    dt = time.perf_counter() - t0
    fs = len(cols)
    if dt * 4 < budget_s:  # widen the sample to use the time budget
        more = int(min(F, max(fs, fs * budget_s / max(dt, 1e-3) / 2)))
        cols = np.linspace(0, F - 1, more).astype(np.int64)
        Xs = synth.log1p_poisson(n, F, seed=seed, cols=cols)
        t0 = time.perf_counter()
        orc.diffuse_matrix_f32(ip, ix, p, Xs, t)
        dt = time.perf_counter() - t0
        fs = len(cols)
    full_time = dt * F / fs
    # one thread on 4 columns
    c1 = np.arange(0, F, F // 4)[:4]
    X1 = synth.log1p_poisson(n, F, seed=seed, cols=c1)
    orc.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        orc.diffuse_matrix_f32(ip, ix, p, X1, t)
        dt1 = time.perf_counter() - t0
    finally:
        orc.set_num_threads(threads)
    full1 = dt1 * F / len(c1)
    return {
        "value": round(t * bytes_full_step / full_time / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"all {n} rows x {fs} of {F} columns, t={t}, {dt:.2f} s on {threads} "
                   f"OpenMP threads; scaled to the full width (columns are independent): "
                   f"{full_time:.1f} s per diffusion; {orc.load().variant}"),
        "one_thread": {"value": round(t * bytes_full_step / full1 / 1e9, 3), "unit": "GB/s",
                       "sample": f"all rows x {len(c1)} columns, t={t}, {dt1:.2f} s; "
                                 f"{full1:.1f} s per diffusion scaled"},
        "host": host_cpu(),
    }


# ------------------------------------------------------------------- legs --

def leg_c3(steps: int) -> dict:
    """knn30_500k, F=128, t=6: fused fwd+bwd plan vs plain kernels."""
    import torch
    from scl_amd import CsrMatrix, batch, diffusion, synth, torch_ops
    n, F, t, every = 500_000, 128, 6, 2
    g = synth.knn_graph(n, 30, seed=3)
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    diffusion.compute_transition_matrix(P)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    X = torch.from_numpy(synth.normal(n, F, seed=3)).to(dev)
    T = torch.from_numpy(synth.normal(n, F, seed=103)).to(dev)
    torch_ops.autotune(P, X)
    torch_ops.autotune(P, X, transpose=True)   # the plain path's adjoint SpMM
    torch_ops.autotune(P, X, adjoint=True)     # the fused path's adjoint step
    plan = batch.DiffusionPlan(P, n, F, t, every=every)
    plan.X.copy_(X)
    plan.T.copy_(T)
    plan.launch()
    torch.cuda.synchronize()
    fused_ms = timed(plan.launch, steps, stream)
    # plain: every iterate kept, separate adjoint SpMM and edge-gradient passes
    its = [X] + [torch.empty_like(X) for _ in range(t)]
    G = [torch.empty_like(X), torch.empty_like(X)]
    dP = torch.empty(P.nnz, dtype=torch.float32, device=dev)

    from scl_amd import _lib
    lib, sp = _lib.load(), stream.cuda_stream

    def spmm(src, dst, transpose):
        _lib.check(lib.scl_gpu_spmm_device(P.handle, src.data_ptr(), dst.data_ptr(),
                                           n, F, transpose, sp))

    def plain():
        for s in range(1, t + 1):
            spmm(its[s - 1], its[s], 0)
        torch.sub(its[t], T, out=G[0])
        cur = 0
        for s in range(t, 0, -1):
            torch_ops.edge_gradient(P, G[cur], its[s - 1], dP, accumulate=s != t)
            spmm(G[cur], G[cur ^ 1], 1)
            cur ^= 1

    plain()
    torch.cuda.synchronize()
    plain_ms = timed(plain, max(2, steps // 2), stream)
    same_dx = bool(torch.equal(G[t % 2], plan.dX))
    nnz = g.nnz
    B_f = step_bytes(n, nnz, F)
    B_e = 8 * n * F + 12 * nnz
    total = 9 * B_f + 6 * (B_f + B_e) + 3 * n * F * 4 + 2 * n * F * 4
    # error against the FP64 gradient of the whole problem (outside timing)
    from oracle import oracle as orc
    ip, ix, pv = P.export()
    t0 = time.perf_counter()
    _, Y64, dX64, dP64 = orc.grad_f64(ip, ix, pv, synth.normal(n, F, seed=3),
                                      synth.normal(n, F, seed=103), t)
    err_s = time.perf_counter() - t0
    dXh, dPh, Yh = plan.dX.cpu().numpy(), plan.dP.cpu().numpy(), plan.Y.cpu().numpy()
    err = {"Y_max": orc.rel_err(Yh, Y64), "dX_max": orc.rel_err(dXh, dX64),
           "dX_median": orc.median_rel_err(dXh, dX64), "dP_max": orc.rel_err(dPh, dP64),
           "dP_median": orc.median_rel_err(dPh, dP64),
           "how": f"FP64 oracle gradient of all {n} x {F} and {nnz} nonzeros ({err_s:.1f} s)"}
    del Y64, dX64, dP64
    plan.close()
    return {
        "workload": "knn30_500k_f128_t6", "nnz": nnz,
        "fused_ms_per_fwd_bwd": round(fused_ms, 3),
        "fused_steps_per_s": round(1e3 / fused_ms, 2),
        "plain_ms_per_fwd_bwd": round(plain_ms, 3),
        "plain_steps_per_s": round(1e3 / plain_ms, 2),
        "speedup_vs_plain": round(plain_ms / fused_ms, 3),
        "algorithmic_GBps": round(total / fused_ms / 1e6, 1),
        "dX_fused_equals_plain": same_dx,
        "err_vs_fp64": err,
    }


def leg_c4(steps: int) -> dict:
    """uniform15_1m, F=256, t=4: reduced-precision storage vs fp32, each with
    its error against FP64 on 16 sampled columns."""
    import numpy as np
    import torch
    from oracle import oracle as orc
    from scl_amd import CsrMatrix, synth, torch_ops
    n, F, t = 1_000_000, 256, 4
    g = synth.uniform_knn(n, 15, seed=4).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    X = synth.normal(n, F, seed=4)
    Xd = torch.from_numpy(X).to(dev)
    torch_ops.autotune(P, Xd)
    Y32 = torch_ops.diffuse(P, Xd, t)
    torch.cuda.synchronize()
    f32_ms = timed(lambda: torch_ops.diffuse(P, Xd, t), steps, stream)
    cols = np.arange(0, F, 16)
    ct = torch.from_numpy(cols).to(dev)
    y64 = orc.diffuse_f64(g.indptr, g.indices, g.data, X[:, cols], t)

    def errs(Y):
        Yh = Y[:, ct].cpu().numpy()
        return {"max_rel_err_vs_fp64": orc.rel_err(Yh, y64),
                "median_rel_err_vs_fp64": orc.median_rel_err(Yh, y64)}

    out = {"workload": "uniform15_1m_f256_t4", "nnz": g.nnz,
           "fp32_ms": round(f32_ms, 3),
           "fp32_GBps": round(t * step_bytes(n, g.nnz, F) / f32_ms / 1e6, 1),
           "fp32": errs(Y32),
           "error_columns": f"{len(cols)} of {F} (every 16th), FP64 oracle"}
    k16 = torch_ops.lowp_scale_exp(P, Xd, t)
    width = {"bf16": 2, "fp16": 2, "comp": 3}
    for fmt, vb, ri in (("comp", True, True), ("comp", False, False), ("bf16", True, True),
                        ("bf16", False, False), ("bf16", True, False), ("fp16", False, False),
                        ("fp16", False, True)):
        kk = k16 if fmt == "fp16" else 0
        run = lambda: torch_ops.diffuse_lowp(P, Xd, t, fmt, bf16_values=vb, scale_exp=kk,
                                             round_input=ri)
        Yb = run()
        ms = timed(run, steps, stream)
        w = width[fmt]
        vbytes = w if vb else 4
        if ri:  # X stored reduced once, then t steps gathering reduced rows
            bytes_ = (n * F * (4 + w) + (t - 1) * step_bytes(n, g.nnz, F, vbytes, w, w)
                      + step_bytes(n, g.nnz, F, vbytes, w, 4))
        else:
            bytes_ = (step_bytes(n, g.nnz, F, vbytes, 4, w)
                      + (t - 2) * step_bytes(n, g.nnz, F, vbytes, w, w)
                      + step_bytes(n, g.nnz, F, vbytes, w, 4))
        name = {"comp": "compensated_bf16", "bf16": "bf16", "fp16": "fp16_2k"}[fmt]
        key = (f"{name}_iterates_{('compensated' if fmt == 'comp' else 'bf16') if vb else 'fp32'}"
               f"_values" + ("_stored_input" if ri else ""))
        rec = {"ms": round(ms, 3), "GBps": round(bytes_ / ms / 1e6, 1),
               "speedup_vs_fp32": round(f32_ms / ms, 3)}
        rec.update(errs(Yb))
        rec["meets_1e-3"] = rec["max_rel_err_vs_fp64"] <= 1e-3
        out[key] = rec
    out["spec"] = ("P values and X stored in bf16: compensated_bf16_iterates_compensated_"
                   "values_stored_input (bf16 plus an 8-bit residual per value, 3 bytes)")
    return out


def leg_mp(steps: int) -> dict:
    """Skewed pattern (Pareto row lengths, 4 hub rows of 200k nonzeros),
    n=1M, F=64: merge-path SpMM vs the autotuned row-split SpMM."""
    import numpy as np
    import torch
    from scl_amd import CsrMatrix, torch_ops
    n, F = 1_000_000, 64
    rng = np.random.Generator(np.random.PCG64(6))
    deg = np.minimum((rng.pareto(1.1, n) * 4).astype(np.int64) + 1, 20_000)
    deg[rng.integers(0, n, 4)] = 200_000
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    key = np.unique(rows * n + rng.integers(0, n, len(rows)))
    rows, cols = key // n, key % n
    ip = np.zeros(n + 1, np.int64)
    np.add.at(ip, rows + 1, 1)
    ip = np.cumsum(ip)
    vals = rng.random(len(cols)).astype(np.float32)
    P = CsrMatrix.from_arrays(ip, cols, vals, (n, n))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    X = torch.from_numpy(rng.standard_normal((n, F), dtype=np.float32)).to(dev)
    row_ms = torch_ops.autotune(P, X)
    row_shape = torch_ops.get_shape(P, F)
    Yr = torch_ops.spmm(P, X)
    row_ms = timed(lambda: torch_ops.spmm(P, X), steps, stream)
    # FP64 oracle on 4 columns (outside timing), in the max-norm convention
    from oracle import oracle as orc
    ecols = [0, 1, 31, 63]
    Xh = X[:, ecols].cpu().numpy()
    y64 = orc.diffuse_f64(ip, cols, vals, Xh, 1)

    def err64(Y):
        Yh = Y[:, ecols].cpu().numpy().astype(np.float64)
        return float(np.abs(Yh - y64).max() / np.abs(y64).max())

    out = {"workload": "pareto_1m_f64", "nnz": int(ip[-1]), "max_row": int(deg.max()),
           "rowsplit_ms": round(row_ms, 3), "rowsplit_shape": row_shape,
           "rowsplit_err_vs_fp64": err64(Yr), "error_columns": ecols}
    best = None
    for m in (1, 3, 5):
        torch_ops.set_shape(P, F, 4, 64, 8, merge=m)
        Ym = torch_ops.spmm(P, X)
        ms = timed(lambda: torch_ops.spmm(P, X), steps, stream)
        err = float((Ym - Yr).abs().max()) / float(Yr.abs().max())
        out[f"merge{m}_ms"] = round(ms, 3)
        out[f"merge{m}_max_rel_diff"] = err
        out[f"merge{m}_err_vs_fp64"] = err64(Ym)
        if best is None or ms < best[0]:
            best = (ms, m)
    out["speedup_merge_vs_rowsplit"] = round(row_ms / best[0], 2)
    out["GBps_best"] = round(step_bytes(n, int(ip[-1]), F) / min(best[0], row_ms) / 1e6, 1)
    return out


def leg_wk(steps: int, with_cpu: bool = True) -> dict:
    """The vector-side diffusion entry points (host arrays in and out, as at
    the C ABI) on knn30_200k; with the CPU baseline on, the CPU restatement
    of each is timed beside it on this host."""
    from scl_amd import CsrMatrix, diffusion, synth
    if with_cpu:
        from oracle import oracle as orc
    n = 200_000
    g = synth.knn_graph(n, 30, seed=7).normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    P.prepare_device()
    x = synth.normal(n, 1, seed=7)[:, 0]
    ip, ix, v = g.indptr, g.indices, g.data
    if not with_cpu:
        orc = None
    cases = {
        "lazy_walk_t20": (lambda: diffusion.lazy_random_walk(P, x.copy(), 20, 0.5),
                          lambda: orc.lazy_walk_f32(ip, ix, v, x, 20, 0.5)),
        "heat_kernel_10": (lambda: diffusion.heat_kernel_signature(P, 1.0, 10),
                           lambda: orc.heat_kernel_f32(ip, ix, v, 1.0, 10)),
        "rwr_100": (lambda: diffusion.random_walk_with_restart(P, [0, 50, 900], 0.85, 100, 1e-30),
                    lambda: orc.rwr_f32(ip, ix, v, [0, 50, 900], 0.85, 100, -1.0)),
        "dpt_100": (lambda: diffusion.compute_dpt(P, 0, 100, 1e-30),
                    lambda: orc.dpt_f32(ip, ix, v, 0, 100, 1e-30)),
        "diffusion_map_k10_i20": (lambda: diffusion.diffusion_map_embedding(P, 10, 20),
                                  lambda: orc.dmap_f32(ip, ix, v, 10, 20)),
    }
    out = {"workload": "knn30_200k walks (host arrays in/out)"}
    if with_cpu:
        out["cpu_threads"] = orc.num_threads()

    def record(name, gpu_fn, cpu_fn, reps):
        g_ms = host_median_ms(gpu_fn, reps)
        out[name] = {"gpu_ms": round(g_ms, 2)}
        if with_cpu:
            t0 = time.perf_counter()
            cpu_fn()
            c_ms = (time.perf_counter() - t0) * 1e3
            out[name].update({"cpu_ms": round(c_ms, 1), "speedup": round(c_ms / g_ms, 1)})

    for name, (gpu_fn, cpu_fn) in cases.items():
        record(name, gpu_fn, cpu_fn, max(3, steps // 2))
    # one lazy-walk step (F = 1 SpMV on the renumbered copy, x gathered into
    # its order, blend) from the difference of a 101-step and a 1-step walk
    reps = max(3, steps // 2)
    t1 = host_median_ms(lambda: diffusion.lazy_random_walk(P, x.copy(), 1, 0.5), reps)
    t101 = host_median_ms(lambda: diffusion.lazy_random_walk(P, x.copy(), 101, 0.5), reps)
    out["lazy_walk_step_us"] = round((t101 - t1) * 10.0, 2)
    # dense output: diffusion distance on a 4000-node graph
    gs = synth.knn_graph(4000, 15, seed=7).normalized()
    Ps = CsrMatrix.from_arrays(gs.indptr, gs.indices, gs.data, (4000, 4000))
    record("distance_n4000_t3", lambda: diffusion.diffusion_distance(Ps, 3),
           lambda: orc.diffusion_distance_f32(gs.indptr, gs.indices, gs.data, 3), 3)
    return out


def leg_cl(steps: int, with_cpu: bool = True) -> dict:
    """The other SpMM callers (imputation, MAGIC, diffusion-component
    pseudotime, Y = alpha A X + beta Y) through the C ABI on knn30_200k with
    a 200000 x 500 expression matrix at 10% density; with the CPU baseline
    on, the CPU restatement of each is timed beside it on this host."""
    import numpy as np
    from scl_amd import CsrMatrix, callers, synth
    orc = None
    if with_cpu:
        from oracle import oracle as orc
    n, G = 200_000, 500
    raw = synth.knn_graph(n, 30, seed=7)
    g = raw.normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    A = CsrMatrix.from_arrays(raw.indptr, raw.indices, raw.data, (n, n))
    P.prepare_device()
    A.prepare_device()
    rng = np.random.Generator(np.random.PCG64(8))
    rows, cols = [], []
    for r0 in range(0, n, 20_000):
        r, c = np.nonzero(rng.random((min(20_000, n - r0), G), dtype=np.float32) < 0.1)
        rows.append(r + r0)
        cols.append(c)
    rows, cols = np.concatenate(rows), np.concatenate(cols).astype(np.int64)
    ip = np.zeros(n + 1, np.int64)
    np.add.at(ip, rows + 1, 1)
    ip = np.cumsum(ip)
    ev = rng.gamma(1.0, 1.0, len(cols)).astype(np.float32)
    M = CsrMatrix.from_arrays(ip, cols, ev, (n, G))
    F = 64
    Xa = synth.normal(n, F, seed=9)
    Ya = synth.normal(n, F, seed=10)
    out = {"workload": f"knn30_200k, expression {n} x {G} ({len(cols)} stored), "
                       "host arrays in/out"}
    if with_cpu:
        out["cpu_threads"] = orc.num_threads()
    cases = {
        "impute_diffusion_t3": (
            lambda: callers.impute_diffusion(M, P, n, G, 3),
            lambda: orc.impute_diffusion_f32(g.indptr, g.indices, g.data,
                                             orc.densify_f32(ip, cols, ev, n, G), 3)),
        "impute_magic_t3": (
            lambda: callers.impute_magic(M, A, n, G, 3),
            lambda: orc.magic_f32(raw.indptr, raw.indices, raw.data,
                                  orc.densify_f32(ip, cols, ev, n, G), 3)),
        "pseudotime_k10": (
            lambda: callers.pseudotime_diffusion(A, 0, 10),
            lambda: orc.pseudotime_f32(raw.indptr, raw.indices, raw.data, 0, 10, 100)),
        "algebra_spmm_f64": (
            lambda: callers.algebra_spmm(P, Xa, Ya.copy(), 0.5, 2.0),
            lambda: orc.algebra_spmm_f32(g.indptr, g.indices, g.data, Xa, Ya, 0.5, 2.0)),
    }
    for name, (gpu_fn, cpu_fn) in cases.items():
        g_ms = host_median_ms(gpu_fn, max(3, steps // 2))
        out[name] = {"gpu_ms": round(g_ms, 2)}
        if with_cpu:
            t0 = time.perf_counter()
            cpu_fn()
            c_ms = (time.perf_counter() - t0) * 1e3
            out[name].update({"cpu_ms": round(c_ms, 1), "speedup": round(c_ms / g_ms, 1)})
    return out


def leg_kn(steps: int, with_cpu: bool = True) -> dict:
    """Building the operator through the C ABI: exact kNN of 20000 sparse
    samples x 500 features (10% stored) with k = 15 (beside the CPU
    restatement), the same scan at 60000 samples (GPU only, with its
    all-pairs FMA rate), then kernel weights, row normalisation and the
    stationary distribution on knn30_200k."""
    import numpy as np
    from scl_amd import CsrMatrix, construct, synth
    orc = None
    if with_cpu:
        from oracle import oracle as orc
    out = {"workload": "exact kNN (k=15) of sparse samples; construction steps on knn30_200k"}
    if with_cpu:
        out["cpu_threads"] = orc.num_threads()

    def samples(n, f, seed):
        rng = np.random.Generator(np.random.PCG64(seed))
        mask = rng.random((n, f), dtype=np.float32) < 0.1
        r, c = np.nonzero(mask)
        ip = np.zeros(n + 1, np.int64)
        np.add.at(ip, r + 1, 1)
        return np.cumsum(ip), c.astype(np.int64), rng.gamma(1.0, 1.0, len(c)).astype(np.float32)

    def timed(fn, reps):
        return host_median_ms(fn, reps)

    for n, f, cpu in ((20_000, 500, with_cpu), (60_000, 500, False)):
        ip, ix, v = samples(n, f, 11 + n)
        M = CsrMatrix.from_arrays(ip, ix, v, (n, f))
        g_ms = timed(lambda: construct.knn(M, 15), 3)
        f_pad = (f + 15) // 16 * 16
        # every pair's dot product over the padded features, 2 flops per FMA
        rec = {"gpu_ms": round(g_ms, 1),
               "tflops_fp32": round(2.0 * n * n * f_pad / (g_ms * 1e-3) / 1e12, 1)}
        if cpu:
            t0 = time.perf_counter()
            orc.knn_f32(ip, ix, v, 15)
            c_ms = (time.perf_counter() - t0) * 1e3
            rec.update({"cpu_ms": round(c_ms, 1), "speedup": round(c_ms / g_ms, 1)})
        out[f"knn_{n}x{f}_k15"] = rec
    raw = synth.knn_graph(200_000, 30, seed=7)
    D = CsrMatrix.from_arrays(raw.indptr, raw.indices, raw.data, (raw.n, raw.n))
    g = raw.normalized()
    P = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (g.n, g.n))
    cases = {
        "kernel_gaussian": (lambda: construct.kernel_matrix(D, 0.7, "gaussian"),
                            lambda: orc.kernel_matrix_f32(raw.data, 0.7, 0)),
        "row_normalize": (lambda: construct.row_normalize(D),
                          lambda: orc.row_normalize_f32(raw.indptr, raw.data)),
        "stationary_1e-6": (lambda: construct.stationary_distribution(P, 1000, 1e-6),
                            lambda: orc.stationary_f32(g.indptr, g.indices, g.data, 1000, 1e-6)),
    }
    for name, (gpu_fn, cpu_fn) in cases.items():
        rec = {"gpu_ms": round(timed(gpu_fn, max(3, steps // 2)), 2)}
        if with_cpu:
            t0 = time.perf_counter()
            cpu_fn()
            c_ms = (time.perf_counter() - t0) * 1e3
            rec.update({"cpu_ms": round(c_ms, 1), "speedup": round(c_ms / rec["gpu_ms"], 1)})
        out[name] = rec
    return out


def leg_c5(steps: int) -> dict:
    """256 uneven graphs, F=64, t=5, fwd+bwd: one replayed hipGraph and one
    persistent launch, each against 256 sequential calls."""
    import numpy as np
    import torch
    from scl_amd import CsrMatrix, batch, synth, torch_ops, _lib
    F, t, every = 64, 5, 1   # every iterate stored, no recompute
    sizes, ks = synth.batch_sizes(256, seed=5)
    graphs = [synth.uniform_knn(int(n), int(k), seed=5000 + i).normalized()
              for i, (n, k) in enumerate(zip(sizes, ks))]
    mats = [CsrMatrix.from_arrays(gr.indptr, gr.indices, gr.data, (gr.n, gr.n))
            for gr in graphs]
    B = batch.block_diag(mats)
    off = batch.offsets(sizes)
    N = int(off[-1])
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    X = torch.from_numpy(synth.normal(N, F, seed=5)).to(dev)
    T = torch.from_numpy(synth.normal(N, F, seed=105)).to(dev)
    torch_ops.autotune(B, X)
    torch_ops.autotune(B, X, adjoint=True)
    plan = batch.DiffusionPlan(B, N, F, t, every=every)
    plan.X.copy_(X)
    plan.T.copy_(T)
    plan.launch()
    torch.cuda.synchronize()
    plan_ms = timed(plan.launch, steps, stream)
    # the same work as ONE persistent launch (every graph pinned to an XCD,
    # its steps ordered by its own progress)
    t0 = time.perf_counter()
    bp = batch.BatchedPlan(B, [int(x) for x in sizes], F, t)
    bp_build_s = time.perf_counter() - t0
    bp.X.copy_(X)
    bp.T.copy_(T)
    bp.launch()
    bp.check()
    persist_ms = timed(bp.launch, steps, stream)
    bp.check()
    persist_same = {"Y": bool(torch.equal(bp.Y, plan.Y)), "dX": bool(torch.equal(bp.dX, plan.dX)),
                    "dP": bool(torch.equal(bp.dP[:B.nnz], plan.dP[:B.nnz]))}
    bp.close()
    # sequential: per graph, the same fwd + dY + reverse with plain launches
    lib = _lib.load()
    import ctypes
    per = []
    for i, m in enumerate(mats):
        a, b = int(off[i]), int(off[i + 1])
        n = b - a
        nb = ctypes.c_size_t(0)
        _lib.check(lib.scl_gpu_diffuse_backward_workspace(m.handle, n, F, every, 1,
                                                          ctypes.byref(nb)))
        m.prepare_device()
        per.append((m, a, b, n, torch.empty(nb.value, dtype=torch.uint8, device=dev),
                    nb.value))
    Y = torch.empty_like(X)
    dY = torch.empty_like(X)
    dX = torch.empty_like(X)
    dP = torch.empty(B.nnz, dtype=torch.float32, device=dev)
    nz = np.concatenate([[0], np.cumsum([gr.nnz for gr in graphs])])
    n_ck = (t + every - 1) // every - 1
    ck = torch.empty((n_ck, int(max(sizes)), F), dtype=torch.float32, device=dev)
    fw = torch.empty((2, int(max(sizes)), F), dtype=torch.float32, device=dev)
    sp = stream.cuda_stream

    def sequential():
        for i, (m, a, b, n, work, nbytes) in enumerate(per):
            _lib.check(lib.scl_gpu_diffuse_forward_ckpt_device(
                m.handle, X[a:b].data_ptr(), Y[a:b].data_ptr(), ck.data_ptr(),
                fw.data_ptr(), n, F, t, every, sp))
            torch.sub(Y[a:b], T[a:b], out=dY[a:b])
            _lib.check(lib.scl_gpu_diffuse_backward_device(
                m.handle, X[a:b].data_ptr(), ck.data_ptr(), dY[a:b].data_ptr(),
                dX[a:b].data_ptr(), dP[int(nz[i]):int(nz[i + 1])].data_ptr(),
                work.data_ptr(), nbytes, n, F, t, every, sp))

    sequential()
    torch.cuda.synchronize()
    seq_ms = timed(sequential, max(2, steps // 2), stream)
    same = bool(torch.equal(dX, plan.dX)) and bool(torch.equal(Y, plan.Y))
    shape = torch_ops.get_shape(B, F)
    # FP64 gradients of 8 sampled graphs (outside timing)
    from oracle import oracle as orc
    Xh, Th = X.cpu().numpy(), T.cpu().numpy()
    pdx, pdp, pY = plan.dX.cpu().numpy(), plan.dP.cpu().numpy(), plan.Y.cpu().numpy()
    rng = np.random.Generator(np.random.PCG64(55))
    ex = ep = ey = 0.0
    for i in rng.choice(256, 8, replace=False):
        a, b = int(off[i]), int(off[i + 1])
        gr = graphs[i]
        _, y64, dx64, dp64 = orc.grad_f64(gr.indptr, gr.indices, gr.data, Xh[a:b], Th[a:b], t)
        ey = max(ey, orc.rel_err(pY[a:b], y64))
        ex = max(ex, orc.rel_err(pdx[a:b], dx64))
        ep = max(ep, orc.rel_err(pdp[int(nz[i]):int(nz[i + 1])], dp64))
    plan.close()
    fwd_bytes = t * step_bytes(N, B.nnz, F)
    return {"workload": "batch256_f64_t5", "graphs": 256, "rows": N, "nnz": B.nnz,
            "plan_ms_per_fwd_bwd": round(plan_ms, 3),
            "sequential_ms_per_fwd_bwd": round(seq_ms, 3),
            "speedup_vs_sequential": round(seq_ms / plan_ms, 2),
            "persistent_ms_per_fwd_bwd": round(persist_ms, 3),
            "persistent_speedup_vs_sequential": round(seq_ms / persist_ms, 2),
            "persistent_equals_plan": persist_same,
            "persistent_plan_build_s": round(bp_build_s, 3),
            "forward_GBps_equiv": round(fwd_bytes / plan_ms / 1e6, 1),
            "autotuned_shape": shape, "results_identical": same,
            "err_vs_fp64": {"Y_max": ey, "dX_max": ex, "dP_max": ep,
                            "how": "max over 8 sampled graphs, FP64 oracle gradient"}}


def leg_cr(steps: int, shape: dict, n: int, nnz_rows, F: int,
           ranges=(2048, 16384, 32768, 65536, 200000)) -> dict:
    """The headline kernel on operators with the headline's row lengths whose
    columns are random (distinct, any spacing) within the first M rows: at
    M = 2048 every slab of them fits an XCD's L2, at 16384 all eight slabs
    fit the 256 MB Infinity Cache (125 MiB), from 65536 on they do not.  The
    times show what the kernel's L2 misses cost when served from each level."""
    import numpy as np
    import torch
    from scl_amd import CsrMatrix, synth, torch_ops
    lens = np.diff(nnz_rows)
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    Xd = torch.from_numpy(synth.normal(n, F, seed=12)).to(dev)
    out = {"workload": "headline row lengths, columns random within the first M rows, t=8",
           "ms_per_launch": {}}
    for M in ranges:
        rng = np.random.Generator(np.random.PCG64(12))
        base = rng.integers(0, M, n)
        step = 1 + 2 * rng.integers(0, max(1, M // 2), n)
        ix = np.empty(int(ip[-1]), np.int64)
        for i in range(n):
            st = int(step[i])
            while math.gcd(st, M) != 1:   # coprime to M: the columns are distinct
                st += 1
            ix[ip[i]:ip[i + 1]] = np.sort((base[i] + st * np.arange(lens[i])) % M)
        ok = all(ip[i + 1] - ip[i] == len(np.unique(ix[ip[i]:ip[i + 1]])) for i in range(0, n, 997))
        P = CsrMatrix.from_arrays(ip, ix, np.full(len(ix), 1.0 / 64, np.float32), (n, n))
        kw = dict(shape)
        torch_ops.set_shape(P, F, kw.pop("vec"), kw.pop("group"), kw.pop("unroll"), **kw)
        torch_ops.diffuse(P, Xd, 8)
        torch.cuda.synchronize()
        ms = timed(lambda: torch_ops.diffuse(P, Xd, 8), steps, stream) / 8
        out["ms_per_launch"][str(M)] = round(ms, 4)
        out.setdefault("distinct_columns_checked", True)
        out["distinct_columns_checked"] = out["distinct_columns_checked"] and bool(ok)
        log(f"leg cr: M={M} {ms:.4f} ms per launch")
        del P
        torch.cuda.empty_cache()
    out["shape"] = dict(shape)
    return out


def leg_l2(steps: int, shape: dict, n: int, nnz_rows, F: int) -> dict:
    """The headline kernel with every gather L2-resident: an operator with
    the headline's row lengths whose columns all fall in 2048 rows (a slab of
    them is 2 MiB, well inside one XCD's 4 MiB L2), run with the headline's
    shape.  Its time bounds what the kernel's gather stream can reach with
    perfect locality."""
    import numpy as np
    import torch
    from scl_amd import CsrMatrix, synth, torch_ops
    lens = np.diff(nnz_rows)
    rng = np.random.Generator(np.random.PCG64(12))
    ip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    base = rng.integers(0, 2048, n)
    step = 1 + 2 * rng.integers(0, 16, n)   # odd: distinct mod 2048
    ix = np.empty(int(ip[-1]), np.int64)
    for i in range(n):
        ix[ip[i]:ip[i + 1]] = np.sort((base[i] + step[i] * np.arange(lens[i])) % 2048)
    P = CsrMatrix.from_arrays(ip, ix, np.full(len(ix), 1.0 / 64, np.float32), (n, n))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    Xd = torch.from_numpy(synth.normal(n, F, seed=12)).to(dev)
    kw = dict(shape)
    torch_ops.set_shape(P, F, kw.pop("vec"), kw.pop("group"), kw.pop("unroll"), **kw)
    # timed as the headline is: a t = 8 diffusion (its inner steps read and
    # write slab-planar iterates), per launch
    torch_ops.diffuse(P, Xd, 8)
    torch.cuda.synchronize()
    ms = timed(lambda: torch_ops.diffuse(P, Xd, 8), steps, stream) / 8
    return {"workload": "headline row lengths, columns within 2048 rows, t=8", "nnz": len(ix),
            "shape": torch_ops.get_shape(P, F), "ms_per_launch": round(ms, 4)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", default="c2,l2,c3,c4,c5,mp,wk,cl,kn")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="vec,group,unroll,flags for the headline SpMM instead of the "
                         "autotuner (flags as SCL_GPU_SHAPE_*)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (the reference kernel "
                         "for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the "
              "multi-GPU path is a design note, not a measured configuration.",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import scl_amd
    from scl_amd import CsrMatrix, diffusion, synth, torch_ops

    n, F, t, k, seed = 200_000, 2000, 8, 30, 2
    log(f"workload knn{k}_{n} F={F} t={t}: generating graph")
    t0 = time.perf_counter()
    g = synth.knn_graph(n, k, seed=seed)
    log(f"graph nnz={g.nnz} in {time.perf_counter() - t0:.1f}s; normalising (FP64, GPU)")
    A = CsrMatrix.from_arrays(g.indptr, g.indices, g.data, (n, n))
    once = {}   # one-time costs of a new operator, in seconds
    t0 = time.perf_counter()
    A.prepare_device()          # upload + locality schedule
    once["upload_and_schedule_s"] = round(time.perf_counter() - t0, 3)
    t0 = time.perf_counter()
    deg, residual = diffusion.compute_transition_matrix_ex(A, False)
    once["normalise_and_residual_s"] = round(time.perf_counter() - t0, 3)
    log(f"stationarity residual {residual:.3e}; generating X")
    X = synth.log1p_poisson(n, F, seed=seed)
    zeros = float((X == 0).mean())
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Yd = torch.empty_like(Xd)
    Wd = torch.empty_like(Xd)
    lib = scl_amd.load()
    stream = torch.cuda.current_stream(dev)
    tune_ms = None
    if args.calibration_copy:
        # a runtime device-to-device copy (the __amd_rocclr_copyBuffer
        # kernel) of known traffic, for scaling the byte counters
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        if hip.hipMemcpy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()),
                         ctypes.c_size_t(a.numel() * 4), 3) != 0:
            raise RuntimeError("calibration copy failed")
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read, 1 GiB written")
        del a, b
    if args.shape:
        v, gr, u, fl = (int(x) for x in args.shape.split(","))
        torch_ops.set_shape(A, F, v, gr, u, nontemporal=bool(fl & 1),
                            schedule=not (fl & 2), xcd_rows=bool(fl & 4),
                            scalar_rows=bool(fl & 8), wg_per_cu=(fl >> 8) & 0xF,
                            merge=(fl >> 12) & 0xF,
                            rows_per_group=(torch_ops.TILED if fl & (1 << 18)
                                            else (0, 4, 8, 64)[(fl >> 16) & 0x3]),
                            xcd_block=(fl >> 20) & 0xF)
        log(f"shape set to {torch_ops.get_shape(A, F)}")

    def one():
        scl_amd.check(lib.scl_gpu_diffuse_matrix_device(
            A.handle, Xd.data_ptr(), Yd.data_ptr(), Wd.data_ptr(), n, F, t,
            stream.cuda_stream))

    default = None
    if not args.shape:
        # the drop-in's own choice, before any tuning: the first call builds
        # the row-grouped form of a wide operator
        t0 = time.perf_counter()
        one()
        torch.cuda.synchronize()
        once["first_call_with_form_build_s"] = round(time.perf_counter() - t0, 3)
        one()
        torch.cuda.synchronize()
        d_ms = timed(one, max(2, args.steps // 2), stream) / t
        default = {"shape": torch_ops.get_shape(A, F), "ms_per_launch": round(d_ms, 4),
                   "GBps": round(step_bytes(n, g.nnz, F) / d_ms / 1e6, 1)}
        log(f"default shape {default['shape']}: {d_ms:.4f} ms/launch")
    if not args.shape and not args.no_autotune:
        t0 = time.perf_counter()
        tune_ms = torch_ops.autotune(A, Xd)
        once["autotune_s"] = round(time.perf_counter() - t0, 2)
        log(f"autotune {once['autotune_s']}s -> {torch_ops.get_shape(A, F)}")

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        one()
    torch.cuda.synchronize()
    log(f"timing {args.steps} diffusions ({args.steps * t} SpMM launches)")
    ms_per_step = timed(one, args.steps, stream)
    ms_per_launch = ms_per_step / t
    B = step_bytes(n, g.nnz, F)
    value = t * B / (ms_per_step / 1e3) / 1e9
    achieved = B / (ms_per_launch / 1e3) / 1e9
    log(f"{ms_per_step:.3f} ms/diffusion, {ms_per_launch:.4f} ms/launch, {value:.1f} GB/s")
    shape_used = torch_ops.get_shape(A, F)

    # the drop-in host-pointer call (PCIe included), for reference: one call
    # to set up its buffers, then the median of three on fresh copies of X
    # (X is written in place)
    Xh = X.copy()
    diffusion.diffuse_matrix(A, Xh, t)
    same = bool(np.array_equal(Xh, Yd.cpu().numpy()))
    host_times = []
    for _ in range(3):
        Xh = X.copy()
        t0 = time.perf_counter()
        diffusion.diffuse_matrix(A, Xh, t)
        host_times.append((time.perf_counter() - t0) * 1e3)
    host_ms = float(np.median(host_times))
    same = same and bool(np.array_equal(Xh, Yd.cpu().numpy()))
    del Xh

    # error against FP64 and bit-identity to the CPU restatement on 8 columns
    from oracle import oracle as orc
    ip, ix, pv = A.export()
    ecols = np.array([0, 1, 250, 999, 1000, 1500, 1998, 1999])
    Yh = Yd[:, torch.from_numpy(ecols).to(dev)].cpu().numpy()
    y64 = orc.diffuse_f64(ip, ix, pv, X[:, ecols], t)
    err = {"max_rel_err_vs_fp64": orc.rel_err(Yh, y64),
           "median_rel_err_vs_fp64": orc.median_rel_err(Yh, y64),
           "bit_identical_to_cpu_restatement": bool(np.array_equal(
               Yh, orc.diffuse_matrix_f32(ip, ix, pv, X[:, ecols], t))),
           "columns": ecols.tolist()}
    del y64

    traffic = counters = None
    if TRAFFIC_FILE.exists():
        counters = json.loads(TRAFFIC_FILE.read_text())
        traffic = counters.get("hbm_bytes_per_launch")
    l2 = None
    if "l2" in args.legs.split(","):
        log("leg l2 (L2-resident ceiling of the headline kernel)")
        try:
            l2 = leg_l2(max(3, args.steps // 2), torch_ops.get_shape(A, F), n, g.indptr, F)
            l2["bound_GBps"] = round(B / l2["ms_per_launch"] / 1e6, 1)
            l2["frac_of_bound"] = round(achieved / l2["bound_GBps"], 4)
            if counters and counters.get("l2_request_bytes_per_launch"):
                l2["l2_request_bytes_per_launch"] = counters["l2_request_bytes_per_launch"]
            torch.cuda.empty_cache()
        except Exception as e:
            l2 = {"error": f"{type(e).__name__}: {e}"}
        log(f"leg l2: {l2}")

    cr = None
    if "cr" in args.legs.split(","):
        log("leg cr (column ranges: L2, Infinity Cache and HBM-served misses)")
        try:
            cr = leg_cr(max(3, args.steps // 2), torch_ops.get_shape(A, F), n, g.indptr, F)
        except Exception as e:
            cr = {"error": f"{type(e).__name__}: {e}"}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (restatement, OpenMP)")
        ip, ix, p = A.export()
        cpu = cpu_baseline((ip, ix, p), n, F, t, seed, B)

    legs = {}
    del Xd, Yd, Wd, A
    torch.cuda.empty_cache()
    for name, fn in (("c3", leg_c3), ("c4", leg_c4), ("c5", leg_c5), ("mp", leg_mp),
                     ("wk", leg_wk), ("cl", leg_cl), ("kn", leg_kn)):
        if name not in args.legs.split(","):
            continue
        log(f"leg {name}")
        try:
            if name in ("wk", "cl", "kn"):
                legs[name] = fn(max(3, args.steps // 2), not args.no_cpu_baseline)
            else:
                legs[name] = fn(max(3, args.steps // 2))
        except Exception as e:  # a failed leg is reported, not fatal
            legs[name] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
        log(f"leg {name}: {legs[name]}")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"knn{k}_{n // 1000}k_f{F}_t{t}",
            "n": n, "k": k, "nnz": g.nnz, "F": F, "t": t,
            "X": f"log1p(Poisson), {zeros:.1%} zeros",
            "bytes_per_launch": B,
            "pct_peak": round(100 * value / PEAK_HBM_GBPS, 2),
            "stationarity_residual_fp64": residual,
            "autotuned_ms_per_launch": tune_ms,
            "shape": shape_used,
            "default_without_autotune": default,
            "err": err,
            "one_time_costs": once,
            "drop_in_host_ms": round(host_ms, 1),
            "drop_in_matches_device": same,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBPS, 4),
            "traffic": traffic,
        },
        "l2_bound": l2,
        **({"column_ranges": cr} if cr is not None else {}),
        "cpu_baseline": cpu,
        "legs": legs,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
