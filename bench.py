#!/usr/bin/env python3
"""Benchmark: voxel time series -> 1%-density CSR at N=58,000, T=240 on one MI355X.

A step is one subject through the construction path: z-score, histogram Gram
pass, emitting Gram pass (BF16x3 MFMA, FP64 recheck of pairs near the
threshold), exact k-th-largest selection, CSR build.  Inputs (the synthetic
time series) are resident in HBM before the timed region.

Side measurements on the same run (extra JSON fields):
  * the FP32 plain form of the same step and the BF16x3/FP32 speed ratio, with
    a check that both forms produce the identical edge set;
  * Brandes betweenness on the resulting graph from a fixed seeded sample of
    sources (GTEPS and the HBM-traffic model fraction);
  * the CPU baseline: the FP64 CPU restatement of the reference's correlation
    + threshold on a bounded sample of voxel rows, timed on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "pagani-toolkit_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

BF16_PEAK = 2516.6e12   # MI355X dense BF16 MFMA FLOP/s
FP32_PEAK = 157.3e12    # MI355X FP32 matrix FLOP/s
HBM_PEAK = 8.0e12       # bytes/s
# Floyd-Warshall updates/s at the FP32 VALU peak without packed math: 256 CUs x
# 64 lanes x 2.4 GHz = 39.3e12 ops/s, 1.5 ops per update (v_add_f32 +
# half a v_min3_f32)
FW_UPDATES_PEAK = 39.3e12 / 1.5
STAGE_STEPS = 2  # untimed steps after the timed ones, every stage timed
# HBM bytes per launch of the emitting BF16x3 Gram kernel at this workload
# (tuned supertile edge 4), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# (KiB per dispatch), each scaled by what the counter reads for a 1 GiB device
# copy (FETCH_SIZE x2, WRITE_SIZE x1); profiles/gram2_emit_n58000_t240_counters.md
GRAM_EMIT_TRAFFIC = 2683291.0 * 1024 * 2 + 253290.9 * 1024

METRIC = "per-subject wall time N=58k; corr TFLOP/s+MFMA%; BC GTEPS; BF16x3 vs FP32"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(x, n, t, thr_r, budget_s=15.0):
    """FP64 CPU restatement of the reference correlation + threshold (oracle
    port), timed on a bounded sample of voxel rows."""
    from oracle import oracle as O
    import numpy as np
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    z64, _ = O.zscore(x)
    tz = time.perf_counter() - t0
    # consecutive row blocks, each sized from the previous block's rate to
    # about 2 s, until the budget is spent
    rows, dt, chunk = 0, 0.0, 16
    while dt < budget_s and rows < n:
        chunk = min(chunk, n - rows)
        t0 = time.perf_counter()
        O.corr_count_rows(z64, thr_r, rows, rows + chunk)
        step = max(time.perf_counter() - t0, 1e-4)
        dt += step
        rows += chunk
        chunk = int(max(16, min(chunk * 2.0 / step, chunk * 8)))
    pairs = rows * (n - 1) - rows * (rows - 1) // 2
    pair_rate = pairs / dt
    full_pairs = n * (n - 1) // 2
    est_s = full_pairs / pair_rate + tz * 1.0
    return {
        "value": 2.0 * t * pair_rate / 1e12,
        "unit": "TFLOP/s",
        "cores": cores,
        "kind": "port",
        "sample": f"FP64 r + threshold for voxel rows 0..{rows} of N={n}, T={t} "
                  f"({pairs} of {full_pairs} pairs, {dt:.1f} s); z-score of all voxels {tz:.2f} s",
        "est_full_subject_s": est_s,
    }


def bc_cpu_baseline(row, col, n, budget_s=20.0):
    """The oracle's Brandes BC (the reference's BC_CPU.cpp algorithm with
    uint64 sigma and FP64 dependencies) on seeded source samples of the graph,
    one thread and every thread, extrapolated linearly to all n sources (the
    per-source cost is the same for every source of a connected graph)."""
    from oracle import oracle as O
    from pagani_amd import synth
    out = {}
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    for label, threads, count in (("1_thread", 1, 8), ("all_threads", cores, 16 * cores)):
        O.set_threads(threads)
        src = synth.sample_sources(n, count, 4242)
        t0 = time.perf_counter()
        O.bc(row, col, src)
        dt = time.perf_counter() - t0
        out[label] = {"cores": threads, "sources": count, "seconds": dt,
                      "gteps": count * float(row[-1]) / dt / 1e9, "all_sources_est_s": dt * n / count}
        if dt > budget_s:
            break
    O.set_threads(0)
    return out


def _timed(ctx, fn):
    ctx.synchronize()
    t0 = time.perf_counter()
    r = fn()
    ctx.synchronize()
    return r, time.perf_counter() - t0


def _device_csr(ctx, n, nnz, weighted=False):
    import torch
    import pagani_amd
    row = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
    val = torch.empty(max(nnz, 1), dtype=torch.float32, device="cuda") if weighted else None
    ctx._check(ctx._lib.pg_corr_csr(ctx._h, pagani_amd.lib._ptr(row), pagani_amd.lib._ptr(col),
                                    pagani_amd.lib._ptr(val)))
    return row, col, val


def run_metrics(a):
    """Graph metrics on the N=58,000 1%-density graph: exact BC from every
    source, degree, binary clustering, efficiency, eigenvector centrality,
    participation, one degree-preserving rewiring; Floyd-Warshall on the weighted graph of
    the first 8,192 voxels (w = 1 - r for r above the 1% threshold)."""
    import torch
    import pagani_amd
    from pagani_amd import synth
    w = synth.WORKLOADS["corr_n58000_t240"]
    n, t = w["n"], w["t"]
    x = synth.bold(w["seed"], n, t, w["k"])
    xd = torch.from_numpy(x).cuda().contiguous()
    ctx = pagani_amd.Context(0)
    info = ctx.corr_build(xd, n, t, 1.0, sparsity=True)
    nnz = info.nnz
    row, col, _ = _device_csr(ctx, n, nnz)
    out = {"metric": "graph metrics on the N=58000 1% graph", "n_voxels": n, "nnz": nnz, "n_gpus": 1}
    deg = torch.empty(n, dtype=torch.int32, device="cuda")
    dref = torch.empty(n, dtype=torch.float32, device="cuda")
    _, dt = _timed(ctx, lambda: ctx.degree(row, deg, dref))
    out["degree_ms"] = dt * 1e3
    tri = torch.empty(n, dtype=torch.int64, device="cuda")
    _, dt = _timed(ctx, lambda: ctx.triangles(row, col, tri))
    out["triangles_ms"] = dt * 1e3
    cp = torch.empty(n, dtype=torch.float32, device="cuda")
    (_, mean), dt = _timed(ctx, lambda: ctx.cp_binary(row, col, cp))
    out["cp_binary_ms"] = dt * 1e3
    out["cp_mean"] = mean
    log(f"degree {out['degree_ms']:.3f} ms, triangles {out['triangles_ms']:.1f} ms, cp {out['cp_binary_ms']:.1f} ms")
    bc = torch.empty(n, dtype=torch.float64, device="cuda")
    _, dt = _timed(ctx, lambda: ctx.bc(row, col, None, out=bc))
    te = float(n) * nnz
    out["bc_all_sources"] = {"sources": n, "seconds": dt, "gteps": te / dt / 1e9,
                             "hbm_model_frac": n * (16.0 * nnz + 40.0 * n) / dt / HBM_PEAK}
    log(f"bc all {n} sources: {dt:.2f} s = {te / dt / 1e9:.1f} GTEPS")
    eff = torch.empty(n, dtype=torch.float32, device="cuda")
    ctx.bfs_efficiency(row, col, eff)  # first call sizes the context's buffers
    (_, lp, iso, unc), dt = _timed(ctx, lambda: ctx.bfs_efficiency(row, col, eff))
    out["bfs_efficiency_all_sources"] = {"seconds": dt, "gteps": te / dt / 1e9, "lp": lp, "isolated": iso,
                                         "unconnected": unc}
    log(f"bfs efficiency all {n} sources: {dt * 1e3:.1f} ms = {te / dt / 1e9:.1f} GTEPS, Lp {lp:.6g}")
    ctx.eigenvector_centrality(row, col)
    (_, _, lam, it), dt = _timed(ctx, lambda: ctx.eigenvector_centrality(row, col))
    # per iteration: the column indices (4 B per stored entry) and one FP64
    # gather per entry from an L2-resident vector; 3 FP64 vectors of n
    spmv_bytes = 4.0 * nnz + 8.0 * 3 * n
    out["eigenvector_centrality"] = {"seconds": dt, "iterations": it, "eigenvalue": lam,
                                     "us_per_iteration": dt * 1e6 / max(it, 1),
                                     "hbm_frac": spmv_bytes * it / dt / HBM_PEAK}
    log(f"eigenvector centrality: {it} iterations, {dt * 1e3:.1f} ms, eigenvalue {lam:.6g}")
    lab = (torch.arange(n, dtype=torch.int32, device="cuda") % 64) + 1
    ctx.participation(row, col, lab)
    (_, _, k), dt = _timed(ctx, lambda: ctx.participation(row, col, lab))
    out["participation_ms"] = dt * 1e3
    log(f"participation coefficient, 64 modules: {dt * 1e3:.2f} ms")
    # one random reference network: 2M degree-preserving swap attempts
    rc = torch.empty_like(col)
    ctx.rewire(row, col, seed=1, out_col=rc)
    (_, _, sw), dt = _timed(ctx, lambda: ctx.rewire(row, col, seed=2, out_col=rc))
    m = nnz // 2
    out["rewire"] = {"seconds": dt, "attempts": 2 * m, "swaps": sw, "attempts_per_s": 2 * m / dt}
    log(f"rewire: {2 * m} attempts, {sw} swaps in {dt * 1e3:.1f} ms")

    # Floyd-Warshall on the first 8192 voxels: weighted graph r > t*, d = 1 - r
    nf = 8192
    xs = xd[:, :nf].contiguous()
    thr = float(info.t_star)
    winfo = ctx.corr_build(xs, nf, t, thr, weighted=True)
    wr, wc, wv = _device_csr(ctx, nf, winfo.nnz, True)
    d = torch.empty((nf, nf), dtype=torch.float32, device="cuda")
    fw = {"n": nf, "nnz": winfo.nnz, "weights": "1 - r"}
    ctx.fw_init(wr, wc, wv, 1, out=d)
    trf = ctx.autotune_fw(d)  # cached, or searched on a copy of this matrix
    fw["tuned_block"] = {"value": trf.value, "candidates_timed": trf.candidates}
    log(f"fw tile {trf.value} ({trf.candidates} candidates timed)")
    for block in (64, 128):
        ctx.fw_init(wr, wc, wv, 1, out=d)
        _, dt = _timed(ctx, lambda: ctx.fw(d, block))
        ups = float(nf) ** 3 / dt
        fw[f"block{block}"] = {"ms": dt * 1e3, "updates_per_s": ups, "valu_frac": ups / FW_UPDATES_PEAK}
        log(f"fw n={nf} block={block}: {dt * 1e3:.1f} ms")
    eff = torch.empty(nf, dtype=torch.float32, device="cuda")
    (_, lp, iso), dt = _timed(ctx, lambda: ctx.fw_efficiency(d, eff))
    fw["efficiency_ms"] = dt * 1e3
    fw["lp"] = lp
    out["fw"] = fw
    ctx.close()
    return out


def _cohort_bold(s):
    from pagani_amd import synth
    seed, n, t = synth.cohort_subject(s)
    return synth.bold(seed, n, t, 90)


def run_cohort(a):
    """A cohort of subjects, subject s with seed 4000 + s, N ~ U[20000, 70000],
    T ~ U[150, 300], K=90, 1% density each: BF16x3 correlation -> CSR ->
    degree + clustering + BC from 4096 sources (seed 4242), every subject's
    launches captured into ONE graph (pg_cohort_run) and launched as a unit.
    Per-subject latency comes from device wall-clock stamps between subjects
    inside the graph."""
    import concurrent.futures as cf
    import numpy as np
    from pagani_amd import synth
    log(f"generating {a.subjects} subjects")
    with cf.ProcessPoolExecutor(max_workers=8) as ex:  # before the GPU is touched
        xs = list(ex.map(_cohort_bold, range(a.subjects)))
    import torch
    import pagani_amd
    ctx = pagani_amd.Context(0)
    subs, keep = [], []
    flop = te = 0.0
    for s, x in enumerate(xs):
        t, n = x.shape
        k = ctx.sparsity_pairs(n, 1.0)
        src = synth.sample_sources(n, 4096, 4242)
        b = torch.from_numpy(x).cuda()
        row = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(max(2 * k, 1), dtype=torch.int32, device="cuda")
        deg = torch.empty(n, dtype=torch.int32, device="cuda")
        cp = torch.empty(n, dtype=torch.float32, device="cuda")
        bc = torch.empty(n, dtype=torch.float64, device="cuda")
        subs.append(pagani_amd.Subject(b.data_ptr(), n, t, 1.0, src.ctypes.data, src.size, row.data_ptr(),
                                       col.data_ptr(), deg.data_ptr(), cp.data_ptr(), bc.data_ptr(), 0, 0.0))
        keep.append((b, row, col, deg, cp, bc, src))
        flop += float(n) * (n - 1) * t
        te += 4096.0 * 2 * k
    del xs
    log("capture + warmup launch")
    info0 = ctx.cohort_run(subs, 1)
    log(f"{info0.graphs} graph(s) of {info0.graph_nodes} nodes built in {info0.ms_build:.0f} ms; "
        f"launch {info0.ms_graph:.1f} ms")
    info = ctx.cohort_run(subs, a.launches)
    lat = np.array([x.ms for x in subs])
    sec = info.ms_graph / 1e3
    out = {"metric": "cohort per-subject latency and throughput (captured HIP graphs)", "subjects": a.subjects,
           "n_gpus": 1, "seeds": "4000+s", "graphs": info.graphs, "graph_nodes": info.graph_nodes,
           "graph_launches": a.launches,
           "capture_instantiate_ms": info.ms_build, "graph_ms": info.ms_graph,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "subjects_per_s": a.subjects / sec, "reruns": info.reruns,
           "statuses": sorted(set(int(x.status) for x in subs)),
           "useful_corr_tflops": flop / sec / 1e12, "bc_gteps": te / sec / 1e9,
           "launch": "pg_cohort_run: captured HIP graphs of consecutive subjects, launched back to back"}
    log(f"cohort: {a.subjects} subjects in {info.ms_graph:.1f} ms = {out['subjects_per_s']:.2f} subjects/s; "
        f"p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms; {info.reruns} reruns")
    ctx.close()
    return out


def run_weighted(a):
    """N=120000, T=300, K=200: weighted network r > 0.3, Onnela clustering and
    leading-eigenvector modularity."""
    import torch
    import pagani_amd
    from pagani_amd import synth
    w = synth.WORKLOADS["weighted_n120000_t300"]
    n, t = (a.voxels or w["n"]), w["t"]
    xd = torch.from_numpy(synth.bold(w["seed"], n, t, w["k"])).cuda()
    ctx = pagani_amd.Context(0)
    info, dt = _timed(ctx, lambda: ctx.corr_build(xd, n, t, 0.3, weighted=True))
    out = {"metric": f"weighted N={n} r>0.3: construction, clustering, modularity", "n_gpus": 1,
           "n_voxels": n, "t": t, "nnz": info.nnz, "corr_ms": dt * 1e3}
    log(f"weighted corr: nnz={info.nnz} {dt * 1e3:.1f} ms")
    row, col, val = _device_csr(ctx, n, info.nnz, True)
    cpw = torch.empty(n, dtype=torch.float64, device="cuda")
    (_, mean), dt = _timed(ctx, lambda: ctx.cp_weighted(row, col, val, 2, 0, out=cpw))
    out["cp_onnela_ms"] = dt * 1e3
    out["cp_onnela_mean"] = mean
    log(f"onnela {dt * 1e3:.1f} ms")
    stops = {"rayleigh": pagani_amd.MOD_STOP_RAYLEIGH, "vector": pagani_amd.MOD_STOP_VECTOR}
    for name in a.mod_stop.split(","):
        log(f"modularity ({name} stop) ...")
        (lab, q, lam, eig, rounds), dt = _timed(
            ctx, lambda: ctx.modularity(row, col, val, pagani_amd.MOD_STRENGTH, stop=stops[name], progress=True))
        out[f"modularity_{name}"] = {"seconds": dt, "q": q, "lambda0": lam, "rounds": rounds,
                                     "communities": int(lab.max())}
        log(f"modularity ({name} stop) {dt:.2f} s, Q={q:.4f}, {int(lab.max())} communities, {rounds} rounds")
    ctx.close()
    return out


FP64_PEAK = 78.6e12  # MI355X dense FP64 MFMA FLOP/s


def run_average(a):
    """Subject averaging of the dense matrix at N=58,000, T=240: per subject the
    FP64 MFMA Gram with the exact FP32 rounding and the packed accumulate (Fisher
    z), then the finish and both threshold kinds on the averaged matrix."""
    import torch
    import pagani_amd
    from pagani_amd import synth
    w = synth.WORKLOADS["corr_n58000_t240"]
    n, t = w["n"], w["t"]
    ns = max(2, min(a.subjects, 4))
    ctx = pagani_amd.Context(0)
    ctx.cormat_begin(n, True)
    per = []
    for s in range(ns):
        x = torch.from_numpy(synth.bold(w["seed"] + s, n, t, w["k"])).cuda().contiguous()
        torch.cuda.synchronize()
        _, dt = _timed(ctx, lambda: ctx.cormat_add(x, t))
        per.append(dt)
        log(f"subject {s}: FP64 Gram + accumulate {dt * 1e3:.1f} ms")
    flop = float(n) * (n - 1) / 2 * t * 2
    _, dfin = _timed(ctx, ctx.cormat_finish)
    _, dr = _timed(ctx, lambda: ctx.cormat_threshold(0.3))
    info_r = ctx.cormat_threshold(0.3)
    _, ds = _timed(ctx, lambda: ctx.cormat_threshold(1.0, sparsity=True))
    info_s = ctx.cormat_threshold(1.0, sparsity=True)
    best = min(per[1:]) if len(per) > 1 else per[0]
    out = {"metric": "subject averaging of the dense matrix (Fisher z), N=58000, T=240", "n_voxels": n,
           "subjects": ns, "ms_per_subject": [p * 1e3 for p in per], "gram64_tflops": flop / best / 1e12,
           "gram64_fp64_mfma_frac": flop / best / FP64_PEAK, "finish_ms": dfin * 1e3,
           "threshold_r_ms": dr * 1e3, "threshold_r_nnz": info_r.nnz, "threshold_s_ms": ds * 1e3,
           "threshold_s_r": info_s.r_thresh, "n_gpus": 1}
    log(f"finish {dfin * 1e3:.1f} ms, r-threshold {dr * 1e3:.1f} ms, 1% sparsity {ds * 1e3:.1f} ms")
    ctx.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--bc-sources", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="skip FP32 and BC side measurements")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only copy 1 GiB device to device 5 times (scale for HBM byte counters)")
    ap.add_argument("--config", choices=["corr", "metrics", "cohort", "weighted", "average"], default="corr",
                    help="corr: the headline line (default); metrics: BC from all sources, clustering "
                         "and N=8192 Floyd-Warshall on the 1%% graph; cohort: 64 subjects of uneven N, T; "
                         "weighted: N=120000 r>0.3 weighted clustering + modularity")
    ap.add_argument("--subjects", type=int, default=64, help="cohort size (--config cohort)")
    ap.add_argument("--launches", type=int, default=1, help="graph launches timed (--config cohort)")
    ap.add_argument("--voxels", type=int, default=0, help="voxel count override (--config weighted)")
    ap.add_argument("--mod-stop", default="rayleigh",
                    help="modularity stopping rules to run, comma separated: rayleigh, vector (--config weighted)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the construction path "
              "is single-GPU in this build.", file=sys.stderr)
        return 2
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py is single-process; do not launch it under torch.distributed", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    if a.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 5
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                          "copy_ms": dt * 1e3, "copy_GBps_rw": 2.0 * (1 << 30) / dt / 1e9}), flush=True)
        return 0
    import pagani_amd
    from pagani_amd import synth

    if a.config != "corr":
        out = {"metrics": run_metrics, "cohort": run_cohort, "weighted": run_weighted,
               "average": run_average}[a.config](a)
        print(json.dumps(out), flush=True)
        return 0

    w = synth.WORKLOADS["corr_n58000_t240"]
    n, t, k_comm, seed = w["n"], w["t"], w["k"], w["seed"]
    sparsity = 1.0
    log(f"generating synthetic series N={n} T={t} K={k_comm} seed={seed}")
    x = synth.bold(seed, n, t, k_comm)
    xd = torch.from_numpy(x).cuda().contiguous()
    torch.cuda.synchronize()
    ctx = pagani_amd.Context(0)
    useful = float(n) * (n - 1) * t  # FLOP of one pass over all i<j pairs

    def step(prec):
        return ctx.corr_build(xd, n, t, sparsity, sparsity=True, prec=prec)

    # launch shape of the Gram walk: cached, or searched on this input
    tr = ctx.autotune_corr(xd, n, t, sparsity, sparsity=True)
    log(f"gram supertile edge {tr.value} ({tr.candidates} candidates timed)")
    tuned = {"gram_super": {"value": tr.value, "candidates_timed": tr.candidates,
                            "emit_ms": tr.ms if tr.candidates else None,
                            "default_emit_ms": tr.ms_default if tr.candidates else None}}
    for i in range(a.warmup):
        info = step(pagani_amd.PREC_BF16X3)
        log(f"warmup {i}: nnz={info.nnz} gram_emit={info.ms_gram_emit:.3f} ms")
    ctx.synchronize()
    torch.cuda.synchronize()
    # timed steps record only the emitting Gram kernel's events (each other
    # event idles the device for a few microseconds); the stage breakdown
    # comes from untimed steps with every stage timed
    ctx.set_timing(1)
    infos = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        infos.append(step(pagani_amd.PREC_BF16X3))
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_timing(2)
    stage_infos = [step(pagani_amd.PREC_BF16X3) for _ in range(STAGE_STEPS)]
    ms_step = dt / a.steps * 1e3
    last = infos[-1]
    avg = lambda f: sum(getattr(i, f) for i in infos) / len(infos)  # noqa: E731
    savg = lambda f: sum(getattr(i, f) for i in stage_infos) / len(stage_infos)  # noqa: E731
    emit_ms, hist_ms = avg("ms_gram_emit"), savg("ms_gram_hist")
    mfma_frac = 3.0 * useful / (emit_ms * 1e-3) / BF16_PEAK
    log(f"bf16x3: {ms_step:.3f} ms/subject, emit {emit_ms:.3f} ms, hist {hist_ms:.3f} ms, "
        f"nnz={last.nnz}, band_pairs={last.band_pairs}, flips={last.band_flips}")

    out = {
        "metric": METRIC,
        "value": useful / (ms_step * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+fp64",
        "data": "synthetic latent-community series (seeded), resident in HBM",
        "config": {"workload": "corr_n58000_t240_k90_sparsity1pct", "n_voxels": n, "t": t,
                   "communities": k_comm, "sparsity_pct": sparsity, "precision": "bf16x3",
                   "edges_undirected": last.nnz // 2, "parallelism": "single-gpu"},
        "per_subject_ms": ms_step,
        "stage_ms": {"zscore": savg("ms_zscore"), "gram_hist": hist_ms, "gram_emit": emit_ms,
                     "select": savg("ms_select"), "csr": savg("ms_csr")},
        "stage_ms_source": f"gram_emit: the timed steps' own kernel events; the others: "
                           f"{STAGE_STEPS} untimed steps with every stage timed",
        "mfma_pct": 100.0 * mfma_frac,
        "band_pairs": last.band_pairs,
        "band_flips": last.band_flips,
        "t_star": last.t_star,
        "roofline": {"bound": "mfma", "kernel": "gram2_kernel<EMIT> (BF16x3)",
                     "achieved": 3.0 * useful / (emit_ms * 1e-3) / 1e12, "peak": BF16_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": mfma_frac, "traffic": GRAM_EMIT_TRAFFIC,
                     "traffic_unit": "bytes per launch",
                     "algorithmic_flop": 3.0 * useful, "duration_ms": emit_ms},
        "tuned": tuned,
    }

    if not a.no_side:
        row_b = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col_b = torch.empty(last.nnz, dtype=torch.int32, device="cuda")
        ctx._check(ctx._lib.pg_corr_csr(ctx._h, pagani_amd.lib._ptr(row_b), pagani_amd.lib._ptr(col_b), None))
        f_infos = []
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.fp32_steps):
            f_infos.append(step(pagani_amd.PREC_FP32))
        ctx.synchronize()
        f_ms = (time.perf_counter() - t0) / a.fp32_steps * 1e3
        f_emit = sum(i.ms_gram_emit for i in f_infos) / len(f_infos)
        row_f = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col_f = torch.empty(f_infos[-1].nnz, dtype=torch.int32, device="cuda")
        ctx._check(ctx._lib.pg_corr_csr(ctx._h, pagani_amd.lib._ptr(row_f), pagani_amd.lib._ptr(col_f), None))
        same = bool(torch.equal(row_b, row_f) and torch.equal(col_b, col_f))
        log(f"fp32: {f_ms:.3f} ms/subject, emit {f_emit:.3f} ms, identical edge set: {same}")
        out["fp32_plain"] = {"per_subject_ms": f_ms, "corr_tflops": useful / (f_ms * 1e-3) / 1e12,
                             "gram_emit_ms": f_emit,
                             "gram_emit_frac_of_fp32_peak": useful / (f_emit * 1e-3) / FP32_PEAK}
        out["bf16x3_vs_fp32"] = {"per_subject": f_ms / ms_step, "gram_emit_kernel": f_emit / emit_ms,
                                 "edge_set_identical": same}

        # betweenness on this graph from a seeded sample of sources
        src = synth.sample_sources(n, a.bc_sources, 4242)
        bc = torch.empty(n, dtype=torch.float64, device="cuda")
        nnz = int(row_b[-1].item())
        srcd = torch.from_numpy(src).cuda()
        trb = ctx.autotune_bc(row_b, col_b, srcd)  # cached, or searched on this graph
        out["tuned"]["bc_batches"] = {"value": trb.value, "candidates_timed": trb.candidates,
                                      "best_s": trb.ms / 1e3 if trb.candidates else None}
        ctx.bc(row_b, col_b, srcd, out=bc)  # warm
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.bc(row_b, col_b, srcd, out=bc)
        ctx.synchronize()
        bdt = time.perf_counter() - t0
        te = float(a.bc_sources) * nnz
        bytes_model = a.bc_sources * (16.0 * nnz + 40.0 * n)
        out["bc"] = {"sources": a.bc_sources, "seconds": bdt, "gteps": te / bdt / 1e9,
                     "hbm_model_frac": bytes_model / bdt / HBM_PEAK,
                     "all_sources_est_s": bdt * n / a.bc_sources}
        log(f"bc: {a.bc_sources} sources in {bdt:.3f} s = {te / bdt / 1e9:.1f} GTEPS")
        if not a.no_cpu_baseline:
            log("bc cpu baseline (bounded source samples)")
            out["bc"]["cpu_baseline"] = bc_cpu_baseline(row_b.cpu().numpy(), col_b.cpu().numpy(), n)

    if not a.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        cb = cpu_baseline(x, n, t, float(last.t_star))
        out["cpu_baseline"] = cb
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
