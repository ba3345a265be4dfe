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
    while dt < budget_s and rows < n // 2:
        chunk = min(chunk, n // 2 - rows)
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--bc-sources", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-side", action="store_true", help="skip FP32 and BC side measurements")
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
    import pagani_amd
    from pagani_amd import synth

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

    for i in range(a.warmup):
        info = step(pagani_amd.PREC_BF16X3)
        log(f"warmup {i}: nnz={info.nnz} gram_emit={info.ms_gram_emit:.3f} ms")
    ctx.synchronize()
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        infos.append(step(pagani_amd.PREC_BF16X3))
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / a.steps * 1e3
    last = infos[-1]
    avg = lambda f: sum(getattr(i, f) for i in infos) / len(infos)  # noqa: E731
    emit_ms, hist_ms = avg("ms_gram_emit"), avg("ms_gram_hist")
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
        "stage_ms": {"zscore": avg("ms_zscore"), "gram_hist": hist_ms, "gram_emit": emit_ms,
                     "select": avg("ms_select"), "csr": avg("ms_csr")},
        "mfma_pct": 100.0 * mfma_frac,
        "band_pairs": last.band_pairs,
        "band_flips": last.band_flips,
        "t_star": last.t_star,
        "roofline": {"bound": "mfma", "kernel": "gram_kernel<BF16X3,EMIT>",
                     "achieved": 3.0 * useful / (emit_ms * 1e-3) / 1e12, "peak": BF16_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": mfma_frac, "traffic": None},
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
        log("cpu baseline (bounded sample)")
        cb = cpu_baseline(x, n, t, float(last.t_star))
        out["cpu_baseline"] = cb
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
