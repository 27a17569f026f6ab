"""Benchmark: colorings/s of the PGBSC color-coding DP on one MI355X.

Workload (one "step" = one coloring through the whole DP, the reference's
plain method): R-MAT scale 20, edge factor 8 (a=.57 b=.19 c=.19, seed 1,
self loops and duplicates removed, symmetrized into CSC), 7-vertex tree
0-1,1-2,2-3,1-4,4-5,4-6, FP32 tables with the reference's numerics.
Inputs are resident in HBM before the timed region; each step draws a fresh
coloring on the device.

Printed JSON adds:
  roofline      dominant kernel's algorithmic bytes / its HIP-event time
  cpu_baseline  the compiled reference (oracle/_ref/pg_ref) on this host's
                cores, same graph and colorings, with per-coloring parity
  batched       the same graph with 64 colorings per call (extra RHS columns)
  fp64          the plain line with FP64 tables (--no-batched skips both)

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batched]
                    [--plain-settings KEY=VAL,...] [--autotune 0|1]

The plain line runs with fixed context options (--plain-settings; by default
the ones the library's autotuner settles on for this shape: staged neighbour
rows, hubs above degree 512 in 1024-entry chunks on 64-lane teams, 16-lane
teams from degree 64, 8-lane teams below), so that its
kernel mix, and the rocprofv3 summaries taken of it, do not change with the
tuner's timing noise.  --autotune 1 tunes the plain shape instead, in one call
before the warmup steps.  The batched line is autotuned.  config.settings
records what ran.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TREE7 = [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def physical_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(n_decl, src, dst, seed, ncol, gpu_raw_fn):
    """The compiled reference timed on host cores, or the C oracle if absent."""
    import oracle
    threads = physical_cores()
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        gp, tp = os.path.join(tmp, "g.el"), os.path.join(tmp, "t.fascia")
        with open(gp, "w") as f:
            f.write(f"{n_decl}\n{len(src)}\n")
            step = 1 << 20
            for i in range(0, len(src), step):
                f.write("".join(f"{a} {b}\n" for a, b in zip(src[i:i + step].tolist(), dst[i:i + step].tolist())))
        with open(tp, "w") as f:
            f.write(f"7\n6\n" + "".join(f"{a} {b}\n" for a, b in TREE7))
        if oracle.have_ref():
            raws, secs = oracle.run_ref(gp, tp, "-", ncol, threads=threads, seed=seed, omp_threads=threads,
                                        timeout=600)
            kind = "reference"
        else:
            import subgraph_amd as sg
            n, s2, d2 = sg.remove_holes(n_decl, src, dst)
            colptr, rowidx = oracle.csc_build(n, s2, d2)
            plan = oracle.Plan(7, TREE7)
            cols = sg.colors_fill(seed, 0, ncol, n, 7)
            raws, secs = [], []
            for i in range(ncol):
                t0 = time.perf_counter()
                raws.append(oracle.count_f32(plan, colptr, rowidx, cols[i]))
                secs.append(time.perf_counter() - t0)
            kind = "port"
    gpu = gpu_raw_fn(ncol)
    err = max(abs(g - r) / abs(r) for g, r in zip(gpu, raws))
    return {"value": ncol / sum(secs), "unit": "colorings/s", "cores": threads, "kind": kind,
            "sample": f"{ncol} colorings (stream seed {seed}) of the bench workload, per-coloring time as "
                      f"CountMat::colorCounting; graph load excluded",
            "seconds_per_coloring": sum(secs) / ncol, "parity_max_rel_err": err}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_plain_rmat_s20_ef8_tree7.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy kernel
    of known size), or None when that file does not cover the kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k["hbm_bytes_per_launch"]


def _timed(fn, reps, sync):
    fn()
    sync()
    t0 = time.perf_counter()
    for r in range(reps):
        fn(r)
    sync()
    return (time.perf_counter() - t0) / reps


def fwd_bwd_variant(sync):
    """Differentiation workload: R-MAT s18/ef16 with symmetric weights
    U(0.5,1.5), 5-star and 5-path, 16 colorings per call, forward + adjoint
    (per-entry gradient), beside the plain one-coloring forward on the same
    graph, autotuned.  FP32 tables."""
    import subgraph_amd as sg
    rng = np.random.default_rng(18)
    n = 1 << 18
    src, dst = sg.rmat_edges(18, 16, seed=1)
    w_und = rng.uniform(0.5, 1.5, len(src)).astype(np.float32)
    colptr, rowidx = sg.csc_from_edges(n, src, dst)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    key = np.minimum(rows, rowidx).astype(np.int64) * n + np.maximum(rows, rowidx)
    ukey = src.astype(np.int64) * n + dst
    order = np.argsort(ukey)
    w = w_und[order[np.searchsorted(ukey[order], key)]]
    ctx = sg.Context(0)
    ctx.upload_graph(colptr, rowidx, vals=w)
    ctx.set_option("autotune", 1)  # tuned per (plan, tile) in each measurement's untimed first call
    out = {"graph": "R-MAT scale 18 edge factor 16 (seed 1), weights U(0.5,1.5) symmetric",
           "nnz": int(len(rowidx)), "colorings_per_call": 16, "dtype": "fp32", "templates": {}}
    import torch
    grad = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")  # the autograd path's output
    for name, te in (("star5", [(0, 1), (0, 2), (0, 3), (0, 4)]), ("u5", [(0, 1), (1, 2), (2, 3), (3, 4)])):
        plan = ctx.plan(5, te)
        t_plain = _timed(lambda r=0: ctx.count(plan, 1, seed=50 + r), 10, sync)
        t_fwd = _timed(lambda r=0: ctx.count(plan, 16, seed=60 + r), 5, sync)
        t_fb1 = _timed(lambda r=0: ctx.count_grad_device(plan, 1, grad.data_ptr(), seed=80 + r), 10, sync)
        t_fb = _timed(lambda r=0: ctx.count_grad_device(plan, 16, grad.data_ptr(), seed=70 + r), 5, sync)
        t_fbh = _timed(lambda r=0: ctx.count_grad(plan, 16, seed=90 + r, precision="fp32"), 3, sync)
        rec = {"plain_fwd_colorings_per_s": 1.0 / t_plain, "batched_fwd_colorings_per_s": 16 / t_fwd,
               "plain_fwd_bwd_colorings_per_s": 1.0 / t_fb1, "fwd_bwd_colorings_per_s": 16 / t_fb,
               "fwd_bwd_host_grad_colorings_per_s": 16 / t_fbh}
        rec["fwd_bwd_vs_plain_fwd_bwd"] = rec["fwd_bwd_colorings_per_s"] / rec["plain_fwd_bwd_colorings_per_s"]
        rec["fwd_bwd_vs_plain_fwd"] = rec["fwd_bwd_colorings_per_s"] / rec["plain_fwd_colorings_per_s"]
        out["templates"][name] = rec
        log(f"fwd+bwd s18 {name}: plain fwd {rec['plain_fwd_colorings_per_s']:.0f}/s, batched fwd "
            f"{rec['batched_fwd_colorings_per_s']:.0f}/s, plain fwd+bwd {rec['plain_fwd_bwd_colorings_per_s']:.0f}/s, "
            f"fwd+bwd {rec['fwd_bwd_colorings_per_s']:.0f}/s (gradient to host memory: "
            f"{rec['fwd_bwd_host_grad_colorings_per_s']:.0f}/s)")
    out["note"] = ("fwd+bwd: the gradient stays on the device as float32 (the torch.autograd path); "
                   "host_grad: the same call returning FP64 gradients in host memory (a PCIe copy of 8 B per "
                   "entry per call)")
    return out


TREE10 = [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]
TREE12 = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]
U5 = [(0, 1), (1, 2), (2, 3), (3, 4)]


def hashed_weights(n, colptr, rowidx):
    """Symmetric entry weights U(0.5, 1.5) in FP32: a splitmix64 hash of the
    undirected edge {u, v}, so both entries of an edge get the same value."""
    rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(colptr).astype(np.int64))
    cols = rowidx.astype(np.uint64)
    x = np.minimum(rows, cols) * np.uint64(n) + np.maximum(rows, cols)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (0.5 + (x >> np.uint64(40)).astype(np.float64) / float(1 << 24)).astype(np.float32)


def dominant_kernel(ctx, run):
    """Run once with per-launch HIP events; the dominant kernel's name,
    launches, ms per launch and algorithmic GB/s."""
    ctx.set_option("timing", 1)
    try:
        run()
        kern = ctx.last_timing_kernels()
    finally:
        ctx.set_option("timing", 0)
    name, (ms, nbytes, nl) = max(kern.items(), key=lambda kv: kv[1][0])
    return {"kernel": name, "launches": nl, "avg_launch_ms": ms / nl,
            "achieved_gbs": nbytes / (ms * 1e-3) / 1e9, "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "call_bytes": sum(v[1] for v in kern.values())}


def call_roofline(ctx, run, seconds_per_call):
    """Whole-call figure: the algorithmic bytes of every launch of one call
    (recorded in an eager run with per-launch events) over the call's time
    when replayed from its captured graphs, and the dominant kernel."""
    dom = dominant_kernel(ctx, run)
    gbs = dom["call_bytes"] / seconds_per_call / 1e9
    return {"algorithmic_bytes_per_call": dom["call_bytes"], "call_gbs": gbs, "call_frac": gbs / HBM_PEAK_GBS,
            "dominant": {k: v for k, v in dom.items() if k != "call_bytes"}}


TUNED_KEYS = ("reg_cols", "stage_rows", "aux_rows", "bucket_high", "bucket_mid", "hub_degree", "hub_chunk",
              "batch_tile")


def s22_variant(sync, with_grad=True):
    """Batching on R-MAT scale 22, edge factor 16 (seed 1, dedup,
    symmetrized): 64 colorings per pass of the 5-path, the 7-vertex tree and
    the 10-vertex tree (autotuned), each beside its plain one-coloring pass,
    then an uneven mixed batch of the three templates in one call; then the
    12-vertex tree forward and forward+backward at unit and U(0.5,1.5)
    weights, 16 colorings per call and one (the plain method), autotuned,
    with the %HBM of whole calls.  FP32 tables."""
    import subgraph_amd as sg
    t0 = time.perf_counter()
    n = 1 << 22
    src, dst = sg.rmat_edges(22, 16, seed=1)
    colptr, rowidx = sg.csc_from_edges(n, src, dst)
    del src, dst
    ctx = sg.Context(0)
    ctx.upload_graph(colptr, rowidx)
    st = ctx.stats()
    log(f"s22: n={n} nnz={len(rowidx)} n_active={st['n_active']} ({time.perf_counter() - t0:.1f}s)")
    out = {"graph": "R-MAT scale 22 edge factor 16 (a=.57 b=.19 c=.19 seed 1), dedup, symmetrized",
           "n": n, "nnz": int(len(rowidx)), "n_active": st["n_active"], "dtype": "fp32", "templates": {}}
    ctx.set_option("autotune", 1)
    plans = {}
    for name, nv, te, reps in (("u5", 5, U5, 3), ("tree7", 7, TREE7, 2), ("tree10", 10, TREE10, 1)):
        plan = plans[name] = ctx.plan(nv, te)
        B = 64
        t1 = time.perf_counter()
        ctx.count(plan, B, seed=11)  # first call: autotuning of this (plan, tile)
        sync()
        tune_s = time.perf_counter() - t1
        tuned = {k: ctx.get_option(k) for k in TUNED_KEYS}
        tuned["colorings_per_pass"] = ctx.get_option("last_tile")
        t_b = _timed(lambda r=0: ctx.count(plan, B, seed=200 + r), reps, sync)
        ctx.set_option("autotune", 0)
        t_p = _timed(lambda r=0: ctx.count(plan, 1, seed=300 + r), 3, sync)
        ctx.set_option("autotune", 1)
        dom = dominant_kernel(ctx, lambda: ctx.count(plan, B, seed=12))
        rec = {"batched_colorings_per_call": B, "batched_colorings_per_s": B / t_b,
               "plain_colorings_per_s": 1.0 / t_p, "batched_vs_plain": (B / t_b) * t_p,
               "tuned": tuned, "tuning_seconds": tune_s, "roofline_dominant": dom}
        out["templates"][name] = rec
        log(f"s22 {name}: batched {B / t_b:.1f}/s, plain {1 / t_p:.1f}/s ({rec['batched_vs_plain']:.2f}x), "
            f"dominant {dom['kernel']} {dom['achieved_gbs']:.0f} GB/s, tuned {tuned} in {tune_s:.1f}s")
    counts = [64, 40, 16]
    pl = [plans["u5"], plans["tree7"], plans["tree10"]]
    ctx.count_multi(pl, counts, [21, 22, 23])
    t_m = _timed(lambda r=0: ctx.count_multi(pl, counts, [31 + r, 41 + r, 51 + r]), 1, sync)
    out["mixed"] = {"templates": ["u5", "tree7", "tree10"], "colorings": counts,
                    "colorings_per_s": sum(counts) / t_m, "seconds": t_m}
    log(f"s22 mixed {counts}: {sum(counts) / t_m:.1f} colorings/s")
    if with_grad:
        import torch
        ctx.set_option("autotune", 1)  # forward and gradient calls: settings tuned per (plan, tile)
        plan = ctx.plan(12, TREE12)
        rec = {}
        for wname in ("unit", "uniform_0.5_1.5"):
            if wname == "unit":
                ctx.set_weights(None)
            else:
                ctx.set_weights(hashed_weights(n, colptr, rowidx))
            t_f = _timed(lambda r=0: ctx.count(plan, 16, seed=400 + r), 1, sync)
            fwd_tile = ctx.get_option("last_tile")
            tuned = {k: ctx.get_option(k) for k in TUNED_KEYS}
            gbuf = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")  # the autograd path's output
            t_fb = _timed(lambda r=0: ctx.count_grad_device(plan, 16, gbuf.data_ptr(), seed=500 + r), 1, sync)
            rec[wname] = {"fwd_colorings_per_s": 16 / t_f, "fwd_bwd_colorings_per_s": 16 / t_fb,
                          "fwd_bwd_vs_fwd": t_f / t_fb, "fwd_colorings_per_pass": fwd_tile,
                          "fwd_bwd_colorings_per_pass": ctx.get_option("last_tile"),
                          "adjoint_tables_recomputed": bool(ctx.get_option("last_recompute")),
                          "tuned_fwd": tuned}
            # the same with every table rebuilt from its subtree in the reverse sweep
            ctx.set_option("adjoint_recompute", 1)
            t_rc = _timed(lambda r=0: ctx.count_grad_device(plan, 16, gbuf.data_ptr(), seed=600 + r), 1, sync)
            rec[wname]["fwd_bwd_recompute_colorings_per_s"] = 16 / t_rc
            rec[wname]["fwd_bwd_recompute_colorings_per_pass"] = ctx.get_option("last_tile")
            ctx.set_option("adjoint_recompute", -1)
            # the plain method: one coloring per call, forward and forward+backward
            t_pf = _timed(lambda r=0: ctx.count(plan, 1, seed=700 + r), 2, sync)
            t_pfb = _timed(lambda r=0: ctx.count_grad_device(plan, 1, gbuf.data_ptr(), seed=800 + r), 2, sync)
            rec[wname].update({"plain_fwd_colorings_per_s": 1 / t_pf, "plain_fwd_bwd_colorings_per_s": 1 / t_pfb,
                               "fwd_vs_plain_fwd": t_pf * 16 / t_f, "fwd_bwd_vs_plain_fwd_bwd": t_pfb * 16 / t_fb})
            log(f"s22 tree12 {wname}: fwd {16 / t_f:.2f}/s ({fwd_tile}/pass; plain {1 / t_pf:.2f}/s), "
                f"fwd+bwd {16 / t_fb:.2f}/s ({rec[wname]['fwd_bwd_colorings_per_pass']}/pass, recompute "
                f"{rec[wname]['adjoint_tables_recomputed']}; plain {1 / t_pfb:.2f}/s); tables rebuilt: "
                f"{16 / t_rc:.2f}/s ({rec[wname]['fwd_bwd_recompute_colorings_per_pass']}/pass)")
            if wname == "unit":
                # %HBM of whole calls, batched and plain
                roof = {"fwd": call_roofline(ctx, lambda: ctx.count(plan, 16, seed=401), t_f),
                        "fwd_bwd": call_roofline(
                            ctx, lambda: ctx.count_grad_device(plan, 16, gbuf.data_ptr(), seed=501), t_fb),
                        "plain_fwd": call_roofline(ctx, lambda: ctx.count(plan, 1, seed=701), t_pf),
                        "plain_fwd_bwd": call_roofline(
                            ctx, lambda: ctx.count_grad_device(plan, 1, gbuf.data_ptr(), seed=801), t_pfb)}
                rec[wname]["roofline"] = roof
                log("s22 tree12 unit %HBM of whole calls: " + ", ".join(
                    f"{k} {v['call_gbs']:.0f} GB/s ({100 * v['call_frac']:.2f}%)" for k, v in roof.items()))
        out["tree12_fwd_bwd"] = {"colorings_per_call": 16, **rec,
                                 "note": "roofline: algorithmic bytes of one call (forward: each table written "
                                         "and read once, the CSC once per gathering launch; backward: the same "
                                         "units plus the FP64 per-entry gradient read and written by each launch "
                                         "that adds to it) over the call's graph-replayed time"}
    return out


def copy_calibration(mib):
    """One elementwise kernel of known traffic: reads and writes `mib` MiB."""
    import torch
    src_t = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst_t = torch.empty_like(src_t)
    torch.add(src_t, 1.0, out=dst_t)
    torch.cuda.synchronize()
    log(f"copy calibration: {mib} MiB read + {mib} MiB written")


# settings the autotuner picked for the 12-vertex tree in an s22 measurement; between
# runs it also picks direct loads or hubs from 1024, within about 1% (every gather of
# this template runs 64-lane teams, so the bucket thresholds do not matter)
TREE12_SETTINGS = {"stage_rows": 1, "bucket_high": 32, "bucket_mid": 8, "hub_degree": 512, "hub_chunk": 1024}


def tree12_variant(sync):
    """The 12-vertex tree on R-MAT s22/ef16 alone, unit weights, with fixed
    settings (TREE12_SETTINGS; 4 colorings per forward pass, 3 per adjoint
    pass, as tuned): 16 colorings forward, 16 forward+backward, and one of
    each (the plain method).  A short run with no tuning passes, for
    profilers.  FP32 tables."""
    import torch
    import subgraph_amd as sg
    n = 1 << 22
    src, dst = sg.rmat_edges(22, 16, seed=1)
    colptr, rowidx = sg.csc_from_edges(n, src, dst)
    del src, dst
    ctx = sg.Context(0, **TREE12_SETTINGS)
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(12, TREE12)
    gbuf = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")
    out = {"graph": "R-MAT scale 22 edge factor 16 (a=.57 b=.19 c=.19 seed 1), dedup, symmetrized",
           "template": "12-vertex tree", "settings": dict(TREE12_SETTINGS)}
    for name, tile, N, fn in (
            ("fwd", 4, 16, lambda r=0: ctx.count(plan, 16, seed=400 + r)),
            ("fwd_bwd", 3, 16, lambda r=0: ctx.count_grad_device(plan, 16, gbuf.data_ptr(), seed=500 + r)),
            ("plain_fwd", 1, 1, lambda r=0: ctx.count(plan, 1, seed=700 + r)),
            ("plain_fwd_bwd", 1, 1, lambda r=0: ctx.count_grad_device(plan, 1, gbuf.data_ptr(), seed=800 + r))):
        ctx.set_option("batch_tile", tile)
        t = _timed(fn, 1 if N > 1 else 2, sync)
        out[name] = {"colorings_per_s": N / t, "colorings_per_pass": ctx.get_option("last_tile")}
        log(f"tree12 {name}: {N / t:.2f} colorings/s ({ctx.get_option('last_tile')}/pass)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--cpu-colorings", type=int, default=12)
    ap.add_argument("--variants", default="grad,s22",
                    help="comma list of extra measurements: grad, s22, tree12 (the s22 12-tree part alone), none")
    ap.add_argument("--plain-settings",
                    default="stage_rows=1,bucket_high=512,bucket_mid=64,hub_degree=512,hub_chunk=1024",
                    metavar="KEY=VAL,...", help="context options of the plain line ('' = library defaults)")
    ap.add_argument("--variants-only", action="store_true",
                    help="run only the --variants measurements (a short run for profilers; no primary line)")
    ap.add_argument("--autotune", type=int, default=0, choices=(0, 1),
                    help="1: the autotuner picks the plain line's settings, in a call before the warmup steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one elementwise kernel that reads and writes MIB MiB (the known-size "
                         "kernel rocprofv3 FETCH_SIZE / WRITE_SIZE readings are scaled by)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU: bench.py supports --gpus 1 only; "
                                   "colorings would shard across GPUs, each rank counting its share of the coloring stream"}))
        sys.exit(2)

    import subgraph_amd as sg

    if args.variants_only:
        import torch
        result = {"metric": "colorings/sec + %HBM roofline; batched and fwd+bwd vs plain DP, same GPU",
                  "note": "--variants-only: extra measurements only, no primary line"}
        if args.copy_calibration > 0:
            copy_calibration(args.copy_calibration)
        run_variants(result, args.variants.split(","), torch.cuda.synchronize)
        print(json.dumps(result), flush=True)
        return

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)

    scale, ef = 20, 8
    t0 = time.perf_counter()
    src, dst = sg.rmat_edges(scale, ef, seed=1)
    n_decl = 1 << scale
    n, s2, d2 = sg.remove_holes(n_decl, src, dst)  # the reference text loader's rule
    colptr, rowidx = sg.csc_from_edges(n, s2, d2)
    log(f"graph: n={n} undirected={len(src)} nnz={len(rowidx)} ({time.perf_counter() - t0:.1f}s)")

    ctx = sg.Context(0)
    for kv in filter(None, args.plain_settings.split(",")):
        key, val = kv.split("=")
        ctx.set_option(key.strip(), int(val))
    ctx.upload_graph(colptr, rowidx)
    st = ctx.stats()
    plan = ctx.plan(7, TREE7)
    log(f"uploaded: n_active={st['n_active']} max_degree={st['max_degree']}")

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    tuning_s = 0.0
    if args.autotune:
        ctx.set_option("autotune", 1)
        t0 = time.perf_counter()
        ctx.count(plan, 1, seed=999)  # tunes this shape once, before the warmup and the timed region
        sync()
        tuning_s = time.perf_counter() - t0
    settings = {k: ctx.get_option(k) for k in TUNED_KEYS}
    log(f"plain settings ({'autotuned in %.1fs' % tuning_s if args.autotune else 'as set'}): {settings}")
    for w in range(args.warmup):
        ctx.count(plan, 1, seed=1000 + w)
    # timed region: the product path (each pass replays its captured hipGraph)
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.count(plan, 1, seed=2000 + k)
    sync()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    # roofline region: the same K steps launched eagerly with a HIP event pair
    # around every kernel on the library's stream (graphs carry no events)
    ctx.set_option("timing", 1)
    agg = {}
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.count(plan, 1, seed=2000 + k)
        for name, (ms, nbytes, launches) in ctx.last_timing_kernels().items():
            a = agg.setdefault(name, [0.0, 0.0, 0])
            a[0] += ms
            a[1] += nbytes
            a[2] += launches
    sync()
    elapsed_ev = time.perf_counter() - t0
    ctx.set_option("timing", 0)
    if not agg or max(v[0] for v in agg.values()) <= 0:
        raise RuntimeError("per-launch HIP event timing returned no durations")
    kern, (kms, kbytes, klaunch) = max(agg.items(), key=lambda kv: kv[1][0])
    achieved = kbytes / (kms * 1e-3) / 1e9
    log(f"plain: {value:.1f} colorings/s ({1e3 * elapsed / args.steps:.3f} ms/step; eager with events "
        f"{1e3 * elapsed_ev / args.steps:.3f} ms/step); dominant {kern} {kms / klaunch:.4f} ms/launch, "
        f"{achieved:.0f} GB/s algorithmic")

    result = {
        "metric": "colorings/sec + %HBM roofline; batched and fwd+bwd vs plain DP, same GPU",
        "value": value, "unit": "colorings/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "plain_rmat_s20_ef8_tree7",
                   "graph": "R-MAT scale 20 edge factor 8 (a=.57 b=.19 c=.19 seed 1), dedup, symmetrized",
                   "template": "7-vertex tree 0-1,1-2,2-3,1-4,4-5,4-6",
                   "colorings_per_step": 1, "n": int(n), "n_active": st["n_active"], "nnz": int(len(rowidx)),
                   "settings": settings, "autotuned": bool(args.autotune), "tuning_seconds": tuning_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kern), "kernel": kern,
                     "launches": klaunch, "avg_launch_ms": kms / klaunch,
                     "algorithmic_bytes_per_launch": kbytes / klaunch},
    }
    # the reference's cost models (CountMat::estimate*, GB/Gflop = 2^30) over
    # the measured time per coloring, beside this library's own byte model
    est = sg.estimate_model(n, len(rowidx), 7, TREE7)
    fp = ctx.footprint(plan, 1)
    sec = elapsed / args.steps
    gib = float(1 << 30)
    result["estimates"] = {
        "reference_model": {k: est[k] for k in ("peak_mem_gb", "pgbsc_spmv_gbytes", "pgbsc_fma_gbytes",
                                                "pgbsc_spmv_gflop", "pgbsc_fma_gflop")},
        "reference_model_gbs": (est["pgbsc_spmv_gbytes"] + est["pgbsc_fma_gbytes"]) * gib / sec / 1e9,
        "reference_model_gflops": (est["pgbsc_spmv_gflop"] + est["pgbsc_fma_gflop"]) * gib / sec / 1e9,
        "algorithmic_bytes_per_coloring": fp["algorithmic_bytes_per_coloring"],
        "algorithmic_gbs": fp["algorithmic_bytes_per_coloring"] / sec / 1e9,
        "device_bytes_one_coloring_pass": fp["device_bytes"],
        "note": "reference_model_*: the reference's PGBSC SpMV + eMA byte and flop models per coloring over the "
                "measured ms_per_step (decimal GB/s and Gflop/s); algorithmic_*: each table written and read "
                "once, CSC once per gathering launch"}
    log(f"models: reference PGBSC {result['estimates']['reference_model_gbs']:.0f} GB/s, "
        f"{result['estimates']['reference_model_gflops']:.0f} Gflop/s; algorithmic "
        f"{result['estimates']['algorithmic_gbs']:.0f} GB/s over the whole pass")

    if not args.no_batched:
        # the FP64 tables (exact counts below 2^53), same plain settings
        ctx.count(plan, 1, seed=900, precision="fp64")
        sync()
        t0 = time.perf_counter()
        for k in range(args.steps):
            ctx.count(plan, 1, seed=2000 + k, precision="fp64")
        sync()
        f64 = args.steps / (time.perf_counter() - t0)
        result["fp64"] = {"value": f64, "unit": "colorings/s", "vs_fp32": f64 / value,
                          "workload": "the plain line with FP64 tables"}
        log(f"fp64 plain: {f64:.1f} colorings/s ({f64 / value:.2f}x the FP32 line)")

        B = 64
        ctx.set_option("autotune", 1)
        t0 = time.perf_counter()
        ctx.count(plan, B, seed=7)  # tunes this shape (untimed)
        sync()
        btune_s = time.perf_counter() - t0
        bsettings = {k: ctx.get_option(k) for k in TUNED_KEYS}
        bsettings["colorings_per_pass"] = ctx.get_option("last_tile")
        t0 = time.perf_counter()
        reps = 3
        for r in range(reps):
            ctx.count(plan, B, seed=3000 + r)
        sync()
        bel = time.perf_counter() - t0
        bval = reps * B / bel
        result["batched"] = {"colorings_per_call": B, "value": bval, "unit": "colorings/s",
                             "vs_plain": bval / value, "workload": "same graph and template as the plain line",
                             "settings": bsettings, "tuning_seconds": btune_s}
        log(f"batched: {bval:.1f} colorings/s ({bval / value:.2f}x plain)")


    if not args.no_cpu:
        seed = 4242
        log("cpu baseline: running the compiled reference on host cores ...")

        def gpu_raw(m):
            return ctx.count(plan, m, seed=seed).tolist()

        try:
            result["cpu_baseline"] = cpu_baseline(n_decl, src, dst, seed, args.cpu_colorings, gpu_raw)
            cb = result["cpu_baseline"]
            log(f"cpu: {cb['value']:.3f} colorings/s on {cb['cores']} threads ({cb['kind']}), "
                f"parity max rel err {cb['parity_max_rel_err']:.2e}")
        except (subprocess.SubprocessError, OSError) as e:
            result["cpu_baseline"] = {"value": None, "unit": "colorings/s", "cores": physical_cores(),
                                      "kind": "reference", "sample": f"failed: {e}"}
    variants = args.variants.split(",")
    if "grad" in variants or "s22" in variants or "tree12" in variants:
        del ctx, plan
    run_variants(result, variants, sync)
    print(json.dumps(result), flush=True)


def run_variants(result, variants, sync):
    """The extra measurements, each under its key of `result`; they never
    cost the primary line: a failure is recorded."""
    for name, key, fn in (("grad", "fwd_bwd", fwd_bwd_variant), ("s22", "s22", s22_variant),
                          ("tree12", "s22_tree12", tree12_variant)):
        if name not in variants:
            continue
        try:
            result[key] = fn(sync)
        except Exception as e:  # noqa: BLE001
            log(f"variant {name} failed: {e!r}")
            result[key] = {"error": repr(e)}


if __name__ == "__main__":
    main()
