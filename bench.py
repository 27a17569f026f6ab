"""Benchmark: colorings/s of the PGBSC color-coding DP on one MI355X.

Primary line (the north-star workload): R-MAT scale 22, edge factor 16
(a=.57 b=.19 c=.19, seed 1, self loops and duplicates removed, symmetrized
into CSC with the reference text loader's hole rule: 4,194,304 vertices,
128M entries), the 12-vertex tree 0-1,1-2,2-3,3-4,4-5,2-6,6-7,7-8,6-9,9-10,
10-11, unit weights, FP32 tables with the reference's numerics.  One step is
one batched forward+backward call of 16 colorings drawn on the device: the
counts and, through the hand-written adjoint, the gradient of their sum
with respect to every entry weight (left on the device as float32, as the
torch.autograd path takes it).  Inputs are resident in HBM before the timed
region.  value = colorings/s over the K timed steps.

Printed JSON adds:
  roofline      the step's dominant kernel: its algorithmic bytes per launch
                over its average HIP-event launch time (eager run of one step)
  cpu_baseline  the compiled reference (oracle/_ref/pg_ref) on this host's
                physical cores, one coloring of the same graph and color
                stream, with parity against the device count
  comparison    the same graph and template: batched forward, plain (one
                coloring per call) forward and forward+backward, the whole
                calls' %HBM, U(0.5,1.5) weights, and parity against the
                reference's golden counts
  plain_s20     the reference's plain method on R-MAT s20/ef8 with the
                7-vertex tree, one coloring per step, with its own roofline
                and reference CPU baseline; fp64 and 64-per-pass beside it
  s22_batching  u5, the 7- and 10-vertex trees at 64 colorings per call
                (autotuned) beside their plain passes, and the mixed batch
  fwd_bwd       s18/ef16 with U(0.5,1.5) weights: 5-star and 5-path

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
                    [--variants plain,s22,grad] [--autotune 0|1]
                    [--settings KEY=VAL,...]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TREE7 = [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]
GOLDEN_S22 = os.path.join(ROOT, "tests", "golden", "golden_s22.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """(physical cores among the CPUs this process may run on, logical
    CPUs in that set, CPU model name)."""
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    model, phys = "unknown", set()
    try:
        cpu, pid, core = None, None, None
        with open("/proc/cpuinfo") as f:
            for line in f.read().splitlines() + [""]:
                if line.startswith("processor"):
                    cpu = int(line.split(":")[1])
                elif line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                elif not line.strip() and cpu is not None:
                    if cpu in allowed:
                        phys.add((pid, core))
                    cpu, pid, core = None, None, None
    except OSError:
        pass
    return max(1, len(phys) or len(allowed)), len(allowed), model


def physical_cores() -> int:
    """Threads of the CPU baseline: the physical cores this process may use,
    at most 16 (one GPU's share of the host)."""
    return max(1, min(16, host_cpu()[0]))


def write_el(path, n_decl, src, dst):
    with open(path, "w") as f:
        f.write(f"{n_decl}\n{len(src)}\n")
        step = 1 << 20
        for i in range(0, len(src), step):
            f.write("".join(f"{a} {b}\n" for a, b in zip(src[i:i + step].tolist(), dst[i:i + step].tolist())))


def write_graph_data(path, m, colptr, rowidx):
    """The reference's binary CSC (CSCGraph::serialize, CSCGraph.hpp:426-434):
    edges, vertices, degrees, column pointers, row indices, 1.0f values."""
    n = len(colptr) - 1
    with open(path, "wb") as f:
        np.array([m, n], np.int32).tofile(f)
        np.diff(colptr).astype(np.int32).tofile(f)
        np.ascontiguousarray(colptr, np.int32).tofile(f)
        np.ascontiguousarray(rowidx, np.int32).tofile(f)
        np.ones(len(rowidx), np.float32).tofile(f)


def run_reference(graph_path, tv, tedges, seed, ncol, threads, timeout):
    """pg_ref on ncol colorings of stream `seed`; its lines are echoed as
    they come (loading, then one per coloring).  Returns ([raw], [s])."""
    import oracle
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        tp = os.path.join(tmp, "t.fascia")
        with open(tp, "w") as f:
            f.write(f"{tv}\n{len(tedges)}\n" + "".join(f"{a} {b}\n" for a, b in tedges))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        proc = subprocess.Popen([oracle.REF_BIN, graph_path, tp, "-", str(ncol), str(threads), str(seed)],
                                stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
        raws, secs = [], []
        t0 = time.perf_counter()
        for line in proc.stdout:
            log("  reference: " + line.rstrip())
            if line.startswith("PGREF coloring"):
                f = line.split()
                raws.append(float(f[4]))
                secs.append(float(f[6]))
            if time.perf_counter() - t0 > timeout:
                proc.kill()
                raise subprocess.TimeoutExpired(oracle.REF_BIN, timeout)
        if proc.wait() != 0 or len(raws) != ncol:
            raise subprocess.CalledProcessError(proc.returncode, oracle.REF_BIN)
    return raws, secs


def cpu_baseline(graph_path, tv, tedges, seed, ncol, gpu_raw_fn, sample, timeout=900):
    """The compiled reference timed on host cores (per-coloring time, the
    scope of the reference's "Time for count per iter", CountMat.cpp:200-213;
    graph load excluded), with per-coloring parity against the device."""
    import oracle
    if not oracle.have_ref():
        return {"value": None, "unit": "colorings/s", "cores": physical_cores(), "kind": "reference",
                "sample": "skipped: oracle/_ref/pg_ref is not built here"}
    threads = physical_cores()
    phys, logical, model = host_cpu()
    raws, secs = run_reference(graph_path, tv, tedges, seed, ncol, threads, timeout)
    gpu = gpu_raw_fn(ncol)
    err = max(abs(g - r) / abs(r) for g, r in zip(gpu, raws))
    return {"value": ncol / sum(secs), "unit": "colorings/s", "cores": threads, "kind": "reference",
            "cpu_model": model, "physical_cores_available": phys, "logical_cpus_available": logical,
            "sample": sample, "seconds_per_coloring": sum(secs) / ncol, "parity_max_rel_err": err,
            "reference_raw": raws, "device_raw": gpu}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` in `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic_<workload>.json,
    calibrated against a copy kernel of known size), or None when no such
    record covers the kernel."""
    try:
        with open(os.path.join(ROOT, "profiles", f"traffic_{workload}.json")) as f:
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


TREE10 = [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]
TREE12 = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]
U5 = [(0, 1), (1, 2), (2, 3), (3, 4)]
TUNED_KEYS = ("reg_cols", "stage_rows", "split_lds", "aux_rows", "bucket_high", "bucket_mid", "hub_degree", "hub_chunk",
              "batch_tile")


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


def kernel_times(ctx, run, reps=1):
    """{kernel: [ms, algorithmic bytes, launches]} summed over `reps` eager
    runs with a HIP event pair around every launch on the library's stream."""
    agg = {}
    ctx.set_option("timing", 1)
    try:
        for r in range(reps):
            run()
            for name, (ms, nbytes, nl) in ctx.last_timing_kernels().items():
                a = agg.setdefault(name, [0.0, 0.0, 0])
                a[0] += ms
                a[1] += nbytes
                a[2] += nl
    finally:
        ctx.set_option("timing", 0)
    if not agg or max(v[0] for v in agg.values()) <= 0:
        raise RuntimeError("per-launch HIP event timing returned no durations")
    return agg


def roofline_of(agg, workload):
    """The dominant kernel (largest summed event time) of kernel_times()."""
    kern, (kms, kbytes, klaunch) = max(agg.items(), key=lambda kv: kv[1][0])
    achieved = kbytes / (kms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kern, workload), "kernel": kern,
            "launches": klaunch, "avg_launch_ms": kms / klaunch, "algorithmic_bytes_per_launch": kbytes / klaunch,
            "share_of_kernel_time": kms / sum(v[0] for v in agg.values())}


def dominant_kernel(ctx, run):
    """Run once with per-launch HIP events; the dominant kernel's name,
    launches, ms per launch and algorithmic GB/s."""
    agg = kernel_times(ctx, run)
    name, (ms, nbytes, nl) = max(agg.items(), key=lambda kv: kv[1][0])
    return {"kernel": name, "launches": nl, "avg_launch_ms": ms / nl,
            "achieved_gbs": nbytes / (ms * 1e-3) / 1e9, "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "call_bytes": sum(v[1] for v in agg.values())}


def call_roofline(ctx, run, seconds_per_call):
    """Whole-call figure: the algorithmic bytes of every launch of one call
    (recorded in an eager run with per-launch events) over the call's time
    when replayed from its captured graphs, and the dominant kernel."""
    dom = dominant_kernel(ctx, run)
    gbs = dom["call_bytes"] / seconds_per_call / 1e9
    return {"algorithmic_bytes_per_call": dom["call_bytes"], "call_gbs": gbs, "call_frac": gbs / HBM_PEAK_GBS,
            "dominant": {k: v for k, v in dom.items() if k != "call_bytes"}}


def golden_s22():
    """The compiled reference's per-coloring raw counts at R-MAT s22/ef16
    (color stream seed 101), by template."""
    try:
        with open(GOLDEN_S22) as f:
            return {c["template"]: c for c in json.load(f)["counts"]}
    except (OSError, ValueError):
        return {}


def parity_vs_golden(ctx, plan, rec):
    """Worst relative error of the device's FP32 counts of the golden
    colorings against the reference's."""
    want = np.array([float.fromhex(r) for r in rec["raw"]])
    got = ctx.count(plan, len(want), seed=rec["colors"]["seed"])
    return float(np.max(np.abs(got - want) / np.abs(want))), len(want)


def s22_graph():
    """R-MAT s22/ef16, seed 1, with the reference text loader's hole rule."""
    import subgraph_amd as sg
    t0 = time.perf_counter()
    n_decl = 1 << 22
    src, dst = sg.rmat_edges(22, 16, seed=1)
    n, s2, d2 = sg.remove_holes(n_decl, src, dst)
    colptr, rowidx = sg.csc_from_edges(n, s2, d2)
    log(f"s22 graph: n={n} undirected={len(src)} nnz={len(rowidx)} ({time.perf_counter() - t0:.1f}s)")
    return n, int(len(src)), colptr, rowidx


# the 12-vertex tree's fixed settings (--autotune 0, the default).  Measured
# at s22, forward+backward of 16 colorings replayed from captured graphs:
# 4 colorings per pass 4.85-4.89 s, 2 per pass 4.95-4.98 s, 3 per pass
# 5.07-5.09 s (five passes of 3 and a remainder pass of 1), with the earlier
# kernels; 4 per pass is also the most the kept tables fit.  The autotuner
# settles on 3 per pass (it times forward passes per coloring, not the call's
# remainder), so the line fixes the tile.  Hubs from degree 4096: 4.18 s
# against 4.26 s from 2048, 4.29 s from 1024 and 4.21 s from 8192; hub
# chunks of 1024 entries (512: 4.20 s, 2048: 4.19 s, 4096: 4.24 s).  Degree
# buckets 16-128 / 4-16 all within 0.2%.
TREE12_SETTINGS = {"stage_rows": 1, "bucket_high": 32, "bucket_mid": 8, "hub_degree": 4096, "hub_chunk": 1024,
                   "batch_tile": 4}
HEADLINE = "tree12_fwd_bwd_rmat_s22_ef16"


def headline(args, sync):
    """The primary line: batched forward+backward of the 12-vertex tree on
    R-MAT s22/ef16, 16 colorings per step, and what it is compared with."""
    import torch
    import subgraph_amd as sg
    n, m, colptr, rowidx = s22_graph()
    ctx = sg.Context(0)
    if args.autotune:
        ctx.set_option("autotune", 1)
    else:
        for k, v in TREE12_SETTINGS.items():
            ctx.set_option(k, v)
    for kv in filter(None, args.settings.split(",")):
        key, val = kv.split("=")
        ctx.set_option(key.strip(), int(val))
    ctx.upload_graph(colptr, rowidx)
    st = ctx.stats()
    plan = ctx.plan(12, TREE12)
    B = args.colorings
    gbuf = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")  # the autograd path's output

    def step(k):
        ctx.count_grad_device(plan, B, gbuf.data_ptr(), seed=5000 + k)

    t0 = time.perf_counter()
    if args.autotune:
        ctx.count(plan, B, seed=4999)    # the forward tile's tuning, untimed
    # the gradient call's own tile and captured graphs, untimed.  It plans
    # first: a forward call's buffers beside it would leave its kept-table
    # plan 3 colorings per pass instead of 4
    step(-1)
    sync()
    tune_s = time.perf_counter() - t0
    settings = {k: ctx.get_option(k) for k in TUNED_KEYS}
    for w in range(args.warmup):
        step(-2 - w)
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    sync()
    elapsed = time.perf_counter() - t0
    value = args.steps * B / elapsed
    tile = ctx.get_option("last_tile")
    log(f"headline: 12-tree fwd+bwd {value:.2f} colorings/s ({1e3 * elapsed / args.steps:.1f} ms per step of {B}, "
        f"{tile} per pass; settings {settings}; first calls {tune_s:.1f}s)")
    agg = kernel_times(ctx, lambda: step(0))
    roof = roofline_of(agg, HEADLINE)
    log(f"headline dominant {roof['kernel']}: {roof['avg_launch_ms']:.3f} ms/launch, {roof['achieved']:.0f} GB/s "
        f"algorithmic ({100 * roof['frac']:.2f}%), {100 * roof['share_of_kernel_time']:.1f}% of kernel time")
    result = {
        "metric": "colorings/sec + %HBM roofline; batched and fwd+bwd vs plain DP, same GPU",
        "value": value, "unit": "colorings/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": HEADLINE,
                   "graph": "R-MAT scale 22 edge factor 16 (a=.57 b=.19 c=.19 seed 1), dedup, symmetrized, "
                            "the reference loader's hole rule",
                   "template": "12-vertex tree 0-1,1-2,2-3,3-4,4-5,2-6,6-7,7-8,6-9,9-10,10-11",
                   "step": f"one forward+backward call of {B} colorings (counts and the per-entry gradient, "
                           "unit weights, gradient kept on the device)",
                   "colorings_per_step": B, "colorings_per_pass": tile,
                   "adjoint_tables_recomputed": bool(ctx.get_option("last_recompute")),
                   "n": int(n), "n_active": st["n_active"], "nnz": int(len(rowidx)),
                   "settings": settings, "autotuned": bool(args.autotune), "first_calls_seconds": tune_s},
        "roofline": roof,
    }
    # ---- what the line is compared with, same graph and template
    cmp_ = {}
    t_fb = elapsed / args.steps
    t_f = _timed(lambda r=0: ctx.count(plan, B, seed=6000 + r), 1, sync)
    cmp_["fwd_colorings_per_s"] = B / t_f
    cmp_["fwd_colorings_per_pass"] = ctx.get_option("last_tile")
    t_pf = _timed(lambda r=0: ctx.count(plan, 1, seed=7000 + r), 2, sync)
    t_pfb = _timed(lambda r=0: ctx.count_grad_device(plan, 1, gbuf.data_ptr(), seed=8000 + r), 2, sync)
    cmp_.update({"fwd_bwd_colorings_per_s": value, "plain_fwd_colorings_per_s": 1 / t_pf,
                 "plain_fwd_bwd_colorings_per_s": 1 / t_pfb, "fwd_vs_plain_fwd": t_pf * B / t_f,
                 "fwd_bwd_vs_plain_fwd_bwd": t_pfb * B / t_fb, "fwd_bwd_vs_fwd": t_f / t_fb})
    log(f"12-tree: fwd {B / t_f:.2f}/s (plain {1 / t_pf:.2f}/s, {cmp_['fwd_vs_plain_fwd']:.2f}x); fwd+bwd "
        f"{value:.2f}/s (plain {1 / t_pfb:.2f}/s, {cmp_['fwd_bwd_vs_plain_fwd_bwd']:.2f}x)")
    roofs = {"fwd": call_roofline(ctx, lambda: ctx.count(plan, B, seed=6001), t_f),
             "fwd_bwd": call_roofline(ctx, lambda: step(1), t_fb),
             "plain_fwd": call_roofline(ctx, lambda: ctx.count(plan, 1, seed=7001), t_pf),
             "plain_fwd_bwd": call_roofline(ctx, lambda: ctx.count_grad_device(plan, 1, gbuf.data_ptr(), seed=8001),
                                            t_pfb)}
    cmp_["whole_call_roofline"] = roofs
    log("12-tree %HBM of whole calls: " + ", ".join(
        f"{k} {v['call_gbs']:.0f} GB/s ({100 * v['call_frac']:.2f}%)" for k, v in roofs.items()))
    gold = golden_s22().get("tree12")
    if gold:
        err, nc = parity_vs_golden(ctx, plan, gold)
        cmp_["parity_vs_reference"] = {"max_rel_err": err, "colorings": nc,
                                       "source": "tests/golden/golden_s22.json (compiled reference, stream seed 101)"}
        log(f"12-tree parity vs the reference's s22 counts: max rel err {err:.2e} over {nc} colorings")
    ctx.set_weights(hashed_weights(n, colptr, rowidx))
    t_wf = _timed(lambda r=0: ctx.count(plan, B, seed=9000 + r), 1, sync)
    t_wfb = _timed(lambda r=0: ctx.count_grad_device(plan, B, gbuf.data_ptr(), seed=9100 + r), 1, sync)
    cmp_["weighted_uniform_0.5_1.5"] = {"fwd_colorings_per_s": B / t_wf, "fwd_bwd_colorings_per_s": B / t_wfb}
    log(f"12-tree U(0.5,1.5) weights: fwd {B / t_wf:.2f}/s, fwd+bwd {B / t_wfb:.2f}/s")
    ctx.set_weights(None)
    cmp_["note"] = ("whole_call_roofline: algorithmic bytes of one call (forward: each table written and read once, "
                    "the CSC once per gathering launch; backward: the same units plus the FP64 per-entry gradient "
                    "read and written by each launch that adds to it) over the call's graph-replayed time")
    result["comparison"] = cmp_
    if not args.no_cpu:
        seed = 4242
        log("cpu baseline: the compiled reference, one coloring of the 12-vertex tree at s22 ...")
        try:
            with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
                gp = os.path.join(tmp, "graph.data")
                write_graph_data(gp, m, colptr, rowidx)
                result["cpu_baseline"] = cpu_baseline(
                    gp, 12, TREE12, seed, 1, lambda c: ctx.count(plan, c, seed=seed).tolist(),
                    "one coloring (stream seed 4242, coloring 0) of the headline graph and template: the whole "
                    "12-vertex-tree DP of CountMat::colorCounting (forward only; the reference has no adjoint)")
            cb = result["cpu_baseline"]
            log(f"cpu: {cb['value']:.4f} colorings/s on {cb['cores']} threads ({cb['cpu_model']}), parity max rel "
                f"err {cb['parity_max_rel_err']:.2e}")
        except (subprocess.SubprocessError, OSError) as e:
            result["cpu_baseline"] = {"value": None, "unit": "colorings/s", "cores": physical_cores(),
                                      "kind": "reference", "sample": f"failed: {e}"}
    return result, ctx, (n, colptr, rowidx)


def s22_batching(ctx, graph, sync):
    """Batching on the headline graph: 64 colorings per call of the 5-path,
    the 7-vertex tree and the 10-vertex tree (autotuned), each beside its
    plain one-coloring pass and with parity against the reference's s22
    counts, then an uneven mixed batch of the three in one call."""
    n, colptr, rowidx = graph
    gold = golden_s22()
    out = {"graph": "R-MAT scale 22 edge factor 16 (the headline graph)", "dtype": "fp32", "templates": {}}
    ctx.set_option("batch_tile", 0)  # the headline's fixed tile off: the tuner picks each template's
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
        if name in gold:
            err, nc = parity_vs_golden(ctx, plan, gold[name])
            rec["parity_vs_reference"] = {"max_rel_err": err, "colorings": nc}
        out["templates"][name] = rec
        log(f"s22 {name}: batched {B / t_b:.1f}/s, plain {1 / t_p:.1f}/s ({rec['batched_vs_plain']:.2f}x), "
            f"dominant {dom['kernel']} {dom['achieved_gbs']:.0f} GB/s, parity "
            f"{rec.get('parity_vs_reference', {}).get('max_rel_err', float('nan')):.2e}")
    ctx.set_option("autotune", 0)
    counts = [64, 40, 16]
    pl = [plans["u5"], plans["tree7"], plans["tree10"]]
    ctx.count_multi(pl, counts, [21, 22, 23])
    t_m = _timed(lambda r=0: ctx.count_multi(pl, counts, [31 + r, 41 + r, 51 + r]), 1, sync)
    out["mixed"] = {"templates": ["u5", "tree7", "tree10"], "colorings": counts,
                    "colorings_per_s": sum(counts) / t_m, "seconds": t_m}
    log(f"s22 mixed {counts}: {sum(counts) / t_m:.1f} colorings/s")
    return out


PLAIN_S20 = "plain_rmat_s20_ef8_tree7"


def plain_s20(args, sync):
    """The reference's plain method on the single-GPU drop-in configuration:
    R-MAT s20/ef8 (seed 1), the 7-vertex tree, one coloring per step drawn
    on the device, fixed settings (the autotuner's choice for this shape),
    with its dominant kernel's roofline and the compiled reference on host
    cores; FP64 tables and 64 colorings per pass beside it.  The neighbour
    sums run in the reordered (hub chunk, neighbour group) order: this
    workload's FP32 counts stay within 2e-6 of the reference's (its golden
    vectors, and the parity printed beside the CPU baseline), so the
    reference's own summation order is not needed for the 1e-5 bar here."""
    import subgraph_amd as sg
    t0 = time.perf_counter()
    src, dst = sg.rmat_edges(20, 8, seed=1)
    n_decl = 1 << 20
    n, s2, d2 = sg.remove_holes(n_decl, src, dst)
    colptr, rowidx = sg.csc_from_edges(n, s2, d2)
    log(f"s20 graph: n={n} nnz={len(rowidx)} ({time.perf_counter() - t0:.1f}s)")
    ctx = sg.Context(0)
    for kv in filter(None, args.plain_settings.split(",")):
        key, val = kv.split("=")
        ctx.set_option(key.strip(), int(val))
    ctx.upload_graph(colptr, rowidx)
    st = ctx.stats()
    plan = ctx.plan(7, TREE7)
    settings = {k: ctx.get_option(k) for k in TUNED_KEYS}
    steps = max(args.steps, 20)
    for w in range(3):
        ctx.count(plan, 1, seed=1000 + w)
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.count(plan, 1, seed=2000 + k)
    sync()
    elapsed = time.perf_counter() - t0
    value = steps / elapsed
    agg = kernel_times(ctx, lambda: ctx.count(plan, 1, seed=2000), reps=steps)
    roof = roofline_of(agg, PLAIN_S20)
    out = {"value": value, "unit": "colorings/s", "steps": steps, "ms_per_step": 1e3 * elapsed / steps,
           "workload": PLAIN_S20,
           "graph": "R-MAT scale 20 edge factor 8 (a=.57 b=.19 c=.19 seed 1), dedup, symmetrized",
           "template": "7-vertex tree 0-1,1-2,2-3,1-4,4-5,4-6", "colorings_per_step": 1,
           "n": int(n), "n_active": st["n_active"], "nnz": int(len(rowidx)), "settings": settings,
           "roofline": roof}
    fp = ctx.footprint(plan, 1)
    out["algorithmic_gbs_whole_pass"] = fp["algorithmic_bytes_per_coloring"] * value / 1e9
    log(f"plain s20: {value:.1f} colorings/s; dominant {roof['kernel']} {roof['avg_launch_ms']:.4f} ms/launch, "
        f"{roof['achieved']:.0f} GB/s algorithmic ({100 * roof['frac']:.2f}%)")
    if args.only == "plain":  # a profiler's run: the line's own launches only
        return out
    ctx.count(plan, 1, seed=900, precision="fp64")
    f64 = _timed(lambda r=0: ctx.count(plan, 1, seed=2000 + r, precision="fp64"), steps, sync)
    out["fp64"] = {"value": 1 / f64, "vs_fp32": (1 / f64) / value}
    B = 64
    ctx.set_option("autotune", 1)
    ctx.count(plan, B, seed=7)  # tunes this shape (untimed)
    sync()
    bsettings = {k: ctx.get_option(k) for k in TUNED_KEYS}
    bsettings["colorings_per_pass"] = ctx.get_option("last_tile")
    t_b = _timed(lambda r=0: ctx.count(plan, B, seed=3000 + r), 3, sync)
    out["batched"] = {"colorings_per_call": B, "value": B / t_b, "vs_plain": (B / t_b) / value,
                      "settings": bsettings}
    ctx.set_option("autotune", 0)
    log(f"plain s20 fp64 {1 / f64:.1f}/s; 64 per pass {B / t_b:.1f}/s ({(B / t_b) / value:.2f}x)")
    if not args.no_cpu:
        seed = 4242
        try:
            with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
                gp = os.path.join(tmp, "g.el")
                write_el(gp, n_decl, src, dst)
                out["cpu_baseline"] = cpu_baseline(
                    gp, 7, TREE7, seed, args.cpu_colorings, lambda c: ctx.count(plan, c, seed=seed).tolist(),
                    f"{args.cpu_colorings} colorings (stream seed {seed}) of this workload")
            cb = out["cpu_baseline"]
            log(f"plain s20 cpu: {cb['value']:.3f} colorings/s on {cb['cores']} threads, parity "
                f"{cb['parity_max_rel_err']:.2e}")
        except (subprocess.SubprocessError, OSError) as e:
            out["cpu_baseline"] = {"value": None, "sample": f"failed: {e}"}
    return out


def fwd_bwd_variant(sync):
    """Differentiation workload: R-MAT s18/ef16 with symmetric weights
    U(0.5,1.5), 5-star and 5-path, 16 colorings per call, forward + adjoint
    (per-entry gradient), beside the plain one-coloring forward on the same
    graph, autotuned.  FP32 tables."""
    import torch
    import subgraph_amd as sg
    n = 1 << 18
    src, dst = sg.rmat_edges(18, 16, seed=1)
    colptr, rowidx = sg.csc_from_edges(n, src, dst)
    ctx = sg.Context(0)
    ctx.upload_graph(colptr, rowidx, vals=hashed_weights(n, colptr, rowidx))
    ctx.set_option("autotune", 1)  # tuned per (plan, tile) in each measurement's untimed first call
    out = {"graph": "R-MAT scale 18 edge factor 16 (seed 1), weights U(0.5,1.5) symmetric",
           "nnz": int(len(rowidx)), "colorings_per_call": 16, "dtype": "fp32", "templates": {}}
    grad = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")  # the autograd path's output
    for name, te in (("star5", [(0, 1), (0, 2), (0, 3), (0, 4)]), ("u5", U5)):
        plan = ctx.plan(5, te)
        t_plain = _timed(lambda r=0: ctx.count(plan, 1, seed=50 + r), 40, sync)
        t_fwd = _timed(lambda r=0: ctx.count(plan, 16, seed=60 + r), 20, sync)
        t_fb1 = _timed(lambda r=0: ctx.count_grad_device(plan, 1, grad.data_ptr(), seed=80 + r), 40, sync)
        t_fb = _timed(lambda r=0: ctx.count_grad_device(plan, 16, grad.data_ptr(), seed=70 + r), 20, sync)
        rec = {"plain_fwd_colorings_per_s": 1.0 / t_plain, "batched_fwd_colorings_per_s": 16 / t_fwd,
               "plain_fwd_bwd_colorings_per_s": 1.0 / t_fb1, "fwd_bwd_colorings_per_s": 16 / t_fb}
        rec["fwd_bwd_vs_plain_fwd_bwd"] = rec["fwd_bwd_colorings_per_s"] / rec["plain_fwd_bwd_colorings_per_s"]
        rec["fwd_bwd_vs_plain_fwd"] = rec["fwd_bwd_colorings_per_s"] / rec["plain_fwd_colorings_per_s"]
        out["templates"][name] = rec
        log(f"fwd+bwd s18 {name}: plain fwd {rec['plain_fwd_colorings_per_s']:.0f}/s, batched fwd "
            f"{rec['batched_fwd_colorings_per_s']:.0f}/s, plain fwd+bwd {rec['plain_fwd_bwd_colorings_per_s']:.0f}/s, "
            f"fwd+bwd {rec['fwd_bwd_colorings_per_s']:.0f}/s")
    out["note"] = "the gradient stays on the device as float32 (the torch.autograd path)"
    return out


def copy_calibration(mib):
    """One elementwise kernel of known traffic: reads and writes `mib` MiB."""
    import torch
    src_t = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst_t = torch.empty_like(src_t)
    torch.add(src_t, 1.0, out=dst_t)
    torch.cuda.synchronize()
    log(f"copy calibration: {mib} MiB read + {mib} MiB written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--colorings", type=int, default=16, help="colorings per step of the primary line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-colorings", type=int, default=12, help="reference colorings of the plain_s20 baseline")
    ap.add_argument("--variants", default="plain,s22,grad",
                    help="comma list of extra measurements: plain (s20 plain method), s22 (batching), "
                         "grad (s18 differentiation), none")
    ap.add_argument("--autotune", type=int, default=0, choices=(0, 1),
                    help="primary line: 0 = fixed settings (TREE12_SETTINGS, plus --settings); 1 = the autotuner "
                         "picks settings and colorings per pass in the untimed first calls")
    ap.add_argument("--settings", default="", metavar="KEY=VAL,...", help="primary line context options")
    ap.add_argument("--plain-settings",
                    default="stage_rows=1,bucket_high=512,bucket_mid=64,hub_degree=512,hub_chunk=1024,exact_order=0",
                    metavar="KEY=VAL,...", help="context options of the plain_s20 line")
    ap.add_argument("--only", default="", choices=("", "headline", "plain", "grad"),
                    help="run just that measurement (a short run for profilers); headline and plain skip their "
                         "comparisons; grad is the s18 forward+backward workload of the line's fwd_bwd entry")
    ap.add_argument("--eager", action="store_true",
                    help="with --only headline: run the steps eagerly with a HIP event pair around every launch "
                         "(one stream), as the roofline figure is measured, for a kernel trace that matches it")
    ap.add_argument("--summarize-stats", default="", metavar="CSV",
                    help="print a rocprofv3 --stats kernel summary (kernel_stats.csv) as a markdown table by the "
                         "library's kernel names, and exit")
    ap.add_argument("--summarize-pmc", nargs=2, default=None, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 --pmc counter_collection.csv files (FETCH_SIZE, WRITE_SIZE; runs with "
                         "--copy-calibration MIB, the same MIB given here) into the traffic JSON for --workload, "
                         "and exit")
    ap.add_argument("--workload", default=HEADLINE, help="workload name of --summarize-pmc")
    ap.add_argument("--summarize-sq", default="", metavar="CSV",
                    help="print a rocprofv3 --pmc counter_collection.csv of the SQ counters SQ_WAVES, "
                         "SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_INSTS_VMEM_RD, SQ_WAIT_INST_LDS, SQ_LDS_BANK_CONFLICT, "
                         "SQ_WAVE_CYCLES, SQ_BUSY_CYCLES as a markdown table by the library's kernel names, and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one elementwise kernel that reads and writes MIB MiB (the known-size "
                         "kernel rocprofv3 FETCH_SIZE / WRITE_SIZE readings are scaled by)")
    args = ap.parse_args()
    if args.summarize_stats:
        print(stats_table(args.summarize_stats))
        return
    if args.summarize_sq:
        print(sq_table(args.summarize_sq))
        return
    if args.summarize_pmc:
        if args.copy_calibration <= 0:
            raise SystemExit("--summarize-pmc needs --copy-calibration MIB, the size the counter runs used")
        print(json.dumps(pmc_traffic(args.summarize_pmc[0], args.summarize_pmc[1], args.workload,
                                     args.copy_calibration), indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU: bench.py supports --gpus 1 only; "
                                   "colorings would shard across GPUs, each rank counting its share of the "
                                   "coloring stream"}))
        sys.exit(2)
    import torch
    sync = torch.cuda.synchronize
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    if args.only == "plain":
        args.no_cpu = True
        print(json.dumps({"plain_s20": plain_s20(args, sync)}), flush=True)
        return
    if args.only == "grad":
        print(json.dumps({"fwd_bwd": fwd_bwd_variant(sync)}), flush=True)
        return
    if args.only == "headline":
        args.no_cpu = True
        result = headline_only(args, sync)
        print(json.dumps(result), flush=True)
        return
    result, ctx, graph = headline(args, sync)
    variants = args.variants.split(",")
    for name, key, fn in (("s22", "s22_batching", lambda: s22_batching(ctx, graph, sync)),
                          ("plain", "plain_s20", lambda: plain_s20(args, sync)),
                          ("grad", "fwd_bwd", lambda: fwd_bwd_variant(sync))):
        if name not in variants:
            continue
        try:
            result[key] = fn()
        except Exception as e:  # noqa: BLE001  (an extra never costs the primary line)
            log(f"variant {name} failed: {e!r}")
            result[key] = {"error": repr(e)}
        if name == "s22":
            del ctx
            graph = None
    print(json.dumps(result), flush=True)


def headline_only(args, sync):
    """The primary line's timed steps alone (no comparisons, no CPU
    baseline): a short run whose kernel trace and counters hold only the
    headline's work plus its untimed first calls."""
    import torch
    import subgraph_amd as sg
    n, m, colptr, rowidx = s22_graph()
    ctx = sg.Context(0)
    for k, v in TREE12_SETTINGS.items():
        ctx.set_option(k, v)
    for kv in filter(None, args.settings.split(",")):
        key, val = kv.split("=")
        ctx.set_option(key.strip(), int(val))
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(12, TREE12)
    gbuf = torch.empty(len(rowidx), dtype=torch.float32, device="cuda")
    B = args.colorings
    if args.eager:
        ctx.set_option("timing", 1)
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        ctx.count_grad_device(plan, B, gbuf.data_ptr(), seed=5000 + k)
        sync()
        log(f"step {k}: {B / (time.perf_counter() - t0):.2f} colorings/s ({ctx.get_option('last_tile')} per pass)")
    ctx.set_option("timing", 0)
    return {"workload": HEADLINE, "eager": bool(args.eager), "settings": {k: ctx.get_option(k) for k in TUNED_KEYS}}


def record_name(symbol):
    """The library's timing name of a kernel symbol as rocprofv3 prints it
    (pg_capi.hip names launches by kernel family and the template arguments
    that matter: k_sub's values per lane only when 1, k_hub_seq without its
    batch depth, the colour-run kernels by mode and main-table kind, the node
    kernels' variants as one)."""
    s = symbol.replace("(anonymous namespace)::", "").split("(")[0]
    for junk in ("void ", "pgbsc::", " "):
        s = s.replace(junk, "")
    m = re.match(r"(k_\w+)<(.*)>$", s)
    if not m:
        return s
    name, a = m.group(1), m.group(2).split(",")
    if name == "k_sub" and len(a) == 5:
        return "k_sub<" + ",".join(a[:4] + ([] if a[4] == "4" else ["1"])) + ">"
    if name == "k_hub_seq":
        return f"k_hub_seq<{a[0]}>"
    if name == "k_runs" and len(a) == 5:
        return f"k_runs<{a[0]},{ {'0': 'fwd', '1': 't'}.get(a[3], a[3]) },{a[4]}>"
    if name in ("k_runs_g", "k_runs_tg"):
        return f"k_runs<{a[0]},{name[7:]},1>"
    if name == "k_runs_seq" and len(a) == 4:
        return f"k_runs<{a[0]},seq,{a[3]}>"
    if name in ("k_bwd_node", "k_bwd_node_lds", "k_bwd_node_rows"):
        return f"k_bwd_node<{a[0]}>"
    if name in ("k_edge_grad", "k_fascia_step"):
        return f"{name}<{a[0]}>"
    return s


def sq_table(path, rows=12):
    """Per library kernel name: kernel time, waves, and per wave the VALU,
    LDS and vector-memory read instructions; LDS wait cycles over wave
    cycles; LDS bank-conflict cycles per LDS instruction; and the VALU
    instructions issued per second over the kernel's time."""
    import collections
    import csv
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    dur = collections.defaultdict(float)
    seen = set()
    with open(path) as f:
        for row in csv.DictReader(f):
            name = record_name(row["Kernel_Name"])
            if not name.startswith("k_"):
                continue
            agg[name][row["Counter_Name"]] += float(row["Counter_Value"])
            if (name, row["Dispatch_Id"]) not in seen:
                seen.add((name, row["Dispatch_Id"]))
                dur[name] += (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e6
    out = ["| kernel | ms | waves | VALU/wave | LDS/wave | VMEM rd/wave | LDS wait/wave cycles | "
           "bank-conflict cycles/LDS inst | VALU inst/s |", "|---|---|---|---|---|---|---|---|---|"]
    for name in sorted(agg, key=lambda k: -dur[k])[:rows]:
        c = agg[name]
        w = max(c["SQ_WAVES"], 1.0)
        out.append(f"| `{name}` | {dur[name]:.0f} | {c['SQ_WAVES']:.0f} | {c['SQ_INSTS_VALU'] / w:.0f} | "
                   f"{c['SQ_INSTS_LDS'] / w:.0f} | {c['SQ_INSTS_VMEM_RD'] / w:.0f} | "
                   f"{c['SQ_WAIT_INST_LDS'] / max(c['SQ_WAVE_CYCLES'], 1.0):.3f} | "
                   f"{c['SQ_LDS_BANK_CONFLICT'] / max(c['SQ_INSTS_LDS'], 1.0):.2f} | "
                   f"{c['SQ_INSTS_VALU'] / max(dur[name] * 1e-3, 1e-9):.2e} |")
    return "\n".join(out)


def stats_table(path):
    """A rocprofv3 kernel_stats.csv as markdown rows by record_name(),
    largest total first: | `name` | calls | avg us | total ms | share |."""
    import csv
    agg = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = record_name(r["Name"])
            a = agg.setdefault(name, [0, 0.0])
            a[0] += int(r["Calls"])
            a[1] += float(r["TotalDurationNs"])
    tot = sum(v[1] for v in agg.values())
    lines = ["| kernel | calls | avg us | total ms | share |", "|---|---|---|---|---|"]
    for name, (calls, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{name}` | {calls} | {ns / calls / 1e3:.1f} | {ns / 1e6:.3f} | {100 * ns / tot:.1f}% |")
    return "\n".join(lines)


def pmc_traffic(fetch_csv, write_csv, workload, mib):
    """HBM bytes per launch by record_name() from one FETCH_SIZE and one
    WRITE_SIZE rocprofv3 --pmc run of the same command, each counter scaled
    by what it reads for the --copy-calibration kernel of that command, which
    reads and writes `mib` MiB."""
    import csv
    runs = []
    for path in (fetch_csv, write_csv):
        with open(path) as f:
            rows = sorted(csv.DictReader(f), key=lambda r: int(r["Dispatch_Id"]))
        runs.append(rows)
    fetch, write = runs
    cal = [i for i, r in enumerate(fetch) if "add" in r["Kernel_Name"] and "elementwise" in r["Kernel_Name"]]
    if not cal:
        raise SystemExit("no copy-calibration dispatch in the FETCH_SIZE run")
    ci = cal[0]
    mib_kib = float(mib) * 1024
    ff = mib_kib / float(fetch[ci]["Counter_Value"])
    wf = mib_kib / float(write[ci]["Counter_Value"])
    agg = {}
    for fr, wr in zip(fetch[ci + 1:], write[ci + 1:]):
        if fr["Kernel_Name"] != wr["Kernel_Name"]:
            raise SystemExit("the two counter runs dispatched different kernels")
        a = agg.setdefault(record_name(fr["Kernel_Name"]), [0, 0.0, 0.0])
        a[0] += 1
        a[1] += float(fr["Counter_Value"])
        a[2] += float(wr["Counter_Value"])
    kernels = {}
    for name, (n, fk, wk) in sorted(agg.items(), key=lambda kv: -(kv[1][1] * ff + kv[1][2] * wf)):
        kernels[name] = {"fetch_kib_per_launch": round(fk / n, 1), "write_kib_per_launch": round(wk / n, 1),
                         "hbm_bytes_per_launch": int((fk * ff + wk * wf) * 1024 / n), "launches_in_run": n}
    return {"workload": workload,
            "calibration": {"kernel": f"torch.add of a float32 tensor of {mib_kib / 1024:.0f} MiB (reads it, "
                                      "writes as much)",
                            "fetch_kib": float(fetch[ci]["Counter_Value"]), "write_kib": float(write[ci]["Counter_Value"]),
                            "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)},
            "note": "hbm_bytes_per_launch = (FETCH_SIZE * fetch_factor + WRITE_SIZE * write_factor) KiB * 1024, "
                    "averaged over every launch of that kernel name in the run",
            "kernels": kernels}


if __name__ == "__main__":
    main()
