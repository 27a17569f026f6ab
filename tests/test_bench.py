"""bench.py's contract, on the CPU: the one-GPU guard, the hashed weights,
and the committed JSON line (profiles/bench_default.json) against the keys
the line must carry and the rocprofv3 summaries it cites."""
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

LINE = os.path.join(ROOT, "profiles", "bench_default.json")


def trace_md(kernel, workload):
    """The kernel-trace summary named for the dominant kernel's family and the workload."""
    return os.path.join(ROOT, "profiles", f"{kernel.split('<')[0]}_{workload}.md")


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in json.loads(out.stdout.strip().splitlines()[-1])["error"]


def test_record_names_of_kernel_symbols():
    """rocprofv3's kernel symbols map to the library's timing names (the
    trace summaries and traffic files are keyed by them)."""
    cases = {
        "void pgbsc::(anonymous namespace)::k_sub<float, 64, 0, 1, 4>(pgbsc::SubArgs<float>)": "k_sub<float,64,0,1>",
        "void pgbsc::(anonymous namespace)::k_sub<float, 8, 0, 1, 1>(pgbsc::SubArgs<float>)": "k_sub<float,8,0,1,1>",
        "void pgbsc::(anonymous namespace)::k_hub_seq<float, 56>(pgbsc::SubArgs<float>)": "k_hub_seq<float>",
        "void pgbsc::(anonymous namespace)::k_runs<float, 8, 4, 1, 1>(pgbsc::SubArgs<float>)": "k_runs<float,t,1>",
        "void pgbsc::(anonymous namespace)::k_runs_tg<float, 8, 4>(pgbsc::SubArgs<float>)": "k_runs<float,tg,1>",
        "void pgbsc::(anonymous namespace)::k_runs_g<float, 1, 8>(pgbsc::SubArgs<float>)": "k_runs<float,g,1>",
        "void pgbsc::(anonymous namespace)::k_runs_seq<float, 1, 8, 0>(pgbsc::SubArgs<float>)": "k_runs<float,seq,0>",
        "void pgbsc::k_bwd_node_rows<float, float>(pgbsc::NodeArgs<float>, int, int)": "k_bwd_node<float>",
    }
    for sym, want in cases.items():
        assert bench.record_name(sym) == want, sym


def test_hashed_weights_symmetric_and_in_range():
    import subgraph_amd as sg
    n = 1 << 10
    s, d = sg.rmat_edges(10, 8, seed=3)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    w = bench.hashed_weights(n, colptr, rowidx)
    assert w.dtype == np.float32 and len(w) == len(rowidx)
    assert w.min() >= 0.5 and w.max() < 1.5
    rows = np.repeat(np.arange(n), np.diff(colptr))
    key = dict(((int(r), int(c)), float(x)) for r, c, x in zip(rows, rowidx, w))
    assert all(key[(c, r)] == x for (r, c), x in key.items())  # both entries of an edge agree
    assert len(np.unique(w)) > len(w) // 4  # not degenerate


@pytest.fixture(scope="module")
def line():
    with open(LINE) as f:
        return json.load(f)


def test_committed_line_keeps_the_contract(line):
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in line, k
    assert line["n_gpus"] == 1 and line["higher_is_better"] is True and line["vs_baseline"] is None
    assert line["unit"] == "colorings/s" and line["dtype"] == "fp32" and line["config"]["workload"]
    per_step = line["config"].get("colorings_per_step", 1)
    assert line["value"] == pytest.approx(per_step * 1e3 / line["ms_per_step"], rel=1e-9)
    r = line["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], rel=1e-12)
    assert r["achieved"] == pytest.approx(r["algorithmic_bytes_per_launch"] / (r["avg_launch_ms"] * 1e-3) / 1e9,
                                          rel=1e-9)
    # traffic comes from the committed calibrated counter runs of that kernel
    assert r["traffic"] == bench.measured_traffic(r["kernel"], line["config"]["workload"]) and r["traffic"] > 0
    cb = line["cpu_baseline"]
    assert cb["kind"] in ("reference", "port") and cb["cores"] >= 1 and cb["value"] > 0 and cb["sample"]
    assert cb["parity_max_rel_err"] <= 1e-5


def test_trace_summary_agrees_with_the_line(line):
    """The committed rocprofv3 kernel-trace summary of the line's workload and
    the line's live HIP-event figure name the same dominant kernel, and its
    average duration agrees within 5%."""
    kern = line["roofline"]["kernel"]
    text = open(trace_md(kern, line["config"]["workload"])).read()
    rows = [ln for ln in text.splitlines() if ln.startswith(f"| `{kern}` |")]
    assert len(rows) == 1, kern
    avg_us = float(rows[0].split("|")[3])
    assert avg_us / 1e3 == pytest.approx(line["roofline"]["avg_launch_ms"], rel=0.05)
    # the dominant kernel of the trace table (first row by total time) is the line's
    first = next(ln for ln in text.splitlines() if re.match(r"\| `k_", ln))
    assert first.startswith(f"| `{kern}` |")
    # settings the summary names are the line's settings
    s = line["config"]["settings"]
    for key in ("stage_rows", "bucket_high", "bucket_mid", "hub_degree", "hub_chunk"):
        assert f"{key}={s[key]}" in text, key
