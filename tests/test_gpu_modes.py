"""The reference's other counting modes and the drop-in CLI end to end on the
device, against fixtures of the compiled reference (tests/golden/modes.json).

Bars: the unpruned mode (isPruned = 0) reproduces the reference's FP32
operation sequence, so its counts are equal bit for bit (NaN where the
reference's are NaN); the pruned modes are within 1e-5 relative per coloring.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import subgraph_amd as sg
from conftest import golden_graph

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SC_BIN = os.path.join(ROOT, "subgraph_amd", "sc-bin")
MODES = os.path.join(ROOT, "tests", "golden", "modes.json")


@pytest.fixture(scope="module")
def modes():
    with open(MODES) as f:
        return json.load(f)


def _same_bits(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    return bool(np.all((got == want) | (np.isnan(got) & np.isnan(want))))


def test_unpruned_mode_matches_reference(modes):
    """PG_FLAG_UNPRUNED on a fresh context (accumulator zero, as after the
    reference's initialization) gives the reference's unpruned raw counts for
    three colorings in a row, bit for bit; counted one call per coloring the
    accumulator carries over the same way."""
    checked = 0
    for c in modes["counts"]:
        if c["isPruned"] != 0 or c["threads"] != 1:
            continue
        t = modes["templates"][c["template"]]
        n, colptr, rowidx = golden_graph(modes["graphs"][c["graph"]])
        print(f"[unpruned] {c['graph']} {c['template']} n={n}", flush=True)
        want = [float.fromhex(r) for r in c["raw"]]
        ctx = sg.Context(0)
        ctx.upload_graph(colptr, rowidx)
        plan = ctx.plan(t["nv"], t["edges"])
        got = ctx.count(plan, 3, seed=101, unpruned=True)
        assert _same_bits(got, want), (c["graph"], c["template"], got.tolist(), want)
        ctx.upload_graph(colptr, rowidx)                     # zeroes the accumulator
        cols = sg.colors_fill(101, 0, 3, n, t["nv"])
        one = [ctx.count(plan, 1, colors=cols[i:i + 1], unpruned=True)[0] for i in range(3)]
        assert _same_bits(one, want), (c["graph"], c["template"], one, want)
        checked += 1
    assert checked >= 8


def test_spmv_flag_counts_match_reference(modes):
    """useSPMM = 0 runs the SpMM kernels; the reference's own useSPMM = 0
    counts (equal to its SpMM counts) are met within 1e-5."""
    ctx = sg.Context(0)
    for c in modes["counts"]:
        if c["isPruned"] != 1 or c["useSPMM"] != 0:
            continue
        t = modes["templates"][c["template"]]
        n, colptr, rowidx = golden_graph(modes["graphs"][c["graph"]])
        ctx.upload_graph(colptr, rowidx)
        got = ctx.count(ctx.plan(t["nv"], t["edges"]), 3, seed=101)
        want = np.array([float.fromhex(r) for r in c["raw"]])
        assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want)), (c["graph"], c["template"])


def _write_el(path, g):
    with open(path, "w") as f:
        f.write(f"{g['n']}\n{len(g['src'])}\n" + "".join(f"{a} {b}\n" for a, b in zip(g["src"], g["dst"])))


def _rmat_el(path, g):
    s, d = sg.rmat_edges(g["scale"], g["edge_factor"], seed=g["seed"])
    _write_el(path, {"n": 1 << g["scale"], "src": s.tolist(), "dst": d.tolist()})


def _lines(text):
    keep = ("Final raw count is", "Prob is", "Final count is")
    return {ln.split(" is")[0]: ln for ln in text.splitlines() if ln.startswith(keep)}


def _value(line):
    return float(line.rsplit(" ", 1)[1])


def test_cli_end_to_end_matches_reference(modes, tmp_path):
    """sc-bin on the reference's inputs: a text edge list, then the
    graph.data it wrote (load_binary = 1), a Fascia template whose ids have
    holes (Graph.cpp:40-74), every mode and the automorphism switch.  The
    printed results match the reference's compute() (CountMat.cpp:247-259):
    probability line exact, raw and final counts within 1e-5 (exact, NaN or
    inf alike, in the unpruned mode)."""
    gp = tmp_path / "g.el"
    _rmat_el(gp, modes["graphs"]["rmat_s10_ef8"])
    tdir = tmp_path / "t"
    tdir.mkdir()
    for name, t in modes["templates"].items():
        (tdir / f"{name}.fascia").write_text(f"{t['nv']}\n{len(t['edges'])}\n" +
                                             "".join(f"{a} {b}\n" for a, b in t["edges"]))
    # u5 with vertex ids 0,2,4,6,8: renumbered to the path 0-1-2-3-4
    (tdir / "u5_holes.fascia").write_text("5\n4\n0 2\n2 4\n4 6\n6 8\n")
    checked = 0
    for rec in modes["compute"]:
        tnames = [rec["template"]] + (["u5_holes"] if rec["template"] == "u5" else [])
        for tname in tnames:
            for load_binary in (0, 1):
                graph = str(tmp_path / "graph.data") if load_binary else str(gp)
                args = [SC_BIN, graph, str(tdir / f"{tname}.fascia"), str(rec["iterations"]), "2", str(load_binary),
                        str(1 - load_binary), str(rec["isPruned"]), str(rec["useSPMM"]), "-1", "1", str(rec["seed"])]
                env = dict(os.environ, PG_AUTOMORPHISMS=str(rec["autom"]))
                print(f"[cli] {tname} pruned={rec['isPruned']} spmm={rec['useSPMM']} autom={rec['autom']} "
                      f"binary={load_binary}", flush=True)
                r = subprocess.run(args, cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
                assert r.returncode == 0, r.stderr
                got = _lines(r.stdout)
                assert got["Prob"] == rec["prob_line"]
                for key, want_line in (("Final raw count", rec["final_raw_line"]), ("Final count", rec["final_line"])):
                    if rec["isPruned"] == 0:
                        # the same digits; a NaN may print with either sign ("-nan" on x86 hosts)
                        assert got[key].replace("-nan", "nan") == want_line.replace("-nan", "nan"), (got[key],
                                                                                                     want_line)
                    else:
                        g, w = _value(got[key]), _value(want_line)
                        assert abs(g - w) <= 1e-5 * abs(w), (got[key], want_line)
                if rec["isPruned"] == 1:       # the final count against compute()'s full-precision return
                    g_fin, w_fin = _value(got["Final count"]), float.fromhex(rec["final"])
                    assert abs(g_fin - w_fin) <= 1e-5 * abs(w_fin)
                checked += 1
    assert checked >= 40


def test_cli_per_subtemplate_report(tmp_path):
    """PG_VERBOSE=1 prints the reference's per-subtemplate "Sub %d, counting
    time" line (CountMat.cpp:833-840) for every non-leaf subtemplate, each
    followed by its kernels, and "Mem utilization ... count val" for every
    subtemplate of every coloring (CountMat.cpp:306-311), the root's value
    being the coloring's raw count."""
    tp = tmp_path / "t.fascia"
    tp.write_text("7\n6\n0 1\n1 2\n2 3\n1 4\n4 5\n4 6\n")
    r = subprocess.run([SC_BIN, "rmat:10:8:1", str(tp), "2", "1", "0", "0", "1", "1", "-1", "1", "5"],
                       cwd=tmp_path, env=dict(os.environ, PG_VERBOSE="1"), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    tables = sg.template_tables(7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])
    nonleaf = [s for s, size in enumerate(tables["sub_size"]) if size > 1]
    subs = [int(ln.split()[1].rstrip(",")) for ln in r.stdout.splitlines() if ln.startswith("Sub ")]
    assert sorted(subs) == sorted(nonleaf)
    assert "    k_sub<float," in r.stdout
    mem = [ln for ln in r.stdout.splitlines() if ln.startswith("Mem utilization compute step sub ")]
    assert len(mem) == 2 * len(tables["sub_size"])
    roots = [float(ln.rsplit(" ", 1)[1]) for ln in mem if ln.startswith("Mem utilization compute step sub 0:")]
    raw = _value(_lines(r.stdout)["Final raw count"])
    assert len(roots) == 2 and abs(sum(roots) / 2 - raw) <= 1e-5 * raw


def test_timing_by_subtemplate(gpu_ctx):
    """The (subtemplate, kernel) breakdown adds up to the per-kernel one."""
    s, d = sg.rmat_edges(12, 8, seed=3)
    n, s, d = sg.remove_holes(1 << 12, s, d)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])
    gpu_ctx.set_option("timing", 1)
    try:
        gpu_ctx.count(plan, 4, seed=1)
        per_kernel = gpu_ctx.last_timing_kernels()
        per_sub = gpu_ctx.last_timing_subs()
    finally:
        gpu_ctx.set_option("timing", 0)
    subs = {s for s, _ in per_sub}
    tables = sg.template_tables(7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])
    assert subs - {-1} == {s for s, size in enumerate(tables["sub_size"]) if size > 1}
    for name, (ms, nbytes, launches) in per_kernel.items():
        rows = [v for (s, k), v in per_sub.items() if k == name]
        assert sum(r[2] for r in rows) == launches
        assert sum(r[0] for r in rows) == pytest.approx(ms, rel=1e-9)
        assert sum(r[1] for r in rows) == pytest.approx(nbytes, rel=1e-9)
