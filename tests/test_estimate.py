"""The reference's cost models (CountMat.cpp:1067-1496) as restated by
pg_estimate_model, against the compiled reference's own figures
(tests/golden/estimate.json, written by make_estimate_golden.py) and a
hand-worked case; then, on the GPU, the CLI's VERBOSE-style report."""
import json
import os
import subprocess

import numpy as np
import pytest

import subgraph_amd as sg
from conftest import golden_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SC_BIN = os.path.join(ROOT, "subgraph_amd", "sc-bin")
FIXTURE = os.path.join(ROOT, "tests", "golden", "estimate.json")

# figures the fixture carries at full precision (%.17g), and the name of
# the matching pg_estimate field or sum of fields
FULL = {
    "pgbsc_spmv_gflop": ("pgbsc_spmv_gflop",), "pgbsc_fma_gflop": ("pgbsc_fma_gflop",),
    "pgbsc_spmv_gbytes": ("pgbsc_spmv_gbytes",), "pgbsc_fma_gbytes": ("pgbsc_fma_gbytes",),
    "pgbsc_gflop": ("pgbsc_spmv_gflop", "pgbsc_fma_gflop"),
    "pgbsc_gbytes": ("pgbsc_spmv_gbytes", "pgbsc_fma_gbytes"),
    "fascia_gflop": ("fascia_gflop",), "fascia_gbytes": ("fascia_gbytes",),
    "pruned_fascia_gflop": ("pruned_fascia_gflop",), "pruned_fascia_gbytes": ("pruned_fascia_gbytes",),
    "workload_n": ("workload_n",), "workload_nnz": ("workload_nnz",),
    "memload_n": ("memload_n",), "memload_nnz": ("memload_nnz",),
}


@pytest.fixture(scope="module")
def fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def test_models_match_reference(fixture):
    assert len(fixture["cases"]) >= 30
    for case in fixture["cases"]:
        t = fixture["templates"][case["template"]]
        e = sg.estimate_model(case["n"], case["nnz"], t["nv"], t["edges"])
        fig = case["figures"]
        for name, fields in FULL.items():
            ours = sum(e[f] for f in fields)
            ref = fig[name]
            assert ours == pytest.approx(ref, rel=1e-13, abs=1e-300), (case["graph"], case["template"], name)
        # the reference prints the peak with %f only
        assert f"{e['peak_mem_gb']:f}" == fig["peak_mem_gb_printed"], (case["graph"], case["template"])


def test_graph_sizes_and_degrees_match_reference(fixture):
    """The models take n and nnz as the reference's CSC holds them; the
    library's loaders give the same, and the same degree histogram as
    degDistri.data."""
    seen = {}
    for case in fixture["cases"]:
        g = case["graph"]
        if g not in seen:
            n, colptr, rowidx = golden_graph(fixture["graphs"][g])
            seen[g] = (n, colptr, rowidx)
        n, colptr, rowidx = seen[g]
        assert (n, len(rowidx)) == (case["n"], case["nnz"]), g
        deg = np.diff(colptr)
        assert int(deg.max()) == case["max_deg"]
        if "deg_distribution" in case:
            hist = np.bincount(deg)
            ours = [[d, int(c)] for d, c in enumerate(hist) if d >= 1 and c > 0]
            assert ours == case["deg_distribution"], g


def test_hand_worked_3path():
    """u3-1 (0-1, 1-2) on n = 10, nnz = 40: subtemplates (3;1,2) and (2;1,1),
    k = 3.  Tables: len(3) = 1, len(2) = 3, len(1) = 3; splits P = C(3,2) = 3
    and C(2,1) = 2."""
    n, z = 10, 40
    e = sg.estimate_model(n, z, 3, [(0, 1), (1, 2)])
    G = float(1 << 30)
    # only the root's aux (size 2) is a table: SpMV over 3 columns
    assert e["pgbsc_spmv_gflop"] * G == pytest.approx(z * 3)
    # eMA: root 1*3 products, (2;1,1) 3*2 products, 2n flops each
    assert e["pgbsc_fma_gflop"] * G == pytest.approx(9 * 2 * n)
    assert e["pgbsc_spmv_gbytes"] * G == pytest.approx(3 * (8 * z + 0.5 * z))
    assert e["pgbsc_fma_gbytes"] * G == pytest.approx(9 * 12 * n)
    assert e["fascia_gflop"] * G == pytest.approx(9 * (z + 2 * n))
    assert e["fascia_gbytes"] * G == pytest.approx((1 + 3) * 8 * n + 9 * (4 * (n + z) + 8 * n))
    assert e["pruned_fascia_gflop"] * G == pytest.approx(3 * z + 9 * 2 * n)
    assert e["pruned_fascia_gbytes"] * G == pytest.approx(3 * 4 * z + (1 + 3) * 8 * n + 9 * 16 * n)
    assert (e["workload_n"], e["workload_nnz"], e["memload_n"], e["memload_nnz"]) == (18, 6, 168, 36)
    # peak: 2 + 16 + 3 buffer columns, then (2;1,1)'s 3 columns; the root is
    # not counted; plus the graph's n + 2 nnz ints
    assert e["peak_mem_gb"] * G == pytest.approx((21 + 3) * 4 * n + (n + 2 * z) * 4)


def test_estimate_rejects_bad_chains():
    from subgraph_amd import _lib
    import ctypes as C
    L = _lib.lib()
    est = _lib.PgEstimate()
    size = np.array([3, 1, 2, 1, 1], np.int32)
    main = np.array([1, -1, 3, -1, -1], np.int32)
    aux = np.array([2, -1, 4, -1, -1], np.int32)
    assert L.pg_estimate_model(10, 20, 3, 5, size, main, aux, C.byref(est)) == 0
    bad = main.copy()
    bad[0] = 0  # a child must come after its parent
    assert L.pg_estimate_model(10, 20, 3, 5, size, bad, aux, C.byref(est)) == -1  # PG_ERR_INVALID
    bad_size = size.copy()
    bad_size[2] = 3  # children's sizes must add up
    assert L.pg_estimate_model(10, 20, 3, 5, bad_size, main, aux, C.byref(est)) == -1  # PG_ERR_INVALID
    assert L.pg_estimate_model(-1, 20, 3, 5, size, main, aux, C.byref(est)) == -1  # PG_ERR_INVALID


@pytest.mark.gpu
def test_cli_estimate_report(fixture, tmp_path):
    """sc-bin with PG_ESTIMATE=1 prints the reference's estimate lines and
    writes its degDistri.data, then stops before counting."""
    case = next(c for c in fixture["cases"] if c["graph"] == "rmat_s10_ef8" and c["template"] == "tree7")
    n, colptr, rowidx = golden_graph(fixture["graphs"]["rmat_s10_ef8"])
    g = fixture["graphs"]["rmat_s10_ef8"]
    s, d = sg.rmat_edges(g["scale"], g["edge_factor"], seed=g["seed"])
    gp, tp = tmp_path / "g.el", tmp_path / "t.fascia"
    gp.write_text(f"{1 << g['scale']}\n{len(s)}\n" + "".join(f"{a} {b}\n" for a, b in zip(s.tolist(), d.tolist())))
    t = fixture["templates"]["tree7"]
    tp.write_text(f"{t['nv']}\n{len(t['edges'])}\n" + "".join(f"{a} {b}\n" for a, b in t["edges"]))
    out = subprocess.run([SC_BIN, str(gp), str(tp), "1", "1", "0", "0", "1", "1", "-1", "1", "5"], cwd=tmp_path,
                         env=dict(os.environ, PG_ESTIMATE="1"), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    fig = case["figures"]
    assert f"Peak memory usage estimated : {fig['peak_mem_gb_printed']} GB" in text
    assert (f"Flops PGBSC estimated : SpMV {fig['pgbsc_spmv_gflop']:f} Gflop, FMA {fig['pgbsc_fma_gflop']:f} Gflop"
            in text)
    assert f"Comm Bytes Fascia estimated:  {fig['fascia_gbytes']:f} GBytes" in text
    assert f"Workload N: {fig['workload_n']:f}, NNZ: {fig['workload_nnz']:f}" in text
    assert f"maxDeg: {case['max_deg']}" in text
    assert "Final raw count" not in text
    dist = [[int(x) for x in ln.split()] for ln in (tmp_path / "degDistri.data").read_text().splitlines()]
    ref = next(c for c in fixture["cases"] if c["graph"] == "rmat_s10_ef8" and "deg_distribution" in c)
    assert dist == ref["deg_distribution"]


@pytest.mark.gpu
def test_footprint_matches_pass_plan():
    """pg_plan_footprint: device bytes grow with the colorings per pass, and
    the algorithmic bytes per coloring follow the byte model (each table
    written and read once, the CSC once per gathering launch) exactly."""
    n = 1 << 10
    s, d = sg.rmat_edges(10, 8, seed=1)
    n, s, d = sg.remove_holes(n, s, d)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    ctx = sg.Context(0)
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])
    f1 = ctx.footprint(plan, 1)
    f16 = ctx.footprint(plan, 16)
    f16d = ctx.footprint(plan, 16, "fp64")
    assert 0 < f1["device_bytes"] < f16["device_bytes"] < f16d["device_bytes"]
    na = ctx.stats()["n_active"]
    nnz = len(rowidx)
    csc = 4.0 * (na + 1) + 4.0 * nnz
    # 7-tree, k = 7: (2;1,1) x2 and (3;2,1) read the histogram; (4;1,3),
    # (6;4,2) and the root gather tables; widths C(7, size)
    from math import comb
    tables = 4.0 * na * ((comb(7, 2) + comb(7, 2)) + (comb(7, 2) + comb(7, 3)) + (comb(7, 3) + comb(7, 4))
                         + (comb(7, 2) + comb(7, 4) + comb(7, 6)) + comb(7, 6))
    for B, f in ((1, f1), (16, f16)):
        want = (tables * B + 3 * csc + (csc + na * B)) / B
        assert f["algorithmic_bytes_per_coloring"] == pytest.approx(want, rel=1e-12)


@pytest.mark.gpu
def test_adjoint_byte_model():
    """The algorithmic bytes the library records per launch (option
    "timing"), summed over a call, against the byte model worked out by hand
    for the 7-vertex tree: the forward sweep; with kept tables the forward
    plus the reverse sweep; with tables rebuilt, also every rebuilt
    subtree's forward steps.  Units: full-width FP32 columns of n_act x B
    values, the CSC once per gathering launch, the FP64 per-entry gradient
    read and written by each launch that adds to it."""
    from math import comb
    s, d = sg.rmat_edges(10, 8, seed=1)
    n, s, d = sg.remove_holes(1 << 10, s, d)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    ctx = sg.Context(0)
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])
    k, B = 7, 3
    na = ctx.stats()["n_active"]
    Z = float(len(rowidx))
    csc = 4.0 * (na + 1) + 4.0 * Z
    col = 4.0 * na * B
    Ck = [comb(k, j) for j in range(k + 1)]
    # subtemplates (size; main, aux): R (7;1,6) root, A (6;4,2), B4 (4;1,3),
    # C3 (3;2,1), D2 (2;1,1) = {2,3}, E2 (2;1,1) = {4,6}
    fwd = {"R": csc + col * Ck[6], "A": csc + col * (Ck[2] + Ck[4] + Ck[6]), "B4": csc + col * (Ck[3] + Ck[4]),
           "C3": col * (Ck[2] + Ck[3]), "D2": col * Ck[2], "E2": col * Ck[2]}
    hist = csc + na * B
    grad_rw = 16.0 * Z

    def gathered(Ca, Cs, Cm, with_main_table):
        """aux child a table: dY, the gradient gather (T_aux, dY in; Y out),
        G_main when the main child is a table, the transposed gather (dY in,
        G_aux out).  Below the root (Cs > 0) a main leaf's dY is its G, read
        in place: no node launch."""
        node = col * (Cs + Cm + Ca) if (with_main_table or Cs == 0) else 0.0
        b = node + (csc + 3 * col * Ca + grad_rw) + (csc + 2 * col * Ca)
        return b + (col * (Cs + Ca + Cm) if with_main_table else 0.0)

    def leaf_aux(Cs, Cm, scaled=False):
        """aux child a leaf: one node launch (G, main row, H in; dY, G_main
        out; none for a main leaf, whose dY is its G, unless the step is the
        FP32 scaled one, whose dY is G * 1e-12), then the per-entry launch
        (entry rows and columns, neighbour colors, vertex colors, dY,
        gradient)"""
        node = col * (Cs + 2 * Cm + (k if Cm else 0) + k) if (Cm or scaled) else 0.0
        return node + 8 * Z + Z * B + na * B + col * k + grad_rw

    # the scaled step is the first non-leaf the DP processes: one of the two
    # (2;1,1) steps, whose bytes are the same, so D2 stands for it
    assert plan.scaled_sub >= 0

    bwd = {"R": gathered(Ck[6], 0, 0, False), "A": gathered(Ck[2], Ck[6], Ck[4], True),
           "B4": gathered(Ck[3], Ck[4], 0, False), "C3": leaf_aux(Ck[3], Ck[2]),
           "D2": leaf_aux(Ck[2], 0, scaled=True), "E2": leaf_aux(Ck[2], 0)}
    subtree = {"A": ["A", "B4", "C3", "D2", "E2"], "B4": ["B4", "C3", "E2"], "C3": ["C3", "E2"],
               "D2": ["D2"], "E2": ["E2"]}
    rebuilt = {"R": ["A"], "A": ["B4", "D2"], "B4": ["C3"], "C3": ["E2"]}  # tables each backward step rebuilds
    want_fwd = sum(fwd.values()) + hist
    want_kept = want_fwd + sum(bwd.values())
    want_rec = want_kept + sum(fwd[t] for kids in rebuilt.values() for c in kids for t in subtree[c])

    def recorded():
        return sum(v[1] for v in ctx.last_timing_kernels().values())

    ctx.set_option("timing", 1)
    ctx.count(plan, B, seed=5)
    assert recorded() == pytest.approx(want_fwd, rel=1e-12)
    for mode, want in ((0, want_kept), (1, want_rec)):
        ctx.set_option("adjoint_recompute", mode)
        ctx.count_grad(plan, B, seed=5, precision="fp32")
        assert (ctx.get_option("last_tile"), ctx.get_option("last_recompute")) == (B, mode)
        assert recorded() == pytest.approx(want, rel=1e-12), mode
    ctx.set_option("adjoint_recompute", -1)
    ctx.set_option("timing", 0)


@pytest.mark.gpu
def test_cli_verbose_report(tmp_path):
    """PG_VERBOSE=1: the estimates, then the count with its per-kernel device
    times and the model's bandwidth over the measured device time."""
    out = subprocess.run([SC_BIN, "rmat:10:8:1", "-", "2", "1", "0", "0", "1", "1", "-1", "1", "5"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode != 0  # "-" is not a template file: the loader's error path
    tp = tmp_path / "t.fascia"
    tp.write_text("5\n4\n0 1\n1 2\n2 3\n3 4\n")
    out = subprocess.run([SC_BIN, "rmat:10:8:1", str(tp), "2", "1", "0", "0", "1", "1", "-1", "1", "5"],
                         cwd=tmp_path, env=dict(os.environ, PG_VERBOSE="1"), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    text = out.stdout
    for line in ("Peak memory usage estimated : ", "PGBSC Arith Intensity: ", "Pruned Fascia Arith Intensity: ",
                 "Device kernel time per iter : ", "Kernel k_sub<float,", "Total Memory bandwidth is : ",
                 "Total Throughput is : ", "Final raw count is "):
        assert line in text, line
    plain = subprocess.run([SC_BIN, "rmat:10:8:1", str(tp), "2", "1", "0", "0", "1", "1", "-1", "1", "5"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert plain.returncode == 0, plain.stderr
    raw = [ln for ln in text.splitlines() if ln.startswith("Final raw count is ")]
    assert raw == [ln for ln in plain.stdout.splitlines() if ln.startswith("Final raw count is ")]


@pytest.mark.gpu
def test_cli_autotune_switch(tmp_path):
    """sc-bin with PG_AUTOTUNE=1 tunes before counting; FP64 counts (exact
    below 2^53) equal the untuned run's for the same seed."""
    tp = tmp_path / "t.fascia"
    tp.write_text("7\n6\n0 1\n1 2\n2 3\n1 4\n4 5\n4 6\n")
    args = [SC_BIN, "rmat:12:8:1", str(tp), "4", "1", "0", "0", "1", "1", "-1", "1", "9", "fp64"]
    plain = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    tuned = subprocess.run(args, cwd=tmp_path, env=dict(os.environ, PG_AUTOTUNE="1"), capture_output=True,
                           text=True, timeout=120)
    assert plain.returncode == 0 and tuned.returncode == 0, (plain.stderr, tuned.stderr)
    raw = [ln for ln in plain.stdout.splitlines() if ln.startswith("Final raw count is ")]
    assert raw and raw == [ln for ln in tuned.stdout.splitlines() if ln.startswith("Final raw count is ")]
