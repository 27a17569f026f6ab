"""Parity at the full north-star size: R-MAT scale 22, edge factor 16 (seed 1,
dedup, symmetrized: 4,194,304 vertices, 128M CSC entries), colorings 0.. of
the color stream with seed 101, against per-coloring raw counts of the
compiled reference (tests/golden/golden_s22.json, CountMat::colorCounting):
u5 over one 64-coloring pass, the 7- and 10-vertex trees, the mixed batch of
the three, and the 12-vertex tree forward and forward+backward.

Bar: FP32 within 1e-5 relative per coloring (the reference's FP32 tables).
The reference's own FP32 rounding is up to 2e-5 from the exact count here
(its long sequential neighbour sums at hubs of degree 166k), so the device
forms the FP32 sums in the reference's order (option exact_order, the
default); the FP64 tables' counts are printed beside them.
"""
import json
import os

import numpy as np
import pytest

import subgraph_amd as sg
from golden_util import edge_checksum

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_s22.json")
TOL = 1e-5


@pytest.fixture(scope="module")
def s22():
    with open(GOLDEN) as f:
        gold = {c["template"]: c for c in json.load(f)["counts"]}
    g = next(iter(gold.values()))["graph"]
    print("[s22] generating R-MAT s22/ef16", flush=True)
    s, d = sg.rmat_edges(g["scale"], g["edge_factor"], seed=g["seed"])
    assert len(s) == g["m"] and edge_checksum(s, d) == g["checksum"], "R-MAT generator changed"
    n, s, d = sg.remove_holes(1 << g["scale"], s, d)
    assert n == g["n_after_holes"]
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    del s, d
    ctx = sg.Context(0)
    print(f"[s22] uploading n={n} nnz={len(rowidx)}", flush=True)
    ctx.upload_graph(colptr, rowidx)
    return ctx, gold, n, len(rowidx)


def _want(rec):
    return np.array([float.fromhex(r) for r in rec["raw"]])


def _check(tag, got, want, tol=TOL):
    rel = np.abs(got - want) / np.abs(want)
    print(f"[s22] {tag}: {len(want)} colorings, worst relative error {rel.max():.3e}", flush=True)
    assert np.all(rel <= tol), (tag, rel.tolist())
    return float(rel.max())


@pytest.mark.parametrize("tname", ["u5", "tree7", "tree10", "tree12"])
def test_s22_counts_match_reference(s22, tname):
    """FP32 tables with the reference's operation order (option exact_order,
    the default) against the reference's counts; the FP64 tables' counts
    show how far the reference's own FP32 rounding is from the exact count
    (up to 2e-5 for the 12-vertex tree), which is why the order matters."""
    ctx, gold, _, _ = s22
    rec = gold[tname]
    plan = ctx.plan(rec["nv"], rec["template_edges"])
    want = _want(rec)
    got64 = ctx.count(plan, len(want), seed=rec["colors"]["seed"], precision="fp64")
    ref_err = np.abs(want - got64) / got64
    print(f"[s22] {tname}: the reference's FP32 counts vs exact (FP64 tables): worst {ref_err.max():.3e}", flush=True)
    got = ctx.count(plan, len(want), seed=rec["colors"]["seed"], precision="fp32")
    tile = ctx.get_option("last_tile")
    _check(f"{tname} fp32 ({tile} per pass)", got, want)
    if tname == "u5":
        assert tile == 64                      # one 64-coloring pass


def test_s22_autotuned_batch_matches_reference(s22):
    """The batching configuration with the autotuner choosing kernel
    variants, degree buckets and the column tile."""
    ctx, gold, _, _ = s22
    rec = gold["u5"]
    plan = ctx.plan(rec["nv"], rec["template_edges"])
    ctx.set_option("autotune", 1)
    try:
        got = ctx.count(plan, 64, seed=101, precision="fp32")
    finally:
        ctx.set_option("autotune", 0)
    _check(f"u5 autotuned ({ctx.get_option('last_tile')} per pass)", got, _want(rec))


def test_s22_mixed_batch_matches_reference(s22):
    """The uneven mixed batch: 64 u5 + 8 tree7 + 4 tree10 colorings in one
    call, each template on the same color stream as its golden counts."""
    ctx, gold, _, _ = s22
    names = ["u5", "tree7", "tree10"]
    plans = [ctx.plan(gold[t]["nv"], gold[t]["template_edges"]) for t in names]
    ncol = [len(gold[t]["raw"]) for t in names]
    outs = ctx.count_multi(plans, ncol, [101] * 3, precision="fp32")
    for t, got in zip(names, outs):
        _check(f"mixed {t}", np.asarray(got), _want(gold[t]))


def test_s22_tree12_forward_backward(s22):
    """The 12-vertex tree through the adjoint at unit weights: the forward
    half of the gradient call matches the reference, and the per-entry
    gradient (colorful maps through each entry) sums to (k-1) x the raw count,
    each map using one entry per template edge."""
    ctx, gold, _, nnz = s22
    rec = gold["tree12"]
    plan = ctx.plan(rec["nv"], rec["template_edges"])
    raw, grad = ctx.count_grad(plan, len(rec["raw"]), seed=101, precision="fp32")
    _check("tree12 fwd+bwd (forward half)", raw, _want(rec))
    assert grad.shape == (nnz,) and np.all(grad >= 0)
    total = float(np.sum(grad, dtype=np.float64))
    want = 11.0 * float(np.sum(raw))
    print(f"[s22] tree12 gradient sum {total:.6e} vs 11 x raw {want:.6e}", flush=True)
    assert abs(total - want) <= 1e-4 * want
