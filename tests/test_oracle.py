"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

* against the compiled reference: golden vectors from oracle/_ref (CSC,
  decomposition and index tables, per-coloring raw counts, bit for bit);
* against what the mathematics fixes: brute-force enumeration, closed forms
  for the 3-path and the star, color-permutation invariance, homogeneity and
  multilinearity in the edge weights.
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from conftest import golden_colors, golden_graph

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")


def test_plans_match_reference(golden):
    for p in golden["plans"]:
        op = oracle.Plan(p["nv"], p["edges"])
        assert op.nsub == len(p["sub"]) == 2 * p["nv"] - 1
        for s in range(op.nsub):
            d = op.sub(s)
            assert [d["size"], d["main"], d["aux"]] == p["sub"][s], (p["name"], s)
            c2c, sm, sa = op.tables(s)
            assert c2c.tolist() == p["c2c"][s]
            if sm is not None:
                assert sm.ravel().tolist() == p["smain"][str(s)]
                assert sa.ravel().tolist() == p["saux"][str(s)]


def test_csc_and_holes_match_reference(golden, tmp_path):
    for c in golden["csc"]:
        path = tmp_path / "g.el"
        path.write_text(f"{c['n_decl']}\n{len(c['src'])}\n" +
                        "".join(f"{a} {b}\n" for a, b in zip(c["src"], c["dst"])))
        n, s, d = oracle.read_edgelist(str(path))
        colptr, rowidx = oracle.csc_build(n, s, d)
        assert colptr.tolist() == c["colptr"], c["name"]
        assert rowidx.tolist() == c["rowidx"], c["name"]


def _counts(golden, max_scale=16):
    for c in golden["counts"]:
        g = c["graph"]
        if g["kind"] == "rmat" and g["scale"] > max_scale:
            continue
        yield c


def test_f32_bit_exact_vs_reference(golden):
    """or_count_f32 reproduces colorCounting's double bit for bit."""
    checked = 0
    for c in _counts(golden):
        nv, edges = golden["templates"][c["template"]]["nv"], golden["templates"][c["template"]]["edges"]
        n, colptr, rowidx = golden_graph(c["graph"])
        cols = golden_colors(c["colors"], n, nv)
        plan = oracle.Plan(nv, edges)
        for i, want in enumerate(c["raw"]):
            got = oracle.count_f32(plan, colptr, rowidx, cols[i])
            assert got == float.fromhex(want), (c["name"], i, got, float.fromhex(want))
            checked += 1
    assert checked > 60


def test_f64_within_1e5_of_reference(golden):
    for c in _counts(golden):
        nv, edges = golden["templates"][c["template"]]["nv"], golden["templates"][c["template"]]["edges"]
        n, colptr, rowidx = golden_graph(c["graph"])
        cols = golden_colors(c["colors"], n, nv)
        plan = oracle.Plan(nv, edges)
        for i, want in enumerate(c["raw"]):
            want = float.fromhex(want)
            got = oracle.count_f64(plan, colptr, rowidx, cols[i])
            assert abs(got - want) <= 1e-5 * max(abs(want), 1.0), (c["name"], i, got, want)


def _random_simple_graph(rng, n, m):
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    keep = s != d
    pairs = {(int(min(a, b)), int(max(a, b))) for a, b in zip(s[keep], d[keep])}
    e = np.array(sorted(pairs), np.int32).reshape(-1, 2)
    return e[:, 0].copy(), e[:, 1].copy()


TREES = [(3, [(0, 1), (1, 2)]), (4, [(0, 1), (0, 2), (0, 3)]), (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
         (5, [(0, 1), (0, 2), (0, 3), (0, 4)]), (6, [(0, 1), (1, 2), (1, 3), (3, 4), (3, 5)]),
         (7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)])]


@pytest.mark.parametrize("tv,tedges", TREES)
def test_f64_equals_brute_force(tv, tedges):
    rng = np.random.default_rng(tv)
    for trial in range(4):
        n = int(rng.integers(8, 40))
        s, d = _random_simple_graph(rng, n, 3 * n)
        colptr, rowidx = oracle.csc_build(n, s, d)
        cols = rng.integers(0, tv, n).astype(np.uint8)
        plan = oracle.Plan(tv, tedges)
        bf = oracle.brute_force(tv, tedges, colptr, rowidx, cols)
        assert oracle.count_f64(plan, colptr, rowidx, cols) == bf
        vals = rng.uniform(0.5, 1.5, len(rowidx))
        bfw = oracle.brute_force(tv, tedges, colptr, rowidx, cols, vals=vals)
        dpw = oracle.count_f64(plan, colptr, rowidx, cols, vals=vals)
        assert abs(dpw - bfw) <= 1e-12 * max(abs(bfw), 1.0)


def _hist(n, colptr, rowidx, cols, k):
    h = np.zeros((n, k), np.int64)
    for u in range(n):
        for e in range(colptr[u], colptr[u + 1]):
            h[rowidx[e], cols[u]] += 1
    return h


def test_closed_form_three_path():
    """raw = 2 sum_u sum_{c1<c2, both != col(u)} h_c1(u) h_c2(u) for a 3-path."""
    rng = np.random.default_rng(3)
    n = 60
    s, d = _random_simple_graph(rng, n, 200)
    colptr, rowidx = oracle.csc_build(n, s, d)
    cols = rng.integers(0, 3, n).astype(np.uint8)
    h = _hist(n, colptr, rowidx, cols, 3)
    want = 0
    for u in range(n):
        for c1, c2 in itertools.combinations(range(3), 2):
            if cols[u] not in (c1, c2):
                want += h[u, c1] * h[u, c2]
    plan = oracle.Plan(3, [(0, 1), (1, 2)])
    assert oracle.count_f64(plan, colptr, rowidx, cols) == 2 * want


@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_closed_form_star(k):
    """raw = (k-1)! sum_v e_{k-1}({h_c(v)}_{c != col(v)}) for a star centred at 0."""
    rng = np.random.default_rng(k)
    n = 50
    s, d = _random_simple_graph(rng, n, 250)
    colptr, rowidx = oracle.csc_build(n, s, d)
    cols = rng.integers(0, k, n).astype(np.uint8)
    h = _hist(n, colptr, rowidx, cols, k)
    want = 0
    for v in range(n):
        others = [h[v, c] for c in range(k) if c != cols[v]]
        want += math.prod(others)  # e_{k-1} of k-1 values is their product
    plan = oracle.Plan(k, [(0, j) for j in range(1, k)])
    assert oracle.count_f64(plan, colptr, rowidx, cols) == math.factorial(k - 1) * want


def test_invariants():
    rng = np.random.default_rng(11)
    n = 80
    s, d = _random_simple_graph(rng, n, 320)
    tv, te = TREES[-1]
    plan = oracle.Plan(tv, te)
    colptr, rowidx = oracle.csc_build(n, s, d)
    cols = rng.integers(0, tv, n).astype(np.uint8)
    base = oracle.count_f64(plan, colptr, rowidx, cols)
    # permuting color names leaves the count unchanged
    perm = rng.permutation(tv).astype(np.uint8)
    assert oracle.count_f64(plan, colptr, rowidx, perm[cols]) == base
    # isolated vertices contribute nothing
    colptr2 = np.concatenate([colptr, np.full(10, colptr[-1], np.int32)])
    cols2 = np.concatenate([cols, rng.integers(0, tv, 10).astype(np.uint8)])
    assert oracle.count_f64(plan, colptr2, rowidx, cols2) == base
    # homogeneous of degree k-1 and affine in each single weight
    w = rng.uniform(0.5, 1.5, len(rowidx))
    f = oracle.count_f64(plan, colptr, rowidx, cols, vals=w)
    f2 = oracle.count_f64(plan, colptr, rowidx, cols, vals=2.0 * w)
    assert abs(f2 - 2.0 ** (tv - 1) * f) <= 1e-11 * abs(f2)
    e = int(rng.integers(0, len(w)))
    vals = [oracle.count_f64(plan, colptr, rowidx, cols, vals=np.where(np.arange(len(w)) == e, x, w))
            for x in (0.0, 1.0, 2.0)]
    assert abs(vals[2] - 2 * vals[1] + vals[0]) <= 1e-9 * max(abs(vals[1]), 1.0)


def test_brute_force_per_edge_counts_sum():
    """Each map uses k-1 entries: per-entry counts add up to (k-1) * total."""
    rng = np.random.default_rng(5)
    n = 30
    s, d = _random_simple_graph(rng, n, 90)
    colptr, rowidx = oracle.csc_build(n, s, d)
    tv, te = TREES[2]
    cols = rng.integers(0, tv, n).astype(np.uint8)
    tot, pe = oracle.brute_force(tv, te, colptr, rowidx, cols, per_edge=True)
    assert pe.sum() == (tv - 1) * tot


def test_reference_scale_bias(golden):
    """The reference scales by float(1e-12) and unscales by the double 1e12:
    raw counts carry a -4e-9 relative bias (8 colorful maps print as 7.99999997)."""
    c = next(c for c in golden["counts"] if c["name"] == "tiny_u3-1")
    want = float.fromhex(c["raw"][0])
    assert abs(want - 8.0 * float(np.float32(1e-12)) * 1e12) <= 1e-14
    assert want < 8.0


def test_planted_mistake_is_caught(golden):
    """An implementation that skipped the float(1e-12) scale (the unscaled FP64
    DP stands in for it here) lands within 1e-5 of the reference but never on
    its bits: the bit-exact golden check is sensitive to that mistake."""
    c = next(c for c in golden["counts"] if c["name"] == "rmat_s10_ef8_tree7")
    nv, edges = golden["templates"]["tree7"]["nv"], golden["templates"]["tree7"]["edges"]
    n, colptr, rowidx = golden_graph(c["graph"])
    cols = golden_colors(c["colors"], n, nv)
    plan = oracle.Plan(nv, edges)
    for i, want in enumerate(c["raw"]):
        want = float.fromhex(want)
        wrong = oracle.count_f64(plan, colptr, rowidx, cols[i])
        assert wrong != want and abs(wrong - want) <= 1e-5 * want
