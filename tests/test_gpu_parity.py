"""Parity of the HIP path (libpgbsc.so through its C ABI) with the oracle and
with golden vectors of the compiled reference.

Bar: FP32 tables within 1e-5 relative of the reference's raw count per
coloring; FP64 tables exactly equal to the oracle's FP64 DP (integer counts
below 2^53) and to brute force.
"""
import numpy as np
import pytest

import oracle
import subgraph_amd as sg
from conftest import golden_colors, golden_graph

pytestmark = pytest.mark.gpu

TOL32 = 1e-5

TREE7 = [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300) if b != 0 else abs(a)


def _same64(got, want) -> bool:
    """FP64 counts summed in different orders: equal below 2^53 (integers
    there are exact in any order), within 1e-12 relative above."""
    got, want = np.atleast_1d(np.asarray(got, np.float64)), np.atleast_1d(np.asarray(want, np.float64))
    small = np.abs(want) < 2.0 ** 53
    return bool(np.array_equal(got[small], want[small]) and
                np.allclose(got[~small], want[~small], rtol=1e-12, atol=0.0))


def _simple_graph(rng, n, m):
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    keep = s != d
    pairs = {(int(min(a, b)), int(max(a, b))) for a, b in zip(s[keep], d[keep])}
    e = np.array(sorted(pairs), np.int32).reshape(-1, 2)
    return e[:, 0].copy(), e[:, 1].copy()


def _random_tree(rng, k):
    e = [(int(rng.integers(0, j)), j) for j in range(1, k)]
    perm = rng.permutation(k)
    e = [(int(perm[a]), int(perm[b])) for a, b in e]
    rng.shuffle(e)
    return e


def test_golden_vectors(golden, gpu_ctx):
    """Every reference golden vector, both precisions (incl. s16/ef16 u3-1 and
    s20/ef8 7-tree at full size)."""
    worst = 0.0
    for c in golden["counts"]:
        t = golden["templates"][c["template"]]
        n, colptr, rowidx = golden_graph(c["graph"])
        print(f"[golden] {c['name']} n={n} nnz={len(rowidx)}", flush=True)
        cols = golden_colors(c["colors"], n, t["nv"])
        gpu_ctx.upload_graph(colptr, rowidx)
        plan = gpu_ctx.plan(t["nv"], t["edges"])
        got32 = gpu_ctx.count(plan, len(cols), colors=cols, precision="fp32")
        want = np.array([float.fromhex(r) for r in c["raw"]])
        for i in range(len(want)):
            r = _rel(got32[i], want[i])
            worst = max(worst, r)
            assert r <= TOL32, (c["name"], i, got32[i], want[i], r)
        if c["graph"]["kind"] == "edges" or c["graph"]["scale"] <= 12:
            op = oracle.Plan(t["nv"], t["edges"])
            got64 = gpu_ctx.count(plan, len(cols), colors=cols, precision="fp64")
            for i in range(len(want)):
                o64 = oracle.count_f64(op, colptr, rowidx, cols[i])
                if abs(o64) < 2.0 ** 53:
                    assert got64[i] == o64, (c["name"], i, got64[i], o64)
                else:
                    assert _rel(got64[i], o64) <= 1e-12
    print(f"[golden] worst FP32 relative error vs reference: {worst:.3e}", flush=True)


@pytest.mark.parametrize("hub", [None, 16])
def test_random_graphs_and_trees(gpu_ctx, hub):
    """Random simple graphs, random trees k=2..12; hub=16 forces the split
    (partial-row) path for every vertex of degree > 16."""
    rng = np.random.default_rng(1234 if hub is None else 99)
    saved = [gpu_ctx.get_option(k) for k in ("hub_degree", "hub_chunk", "bucket_high", "bucket_mid")]
    gpu_ctx.set_option("bucket_high", 32)  # all three team shapes on these small graphs
    gpu_ctx.set_option("bucket_mid", 8)
    if hub is not None:
        gpu_ctx.set_option("hub_degree", hub)
        gpu_ctx.set_option("hub_chunk", 16)
    try:
        for trial in range(12):
            n = int(rng.integers(30, 400))
            s, d = _simple_graph(rng, n, int(rng.integers(2 * n, 10 * n)))
            colptr, rowidx = sg.csc_from_edges(n, s, d)
            gpu_ctx.upload_graph(colptr, rowidx)
            k = int(rng.integers(2, 13))
            te = _random_tree(rng, k)
            print(f"[random] trial={trial} n={n} nnz={len(rowidx)} k={k} hub={hub}", flush=True)
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            cols = rng.integers(0, k, size=(3, n)).astype(np.uint8)
            g32 = gpu_ctx.count(plan, 3, colors=cols, precision="fp32")
            g64 = gpu_ctx.count(plan, 3, colors=cols, precision="fp64")
            for i in range(3):
                o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                o64 = oracle.count_f64(op, colptr, rowidx, cols[i])
                assert _rel(g32[i], o32) <= TOL32, (trial, i, g32[i], o32)
                assert _same64(g64[i], o64), (trial, i, g64[i], o64)
    finally:
        for key, v in zip(("hub_degree", "hub_chunk", "bucket_high", "bucket_mid"), saved):
            gpu_ctx.set_option(key, v)


VARIANTS = [  # (aux_rows, stage_rows, reg_cols): every neighbour-sum kernel the library can pick
    (0, 0, 0),   # direct gathers, one value per lane for the root's single value
    (0, 0, 1),   # direct gathers, one value per lane
    (0, 0, 4),   # direct gathers, four values per lane
    (0, 1, 4),   # gathers from rows staged in LDS
    (1, 0, 0),   # row-wise sums with LDS atomics
    (-1, 1, 4),  # per-subtemplate choice
]


@pytest.mark.parametrize("aux_rows,stage_rows,reg_cols", VARIANTS)
@pytest.mark.parametrize("B", [1, 5, 16])
def test_kernel_variants_match_oracle(gpu_ctx, aux_rows, stage_rows, reg_cols, B):
    """Every neighbour-sum variant, at several batch widths (segments of
    one or several colorings per chunk), against the oracle: FP64 exact,
    FP32 within the tolerance; the hub split path included."""
    rng = np.random.default_rng(500 + 10 * B + 3 * aux_rows + stage_rows + reg_cols)
    keys = ("aux_rows", "stage_rows", "reg_cols", "hub_degree", "hub_chunk", "bucket_high", "bucket_mid",
            "exact_order")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("exact_order", 0)   # the variants of the fast (reordered) neighbour sums
    gpu_ctx.set_option("bucket_high", 32)  # all three team shapes on these small graphs
    gpu_ctx.set_option("bucket_mid", 8)
    gpu_ctx.set_option("aux_rows", aux_rows)
    gpu_ctx.set_option("stage_rows", stage_rows)
    gpu_ctx.set_option("reg_cols", reg_cols)
    gpu_ctx.set_option("hub_degree", 24)
    gpu_ctx.set_option("hub_chunk", 24)
    try:
        for trial, (k, te) in enumerate([(7, TREE7), (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
                                         (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8),
                                               (7, 9)]),
                                         # 462-value rows: slots padded to whole chunks, wide staging
                                         (12, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8),
                                               (6, 9), (9, 10), (10, 11)])]):
            n = int(rng.integers(60, 300))
            s, d = _simple_graph(rng, n, int(rng.integers(4 * n, 10 * n)))
            colptr, rowidx = sg.csc_from_edges(n, s, d)
            gpu_ctx.upload_graph(colptr, rowidx)
            print(f"[variants] aux_rows={aux_rows} stage_rows={stage_rows} reg_cols={reg_cols} B={B} k={k} n={n} "
                  f"nnz={len(rowidx)}", flush=True)
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            cols = rng.integers(0, k, size=(B, n)).astype(np.uint8)
            g32 = gpu_ctx.count(plan, B, colors=cols, precision="fp32")
            g64 = gpu_ctx.count(plan, B, colors=cols, precision="fp64")
            for i in range(B):
                o64 = oracle.count_f64(op, colptr, rowidx, cols[i])
                assert _same64(g64[i], o64), (trial, i, g64[i], o64)
                if i < 3:
                    o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                    assert _rel(g32[i], o32) <= TOL32, (trial, i, g32[i], o32)
    finally:
        for key, v in zip(keys, saved):
            gpu_ctx.set_option(key, v)


def test_autotune_keeps_results(gpu_ctx):
    """Option "autotune": the tuned pass equals the oracle (FP64 exact, FP32
    within the tolerance), the chosen settings are among the candidates,
    and the second call of the same shape reuses them."""
    rng = np.random.default_rng(77)
    keys = ("autotune", "reg_cols", "stage_rows", "aux_rows", "bucket_high", "bucket_mid", "hub_degree", "hub_chunk")
    saved = [gpu_ctx.get_option(k) for k in keys]
    s, d = sg.rmat_edges(12, 8, seed=4)
    n = 1 << 12
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, TREE7)
    op = oracle.Plan(7, TREE7)
    cols = rng.integers(0, 7, size=(4, n)).astype(np.uint8)
    try:
        gpu_ctx.set_option("autotune", 1)
        print("[autotune] first call (tunes)", flush=True)
        g64 = gpu_ctx.count(plan, 4, colors=cols, precision="fp64")
        chosen64 = {k: gpu_ctx.get_option(k) for k in keys[1:]}
        tile64 = gpu_ctx.get_option("last_tile")
        assert tile64 in (1, 2, 3, 4)  # the column tile is tuned too: 3/4, 1/2, 1/4 of 4
        g32 = gpu_ctx.count(plan, 4, colors=cols, precision="fp32")  # its own key: tunes again
        chosen32 = {k: gpu_ctx.get_option(k) for k in keys[1:]}
        print(f"[autotune] chosen fp64 {chosen64} fp32 {chosen32}", flush=True)
        for chosen in (chosen64, chosen32):
            assert chosen["stage_rows"] in (0, 1) and chosen["aux_rows"] in (0, -1) and chosen["reg_cols"] in (0, 1)
            assert chosen["bucket_high"] in (16, 32, 64, 128, 256, 512)
            assert chosen["bucket_mid"] in (4, 8, 16, 32, 64, 128)
            assert chosen["hub_degree"] in (128, 256, 512, 1024) and chosen["hub_chunk"] in (256, 1024, 4096)
        for i in range(4):
            assert _same64(g64[i], oracle.count_f64(op, colptr, rowidx, cols[i]))
            assert _rel(g32[i], oracle.count_f32(op, colptr, rowidx, cols[i])) <= TOL32
        again = gpu_ctx.count(plan, 4, colors=cols, precision="fp64")  # reuses the fp64 choice
        assert np.array_equal(again, g64)
        assert {k: gpu_ctx.get_option(k) for k in keys[1:]} == chosen64
        assert gpu_ctx.get_option("last_tile") == tile64
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


def test_fp64_equals_brute_force(gpu_ctx):
    rng = np.random.default_rng(8)
    trees = [(3, [(0, 1), (1, 2)]), (5, [(0, 1), (0, 2), (0, 3), (0, 4)]), (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
             (6, [(0, 1), (1, 2), (1, 3), (3, 4), (3, 5)])]
    for tv, te in trees:
        n = 40
        s, d = _simple_graph(rng, n, 160)
        colptr, rowidx = sg.csc_from_edges(n, s, d)
        gpu_ctx.upload_graph(colptr, rowidx)
        plan = gpu_ctx.plan(tv, te)
        cols = rng.integers(0, tv, size=(2, n)).astype(np.uint8)
        got = gpu_ctx.count(plan, 2, colors=cols, precision="fp64")
        for i in range(2):
            assert got[i] == oracle.brute_force(tv, te, colptr, rowidx, cols[i])


def test_batch_tiles_agree(gpu_ctx):
    """One call with 37 colorings equals any tiling of them (and per-coloring
    calls): FP64 exactly; FP32 within 1e-6, since the neighbour-group shape of
    a team (and so the order of a vertex's FP32 neighbour sum) follows the
    number of values per row, i.e. the tile."""
    n = 1 << 12
    s, d = sg.rmat_edges(12, 8, seed=4)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, TREE7)
    cols = sg.colors_fill(5, 0, 37, n, 7)
    ref64 = None
    for tile in (0, 1, 5, 16):
        gpu_ctx.set_option("batch_tile", tile)
        print(f"[tiles] batch_tile={tile}", flush=True)
        g64 = gpu_ctx.count(plan, 37, colors=cols, precision="fp64")
        g32 = gpu_ctx.count(plan, 37, colors=cols, precision="fp32")
        if ref64 is None:
            ref64, ref32 = g64, g32
        assert _same64(g64, ref64)  # other tiles sum in other orders
        assert np.all(np.abs(g32 - ref32) <= 1e-6 * np.abs(ref64))
        assert np.all(np.abs(g32 - g64) <= TOL32 * np.abs(g64))
    gpu_ctx.set_option("batch_tile", 0)
    one = gpu_ctx.count(plan, 1, colors=cols[3:4], precision="fp64")
    assert _same64(one[0], ref64[3])
    again = gpu_ctx.count(plan, 37, colors=cols, precision="fp32")
    assert np.array_equal(again, ref32)  # run to run: bit-identical


def test_graph_replays_match_eager(gpu_ctx):
    """Each pass runs as a captured hipGraph, replayed for every later pass of
    the same shape: several replays, with eager passes in between, all equal
    the eager launches bit for bit."""
    n = 1 << 12
    s, d = sg.rmat_edges(12, 16, seed=6)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, TREE7)
    gpu_ctx.set_option("graphs", 0)
    try:
        eager = [gpu_ctx.count(plan, 5, seed=40 + r, precision="fp32") for r in range(3)]
        eager64 = gpu_ctx.count(plan, 5, seed=40, precision="fp64")
    finally:
        gpu_ctx.set_option("graphs", 1)
    for rep in range(3):
        print(f"[graphs] replay round {rep}", flush=True)
        for r in range(3):
            assert np.array_equal(gpu_ctx.count(plan, 5, seed=40 + r, precision="fp32"), eager[r])
        gpu_ctx.set_option("graphs", 0)
        assert np.array_equal(gpu_ctx.count(plan, 5, seed=40, precision="fp64"), eager64)
        gpu_ctx.set_option("graphs", 1)
        assert np.array_equal(gpu_ctx.count(plan, 5, seed=40, precision="fp64"), eager64)


def test_timing_breakdown(gpu_ctx):
    """Option "timing" runs passes eagerly with an event pair per launch: the
    per-kernel breakdown is non-empty, its durations positive, its total no
    more than the call's device time, and the counts unchanged."""
    n = 1 << 12
    s, d = sg.rmat_edges(12, 16, seed=6)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, TREE7)
    want = gpu_ctx.count(plan, 3, seed=11)
    gpu_ctx.set_option("timing", 1)
    try:
        got = gpu_ctx.count(plan, 3, seed=11)
        kern = gpu_ctx.last_timing_kernels()
        tot = gpu_ctx.last_timing()
    finally:
        gpu_ctx.set_option("timing", 0)
    assert np.array_equal(got, want)
    assert kern and all(ms > 0 and nl > 0 for ms, _, nl in kern.values()), kern
    assert any(name.startswith("k_sub<float") and nb > 0 for name, (_, nb, _) in kern.items())
    assert sum(ms for ms, _, _ in kern.values()) <= tot["total_ms"] * 1.01
    assert tot["kernel"] in kern and tot["kernel_ms"] == max(ms for ms, _, _ in kern.values())


def test_device_colors_equal_host_stream(gpu_ctx):
    n = 1 << 11
    s, d = sg.rmat_edges(11, 8, seed=9)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    dev = gpu_ctx.count(plan, 6, colors=None, seed=77, precision="fp64")
    host = gpu_ctx.count(plan, 6, colors=sg.colors_fill(77, 0, 6, n, 5), precision="fp64")
    assert np.array_equal(dev, host)


def test_two_vertex_template_quirk(gpu_ctx):
    """u2: the reference reports 1e12 x the count (CountMat.cpp:811-824);
    PG_FLAG_FIX_2VERTEX gives the count itself."""
    colptr, rowidx = sg.csc_from_edges(4, np.array([0, 1, 2, 2], np.int32), np.array([1, 2, 0, 3], np.int32))
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(2, [(0, 1)])
    cols = np.array([[0, 1, 0, 1]], np.uint8)
    assert gpu_ctx.count(plan, 1, colors=cols, precision="fp32")[0] == 6.0e12
    assert gpu_ctx.count(plan, 1, colors=cols, precision="fp64")[0] == 6.0e12
    assert gpu_ctx.count(plan, 1, colors=cols, precision="fp32", fix_2vertex=True)[0] == 6.0
    assert gpu_ctx.count(plan, 1, colors=cols, precision="fp64", fix_2vertex=True)[0] == 6.0


def test_errors_are_reported(gpu_ctx):
    colptr, rowidx = sg.csc_from_edges(4, np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32))
    fresh = sg.Context(0)
    plan = fresh.plan(3, [(0, 1), (1, 2)])
    with pytest.raises(sg.PGError):
        fresh.count(plan, 1)  # no graph yet
    fresh.upload_graph(colptr, rowidx)
    with pytest.raises(sg.PGError):
        fresh.count(plan, 1, colors=np.array([[0, 1, 3, 0]], np.uint8))  # color >= k
    with pytest.raises(sg.PGError):
        fresh.upload_graph(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32))  # colptr not monotone
    with pytest.raises(sg.PGError):
        fresh.plan(3, [(0, 1), (0, 1)])  # not a tree
    with pytest.raises(sg.PGError):
        fresh.set_option("hub_degree", 1)
    bad = sg.template_tables(3, [(0, 1), (1, 2)])
    bad["split_aux"] = bad["split_aux"].copy()
    bad["split_aux"][0] = 99
    with pytest.raises(sg.PGError):
        fresh.plan_from_tables(3, bad["sub_size"], bad["main_idx"], bad["aux_idx"], bad["comb_to_count"],
                               bad["split_main"], bad["split_aux"])


def test_isolated_vertices_and_empty_graph(gpu_ctx):
    # vertices 5..9 isolated: same counts as without them
    s = np.array([0, 1, 2, 3, 0], np.int32)
    d = np.array([1, 2, 3, 4, 2], np.int32)
    c5 = sg.csc_from_edges(5, s, d)
    c10 = sg.csc_from_edges(10, s, d)
    rng = np.random.default_rng(2)
    cols = rng.integers(0, 3, 10).astype(np.uint8)
    plan = gpu_ctx.plan(3, [(0, 1), (1, 2)])
    gpu_ctx.upload_graph(*c5)
    a = gpu_ctx.count(plan, 1, colors=cols[None, :5], precision="fp64")
    gpu_ctx.upload_graph(*c10)
    b = gpu_ctx.count(plan, 1, colors=cols[None, :], precision="fp64")
    assert a[0] == b[0]
    # a graph with no edges counts zero
    gpu_ctx.upload_graph(np.zeros(8, np.int32), np.zeros(0, np.int32))
    assert gpu_ctx.count(plan, 2, precision="fp32").tolist() == [0.0, 0.0]


def test_full_size_properties_s20(gpu_ctx):
    """s20/ef8 7-tree (the drop-in workload): color renaming leaves every count
    unchanged (FP64, same summation order: exact), FP32 within 1e-5 of FP64,
    hubs split or not (another summation order: the raw counts are ~1e21, far
    above 2^53, so FP64 agrees to rounding)."""
    n = 1 << 20
    s, d = sg.rmat_edges(20, 8, seed=1)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(7, TREE7)
    cols = sg.colors_fill(3, 0, 2, n, 7)
    perm = np.array([3, 6, 0, 1, 5, 2, 4], np.uint8)
    g64 = gpu_ctx.count(plan, 2, colors=cols, precision="fp64")
    p64 = gpu_ctx.count(plan, 2, colors=perm[cols], precision="fp64")
    g32 = gpu_ctx.count(plan, 2, colors=cols, precision="fp32")
    assert np.array_equal(g64, p64)
    for i in range(2):
        assert _rel(g32[i], g64[i]) <= TOL32
    saved = gpu_ctx.get_option("hub_degree")
    gpu_ctx.set_option("hub_degree", 1 << 30)
    try:
        nohub = gpu_ctx.count(plan, 2, colors=cols, precision="fp64")
    finally:
        gpu_ctx.set_option("hub_degree", saved)
    if np.abs(g64).max() < 2.0 ** 53:
        assert np.array_equal(nohub, g64)
    else:
        assert np.allclose(nohub, g64, rtol=1e-12, atol=0.0)


def test_mixed_batch_matches_single_counts(gpu_ctx):
    """pg_count_multi runs a pass of every template in one graph launch (a
    branch per template).  Its counts equal each template counted alone with
    the same seeds, bit for bit in both precisions: over several rounds of
    passes (a small batch tile), on replays of the cached graph, with the
    autotuner, and with the oracle for one coloring of each."""
    keys = ("batch_tile", "autotune", "reg_cols", "stage_rows", "aux_rows", "bucket_high", "bucket_mid",
            "hub_degree", "hub_chunk")
    saved = [gpu_ctx.get_option(k) for k in keys]
    n = 1 << 12
    s, d = sg.rmat_edges(12, 8, seed=6)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    u5 = [(0, 1), (1, 2), (2, 3), (3, 4)]
    tree10 = [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]
    try:
        gpu_ctx.upload_graph(colptr, rowidx)
        plans = [gpu_ctx.plan(5, u5), gpu_ctx.plan(7, TREE7), gpu_ctx.plan(10, tree10)]
        counts, seeds = [5, 3, 2], [11, 12, 13]
        for tile, tune in ((0, 0), (2, 0), (0, 1)):
            gpu_ctx.set_option("batch_tile", tile)
            gpu_ctx.set_option("autotune", tune)
            for prec in ("fp64", "fp32"):
                print(f"[mixed] tile={tile} autotune={tune} {prec}", flush=True)
                multi = gpu_ctx.count_multi(plans, counts, seeds, precision=prec)
                again = gpu_ctx.count_multi(plans, counts, seeds, precision=prec)  # the cached graph
                for i, p in enumerate(plans):
                    single = gpu_ctx.count(p, counts[i], seed=seeds[i], precision=prec)
                    assert np.array_equal(multi[i], single), (tile, tune, prec, i, multi[i], single)
                    assert np.array_equal(again[i], multi[i])
        # the oracle on the first coloring of each template (the color stream of its seed)
        gpu_ctx.set_option("batch_tile", 0)
        gpu_ctx.set_option("autotune", 0)
        multi = gpu_ctx.count_multi(plans, counts, seeds, precision="fp64")
        for i, (k, te) in enumerate(((5, u5), (7, TREE7), (10, tree10))):
            cols = sg.colors_fill(seeds[i], 0, 1, n, k)
            assert _same64(multi[i][0], oracle.count_f64(oracle.Plan(k, te), colptr, rowidx, cols[0]))
        # the same plan twice in one batch: two branches, distinct seeds
        two = gpu_ctx.count_multi([plans[1], plans[1]], [2, 2], [21, 22], precision="fp64")
        assert np.array_equal(two[0], gpu_ctx.count(plans[1], 2, seed=21, precision="fp64"))
        assert np.array_equal(two[1], gpu_ctx.count(plans[1], 2, seed=22, precision="fp64"))
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


def test_vertex_order_option(gpu_ctx):
    """Option "vertex_order" (read at upload) changes only the internal order:
    degree classes with breadth-first order inside give the counts of the
    default descending-degree order (FP64 equal below 2^53, FP32 within the
    tolerance), also with the hub split on and weights set."""
    n = 1 << 12
    s, d = sg.rmat_edges(12, 16, seed=9)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    cols = sg.colors_fill(4, 0, 3, n, 7)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 31 + hi * 17) % 100) / 100.0).astype(np.float32)
    keys = ("vertex_order", "hub_degree", "hub_chunk")
    saved = [gpu_ctx.get_option(k) for k in keys]
    try:
        gpu_ctx.set_option("hub_degree", 64)
        gpu_ctx.set_option("hub_chunk", 64)
        out = {}
        for order in (0, 1):
            gpu_ctx.set_option("vertex_order", order)
            gpu_ctx.upload_graph(colptr, rowidx)
            plan = gpu_ctx.plan(7, TREE7)
            print(f"[vertex-order] {order}", flush=True)
            r64 = gpu_ctx.count(plan, 3, colors=cols, precision="fp64")
            r32 = gpu_ctx.count(plan, 3, colors=cols, precision="fp32")
            gpu_ctx.set_weights(w)
            w64 = gpu_ctx.count(plan, 3, colors=cols, precision="fp64")
            gpu_ctx.set_weights(None)
            out[order] = (r64, r32, w64)
        assert _same64(out[1][0], out[0][0])
        assert np.all(np.abs(out[1][1] - out[0][0]) <= TOL32 * np.abs(out[0][0]))
        assert np.allclose(out[1][2], out[0][2], rtol=1e-12, atol=0.0)
        op = oracle.Plan(7, TREE7)
        assert _same64(out[1][0][0], oracle.count_f64(op, colptr, rowidx, cols[0]))
        with pytest.raises(sg.PGError):
            gpu_ctx.set_option("vertex_order", 2)
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)
        gpu_ctx.upload_graph(colptr, rowidx)


BUCKET_SHAPES = [
    # hub_degree, bucket_high, bucket_mid, hub_chunk
    (512, 512, 64, 1024),   # no 64-lane bucket below the hubs; 16-lane teams from degree 64
    (512, 512, 128, 1024),
    (128, 512, 256, 256),   # bucket_high above hub_degree: its range is empty
    (1024, 16, 128, 4096),  # bucket_mid above bucket_high: every non-high vertex on 8-lane teams
    (256, 256, 256, 1024),
    (128, 16, 4, 256),      # the smallest candidates
]


@pytest.mark.parametrize("hub_degree,bucket_high,bucket_mid,hub_chunk", BUCKET_SHAPES)
def test_bucket_shapes_match_oracle(gpu_ctx, hub_degree, bucket_high, bucket_mid, hub_chunk):
    """The degree buckets the autotuner may choose, including orderings that
    leave a bucket empty, on a power-law graph with hubs: FP64 exact, FP32
    within the tolerance, one and several colorings per pass."""
    keys = ("hub_degree", "bucket_high", "bucket_mid", "hub_chunk", "exact_order")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("exact_order", 0)   # bucket shapes of the fast neighbour sums
    s, d = sg.rmat_edges(12, 16, seed=7)
    n = 1 << 12
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    assert np.diff(colptr).max() > 1024
    try:
        for key, v in zip(keys[:4], (hub_degree, bucket_high, bucket_mid, hub_chunk)):
            gpu_ctx.set_option(key, v)
        gpu_ctx.upload_graph(colptr, rowidx)
        rng = np.random.default_rng(hub_degree + bucket_high + bucket_mid)
        for k, te in ((7, TREE7), (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)])):
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            for B in (1, 4):
                print(f"[buckets] hd={hub_degree} bh={bucket_high} bm={bucket_mid} hc={hub_chunk} k={k} B={B}",
                      flush=True)
                cols = rng.integers(0, k, size=(B, n)).astype(np.uint8)
                g64 = gpu_ctx.count(plan, B, colors=cols, precision="fp64")
                g32 = gpu_ctx.count(plan, B, colors=cols, precision="fp32")
                for i in range(min(B, 2)):
                    o64 = oracle.count_f64(op, colptr, rowidx, cols[i])
                    assert _same64(g64[i], o64), (k, B, i, g64[i], o64)
                    o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                    assert _rel(g32[i], o32) <= TOL32, (k, B, i, g32[i], o32)
    finally:
        for key, v in zip(keys, saved):
            gpu_ctx.set_option(key, v)


@pytest.mark.parametrize("aux_rows,stage_rows", [(0, 0), (0, 1), (1, 0)])
def test_64_colorings_per_pass_match_oracle(gpu_ctx, aux_rows, stage_rows):
    """The batching configuration's 64 colorings in ONE pass (extra SpMM
    right-hand-side columns) on an R-MAT scale-13 graph, every coloring
    against the oracle: FP32 within the tolerance, FP64 exact below 2^53;
    the pass width is asserted, not assumed."""
    s, d = sg.rmat_edges(13, 8, seed=4)
    n, s, d = sg.remove_holes(1 << 13, s, d)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    keys = ("aux_rows", "stage_rows", "batch_tile", "exact_order")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("exact_order", 0 if aux_rows or stage_rows else 1)
    gpu_ctx.upload_graph(colptr, rowidx)
    gpu_ctx.set_option("aux_rows", aux_rows)
    gpu_ctx.set_option("stage_rows", stage_rows)
    gpu_ctx.set_option("batch_tile", 64)
    try:
        for k, te in ((5, [(0, 1), (1, 2), (2, 3), (3, 4)]), (7, TREE7)):
            print(f"[tile64] aux_rows={aux_rows} stage_rows={stage_rows} k={k} n={n} nnz={len(rowidx)}", flush=True)
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            cols = sg.colors_fill(77, 0, 64, n, k)
            g32 = gpu_ctx.count(plan, 64, colors=cols, precision="fp32")
            assert gpu_ctx.get_option("last_tile") == 64
            # FP64 rows are twice as wide: the widest pass that fits LDS
            gpu_ctx.set_option("batch_tile", 0)
            g64 = gpu_ctx.count(plan, 64, colors=cols, precision="fp64")
            gpu_ctx.set_option("batch_tile", 64)
            worst = 0.0
            for i in range(64):
                o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                worst = max(worst, _rel(g32[i], o32))
                assert _rel(g32[i], o32) <= TOL32, (k, i, g32[i], o32)
                assert _same64(g64[i], oracle.count_f64(op, colptr, rowidx, cols[i])), (k, i)
            print(f"[tile64] k={k}: worst FP32 relative error {worst:.3e}", flush=True)
    finally:
        for key, v in zip(keys, saved):
            gpu_ctx.set_option(key, v)


EXACT_TEMPLATES = [(5, [(0, 1), (1, 2), (2, 3), (3, 4)]), (7, TREE7),
                   (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]),
                   (12, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)])]


@pytest.mark.parametrize("hub_degree,color_runs", [(1 << 30, -1), (16, -1), (16, 0), (16, 1)])
def test_exact_order_is_the_reference_arithmetic(gpu_ctx, hub_degree, color_runs):
    """Option exact_order (the default): every FP32 operation of the DP is the
    reference's, in its order (neighbour sums one after another in column
    order, splits in IndexSys order, the scaled step in FP64), so the counts
    equal the oracle's bit-exact restatement of the reference up to the order
    of the final FP64 sum over vertices (1e-12 relative here); hubs (split by
    columns when hub_degree is small) and one or several colorings per pass,
    with the neighbour sums gathered per lane (color_runs 0), by whole rows
    (color_runs 1: every row of up to 512 values) or as chosen by width, and
    hub rows by the batched hub kernel."""
    keys = ("hub_degree", "exact_order", "color_runs")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("exact_order", 1)
    gpu_ctx.set_option("hub_degree", hub_degree)
    gpu_ctx.set_option("color_runs", color_runs)
    rng = np.random.default_rng(91 + (hub_degree & 0xff))
    try:
        s, d = sg.rmat_edges(12, 16, seed=7)
        n = 1 << 12
        colptr, rowidx = sg.csc_from_edges(n, s, d)
        gpu_ctx.upload_graph(colptr, rowidx)
        worst = 0.0
        for k, te in EXACT_TEMPLATES:
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            for B in (1, 5):
                print(f"[exact] hub_degree={hub_degree} color_runs={color_runs} k={k} B={B}", flush=True)
                cols = rng.integers(0, k, size=(B, n)).astype(np.uint8)
                g32 = gpu_ctx.count(plan, B, colors=cols, precision="fp32")
                for i in range(B):
                    o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                    worst = max(worst, _rel(g32[i], o32))
                    assert _rel(g32[i], o32) <= 1e-12, (k, B, i, g32[i], o32)
        print(f"[exact] worst relative difference to the oracle: {worst:.3e}", flush=True)
    finally:
        for key, v in zip(keys, saved):
            gpu_ctx.set_option(key, v)


def _giant_hub_graph():
    """R-MAT s15/ef8 plus a hub joined to 20,000 vertices: one neighbour list
    of more than 16,384 entries."""
    s, d = sg.rmat_edges(15, 8, seed=5)
    rng = np.random.default_rng(23)
    nb = rng.choice(np.arange(1, 1 << 15), size=20000, replace=False)
    s = np.concatenate([s, np.zeros(len(nb), s.dtype)])
    d = np.concatenate([d, nb.astype(d.dtype)])
    n = 1 << 15
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    return n, colptr, rowidx


def test_giant_hub_rows_keep_the_reference_arithmetic(gpu_ctx):
    """A hub of degree above 16,384 (20,000 added neighbours on R-MAT s15):
    its rows are chains of that many ordered adds in the hub kernel, and the
    FP32 counts still equal the oracle's restatement of the reference, one
    and several colorings per pass."""
    keys = ("hub_degree", "exact_order")
    saved = [gpu_ctx.get_option(k) for k in keys]
    try:
        n, colptr, rowidx = _giant_hub_graph()
        assert np.diff(colptr).max() >= 16384
        gpu_ctx.set_option("exact_order", 1)
        gpu_ctx.set_option("hub_degree", 2048)
        gpu_ctx.upload_graph(colptr, rowidx)
        rng = np.random.default_rng(29)
        for k, te in EXACT_TEMPLATES[:2]:
            plan = gpu_ctx.plan(k, te)
            op = oracle.Plan(k, te)
            for B in (1, 3):
                print(f"[giant hub] k={k} B={B}", flush=True)
                cols = rng.integers(0, k, size=(B, n)).astype(np.uint8)
                g32 = gpu_ctx.count(plan, B, colors=cols, precision="fp32")
                for i in range(B):
                    o32 = oracle.count_f32(op, colptr, rowidx, cols[i])
                    assert _rel(g32[i], o32) <= 1e-12, (k, B, i, g32[i], o32)
    finally:
        for key, v in zip(keys, saved):
            gpu_ctx.set_option(key, v)
