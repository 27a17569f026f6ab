"""The adjoint DP (pg_count_grad) against brute force and FP64 central
finite differences, and the torch.autograd wrapper around it.

* per-entry gradients at unit weights equal brute-force per-entry counts of
  colorful maps (exact, FP64);
* with weights U(0.5, 1.5) the count is affine in every single entry weight,
  so FP64 central differences (h = 1e-3) equal the gradient up to rounding;
* FP32 gradients agree with FP64 to 1e-5 relative (norm-wise).
"""
import numpy as np
import pytest

import oracle
import subgraph_amd as sg

pytestmark = pytest.mark.gpu

STAR5 = [(0, 1), (0, 2), (0, 3), (0, 4)]
U5 = [(0, 1), (1, 2), (2, 3), (3, 4)]
TREE7 = [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]


def _simple_graph(rng, n, m):
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    keep = s != d
    pairs = {(int(min(a, b)), int(max(a, b))) for a, b in zip(s[keep], d[keep])}
    e = np.array(sorted(pairs), np.int32).reshape(-1, 2)
    return e[:, 0].copy(), e[:, 1].copy()


@pytest.mark.parametrize("tv,te", [(3, [(0, 1), (1, 2)]), (5, STAR5), (5, U5), (6, [(0, 1), (1, 2), (1, 3), (3, 4), (3, 5)]),
                                   (7, TREE7)])
def test_grad_equals_brute_force_per_entry(gpu_ctx, tv, te):
    rng = np.random.default_rng(tv * 7 + len(te))
    for trial in range(3):
        n = int(rng.integers(20, 60))
        s, d = _simple_graph(rng, n, 4 * n)
        colptr, rowidx = sg.csc_from_edges(n, s, d)
        gpu_ctx.upload_graph(colptr, rowidx)
        plan = gpu_ctx.plan(tv, te)
        cols = rng.integers(0, tv, size=(2, n)).astype(np.uint8)
        print(f"[bf-grad] k={tv} trial={trial} n={n} nnz={len(rowidx)}", flush=True)
        raw, grad = gpu_ctx.count_grad(plan, 2, colors=cols, precision="fp64")
        want = np.zeros(len(rowidx))
        for i in range(2):
            tot, pe = oracle.brute_force(tv, te, colptr, rowidx, cols[i], per_edge=True)
            assert raw[i] == tot
            want += pe
        assert np.array_equal(grad, want), np.flatnonzero(grad != want)[:10]


def test_weighted_grad_brute_force_and_fd(gpu_ctx):
    rng = np.random.default_rng(3)
    n = 40
    s, d = _simple_graph(rng, n, 170)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    # symmetric weights: one value per undirected edge, both entries
    w = np.zeros(len(rowidx))
    wu = {}
    for u in range(n):
        for e in range(colptr[u], colptr[u + 1]):
            key = (min(u, rowidx[e]), max(u, rowidx[e]))
            wu.setdefault(key, rng.uniform(0.5, 1.5))
            w[e] = np.float32(wu[key])
    gpu_ctx.upload_graph(colptr, rowidx, vals=w.astype(np.float32))
    plan = gpu_ctx.plan(5, U5)
    cols = rng.integers(0, 5, size=(3, n)).astype(np.uint8)
    raw, grad = gpu_ctx.count_grad(plan, 3, colors=cols, precision="fp64")
    want = np.zeros(len(rowidx))
    for i in range(3):
        tot, pe = oracle.brute_force(5, U5, colptr, rowidx, cols[i], vals=w, per_edge=True)
        assert abs(raw[i] - tot) <= 1e-12 * abs(tot)
        want += pe
    assert np.allclose(grad, want, rtol=1e-12, atol=1e-12)
    # central differences of the FP64 forward on single entries (affine in each)
    _, g1 = gpu_ctx.count_grad(plan, 1, colors=cols[:1], precision="fp64")
    op = oracle.Plan(5, U5)
    h = 1e-3
    for e in rng.choice(len(rowidx), 8, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fp = oracle.count_f64(op, colptr, rowidx, cols[0], vals=wp)
        fm = oracle.count_f64(op, colptr, rowidx, cols[0], vals=wm)
        assert abs((fp - fm) / (2 * h) - g1[e]) <= 1e-7 * max(1.0, abs(g1[e]))


def test_fp32_grad_close_to_fp64(gpu_ctx):
    n = 1 << 11
    s, d = sg.rmat_edges(11, 8, seed=2)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    for te, tv in ((TREE7, 7), (STAR5, 5)):
        plan = gpu_ctx.plan(tv, te)
        r64, g64 = gpu_ctx.count_grad(plan, 4, seed=9, precision="fp64")
        r32, g32 = gpu_ctx.count_grad(plan, 4, seed=9, precision="fp32")
        assert np.all(np.abs(r32 - r64) <= 1e-5 * np.abs(r64))
        assert np.linalg.norm(g32 - g64) <= 1e-5 * np.linalg.norm(g64)
        # the gradient's forward equals the plain forward
        assert np.array_equal(r64, gpu_ctx.count(plan, 4, seed=9, precision="fp64"))


def test_config4_fd_s18(gpu_ctx):
    """R-MAT s18/ef16, weights U(0.5,1.5) (symmetric), 5-star and 5-path, 16
    colorings: FP64 central differences (h=1e-3) of the device forward on 32
    sampled edges (both entries of each) against the adjoint."""
    rng = np.random.default_rng(18)
    n = 1 << 18
    s, d = sg.rmat_edges(18, 16, seed=1)
    w_und = rng.uniform(0.5, 1.5, len(s)).astype(np.float32)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    # entry weights: look up the undirected edge of every entry
    key = np.minimum(np.repeat(np.arange(n), np.diff(colptr)), rowidx).astype(np.int64) * n + \
        np.maximum(np.repeat(np.arange(n), np.diff(colptr)), rowidx)
    ukey = s.astype(np.int64) * n + d
    order = np.argsort(ukey)
    w = w_und[order[np.searchsorted(ukey[order], key)]]
    gpu_ctx.upload_graph(colptr, rowidx, vals=w)
    for name, te in (("star5", STAR5), ("u5", U5)):
        plan = gpu_ctx.plan(5, te)
        print(f"[config4] {name}: adjoint", flush=True)
        raw, grad = gpu_ctx.count_grad(plan, 16, seed=4, precision="fp64")
        for t in range(32):
            i = int(rng.integers(0, len(s)))
            a, b = int(s[i]), int(d[i])
            ents = [colptr[a] + int(np.searchsorted(rowidx[colptr[a]:colptr[a + 1]], b)),
                    colptr[b] + int(np.searchsorted(rowidx[colptr[b]:colptr[b + 1]], a))]
            # weights are FP32: perturb to the FP32 values nearest w +- h and
            # divide by the step actually taken (the count is affine in w_e)
            h = 1e-3
            wp = w.copy()
            wm = w.copy()
            wp[ents] = (w[ents].astype(np.float64) + h).astype(np.float32)
            wm[ents] = (w[ents].astype(np.float64) - h).astype(np.float32)
            gpu_ctx.set_weights(wp)
            fp = gpu_ctx.count(plan, 16, seed=4, precision="fp64").sum()
            gpu_ctx.set_weights(wm)
            fm = gpu_ctx.count(plan, 16, seed=4, precision="fp64").sum()
            step = float(wp[ents[0]]) - float(wm[ents[0]])
            fd = (fp - fm) / step
            g = grad[ents].sum()
            tol = 1e-6 * abs(g) + 1e-13 * abs(raw.sum()) / step
            assert abs(fd - g) <= tol, (name, t, fd, g)
        gpu_ctx.set_weights(w)


def test_autograd_function(gpu_ctx):
    torch = pytest.importorskip("torch")
    from subgraph_amd.autograd import colorful_count
    n = 1 << 10
    s, d = sg.rmat_edges(10, 8, seed=3)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    gpu_ctx.upload_graph(colptr, rowidx)
    plan = gpu_ctx.plan(5, U5)
    w = torch.rand(len(rowidx), dtype=torch.float32, device="cuda") + 0.5
    w.requires_grad_(True)
    raw = colorful_count(w, gpu_ctx, plan, 4, seed=2, precision="fp64")
    coef = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64, device="cuda")
    (raw * coef).sum().backward()
    gpu_ctx.set_weights(w.detach().cpu().numpy())
    r2, g2 = gpu_ctx.count_grad(plan, 4, seed=2, precision="fp64", cotangent=coef.cpu().numpy())
    assert np.allclose(raw.detach().cpu().numpy(), r2, rtol=1e-12)
    assert np.allclose(w.grad.cpu().numpy(), g2.astype(np.float32), rtol=1e-6, atol=1e-6)


TREE10 = [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]
TREE12 = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]


@pytest.mark.parametrize("tv,te", [(7, TREE7), (10, TREE10), (12, TREE12)])
def test_recompute_matches_kept_tables(gpu_ctx, tv, te):
    """The adjoint that rebuilds tables from their subtrees (option
    "adjoint_recompute" = 1) equals the one that keeps every table: counts
    bit for bit, gradients exactly for integer contributions (unit weights,
    FP64) and to FP64 rounding otherwise (the two sweeps visit the
    subtemplates in different orders)."""
    rng = np.random.default_rng(tv)
    n = 1 << 10
    s, d = sg.rmat_edges(10, 8, seed=tv)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0).astype(np.float32)  # symmetric
    saved = gpu_ctx.get_option("adjoint_recompute")
    gpu_ctx.set_option("fuse_grad", 0)  # the same gradient gathers in both sweeps
    try:
        for weighted in (False, True):
            gpu_ctx.upload_graph(colptr, rowidx, vals=w if weighted else None)
            plan = gpu_ctx.plan(tv, te)
            for prec in ("fp64", "fp32"):
                out = {}
                for mode in (0, 1):
                    gpu_ctx.set_option("adjoint_recompute", mode)
                    print(f"[recompute] k={tv} weighted={weighted} {prec} mode={mode}", flush=True)
                    out[mode] = gpu_ctx.count_grad(plan, 5, seed=tv, precision=prec)
                    assert gpu_ctx.get_option("last_recompute") == mode
                (r0, g0), (r1, g1) = out[0], out[1]
                assert np.array_equal(r0, r1)
                if prec == "fp64" and not weighted and np.abs(g0).max() < 2.0 ** 50:
                    assert np.array_equal(g0, g1)  # integer partial sums: exact in any order
                else:
                    assert np.allclose(g0, g1, rtol=1e-12, atol=0.0)
                assert np.any(g0 != 0)
    finally:
        gpu_ctx.set_option("adjoint_recompute", saved)
        gpu_ctx.set_option("fuse_grad", 1)


@pytest.mark.parametrize("color_runs", [-1, 1])
def test_fused_gradient_matches_gradient_gather(gpu_ctx, color_runs):
    """Option fuse_grad (default 1, kept tables): the forward keeps Y of the
    steps whose transposed gather runs by colour runs, and that gather also
    takes the entries' gradients (dY of each neighbour dotted with the
    vertex's own row).  Against the gradient gather that recomputes Y: FP64
    gradients exactly equal for unit weights (integer partial sums), to FP64
    rounding with weights; FP32 within 1e-6 in norm (the dot products sum in
    another order); counts bit for bit.  color_runs 1 fuses every step with a
    wide enough row; small hubs exercise the hub chunks of the fused gather."""
    n = 1 << 11
    s, d = sg.rmat_edges(11, 16, seed=23)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0).astype(np.float32)  # symmetric
    keys = ("adjoint_recompute", "fuse_grad", "color_runs", "hub_degree", "hub_chunk")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("adjoint_recompute", 0)
    gpu_ctx.set_option("color_runs", color_runs)
    gpu_ctx.set_option("hub_degree", 48)
    gpu_ctx.set_option("hub_chunk", 32)
    try:
        for weighted in (False, True):
            gpu_ctx.upload_graph(colptr, rowidx, vals=w if weighted else None)
            for tv, te in ((12, TREE12), (10, TREE10)):
                plan = gpu_ctx.plan(tv, te)
                for prec in ("fp64", "fp32"):
                    out = {}
                    for fuse in (0, 1):
                        gpu_ctx.set_option("fuse_grad", fuse)
                        print(f"[fused-grad] color_runs={color_runs} weighted={weighted} k={tv} {prec} fuse={fuse}",
                              flush=True)
                        out[fuse] = gpu_ctx.count_grad(plan, 5, seed=tv, precision=prec)
                        if fuse and (tv == 12 or color_runs == 1):
                            gpu_ctx.set_option("timing", 1)
                            gpu_ctx.count_grad(plan, 5, seed=tv, precision=prec)
                            gpu_ctx.set_option("timing", 0)
                            names = list(gpu_ctx.last_timing_kernels())
                            assert any(",tg," in k for k in names), names
                    (r0, g0), (r1, g1) = out[0], out[1]
                    assert np.array_equal(r0, r1)
                    assert np.any(g0 != 0)
                    if prec == "fp64" and not weighted and np.abs(g0).max() < 2.0 ** 50:
                        assert np.array_equal(g0, g1), np.flatnonzero(g0 != g1)[:10]
                    elif prec == "fp64":
                        assert np.allclose(g0, g1, rtol=1e-12, atol=0.0)
                    else:
                        assert np.linalg.norm(g1 - g0) <= 1e-6 * np.linalg.norm(g0)
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


def test_default_adjoint_mode(gpu_ctx):
    """Option "adjoint_recompute" = -1 (default) keeps the tables while they
    fit, and rebuilds them once a memory budget leaves no room to keep one
    coloring's tables; both give the kept-table results."""
    n = 1 << 16
    s, d = sg.rmat_edges(16, 8, seed=3)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    keys = ("adjoint_recompute", "mem_budget_mb")
    saved = [gpu_ctx.get_option(k) for k in keys]
    try:
        gpu_ctx.upload_graph(colptr, rowidx)
        plan = gpu_ctx.plan(10, TREE10)
        gpu_ctx.set_option("adjoint_recompute", -1)
        print("[default-adjoint] ample memory", flush=True)
        r_keep, g_keep = gpu_ctx.count_grad(plan, 2, seed=9, precision="fp64")
        assert gpu_ctx.get_option("last_recompute") == 0
        chosen = None
        for mb in (512, 400, 320, 256, 200, 160, 128, 100, 80, 64, 50, 40, 32):
            gpu_ctx.set_option("mem_budget_mb", mb)
            print(f"[default-adjoint] budget {mb} MB", flush=True)
            try:
                r, g = gpu_ctx.count_grad(plan, 2, seed=9, precision="fp64")
            except sg.PGError:
                break  # not even the rebuilt tables fit
            if gpu_ctx.get_option("last_recompute") == 1:
                chosen = mb
                # the two runs may differ in colorings per pass, hence in the
                # order of their FP64 sums: exact only below 2^53
                if np.abs(r_keep).max() < 2.0 ** 53:
                    assert np.array_equal(r, r_keep)
                else:
                    assert np.allclose(r, r_keep, rtol=1e-12, atol=0.0)
                assert np.allclose(g, g_keep, rtol=1e-12, atol=0.0)
                break
        assert chosen is not None, "no budget made the default rebuild tables"
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


def test_grad_graph_replays_match_eager(gpu_ctx):
    """Gradient passes (forward + reverse sweep) replay from a captured graph
    per shape: every replay equals the eager launches bit for bit, in both
    table modes and precisions, and after new weight values (the same
    buffers, no new capture) the count and gradient replays use them."""
    n = 1 << 11
    s, d = sg.rmat_edges(11, 8, seed=31)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)

    def weights(salt):
        return (0.5 + ((lo * 7919 + hi * 104729 + salt) % 1000) / 1000.0).astype(np.float32)

    keys = ("graphs", "adjoint_recompute")
    saved = [gpu_ctx.get_option(k) for k in keys]
    try:
        gpu_ctx.upload_graph(colptr, rowidx, vals=weights(0))
        plan = gpu_ctx.plan(7, TREE7)
        for mode in (0, 1):
            gpu_ctx.set_option("adjoint_recompute", mode)
            for prec in ("fp64", "fp32"):
                for salt in (0, 17):
                    gpu_ctx.set_weights(weights(salt))
                    gpu_ctx.set_option("graphs", 0)
                    r_e, g_e = gpu_ctx.count_grad(plan, 3, seed=5, precision=prec)
                    c_e = gpu_ctx.count(plan, 3, seed=5, precision=prec)
                    gpu_ctx.set_option("graphs", 1)
                    for rep in range(3):
                        print(f"[grad-graphs] mode={mode} {prec} salt={salt} replay {rep}", flush=True)
                        r_g, g_g = gpu_ctx.count_grad(plan, 3, seed=5, precision=prec)
                        assert np.array_equal(r_g, r_e) and np.array_equal(g_g, g_e)
                        assert np.array_equal(gpu_ctx.count(plan, 3, seed=5, precision=prec), c_e)
                    assert np.any(g_e != 0)
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


def test_autotuned_gradient(gpu_ctx):
    """Option "autotune" = 1 on a gradient call: the kernel variants, degree
    buckets and (when smaller) column tile tuned on forward passes of the
    adjoint's tile are applied, and the adjoint's memory plan made again with
    them.  Counts and gradients equal the untuned call's: exactly for unit
    weights in FP64 (integer sums below 2^50), to FP64 rounding with
    weights."""
    n = 1 << 12
    s, d = sg.rmat_edges(12, 16, seed=41)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    assert np.diff(colptr).max() > 1024  # hubs under every hub_degree candidate
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0).astype(np.float32)
    keys = ("autotune", "adjoint_recompute", "hub_degree")
    saved = [gpu_ctx.get_option(k) for k in keys]
    try:
        for weighted in (False, True):
            gpu_ctx.upload_graph(colptr, rowidx, vals=w if weighted else None)
            for tv, te in ((7, TREE7), (10, TREE10)):
                plan = gpu_ctx.plan(tv, te)
                for mode in (0, 1):
                    gpu_ctx.set_option("adjoint_recompute", mode)
                    gpu_ctx.set_option("autotune", 0)
                    gpu_ctx.set_option("hub_degree", 1024)  # few hub items before tuning
                    r0, g0 = gpu_ctx.count_grad(plan, 4, seed=tv, precision="fp64")
                    tile0 = gpu_ctx.get_option("last_tile")
                    gpu_ctx.set_option("autotune", 1)
                    print(f"[tuned-grad] k={tv} weighted={weighted} mode={mode}", flush=True)
                    for rep in range(2):  # tuning, then the tuned settings recalled
                        r1, g1 = gpu_ctx.count_grad(plan, 4, seed=tv, precision="fp64")
                        # the tuned settings change the LDS a pass needs, so the widest
                        # tile that fits can differ from the untuned call's
                        assert 1 <= gpu_ctx.get_option("last_tile") <= 4
                        if not weighted and np.abs(g0).max() < 2.0 ** 50:
                            assert np.array_equal(r1, r0) and np.array_equal(g1, g0)
                        else:
                            assert np.allclose(r1, r0, rtol=1e-12, atol=0.0)
                            assert np.allclose(g1, g0, rtol=1e-12, atol=1e-9 * np.abs(g0).max())
                    # a forward call of the same tile reuses the tuned settings
                    c1 = gpu_ctx.count(plan, tile0, seed=tv, precision="fp64")
                    assert np.allclose(c1, r0[:tile0], rtol=1e-12, atol=0.0)
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)


@pytest.mark.parametrize("tv,te", [(10, TREE10), (12, TREE12)])
def test_wide_template_grad_fd(gpu_ctx, tv, te):
    """The wide templates' adjoint (main-table steps whose splits read each
    row value many times take the LDS-staged node kernel) against FP64
    central differences of the oracle's weighted count (affine in each entry
    weight), the forward against the oracle, and FP32 against FP64, on a
    small dense graph."""
    rng = np.random.default_rng(tv + 100)
    n = 48
    s, d = _simple_graph(rng, n, 420)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0).astype(np.float32).astype(np.float64)  # symmetric
    gpu_ctx.upload_graph(colptr, rowidx, vals=w.astype(np.float32))
    plan = gpu_ctx.plan(tv, te)
    op = oracle.Plan(tv, te)
    cols = rng.integers(0, tv, size=(1, n)).astype(np.uint8)
    print(f"[wide-fd] k={tv} n={n} nnz={len(rowidx)}", flush=True)
    raw, g = gpu_ctx.count_grad(plan, 1, colors=cols, precision="fp64")
    want = oracle.count_f64(op, colptr, rowidx, cols[0], vals=w)
    assert want > 0 and abs(raw[0] - want) <= 1e-12 * want
    assert np.count_nonzero(g) > 0
    h = 1e-3
    picks = np.concatenate([rng.choice(np.flatnonzero(g), 6, replace=False), rng.choice(len(rowidx), 4, replace=False)])
    for e in picks:
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (oracle.count_f64(op, colptr, rowidx, cols[0], vals=wp)
              - oracle.count_f64(op, colptr, rowidx, cols[0], vals=wm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-7 * max(1.0, abs(g).max()), (e, fd, g[e])
    # FP32 tables (the staged node kernel in float): within 1e-5 of FP64, norm-wise
    r32, g32 = gpu_ctx.count_grad(plan, 1, colors=cols, precision="fp32")
    assert abs(r32[0] - raw[0]) <= 1e-5 * raw[0]
    assert np.linalg.norm(g32 - g) <= 1e-5 * np.linalg.norm(g)


def _random_tree(rng, k):
    return [(int(rng.integers(0, j)), j) for j in range(1, k)]


def test_random_tree_grads_fd(gpu_ctx):
    """Adjoint of random trees (k = 6..12: split shapes on both sides of the
    staged node kernel's threshold, main and aux children of every size)
    against FP64 central differences of the oracle's weighted count."""
    rng = np.random.default_rng(2026)
    n = 40
    s, d = _simple_graph(rng, n, 360)
    colptr, rowidx = sg.csc_from_edges(n, s, d)
    rows = np.repeat(np.arange(n), np.diff(colptr))
    lo, hi = np.minimum(rows, rowidx), np.maximum(rows, rowidx)
    w = (0.5 + ((lo * 7919 + hi * 104729) % 1000) / 1000.0).astype(np.float32).astype(np.float64)
    gpu_ctx.upload_graph(colptr, rowidx, vals=w.astype(np.float32))
    h = 1e-3
    for trial in range(8):
        k = int(rng.integers(6, 13))
        te = _random_tree(rng, k)
        plan = gpu_ctx.plan(k, te)
        op = oracle.Plan(k, te)
        cols = rng.integers(0, k, size=(1, n)).astype(np.uint8)
        print(f"[random-grad] trial={trial} k={k} edges={te}", flush=True)
        raw, g = gpu_ctx.count_grad(plan, 1, colors=cols, precision="fp64")
        want = oracle.count_f64(op, colptr, rowidx, cols[0], vals=w)
        assert abs(raw[0] - want) <= 1e-12 * max(abs(want), 1.0), (trial, raw[0], want)
        nz = np.flatnonzero(g)
        picks = rng.choice(nz, min(4, len(nz)), replace=False) if len(nz) else rng.choice(len(rowidx), 2)
        for e in picks:
            wp, wm = w.copy(), w.copy()
            wp[e] += h
            wm[e] -= h
            fd = (oracle.count_f64(op, colptr, rowidx, cols[0], vals=wp)
                  - oracle.count_f64(op, colptr, rowidx, cols[0], vals=wm)) / (2 * h)
            assert abs(fd - g[e]) <= 1e-7 * max(1.0, abs(g).max()), (trial, k, e, fd, g[e])


def test_node_kernel_variants_agree_bit_for_bit(gpu_ctx):
    """The adjoint's node launches of the 12- and 10-vertex trees' main-table
    steps run the staged kernel (a thread per (row, value), items by kind)
    or, for rows of 32+ values with option node_rows 1, the rows-inner
    kernel (a thread per value over a group's rows).  Every value is the same
    sum in the same order either way: counts and gradients equal bit for bit,
    FP32 (FP32 split sums) and FP64.  With 2045 vertices the rows-inner
    kernel's last group of every launch is a partial one."""
    s, d = sg.rmat_edges(11, 16, seed=31)
    keys = ("adjoint_recompute", "node_rows")
    saved = [gpu_ctx.get_option(k) for k in keys]
    gpu_ctx.set_option("adjoint_recompute", 0)
    try:
        for n in (1 << 11, (1 << 11) - 3):
            keep = (s < n) & (d < n)
            colptr, rowidx = sg.csc_from_edges(n, s[keep], d[keep])
            gpu_ctx.upload_graph(colptr, rowidx)
            for tv, te in ((12, TREE12), (10, TREE10)):
                plan = gpu_ctx.plan(tv, te)
                for prec in ("fp32", "fp64"):
                    out = {}
                    for nr in (0, 1):
                        gpu_ctx.set_option("node_rows", nr)
                        print(f"[node variants] n={n} k={tv} {prec} node_rows={nr}", flush=True)
                        out[nr] = gpu_ctx.count_grad(plan, 3, seed=tv + 40, precision=prec)
                    (r0, g0), (r1, g1) = out[0], out[1]
                    assert np.array_equal(r0, r1)
                    assert np.any(g0 != 0)
                    assert np.array_equal(g0, g1), np.flatnonzero(g0 != g1)[:10]
    finally:
        for k, v in zip(keys, saved):
            gpu_ctx.set_option(k, v)
