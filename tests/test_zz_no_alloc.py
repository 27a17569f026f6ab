"""pg_count_noalloc (no reference counterpart: the entry point for callers
that cannot allocate device memory) against pg_count, through the C ABI.

Two calls in a row, the second on a smaller graph (the static scratch then
holds the first call's contents), each checked against pg_count on the same
inputs and pass width: FP64 counts exactly equal (integers below 2^53),
FP32 within 1e-5 relative.  A shape beyond the scratch is refused.
"""
import ctypes as C

import numpy as np
import pytest

import subgraph_amd as sg
from subgraph_amd import _lib

pytestmark = pytest.mark.gpu

TREE7 = [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]
TREE12 = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]


def _noalloc(ctx, plan, ncol, colors=None, seed=0, precision="fp32"):
    out = np.zeros(ncol, np.float64)
    cp = None
    if colors is not None:
        colors = np.ascontiguousarray(colors, np.uint8)
        cp = colors.ctypes.data
    prec = sg.PG_FP32 if precision == "fp32" else sg.PG_FP64
    rc = _lib.lib().pg_count_noalloc(ctx.ptr, plan.ptr, ncol, cp, seed, prec, 0,
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def _graph(scale, ef, seed):
    s, d = sg.rmat_edges(scale, ef, seed=seed)
    n, s, d = sg.remove_holes(1 << scale, s, d)
    return (n,) + sg.csc_from_edges(n, s, d)


def _check(ctx, plan, ncol, n, k, tag):
    cols = sg.colors_fill(31, 0, ncol, n, k)
    for prec in ("fp64", "fp32"):
        for colors, seed in ((cols, 0), (None, 77)):
            want = ctx.count(plan, ncol, colors=colors, seed=seed, precision=prec)
            tile = ctx.get_option("last_tile")
            rc, got = _noalloc(ctx, plan, ncol, colors=colors, seed=seed, precision=prec)
            assert rc == 0, (tag, rc, _lib.lib().pg_last_error(ctx.ptr))
            assert ctx.get_option("last_tile") == tile, tag
            if prec == "fp64":
                assert np.array_equal(got, want), (tag, got.tolist(), want.tolist())
            else:
                assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want)), (tag, got.tolist(), want.tolist())
            print(f"[noalloc] {tag} {prec} colors={'host' if colors is not None else 'device'}: {ncol} colorings "
                  f"at {tile} per pass, bit-equal {np.array_equal(got, want)}", flush=True)


def test_noalloc_matches_count_twice():
    ctx = sg.Context(0)
    ctx.set_option("batch_tile", 4)  # the same pass width for both entry points
    n, colptr, rowidx = _graph(13, 8, 3)
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(7, TREE7)
    _check(ctx, plan, 10, n, 7, "s13/ef8 tree7")
    # second call: a smaller graph and template, fewer colorings
    n2, colptr2, rowidx2 = _graph(10, 8, 5)
    ctx.upload_graph(colptr2, rowidx2)
    plan2 = ctx.plan(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    _check(ctx, plan2, 3, n2, 5, "s10/ef8 u5")


def test_noalloc_refuses_beyond_scratch():
    ctx = sg.Context(0)
    n, colptr, rowidx = _graph(16, 16, 1)
    ctx.upload_graph(colptr, rowidx)
    plan = ctx.plan(12, TREE12)
    rc, _ = _noalloc(ctx, plan, 1, seed=1)
    assert rc == -6  # PG_ERR_RANGE
    assert b"static" in _lib.lib().pg_last_error(ctx.ptr)
