"""subgraph_amd -- PGBSC color-coding subgraph counting on AMD MI355X.

Python face of libpgbsc.so.  The DP runs in hand-written HIP kernels; this
module only moves arrays across the C ABI (include/pgbsc.h).

Reference surface (dsc-nec/subgraph) and where it lives here:
  CSCGraph::createFromEdgeListFile  -> csc_from_edges()
  EdgeList hole removal             -> remove_holes()
  DivideTemplates + IndexSys        -> Context.plan()
  CountMat::colorCounting loop      -> Context.count()
  CountMat::initialization/compute  -> CountMat
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import (ESTIMATE_FIELDS, PG_FLAG_FIX_2VERTEX, PG_FLAG_UNPRUNED, PG_FP32, PG_FP64, PGError,
                   PgEstimate, exported_symbols, lib)

__all__ = ["Context", "Plan", "CountMat", "PGError", "rmat_edges", "csc_from_edges", "colors_fill",
           "remove_holes", "exported_symbols", "estimate_model", "PG_FP32", "PG_FP64"]

_PREC = {"fp32": PG_FP32, "fp64": PG_FP64, PG_FP32: PG_FP32, PG_FP64: PG_FP64}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def rmat_edges(scale: int, edge_factor: int, seed: int = 1, a: float = 0.57, b: float = 0.19,
               c: float = 0.19):
    """Undirected R-MAT edge list (src, dst) with src < dst, deduplicated."""
    L = lib()
    m = C.c_int64()
    rc = L.pg_rmat_edges(scale, edge_factor, a, b, c, seed, 0, None, None, C.byref(m))
    if rc:
        raise PGError(rc, "pg_rmat_edges")
    src = np.empty(m.value, np.int32)
    dst = np.empty(m.value, np.int32)
    rc = L.pg_rmat_edges(scale, edge_factor, a, b, c, seed, m.value, src.ctypes.data, dst.ctypes.data,
                         C.byref(m))
    if rc:
        raise PGError(rc, "pg_rmat_edges")
    return src, dst


def csc_from_edges(n: int, src, dst):
    """The reference's CSC (CSCGraph.hpp:96-178): colptr[n+1], rowidx[2m]."""
    src, dst = _i32(src), _i32(dst)
    colptr = np.empty(n + 1, np.int32)
    rowidx = np.empty(max(2 * len(src), 1), np.int32)
    rc = lib().pg_csc_from_edges(n, len(src), src, dst, colptr, rowidx.ctypes.data)
    if rc:
        raise PGError(rc, "pg_csc_from_edges")
    return colptr, rowidx[: 2 * len(src)]


def remove_holes(n: int, src, dst):
    """EdgeList.cpp:70-112 on an in-memory list; returns (n, src, dst)."""
    src, dst = _i32(src).copy(), _i32(dst).copy()
    r = lib().pg_edges_remove_holes(n, len(src), src, dst)
    if r < 0:
        raise PGError(int(r), "pg_edges_remove_holes")
    return int(r), src, dst


def template_tables(nv: int, edges) -> dict:
    """Host-only DivideTemplates + IndexSys output (what pg_plan_create takes)."""
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    s, d = _i32(e[:, 0]), _i32(e[:, 1])
    L = lib()
    nsub, ncomb, nsplit = C.c_int32(), C.c_int64(), C.c_int64()
    rc = L.pg_template_sizes(nv, len(e), s, d, C.byref(nsub), C.byref(ncomb), C.byref(nsplit))
    if rc:
        raise PGError(rc, "pg_template_sizes")
    out = {"k": nv, "sub_size": np.zeros(nsub.value, np.int32), "main_idx": np.zeros(nsub.value, np.int32),
           "aux_idx": np.zeros(nsub.value, np.int32), "comb_to_count": np.zeros(ncomb.value, np.int32),
           "split_main": np.zeros(max(nsplit.value, 1), np.int32),
           "split_aux": np.zeros(max(nsplit.value, 1), np.int32)}
    rc = L.pg_template_tables(nv, len(e), s, d, out["sub_size"], out["main_idx"], out["aux_idx"],
                              out["comb_to_count"], out["split_main"], out["split_aux"])
    if rc:
        raise PGError(rc, "pg_template_tables")
    out["split_main"] = out["split_main"][: nsplit.value]
    out["split_aux"] = out["split_aux"][: nsplit.value]
    return out


def estimate_model(n: int, nnz: int, nv: int, edges) -> dict:
    """The reference's cost models (CountMat::estimatePeakMemUsage,
    estimateFlops*/estimateMemComm* and estimateTemplate, CountMat.cpp:
    1067-1496) for a graph of n vertices and nnz directed entries and a
    template tree: GB / GBytes / Gflop are 2^30.  Host only."""
    t = template_tables(nv, edges)
    e = PgEstimate()
    rc = lib().pg_estimate_model(int(n), int(nnz), int(nv), len(t["sub_size"]), t["sub_size"], t["main_idx"],
                                 t["aux_idx"], C.byref(e))
    if rc:
        raise PGError(rc, "pg_estimate_model")
    return {f: getattr(e, f) for f in ESTIMATE_FIELDS}


def colors_fill(seed: int, first: int, ncolorings: int, n: int, k: int) -> np.ndarray:
    """The device color stream, reproduced on the host: [ncolorings][n] uint8."""
    out = np.empty((ncolorings, n), np.uint8)
    lib().pg_colors_fill(seed, first, ncolorings, n, k, out.reshape(-1))
    return out


class Plan:
    """A template's subtemplate chain and index tables on the device."""

    def __init__(self, ctx: "Context", ptr: int, nv: int):
        self._ctx = ctx
        self.ptr = ptr
        # bound now: at interpreter exit the module's globals may already be gone
        self._destroy = lib().pg_plan_destroy
        k, nsub, nonleaf, scaled = (C.c_int() for _ in range(4))
        lib().pg_plan_info(ptr, C.byref(k), C.byref(nsub), C.byref(nonleaf), C.byref(scaled))
        self.k, self.nsub, self.n_nonleaf, self.scaled_sub = k.value, nsub.value, nonleaf.value, scaled.value

    def __del__(self):
        if getattr(self, "ptr", None):
            self._destroy(self.ptr)
            self.ptr = None


class Context:
    """One device, one graph, any number of plans (pg_ctx)."""

    def __init__(self, device: int = 0, **options):
        L = lib()
        p = C.c_void_p()
        rc = L.pg_create(C.byref(p), device)
        if rc:
            raise PGError(rc, f"pg_create(device={device}): no usable HIP device")
        self.ptr = p.value
        self._destroy = L.pg_destroy  # bound now, as for Plan
        self.n = self.nnz = 0
        for k, v in options.items():
            self.set_option(k, v)

    def __del__(self):
        if getattr(self, "ptr", None):
            self._destroy(self.ptr)
            self.ptr = None

    def _check(self, rc: int, what: str):
        if rc:
            raise PGError(rc, f"{what}: {lib().pg_last_error(self.ptr).decode()}")

    def set_option(self, key: str, value: int):
        self._check(lib().pg_set_option(self.ptr, key.encode(), int(value)), f"pg_set_option({key})")

    def get_option(self, key: str) -> int:
        v = C.c_int64()
        self._check(lib().pg_get_option(self.ptr, key.encode(), C.byref(v)), f"pg_get_option({key})")
        return v.value

    def upload_graph(self, colptr, rowidx, vals=None):
        colptr, rowidx = _i32(colptr), _i32(rowidx)
        vp = None
        if vals is not None:
            vals = np.ascontiguousarray(vals, np.float32)
            vp = vals.ctypes.data
        self._check(lib().pg_graph_upload(self.ptr, len(colptr) - 1, len(rowidx), colptr, rowidx, vp),
                    "pg_graph_upload")
        self.n, self.nnz = len(colptr) - 1, len(rowidx)

    def set_weights(self, vals=None, device_ptr: int | None = None):
        if device_ptr is not None:
            self._check(lib().pg_graph_set_weights(self.ptr, device_ptr, 1), "pg_graph_set_weights")
        elif vals is None:
            self._check(lib().pg_graph_set_weights(self.ptr, None, 0), "pg_graph_set_weights")
        else:
            vals = np.ascontiguousarray(vals, np.float32)
            self._check(lib().pg_graph_set_weights(self.ptr, vals.ctypes.data, 0), "pg_graph_set_weights")

    def footprint(self, plan: Plan, colorings_per_pass: int = 1, precision="fp32") -> dict:
        """Device bytes of a pass of `colorings_per_pass` colorings (memory
        plan + graph) and the algorithmic HBM bytes per coloring that the
        roofline figures use."""
        dev = C.c_int64()
        alg = C.c_double()
        self._check(lib().pg_plan_footprint(self.ptr, plan.ptr, int(colorings_per_pass), _PREC[precision],
                                            C.byref(dev), C.byref(alg)), "pg_plan_footprint")
        return {"device_bytes": dev.value, "algorithmic_bytes_per_coloring": alg.value}

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        self._check(lib().pg_graph_stats(self.ptr, *[C.byref(x) for x in v]), "pg_graph_stats")
        return dict(zip(("n", "nnz", "n_active", "max_degree"), (x.value for x in v)))

    def plan(self, nv: int, edges) -> Plan:
        e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
        p = C.c_void_p()
        self._check(lib().pg_plan_from_edges(self.ptr, nv, len(e), _i32(e[:, 0]), _i32(e[:, 1]), C.byref(p)),
                    "pg_plan_from_edges")
        return Plan(self, p.value, nv)

    def plan_from_tables(self, k, sub_size, main_idx, aux_idx, comb_to_count, split_main, split_aux) -> Plan:
        p = C.c_void_p()
        self._check(lib().pg_plan_create(self.ptr, k, len(sub_size), _i32(sub_size), _i32(main_idx),
                                         _i32(aux_idx), _i32(comb_to_count), _i32(split_main),
                                         _i32(split_aux), C.byref(p)), "pg_plan_create")
        return Plan(self, p.value, k)

    def count(self, plan: Plan, ncolorings: int, colors=None, seed: int = 0, precision="fp32",
              fix_2vertex: bool = False, unpruned: bool = False) -> np.ndarray:
        """Raw colorful count of each coloring (CountMat::colorCounting).

        unpruned=True runs the reference's isPruned = 0 mode with its
        numerics (countNonBottomeOriginal, CountMat.cpp:851-977), whose
        accumulator carries over from one coloring and call to the next until
        the graph is uploaded again."""
        out = np.zeros(ncolorings, np.float64)
        cp = None
        if colors is not None:
            colors = np.ascontiguousarray(colors, np.uint8)
            if colors.size != ncolorings * self.n:
                raise ValueError("colors must be [ncolorings][n]")
            cp = colors.ctypes.data
        flags = (PG_FLAG_FIX_2VERTEX if fix_2vertex else 0) | (PG_FLAG_UNPRUNED if unpruned else 0)
        self._check(lib().pg_count(self.ptr, plan.ptr, ncolorings, cp, seed, _PREC[precision], flags, out),
                    "pg_count")
        return out

    def count_grad(self, plan: Plan, ncolorings: int, colors=None, seed: int = 0, precision="fp64",
                   fix_2vertex: bool = False, cotangent=None):
        """Raw counts and d(sum_c cot_c * raw_c)/d(entry weight), caller entry order."""
        raw = np.zeros(ncolorings, np.float64)
        grad = np.zeros(max(self.nnz, 1), np.float64)
        cp = None
        if colors is not None:
            colors = np.ascontiguousarray(colors, np.uint8)
            if colors.size != ncolorings * self.n:
                raise ValueError("colors must be [ncolorings][n]")
            cp = colors.ctypes.data
        cot = None
        if cotangent is not None:
            cotangent = np.ascontiguousarray(cotangent, np.float64)
            cot = cotangent.ctypes.data
        flags = PG_FLAG_FIX_2VERTEX if fix_2vertex else 0
        self._check(lib().pg_count_grad(self.ptr, plan.ptr, ncolorings, cp, seed, _PREC[precision], flags, cot,
                                        raw, grad.ctypes.data, 0), "pg_count_grad")
        return raw, grad[: self.nnz]

    def count_grad_device(self, plan: Plan, ncolorings: int, grad_ptr: int, seed: int = 0, precision="fp32",
                          cotangent=None, fix_2vertex: bool = False) -> np.ndarray:
        """As count_grad, writing the gradient to a device buffer (float for fp32, double for fp64)."""
        raw = np.zeros(ncolorings, np.float64)
        cot = None
        if cotangent is not None:
            cotangent = np.ascontiguousarray(cotangent, np.float64)
            cot = cotangent.ctypes.data
        flags = PG_FLAG_FIX_2VERTEX if fix_2vertex else 0
        self._check(lib().pg_count_grad(self.ptr, plan.ptr, ncolorings, None, seed, _PREC[precision], flags, cot,
                                        raw, grad_ptr, 1), "pg_count_grad")
        return raw

    def count_multi(self, plans, ncolorings, seeds, precision="fp32") -> list[np.ndarray]:
        arr = (C.c_void_p * len(plans))(*[p.ptr for p in plans])
        nc = _i32(ncolorings)
        sd = (C.c_uint64 * len(plans))(*seeds)
        out = np.zeros(int(nc.sum()), np.float64)
        self._check(lib().pg_count_multi(self.ptr, len(plans), C.cast(arr, C.c_void_p), nc,
                                         C.cast(sd, C.c_void_p), _PREC[precision], 0, out), "pg_count_multi")
        return np.split(out, np.cumsum(nc)[:-1])

    def last_timing(self) -> dict:
        t, km, kb = C.c_double(), C.c_double(), C.c_double()
        kl = C.c_int64()
        lib().pg_last_timing(self.ptr, C.byref(t), C.byref(km), C.byref(kl), C.byref(kb))
        buf = C.create_string_buffer(128)
        lib().pg_last_timing_kernel(self.ptr, buf, 128)
        return {"total_ms": t.value, "kernel": buf.value.decode(), "kernel_ms": km.value,
                "kernel_launches": kl.value, "kernel_bytes": kb.value}

    def last_timing_kernels(self) -> dict:
        """{kernel name: (summed ms, launches, algorithmic bytes)} of the last
        count call made with option "timing" = 1."""
        out = {}
        buf = C.create_string_buffer(128)
        ms, nl, nb = C.c_double(), C.c_int64(), C.c_double()
        idx = 0
        while True:
            rc = lib().pg_last_timing_entry(self.ptr, idx, buf, 128, C.byref(ms), C.byref(nl), C.byref(nb))
            if rc == -6:          # PG_ERR_RANGE: past the last kernel
                break
            if rc < 0:
                self._check(rc, "pg_last_timing_entry")
            out[buf.value.decode()] = (ms.value, nb.value, nl.value)
            idx += 1
            if idx >= rc:
                break
        return out


    def last_timing_subs(self) -> dict:
        """{(subtemplate, kernel): (ms, algorithmic bytes, launches)} of the
        last call with option timing on; subtemplate -1 = pass-level launches."""
        out = {}
        buf = C.create_string_buffer(128)
        sub, ms, nl, nb = C.c_int(), C.c_double(), C.c_int64(), C.c_double()
        idx = 0
        while True:
            rc = lib().pg_last_timing_sub(self.ptr, idx, C.byref(sub), buf, 128, C.byref(ms), C.byref(nl),
                                          C.byref(nb))
            if rc == -6:
                break
            if rc < 0:
                self._check(rc, "pg_last_timing_sub")
            out[(sub.value, buf.value.decode())] = (ms.value, nb.value, nl.value)
            idx += 1
            if idx >= rc:
                break
        return out


class CountMat:
    """CountMat.hpp:19-191: initialization(graph, ...) then compute(template)."""

    def initialization(self, colptr, rowidx, thd_num: int, itr_num: int, isPruned: int = 1,
                       useSPMM: int = 1, vtuneStart: int = -1, calculate_automorphisms: bool = False,
                       device: int = 0, precision="fp32", seed: int = 0):
        self.ctx = Context(device)
        self.ctx.upload_graph(colptr, rowidx)
        self.n, self.nnz = len(colptr) - 1, len(rowidx)
        self.itr_num, self.precision, self.seed = itr_num, precision, seed
        self.calculate_automorphisms = calculate_automorphisms
        # useSPMM = 0 (countNonBottomePruned, CountMat.cpp:325-554) forms the
        # same FP32 sums in the same order as the SpMM path: one mode here
        self.isPruned, self.useSPMM = isPruned, useSPMM

    def compute(self, nv: int, edges, isEstimate: bool = False) -> float:
        self.template = (nv, edges)
        plan = self.ctx.plan(nv, edges)
        if isEstimate:
            return 0.0
        self.raw = self.ctx.count(plan, self.itr_num, seed=self.seed, precision=self.precision,
                                  unpruned=self.isPruned == 0)
        k = nv
        prob = math.factorial(k) / float(k) ** k  # CountMat.cpp:248-249
        aut = automorphisms(nv, edges) if self.calculate_automorphisms else 1
        return math.floor(float(np.mean(self.raw)) / (prob * aut) + 0.5)

    def estimates(self) -> dict:
        """Every figure of the reference's estimate* models (CountMat.cpp:
        1067-1496) for the graph and the template of the last compute()."""
        return estimate_model(self.n, self.nnz, *self.template)


def automorphisms(nv: int, edges) -> int:
    """CountMat::automorphismNum (CountMat.cpp:1498-1572): permutations s with
    deg(i) = deg(s(i)) and N(i) a subset of s(N(s(i))) for every vertex i, the
    reference's acceptance rule (pg_template_automorphisms)."""
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    r = lib().pg_template_automorphisms(nv, len(e), np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1]))
    if r < 0:
        raise PGError(int(r), "pg_template_automorphisms")
    return int(r)
