"""Python access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker.  The product
(``subgraph_amd``) never imports it.

* ``liboracle.so`` (built from ``oracle.c`` by ``make -C oracle oracle``) is a
  plain-C restatement of the reference's DP; ``oracle.h`` cites the reference
  file:line behind every function.
* ``_ref/pg_ref`` (built by ``make -C oracle ref`` where the reference sources
  exist) is the reference itself, compiled from its own sources and driven by
  ``ref_driver.cpp``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "pg_ref")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_lib = None


def build() -> None:
    """Compile liboracle.so (and _ref/pg_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle oracle)")
        L = C.CDLL(LIB_PATH)
        L.or_plan_create.restype = C.c_void_p
        L.or_plan_create.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        L.or_plan_free.argtypes = [C.c_void_p]
        L.or_plan_k.argtypes = [C.c_void_p]
        L.or_plan_nsub.argtypes = [C.c_void_p]
        L.or_plan_sub.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 6
        L.or_plan_tables.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _i32p]
        L.or_comb.argtypes = [C.c_int, C.c_int]
        for f in (L.or_read_edgelist, L.or_read_template):
            f.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int))]
        L.or_free.argtypes = [C.c_void_p]
        L.or_csc_build.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p]
        L.or_count_f32.restype = C.c_double
        L.or_count_f32.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, _u8p]
        L.or_count_f64.restype = C.c_double
        L.or_count_f64.argtypes = [C.c_void_p, C.c_int, _i32p, _i32p, C.c_void_p, _u8p, C.c_int]
        L.or_brute_force.restype = C.c_double
        L.or_brute_force.argtypes = [C.c_int, C.c_int, _i32p, _i32p, C.c_int, _i32p, _i32p,
                                     C.c_void_p, _u8p, C.c_void_p]
        _lib = L
    return _lib


def comb(n: int, k: int) -> int:
    return lib().or_comb(n, k)


class Plan:
    """or_plan: DivideTemplates + IndexSys restated (oracle.h)."""

    def __init__(self, nv: int, edges):
        e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
        self.nv = nv
        self.edges = e
        self._src = np.ascontiguousarray(e[:, 0])
        self._dst = np.ascontiguousarray(e[:, 1])
        self.ptr = lib().or_plan_create(nv, len(e), self._src, self._dst)
        if not self.ptr:
            raise ValueError("oracle rejected the template (needs a tree with 2..16 vertices)")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().or_plan_free(self.ptr)
            self.ptr = None

    @property
    def k(self) -> int:
        return lib().or_plan_k(self.ptr)

    @property
    def nsub(self) -> int:
        return lib().or_plan_nsub(self.ptr)

    def sub(self, s: int) -> dict:
        v = [C.c_int() for _ in range(6)]
        lib().or_plan_sub(self.ptr, s, *[C.byref(x) for x in v])
        keys = ("size", "main", "aux", "parent", "ncomb", "nsplit")
        return dict(zip(keys, (x.value for x in v)))

    def tables(self, s: int):
        d = self.sub(s)
        c2c = np.zeros(d["ncomb"], np.int32)
        n = max(d["ncomb"] * d["nsplit"], 1)
        sm = np.zeros(n, np.int32)
        sa = np.zeros(n, np.int32)
        lib().or_plan_tables(self.ptr, s, c2c, sm, sa)
        if d["nsplit"] == 0:
            return c2c, None, None
        return c2c, sm.reshape(d["ncomb"], d["nsplit"]), sa.reshape(d["ncomb"], d["nsplit"])


def _read(fn, path):
    n, m = C.c_int(), C.c_int()
    s, d = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    rc = fn(path.encode(), C.byref(n), C.byref(m), C.byref(s), C.byref(d))
    if rc:
        raise OSError(f"oracle could not read {path} ({rc})")
    src = np.ctypeslib.as_array(s, shape=(max(m.value, 1),))[: m.value].copy()
    dst = np.ctypeslib.as_array(d, shape=(max(m.value, 1),))[: m.value].copy()
    lib().or_free(C.cast(s, C.c_void_p))
    lib().or_free(C.cast(d, C.c_void_p))
    return n.value, src.astype(np.int32), dst.astype(np.int32)


def read_edgelist(path: str):
    return _read(lib().or_read_edgelist, path)


def read_template(path: str):
    return _read(lib().or_read_template, path)


def csc_build(n: int, src, dst):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    colptr = np.zeros(n + 1, np.int32)
    rowidx = np.zeros(max(2 * len(src), 1), np.int32)
    lib().or_csc_build(n, len(src), src, dst, colptr, rowidx)
    return colptr, rowidx[: 2 * len(src)]


def count_f32(plan: Plan, colptr, rowidx, colors) -> float:
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    colors = np.ascontiguousarray(colors, np.uint8)
    return lib().or_count_f32(plan.ptr, len(colptr) - 1, colptr, rowidx, colors)


def count_f64(plan: Plan, colptr, rowidx, colors, vals=None, compat_2vertex=True) -> float:
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    colors = np.ascontiguousarray(colors, np.uint8)
    vp = None
    if vals is not None:
        vals = np.ascontiguousarray(vals, np.float64)
        vp = vals.ctypes.data_as(C.c_void_p)
    return lib().or_count_f64(plan.ptr, len(colptr) - 1, colptr, rowidx, vp, colors,
                              1 if compat_2vertex else 0)


def brute_force(nv: int, edges, colptr, rowidx, colors, vals=None, per_edge=False):
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    colors = np.ascontiguousarray(colors, np.uint8)
    vp = None
    if vals is not None:
        vals = np.ascontiguousarray(vals, np.float64)
        vp = vals.ctypes.data_as(C.c_void_p)
    ec = np.zeros(len(rowidx), np.float64) if per_edge else None
    tot = lib().or_brute_force(nv, len(e), np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1]),
                               len(colptr) - 1, colptr, rowidx, vp, colors,
                               ec.ctypes.data_as(C.c_void_p) if per_edge else None)
    return (tot, ec) if per_edge else tot


def have_ref() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(graph_path: str, tmpl_path: str, colors_path: str, ncol: int, threads: int = 1,
            seed: int = 1, omp_threads: int | None = None, timeout: float | None = None):
    """Run the compiled reference; returns ([raw per coloring], [seconds per coloring])."""
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(omp_threads if omp_threads is not None else threads)
    out = subprocess.run([REF_BIN, graph_path, tmpl_path, colors_path, str(ncol), str(threads), str(seed)],
                         check=True, capture_output=True, text=True, env=env, timeout=timeout).stdout
    raws, secs = [], []
    for line in out.splitlines():
        if line.startswith("PGREF coloring"):
            f = line.split()
            raws.append(float(f[4]))
            secs.append(float(f[6]))
    return raws, secs


def estimate_ref(graph_path: str, tmpl_path: str, workdir: str) -> dict:
    """The reference's cost models (CountMat.cpp:1067-1400) for a graph /
    template pair: the figures at full precision, the printed peak-memory
    line's value, and degDistri.data as [[degree, count], ...]."""
    out = subprocess.run([REF_BIN, "estimate", graph_path, tmpl_path], check=True, cwd=workdir,
                         capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
    res: dict = {}
    for line in out.splitlines():
        if line.startswith("PGREF est "):
            f = line.split()
            res[f[2]] = float(f[3])
        elif line.startswith("PGREF n "):
            f = line.split()
            res["n"], res["nnz"] = int(f[2]), int(f[4])
        elif line.startswith("Peak memory usage estimated : "):
            res["peak_mem_gb_printed"] = line.split(":")[1].split()[0]
        elif line.startswith("maxDeg: "):
            res["max_deg"] = int(line.split()[1])
    with open(os.path.join(workdir, "degDistri.data")) as f:
        res["deg_distribution"] = [[int(a) for a in ln.split()] for ln in f if ln.strip()]
    return res


def estimate_template_ref(graph_path: str, tmpl_path: str, timeout: float = 60.0) -> dict:
    """estimateTemplate's two printed lines (CountMat.cpp:1448-1496).  The
    function lacks a return statement and, built at -O3, does not return:
    the child is ended once both lines have been read."""
    proc = subprocess.Popen([REF_BIN, "estimate_template", graph_path, tmpl_path], stdout=subprocess.PIPE,
                            stderr=subprocess.DEVNULL, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    res: dict = {}
    deadline = time.monotonic() + timeout
    try:
        for line in proc.stdout:
            if line.startswith("Workload N: "):
                a, b = line.split(":", 1)[1].split(",")
                res["workload_n"], res["workload_nnz"] = float(a), float(b.split(":")[1])
            elif line.startswith("Memload N: "):
                a, b = line.split(":", 1)[1].split(",")
                res["memload_n"], res["memload_nnz"] = float(a), float(b.split(":")[1])
                break
            if time.monotonic() > deadline:
                break
    finally:
        proc.kill()
        proc.wait()
    if len(res) != 4:
        raise RuntimeError("estimateTemplate printed no figures")
    return res


def dump_ref(graph_path: str, tmpl_path: str) -> dict:
    """The reference's CSC and plan tables for a graph/template pair."""
    out = subprocess.run([REF_BIN, "dump", graph_path, tmpl_path], check=True,
                         capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
    res: dict = {"sub": {}, "c2c": {}, "smain": {}, "saux": {}}
    for line in out.splitlines():
        if not line.startswith("PGREF "):
            continue
        f = line.split()
        tag, s, vals = f[1], int(f[2]), [int(x) for x in f[3:]]
        if tag in ("colptr", "rowidx"):
            res[tag] = vals
        elif tag == "sub":
            res["sub"][s] = vals
        else:
            res[tag][s] = vals
    return res
