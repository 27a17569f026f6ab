"""Host surface of the product (no GPU): template split and index tables,
CSC build and loaders against the compiled reference's outputs, the R-MAT
source, the color stream, the CLI's error paths and binary graph files, and
the C ABI's exported symbols."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import subgraph_amd as sg
from subgraph_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SC_BIN = os.path.join(ROOT, "subgraph_amd", "sc-bin")


def test_template_tables_match_reference(golden):
    for p in golden["plans"]:
        t = sg.template_tables(p["nv"], p["edges"])
        assert len(t["sub_size"]) == len(p["sub"])
        off = offs = 0
        for s, (size, mi, ai) in enumerate(p["sub"]):
            assert (t["sub_size"][s], t["main_idx"][s], t["aux_idx"][s]) == (size, mi, ai), (p["name"], s)
            c2c = p["c2c"][s]
            assert t["comb_to_count"][off:off + len(c2c)].tolist() == c2c
            off += len(c2c)
            if str(s) in p["smain"]:
                sm, sa = p["smain"][str(s)], p["saux"][str(s)]
                assert t["split_main"][offs:offs + len(sm)].tolist() == sm
                assert t["split_aux"][offs:offs + len(sa)].tolist() == sa
                offs += len(sm)
        assert off == len(t["comb_to_count"]) and offs == len(t["split_main"])


def test_csc_and_hole_rule_match_reference(golden):
    for c in golden["csc"]:
        n, s, d = sg.remove_holes(c["n_decl"], np.array(c["src"], np.int32), np.array(c["dst"], np.int32))
        colptr, rowidx = sg.csc_from_edges(n, s, d)
        assert colptr.tolist() == c["colptr"], c["name"]
        assert rowidx.tolist() == c["rowidx"], c["name"]


def test_rmat_invariants():
    s, d = sg.rmat_edges(12, 8, seed=7)
    assert np.all(s < d) and len(s) > 0
    assert len(set(zip(s.tolist(), d.tolist()))) == len(s)   # deduplicated
    assert s.max() < 4096 and d.max() < 4096
    key = s.astype(np.int64) * 4096 + d
    assert np.all(np.diff(key) > 0)                           # sorted
    s2, d2 = sg.rmat_edges(12, 8, seed=7)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)   # deterministic
    s3, _ = sg.rmat_edges(12, 8, seed=8)
    assert not np.array_equal(s, s3[: len(s)])
    # skew: a=.57 concentrates edges; the top 1% of vertices hold > 10% of entries
    deg = np.bincount(np.concatenate([s, d]), minlength=4096)
    top = np.sort(deg)[::-1][:41].sum()
    assert top > 0.1 * deg.sum()


def test_color_stream():
    c = sg.colors_fill(5, 0, 3, 1000, 7)
    assert c.shape == (3, 1000) and c.max() < 7
    assert np.array_equal(c[1:], sg.colors_fill(5, 1, 2, 1000, 7))   # coloring index, not call, keys it
    counts = np.bincount(c.ravel(), minlength=7)
    assert counts.min() > 300                                        # roughly uniform


def test_header_symbols_are_exported():
    with open(os.path.join(ROOT, "include", "pgbsc.h")) as f:
        decl = set(re.findall(r"\b(pg_[a-z_]+)\s*\(", f.read()))
    L = _lib.lib()
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert decl == set(_lib.exported_symbols())


def test_one_hip_runtime_in_process():
    """libpgbsc.so and PyTorch share one HIP runtime (two HSA runtimes in one
    process leave the second without a GPU)."""
    code = ("import re; from subgraph_amd import _lib; _lib.lib(); import torch; "
            "m = open('/proc/self/maps').read(); "
            "print(len(set(re.findall(r'/\\S*libamdhip64\\S*', m))), "
            "len(set(re.findall(r'/\\S*libhsa-runtime64\\S*', m))))")
    r = subprocess.run([os.sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["1", "1"], r.stdout


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def _run_cli(args, cwd):
    return subprocess.run([SC_BIN] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.fixture
def small_graph(tmp_path):
    g = tmp_path / "g.el"
    _write(g, "6\n7\n0 1\n1 2\n2 0\n2 3\n3 4\n4 5\n5 3\n")
    return tmp_path, g


def test_cli_rejects_bad_template(small_graph):
    tmp, g = small_graph
    bad = tmp / "bad.fascia"
    _write(bad, "3\n3\n0 1\n1 2\n2 0\n")            # a triangle is not a tree
    r = _run_cli([str(g), str(bad), "1", "1", "0", "0"], tmp)
    assert r.returncode != 0 and "tree" in r.stderr
    r = _run_cli([str(g), str(tmp / "missing.fascia"), "1", "1", "0", "0"], tmp)
    assert r.returncode != 0 and "cannot open" in r.stderr


def test_cli_missing_graph_is_an_error(small_graph):
    tmp, _ = small_graph
    t = tmp / "u3.fascia"
    _write(t, "3\n2\n0 1\n1 2\n")
    for load_binary in ("0", "1"):
        r = _run_cli([str(tmp / "nope.data"), str(t), "1", "1", load_binary, "0"], tmp)
        assert r.returncode != 0
        assert not (tmp / "nope.data").exists()     # never created (the reference creates it)


def test_cli_binary_graph_round_trip(small_graph):
    """write_binary=1 writes graph.data in the reference layout
    (CSCGraph.hpp:428-436); load_binary=1 reads it back."""
    tmp, g = small_graph
    bad = tmp / "bad.fascia"
    _write(bad, "3\n3\n0 1\n1 2\n2 0\n")            # stop before counting (no GPU needed)
    _run_cli([str(g), str(bad), "1", "1", "0", "1"], tmp)
    data = (tmp / "graph.data").read_bytes()
    ne, n = struct.unpack("<ii", data[:8])
    assert (ne, n) == (7, 6)
    deg = np.frombuffer(data, np.int32, n, 8)
    colptr = np.frombuffer(data, np.int32, n + 1, 8 + 4 * n)
    nnz = colptr[-1]
    rowidx = np.frombuffer(data, np.int32, nnz, 8 + 4 * (2 * n + 1))
    vals = np.frombuffer(data, np.float32, nnz, 8 + 4 * (2 * n + 1 + nnz))
    want_colptr, want_rowidx = sg.csc_from_edges(6, np.array([0, 1, 2, 2, 3, 4, 5], np.int32),
                                                 np.array([1, 2, 0, 3, 4, 5, 3], np.int32))
    assert colptr.tolist() == want_colptr.tolist() and rowidx.tolist() == want_rowidx.tolist()
    assert deg.tolist() == np.diff(want_colptr).tolist() and np.all(vals == 1.0)
    assert len(data) == 8 + 4 * (2 * n + 1 + 2 * nnz)
    r = _run_cli([str(tmp / "graph.data"), str(bad), "1", "1", "1", "0"], tmp)
    assert "tree" in r.stderr                       # the graph loaded, the template failed
    (tmp / "trunc.data").write_bytes(data[:40])
    r = _run_cli([str(tmp / "trunc.data"), str(bad), "1", "1", "1", "0"], tmp)
    assert r.returncode != 0 and "truncated" in r.stderr


def test_python_errors_without_gpu_are_clean():
    with pytest.raises(sg.PGError):
        sg.template_tables(3, [(0, 1), (1, 0)])     # not a tree
    with pytest.raises(sg.PGError):
        sg.csc_from_edges(3, np.array([0, 5], np.int32), np.array([1, 1], np.int32))
