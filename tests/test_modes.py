"""The reference surface around the pruned path, on the CPU, against fixtures
of the compiled reference (tests/golden/modes.json): its other counting
modes, automorphismNum, the tail of CountMat::compute, and graph.data bytes.
"""
import hashlib
import itertools
import json
import math
import os
import subprocess

import numpy as np
import pytest

import subgraph_amd as sg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SC_BIN = os.path.join(ROOT, "subgraph_amd", "sc-bin")
MODES = os.path.join(ROOT, "tests", "golden", "modes.json")


@pytest.fixture(scope="module")
def modes():
    with open(MODES) as f:
        return json.load(f)


def test_spmv_mode_is_the_spmm_path(modes):
    """useSPMM = 0 (countNonBottomePruned) gives the SpMM path's counts bit
    for bit in the reference itself: spmvNaiveSplit adds 1.0f * x in the same
    row order as spmmSplit (CSCGraph.hpp:220-310).  This is why the library
    runs one set of kernels for both flags."""
    by = {}
    for c in modes["counts"]:
        by.setdefault((c["graph"], c["template"]), {})[(c["isPruned"], c["useSPMM"])] = c["raw"]
    checked = 0
    for key, v in by.items():
        if (1, 0) in v and (1, 1) in v:
            assert v[(1, 0)] == v[(1, 1)], key
            checked += 1
    assert checked >= 9


def test_unpruned_mode_is_thread_independent(modes):
    """The unpruned mode's counts (an accumulator that is never cleared) do
    not depend on the reference's thread count: each vertex's operations
    follow one fixed sequence, which the device replays per vertex."""
    by = {}
    for c in modes["counts"]:
        if c["isPruned"] == 0:
            by.setdefault((c["graph"], c["template"]), {})[c["threads"]] = c["raw"]
    assert any(len(v) == 2 for v in by.values())
    for key, v in by.items():
        if len(v) == 2:
            assert v[1] == v[3], key


def _literal_autom(nv, edges):
    """calcAutomorphismRecursive + calcAutomorphismZero (CountMat.cpp:
    1510-1572) over all nv! permutations: s accepted when, for every i,
    deg(i) = deg(s(i)) and every neighbour u of i is s(w) for a neighbour w
    of s(i)."""
    nbr = [[] for _ in range(nv)]
    for u, v in edges:
        nbr[u].append(v)
        nbr[v].append(u)
    count = 0
    for s in itertools.permutations(range(nv)):
        ok = all(len(nbr[i]) == len(nbr[s[i]]) and all(any(u == s[w] for w in nbr[s[i]]) for u in nbr[i])
                 for i in range(nv))
        count += ok
    return count


def test_automorphisms_match_reference(modes):
    assert len(modes["autom"]) >= 20
    for a in modes["autom"]:
        assert sg.automorphisms(a["nv"], a["edges"]) == a["autom"], a["name"]


def test_automorphisms_literal_rule_random_trees():
    """The pruned search counts the reference's predicate over all k!
    permutations, also where that predicate is not the automorphism group."""
    rng = np.random.default_rng(3)
    for _ in range(120):
        k = int(rng.integers(2, 8))
        e = [(int(rng.integers(0, j)), j) for j in range(1, k)]
        perm = rng.permutation(k)
        e = [(int(perm[a]), int(perm[b])) for a, b in e]
        assert sg.automorphisms(k, e) == _literal_autom(k, e), e


def test_automorphisms_known_values():
    assert sg.automorphisms(5, [(0, 1), (0, 2), (0, 3), (0, 4)]) == 24       # star: 4!
    assert sg.automorphisms(5, [(0, 1), (1, 2), (2, 3), (3, 4)]) == 2        # path
    assert sg.automorphisms(9, [(0, j) for j in range(1, 9)]) == math.factorial(8)


def _golden_raws(golden, name, n):
    for c in golden["counts"]:
        if c["name"] == name:
            return [float.fromhex(r) for r in c["raw"]][:n]
    raise KeyError(name)


def test_compute_tail_matches_reference(modes, golden):
    """CountMat::compute's tail (CountMat.cpp:247-259) on the reference's own
    per-coloring counts: mean / (k!/k^k * aut), rounded, equals what the
    reference's compute() returned end to end, and the printed probability
    line is the reference's."""
    tmpl = modes["templates"]
    checked = 0
    for rec in modes["compute"]:
        if rec["isPruned"] != 1 or rec["template"] not in ("u5", "star5", "tree7"):
            continue
        k = tmpl[rec["template"]]["nv"]
        raws = _golden_raws(golden, f"rmat_s10_ef8_{rec['template']}", rec["iterations"])
        mean = sum(raws) / len(raws)
        prob = math.factorial(k) / float(k) ** k
        aut = sg.automorphisms(k, tmpl[rec["template"]]["edges"]) if rec["autom"] else 1
        assert f"Prob is {prob:f}" == rec["prob_line"]
        assert f"Final raw count is {mean:e}" == rec["final_raw_line"]
        assert math.floor(mean / (prob * aut) + 0.5) == float.fromhex(rec["final"]), rec
        checked += 1
    assert checked >= 6


def _write_el(path, g):
    with open(path, "w") as f:
        f.write(f"{g['n']}\n{len(g['src'])}\n" + "".join(f"{a} {b}\n" for a, b in zip(g["src"], g["dst"])))


def test_graph_data_bytes_match_reference_serialize(modes, tmp_path):
    """sc-bin's write_binary file is byte for byte the reference's
    CSCGraph::serialize (CSCGraph.hpp:426-434), holes and duplicates
    included.  A bad template stops the CLI after the write (no GPU)."""
    bad = tmp_path / "bad.fascia"
    bad.write_text("3\n3\n0 1\n1 2\n2 0\n")
    for rec in modes["serialize"]:
        g = modes["graphs"][rec["graph"]]
        gp = tmp_path / f"{rec['graph']}.el"
        _write_el(gp, g)
        r = subprocess.run([SC_BIN, str(gp), str(bad), "1", "1", "0", "1"], cwd=tmp_path, capture_output=True,
                           text=True, timeout=120)
        assert "tree" in r.stderr, r.stderr
        data = (tmp_path / "graph.data").read_bytes()
        assert len(data) == rec["bytes"] and hashlib.sha256(data).hexdigest() == rec["sha256"], rec["graph"]
