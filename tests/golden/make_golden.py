"""Writes tests/golden/golden.json from the COMPILED REFERENCE (oracle/_ref/pg_ref).

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_golden.py

Every expected value in the fixture comes from the reference's own code
(dsc-nec/subgraph built -DNEC, driven by oracle/ref_driver.cpp): per-coloring
raw counts of CountMat::colorCounting, the CSC of
CSCGraph::createFromEdgeListFile after EdgeList's hole removal, and the
DivideTemplates / IndexSys tables.  Inputs are synthetic: small explicit
graphs, and R-MAT graphs recorded by their generator parameters plus an edge
checksum (subgraph_amd.rmat_edges regenerates them).
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from golden_util import edge_checksum  # noqa: E402
import subgraph_amd as sg  # noqa: E402

TEMPLATES = {
    "u2": (2, [(0, 1)]),
    "u3-1": (3, [(0, 1), (1, 2)]),
    "u5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    "star5": (5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
    "tree7": (7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]),
    "tree10": (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]),
    "tree12": (12, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]),
}


def write_el(path, n, src, dst):
    with open(path, "w") as f:
        f.write(f"{n}\n{len(src)}\n")
        f.write("".join(f"{a} {b}\n" for a, b in zip(src, dst)))


def run_case(tmp, n_decl, src, dst, tname, tv, tedges, colors=None, stream=None, ncol=1, threads=1):
    gp, tp = os.path.join(tmp, "g.el"), os.path.join(tmp, "t.fascia")
    write_el(gp, n_decl, src, dst)
    with open(tp, "w") as f:
        f.write(f"{tv}\n{len(tedges)}\n" + "".join(f"{a} {b}\n" for a, b in tedges))
    if colors is not None:
        cp = os.path.join(tmp, "c.u8")
        np.ascontiguousarray(colors, np.uint8).tofile(cp)
        raws, secs = oracle.run_ref(gp, tp, cp, ncol, threads=threads, omp_threads=1)
    else:
        raws, secs = oracle.run_ref(gp, tp, "-", ncol, threads=threads, seed=stream, omp_threads=threads)
    return [float(r).hex() for r in raws], secs


def main():
    if not oracle.have_ref():
        sys.exit("oracle/_ref/pg_ref missing: build it with make -C oracle ref (needs the reference sources)")
    rng = np.random.default_rng(20260101)
    out = {"source": "dsc-nec/subgraph compiled -DNEC, driven by oracle/ref_driver.cpp",
           "templates": {k: {"nv": v[0], "edges": v[1]} for k, v in TEMPLATES.items()},
           "plans": [], "csc": [], "counts": []}
    with tempfile.TemporaryDirectory() as tmp:
        # ---- decomposition tables: BASELINE templates and random trees
        trees = list(TEMPLATES.items())
        for i in range(24):
            k = int(rng.integers(3, 13))
            e = [(int(rng.integers(0, j)), j) for j in range(1, k)]
            perm = rng.permutation(k)
            e = [(int(perm[a]), int(perm[b])) for a, b in e]
            rng.shuffle(e)
            trees.append((f"rand{i}", (k, e)))
        gp = os.path.join(tmp, "g.el")
        write_el(gp, 4, [0, 1, 2, 2], [1, 2, 0, 3])
        for name, (k, e) in trees:
            tp = os.path.join(tmp, "t.fascia")
            with open(tp, "w") as f:
                f.write(f"{k}\n{len(e)}\n" + "".join(f"{a} {b}\n" for a, b in e))
            d = oracle.dump_ref(gp, tp)
            out["plans"].append({"name": name, "nv": k, "edges": e,
                                 "sub": [d["sub"][s] for s in sorted(d["sub"])],
                                 "c2c": [d["c2c"][s] for s in sorted(d["c2c"])],
                                 "smain": {str(s): v for s, v in d["smain"].items()},
                                 "saux": {str(s): v for s, v in d["saux"].items()}})

        # ---- CSC + hole removal on edge-case lists (declared n, edges)
        csc_cases = [
            ("triangle_tail", 4, [0, 1, 2, 2], [1, 2, 0, 3]),
            ("dups_selfloops", 6, [0, 0, 1, 2, 3, 3, 4, 5, 1], [1, 1, 2, 2, 4, 3, 5, 0, 0]),
            ("holes", 9, [0, 2, 5, 5, 7], [2, 5, 7, 2, 0]),          # max id 7 != 8: renumbered, n -> 4
            ("isolated_tail", 8, [0, 1, 2, 7], [1, 2, 3, 6]),         # max id 7 == n-1: no renumbering
        ]
        for name, n, s, d in csc_cases:
            write_el(gp, n, s, d)
            tp = os.path.join(tmp, "t.fascia")
            with open(tp, "w") as f:
                f.write("3\n2\n0 1\n1 2\n")
            dd = oracle.dump_ref(gp, tp)
            out["csc"].append({"name": name, "n_decl": n, "src": s, "dst": d,
                               "colptr": dd["colptr"], "rowidx": dd["rowidx"]})

        # ---- counts: tiny hand-checked cases
        for tname, colors in (("u3-1", [0, 1, 2, 0]), ("u2", [0, 1, 0, 1])):
            tv, te = TEMPLATES[tname]
            raws, _ = run_case(tmp, 4, [0, 1, 2, 2], [1, 2, 0, 3], tname, tv, te, colors=np.array([colors]))
            out["counts"].append({"name": f"tiny_{tname}", "graph": {"kind": "edges", "n": 4,
                                  "src": [0, 1, 2, 2], "dst": [1, 2, 0, 3]}, "template": tname,
                                  "colors": {"kind": "explicit", "data": [colors]}, "raw": raws})

        # ---- edge-case graphs with several templates, explicit colors
        for name, n, s, d in csc_cases[1:]:
            nn, ss, dd = sg.remove_holes(n, s, d)
            for tname in ("u3-1", "u5", "star5"):
                tv, te = TEMPLATES[tname]
                if nn < 4:
                    continue
                cols = rng.integers(0, tv, size=(2, nn)).astype(np.uint8)
                raws, _ = run_case(tmp, n, s, d, tname, tv, te, colors=cols, ncol=2)
                out["counts"].append({"name": f"{name}_{tname}", "graph": {"kind": "edges", "n": n, "src": s,
                                      "dst": d}, "template": tname,
                                      "colors": {"kind": "explicit", "data": cols.tolist()}, "raw": raws})

        # ---- random small multigraphs (duplicates and self loops allowed)
        for i in range(10):
            n = int(rng.integers(20, 200))
            m = int(rng.integers(n, 6 * n))
            s = rng.integers(0, n, m).astype(np.int32)
            d = rng.integers(0, n, m).astype(np.int32)
            s[0], d[0] = 0, n - 1
            tname = ["u3-1", "u5", "star5", "tree7", "tree10"][i % 5]
            tv, te = TEMPLATES[tname]
            cols = rng.integers(0, tv, size=(3, n)).astype(np.uint8)
            raws, _ = run_case(tmp, n, s.tolist(), d.tolist(), tname, tv, te, colors=cols, ncol=3)
            out["counts"].append({"name": f"random{i}_{tname}", "graph": {"kind": "edges", "n": n,
                                  "src": s.tolist(), "dst": d.tolist()}, "template": tname,
                                  "colors": {"kind": "explicit", "data": cols.tolist()}, "raw": raws})

        # ---- R-MAT graphs, stream colors (pgbsc.h color stream, seed 101)
        rmat_cases = [(10, 8, 1, ["u3-1", "u5", "star5", "tree7", "tree10", "tree12"], 4),
                      (12, 16, 2, ["tree7", "tree10", "tree12"], 2),
                      (16, 16, 1, ["u3-1", "tree7"], 2),
                      (20, 8, 1, ["tree7"], 2)]
        for scale, ef, gseed, tnames, ncol in rmat_cases:
            s, d = sg.rmat_edges(scale, ef, seed=gseed)
            n = 1 << scale
            nn, ss, dd = sg.remove_holes(n, s, d)
            for tname in tnames:
                tv, te = TEMPLATES[tname]
                raws, secs = run_case(tmp, n, s.tolist(), d.tolist(), tname, tv, te, stream=101, ncol=ncol,
                                      threads=8)
                out["counts"].append({"name": f"rmat_s{scale}_ef{ef}_{tname}",
                                      "graph": {"kind": "rmat", "scale": scale, "edge_factor": ef,
                                                "seed": gseed, "m": int(len(s)),
                                                "checksum": edge_checksum(s, d), "n_after_holes": int(nn)},
                                      "template": tname,
                                      "colors": {"kind": "stream", "seed": 101, "first": 0, "count": ncol},
                                      "raw": raws, "ref_seconds": secs})
                print(f"s{scale}/ef{ef} {tname}: {[float.fromhex(r) for r in raws]}", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
