"""Writes tests/golden/estimate.json from the COMPILED REFERENCE (oracle/_ref/pg_ref).

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_estimate_golden.py

The fixture holds the reference's own cost-model figures
(CountMat::estimatePeakMemUsage, estimateFlops{PGBSC,Fascia,PrunedFascia},
estimateMemComm{PGBSC,Fascia,PrunedFascia}, estimateTemplate and the
degDistri.data file of degreeDistribution, CountMat.cpp:1067-1496) for
synthetic graphs and template trees.  The models read only n, nnz, the
template's subtemplate chain and the degree list, so each case records the
graph by its parameters.
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


def main():
    if not oracle.have_ref():
        sys.exit("oracle/_ref/pg_ref missing: build it with make -C oracle ref (needs the reference sources)")
    rng = np.random.default_rng(20261202)
    trees = list(TEMPLATES.items())
    for i in range(6):
        k = int(rng.integers(3, 13))
        e = [(int(rng.integers(0, j)), j) for j in range(1, k)]
        perm = rng.permutation(k)
        trees.append((f"rand{i}", (k, [(int(perm[a]), int(perm[b])) for a, b in e])))

    # graphs: a 4-vertex one, an R-MAT graph, and a large sparse graph whose
    # last vertex is touched (no hole renumbering), so that the printed
    # peak-memory line carries enough digits
    s10, d10 = sg.rmat_edges(10, 8, seed=1)
    n_big = 1 << 21
    m_big = 3000
    sb = rng.integers(0, n_big, m_big).astype(np.int32)
    db = rng.integers(0, n_big, m_big).astype(np.int32)
    sb[0], db[0] = 0, n_big - 1
    graphs = [
        ("tiny", {"kind": "edges", "n": 4, "src": [0, 1, 2, 2], "dst": [1, 2, 0, 3]}, 4, [0, 1, 2, 2], [1, 2, 0, 3]),
        ("rmat_s10_ef8", {"kind": "rmat", "scale": 10, "edge_factor": 8, "seed": 1, "m": int(len(s10)),
                          "checksum": edge_checksum(s10, d10)}, 1 << 10, s10.tolist(), d10.tolist()),
        ("sparse_2e21", {"kind": "edges", "n": n_big, "src": sb.tolist(), "dst": db.tolist()}, n_big,
         sb.tolist(), db.tolist()),
    ]
    out = {"source": "dsc-nec/subgraph compiled -DNEC, driven by oracle/ref_driver.cpp (estimate modes)",
           "templates": {k: {"nv": v[0], "edges": v[1]} for k, v in trees}, "cases": []}
    with tempfile.TemporaryDirectory() as tmp:
        gp, tp = os.path.join(tmp, "g.el"), os.path.join(tmp, "t.fascia")
        for gname, gspec, n, src, dst in graphs:
            write_el(gp, n, src, dst)
            deg_written = False
            for tname, (k, e) in trees:
                with open(tp, "w") as f:
                    f.write(f"{k}\n{len(e)}\n" + "".join(f"{a} {b}\n" for a, b in e))
                r = oracle.estimate_ref(gp, tp, tmp)
                r.update(oracle.estimate_template_ref(gp, tp))
                case = {"graph": gname, "template": tname, "n": r.pop("n"), "nnz": r.pop("nnz"),
                        "max_deg": r.pop("max_deg")}
                dist = r.pop("deg_distribution")
                if not deg_written:
                    case["deg_distribution"] = dist
                    deg_written = True
                case["figures"] = r
                out["cases"].append(case)
                print(gname, tname, r["pgbsc_gbytes"], r["peak_mem_gb_printed"], flush=True)
            out.setdefault("graphs", {})[gname] = gspec
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "estimate.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
