"""Writes tests/golden/golden_s22.json from the COMPILED REFERENCE (oracle/_ref/pg_ref).

Per-coloring raw counts of CountMat::colorCounting (CountMat.cpp:268-322) on
R-MAT scale 22, edge factor 16 (a=.57 b=.19 c=.19, seed 1, self loops and
duplicates removed, symmetrized by CSCGraph::createFromEdgeListFile), with
colorings from the library's color stream (seed 101, colorings 0..N-1):

    u5      64 colorings  (one full 64-coloring pass)
    tree7    8 colorings
    tree10   4 colorings
    tree12   2 colorings  (about 80 s each on 8 cores, 34 GB resident)

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_golden_s22.py [--threads 8] [--only tree12,...]

The graph is recorded by its generator parameters plus an edge checksum
(subgraph_amd.rmat_edges regenerates it); no reference source is stored.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from golden_util import edge_checksum  # noqa: E402
import subgraph_amd as sg  # noqa: E402

TEMPLATES = {
    "u5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    "tree7": (7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]),
    "tree10": (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]),
    "tree12": (12, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6), (6, 7), (7, 8), (6, 9), (9, 10), (10, 11)]),
}
NCOL = {"u5": 64, "tree7": 8, "tree10": 4, "tree12": 2}
SCALE, EF, GSEED, CSEED = 22, 16, 1, 101
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_s22.json")


def write_el(path, n, src, dst):
    with open(path, "w") as f:
        f.write(f"{n}\n{len(src)}\n")
        step = 1 << 20
        for i in range(0, len(src), step):
            f.write("".join(f"{a} {b}\n" for a, b in zip(src[i:i + step].tolist(), dst[i:i + step].tolist())))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", default=",".join(TEMPLATES))
    args = ap.parse_args()
    if not oracle.have_ref():
        sys.exit("oracle/_ref/pg_ref missing: build it with make -C oracle ref (needs the reference sources)")
    out = {"source": "dsc-nec/subgraph compiled -DNEC, driven by oracle/ref_driver.cpp", "counts": []}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    s, d = sg.rmat_edges(SCALE, EF, seed=GSEED)
    n = 1 << SCALE
    nn, _, _ = sg.remove_holes(n, s, d)
    graph = {"kind": "rmat", "scale": SCALE, "edge_factor": EF, "seed": GSEED, "m": int(len(s)),
             "checksum": edge_checksum(s, d), "n_after_holes": int(nn)}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        gp, tp = os.path.join(tmp, "g.el"), os.path.join(tmp, "t.fascia")
        t0 = time.time()
        write_el(gp, n, s, d)
        print(f"wrote edge list ({len(s)} edges) in {time.time() - t0:.1f} s", flush=True)
        del s, d
        for tname in args.only.split(","):
            tv, te = TEMPLATES[tname]
            with open(tp, "w") as f:
                f.write(f"{tv}\n{len(te)}\n" + "".join(f"{a} {b}\n" for a, b in te))
            t0 = time.time()
            raws, secs = oracle.run_ref(gp, tp, "-", NCOL[tname], threads=args.threads, seed=CSEED,
                                        omp_threads=args.threads)
            if len(raws) != NCOL[tname]:
                sys.exit(f"{tname}: reference printed {len(raws)} counts, expected {NCOL[tname]}")
            name = f"rmat_s{SCALE}_ef{EF}_{tname}"
            out["counts"] = [c for c in out["counts"] if c["name"] != name]
            out["counts"].append({"name": name, "graph": graph, "template": tname,
                                  "template_edges": te, "nv": tv,
                                  "colors": {"kind": "stream", "seed": CSEED, "first": 0, "count": NCOL[tname]},
                                  "raw": [float(r).hex() for r in raws], "ref_seconds": secs,
                                  "ref_threads": args.threads})
            with open(OUT, "w") as f:
                json.dump(out, f, indent=1)
            print(f"{name}: {len(raws)} colorings in {time.time() - t0:.1f} s "
                  f"(mean {np.mean(secs):.2f} s per coloring); first {raws[0]!r}", flush=True)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
