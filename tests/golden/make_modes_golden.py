"""Writes tests/golden/modes.json from the COMPILED REFERENCE (oracle/_ref/pg_ref).

Fixtures for the reference surface around the pruned SpMM path:
  counts    per-coloring raw counts of colorCounting in the other modes of
            CountMat::initialization: isPruned=1 useSPMM=0
            (countNonBottomePruned, CountMat.cpp:325-554) and isPruned=0
            (countNonBottomeOriginal, CountMat.cpp:851-977), colorings counted
            one after another by one CountMat (the unpruned mode's _bufVec
            carries over between them)
  compute   CountMat::compute(t, false) end to end (CountMat.cpp:79-266):
            the reference's own "Final raw count is", "Prob is" and "Final
            count is" lines and its return value, with and without
            automorphisms, in each mode
  autom     CountMat::automorphismNum (CountMat.cpp:1498-1572) of templates
            up to 10 vertices (the reference enumerates all k! permutations)
  serialize the reference's CSCGraph::serialize bytes (size and sha256) of
            edge lists with duplicates, self loops and holes

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_modes_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
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
    "u3-1": (3, [(0, 1), (1, 2)]),
    "u5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    "star5": (5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
    "tree7": (7, [(0, 1), (1, 2), (2, 3), (1, 4), (4, 5), (4, 6)]),
    "spider7": (7, [(0, 1), (1, 2), (0, 3), (3, 4), (0, 5), (5, 6)]),
    "caterpillar8": (8, [(0, 1), (1, 2), (2, 3), (1, 4), (2, 5), (3, 6), (3, 7)]),
    "tree10": (10, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5), (5, 6), (5, 7), (7, 8), (7, 9)]),
}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "modes.json")


def write_el(path, n, src, dst):
    with open(path, "w") as f:
        f.write(f"{n}\n{len(src)}\n")
        f.write("".join(f"{a} {b}\n" for a, b in zip(src, dst)))


def write_t(path, tv, te):
    with open(path, "w") as f:
        f.write(f"{tv}\n{len(te)}\n" + "".join(f"{a} {b}\n" for a, b in te))


def ref(*args, threads=1):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    return subprocess.run([oracle.REF_BIN, *map(str, args)], check=True, capture_output=True, text=True,
                          env=env, timeout=900).stdout


def parse_counts(out):
    return [float(line.split()[4]).hex() for line in out.splitlines() if line.startswith("PGREF coloring")]


def main():
    if not oracle.have_ref():
        sys.exit("oracle/_ref/pg_ref missing: build it with make -C oracle ref (needs the reference sources)")
    rng = np.random.default_rng(20261203)
    out = {"source": "dsc-nec/subgraph compiled -DNEC, driven by oracle/ref_driver.cpp",
           "templates": {k: {"nv": v[0], "edges": v[1]} for k, v in TEMPLATES.items()},
           "counts": [], "compute": [], "autom": [], "serialize": []}
    s10, d10 = sg.rmat_edges(10, 8, seed=1)
    n_rand = 60
    sr = rng.integers(0, n_rand, 240).astype(np.int32)
    dr = rng.integers(0, n_rand, 240).astype(np.int32)
    sr[0], dr[0] = 0, n_rand - 1
    graphs = {
        "rmat_s10_ef8": {"kind": "rmat", "scale": 10, "edge_factor": 8, "seed": 1, "m": int(len(s10)),
                         "checksum": edge_checksum(s10, d10), "n": 1 << 10, "src": s10, "dst": d10},
        "random60": {"kind": "edges", "n": n_rand, "src": sr, "dst": dr},
        "holes": {"kind": "edges", "n": 12, "src": np.array([0, 2, 5, 5, 7, 9, 2, 2], np.int32),
                  "dst": np.array([2, 5, 7, 2, 0, 5, 9, 9], np.int32)},
    }
    with tempfile.TemporaryDirectory() as tmp:
        gp, tp = os.path.join(tmp, "g.el"), os.path.join(tmp, "t.fascia")

        # ---- counts in the other modes
        for gname, tnames in (("rmat_s10_ef8", ["u3-1", "u5", "star5", "tree7"]),
                              ("random60", ["u3-1", "u5", "star5", "tree7", "tree10"]),
                              ("holes", ["u3-1", "u5"])):
            g = graphs[gname]
            write_el(gp, g["n"], g["src"].tolist(), g["dst"].tolist())
            for tname in tnames:
                tv, te = TEMPLATES[tname]
                write_t(tp, tv, te)
                for pruned, spmm in ((1, 0), (0, 1), (1, 1)):
                    # splitCSC needs n >= 4 * threads (CSCGraph.hpp:199, SIGFPE otherwise)
                    thr = (1, 3) if pruned == 0 else (2,)
                    for threads in (t for t in thr if 4 * t <= len(np.unique(np.r_[g["src"], g["dst"]]))):
                        raws = parse_counts(ref(gp, tp, "-", 3, threads, 101, pruned, spmm, threads=threads))
                        out["counts"].append({"graph": gname, "template": tname, "isPruned": pruned,
                                              "useSPMM": spmm, "threads": threads,
                                              "colors": {"kind": "stream", "seed": 101, "first": 0, "count": 3},
                                              "raw": raws})
                print(gname, tname, "counts done", flush=True)

        # ---- compute end to end
        g = graphs["rmat_s10_ef8"]
        write_el(gp, g["n"], g["src"].tolist(), g["dst"].tolist())
        for tname in ("u5", "star5", "tree7", "spider7"):
            tv, te = TEMPLATES[tname]
            write_t(tp, tv, te)
            for pruned, spmm, autom in ((1, 1, 0), (1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1)):
                txt = ref("compute", gp, tp, 3, 2, 101, pruned, spmm, autom, threads=2)
                rec = {"graph": "rmat_s10_ef8", "template": tname, "isPruned": pruned, "useSPMM": spmm,
                       "autom": autom, "iterations": 3, "seed": 101}
                for line in txt.splitlines():
                    if line.startswith("Final raw count is"):
                        rec["final_raw_line"] = line
                    elif line.startswith("Prob is"):
                        rec["prob_line"] = line
                    elif line.startswith("Final count is"):
                        rec["final_line"] = line
                    elif line.startswith("PGREF final"):
                        rec["final"] = float(line.split()[2]).hex()
                out["compute"].append(rec)
            print(tname, "compute done", flush=True)

        # ---- automorphisms (the reference tries all k! permutations)
        with open(os.path.join(os.path.dirname(OUT), "golden.json")) as f:
            plans = json.load(f)["plans"]
        seen = set()
        cands = [(n, v[0], v[1]) for n, v in TEMPLATES.items()] + [(p["name"], p["nv"], p["edges"]) for p in plans]
        for name, tv, te in cands:
            if tv > 10 or name in seen:
                continue
            seen.add(name)
            write_t(tp, tv, te)
            a = int(ref("autom", tp).split()[2])
            out["autom"].append({"name": name, "nv": tv, "edges": [list(map(int, e)) for e in te], "autom": a})
        print("autom done", len(out["autom"]), flush=True)

        # ---- graph.data bytes of the reference's serialize
        for gname in ("random60", "holes"):
            g = graphs[gname]
            write_el(gp, g["n"], g["src"].tolist(), g["dst"].tolist())
            dp = os.path.join(tmp, "ref.data")
            ref("serialize", gp, dp)
            data = open(dp, "rb").read()
            out["serialize"].append({"graph": gname, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()})

    out["graphs"] = {k: gdesc_static(v) for k, v in graphs.items()}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


def gdesc_static(g):
    if g["kind"] == "rmat":
        return {k: g[k] for k in ("kind", "scale", "edge_factor", "seed", "m", "checksum", "n")}
    return {"kind": "edges", "n": g["n"], "src": g["src"].tolist(), "dst": g["dst"].tolist()}


if __name__ == "__main__":
    main()
