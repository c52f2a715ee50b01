"""Writes the committed fixtures of this directory.

sbm1024_seed1.bin
    The plain small workload: planted partition, 1,024 vertices in 16 blocks
    of 64, p_in=0.30, p_out=0.005, unit weights, seed 1, in the reference's
    binary format (graph.cpp:322-495: int64 nv, int64 ne, (nv+1) int64
    offsets, ne x {int64 tail, double weight}).
sbm1024_seed1_expected.json
    The oracle's run on it with -b 2 -t 1e-6 -i 500: per-phase vertex/arc
    counts, iteration counts, modularity after every iteration, per-phase
    communities and the final original-vertex clustering.  A regression
    fixture: the oracle itself is pinned by the closed-form, networkx,
    brute-force and independent-restatement tests in tests/.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import nido_amd as na  # noqa: E402


def main():
    g = na.gen_sbm(1024, 64, 0.30, 0.005, False, 1)
    path = os.path.join(HERE, "sbm1024_seed1.bin")
    g.write(path)
    off, col, w = na.Graph.read(path).arrays()
    o = oracle.Oracle(off, col, w, 2, 1)
    r = o.run(500, 1e-6, record=True)
    out = {
        "graph": "sbm1024_seed1.bin",
        "params": {"nbatches": 2, "tol": 1e-6, "max_loops": 500},
        "nv": int(len(off) - 1), "ne": int(len(col)),
        "phases": [{
            "nv": p["nv"], "ne": p["ne"], "loops": p["loops"], "q0": p["q0"], "q": p["q"],
            "iter_q": [float(x) for x in p["iter_q"]],
            "comm": [int(x) for x in p["comm"]],
        } for p in r["phases"]],
        "q": r["q"],
        "clustering": [int(x) for x in o.clustering()],
    }
    with open(os.path.join(HERE, "sbm1024_seed1_expected.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, "nv", out["nv"], "ne", out["ne"], "phases", len(out["phases"]), "Q", out["q"])


if __name__ == "__main__":
    main()
