"""Helpers shared by the parity tests: run the oracle and the HIP path on the
same graph and compare per-phase communities (bit-exact) and modularity."""
from __future__ import annotations

import numpy as np

import oracle

Q_TOL = 1e-12   # |dQ| bound; the canonical arithmetic makes it 0 in practice


def run_oracle(off, col, w, nb, tol=1e-6, max_loops=500, nthreads=0, tol_phase=None):
    o = oracle.Oracle(off, col, w, nb, nthreads)
    r = o.run(max_loops, tol, tol_phase, record=True)
    r["clustering"] = o.clustering()
    o.close()
    return r


def run_gpu(off, col, w, nb, tol=1e-6, max_loops=500, **knobs):
    import nido_amd as na
    g = na.GraphGPU((off, col, w), nb)
    if knobs:
        g.set_tuning(**knobs)
    st = na.LouvainGPU(max_loops, tol, nb).run(g, record=True)
    phases = []
    for i in range(st.phases):
        p = g.phase(i).as_dict()
        p["iter_q"] = g.iteration_q(i)
        p["comm"] = g.phase_communities(i)
        phases.append(p)
    out = {"phases": phases, "q": st.q, "clustering": g.clustering(), "stats": st.as_dict()}
    g.close()
    return out


def assert_runs_equal(ref, got):
    assert len(got["phases"]) == len(ref["phases"]), (len(got["phases"]), len(ref["phases"]))
    for i, (a, b) in enumerate(zip(ref["phases"], got["phases"])):
        assert a["nv"] == b["nv"], f"phase {i}: nv {a['nv']} vs {b['nv']}"
        assert a["ne"] == b["ne"], f"phase {i}: ne {a['ne']} vs {b['ne']}"
        assert a["loops"] == b["loops"], f"phase {i}: loops {a['loops']} vs {b['loops']}"
        assert abs(a["q0"] - b["q0"]) <= Q_TOL and abs(a["q"] - b["q"]) <= Q_TOL
        assert np.array_equal(a["iter_q"], b["iter_q"]), f"phase {i}: iteration Q differs"
        assert np.array_equal(a["comm"], b["comm"]), (
            f"phase {i}: {int(np.sum(a['comm'] != b['comm']))} vertices differ")
    assert np.array_equal(ref["clustering"], got["clustering"])
    assert abs(ref["q"] - got["q"]) <= Q_TOL


def sorted_rows(off, col, w):
    """CSR with every row sorted by neighbour id (coarse rows from the GPU
    come out in hash order)."""
    col2 = col.copy()
    w2 = w.copy()
    for v in range(len(off) - 1):
        s, e = off[v], off[v + 1]
        o = np.argsort(col[s:e], kind="stable")
        col2[s:e] = col[s:e][o]
        w2[s:e] = w[s:e][o]
    return off, col2, w2
