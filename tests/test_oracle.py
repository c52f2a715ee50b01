"""Pins the CPU oracle (oracle/louvain_oracle.c) without a GPU:
- Q64.64 conversions against exact rational arithmetic;
- closed-form tiny graphs whose outcome is derived by hand from the
  reference's rules (louvain_gpu.cpp:53-118, graph_cuda.cu:1481-1537);
- modularity against networkx and a brute-force rational evaluation;
- whole runs against an independent pure-Python restatement (tests/pyref.py);
- -b batch boundaries (graph_gpu.cpp:303-337) including the fix-up branches;
- coarsening invariants (graph_gpu.cpp:1382-1545).
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle
import pyref

ONE = 1 << 64


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR, rows sorted by neighbour; edges are (u, v) with u != v
    or u == v for a self-loop (stored once, like coarse self-loops)."""
    adj = [dict() for _ in range(n)]
    for i, (u, v) in enumerate(edges):
        w = 1.0 if weights is None else weights[i]
        adj[u][v] = adj[u].get(v, 0.0) + w
        if u != v:
            adj[v][u] = adj[v].get(u, 0.0) + w
    off, col, wt = [0], [], []
    for u in range(n):
        for v in sorted(adj[u]):
            col.append(v)
            wt.append(adj[u][v])
        off.append(len(col))
    return np.array(off, np.int64), np.array(col, np.int64), np.array(wt, np.float64)


# ---------------------------------------------------------------------------
# fixed point
# ---------------------------------------------------------------------------
def _rne_fx(w):
    return round(Fraction(w) * ONE)


@pytest.mark.parametrize("seed", range(4))
def test_fx_from_double_matches_exact_rounding(seed):
    rng = random.Random(seed)
    vals = [0.0, 1.0, -1.0, 0.5, 2.0 ** -64, 2.0 ** -65, 3 * 2.0 ** -66, 2.0 ** -70, 5e-324,
            2.0 ** 61, -(2.0 ** 60) + 1.0, 1.0 - 2.0 ** -53]
    for _ in range(400):
        e = rng.randint(-80, 61)
        vals.append(rng.choice([-1, 1]) * rng.random() * 2.0 ** e)
    for v in vals:
        assert oracle.fx_from_double(v) == _rne_fx(v), v


def test_fx_ties_round_to_even():
    # 2^-65 is exactly half a unit: rounds to 0 (even); 3*2^-65 -> 2 units
    assert oracle.fx_from_double(2.0 ** -65) == 0
    assert oracle.fx_from_double(3 * 2.0 ** -65) == 2
    assert oracle.fx_from_double(5 * 2.0 ** -65) == 2
    assert oracle.fx_from_double(-(3 * 2.0 ** -65)) == -2


@pytest.mark.parametrize("seed", range(4))
def test_fx_to_double_matches_exact_rounding(seed):
    rng = random.Random(100 + seed)
    xs = [0, 1, -1, ONE, -ONE, (1 << 127) - 1, -(1 << 127), (1 << 53) + 1, ((1 << 54) + 1) << 10,
          ((1 << 54) + 2) << 10, ((1 << 54) + 3) << 10]
    for _ in range(500):
        xs.append(rng.randint(-(1 << 126), 1 << 126) >> rng.randint(0, 120))
    for x in xs:
        assert oracle.fx_to_double(x) == float(Fraction(x, ONE)), x


def test_fx_grid_weights_are_exact():
    rng = random.Random(7)
    for _ in range(1000):
        w = (2 ** 53 - rng.getrandbits(53)) * 2.0 ** -53      # the generators' (0,1] grid
        x = oracle.fx_from_double(w)
        assert Fraction(x, ONE) == Fraction(w)
        assert oracle.fx_to_double(x) == w


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def run(off, col, w, nb, **kw):
    o = oracle.Oracle(off, col, w, nb, 2)
    r = o.run(kw.get("max_loops", 500), kw.get("tol", 1e-6), record=True)
    r["clustering"] = o.clustering()
    o.close()
    return r


def test_single_edge_swaps_when_both_vertices_share_a_batch():
    # -b 1 (and -b 2: the cut rule is a strict '>', so both vertices stay in
    # batch 0): both move to each other's community, Q stays -0.5, dQ = 0.
    off, col, w = csr_from_edges(2, [(0, 1)])
    for nb in (1, 2):
        r = run(off, col, w, nb)
        p = r["phases"]
        assert len(p) == 1 and p[0]["loops"] == 1
        assert p[0]["q0"] == -0.5 and list(p[0]["iter_q"]) == [-0.5]
        assert list(p[0]["comm"]) == [1, 0]
        assert list(r["clustering"]) == [0, 1]           # no aggregation happened


def test_single_edge_merges_with_one_vertex_per_batch():
    # -b 3: floor(A/b) = 0, so each vertex is its own batch: vertex 0 joins
    # vertex 1, vertex 1 then sees only its own community.  Q: -0.5 -> 0.
    off, col, w = csr_from_edges(2, [(0, 1)])
    r = run(off, col, w, 3)
    p = r["phases"]
    assert list(p[0]["iter_q"]) == [0.0, 0.0]
    assert list(p[0]["comm"]) == [1, 1]
    assert list(r["clustering"]) == [0, 0]              # K == 1 stops after renumbering


def test_path_rolls_back():
    # a-b-c, -b 1: all three move at once ({a,c}->b's old id..., b->a's),
    # Q drops from -0.375 to -0.5, the iteration is restored and the run ends.
    off, col, w = csr_from_edges(3, [(0, 1), (1, 2)])
    r = run(off, col, w, 1)
    p = r["phases"]
    assert p[0]["q0"] == -0.375
    assert list(p[0]["iter_q"]) == [-0.5]
    assert p[0]["q"] == -0.375
    assert list(p[0]["comm"]) == [0, 1, 2]


def test_two_cliques_with_a_bridge():
    # Two K4 joined by one edge: the optimum splits at the bridge,
    # Q = 2 * (13/26 - (13/26)^2)... computed exactly below.
    edges = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    edges += [(i + 4, j + 4) for i in range(4) for j in range(i + 1, 4)]
    edges += [(3, 4)]
    off, col, w = csr_from_edges(8, edges)
    r = run(off, col, w, 8)
    comm = r["clustering"]
    assert len(set(comm[:4])) == 1 and len(set(comm[4:])) == 1 and comm[0] != comm[4]
    m = 13
    q_exact = 2 * (Fraction(12, 2 * m) - Fraction(13, 2 * m) ** 2)
    assert abs(r["q"] - float(q_exact)) < 1e-15


def _tie_graph():
    """Vertex 0 joined to vertices 5 and 3 (in that arc order) by unit
    edges; 3 and 5 are also joined to 1 and 2, one each, and 4 to 6 (so no
    vertex is isolated).  In the singleton partition (-b 1, every vertex
    decides on the same snapshot) k_0 = k_3 = k_5 = 2, 2m = 10, m = 5, and
    vertex 0 sees two candidates with the same e = 1 and a = 2:
        g_3 = g_5 = (1 - 2 * 2 / 10) / 5 = 3/25,   s' = 0,
    an exact tie (both gains are computed from identical operands, so they
    round identically).  The smallest id wins: C'[0] = 3, not 5 (its first
    arc) and not the larger id."""
    off = np.array([0, 2, 3, 4, 6, 7, 9, 10], np.int64)
    col = np.array([5, 3, 3, 5, 0, 1, 6, 0, 2, 4], np.int64)
    w = np.ones(len(col))
    return off, col, w


def test_min_id_tie_break_closed_form():
    off, col, w = _tie_graph()
    # the CSR is symmetric and k, m are as stated
    pairs = {(v, int(u)) for v in range(7) for u in col[off[v]:off[v + 1]]}
    assert all((u, v) in pairs for v, u in pairs)
    o = oracle.Oracle(off, col, w, 1)
    o.singleton_partition()
    assert list(o.vertex_weights()) == [2, 1, 1, 2, 1, 2, 1] and o.mass == 5.0
    o.louvain_update(0)
    cn = o.new_communities()
    assert cn[0] == 3
    # vertex 4's only candidate is 6 (e = 1, a = 1): g = (1 - 1/10) / 5 > 0;
    # 1 and 2 have one candidate each
    assert cn[4] == 6 and cn[1] == 3 and cn[2] == 5
    # relabelling 3 <-> 5 puts the same tie in the other arc order ([3, 5]):
    # 3 again, so neither the first nor the last arc decides
    perm = np.array([0, 1, 2, 5, 4, 3, 6])
    inv = np.argsort(perm)
    off2 = np.zeros(8, np.int64)
    rows = [col[off[inv[v]]:off[inv[v] + 1]] for v in range(7)]
    off2[1:] = np.cumsum([len(r) for r in rows])
    col2 = np.concatenate([perm[r] for r in rows])
    o2 = oracle.Oracle(off2, col2, np.ones(len(col2)), 1)
    o2.singleton_partition()
    o2.louvain_update(0)
    assert o2.new_communities()[0] == 3
    o.close()
    o2.close()


def test_higher_gain_beats_a_smaller_id():
    # vertex 0 with a weight-2 edge to 5 and a weight-1 edge to 3: k_0 =
    # k_5 = 3, 2m = 12, g_5 = (2 - 9/12) / 6 > g_3 = (1 - 6/12) / 6, so 5
    # wins although 3 is smaller
    off, col, w = _tie_graph()
    w = w.copy()
    w[0] = 2.0          # arc 0 -> 5
    w[7] = 2.0          # arc 5 -> 0
    o = oracle.Oracle(off, col, w, 1)
    o.singleton_partition()
    o.louvain_update(0)
    assert o.new_communities()[0] == 5
    o.close()


def test_empty_and_isolated_graphs():
    for nv in (0, 5):
        off = np.zeros(nv + 1, np.int64)
        r = run(off, np.zeros(0, np.int64), np.zeros(0), 2)
        assert r["q"] == 0.0 and r["phases"][0]["loops"] == 1
        assert list(r["clustering"]) == list(range(nv))


# ---------------------------------------------------------------------------
# modularity
# ---------------------------------------------------------------------------
def brute_q(off, col, w, comm):
    twom = sum(Fraction(x) for x in w)
    internal = Fraction(0)
    a = {}
    for v in range(len(off) - 1):
        for e in range(off[v], off[v + 1]):
            a[comm[v]] = a.get(comm[v], 0) + Fraction(w[e])
            if comm[col[e]] == comm[v]:
                internal += Fraction(w[e])
    return internal / twom - sum((x / twom) ** 2 for x in a.values())


@pytest.mark.parametrize("seed", range(3))
def test_modularity_matches_networkx_and_brute_force(seed):
    import networkx as nx
    rng = np.random.default_rng(seed)
    n = 60
    edges, weights = [], []
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < (0.3 if u // 15 == v // 15 else 0.03):
                edges.append((u, v))
                weights.append(float((2 ** 53 - int(rng.integers(0, 2 ** 53))) * 2.0 ** -53))
    off, col, w = csr_from_edges(n, edges, weights)
    o = oracle.Oracle(off, col, w, 2, 1)
    G = nx.Graph()
    G.add_nodes_from(range(n))
    for (u, v), x in zip(edges, weights):
        G.add_edge(u, v, weight=x)
    for trial in range(5):
        comm = rng.integers(0, 8, size=n) if trial else np.arange(n) // 15
        o.set_communities(comm)
        q = o.compute_modularity()
        sets = [set(np.flatnonzero(comm == c)) for c in np.unique(comm)]
        assert abs(q - nx.community.modularity(G, sets, weight="weight")) < 1e-13
        assert abs(q - float(brute_q(off, col, w, comm))) < 1e-13
    o.close()


def test_modularity_counts_self_loops_as_internal():
    # one self-loop of weight 2 at vertex 0 plus edge 0-1: k0 = 3, k1 = 1, 2m = 4
    off, col, w = np.array([0, 2, 3]), np.array([0, 1, 0]), np.array([2.0, 1.0, 1.0])
    o = oracle.Oracle(off, col, w, 1, 1)
    o.singleton_partition()
    assert o.compute_modularity() == float(Fraction(2, 4) - Fraction(3, 4) ** 2 - Fraction(1, 4) ** 2)
    o.close()


# ---------------------------------------------------------------------------
# whole runs vs the independent restatement
# ---------------------------------------------------------------------------
def _random_graph(n, p, weighted, seed, self_loops=False):
    rng = np.random.default_rng(seed)
    edges, weights = [], []
    for u in range(n):
        for v in range(u + (0 if self_loops else 1), n):
            if rng.random() < p * (4 if u // 16 == v // 16 else 1):
                edges.append((u, v))
                weights.append(float((2 ** 53 - int(rng.integers(0, 2 ** 53))) * 2.0 ** -53) if weighted else 1.0)
    return csr_from_edges(n, edges, weights)


@pytest.mark.parametrize("n,p,weighted,nb,seed", [
    (64, 0.05, False, 2, 1), (96, 0.04, True, 3, 2), (128, 0.03, True, 8, 3), (80, 0.06, False, 1, 4),
    (100, 0.05, True, 5, 5),
])
def test_oracle_matches_independent_restatement(n, p, weighted, nb, seed):
    off, col, w = _random_graph(n, p, weighted, seed)
    got = run(off, col, w, nb)
    ref = pyref.run(off, col, w, nb)
    assert len(got["phases"]) == len(ref["phases"])
    for a, b in zip(got["phases"], ref["phases"]):
        assert a["nv"] == b["nv"] and a["ne"] == b["ne"] and a["loops"] == b["loops"]
        assert a["q0"] == b["q0"] and a["q"] == b["q"]
        assert list(a["iter_q"]) == b["iter_q"]
        assert list(a["comm"]) == b["comm"]
    assert list(got["clustering"]) == ref["clustering"]


def test_oracle_matches_restatement_with_self_loops():
    off, col, w = _random_graph(70, 0.05, True, 9, self_loops=True)
    assert any(col[e] == v for v in range(70) for e in range(off[v], off[v + 1]))
    got = run(off, col, w, 4)
    ref = pyref.run(off, col, w, 4)
    assert got["q"] == ref["q"]
    for a, b in zip(got["phases"], ref["phases"]):
        assert list(a["comm"]) == b["comm"]


def test_golden_fixture_regression(golden_dir):
    import json
    import os
    import nido_amd as na
    off, col, w = na.Graph.read(os.path.join(golden_dir, "sbm1024_seed1.bin")).arrays()
    with open(os.path.join(golden_dir, "sbm1024_seed1_expected.json")) as f:
        exp = json.load(f)
    got = run(off, col, w, 2)
    assert len(got["phases"]) == len(exp["phases"])
    for a, b in zip(got["phases"], exp["phases"]):
        assert list(a["comm"]) == b["comm"] and list(a["iter_q"]) == b["iter_q"]
    assert list(got["clustering"]) == exp["clustering"]


# ---------------------------------------------------------------------------
# batch boundaries
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("degs,nb,expect", [
    ([4, 4, 4, 4], 2, [0, 3, 4]),              # cut when the running count exceeds floor(16/2)=8
    ([1, 1], 2, [0, 2, 2]),                    # '>' keeps both in batch 0, batch 1 empty
    ([1, 1], 3, [0, 1, 2, 2]),                 # ave = 0: every non-empty row cuts
    ([0, 0, 0], 2, [0, 3, 3]),                 # no arcs
    ([5, 1, 1, 1, 1, 1], 4, [0, 1, 4, 6, 6]),  # ave = 2
    ([2, 2, 2, 2, 2, 2, 2, 2], 2, [0, 5, 8]),
])
def test_batch_bounds_table(degs, nb, expect):
    off = np.concatenate([[0], np.cumsum(degs)]).astype(np.int64)
    assert list(oracle.batch_bounds(off, nb)) == expect
    assert pyref.batch_bounds(list(off), nb) == expect


def _raw_cuts(off, nb):
    nv = len(off) - 1
    ave = (off[nv] - off[0]) // nb
    parts, start = [0], 0
    for i in range(1, nv + 1):
        if off[i] - off[start] > ave:
            parts.append(i)
            start = i
        elif i == nv:
            parts.append(nv)
    return parts


@pytest.mark.parametrize("seed", range(20))
def test_batch_cut_list_never_needs_the_fixup(seed):
    # Every batch before the last holds >= floor(A/nb)+1 arcs, so with one
    # device the greedy cut list has at most nb+1 entries and the "more cuts
    # than batches" branch (graph_gpu.cpp:324-331) cannot trigger; only the
    # padding branch (:332-336) does.
    rng = np.random.default_rng(seed)
    degs = rng.integers(0, int(rng.integers(1, 50)), size=int(rng.integers(1, 400)))
    off = np.concatenate([[0], np.cumsum(degs)]).astype(np.int64)
    for nb in (1, 2, 3, 5, 8, 32, 500):
        raw = _raw_cuts(list(off), nb)
        assert len(raw) <= nb + 1
        assert list(oracle.batch_bounds(off, nb)) == raw + [len(degs)] * (nb + 1 - len(raw))


@pytest.mark.parametrize("seed", range(5))
def test_batch_bounds_random_agree(seed):
    rng = np.random.default_rng(seed)
    degs = rng.integers(0, 20, size=int(rng.integers(1, 300)))
    off = np.concatenate([[0], np.cumsum(degs)]).astype(np.int64)
    for nb in (1, 2, 3, 7, 32, 1000):
        b = list(oracle.batch_bounds(off, nb))
        assert b == pyref.batch_bounds(list(off), nb)
        assert b[0] == 0 and b[-1] == len(degs) and all(x <= y for x, y in zip(b, b[1:]))


# ---------------------------------------------------------------------------
# coarsening invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("weighted", [False, True])
def test_aggregation_invariants(weighted):
    off, col, w = _random_graph(200, 0.03, weighted, 21)
    o = oracle.Oracle(off, col, w, 4, 2)
    o.singleton_partition()
    q = o.compute_modularity()
    for _ in range(3):
        q, dq, _qt = o.louvain_iteration(q)
    C = o.communities()
    fine_twom = o.twom_fx
    done = o.aggregation()
    assert not done
    clus = o.clustering()
    used = np.unique(C)
    # dense renumbering, ascending by old id (graph_gpu.cpp:1178-1218)
    assert list(clus) == [int(np.searchsorted(used, C[v])) for v in range(len(C))]
    coff, ccol, cw = o.graph()
    K = len(used)
    assert len(coff) == K + 1
    # symmetric, rows sorted, weights match the exact sums of member arcs
    exp = {}
    for v in range(len(off) - 1):
        for e in range(off[v], off[v + 1]):
            key = (clus[v], clus[col[e]])
            exp[key] = exp.get(key, 0) + pyref.fx(w[e])
    got = {}
    for i in range(K):
        row = ccol[coff[i]:coff[i + 1]]
        assert list(row) == sorted(row)
        for e in range(coff[i], coff[i + 1]):
            got[(i, int(ccol[e]))] = cw[e]
    assert set(got) == set(exp)
    for key, x in exp.items():
        assert got[key] == pyref.fxd(x)
        assert got[(key[1], key[0])] == got[key]
    if not weighted:
        # unit weights: every sum is exact, so 2m is preserved and the coarse
        # singleton partition has the fine partition's modularity
        assert o.twom_fx == fine_twom
        o.singleton_partition()
        assert o.compute_modularity() == q
    o.close()
