"""The reference's -r generator and -c colouring: CPU checks.

* The Python restatements of the standard-library engines (stdlib_rng.py)
  against the values the C++ standard requires of them, and the library's
  BA generator (nido_gen_ba) against a restatement of graph.cpp:76-110 built
  on them: the same arcs with the same weights, bit for bit.
* The oracle's colouring (the reference's loop, graph.cpp:573-605, run by
  one thread) against closed forms -- a clique whose random weights fall
  with the id takes one colour per vertex, one whose weights rise takes a
  single colour, equal weights give the sequential greedy colouring (a path
  alternates) -- and against the rule itself on random graphs: every colour
  is the smallest one missing among the vertex's earlier neighbours of
  larger or equal weight.
* The oracle's renumbering (graph.cpp:607-691): (colour, id) order, and the
  renumbered graph is the old one relabelled, row arc order kept.
"""
import numpy as np
import pytest

import oracle
import stdlib_rng as R


def test_engines_meet_the_standard_values():
    g = R.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042       # [rand.predef]: mt19937_64
    m = R.MinStdRand0()
    for _ in range(9999):
        m()
    assert m() == 1043618065                # [rand.predef]: minstd_rand0


def test_uniform_int_lemire_matches_a_direct_rejection_count():
    # Lemire's method maps the 128-bit product's high word; over a tiny range
    # every value is hit about equally often
    e = R.MT19937_64(1)
    xs = [R.uniform_int(e, 0, 6) for _ in range(7000)]
    counts = np.bincount(xs, minlength=7)
    assert counts.min() > 850 and counts.max() < 1150


@pytest.mark.parametrize("nv,m0", [(60, 1), (500, 3), (3000, 5)])
def test_ba_generator_is_the_reference_construction(nv, m0):
    import nido_amd as na
    off, col, w = na.gen_ba(nv, m0).arrays()
    el = sorted(R.ba_edges(nv, m0), key=lambda t: (t[0], t[1]))   # stable: draw order among duplicates
    assert len(col) == len(el) == 2 * m0 * (nv - m0)
    x = np.array([t[0] for t in el])
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.bincount(x, minlength=nv))]))
    assert np.array_equal(col, np.array([t[1] for t in el]))
    assert np.array_equal(w, np.array([t[2] for t in el]))
    # the construction: a star on 0..m0, then m0 attachments to earlier vertices
    src = np.repeat(np.arange(nv), np.diff(off))
    assert not np.any(src == col)
    earlier = np.bincount(src[col < src], minlength=nv)
    assert np.all(earlier[1:m0 + 1] == 1) and np.all(earlier[m0 + 1:] == m0)
    assert np.all(w > 0) and np.all(np.isfinite(w))


def _csr(nv, edges):
    rows = [[] for _ in range(nv)]
    for u, v in edges:
        rows[u].append(v)
        rows[v].append(u)
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    col = np.array([c for r in rows for c in r], np.int64)
    return off, col


def test_coloring_closed_forms():
    n = 7
    off, col = _csr(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
    c, K = oracle.coloring(off, col, np.arange(n, 0, -1))        # weights fall with the id
    assert K == n and np.array_equal(c, np.arange(n))
    c, K = oracle.coloring(off, col, np.arange(n))               # weights rise with the id
    assert K == 1 and np.all(c == 0)
    off, col = _csr(6, [(i, i + 1) for i in range(5)])           # path, equal weights
    c, K = oracle.coloring(off, col, np.zeros(6, np.int64))
    assert np.array_equal(c, [0, 1, 0, 1, 0, 1])
    # a star whose centre (id 0) has the smallest weight: the leaves see it
    # only as a neighbour with a smaller weight, so every leaf keeps colour 0
    off, col = _csr(5, [(0, i) for i in range(1, 5)])
    c, _ = oracle.coloring(off, col, np.array([0, 5, 5, 5, 5]))
    assert np.all(c == 0)
    c, _ = oracle.coloring(off, col, np.array([9, 5, 5, 5, 5]))  # now they avoid it
    assert np.array_equal(c, [0, 1, 1, 1, 1])


def _check_rule(off, col, r, c):
    nv = len(off) - 1
    for i in range(nv):
        nb = col[off[i]:off[i + 1]]
        dep = nb[(nb < i) & (r[nb] >= r[i])]
        held = set(c[dep].tolist())
        assert c[i] not in held
        assert all(k in held for k in range(c[i]))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_coloring_rule_on_random_graphs(seed):
    rng = np.random.default_rng(seed)
    nv = 300
    edges = set()
    while len(edges) < 1500:
        u, v = rng.integers(0, nv, 2)
        if u != v:
            edges.add((min(u, v), max(u, v)))
    off, col = _csr(nv, sorted(edges))
    r = rng.integers(0, 4 * nv + 1, nv)
    r[:20] = r[20]                                               # ties count as dependencies
    c, K = oracle.coloring(off, col, r)
    assert K == c.max() + 1
    _check_rule(off, col, r, c)


def test_reorder_relabels_the_graph():
    rng = np.random.default_rng(5)
    nv = 200
    edges = [tuple(sorted(rng.choice(nv, 2, replace=False))) for _ in range(800)]
    off, col = _csr(nv, edges)
    w = rng.uniform(0.1, 1.0, len(col))
    c, _ = oracle.coloring(off, col, rng.integers(0, 800, nv))
    no, noff, ncol, nw = oracle.reorder(off, col, w, c)
    keys = [(c[v], v) for v in no]
    assert keys == sorted(keys) and sorted(no.tolist()) == list(range(nv))
    orders = np.empty(nv, np.int64)
    orders[no] = np.arange(nv)
    for p in range(nv):
        v = no[p]
        assert np.array_equal(ncol[noff[p]:noff[p + 1]], orders[col[off[v]:off[v + 1]]])
        assert np.array_equal(nw[noff[p]:noff[p + 1]], w[off[v]:off[v + 1]])
