"""HIP path vs the oracle on the same seeded graphs: per-phase communities
bit-exact, modularity within 1e-12 (identical in practice), through the
C ABI (nido_amd -> libnido_amd.so)."""
import json
import os

import numpy as np
import pytest

import oracle
from parity import Q_TOL, assert_runs_equal, run_gpu, run_oracle, sorted_rows

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def _fixture(golden_dir):
    na = _na()
    off, col, w = na.Graph.read(os.path.join(golden_dir, "sbm1024_seed1.bin")).arrays()
    with open(os.path.join(golden_dir, "sbm1024_seed1_expected.json")) as f:
        exp = json.load(f)
    return off, col, w, exp


def test_sbm1024_full_run_matches_golden_and_oracle(golden_dir):
    off, col, w, exp = _fixture(golden_dir)
    print("case: sbm1024 seed1 -b 2 -t 1e-6", flush=True)
    got = run_gpu(off, col, w, 2)
    ref = run_oracle(off, col, w, 2)
    assert_runs_equal(ref, got)
    assert len(got["phases"]) == len(exp["phases"])
    for p, e in zip(got["phases"], exp["phases"]):
        assert p["loops"] == e["loops"]
        assert np.array_equal(p["comm"], np.asarray(e["comm"], dtype=np.int64))
        assert np.array_equal(p["iter_q"], np.asarray(e["iter_q"]))
    assert np.array_equal(got["clustering"], np.asarray(exp["clustering"], dtype=np.int64))


def _step_compare(off, col, w, nb, phases=2):
    """Drive GraphGPU and the oracle call by call; compare after every step."""
    na = _na()
    g = na.GraphGPU((off, col, w), nb)
    o = oracle.Oracle(off, col, w, nb)
    for ph in range(phases):
        assert np.array_equal(g.batch_bounds(), o.batch_bounds())
        assert np.array_equal(g.vertex_weights(), o.vertex_weights())
        assert g.info()["mass"] == o.mass
        g.singleton_partition()
        o.singleton_partition()
        q = g.compute_modularity()
        assert q == o.compute_modularity()
        q_old = q
        dq, loops = 1e308, 0
        while 1e-6 < dq and loops < 50:
            for b in range(nb):
                g.louvain_update(b)
                o.louvain_update(b)
                assert np.array_equal(g.new_communities(), o.new_communities()), f"phase {ph} loop {loops} batch {b}"
                g.update_community_weights(b)
                g.update_community_ids(b)
                o.update_community_weights(b)
                o.update_community_ids(b)
                assert np.array_equal(g.communities(), o.communities())
                assert np.array_equal(g.community_weights(), o.community_weights())
            qt = g.compute_modularity()
            assert qt == o.compute_modularity()
            dq = qt - q
            loops += 1
            if dq < 0:
                g.restore_community()
                o.restore_community()
                assert np.array_equal(g.communities(), o.communities())
                assert np.array_equal(g.community_weights(), o.community_weights())
            else:
                q = qt
        if not (q - q_old > 1e-6):
            break
        d1 = g.aggregation()
        d2 = o.aggregation()
        assert d1 == d2
        assert np.array_equal(g.clustering(), o.clustering())
        if d1:
            break
        go = sorted_rows(*g.graph())
        oo = o.graph()
        for x, y in zip(go, oo):
            assert np.array_equal(x, y)
    g.close()
    o.close()


def test_sbm1024_step_by_step(golden_dir):
    off, col, w, _ = _fixture(golden_dir)
    print("case: sbm1024 stepwise", flush=True)
    _step_compare(off, col, w, 2)


@pytest.mark.parametrize("nb", [1, 3, 8])
def test_weighted_sbm_step_by_step(nb):
    na = _na()
    off, col, w = na.gen_sbm(3000, 50, 0.2, 0.004, True, 7).arrays()
    print(f"case: weighted sbm3000 nb={nb}", flush=True)
    _step_compare(off, col, w, nb)


@pytest.mark.parametrize("scale,ef,weighted,nb", [(12, 16, True, 8), (14, 8, False, 32)])
def test_rmat_full_run(scale, ef, weighted, nb):
    na = _na()
    off, col, w = na.gen_rmat(scale, ef, 0.57, 0.19, 0.19, weighted, 7).arrays()
    print(f"case: rmat s{scale} ef{ef} w{int(weighted)} nb={nb} maxdeg={int(np.max(np.diff(off)))}", flush=True)
    got = run_gpu(off, col, w, nb)
    ref = run_oracle(off, col, w, nb)
    assert_runs_equal(ref, got)


KNOBS = [
    {"thread_max_deg": 0},
    {"thread_max_deg": 1},
    {"big_min_deg": 16, "big_part_arcs": 4},                        # many partitions per row
    {"big_min_deg": 0, "big_part_arcs": 1},                         # every row partitioned
    {"big_min_deg": 32, "big_part_arcs": 64, "big_key_limit": 2},   # forces the HBM fallback
    {"split_small": 1},                                             # one launch per short-row bin
    {"overlap_long": 0},                                            # long rows on the main stream
    {"overlap_long": 0, "split_small": 1},
    {"use_graph": 2},                                               # iterations replayed from a hipGraph
    {"use_graph": 2, "overlap_long": 1},                            # ... with the long rows on a second stream
    {"use_graph": 0, "overlap_long": 0},                            # stream launches only
    {"big_slots": 1024, "big_key_limit": 512, "big_part_arcs": 384},
    {"big_slots": 1024, "big_key_limit": 512, "big_part_arcs": 384, "overlap_long": 1},
    {"full_modularity": 1},                                         # every iteration sweeps all arcs for Q
    {"full_modularity": 1, "use_graph": 2},
    {"full_modularity": 2},                                         # incremental Q after every first iteration
    {"full_modularity": 2, "use_graph": 2},
    {"fixed_short": 0},                                             # short rows in Q64.64 instead of 64-bit fixed point
]


@pytest.mark.parametrize("knobs", KNOBS, ids=[",".join(f"{k}={v}" for k, v in d.items()) for d in KNOBS])
def test_launch_shape_does_not_change_results(knobs):
    na = _na()
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 3).arrays()
    print(f"case: rmat s12 knobs {knobs}", flush=True)
    ref = run_oracle(off, col, w, 4)
    g = na.GraphGPU((off, col, w), 4)
    g.set_tuning(**knobs)
    assert all(g.get_tuning()[k] == v for k, v in knobs.items())
    st = na.LouvainGPU(500, 1e-6, 4).run(g, record=True)
    assert st.phases == len(ref["phases"])
    assert st.q == ref["q"]
    for i, p in enumerate(ref["phases"]):
        assert np.array_equal(g.phase_communities(i), p["comm"]), f"phase {i}"
    g.close()


@pytest.mark.parametrize("scale_w", [2.0 ** 33, 1.0 / 1024])
def test_weight_ranges(scale_w):
    """Row sums >= 2^31 take the 64-bit-high-word long-row tables; tiny
    weights exercise the low end of the Q64.64 grid.  Scaling by a power of
    two keeps every weight exact."""
    na = _na()
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 5).arrays()
    w = w * scale_w
    print(f"case: rmat s12 weights x {scale_w}", flush=True)
    got = run_gpu(off, col, w, 4)
    ref = run_oracle(off, col, w, 4)
    assert_runs_equal(ref, got)


@pytest.mark.parametrize("fixed_short", [1, 0])
def test_signed_and_large_integer_weights(fixed_short):
    """Short-row sums in 64-bit fixed point with integer weights (F = 0) of
    magnitude up to 97 * 2^38, a seventh of them negative (two's-complement
    sums that change sign), against the oracle; and the same graph with the
    short rows in Q64.64."""
    na = _na()
    off, col, w = na.gen_rmat(11, 16, 0.57, 0.19, 0.19, True, 8).arrays()
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    lo, hi = np.minimum(src, col), np.maximum(src, col)
    h = (lo * 2654435761 + hi * 40503) % 97                       # one value per undirected edge
    w = np.ldexp(1.0 + h, 38) * np.where(h % 7 == 0, -1.0, 1.0)    # exact integers up to 97 * 2^38
    print(f"case: rmat s11 signed integer weights, fixed_short={fixed_short}", flush=True)
    ref = run_oracle(off, col, w, 4)
    g = na.GraphGPU((off, col, w), 4)
    g.set_tuning(fixed_short=fixed_short)
    st = na.LouvainGPU(500, 1e-6, 4).run(g, record=True)
    assert st.phases == len(ref["phases"]) and st.q == ref["q"]
    for i, p in enumerate(ref["phases"]):
        assert np.array_equal(g.phase_communities(i), p["comm"]), f"phase {i}"
    assert np.array_equal(g.clustering(), ref["clustering"])
    g.close()


def test_repeat_runs_are_bit_identical():
    na = _na()
    off, col, w = na.gen_sbm(4000, 64, 0.25, 0.003, True, 11).arrays()
    print("case: repeat sbm4000", flush=True)
    a = run_gpu(off, col, w, 4)
    b = run_gpu(off, col, w, 4)
    assert_runs_equal(a, b)


@pytest.mark.parametrize("case", ["empty", "isolated", "one_edge_b1", "one_edge_b3", "path_rollback"])
def test_tiny_graphs(case):
    na = _na()
    if case == "empty":
        off, col, w, nb = np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), 2
    elif case == "isolated":
        off, col, w, nb = np.zeros(6, np.int64), np.zeros(0, np.int64), np.zeros(0), 2
    elif case.startswith("one_edge"):
        off, col, w = np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, 1.0])
        nb = 1 if case.endswith("b1") else 3
    else:
        off, col, w = np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]), np.ones(4)
        nb = 1
    print(f"case: tiny {case}", flush=True)
    got = run_gpu(off, col, w, nb)
    ref = run_oracle(off, col, w, nb)
    assert_runs_equal(ref, got)
    assert abs(got["q"] - ref["q"]) <= Q_TOL


def test_autotune_keeps_results():
    na = _na()
    off, col, w = na.gen_rmat(13, 16, 0.57, 0.19, 0.19, True, 9).arrays()
    print("case: autotune rmat s13", flush=True)
    ref = run_oracle(off, col, w, 8)
    g = na.GraphGPU((off, col, w), 8)
    chosen = g.autotune(1)
    assert g.get_tuning() == chosen
    st = na.LouvainGPU(500, 1e-6, 8).run(g, record=True)
    assert st.phases == len(ref["phases"]) and st.q == ref["q"]
    assert np.array_equal(g.clustering(), ref["clustering"])
    g.close()


def test_chung_lu_full_run():
    """The large-graph family (Chung-Lu, Pareto degrees) at a size the oracle
    finishes in seconds."""
    na = _na()
    off, col, w = na.gen_chung_lu(30000, 2.1, 2.0, 3000.0, 16.0, False, 11).arrays()
    print(f"case: chung-lu 30000 maxdeg={int(np.max(np.diff(off)))}", flush=True)
    got = run_gpu(off, col, w, 32)
    ref = run_oracle(off, col, w, 32)
    assert_runs_equal(ref, got)


def test_cli_matches_oracle(golden_dir, tmp_path):
    """nido_louvain -f <file> -b 2 -t 1e-6 -o <out>: the reference's command
    line (main.cpp:15-177); the written partition is the oracle's."""
    import subprocess
    na = _na()
    exe = os.path.join(os.path.dirname(na.LIB_PATH), "nido_louvain")
    src = os.path.join(golden_dir, "sbm1024_seed1.bin")
    out = tmp_path / "comm.txt"
    print("case: CLI sbm1024", flush=True)
    r = subprocess.run([exe, "-f", src, "-b", "2", "-t", "1e-6", "-o", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "#Batches: 2" in r.stdout
    got = np.loadtxt(out, dtype=np.int64)
    off, col, w = na.Graph.read(src).arrays()
    ref = run_oracle(off, col, w, 2)
    assert np.array_equal(got, ref["clustering"])


def test_profiling_events_sample_every_stride_th_launch():
    """Per-class HIP-event timing: every launch with stride 1, every third
    with stride 3 (the bench's default), and the results are unchanged."""
    na = _na()
    off, col, w = na.gen_rmat(14, 16, 0.57, 0.19, 0.19, True, 5).arrays()
    nb = 8
    print(f"case: profiling strides, rmat14 -b {nb}", flush=True)
    runs = []
    for stride in (1, 3):
        g = na.GraphGPU((off, col, w), nb)
        g.singleton_partition()
        q = g.compute_modularity()
        g.set_profiling(True, classes=["lm_short", "lm_long"], stride=stride)
        qa, dq, qe = g.louvain_iteration(q)
        ks = g.kernel_stats()
        assert ks["lm_short"]["launches"] == -(-nb // stride)
        assert ks["lm_short"]["ms"] > 0 and ks["lm_short"]["bytes"] > 0
        assert ks["commit"]["launches"] == 0
        runs.append((qa, dq, qe, g.communities()))
        g.close()
    assert runs[0][:3] == runs[1][:3]
    assert np.array_equal(runs[0][3], runs[1][3])
    g = na.GraphGPU((off, col, w), nb)
    with pytest.raises(na.NidoError):
        g.set_profiling(True, stride=0)
    g.close()


@pytest.mark.parametrize("use_graph", [0, 2])
def test_modularity_after_singleton_iteration_and_restore(use_graph):
    """compute_modularity right after singleton_partition counts self arcs
    alone (C is the identity there); after an iteration, a restore or a
    second singleton reset it must match the oracle's full sweep again."""
    na = _na()
    off, col, w = na.gen_rmat(13, 8, 0.57, 0.19, 0.19, True, 9).arrays()
    nb = 4
    print(f"case: modularity after singleton/iteration/restore, rmat13 -b {nb} use_graph={use_graph}", flush=True)
    g = na.GraphGPU((off, col, w), nb)
    g.set_tuning(use_graph=use_graph)
    o = oracle.Oracle(off, col, w, nb)
    for _ in range(2):
        g.singleton_partition()
        o.singleton_partition()
        q = g.compute_modularity()
        assert q == o.compute_modularity()
        qa, dq, qe = g.louvain_iteration(q)
        ra, rdq, rqe = o.louvain_iteration(q)
        assert (qa, dq, qe) == (ra, rdq, rqe)
        assert g.compute_modularity() == o.compute_modularity()
        for b in range(nb):
            g.louvain_update(b)
            o.louvain_update(b)
            g.update_community_weights(b)
            g.update_community_ids(b)
            o.update_community_weights(b)
            o.update_community_ids(b)
        assert g.compute_modularity() == o.compute_modularity()
        g.restore_community()
        o.restore_community()
        assert np.array_equal(g.communities(), o.communities())
        assert g.compute_modularity() == o.compute_modularity()
    g.close()


def _hubs_at_bin_edges(seed=17):
    """Hubs whose degrees sit on both sides of every bin and partition edge,
    each joined to its own leaves (so in the first iteration every hub sees
    d distinct communities: the tables' fullest case), plus a sparse random
    layer among the leaves; U(0,1] weights on the 2^-53 grid."""
    rng = np.random.default_rng(seed)
    degs = [8, 9, 16, 17, 32, 33, 64, 65, 128, 129, 256, 257, 511, 512, 513, 1023, 1024, 1025, 2048, 2049, 4097]
    nh = len(degs)
    n = nh + sum(degs)
    edges = set()
    nxt = nh
    for h, d in enumerate(degs):
        for j in range(d):
            edges.add((h, nxt + j))
        nxt += d
    leaves = np.arange(nh, n)
    for _ in range(3 * n):
        a, b = rng.choice(leaves, 2, replace=False)
        edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), dtype=np.int64)
    wt = (np.int64(1 << 53) - rng.integers(0, 1 << 53, len(e))) * 2.0 ** -53
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    ww = np.concatenate([wt, wt])
    o = np.lexsort((dst, src))
    src, dst, ww = src[o], dst[o], ww[o]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst, ww


@pytest.mark.parametrize("knobs", [{}, {"big_slots": 1024, "big_part_arcs": 512, "big_key_limit": 512},
                                   {"big_slots": 4096, "big_part_arcs": 2048, "big_key_limit": 2048}])
def test_rows_at_bin_and_partition_edges(knobs):
    """Every short-row bin at its capacity (keys = 2 x lanes) and long rows
    just below and above each partition boundary, under three table sizes."""
    na = _na()
    off, col, w = _hubs_at_bin_edges()
    nb = 3
    print(f"case: hubs at bin edges, -b {nb}, {knobs}", flush=True)
    ref = run_oracle(off, col, w, nb)
    g = na.GraphGPU((off, col, w), nb)
    if knobs:
        g.set_tuning(**knobs)
    st = na.LouvainGPU(500, 1e-6, nb).run(g, record=True)
    got = {"phases": [], "q": st.q, "clustering": g.clustering()}
    for i in range(st.phases):
        p = g.phase(i).as_dict()
        p["iter_q"] = g.iteration_q(i)
        p["comm"] = g.phase_communities(i)
        got["phases"].append(p)
    g.close()
    assert_runs_equal(ref, got)


def test_min_id_tie_break_closed_form_on_the_device():
    """Vertex 0 of the oracle tests' tie graph has two candidates (5 then 3
    in its row) with exactly equal gains, (1 - 2*2/10)/5; the smallest id,
    3, must win on the device too, and again with 3 and 5 relabelled (the
    tie in the other arc order).  A weight-2 edge 0-5 breaks the tie in 5's
    favour."""
    na = _na()
    off = np.array([0, 2, 3, 4, 6, 7, 9, 10], np.int64)
    col = np.array([5, 3, 3, 5, 0, 1, 6, 0, 2, 4], np.int64)
    col_swapped = np.array([3, 5, 5, 3, 0, 2, 6, 0, 1, 4], np.int64)
    w2 = np.ones(10)
    w2[[0, 7]] = 2.0                     # arcs 0 -> 5 and 5 -> 0
    for c, w, want in ((col, np.ones(10), 3), (col_swapped, np.ones(10), 3), (col, w2, 5)):
        g = na.GraphGPU((off, c, w), 1)
        g.singleton_partition()
        g.louvain_update(0)
        assert g.new_communities()[0] == want, (list(c), list(w))
        g.close()


def _asymmetric_graphs(na):
    """Graphs where an arc's reverse carries another weight, or is missing."""
    rng = np.random.default_rng(11)
    n = 5000
    src = rng.integers(0, n // 4, 3000)
    dst = rng.integers(0, n // 4, 3000)
    keep = src != dst
    e = np.unique(np.concatenate([np.stack([src[keep], dst[keep]], 1), np.stack([dst[keep], src[keep]], 1)]), axis=0)
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, e[:, 0] + 1, 1)
    sparse = (np.cumsum(off), e[:, 1].astype(np.int64), rng.uniform(0.1, 1.0, len(e)))   # weights per arc
    off, col, w = na.gen_rmat(11, 16, 0.57, 0.19, 0.19, True, 9).arrays()
    rmat = (off, col, w * rng.uniform(0.5, 1.5, len(w)))                                   # weights per arc
    one_way = (off, col, w.copy())
    one_way[2][::7] = 0.0      # some arcs weigh nothing one way only
    return [sparse, rmat, one_way]


@pytest.mark.parametrize("nb", [1, 2, 4, 32])
def test_asymmetric_weights(nb):
    """The incremental modularity counts an arc whose state changes for
    itself and for its reverse arc; where the reverse arc weighs something
    else every iteration sweeps all arcs, and runs equal the oracle's."""
    na = _na()
    for i, (off, col, w) in enumerate(_asymmetric_graphs(na)):
        print(f"case: asymmetric graph {i} nb={nb}", flush=True)
        assert_runs_equal(run_oracle(off, col, w, nb), run_gpu(off, col, w, nb))
