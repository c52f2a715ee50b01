"""Many graphs in one handle (nido_batch_run): every graph's run must equal
its run alone -- same phases, iterations per phase, modularity and final
clustering -- checked against the oracle, graph by graph."""
import numpy as np
import pytest

from parity import Q_TOL, run_oracle

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def _mixed_graphs(na):
    gs = []
    # planted partitions of different sizes and densities, weighted and not
    for i, (n, blk, pin, pout, wtd) in enumerate([(256, 64, 12 / 63, 2 / 192, True), (700, 64, 0.2, 0.01, False),
                                                   (1500, 50, 0.25, 0.004, True), (3000, 64, 12 / 63, 2 / 2936, True),
                                                   (330, 30, 0.4, 0.05, True)]):
        gs.append(na.gen_sbm(n, blk, pin, pout, wtd, 1000 + i).arrays())
    gs.append(na.gen_rmat(10, 8, 0.57, 0.19, 0.19, True, 3).arrays())        # skewed, has long rows
    gs.append((np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]), np.ones(4)))   # path: rolls back
    gs.append((np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, 1.0])))  # one edge
    gs.append((np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0)))    # no vertices
    gs.append((np.zeros(6, np.int64), np.zeros(0, np.int64), np.zeros(0)))    # isolated vertices
    gs.append((np.array([0, 1, 2]), np.array([1, 0]), np.array([0.0, 0.0])))  # arcs of weight 0
    return gs


def _check(na, graphs, nb, max_loops=500, tol=1e-6, tol_phase=None, **tuning):
    ref = [run_oracle(off, col, w, nb, tol, max_loops, tol_phase=tol_phase) for off, col, w in graphs]
    bg = na.BatchGPU(graphs, nb)
    if tuning:
        bg.set_tuning(**tuning)
    st = bg.run(max_loops, tol, tol_phase)
    assert st.phases == max(len(r["phases"]) for r in ref)
    for i, r in enumerate(ref):
        gs = bg.graph_stats(i)
        assert gs.phases == len(r["phases"]), f"graph {i}: phases {gs.phases} vs {len(r['phases'])}"
        for p, rp in enumerate(r["phases"]):
            ps = bg.graph_phase(i, p)
            assert ps.nv == rp["nv"] and ps.ne == rp["ne"], f"graph {i} phase {p}"
            assert ps.loops == rp["loops"], f"graph {i} phase {p}: loops {ps.loops} vs {rp['loops']}"
            assert abs(ps.q - rp["q"]) <= Q_TOL and abs(ps.q0 - rp["q0"]) <= Q_TOL, f"graph {i} phase {p}"
        assert abs(gs.q - r["q"]) <= Q_TOL, f"graph {i}: Q {gs.q} vs {r['q']}"
        assert np.array_equal(bg.clustering(i), r["clustering"]), f"graph {i}: clustering differs"
    rebuilds = bg.list_rebuilds()
    bg.close()
    return rebuilds


@pytest.mark.parametrize("nb", [1, 4, 32])
def test_batch_equals_each_graph_alone(nb):
    na = _na()
    print(f"case: mixed batch nb={nb}", flush=True)
    _check(na, _mixed_graphs(na), nb)


def test_batch_replayed_from_hip_graphs():
    na = _na()
    print("case: mixed batch nb=8 use_graph=2", flush=True)
    _check(na, _mixed_graphs(na), 8, use_graph=2)


def test_batch_of_one_graph_is_the_plain_run():
    na = _na()
    off, col, w = na.gen_sbm(2000, 64, 0.2, 0.003, True, 5).arrays()
    print("case: batch of one", flush=True)
    _check(na, [(off, col, w)], 8)


def test_batch_many_small_graphs():
    na = _na()
    rng = np.random.default_rng(7)
    graphs = []
    for i in range(64):
        n = int(np.exp(rng.uniform(np.log(128), np.log(2048))))
        graphs.append(na.gen_sbm(n, 64, 12 / 63, 2 / max(n - 64, 1), True, 2000 + i).arrays())
    print("case: 64 small sbms nb=32", flush=True)
    _check(na, graphs, 32)



def test_batch_autotune_keeps_results():
    na = _na()
    graphs = _mixed_graphs(na)
    print("case: mixed batch nb=4 autotuned", flush=True)
    ref = [run_oracle(off, col, w, 4) for off, col, w in graphs]
    bg = na.BatchGPU(graphs, 4)
    bg.autotune(1)
    bg.run(500, 1e-6)
    for i, r in enumerate(ref):
        assert abs(bg.graph_stats(i).q - r["q"]) <= Q_TOL, f"graph {i}"
        assert np.array_equal(bg.clustering(i), r["clustering"]), f"graph {i}"
    bg.close()


def test_batch_with_partitioned_long_rows():
    """Graphs whose hubs exceed the long-row threshold and the partition
    size (rows split over several workgroups) next to small graphs."""
    na = _na()
    graphs = [na.gen_rmat(14, 16, 0.57, 0.19, 0.19, True, 21).arrays(),
              na.gen_sbm(500, 50, 0.3, 0.02, True, 22).arrays(),
              na.gen_rmat(13, 32, 0.6, 0.15, 0.15, False, 23).arrays()]
    for off, col, w in graphs[::2]:
        assert int(np.max(np.diff(off))) > 1536
    print("case: batch with long rows nb=8", flush=True)
    _check(na, graphs, 8)
    print("case: batch with long rows nb=8, small partitions", flush=True)
    _check(na, graphs, 8, big_part_arcs=64, big_min_deg=128)


def _history(bg, n):
    out = []
    for i in range(n):
        gs = bg.graph_stats(i)
        ph = [bg.graph_phase(i, p) for p in range(gs.phases)]
        out.append(([(p.nv, p.ne, p.loops, p.q0, p.q) for p in ph], gs.q, bg.clustering(i).tolist()))
    return out


def test_batch_incremental_modularity_equals_the_sweep():
    """Per-graph Q from the tracked commits and per-batch deltas (always
    after a phase's first iteration, or when few arcs moved) must equal a
    sweep over every arc bit for bit: same loops, same Q in every phase of
    every graph."""
    na = _na()
    graphs = _mixed_graphs(na) + [na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 31).arrays()]
    runs = []
    for full in (1, 0, 2):
        print(f"case: mixed batch nb=8 full_modularity={full}", flush=True)
        bg = na.BatchGPU(graphs, 8)
        bg.set_tuning(full_modularity=full)
        bg.run(500, 1e-6)
        runs.append(_history(bg, len(graphs)))
        bg.close()
    for i, (a, b, c) in enumerate(zip(*runs)):
        assert a == b == c, f"graph {i}: the incremental Q differs from the sweep"


def test_batch_rollbacks_that_go_on():
    """tol < 0: a graph whose iteration lowers Q rolls back and keeps
    iterating (the loop runs while dQ > tol, louvain_gpu.cpp:53), so its Q
    must be evaluated afresh from the restored partition."""
    na = _na()
    print("case: mixed batch nb=4 tol=-1 max_loops=5", flush=True)
    _check(na, _mixed_graphs(na), 4, max_loops=5, tol=-1.0, tol_phase=1e-6)


def test_batch_with_asymmetric_weights():
    """One graph whose reverse arcs weigh something else: the handle sweeps
    every iteration, and every graph still equals its oracle run."""
    from test_gpu_parity import _asymmetric_graphs
    na = _na()
    graphs = _mixed_graphs(na)[:6] + _asymmetric_graphs(na)[:1]
    print("case: mixed batch with an asymmetric graph nb=4", flush=True)
    _check(na, graphs, 4)
