"""Out-of-core edge windows (graph_gpu.cpp:341-417, 967-1027): the device
holds at most `window` arcs at a time.  Runs with windows small enough to cut
every batch into several units, and to keep the first coarse graphs on the
host, must equal the oracle's -- every phase's communities, every
iteration's modularity, the final clustering -- and the step-by-step API
must see the graph it was given."""
import numpy as np
import pytest

from parity import assert_runs_equal, run_oracle

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def _run(na, off, col, w, nb, window, **knobs):
    g = na.GraphGPU((off, col, w), nb, window=window)
    if knobs:
        g.set_tuning(**knobs)
    st = na.LouvainGPU(500, 1e-6, nb).run(g, record=True)
    phases = []
    for i in range(st.phases):
        p = g.phase(i).as_dict()
        p["iter_q"] = g.iteration_q(i)
        p["comm"] = g.phase_communities(i)
        phases.append(p)
    out = {"phases": phases, "q": st.q, "clustering": g.clustering()}
    g.close()
    return out


@pytest.mark.parametrize("window", [500, 3000, 20000])
def test_windowed_sbm_matches_oracle(window):
    na = _na()
    off, col, w = na.gen_sbm(4000, 64, 0.25, 0.004, True, 21).arrays()
    deg = np.diff(off)
    print(f"case: sbm 4000 weighted, {len(col)} arcs, max degree {deg.max()}, window {window}, -b 3", flush=True)
    assert deg.max() <= window < len(col)
    ref = run_oracle(off, col, w, 3)
    got = _run(na, off, col, w, 3, window)
    assert_runs_equal(ref, got)
    if window == 500:   # the first coarse graph has more arcs than the window: it stayed on the host
        assert ref["phases"][1]["ne"] > window


@pytest.mark.parametrize("window", [8192, 60000])
def test_windowed_rmat_with_long_rows_matches_oracle(window):
    """R-MAT rows of more than 512 arcs take the partitioned long-row path
    (key fill into buckets, partial combine) inside the window."""
    na = _na()
    off, col, w = na.gen_rmat(13, 16, 0.57, 0.19, 0.19, True, 4).arrays()
    deg = np.diff(off)
    print(f"case: rmat s13, {len(col)} arcs, max degree {deg.max()}, window {window}, -b 4", flush=True)
    assert deg.max() <= window < len(col) and int(np.sum(deg > 2048)) > 0
    ref = run_oracle(off, col, w, 4)
    got = _run(na, off, col, w, 4, window)
    assert_runs_equal(ref, got)


def test_windowed_step_api_and_graph_readback():
    na = _na()
    off, col, w = na.gen_sbm(2000, 50, 0.3, 0.01, True, 5).arrays()
    g = na.GraphGPU((off, col, w), 2, window=4000)
    o2, c2, w2 = g.graph()
    assert np.array_equal(o2, off) and np.array_equal(c2, col) and np.array_equal(w2, w)
    g.singleton_partition()
    q = g.compute_modularity()
    import oracle
    orc = oracle.Oracle(off, col, w, 2)
    orc.singleton_partition()
    assert q == orc.compute_modularity()
    q1, dq1, qe1 = g.louvain_iteration(q)
    q2, dq2, qe2 = orc.louvain_iteration(q)
    assert qe1 == qe2 and np.array_equal(g.communities(), orc.communities())
    orc.close()
    g.close()


def test_row_longer_than_the_window_is_refused():
    na = _na()
    off, col, w = na.gen_rmat(10, 16, 0.57, 0.19, 0.19, True, 2).arrays()
    with pytest.raises(na.NidoError, match="window"):
        na.GraphGPU((off, col, w), 2, window=int(np.diff(off).max()) - 1)


@pytest.mark.parametrize("fit", [1, 2, 3])
@pytest.mark.parametrize("full_modularity", [0, 2])
def test_windowed_batches_that_fit_update_q_incrementally(full_modularity, fit):
    """A window that holds every batch at once (but not the graph): each
    batch's modularity delta is taken right after its commit, from the
    arcs still in the window, instead of a second pass over all arcs.
    With room for two batches (fit 2, and fit 3 for an odd window) the
    next batch is copied into the other half while the current one runs."""
    import oracle
    na = _na()
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 6).arrays()
    parts = oracle.batch_bounds(off, 8)
    largest = int(max(off[parts[b + 1]] - off[parts[b]] for b in range(8)))
    window = {1: largest, 2: 2 * largest, 3: 2 * largest + 1}[fit]
    print(f"case: rmat s12, {len(col)} arcs, window {window} ({fit} x the largest batch), -b 8, "
          f"full_modularity {full_modularity}", flush=True)
    assert window < len(col)
    ref = run_oracle(off, col, w, 8)
    got = _run(na, off, col, w, 8, window, full_modularity=full_modularity)
    assert_runs_equal(ref, got)
