"""Request loop (nido_serve_*): requests that arrive while others run are
admitted as slots and the HBM budget free up and overlap on the device,
small ones grouped into one batched handle per worker when `group` > 1; each
result must equal the result of the same graph run alone."""
import numpy as np
import pytest

from parity import Q_TOL, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def _requests(na):
    gs = []
    for i, (s, ef) in enumerate([(10, 8), (12, 16), (11, 32), (13, 8), (10, 16), (12, 8), (14, 16), (11, 8)]):
        off, col, _ = na.gen_rmat(s, ef, 0.57, 0.19, 0.19, False, 500 + i).arrays()
        gs.append((off, col, np.ones(len(col))))
    gs.append(na.gen_sbm(3000, 64, 12 / 63, 2 / 2936, True, 77).arrays())
    gs.append((np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]), np.ones(4)))
    return gs


@pytest.mark.parametrize("slots,budget,group,borrow", [(4, 200 << 30, 1, False), (3, 150 << 20, 1, False),
                                                       (2, 200 << 30, 64, False), (2, 200 << 30, 64, True),
                                                       (3, 150 << 20, 8, False), (1, 200 << 30, 4, True)])
def test_each_result_is_the_same_alone_and_in_the_mix(slots, budget, group, borrow):
    na = _na()
    reqs = _requests(na)
    print(f"case: serve {len(reqs)} requests, {slots} slots, budget {budget}, group {group}, borrow {borrow}",
          flush=True)
    srv = na.Server(slots=slots, nbatches=32, hbm_budget=budget, group=group)
    ids = [srv.submit(g, borrow=borrow) for g in reqs]
    got = []
    for rid in ids:
        st, lat = srv.wait(rid)
        assert lat > 0
        got.append((st.q, st.phases, st.total_loops, srv.clustering(rid)))
        srv.release(rid)
    status = srv.status()
    assert status["completed"] == len(reqs) and status["queued"] == 0
    # running counts requests: a worker's group holds up to `group` of them
    assert 1 <= status["peak_running"] <= slots * group
    srv.close()
    for i, (off, col, w) in enumerate(reqs):
        alone = run_gpu(off, col, w, 32)
        q, phases, loops, clus = got[i]
        assert phases == len(alone["phases"]), f"request {i}"
        assert loops == alone["stats"]["total_loops"], f"request {i}"
        assert abs(q - alone["q"]) <= Q_TOL, f"request {i}"
        assert np.array_equal(clus, alone["clustering"]), f"request {i}"
    # and the plain run itself is the oracle's
    off, col, w = reqs[1]
    ref = run_oracle(off, col, w, 32)
    assert np.array_equal(got[1][3], ref["clustering"])


def test_request_larger_than_the_budget_is_refused():
    na = _na()
    srv = na.Server(slots=2, nbatches=4, hbm_budget=1 << 20)
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, False, 1).arrays()
    with pytest.raises(na.NidoError):
        srv.submit((off, col, w))
    srv.close()


def test_malformed_requests_are_refused_and_the_group_still_runs():
    """Each request is checked on its own at submission: a column outside
    the graph, a weight that is not finite, or weights whose total would
    wrap the exact sums are refused there, and the well-formed requests queued
    around them (grouped into one batched handle) get their own results."""
    na = _na()
    reqs = _requests(na)[:5]
    srv = na.Server(slots=1, nbatches=8, group=64)
    off, col, w = reqs[0]
    bad_col = col.copy()
    bad_col[3] = len(off) + 5
    bad_w = w.copy()
    bad_w[1] = np.nan
    path = (np.array([0, 1, 3, 4]), np.array([1, 0, 2, 1]), np.full(4, 3e18))
    ids = []
    for i, g in enumerate(reqs):
        ids.append(srv.submit(g))
        if i == 1:
            for bad in ((off, bad_col, w), (off, col, bad_w), path):
                with pytest.raises(na.NidoError) as e:
                    srv.submit(bad)
                assert e.value.code == -1
    got = []
    for rid in ids:
        st, _ = srv.wait(rid)
        got.append((st.q, srv.clustering(rid)))
        srv.release(rid)
    srv.close()
    for i, (o, c, x) in enumerate(reqs):
        alone = run_gpu(o, c, x, 8)
        assert abs(got[i][0] - alone["q"]) <= Q_TOL, f"request {i}"
        assert np.array_equal(got[i][1], alone["clustering"]), f"request {i}"
