"""Work lists rebuilt for the graphs still iterating in a batched handle
(nido_set_list_rebuild): the stopped graphs' rows leave the short- and
long-row launches, and every graph's run must stay its run alone (phases,
iterations per phase, Q, clustering against the oracle)."""
import pytest

from test_gpu_batch import _check, _mixed_graphs

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


@pytest.fixture
def eager_rebuild():
    na = _na()
    na.set_list_rebuild(100)   # rebuild as soon as one listed graph stops
    yield na
    na.set_list_rebuild(-1)


@pytest.mark.parametrize("nb", [1, 4, 32])
def test_rebuilt_lists_keep_results(eager_rebuild, nb):
    na = eager_rebuild
    print(f"case: mixed batch nb={nb}, lists rebuilt at every stop", flush=True)
    assert _check(na, _mixed_graphs(na), nb) > 0


@pytest.mark.parametrize("knobs", [{"use_graph": 2}, {"full_modularity": 2}, {"full_modularity": 1},
                                   {"big_min_deg": 64, "big_part_arcs": 64}],
                         ids=["replayed", "incremental_q", "swept_q", "long_rows_from_64"])
def test_rebuilt_lists_with_knobs(eager_rebuild, knobs):
    na = eager_rebuild
    print(f"case: mixed batch nb=8 knobs {knobs}, lists rebuilt at every stop", flush=True)
    assert _check(na, _mixed_graphs(na), 8, **knobs) > 0


def test_rebuilt_device_lists(eager_rebuild):
    na = eager_rebuild
    na.set_device_lists(1)
    try:
        print("case: mixed batch nb=8, device lists rebuilt at every stop", flush=True)
        assert _check(na, _mixed_graphs(na), 8) > 0
        print("case: long-row batch nb=8, device lists rebuilt at every stop", flush=True)
        graphs = [na.gen_rmat(13, 16, 0.57, 0.19, 0.19, True, 21).arrays(),
                  na.gen_sbm(500, 50, 0.3, 0.02, True, 22).arrays(),
                  na.gen_rmat(12, 32, 0.6, 0.15, 0.15, False, 23).arrays()]
        _check(na, graphs, 8, big_part_arcs=64, big_min_deg=128)
    finally:
        na.set_device_lists(0)


def test_rollbacks_with_rebuilt_lists(eager_rebuild):
    """tol < 0: graphs that roll back go on iterating after others stop."""
    na = eager_rebuild
    print("case: mixed batch nb=4 tol=-1 max_loops=12, lists rebuilt at every stop", flush=True)
    _check(na, _mixed_graphs(na), 4, max_loops=12, tol=-1.0, tol_phase=1e-6)


def test_no_rebuild_when_switched_off():
    na = _na()
    na.set_list_rebuild(0)
    try:
        print("case: mixed batch nb=4, no rebuilds", flush=True)
        assert _check(na, _mixed_graphs(na), 4) == 0
    finally:
        na.set_list_rebuild(-1)
