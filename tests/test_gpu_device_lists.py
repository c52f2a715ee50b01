"""Degree bins and sweep row lists built on the device (bins.hip) instead of
the host: the device builder is forced for every phase here, and the runs
must equal the oracle's as with the host lists (every batch's C', the
totals, Q, the clustering)."""
import numpy as np
import pytest

from parity import assert_runs_equal, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


@pytest.fixture
def device_lists():
    na = _na()
    na.set_device_lists(1)
    yield na
    na.set_device_lists(0)


def test_step_by_step_with_device_lists(device_lists):
    from test_gpu_parity import _step_compare
    na = device_lists
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 5).arrays()
    print("case: rmat s12 step compare nb=8, device lists", flush=True)
    _step_compare(off, col, w, 8)


@pytest.mark.parametrize("nb", [1, 7, 32])
@pytest.mark.parametrize("knobs", [{}, {"big_min_deg": 64, "big_part_arcs": 64}, {"thread_max_deg": 0}],
                         ids=["default", "long_rows_from_64", "no_thread_bin"])
def test_full_runs_with_device_lists(device_lists, nb, knobs):
    na = device_lists
    off, col, w = na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 3).arrays()
    print(f"case: rmat s12 nb={nb} knobs {knobs}, device lists", flush=True)
    assert_runs_equal(run_oracle(off, col, w, nb), run_gpu(off, col, w, nb, **knobs))


def test_isolated_and_tiny_graphs_with_device_lists(device_lists):
    from test_gpu_parity import _asymmetric_graphs
    na = device_lists
    sparse = _asymmetric_graphs(na)[0]   # three quarters of its vertices isolated
    cases = [sparse, (np.zeros(9, np.int64), np.zeros(0, np.int64), np.zeros(0)),
             (np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, 1.0]))]
    for i, (o, c, x) in enumerate(cases):
        for nb in (1, 4, 32):
            print(f"case: sparse/tiny graph {i} nb={nb}, device lists", flush=True)
            assert_runs_equal(run_oracle(o, c, x, nb), run_gpu(o, c, x, nb))


@pytest.mark.parametrize("nb", [4, 32])
def test_batch_with_device_lists(device_lists, nb):
    from test_gpu_batch import _check, _mixed_graphs
    na = device_lists
    print(f"case: mixed batch nb={nb}, device lists", flush=True)
    _check(na, _mixed_graphs(na), nb)
    print(f"case: mixed batch nb={nb}, device lists, long rows from 64", flush=True)
    _check(na, _mixed_graphs(na), nb, big_min_deg=64, big_part_arcs=64)
