"""Host-side paths of the device entry points that only large inputs reach,
forced at small sizes: the pinned, double-buffered upload that narrows arc
ids to int32 while staging (several chunks per array), and the checks of
the tuning knobs on a live handle."""
import numpy as np
import pytest

from parity import assert_runs_equal, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def test_pinned_staged_upload_with_many_chunks():
    na = _na()
    off, col, w = na.gen_rmat(13, 16, 0.57, 0.19, 0.19, True, 3).arrays()
    # ids cross as int32: ~0.5 MB of them, weights ~1 MB, offsets 64 KB; 16 KiB
    # chunks give dozens of hand-offs between the two pinned buffers
    na.set_upload_staging(4096, 16384)
    try:
        print(f"case: pinned upload, rmat s13 ({len(col)} arcs) in 16 KiB chunks", flush=True)
        g = na.GraphGPU((off, col, w), 8)
        o2, c2, w2 = g.graph()
        assert np.array_equal(o2, off) and np.array_equal(c2, col) and np.array_equal(w2, w)
        g.close()
        assert_runs_equal(run_oracle(off, col, w, 8), run_gpu(off, col, w, 8))
        # a batched handle uploads its concatenated CSR the same way
        parts = [na.gen_sbm(n, 64, 12 / 63, 2 / (n - 64), True, 40 + n).arrays() for n in (700, 1500, 300)]
        vptr, boff, bcol, bw = na.concat_graphs(parts)
        bg = na.BatchGPU(csr=(vptr, boff, bcol, bw), nbatches=4)
        bg.run(500, 1e-6)
        for i, (o, c, x) in enumerate(parts):
            assert np.array_equal(bg.clustering(i), run_oracle(o, c, x, 4)["clustering"]), f"graph {i}"
        bg.close()
    finally:
        na.set_upload_staging(0, 0)
    with pytest.raises(na.NidoError):
        na.set_upload_staging(4096, 1000)          # chunks are multiples of 4096


def test_batched_ids_sent_as_16_bit_graph_local_ids():
    """A batched handle whose graphs all have at most 2^16 vertices sends
    its ids as 16-bit graph-local ids while staging, and the device puts
    the graph bases back: the device CSR must equal the input exactly (one
    graph of exactly 2^16 vertices, the largest local id 65535), and a
    handle with a larger graph takes the int32 form."""
    import ctypes as C
    na = _na()

    def device_csr(bg, vptr, off, col):
        o = np.zeros(len(off), np.int64)
        c = np.zeros(len(col), np.int64)
        x = np.zeros(len(col), np.float64)
        assert na.lib.nido_get_graph(bg._h, o.ctypes.data_as(C.POINTER(C.c_int64)),
                                     c.ctypes.data_as(C.POINTER(C.c_int64)),
                                     x.ctypes.data_as(C.POINTER(C.c_double))) == 0
        return o, c, x

    small = [na.gen_sbm(n, 64, 12 / 63, 2 / (n - 64), True, 70 + n).arrays() for n in (300, 65536, 1000)]
    large = small + [na.gen_sbm(70000, 64, 6 / 63, 1 / 70000, True, 77).arrays()]
    na.set_upload_staging(4096, 16384)
    try:
        for name, parts in (("all graphs <= 2^16 vertices", small), ("one graph above 2^16", large)):
            print(f"case: batched upload, {name}", flush=True)
            vptr, off, col, w = na.concat_graphs(parts)
            bg = na.BatchGPU(csr=(vptr, off, col, w), nbatches=4)
            o2, c2, w2 = device_csr(bg, vptr, off, col)
            assert np.array_equal(o2, off) and np.array_equal(c2, col) and np.array_equal(w2, w)
            bg.close()
    finally:
        na.set_upload_staging(0, 0)


def test_pinned_staged_download_of_clusterings():
    """A batched run's clusterings come back through the two pinned buffers
    (graph-local ids made on the device): forced with 16 KiB chunks, dozens
    of hand-offs for 64 graphs, and the one-graph form of the batched API."""
    from test_gpu_batch import _check
    na = _na()
    rng = np.random.default_rng(11)
    graphs = []
    for i in range(64):
        n = int(rng.integers(200, 1500))
        graphs.append(na.gen_sbm(n, 64, 12 / 63, 2 / max(n - 64, 1), True, 3000 + i).arrays())
    na.set_upload_staging(4096, 16384)
    try:
        print("case: 64 sbms nb=8, clusterings through 16 KiB pinned chunks", flush=True)
        _check(na, graphs, 8)
        print("case: batch of one graph, clustering through 16 KiB pinned chunks", flush=True)
        _check(na, [na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 9).arrays()], 8)
    finally:
        na.set_upload_staging(0, 0)


def test_checks_made_while_staging_match_the_host_checks():
    """Inputs that take the pinned staging path are checked while they pass
    through it (arc columns against their graph, weights finite and below
    2^62, a bound on each graph's total); forced at small sizes, every input
    the host check refuses is refused with the same message, and inputs
    that need its exact totals (negative weights, totals near 2^62) still
    get them."""
    na = _na()
    off = np.array([0, 1, 3, 4])
    col = np.array([1, 0, 2, 1])
    G = 100
    vptr100 = np.arange(0, 2 * G + 1, 2)
    off100 = np.arange(0, 2 * G + 1)
    col100 = np.arange(2 * G) ^ 1
    bad_col100 = col100.copy()
    bad_col100[2 * 80] = 2 * 80 + 5
    bad_w100 = np.ones(2 * G)
    bad_w100[2 * 70 + 1] = np.inf
    vptr2 = np.array([0, 3, 6])
    off2 = np.array([0, 1, 3, 4, 5, 7, 8])
    col2 = np.array([1, 0, 2, 1, 4, 3, 5, 4])
    cases = {
        "column out of range": lambda: na.GraphGPU((off, np.array([1, 0, 7, 1]), np.ones(4)), 2),
        "NaN weight": lambda: na.GraphGPU((off, col, np.array([1.0, np.nan, 1.0, 1.0])), 2),
        "weight 2^62": lambda: na.GraphGPU((off, col, np.array([1.0, 2.0 ** 62, 1.0, 1.0])), 2),
        "total 2^62": lambda: na.GraphGPU((off, col, np.full(4, 3e18)), 2),
        "cancelling negatives": lambda: na.GraphGPU(
            (off, col, np.array([1.0, 1.0, -1.0 + 2.0 ** -40, -1.0 + 2.0 ** -40])), 2),
        "graph total 2^62": lambda: na.BatchGPU(
            csr=(vptr2, off2, col2, np.concatenate([np.ones(4), np.full(4, 3e18)])), nbatches=2),
        "arc leaves its graph": lambda: na.BatchGPU(
            csr=(vptr2, off2, np.array([1, 0, 2, 1, 4, 3, 5, 0]), np.ones(8)), nbatches=2),
        "first bad graph weight": lambda: na.BatchGPU(csr=(vptr100, off100, bad_col100, bad_w100), nbatches=2),
        "first bad graph column": lambda: na.BatchGPU(
            csr=(vptr100, off100, bad_col100, np.ones(2 * G)), nbatches=2),
    }

    def refusal(fn):
        try:
            h = fn()
        except na.NidoError as e:
            return str(e)
        h.close()
        return None

    expected = {k: refusal(f) for k, f in cases.items()}
    assert all(expected.values()), expected
    na.set_upload_staging(16, 4096)   # weights of two or more arcs take the staging path
    try:
        for k, f in cases.items():
            print(f"case: staged checks, {k}", flush=True)
            assert refusal(f) == expected[k], k
        # negative weights that pass: the exact check runs after the upload
        mild = np.array([1.0, 1.0, -0.5, -0.5])
        print("case: staged checks, mild negative weights", flush=True)
        assert_runs_equal(run_oracle(off, col, mild, 2), run_gpu(off, col, mild, 2))
    finally:
        na.set_upload_staging(0, 0)


def test_tuning_switches_reject_values_out_of_range():
    na = _na()
    off, col, w = na.gen_sbm(500, 50, 0.2, 0.01, True, 3).arrays()
    g = na.GraphGPU((off, col, w), 2)
    for knob, bad in (("split_small", 2), ("overlap_long", 2), ("full_modularity", 3)):
        with pytest.raises(na.NidoError):
            g.set_tuning(**{knob: bad})
    g.set_tuning(split_small=1, overlap_long=0, full_modularity=2)
    assert g.get_tuning()["split_small"] == 1 and g.get_tuning()["full_modularity"] == 2
    g.close()
