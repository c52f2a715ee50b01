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
