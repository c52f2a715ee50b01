"""nido_modularity_noalloc against nido_compute_modularity.

The allocation-free entry point takes the CSR and the partition as device
arrays (allocated here with hipMalloc from the HIP runtime the library is
linked against, only as the caller's memory) and must
give the same modularity, bit for bit, as the handle's own evaluation of the
same partition: the partition after one Louvain iteration (and the singleton
partition) of a weighted R-MAT graph, then of a smaller graph on the next
call, which finds the larger call's sums still in the shared working memory.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libnido_amd.so loaded (the same library and device
    context), with the few calls the test needs as the caller."""
    import ctypes as C
    path = next((ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln), None)
    assert path, "libamdhip64 is not loaded"
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _setup():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na, _hip()


def _reference(na, off, col, w, nb):
    """Q from the handle for the singleton partition and after one iteration,
    with the iteration's partition."""
    g = na.GraphGPU((off, col, w), nb)
    g.singleton_partition()
    q0 = g.compute_modularity()
    q, dq, qe = g.louvain_iteration(q0)
    q1 = g.compute_modularity()
    C = g.communities()
    g.close()
    return q0, q1, C


def _noalloc(na, hip, off, col, w, C):
    import ctypes as Ct
    host = [np.ascontiguousarray(x) for x in (off, col, w, C, np.array([-7.0]))]
    dev = []
    try:
        for x in host:
            p = Ct.c_void_p()
            assert hip.hipMalloc(Ct.byref(p), max(x.nbytes, 8)) == 0
            dev.append(p.value)
            assert hip.hipMemcpy(p.value, x.ctypes.data, x.nbytes, 1) == 0        # host to device
        assert hip.hipDeviceSynchronize() == 0
        na.modularity_noalloc(len(off) - 1, len(col), dev[0], dev[1], dev[2], dev[3], dev[4], 0)
        assert hip.hipDeviceSynchronize() == 0
        q = np.zeros(1)
        assert hip.hipMemcpy(q.ctypes.data, dev[4], 8, 2) == 0                    # device to host
        return float(q[0])
    finally:
        for p in dev:
            hip.hipFree(p)


def test_noalloc_matches_compute_modularity_on_two_calls():
    na, hip = _setup()
    big = na.gen_rmat(14, 16, 0.57, 0.19, 0.19, True, 7).arrays()
    small = na.gen_sbm(1000, 50, 0.3, 0.01, True, 3).arrays()
    for name, (off, col, w), nb in (("rmat s14 weighted", big, 4), ("sbm 1000 weighted", small, 2)):
        print(f"case: {name}: {len(off) - 1} vertices, {len(col)} arcs", flush=True)
        q0, q1, C = _reference(na, off, col, w, nb)
        assert len(np.unique(C)) < len(C)               # the iteration moved vertices
        got1 = _noalloc(na, hip, off, col, w, C)
        assert got1 == q1, (got1, q1)
        ident = np.arange(len(off) - 1, dtype=np.int64)
        got0 = _noalloc(na, hip, off, col, w, ident)
        assert got0 == q0, (got0, q0)


def test_noalloc_refuses_and_flags_bad_input():
    na, hip = _setup()
    with pytest.raises(na.NidoError):
        na.modularity_noalloc(na.NOALLOC_MAX_VERTICES + 1, 0, 0, 0, 0, 0, 0, 0)   # refused on the host
    off, col, w = na.gen_sbm(256, 64, 0.3, 0.01, False, 5).arrays()
    C = np.zeros(len(off) - 1, np.int64)
    C[7] = len(off) - 1                                  # one id outside [0, nv): flagged on the device
    assert np.isnan(_noalloc(na, hip, off, col, w, C))
    C[7] = 0
    ok = _noalloc(na, hip, off, col, w, C)             # every vertex in one community: Q = 1 - 1 = 0
    assert ok == 0.0
