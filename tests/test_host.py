"""Host-side parts of libnido_amd.so that need no GPU: the C ABI exports,
the reference binary format (graph.cpp:322-495), the seeded generators, and
the argument checks of the device entry points."""
import os
import re
import subprocess

import numpy as np
import pytest

import nido_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    with open(os.path.join(ROOT, "include", "nido_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\**\s*(nido_[a-z_0-9]+)\s*\(", text, re.M):
        names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", na.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    for n in names:
        getattr(na.lib, n)


def test_abi_version():
    assert na.lib.nido_abi_version() == na.ABI_VERSION == 3


def write_reference_file(path, off, col, w, base=0, ne_header=None):
    """Writes the reference format with numpy: header, offsets (optionally
    shifted by `base`), then {int64 tail, double weight} records."""
    nv = len(off) - 1
    rec = np.zeros(len(col), dtype=[("t", "<i8"), ("w", "<f8")])
    rec["t"], rec["w"] = col, w
    with open(path, "wb") as f:
        np.array([nv, len(col) if ne_header is None else ne_header], "<i8").tofile(f)
        (np.asarray(off, np.int64) + base).astype("<i8").tofile(f)
        rec.tofile(f)


def test_reader_matches_numpy_writer_and_roundtrips(tmp_path):
    off = np.array([0, 2, 3, 5, 6])
    col = np.array([1, 2, 0, 0, 3, 2])
    w = np.array([0.5, 1.5, 0.5, 1.5, 2.0, 2.0])
    p = tmp_path / "g.bin"
    write_reference_file(p, off, col, w)
    g = na.Graph.read(str(p))
    o2, c2, w2 = g.arrays()
    assert g.num_vertices == 4 and g.num_edges == 6
    assert list(o2) == list(off) and list(c2) == list(col) and list(w2) == list(w)
    q = tmp_path / "g2.bin"
    g.write(str(q))
    assert p.read_bytes() == q.read_bytes()


def test_reader_rebases_offsets_and_trusts_them_over_the_header(tmp_path):
    # graph.cpp:376-382 and 470-472: the arc count comes from the offsets and
    # offsets are rebased to start at 0
    off = np.array([0, 1, 2])
    p = tmp_path / "g.bin"
    write_reference_file(p, off, np.array([1, 0]), np.array([1.0, 1.0]), base=0, ne_header=99)
    g = na.Graph.read(str(p))
    assert g.num_edges == 2
    o2, _, _ = g.arrays()
    assert list(o2) == [0, 1, 2]


def test_reader_errors(tmp_path):
    with pytest.raises(na.NidoError):
        na.Graph.read(str(tmp_path / "missing.bin"))
    p = tmp_path / "short.bin"
    p.write_bytes(np.array([10, 5], "<i8").tobytes())
    with pytest.raises(na.NidoError):
        na.Graph.read(str(p))


def test_reader_refuses_a_header_no_vector_can_hold(tmp_path):
    # a corrupted header asking for 2^62 vertices: an error, not an abort
    p = tmp_path / "huge.bin"
    p.write_bytes(np.array([1 << 62, 5], "<i8").tobytes())
    with pytest.raises(na.NidoError):
        na.Graph.read(str(p))


def check_simple_symmetric(off, col, w):
    nv = len(off) - 1
    assert off[0] == 0 and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(nv), np.diff(off))
    assert np.all(src != col), "self-loop"
    for v in range(min(nv, 2000)):
        row = col[off[v]:off[v + 1]]
        assert np.all(np.diff(row) > 0), "unsorted or duplicate row"
    fwd = {(int(a), int(b)): float(x) for a, b, x in zip(src, col, w)}
    for (a, b), x in list(fwd.items())[:50000]:
        assert fwd[(b, a)] == x


def test_sbm_generator():
    g = na.gen_sbm(1024, 64, 0.30, 0.005, False, 1)
    off, col, w = g.arrays()
    check_simple_symmetric(off, col, w)
    assert np.all(w == 1.0)
    exp = 2 * (16 * 64 * 63 / 2 * 0.30 + 120 * 64 * 64 * 0.005)
    assert abs(len(col) - exp) < 0.05 * exp
    src = np.repeat(np.arange(1024), np.diff(off))
    intra = np.mean(src // 64 == col // 64)
    exp_intra = (16 * 64 * 63 * 0.30) / exp          # ~0.797
    assert abs(intra - exp_intra) < 0.02
    off2, col2, _ = na.gen_sbm(1024, 64, 0.30, 0.005, False, 1).arrays()
    assert np.array_equal(off, off2) and np.array_equal(col, col2)       # seeded
    off3, _, _ = na.gen_sbm(1024, 64, 0.30, 0.005, False, 2).arrays()
    assert not np.array_equal(off, off3)


def test_weighted_generators_use_the_unit_interval_grid():
    off, col, w = na.gen_sbm(500, 50, 0.2, 0.01, True, 3).arrays()
    check_simple_symmetric(off, col, w)
    assert np.all((w > 0) & (w <= 1))
    assert np.all(w * 2.0 ** 53 == np.round(w * 2.0 ** 53))


def test_rmat_generator():
    off, col, w = na.gen_rmat(12, 8, 0.57, 0.19, 0.19, True, 7).arrays()
    check_simple_symmetric(off, col, w)
    assert len(off) == 4097
    assert 0.6 * 2 * 8 * 4096 < len(col) <= 2 * 8 * 4096
    deg = np.diff(off)
    assert deg[:64].mean() > 4 * deg.mean()          # skew towards low ids


def test_chung_lu_generator():
    off, col, w = na.gen_chung_lu(20000, 2.1, 2.0, 500.0, 16.0, False, 11).arrays()
    check_simple_symmetric(off, col, w)
    mean = len(col) / 20000
    assert 10 < mean < 17
    assert np.diff(off).max() > 100


def test_generator_argument_errors():
    with pytest.raises(na.NidoError):
        na.gen_sbm(-1, 64, 0.3, 0.1, False, 1)
    with pytest.raises(na.NidoError):
        na.gen_rmat(40, 16, 0.57, 0.19, 0.19, False, 1)


def test_device_entry_points_validate_before_touching_the_gpu():
    # rejected on the host: out-of-range column, NaN weight, zero batches
    off = np.array([0, 1, 2])
    with pytest.raises(na.NidoError) as e:
        na.GraphGPU((off, np.array([1, 5]), np.array([1.0, 1.0])), 2)
    assert e.value.code == -1
    with pytest.raises(na.NidoError):
        na.GraphGPU((off, np.array([1, 0]), np.array([1.0, np.nan])), 2)
    with pytest.raises(na.NidoError):
        na.GraphGPU((off, np.array([1, 0]), np.array([1.0, 1.0])), 0)


def test_process_wide_settings_validate_their_arguments():
    import ctypes as C
    with pytest.raises(na.NidoError):
        na.set_list_rebuild(101)
    with pytest.raises(na.NidoError):
        na.set_device_lists(-1)
    na.set_list_rebuild(0)
    na.set_list_rebuild(-1)    # back to the default
    n = C.c_int64()
    assert na.lib.nido_batch_list_rebuilds(None, C.byref(n)) == -1
    assert b"null argument" in na.lib.nido_last_error()


def test_weight_totals_that_would_wrap_the_exact_sums_are_refused():
    """Each |w| below 2^62 is not enough: the Q64.64 sums (2m, k_v, a_c) need
    the total of |w| below 2^62 too.  The path 0-1-2 with two edges of
    weight 3e18 (2m = 1.2e19) used to wrap to a negative mass; both the
    library and the oracle now refuse it, before any device work.  With
    negative weights, sum|w| / |2m| must also stay below 2^31 (square terms
    of the modularity)."""
    import oracle
    off = np.array([0, 1, 3, 4])
    col = np.array([1, 0, 2, 1])
    big = np.full(4, 3e18)
    with pytest.raises(na.NidoError) as e:
        na.GraphGPU((off, col, big), 2)
    assert e.value.code == -1 and "2^62" in str(e.value)
    with pytest.raises(ValueError):
        oracle.Oracle(off, col, big, 2)
    # just below the bound on both sides: accepted by the oracle, and the
    # library gets past its checks (it then needs a device)
    ok = np.full(4, 2.0 ** 60 - 2.0 ** 8)
    o = oracle.Oracle(off, col, ok, 2)
    assert o.mass == 2.0 * ok[0] * 2 / 2
    o.close()
    # negative weights that nearly cancel: |2m| tiny against sum|w|
    neg = np.array([1.0, 1.0, -1.0 + 2.0 ** -40, -1.0 + 2.0 ** -40])
    with pytest.raises(na.NidoError) as e:
        na.GraphGPU((off, col, neg), 2)
    assert "negative" in str(e.value)
    with pytest.raises(ValueError):
        oracle.Oracle(off, col, neg, 2)
    # the same shape with a moderate ratio is accepted by the oracle
    mild = np.array([1.0, 1.0, -0.5, -0.5])
    oracle.Oracle(off, col, mild, 2).close()
    # per graph in a batched handle: graph 1 alone exceeds the bound
    vptr = np.array([0, 3, 6])
    off2 = np.array([0, 1, 3, 4, 5, 7, 8])
    col2 = np.array([1, 0, 2, 1, 4, 3, 5, 4])
    w2 = np.concatenate([np.ones(4), big])
    with pytest.raises(na.NidoError) as e:
        na.BatchGPU(csr=(vptr, off2, col2, w2), nbatches=2)
    assert "graph 1" in str(e.value)


def test_tuning_flags_accept_only_their_values():
    """split_small, overlap_long and fixed_short are 0/1 switches,
    full_modularity takes 0, 1 or 2; the knob values are checked before the
    handle, so a null handle shows which check refused the struct without a
    device."""
    import ctypes as C
    base = dict(thread_max_deg=2, big_min_deg=512, big_part_arcs=1024, big_key_limit=1024, use_graph=1,
                big_slots=2048, split_small=0, overlap_long=1, full_modularity=0, fixed_short=1)
    t = na.Tuning(**base)
    assert na.lib.nido_graph_set_tuning(None, C.byref(t)) == -1
    assert b"null handle" in na.lib.nido_last_error()
    for knob in ("split_small", "overlap_long", "full_modularity", "fixed_short"):
        for bad in ((3, -1, 7) if knob == "full_modularity" else (2, -1, 7)):
            t = na.Tuning(**dict(base, **{knob: bad}))
            assert na.lib.nido_graph_set_tuning(None, C.byref(t)) == -1
            assert knob.encode() in na.lib.nido_last_error(), (knob, bad)
    t = na.Tuning(**dict(base, full_modularity=2))
    assert na.lib.nido_graph_set_tuning(None, C.byref(t)) == -1
    assert b"null handle" in na.lib.nido_last_error()


def test_batch_bounds_host_matches_oracle():
    import oracle
    off, _, _ = na.gen_rmat(10, 8, 0.57, 0.19, 0.19, False, 1).arrays()
    for nb in (1, 2, 7, 32):
        assert np.array_equal(na.batch_bounds_host(off, nb), oracle.batch_bounds(off, nb))


def test_batch_create_validates_on_the_host():
    # arcs must stay inside their graph's vertex range; checked before any device work
    vptr = np.array([0, 2, 4])
    off = np.array([0, 1, 2, 3, 4])
    with pytest.raises(na.NidoError) as e:
        na.BatchGPU(csr=(vptr, off, np.array([1, 2, 3, 2]), np.ones(4)), nbatches=2)
    assert e.value.code == -1
    with pytest.raises(na.NidoError):
        na.BatchGPU(csr=(np.array([0, 3, 2]), np.array([0, 0, 0, 0]), np.zeros(0, np.int64), np.zeros(0)),
                    nbatches=2)


def test_batch_validation_names_the_first_bad_graph_among_many():
    """With 64+ graphs the checks run one graph per task: the error still
    names the first graph (in order) with a bad arc or weight."""
    G = 100
    vptr = np.arange(0, 2 * G + 1, 2)
    off = np.arange(0, 2 * G + 1)
    col = np.arange(2 * G) ^ 1                  # each graph: one edge, both directions
    w = np.ones(2 * G)
    bad_col, bad_w = col.copy(), w.copy()
    bad_col[2 * 80] = 2 * 80 + 5                # graph 80 points outside itself
    bad_w[2 * 70 + 1] = np.inf                  # graph 70 has a weight that is not finite
    with pytest.raises(na.NidoError) as e:
        na.BatchGPU(csr=(vptr, off, bad_col, bad_w), nbatches=2)
    assert "weight" in str(e.value)
    with pytest.raises(na.NidoError) as e:
        na.BatchGPU(csr=(vptr, off, bad_col, w), nbatches=2)
    assert "graph 80" in str(e.value)


def test_concat_graphs_layout():
    g1 = (np.array([0, 1, 2]), np.array([1, 0]), np.array([0.5, 0.5]))
    g2 = (np.array([0, 2, 3, 4]), np.array([1, 2, 0, 0]), np.array([1.0, 2.0, 1.0, 2.0]))
    vptr, off, col, w = na.concat_graphs([g1, g2])
    assert list(vptr) == [0, 2, 5]
    assert list(off) == [0, 1, 2, 4, 5, 6]
    assert list(col) == [1, 0, 3, 4, 2, 2]
    assert list(w) == [0.5, 0.5, 1.0, 2.0, 1.0, 2.0]


def _lib_fx(w):
    import ctypes as C
    lo, hi = C.c_uint64(), C.c_int64()
    na.lib.nido_fx_from_double(C.c_double(w), C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


def test_device_conversions_match_the_oracle():
    """fixed.h (the conversions the kernels run, built for the host) against
    the oracle's independent restatement, itself pinned by exact rational
    arithmetic in test_oracle.py: weights of every magnitude and sign,
    subnormals, exact ties at 2^-65, and 128-bit values spanning all widths."""
    import ctypes as C
    import oracle
    na.lib.nido_fx_from_double.argtypes = [C.c_double, C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]
    na.lib.nido_fx_to_double.argtypes = [C.c_uint64, C.c_int64]
    na.lib.nido_fx_to_double.restype = C.c_double
    rng = np.random.default_rng(3)
    ws = [0.0, -0.0, 1.0, -1.0, 2.0 ** -64, 2.0 ** -65, 3 * 2.0 ** -65, 5 * 2.0 ** -65, -(3 * 2.0 ** -65),
          2.0 ** -1074, 2.0 ** -1022, 2.0 ** 61, -(2.0 ** 61), np.nextafter(2.0 ** 62, 0), 1.0 - 2.0 ** -53]
    ws += list(rng.uniform(-1, 1, 20000) * 2.0 ** rng.integers(-80, 62, 20000))
    ws += list(np.ldexp(rng.integers(1, 1 << 20, 5000), rng.integers(-90, -60, 5000)) + 0.0)
    for w in ws:
        got = _lib_fx(float(w))
        got = got - (1 << 128) if got >= (1 << 127) else got
        assert got == oracle.fx_from_double(float(w)), w
    xs = [0, 1, -1, (1 << 64), -(1 << 64), (1 << 127) - 1, -(1 << 127), (1 << 53) + 1, 3 << 70]
    for bits in rng.integers(1, 127, 20000):
        x = int(rng.integers(0, 1 << 62)) << max(0, int(bits) - 62) | int(rng.integers(0, 1 << 62))
        x &= (1 << int(bits)) - 1
        xs.append(x if rng.random() < 0.5 else -x)
    for x in xs:
        lo, hi = x & ((1 << 64) - 1), x >> 64
        assert na.lib.nido_fx_to_double(lo, hi) == oracle.fx_to_double(x), x


def test_reciprocal_division_is_ieee_division():
    """div_rcp (fixed.h, built for the host), which replaces both divisions
    of the gain formulas, against IEEE division inside its stated range:
    random quotients of every magnitude, quotients of Q64.64-grid values as
    the gains produce them, quotients near a rounding midpoint (and one ulp
    either side), and zero numerators."""
    import ctypes as C
    f = na.lib.nido_div_rcp
    f.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]

    def check(a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty_like(a)
        f(a.size, a.ctypes.data, b.ctypes.data, out.ctypes.data)
        same = out.view(np.int64) == (a / b).view(np.int64)
        assert same.all(), (a[~same][:4], b[~same][:4])

    rng = np.random.default_rng(11)
    n = 400_000
    check(rng.uniform(-2, 2, n), rng.uniform(0.1, 3, n))
    check(np.exp2(rng.uniform(-399, 399, n)) * rng.choice([-1.0, 1.0], n), np.exp2(rng.uniform(-398, 398, n)))
    check(rng.integers(-10 ** 6, 10 ** 6, n).astype(float), rng.integers(1, 10 ** 7, n) / 2.0)
    # e - k a / (2m) and its quotient by m with Q64.64-grid sums (multiples of 2^-64)
    m = np.ldexp(rng.integers(1, 1 << 53, n).astype(float), rng.integers(-64, 40, n))
    k = np.ldexp(rng.integers(1, 1 << 53, n).astype(float), rng.integers(-64, 30, n))
    a = np.ldexp(rng.integers(1, 1 << 53, n).astype(float), rng.integers(-64, 30, n))
    check(k * a, 2 * m)
    check(k - (k * a) / (2 * m), m)
    b = rng.uniform(1, 2, n)
    q = rng.uniform(1, 2, n)
    mid = b * (q + np.spacing(q) / 2)
    for x in (mid, np.nextafter(mid, np.inf), np.nextafter(mid, -np.inf)):
        check(x, b)
    check([0.0, 0.0], [3.0, 2.0 ** -65])
