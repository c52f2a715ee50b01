"""The -c colouring on the device and Louvain on the renumbered graph.

nido_hgraph_coloring (random weights on the host, the greedy colouring on the
device, renumbering) against the oracle's colouring loop and renumbering fed
the random weights of the Python restatement of the reference's engines;
then Louvain on the renumbered graph against the oracle, and the community
file written through the new order (main.cpp:162-163, clustering.cpp:25-37)
against the oracle's clustering mapped back by hand.  Also the CLI with
-r (the BA generator) and -c.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import stdlib_rng as R
from parity import assert_runs_equal, run_gpu, run_oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _na():
    import nido_amd as na
    assert na.device_count() >= 1, "no HIP device visible"
    return na


def _graphs(na):
    yield "sbm 1024 fixture", na.Graph.read(os.path.join(ROOT, "tests", "golden", "sbm1024_seed1.bin")), 1
    yield "rmat s12 weighted", na.gen_rmat(12, 16, 0.57, 0.19, 0.19, True, 3), 1
    yield "ba 3000 x 3", na.gen_ba(3000, 3), 1
    yield "rmat s10, 3 engines", na.gen_rmat(10, 8, 0.57, 0.19, 0.19, False, 4), 3


@pytest.mark.parametrize("case", range(4))
def test_device_coloring_matches_oracle(case):
    na = _na()
    name, g, engines = list(_graphs(na))[case]
    off, col, w = g.arrays()
    nv = len(off) - 1
    print(f"case: {name}: {nv} vertices, {len(col)} arcs, {engines} engine(s)", flush=True)
    r = np.array(R.coloring_weights(nv, engines), np.int64)
    c_ref, K_ref = oracle.coloring(off, col, r)
    no, colors, K = g.coloring(engines)
    assert K == K_ref
    assert np.array_equal(colors, c_ref)
    no_ref, noff, ncol, nw = oracle.reorder(off, col, w, c_ref)
    assert np.array_equal(no, no_ref)
    off2, col2, w2 = g.arrays()
    assert np.array_equal(off2, noff) and np.array_equal(col2, ncol) and np.array_equal(w2, nw)


def test_isolated_vertices_and_tiny_graphs():
    na = _na()
    off = np.array([0, 1, 2, 2, 2, 3, 4], np.int64)     # 0-1, 4-5, two isolated vertices
    col = np.array([1, 0, 5, 4], np.int64)
    g = na.Graph.from_csr(off, col, np.ones(4))
    no, colors, K = g.coloring()
    r = np.array(R.coloring_weights(6), np.int64)
    c_ref, K_ref = oracle.coloring(off, col, r)
    assert K == K_ref and np.array_equal(colors, c_ref)


def test_louvain_on_colored_graph_and_mapped_dump(tmp_path):
    na = _na()
    g = na.gen_rmat(13, 16, 0.57, 0.19, 0.19, True, 9)
    off0, col0, w0 = g.arrays()
    no, colors, K = g.coloring()
    off, col, w = g.arrays()
    print(f"case: rmat s13 colored, {K} colours", flush=True)
    got = run_gpu(off, col, w, 4)
    ref = run_oracle(off, col, w, 4)
    assert_runs_equal(ref, got)
    gg = na.GraphGPU((off, col, w), 4)
    na.LouvainGPU(500, 1e-6, 4).run(gg)
    p = tmp_path / "comm.txt"
    gg.dump_partition(str(p), no)
    gg.close()
    lines = np.loadtxt(p, dtype=np.int64)
    want = np.empty(len(no), np.int64)
    want[no] = ref["clustering"]            # line of old vertex no[i] = community of new vertex i
    assert np.array_equal(lines, want)


def test_cli_random_graph_colored(tmp_path):
    na = _na()
    out = tmp_path / "c.txt"
    exe = os.path.join(ROOT, "nido_amd", "nido_louvain")
    print("case: nido_louvain -r '4000 4' -c -b 3", flush=True)
    r = subprocess.run([exe, "-r", "4000 4", "-c", "-b", "3", "-o", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Coloring Time" in r.stdout
    g = na.gen_ba(4000, 4)
    no, _, _ = g.coloring()
    off, col, w = g.arrays()
    ref = run_oracle(off, col, w, 3)
    want = np.empty(len(no), np.int64)
    want[no] = ref["clustering"]
    assert np.array_equal(np.loadtxt(out, dtype=np.int64), want)
