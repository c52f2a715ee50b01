"""nido_amd -- MI355X-native batched Louvain (local moving + coarsening).

Python view of the C ABI in include/nido_amd.h (libnido_amd.so, built from
nido_amd/csrc by `make -C nido_amd`).  The classes mirror the reference's
host API: `Graph` (graph.hpp:14-102), `GraphGPU` (graph_gpu.hpp:168-247) and
`LouvainGPU` (louvain_gpu.hpp:15-23).  Every call goes to the HIP library;
there is no CPU fallback, and importing this module fails loudly when the
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

__all__ = [
    "lib", "LIB_PATH", "NidoError", "set_upload_staging", "Graph", "GraphGPU", "LouvainGPU", "PhaseStats", "RunStats",
    "gen_sbm", "gen_rmat", "gen_chung_lu", "gen_ba", "device_count", "batch_bounds_host", "modularity_noalloc",
    "DEFAULT_BATCHES", "DEFAULT_ITERATIONS", "DEFAULT_THRESHOLD", "MAX_PHASES",
]

DEFAULT_BATCHES = 2          # types.hpp:32-34
DEFAULT_ITERATIONS = 500     # types.hpp:28-30
DEFAULT_THRESHOLD = 1.0e-6   # types.hpp:36-38
MAX_PHASES = 200             # types.hpp:24-26

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnido_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (hipcc, gfx950). "
        "nido_amd has no CPU fallback.")

lib = C.CDLL(LIB_PATH)

_i64 = C.c_int64
_pi64 = C.POINTER(C.c_int64)
_pdbl = C.POINTER(C.c_double)


class PhaseStats(C.Structure):
    _fields_ = [("nv", C.c_int64), ("ne", C.c_int64), ("loops", C.c_int64), ("q0", C.c_double),
                ("q", C.c_double), ("t_loop", C.c_double), ("t_agg", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunStats(C.Structure):
    _fields_ = [("phases", C.c_int32), ("done", C.c_int32), ("total_loops", C.c_int64),
                ("q", C.c_double), ("t_total", C.c_double), ("t_loop", C.c_double),
                ("t_agg", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("thread_max_deg", C.c_int32), ("big_min_deg", C.c_int32),
                ("big_part_arcs", C.c_int32), ("big_key_limit", C.c_int32),
                ("use_graph", C.c_int32), ("big_slots", C.c_int32), ("split_small", C.c_int32),
                ("overlap_long", C.c_int32), ("full_modularity", C.c_int32), ("fixed_short", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("nido_last_error", C.c_char_p)
_sig("nido_abi_version", C.c_int)
ABI_VERSION = 3   # the struct layouts below
if lib.nido_abi_version() != ABI_VERSION:
    raise ImportError(f"{LIB_PATH} has ABI {lib.nido_abi_version()}, this binding needs {ABI_VERSION}: "
                      f"rebuild it with `make -C {_HERE}`")
_sig("nido_device_count", C.c_int, C.POINTER(C.c_int))
_sig("nido_set_upload_staging", C.c_int, _i64, _i64)
_sig("nido_set_device_lists", C.c_int, _i64)
_sig("nido_set_list_rebuild", C.c_int, C.c_int)
_sig("nido_hgraph_read", C.c_int, C.c_char_p, C.POINTER(C.c_void_p))
_sig("nido_hgraph_write", C.c_int, C.c_void_p, C.c_char_p)
_sig("nido_hgraph_from_csr", C.c_int, _i64, _i64, _pi64, _pi64, _pdbl, C.POINTER(C.c_void_p))
_sig("nido_hgraph_info", C.c_int, C.c_void_p, _pi64, _pi64)
_sig("nido_hgraph_arrays", C.c_int, C.c_void_p, C.POINTER(_pi64), C.POINTER(_pi64), C.POINTER(_pdbl))
_sig("nido_hgraph_destroy", None, C.c_void_p)
_sig("nido_gen_sbm", C.c_int, _i64, _i64, C.c_double, C.c_double, C.c_int, C.c_uint64,
     C.POINTER(C.c_void_p))
_sig("nido_gen_rmat", C.c_int, C.c_int, _i64, C.c_double, C.c_double, C.c_double, C.c_int,
     C.c_uint64, C.POINTER(C.c_void_p))
_sig("nido_gen_chung_lu", C.c_int, _i64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
     C.c_uint64, C.POINTER(C.c_void_p))
_sig("nido_gen_ba", C.c_int, _i64, _i64, C.POINTER(C.c_void_p))
_sig("nido_hgraph_coloring", C.c_int, C.c_void_p, C.c_int, _pi64, _pi64, _pi64)
_sig("nido_dump_partition_orders", C.c_int, C.c_void_p, C.c_char_p, _pi64)
_sig("nido_graph_create", C.c_int, _i64, _i64, _pi64, _pi64, _pdbl, C.c_int, C.c_void_p,
     C.POINTER(C.c_void_p))
_sig("nido_graph_create_windowed", C.c_int, _i64, _i64, _pi64, _pi64, _pdbl, C.c_int, _i64, C.c_void_p,
     C.POINTER(C.c_void_p))
_sig("nido_graph_destroy", C.c_int, C.c_void_p)
_sig("nido_graph_info", C.c_int, C.c_void_p, _pi64, _pi64, _pi64, _pdbl)
_sig("nido_graph_batch_bounds", C.c_int, C.c_void_p, _pi64)
_sig("nido_graph_set_tuning", C.c_int, C.c_void_p, C.POINTER(Tuning))
_sig("nido_graph_get_tuning", C.c_int, C.c_void_p, C.POINTER(Tuning))
_sig("nido_graph_sync", C.c_int, C.c_void_p)
_sig("nido_singleton_partition", C.c_int, C.c_void_p)
_sig("nido_compute_modularity", C.c_int, C.c_void_p, _pdbl)
_sig("nido_louvain_update", C.c_int, C.c_void_p, C.c_int)
_sig("nido_update_community_weights", C.c_int, C.c_void_p, C.c_int)
_sig("nido_update_community_ids", C.c_int, C.c_void_p, C.c_int)
_sig("nido_restore_community", C.c_int, C.c_void_p)
_sig("nido_aggregation", C.c_int, C.c_void_p, C.POINTER(C.c_int))
_sig("nido_get_communities", C.c_int, C.c_void_p, _pi64)
_sig("nido_get_new_communities", C.c_int, C.c_void_p, _pi64)
_sig("nido_get_vertex_weights", C.c_int, C.c_void_p, _pdbl)
_sig("nido_get_community_weights", C.c_int, C.c_void_p, _pdbl)
_sig("nido_get_graph", C.c_int, C.c_void_p, _pi64, _pi64, _pdbl)
_sig("nido_get_clustering", C.c_int, C.c_void_p, _pi64)
_sig("nido_dump_partition", C.c_int, C.c_void_p, C.c_char_p)
_sig("nido_louvain_iteration", C.c_int, C.c_void_p, _pdbl, _pdbl, _pdbl)
_sig("nido_modularity_noalloc", C.c_int, _i64, _i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
     C.c_void_p, C.c_void_p)
NOALLOC_MAX_VERTICES = 1 << 22   # NIDO_NOALLOC_MAX_VERTICES
_sig("nido_graph_set_profiling", C.c_int, C.c_void_p, C.c_int)
_sig("nido_graph_set_profiling_classes", C.c_int, C.c_void_p, C.c_uint)
_sig("nido_graph_set_profiling_stride", C.c_int, C.c_void_p, C.c_int)
_sig("nido_graph_kernel_stats", C.c_int, C.c_void_p, C.c_int, _pi64, _pdbl, _pdbl)
_sig("nido_louvain_run", C.c_int, C.c_void_p, _i64, C.c_double, C.c_double, C.c_int, C.c_int,
     C.POINTER(RunStats))

KERNEL_CLASSES = ("lm_short", "lm_long", "commit", "modularity")   # NIDO_K_* order
_sig("nido_run_phase", C.c_int, C.c_void_p, C.c_int, C.POINTER(PhaseStats))
_sig("nido_run_iteration_q", C.c_int, C.c_void_p, C.c_int, _pdbl)
_sig("nido_run_phase_communities", C.c_int, C.c_void_p, C.c_int, _pi64)
_sig("nido_autotune", C.c_int, C.c_void_p, C.c_int, C.POINTER(Tuning))
_sig("nido_batch_create", C.c_int, C.c_int, _pi64, _pi64, _pi64, _pdbl, C.c_int, C.c_void_p,
     C.POINTER(C.c_void_p))
_sig("nido_batch_run", C.c_int, C.c_void_p, _i64, C.c_double, C.c_double, C.POINTER(RunStats))
_sig("nido_batch_size", C.c_int, C.c_void_p, C.POINTER(C.c_int))
_sig("nido_batch_list_rebuilds", C.c_int, C.c_void_p, _pi64)
_sig("nido_batch_graph_stats", C.c_int, C.c_void_p, C.c_int, C.POINTER(RunStats))
_sig("nido_batch_graph_phase", C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(PhaseStats))
_sig("nido_batch_clustering", C.c_int, C.c_void_p, C.c_int, _pi64)


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("nbatches", C.c_int32), ("hbm_budget", C.c_int64),
                ("max_loops", C.c_int64), ("tol", C.c_double), ("tol_phase", C.c_double),
                ("group", C.c_int32)]


_sig("nido_serve_create", C.c_int, C.POINTER(ServeConfig), C.POINTER(C.c_void_p))
_sig("nido_serve_footprint", C.c_int64, _i64, _i64)
_sig("nido_serve_submit", C.c_int, C.c_void_p, _i64, _pi64, _pi64, _pdbl, C.c_int, _pi64)
_sig("nido_serve_wait", C.c_int, C.c_void_p, _i64, C.POINTER(RunStats), _pdbl)
_sig("nido_serve_clustering", C.c_int, C.c_void_p, _i64, _pi64)
_sig("nido_serve_release", C.c_int, C.c_void_p, _i64)
_sig("nido_serve_status", C.c_int, C.c_void_p, _pi64, _pi64, _pi64, _pi64)
_sig("nido_serve_destroy", C.c_int, C.c_void_p)


class NidoError(RuntimeError):
    """A negative status from libnido_amd, with the library's message."""

    def __init__(self, code: int, where: str):
        msg = (lib.nido_last_error() or b"").decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise NidoError(rc, where)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(_pi64)


def _pd(a: np.ndarray):
    return a.ctypes.data_as(_pdbl)


def device_count() -> int:
    n = C.c_int(0)
    _check(lib.nido_device_count(C.byref(n)), "nido_device_count")
    return n.value


def set_upload_staging(min_bytes: int = 0, chunk_bytes: int = 0) -> None:
    """Pinned staging of graph uploads from min_bytes in chunk_bytes pieces
    (0: the defaults, 256 MiB and 64 MiB); process-wide."""
    _check(lib.nido_set_upload_staging(int(min_bytes), int(chunk_bytes)), "nido_set_upload_staging")


def set_device_lists(min_vertices: int = 0) -> None:
    """Phases of at least min_vertices vertices build their degree bins and
    row lists on the device (0: the default, 2^17); process-wide."""
    _check(lib.nido_set_device_lists(int(min_vertices)), "nido_set_device_lists")


def set_list_rebuild(percent: int = -1) -> None:
    """A batched handle rebuilds its work lists for the graphs still
    iterating once they hold at most `percent` of the listed vertices (0
    never, 100 as soon as one stops; negative: the default, 50);
    process-wide."""
    _check(lib.nido_set_list_rebuild(int(percent)), "nido_set_list_rebuild")


def modularity_noalloc(nv: int, ne: int, offsets: int, cols: int, weights: int, comm: int, q: int,
                       stream: int = 0) -> None:
    """nido_modularity_noalloc: every array argument is a device address
    (int); *q (a device double) receives the modularity of partition comm
    when `stream` reaches it.  Allocates no device memory."""
    _check(lib.nido_modularity_noalloc(nv, ne, offsets, cols, weights, comm, q, stream or None),
           "nido_modularity_noalloc")


def batch_bounds_host(off: np.ndarray, nbatches: int) -> np.ndarray:
    """partition_graph_edge_batch (graph_gpu.cpp:303-339) on the host, for
    callers that want the boundaries before creating device state."""
    nv = len(off) - 1
    ave = int(off[nv] - off[0]) // nbatches
    parts = [0]
    start = 0
    for i in range(1, nv + 1):
        if off[i] - off[start] > ave:
            parts.append(i)
            start = i
        elif i == nv:
            parts.append(nv)
    if len(parts) > nbatches + 1:
        nnv = (nv - parts[nbatches]) // nbatches
        for i in range(1, nbatches):
            parts[i] += nnv
        parts[nbatches] = nv
    while len(parts) < nbatches + 1:
        parts.append(nv)
    return np.asarray(parts[: nbatches + 1], dtype=np.int64)


class _View:
    """A numpy array interface over memory `owner` holds; the array made from
    it keeps `owner` alive."""

    def __init__(self, owner, ptr, n: int, typestr: str):
        self._owner = owner
        addr = C.cast(ptr, C.c_void_p).value
        self.__array_interface__ = {"shape": (n,), "typestr": typestr, "data": (addr, False), "version": 3}


class Graph:
    """Host CSR graph (the reference's Graph, graph.hpp:14-102)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def read(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(lib.nido_hgraph_read(os.fsencode(path), C.byref(h)), "nido_hgraph_read")
        return cls(h)

    @classmethod
    def from_csr(cls, off, col, w) -> "Graph":
        off = np.ascontiguousarray(off, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        h = C.c_void_p()
        _check(lib.nido_hgraph_from_csr(len(off) - 1, len(col), _p64(off), _p64(col), _pd(w),
                                        C.byref(h)), "nido_hgraph_from_csr")
        return cls(h)

    def write(self, path: str) -> None:
        _check(lib.nido_hgraph_write(self._h, os.fsencode(path)), "nido_hgraph_write")

    @property
    def num_vertices(self) -> int:
        nv = C.c_int64()
        _check(lib.nido_hgraph_info(self._h, C.byref(nv), None), "nido_hgraph_info")
        return nv.value

    @property
    def num_edges(self) -> int:
        ne = C.c_int64()
        _check(lib.nido_hgraph_info(self._h, None, C.byref(ne)), "nido_hgraph_info")
        return ne.value

    def arrays(self, copy: bool = True):
        """(offsets, cols, weights) as numpy arrays; with copy=False, views of
        the graph's own memory that keep the graph alive."""
        nv, ne = self.num_vertices, self.num_edges
        po, pc, pw = _pi64(), _pi64(), _pdbl()
        _check(lib.nido_hgraph_arrays(self._h, C.byref(po), C.byref(pc), C.byref(pw)),
               "nido_hgraph_arrays")
        off = np.asarray(_View(self, po, nv + 1, "<i8"))
        col = np.asarray(_View(self, pc, ne, "<i8")) if ne else np.zeros(0, np.int64)
        w = np.asarray(_View(self, pw, ne, "<f8")) if ne else np.zeros(0, np.float64)
        if copy:
            return off.copy(), col.copy(), w.copy()
        return off, col, w

    def coloring(self, engines: int = 1):
        """Graph::coloring (graph.cpp:547-694), in place: returns
        (new_orders, colors, num_colors); new_orders[i] is the old id of new
        vertex i, colors[v] the colour of old vertex v."""
        nv = self.num_vertices
        no = np.empty(max(nv, 1), np.int64)
        colors = np.empty(max(nv, 1), np.int64)
        k = C.c_int64()
        _check(lib.nido_hgraph_coloring(self._h, engines, _p64(colors), _p64(no), C.byref(k)),
               "nido_hgraph_coloring")
        return no[:nv], colors[:nv], k.value

    def close(self) -> None:
        if self._h:
            lib.nido_hgraph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_sbm(nv: int, block: int, p_in: float, p_out: float, weighted: bool, seed: int) -> Graph:
    h = C.c_void_p()
    _check(lib.nido_gen_sbm(nv, block, p_in, p_out, int(weighted), seed, C.byref(h)), "nido_gen_sbm")
    return Graph(h)


def gen_rmat(scale: int, edge_factor: int, a: float, b: float, c: float, weighted: bool,
             seed: int) -> Graph:
    h = C.c_void_p()
    _check(lib.nido_gen_rmat(scale, edge_factor, a, b, c, int(weighted), seed, C.byref(h)),
           "nido_gen_rmat")
    return Graph(h)


def gen_ba(nv: int, m0: int) -> Graph:
    """The reference's -r generator (graph.cpp:76-110): Barabasi-Albert,
    nv vertices, m0 attachments per new vertex, weights |N(1, 1)|."""
    h = C.c_void_p()
    _check(lib.nido_gen_ba(nv, m0, C.byref(h)), "nido_gen_ba")
    return Graph(h)


def gen_chung_lu(nv: int, alpha: float, dmin: float, dmax: float, mean: float, weighted: bool,
                 seed: int) -> Graph:
    h = C.c_void_p()
    _check(lib.nido_gen_chung_lu(nv, alpha, dmin, dmax, mean, int(weighted), seed, C.byref(h)),
           "nido_gen_chung_lu")
    return Graph(h)


class GraphGPU:
    """Device-resident Louvain state (the reference's GraphGPU,
    graph_gpu.hpp:168-247), one graph in HBM."""

    def __init__(self, graph, nbatches: int = DEFAULT_BATCHES, stream: int | None = None, window: int = 0):
        """window > 0: the device holds at most `window` arcs at a time
        (nido_graph_create_windowed); 0: nido_graph_create."""
        if isinstance(graph, Graph):
            off, col, w = graph.arrays(copy=False)
        else:
            off, col, w = graph
            off = np.ascontiguousarray(off, dtype=np.int64)
            col = np.ascontiguousarray(col, dtype=np.int64)
            w = np.ascontiguousarray(w, dtype=np.float64)
        self._keep = (graph,)
        h = C.c_void_p()
        st = C.c_void_p(stream) if stream else None
        if window:
            _check(lib.nido_graph_create_windowed(len(off) - 1, len(col), _p64(off), _p64(col), _pd(w), nbatches,
                                                  window, st, C.byref(h)), "nido_graph_create_windowed")
        else:
            _check(lib.nido_graph_create(len(off) - 1, len(col), _p64(off), _p64(col), _pd(w), nbatches, st,
                                         C.byref(h)), "nido_graph_create")
        self._h = h
        self.nbatches = nbatches
        self.n0 = len(off) - 1

    # ---- info ----
    def info(self):
        nv, ne, mo, m = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        _check(lib.nido_graph_info(self._h, C.byref(nv), C.byref(ne), C.byref(mo), C.byref(m)),
               "nido_graph_info")
        return {"nv": nv.value, "ne": ne.value, "max_order": mo.value, "mass": m.value}

    @property
    def num_vertices(self) -> int:
        return self.info()["nv"]

    def batch_bounds(self) -> np.ndarray:
        out = np.zeros(self.nbatches + 1, dtype=np.int64)
        _check(lib.nido_graph_batch_bounds(self._h, _p64(out)), "nido_graph_batch_bounds")
        return out

    def autotune(self, iterations: int = 2) -> dict:
        """Times candidate launch knobs on the first iterations of the phase
        (nido_autotune), applies the fastest and returns it."""
        t = Tuning()
        _check(lib.nido_autotune(self._h, iterations, C.byref(t)), "nido_autotune")
        return t.as_dict()

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib.nido_graph_get_tuning(self._h, C.byref(t)), "nido_graph_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        """Launch-shape knobs (see nido_tuning in include/nido_amd.h); the
        results do not depend on them."""
        t = Tuning()
        _check(lib.nido_graph_get_tuning(self._h, C.byref(t)), "nido_graph_get_tuning")
        for k, v in kw.items():
            if k not in t.as_dict():
                raise KeyError(k)
            setattr(t, k, int(v))
        _check(lib.nido_graph_set_tuning(self._h, C.byref(t)), "nido_graph_set_tuning")

    def sync(self) -> None:
        _check(lib.nido_graph_sync(self._h), "nido_graph_sync")

    # ---- GraphGPU methods ----
    def singleton_partition(self) -> None:
        _check(lib.nido_singleton_partition(self._h), "nido_singleton_partition")

    def compute_modularity(self) -> float:
        q = C.c_double()
        _check(lib.nido_compute_modularity(self._h, C.byref(q)), "nido_compute_modularity")
        return q.value

    def louvain_update(self, batch: int) -> None:
        _check(lib.nido_louvain_update(self._h, batch), "nido_louvain_update")

    def update_community_weights(self, batch: int) -> None:
        _check(lib.nido_update_community_weights(self._h, batch), "nido_update_community_weights")

    def update_community_ids(self, batch: int) -> None:
        _check(lib.nido_update_community_ids(self._h, batch), "nido_update_community_ids")

    def restore_community(self) -> None:
        _check(lib.nido_restore_community(self._h), "nido_restore_community")

    def aggregation(self) -> bool:
        done = C.c_int()
        _check(lib.nido_aggregation(self._h, C.byref(done)), "nido_aggregation")
        return bool(done.value)

    def louvain_iteration(self, q: float):
        """One loop-body pass (all batches, modularity, accept/rollback):
        returns (accepted Q, dQ, evaluated Q)."""
        qq, dq, qe = C.c_double(q), C.c_double(), C.c_double()
        _check(lib.nido_louvain_iteration(self._h, C.byref(qq), C.byref(dq), C.byref(qe)),
               "nido_louvain_iteration")
        return qq.value, dq.value, qe.value

    def set_profiling(self, on: bool, classes=None, stride: int = 1) -> None:
        """classes: names from KERNEL_CLASSES to bracket with events (all by
        default); stride: events on every stride-th launch of each class."""
        _check(lib.nido_graph_set_profiling_stride(self._h, int(stride)), "nido_graph_set_profiling_stride")
        if classes is None:
            _check(lib.nido_graph_set_profiling(self._h, int(on)), "nido_graph_set_profiling")
        else:
            mask = 0
            for c in classes:
                mask |= 1 << KERNEL_CLASSES.index(c)
            _check(lib.nido_graph_set_profiling_classes(self._h, mask if on else 0),
                   "nido_graph_set_profiling_classes")

    def kernel_stats(self) -> dict:
        out = {}
        for i, name in enumerate(KERNEL_CLASSES):
            n, ms, b = C.c_int64(), C.c_double(), C.c_double()
            _check(lib.nido_graph_kernel_stats(self._h, i, C.byref(n), C.byref(ms), C.byref(b)),
                   "nido_graph_kernel_stats")
            out[name] = {"launches": n.value, "ms": ms.value, "bytes": b.value}
        return out

    # ---- readers ----
    def communities(self) -> np.ndarray:
        out = np.zeros(self.num_vertices, dtype=np.int64)
        _check(lib.nido_get_communities(self._h, _p64(out)), "nido_get_communities")
        return out

    def new_communities(self) -> np.ndarray:
        out = np.zeros(self.num_vertices, dtype=np.int64)
        _check(lib.nido_get_new_communities(self._h, _p64(out)), "nido_get_new_communities")
        return out

    def vertex_weights(self) -> np.ndarray:
        out = np.zeros(self.num_vertices, dtype=np.float64)
        _check(lib.nido_get_vertex_weights(self._h, _pd(out)), "nido_get_vertex_weights")
        return out

    def community_weights(self) -> np.ndarray:
        out = np.zeros(self.num_vertices, dtype=np.float64)
        _check(lib.nido_get_community_weights(self._h, _pd(out)), "nido_get_community_weights")
        return out

    def graph(self):
        i = self.info()
        off = np.zeros(i["nv"] + 1, dtype=np.int64)
        col = np.zeros(i["ne"], dtype=np.int64)
        w = np.zeros(i["ne"], dtype=np.float64)
        _check(lib.nido_get_graph(self._h, _p64(off), _p64(col), _pd(w)), "nido_get_graph")
        return off, col, w

    def clustering(self) -> np.ndarray:
        out = np.zeros(self.n0, dtype=np.int64)
        _check(lib.nido_get_clustering(self._h, _p64(out)), "nido_get_clustering")
        return out

    def dump_partition(self, path: str, new_orders=None) -> None:
        if new_orders is None:
            _check(lib.nido_dump_partition(self._h, os.fsencode(path)), "nido_dump_partition")
        else:
            no = np.ascontiguousarray(new_orders, np.int64)
            _check(lib.nido_dump_partition_orders(self._h, os.fsencode(path), _p64(no)),
                   "nido_dump_partition_orders")

    # ---- run history ----
    def phase(self, i: int) -> PhaseStats:
        st = PhaseStats()
        _check(lib.nido_run_phase(self._h, i, C.byref(st)), "nido_run_phase")
        return st

    def iteration_q(self, i: int) -> np.ndarray:
        out = np.zeros(self.phase(i).loops, dtype=np.float64)
        _check(lib.nido_run_iteration_q(self._h, i, _pd(out)), "nido_run_iteration_q")
        return out

    def phase_communities(self, i: int) -> np.ndarray:
        out = np.zeros(self.phase(i).nv, dtype=np.int64)
        _check(lib.nido_run_phase_communities(self._h, i, _p64(out)), "nido_run_phase_communities")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.nido_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LouvainGPU:
    """The reference's LouvainGPU (louvain_gpu.hpp:15-23): the driver loop
    runs inside the library (nido_louvain_run)."""

    def __init__(self, max_loops: int = DEFAULT_ITERATIONS, tol: float = DEFAULT_THRESHOLD,
                 nbatches: int = DEFAULT_BATCHES, tol_phase: float | None = None):
        self.max_loops = max_loops
        self.tol = tol
        self.tol_phase = tol if tol_phase is None else tol_phase
        self.nbatches = nbatches

    def run(self, graph: GraphGPU, record: bool = False, verbose: bool = False) -> RunStats:
        st = RunStats()
        _check(lib.nido_louvain_run(graph._h, self.max_loops, self.tol, self.tol_phase, int(record),
                                    int(verbose), C.byref(st)), "nido_louvain_run")
        return st


def concat_graphs(graphs):
    """One CSR for several graphs: (vptr, offsets, cols, weights) with
    handle-wide column ids, graph i holding vertices [vptr[i], vptr[i+1])."""
    vptr = [0]
    offs, cols, ws = [], [], []
    abase = 0
    for g in graphs:
        if isinstance(g, Graph):
            off, col, w = g.arrays(copy=False)
        else:
            off, col, w = g
        off = np.asarray(off, dtype=np.int64)
        n = len(off) - 1
        offs.append(off[:-1] - off[0] + abase)
        cols.append(np.asarray(col, dtype=np.int64)[off[0]:off[-1]] + vptr[-1])
        ws.append(np.asarray(w, dtype=np.float64)[off[0]:off[-1]])
        abase += int(off[-1] - off[0])
        vptr.append(vptr[-1] + n)
    offs.append(np.array([abase], dtype=np.int64))
    return (np.asarray(vptr, dtype=np.int64), np.concatenate(offs),
            np.concatenate(cols) if cols else np.zeros(0, np.int64),
            np.concatenate(ws) if ws else np.zeros(0))


class BatchGPU:
    """Many independent graphs in one device handle; run() executes
    LouvainGPU::run on all of them at once (nido_batch_run).  Every graph's
    result equals its result alone."""

    def __init__(self, graphs=None, nbatches: int = DEFAULT_BATCHES, csr=None, stream: int | None = None):
        vptr, off, col, w = csr if csr is not None else concat_graphs(graphs)
        vptr, off = np.ascontiguousarray(vptr, np.int64), np.ascontiguousarray(off, np.int64)
        col, w = np.ascontiguousarray(col, np.int64), np.ascontiguousarray(w, np.float64)
        self._keep = (vptr, off, col, w)
        self.vptr = vptr
        self.ngraphs = len(vptr) - 1
        h = C.c_void_p()
        _check(lib.nido_batch_create(self.ngraphs, _p64(vptr), _p64(off), _p64(col), _pd(w), nbatches,
                                     C.c_void_p(stream) if stream else None, C.byref(h)), "nido_batch_create")
        self._h = h
        self.nbatches = nbatches

    def run(self, max_loops: int = DEFAULT_ITERATIONS, tol: float = DEFAULT_THRESHOLD,
            tol_phase: float | None = None) -> RunStats:
        st = RunStats()
        _check(lib.nido_batch_run(self._h, max_loops, tol, tol if tol_phase is None else tol_phase,
                                  C.byref(st)), "nido_batch_run")
        return st

    def list_rebuilds(self) -> int:
        """Work-list rebuilds for the graphs still iterating in the last run."""
        n = C.c_int64()
        _check(lib.nido_batch_list_rebuilds(self._h, C.byref(n)), "nido_batch_list_rebuilds")
        return n.value

    def graph_stats(self, i: int) -> RunStats:
        st = RunStats()
        _check(lib.nido_batch_graph_stats(self._h, i, C.byref(st)), "nido_batch_graph_stats")
        return st

    def autotune(self, iterations: int = 2) -> dict:
        t = Tuning()
        _check(lib.nido_autotune(self._h, iterations, C.byref(t)), "nido_autotune")
        return t.as_dict()

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib.nido_graph_get_tuning(self._h, C.byref(t)), "nido_graph_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        t = Tuning()
        _check(lib.nido_graph_get_tuning(self._h, C.byref(t)), "nido_graph_get_tuning")
        for k, v in kw.items():
            if k not in t.as_dict():
                raise KeyError(k)
            setattr(t, k, int(v))
        _check(lib.nido_graph_set_tuning(self._h, C.byref(t)), "nido_graph_set_tuning")

    def graph_phase(self, i: int, phase: int) -> PhaseStats:
        ps = PhaseStats()
        _check(lib.nido_batch_graph_phase(self._h, i, phase, C.byref(ps)), "nido_batch_graph_phase")
        return ps

    def clustering(self, i: int) -> np.ndarray:
        out = np.zeros(int(self.vptr[i + 1] - self.vptr[i]), dtype=np.int64)
        _check(lib.nido_batch_clustering(self._h, i, _p64(out)), "nido_batch_clustering")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.nido_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (nido_serve_*): submit() copies
    the graph and returns at once; wait() blocks for that request."""

    def __init__(self, slots: int = 8, nbatches: int = DEFAULT_BATCHES, hbm_budget: int = 200 << 30,
                 max_loops: int = DEFAULT_ITERATIONS, tol: float = DEFAULT_THRESHOLD, tol_phase: float | None = None,
                 group: int = 1):
        cfg = ServeConfig(slots, nbatches, hbm_budget, max_loops, tol, tol if tol_phase is None else tol_phase,
                          group)
        h = C.c_void_p()
        _check(lib.nido_serve_create(C.byref(cfg), C.byref(h)), "nido_serve_create")
        self._h = h
        self._n = {}

    def submit(self, graph, borrow: bool = False) -> int:
        """borrow=True: the arrays are not copied; keep `graph` alive and
        unchanged until wait() returns for this request."""
        if isinstance(graph, Graph):
            off, col, w = graph.arrays(copy=False)
        else:
            off, col, w = graph
        off = np.ascontiguousarray(off, np.int64)
        col = np.ascontiguousarray(col, np.int64)
        w = np.ascontiguousarray(w, np.float64)
        rid = C.c_int64()
        _check(lib.nido_serve_submit(self._h, len(off) - 1, _p64(off), _p64(col), _pd(w), 1 if borrow else 0,
                                     C.byref(rid)), "nido_serve_submit")
        self._n[rid.value] = len(off) - 1
        if borrow:
            self._held = getattr(self, "_held", {})
            self._held[rid.value] = (graph, off, col, w)
        return rid.value

    def wait(self, rid: int):
        st, lat = RunStats(), C.c_double()
        _check(lib.nido_serve_wait(self._h, rid, C.byref(st), C.byref(lat)), "nido_serve_wait")
        return st, lat.value

    def clustering(self, rid: int) -> np.ndarray:
        out = np.zeros(max(self._n[rid], 1), dtype=np.int64)
        _check(lib.nido_serve_clustering(self._h, rid, _p64(out)), "nido_serve_clustering")
        return out[:self._n[rid]]

    def release(self, rid: int) -> None:
        _check(lib.nido_serve_release(self._h, rid), "nido_serve_release")
        self._n.pop(rid, None)
        getattr(self, "_held", {}).pop(rid, None)

    def status(self) -> dict:
        q, r, c, p = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib.nido_serve_status(self._h, C.byref(q), C.byref(r), C.byref(c), C.byref(p)), "nido_serve_status")
        return {"queued": q.value, "running": r.value, "completed": c.value, "peak_running": p.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.nido_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
