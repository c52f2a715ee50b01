"""CPU oracle for nido's batched Louvain path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  It wraps liboracle.so, built from
louvain_oracle.c by oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).

The oracle restates the reference's semantics (citations in
louvain_oracle.c) with canonical arithmetic: exact Q64.64 sums, min-id
gain ties, coarse rows sorted by neighbour id.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


if not os.path.exists(LIB_PATH):
    build()

_lib = C.CDLL(LIB_PATH)
_i64 = C.c_int64
_pi64 = C.POINTER(C.c_int64)
_pdbl = C.POINTER(C.c_double)
_pu64 = C.POINTER(C.c_uint64)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("orc_create", C.c_void_p, _i64, _i64, _pi64, _pi64, _pdbl, C.c_int, C.c_int)
_sig("orc_destroy", None, C.c_void_p)
_sig("orc_error", C.c_int, C.c_void_p)
_sig("orc_nv", _i64, C.c_void_p)
_sig("orc_ne", _i64, C.c_void_p)
_sig("orc_max_order", _i64, C.c_void_p)
_sig("orc_mass", C.c_double, C.c_void_p)
_sig("orc_twom_words", None, C.c_void_p, _pu64, _pi64)
_sig("orc_batch_bounds", None, _i64, _pi64, C.c_int, _pi64)
_sig("orc_get_batch_bounds", None, C.c_void_p, _pi64)
_sig("orc_get_graph", None, C.c_void_p, _pi64, _pi64, _pdbl)
_sig("orc_get_communities", None, C.c_void_p, _pi64)
_sig("orc_get_new_communities", None, C.c_void_p, _pi64)
_sig("orc_set_communities", None, C.c_void_p, _pi64, _pi64)
_sig("orc_get_vertex_weights", None, C.c_void_p, _pdbl)
_sig("orc_get_comm_weights", None, C.c_void_p, _pdbl)
_sig("orc_get_comm_weights_words", None, C.c_void_p, _pu64, _pi64)
_sig("orc_get_clustering", None, C.c_void_p, _pi64)
_sig("orc_singleton_partition", None, C.c_void_p)
_sig("orc_compute_modularity", C.c_double, C.c_void_p)
_sig("orc_louvain_update", None, C.c_void_p, C.c_int)
_sig("orc_update_community_weights", None, C.c_void_p, C.c_int)
_sig("orc_update_community_ids", None, C.c_void_p, C.c_int)
_sig("orc_restore_community", None, C.c_void_p)
_sig("orc_aggregation", C.c_int, C.c_void_p)
_sig("orc_run", None, C.c_void_p, _i64, C.c_double, C.c_double, C.c_int)
_sig("orc_louvain_iteration", C.c_double, C.c_void_p, _pdbl, _pdbl)
_sig("orc_hist_phases", C.c_int, C.c_void_p)
_sig("orc_hist_phase", None, C.c_void_p, C.c_int, _pi64, _pi64, _pi64, _pdbl, _pdbl, _pdbl, _pdbl)
_sig("orc_hist_iter_q", None, C.c_void_p, C.c_int, _pdbl)
_sig("orc_hist_comm", C.c_int, C.c_void_p, C.c_int, _pi64)
_sig("orc_times", None, C.c_void_p, _pdbl, _pdbl, _pdbl, _pi64)
_sig("orc_threads", C.c_int, C.c_void_p)
_sig("orc_fx_from_double_words", None, C.c_double, _pu64, _pi64, C.POINTER(C.c_int))
_sig("orc_fx_words_to_double", C.c_double, C.c_uint64, C.c_int64)
_sig("orc_coloring", _i64, _i64, _pi64, _pi64, _pi64, _pi64)
_sig("orc_reorder", None, _i64, _pi64, _pi64, _pdbl, _pi64, _pi64, _pi64, _pi64, _pdbl)


def fx_from_double(w: float) -> int:
    """Q64.64 value (a Python int in units of 2^-64) of an FP64 weight."""
    lo, hi, err = C.c_uint64(), C.c_int64(), C.c_int()
    _lib.orc_fx_from_double_words(float(w), C.byref(lo), C.byref(hi), C.byref(err))
    if err.value:
        raise ValueError(f"weight {w!r} is not representable")
    return (hi.value << 64) | lo.value


def fx_to_double(x: int) -> float:
    x &= (1 << 128) - 1
    lo = x & ((1 << 64) - 1)
    hi = x >> 64
    if hi >= 1 << 63:
        hi -= 1 << 64
    return _lib.orc_fx_words_to_double(lo, hi)


def batch_bounds(off: np.ndarray, nbatches: int) -> np.ndarray:
    off = np.ascontiguousarray(off, dtype=np.int64)
    out = np.zeros(nbatches + 1, dtype=np.int64)
    _lib.orc_batch_bounds(len(off) - 1, off.ctypes.data_as(_pi64), nbatches, out.ctypes.data_as(_pi64))
    return out


class Oracle:
    """CPU restatement of GraphGPU + LouvainGPU on one graph."""

    def __init__(self, off, col, w, nbatches: int = 2, nthreads: int = 0):
        off = np.ascontiguousarray(off, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        self.n0 = len(off) - 1
        self.nbatches = nbatches
        self._h = _lib.orc_create(self.n0, len(col), off.ctypes.data_as(_pi64),
                                  col.ctypes.data_as(_pi64), w.ctypes.data_as(_pdbl), nbatches,
                                  nthreads)
        if not self._h:
            raise ValueError("orc_create rejected the graph")
        if _lib.orc_error(self._h):
            code = _lib.orc_error(self._h)
            self.close()
            raise ValueError(f"oracle input error {code}")

    # info
    @property
    def nv(self) -> int:
        return _lib.orc_nv(self._h)

    @property
    def ne(self) -> int:
        return _lib.orc_ne(self._h)

    @property
    def max_order(self) -> int:
        return _lib.orc_max_order(self._h)

    @property
    def mass(self) -> float:
        return _lib.orc_mass(self._h)

    @property
    def twom_fx(self) -> int:
        lo, hi = C.c_uint64(), C.c_int64()
        _lib.orc_twom_words(self._h, C.byref(lo), C.byref(hi))
        return (hi.value << 64) | lo.value

    @property
    def threads(self) -> int:
        return _lib.orc_threads(self._h)

    def batch_bounds(self) -> np.ndarray:
        out = np.zeros(self.nbatches + 1, dtype=np.int64)
        _lib.orc_get_batch_bounds(self._h, out.ctypes.data_as(_pi64))
        return out

    def graph(self):
        off = np.zeros(self.nv + 1, dtype=np.int64)
        col = np.zeros(self.ne, dtype=np.int64)
        w = np.zeros(self.ne, dtype=np.float64)
        _lib.orc_get_graph(self._h, off.ctypes.data_as(_pi64), col.ctypes.data_as(_pi64),
                           w.ctypes.data_as(_pdbl))
        return off, col, w

    def communities(self) -> np.ndarray:
        out = np.zeros(self.nv, dtype=np.int64)
        _lib.orc_get_communities(self._h, out.ctypes.data_as(_pi64))
        return out

    def new_communities(self) -> np.ndarray:
        out = np.zeros(self.nv, dtype=np.int64)
        _lib.orc_get_new_communities(self._h, out.ctypes.data_as(_pi64))
        return out

    def set_communities(self, comm, new_comm=None) -> None:
        comm = np.ascontiguousarray(comm, dtype=np.int64)
        nc = None if new_comm is None else np.ascontiguousarray(new_comm, dtype=np.int64)
        _lib.orc_set_communities(self._h, comm.ctypes.data_as(_pi64),
                                 None if nc is None else nc.ctypes.data_as(_pi64))

    def vertex_weights(self) -> np.ndarray:
        out = np.zeros(self.nv, dtype=np.float64)
        _lib.orc_get_vertex_weights(self._h, out.ctypes.data_as(_pdbl))
        return out

    def community_weights(self) -> np.ndarray:
        out = np.zeros(self.nv, dtype=np.float64)
        _lib.orc_get_comm_weights(self._h, out.ctypes.data_as(_pdbl))
        return out

    def community_weights_fx(self) -> list[int]:
        lo = np.zeros(self.nv, dtype=np.uint64)
        hi = np.zeros(self.nv, dtype=np.int64)
        _lib.orc_get_comm_weights_words(self._h, lo.ctypes.data_as(_pu64), hi.ctypes.data_as(_pi64))
        return [(int(h) << 64) | int(l) for l, h in zip(lo, hi)]

    def clustering(self) -> np.ndarray:
        out = np.zeros(self.n0, dtype=np.int64)
        _lib.orc_get_clustering(self._h, out.ctypes.data_as(_pi64))
        return out

    # GraphGPU method set
    def singleton_partition(self) -> None:
        _lib.orc_singleton_partition(self._h)

    def compute_modularity(self) -> float:
        return _lib.orc_compute_modularity(self._h)

    def louvain_update(self, b: int) -> None:
        _lib.orc_louvain_update(self._h, b)

    def update_community_weights(self, b: int) -> None:
        _lib.orc_update_community_weights(self._h, b)

    def update_community_ids(self, b: int) -> None:
        _lib.orc_update_community_ids(self._h, b)

    def restore_community(self) -> None:
        _lib.orc_restore_community(self._h)

    def aggregation(self) -> bool:
        return bool(_lib.orc_aggregation(self._h))

    def louvain_iteration(self, q: float):
        """One loop-body pass; returns (accepted Q, dQ, evaluated Q)."""
        qq, dq = C.c_double(q), C.c_double()
        qt = _lib.orc_louvain_iteration(self._h, C.byref(qq), C.byref(dq))
        return qq.value, dq.value, qt

    # LouvainGPU::run
    def run(self, max_loops: int = 500, tol: float = 1e-6, tol_phase: float | None = None,
            record: bool = True) -> dict:
        _lib.orc_run(self._h, max_loops, tol, tol if tol_phase is None else tol_phase, int(record))
        phases = []
        for i in range(_lib.orc_hist_phases(self._h)):
            nv, ne, loops = C.c_int64(), C.c_int64(), C.c_int64()
            q0, q, tl, ta = C.c_double(), C.c_double(), C.c_double(), C.c_double()
            _lib.orc_hist_phase(self._h, i, C.byref(nv), C.byref(ne), C.byref(loops), C.byref(q0),
                                C.byref(q), C.byref(tl), C.byref(ta))
            iq = np.zeros(loops.value, dtype=np.float64)
            _lib.orc_hist_iter_q(self._h, i, iq.ctypes.data_as(_pdbl))
            comm = None
            if record:
                comm = np.zeros(nv.value, dtype=np.int64)
                _lib.orc_hist_comm(self._h, i, comm.ctypes.data_as(_pi64))
            phases.append({"nv": nv.value, "ne": ne.value, "loops": loops.value, "q0": q0.value,
                           "q": q.value, "t_loop": tl.value, "t_agg": ta.value, "iter_q": iq,
                           "comm": comm})
        tt, tl, ta, nl = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        _lib.orc_times(self._h, C.byref(tt), C.byref(tl), C.byref(ta), C.byref(nl))
        return {"phases": phases, "t_total": tt.value, "t_loop": tl.value, "t_agg": ta.value,
                "total_loops": nl.value, "q": phases[-1]["q"] if phases else 0.0}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coloring(off, col, r):
    """Colour of every vertex by the reference's colouring loop run by one
    thread (graph.cpp:573-605), for random weights r; returns (colors, K)."""
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    r = np.ascontiguousarray(r, np.int64)
    nv = len(off) - 1
    color = np.empty(nv, np.int64)
    K = _lib.orc_coloring(nv, _p(off), _p(col), _p(r), _p(color))
    return color, int(K)


def reorder(off, col, w, color):
    """(new_orders, off, col, w) of the graph renumbered in (colour, id)
    order, graph.cpp:607-691."""
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    w = np.ascontiguousarray(w, np.float64)
    color = np.ascontiguousarray(color, np.int64)
    nv = len(off) - 1
    no = np.empty(nv, np.int64)
    noff = np.empty(nv + 1, np.int64)
    ncol = np.empty(len(col), np.int64)
    nw = np.empty(len(col), np.float64)
    _lib.orc_reorder(nv, _p(off), _p(col), w.ctypes.data_as(_pdbl), _p(color), _p(no), _p(noff), _p(ncol),
                     nw.ctypes.data_as(_pdbl))
    return no, noff, ncol, nw


def _p(a):
    return a.ctypes.data_as(_pi64)
