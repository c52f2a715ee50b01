"""Independent pure-Python restatement of the canonical Louvain semantics,
used only to pin the C oracle on small graphs.

Written from the same reference sources as the oracle but with different
machinery: Python integers for the Q64.64 sums, fractions.Fraction for every
rounding step (float(Fraction) is correctly rounded, half to even), dicts for
the neighbour-community tables.  Citations are sg0/nido file:line.
"""
from __future__ import annotations

from fractions import Fraction

ONE = 1 << 64
MAX_FLOAT = 1.0e308          # cuda_wrapper.hpp:112


def fx(w: float) -> int:
    """FP64 -> Q64.64 integer, round half to even at 2^-64."""
    return round(Fraction(w) * ONE)          # round() on Fraction is half-to-even


def fxd(x: int) -> float:
    """Q64.64 integer -> FP64, round half to even."""
    return float(Fraction(x, ONE))


def batch_bounds(off, nb):
    """partition_graph_edge_batch, graph_gpu.cpp:303-337 (one device, V0=0)."""
    nv = len(off) - 1
    ave = (off[nv] - off[0]) // nb
    parts, start = [0], 0
    for i in range(1, nv + 1):
        if off[i] - off[start] > ave:
            parts.append(i)
            start = i
        elif i == nv:
            parts.append(nv)
    if len(parts) > nb + 1:
        step = (nv - parts[nb]) // nb
        for i in range(1, nb):
            parts[i] += step
        parts[nb] = nv
    while len(parts) < nb + 1:
        parts.append(nv)
    return parts[: nb + 1]


class State:
    def __init__(self, off, col, w, nb):
        self.off = [int(x) - int(off[0]) for x in off]
        self.col = [int(x) for x in col]
        self.w = [float(x) for x in w]
        self.wf = [fx(x) for x in self.w]
        self.nv = len(self.off) - 1
        self.nb = nb
        self.k = [sum(self.wf[self.off[v]:self.off[v + 1]]) for v in range(self.nv)]   # graph_cuda.cu:332-343
        self.kd = [fxd(x) for x in self.k]
        self.twom = sum(self.k)
        self.m = fxd(self.twom) * 0.5                                                    # graph_cuda.cu:1855
        self.parts = batch_bounds(self.off, nb)
        self.C = list(range(self.nv))
        self.Cn = list(range(self.nv))
        self.a = list(self.k)

    def singleton(self):                                                               # graph_cuda.cu:422-428
        self.C = list(range(self.nv))
        self.Cn = list(range(self.nv))
        self.a = list(self.k)

    def modularity(self):                                                              # graph_cuda.cu:2103-2132
        if self.twom == 0:
            return 0.0
        internal = 0
        for v in range(self.nv):
            for e in range(self.off[v], self.off[v + 1]):
                if self.C[self.col[e]] == self.C[v]:
                    internal += self.wf[e]
        twom_d = fxd(self.twom)
        sq = 0
        for c in range(self.nv):
            x = fxd(self.a[c]) / twom_d
            sq += fx(x * x)
        return fxd(internal) / twom_d - fxd(sq)

    def decide(self, v):                                                               # graph_cuda.cu:1481-1537
        own = self.C[v]
        e_vc = {}
        for e in range(self.off[v], self.off[v + 1]):
            u = self.col[e]
            if u == v:
                continue
            c = self.C[u]
            e_vc[c] = e_vc.get(c, 0) + self.wf[e]
        m = self.m
        ki = self.kd[v]
        cands = []
        for c, s in e_vc.items():
            if c == own:
                continue
            g = fxd(s)
            g -= ki * fxd(self.a[c]) / (2. * m)
            g /= m
            cands.append((-g, c))          # max gain, then min id
        if not cands:
            return own
        neg_g, best = min(cands)
        gbest = -neg_g
        self_w = fxd(e_vc.get(own, 0))
        self_w -= ki * fxd(self.a[own] - self.k[v]) / (2. * m)
        self_w /= m
        gain = gbest
        gain -= self_w
        return best if gain > 0 else own

    def batch(self, b):
        v0, v1 = self.parts[b], self.parts[b + 1]
        if self.twom == 0:
            for v in range(v0, v1):
                self.Cn[v] = self.C[v]
        else:
            dec = [self.decide(v) for v in range(v0, v1)]
            for v, c in zip(range(v0, v1), dec):
                self.Cn[v] = c
        self.commit(v0, v1)

    def commit(self, v0, v1):                                                          # graph_cuda.cu:1234-1249, 1191-1196
        for v in range(v0, v1):
            s, d = self.C[v], self.Cn[v]
            if s != d:
                self.a[s] -= self.k[v]
                self.a[d] += self.k[v]
        for v in range(v0, v1):
            self.C[v], self.Cn[v] = self.Cn[v], self.C[v]

    def aggregate(self, clustering):                                                   # graph_gpu.cpp:1382-1545
        used = sorted(set(self.C))
        newid = {c: i for i, c in enumerate(used)}
        K = len(used)
        for i in range(len(clustering)):
            clustering[i] = newid[self.C[clustering[i]]]
        if K == self.nv or K == 1:
            return None
        rows = [dict() for _ in range(K)]
        for v in range(self.nv):
            i = newid[self.C[v]]
            for e in range(self.off[v], self.off[v + 1]):
                j = newid[self.C[self.col[e]]]
                rows[i][j] = rows[i].get(j, 0) + self.wf[e]
        off, col, w = [0], [], []
        for i in range(K):
            for j in sorted(rows[i]):
                col.append(j)
                w.append(fxd(rows[i][j]))
            off.append(len(col))
        return off, col, w


def run(off, col, w, nb, tol=1e-6, max_loops=500, max_phases=200):
    """LouvainGPU::run, louvain_gpu.cpp:28-124; per-phase records."""
    st = State(off, col, w, nb)
    clustering = list(range(st.nv))
    phases = []
    done = False
    while not done and len(phases) < max_phases:
        st.singleton()
        q = st.modularity()
        q_old = q
        rec = {"nv": st.nv, "ne": len(st.col), "q0": q, "iter_q": []}
        dq, loops = MAX_FLOAT, 0
        while tol < dq and loops < max_loops:
            for b in range(nb):
                st.batch(b)
            qt = st.modularity()
            dq = qt - q
            rec["iter_q"].append(qt)
            loops += 1
            if dq < 0:
                st.commit(0, st.nv)         # restore_community, graph_gpu.cpp:822-842
            else:
                q = qt
        rec.update(loops=loops, q=q, comm=list(st.C))
        phases.append(rec)
        if q - q_old > tol:
            g = st.aggregate(clustering)
            if g is None:
                done = True
            else:
                st = State(*g, nb)
        else:
            done = True
    return {"phases": phases, "q": phases[-1]["q"], "clustering": clustering}
