// device_common.h -- device helpers shared by the kernel translation units.
#pragma once
#include <climits>
#include "internal.h"

namespace nido {

constexpr double kMaxFloat = 1.0e308;   // MAX_FLOAT, cuda_wrapper.hpp:112
constexpr int kNone = INT_MAX;          // "no candidate community"
constexpr int kBlock = 256;
static_assert(kMovedSlots == kBlock, "one moved-arc counter per thread of a workgroup");
constexpr unsigned kHashMul = 0x9E3779B1u;    // slot hash (Fibonacci)
constexpr unsigned kPartMul = 0x85EBCA6Bu;    // partition hash, independent of the slot hash

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

inline unsigned grid_for(int64_t items, int64_t per_block, unsigned cap = 8192)
{
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Candidate order: higher gain first, then smaller community id; kNone is
// below everything.  A total order, so any reduction tree picks the same
// winner (the reference keeps the leftmost lane instead, graph_cuda.cu:1507-1526).
__device__ __forceinline__ bool better(double g, int c, double bg, int bc)
{
    return c != kNone && (bc == kNone || g > bg || (g == bg && c < bc));
}

// Makes one lane's LDS writes visible to the other lanes of its wavefront.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ fx_t wave_sum_fx(fx_t x)
{
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) x += fx_shfl_xor(x, msk, 64);
    return x;
}

__device__ __forceinline__ void acc_add(fx_t* acc, fx_t x)
{
    unsigned long long* p = reinterpret_cast<unsigned long long*>(acc);
    fx_atomic_add(p, p + 1, x);
}

// The arcs of a workgroup's moved vertices, added with one atomic into one
// of kMovedSlots counters (the same address from every wavefront of a
// launch serialises: it cost a commit 40% of its time).  Every thread of the
// workgroup calls it; `red` holds kBlock/64 values of LDS.
__device__ __forceinline__ void add_moved_arcs(unsigned long long* slots, unsigned long long x,
                                               unsigned long long* red)
{
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) x += __shfl_xor(x, msk, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
        if (t != 0) atomicAdd(slots + (blockIdx.x & (kMovedSlots - 1)), t);
    }
}

// Sum of the kMovedSlots counters (kBlock == kMovedSlots threads, one
// workgroup); valid in thread 0.
__device__ __forceinline__ unsigned long long sum_moved_slots(const unsigned long long* slots,
                                                              unsigned long long* red)
{
    unsigned long long x = slots[threadIdx.x];
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) x += __shfl_xor(x, msk, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    return t;
}

// Loads that bypass the (non-coherent) vector L1, for data written by other
// workgroups through L2 atomics or stores in the same launch.
__device__ __forceinline__ int ld_i32_agent(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_u64_agent(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ fx_t ld_fx_agent(const fx_t* p)
{
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
    return fx_make(ld_u64_agent(q), (int64_t)ld_u64_agent(q + 1));
}

// Per-block partial sums of Q64.64 values: one global store per block and a
// single-block finish, instead of same-address atomics from every wavefront.
__device__ __forceinline__ fx_t block_sum_fx(fx_t x, fx_t* red)
{
    x = wave_sum_fx(x);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    fx_t t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    __syncthreads();
    return t;   // valid in thread 0
}

__device__ __forceinline__ void delta_graphs(int b, const int64_t* __restrict__ Cn, const int* __restrict__ C32,
                                             const int* __restrict__ col, const double* __restrict__ w,
                                             const fx_t* __restrict__ a, const ModDelta& D, int bi, int nbk);

__device__ __forceinline__ void delta_flush(const ModDelta& D, int g, int which, fx_t& x)
{
    x = wave_sum_fx(x);
    if ((threadIdx.x & 63) == 0 && g >= 0 && x != 0) acc_add(D.acc + 2 * (int64_t)g + which, x);
    x = 0;
}

// Lane value t of graph g (-1: none; lane 0 always has one) joins the sums
// carried for graph cg, or, in a window that spans graphs, goes to its
// graph's accumulator at once.
__device__ __forceinline__ void delta_carry(const ModDelta& D, int which, int g, fx_t t, int& cg, fx_t& carried)
{
    const int g0 = __shfl(g, 0, 64);
    if (__ballot(g >= 0 && g != g0) == 0) {   // uniform
        if (g0 != cg) {
            delta_flush(D, cg, which, carried);
            cg = g0;
        }
        carried += t;
    } else {
        delta_flush(D, cg, which, carried);
        cg = -1;
        if (t != 0) acc_add(D.acc + 2 * (int64_t)g + which, t);
    }
}

// Listed chunks [j, j1) (j1 - j <= 64) of batch b, their arcs flattened:
// lane l holds chunk j + l and the inclusive prefix of the chunks' lengths;
// each step then takes the next 64 arcs of their concatenation, one per
// lane, finds the arc's chunk by a binary search over the prefix and reads
// that chunk's context from its lane.  For a moved vertex v and arc (v, u):
// old = [C'(v) == old(u)], new = [C(v) == C(u)], where old(u) is C'(u) when
// u is in the batch ([lo, hi): [v0, v1) for one graph, the bounds of batch b
// of v's graph otherwise; the commit left every batch vertex's previous
// community in C') and C(u) outside it; the arc (u, v) changes the same way,
// so the difference counts twice unless u moved too (then u's row counts
// it).  Each step's values join the sums carried for the wavefront's graph.
__device__ __forceinline__ void delta_span(int b, int64_t j, int64_t j1, int64_t v0, int64_t v1,
                                           const int64_t* __restrict__ Cn, const int* __restrict__ C32,
                                           const int* __restrict__ col, const double* __restrict__ w,
                                           const ModDelta& D, int& cg, fx_t& carried)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = j + lane;
    int64_t e0 = 0, lo = v0, hi = v1;
    int len = 0, cvn = 0, cvo = 0, g = 0;
    if (i < j1) {
        const RowItem ch = D.chunks[i];
        const int v = (int)ch.v;
        e0 = ch.e0;
        len = (int)(ch.e1 - ch.e0);   // at most kDeltaArcs
        if (D.gid) {
            g = D.gid[v];
            lo = D.gp[(int64_t)g * D.nb1 + b];
            hi = D.gp[(int64_t)g * D.nb1 + b + 1];
        }
        cvn = C32[v];
        cvo = (int)Cn[v];
    }
    int inc = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    const int total = __shfl(inc, 63, 64);
    for (int p0 = 0; p0 < total; p0 += 64) {   // uniform
        const int p = p0 + lane;
        int k = 0;   // chunks that end at or before p (inc is nondecreasing)
#pragma unroll
        for (int sft = 32; sft > 0; sft >>= 1)
            if (__shfl(inc, k + sft - 1, 64) <= p) k += sft;
        const int start = __shfl(inc, k, 64) - __shfl(len, k, 64);
        const int64_t ek = __shfl(e0, k, 64) + (p - start);
        const int vn = __shfl(cvn, k, 64), vo = __shfl(cvo, k, 64);
        int64_t l = v0, h = v1;
        int gk = 0;
        if (D.gid) {   // uniform
            l = __shfl(lo, k, 64);
            h = __shfl(hi, k, 64);
            gk = __shfl(g, k, 64);
        }
        fx_t t = 0;
        if (p < total) {
            const int u = col[ek];
            const int cun = C32[u];
            const bool inb = u >= l && u < h;
            const int cuo = inb ? (int)Cn[u] : cun;
            const int dlt = (int)(vn == cun) - (int)(vo == cuo);
            if (dlt) {
                fx_t f = fx_from_double(w[ek]);
                if (!(inb && cuo != cun)) f += f;
                t = dlt > 0 ? f : -f;
            }
        } else {
            gk = -1;
        }
        delta_carry(D, 0, gk, t, cg, carried);
    }
}

// The exact change of the internal weight and of the square terms caused by
// batch [v0, v1)'s moves (after its commit), on workgroups [bi, bi + nbk) of
// a launch.  For a moved vertex v and arc (v, u): old = [C'(v) == old(u)],
// new = [C(v) == C(u)] where old(u) is C'(u) inside the batch (the commit
// left every batch vertex's previous community there) and C(u) outside it;
// the arc (u, v) changes the same way, so the difference counts twice unless
// u moved too (then u's row counts it).  Touched communities re-round their
// square term.  `red` holds kBlock/64 values of LDS.
__device__ __forceinline__ void delta_body(int b, int64_t v0, int64_t v1, const int64_t* __restrict__ Cn,
                                           const int* __restrict__ C32, const int* __restrict__ col,
                                           const double* __restrict__ w, const fx_t* __restrict__ a,
                                           double twom_d, const ModDelta& D, int bi, int nbk, fx_t* red)
{
    if (D.gid) {   // uniform
        delta_graphs(b, Cn, C32, col, w, a, D, bi, nbk);
        return;
    }
    const int nch = D.counts[2 * b], ncm = D.counts[2 * b + 1];
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)bi * (kBlock / 64) + (threadIdx.x >> 6);
    const int64_t nwv = (int64_t)nbk * (kBlock / 64);
    fx_t dI = 0, dT = 0;
    for (int64_t i = wid; i < nch; i += nwv) {
        const RowItem ch = D.chunks[i];
        const int v = (int)ch.v;
        const int cvn = C32[v], cvo = (int)Cn[v];
        for (int64_t e = ch.e0 + lane; e < ch.e1; e += 64) {
            const int u = col[e];
            const int cun = C32[u];
            const bool inb = u >= v0 && u < v1;
            const int cuo = inb ? (int)Cn[u] : cun;
            const int dlt = (int)(cvn == cun) - (int)(cvo == cuo);
            if (dlt) {
                fx_t f = fx_from_double(w[e]);
                if (!(inb && cuo != cun)) f += f;
                dI += dlt > 0 ? f : -f;
            }
        }
    }
    for (int64_t j = (int64_t)bi * kBlock + threadIdx.x; j < ncm; j += (int64_t)nbk * kBlock) {
        const int c = D.comms[j];
        const double x = fx_to_double(a[c]) / twom_d;
        const fx_t nt = fx_from_double(x * x);
        dT += nt - D.term[c];
        D.term[c] = nt;
        D.touched[c] = 0;
    }
    dI = block_sum_fx(dI, red);
    dT = block_sum_fx(dT, red);
    if (threadIdx.x == 0 && (dI != 0 || dT != 0)) {
        acc_add(D.acc, dI);
        acc_add(D.acc + 1, dT);
    }
}

// The delta of batch b in a handle of several graphs.  Arcs stay inside
// their graph, so u is in the batch when it lies inside its graph's bounds
// of batch b, and each graph's changes go to its own accumulator pair.  A
// wavefront takes a contiguous run of each list (a commit lists its rows and
// communities in runs of one graph), carries its sums while the graph stays
// the same and adds them with one atomic pair when it changes; a window of 64
// communities that spans graphs adds lane by lane.
__device__ __forceinline__ void delta_graphs(int b, const int64_t* __restrict__ Cn, const int* __restrict__ C32,
                                             const int* __restrict__ col, const double* __restrict__ w,
                                             const fx_t* __restrict__ a, const ModDelta& D, int bi, int nbk)
{
    const int64_t nch = D.counts[2 * b], ncm = D.counts[2 * b + 1];
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)bi * (kBlock / 64) + (threadIdx.x >> 6);
    const int64_t nwv = (int64_t)nbk * (kBlock / 64);
    {
        int cg = -1;
        fx_t dI = 0;
        const int64_t i1 = nch * (wid + 1) / nwv;
        for (int64_t j = nch * wid / nwv; j < i1; j += 64)
            delta_span(b, j, j + 64 < i1 ? j + 64 : i1, 0, 0, Cn, C32, col, w, D, cg, dI);
        delta_flush(D, cg, 0, dI);
    }
    {
        int cg = -1;
        fx_t dT = 0;
        const int64_t j1 = ncm * (wid + 1) / nwv;
        for (int64_t j = ncm * wid / nwv; j < j1; j += 64) {
            const int64_t jj = j + lane;
            int g = -1;
            fx_t t = 0;
            if (jj < j1) {
                const int c = D.comms[jj];
                g = D.gid[c];
                const double x = fx_to_double(a[c]) / D.twomg[g];
                const fx_t nt = fx_from_double(x * x);
                t = nt - D.term[c];
                D.term[c] = nt;
                D.touched[c] = 0;
            }
            delta_carry(D, 1, g, t, cg, dT);
        }
        delta_flush(D, cg, 1, dT);
    }
}

// Workgroup-wide exclusive scan of one int per thread; *total receives the
// sum.  `wtot` holds kBlock/64 ints of LDS; the caller synchronises before
// wtot is used again.
__device__ __forceinline__ int block_excl_scan_i32(int x, int* wtot, int* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
    int pre = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        if (i < wv) pre += wtot[i];
        all += wtot[i];
    }
    *total = all;
    return pre + inc - x;
}

// Workgroup-wide argmax + self-weight reduction; result valid in thread 0.
struct BlockRed {
    double g[kBlock / 64];
    int c[kBlock / 64];
    unsigned long long slo[kBlock / 64];
    unsigned long long shi[kBlock / 64];
};

// Workgroup-wide argmax alone; result valid in thread 0.
__device__ __forceinline__ void block_argmax(BlockRed& R, double& bg, int& bc)
{
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) {
        const double og = __shfl_xor(bg, msk, 64);
        const int oc = __shfl_xor(bc, msk, 64);
        if (better(og, oc, bg, bc)) {
            bg = og;
            bc = oc;
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        R.g[wv] = bg;
        R.c[wv] = bc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / 64; ++i) {
            if (better(R.g[i], R.c[i], bg, bc)) {
                bg = R.g[i];
                bc = R.c[i];
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void block_reduce(BlockRed& R, fx_t& self, double& bg, int& bc)
{
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) {
        const double og = __shfl_xor(bg, msk, 64);
        const int oc = __shfl_xor(bc, msk, 64);
        self += fx_shfl_xor(self, msk, 64);
        if (better(og, oc, bg, bc)) {
            bg = og;
            bc = oc;
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        R.g[wv] = bg;
        R.c[wv] = bc;
        R.slo[wv] = fx_lo(self);
        R.shi[wv] = (unsigned long long)fx_hi(self);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / 64; ++i) {
            self += fx_make(R.slo[i], (int64_t)R.shi[i]);
            if (better(R.g[i], R.c[i], bg, bc)) {
                bg = R.g[i];
                bc = R.c[i];
            }
        }
    }
    __syncthreads();
}

}  // namespace nido
