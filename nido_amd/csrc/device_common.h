// device_common.h -- device helpers shared by the kernel translation units.
#pragma once
#include <climits>
#include "internal.h"

namespace nido {

constexpr double kMaxFloat = 1.0e308;   // MAX_FLOAT, cuda_wrapper.hpp:112
constexpr int kNone = INT_MAX;          // "no candidate community"
constexpr int kBlock = 256;
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
