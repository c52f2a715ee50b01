// kernels.hip -- CDNA4 (gfx950) kernels of the batched Louvain path.
//
// Reference path being replaced (sg0/nido):
//   fill_edges_community_ids + copy_weights_exclude_self_loops + Thrust
//   stable_sort_by_key/reduce_by_key + fill_unique_community_counts +
//   inclusive_scan + louvain_update_kernel  (graph_gpu.cpp:457-517,
//   graph_cuda.cu:49-98, 1450-1611, 2289-2348)
// is one pass here: per vertex, the row is scanned once (coalesced over
// CSR), C[col] gathered, the weight converted to Q64.64 and accumulated in
// an open-addressing table keyed by community id (LDS for rows up to 512
// arcs, HBM above), then the FP64 gain argmax with min-community-id ties is
// taken over the table.  Nothing is materialised per arc.
#include "internal.h"
#include <climits>

namespace nido {
namespace {

constexpr double kMaxFloat = 1.0e308;   // MAX_FLOAT, cuda_wrapper.hpp:112
constexpr int kNone = INT_MAX;          // "no candidate community"
constexpr int kBlock = 256;
constexpr unsigned kHashMul = 0x9E3779B1u;

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

__host__ inline unsigned grid_for(int64_t items, int64_t per_block, unsigned cap = 8192)
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

// g_c = (e_{v,c} - k_v a_c / (2m)) / m, graph_cuda.cu:1505-1506.
__device__ __forceinline__ double gain_of(const LmArgs& A, double ki, fx_t e, int c)
{
    double g = fx_to_double(e);
    const double ac = fx_to_double(A.a[c]);
    g -= ki * ac / (2. * A.m);
    g /= A.m;
    return g;
}

// Move iff g_best - s' > 0 with s' = (e_{v,own} - k_v (a_own - k_v)/(2m))/m,
// graph_cuda.cu:1527-1536.
__device__ __forceinline__ int64_t decide(const LmArgs& A, int v, int own, fx_t self,
                                          double best_g, int best_c)
{
    if (best_c == kNone) return own;
    const double m = A.m;
    const double ki = A.kd[v];
    double s = fx_to_double(self);
    const double ar = fx_to_double(A.a[own] - A.k[v]);
    s -= ki * ar / (2. * m);
    s /= m;
    double gain = best_g;
    gain -= s;
    return gain > 0 ? (int64_t)best_c : (int64_t)own;
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

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, T val, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = val;
}

// ---------------------------------------------------------------------------
// Local moving
// ---------------------------------------------------------------------------

// One lane per vertex, rows of at most kThreadMaxDeg arcs held in registers.
__global__ void __launch_bounds__(kBlock)
lm_thread_kernel(LmArgs A, const int* __restrict__ verts, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int v = verts[i];
        const int own = (int)A.C[v];
        const int64_t e0 = A.off[v];
        const int d = (int)(A.off[v + 1] - e0);
        int cc[kThreadMaxDeg];
        fx_t ww[kThreadMaxDeg];
#pragma unroll
        for (int j = 0; j < kThreadMaxDeg; ++j) {
            cc[j] = -1;
            ww[j] = 0;
            if (j < d) {
                const int u = (int)A.col[e0 + j];
                if (u != v) {                      // self arcs weigh 0 (graph_cuda.cu:2312-2315)
                    cc[j] = (int)A.C[u];
                    ww[j] = fx_from_double(A.w[e0 + j]);
                }
            }
        }
        const double ki = A.kd[v];
        fx_t self = 0;
        double bg = -kMaxFloat;
        int bc = kNone;
#pragma unroll
        for (int j = 0; j < kThreadMaxDeg; ++j) {
            bool first = cc[j] >= 0;
#pragma unroll
            for (int q = 0; q < j; ++q) first = first && (cc[q] != cc[j]);
            if (!first) continue;
            fx_t e = ww[j];
#pragma unroll
            for (int q = j + 1; q < kThreadMaxDeg; ++q)
                if (cc[q] == cc[j]) e += ww[q];
            if (cc[j] == own) {
                self = e;
                continue;
            }
            const double g = gain_of(A, ki, e, cc[j]);
            if (better(g, cc[j], bg, bc)) {
                bg = g;
                bc = cc[j];
            }
        }
        A.Cn[v] = decide(A, v, own, self, bg, bc);
    }
}

// Insert (c, w) into an LDS open-addressing table of S slots (S a power of 2).
template <int S>
__device__ __forceinline__ void lds_insert(int* key, unsigned long long* lo, unsigned long long* hi,
                                           int c, fx_t wf)
{
    unsigned h = ((unsigned)c * kHashMul) >> (32 - ilog2c(S));
    for (;;) {
        const int kk = key[h];
        if (kk == c) break;
        if (kk < 0) {
            const int prev = atomicCAS(&key[h], -1, c);
            if (prev < 0 || prev == c) break;
        }
        h = (h + 1) & (S - 1);
    }
    fx_atomic_add(&lo[h], &hi[h], wf);
}

// G lanes per vertex (G <= 64), S-slot LDS table per vertex; 256/G vertices
// per workgroup.  LDS: (256/G)*S*20 bytes.
template <int G, int S>
__global__ void __launch_bounds__(kBlock)
lm_group_kernel(LmArgs A, const int* __restrict__ verts, int64_t n)
{
    constexpr int NG = kBlock / G;
    __shared__ int skey[NG * S];
    __shared__ unsigned long long slo[NG * S];
    __shared__ unsigned long long shi[NG * S];
    const int grp = threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    int* key = skey + grp * S;
    unsigned long long* lo = slo + grp * S;
    unsigned long long* hi = shi + grp * S;

    for (int64_t base = (int64_t)blockIdx.x * NG; base < n; base += (int64_t)gridDim.x * NG) {
        for (int s = lane; s < S; s += G) {
            key[s] = -1;
            lo[s] = 0;
            hi[s] = 0;
        }
        __syncthreads();
        const int64_t idx = base + grp;
        const bool act = idx < n;
        int v = 0, own = 0;
        if (act) {
            v = verts[idx];
            own = (int)A.C[v];
            const int64_t e1 = A.off[v + 1];
            for (int64_t e = A.off[v] + lane; e < e1; e += G) {
                const int u = (int)A.col[e];
                if (u == v) continue;
                lds_insert<S>(key, lo, hi, (int)A.C[u], fx_from_double(A.w[e]));
            }
        }
        __syncthreads();
        if (act) {   // uniform within the group
            const double ki = A.kd[v];
            fx_t self = 0;
            double bg = -kMaxFloat;
            int bc = kNone;
            for (int s = lane; s < S; s += G) {
                const int c = key[s];
                if (c < 0) continue;
                const fx_t e = fx_make(lo[s], (int64_t)hi[s]);
                if (c == own) {
                    self = e;
                    continue;
                }
                const double g = gain_of(A, ki, e, c);
                if (better(g, c, bg, bc)) {
                    bg = g;
                    bc = c;
                }
            }
#pragma unroll
            for (int msk = G / 2; msk > 0; msk >>= 1) {
                const double og = __shfl_xor(bg, msk, G);
                const int oc = __shfl_xor(bc, msk, G);
                self += fx_shfl_xor(self, msk, G);
                if (better(og, oc, bg, bc)) {
                    bg = og;
                    bc = oc;
                }
            }
            if (lane == 0) A.Cn[v] = decide(A, v, own, self, bg, bc);
        }
        __syncthreads();
    }
}

// Workgroup-wide argmax + self-weight reduction; result valid in thread 0.
struct BlockRed {
    double g[kBlock / 64];
    int c[kBlock / 64];
    unsigned long long slo[kBlock / 64];
    unsigned long long shi[kBlock / 64];
};

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
}

// One workgroup per vertex, S-slot LDS table.
template <int S>
__global__ void __launch_bounds__(kBlock)
lm_block_kernel(LmArgs A, const int* __restrict__ verts, int64_t n)
{
    __shared__ int key[S];
    __shared__ unsigned long long lo[S];
    __shared__ unsigned long long hi[S];
    __shared__ BlockRed R;
    for (int64_t idx = blockIdx.x; idx < n; idx += gridDim.x) {
        for (int s = threadIdx.x; s < S; s += kBlock) {
            key[s] = -1;
            lo[s] = 0;
            hi[s] = 0;
        }
        __syncthreads();
        const int v = verts[idx];
        const int own = (int)A.C[v];
        const int64_t e1 = A.off[v + 1];
        for (int64_t e = A.off[v] + threadIdx.x; e < e1; e += kBlock) {
            const int u = (int)A.col[e];
            if (u == v) continue;
            lds_insert<S>(key, lo, hi, (int)A.C[u], fx_from_double(A.w[e]));
        }
        __syncthreads();
        const double ki = A.kd[v];
        fx_t self = 0;
        double bg = -kMaxFloat;
        int bc = kNone;
        for (int s = threadIdx.x; s < S; s += kBlock) {
            const int c = key[s];
            if (c < 0) continue;
            const fx_t e = fx_make(lo[s], (int64_t)hi[s]);
            if (c == own) {
                self = e;
                continue;
            }
            const double g = gain_of(A, ki, e, c);
            if (better(g, c, bg, bc)) {
                bg = g;
                bc = c;
            }
        }
        block_reduce(R, self, bg, bc);
        if (threadIdx.x == 0) A.Cn[v] = decide(A, v, own, self, bg, bc);
        __syncthreads();
    }
}

__device__ __forceinline__ int ld_key_agent(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_u64_agent(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per long row; the open-addressing table lives in HBM
// (tab_off[idx] .. tab_off[idx+1], a power of two >= 2 * row length).
__global__ void __launch_bounds__(kBlock)
lm_huge_kernel(LmArgs A, const int* __restrict__ verts, int64_t n, const int64_t* __restrict__ tab_off,
               int* __restrict__ tab_key, fx_t* __restrict__ tab_val)
{
    __shared__ BlockRed R;
    for (int64_t idx = blockIdx.x; idx < n; idx += gridDim.x) {
        const int64_t base = tab_off[idx];
        const int64_t size = tab_off[idx + 1] - base;
        const int sh = 32 - (63 - __builtin_clzll((unsigned long long)size));
        int* key = tab_key + base;
        unsigned long long* val = reinterpret_cast<unsigned long long*>(tab_val + base);
        for (int64_t s = threadIdx.x; s < size; s += kBlock) {
            key[s] = -1;
            val[2 * s] = 0;
            val[2 * s + 1] = 0;
        }
        __threadfence();
        __syncthreads();
        const int v = verts[idx];
        const int own = (int)A.C[v];
        const int64_t e1 = A.off[v + 1];
        for (int64_t e = A.off[v] + threadIdx.x; e < e1; e += kBlock) {
            const int u = (int)A.col[e];
            if (u == v) continue;
            const int c = (int)A.C[u];
            const fx_t wf = fx_from_double(A.w[e]);
            uint64_t h = sh >= 32 ? 0 : (((unsigned)c * kHashMul) >> sh);
            for (;;) {
                const int kk = ld_key_agent(&key[h]);
                if (kk == c) break;
                if (kk < 0) {
                    const int prev = atomicCAS(&key[h], -1, c);
                    if (prev < 0 || prev == c) break;
                }
                h = (h + 1) & (uint64_t)(size - 1);
            }
            fx_atomic_add(&val[2 * h], &val[2 * h + 1], wf);
        }
        __threadfence();
        __syncthreads();
        const double ki = A.kd[v];
        fx_t self = 0;
        double bg = -kMaxFloat;
        int bc = kNone;
        for (int64_t s = threadIdx.x; s < size; s += kBlock) {
            const int c = ld_key_agent(&key[s]);
            if (c < 0) continue;
            const fx_t e = fx_make(ld_u64_agent(&val[2 * s]), (int64_t)ld_u64_agent(&val[2 * s + 1]));
            if (c == own) {
                self = e;
                continue;
            }
            const double g = gain_of(A, ki, e, c);
            if (better(g, c, bg, bc)) {
                bg = g;
                bc = c;
            }
        }
        block_reduce(R, self, bg, bc);
        if (threadIdx.x == 0) A.Cn[v] = decide(A, v, own, self, bg, bc);
        __syncthreads();
    }
}

// update_community_weights + update_community_ids of one batch, fused
// (graph_cuda.cu:1234-1249 then 1191-1196): each thread touches only its own
// vertex's C/C' entries, so the swap needs no second launch.
__global__ void __launch_bounds__(kBlock)
commit_kernel(int64_t v0, int64_t v1, int64_t* __restrict__ C, int64_t* __restrict__ Cn,
              const fx_t* __restrict__ k, fx_t* __restrict__ a)
{
    for (int64_t v = v0 + blockIdx.x * (int64_t)kBlock + threadIdx.x; v < v1;
         v += (int64_t)gridDim.x * kBlock) {
        const int64_t src = C[v];
        const int64_t dst = Cn[v];
        if (src != dst) {
            const fx_t kv = k[v];
            acc_add(a + src, -kv);
            acc_add(a + dst, kv);
        }
        C[v] = dst;
        Cn[v] = src;
    }
}

// ---------------------------------------------------------------------------
// Vertex weights (graph_cuda.cu:302-346), singleton (graph_cuda.cu:410-429)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
vw_small_kernel(const int64_t* __restrict__ off, const double* __restrict__ w,
                const int* __restrict__ small, int64_t n, fx_t* __restrict__ k, fx_t* acc)
{
    fx_t tot = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int v = small[i];
        fx_t s = 0;
        for (int64_t e = off[v]; e < off[v + 1]; ++e) s += fx_from_double(w[e]);
        k[v] = s;
        tot += s;
    }
    tot = wave_sum_fx(tot);
    if ((threadIdx.x & 63) == 0 && tot != 0) acc_add(acc, tot);
}

__global__ void __launch_bounds__(kBlock)
vw_items_kernel(const double* __restrict__ w, const RowItem* __restrict__ items, int64_t n,
                fx_t* __restrict__ k, fx_t* acc)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = wid; i < n; i += nw) {
        const RowItem it = items[i];
        fx_t s = 0;
        for (int64_t e = it.e0 + lane; e < it.e1; e += 64) s += fx_from_double(w[e]);
        s = wave_sum_fx(s);
        if (lane == 0) {
            acc_add(k + it.v, s);
            acc_add(acc, s);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
round_kernel(const fx_t* __restrict__ k, double* __restrict__ kd, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        kd[i] = fx_to_double(k[i]);
}

__global__ void __launch_bounds__(kBlock)
singleton_kernel(int64_t* __restrict__ C, int64_t* __restrict__ Cn, fx_t* __restrict__ a,
                 const fx_t* __restrict__ k, int64_t nv)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock) {
        C[i] = i;
        Cn[i] = i;
        a[i] = k[i];
    }
}

// ---------------------------------------------------------------------------
// Modularity (graph_cuda.cu:2078-2134): Q = I/(2m) - sum_c (a_c/(2m))^2
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
mod_small_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ col,
                 const double* __restrict__ w, const int64_t* __restrict__ C,
                 const int* __restrict__ small, int64_t n, fx_t* acc)
{
    fx_t tot = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int v = small[i];
        const int64_t cv = C[v];
        for (int64_t e = off[v]; e < off[v + 1]; ++e)
            if (C[col[e]] == cv) tot += fx_from_double(w[e]);
    }
    tot = wave_sum_fx(tot);
    if ((threadIdx.x & 63) == 0 && tot != 0) acc_add(acc, tot);
}

__global__ void __launch_bounds__(kBlock)
mod_items_kernel(const int64_t* __restrict__ col, const double* __restrict__ w,
                 const int64_t* __restrict__ C, const RowItem* __restrict__ items, int64_t n,
                 fx_t* acc)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    fx_t tot = 0;
    for (int64_t i = wid; i < n; i += nw) {
        const RowItem it = items[i];
        const int64_t cv = C[it.v];
        for (int64_t e = it.e0 + lane; e < it.e1; e += 64)
            if (C[col[e]] == cv) tot += fx_from_double(w[e]);
    }
    tot = wave_sum_fx(tot);
    if (lane == 0 && tot != 0) acc_add(acc, tot);
}

__global__ void __launch_bounds__(kBlock)
mod_comm_kernel(const fx_t* __restrict__ a, int64_t nv, double twom_d, fx_t* acc)
{
    fx_t tot = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock) {
        const double x = fx_to_double(a[i]) / twom_d;
        tot += fx_from_double(x * x);
    }
    tot = wave_sum_fx(tot);
    if ((threadIdx.x & 63) == 0 && tot != 0) acc_add(acc, tot);
}

__global__ void mod_finalize_kernel(fx_t* acc, double twom_d, double* q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *q = fx_to_double(acc[0]) / twom_d - fx_to_double(acc[1]);
        acc[0] = 0;
        acc[1] = 0;
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan (int64), tiles of 2048 values per 256-thread workgroup.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int64_t block_excl_scan(int64_t x, int64_t* warp_tot, int64_t* total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) warp_tot[wv] = inc;
    __syncthreads();
    int64_t pre = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        if (i < wv) pre += warp_tot[i];
        all += warp_tot[i];
    }
    *total = all;
    return pre + inc - x;
}

__global__ void __launch_bounds__(kBlock)
scan_tile_sums_kernel(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ sums)
{
    __shared__ int64_t wt[kBlock / 64];
    const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = t0 + (int64_t)j * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    int64_t tot;
    block_excl_scan(s, wt, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock)
scan_tile_apply_kernel(const int64_t* in, int64_t* out, int64_t n,
                       const int64_t* __restrict__ tile_off, int64_t* total)
{
    __shared__ int64_t wt[kBlock / 64];
    const int64_t t0 = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = t0 + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = block_excl_scan(s, wt, &tot) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = t0 + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) *total = run;
}

// ---------------------------------------------------------------------------
// Coarsening (graph_gpu.cpp:1151-1304, 1382-1545), all on the device.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
mark_kernel(const int64_t* __restrict__ C, int64_t nv, int64_t* __restrict__ flag)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock)
        flag[C[i]] = 1;
}

__global__ void __launch_bounds__(kBlock)
clustering_kernel(int64_t* __restrict__ clus, int64_t n0, const int64_t* __restrict__ C,
                  const int64_t* __restrict__ newid)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n0;
         i += (int64_t)gridDim.x * kBlock)
        clus[i] = newid[C[clus[i]]];
}

__global__ void __launch_bounds__(kBlock)
member_counts_kernel(const int64_t* __restrict__ C, const int64_t* __restrict__ newid,
                     const int64_t* __restrict__ off, int64_t nv, int64_t* __restrict__ cnt,
                     int64_t* __restrict__ arcs)
{
    for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * kBlock) {
        const int64_t i = newid[C[v]];
        if (cnt) atomicAdd(reinterpret_cast<unsigned long long*>(cnt + i), 1ull);
        const int64_t d = off[v + 1] - off[v];
        if (d) atomicAdd(reinterpret_cast<unsigned long long*>(arcs + i), (unsigned long long)d);
    }
}

__global__ void __launch_bounds__(kBlock)
table_sizes_kernel(const int64_t* __restrict__ arcs, int64_t K, int64_t* __restrict__ tsz)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < K;
         i += (int64_t)gridDim.x * kBlock) {
        int64_t need = arcs[i] < K ? arcs[i] : K;
        int64_t s = 0;
        if (need > 0) {
            s = 2;
            while (s < 2 * need) s <<= 1;
        }
        tsz[i] = s;
    }
}

__device__ __forceinline__ void coarse_insert_one(int64_t i, int j, fx_t wf,
                                                  const int64_t* __restrict__ toff,
                                                  int* __restrict__ tkey, fx_t* __restrict__ tval)
{
    const int64_t base = toff[i];
    const int64_t size = toff[i + 1] - base;
    const int sh = 32 - (63 - __builtin_clzll((unsigned long long)size));
    int* key = tkey + base;
    unsigned long long* val = reinterpret_cast<unsigned long long*>(tval + base);
    uint64_t h = sh >= 32 ? 0 : (((unsigned)j * kHashMul) >> sh);
    for (;;) {
        const int kk = ld_key_agent(&key[h]);
        if (kk == j) break;
        if (kk < 0) {
            const int prev = atomicCAS(&key[h], -1, j);
            if (prev < 0 || prev == j) break;
        }
        h = (h + 1) & (uint64_t)(size - 1);
    }
    fx_atomic_add(&val[2 * h], &val[2 * h + 1], wf);
}

// Every fine arc (u->x, w) goes into coarse row new(u) under key new(x);
// internal arcs and old self-loops become the coarse self-loop
// (graph_gpu.cpp:1233-1250).  Grid-stride over vertices, one wavefront per
// vertex so that long rows stay coalesced.
__global__ void __launch_bounds__(kBlock)
coarse_insert_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ col,
                     const double* __restrict__ w, const int64_t* __restrict__ C,
                     const int64_t* __restrict__ newid, int64_t nv,
                     const int64_t* __restrict__ toff, int* __restrict__ tkey,
                     fx_t* __restrict__ tval)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t v = wid; v < nv; v += nw) {
        const int64_t i = newid[C[v]];
        const int64_t e1 = off[v + 1];
        for (int64_t e = off[v] + lane; e < e1; e += 64)
            coarse_insert_one(i, (int)newid[C[col[e]]], fx_from_double(w[e]), toff, tkey, tval);
    }
}

__global__ void __launch_bounds__(kBlock)
coarse_count_kernel(const int64_t* __restrict__ toff, const int* __restrict__ tkey, int64_t K,
                    int64_t* __restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = wid; i < K; i += nw) {
        int64_t c = 0;
        for (int64_t s = toff[i] + lane; s < toff[i + 1]; s += 64) c += ld_key_agent(&tkey[s]) >= 0;
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) c += __shfl_xor(c, msk, 64);
        if (lane == 0) cnt[i] = c;
    }
}

__global__ void __launch_bounds__(kBlock)
coarse_emit_kernel(const int64_t* __restrict__ toff, const int* __restrict__ tkey,
                   const fx_t* __restrict__ tval, int64_t K, const int64_t* __restrict__ noff,
                   int64_t* __restrict__ ncol, double* __restrict__ nw_out)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = wid; i < K; i += nw) {
        int64_t pos = noff[i];
        for (int64_t s0 = toff[i]; s0 < toff[i + 1]; s0 += 64) {
            const int64_t s = s0 + lane;
            const int c = s < toff[i + 1] ? ld_key_agent(&tkey[s]) : -1;
            const unsigned long long m = __ballot(c >= 0);
            if (c >= 0) {
                const unsigned long long below = lane ? (m & ((1ull << lane) - 1)) : 0ull;
                const int64_t p = pos + __popcll(below);
                const unsigned long long* vv = reinterpret_cast<const unsigned long long*>(tval + s);
                ncol[p] = c;
                nw_out[p] = fx_to_double(fx_make(ld_u64_agent(vv), (int64_t)ld_u64_agent(vv + 1)));
            }
            pos += __popcll(m);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
void launch_fill_i64(int64_t* p, int64_t val, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    fill_kernel<int64_t><<<grid_for(n, kBlock * 4), kBlock, 0, s>>>(p, val, n);
}

void launch_fill_i32(int32_t* p, int32_t val, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    fill_kernel<int32_t><<<grid_for(n, kBlock * 4), kBlock, 0, s>>>(p, val, n);
}

void launch_fill_fx(fx_t* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    fill_kernel<fx_t><<<grid_for(n, kBlock * 4), kBlock, 0, s>>>(p, (fx_t)0, n);
}

void launch_vertex_weights(const int64_t* off, const double* w, const int* small, int64_t nsmall,
                           const RowItem* items, int64_t nitems, fx_t* k, double* kd,
                           int64_t nv, fx_t* acc, hipStream_t s)
{
    if (nv <= 0) return;
    launch_fill_fx(k, nv, s);
    if (nsmall > 0)
        vw_small_kernel<<<grid_for(nsmall, kBlock), kBlock, 0, s>>>(off, w, small, nsmall, k, acc);
    if (nitems > 0)
        vw_items_kernel<<<grid_for(nitems, kBlock / 64), kBlock, 0, s>>>(w, items, nitems, k, acc);
    round_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(k, kd, nv);
}

void launch_singleton(int64_t* C, int64_t* Cn, fx_t* a, const fx_t* k, int64_t nv, hipStream_t s)
{
    if (nv <= 0) return;
    singleton_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, Cn, a, k, nv);
}

void launch_lm_bin(int bin, const LmArgs& A, const int* verts, int64_t n,
                   const int64_t* tab_off, int* tab_key, fx_t* tab_val, hipStream_t s)
{
    if (n <= 0) return;
    switch (bin) {
    case BIN_THREAD:
        lm_thread_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_G8:
        lm_group_kernel<8, 32><<<grid_for(n, kBlock / 8), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_G16:
        lm_group_kernel<16, 64><<<grid_for(n, kBlock / 16), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_G32:
        lm_group_kernel<32, 128><<<grid_for(n, kBlock / 32), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_G64:
        lm_group_kernel<64, 256><<<grid_for(n, kBlock / 64), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_B256:
        lm_block_kernel<1024><<<grid_for(n, 1), kBlock, 0, s>>>(A, verts, n);
        break;
    case BIN_HUGE:
        lm_huge_kernel<<<grid_for(n, 1), kBlock, 0, s>>>(A, verts, n, tab_off, tab_key, tab_val);
        break;
    default:
        break;
    }
}

void launch_commit(int64_t v0, int64_t v1, int64_t* C, int64_t* Cn, const fx_t* k, fx_t* a,
                   hipStream_t s)
{
    if (v1 <= v0) return;
    commit_kernel<<<grid_for(v1 - v0, kBlock), kBlock, 0, s>>>(v0, v1, C, Cn, k, a);
}

void launch_modularity(const int64_t* off, const int64_t* col, const double* w, const int64_t* C,
                       const fx_t* a, int64_t nv, const int* small, int64_t nsmall,
                       const RowItem* items, int64_t nitems, double twom_d, fx_t* acc,
                       double* q, hipStream_t s)
{
    if (nsmall > 0)
        mod_small_kernel<<<grid_for(nsmall, kBlock), kBlock, 0, s>>>(off, col, w, C, small, nsmall, acc);
    if (nitems > 0)
        mod_items_kernel<<<grid_for(nitems, kBlock / 64), kBlock, 0, s>>>(col, w, C, items, nitems, acc);
    if (nv > 0)
        mod_comm_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(a, nv, twom_d, acc + 1);
    mod_finalize_kernel<<<1, 64, 0, s>>>(acc, twom_d, q);
}

size_t scan_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    int64_t m = n;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        bytes += (size_t)m * sizeof(int64_t);
    }
    return bytes + 64;
}

void launch_exclusive_scan(const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                           void* temp, hipStream_t s)
{
    if (n <= 0) {
        if (total) launch_fill_i64(total, 0, 1, s);
        return;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        scan_tile_apply_kernel<<<1, kBlock, 0, s>>>(in, out, n, nullptr, total);
        return;
    }
    int64_t* sums = reinterpret_cast<int64_t*>(temp);
    void* rest = sums + tiles;
    scan_tile_sums_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(in, n, sums);
    launch_exclusive_scan(sums, sums, tiles, nullptr, rest, s);
    scan_tile_apply_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(in, out, n, sums, total);
}

void launch_mark_communities(const int64_t* C, int64_t nv, int64_t* flag, hipStream_t s)
{
    launch_fill_i64(flag, 0, nv, s);
    if (nv > 0) mark_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, nv, flag);
}

void launch_update_clustering(int64_t* clus, int64_t n0, const int64_t* C, const int64_t* newid,
                              hipStream_t s)
{
    if (n0 > 0) clustering_kernel<<<grid_for(n0, kBlock), kBlock, 0, s>>>(clus, n0, C, newid);
}

void launch_member_counts(const int64_t* C, const int64_t* newid, const int64_t* off, int64_t nv,
                          int64_t* cnt, int64_t* arcs, hipStream_t s)
{
    if (nv > 0)
        member_counts_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, newid, off, nv, cnt, arcs);
}

void launch_table_sizes(const int64_t* arcs, int64_t K, int64_t* tsz, hipStream_t s)
{
    if (K > 0) table_sizes_kernel<<<grid_for(K, kBlock), kBlock, 0, s>>>(arcs, K, tsz);
}

void launch_coarse_insert(const int64_t* off, const int64_t* col, const double* w, const int64_t* C,
                          const int64_t* newid, int64_t nv, const int64_t* toff, int* tkey,
                          fx_t* tval, hipStream_t s)
{
    if (nv > 0)
        coarse_insert_kernel<<<grid_for(nv, kBlock / 64), kBlock, 0, s>>>(off, col, w, C, newid, nv,
                                                                          toff, tkey, tval);
}

void launch_coarse_count(const int64_t* toff, const int* tkey, int64_t K, int64_t* cnt,
                         hipStream_t s)
{
    if (K > 0) coarse_count_kernel<<<grid_for(K, kBlock / 64), kBlock, 0, s>>>(toff, tkey, K, cnt);
}

void launch_coarse_emit(const int64_t* toff, const int* tkey, const fx_t* tval, int64_t K,
                        const int64_t* noff, int64_t* ncol, double* nw, hipStream_t s)
{
    if (K > 0)
        coarse_emit_kernel<<<grid_for(K, kBlock / 64), kBlock, 0, s>>>(toff, tkey, tval, K, noff,
                                                                       ncol, nw);
}

}  // namespace nido
