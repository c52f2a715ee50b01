// kernels.hip -- per-phase and per-iteration kernels other than local moving:
// vertex weights and 2m, singleton partition, modularity, exclusive scan and
// the coarsening (renumber + coarse CSR rebuild) of the batched Louvain path.
#include "device_common.h"

namespace nido {
namespace {

// p[i] = i.
__global__ void iota_kernel(int64_t* __restrict__ p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = i;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, T val, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = val;
}

// ---------------------------------------------------------------------------
// Vertex weights (graph_cuda.cu:302-346), singleton (graph_cuda.cu:410-429)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
vw_small_kernel(const int64_t* __restrict__ off, const double* __restrict__ w,
                const int* __restrict__ small, int64_t n, fx_t* __restrict__ k, int* __restrict__ src)
{
    // kRowLanes lanes per row (at most kSmallRowArcs arcs), each taking
    // every kRowLanes-th arc: neighbouring lanes read neighbouring arcs, and
    // the group's first lane adds up their exact partial sums (any order
    // gives the same sum).  The partials meet in LDS: summed with 128-bit
    // lane shuffles of width kRowLanes instead, some k_v came out wrong on
    // gfx950, a different few on every run.
    __shared__ fx_t part[kBlock];
    const int lane = threadIdx.x & (kRowLanes - 1);
    const int64_t stride = (int64_t)gridDim.x * (kBlock / kRowLanes);
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / kRowLanes) + threadIdx.x / kRowLanes; i < n; i += stride) {
        const int v = small[i];   // uniform within the group
        const int64_t e1 = off[v + 1];
        fx_t s = 0;
        for (int64_t e = off[v] + lane; e < e1; e += kRowLanes) {
            s += fx_from_double(w[e]);
            src[e] = v;
        }
        part[threadIdx.x] = s;
        wave_lds_sync();   // a group's lanes share one wavefront
        if (lane == 0) {
            fx_t t = 0;
#pragma unroll
            for (int j = 0; j < kRowLanes; ++j) t += part[threadIdx.x + j];
            k[v] = t;
        }
        wave_lds_sync();   // read before the group's next row overwrites them
    }
}

__global__ void __launch_bounds__(kBlock)
vw_items_kernel(const double* __restrict__ w, const RowItem* __restrict__ items, int64_t n,
                fx_t* __restrict__ k, int* __restrict__ src)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = wid; i < n; i += nw) {
        const RowItem it = items[i];
        fx_t s = 0;
        for (int64_t e = it.e0 + lane; e < it.e1; e += 64) {
            s += fx_from_double(w[e]);
            src[e] = (int)it.v;
        }
        s = wave_sum_fx(s);
        if (lane == 0) acc_add(k + it.v, s);
    }
}

// kd = FP64(k); part[block] = this block's sum of k; *maxabs = max |kd| as
// raw bits (|x| bit patterns order like the values).
__global__ void __launch_bounds__(kBlock)
round_kernel(const fx_t* __restrict__ k, double* __restrict__ kd, int64_t n,
             unsigned long long* maxabs, fx_t* __restrict__ part)
{
    __shared__ fx_t red[kBlock / 64];
    unsigned long long mx = 0;
    fx_t s = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const fx_t ki = k[i];
        const double x = fx_to_double(ki);
        kd[i] = x;
        s += ki;
        const unsigned long long b = nido_dbits(x) & 0x7fffffffffffffffull;
        mx = b > mx ? b : mx;
    }
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) {
        const unsigned long long o = __shfl_xor(mx, msk, 64);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(maxabs, mx);
    s = block_sum_fx(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = sum of part[0..n) (one workgroup).
__global__ void __launch_bounds__(kBlock)
sum_partials_kernel(const fx_t* __restrict__ part, int n, fx_t* out)
{
    __shared__ fx_t red[kBlock / 64];
    fx_t s = 0;
    for (int i = threadIdx.x; i < n; i += kBlock) s += part[i];
    s = block_sum_fx(s, red);
    if (threadIdx.x == 0) *out = s;
}

__global__ void __launch_bounds__(kBlock)
singleton_kernel(int64_t* __restrict__ C, int64_t* __restrict__ Cn, int* __restrict__ C32,
                 fx_t* __restrict__ a, const fx_t* __restrict__ k, int64_t nv)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock) {
        C[i] = i;
        Cn[i] = i;
        C32[i] = (int)i;
        a[i] = k[i];
    }
}

__global__ void __launch_bounds__(kBlock)
narrow_kernel(const int64_t* __restrict__ col, int* __restrict__ col32, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        col32[i] = (int)col[i];
}

__global__ void __launch_bounds__(kBlock)
widen_kernel(const int* __restrict__ col32, int64_t* __restrict__ col, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        col[i] = col32[i];
}

// The grid every weight lies on: *fneed = max over nonzero weights of the
// fractional bits the weight needs (w = sig 2^(ex - 1075), lowest set bit
// 2^(ex - 1075 + ctz(sig)); 2000 for a subnormal or non-finite weight) and
// *wmax = the bits of max |w| (|x| bit patterns order like the values).
__global__ void __launch_bounds__(kBlock)
weight_grid_kernel(const double* __restrict__ w, int64_t ne, int* fneed, unsigned long long* wmax)
{
    int f = INT_MIN;
    unsigned long long mx = 0;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
        const unsigned long long b = nido_dbits(w[e]) & 0x7fffffffffffffffull;
        if (b == 0) continue;
        const int ex = (int)(b >> 52);
        int need = 2000;
        if (ex != 0 && ex != 0x7ff) {
            const unsigned long long sig = (b & 0x000fffffffffffffull) | 0x0010000000000000ull;
            need = 1075 - ex - __builtin_ctzll(sig);
        }
        f = need > f ? need : f;
        mx = b > mx ? b : mx;
    }
#pragma unroll
    for (int msk = 32; msk > 0; msk >>= 1) {
        const int of = __shfl_xor(f, msk, 64);
        const unsigned long long om = __shfl_xor(mx, msk, 64);
        f = of > f ? of : f;
        mx = om > mx ? om : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        if (f != INT_MIN) atomicMax(fneed, f);
        if (mx) atomicMax(wmax, mx);
    }
}

// wq[e] = w[e] * 2^F, exact: every weight is a multiple of 2^-F and
// |w| 2^F < 2^63 (checked by the caller).
__global__ void __launch_bounds__(kBlock)
to_fixed_kernel(const double* __restrict__ w, int64_t* __restrict__ wq, int64_t ne, double scale)
{
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock)
        wq[e] = (int64_t)(w[e] * scale);
}

// ---------------------------------------------------------------------------
// Modularity (graph_cuda.cu:2078-2134): Q = I/(2m) - sum_c (a_c/(2m))^2.
// Arc-parallel over the whole CSR (tiles of 256 x kModUnroll arcs, every load
// coalesced except the two C gathers), then the community sweep; each block
// leaves two partials and mod_finalize_kernel adds them up.  Ident: C is
// the identity (the singleton partition every phase starts from), so an arc
// is internal exactly when it is a self arc and the C gathers are left out;
// the sums are the same integers either way.
// ---------------------------------------------------------------------------
constexpr int kModUnroll = 8;

template <bool Ident>
__global__ void __launch_bounds__(kBlock)
mod_kernel(const int* __restrict__ src, const int* __restrict__ col,
           const double* __restrict__ w, const int* __restrict__ C, int64_t ne,
           const fx_t* __restrict__ a, int64_t nv, double twom_d, fx_t* __restrict__ part,
           fx_t* __restrict__ term)
{
    __shared__ fx_t red[kBlock / 64];
    fx_t I = 0, T = 0;
    const int64_t tile = (int64_t)kBlock * kModUnroll;
    for (int64_t t0 = blockIdx.x * tile; t0 < ne; t0 += (int64_t)gridDim.x * tile) {
        int s[kModUnroll];
        int u[kModUnroll];
#pragma unroll
        for (int j = 0; j < kModUnroll; ++j) {
            const int64_t e = t0 + j * kBlock + threadIdx.x;
            s[j] = e < ne ? src[e] : 0;
            u[j] = e < ne ? col[e] : 0;
        }
        bool in[kModUnroll];
#pragma unroll
        for (int j = 0; j < kModUnroll; ++j) {
            const int64_t e = t0 + j * kBlock + threadIdx.x;
            if constexpr (Ident) in[j] = e < ne && s[j] == u[j];
            else in[j] = e < ne && C[s[j]] == C[u[j]];    // self arcs are internal
        }
#pragma unroll
        for (int j = 0; j < kModUnroll; ++j)
            if (in[j]) I += fx_from_double(w[t0 + j * kBlock + threadIdx.x]);
    }
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock) {
        const double x = fx_to_double(a[i]) / twom_d;
        const fx_t t = fx_from_double(x * x);
        T += t;
        if (term) term[i] = t;
    }
    I = block_sum_fx(I, red);
    T = block_sum_fx(T, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = I;
        part[2 * blockIdx.x + 1] = T;
    }
}

__global__ void __launch_bounds__(kBlock)
mod_finalize_kernel(const fx_t* __restrict__ part, int nblocks, double twom_d, double* q, fx_t* acc)
{
    __shared__ fx_t red[kBlock / 64];
    fx_t I = 0, T = 0;
    for (int i = threadIdx.x; i < nblocks; i += kBlock) {
        I += part[2 * i];
        T += part[2 * i + 1];
    }
    I = block_sum_fx(I, red);
    T = block_sum_fx(T, red);
    if (threadIdx.x == 0) {
        *q = fx_to_double(I) / twom_d - fx_to_double(T);
        if (acc) {
            acc[0] = I;
            acc[1] = T;
        }
    }
}

// Modularity of a phase whose arcs stream through a window: the internal
// weight of rows [v0, v1) (col / w may be based so that the window's arcs
// keep their CSR indices), added exactly into acc[0]; then every
// community's square term into acc[1]; then Q.  Each thread takes runs of
// kModRun consecutive arcs of [off[v0], off[v1]): a binary search over
// off[v0 .. v1 - 1] finds the row of a run's first arc, then the run walks
// the rows forward, so a long row is spread over many threads and short
// rows leave no lane idle.
constexpr int kModRun = 16;

__global__ void __launch_bounds__(kBlock)
mod_rows_kernel(const int64_t* __restrict__ off, const int* __restrict__ col, const double* __restrict__ w,
                const int* __restrict__ C, int64_t v0, int64_t v1, fx_t* __restrict__ acc)
{
    __shared__ fx_t red[kBlock / 64];
    const int64_t e0 = off[v0], e1 = off[v1];
    fx_t I = 0;
    for (int64_t a = e0 + (blockIdx.x * (int64_t)kBlock + threadIdx.x) * kModRun; a < e1;
         a += (int64_t)gridDim.x * kBlock * kModRun) {
        // the last row r of [v0, v1) with off[r] <= a: off[v0] <= a < off[v1]
        int64_t lo = v0, hi = v1 - 1;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo + 1) / 2;
            if (off[mid] <= a) lo = mid;
            else hi = mid - 1;
        }
        int64_t r = lo, rend = off[r + 1];
        int cv = C[r];
        const int64_t z = a + kModRun < e1 ? a + kModRun : e1;
        for (int64_t e = a; e < z; ++e) {
            while (rend <= e) {   // e < off[v1]: r stays below v1
                ++r;
                rend = off[r + 1];
                cv = C[r];
            }
            if (C[col[e]] == cv) I += fx_from_double(w[e]);   // self arcs are internal
        }
    }
    I = block_sum_fx(I, red);
    if (threadIdx.x == 0 && I != 0) acc_add(acc, I);
}

__global__ void __launch_bounds__(kBlock)
mod_terms_kernel(const fx_t* __restrict__ a, int64_t nv, double twom_d, fx_t* __restrict__ acc,
                 fx_t* __restrict__ term)
{
    __shared__ fx_t red[kBlock / 64];
    fx_t T = 0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv; i += (int64_t)gridDim.x * kBlock) {
        const double x = fx_to_double(a[i]) / twom_d;
        const fx_t t = fx_from_double(x * x);
        if (term) term[i] = t;
        T += t;
    }
    T = block_sum_fx(T, red);
    if (threadIdx.x == 0 && T != 0) acc_add(acc + 1, T);
}

__global__ void mod_acc_finish_kernel(const fx_t* __restrict__ acc, double twom_d, double* q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *q = fx_to_double(acc[0]) / twom_d - fx_to_double(acc[1]);
}

// The multisets {(src, dst, w)} and {(dst, src, w)} of the arcs, compared
// through sums (mod 2^64) of two keyed hashes of each form: equal when every
// arc has a reverse arc of the same weight, and unequal otherwise except
// with probability about 2^-128.  The mixer is the output function of
// splitmix64 (Vigna's public-domain splitmix64.c; Stafford's "Mix13"
// constants); the two keys are the first 64 fraction bits of sqrt(2) and
// sqrt(3).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long arc_hash(unsigned a, unsigned b, unsigned long long wb,
                                                       unsigned long long key)
{
    return mix64(mix64((((unsigned long long)a << 32) | b) ^ key) ^ wb);
}

__global__ void __launch_bounds__(kBlock)
sym_hash_kernel(const int* __restrict__ src, const int* __restrict__ col, const double* __restrict__ w, int64_t ne,
                unsigned long long* __restrict__ out)
{
    constexpr unsigned long long k1 = 0x6a09e667f3bcc909ull, k2 = 0xbb67ae8584caa73bull;
    unsigned long long h[4] = {0, 0, 0, 0};
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
        const unsigned a = (unsigned)src[e], b = (unsigned)col[e];
        const unsigned long long wb = (unsigned long long)__double_as_longlong(w[e]);
        h[0] += arc_hash(a, b, wb, k1);
        h[1] += arc_hash(b, a, wb, k1);
        h[2] += arc_hash(a, b, wb, k2);
        h[3] += arc_hash(b, a, wb, k2);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) h[i] += __shfl_xor(h[i], msk, 64);
        if ((threadIdx.x & 63) == 0 && h[i] != 0) atomicAdd(out + i, h[i]);
    }
}

__global__ void __launch_bounds__(kBlock)
mod_delta_kernel(int b, int64_t v0, int64_t v1, const int64_t* __restrict__ Cn, const int* __restrict__ C32,
                 const int* __restrict__ col, const double* __restrict__ w, const fx_t* __restrict__ a,
                 double twom_d, ModDelta D)
{
    __shared__ fx_t red[kBlock / 64];
    delta_body(b, v0, v1, Cn, C32, col, w, a, twom_d, D, (int)blockIdx.x, (int)gridDim.x, red);
}

__global__ void __launch_bounds__(kBlock)
mod_finish_kernel(ModDelta D, int nb, double twom_d, double* q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *q = fx_to_double(D.acc[0]) / twom_d - fx_to_double(D.acc[1]);
    for (int i = threadIdx.x; i < 2 * nb; i += kBlock) D.counts[i] = 0;
}

// Per graph: the graphs that sat out keep their last q[g] (and their
// accumulators describe no partition once they have been rolled back).
__global__ void __launch_bounds__(kBlock)
mod_finish_graphs_kernel(ModDelta D, int nb, int G, double* __restrict__ q, const unsigned char* __restrict__ act)
{
    __shared__ unsigned long long red[kBlock / 64];
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock) {
        if (!act[g]) continue;
        const double tw = D.twomg[g];
        q[g] = tw == 0 ? 0.0 : fx_to_double(D.acc[2 * g]) / tw - fx_to_double(D.acc[2 * g + 1]);
    }
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < 2 * nb; i += kBlock) D.counts[i] = 0;
        if (D.moved) {
            const unsigned long long t = sum_moved_slots(D.moved, red);
            if (threadIdx.x == 0) reinterpret_cast<unsigned long long*>(q)[G] = t;
        }
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
mark_kernel(const int64_t* __restrict__ C, int64_t nv, int64_t* __restrict__ flag,
            const int* __restrict__ gid, const unsigned char* __restrict__ keep)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * kBlock)
        if (!gid || keep[gid[i]]) flag[C[i]] = 1;
}

__global__ void __launch_bounds__(kBlock)
clustering_kernel(int64_t* __restrict__ clus, int64_t n0, const int64_t* __restrict__ C,
                  const int64_t* __restrict__ newid, const int* __restrict__ gid0,
                  const unsigned char* __restrict__ keep)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n0;
         i += (int64_t)gridDim.x * kBlock)
        if (!gid0 || keep[gid0[i]]) clus[i] = newid[C[clus[i]]];
}

__global__ void __launch_bounds__(kBlock)
widen_local_kernel(const ArcPiece* __restrict__ pieces, int64_t n, const uint16_t* __restrict__ local,
                   int* __restrict__ col32)
{
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const ArcPiece p = pieces[i];
        for (int64_t e = p.e0 + threadIdx.x; e < p.e1; e += kBlock) col32[e] = (int)(p.base + local[e]);
    }
}

// out[i] = clus[i] - base[graph of i]: every graph's clustering in its own
// ids.
__global__ void __launch_bounds__(kBlock)
local_ids_kernel(const int64_t* __restrict__ clus, int64_t n0, const int* __restrict__ gid0,
                 const int64_t* __restrict__ base, int64_t* __restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n0; i += (int64_t)gridDim.x * kBlock)
        out[i] = clus[i] - base[gid0[i]];
}

// The coarse rows of a workgroup's kBlock consecutive vertices, grouped in
// an LDS table of 2 * kBlock slots (at most kBlock keys: half full at most,
// so probing ends): vertices of one community are often neighbours in id,
// and each distinct row then takes one global atomic instead of one per
// vertex.  Returns the slot of row i (i >= 0).
constexpr int kGroupSlots = 2 * kBlock;
__device__ __forceinline__ int group_slot(int* key, int i)
{
    unsigned h = ((unsigned)i * kHashMul) >> (32 - ilog2c(kGroupSlots));
    for (;;) {
        const int kk = key[h];
        if (kk == i) return (int)h;
        if (kk < 0) {
            const int prev = atomicCAS(&key[h], -1, i);
            if (prev < 0 || prev == i) return (int)h;
        }
        h = (h + 1) & (kGroupSlots - 1);
    }
}

__global__ void __launch_bounds__(kBlock)
member_counts_kernel(const int64_t* __restrict__ C, const int64_t* __restrict__ newid,
                     const int64_t* __restrict__ off, int64_t nv, int64_t* __restrict__ cnt,
                     int64_t* __restrict__ arcs, const int* __restrict__ gid,
                     const unsigned char* __restrict__ keep)
{
    __shared__ int key[kGroupSlots];
    __shared__ unsigned long long nm[kGroupSlots], na[kGroupSlots];
    for (int64_t base = blockIdx.x * (int64_t)kBlock; base < nv; base += (int64_t)gridDim.x * kBlock) {
        for (int s = threadIdx.x; s < kGroupSlots; s += kBlock) {
            key[s] = -1;
            nm[s] = 0;
            na[s] = 0;
        }
        __syncthreads();
        const int64_t v = base + threadIdx.x;
        if (v < nv && (!gid || keep[gid[v]])) {
            const int h = group_slot(key, (int)newid[C[v]]);
            atomicAdd(&nm[h], 1ull);
            atomicAdd(&na[h], (unsigned long long)(off[v + 1] - off[v]));
        }
        __syncthreads();
        for (int s = threadIdx.x; s < kGroupSlots; s += kBlock) {
            const int i = key[s];
            if (i < 0) continue;
            if (cnt) atomicAdd(reinterpret_cast<unsigned long long*>(cnt + i), nm[s]);
            if (na[s]) atomicAdd(reinterpret_cast<unsigned long long*>(arcs + i), na[s]);
        }
        __syncthreads();   // the table is cleared for the next chunk
    }
}

// Coarse row i with arcs[i] <= kLdsRowArcs is aggregated in LDS and needs
// csz[i] = min(arcs, K) compact entries; a longer one gets an HBM hash table
// of tsz[i] slots (a power of two >= twice its possible keys).
__global__ void __launch_bounds__(kBlock)
coarse_sizes_kernel(const int64_t* __restrict__ arcs, int64_t K, int64_t* __restrict__ tsz,
                    int64_t* __restrict__ csz)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < K;
         i += (int64_t)gridDim.x * kBlock) {
        const int64_t need = arcs[i] < K ? arcs[i] : K;
        int64_t s = 0;
        if (arcs[i] > kLdsRowArcs) {
            s = 2;
            while (s < 2 * need) s <<= 1;
        }
        tsz[i] = s;
        csz[i] = arcs[i] > kLdsRowArcs ? 0 : need;
    }
}

// members[mstart[i] + r] = the vertices of coarse row i, in no set order,
// and hlist[0, *nh) the vertices whose row has an HBM table (toff[i + 1] >
// toff[i]).  Each workgroup reserves a block of its rows' member slots with
// one atomic per distinct row (grouped in LDS as in member_counts_kernel).
__global__ void __launch_bounds__(kBlock)
member_scatter_kernel(const int64_t* __restrict__ C, const int64_t* __restrict__ newid, int64_t nv,
                      const int64_t* __restrict__ mstart, unsigned long long* __restrict__ mfill,
                      int* __restrict__ members, const int64_t* __restrict__ toff, int* __restrict__ hlist,
                      unsigned long long* __restrict__ nh, const int* __restrict__ gid,
                      const unsigned char* __restrict__ keep)
{
    __shared__ int key[kGroupSlots];
    __shared__ unsigned nm[kGroupSlots];
    __shared__ unsigned long long at[kGroupSlots];
    __shared__ unsigned s_nh;
    __shared__ unsigned long long s_hbase;
    for (int64_t base = blockIdx.x * (int64_t)kBlock; base < nv; base += (int64_t)gridDim.x * kBlock) {
        for (int s = threadIdx.x; s < kGroupSlots; s += kBlock) {
            key[s] = -1;
            nm[s] = 0;
        }
        if (threadIdx.x == 0) s_nh = 0;
        __syncthreads();
        const int64_t v = base + threadIdx.x;
        int64_t i = -1;
        int h = 0;
        unsigned rank = 0, hrank = 0;
        bool hbm = false;
        if (v < nv && (!gid || keep[gid[v]])) {
            i = newid[C[v]];
            h = group_slot(key, (int)i);
            rank = atomicAdd(&nm[h], 1u);
            hbm = toff[i + 1] > toff[i];
            if (hbm) hrank = atomicAdd(&s_nh, 1u);
        }
        __syncthreads();
        for (int s = threadIdx.x; s < kGroupSlots; s += kBlock)
            if (key[s] >= 0) at[s] = atomicAdd(mfill + key[s], (unsigned long long)nm[s]);
        if (threadIdx.x == 0 && s_nh) s_hbase = atomicAdd(nh, (unsigned long long)s_nh);
        __syncthreads();
        if (i >= 0) members[mstart[i] + (int64_t)(at[h] + rank)] = (int)v;
        if (hbm) hlist[s_hbase + hrank] = (int)v;
        __syncthreads();   // the table is cleared for the next chunk
    }
}

// Coarse rows of at most kLdsRowArcs arcs, one workgroup each: the arcs of
// its members (their concatenation, 256 at a time, each thread locating its
// member by binary search over the scanned lengths) go into an LDS table
// keyed by the coarse target, with exact Q64.64 sums; the arcs into the row
// itself are summed per wavefront first and inserted once.  The row's
// distinct targets land compactly at ckey/cw[cpos[i] ...) and their count
// in cnt[i].  Keys whose weights sum to zero are kept, as in the HBM tables.
constexpr int kLdsSlots = 2 * kLdsRowArcs;
__global__ void __launch_bounds__(kBlock)
coarse_lds_kernel(const int64_t* __restrict__ arcs, const int64_t* __restrict__ mstart,
                  const int* __restrict__ members, const int64_t* __restrict__ off,
                  const int64_t* __restrict__ col, const double* __restrict__ w, const int64_t* __restrict__ C,
                  const int64_t* __restrict__ newid, int64_t K, const int64_t* __restrict__ cpos,
                  int* __restrict__ ckey, double* __restrict__ cw, int64_t* __restrict__ cnt)
{
    __shared__ int key[kLdsSlots];
    __shared__ unsigned long long lo[kLdsSlots], hi[kLdsSlots];
    __shared__ int klist[kLdsRowArcs];
    __shared__ int s_pre[kBlock], wtot[kBlock / 64];
    __shared__ int64_t s_e0[kBlock];
    __shared__ int s_nk;
    const int lane = threadIdx.x & 63;
    for (int64_t i = blockIdx.x; i < K; i += gridDim.x) {
        const int64_t a = arcs[i];
        if (a > kLdsRowArcs) continue;   // uniform: an HBM-table row
        const int64_t need = a < K ? a : K;
        int lg = 6;
        while ((1ll << lg) < 2 * need) ++lg;
        const int ts = 1 << lg;          // <= kLdsSlots
        for (int s = threadIdx.x; s < ts; s += kBlock) {
            key[s] = -1;
            lo[s] = 0;
            hi[s] = 0;
        }
        if (threadIdx.x == 0) s_nk = 0;
        __syncthreads();
        // ends: at most need <= ts / 2 distinct keys
        auto insert = [&](int j, fx_t x) {
            unsigned h = ((unsigned)j * kHashMul) >> (32 - lg);
            for (;;) {
                const int kk = key[h];
                if (kk == j) break;
                if (kk < 0) {
                    const int prev = atomicCAS(&key[h], -1, j);
                    if (prev < 0) {
                        klist[atomicAdd(&s_nk, 1)] = (int)h;
                        break;
                    }
                    if (prev == j) break;
                }
                h = (h + 1) & (unsigned)(ts - 1);
            }
            fx_atomic_add(&lo[h], &hi[h], x);
        };
        fx_t self = 0;
        bool any_self = false;           // uniform in the wavefront
        const int64_t m1 = mstart[i + 1];
        for (int64_t m0 = mstart[i]; m0 < m1; m0 += kBlock) {   // uniform
            const int64_t mi = m0 + threadIdx.x;
            int d = 0;
            int64_t e0 = 0;
            if (mi < m1) {
                const int v = members[mi];
                e0 = off[v];
                d = (int)(off[v + 1] - e0);   // <= a <= kLdsRowArcs
            }
            int tot;
            const int pre = block_excl_scan_i32(d, wtot, &tot);
            s_pre[threadIdx.x] = pre;
            s_e0[threadIdx.x] = e0;
            __syncthreads();
            const int nm = m1 - m0 < kBlock ? (int)(m1 - m0) : kBlock;
            int top = 1;
            while (2 * top < nm) top *= 2;
            for (int p0 = 0; p0 < tot; p0 += kBlock) {   // uniform
                const int p = p0 + threadIdx.x;
                int j = -1;
                fx_t x = 0;
                if (p < tot) {
                    int l = 0;                 // the last member whose arcs start at or before p
                    for (int st = top; st > 0; st >>= 1)
                        if (l + st < nm && s_pre[l + st] <= p) l += st;
                    const int64_t e = s_e0[l] + (p - s_pre[l]);
                    j = (int)newid[C[col[e]]];
                    x = fx_from_double(w[e]);
                }
                const bool own = j == (int)i;
                any_self |= __ballot(own) != 0;
                if (own) self += x;
                else if (j >= 0) insert(j, x);
            }
            __syncthreads();               // s_pre / s_e0 / wtot are rewritten by the next chunk
        }
        self = wave_sum_fx(self);
        if (lane == 0 && any_self) insert((int)i, self);
        __syncthreads();
        const int nk = s_nk;
        const int64_t base = cpos[i];
        for (int r = threadIdx.x; r < nk; r += kBlock) {
            const int sl = klist[r];
            ckey[base + r] = key[sl];
            cw[base + r] = fx_to_double(fx_make(lo[sl], (int64_t)hi[sl]));
        }
        if (threadIdx.x == 0) cnt[i] = nk;
        __syncthreads();                   // the table is cleared for the next row
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
        const int kk = ld_i32_agent(&key[h]);
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
// vertex so that long rows stay coalesced.  The arcs of a row that land in
// its own coarse vertex (most of them once communities have formed) are
// summed across the wavefront first and inserted once: otherwise they all
// take device atomics on the same self-loop slot.  The sums are exact
// integers, so the grouping does not change them.
__global__ void __launch_bounds__(kBlock)
coarse_insert_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ col,
                     const double* __restrict__ w, const int64_t* __restrict__ C,
                     const int64_t* __restrict__ newid, const int* __restrict__ hlist,
                     const unsigned long long* __restrict__ nh, const int64_t* __restrict__ toff,
                     int* __restrict__ tkey, fx_t* __restrict__ tval)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    const int64_t n = (int64_t)*nh;
    for (int64_t t = wid; t < n; t += nw) {
        const int64_t v = hlist[t];              // a vertex whose coarse row has an HBM table
        const int64_t i = newid[C[v]];
        const int64_t e1 = off[v + 1];
        for (int64_t e0 = off[v]; e0 < e1; e0 += 64) {   // uniform trip count in the wavefront
            const int64_t e = e0 + lane;
            const bool on = e < e1;
            const int j = on ? (int)newid[C[col[e]]] : -1;
            const fx_t x = on ? fx_from_double(w[e]) : fx_t(0);
            const bool self = on && j == (int)i;
            const fx_t sx = wave_sum_fx(self ? x : fx_t(0));   // every lane takes part
            const bool any_self = __ballot(self) != 0;
            if (on && !self) coarse_insert_one(i, j, x, toff, tkey, tval);
            if (lane == 0 && any_self) coarse_insert_one(i, (int)i, sx, toff, tkey, tval);
        }
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
        if (toff[i + 1] == toff[i]) continue;   // uniform: counted by the LDS aggregation
        int64_t c = 0;
        for (int64_t s = toff[i] + lane; s < toff[i + 1]; s += 64) c += ld_i32_agent(&tkey[s]) >= 0;
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) c += __shfl_xor(c, msk, 64);
        if (lane == 0) cnt[i] = c;
    }
}

__global__ void __launch_bounds__(kBlock)
coarse_emit_kernel(const int64_t* __restrict__ toff, const int* __restrict__ tkey,
                   const fx_t* __restrict__ tval, int64_t K, const int64_t* __restrict__ noff,
                   int64_t* __restrict__ ncol, double* __restrict__ nw_out, const int64_t* __restrict__ cpos,
                   const int* __restrict__ ckey, const double* __restrict__ cw)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = wid; i < K; i += nw) {
        int64_t pos = noff[i];
        if (toff[i + 1] == toff[i]) {   // uniform: compact entries of an LDS-aggregated row
            const int64_t n = noff[i + 1] - pos, c0 = cpos[i];
            for (int64_t r = lane; r < n; r += 64) {
                ncol[pos + r] = ckey[c0 + r];
                nw_out[pos + r] = cw[c0 + r];
            }
            continue;
        }
        for (int64_t s0 = toff[i]; s0 < toff[i + 1]; s0 += 64) {
            const int64_t s = s0 + lane;
            const int c = s < toff[i + 1] ? ld_i32_agent(&tkey[s]) : -1;
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

// ---------------------------------------------------------------------------
// Several graphs in one CSR: per-graph modularity, graph ids, vertex bases.
// ---------------------------------------------------------------------------
// Item = arcs [e0, e1) and vertices [v0, v1) of graph g; each workgroup sums
// its item exactly and adds the two sums to graph g's accumulators with
// integer atomics (order-free, so the result is exact and deterministic).
__global__ void __launch_bounds__(kBlock)
mod_seg_kernel(const SegItem* __restrict__ items, int64_t nitems, const int* __restrict__ src,
               const int* __restrict__ col, const double* __restrict__ w, const int* __restrict__ C,
               const fx_t* __restrict__ a, const double* __restrict__ twom_g, fx_t* __restrict__ acc,
               const unsigned char* __restrict__ act, fx_t* __restrict__ term)
{
    __shared__ fx_t red[kBlock / 64];
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const SegItem s = items[it];
        if (act && !act[s.g]) continue;   // uniform: a graph that sat out kept its partition
        fx_t I = 0, T = 0;
        for (int64_t e = s.e0 + threadIdx.x; e < s.e1; e += kBlock)
            if (C[src[e]] == C[col[e]]) I += fx_from_double(w[e]);    // self arcs are internal
        const double tw = twom_g[s.g];
        for (int64_t i = s.v0 + threadIdx.x; i < s.v1; i += kBlock) {
            const double x = fx_to_double(a[i]) / tw;
            const fx_t t = fx_from_double(x * x);
            if (term) term[i] = t;
            T += t;
        }
        I = block_sum_fx(I, red);
        T = block_sum_fx(T, red);
        if (threadIdx.x == 0) {
            acc_add(acc + 2 * s.g, I);
            acc_add(acc + 2 * s.g + 1, T);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
mod_seg_finalize_kernel(const fx_t* __restrict__ acc, const double* __restrict__ twom_g, int G,
                        double* __restrict__ q, const unsigned char* __restrict__ act,
                        const unsigned long long* __restrict__ moved)
{
    __shared__ unsigned long long red[kBlock / 64];
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock) {
        if (act && !act[g]) continue;     // q[g] keeps the value of its last evaluation
        const double tw = twom_g[g];
        q[g] = tw == 0 ? 0.0 : fx_to_double(acc[2 * g]) / tw - fx_to_double(acc[2 * g + 1]);
    }
    if (moved && blockIdx.x == 0) {   // uniform
        const unsigned long long t = sum_moved_slots(moved, red);
        if (threadIdx.x == 0) reinterpret_cast<unsigned long long*>(q)[G] = t;
    }
}

// out[g] = newid[vb[g]] (the scan total where vb[g] == nv), g = 0..G.
__global__ void __launch_bounds__(kBlock)
gather_bases_kernel(const int64_t* __restrict__ newid, const int64_t* __restrict__ total,
                    const int64_t* __restrict__ vb, int G, int64_t nv, int64_t* __restrict__ out)
{
    for (int g = blockIdx.x * kBlock + threadIdx.x; g <= G; g += gridDim.x * kBlock)
        out[g] = vb[g] < nv ? newid[vb[g]] : *total;
}

// out[g] = the exact sum of k over graph g's vertices; one workgroup per graph.
__global__ void __launch_bounds__(kBlock)
graph_sums_kernel(const fx_t* __restrict__ k, const int64_t* __restrict__ vb, int G, fx_t* __restrict__ out)
{
    __shared__ fx_t red[kBlock / 64];
    for (int g = blockIdx.x; g < G; g += gridDim.x) {
        fx_t t = 0;
        for (int64_t v = vb[g] + threadIdx.x; v < vb[g + 1]; v += kBlock) t += k[v];
        t = block_sum_fx(t, red);
        if (threadIdx.x == 0) out[g] = t;
    }
}

// gid[i] = g for i in [vb[g], vb[g+1]); one workgroup per graph.
__global__ void __launch_bounds__(kBlock)
fill_gid_kernel(const int64_t* __restrict__ vb, int G, int* __restrict__ gid)
{
    for (int g = blockIdx.x; g < G; g += gridDim.x)
        for (int64_t i = vb[g] + threadIdx.x; i < vb[g + 1]; i += kBlock) gid[i] = g;
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

void launch_weight_grid(const double* w, int64_t ne, int* fneed, unsigned long long* wmax, hipStream_t s)
{
    launch_fill_i32(fneed, INT_MIN, 1, s);
    launch_fill_i64(reinterpret_cast<int64_t*>(wmax), 0, 1, s);
    if (ne > 0) weight_grid_kernel<<<grid_for(ne, kBlock * 8, 2048), kBlock, 0, s>>>(w, ne, fneed, wmax);
}

void launch_to_fixed(const double* w, int64_t* wq, int64_t ne, double scale, hipStream_t s)
{
    if (ne > 0) to_fixed_kernel<<<grid_for(ne, kBlock * 4), kBlock, 0, s>>>(w, wq, ne, scale);
}

void launch_iota_i64(int64_t* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    iota_kernel<<<grid_for(n, kBlock * 4), kBlock, 0, s>>>(p, n);
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
                           int64_t nv, fx_t* acc, unsigned long long* maxabs, fx_t* part, int* src,
                           hipStream_t s)
{
    launch_fill_i64(reinterpret_cast<int64_t*>(maxabs), 0, 1, s);
    launch_fill_fx(acc, 1, s);
    if (nv <= 0) return;
    launch_fill_fx(k, nv, s);
    if (nsmall > 0)
        vw_small_kernel<<<grid_for(nsmall * kRowLanes, kBlock), kBlock, 0, s>>>(off, w, small, nsmall, k, src);
    if (nitems > 0)
        vw_items_kernel<<<grid_for(nitems, kBlock / 64), kBlock, 0, s>>>(w, items, nitems, k, src);
    const unsigned g = grid_for(nv, kBlock, kPartBlocks);
    round_kernel<<<g, kBlock, 0, s>>>(k, kd, nv, maxabs, part);
    sum_partials_kernel<<<1, kBlock, 0, s>>>(part, (int)g, acc);
}

void launch_singleton(int64_t* C, int64_t* Cn, int* C32, fx_t* a, const fx_t* k, int64_t nv,
                      hipStream_t s)
{
    if (nv <= 0) return;
    singleton_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, Cn, C32, a, k, nv);
}

void launch_narrow_ids(const int64_t* col, int* col32, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    narrow_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(col, col32, n);
}

void launch_widen_ids(const int* col32, int64_t* col, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    widen_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(col32, col, n);
}

void launch_modularity(const int* src, const int* col, const double* w, const int* C,
                       int64_t ne, const fx_t* a, int64_t nv, double twom_d, fx_t* part, double* q,
                       hipStream_t s, fx_t* term, fx_t* acc, bool ident)
{
    const int64_t tiles = (ne + (int64_t)kBlock * kModUnroll - 1) / ((int64_t)kBlock * kModUnroll);
    const int64_t need = tiles > (nv + kBlock - 1) / kBlock ? tiles : (nv + kBlock - 1) / kBlock;
    const unsigned g = grid_for(need, 1, kPartBlocks);
    if (ident) mod_kernel<true><<<g, kBlock, 0, s>>>(src, col, w, C, ne, a, nv, twom_d, part, term);
    else mod_kernel<false><<<g, kBlock, 0, s>>>(src, col, w, C, ne, a, nv, twom_d, part, term);
    mod_finalize_kernel<<<1, kBlock, 0, s>>>(part, (int)g, twom_d, q, acc);
}

void launch_modularity_rows(const int64_t* off, const int* col, const double* w, const int* C, int64_t v0,
                            int64_t v1, int64_t arcs, fx_t* acc, hipStream_t s)
{
    if (v1 > v0 && arcs > 0)
        mod_rows_kernel<<<grid_for(arcs, (int64_t)kBlock * kModRun, 8192), kBlock, 0, s>>>(off, col, w, C, v0, v1,
                                                                                            acc);
}

void launch_modularity_terms(const fx_t* a, int64_t nv, double twom_d, fx_t* acc, double* q, hipStream_t s,
                             fx_t* term)
{
    if (nv > 0) mod_terms_kernel<<<grid_for(nv, kBlock, kPartBlocks), kBlock, 0, s>>>(a, nv, twom_d, acc, term);
    mod_acc_finish_kernel<<<1, 64, 0, s>>>(acc, twom_d, q);
}

void launch_sym_hash(const int* src, const int* col, const double* w, int64_t ne, unsigned long long* out,
                     hipStream_t s)
{
    if (ne > 0) sym_hash_kernel<<<grid_for(ne, kBlock, 2048), kBlock, 0, s>>>(src, col, w, ne, out);
}

void launch_mod_delta(int b, int64_t v0, int64_t v1, const int64_t* Cn, const int* C32, const int* col,
                      const double* w, const fx_t* a, double twom_d, const ModDelta& D, hipStream_t s)
{
    mod_delta_kernel<<<512, kBlock, 0, s>>>(b, v0, v1, Cn, C32, col, w, a, twom_d, D);
}

void launch_mod_finish_graphs(const ModDelta& D, int nb, int G, double* q, const unsigned char* act,
                              hipStream_t s)
{
    mod_finish_graphs_kernel<<<grid_for(G, kBlock), kBlock, 0, s>>>(D, nb, G, q, act);
}

void launch_mod_finish(const ModDelta& D, int nb, double twom_d, double* q, hipStream_t s)
{
    mod_finish_kernel<<<1, kBlock, 0, s>>>(D, nb, twom_d, q);
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

void launch_mark_communities(const int64_t* C, int64_t nv, int64_t* flag, hipStream_t s,
                             const int* gid, const unsigned char* keep)
{
    launch_fill_i64(flag, 0, nv, s);
    if (nv > 0) mark_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, nv, flag, gid, keep);
}

void launch_update_clustering(int64_t* clus, int64_t n0, const int64_t* C, const int64_t* newid,
                              hipStream_t s, const int* gid0, const unsigned char* keep)
{
    if (n0 > 0) clustering_kernel<<<grid_for(n0, kBlock), kBlock, 0, s>>>(clus, n0, C, newid, gid0, keep);
}

void launch_widen_local(const ArcPiece* pieces, int64_t n, const uint16_t* local, int* col32, hipStream_t s)
{
    if (n > 0) widen_local_kernel<<<grid_for(n, 1, 16384), kBlock, 0, s>>>(pieces, n, local, col32);
}

void launch_local_ids(const int64_t* clus, int64_t n0, const int* gid0, const int64_t* base, int64_t* out,
                      hipStream_t s)
{
    if (n0 > 0) local_ids_kernel<<<grid_for(n0, kBlock), kBlock, 0, s>>>(clus, n0, gid0, base, out);
}

void launch_member_counts(const int64_t* C, const int64_t* newid, const int64_t* off, int64_t nv,
                          int64_t* cnt, int64_t* arcs, hipStream_t s, const int* gid,
                          const unsigned char* keep)
{
    if (nv > 0)
        member_counts_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, newid, off, nv, cnt, arcs, gid,
                                                                      keep);
}

void launch_coarse_sizes(const int64_t* arcs, int64_t K, int64_t* tsz, int64_t* csz, hipStream_t s)
{
    if (K > 0) coarse_sizes_kernel<<<grid_for(K, kBlock), kBlock, 0, s>>>(arcs, K, tsz, csz);
}

void launch_member_scatter(const int64_t* C, const int64_t* newid, int64_t nv, const int64_t* mstart,
                           unsigned long long* mfill, int* members, const int64_t* toff, int* hlist,
                           unsigned long long* nh, hipStream_t s, const int* gid, const unsigned char* keep)
{
    if (nv > 0)
        member_scatter_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(C, newid, nv, mstart, mfill, members, toff,
                                                                       hlist, nh, gid, keep);
}

void launch_coarse_lds(const int64_t* arcs, const int64_t* mstart, const int* members, const int64_t* off,
                       const int64_t* col, const double* w, const int64_t* C, const int64_t* newid, int64_t K,
                       const int64_t* cpos, int* ckey, double* cw, int64_t* cnt, hipStream_t s)
{
    if (K > 0)
        coarse_lds_kernel<<<grid_for(K, 1, 8192), kBlock, 0, s>>>(arcs, mstart, members, off, col, w, C, newid, K,
                                                                  cpos, ckey, cw, cnt);
}

void launch_coarse_insert(const int64_t* off, const int64_t* col, const double* w, const int64_t* C,
                          const int64_t* newid, const int* hlist, const unsigned long long* nh, int64_t nmax,
                          const int64_t* toff, int* tkey, fx_t* tval, hipStream_t s)
{
    if (nmax > 0)
        coarse_insert_kernel<<<grid_for(nmax, kBlock / 64), kBlock, 0, s>>>(off, col, w, C, newid, hlist, nh,
                                                                            toff, tkey, tval);
}

void launch_modularity_seg(const SegItem* items, int64_t nitems, const int* src, const int* col,
                           const double* w, const int* C, const fx_t* a, const double* twom_g, int G,
                           fx_t* acc, double* q, hipStream_t s, const unsigned char* act, fx_t* term,
                           const unsigned long long* moved)
{
    if (G <= 0) return;
    launch_fill_fx(acc, 2 * (int64_t)G, s);
    if (nitems > 0)
        mod_seg_kernel<<<grid_for(nitems, 1, 8192), kBlock, 0, s>>>(items, nitems, src, col, w, C, a,
                                                                    twom_g, acc, act, term);
    mod_seg_finalize_kernel<<<grid_for(G, kBlock), kBlock, 0, s>>>(acc, twom_g, G, q, act, moved);
}

void launch_graph_sums(const fx_t* k, const int64_t* vb, int G, fx_t* out, hipStream_t s)
{
    if (G > 0) graph_sums_kernel<<<grid_for(G, 1, 8192), kBlock, 0, s>>>(k, vb, G, out);
}

void launch_gather_bases(const int64_t* newid, const int64_t* total, const int64_t* vb, int G, int64_t nv,
                         int64_t* out, hipStream_t s)
{
    gather_bases_kernel<<<grid_for(G + 1, kBlock), kBlock, 0, s>>>(newid, total, vb, G, nv, out);
}

void launch_fill_gid(const int64_t* vb, int G, int* gid, hipStream_t s)
{
    if (G > 0) fill_gid_kernel<<<grid_for(G, 1, 8192), kBlock, 0, s>>>(vb, G, gid);
}

void launch_coarse_count(const int64_t* toff, const int* tkey, int64_t K, int64_t* cnt,
                         hipStream_t s)
{
    if (K > 0) coarse_count_kernel<<<grid_for(K, kBlock / 64), kBlock, 0, s>>>(toff, tkey, K, cnt);
}

void launch_coarse_emit(const int64_t* toff, const int* tkey, const fx_t* tval, int64_t K,
                        const int64_t* noff, int64_t* ncol, double* nw, const int64_t* cpos, const int* ckey,
                        const double* cw, hipStream_t s)
{
    if (K > 0)
        coarse_emit_kernel<<<grid_for(K, kBlock / 64), kBlock, 0, s>>>(toff, tkey, tval, K, noff,
                                                                       ncol, nw, cpos, ckey, cw);
}

}  // namespace nido
