// bins.hip -- per-phase work lists built on the device for large phases:
// the degree bins of every batch (rows grouped by (batch, bin), ascending id
// inside a group, isolated vertices left out) and the row lists of the
// degree and modularity sweeps.  The same lists as the host builder
// (host_bins.cpp), by stable counting sorts over tiles of consecutive
// vertices.
#include "device_common.h"

namespace nido {
namespace {

// The bin of a row of d > 0 arcs (bin_of in host_bins.cpp).
__device__ __forceinline__ int dev_bin(int64_t d, int thread_max, int64_t big_min)
{
    if (d > big_min) return BIN_BIG;
    if (d <= thread_max) return BIN_THREAD;
    if (d <= kBinMaxDeg[BIN_G8]) return BIN_G8;
    if (d <= kBinMaxDeg[BIN_G16]) return BIN_G16;
    if (d <= kBinMaxDeg[BIN_G32]) return BIN_G32;
    if (d <= kBinMaxDeg[BIN_G64]) return BIN_G64;
    if (d <= kBinMaxDeg[BIN_B256]) return BIN_B256;
    return BIN_BIG;
}

// Per tile: its rows of each bin, at cnt[at + k * stride].
__global__ void __launch_bounds__(kBlock)
bin_count_kernel(const BinTile* __restrict__ tiles, int64_t ntiles, const int64_t* __restrict__ off, int thread_max,
                 int64_t big_min, int64_t* __restrict__ cnt)
{
    __shared__ int c[NUM_BINS];
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const BinTile T = tiles[t];
        if (threadIdx.x < NUM_BINS) c[threadIdx.x] = 0;
        __syncthreads();
        for (int64_t v = T.v0 + threadIdx.x; v < T.v1; v += kBlock) {
            const int64_t d = off[v + 1] - off[v];
            if (d > 0) atomicAdd(&c[dev_bin(d, thread_max, big_min)], 1);
        }
        __syncthreads();
        if (threadIdx.x < NUM_BINS) cnt[T.at + threadIdx.x * T.stride] = c[threadIdx.x];
        __syncthreads();   // c is cleared for the next tile
    }
}

// Per tile: every row written at base[at + k * stride] plus its rank among
// the tile's earlier rows of bin k, so each (batch, bin) group lists its rows
// in ascending id.
__global__ void __launch_bounds__(kBlock)
bin_scatter_kernel(const BinTile* __restrict__ tiles, int64_t ntiles, const int64_t* __restrict__ off, int thread_max,
                   int64_t big_min, const int64_t* __restrict__ base, BinRow* __restrict__ binv)
{
    __shared__ int wc[kBlock / 64][NUM_BINS];   // this pass's rows per wavefront and bin
    __shared__ int run[NUM_BINS];               // rows of the tile's earlier passes per bin
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const BinTile T = tiles[t];
        if (threadIdx.x < NUM_BINS) run[threadIdx.x] = 0;
        __syncthreads();
        for (int64_t r0 = T.v0; r0 < T.v1; r0 += kBlock) {   // uniform; one pass per kBlock rows
            const int64_t v = r0 + threadIdx.x;
            int k = -1;
            int64_t d = 0, e0 = 0;
            if (v < T.v1) {
                e0 = off[v];
                d = off[v + 1] - e0;
                if (d > 0) k = dev_bin(d, thread_max, big_min);
            }
            int rank = 0;
#pragma unroll
            for (int kk = 0; kk < NUM_BINS; ++kk) {
                const unsigned long long m = __ballot(k == kk);
                if (k == kk) rank = __popcll(m & below);
                if (lane == 0) wc[wv][kk] = __popcll(m);
            }
            __syncthreads();
            if (k >= 0) {
                int before = run[k];
                for (int i = 0; i < wv; ++i) before += wc[i][k];
                binv[base[T.at + k * T.stride] + before + rank] =
                    BinRow{(int)v, (int)(d < INT32_MAX ? d : INT32_MAX), e0};
            }
            __syncthreads();   // run and wc read before they change
            if (threadIdx.x < NUM_BINS) {
                int s = 0;
                for (int i = 0; i < kBlock / 64; ++i) s += wc[i][threadIdx.x];
                run[threadIdx.x] += s;
            }
            __syncthreads();
        }
    }
}

// out[b * NUM_BINS + k] = the first position of group (b, k), or -1 for a
// batch without tiles.
__global__ void __launch_bounds__(kBlock)
bin_starts_kernel(const int64_t* __restrict__ base, const int64_t* __restrict__ tb0, int nb, int64_t* __restrict__ out)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nb * NUM_BINS; i += gridDim.x * kBlock) {
        const int b = i / NUM_BINS, k = i % NUM_BINS;
        const int64_t n = tb0[b + 1] - tb0[b];
        out[i] = n > 0 ? base[NUM_BINS * tb0[b] + k * n] : -1;
    }
}

// Row lists over tiles of kBinTileVerts consecutive vertices: rows of at
// most kSmallRowArcs arcs (one lane each) and kItemRowArcs-arc slices of
// the longer rows.
__device__ __forceinline__ int64_t slices_of(int64_t d)
{
    return d > kSmallRowArcs ? (d + kItemRowArcs - 1) / kItemRowArcs : 0;
}

__global__ void __launch_bounds__(kBlock)
rowlist_count_kernel(const int64_t* __restrict__ off, int64_t nv, int64_t ntiles, int64_t* __restrict__ cs,
                     int64_t* __restrict__ ci)
{
    __shared__ int64_t red[2][kBlock / 64];
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t v0 = t * kBinTileVerts, v1 = v0 + kBinTileVerts < nv ? v0 + kBinTileVerts : nv;
        int64_t s = 0, it = 0;
        for (int64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
            const int64_t d = off[v + 1] - off[v];
            s += d <= kSmallRowArcs;
            it += slices_of(d);
        }
#pragma unroll
        for (int msk = 32; msk > 0; msk >>= 1) {
            s += __shfl_xor(s, msk, 64);
            it += __shfl_xor(it, msk, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = s;
            red[1][threadIdx.x >> 6] = it;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t a = 0, b = 0;
            for (int i = 0; i < kBlock / 64; ++i) {
                a += red[0][i];
                b += red[1][i];
            }
            cs[t] = a;
            ci[t] = b;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
rowlist_scatter_kernel(const int64_t* __restrict__ off, int64_t nv, int64_t ntiles, const int64_t* __restrict__ cs,
                       const int64_t* __restrict__ ci, int* __restrict__ small, RowItem* __restrict__ items)
{
    __shared__ int wts[kBlock / 64];
    __shared__ int64_t wti[kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t v0 = t * kBinTileVerts, v1 = v0 + kBinTileVerts < nv ? v0 + kBinTileVerts : nv;
        int64_t ps = cs[t], pi = ci[t];   // positions of the pass's first row and slice
        for (int64_t r0 = v0; r0 < v1; r0 += kBlock) {   // uniform
            const int64_t v = r0 + threadIdx.x;
            int64_t e0 = 0, e1 = 0, fi = 0;
            int fs = 0;
            if (v < v1) {
                e0 = off[v];
                e1 = off[v + 1];
                fs = e1 - e0 <= kSmallRowArcs;
                fi = slices_of(e1 - e0);
            }
            // inclusive scans of both counts across the wavefront
            int is = fs;
            int64_t ii = fi;
#pragma unroll
            for (int dd = 1; dd < 64; dd <<= 1) {
                const int ys = __shfl_up(is, dd, 64);
                const int64_t yi = __shfl_up(ii, dd, 64);
                if (lane >= dd) {
                    is += ys;
                    ii += yi;
                }
            }
            if (lane == 63) {
                wts[wv] = is;
                wti[wv] = ii;
            }
            __syncthreads();
            int64_t bs = ps, bi = pi, ts = 0, ti = 0;
            for (int i = 0; i < kBlock / 64; ++i) {
                if (i < wv) {
                    bs += wts[i];
                    bi += wti[i];
                }
                ts += wts[i];
                ti += wti[i];
            }
            if (fs) small[bs + is - 1] = (int)v;
            for (int64_t j = 0; j < fi; ++j) {
                const int64_t s0 = e0 + (int64_t)j * kItemRowArcs;
                items[bi + ii - fi + j] = RowItem{v, s0, s0 + kItemRowArcs < e1 ? s0 + kItemRowArcs : e1};
            }
            ps += ts;
            pi += ti;
            __syncthreads();   // wts / wti read before the next pass writes them
        }
    }
}

}  // namespace

void launch_bin_count(const BinTile* tiles, int64_t ntiles, const int64_t* off, int thread_max, int64_t big_min,
                      int64_t* cnt, hipStream_t s)
{
    if (ntiles > 0)
        bin_count_kernel<<<grid_for(ntiles, 1, 8192), kBlock, 0, s>>>(tiles, ntiles, off, thread_max, big_min, cnt);
}

void launch_bin_scatter(const BinTile* tiles, int64_t ntiles, const int64_t* off, int thread_max, int64_t big_min,
                        const int64_t* base, BinRow* binv, hipStream_t s)
{
    if (ntiles > 0)
        bin_scatter_kernel<<<grid_for(ntiles, 1, 8192), kBlock, 0, s>>>(tiles, ntiles, off, thread_max, big_min,
                                                                          base, binv);
}

void launch_bin_starts(const int64_t* base, const int64_t* tb0, int nb, int64_t* out, hipStream_t s)
{
    if (nb > 0) bin_starts_kernel<<<grid_for((int64_t)nb * NUM_BINS, kBlock), kBlock, 0, s>>>(base, tb0, nb, out);
}

void launch_rowlist_count(const int64_t* off, int64_t nv, int64_t* cs, int64_t* ci, hipStream_t s)
{
    const int64_t nt = (nv + kBinTileVerts - 1) / kBinTileVerts;
    if (nt > 0) rowlist_count_kernel<<<grid_for(nt, 1, 8192), kBlock, 0, s>>>(off, nv, nt, cs, ci);
}

void launch_rowlist_scatter(const int64_t* off, int64_t nv, const int64_t* cs, const int64_t* ci, int* small,
                            RowItem* items, hipStream_t s)
{
    const int64_t nt = (nv + kBinTileVerts - 1) / kBinTileVerts;
    if (nt > 0) rowlist_scatter_kernel<<<grid_for(nt, 1, 8192), kBlock, 0, s>>>(off, nv, nt, cs, ci, small, items);
}

}  // namespace nido
