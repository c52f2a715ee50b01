// coloring.hip -- the colouring behind the reference's -c option
// (Graph::coloring, graph.cpp:547-605) on the device.
//
// With one random-number engine the reference visits the vertices in id
// order, and vertex i takes the smallest colour that none of its
// already-coloured neighbours j (so j < i) with r_j >= r_i holds:
//   colour(i) = mex { colour(j) : j in N(i), j < i, r_j >= r_i }.
// That sequential rule is computed here exactly, in rounds: a vertex is
// ready once every such j (its predecessors) has its colour, and then its
// colour is the same value whichever round computes it.  Each vertex counts
// its predecessors once; the dependents of every vertex (the x whose
// predecessor it is) are listed in a CSR of their own, built on the device;
// a round colours the ready vertices and counts each one off its
// dependents, and a dependent whose count reaches zero is ready in the next
// round.  Every arc is visited a fixed number of times however many rounds
// there are (rounds: the longest predecessor chain).  Predecessors have
// smaller ids, so the uncoloured vertex of smallest id is always ready and
// the rounds end.
#include <vector>

#include "../../include/nido_amd.h"
#include "device_common.h"
#include "errors.h"
#include "host_check.h"
#include "host_copy.h"

namespace nido {
namespace {

constexpr int kWindow = 4096;   // colours marked per pass (64 lanes x 64 bits)

// One wavefront per vertex x: pending[x] = its predecessors, each
// predecessor j counted in dcnt[j] (its dependents), ready[x] = 1 when x
// has none.
__global__ void __launch_bounds__(kBlock)
pred_count_kernel(const int64_t* __restrict__ off, const int* __restrict__ col, const int64_t* __restrict__ r,
                  int64_t nv, int* __restrict__ pending, unsigned long long* __restrict__ dcnt,
                  int64_t* __restrict__ ready)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t x = wid; x < nv; x += nw) {   // uniform in the wavefront
        const int64_t rx = r[x];
        int cnt = 0;
        for (int64_t e = off[x] + lane; e < off[x + 1]; e += 64) {
            const int j = col[e];
            if (j < x && r[j] >= rx) {
                ++cnt;
                atomicAdd(&dcnt[j], 1ull);
            }
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) cnt += __shfl_xor(cnt, m, 64);
        if (lane == 0) {
            pending[x] = cnt;
            ready[x] = cnt == 0;
        }
    }
}

// Dependents' CSR: x appended to the list of each of its predecessors j.
__global__ void __launch_bounds__(kBlock)
dep_fill_kernel(const int64_t* __restrict__ off, const int* __restrict__ col, const int64_t* __restrict__ r,
                int64_t nv, const int64_t* __restrict__ dstart, unsigned long long* __restrict__ dfill,
                int* __restrict__ deps)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t x = wid; x < nv; x += nw) {
        const int64_t rx = r[x];
        for (int64_t e = off[x] + lane; e < off[x + 1]; e += 64) {
            const int j = col[e];
            if (j < x && r[j] >= rx) deps[dstart[j] + (int64_t)atomicAdd(&dfill[j], 1ull)] = (int)x;
        }
    }
}

// The first round's list: the vertices without predecessors, ascending.
__global__ void __launch_bounds__(kBlock)
ready_scatter_kernel(const int64_t* __restrict__ ready, const int64_t* __restrict__ pos, int64_t nv,
                     int* __restrict__ list)
{
    for (int64_t x = blockIdx.x * (int64_t)kBlock + threadIdx.x; x < nv; x += (int64_t)gridDim.x * kBlock)
        if (ready[x]) list[pos[x]] = (int)x;
}

// One wavefront per ready vertex i: its colour (every predecessor has one),
// then each dependent counted off; those that reach zero go to `next`.
__global__ void __launch_bounds__(kBlock)
color_round_kernel(const int64_t* __restrict__ off, const int* __restrict__ col, const int64_t* __restrict__ r,
                   int* __restrict__ color, const int* __restrict__ list, int64_t n,
                   const int64_t* __restrict__ dstart, const int* __restrict__ deps, int* __restrict__ pending,
                   int* __restrict__ next, int* __restrict__ nnext)
{
    __shared__ unsigned long long bits[kBlock / 64][64];
    const int lane = threadIdx.x & 63;
    unsigned long long* B = bits[threadIdx.x >> 6];
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int64_t li = wid; li < n; li += nw) {   // uniform in the wavefront
        const int i = list[li];
        const int64_t ri = r[i];
        const int64_t e0 = off[i], e1 = off[i + 1];
        int mex = -1;
        for (int base = 0; mex < 0; base += kWindow) {
            B[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            for (int64_t e = e0 + lane; e < e1; e += 64) {
                const int j = col[e];
                if (j < i && r[j] >= ri) {
                    const int c = color[j];   // coloured in an earlier round
                    if (c >= base && c < base + kWindow) atomicOr(&B[(c - base) >> 6], 1ull << ((c - base) & 63));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const unsigned long long freew = ~B[lane];
            const unsigned long long m = __ballot(freew != 0);
            if (m) {
                const int l = __builtin_ctzll(m);
                const unsigned long long fw = __shfl(freew, l, 64);
                mex = base + l * 64 + __builtin_ctzll(fw);
            }
            __builtin_amdgcn_wave_barrier();   // B is cleared for the next window only after every lane read it
        }
        if (lane == 0) color[i] = mex;
        const int64_t d1 = dstart[i + 1];
        for (int64_t d0 = dstart[i]; d0 < d1; d0 += 64) {   // uniform
            const int64_t d = d0 + lane;
            int x = -1;
            if (d < d1) {
                x = deps[d];
                if (atomicSub(&pending[x], 1) != 1) x = -1;
            }
            const unsigned long long rel = __ballot(x >= 0);
            if (rel) {
                int at = 0;
                if (lane == 0) at = atomicAdd(nnext, __popcll(rel));
                at = __shfl(at, 0, 64);
                if (x >= 0) next[at + __popcll(rel & below)] = x;
            }
        }
    }
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(int64_t n) { return hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)); }
};

}  // namespace

int device_coloring(int64_t nv, const int64_t* off, const int64_t* col, const int64_t* r, int* colors,
                    int64_t* rounds)
{
    const int64_t ne = off[nv];
    DevBuf<int64_t> d_off, d_r, d_dstart, d_ready, d_total;
    DevBuf<unsigned long long> d_dfill;
    DevBuf<int> d_col, d_color, d_pending, d_deps, d_list[2], d_n;
    DevBuf<char> d_tmp;
    hipStream_t s = nullptr;
#define NIDO_CTRY(expr)                                                                                 \
    do {                                                                                               \
        const hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                        \
            if (s) (void)hipStreamDestroy(s);                                                          \
            return set_error(NIDO_ERR_HIP, "coloring: %s: %s", #expr, hipGetErrorString(e_));          \
        }                                                                                              \
    } while (0)
    NIDO_CTRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    NIDO_CTRY(d_off.alloc(nv + 1));
    NIDO_CTRY(d_r.alloc(nv));
    NIDO_CTRY(d_col.alloc(ne));
    NIDO_CTRY(d_color.alloc(nv));
    NIDO_CTRY(d_pending.alloc(nv));
    NIDO_CTRY(d_dstart.alloc(nv + 1));
    NIDO_CTRY(d_dfill.alloc(nv));
    NIDO_CTRY(d_ready.alloc(nv + 1));
    NIDO_CTRY(d_total.alloc(1));
    NIDO_CTRY(d_deps.alloc(ne));
    NIDO_CTRY(d_list[0].alloc(nv));
    NIDO_CTRY(d_list[1].alloc(nv));
    NIDO_CTRY(d_n.alloc(1));
    NIDO_CTRY(d_tmp.alloc((int64_t)scan_temp_bytes(nv + 1)));
    NIDO_CTRY(hipMemcpyAsync(d_off.p, off, sizeof(int64_t) * (size_t)(nv + 1), hipMemcpyHostToDevice, s));
    NIDO_CTRY(hipMemcpyAsync(d_r.p, r, sizeof(int64_t) * (size_t)nv, hipMemcpyHostToDevice, s));
    if (ne > 0) NIDO_CTRY(staged_h2d_ids(d_col.p, col, ne, s));   // narrowed to int32 by the host threads
    // predecessor counts and the dependents' CSR (dstart = exclusive scan of
    // the dependents counts, nv + 1 entries)
    launch_fill_i64(reinterpret_cast<int64_t*>(d_dfill.p), 0, nv, s);
    launch_fill_i64(d_dstart.p, 0, nv + 1, s);
    const unsigned gw = grid_for(nv, kBlock / 64, 16384);
    pred_count_kernel<<<gw, kBlock, 0, s>>>(d_off.p, d_col.p, d_r.p, nv, d_pending.p,
                                            reinterpret_cast<unsigned long long*>(d_dstart.p), d_ready.p);
    NIDO_CTRY(hipGetLastError());
    launch_exclusive_scan(d_dstart.p, d_dstart.p, nv + 1, nullptr, d_tmp.p, s);
    dep_fill_kernel<<<gw, kBlock, 0, s>>>(d_off.p, d_col.p, d_r.p, nv, d_dstart.p, d_dfill.p, d_deps.p);
    NIDO_CTRY(hipGetLastError());
    // the first round's list: the vertices without predecessors
    launch_fill_i64(d_ready.p + nv, 0, 1, s);
    launch_exclusive_scan(d_ready.p, reinterpret_cast<int64_t*>(d_dfill.p), nv, d_total.p, d_tmp.p, s);
    ready_scatter_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(d_ready.p, reinterpret_cast<int64_t*>(d_dfill.p),
                                                                 nv, d_list[0].p);
    NIDO_CTRY(hipGetLastError());
    int64_t n = 0, k = 0, done = 0;
    NIDO_CTRY(hipMemcpyAsync(&n, d_total.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    NIDO_CTRY(hipStreamSynchronize(s));
    int cur = 0;
    int nnext = 0;
    while (n > 0) {
        launch_fill_i32(d_n.p, 0, 1, s);
        color_round_kernel<<<grid_for(n, kBlock / 64, 16384), kBlock, 0, s>>>(
            d_off.p, d_col.p, d_r.p, d_color.p, d_list[cur].p, n, d_dstart.p, d_deps.p, d_pending.p,
            d_list[1 - cur].p, d_n.p);
        NIDO_CTRY(hipGetLastError());
        NIDO_CTRY(hipMemcpyAsync(&nnext, d_n.p, sizeof(int), hipMemcpyDeviceToHost, s));
        NIDO_CTRY(hipStreamSynchronize(s));
        ++k;
        done += n;
        if (done + nnext > nv) {   // cannot happen: each vertex is released once
            (void)hipStreamDestroy(s);
            return set_error(NIDO_ERR_STATE, "coloring listed more vertices than it has in round %lld", (long long)k);
        }
        n = nnext;
        cur = 1 - cur;
    }
    if (done != nv) {   // cannot happen: predecessors have smaller ids
        (void)hipStreamDestroy(s);
        return set_error(NIDO_ERR_STATE, "coloring stopped with %lld of %lld vertices coloured", (long long)done,
                         (long long)nv);
    }
    NIDO_CTRY(hipMemcpyAsync(colors, d_color.p, sizeof(int) * (size_t)nv, hipMemcpyDeviceToHost, s));
    NIDO_CTRY(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
#undef NIDO_CTRY
    if (rounds) *rounds = k;
    return NIDO_OK;
}

}  // namespace nido
