// coloring.hip -- the colouring behind the reference's -c option
// (Graph::coloring, graph.cpp:547-605) on the device.
//
// With one random-number engine the reference visits the vertices in id
// order, and vertex i takes the smallest colour that none of its
// already-coloured neighbours j (so j < i) with r_j >= r_i holds:
//   colour(i) = mex { colour(j) : j in N(i), j < i, r_j >= r_i }.
// That sequential rule is computed here exactly, in rounds: a vertex is
// ready once every such j has its colour, and then its colour is the same
// value whichever round computes it, so rounds can colour in place.  The
// uncoloured vertex of smallest id is always ready (its j are all smaller),
// so every round colours at least one vertex and the rounds end.
#include <vector>

#include "../../include/nido_amd.h"
#include "device_common.h"
#include "errors.h"
#include "host_check.h"

namespace nido {
namespace {

constexpr int kWindow = 4096;   // colours marked per pass (64 lanes x 64 bits)

// One wavefront per listed vertex; vertices not yet ready go to `next`.
__global__ void __launch_bounds__(kBlock)
color_round_kernel(const int64_t* __restrict__ off, const int* __restrict__ col, const int64_t* __restrict__ r,
                   int* color, const int* __restrict__ list, int64_t n, int* __restrict__ next, int* nnext)
{
    __shared__ unsigned long long bits[kBlock / 64][64];
    const int lane = threadIdx.x & 63;
    unsigned long long* B = bits[threadIdx.x >> 6];
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t li = wid; li < n; li += nw) {   // uniform in the wavefront
        const int i = list[li];
        const int64_t ri = r[i];
        const int64_t e0 = off[i], e1 = off[i + 1];
        int mex = -1;
        bool ready = true;
        for (int base = 0; mex < 0; base += kWindow) {
            B[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            bool wait = false;
            for (int64_t e = e0 + lane; e < e1; e += 64) {
                const int j = col[e];
                if (j < i && r[j] >= ri) {
                    // a colour written by another wavefront of this round may
                    // not be seen yet: the vertex then waits one more round
                    const int c = __hip_atomic_load(color + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (c < 0) wait = true;
                    else if (c >= base && c < base + kWindow)
                        atomicOr(&B[(c - base) >> 6], 1ull << ((c - base) & 63));
                }
            }
            if (__ballot(wait)) {   // uniform
                ready = false;
                break;
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
        if (lane == 0) {
            if (ready) color[i] = mex;
            else next[atomicAdd(nnext, 1)] = i;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
color_init_kernel(int* __restrict__ color, int* __restrict__ list, int64_t nv)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv; i += (int64_t)gridDim.x * kBlock) {
        color[i] = -1;
        list[i] = (int)i;
    }
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(int64_t n) { return hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)); }
};

}  // namespace

int device_coloring(int64_t nv, const int64_t* off, const int* col, const int64_t* r, int* colors,
                    int64_t* rounds)
{
    const int64_t ne = off[nv];
    DevBuf<int64_t> d_off, d_r;
    DevBuf<int> d_col, d_color, d_list[2], d_n;
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
    NIDO_CTRY(d_list[0].alloc(nv));
    NIDO_CTRY(d_list[1].alloc(nv));
    NIDO_CTRY(d_n.alloc(1));
    NIDO_CTRY(hipMemcpyAsync(d_off.p, off, sizeof(int64_t) * (size_t)(nv + 1), hipMemcpyHostToDevice, s));
    NIDO_CTRY(hipMemcpyAsync(d_r.p, r, sizeof(int64_t) * (size_t)nv, hipMemcpyHostToDevice, s));
    if (ne > 0) NIDO_CTRY(hipMemcpyAsync(d_col.p, col, sizeof(int) * (size_t)ne, hipMemcpyHostToDevice, s));
    color_init_kernel<<<grid_for(nv, kBlock), kBlock, 0, s>>>(d_color.p, d_list[0].p, nv);
    NIDO_CTRY(hipGetLastError());
    int64_t n = nv, k = 0;
    int cur = 0;
    int nnext = 0;
    while (n > 0) {
        NIDO_CTRY(hipMemsetAsync(d_n.p, 0, sizeof(int), s));
        color_round_kernel<<<grid_for(n, kBlock / 64, 16384), kBlock, 0, s>>>(d_off.p, d_col.p, d_r.p, d_color.p,
                                                                            d_list[cur].p, n, d_list[1 - cur].p,
                                                                            d_n.p);
        NIDO_CTRY(hipGetLastError());
        NIDO_CTRY(hipMemcpyAsync(&nnext, d_n.p, sizeof(int), hipMemcpyDeviceToHost, s));
        NIDO_CTRY(hipStreamSynchronize(s));
        ++k;
        if (nnext >= n) {   // cannot happen: the smallest uncoloured id is always ready
            (void)hipStreamDestroy(s);
            return set_error(NIDO_ERR_STATE, "coloring made no progress in round %lld", (long long)k);
        }
        n = nnext;
        cur = 1 - cur;
    }
    NIDO_CTRY(hipMemcpyAsync(colors, d_color.p, sizeof(int) * (size_t)nv, hipMemcpyDeviceToHost, s));
    NIDO_CTRY(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
#undef NIDO_CTRY
    if (rounds) *rounds = k;
    return NIDO_OK;
}

}  // namespace nido
