// noalloc.hip -- modularity of a caller-supplied partition without any
// device allocation.
//
// nido_compute_modularity evaluates Q for the partition a handle holds; the
// handle owns its state and its reduction buffers, all allocated when it was
// created.  Callers that cannot allocate device memory at all -- code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc --
// cannot create a handle.  nido_modularity_noalloc takes the CSR and the
// community of every vertex as device arrays and computes the same Q (the
// same exact Q64.64 sums, rounded once, so the same bits) with its working
// memory in one file-scope device array.  That array is sized for
// kNoAllocMaxVertices communities; every GPU the library runs on holds it
// until the process ends.
//
// The array's name is a host-side stand-in for the symbol, not a device
// address: the first call on a device, with that device current, resolves
// it with hipGetSymbolAddress and keeps the address per device; every kernel
// argument and fill after that uses the kept address.
//
// Calls on one GPU share its array: each call clears what it sums into
// before using it (with a fill kernel, never hipMemsetAsync, so the launches
// can be captured into a graph), and two calls whose launches overlap on one
// GPU -- on different streams without ordering between them -- race on it.
#include <mutex>

#include "../../include/nido_amd.h"
#include "device_common.h"
#include "errors.h"

using namespace nido;

namespace {

constexpr int64_t kNoAllocMaxVertices = NIDO_NOALLOC_MAX_VERTICES;
// [0, max) community totals a_c, then four accumulators: internal weight,
// sum of square terms, 2m, and a flag set when an id lies outside [0, nv).
constexpr int64_t kScratchFx = kNoAllocMaxVertices + 4;

__device__ __align__(256) fx_t scratch[kScratchFx];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
fx_t* g_scratch_ptr[kMaxDevices] = {};

// Clears a[0, nv) and the four accumulators.
__global__ void __launch_bounds__(kBlock)
na_clear_kernel(fx_t* __restrict__ a, int64_t nv, fx_t* __restrict__ acc)
{
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv; i += (int64_t)gridDim.x * kBlock) a[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 4) acc[threadIdx.x] = 0;
}

// One wavefront per row: k_v = sum of the row's weights (self arcs
// included, graph_cuda.cu:302-346) goes into a[comm[v]]; arcs whose ends
// share a community (self arcs included) add their weight to the internal
// sum.  Exact integer sums, so the order of the atomics does not matter.
// An id outside [0, nv), or a row outside [0, ne), is never used as an
// index: it sets the flag.
__global__ void __launch_bounds__(kBlock)
na_rows_kernel(const int64_t* __restrict__ off, const int64_t* __restrict__ col, const double* __restrict__ w,
               const int64_t* __restrict__ comm, int64_t nv, int64_t ne, fx_t* __restrict__ a,
               fx_t* __restrict__ acc)
{
    __shared__ fx_t red[kBlock / 64];
    const int lane = threadIdx.x & 63;
    const int64_t wid = (blockIdx.x * (int64_t)kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    fx_t I = 0, M = 0;
    bool bad = false;
    for (int64_t v = wid; v < nv; v += nw) {
        const int64_t cv = comm[v];
        const bool cv_ok = cv >= 0 && cv < nv;
        int64_t e0 = off[v], e1 = off[v + 1];
        if (e0 < 0 || e1 > ne || e0 > e1) {   // not a CSR row: flagged, not read
            bad = true;
            e0 = e1 = 0;
        }
        fx_t k = 0;
        for (int64_t e = e0 + lane; e < e1; e += 64) {
            const fx_t x = fx_from_double(w[e]);
            k += x;
            const int64_t u = col[e];
            if (u < 0 || u >= nv) bad = true;
            else if (comm[u] == cv) I += x;
        }
        k = wave_sum_fx(k);
        if (!cv_ok) bad = true;
        else if (lane == 0 && k != 0) acc_add(a + cv, k);
        M += k;   // every lane holds the row's sum; lane 0's copy is kept below
    }
    if (lane != 0) M = 0;
    I = block_sum_fx(I, red);
    M = block_sum_fx(M, red);
    if (threadIdx.x == 0) {
        if (I != 0) acc_add(acc, I);
        if (M != 0) acc_add(acc + 2, M);
    }
    if (bad) acc_add(acc + 3, fx_t(1));
}

// T = sum over communities of fx((RN(a_c) / 2m)^2), with 2m = RN(acc[2]).
__global__ void __launch_bounds__(kBlock)
na_terms_kernel(const fx_t* __restrict__ a, int64_t nv, fx_t* __restrict__ acc)
{
    __shared__ fx_t red[kBlock / 64];
    const double twom_d = fx_to_double(acc[2]);
    fx_t T = 0;
    if (twom_d != 0.0) {
        for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nv; i += (int64_t)gridDim.x * kBlock) {
            const double x = fx_to_double(a[i]) / twom_d;
            T += fx_from_double(x * x);
        }
    }
    T = block_sum_fx(T, red);
    if (threadIdx.x == 0 && T != 0) acc_add(acc + 1, T);
}

// Q = RN(I) / 2m - RN(T); Q = 0 for a graph without weight, NaN when an id
// was out of range.
__global__ void na_finish_kernel(const fx_t* __restrict__ acc, double* __restrict__ q)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double twom_d = fx_to_double(acc[2]);
        double r = twom_d == 0.0 ? 0.0 : fx_to_double(acc[0]) / twom_d - fx_to_double(acc[1]);
        if (acc[3] != 0) r = __builtin_nan("");
        *q = r;
    }
}

// The device address of `scratch` on the current device, resolved once per
// device.
int scratch_address(fx_t** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(NIDO_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(NIDO_ERR_INVALID, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NIDO_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<fx_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return NIDO_OK;
}

}  // namespace

extern "C" int nido_modularity_noalloc(int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* cols,
                                       const double* weights, const int64_t* comm, double* q, void* stream)
{
    if (nv < 0 || ne < 0) return set_error(NIDO_ERR_INVALID, "negative size (nv %lld, ne %lld)", (long long)nv,
                                           (long long)ne);
    if (nv > kNoAllocMaxVertices)
        return set_error(NIDO_ERR_CAPACITY, "%lld vertices exceed the %lld this entry point's fixed working "
                         "memory holds", (long long)nv, (long long)kNoAllocMaxVertices);
    if (!q || (nv > 0 && (!offsets || !comm)) || (ne > 0 && (!cols || !weights)))
        return set_error(NIDO_ERR_INVALID, "null array");
    fx_t* base = nullptr;
    int rc = scratch_address(&base);
    if (rc != NIDO_OK) return rc;
    fx_t* a = base;
    fx_t* acc = base + kNoAllocMaxVertices;
    hipStream_t s = static_cast<hipStream_t>(stream);
    na_clear_kernel<<<grid_for(nv, kBlock, 2048), kBlock, 0, s>>>(a, nv, acc);
    if (nv > 0) {
        na_rows_kernel<<<grid_for(nv, kBlock / 64, 4096), kBlock, 0, s>>>(offsets, cols, weights, comm, nv, ne, a, acc);
        na_terms_kernel<<<grid_for(nv, kBlock, 1024), kBlock, 0, s>>>(a, nv, acc);
    }
    na_finish_kernel<<<1, 64, 0, s>>>(acc, q);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NIDO_ERR_HIP, "nido_modularity_noalloc launch: %s", hipGetErrorString(e));
    return NIDO_OK;
}
