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
