// fixed.h -- exact Q64.64 accumulators for the Louvain sums (host + device).
//
// Every sum of input weights on the local-moving path (vertex weight k_v,
// neighbour-community weight e_{v,c}, community total a_c, 2m, the internal
// weight of the modularity and the coarse arc weights) is held as a
// two's-complement 128-bit integer in units of 2^-64.  Integer addition is
// associative, so LDS/global atomics in any order, any reduction tree and
// any launch shape give identical bits; each sum is rounded to FP64 once,
// round-to-nearest-even, where the FP64 gain/modularity formulas read it.
//
// The reference sums the same quantities in FP64 in launch-dependent order
// (graph_cuda.cu:335-340 warp trees, graph_gpu.cpp:499 reduce_by_key,
// graph_cuda.cu:1244-1248 FP64 atomics), which is what this replaces.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NIDO_HD __host__ __device__ __forceinline__
#else
#define NIDO_HD static inline
#endif

typedef __int128 fx_t;
typedef unsigned __int128 ufx_t;

NIDO_HD uint64_t nido_dbits(double d) { return __builtin_bit_cast(uint64_t, d); }
NIDO_HD double nido_bitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

NIDO_HD fx_t fx_make(uint64_t lo, int64_t hi)
{
    return (fx_t)(((ufx_t)(uint64_t)hi << 64) | (ufx_t)lo);
}
NIDO_HD uint64_t fx_lo(fx_t x) { return (uint64_t)x; }
NIDO_HD int64_t fx_hi(fx_t x) { return (int64_t)((ufx_t)x >> 64); }

// FP64 -> Q64.64, round-to-nearest-even at 2^-64.  Exact for every weight
// on the 2^-64 grid (all |w| >= 2^-12).  The host checks |w| < 2^62 and
// finiteness once, when a graph is created.
NIDO_HD fx_t fx_from_double(double w)
{
    uint64_t b = nido_dbits(w);
    uint32_t ex = (uint32_t)(b >> 52) & 0x7ffu;
    uint64_t mt = b & 0x000fffffffffffffull;
    if (ex == 0) {
        if (mt == 0) return 0;
        ex = 1;                                  // subnormal: below 2^-1022
    } else {
        mt |= 0x0010000000000000ull;
    }
    int sh = (int)ex - 1011;                     // w * 2^64 = mt * 2^sh
    uint64_t lo, hi;                             // |w| * 2^64 as two 64-bit words
    if (sh >= 64) {                              // |w| < 2^62 keeps sh < 74
        lo = 0;
        hi = mt << (sh - 64);
    } else if (sh > 0) {
        lo = mt << sh;
        hi = mt >> (64 - sh);
    } else if (sh == 0) {
        lo = mt;
        hi = 0;
    } else {
        int r = -sh;
        if (r >= 54) return 0;
        uint64_t q = mt >> r;
        uint64_t rest = mt & ((1ull << r) - 1);
        uint64_t half = 1ull << (r - 1);
        q += (rest > half) || (rest == half && (q & 1ull));
        lo = q;
        hi = 0;
    }
    if (b >> 63) {                               // two's-complement negation
        hi = ~hi + (lo == 0);
        lo = 0 - lo;
    }
    return fx_make(lo, (int64_t)hi);
}

// Q64.64 -> FP64, round-to-nearest-even.  The magnitude is normalised
// into one 64-bit word t (its leading 1 at bit 63 when the high word is
// nonzero) with every bit shifted out ORed into t's lowest bit: those bits
// lie below the rounding position (bit 10), so the sticky bit keeps the
// above/at/below-half decision intact and the correctly rounded u64 -> f64
// conversion does the RNE.  Scaling by a power of two is exact.
NIDO_HD double fx_to_double(fx_t x)
{
    uint64_t lo = fx_lo(x);
    uint64_t hi = (uint64_t)fx_hi(x);
    const bool neg = (int64_t)hi < 0;
    if (neg) {
        hi = ~hi + (lo == 0);
        lo = 0 - lo;
    }
    double r;
    if (hi == 0) {
        r = (double)lo * nido_bitsd((uint64_t)(1023 - 64) << 52);     // lo * 2^-64
    } else {
        const int n = __builtin_clzll(hi);                           // 0..63 (hi < 2^63 here)
        const uint64_t t = n ? (hi << n) | (lo >> (64 - n)) : hi;
        const uint64_t sticky = (n ? lo << n : lo) != 0;
        r = (double)(t | sticky) * nido_bitsd((uint64_t)(1023 - n) << 52);   // * 2^-n
    }
    return neg ? -r : r;
}

// ---- division by a per-graph constant, rounded as the division is ---------
// RN(a / b) from y = RN(1 / b) (or y / 2 for b = 2 b', exactly RN(1 / (2 b'))),
// without the division: q0 = RN(a y) is within 1.5 ulp of a / b, the
// correction q1 = RN(q0 + (a - b q0) y) is within one ulp (faithful), and
// with the residual a - b q1, exact in one FMA, q2 = RN(q1 + (a - b q1) y) is
// the correctly rounded quotient (P. Markstein, IBM J. Res. Dev. 34(1), 1990).
// The argument needs the quotient and the residuals normal: b in
// [2^-399, 2^399] and a = 0 or |a| in [2^-400, 2^400].  Every division of
// the gain formulas is in that range: the sums are Q64.64 values (0 or at
// least 2^-64 in magnitude, below 2^63), so m is in [2^-65, 2^100], k_v a_c
// in [2^-128, 2^200], its quotient by 2m at least 2^-229, and a difference
// e - k_v a_c / (2m) is 0 or at least one ulp of 2^-229.
NIDO_HD double div_rcp(double a, double b, double y)
{
    double q = a * y;
    double r = fma(-b, q, a);
    q = fma(r, y, q);
    r = fma(-b, q, a);
    return fma(r, y, q);
}

#if defined(__HIPCC__)
// Exact atomic add of a Q64.64 value held as two 64-bit words {lo, hi}.
// The low word wraps modulo 2^64 and its carry is added to the high word, so
// the final 128-bit value is the exact sum whatever the interleaving.
// Works on LDS and global addresses alike (the compiler picks ds_/global_
// atomics when the address space is known after inlining).
__device__ __forceinline__ void fx_atomic_add(unsigned long long* lo, unsigned long long* hi, fx_t v)
{
    unsigned long long vl = (unsigned long long)fx_lo(v);
    unsigned long long vh = (unsigned long long)fx_hi(v);
    if (vl) {
        unsigned long long old = atomicAdd(lo, vl);
        vh += (old + vl) < old;
    }
    if (vh) atomicAdd(hi, vh);
}

// Same, with a 32-bit high word: exact while |value| < 2^31 (in units of 1),
// which holds for every partial sum of a row whose weights are non-negative
// and sum to less than 2^31.
__device__ __forceinline__ void fx_atomic_add_narrow(unsigned long long* lo, int* hi, fx_t v)
{
    unsigned long long vl = (unsigned long long)fx_lo(v);
    int vh = (int)fx_hi(v);
    if (vl) {
        unsigned long long old = atomicAdd(lo, vl);
        vh += (old + vl) < old;
    }
    if (vh) atomicAdd(hi, vh);
}

// Lane-exchange helpers for 128-bit values within a wavefront.
__device__ __forceinline__ fx_t fx_shfl_xor(fx_t x, int mask, int width)
{
    uint64_t lo = fx_lo(x);
    int64_t hi = fx_hi(x);
    lo = __shfl_xor(lo, mask, width);
    hi = __shfl_xor(hi, mask, width);
    return fx_make(lo, hi);
}
#endif
