// lm_kernels.hip -- local-moving kernels (one -b batch per launch pair).
//
// Reference path being replaced (sg0/nido): fill_edges_community_ids +
// copy_weights_exclude_self_loops + Thrust stable_sort_by_key/reduce_by_key
// + fill_unique_community_counts + inclusive_scan + louvain_update_kernel
// (graph_gpu.cpp:457-517, graph_cuda.cu:49-98, 1450-1611, 2289-2348).
// Here each row is scanned coalesced over CSR; C[col] is gathered, the
// weight converted to Q64.64 and summed in an open-addressing table keyed by
// community id; the FP64 gain argmax (min-id ties) is taken over the table.
//
// Degree bins (rows of d arcs):
//   d <= 2      one lane per vertex, neighbours in registers
//   d <= 16/32/64/128   8/16/32/64 lanes per vertex, LDS table of 2d..4d slots
//   d <= 512    one workgroup per vertex, 1024-slot LDS table
//   d >  512    long rows, split by a hash of the community id into
//               P = 2^ceil(log2(d / part_arcs)) partitions, so that every
//               e_{v,c} is complete and exact in one workgroup's LDS table;
//               the last of the P workgroups to finish combines the P
//               partial argmaxes.  For P > 2, launch 1 buckets the row's
//               arcs by partition (a counting sort per 2048-arc chunk) and
//               workgroup (v, p) of launch 2 reads only its own buckets;
//               with P <= 2 each workgroup reads the row and filters it.
// Launch 1 = short-row bins + long-row key fill (20 KB LDS per workgroup,
// 12 KB with fixed-point sums);
// launch 2 = long-row partitions (40 KB LDS with the default 2048-slot
// table).  Row loops keep two to eight independent loads in flight per lane.
#include <type_traits>

#include "device_common.h"

namespace nido {
namespace {

// A graph's mass m with y = RN(1/m).  Both divisions of a gain are by m or
// 2m (with 1/(2m) = y/2 exactly), so each is a multiplication by y plus two
// FMA corrections (div_rcp), rounded as the division is.
struct Mass {
    double m, y;
};

// g_c = (e_{v,c} - k_v a_c / (2m)) / m, graph_cuda.cu:1505-1506.
__device__ __forceinline__ double gain_of(const LmArgs& A, Mass m, double ki, fx_t e, int c)
{
    double g = fx_to_double(e);
    const double ac = fx_to_double(A.a[c]);
    g -= div_rcp(ki * ac, 2. * m.m, 0.5 * m.y);
    return div_rcp(g, m.m, m.y);
}

// Mass of v's graph; false when that graph sits out the iteration.  The
// kernels are compiled twice: Multi = false (one graph) leaves out the
// graph lookups altogether.
template <bool Multi>
__device__ __forceinline__ bool graph_mass(const LmArgs& A, int v, Mass& m)
{
    if constexpr (!Multi) {
        m = Mass{A.m, A.ym};
        return true;
    }
    const int gg = A.gid[v];
    m = Mass{A.mg[2 * gg], A.mg[2 * gg + 1]};
    return A.act[gg] != 0;
}

// Move iff g_best - s' > 0 with s' = (e_{v,own} - k_v (a_own - k_v)/(2m))/m,
// graph_cuda.cu:1527-1536.
// `rest` = a_own - k_v, loaded by the caller ahead of the row scan.
// `self_d` = e_{v,own} rounded to FP64.
__device__ __forceinline__ int64_t decide_d(Mass m, double ki, fx_t rest, int own, double self_d,
                                            double best_g, int best_c)
{
    if (best_c == kNone) return own;    // gain stays -MAX_FLOAT, graph_cuda.cu:1489
    double s = self_d;
    const double ar = fx_to_double(rest);
    s -= div_rcp(ki * ar, 2. * m.m, 0.5 * m.y);
    s = div_rcp(s, m.m, m.y);
    double gain = best_g;
    gain -= s;
    return gain > 0 ? (int64_t)best_c : (int64_t)own;
}

__device__ __forceinline__ int64_t decide(Mass m, double ki, fx_t rest, int own, fx_t self, double best_g,
                                          int best_c)
{
    return decide_d(m, ki, rest, own, fx_to_double(self), best_g, best_c);
}

__device__ __forceinline__ int64_t decide(const LmArgs& A, Mass m, int v, int own, fx_t self,
                                          double best_g, int best_c)
{
    return decide(m, A.kd[v], A.a[own] - A.k[v], own, self, best_g, best_c);
}

// Candidate update for one table entry whose a_c the caller has loaded.
__device__ __forceinline__ void consider_a(Mass m, double ki, int own, int c, fx_t e, fx_t ac,
                                           fx_t& self, double& bg, int& bc)
{
    if (c == own) {
        self = e;
        return;
    }
    double g = fx_to_double(e);
    g -= div_rcp(ki * fx_to_double(ac), 2. * m.m, 0.5 * m.y);
    g = div_rcp(g, m.m, m.y);
    if (better(g, c, bg, bc)) {
        bg = g;
        bc = c;
    }
}

// Candidate update for one table entry.
__device__ __forceinline__ void consider(const LmArgs& A, Mass m, double ki, int own, int c, fx_t e,
                                         fx_t& self, double& bg, int& bc)
{
    if (c == own) {
        self = e;
        return;
    }
    const double g = gain_of(A, m, ki, e, c);
    if (better(g, c, bg, bc)) {
        bg = g;
        bc = c;
    }
}

// Candidate update for a candidate other than the own community, from
// e_{v,c} already rounded to FP64.
__device__ __forceinline__ void consider_e(Mass m, double ki, int c, double e, fx_t ac, double& bg, int& bc)
{
    double g = e;
    g -= div_rcp(ki * fx_to_double(ac), 2. * m.m, 0.5 * m.y);
    g = div_rcp(g, m.m, m.y);
    if (better(g, c, bg, bc)) {
        bg = g;
        bc = c;
    }
}

// e_{v,c} of table slot s rounded to FP64: a Q64.64 sum in {lo, hi}, or
// (Fixed) a 64-bit sum in units of 2^-F in lo, scaled by wscale = 2^-F
// (the int64 -> FP64 conversion rounds once; the scaling is exact).
template <bool Fixed>
__device__ __forceinline__ double slot_sum(const LmArgs& A, const unsigned long long* lo,
                                           const unsigned long long* hi, int s)
{
    if constexpr (Fixed) return (double)(int64_t)lo[s] * A.wscale;
    else return fx_to_double(fx_make(lo[s], (int64_t)hi[s]));
}

// The keys a lane created in a short row's table: their slots and
// communities, and a_c, loaded as soon as the key was new -- the lane that
// creates a key scores it, so the load travels with the row's inserts
// instead of after them.  A short-row lane handles at most two arcs.
struct OwnKeys {
    int s[2], c[2];
    fx_t a[2];
};

// Scores a lane's own keys (after the table is complete); the lane holding
// the row's own community stores e_{v,own} (rounded) in *self_d, so only the
// argmax needs a cross-lane reduction.
template <bool Fixed>
__device__ __forceinline__ void sweep_own(const LmArgs& A, Mass m, double ki, int own, const OwnKeys& k,
                                          const unsigned long long* lo, const unsigned long long* hi,
                                          double* self_d, double& bg, int& bc)
{
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (k.s[j] < 0) continue;
        const double e = slot_sum<Fixed>(A, lo, hi, k.s[j]);
        if (k.c[j] == own) *self_d = e;
        else consider_e(m, ki, k.c[j], e, k.a[j], bg, bc);
    }
}

// Set bits of `mask` below this lane.
__device__ __forceinline__ int lane_rank(unsigned long long mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// The slot of key c in an LDS open-addressing table of S slots (S a power
// of 2, never more than half full by construction of the bins), creating
// it if absent; *fresh tells whether this lane created it.
template <int S>
__device__ __forceinline__ unsigned lds_claim(int* key, int c, bool* fresh)
{
    unsigned h = ((unsigned)c * kHashMul) >> (32 - ilog2c(S));
    *fresh = false;
    for (;;) {
        const int kk = key[h];
        if (kk == c) break;
        if (kk < 0) {
            const int prev = atomicCAS(&key[h], -1, c);
            if (prev < 0) {
                *fresh = true;
                break;
            }
            if (prev == c) break;
        }
        h = (h + 1) & (S - 1);
    }
    return h;
}

// Inserts arc (c, weight) of a short row: the weight joins slot c's sum
// (Q64.64 in {lo, hi}, or Fixed: one 64-bit word in units of 2^-F); a new
// key goes to the lane's own list j with its a_c load issued at once.
template <int S, bool Fixed>
__device__ __forceinline__ void insert_own(const LmArgs& A, int* key, unsigned long long* lo, unsigned long long* hi,
                                           int c, double w, int64_t wq, OwnKeys& k, int j)
{
    bool fresh;
    const unsigned h = lds_claim<S>(key, c, &fresh);
    if constexpr (Fixed) atomicAdd(&lo[h], (unsigned long long)wq);
    else fx_atomic_add(&lo[h], &hi[h], fx_from_double(w));
    if (fresh) {
        k.s[j] = (int)h;
        k.c[j] = c;
        k.a[j] = A.a[c];
    }
}

// Scans short row v with `lanes` lanes (lane index `lane`): the row has at
// most 2 * lanes arcs, so each lane takes arcs e0 + lane and e0 + lane +
// lanes, inserting every non-self arc into the S-slot table and keeping the
// keys it creates in k.
template <int S, bool Fixed>
__device__ __forceinline__ void scan_row(const LmArgs& A, int v, int64_t e0, int64_t e1, int lane, int lanes, int* key,
                                         unsigned long long* lo, unsigned long long* hi, OwnKeys& k)
{
    k.s[0] = k.s[1] = -1;
    const int64_t e = e0 + lane, f = e + lanes;
    const bool one = e < e1, two = f < e1;
    const int u1 = one ? (int)A.col[e] : v;
    const int u2 = two ? (int)A.col[f] : v;
    double w1 = 0.0, w2 = 0.0;
    int64_t q1 = 0, q2 = 0;
    if constexpr (Fixed) {
        q1 = one ? A.wq[e] : 0;
        q2 = two ? A.wq[f] : 0;
    } else {
        w1 = one ? A.w[e] : 0.0;
        w2 = two ? A.w[f] : 0.0;
    }
    const int c1 = u1 != v ? (int)A.C[u1] : -1;     // self arcs weigh 0 (graph_cuda.cu:2312-2315)
    const int c2 = u2 != v ? (int)A.C[u2] : -1;
    if (c1 >= 0) insert_own<S, Fixed>(A, key, lo, hi, c1, w1, q1, k, 0);
    if (c2 >= 0) insert_own<S, Fixed>(A, key, lo, hi, c2, w2, q2, k, 1);
}

// ---- launch 1 bodies (blocks [bi, bi+nbk) of the fused launch) -------------

template <bool Multi, bool Fixed>
__device__ __forceinline__ void thread_body(const LmArgs& A, const BinRow* __restrict__ verts, int64_t n, int bi, int nbk)
{
    for (int64_t i = (int64_t)bi * kBlock + threadIdx.x; i < n; i += (int64_t)nbk * kBlock) {
        const BinRow r = verts[i];
        const int v = r.v;
        Mass m;
        if (!graph_mass<Multi>(A, v, m)) continue;
        const int own = A.C[v];
        const double ki = A.kd[v];
        const fx_t kv = A.k[v];
        const int64_t e0 = r.e0;
        const int d = r.d;
        using W = typename std::conditional<Fixed, int64_t, double>::type;
        int uu[kThreadMaxDeg];
        W wd[kThreadMaxDeg];
#pragma unroll
        for (int j = 0; j < kThreadMaxDeg; ++j) {
            uu[j] = j < d ? (int)A.col[e0 + j] : v;
            if constexpr (Fixed) wd[j] = j < d ? A.wq[e0 + j] : 0;
            else wd[j] = j < d ? A.w[e0 + j] : 0.0;
        }
        int cc[kThreadMaxDeg];
#pragma unroll
        for (int j = 0; j < kThreadMaxDeg; ++j) cc[j] = uu[j] != v ? (int)A.C[uu[j]] : -1;
        // a_own is loaded once the row's gathers are on their way: it waits
        // for C[v], which would otherwise hold up the row's loads
        const fx_t rest = A.a[own] - kv;
        double self_d = 0.0;
        double bg = -kMaxFloat;
        int bc = kNone;
#pragma unroll
        for (int j = 0; j < kThreadMaxDeg; ++j) {
            bool first = cc[j] >= 0;
#pragma unroll
            for (int q = 0; q < j; ++q) first = first && (cc[q] != cc[j]);
            if (!first) continue;
            double e_d;
            if constexpr (Fixed) {
                int64_t e = wd[j];
#pragma unroll
                for (int q = j + 1; q < kThreadMaxDeg; ++q)
                    if (cc[q] == cc[j]) e += wd[q];
                e_d = (double)e * A.wscale;
            } else {
                fx_t e = fx_from_double(wd[j]);
#pragma unroll
                for (int q = j + 1; q < kThreadMaxDeg; ++q)
                    if (cc[q] == cc[j]) e += fx_from_double(wd[q]);
                e_d = fx_to_double(e);
            }
            if (cc[j] == own) self_d = e_d;
            else {
                const fx_t ac = A.a[cc[j]];
                consider_e(m, ki, cc[j], e_d, ac, bg, bc);
            }
        }
        A.Cn[v] = decide_d(m, ki, rest, own, self_d, bg, bc);
    }
}

// G lanes per vertex (G <= 64), S-slot LDS table per vertex; 256/G vertices
// per workgroup; LDS (256/G)*S*20 bytes carved from `lds` (12 per slot with
// fixed-point sums, which leave the high words out).
template <int G, int S, bool Multi, bool Fixed>
__device__ __forceinline__ void group_body(const LmArgs& A, const BinRow* __restrict__ verts, int64_t n, int bi, int nbk,
                           char* lds, double* sself)
{
    static_assert(S == 4 * G, "a row of at most 2G arcs fills at most half the table");
    constexpr int NG = kBlock / G;
    int* skey = reinterpret_cast<int*>(lds);
    unsigned long long* slo = reinterpret_cast<unsigned long long*>(lds + NG * S * 4);
    unsigned long long* shi = slo + NG * S;
    const int grp = threadIdx.x / G;
    const int lane = threadIdx.x & (G - 1);
    int* key = skey + grp * S;
    unsigned long long* lo = slo + grp * S;
    unsigned long long* hi = shi + grp * S;
    double* const self_d = sself + grp;

    for (int64_t base = (int64_t)bi * NG; base < n; base += (int64_t)nbk * NG) {
        for (int s = lane; s < S; s += G) {
            key[s] = -1;
            lo[s] = 0;
            if constexpr (!Fixed) hi[s] = 0;
        }
        if (lane == 0) *self_d = 0.0;
        const int64_t idx = base + grp;
        bool act = idx < n;
        const BinRow r = act ? verts[idx] : BinRow{0, 0, 0};
        const int v = r.v;
        Mass m{0.0, 0.0};
        act = act && graph_mass<Multi>(A, v, m);
        const int own = act ? A.C[v] : 0;
        const double ki = act ? A.kd[v] : 0.0;
        wave_lds_sync();
        OwnKeys k;
        k.s[0] = k.s[1] = -1;
        if (act) scan_row<S, Fixed>(A, v, r.e0, r.e0 + r.d, lane, G, key, lo, hi, k);
        // a_own after the row's loads (it waits for C[v])
        const fx_t rest = act ? A.a[own] - A.k[v] : fx_t(0);
        wave_lds_sync();
        if (act) {   // uniform within the group
            double bg = -kMaxFloat;
            int bc = kNone;
            sweep_own<Fixed>(A, m, ki, own, k, lo, hi, self_d, bg, bc);
#pragma unroll
            for (int msk = G / 2; msk > 0; msk >>= 1) {
                const double og = __shfl_xor(bg, msk, G);
                const int oc = __shfl_xor(bc, msk, G);
                if (better(og, oc, bg, bc)) {
                    bg = og;
                    bc = oc;
                }
            }
            wave_lds_sync();
            if (lane == 0) A.Cn[v] = decide_d(m, ki, rest, own, *self_d, bg, bc);
        }
        wave_lds_sync();
    }
}

// One workgroup per vertex, S-slot LDS table.
template <int S, bool Multi, bool Fixed>
__device__ __forceinline__ void block_body(const LmArgs& A, const BinRow* __restrict__ verts, int64_t n, int bi, int nbk,
                           char* lds, BlockRed& R, double* self_d)
{
    static_assert(S == 4 * kBlock, "a row of at most 2 * kBlock arcs fills at most half the table");
    int* key = reinterpret_cast<int*>(lds);
    unsigned long long* lo = reinterpret_cast<unsigned long long*>(lds + S * 4);
    unsigned long long* hi = lo + S;
    for (int64_t idx = bi; idx < n; idx += nbk) {
        for (int s = threadIdx.x; s < S; s += kBlock) {
            key[s] = -1;
            lo[s] = 0;
            if constexpr (!Fixed) hi[s] = 0;
        }
        if (threadIdx.x == 0) *self_d = 0.0;
        const BinRow r = verts[idx];
        const int v = r.v;
        Mass m;
        if (!graph_mass<Multi>(A, v, m)) continue;   // uniform: the workgroup holds one vertex
        const int own = A.C[v];
        const double ki = A.kd[v];
        __syncthreads();
        OwnKeys k;
        scan_row<S, Fixed>(A, v, r.e0, r.e0 + r.d, threadIdx.x, kBlock, key, lo, hi, k);
        const fx_t rest = A.a[own] - A.k[v];   // after the row's loads (it waits for C[v])
        __syncthreads();
        double bg = -kMaxFloat;
        int bc = kNone;
        sweep_own<Fixed>(A, m, ki, own, k, lo, hi, self_d, bg, bc);
        block_argmax(R, bg, bc);         // its barriers also publish *self_d
        if (threadIdx.x == 0) A.Cn[v] = decide_d(m, ki, rest, own, *self_d, bg, bc);
    }
}

// Key fill of one chunk of a partitioned long row: an LDS counting sort of
// the chunk's arcs by partition.  Each arc takes its rank inside its
// partition from an LDS atomic (the order inside a bucket is immaterial: the
// sums are exact and ties are broken by id); the exclusive scan of the P + 1
// counters -- the last one counts self arcs, which are dropped -- gives the
// segment starts.
constexpr int kFillUnroll = (int)(kKeyChunkArcs / kBlock);
constexpr int kFillScan = (kMaxParts + 1 + kBlock - 1) / kBlock;   // counters per thread in the scan

template <bool Multi>
__device__ __forceinline__ void keyfill_body(const LmArgs& A, const SmallBins& L, int bi, int nbk, char* lds)
{
    int* cnt = reinterpret_cast<int*>(lds);      // kMaxParts + 1 counters
    int* wtot = cnt + kMaxParts + 1;
    for (int64_t i = bi; i < L.nchunks; i += nbk) {
        const KeyChunk ch = L.chunks[i];
        Mass m;
        if (!graph_mass<Multi>(A, ch.v, m)) continue;   // uniform over the workgroup
        const int P = ch.P;
        const int psh = 32 - (31 - __builtin_clz((unsigned)P));   // top log2(P) bits of the hash
        for (int s = threadIdx.x; s <= P; s += kBlock) cnt[s] = 0;
        __syncthreads();
        // cc: community key (-1 self arc, -2 past the chunk); rk: rank in
        // its partition.  Weights are read again after the scan (L2-warm)
        // rather than held in registers across it.
        int cc[kFillUnroll], rk[kFillUnroll];
#pragma unroll
        for (int j = 0; j < kFillUnroll; ++j) {
            const int64_t e = ch.e0 + j * kBlock + threadIdx.x;
            cc[j] = e < ch.e1 ? A.col[e] : -2;
        }
#pragma unroll
        for (int j = 0; j < kFillUnroll; ++j) cc[j] = cc[j] >= 0 ? (cc[j] != ch.v ? A.C[cc[j]] : -1) : -2;
        auto part_of = [&](int c) { return c == -2 ? -1 : c < 0 ? P : (int)(((unsigned)c * kPartMul) >> psh); };
#pragma unroll
        for (int j = 0; j < kFillUnroll; ++j) {
            const int pj = part_of(cc[j]);
            rk[j] = pj >= 0 ? atomicAdd(&cnt[pj], 1) : 0;
        }
        __syncthreads();
        // thread t scans counters [t * kFillScan, t * kFillScan + kFillScan)
        int loc[kFillScan];
        int sum = 0;
#pragma unroll
        for (int q = 0; q < kFillScan; ++q) {
            const int idx = threadIdx.x * kFillScan + q;
            loc[q] = idx <= P ? cnt[idx] : 0;
            sum += loc[q];
        }
        int tot;
        int pre = block_excl_scan_i32(sum, wtot, &tot);   // synchronises after every read of cnt
#pragma unroll
        for (int q = 0; q < kFillScan; ++q) {
            const int idx = threadIdx.x * kFillScan + q;
            if (idx <= P) {
                cnt[idx] = pre;
                L.cpo[ch.c0 + idx] = pre;
            }
            pre += loc[q];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kFillUnroll; ++j) {
            const int pj = part_of(cc[j]);
            if (pj >= 0 && pj < P) {
                const int64_t pos = ch.k0 + cnt[pj] + rk[j];
                L.bkey[pos] = cc[j];
                L.bw[pos] = A.w[ch.e0 + j * kBlock + threadIdx.x];
            }
        }
        __syncthreads();          // cnt and wtot are reused by the next chunk
    }
}

constexpr int kSmallLds = 20480;   // (256/G)*S*20 for every short-row bin
constexpr int kSmallLdsFixed = 12288;   // (256/G)*S*12: fixed-point sums need no high words
static_assert((kMaxParts + 1 + kBlock / 64) * 4 <= kSmallLdsFixed, "key-fill counters exceed the LDS buffer");
static_assert((kMaxParts + 1 + kBlock / 64) * 4 <= kSmallLds, "key-fill counters exceed the LDS buffer");

// Seven waves per SIMD; eight for a handle of many graphs with fixed-point
// sums (12 KB of tables instead of 20 KB, registers held to 64 with a few
// spilled): on 4,096 SBMs, whose rows are short, their iterations took
// 0.339 s against 0.350 s, while the same form cost R-MAT scale 22 2%.
template <bool Multi, bool Fixed>
__global__ void __launch_bounds__(kBlock, Multi && Fixed ? 8 : 7)
lm_small_kernel(LmArgs A, SmallBins L)
{
    __shared__ __align__(16) char lds[Fixed ? kSmallLdsFixed : kSmallLds];
    __shared__ BlockRed R;
    __shared__ double sself[kBlock / 8];       // e_{v,own} per row, rounded
    // Every index into L is a constant (a dynamically indexed kernel
    // argument would be copied to scratch memory).
    // Blocks are numbered backwards, so the dispatcher starts the longest
    // workgroups (delta, key fill, whole-workgroup rows) first.
    const int b = (int)(gridDim.x - 1 - blockIdx.x);
    int bin = 0;   // 0..5 short-row bins, 6 key fill, 7 the previous batch's modularity delta
#pragma unroll
    for (int k = 1; k <= NUM_SMALL_BINS + 1; ++k) bin += b >= L.blk0[k];
#define NIDO_BIN_ARGS(K) L.verts[K], L.n[K], b - L.blk0[K], L.blk0[K + 1] - L.blk0[K]
    switch (bin) {
    case BIN_THREAD: thread_body<Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_THREAD)); break;
    case BIN_G8: group_body<8, 32, Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_G8), lds, sself); break;
    case BIN_G16: group_body<16, 64, Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_G16), lds, sself); break;
    case BIN_G32: group_body<32, 128, Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_G32), lds, sself); break;
    case BIN_G64: group_body<64, 256, Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_G64), lds, sself); break;
    case BIN_B256: block_body<1024, Multi, Fixed>(A, NIDO_BIN_ARGS(BIN_B256), lds, R, sself); break;
    case NUM_SMALL_BINS:
        keyfill_body<Multi>(A, L, b - L.blk0[NUM_SMALL_BINS], L.blk0[NUM_SMALL_BINS + 1] - L.blk0[NUM_SMALL_BINS],
                            lds);
        break;
    default:
        delta_body(L.db, L.dv0, L.dv1, L.Cn64, A.C, A.col, A.w, A.a, L.twom_d, L.D, b - L.blk0[NUM_SMALL_BINS + 1],
                   L.blk0[NUM_SMALL_BINS + 2] - L.blk0[NUM_SMALL_BINS + 1], reinterpret_cast<fx_t*>(lds));
        break;
    }
#undef NIDO_BIN_ARGS
}

// ---- launch 2: partitioned long rows ---------------------------------------

// Whole-row aggregation in an HBM table of `size` slots (>= 2 * row length):
// the exact fallback taken when a partition outgrows its LDS table.
__device__ __forceinline__ void global_row_decide(const LmArgs& A, Mass m, int v, int* __restrict__ key,
                                  fx_t* __restrict__ tval, int64_t size, BlockRed& R)
{
    unsigned long long* val = reinterpret_cast<unsigned long long*>(tval);
    const int sh = 32 - (63 - __builtin_clzll((unsigned long long)size));
    for (int64_t s = threadIdx.x; s < size; s += kBlock) {
        key[s] = -1;
        val[2 * s] = 0;
        val[2 * s + 1] = 0;
    }
    __threadfence();
    __syncthreads();
    const int own = (int)A.C[v];
    const int64_t e1 = A.off[v + 1];
    for (int64_t e = A.off[v] + threadIdx.x; e < e1; e += kBlock) {
        const int u = (int)A.col[e];
        if (u == v) continue;
        const int c = (int)A.C[u];
        const fx_t wf = fx_from_double(A.w[e]);
        uint64_t h = sh >= 32 ? 0 : (((unsigned)c * kHashMul) >> sh);
        for (;;) {
            const int kk = ld_i32_agent(&key[h]);
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
        const int c = ld_i32_agent(&key[s]);
        if (c >= 0) consider(A, m, ki, own, c, ld_fx_agent(tval + s), self, bg, bc);
    }
    block_reduce(R, self, bg, bc);
    if (threadIdx.x == 0) A.Cn[v] = decide(A, m, v, own, self, bg, bc);
}

constexpr int kMaxProbe = 256;

template <bool Narrow> struct HiWord { using type = unsigned long long; };
template <> struct HiWord<true> { using type = int; };

__device__ __forceinline__ void lds_add(unsigned long long* lo, unsigned long long* hi, fx_t w)
{
    fx_atomic_add(lo, hi, w);
}
__device__ __forceinline__ void lds_add(unsigned long long* lo, int* hi, fx_t w)
{
    fx_atomic_add_narrow(lo, hi, w);
}

constexpr int kFill = 6;      // entries per thread loaded ahead of their inserts
constexpr int kOverflow = -2; // partial community of a partition that outgrew its table

__device__ __forceinline__ int rfl(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t rfl64(int64_t x)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// An item read with vector loads: a scalar load would share its wait
// counter with the LDS traffic of the item in progress, and every LDS wait
// would then wait for the prefetch too.
__device__ __forceinline__ BigItem load_item_v(const BigItem* items, int64_t i)
{
    asm volatile("" : "+v"(i));
    return items[i];
}

// Segment start and length of partition x.p in chunk j of item x's row
// (zero past the row's chunks or for an unpartitioned row).
__device__ __forceinline__ void item_segment(const BigArgs& B, const BigItem& x, int j, int& l0, int& len)
{
    l0 = 0;
    len = 0;
    const int nch = x.c0 >= 0 ? (int)((x.d + kKeyChunkArcs - 1) / kKeyChunkArcs) : 0;
    if (j < nch) {
        const int* cp = B.cpo + x.c0 + (int64_t)j * (x.P + 1) + x.p;
        l0 = cp[0];
        len = cp[1] - l0;
    }
}

// S-slot LDS table per workgroup.  Narrow: 32-bit high words (S*16 bytes),
// used when every row's weights are non-negative and sum to less than 2^31.
// Items come longest row first, so the long poles start first.
//
// A row with P <= kDirectParts is read straight from the CSR (for P = 2 each
// of the two workgroups keeps the arcs of its partition).  Otherwise the key fill
// has bucketed each chunk of the row by partition, so workgroup (v, p) reads
// only its own segment of every chunk: the segment lengths of up to 256
// chunks are scanned across the workgroup and the concatenated segments are
// processed 256 entries at a time, each thread locating its chunk by binary
// search in the scanned starts.  Every key a workgroup reads belongs to it.
template <bool Narrow, int S, bool Multi>
__global__ void __launch_bounds__(kBlock)
lm_big_kernel(LmArgs A, BigArgs B)
{
    using Hi = typename HiWord<Narrow>::type;
    constexpr int kList = S / 2 + kBlock;   // key_limit (<= S/2) + one step of every wavefront
    __shared__ int key[S];
    __shared__ unsigned long long lo[S];
    __shared__ Hi hi[S];
    __shared__ int klist[kList];            // occupied slots, in order of first insertion
    __shared__ int s_pre[kBlock], s_lo[kBlock], wtot[kBlock / 64];
    __shared__ BlockRed R;
    __shared__ int s_nkeys, s_ovf, s_last;
    const int lane = threadIdx.x & 63;

    // Software pipeline over the workgroup's items: the next item's fields
    // and its first group of segment starts are loaded while the current
    // item runs, so an item starts with its key reads.
    int64_t it = blockIdx.x;
    BigItem nx{};
    int n_l0 = 0, n_len = 0;
    if (it < B.nitems) {
        nx = load_item_v(B.items, it);
        item_segment(B, nx, threadIdx.x, n_l0, n_len);
    }
    for (; it < B.nitems; it += gridDim.x) {
        const int v = rfl(nx.v), p = rfl(nx.p), P = rfl(nx.P), slot = rfl(nx.slot);
        const int64_t k0 = rfl64(nx.k0), c0 = rfl64(nx.c0), e0 = rfl64(nx.e0), d = rfl64(nx.d);
        const int c_l0 = n_l0, c_len = n_len;
        const int64_t itn = it + gridDim.x;
        if (itn < B.nitems) nx = load_item_v(B.items, itn);
        auto prefetch_next = [&]() {
            if (itn < B.nitems) item_segment(B, nx, threadIdx.x, n_l0, n_len);
        };
        Mass m;
        if (!graph_mass<Multi>(A, v, m)) {   // uniform; every partition of the row skips it
            prefetch_next();
            continue;
        }
        const int own = A.C[v];
        const double ki = A.kd[v];
        const bool whole = c0 < 0;       // the row is read from the CSR (P = 1 or few partitions)
        const int nch = whole ? 0 : (int)((d + kKeyChunkArcs - 1) / kKeyChunkArcs);
        const int psh = 32 - (31 - __builtin_clz((unsigned)P));   // top log2(P) bits of the hash

        // Segments of chunks [g * kBlock, g * kBlock + kBlock) of partition p:
        // s_pre = exclusive scan of their lengths, s_lo = their starts.
        auto scan_group = [&](int g) -> int {
            const int j = g * kBlock + threadIdx.x;
            int l0 = c_l0, len = c_len;
            if (g > 0) {
                l0 = 0;
                len = 0;
                if (j < nch) {
                    const int* cp = B.cpo + c0 + (int64_t)j * (P + 1) + p;
                    l0 = cp[0];
                    len = cp[1] - l0;
                }
            }
            int tot;
            const int pre = block_excl_scan_i32(len, wtot, &tot);
            s_pre[threadIdx.x] = pre;
            s_lo[threadIdx.x] = l0;
            __syncthreads();
            return tot;
        };
        int total = whole ? (int)d : scan_group(0);
        // The table needs at most 2x the partition's entries (itself a bound
        // on its distinct keys): the smallest power of two >= 2 * entries, at
        // least 256, so the clear and the sweep cost O(entries) rather than O(S).
        const int64_t entries = (whole && P > 1) || nch > kBlock ? (int64_t)S : (int64_t)total;
        int lg = 8;
        while ((1ll << lg) < 2 * entries && lg < ilog2c(S)) ++lg;
        const int ts = 1 << lg;
        for (int s = threadIdx.x; s < ts; s += kBlock) {
            key[s] = -1;
            lo[s] = 0;
            hi[s] = 0;
        }
        if (threadIdx.x == 0) {
            s_nkeys = 0;
            s_ovf = 0;
        }
        __syncthreads();

        // Inserts (c, w) when c >= 0.  Probing gives up after kMaxProbe
        // slots and marks the partition overflowed (the row then takes the
        // exact HBM path), so no lane can spin on a full table.  New keys are
        // counted per wavefront step and their slots appended to klist; past
        // key_limit (<= S/2) the partition is marked overflowed.  A wavefront
        // step adds at most 64 keys, so a partition that stays within the
        // limit has every slot in the list; entries past kList are dropped
        // (their partition has overflowed).  Called by all lanes of the
        // workgroup with a uniform trip count.
        auto insert = [&](int c, double w) {
            bool fresh = false;
            unsigned h = 0;
            if (c >= 0) {
                h = ((unsigned)c * kHashMul) >> (32 - lg);
                bool ok = false;
                for (int probes = 0; probes < kMaxProbe; ++probes) {
                    const int kk = key[h];
                    if (kk == c) { ok = true; break; }
                    if (kk < 0) {
                        const int prev = atomicCAS(&key[h], -1, c);
                        if (prev < 0 || prev == c) {
                            ok = true;
                            fresh = prev < 0;
                            break;
                        }
                    }
                    h = (h + 1) & (ts - 1);
                }
                if (ok) lds_add(&lo[h], &hi[h], fx_from_double(w));
                else s_ovf = 1;
            }
            const unsigned long long nf = __ballot(fresh);
            if (nf) {                                   // uniform
                int base = 0;
                if (lane == 0) {
                    base = atomicAdd(&s_nkeys, __popcll(nf));
                    if (base + __popcll(nf) > B.key_limit) s_ovf = 1;
                }
                base = __shfl(base, 0, 64);
                const int at = base + lane_rank(nf);
                if (fresh && at < kList) klist[at] = (int)h;
            }
        };

        // Each pass loads kFill entries per thread before inserting any, so
        // their gathers are in flight together (LDS caps this kernel at four
        // workgroups per CU with 2048 slots; registers are not the limit).
        if (whole) {
            for (int64_t i0 = 0; i0 < d; i0 += kFill * kBlock) {
                if (s_ovf) break;
                int us[kFill], cs[kFill];
                double ws[kFill];
#pragma unroll
                for (int r = 0; r < kFill; ++r) {
                    const int64_t i = i0 + r * kBlock + threadIdx.x;
                    us[r] = i < d ? A.col[e0 + i] : v;
                    ws[r] = i < d ? A.w[e0 + i] : 0.0;
                }
#pragma unroll
                for (int r = 0; r < kFill; ++r) cs[r] = us[r] != v ? A.C[us[r]] : -1;   // self arcs weigh 0 (graph_cuda.cu:2312-2315)
                if (P > 1) {
#pragma unroll
                    for (int r = 0; r < kFill; ++r)
                        if (cs[r] >= 0 && (int)(((unsigned)cs[r] * kPartMul) >> psh) != p) cs[r] = -1;
                }
#pragma unroll
                for (int r = 0; r < kFill; ++r) insert(cs[r], ws[r]);
            }
        } else {
            for (int g = 0; g * kBlock < nch; ++g) {
                if (g > 0) {
                    __syncthreads();              // every lane is done with the previous group
                    if (s_ovf) break;             // uniform after the barrier
                    total = scan_group(g);
                }
                const int ng = min(kBlock, nch - g * kBlock);
                const int64_t cbase = k0 + (int64_t)g * kBlock * kKeyChunkArcs;
                int top = 1;                                 // largest power of two < ng (or 1)
                while (2 * top < ng) top *= 2;
                for (int i0 = 0; i0 < total; i0 += kFill * kBlock) {
                    if (s_ovf) break;
                    int l[kFill], cs[kFill];
                    double ws[kFill];
#pragma unroll
                    for (int r = 0; r < kFill; ++r) l[r] = 0;
                    // last chunk whose start is <= i, all kFill searches in step
                    for (int st = top; st > 0; st >>= 1) {
#pragma unroll
                        for (int r = 0; r < kFill; ++r) {
                            const int i = i0 + r * kBlock + threadIdx.x;
                            const int j = l[r] + st;
                            if (j < ng && s_pre[j] <= i) l[r] = j;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < kFill; ++r) {
                        const int i = i0 + r * kBlock + threadIdx.x;
                        cs[r] = -1;
                        ws[r] = 0.0;
                        if (i < total) {
                            const int64_t at = cbase + (int64_t)l[r] * kKeyChunkArcs + s_lo[l[r]] + (i - s_pre[l[r]]);
                            cs[r] = B.bkey[at];
                            ws[r] = B.bw[at];
                        }
                    }
#pragma unroll
                    for (int r = 0; r < kFill; ++r) insert(cs[r], ws[r]);
                }
            }
        }
        __syncthreads();
        prefetch_next();
        const bool ovf = s_ovf != 0;
        fx_t self = 0;
        double bg = -kMaxFloat;
        int bc = kNone;
        if (!ovf) {
            // score the occupied slots from the list, two per thread at a
            // time with both a_c gathers in flight
            const int nk = s_nkeys;
            for (int i = threadIdx.x; i < nk; i += 2 * kBlock) {
                const int s1 = klist[i];
                const int s2 = i + kBlock < nk ? klist[i + kBlock] : -1;
                const int c1 = key[s1];
                const int c2 = s2 >= 0 ? key[s2] : -1;
                const fx_t a1 = A.a[c1];
                const fx_t a2 = c2 >= 0 ? A.a[c2] : fx_t(0);
                consider_a(m, ki, own, c1, fx_make(lo[s1], (int64_t)hi[s1]), a1, self, bg, bc);
                if (c2 >= 0) consider_a(m, ki, own, c2, fx_make(lo[s2], (int64_t)hi[s2]), a2, self, bg, bc);
            }
        }
        block_reduce(R, self, bg, bc);
        const int64_t tb = B.toff[slot];
        const int64_t tsize = B.toff[slot + 1] - tb;
        if (P == 1) {
            if (ovf) global_row_decide(A, m, v, B.tkey + tb, B.tval + tb, tsize, R);
            else if (threadIdx.x == 0) A.Cn[v] = decide(A, m, v, own, self, bg, bc);
            __syncthreads();
            continue;
        }
        // The partial reaches the row's last workgroup through memory-side
        // atomics only: agent-scope exchanges, whose returned old values are
        // awaited before the counter increment is issued, and agent-scope
        // atomic loads on the reading side.  An agent-scope fence is avoided
        // on purpose: on this GPU it writes back and invalidates the XCD's
        // whole L2 under every other workgroup resident there.  An overflowed
        // partition reports community kOverflow.
        if (threadIdx.x == 0) {
            const int64_t q = B.poff[slot] + p;
            unsigned long long* pgw = reinterpret_cast<unsigned long long*>(B.pg + q);
            unsigned long long* psw = reinterpret_cast<unsigned long long*>(B.pself + q);
            const unsigned long long r0 = __hip_atomic_exchange(
                pgw, (unsigned long long)__double_as_longlong(bg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int r1 = __hip_atomic_exchange(B.pc + q, ovf ? kOverflow : bc, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long r2 = __hip_atomic_exchange(psw, fx_lo(self), __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long r3 = __hip_atomic_exchange(
                psw + 1, (unsigned long long)fx_hi(self), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("" : : "v"(r0), "v"(r1), "v"(r2), "v"(r3) : "memory");
            s_last = __hip_atomic_fetch_add(&B.cnt[slot], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == P - 1;
        }
        __syncthreads();
        if (s_last) {                      // this workgroup finished the row last: combine
            fx_t s2 = 0;                   // the P partials, one per thread
            double g2 = -kMaxFloat;
            int c2 = kNone;
            int ovf2 = 0;
            const int64_t q0 = B.poff[slot];
            for (int i = threadIdx.x; i < P; i += kBlock) {
                const int ci = ld_i32_agent(B.pc + q0 + i);
                s2 += ld_fx_agent(B.pself + q0 + i);
                const double gi = __longlong_as_double(
                    (long long)ld_u64_agent(reinterpret_cast<const unsigned long long*>(B.pg + q0 + i)));
                if (ci == kOverflow) ovf2 = 1;
                else if (better(gi, ci, g2, c2)) {
                    g2 = gi;
                    c2 = ci;
                }
            }
            const int any_ovf = __syncthreads_or(ovf2);
            block_reduce(R, s2, g2, c2);
            if (threadIdx.x == 0) B.cnt[slot] = 0;   // leave the counter clean for the next batch
            if (any_ovf) global_row_decide(A, m, v, B.tkey + tb, B.tval + tb, tsize, R);
            else if (threadIdx.x == 0) A.Cn[v] = decide(A, m, v, own, s2, g2, c2);
        }
        __syncthreads();
    }
}

// Community-total updates of a commit, aggregated per workgroup: the moved
// vertices of a kBlock-vertex chunk add -k_v to their old and +k_v to their
// new community in an LDS table, and each distinct community of the chunk
// then takes one global atomic.  Rows moving into one popular community
// would otherwise serialise on its a_c.  The sums are exact, so the totals
// do not depend on the grouping.  The table is cleared once per workgroup;
// new keys list their slots, and after each chunk only the listed slots are
// cleared, so a chunk costs in proportion to its moves (and nothing beyond
// one barrier when none of its vertices moved, the common case after the
// first iterations of a phase).
constexpr int kCommitSlots = 4 * kBlock;   // at most 2 * kBlock keys: load <= 1/2
struct CommitTable {
    int key[kCommitSlots];
    unsigned long long lo[kCommitSlots];
    unsigned long long hi[kCommitSlots];
    int list[2 * kBlock];                  // slots of this chunk's keys
    int n;                                 // keys listed
};

__device__ __forceinline__ void ctab_init(CommitTable& T)
{
#pragma unroll
    for (int j = 0; j < kCommitSlots / kBlock; ++j) {
        const int s = threadIdx.x + j * kBlock;
        T.key[s] = -1;
        T.lo[s] = 0;
        T.hi[s] = 0;
    }
    if (threadIdx.x == 0) T.n = 0;
}

__device__ __forceinline__ void ctab_add(CommitTable& T, int c, fx_t x)
{
    unsigned h = ((unsigned)c * kHashMul) >> (32 - ilog2c(kCommitSlots));
    for (;;) {   // ends: the table never holds more than half its slots
        const int kk = T.key[h];
        if (kk == c) break;
        if (kk < 0) {
            const int prev = atomicCAS(&T.key[h], -1, c);
            if (prev < 0) {
                T.list[atomicAdd(&T.n, 1)] = (int)h;
                break;
            }
            if (prev == c) break;
        }
        h = (h + 1) & (kCommitSlots - 1);
    }
    fx_atomic_add(&T.lo[h], &T.hi[h], x);
}

// Listed slot i of the chunk (i < T.n): its key and sum, and the slot cleared.
__device__ __forceinline__ int ctab_take(CommitTable& T, int i, fx_t& x)
{
    const int s = T.list[i];
    const int c = T.key[s];
    x = fx_make(T.lo[s], (int64_t)T.hi[s]);
    T.key[s] = -1;
    T.lo[s] = 0;
    T.hi[s] = 0;
    return c;
}

// update_community_weights + update_community_ids of one batch, fused
// (graph_cuda.cu:1234-1249 then 1191-1196): each thread touches only its own
// vertex's C/C' entries, so the swap needs no second launch.  C32 mirrors C.
// The vertices are list[0, n) or, with a null list, [v0, v0 + n); with gid
// set, only vertices of graphs whose flag is set take part.  A vertex that
// did not move has C = C' already, so its swap writes nothing.
__global__ void __launch_bounds__(kBlock)
commit_kernel(const BinRow* __restrict__ list, int64_t v0, int64_t n, int64_t* __restrict__ C,
              int64_t* __restrict__ Cn, int* __restrict__ C32, const fx_t* __restrict__ k,
              fx_t* __restrict__ a, const int* __restrict__ gid, const unsigned char* __restrict__ flag,
              const int64_t* __restrict__ off, unsigned long long* __restrict__ moved_arcs)
{
    __shared__ CommitTable T;
    ctab_init(T);
    unsigned long long marcs = 0;
    for (int64_t base = blockIdx.x * (int64_t)kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
        const int64_t i = base + threadIdx.x;
        int64_t v = 0, src = 0, dst = 0;
        bool moved = false;
        if (i < n) {
            v = list ? (int64_t)list[i].v : v0 + i;
            if (!gid || flag[gid[v]]) {
                src = C[v];
                dst = Cn[v];
                moved = src != dst;
            }
        }
        if (!__syncthreads_or(moved)) continue;   // uniform; also orders the previous chunk's clears
        if (moved) {
            const fx_t kv = k[v];
            ctab_add(T, (int)src, -kv);
            ctab_add(T, (int)dst, kv);
            C32[v] = (int)dst;
            C[v] = dst;
            Cn[v] = src;
            if (moved_arcs) marcs += list ? (unsigned long long)list[i].d : (unsigned long long)(off[v + 1] - off[v]);
        }
        __syncthreads();
        const int nk = T.n;
        for (int j = threadIdx.x; j < nk; j += kBlock) {
            fx_t x;
            const int c = ctab_take(T, j, x);
            if (x != 0) acc_add(a + c, x);
        }
        __syncthreads();                 // every thread has read T.n
        if (threadIdx.x == 0) T.n = 0;
    }
    if (moved_arcs) {   // uniform
        __shared__ unsigned long long red[kBlock / 64];
        add_moved_arcs(moved_arcs, marcs, red);
    }
}

// The commit plus the bookkeeping of the incremental modularity: a moved
// vertex lists its row (in chunks) and its old and new community (each
// community once per batch, through its touched flag).  The vertices are
// list[0, n) or, with a null list, [v0, v0 + n); with D.gid set, only
// vertices of graphs whose flag is set take part.
__global__ void __launch_bounds__(kBlock)
commit_tracked_kernel(const BinRow* __restrict__ list, int64_t v0, int64_t n, int b, int64_t* __restrict__ C,
                      int64_t* __restrict__ Cn, int* __restrict__ C32, const fx_t* __restrict__ k,
                      fx_t* __restrict__ a, const int64_t* __restrict__ off, ModDelta D,
                      const unsigned char* __restrict__ flag)
{
    __shared__ CommitTable T;
    __shared__ int wtot[kBlock / 64];
    __shared__ int s_chunk0, s_comm0;
    ctab_init(T);
    unsigned long long marcs = 0;
    // workgroup-uniform loop: list space is reserved once per workgroup
    // and chunk, both lists with one 64-bit atomic on the batch's counter
    // pair {chunks, communities} (no carry crosses: each count stays below
    // 2^31)
    for (int64_t base = blockIdx.x * (int64_t)kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
        const int64_t i = base + threadIdx.x;
        int64_t v = 0, src = 0, dst = 0;
        bool moved = false;
        if (i < n) {
            v = list ? (int64_t)list[i].v : v0 + i;
            if (!D.gid || flag[D.gid[v]]) {
                src = C[v];
                dst = Cn[v];
                moved = src != dst;
            }
        }
        if (!__syncthreads_or(moved)) continue;   // uniform; also orders the previous chunk's clears
        int nchunk = 0;
        int64_t e0 = 0, e1 = 0;
        if (moved) {
            const fx_t kv = k[v];
            ctab_add(T, (int)src, -kv);
            ctab_add(T, (int)dst, kv);
            C32[v] = (int)dst;
            C[v] = dst;
            Cn[v] = src;
            e0 = off[v];
            e1 = off[v + 1];
            nchunk = (int)((e1 - e0 + kDeltaArcs - 1) / kDeltaArcs);
            marcs += (unsigned long long)(e1 - e0);
        }
        __syncthreads();                  // the table is complete
        // every community a moved vertex left or joined: its total, and a
        // place in the touched list the first time in this batch (at most
        // 2 * kBlock keys: two per thread); a flag already seen set needs no
        // exchange (flags go back to 0 only after the batch's commit)
        const int nk = T.n;
        int cs[2] = {-1, -1};
        int nf = 0;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int i = threadIdx.x + j * kBlock;
            if (i < nk) {
                fx_t x;
                const int c = ctab_take(T, i, x);
                if (x != 0) acc_add(a + c, x);
                if (ld_i32_agent(&D.touched[c]) == 0 && atomicExch(&D.touched[c], 1) == 0) {
                    cs[j] = c;
                    ++nf;
                }
            }
        }
        int tc;
        const int pc = block_excl_scan_i32(nchunk, wtot, &tc);   // its barrier: every thread has read T.n
        __syncthreads();                  // wtot read before the next scan writes it
        int tm;
        int at = block_excl_scan_i32(nf, wtot, &tm);
        if (threadIdx.x == 0) {
            if (tc | tm) {
                const unsigned long long old =
                    atomicAdd(reinterpret_cast<unsigned long long*>(&D.counts[2 * b]),
                              (unsigned long long)(unsigned)tc | ((unsigned long long)(unsigned)tm << 32));
                s_chunk0 = (int)(unsigned)old;
                s_comm0 = (int)(old >> 32);
            }
            T.n = 0;
        }
        __syncthreads();
        for (int i = 0; i < nchunk; ++i) {
            const int64_t s0 = e0 + (int64_t)i * kDeltaArcs;
            D.chunks[s_chunk0 + pc + i] = RowItem{v, s0, s0 + kDeltaArcs < e1 ? s0 + kDeltaArcs : e1};
        }
        at += s_comm0;
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (cs[j] >= 0) D.comms[at++] = cs[j];
    }
    if (D.moved) {   // uniform
        __shared__ unsigned long long red[kBlock / 64];
        add_moved_arcs(D.moved, marcs, red);
    }
}

}  // namespace

// Rows per workgroup of each short-row bin.
static constexpr int kRowsPerBlock[NUM_SMALL_BINS] = {kBlock, kBlock / 8, kBlock / 16, kBlock / 32,
                                                      kBlock / 64, 1};

int small_bins_blocks(SmallBins& L, unsigned cap_per_bin)
{
    int total = 0;
    for (int b = 0; b < NUM_SMALL_BINS; ++b) {
        L.blk0[b] = total;
        if (L.n[b] > 0) total += (int)grid_for(L.n[b], kRowsPerBlock[b], cap_per_bin);
    }
    L.blk0[NUM_SMALL_BINS] = total;
    if (L.nchunks > 0) total += (int)grid_for(L.nchunks, 1, cap_per_bin);
    L.blk0[NUM_SMALL_BINS + 1] = total;
    total += L.dblocks;
    L.blk0[NUM_SMALL_BINS + 2] = total;
    return total;
}

static void small_launch(const LmArgs& A, const SmallBins& L, int blocks, hipStream_t s)
{
    if (A.gid) {
        if (A.wq) lm_small_kernel<true, true><<<blocks, kBlock, 0, s>>>(A, L);
        else lm_small_kernel<true, false><<<blocks, kBlock, 0, s>>>(A, L);
    } else {
        if (A.wq) lm_small_kernel<false, true><<<blocks, kBlock, 0, s>>>(A, L);
        else lm_small_kernel<false, false><<<blocks, kBlock, 0, s>>>(A, L);
    }
}

void launch_lm_small(const LmArgs& A, const SmallBins& L, bool bins, bool fill, bool split, hipStream_t s,
                     bool delta)
{
    // parts: 0..5 short-row bins, 6 key fill, 7 delta (goes with the bins)
    bool want[NUM_SMALL_BINS + 2];
    for (int k = 0; k < NUM_SMALL_BINS + 2; ++k) want[k] = k == NUM_SMALL_BINS ? fill : bins;
    want[NUM_SMALL_BINS + 1] = bins && delta && L.dblocks > 0;
    auto launch_parts = [&](const bool* on) {
        SmallBins one = L;
        for (int k = 0; k < NUM_SMALL_BINS; ++k) one.n[k] = on[k] ? L.n[k] : 0;
        if (!on[NUM_SMALL_BINS]) one.nchunks = 0;
        if (!on[NUM_SMALL_BINS + 1]) one.dblocks = 0;
        const int blocks = small_bins_blocks(one, 4096);
        if (blocks > 0) small_launch(A, one, blocks, s);
    };
    if (!split) {
        launch_parts(want);
        return;
    }
    for (int k = 0; k < NUM_SMALL_BINS + 2; ++k) {
        if (!want[k]) continue;
        bool on[NUM_SMALL_BINS + 2] = {};
        on[k] = true;
        launch_parts(on);
    }
}

void launch_lm_big(const LmArgs& A, const BigArgs& B, hipStream_t s)
{
    if (B.nitems <= 0) return;
    // one workgroup per item: the dispatcher hands items to CUs as earlier
    // ones finish (static round-robin over fewer workgroups was 0.7% slower)
    const unsigned g = grid_for(B.nitems, 1, 65535);
    // a 1024-slot table keeps the workgroup at ~20 KB of LDS, so it can run
    // beside the short-row workgroups (overlap_long)
    const int size = B.slots == 1024 ? 0 : B.slots == 2048 ? 1 : 2;
    const int which = size * 4 + (B.narrow ? 0 : 2) + (A.gid ? 1 : 0);
    switch (which) {
    case 0: lm_big_kernel<true, 1024, false><<<g, kBlock, 0, s>>>(A, B); break;
    case 1: lm_big_kernel<true, 1024, true><<<g, kBlock, 0, s>>>(A, B); break;
    case 2: lm_big_kernel<false, 1024, false><<<g, kBlock, 0, s>>>(A, B); break;
    case 3: lm_big_kernel<false, 1024, true><<<g, kBlock, 0, s>>>(A, B); break;
    case 4: lm_big_kernel<true, 2048, false><<<g, kBlock, 0, s>>>(A, B); break;
    case 5: lm_big_kernel<true, 2048, true><<<g, kBlock, 0, s>>>(A, B); break;
    case 6: lm_big_kernel<false, 2048, false><<<g, kBlock, 0, s>>>(A, B); break;
    case 7: lm_big_kernel<false, 2048, true><<<g, kBlock, 0, s>>>(A, B); break;
    case 8: lm_big_kernel<true, 4096, false><<<g, kBlock, 0, s>>>(A, B); break;
    case 9: lm_big_kernel<true, 4096, true><<<g, kBlock, 0, s>>>(A, B); break;
    case 10: lm_big_kernel<false, 4096, false><<<g, kBlock, 0, s>>>(A, B); break;
    default: lm_big_kernel<false, 4096, true><<<g, kBlock, 0, s>>>(A, B); break;
    }
}

void launch_commit(int64_t v0, int64_t v1, int64_t* C, int64_t* Cn, int* C32, const fx_t* k, fx_t* a,
                   hipStream_t s, const int64_t* off, unsigned long long* moved)
{
    if (v1 <= v0) return;
    commit_kernel<<<grid_for(v1 - v0, kBlock), kBlock, 0, s>>>(nullptr, v0, v1 - v0, C, Cn, C32, k, a,
                                                               nullptr, nullptr, off, moved);
}

void launch_commit_tracked(int64_t v0, int64_t v1, int b, int64_t* C, int64_t* Cn, int* C32, const fx_t* k,
                           fx_t* a, const int64_t* off, const ModDelta& D, hipStream_t s)
{
    if (v1 <= v0) return;
    commit_tracked_kernel<<<grid_for(v1 - v0, kBlock), kBlock, 0, s>>>(nullptr, v0, v1 - v0, b, C, Cn, C32, k, a,
                                                                       off, D, nullptr);
}

void launch_commit_tracked_list(const BinRow* list, int64_t n, int b, int64_t* C, int64_t* Cn, int* C32,
                                const fx_t* k, fx_t* a, const int64_t* off, const ModDelta& D,
                                const unsigned char* flag, hipStream_t s)
{
    if (n <= 0) return;
    commit_tracked_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(list, 0, n, b, C, Cn, C32, k, a, off, D, flag);
}

void launch_commit_filtered(const BinRow* list, int64_t v0, int64_t n, int64_t* C, int64_t* Cn, int* C32,
                            const fx_t* k, fx_t* a, const int* gid, const unsigned char* flag, hipStream_t s,
                            unsigned long long* moved)
{
    if (n <= 0) return;
    commit_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(list, v0, n, C, Cn, C32, k, a, gid, flag, nullptr,
                                                         moved);
}

}  // namespace nido
