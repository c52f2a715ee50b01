/*
 * louvain_oracle.c -- CPU restatement of nido's batched Louvain path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it,
 * and only as the checker.  The product (libnido_amd.so) never links,
 * loads or calls anything here.
 *
 * It restates, in plain C with OpenMP, the semantics of the reference's
 * GPU path (file:line citations are into the sg0/nido sources):
 *
 *   driver loop ............ louvain_gpu.cpp:28-124
 *   -b batch boundaries .... graph_gpu.cpp:280-359 (part_on_batch=1)
 *   vertex weight k_v ...... graph_cuda.cu:302-346
 *   mass m = sum(k)/2 ...... graph_cuda.cu:1786-1856 (0.5*m at 1855)
 *   singleton partition .... graph_cuda.cu:410-429
 *   neighbour-community
 *     aggregation .......... graph_gpu.cpp:457-517, graph_cuda.cu:49-98,
 *                            self arcs weighted 0: graph_cuda.cu:2289-2318
 *   gain / move rule ....... graph_cuda.cu:1450-1540
 *   commit (a update, swap)  graph_cuda.cu:1182-1251, graph_gpu.cpp:762-820
 *   rollback ............... graph_gpu.cpp:822-842
 *   modularity ............. graph_cuda.cu:2078-2134
 *   coarsening ............. graph_gpu.cpp:1088-1304, 1382-1545
 *   clustering map ......... clustering.cpp:12-21
 *   file format ............ graph.cpp:322-495, types.hpp:48-55
 *
 * Canonical arithmetic (deliberate, documented deviations from the
 * reference, which is not reproducible against itself):
 *   - every sum of input weights (k_v, e_{v,c}, a_c, 2m, the internal
 *     weight of the modularity, the coarse arc weights) is an exact
 *     two's-complement Q64.64 integer; each FP64 weight is converted to
 *     Q64.64 with round-to-nearest-even at 2^-64 (exact for |w| >= 2^-12
 *     and for any weight on a 2^-64 grid) and each sum is rounded to FP64
 *     once, round-to-nearest-even, when it is read;
 *   - sum_c a_c^2/(4m^2) is the exact Q64.64 sum of the FP64 squares
 *     x_c*x_c, x_c = a_c/(2m), each square rounded to Q64.64;
 *   - gain ties go to the smallest community id;
 *   - coarse rows are sorted by neighbour id.
 * Gains are evaluated in FP64 in the reference's operation order
 * (graph_cuda.cu:1505-1506, 1530-1531); build with -ffp-contract=off.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <omp.h>

typedef __int128 fx_t;
typedef unsigned __int128 ufx_t;

#define ORC_MAX_PHASES 200            /* types.hpp:24-26 */
#define ORC_MAX_FLOAT  1.0e308        /* cuda_wrapper.hpp:112 */

/* ------------------------------------------------------------------ */
/* Q64.64 <-> FP64                                                    */
/* ------------------------------------------------------------------ */

/* Round-to-nearest-even conversion of a finite double to Q64.64.
 * Returns 0 and sets *err for |w| >= 2^62 or a non-finite value. */
fx_t orc_fx_from_double_e(double w, int* err)
{
    uint64_t bits;
    memcpy(&bits, &w, 8);
    int neg = (int)(bits >> 63);
    int e = (int)((bits >> 52) & 0x7ff);
    uint64_t mant = bits & ((1ull << 52) - 1);
    if (e == 0x7ff) { if (err) *err = 1; return 0; }
    if (e == 0) { if (mant == 0) return 0; e = 1; }
    else mant |= (1ull << 52);
    /* w = mant * 2^(e-1075); w * 2^64 = mant * 2^(e-1011) */
    int sh = e - 1011;
    ufx_t u;
    if (sh >= 0) {
        if (sh > 73) { if (err) *err = 1; return 0; }
        u = ((ufx_t)mant) << sh;
    } else {
        int r = -sh;
        if (r >= 54) return 0;                   /* below half an ulp of 2^-64 */
        uint64_t q = mant >> r;
        uint64_t rem = mant & ((1ull << r) - 1);
        uint64_t half = 1ull << (r - 1);
        if (rem > half || (rem == half && (q & 1))) q++;
        u = q;
    }
    return neg ? -(fx_t)u : (fx_t)u;
}

fx_t orc_fx_from_double(double w) { return orc_fx_from_double_e(w, NULL); }

static inline int orc_clz128(ufx_t u)
{
    uint64_t hi = (uint64_t)(u >> 64), lo = (uint64_t)u;
    if (hi) return __builtin_clzll(hi);
    if (lo) return 64 + __builtin_clzll(lo);
    return 128;
}

static inline double orc_pow2(int e)   /* exact 2^e for -1022 <= e <= 1023 */
{
    uint64_t b = ((uint64_t)(1023 + e)) << 52;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

/* Round-to-nearest-even conversion of a Q64.64 value to FP64. */
double orc_fx_to_double(fx_t x)
{
    if (x == 0) return 0.0;
    int neg = x < 0;
    ufx_t u = neg ? (ufx_t)(-x) : (ufx_t)x;
    int msb = 127 - orc_clz128(u);
    double r;
    if (msb <= 52) {
        r = (double)(uint64_t)u * orc_pow2(-64);
    } else {
        int shift = msb - 52;
        uint64_t mant = (uint64_t)(u >> shift);
        ufx_t rem = u & ((((ufx_t)1) << shift) - 1);
        ufx_t half = ((ufx_t)1) << (shift - 1);
        if (rem > half || (rem == half && (mant & 1))) mant++;
        if (mant == (1ull << 53)) { mant >>= 1; shift++; }
        r = (double)mant * orc_pow2(shift - 64);
    }
    return neg ? -r : r;
}

/* ctypes helpers: Q64.64 passed as (lo, hi) words */
void orc_fx_from_double_words(double w, uint64_t* lo, int64_t* hi, int* err)
{
    int e = 0;
    fx_t x = orc_fx_from_double_e(w, &e);
    *lo = (uint64_t)x;
    *hi = (int64_t)(x >> 64);
    if (err) *err = e;
}

double orc_fx_words_to_double(uint64_t lo, int64_t hi)
{
    fx_t x = (fx_t)(((ufx_t)(uint64_t)hi << 64) | lo);
    return orc_fx_to_double(x);
}

/* ------------------------------------------------------------------ */
/* Batch boundaries: graph_gpu.cpp:303-337 (V0 = 0, one device)       */
/* ------------------------------------------------------------------ */
void orc_batch_bounds(int64_t nv, const int64_t* off, int nb, int64_t* parts)
{
    int64_t ne = off[nv] - off[0];
    int64_t ave = ne / nb;
    int64_t cap = (int64_t)nb + 1 + (nv > 0 ? nv : 0) + 1;
    int64_t* vp = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    int64_t n = 0, start = 0;
    vp[n++] = 0;
    for (int64_t i = 1; i <= nv; ++i) {
        if (off[i] - off[start] > ave) { vp[n++] = i; start = i; }
        else if (i == nv) vp[n++] = nv;
    }
    if (n > nb + 1) {
        int64_t remain = nv - vp[nb];
        int64_t nnv = remain / nb;
        for (int i = 1; i < nb; ++i) vp[i] += nnv;
        vp[nb] = nv;
    } else {
        while (n < nb + 1) vp[n++] = nv;
    }
    for (int i = 0; i <= nb; ++i) parts[i] = vp[i];
    free(vp);
}

/* ------------------------------------------------------------------ */
/* Open-addressing table (community id -> Q64.64), one per thread     */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t  cap;       /* power of two */
    int64_t* key;       /* -1 = empty */
    fx_t*    val;
    int64_t* used;      /* slots touched, for O(touched) clearing */
    int64_t  nused;
} orc_table;

static void tbl_init(orc_table* t, int64_t need)
{
    int64_t cap = 16;
    while (cap < 2 * need) cap <<= 1;
    t->cap = cap;
    t->key = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    t->val = (fx_t*)malloc(sizeof(fx_t) * (size_t)cap);
    t->used = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    for (int64_t i = 0; i < cap; ++i) t->key[i] = -1;
    t->nused = 0;
}

static void tbl_free(orc_table* t)
{
    free(t->key); free(t->val); free(t->used);
    memset(t, 0, sizeof(*t));
}

static void tbl_clear(orc_table* t)
{
    for (int64_t i = 0; i < t->nused; ++i) t->key[t->used[i]] = -1;
    t->nused = 0;
}

static inline uint64_t orc_hash(int64_t k)
{
    uint64_t x = (uint64_t)k * 0x9E3779B97F4A7C15ull;
    return x ^ (x >> 29);
}

static inline void tbl_add(orc_table* t, int64_t k, fx_t v)
{
    uint64_t m = (uint64_t)t->cap - 1, s = orc_hash(k) & m;
    for (;;) {
        if (t->key[s] == k) { t->val[s] += v; return; }
        if (t->key[s] < 0) {
            t->key[s] = k; t->val[s] = v; t->used[t->nused++] = (int64_t)s;
            return;
        }
        s = (s + 1) & m;
    }
}

static inline int tbl_get(const orc_table* t, int64_t k, fx_t* v)
{
    uint64_t m = (uint64_t)t->cap - 1, s = orc_hash(k) & m;
    for (;;) {
        if (t->key[s] == k) { *v = t->val[s]; return 1; }
        if (t->key[s] < 0) return 0;
        s = (s + 1) & m;
    }
}

/* ------------------------------------------------------------------ */
/* State                                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t nv, ne, maxdeg;
    int64_t* off; int64_t* col; double* w;
    fx_t* wf;                 /* Q64.64 copy of w */
    fx_t* k; double* kd;      /* vertex weights (graph_cuda.cu:332-343) */
    fx_t twom;  double m;     /* 2m exact, m as FP64 */
    int64_t* C; int64_t* Cn;  /* commIds_, newCommIds_ */
    fx_t* a;                  /* commWeights_ */
    int nb; int64_t* parts;   /* vertex_per_batch_ */
} orc_phase;

typedef struct {
    int64_t nv, ne, loops;
    double q0, q;
    double* iter_q;          /* Qtmp per iteration */
    int64_t* comm;           /* C at the end of the phase (phase-local ids) */
    double t_loop, t_agg;
} orc_hist;

typedef struct {
    orc_phase ph;
    int64_t n0;
    int64_t* clustering;     /* Clustering::commIds_ (clustering.cpp:12-21) */
    int nthreads;
    int err;                 /* first input error (non-finite / huge weight) */
    orc_table* tbl;          /* one aggregation table per OpenMP thread */
    int ntbl;
    int64_t tbl_need;
    /* run history */
    int nphases;
    orc_hist hist[ORC_MAX_PHASES];
    double t_total, t_loop, t_agg;
    int64_t total_loops;
} orc_ctx;

static void phase_free(orc_phase* p)
{
    free(p->off); free(p->col); free(p->w); free(p->wf);
    free(p->k); free(p->kd); free(p->C); free(p->Cn); free(p->a); free(p->parts);
    memset(p, 0, sizeof(*p));
}

static void hist_free(orc_ctx* c)
{
    for (int i = 0; i < c->nphases; ++i) { free(c->hist[i].iter_q); free(c->hist[i].comm); }
    memset(c->hist, 0, sizeof(c->hist));
    c->nphases = 0;
}

/* k_v = sum of the row's weights (self arcs included), 2m = sum_v k_v,
 * batch boundaries; GraphGPU ctor, graph_gpu.cpp:94-115 and :139. */
static void phase_setup(orc_ctx* c)
{
    orc_phase* p = &c->ph;
    int64_t nv = p->nv;
    p->k = (fx_t*)malloc(sizeof(fx_t) * (size_t)(nv > 0 ? nv : 1));
    p->kd = (double*)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    p->C = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    p->Cn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    p->a = (fx_t*)malloc(sizeof(fx_t) * (size_t)(nv > 0 ? nv : 1));
    p->wf = (fx_t*)malloc(sizeof(fx_t) * (size_t)(p->ne > 0 ? p->ne : 1));
    int err = 0;
    #pragma omp parallel for schedule(static) reduction(|:err)
    for (int64_t e = 0; e < p->ne; ++e) {
        int e1 = 0;
        p->wf[e] = orc_fx_from_double_e(p->w[e], &e1);
        err |= e1;
    }
    if (err && !c->err) c->err = 1;
    /* Every exact sum (k_v, e_vc, a_c, 2m, coarse weights) is bounded by the
     * total of |w|, which must stay below 2^62; with negative weights
     * |a_c / 2m| <= sum|w| / |2m| must stay below 2^31 so that each square
     * term of the modularity stays below 2^62.  Serial and saturating. */
    if (!err && p->ne > 0) {
        const ufx_t cap = (ufx_t)1 << 126;
        ufx_t sabs = 0, ssum = 0;
        int neg = 0;
        for (int64_t e = 0; e < p->ne && sabs < cap; ++e) {
            fx_t f = p->wf[e];
            ufx_t af = f < 0 ? (ufx_t)(-f) : (ufx_t)f;
            neg |= f < 0;
            sabs = af >= cap - sabs ? cap : sabs + af;
            ssum += (ufx_t)f;
        }
        fx_t t = (fx_t)ssum;
        ufx_t tabs = t < 0 ? (ufx_t)(-t) : (ufx_t)t;
        if (sabs >= cap && !c->err) c->err = 3;
        else if (neg && tabs != 0 && (sabs >> 31) >= tabs && !c->err) c->err = 4;
    }
    int64_t maxdeg = 0;
    #pragma omp parallel for schedule(dynamic, 1024) reduction(max:maxdeg)
    for (int64_t v = 0; v < nv; ++v) {
        fx_t s = 0;
        for (int64_t e = p->off[v]; e < p->off[v + 1]; ++e) s += p->wf[e];
        p->k[v] = s;
        p->kd[v] = orc_fx_to_double(s);
        int64_t d = p->off[v + 1] - p->off[v];
        if (d > maxdeg) maxdeg = d;
    }
    p->maxdeg = maxdeg;
    /* exact integer sum: any order gives the same bits */
    fx_t twom = 0;
    for (int64_t v = 0; v < nv; ++v) twom += p->k[v];
    p->twom = twom;
    p->m = orc_fx_to_double(twom) * 0.5;
    p->parts = (int64_t*)malloc(sizeof(int64_t) * (size_t)(p->nb + 1));
    orc_batch_bounds(nv, p->off, p->nb, p->parts);
}

/* ------------------------------------------------------------------ */
/* Public API (mirrors GraphGPU's public methods, graph_gpu.hpp:168-247) */
/* ------------------------------------------------------------------ */
orc_ctx* orc_create(int64_t nv, int64_t ne, const int64_t* off, const int64_t* col,
                    const double* w, int nbatches, int nthreads)
{
    if (nv < 0 || nbatches < 1) return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    omp_set_num_threads(c->nthreads);
    orc_phase* p = &c->ph;
    p->nv = nv;
    /* offsets are rebased to start at 0 (graph.cpp:470-472); the arc count
     * comes from the offsets (graph.cpp:376-382) */
    int64_t base = nv >= 0 ? off[0] : 0;
    p->ne = off[nv] - base;
    (void)ne;
    p->off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv + 1));
    for (int64_t i = 0; i <= nv; ++i) p->off[i] = off[i] - base;
    p->col = (int64_t*)malloc(sizeof(int64_t) * (size_t)(p->ne > 0 ? p->ne : 1));
    p->w = (double*)malloc(sizeof(double) * (size_t)(p->ne > 0 ? p->ne : 1));
    if (p->ne > 0) {
        memcpy(p->col, col, sizeof(int64_t) * (size_t)p->ne);
        memcpy(p->w, w, sizeof(double) * (size_t)p->ne);
    }
    p->nb = nbatches;
    for (int64_t e = 0; e < p->ne; ++e)
        if (p->col[e] < 0 || p->col[e] >= nv) { c->err = 2; break; }
    phase_setup(c);
    c->n0 = nv;
    c->clustering = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    for (int64_t i = 0; i < nv; ++i) c->clustering[i] = i;
    return c;
}

static void tables_free(orc_ctx* c)
{
    for (int i = 0; i < c->ntbl; ++i) tbl_free(&c->tbl[i]);
    free(c->tbl);
    c->tbl = NULL; c->ntbl = 0; c->tbl_need = 0;
}

static void tables_ensure(orc_ctx* c, int64_t need)
{
    if (c->tbl && c->tbl_need >= need && c->ntbl >= c->nthreads) return;
    tables_free(c);
    c->ntbl = c->nthreads;
    c->tbl = (orc_table*)calloc((size_t)c->ntbl, sizeof(orc_table));
    for (int i = 0; i < c->ntbl; ++i) tbl_init(&c->tbl[i], need);
    c->tbl_need = need;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    tables_free(c);
    phase_free(&c->ph);
    hist_free(c);
    free(c->clustering);
    free(c);
}

int orc_error(const orc_ctx* c) { return c->err; }
int64_t orc_nv(const orc_ctx* c) { return c->ph.nv; }
int64_t orc_ne(const orc_ctx* c) { return c->ph.ne; }
int64_t orc_max_order(const orc_ctx* c) { return c->ph.maxdeg; }
double orc_mass(const orc_ctx* c) { return c->ph.m; }
void orc_twom_words(const orc_ctx* c, uint64_t* lo, int64_t* hi)
{
    *lo = (uint64_t)c->ph.twom;
    *hi = (int64_t)(c->ph.twom >> 64);
}

void orc_get_batch_bounds(const orc_ctx* c, int64_t* parts)
{
    memcpy(parts, c->ph.parts, sizeof(int64_t) * (size_t)(c->ph.nb + 1));
}

void orc_get_graph(const orc_ctx* c, int64_t* off, int64_t* col, double* w)
{
    const orc_phase* p = &c->ph;
    if (off) memcpy(off, p->off, sizeof(int64_t) * (size_t)(p->nv + 1));
    if (col && p->ne) memcpy(col, p->col, sizeof(int64_t) * (size_t)p->ne);
    if (w && p->ne) memcpy(w, p->w, sizeof(double) * (size_t)p->ne);
}

void orc_get_communities(const orc_ctx* c, int64_t* out)
{
    memcpy(out, c->ph.C, sizeof(int64_t) * (size_t)c->ph.nv);
}

void orc_get_new_communities(const orc_ctx* c, int64_t* out)
{
    memcpy(out, c->ph.Cn, sizeof(int64_t) * (size_t)c->ph.nv);
}

void orc_set_communities(orc_ctx* c, const int64_t* C, const int64_t* Cn)
{
    /* test hook: install a partition and rebuild a_c = sum_{C[v]=c} k_v
     * (the CHECK-build helper, graph_gpu.cpp:433-455) */
    orc_phase* p = &c->ph;
    memcpy(p->C, C, sizeof(int64_t) * (size_t)p->nv);
    memcpy(p->Cn, Cn ? Cn : C, sizeof(int64_t) * (size_t)p->nv);
    for (int64_t v = 0; v < p->nv; ++v) p->a[v] = 0;
    for (int64_t v = 0; v < p->nv; ++v) p->a[p->C[v]] += p->k[v];
}

void orc_get_vertex_weights(const orc_ctx* c, double* out)
{
    memcpy(out, c->ph.kd, sizeof(double) * (size_t)c->ph.nv);
}

void orc_get_comm_weights(const orc_ctx* c, double* out)
{
    for (int64_t i = 0; i < c->ph.nv; ++i) out[i] = orc_fx_to_double(c->ph.a[i]);
}

void orc_get_comm_weights_words(const orc_ctx* c, uint64_t* lo, int64_t* hi)
{
    for (int64_t i = 0; i < c->ph.nv; ++i) {
        lo[i] = (uint64_t)c->ph.a[i];
        hi[i] = (int64_t)(c->ph.a[i] >> 64);
    }
}

void orc_get_clustering(const orc_ctx* c, int64_t* out)
{
    memcpy(out, c->clustering, sizeof(int64_t) * (size_t)c->n0);
}

/* singleton_partition_kernel, graph_cuda.cu:422-428 */
void orc_singleton_partition(orc_ctx* c)
{
    orc_phase* p = &c->ph;
    #pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < p->nv; ++v) {
        p->C[v] = v; p->Cn[v] = v; p->a[v] = p->k[v];
    }
}

/* Q = I/(2m) - sum_c (a_c/(2m))^2, graph_cuda.cu:2103-2132 (self arcs are
 * internal arcs).  I and the sum of squares are exact Q64.64 sums. */
double orc_compute_modularity(orc_ctx* c)
{
    orc_phase* p = &c->ph;
    double twom_d = orc_fx_to_double(p->twom);
    if (p->twom == 0) return 0.0;
    fx_t I = 0, T = 0;
    #pragma omp parallel
    {
        fx_t li = 0, lt = 0;
        #pragma omp for schedule(dynamic, 4096) nowait
        for (int64_t v = 0; v < p->nv; ++v) {
            int64_t cv = p->C[v];
            for (int64_t e = p->off[v]; e < p->off[v + 1]; ++e)
                if (p->C[p->col[e]] == cv) li += p->wf[e];
            double x = orc_fx_to_double(p->a[v]) / twom_d;
            lt += orc_fx_from_double(x * x);
        }
        #pragma omp critical
        { I += li; T += lt; }
    }
    return orc_fx_to_double(I) / twom_d - orc_fx_to_double(T);
}

/* One vertex's move decision; graph_cuda.cu:1481-1537 with the min-id tie
 * break.  Returns the chosen community. */
static int64_t orc_decide(const orc_phase* p, int64_t v, orc_table* t)
{
    int64_t own = p->C[v];
    tbl_clear(t);
    for (int64_t e = p->off[v]; e < p->off[v + 1]; ++e) {
        int64_t u = p->col[e];
        if (u == v) continue;                       /* graph_cuda.cu:2312-2315 */
        tbl_add(t, p->C[u], p->wf[e]);
    }
    double m = p->m;
    double ki = p->kd[v];
    int64_t best_c = -1;
    double best_g = -ORC_MAX_FLOAT;
    for (int64_t i = 0; i < t->nused; ++i) {
        int64_t s = t->used[i];
        int64_t cc = t->key[s];
        if (cc == own) continue;
        double g = orc_fx_to_double(t->val[s]);
        double ac = orc_fx_to_double(p->a[cc]);
        g -= ki * ac / (2. * m);                   /* graph_cuda.cu:1505 */
        g /= m;                                     /* graph_cuda.cu:1506 */
        if (best_c < 0 || g > best_g || (g == best_g && cc < best_c)) {
            best_g = g; best_c = cc;
        }
    }
    if (best_c < 0) return own;                     /* gain stays -MAX_FLOAT */
    fx_t sv = 0;
    tbl_get(t, own, &sv);
    double self = orc_fx_to_double(sv);
    double ar = orc_fx_to_double(p->a[own] - p->k[v]);
    self -= ki * ar / (2. * m);                     /* graph_cuda.cu:1530 */
    self /= m;                                      /* graph_cuda.cu:1531 */
    double gain = best_g;
    gain -= self;                                   /* graph_cuda.cu:1532 */
    return (gain > 0) ? best_c : own;
}

/* GraphGPU::louvain_update(batch), graph_gpu.cpp:602-624: writes C' for the
 * batch's vertices from one snapshot of C and a. */
void orc_louvain_update(orc_ctx* c, int b)
{
    orc_phase* p = &c->ph;
    int64_t v0 = p->parts[b], v1 = p->parts[b + 1];
    if (v1 <= v0) return;
    if (p->twom == 0) {
        for (int64_t v = v0; v < v1; ++v) p->Cn[v] = p->C[v];
        return;
    }
    tables_ensure(c, p->maxdeg + 1);
    #pragma omp parallel num_threads(c->ntbl)
    {
        orc_table* t = &c->tbl[omp_get_thread_num()];
        #pragma omp for schedule(dynamic, 256)
        for (int64_t v = v0; v < v1; ++v) p->Cn[v] = orc_decide(p, v, t);
    }
}

/* update_community_weights_kernel, graph_cuda.cu:1234-1249 */
static void orc_move_weights(orc_phase* p, int64_t v0, int64_t v1)
{
    /* exact integer updates commute: any order gives the same bits */
    for (int64_t v = v0; v < v1; ++v) {
        int64_t s = p->C[v], d = p->Cn[v];
        if (s != d) { p->a[s] -= p->k[v]; p->a[d] += p->k[v]; }
    }
}

void orc_update_community_weights(orc_ctx* c, int b)
{
    orc_move_weights(&c->ph, c->ph.parts[b], c->ph.parts[b + 1]);
}

/* exchange_vector_kernel, graph_cuda.cu:1191-1196 */
static void orc_swap(orc_phase* p, int64_t v0, int64_t v1)
{
    for (int64_t v = v0; v < v1; ++v) {
        int64_t t = p->C[v]; p->C[v] = p->Cn[v]; p->Cn[v] = t;
    }
}

void orc_update_community_ids(orc_ctx* c, int b)
{
    orc_swap(&c->ph, c->ph.parts[b], c->ph.parts[b + 1]);
}

/* graph_gpu.cpp:822-842 */
void orc_restore_community(orc_ctx* c)
{
    orc_move_weights(&c->ph, 0, c->ph.nv);
    orc_swap(&c->ph, 0, c->ph.nv);
}

static int cmp_i64_pair(const void* x, const void* y)
{
    const int64_t* a = (const int64_t*)x;
    const int64_t* b = (const int64_t*)y;
    return (a[0] > b[0]) - (a[0] < b[0]);
}

/* GraphGPU::aggregation, graph_gpu.cpp:1382-1545.  Returns 1 when the run
 * is done (K == N or K == 1), else builds the coarse graph and returns 0. */
int orc_aggregation(orc_ctx* c)
{
    orc_phase* p = &c->ph;
    int64_t nv = p->nv;
    /* dense renumbering, ascending by old id (graph_gpu.cpp:1178-1218) */
    int64_t* newid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv + 1));
    for (int64_t i = 0; i < nv; ++i) newid[i] = 0;
    for (int64_t v = 0; v < nv; ++v) newid[p->C[v]] = 1;
    int64_t K = 0;
    for (int64_t i = 0; i < nv; ++i) { int64_t f = newid[i]; newid[i] = K; K += f; }
    /* cumulative map, before the termination test (graph_gpu.cpp:1389) */
    for (int64_t i = 0; i < c->n0; ++i) c->clustering[i] = newid[p->C[c->clustering[i]]];
    if (K == nv || K == 1) { free(newid); return 1; }

    /* members of every coarse vertex, ascending fine id */
    int64_t* moff = (int64_t*)calloc((size_t)(K + 1), sizeof(int64_t));
    int64_t* mem = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    for (int64_t v = 0; v < nv; ++v) moff[newid[p->C[v]] + 1]++;
    for (int64_t i = 0; i < K; ++i) moff[i + 1] += moff[i];
    {
        int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)K);
        memcpy(pos, moff, sizeof(int64_t) * (size_t)K);
        for (int64_t v = 0; v < nv; ++v) mem[pos[newid[p->C[v]]]++] = v;
        free(pos);
    }
    /* merged rows: arcs (u->x,w) become (new(u)->new(x)), summed exactly */
    int64_t* cnt = (int64_t*)calloc((size_t)(K + 1), sizeof(int64_t));
    int64_t* rowcap = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
    rowcap[0] = 0;
    for (int64_t i = 0; i < K; ++i) {
        int64_t s = 0;
        for (int64_t j = moff[i]; j < moff[i + 1]; ++j) {
            int64_t v = mem[j];
            s += p->off[v + 1] - p->off[v];
        }
        rowcap[i + 1] = rowcap[i] + s;
    }
    int64_t* tcol = (int64_t*)malloc(sizeof(int64_t) * (size_t)(rowcap[K] > 0 ? rowcap[K] : 1));
    fx_t* tw = (fx_t*)malloc(sizeof(fx_t) * (size_t)(rowcap[K] > 0 ? rowcap[K] : 1));
    int64_t maxrow = 0;
    for (int64_t i = 0; i < K; ++i)
        if (rowcap[i + 1] - rowcap[i] > maxrow) maxrow = rowcap[i + 1] - rowcap[i];
    #pragma omp parallel
    {
        orc_table t;
        tbl_init(&t, maxrow + 1);
        int64_t* pairs = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)(maxrow + 1));
        #pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < K; ++i) {
            tbl_clear(&t);
            for (int64_t j = moff[i]; j < moff[i + 1]; ++j) {
                int64_t v = mem[j];
                for (int64_t e = p->off[v]; e < p->off[v + 1]; ++e)
                    tbl_add(&t, newid[p->C[p->col[e]]], p->wf[e]);
            }
            int64_t n = t.nused;
            for (int64_t x = 0; x < n; ++x) {
                pairs[2 * x] = t.key[t.used[x]];
                pairs[2 * x + 1] = t.used[x];
            }
            qsort(pairs, (size_t)n, 2 * sizeof(int64_t), cmp_i64_pair);
            for (int64_t x = 0; x < n; ++x) {
                tcol[rowcap[i] + x] = pairs[2 * x];
                tw[rowcap[i] + x] = t.val[pairs[2 * x + 1]];
            }
            cnt[i + 1] = n;
        }
        free(pairs);
        tbl_free(&t);
    }
    int64_t* noff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K + 1));
    noff[0] = 0;
    for (int64_t i = 0; i < K; ++i) noff[i + 1] = noff[i] + cnt[i + 1];
    int64_t nne = noff[K];
    int64_t* ncol = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nne > 0 ? nne : 1));
    double* nw = (double*)malloc(sizeof(double) * (size_t)(nne > 0 ? nne : 1));
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < K; ++i)
        for (int64_t x = 0; x < noff[i + 1] - noff[i]; ++x) {
            ncol[noff[i] + x] = tcol[rowcap[i] + x];
            nw[noff[i] + x] = orc_fx_to_double(tw[rowcap[i] + x]);
        }
    free(tcol); free(tw); free(cnt); free(rowcap); free(moff); free(mem); free(newid);

    int nb = p->nb;
    phase_free(p);
    p->nv = K; p->ne = nne; p->off = noff; p->col = ncol; p->w = nw; p->nb = nb;
    phase_setup(c);
    return 0;
}

/* The loop body of LouvainGPU::run, louvain_gpu.cpp:55-90: all batches,
 * modularity, accept or roll back.  Returns the evaluated modularity. */
double orc_louvain_iteration(orc_ctx* c, double* q, double* dq)
{
    for (int b = 0; b < c->ph.nb; ++b) {
        orc_louvain_update(c, b);
        orc_update_community_weights(c, b);
        orc_update_community_ids(c, b);
    }
    double qt = orc_compute_modularity(c);
    *dq = qt - *q;
    if (*dq < 0) orc_restore_community(c);
    else *q = qt;
    return qt;
}

/* LouvainGPU::run, louvain_gpu.cpp:6-139.  Records per-phase history. */
void orc_run(orc_ctx* c, int64_t max_loops, double tol, double tol_phase, int record)
{
    hist_free(c);
    omp_set_num_threads(c->nthreads);
    int done = 0;
    int64_t nphase = 0;
    double Q_old = 0.;
    double t0 = omp_get_wtime();
    c->t_loop = 0; c->t_agg = 0; c->total_loops = 0;
    while (!done && nphase < ORC_MAX_PHASES) {
        orc_hist* h = &c->hist[nphase];
        h->nv = c->ph.nv; h->ne = c->ph.ne;
        orc_singleton_partition(c);
        double Q = orc_compute_modularity(c);
        Q_old = Q;
        h->q0 = Q;
        double dQ = ORC_MAX_FLOAT;
        int64_t loops = 0;
        int64_t cap = 16;
        h->iter_q = (double*)malloc(sizeof(double) * (size_t)cap);
        double ts = omp_get_wtime();
        while (tol < dQ && loops < max_loops) {
            for (int b = 0; b < c->ph.nb; ++b) {
                orc_louvain_update(c, b);
                orc_update_community_weights(c, b);
                orc_update_community_ids(c, b);
            }
            double Qtmp = orc_compute_modularity(c);
            dQ = Qtmp - Q;
            if (loops >= cap) { cap *= 2; h->iter_q = (double*)realloc(h->iter_q, sizeof(double) * (size_t)cap); }
            h->iter_q[loops] = Qtmp;
            loops++;
            if (dQ < 0) orc_restore_community(c);
            else Q = Qtmp;
        }
        double te = omp_get_wtime();
        h->loops = loops; h->q = Q; h->t_loop = te - ts;
        c->t_loop += te - ts;
        c->total_loops += loops;
        if (record) {
            h->comm = (int64_t*)malloc(sizeof(int64_t) * (size_t)(c->ph.nv > 0 ? c->ph.nv : 1));
            memcpy(h->comm, c->ph.C, sizeof(int64_t) * (size_t)c->ph.nv);
        }
        c->nphases = (int)(nphase + 1);
        if (Q - Q_old > tol_phase) {
            double a0 = omp_get_wtime();
            done = orc_aggregation(c);
            h->t_agg = omp_get_wtime() - a0;
            c->t_agg += h->t_agg;
        } else {
            done = 1;
        }
        nphase++;
    }
    c->t_total = omp_get_wtime() - t0;
}

int orc_hist_phases(const orc_ctx* c) { return c->nphases; }

void orc_hist_phase(const orc_ctx* c, int i, int64_t* nv, int64_t* ne, int64_t* loops,
                    double* q0, double* q, double* t_loop, double* t_agg)
{
    const orc_hist* h = &c->hist[i];
    *nv = h->nv; *ne = h->ne; *loops = h->loops; *q0 = h->q0; *q = h->q;
    *t_loop = h->t_loop; *t_agg = h->t_agg;
}

void orc_hist_iter_q(const orc_ctx* c, int i, double* out)
{
    memcpy(out, c->hist[i].iter_q, sizeof(double) * (size_t)c->hist[i].loops);
}

int orc_hist_comm(const orc_ctx* c, int i, int64_t* out)
{
    if (!c->hist[i].comm) return -1;
    memcpy(out, c->hist[i].comm, sizeof(int64_t) * (size_t)c->hist[i].nv);
    return 0;
}

void orc_times(const orc_ctx* c, double* total, double* loop, double* agg, int64_t* loops)
{
    *total = c->t_total; *loop = c->t_loop; *agg = c->t_agg; *loops = c->total_loops;
}

int orc_threads(const orc_ctx* c) { return c->nthreads; }

/* ------------------------------------------------------------------ */
/* Colouring + renumbering (the reference's -c), graph.cpp:547-694    */
/* ------------------------------------------------------------------ */

/* graph.cpp:573-605 as one thread runs it: vertices in id order; vertex i
 * collects the colours of its neighbours j that are already coloured (so
 * j < i) with r[j] >= r[i] (graph.cpp:583-592) and takes the smallest colour
 * not among them (graph.cpp:594-601).  Returns the colour count. */
int64_t orc_coloring(int64_t nv, const int64_t* off, const int64_t* col, const int64_t* r, int64_t* color)
{
    int64_t maxc = -1;
    unsigned char* seen = (unsigned char*)calloc((size_t)nv + 1, 1);
    for (int64_t i = 0; i < nv; ++i) color[i] = -1;
    for (int64_t i = 0; i < nv; ++i) {
        const int64_t d = off[i + 1] - off[i];
        /* the smallest free colour is at most d: mark colours 0..d only */
        for (int64_t k = off[i]; k < off[i + 1]; ++k) {
            const int64_t j = col[k];
            const int64_t jc = color[j];
            if (jc != -1 && r[i] <= r[j] && jc <= d) seen[jc] = 1;
        }
        int64_t c = 0;
        while (seen[c]) ++c;
        color[i] = c;
        if (c > maxc) maxc = c;
        for (int64_t k = off[i]; k < off[i + 1]; ++k) {
            const int64_t jc = color[col[k]];
            if (jc >= 0 && jc <= d) seen[jc] = 0;
        }
        seen[c] = 0;
    }
    free(seen);
    return maxc + 1;
}

/* graph.cpp:607-691: (colour, id) order; new_orders[pos] = old id,
 * orders[old id] = pos; the CSR rebuilt with row pos = old row
 * new_orders[pos] (arc order kept) and every column mapped through orders.
 * noff[nv+1], ncol[ne], nw[ne] out. */
void orc_reorder(int64_t nv, const int64_t* off, const int64_t* col, const double* w, const int64_t* color,
                 int64_t* new_orders, int64_t* noff, int64_t* ncol, double* nw)
{
    int64_t* orders = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nv > 0 ? nv : 1));
    int64_t pos = 0;
    /* a stable pass per colour, ids ascending (the sort of graph.cpp:614) */
    int64_t maxc = -1;
    for (int64_t i = 0; i < nv; ++i)
        if (color[i] > maxc) maxc = color[i];
    for (int64_t c = 0; c <= maxc; ++c)
        for (int64_t i = 0; i < nv; ++i)
            if (color[i] == c) {
                orders[i] = pos;
                new_orders[pos++] = i;
            }
    noff[0] = 0;
    for (int64_t p = 0; p < nv; ++p) {
        const int64_t v = new_orders[p];
        noff[p + 1] = noff[p] + (off[v + 1] - off[v]);
    }
    for (int64_t p = 0; p < nv; ++p) {
        const int64_t v = new_orders[p];
        for (int64_t k = 0; k < off[v + 1] - off[v]; ++k) {
            ncol[noff[p] + k] = orders[col[off[v] + k]];
            nw[noff[p] + k] = w[off[v] + k];
        }
    }
    free(orders);
}
