// graph_gpu.hip -- device-resident Louvain state and the C ABI over it.
//
// Mirrors the reference's GraphGPU (graph_gpu.hpp:168-247) and
// LouvainGPU::run (louvain_gpu.cpp:6-139) on one MI355X: the whole graph
// stays in HBM, every step is a hand-written kernel on the handle's stream,
// and the host only reads back one scalar per modularity evaluation.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/nido_amd.h"
#include "errors.h"
#include "internal.h"

using namespace nido;

namespace {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(NIDO_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                             __FILE__, __LINE__);                                              \
    } while (0)

#define NIDO_TRY(expr)                  \
    do {                                \
        int r_ = (expr);                \
        if (r_ != NIDO_OK) return r_;   \
    } while (0)

constexpr int64_t kSmallRow = 32;      // rows handled by one lane in the degree/modularity sweeps
constexpr int64_t kItemArcs = 1024;    // longer rows are cut into slices of this many arcs

template <typename T>
int dev_alloc(T** p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(NIDO_ERR_NOMEM, "hipMalloc of %lld bytes failed: %s",
                         (long long)(sizeof(T) * (size_t)count), hipGetErrorString(e));
    }
    return NIDO_OK;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

struct PhaseHist {
    nido_phase_stats st{};
    std::vector<double> iter_q;
    std::vector<int64_t> comm;
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct nido_graph {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    int nb = NIDO_DEFAULT_BATCHES;
    nido_tuning tuning{};

    int64_t n0 = 0;        // original vertex count
    int64_t nv = 0, ne = 0, maxdeg = 0;
    std::vector<int64_t> h_off;   // current offsets (host copy)
    std::vector<int64_t> parts;   // batch boundaries, nb+1

    // current CSR
    int64_t* d_off = nullptr;
    int64_t* d_col = nullptr;
    double* d_w = nullptr;
    // state, capacity n0
    int64_t* d_C = nullptr;
    int64_t* d_Cn = nullptr;
    fx_t* d_k = nullptr;
    fx_t* d_a = nullptr;
    double* d_kd = nullptr;
    int64_t* d_clus = nullptr;
    int64_t* d_work = nullptr;     // n0+1 int64 scratch (renumbering)
    void* d_scan_tmp = nullptr;
    // scalars
    fx_t* d_acc = nullptr;         // [0] internal weight, [1] sum of squares, [2] 2m
    double* d_q = nullptr;
    int64_t* d_tot = nullptr;      // [2] scan totals
    double* h_q = nullptr;         // pinned
    // local-moving bins
    int* d_binv = nullptr;
    std::vector<int64_t> binoff;   // nb * (NUM_BINS + 1)
    int64_t* d_htoff = nullptr;
    std::vector<int64_t> htoff_base;
    int* d_tkey = nullptr;
    fx_t* d_tval = nullptr;
    int64_t tab_cap = 0;
    // degree / modularity work lists
    int* d_small = nullptr;
    int64_t nsmall = 0;
    RowItem* d_items = nullptr;
    int64_t nitems = 0;

    fx_t twom = 0;
    double twom_d = 0, m = 0;
    std::vector<int> committed;    // per batch: weights applied since its last update

    std::vector<PhaseHist> hist;
    nido_run_stats last{};
};

namespace {

int bin_of(int64_t d, const nido_tuning& t)
{
    if (d <= t.thread_max_deg) return BIN_THREAD;
    for (int b = BIN_G8; b < NUM_BINS; ++b)
        if (d <= kBinMaxDeg[b]) return b;
    return BIN_HUGE;
}

// partition_graph_edge_batch with part_on_batch=1 on one device,
// graph_gpu.cpp:303-339.
void batch_bounds(const std::vector<int64_t>& off, int64_t nv, int nb, std::vector<int64_t>& parts)
{
    const int64_t ne = off[nv] - off[0];
    const int64_t ave = ne / nb;
    std::vector<int64_t> vp;
    vp.reserve((size_t)nb + 2);
    vp.push_back(0);
    int64_t start = 0;
    for (int64_t i = 1; i <= nv; ++i) {
        if (off[i] - off[start] > ave) {
            vp.push_back(i);
            start = i;
        } else if (i == nv) {
            vp.push_back(nv);
        }
    }
    if ((int64_t)vp.size() > nb + 1) {
        const int64_t remain = nv - vp[nb];
        const int64_t nnv = remain / nb;
        for (int i = 1; i < nb; ++i) vp[i] += nnv;
        vp[nb] = nv;
    } else {
        while ((int64_t)vp.size() < nb + 1) vp.push_back(nv);
    }
    parts.assign(vp.begin(), vp.begin() + nb + 1);
}

int64_t pow2_at_least(int64_t x)
{
    int64_t s = 2;
    while (s < x) s <<= 1;
    return s;
}

template <typename T>
int upload(T** d, const std::vector<T>& h, hipStream_t s)
{
    dev_free(*d);
    NIDO_TRY(dev_alloc(d, (int64_t)h.size()));
    if (!h.empty()) HIP_TRY(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return NIDO_OK;
}

// Per-phase set-up: maximum order, batches, degree bins, the row slices of
// the degree/modularity sweeps, k_v and 2m (GraphGPU ctor steps
// graph_gpu.cpp:94-139 and aggregation steps graph_gpu.cpp:1487-1516).
int setup_phase(nido_graph* g)
{
    const int64_t nv = g->nv;
    const auto& off = g->h_off;
    g->maxdeg = 0;
    for (int64_t v = 0; v < nv; ++v) g->maxdeg = std::max(g->maxdeg, off[v + 1] - off[v]);
    batch_bounds(off, nv, g->nb, g->parts);
    g->committed.assign((size_t)g->nb, 0);

    // vertices grouped by (batch, bin), ascending id inside a group
    const int NB1 = NUM_BINS + 1;
    g->binoff.assign((size_t)g->nb * NB1, 0);
    std::vector<int> binv((size_t)std::max<int64_t>(nv, 1));
    std::vector<int64_t> htoff;
    g->htoff_base.assign((size_t)g->nb, 0);
    int64_t pos = 0, tab_need = 0;
    for (int b = 0; b < g->nb; ++b) {
        int64_t cnt[NUM_BINS] = {0};
        for (int64_t v = g->parts[b]; v < g->parts[b + 1]; ++v) cnt[bin_of(off[v + 1] - off[v], g->tuning)]++;
        int64_t cur[NUM_BINS];
        for (int k = 0; k < NUM_BINS; ++k) {
            g->binoff[(size_t)b * NB1 + k] = pos;
            cur[k] = pos;
            pos += cnt[k];
        }
        g->binoff[(size_t)b * NB1 + NUM_BINS] = pos;
        g->htoff_base[(size_t)b] = (int64_t)htoff.size();
        int64_t tsum = 0;
        htoff.push_back(0);
        for (int64_t v = g->parts[b]; v < g->parts[b + 1]; ++v) {
            const int64_t d = off[v + 1] - off[v];
            const int k = bin_of(d, g->tuning);
            binv[(size_t)cur[k]++] = (int)v;
            if (k == BIN_HUGE) {
                tsum += pow2_at_least(2 * d);
                htoff.push_back(tsum);
            }
        }
        tab_need = std::max(tab_need, tsum);
    }
    NIDO_TRY(upload(&g->d_binv, binv, g->stream));
    if (htoff.empty()) htoff.push_back(0);
    NIDO_TRY(upload(&g->d_htoff, htoff, g->stream));
    if (tab_need > g->tab_cap) {
        dev_free(g->d_tkey);
        dev_free(g->d_tval);
        NIDO_TRY(dev_alloc(&g->d_tkey, tab_need));
        NIDO_TRY(dev_alloc(&g->d_tval, tab_need));
        g->tab_cap = tab_need;
    }

    std::vector<int> small;
    std::vector<RowItem> items;
    for (int64_t v = 0; v < nv; ++v) {
        const int64_t d = off[v + 1] - off[v];
        if (d <= kSmallRow) {
            small.push_back((int)v);
        } else {
            for (int64_t e = off[v]; e < off[v + 1]; e += kItemArcs)
                items.push_back(RowItem{v, e, std::min(off[v + 1], e + kItemArcs)});
        }
    }
    g->nsmall = (int64_t)small.size();
    g->nitems = (int64_t)items.size();
    NIDO_TRY(upload(&g->d_small, small, g->stream));
    NIDO_TRY(upload(&g->d_items, items, g->stream));

    launch_vertex_weights(g->d_off, g->d_w, g->d_small, g->nsmall, g->d_items, g->nitems, g->d_k,
                          g->d_kd, nv, g->d_acc + 2, g->stream);
    HIP_TRY(hipGetLastError());
    fx_t twom = 0;
    HIP_TRY(hipMemcpyAsync(&twom, g->d_acc + 2, sizeof(fx_t), hipMemcpyDeviceToHost, g->stream));
    launch_fill_fx(g->d_acc + 2, 1, g->stream);
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->twom = twom;
    g->twom_d = fx_to_double(twom);
    g->m = g->twom_d * 0.5;                     // mass_ = 0.5 * sum(k), graph_cuda.cu:1855
    return NIDO_OK;
}

int check_batch(const nido_graph* g, int batch)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (batch < 0 || batch >= g->nb) return set_error(NIDO_ERR_INVALID, "batch %d out of range", batch);
    return NIDO_OK;
}

void free_graph_arrays(nido_graph* g)
{
    dev_free(g->d_off);
    dev_free(g->d_col);
    dev_free(g->d_w);
}

// ---- the GraphGPU method set ---------------------------------------------
int do_singleton(nido_graph* g)
{
    launch_singleton(g->d_C, g->d_Cn, g->d_a, g->d_k, g->nv, g->stream);
    HIP_TRY(hipGetLastError());
    std::fill(g->committed.begin(), g->committed.end(), 0);
    return NIDO_OK;
}

int do_modularity(nido_graph* g, double* q)
{
    if (g->twom == 0) {   // no weight at all: nothing is internal, nothing moves
        *q = 0.0;
        return NIDO_OK;
    }
    launch_modularity(g->d_off, g->d_col, g->d_w, g->d_C, g->d_a, g->nv, g->d_small, g->nsmall,
                      g->d_items, g->nitems, g->twom_d, g->d_acc, g->d_q, g->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g->h_q, g->d_q, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *q = *g->h_q;
    return NIDO_OK;
}

int do_lm(nido_graph* g, int b)
{
    const int64_t v0 = g->parts[(size_t)b], v1 = g->parts[(size_t)b + 1];
    g->committed[(size_t)b] = 0;
    if (v1 <= v0) return NIDO_OK;
    if (g->twom == 0) {
        HIP_TRY(hipMemcpyAsync(g->d_Cn + v0, g->d_C + v0, sizeof(int64_t) * (size_t)(v1 - v0),
                               hipMemcpyDeviceToDevice, g->stream));
        return NIDO_OK;
    }
    LmArgs A{g->d_off, g->d_col, g->d_w, g->d_C, g->d_Cn, g->d_kd, g->d_k, g->d_a, g->m};
    const int NB1 = NUM_BINS + 1;
    for (int k = 0; k < NUM_BINS; ++k) {
        const int64_t s = g->binoff[(size_t)b * NB1 + k], e = g->binoff[(size_t)b * NB1 + k + 1];
        if (e > s)
            launch_lm_bin(k, A, g->d_binv + s, e - s, g->d_htoff + g->htoff_base[(size_t)b], g->d_tkey,
                          g->d_tval, g->stream);
    }
    HIP_TRY(hipGetLastError());
    return NIDO_OK;
}

int do_commit(nido_graph* g, int b)
{
    launch_commit(g->parts[(size_t)b], g->parts[(size_t)b + 1], g->d_C, g->d_Cn, g->d_k, g->d_a,
                  g->stream);
    HIP_TRY(hipGetLastError());
    g->committed[(size_t)b] = 1;
    return NIDO_OK;
}

int do_restore(nido_graph* g)
{
    launch_commit(0, g->nv, g->d_C, g->d_Cn, g->d_k, g->d_a, g->stream);
    HIP_TRY(hipGetLastError());
    return NIDO_OK;
}

// aggregation, graph_gpu.cpp:1382-1545, on the device.
int do_aggregation(nido_graph* g, int* done)
{
    hipStream_t s = g->stream;
    const int64_t nv = g->nv;
    int64_t* newid = g->d_work;
    launch_mark_communities(g->d_C, nv, newid, s);
    launch_exclusive_scan(newid, newid, nv, g->d_tot, g->d_scan_tmp, s);
    HIP_TRY(hipGetLastError());
    int64_t K = 0;
    HIP_TRY(hipMemcpyAsync(&K, g->d_tot, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    launch_update_clustering(g->d_clus, g->n0, g->d_C, newid, s);   // graph_gpu.cpp:1389
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    if (K == nv || K == 1) {                                          // graph_gpu.cpp:1391-1392
        *done = 1;
        return NIDO_OK;
    }
    *done = 0;

    int64_t *arcs = nullptr, *toff = nullptr, *cnt = nullptr, *noff = nullptr, *ncol = nullptr;
    double* nw = nullptr;
    int* tkey = nullptr;
    fx_t* tval = nullptr;
    void* tmp = nullptr;
    auto cleanup = [&]() {
        dev_free(arcs); dev_free(toff); dev_free(cnt); dev_free(noff);
        dev_free(ncol); dev_free(nw); dev_free(tkey); dev_free(tval);
        if (tmp) (void)hipFree(tmp);
        tmp = nullptr;
    };
    int rc = NIDO_OK;
    do {
        if ((rc = dev_alloc(&arcs, K + 1)) || (rc = dev_alloc(&toff, K + 1)) ||
            (rc = dev_alloc(&cnt, K + 1)) || (rc = dev_alloc(&noff, K + 1)))
            break;
        if (hipMalloc(&tmp, scan_temp_bytes(K + 1)) != hipSuccess) {
            rc = set_error(NIDO_ERR_NOMEM, "scan scratch allocation failed");
            break;
        }
        launch_fill_i64(arcs, 0, K + 1, s);
        launch_member_counts(g->d_C, newid, g->d_off, nv, nullptr, arcs, s);
        launch_table_sizes(arcs, K, arcs, s);                      // in place: arcs -> table size
        launch_exclusive_scan(arcs, toff, K, toff + K, tmp, s);
        int64_t T = 0;
        HIP_TRY(hipMemcpyAsync(&T, toff + K, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if ((rc = dev_alloc(&tkey, T)) || (rc = dev_alloc(&tval, T))) break;
        launch_fill_i32(tkey, -1, T, s);
        launch_fill_fx(tval, T, s);
        launch_coarse_insert(g->d_off, g->d_col, g->d_w, g->d_C, newid, nv, toff, tkey, tval, s);
        launch_coarse_count(toff, tkey, K, cnt, s);
        launch_exclusive_scan(cnt, noff, K, noff + K, tmp, s);
        HIP_TRY(hipGetLastError());
        std::vector<int64_t> hoff((size_t)K + 1);
        HIP_TRY(hipMemcpyAsync(hoff.data(), noff, sizeof(int64_t) * (size_t)(K + 1), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        const int64_t nne = hoff[(size_t)K];
        if ((rc = dev_alloc(&ncol, nne)) || (rc = dev_alloc(&nw, nne))) break;
        launch_coarse_emit(toff, tkey, tval, K, noff, ncol, nw, s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        free_graph_arrays(g);
        g->d_off = noff; noff = nullptr;
        g->d_col = ncol; ncol = nullptr;
        g->d_w = nw; nw = nullptr;
        g->nv = K;
        g->ne = nne;
        g->h_off.swap(hoff);
    } while (0);
    cleanup();
    if (rc) return rc;
    return setup_phase(g);
}

int do_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, int record, int verbose,
           nido_run_stats* st)
{
    g->hist.clear();
    nido_run_stats S{};
    bool done = false;
    int nphase = 0;
    const double t0 = now_s();
    while (!done && nphase < NIDO_MAX_PHASES) {
        PhaseHist h;
        h.st.nv = g->nv;
        h.st.ne = g->ne;
        if (verbose) printf("PHASE #%d: \n", nphase);
        NIDO_TRY(do_singleton(g));
        double Q = 0;
        NIDO_TRY(do_modularity(g, &Q));
        const double Q_old = Q;
        h.st.q0 = Q;
        if (verbose) {
            printf("LOOP# \tQ \t\tdQ\n----------------------------------------\n");
            printf("0 \t%g \t0\n", Q);
        }
        double dQ = 1.0e308;   // MAX_FLOAT
        int64_t loops = 0;
        const double ts = now_s();
        while (tol < dQ && loops < max_loops) {   // louvain_gpu.cpp:53
            for (int b = 0; b < g->nb; ++b) {
                NIDO_TRY(do_lm(g, b));
                NIDO_TRY(do_commit(g, b));
            }
            double Qtmp = 0;
            NIDO_TRY(do_modularity(g, &Qtmp));
            dQ = Qtmp - Q;
            h.iter_q.push_back(Qtmp);
            loops++;
            if (dQ < 0) NIDO_TRY(do_restore(g));
            else Q = Qtmp;
            if (verbose) printf("%lld \t%g \t%g\n", (long long)loops, Qtmp, dQ);
        }
        HIP_TRY(hipStreamSynchronize(g->stream));
        const double te = now_s();
        h.st.loops = loops;
        h.st.q = Q;
        h.st.t_loop = te - ts;
        S.t_loop += te - ts;
        S.total_loops += loops;
        S.q = Q;
        if (verbose) {
            if (loops >= max_loops) printf("Exceed maximum loop number\n");
            printf("Final Q: %g\nTime elapse %g s\n", Q, te - ts);
        }
        if (record) {
            h.comm.resize((size_t)g->nv);
            if (g->nv)
                HIP_TRY(hipMemcpy(h.comm.data(), g->d_C, sizeof(int64_t) * (size_t)g->nv, hipMemcpyDeviceToHost));
        }
        if (Q - Q_old > tol_phase) {                    // louvain_gpu.cpp:108
            const double a0 = now_s();
            int d = 0;
            // keep the phase record even if aggregation fails
            g->hist.push_back(h);
            NIDO_TRY(do_aggregation(g, &d));
            done = d != 0;
            const double at = now_s() - a0;
            g->hist.back().st.t_agg = at;
            S.t_agg += at;
            if (verbose) printf("Aggregation time: %g s\n", at);
        } else {
            done = true;
            g->hist.push_back(h);
        }
        if (verbose) printf("----------------------------------------\n");
        nphase++;
    }
    S.t_total = now_s() - t0;
    S.phases = nphase;
    S.done = done ? 1 : 0;
    if (verbose) {
        printf("Total time elapse: %g s\nTotal Loops: %lld\n", S.t_total, (long long)S.total_loops);
        printf("Time per loop: %g s/loop\n", S.total_loops ? S.t_loop / (double)S.total_loops : 0.0);
        printf("Aggregation time: %g s\n", S.t_agg);
    }
    g->last = S;
    if (st) *st = S;
    return NIDO_OK;
}

int create_impl(int64_t nv, const int64_t* off, const int64_t* col, const double* w, int nb,
                void* stream, nido_graph** out)
{
    if (!out) return set_error(NIDO_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (nv < 0 || !off) return set_error(NIDO_ERR_INVALID, "invalid vertex count or offsets");
    if (nv >= (1ll << 31) - 1) return set_error(NIDO_ERR_CAPACITY, "nv must be below 2^31-1");
    if (nb < 1) return set_error(NIDO_ERR_INVALID, "nbatches must be >= 1");
    const int64_t base = off[0];
    for (int64_t i = 0; i < nv; ++i)
        if (off[i + 1] < off[i]) return set_error(NIDO_ERR_INVALID, "offsets decrease at %lld", (long long)i);
    const int64_t ne = off[nv] - base;
    if (ne > 0 && (!col || !w)) return set_error(NIDO_ERR_INVALID, "null cols/weights");
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t c = col[base + e];
        if (c < 0 || c >= nv) return set_error(NIDO_ERR_INVALID, "col[%lld]=%lld out of range", (long long)e, (long long)c);
        const double x = w[base + e];
        if (!(x == x) || x >= 0x1p62 || x <= -0x1p62)
            return set_error(NIDO_ERR_INVALID, "weight[%lld] is not finite or |w| >= 2^62", (long long)e);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_error(NIDO_ERR_HIP, "no HIP device available");
    }
    nido_graph* g = new nido_graph;
    g->nb = nb;
    g->tuning.thread_max_deg = kThreadMaxDeg;
    g->tuning.use_graph = 0;
    auto fail = [&](int rc) {
        nido_graph_destroy(g);
        return rc;
    };
    if (hipGetDevice(&g->device) != hipSuccess) return fail(set_error(NIDO_ERR_HIP, "hipGetDevice failed"));
    if (stream) {
        g->stream = reinterpret_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
            return fail(set_error(NIDO_ERR_HIP, "hipStreamCreate failed"));
        g->own_stream = true;
    }
    g->n0 = nv;
    g->nv = nv;
    g->ne = ne;
    g->h_off.resize((size_t)nv + 1);
    for (int64_t i = 0; i <= nv; ++i) g->h_off[(size_t)i] = off[i] - base;
    int rc;
    if ((rc = dev_alloc(&g->d_off, nv + 1)) || (rc = dev_alloc(&g->d_col, ne)) ||
        (rc = dev_alloc(&g->d_w, ne)) || (rc = dev_alloc(&g->d_C, nv)) ||
        (rc = dev_alloc(&g->d_Cn, nv)) || (rc = dev_alloc(&g->d_k, nv)) ||
        (rc = dev_alloc(&g->d_a, nv)) || (rc = dev_alloc(&g->d_kd, nv)) ||
        (rc = dev_alloc(&g->d_clus, nv)) || (rc = dev_alloc(&g->d_work, nv + 1)) ||
        (rc = dev_alloc(&g->d_acc, 4)) || (rc = dev_alloc(&g->d_q, 1)) ||
        (rc = dev_alloc(&g->d_tot, 2)))
        return fail(rc);
    if (hipMalloc(&g->d_scan_tmp, scan_temp_bytes(nv + 1)) != hipSuccess)
        return fail(set_error(NIDO_ERR_NOMEM, "scan scratch allocation failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&g->h_q), sizeof(double)) != hipSuccess)
        return fail(set_error(NIDO_ERR_NOMEM, "pinned allocation failed"));
    hipStream_t s = g->stream;
    if (hipMemcpyAsync(g->d_off, g->h_off.data(), sizeof(int64_t) * (size_t)(nv + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        (ne > 0 && hipMemcpyAsync(g->d_col, col + base, sizeof(int64_t) * (size_t)ne, hipMemcpyHostToDevice, s) != hipSuccess) ||
        (ne > 0 && hipMemcpyAsync(g->d_w, w + base, sizeof(double) * (size_t)ne, hipMemcpyHostToDevice, s) != hipSuccess))
        return fail(set_error(NIDO_ERR_HIP, "upload failed"));
    launch_fill_fx(g->d_acc, 4, s);
    // Clustering::commIds_ starts as the identity (clustering.cpp:4-10)
    std::vector<int64_t> ident((size_t)std::max<int64_t>(nv, 1));
    for (int64_t i = 0; i < nv; ++i) ident[(size_t)i] = i;
    if (nv > 0 && hipMemcpyAsync(g->d_clus, ident.data(), sizeof(int64_t) * (size_t)nv, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(set_error(NIDO_ERR_HIP, "upload failed"));
    if (hipStreamSynchronize(s) != hipSuccess) return fail(set_error(NIDO_ERR_HIP, "upload sync failed"));
    if ((rc = setup_phase(g))) return fail(rc);
    if ((rc = do_singleton(g))) return fail(rc);
    *out = g;
    return NIDO_OK;
}

template <typename T>
int d2h(nido_graph* g, T* dst, const T* src, int64_t n)
{
    if (!dst) return set_error(NIDO_ERR_INVALID, "null output array");
    if (n > 0) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return NIDO_OK;
}

}  // namespace

#define API_GUARD(body)                                                          \
    try {                                                                        \
        body                                                                     \
    } catch (const std::bad_alloc&) {                                            \
        return set_error(NIDO_ERR_NOMEM, "host allocation failed");              \
    }

extern "C" {

int nido_abi_version(void) { return NIDO_ABI_VERSION; }

int nido_device_count(int* count)
{
    if (!count) return set_error(NIDO_ERR_INVALID, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return NIDO_OK;
}

int nido_graph_create(int64_t nv, int64_t ne, const int64_t* off, const int64_t* col,
                      const double* w, int nbatches, void* stream, nido_graph** out)
{
    (void)ne;
    API_GUARD(return create_impl(nv, off, col, w, nbatches, stream, out);)
}

int nido_graph_destroy(nido_graph* g)
{
    if (!g) return NIDO_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    free_graph_arrays(g);
    dev_free(g->d_C); dev_free(g->d_Cn); dev_free(g->d_k); dev_free(g->d_a); dev_free(g->d_kd);
    dev_free(g->d_clus); dev_free(g->d_work); dev_free(g->d_acc); dev_free(g->d_q); dev_free(g->d_tot);
    dev_free(g->d_binv); dev_free(g->d_htoff); dev_free(g->d_tkey); dev_free(g->d_tval);
    dev_free(g->d_small); dev_free(g->d_items);
    if (g->d_scan_tmp) (void)hipFree(g->d_scan_tmp);
    if (g->h_q) (void)hipHostFree(g->h_q);
    if (g->own_stream && g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return NIDO_OK;
}

int nido_graph_info(const nido_graph* g, int64_t* nv, int64_t* ne, int64_t* max_order, double* mass)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    if (max_order) *max_order = g->maxdeg;
    if (mass) *mass = g->m;
    return NIDO_OK;
}

int nido_graph_batch_bounds(const nido_graph* g, int64_t* parts)
{
    if (!g || !parts) return set_error(NIDO_ERR_INVALID, "null argument");
    std::copy(g->parts.begin(), g->parts.end(), parts);
    return NIDO_OK;
}

int nido_graph_set_tuning(nido_graph* g, const nido_tuning* t)
{
    if (!g || !t) return set_error(NIDO_ERR_INVALID, "null argument");
    if (t->thread_max_deg < 0 || t->thread_max_deg > kThreadMaxDeg)
        return set_error(NIDO_ERR_INVALID, "thread_max_deg must be in [0, %d]", kThreadMaxDeg);
    API_GUARD(g->tuning = *t; return setup_phase(g);)
}

int nido_graph_get_tuning(const nido_graph* g, nido_tuning* t)
{
    if (!g || !t) return set_error(NIDO_ERR_INVALID, "null argument");
    *t = g->tuning;
    return NIDO_OK;
}

int nido_graph_sync(nido_graph* g)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return NIDO_OK;
}

int nido_singleton_partition(nido_graph* g)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return do_singleton(g);
}

int nido_compute_modularity(nido_graph* g, double* q)
{
    if (!g || !q) return set_error(NIDO_ERR_INVALID, "null argument");
    return do_modularity(g, q);
}

int nido_louvain_update(nido_graph* g, int batch)
{
    NIDO_TRY(check_batch(g, batch));
    return do_lm(g, batch);
}

int nido_update_community_weights(nido_graph* g, int batch)
{
    NIDO_TRY(check_batch(g, batch));
    return do_commit(g, batch);
}

int nido_update_community_ids(nido_graph* g, int batch)
{
    NIDO_TRY(check_batch(g, batch));
    if (!g->committed[(size_t)batch])
        return set_error(NIDO_ERR_STATE, "update_community_ids(%d) before update_community_weights", batch);
    return NIDO_OK;
}

int nido_restore_community(nido_graph* g)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return do_restore(g);
}

int nido_aggregation(nido_graph* g, int* done)
{
    if (!g || !done) return set_error(NIDO_ERR_INVALID, "null argument");
    API_GUARD(return do_aggregation(g, done);)
}

int nido_get_communities(nido_graph* g, int64_t* comm)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return d2h(g, comm, g->d_C, g->nv);
}

int nido_get_new_communities(nido_graph* g, int64_t* comm)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return d2h(g, comm, g->d_Cn, g->nv);
}

int nido_get_vertex_weights(nido_graph* g, double* k)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return d2h(g, k, g->d_kd, g->nv);
}

int nido_get_community_weights(nido_graph* g, double* a)
{
    if (!g || !a) return set_error(NIDO_ERR_INVALID, "null argument");
    API_GUARD(
        std::vector<fx_t> h((size_t)std::max<int64_t>(g->nv, 1));
        NIDO_TRY(d2h(g, h.data(), g->d_a, g->nv));
        for (int64_t i = 0; i < g->nv; ++i) a[i] = fx_to_double(h[(size_t)i]);
        return NIDO_OK;)
}

int nido_get_graph(nido_graph* g, int64_t* off, int64_t* col, double* w)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (off) NIDO_TRY(d2h(g, off, g->d_off, g->nv + 1));
    if (col) NIDO_TRY(d2h(g, col, g->d_col, g->ne));
    if (w) NIDO_TRY(d2h(g, w, g->d_w, g->ne));
    return NIDO_OK;
}

int nido_get_clustering(nido_graph* g, int64_t* comm)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    return d2h(g, comm, g->d_clus, g->n0);
}

int nido_dump_partition(nido_graph* g, const char* path)
{
    if (!g || !path) return set_error(NIDO_ERR_INVALID, "null argument");
    API_GUARD(
        std::vector<int64_t> c((size_t)std::max<int64_t>(g->n0, 1));
        NIDO_TRY(d2h(g, c.data(), g->d_clus, g->n0));
        FILE* f = fopen(path, "wt");
        if (!f) return set_error(NIDO_ERR_IO, "cannot open %s", path);
        for (int64_t i = 0; i < g->n0; ++i) fprintf(f, "%lld\n", (long long)c[(size_t)i]);
        fclose(f);
        return NIDO_OK;)
}

int nido_louvain_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, int record,
                     int verbose, nido_run_stats* stats)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    API_GUARD(return do_run(g, max_loops, tol, tol_phase, record, verbose, stats);)
}

int nido_run_phase(const nido_graph* g, int phase, nido_phase_stats* out)
{
    if (!g || !out) return set_error(NIDO_ERR_INVALID, "null argument");
    if (phase < 0 || phase >= (int)g->hist.size()) return set_error(NIDO_ERR_INVALID, "no phase %d", phase);
    *out = g->hist[(size_t)phase].st;
    return NIDO_OK;
}

int nido_run_iteration_q(const nido_graph* g, int phase, double* q)
{
    if (!g || !q) return set_error(NIDO_ERR_INVALID, "null argument");
    if (phase < 0 || phase >= (int)g->hist.size()) return set_error(NIDO_ERR_INVALID, "no phase %d", phase);
    const auto& v = g->hist[(size_t)phase].iter_q;
    std::copy(v.begin(), v.end(), q);
    return NIDO_OK;
}

int nido_run_phase_communities(const nido_graph* g, int phase, int64_t* comm)
{
    if (!g || !comm) return set_error(NIDO_ERR_INVALID, "null argument");
    if (phase < 0 || phase >= (int)g->hist.size()) return set_error(NIDO_ERR_INVALID, "no phase %d", phase);
    const auto& v = g->hist[(size_t)phase].comm;
    if (v.size() != (size_t)g->hist[(size_t)phase].st.nv)
        return set_error(NIDO_ERR_STATE, "phase %d was not recorded (run with record=1)", phase);
    std::copy(v.begin(), v.end(), comm);
    return NIDO_OK;
}

}  // extern "C"
