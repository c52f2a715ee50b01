// graph_gpu.hip -- device-resident Louvain state and the C ABI over it.
//
// Mirrors the reference's GraphGPU (graph_gpu.hpp:168-247) and
// LouvainGPU::run (louvain_gpu.cpp:6-139) on one MI355X: the whole graph
// stays in HBM, every step is a hand-written kernel on the handle's stream,
// and the host only reads back one scalar per modularity evaluation.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/nido_amd.h"
#include "errors.h"
#include "host_bins.h"
#include "host_check.h"
#include "host_copy.h"
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


// Device memory is stream-ordered: allocations and frees made while a
// handle's AllocScope is active go through a library-owned memory pool on
// that handle's stream (hipMallocFromPoolAsync / hipFreeAsync), so a free
// never waits for the whole device -- handles running side by side on
// their own streams (a serving loop) do not stall each other.  The pool
// keeps freed memory for reuse.
thread_local hipStream_t t_alloc_stream = nullptr;

struct AllocScope {
    hipStream_t prev;
    explicit AllocScope(hipStream_t s) : prev(t_alloc_stream) { t_alloc_stream = s; }
    ~AllocScope() { t_alloc_stream = prev; }
};

hipMemPool_t device_pool()
{
    static std::mutex mu;
    static std::vector<hipMemPool_t> pools;
    static std::vector<char> tried;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)pools.size() <= dev) {
        pools.resize((size_t)dev + 1, nullptr);
        tried.resize((size_t)dev + 1, 0);
    }
    if (!tried[(size_t)dev]) {
        tried[(size_t)dev] = 1;
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            pools[(size_t)dev] = pool;
        } else {
            (void)hipGetLastError();
        }
    }
    return pools[(size_t)dev];
}

int raw_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipMemPool_t pool = t_alloc_stream ? device_pool() : nullptr;
    hipError_t e = pool ? hipMallocFromPoolAsync(p, bytes, pool, t_alloc_stream) : hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(NIDO_ERR_NOMEM, "device allocation of %lld bytes failed: %s", (long long)bytes,
                         hipGetErrorString(e));
    }
    return NIDO_OK;
}

void raw_free(void* p)
{
    if (!p) return;
    if (t_alloc_stream && device_pool()) (void)hipFreeAsync(p, t_alloc_stream);
    else (void)hipFree(p);
}

template <typename T>
int dev_alloc(T** p, int64_t count)
{
    void* q = nullptr;
    const int rc = raw_alloc(&q, sizeof(T) * (size_t)(count > 0 ? count : 1));
    *p = static_cast<T*>(q);
    return rc;
}

template <typename T>
void dev_free(T*& p)
{
    raw_free(p);
    p = nullptr;
}

struct PhaseHist {
    nido_phase_stats st{};
    std::vector<double> iter_q;
    std::vector<int64_t> comm;
};

// One graph's outcome in a batched run.
struct BatchResult {
    nido_run_stats st{};
    std::vector<nido_phase_stats> phases;
};

constexpr int64_t kSegArcs = 65536;    // arcs per per-graph modularity item
constexpr int64_t kSegVerts = 16384;   // vertices per per-graph modularity item

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Per-kernel-class timing with HIP events on the handle's stream.
struct Prof {
    struct Entry {
        int cls;
        hipEvent_t a, b;
        double bytes;
    };
    bool on = false;
    unsigned mask = ~0u;               // kernel classes that get events
    int stride = 1;                    // events on every stride-th launch of a class
    int64_t seq[NIDO_K_CLASSES] = {};
    std::vector<hipEvent_t> pool;
    std::vector<Entry> pending;
    int64_t launches[NIDO_K_CLASSES] = {};
    double ms[NIDO_K_CLASSES] = {};
    double bytes[NIDO_K_CLASSES] = {};

    hipEvent_t get()
    {
        if (pool.empty()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // Folds every finished interval into the totals (the stream has been
    // synchronised by the caller).
    void harvest()
    {
        for (auto& p : pending) {
            float t = 0;
            if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
                launches[p.cls]++;
                ms[p.cls] += t;
                bytes[p.cls] += p.bytes;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset()
    {
        for (int i = 0; i < NIDO_K_CLASSES; ++i) {
            launches[i] = 0;
            ms[i] = 0;
            bytes[i] = 0;
            seq[i] = 0;
        }
    }
    ~Prof()
    {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Brackets the launches issued in its lifetime when profiling is on.
struct ProfScope {
    Prof& P;
    hipStream_t s;
    int cls;
    double bytes;
    hipEvent_t a = nullptr;
    ProfScope(Prof& p, hipStream_t st, int c, double b) : P(p), s(st), cls(c), bytes(b)
    {
        if (P.on && ((P.mask >> c) & 1u) && P.seq[c]++ % P.stride == 0 && (a = P.get()))
            (void)hipEventRecord(a, s);
    }
    ~ProfScope()
    {
        if (!a) return;
        hipEvent_t b = P.get();
        if (b && hipEventRecord(b, s) == hipSuccess) P.pending.push_back({cls, a, b, bytes});
        else P.pool.push_back(a);
    }
};

}  // namespace

struct nido_graph {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t side = nullptr;    // long rows of a batch, beside the short rows on `stream`
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
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
    int* d_C32 = nullptr;          // (int)d_C, kept in step by singleton and commit
    fx_t* d_k = nullptr;
    fx_t* d_a = nullptr;
    double* d_kd = nullptr;
    int64_t* d_clus = nullptr;
    int64_t* d_work = nullptr;     // n0+1 int64 scratch (renumbering)
    int* d_src = nullptr;          // row of every arc, capacity = phase-0 arcs
    int* d_col32 = nullptr;        // (int)d_col, capacity = phase-0 arcs
    int64_t* d_wq = nullptr;       // d_w in units of 2^-F when the short rows sum in fixed point
    int64_t wq_cap = 0;
    bool fixed = false;            // this phase's short rows sum in 64-bit fixed point (d_wq)
    double wscale = 1.0;           // 2^-F
    fx_t* d_part = nullptr;        // 2 * kPartBlocks reduction partials
    void* d_scan_tmp = nullptr;
    // scalars
    fx_t* d_acc = nullptr;         // [0] internal weight, [1] sum of squares, [2] 2m
    double* d_q = nullptr;
    int64_t* d_tot = nullptr;      // [2] scan totals
    double* h_q = nullptr;         // pinned
    // local-moving bins: per batch, the fused small-row launch and the
    // partitioned long-row launch
    BinRow* d_binv = nullptr;
    std::vector<SmallBins> small_bins;   // nb
    std::vector<BigArgs> big_args;       // nb
    BigItem* d_big_items = nullptr;
    KeyChunk* d_key_chunks = nullptr;
    int* d_bkey = nullptr;         // long-row buckets: keys, weights, segment starts
    double* d_bw = nullptr;
    int64_t key_cap = 0;
    int* d_cpo = nullptr;
    int64_t cpo_cap = 0;
    int64_t* d_big_poff = nullptr;
    int64_t* d_big_toff = nullptr;
    double* d_pg = nullptr;
    int* d_pc = nullptr;
    fx_t* d_pself = nullptr;
    int64_t part_cap = 0;
    int* d_cnt = nullptr;
    int64_t slot_cap = 0;
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
    bool nonneg = true;            // every weight >= 0 (coarse weights are sums of these)
    int narrow = 0;                // long-row tables may keep 32-bit high words
    int64_t listed = 0;            // vertices of the graphs the work lists hold
    int64_t rebuilds = 0;          // work-list rebuilds for the iterating graphs (batched runs)
    std::vector<int> committed;    // per batch: weights applied since its last update

    std::vector<double> short_bytes, long_bytes;   // per batch, algorithmic
    Prof prof;

    std::vector<PhaseHist> hist;
    nido_run_stats last{};

    // ---- several graphs in one handle (a batched run); G == 1 otherwise ----
    int G = 1;
    std::vector<int64_t> vb;             // G+1 vertex bases of the current phase
    std::vector<int64_t> vb0;            // G+1 original vertex bases
    std::vector<int64_t> gparts;         // G * (nb+1) batch bounds (handle vertex ids)
    std::vector<fx_t> twom_g;            // exact 2m per graph
    std::vector<double> twom_dg;
    int* d_gid = nullptr;                // graph of every current vertex (capacity n0)
    int* d_gid0 = nullptr;               // graph of every original vertex
    int64_t* d_vb = nullptr;             // G+1
    int64_t* d_vbn = nullptr;            // G+1 (bases after renumbering)
    double* d_mg = nullptr;              // {mass, 1 / mass} per graph
    double* d_twomg = nullptr;           // 2m per graph, FP64
    unsigned char* d_act = nullptr;      // per graph: moves in this iteration
    unsigned char* d_keep = nullptr;     // per graph: selected for a restore / a coarsening
    fx_t* d_gacc = nullptr;              // 2 accumulators per graph
    double* d_gq = nullptr;              // modularity per graph, then the iteration's moved arcs (bits)
    unsigned long long* d_moved = nullptr;   // kMovedSlots counters of moved arcs
    SegItem* d_seg = nullptr;
    int64_t nseg = 0;
    std::vector<int64_t> blist0, blistn; // per batch: its slice of d_binv (non-isolated vertices)
    hipGraphExec_t iter_exec[2] = {};    // one iteration's launches (full / incremental Q), per phase
    // incremental modularity (per graph with G > 1)
    int64_t* d_gp = nullptr;             // G * (nb+1): gparts on the device
    int64_t gp_cap = 0;
    int* d_dcounts = nullptr;            // 2 per batch
    RowItem* d_dchunks = nullptr;
    int64_t dchunk_cap = 0;
    int* d_dcomms = nullptr;             // capacity n0
    int* d_touched = nullptr;            // capacity n0, kept zero between batches
    fx_t* d_term = nullptr;              // capacity n0
    fx_t* d_iacc = nullptr;              // 2
    bool inc_ok = false;                 // d_term and d_iacc (d_gacc) describe the current partition
    bool ident = false;                  // C is the singleton partition: set only by do_singleton
    int64_t last_moved = -1;             // several graphs: arcs the last iteration moved; -1: unknown
    // Every arc (v, u, w) has a reverse arc (u, v, w): the incremental update
    // counts an arc whose state changes twice, for itself and its reverse.
    // Checked once at creation (coarse graphs of such a graph are such
    // graphs too); otherwise every iteration sweeps.
    bool symmetric = false;
    bool ride_delta = false;             // short-row launches carry the previous batch's delta
    int64_t iters_since_singleton = 0;   // the first iteration moves most vertices: full sweep there
    // several graphs: the iteration's moved-arc counters (their sum is read
    // back with Q, after d_gq[G - 1]); a single graph does not count them
    unsigned long long* moved() const { return G == 1 ? nullptr : d_moved; }
    ModDelta delta() const
    {
        if (G == 1)
            return ModDelta{d_dcounts, d_dchunks, d_dcomms, d_touched, d_term,   d_iacc,
                            nullptr,   nullptr,   nullptr,  0,         moved()};
        return ModDelta{d_dcounts, d_dchunks, d_dcomms, d_touched, d_term, d_gacc,
                        d_gid,     d_gp,      d_twomg,  nb + 1,    moved()};
    }
    std::vector<BatchResult> bres;       // per graph, from the last batched run
    // every graph's clustering from the last batched run (original vertex ->
    // final community, graph-local ids; graph i at [vb0[i], vb0[i+1]), or
    // all n0 vertices for one graph), left uninitialised until copied
    std::unique_ptr<int64_t[]> bclus;
    int64_t bclus_n = 0;
    // ---- arcs streamed through a device window (graph_gpu.cpp:341-417, 967-1027) ----
    int64_t window = 0;                  // arcs the device holds at once; 0: every phase resident
    bool windowed = false;               // this phase has more arcs than the window
    bool win_single = false;             // windowed, and every batch's arcs fit the window at once
    bool win_double = false;             // ... and two batches at once: copies overlap the launches
    hipEvent_t ev_copied[2] = {nullptr, nullptr};   // half h of the window holds its next batch
    hipEvent_t ev_free[2] = {nullptr, nullptr};     // half h's last batch is done with it
    bool win_live = false;               // an iteration is streaming its batches through the halves
    int* h_col32 = nullptr;              // the phase's CSR on the host (pinned) while windowed
    double* h_w = nullptr;
    int64_t h_cap = 0;
    std::vector<int64_t> units;          // window units: consecutive rows of one batch
    std::vector<int> first_unit;         // batch b: units first_unit[b] .. first_unit[b + 1]
    int64_t win_e0 = -1, win_e1 = -1;    // arcs [win_e0, win_e1) sit in the window buffers
};

namespace {

template <typename T>
int d2h(nido_graph* g, T* dst, const T* src, int64_t n);
int host_clustering(nido_graph* g, const int64_t* src);

// Device bytes of a resident graph: per vertex (state, bins, scratch) and per
// arc at the peak of a phase (CSR, 32-bit mirrors, fixed-point weights, the
// coarsening's hash tables and the coarse CSR being built); per arc of a
// window (its CSR slices and long-row buckets).
constexpr double kDeviceBytesPerVertex = 256.0;
constexpr double kDeviceBytesPerArc = 96.0;
constexpr double kWindowBytesPerArc = 48.0;

void default_tuning(nido_tuning& t)
{
    t = nido_tuning{};
    t.thread_max_deg = kThreadMaxDeg;
    t.big_min_deg = (int32_t)kBinMaxDeg[BIN_B256];
    t.big_part_arcs = (int32_t)kBigPartArcs;
    t.big_key_limit = kBigLimit;
    t.use_graph = 1;
    t.big_slots = kBigSlots;
    t.split_small = 0;
    t.overlap_long = 1;
    t.full_modularity = 0;
    t.fixed_short = 1;
}

void batch_bounds(const std::vector<int64_t>& off, int64_t nv, int nb, std::vector<int64_t>& parts)
{
    nido::batch_bounds(off.data(), nv, nb, parts);
}


template <typename T>
int upload(T** d, const std::vector<T>& h, hipStream_t s)
{
    dev_free(*d);
    NIDO_TRY(dev_alloc(d, (int64_t)h.size()));
    if (!h.empty()) HIP_TRY(staged_h2d(*d, h.data(), sizeof(T) * h.size(), s));
    return NIDO_OK;
}

void drop_iteration_graph(nido_graph* g);

// Phases of at least this many vertices (and arcs resident on the device)
// build their degree bins and row lists on the device: for 4,096 graphs of
// 48.5 M vertices in one handle the host built them in ~100 ms and uploaded
// 1 GB of them, and one host thread took ~10 ms for the 0.8 M vertices of
// their first coarse phase; the device builder costs a few round trips.
constexpr int64_t kDeviceListsMin = int64_t(1) << 17;
std::atomic<int64_t> g_device_lists_min{kDeviceListsMin};
// A handle of several graphs rebuilds its work lists for the graphs still
// iterating once their vertices are at most this share (percent) of the
// listed ones: the short-row launches otherwise visit every stopped graph's
// rows, a floor that stays while the last graphs of a phase converge.
constexpr int kListRebuildPct = 50;
std::atomic<int> g_list_rebuild_pct{kListRebuildPct};

// The degree bins of the phase built on the device into d_binv (bins.hip);
// the host gets the group bounds and each batch's long rows, and builds the
// long-row work from those.
int device_bins(nido_graph* g, PhaseBins& pb, const unsigned char* keep)
{
    const int nb = g->nb, G = g->G;
    const auto& off = g->h_off;
    hipStream_t s = g->stream;
    std::vector<BinTile> tiles;
    std::vector<int64_t> tb0;
    bin_tiles(nb, G, g->parts, g->gparts, keep, tiles, tb0);
    const int64_t T = (int64_t)tiles.size(), nc = NUM_BINS * T;
    BinTile* d_tiles = nullptr;
    int64_t *d_tb0 = nullptr, *d_cnt = nullptr, *d_starts = nullptr, *d_total = nullptr;
    void* tmp = nullptr;
    auto cleanup = [&]() {
        dev_free(d_tiles); dev_free(d_tb0); dev_free(d_cnt); dev_free(d_starts); dev_free(d_total);
        raw_free(tmp);
        tmp = nullptr;
    };
    std::vector<int64_t> starts((size_t)nb * NUM_BINS, -1);
    int64_t rows = 0;
    int rc = NIDO_OK;
    do {
        if ((rc = dev_alloc(&d_tiles, std::max<int64_t>(T, 1))) || (rc = dev_alloc(&d_tb0, (int64_t)nb + 1)) ||
            (rc = dev_alloc(&d_cnt, std::max<int64_t>(nc, 1))) ||
            (rc = dev_alloc(&d_starts, std::max<int64_t>((int64_t)nb * NUM_BINS, 1))) ||
            (rc = dev_alloc(&d_total, 1)) || (rc = raw_alloc(&tmp, scan_temp_bytes(std::max<int64_t>(nc, 1)))))
            break;
        if (T > 0 && (hipMemcpyAsync(d_tiles, tiles.data(), sizeof(BinTile) * (size_t)T, hipMemcpyHostToDevice,
                                     s) != hipSuccess ||
                      hipMemcpyAsync(d_tb0, tb0.data(), sizeof(int64_t) * tb0.size(), hipMemcpyHostToDevice,
                                     s) != hipSuccess)) {
            rc = set_error(NIDO_ERR_HIP, "bin tile upload failed");
            break;
        }
        if (T > 0) {
            launch_bin_count(d_tiles, T, g->d_off, g->tuning.thread_max_deg, g->tuning.big_min_deg, d_cnt, s);
            launch_exclusive_scan(d_cnt, d_cnt, nc, d_total, tmp, s);
            if (hipMemcpyAsync(&rows, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = set_error(NIDO_ERR_HIP, "bin count failed");
                break;
            }
        }
        dev_free(g->d_binv);
        if ((rc = dev_alloc(&g->d_binv, std::max<int64_t>(rows, 1)))) break;
        if (T > 0) {
            launch_bin_scatter(d_tiles, T, g->d_off, g->tuning.thread_max_deg, g->tuning.big_min_deg, d_cnt,
                               g->d_binv, s);
            launch_bin_starts(d_cnt, d_tb0, nb, d_starts, s);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(starts.data(), d_starts, sizeof(int64_t) * starts.size(), hipMemcpyDeviceToHost,
                               s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = set_error(NIDO_ERR_HIP, "bin scatter failed");
                break;
            }
        }
    } while (false);
    cleanup();
    if (rc != NIDO_OK) return rc;
    // a batch without tiles starts where the next one does
    int64_t next = rows;
    for (int b = nb - 1; b >= 0; --b) {
        for (int k = NUM_BINS - 1; k >= 0; --k) {
            int64_t& st = starts[(size_t)b * NUM_BINS + k];
            if (st < 0) st = next;
        }
        next = starts[(size_t)b * NUM_BINS];
    }
    std::vector<int64_t> cnt((size_t)nb * NUM_BINS), bverts((size_t)nb, 0), barcs((size_t)nb, 0);
    for (size_t i = 0; i < cnt.size(); ++i) cnt[i] = (i + 1 < cnt.size() ? starts[i + 1] : rows) - starts[i];
    std::vector<std::vector<BinRow>> big((size_t)nb);
    for (int b = 0; b < nb; ++b) {
        const int64_t n = cnt[(size_t)b * NUM_BINS + BIN_BIG];
        big[(size_t)b].resize((size_t)n);
        if (n > 0)
            HIP_TRY(hipMemcpyAsync(big[(size_t)b].data(), g->d_binv + starts[(size_t)b * NUM_BINS + BIN_BIG],
                                   sizeof(BinRow) * (size_t)n, hipMemcpyDeviceToHost, s));
        auto range = [&](int64_t v0, int64_t v1) {
            bverts[(size_t)b] += v1 - v0;
            barcs[(size_t)b] += off[(size_t)v1] - off[(size_t)v0];
        };
        if (G == 1) range(g->parts[(size_t)b], g->parts[(size_t)b + 1]);
        else
            for (int i = 0; i < G; ++i)
                if (!keep || keep[i])
                    range(g->gparts[(size_t)i * (nb + 1) + b], g->gparts[(size_t)i * (nb + 1) + b + 1]);
    }
    HIP_TRY(hipStreamSynchronize(s));
    build_phase_lists(off.data(), nb, cnt, big, bverts, barcs, g->tuning, pb);
    return NIDO_OK;
}

// The row lists of the degree and modularity sweeps built on the device.
int device_row_lists(nido_graph* g)
{
    const int64_t nv = g->nv, nt = (nv + kBinTileVerts - 1) / kBinTileVerts;
    hipStream_t s = g->stream;
    int64_t *cs = nullptr, *ci = nullptr, *tot = nullptr;
    void* tmp = nullptr;
    auto cleanup = [&]() {
        dev_free(cs); dev_free(ci); dev_free(tot);
        raw_free(tmp);
        tmp = nullptr;
    };
    int64_t h_tot[2] = {0, 0};
    int rc = NIDO_OK;
    do {
        if ((rc = dev_alloc(&cs, std::max<int64_t>(nt, 1))) || (rc = dev_alloc(&ci, std::max<int64_t>(nt, 1))) ||
            (rc = dev_alloc(&tot, 2)) || (rc = raw_alloc(&tmp, scan_temp_bytes(std::max<int64_t>(nt, 1)))))
            break;
        if (nt > 0) {
            launch_rowlist_count(g->d_off, nv, cs, ci, s);
            launch_exclusive_scan(cs, cs, nt, tot, tmp, s);
            launch_exclusive_scan(ci, ci, nt, tot + 1, tmp, s);
            if (hipMemcpyAsync(h_tot, tot, sizeof(h_tot), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = set_error(NIDO_ERR_HIP, "row list count failed");
                break;
            }
        }
        dev_free(g->d_small);
        dev_free(g->d_items);
        if ((rc = dev_alloc(&g->d_small, std::max<int64_t>(h_tot[0], 1))) ||
            (rc = dev_alloc(&g->d_items, std::max<int64_t>(h_tot[1], 1))))
            break;
        if (nt > 0) launch_rowlist_scatter(g->d_off, nv, cs, ci, g->d_small, g->d_items, s);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            rc = set_error(NIDO_ERR_HIP, "row list scatter failed");
    } while (false);
    cleanup();
    g->nsmall = h_tot[0];
    g->nitems = h_tot[1];
    return rc;
}

// The work lists of the phase's launches (degree bins, long-row items,
// key-fill chunks, the per-launch arguments), for every graph of the
// handle or, with keep set, only the graphs g with keep[g] set (the others
// sit out the rest of the phase).
int build_lists(nido_graph* g, const unsigned char* keep)
{
    const int64_t nv = g->nv;
    const auto& off = g->h_off;
    const int G = g->G;
    // launch units: the batches, or their window units (a handle of several
    // graphs keeps its batch bounds per graph, in gparts)
    const std::vector<int64_t>& lparts = g->windowed ? g->units : g->parts;
    const int nl = g->windowed ? (int)g->units.size() - 1 : g->nb;
    drop_iteration_graph(g);
    const bool dev_lists = !g->windowed && nv >= g_device_lists_min.load();
    PhaseBins pb;
    if (dev_lists) NIDO_TRY(device_bins(g, pb, keep));
    else build_phase_bins(off.data(), nv, nl, G, lparts, g->gparts, g->tuning, pb, keep);
    const auto& binv = pb.binv;
    const auto& bitems = pb.bitems;
    const auto& chunks = pb.chunks;
    const auto& poff = pb.poff;
    const auto& toff = pb.toff;
    const auto& item_base = pb.item_base;
    const auto& poff_base = pb.poff_base;
    const auto& toff_base = pb.toff_base;
    const auto& chunk_base = pb.chunk_base;
    const auto& small_start = pb.small_start;
    const auto& small_n = pb.small_n;
    const int64_t tab_need = pb.tab_need, part_need = pb.part_need, slot_need = pb.slot_need,
                  key_need = pb.key_need, cpo_need = pb.cpo_need;
    g->short_bytes = pb.short_bytes;
    g->long_bytes = pb.long_bytes;
    g->blist0 = pb.blist0;
    g->blistn = pb.blistn;
    if (!dev_lists) NIDO_TRY(upload(&g->d_binv, binv, g->stream));
    NIDO_TRY(upload(&g->d_big_items, bitems, g->stream));
    NIDO_TRY(upload(&g->d_key_chunks, chunks, g->stream));
    if (key_need > g->key_cap) {
        dev_free(g->d_bkey);
        dev_free(g->d_bw);
        NIDO_TRY(dev_alloc(&g->d_bkey, key_need));
        NIDO_TRY(dev_alloc(&g->d_bw, key_need));
        g->key_cap = key_need;
    }
    if (cpo_need > g->cpo_cap) {
        dev_free(g->d_cpo);
        NIDO_TRY(dev_alloc(&g->d_cpo, cpo_need));
        g->cpo_cap = cpo_need;
    }
    NIDO_TRY(upload(&g->d_big_poff, poff, g->stream));
    NIDO_TRY(upload(&g->d_big_toff, toff, g->stream));
    if (tab_need > g->tab_cap) {
        dev_free(g->d_tkey);
        dev_free(g->d_tval);
        NIDO_TRY(dev_alloc(&g->d_tkey, tab_need));
        NIDO_TRY(dev_alloc(&g->d_tval, tab_need));
        g->tab_cap = tab_need;
    }
    if (part_need > g->part_cap) {
        dev_free(g->d_pg);
        dev_free(g->d_pc);
        dev_free(g->d_pself);
        NIDO_TRY(dev_alloc(&g->d_pg, part_need));
        NIDO_TRY(dev_alloc(&g->d_pc, part_need));
        NIDO_TRY(dev_alloc(&g->d_pself, part_need));
        g->part_cap = part_need;
    }
    if (slot_need > g->slot_cap) {
        dev_free(g->d_cnt);
        NIDO_TRY(dev_alloc(&g->d_cnt, slot_need));
        launch_fill_i32(g->d_cnt, 0, slot_need, g->stream);   // the kernels keep them zero afterwards
        g->slot_cap = slot_need;
    }
    const size_t nbig_items = bitems.size();
    g->small_bins.assign((size_t)nl, SmallBins{});
    g->big_args.assign((size_t)nl, BigArgs{});
    for (int b = 0; b < nl; ++b) {
        SmallBins& L = g->small_bins[(size_t)b];
        for (int k = 0; k < NUM_SMALL_BINS; ++k) {
            L.verts[k] = g->d_binv + small_start[(size_t)b * NUM_SMALL_BINS + k];
            L.n[k] = small_n[(size_t)b * NUM_SMALL_BINS + k];
        }
        L.chunks = g->d_key_chunks + chunk_base[(size_t)b];
        L.nchunks = chunk_base[(size_t)b + 1] - chunk_base[(size_t)b];
        L.bkey = g->d_bkey;
        L.bw = g->d_bw;
        L.cpo = g->d_cpo;
        if (b > 0 && !g->windowed) {   // the previous batch's modularity delta rides in this launch
            L.dblocks = kDeltaBlocks;
            L.db = b - 1;
            L.dv0 = G == 1 ? g->parts[(size_t)b - 1] : 0;
            L.dv1 = G == 1 ? g->parts[(size_t)b] : 0;
            L.D = g->delta();
            L.Cn64 = g->d_Cn;
        }
        small_bins_blocks(L, 4096);
        BigArgs& B = g->big_args[(size_t)b];
        const int64_t i0 = item_base[(size_t)b];
        const int64_t i1 = b + 1 < nl ? item_base[(size_t)b + 1] : (int64_t)nbig_items;
        B.items = g->d_big_items + i0;
        B.nitems = i1 - i0;
        B.bkey = g->d_bkey;
        B.bw = g->d_bw;
        B.cpo = g->d_cpo;
        B.key_limit = g->tuning.big_key_limit;
        B.slots = g->tuning.big_slots;
        B.poff = g->d_big_poff + poff_base[(size_t)b];
        B.pg = g->d_pg;
        B.pc = g->d_pc;
        B.pself = g->d_pself;
        B.cnt = g->d_cnt;
        B.toff = g->d_big_toff + toff_base[(size_t)b];
        B.tkey = g->d_tkey;
        B.tval = g->d_tval;
    }

    for (auto& B : g->big_args) B.narrow = g->narrow;
    for (auto& L : g->small_bins) L.twom_d = g->twom_d;
    g->listed = 0;
    for (int i = 0; i < G; ++i)
        if (!keep || keep[i]) g->listed += G == 1 ? nv : g->vb[(size_t)i + 1] - g->vb[(size_t)i];
    return NIDO_OK;
}

// Per-phase set-up: maximum order, batches, degree bins, the row slices of
// the degree/modularity sweeps, k_v and 2m (GraphGPU ctor steps
// graph_gpu.cpp:94-139 and aggregation steps graph_gpu.cpp:1487-1516).
int setup_phase(nido_graph* g)
{
    const int64_t nv = g->nv;
    const auto& off = g->h_off;
    drop_iteration_graph(g);
    g->inc_ok = false;
    g->ident = false;
    g->maxdeg = max_degree(off.data(), nv);
    g->committed.assign((size_t)g->nb, 0);
    // Batch b of the handle = batch b of every graph it holds (each graph
    // cut by its own rule, graph_gpu.cpp:303-339).
    const int G = g->G;
    if (G == 1) batch_bounds(off, nv, g->nb, g->parts);
    else batch_bounds_all(off.data(), G, g->vb.data(), g->nb, g->gparts);
    // A phase with more arcs than the device window streams them: every
    // batch is cut into units of rows whose arcs fit the window, and each
    // unit is one local-moving launch pair over the arcs copied in for it.
    g->windowed = G == 1 && g->window > 0 && g->ne > g->window;
    g->win_e0 = g->win_e1 = -1;
    if (g->windowed) NIDO_TRY(window_units(off.data(), g->parts, g->nb, g->window, g->units, g->first_unit));
    // A windowed phase whose batches each fit the window at once can update
    // Q incrementally: a batch's arcs are still in the window when its commit
    // has listed the rows that moved.
    g->win_single = g->windowed;
    for (int b = 0; g->win_single && b < g->nb; ++b)
        g->win_single = g->first_unit[(size_t)b + 1] - g->first_unit[(size_t)b] <= 1;
    // Two batches at once: batch b + 1 is copied into one half of the window
    // (on the side stream) while batch b's launches read the other.
    g->win_double = g->win_single;
    for (int b = 0; g->win_double && b < g->nb; ++b)
        g->win_double = 2 * (off[(size_t)g->parts[(size_t)b + 1]] - off[(size_t)g->parts[(size_t)b]]) <= g->window;
    const bool dev_lists = !g->windowed && nv >= g_device_lists_min.load();
    if (!g->windowed || g->win_single) {
        // incremental modularity: room for every row chunk of the largest batch
        int64_t need = 0;
        if (G == 1) {
            for (int b = 0; b < g->nb; ++b) {
                int64_t n = 0;
                for (int64_t v = g->parts[(size_t)b]; v < g->parts[(size_t)b + 1]; ++v)
                    n += (off[v + 1] - off[v] + kDeltaArcs - 1) / kDeltaArcs;
                need = std::max(need, n);
            }
        } else {
            // ceil(d / K) <= 1 + d / K: rows + arcs / K of the batch bound it
            const size_t nb1 = (size_t)g->nb + 1;
            for (int b = 0; b < g->nb; ++b) {
                int64_t rows = 0, arcs = 0;
                for (int i = 0; i < G; ++i) {
                    const int64_t lo = g->gparts[(size_t)i * nb1 + (size_t)b];
                    const int64_t hi = g->gparts[(size_t)i * nb1 + (size_t)b + 1];
                    rows += hi - lo;
                    arcs += off[hi] - off[lo];
                }
                need = std::max(need, rows + arcs / kDeltaArcs);
            }
            if ((int64_t)g->gparts.size() > g->gp_cap) {
                dev_free(g->d_gp);
                NIDO_TRY(dev_alloc(&g->d_gp, (int64_t)g->gparts.size()));
                g->gp_cap = (int64_t)g->gparts.size();
            }
            HIP_TRY(hipMemcpyAsync(g->d_gp, g->gparts.data(), sizeof(int64_t) * g->gparts.size(),
                                   hipMemcpyHostToDevice, g->stream));
        }
        if (need > g->dchunk_cap) {
            dev_free(g->d_dchunks);
            NIDO_TRY(dev_alloc(&g->d_dchunks, need));
            g->dchunk_cap = need;
        }
        dev_free(g->d_dcounts);
        NIDO_TRY(dev_alloc(&g->d_dcounts, 2 * (int64_t)g->nb));
        launch_fill_i32(g->d_dcounts, 0, 2 * (int64_t)g->nb, g->stream);
        launch_fill_i32(g->d_touched, 0, g->n0, g->stream);
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    // after the delta buffers above: the short-row launches keep their
    // addresses (g->delta())
    NIDO_TRY(build_lists(g, nullptr));
    fx_t twom = 0;
    unsigned long long maxk_bits = 0;
    if (g->windowed) {
        // the arcs are on the host: k_v, 2m and max |k_v| from there
        std::vector<fx_t> hk((size_t)nv);
        std::vector<double> hkd((size_t)nv);
        double maxk_d = 0;
        host_vertex_weights(nv, off.data(), g->h_w, hk.data(), hkd.data(), &twom, &maxk_d);
        maxk_bits = nido_dbits(maxk_d);
        if (nv > 0) {
            HIP_TRY(hipMemcpyAsync(g->d_k, hk.data(), sizeof(fx_t) * (size_t)nv, hipMemcpyHostToDevice, g->stream));
            HIP_TRY(hipMemcpyAsync(g->d_kd, hkd.data(), sizeof(double) * (size_t)nv, hipMemcpyHostToDevice,
                                   g->stream));
        }
        HIP_TRY(hipStreamSynchronize(g->stream));   // hk / hkd go out of scope
    } else {
        if (dev_lists) {
            NIDO_TRY(device_row_lists(g));
        } else {
            std::vector<int> small;
            std::vector<RowItem> items;
            build_row_lists(off.data(), nv, small, items);
            g->nsmall = (int64_t)small.size();
            g->nitems = (int64_t)items.size();
            NIDO_TRY(upload(&g->d_small, small, g->stream));
            NIDO_TRY(upload(&g->d_items, items, g->stream));
        }

        launch_narrow_ids(g->d_col, g->d_col32, g->ne, g->stream);
        unsigned long long* d_maxk = reinterpret_cast<unsigned long long*>(g->d_tot + 1);
        launch_vertex_weights(g->d_off, g->d_w, g->d_small, g->nsmall, g->d_items, g->nitems, g->d_k,
                              g->d_kd, nv, g->d_acc + 2, d_maxk, g->d_part, g->d_src, g->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&twom, g->d_acc + 2, sizeof(fx_t), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(&maxk_bits, d_maxk, sizeof(maxk_bits), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    g->twom = twom;
    g->twom_d = fx_to_double(twom);
    g->m = g->twom_d * 0.5;                     // mass_ = 0.5 * sum(k), graph_cuda.cu:1855
    // Long-row tables may use 32-bit high words when no partial sum of a row
    // can reach 2^31: weights non-negative and every k_v < 2^31 (a rounded
    // k_v below 2^31 - 2^8 leaves room for the rounding).
    const double maxk = nido_bitsd(maxk_bits);
    g->narrow = g->nonneg && maxk < 2147483648.0 - 256.0;
    for (auto& B : g->big_args) B.narrow = g->narrow;
    for (auto& L : g->small_bins) L.twom_d = g->twom_d;
    // The short rows sum in 64-bit fixed point when every weight is a
    // multiple of 2^-F (F <= 62) and no short row's sum can reach 2^63 in
    // units of 2^-F: the sums are then the same exact values as in Q64.64
    // (scaled by 2^(64 - F)), one word and no carry per LDS update.
    g->fixed = false;
    if (g->tuning.fixed_short && g->ne > 0 && g->twom != 0 && !g->windowed) {
        int* d_fneed = reinterpret_cast<int*>(g->d_tot);
        unsigned long long* d_wmax = reinterpret_cast<unsigned long long*>(g->d_tot + 1);
        launch_weight_grid(g->d_w, g->ne, d_fneed, d_wmax, g->stream);
        HIP_TRY(hipGetLastError());
        int fneed = 0;
        unsigned long long wbits = 0;
        HIP_TRY(hipMemcpyAsync(&fneed, d_fneed, sizeof(fneed), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipMemcpyAsync(&wbits, d_wmax, sizeof(wbits), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        const int F = std::max(fneed, 0);
        const double wq_max = std::ldexp(nido_bitsd(wbits), F);   // exact: a power-of-two scaling
        const uint64_t dmax = (uint64_t)std::max<int64_t>(g->tuning.big_min_deg, kThreadMaxDeg);
        if (fneed <= 62 && wq_max < 0x1p62 &&
            (unsigned __int128)(uint64_t)wq_max * dmax < ((unsigned __int128)1 << 63)) {
            if (g->ne > g->wq_cap) {
                dev_free(g->d_wq);
                g->wq_cap = 0;
                NIDO_TRY(dev_alloc(&g->d_wq, g->ne));
                g->wq_cap = g->ne;
            }
            launch_to_fixed(g->d_w, g->d_wq, g->ne, std::ldexp(1.0, F), g->stream);
            HIP_TRY(hipGetLastError());
            g->fixed = true;
            g->wscale = std::ldexp(1.0, -F);
        }
    }
    if (G > 1) {
        // per-graph modularity items: arc and vertex chunks of each graph
        std::vector<SegItem> seg;
        for (int gg = 0; gg < G; ++gg) {
            const int64_t v0 = g->vb[(size_t)gg], v1 = g->vb[(size_t)gg + 1];
            const int64_t e0 = off[v0], e1 = off[v1];
            const int64_t n = std::max((e1 - e0 + kSegArcs - 1) / kSegArcs, (v1 - v0 + kSegVerts - 1) / kSegVerts);
            for (int64_t i = 0; i < n; ++i)
                seg.push_back(SegItem{gg, std::min(e1, e0 + i * kSegArcs), std::min(e1, e0 + (i + 1) * kSegArcs),
                                      std::min(v1, v0 + i * kSegVerts), std::min(v1, v0 + (i + 1) * kSegVerts)});
        }
        g->nseg = (int64_t)seg.size();
        NIDO_TRY(upload(&g->d_seg, seg, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    return NIDO_OK;
}

// The GraphGPU method set works on handles that hold one graph.
int single_graph(const nido_graph* g)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (g->G != 1) return set_error(NIDO_ERR_STATE, "this handle holds %d graphs: use nido_batch_run", g->G);
    return NIDO_OK;
}

int check_batch(const nido_graph* g, int batch)
{
    NIDO_TRY(single_graph(g));
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
    g->inc_ok = false;
    g->ident = false;
    g->iters_since_singleton = 0;
    g->last_moved = -1;
    launch_singleton(g->d_C, g->d_Cn, g->d_C32, g->d_a, g->d_k, g->nv, g->stream);
    HIP_TRY(hipGetLastError());
    g->ident = true;
    std::fill(g->committed.begin(), g->committed.end(), 0);
    return NIDO_OK;
}

int window_load(nido_graph* g, int64_t v0, int64_t v1, int64_t* e0);

// Q of a windowed phase into d_q: the internal weight window by window (the
// arcs of each unit copied in), then the square terms.
int windowed_modularity(nido_graph* g)
{
    launch_fill_fx(g->d_acc, 2, g->stream);
    for (size_t u = 0; u + 1 < g->units.size(); ++u) {
        int64_t e0 = 0;
        NIDO_TRY(window_load(g, g->units[u], g->units[u + 1], &e0));
        launch_modularity_rows(g->d_off, g->d_col32 - e0, g->d_w - e0, g->d_C32, g->units[u], g->units[u + 1],
                               g->h_off[(size_t)g->units[u + 1]] - e0, g->d_acc, g->stream);
    }
    launch_modularity_terms(g->d_a, g->nv, g->twom_d, g->d_acc, g->d_q, g->stream, g->d_term);
    HIP_TRY(hipGetLastError());
    // the pair the incremental update continues from
    HIP_TRY(hipMemcpyAsync(g->d_iacc, g->d_acc, 2 * sizeof(fx_t), hipMemcpyDeviceToDevice, g->stream));
    return NIDO_OK;
}

int do_modularity(nido_graph* g, double* q)
{
    if (g->twom == 0) {   // no weight at all: nothing is internal, nothing moves
        *q = 0.0;
        return NIDO_OK;
    }
    {
        ProfScope ps(g->prof, g->stream, NIDO_K_MODULARITY, 24.0 * (double)g->ne + 24.0 * (double)g->nv);
        if (g->windowed)
            NIDO_TRY(windowed_modularity(g));
        else
            launch_modularity(g->d_src, g->d_col32, g->d_w, g->d_C32, g->ne, g->d_a, g->nv, g->twom_d,
                              g->d_part, g->d_q, g->stream, g->G == 1 ? g->d_term : nullptr,
                              g->G == 1 ? g->d_iacc : nullptr, g->ident);
    }
    g->inc_ok = g->G == 1 && (!g->windowed || g->win_single);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g->h_q, g->d_q, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->prof.harvest();
    *q = *g->h_q;
    return NIDO_OK;
}

bool replaying(const nido_graph* g);
int lm_launch(nido_graph* g, const LmArgs& A, int u);

// Copies the arcs of rows [v0, v1) of a windowed phase into the window
// buffers (unless they are there already); *e0 = their first CSR index.
int window_load(nido_graph* g, int64_t v0, int64_t v1, int64_t* e0)
{
    const int64_t a = g->h_off[(size_t)v0], z = g->h_off[(size_t)v1];
    *e0 = a;
    if (a >= g->win_e0 && z <= g->win_e1) return NIDO_OK;
    if (z - a > g->window) return set_error(NIDO_ERR_CAPACITY, "window of %lld arcs too small for %lld",
                                            (long long)g->window, (long long)(z - a));
    if (z > a) {
        HIP_TRY(hipMemcpyAsync(g->d_col32, g->h_col32 + a, sizeof(int) * (size_t)(z - a), hipMemcpyHostToDevice,
                               g->stream));
        HIP_TRY(hipMemcpyAsync(g->d_w, g->h_w + a, sizeof(double) * (size_t)(z - a), hipMemcpyHostToDevice,
                               g->stream));
    }
    g->win_e0 = a;
    g->win_e1 = z;
    return NIDO_OK;
}

// Offset of the window half that batch b's arcs use while an iteration
// streams its batches through the two halves.
int64_t win_half(const nido_graph* g, int b)
{
    return (int64_t)(b & 1) * (g->window / 2);
}

// Copies batch b's arcs into its half of the window on the side stream, once
// the half's previous batch is done with it; ev_copied marks the copy done.
int window_prefetch(nido_graph* g, int b)
{
    const int h = b & 1;
    const int64_t a = g->h_off[(size_t)g->parts[(size_t)b]], z = g->h_off[(size_t)g->parts[(size_t)b + 1]];
    if (2 * (z - a) > g->window) return set_error(NIDO_ERR_CAPACITY, "batch %d does not fit half the window", b);
    HIP_TRY(hipStreamWaitEvent(g->side, g->ev_free[h], 0));
    if (z > a) {
        HIP_TRY(hipMemcpyAsync(g->d_col32 + win_half(g, b), g->h_col32 + a, sizeof(int) * (size_t)(z - a),
                               hipMemcpyHostToDevice, g->side));
        HIP_TRY(hipMemcpyAsync(g->d_w + win_half(g, b), g->h_w + a, sizeof(double) * (size_t)(z - a),
                               hipMemcpyHostToDevice, g->side));
    }
    HIP_TRY(hipEventRecord(g->ev_copied[h], g->side));
    return NIDO_OK;
}

int do_lm(nido_graph* g, int b)
{
    g->committed[(size_t)b] = 0;
    LmArgs A{g->d_off, g->d_col32, g->d_w, g->d_C32, g->d_Cn, g->d_kd, g->d_k, g->d_a, g->m, 1.0 / g->m};
    A.wq = g->fixed ? g->d_wq : nullptr;
    A.wscale = g->wscale;
    if (g->G == 1) {
        const int64_t v0 = g->parts[(size_t)b], v1 = g->parts[(size_t)b + 1];
        // an empty batch still carries the previous batch's modularity delta
        if (v1 <= v0 && g->small_bins[(size_t)b].dblocks == 0) return NIDO_OK;
        if (g->twom == 0) {
            if (v1 > v0)
                HIP_TRY(hipMemcpyAsync(g->d_Cn + v0, g->d_C + v0, sizeof(int64_t) * (size_t)(v1 - v0),
                                       hipMemcpyDeviceToDevice, g->stream));
            return NIDO_OK;
        }
    } else {
        // graphs without weight never take part (their act flag stays 0);
        // an empty batch still carries the previous batch's modularity delta
        if (g->blistn[(size_t)b] == 0 && g->small_bins[(size_t)b].dblocks == 0) return NIDO_OK;
        A.gid = g->d_gid;
        A.mg = g->d_mg;
        A.act = g->d_act;
    }
    if (g->windowed) {
        // each unit of the batch: its arcs copied into the window, then its
        // launches, with the arc arrays based so that CSR indices address
        // the window (the unit's rows use arcs [E0, E1) only)
        for (int u = g->first_unit[(size_t)b]; u < g->first_unit[(size_t)b + 1]; ++u) {
            int64_t e0 = 0, at = 0;
            if (g->win_live) {   // the batch (its one unit) was copied into its half
                e0 = g->h_off[(size_t)g->units[(size_t)u]];
                at = win_half(g, b);
            } else {
                NIDO_TRY(window_load(g, g->units[(size_t)u], g->units[(size_t)u + 1], &e0));
            }
            LmArgs Aw = A;
            Aw.col = g->d_col32 + at - e0;
            Aw.w = g->d_w + at - e0;
            NIDO_TRY(lm_launch(g, Aw, u));
        }
        return NIDO_OK;
    }
    return lm_launch(g, A, b);
}

// The local-moving launches of launch unit u (a batch, or one window unit
// of a batch).
int lm_launch(nido_graph* g, const LmArgs& A, int u)
{
    const int b = u;
    const SmallBins& L = g->small_bins[(size_t)b];
    const BigArgs& B = g->big_args[(size_t)b];
    const bool split = g->tuning.split_small != 0;
    // The second stream pays off only inside a replayed graph, where the
    // fork and join are graph edges (from the stream: -7% on R-MAT s22).
    if (B.nitems > 0 && g->tuning.overlap_long && replaying(g)) {
        // fork: the long rows (key fill, then partitions) run on the side
        // stream while the short rows run on the main stream; join before
        // the commit
        HIP_TRY(hipEventRecord(g->ev_fork, g->stream));
        HIP_TRY(hipStreamWaitEvent(g->side, g->ev_fork, 0));
        {
            ProfScope ps(g->prof, g->side, NIDO_K_LM_LONG, g->long_bytes[(size_t)b]);
            launch_lm_small(A, L, false, true, split, g->side);
            launch_lm_big(A, B, g->side);
        }
        HIP_TRY(hipEventRecord(g->ev_join, g->side));
        {
            ProfScope ps(g->prof, g->stream, NIDO_K_LM_SHORT, g->short_bytes[(size_t)b]);
            launch_lm_small(A, L, true, false, split, g->stream, g->ride_delta);
        }
        HIP_TRY(hipStreamWaitEvent(g->stream, g->ev_join, 0));
    } else {
        // launch 1 fills the long rows' buckets that launch 2 reads
        {
            ProfScope ps(g->prof, g->stream, NIDO_K_LM_SHORT, g->short_bytes[(size_t)b]);
            launch_lm_small(A, L, true, true, split, g->stream, g->ride_delta);
        }
        if (B.nitems > 0) {
            ProfScope ps(g->prof, g->stream, NIDO_K_LM_LONG, g->long_bytes[(size_t)b]);
            launch_lm_big(A, B, g->stream);
        }
    }
    HIP_TRY(hipGetLastError());
    return NIDO_OK;
}

int do_commit(nido_graph* g, int b)
{
    g->inc_ok = false;
    g->ident = false;
    {
        ProfScope ps(g->prof, g->stream, NIDO_K_COMMIT, 0.0);
        if (g->G == 1)
            launch_commit(g->parts[(size_t)b], g->parts[(size_t)b + 1], g->d_C, g->d_Cn, g->d_C32, g->d_k,
                          g->d_a, g->stream, g->d_off, g->moved());
        else
            launch_commit_filtered(g->d_binv + g->blist0[(size_t)b], 0, g->blistn[(size_t)b], g->d_C, g->d_Cn,
                                   g->d_C32, g->d_k, g->d_a, g->d_gid, g->d_act, g->stream, g->moved());
    }
    HIP_TRY(hipGetLastError());
    g->committed[(size_t)b] = 1;
    return NIDO_OK;
}

int do_restore(nido_graph* g)
{
    g->inc_ok = false;
    g->ident = false;
    launch_commit(0, g->nv, g->d_C, g->d_Cn, g->d_C32, g->d_k, g->d_a, g->stream);
    HIP_TRY(hipGetLastError());
    return NIDO_OK;
}

// The loop body of LouvainGPU::run, louvain_gpu.cpp:55-90; *qt receives
// the modularity evaluated after the batches (before any rollback).
// The launches of one iteration: every batch's local moving and commit,
// then the modularity kernels (the value lands in d_q, or d_gq per graph).
// inc: Q by the exact incremental update (tracked commits + per-batch
// deltas) instead of a sweep over every arc.
int iteration_launches(nido_graph* g, bool inc)
{
    g->ident = false;
    const ModDelta D = g->delta();
    if (g->G > 1) launch_fill_i64(reinterpret_cast<int64_t*>(g->d_moved), 0, kMovedSlots, g->stream);
    g->ride_delta = inc;
    // A windowed phase whose batches fit half the window each streams them
    // through the two halves: batch b + 1's copy (side stream) runs beside
    // batch b's launches (main stream).
    struct LiveGuard {
        nido_graph* g;
        ~LiveGuard() { g->win_live = g->ride_delta = false; }
    } live_guard{g};
    g->win_live = g->win_double && g->nb > 0;
    if (g->win_live) {
        // the halves are free once the main stream's earlier work (which may
        // read the window) is done; window_load must copy afresh afterwards
        HIP_TRY(hipEventRecord(g->ev_free[0], g->stream));
        HIP_TRY(hipEventRecord(g->ev_free[1], g->stream));
        g->win_e0 = g->win_e1 = -1;
        NIDO_TRY(window_prefetch(g, 0));
    }
    for (int b = 0; b < g->nb; ++b) {
        if (g->win_live) {
            if (b + 1 < g->nb) NIDO_TRY(window_prefetch(g, b + 1));
            HIP_TRY(hipStreamWaitEvent(g->stream, g->ev_copied[b & 1], 0));
        }
        NIDO_TRY(do_lm(g, b));
        if (!inc) {
            NIDO_TRY(do_commit(g, b));
            if (g->win_live) HIP_TRY(hipEventRecord(g->ev_free[b & 1], g->stream));
            continue;
        }
        {
            ProfScope ps(g->prof, g->stream, NIDO_K_COMMIT, 0.0);
            if (g->G == 1)
                launch_commit_tracked(g->parts[(size_t)b], g->parts[(size_t)b + 1], b, g->d_C, g->d_Cn, g->d_C32,
                                      g->d_k, g->d_a, g->d_off, D, g->stream);
            else
                launch_commit_tracked_list(g->d_binv + g->blist0[(size_t)b], g->blistn[(size_t)b], b, g->d_C,
                                           g->d_Cn, g->d_C32, g->d_k, g->d_a, g->d_off, D, g->d_act, g->stream);
        }
        // A windowed phase (every batch one window load) takes the batch's
        // delta now, while its arcs are still in the window and before the
        // next batch's commit changes any community.
        if (g->windowed && g->parts[(size_t)b + 1] > g->parts[(size_t)b]) {
            const int64_t at = g->win_live ? win_half(g, b) : 0;
            const int64_t e0 = g->win_live ? g->h_off[(size_t)g->parts[(size_t)b]] : g->win_e0;
            launch_mod_delta(b, g->parts[(size_t)b], g->parts[(size_t)b + 1], g->d_Cn, g->d_C32,
                             g->d_col32 + at - e0, g->d_w + at - e0, g->d_a, g->twom_d, D, g->stream);
        }
        if (g->win_live) HIP_TRY(hipEventRecord(g->ev_free[b & 1], g->stream));
        g->committed[(size_t)b] = 1;
    }
    g->ride_delta = false;
    ProfScope ps(g->prof, g->stream, NIDO_K_MODULARITY, 24.0 * (double)g->ne + 24.0 * (double)g->nv);
    if (inc && g->nb > 0 && !g->windowed)   // the earlier batches' deltas rode in the next batch's short-row launch
        launch_mod_delta(g->nb - 1, g->G == 1 ? g->parts[(size_t)g->nb - 1] : 0,
                         g->G == 1 ? g->parts[(size_t)g->nb] : 0, g->d_Cn, g->d_C32, g->d_col32, g->d_w, g->d_a,
                         g->twom_d, D, g->stream);
    if (g->G > 1 && inc)    // the graphs that sat out this iteration keep their modularity
        launch_mod_finish_graphs(D, g->nb, g->G, g->d_gq, g->d_act, g->stream);
    else if (g->G > 1)
        launch_modularity_seg(g->d_seg, g->nseg, g->d_src, g->d_col32, g->d_w, g->d_C32, g->d_a, g->d_twomg,
                              g->G, g->d_gacc, g->d_gq, g->stream, g->d_act, g->d_term, g->d_moved);
    else if (g->twom != 0 && inc)
        launch_mod_finish(D, g->nb, g->twom_d, g->d_q, g->stream);
    else if (g->twom != 0 && g->windowed)
        NIDO_TRY(windowed_modularity(g));
    else if (g->twom != 0)
        launch_modularity(g->d_src, g->d_col32, g->d_w, g->d_C32, g->ne, g->d_a, g->nv, g->twom_d, g->d_part,
                          g->d_q, g->stream, g->d_term, g->d_iacc);
    HIP_TRY(hipGetLastError());
    return NIDO_OK;
}

constexpr int64_t kGraphMinArcs = int64_t(1) << 28;
constexpr int64_t kIncShare = 8;

// Whether this phase's iterations are replayed from a hipGraph.
bool replaying(const nido_graph* g)
{
    const int ug = g->tuning.use_graph;
    return !g->prof.on && !g->windowed && (g->G > 1 || g->twom != 0) &&
           (ug == 2 || (ug == 1 && g->ne >= kGraphMinArcs));
}

void drop_iteration_graph(nido_graph* g)
{
    for (auto& ex : g->iter_exec) {
        if (ex) (void)hipGraphExecDestroy(ex);
        ex = nullptr;
    }
}

int launch_or_replay(nido_graph* g, bool inc);

// With use_graph set (and no per-kernel events being recorded) the
// iteration's launches are captured into a hipGraph once per phase and
// replayed; the kernels' arguments do not change within a phase.
// use_graph = 1 does so for graphs of kGraphMinArcs arcs or more: a capture
// per phase costs more than it saves on small graphs (a serving loop's
// requests ran at 60% of their throughput), and a graph of 128 M arcs
// (R-MAT scale 22) ran 1.8% slower replayed with its long rows beside the
// short rows than from the stream, a graph of 1 G arcs (Chung-Lu 2^25)
// 0.9% faster; 2 always.
int run_iteration_launches(nido_graph* g)
{
    // Q by a sweep of every arc or by the exact incremental update over the
    // rows that moved (the same Q).  The first iteration from the singleton
    // partition moves most vertices (their rows hold ~70% of the arcs on
    // R-MAT) and sweeps; with full_modularity 0 a single graph takes the
    // update in every later iteration (its deltas ride beside the next
    // batch's rows), and a handle of many graphs when the iteration before
    // moved at most 1/kIncShare of the arcs: its batches are all short rows,
    // so the deltas' dependent gathers set the length of those launches and
    // cost several times the sweep's streaming per arc.  2: the update in
    // every later iteration; 1: never.  A windowed phase whose batches each
    // fit the window takes the update from its first iteration on: there a
    // sweep would stream every arc over the host link a second time, while
    // each batch's delta reads arcs already in the window.
    const bool any_w = g->G > 1 || g->twom != 0;
    const int fm = g->tuning.full_modularity;
    bool inc = any_w && g->symmetric && g->inc_ok && fm != 1 && (g->iters_since_singleton > 0 || g->windowed) &&
               (!g->windowed || g->win_single);
    if (inc && fm == 0 && g->G > 1) inc = g->last_moved >= 0 && g->last_moved * kIncShare <= g->ne;
    // the short-row launches must carry the delta buffers' current addresses
    for (const SmallBins& L : g->small_bins)
        if (L.dblocks && (L.D.counts != g->d_dcounts || L.D.chunks != g->d_dchunks || L.D.gp != g->d_gp))
            return set_error(NIDO_ERR_STATE, "work lists hold stale delta buffers");
    ++g->iters_since_singleton;
    NIDO_TRY(launch_or_replay(g, inc));
    g->inc_ok = any_w && (!g->windowed || g->win_single);   // both forms leave term/acc describing the new partition
    return NIDO_OK;
}

int launch_or_replay(nido_graph* g, bool inc)
{
    g->ident = false;
    if (!replaying(g)) return iteration_launches(g, inc);
    hipGraphExec_t& exec = g->iter_exec[inc ? 1 : 0];
    if (!exec) {
        HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        const int rc = iteration_launches(g, inc);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(g->stream, &graph);
        if (rc != NIDO_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(e);
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ei);
        exec = ex;
    }
    HIP_TRY(hipGraphLaunch(exec, g->stream));
    return NIDO_OK;
}

int do_iteration(nido_graph* g, double* q, double* dq, double* qt_out = nullptr)
{
    NIDO_TRY(run_iteration_launches(g));
    double qt = 0;
    if (g->twom != 0) {
        HIP_TRY(hipMemcpyAsync(g->h_q, g->d_q, sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        g->prof.harvest();
        qt = *g->h_q;
    }
    if (qt_out) *qt_out = qt;
    *dq = qt - *q;
    if (*dq < 0) NIDO_TRY(do_restore(g));
    else *q = qt;
    return NIDO_OK;
}

// The coarse graph of communities (graph_gpu.cpp:1396-1516): coarse vertex
// newid[c] for every community c, arcs merged by target with exact sums.
// With gid set, only the graphs g with keep[g] set are carried over.
int coarsen(nido_graph* g, const int64_t* newid, int64_t K, const int* gid, const unsigned char* keep)
{
    hipStream_t s = g->stream;
    const int64_t nv = g->nv;
    int64_t *arcs = nullptr, *toff = nullptr, *cnt = nullptr, *noff = nullptr, *ncol = nullptr;
    int64_t *cpos = nullptr, *mstart = nullptr;
    unsigned long long* mfill = nullptr;
    int *members = nullptr, *ckey = nullptr, *hlist = nullptr;
    double *nw = nullptr, *cw = nullptr;
    int* tkey = nullptr;
    fx_t* tval = nullptr;
    void* tmp = nullptr;
    auto cleanup = [&]() {
        dev_free(arcs); dev_free(toff); dev_free(cnt); dev_free(noff);
        dev_free(ncol); dev_free(nw); dev_free(tkey); dev_free(tval);
        dev_free(cpos); dev_free(mstart); dev_free(mfill); dev_free(members); dev_free(ckey); dev_free(cw);
        dev_free(hlist);
        raw_free(tmp);
        tmp = nullptr;
    };
    int rc = NIDO_OK;
    // a failed HIP call leaves the loop with the error, and cleanup runs
#define COARSEN_HIP(expr)                                                                                  \
    if (hipError_t e_ = (expr); e_ != hipSuccess) {                                                        \
        rc = set_error(NIDO_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
        break;                                                                                             \
    }
    do {
        if ((rc = dev_alloc(&arcs, K + 1)) || (rc = dev_alloc(&toff, K + 1)) ||
            (rc = dev_alloc(&cnt, K + 1)) || (rc = dev_alloc(&noff, K + 1)) ||
            (rc = dev_alloc(&cpos, K + 1)) || (rc = dev_alloc(&mstart, K + 1)) ||
            (rc = dev_alloc(&mfill, K + 1)) || (rc = dev_alloc(&members, std::max<int64_t>(nv, 1))) ||
            (rc = dev_alloc(&hlist, std::max<int64_t>(nv, 1))))
            break;
        if ((rc = raw_alloc(&tmp, scan_temp_bytes(K + 1)))) break;
        // per coarse row: its arcs and members; rows of few arcs are
        // aggregated in LDS (compact entries at cpos), the others in HBM
        // tables (toff)
        launch_fill_i64(arcs, 0, K + 1, s);
        launch_fill_i64(cnt, 0, K + 1, s);
        launch_fill_i64(reinterpret_cast<int64_t*>(mfill), 0, K + 1, s);   // [K]: HBM-row vertices
        launch_member_counts(g->d_C, newid, g->d_off, nv, cnt, arcs, s, gid, keep);
        launch_exclusive_scan(cnt, mstart, K, mstart + K, tmp, s);
        launch_coarse_sizes(arcs, K, toff, cpos, s);
        launch_exclusive_scan(toff, toff, K, toff + K, tmp, s);
        launch_exclusive_scan(cpos, cpos, K, cpos + K, tmp, s);
        int64_t T = 0, Tc = 0;
        COARSEN_HIP(hipMemcpyAsync(&T, toff + K, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        COARSEN_HIP(hipMemcpyAsync(&Tc, cpos + K, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        COARSEN_HIP(hipStreamSynchronize(s));
        if ((rc = dev_alloc(&tkey, std::max<int64_t>(T, 1))) || (rc = dev_alloc(&tval, std::max<int64_t>(T, 1))) ||
            (rc = dev_alloc(&ckey, std::max<int64_t>(Tc, 1))) || (rc = dev_alloc(&cw, std::max<int64_t>(Tc, 1))))
            break;
        if (T > 0) {
            launch_fill_i32(tkey, -1, T, s);
            launch_fill_fx(tval, T, s);
        }
        launch_member_scatter(g->d_C, newid, nv, mstart, mfill, members, toff, hlist, mfill + K, s, gid, keep);
        launch_coarse_lds(arcs, mstart, members, g->d_off, g->d_col, g->d_w, g->d_C, newid, K, cpos, ckey, cw, cnt, s);
        if (T > 0) {
            launch_coarse_insert(g->d_off, g->d_col, g->d_w, g->d_C, newid, hlist, mfill + K, nv, toff, tkey, tval, s);
            launch_coarse_count(toff, tkey, K, cnt, s);
        }
        launch_exclusive_scan(cnt, noff, K, noff + K, tmp, s);
        COARSEN_HIP(hipGetLastError());
        std::vector<int64_t> hoff((size_t)K + 1);
        COARSEN_HIP(hipMemcpyAsync(hoff.data(), noff, sizeof(int64_t) * (size_t)(K + 1), hipMemcpyDeviceToHost, s));
        COARSEN_HIP(hipStreamSynchronize(s));
        const int64_t nne = hoff[(size_t)K];
        if ((rc = dev_alloc(&ncol, nne)) || (rc = dev_alloc(&nw, nne))) break;
        launch_coarse_emit(toff, tkey, tval, K, noff, ncol, nw, cpos, ckey, cw, s);
        COARSEN_HIP(hipGetLastError());
        COARSEN_HIP(hipStreamSynchronize(s));
        free_graph_arrays(g);
        g->d_off = noff; noff = nullptr;
        g->d_col = ncol; ncol = nullptr;
        g->d_w = nw; nw = nullptr;
        g->nv = K;
        g->ne = nne;
        g->h_off.swap(hoff);
    } while (0);
#undef COARSEN_HIP
    cleanup();
    return rc;
}

// The coarse graph of a windowed phase, built on the host from its host
// CSR (as the reference builds every coarse graph, graph_gpu.cpp:1088-1304):
// it stays on the host while it has more arcs than the window, otherwise it
// moves into the device arrays.
int coarsen_host(nido_graph* g, const int64_t* newid, int64_t K)
{
    const int64_t nv = g->nv;
    std::vector<int64_t> hC((size_t)nv), hn((size_t)nv);
    NIDO_TRY(d2h(g, hC.data(), g->d_C, nv));
    NIDO_TRY(d2h(g, hn.data(), newid, nv));
    std::vector<int64_t> noff;
    std::vector<int> ncol;
    std::vector<double> nw;
    host_coarsen(nv, g->h_off.data(), g->h_col32, g->h_w, hC.data(), hn.data(), K, noff, ncol, nw);
    const int64_t nne = noff[(size_t)K];
    hipStream_t s = g->stream;
    HIP_TRY(hipMemcpyAsync(g->d_off, noff.data(), sizeof(int64_t) * (size_t)(K + 1), hipMemcpyHostToDevice, s));
    if (nne > g->window) {
        std::copy(ncol.begin(), ncol.end(), g->h_col32);   // fits: a coarse graph has at most the arcs of the fine one
        std::copy(nw.begin(), nw.end(), g->h_w);
    } else if (nne > 0) {
        HIP_TRY(hipMemcpyAsync(g->d_col32, ncol.data(), sizeof(int) * (size_t)nne, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(g->d_w, nw.data(), sizeof(double) * (size_t)nne, hipMemcpyHostToDevice, s));
        launch_widen_ids(g->d_col32, g->d_col, nne, s);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s));   // the host vectors go out of scope
    g->nv = K;
    g->ne = nne;
    g->h_off.swap(noff);
    return NIDO_OK;
}

// aggregation, graph_gpu.cpp:1382-1545, on the device.
int do_aggregation(nido_graph* g, int* done)
{
    g->ident = false;
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
    if (g->windowed) NIDO_TRY(coarsen_host(g, newid, K));
    else NIDO_TRY(coarsen(g, newid, K, nullptr, nullptr));
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
            double Qtmp = 0;
            NIDO_TRY(do_iteration(g, &Q, &dQ, &Qtmp));
            h.iter_q.push_back(Qtmp);
            loops++;
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
                HIP_TRY(hipMemcpyAsync(h.comm.data(), g->d_C, sizeof(int64_t) * (size_t)g->nv, hipMemcpyDeviceToHost,
                                       g->stream));
            HIP_TRY(hipStreamSynchronize(g->stream));
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

// ---- many graphs in one handle ---------------------------------------------
// Every graph follows exactly the steps LouvainGPU::run would take on it
// alone (louvain_gpu.cpp:6-139); the graphs only share launches.  A graph
// whose iteration loop has ended sits out (its act flag is 0) until the
// phase ends for all; a graph whose run has ended keeps no vertices in the
// next phase's coarse graph.

int batch_modularity(nido_graph* g, std::vector<double>& q)
{
    {
        ProfScope ps(g->prof, g->stream, NIDO_K_MODULARITY, 24.0 * (double)g->ne + 24.0 * (double)g->nv);
        launch_modularity_seg(g->d_seg, g->nseg, g->d_src, g->d_col32, g->d_w, g->d_C32, g->d_a, g->d_twomg,
                              g->G, g->d_gacc, g->d_gq, g->stream);
    }
    HIP_TRY(hipGetLastError());
    q.resize((size_t)g->G);
    HIP_TRY(hipMemcpyAsync(q.data(), g->d_gq, sizeof(double) * (size_t)g->G, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    g->prof.harvest();
    return NIDO_OK;
}

int upload_flags(nido_graph* g, unsigned char* d, const std::vector<unsigned char>& h)
{
    HIP_TRY(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));   // h may change right after
    return NIDO_OK;
}

// Renumbers the communities of the graphs with keep[g] set (ascending, per
// graph contiguous because graphs hold contiguous vertex ranges); returns
// the new vertex bases in nvb (G+1).
int batch_renumber(nido_graph* g, const std::vector<unsigned char>& keep, std::vector<int64_t>& nvb)
{
    hipStream_t s = g->stream;
    int64_t* newid = g->d_work;
    NIDO_TRY(upload_flags(g, g->d_keep, keep));
    launch_mark_communities(g->d_C, g->nv, newid, s, g->d_gid, g->d_keep);
    launch_exclusive_scan(newid, newid, g->nv, g->d_tot, g->d_scan_tmp, s);
    launch_gather_bases(newid, g->d_tot, g->d_vb, g->G, g->nv, g->d_vbn, s);
    HIP_TRY(hipGetLastError());
    nvb.resize((size_t)g->G + 1);
    HIP_TRY(hipMemcpyAsync(nvb.data(), g->d_vbn, sizeof(int64_t) * nvb.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return NIDO_OK;
}

int do_batch_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, nido_run_stats* st)
{
    const int G = g->G;
    const double t0 = now_s();
    g->ident = false;
    g->rebuilds = 0;
    g->bres.assign((size_t)G, BatchResult{});
    std::vector<unsigned char> alive((size_t)G, 1), act((size_t)G), flag((size_t)G);
    std::vector<int64_t> base((size_t)G, 0);   // subtracted from clus to give graph-local ids
    std::vector<double> Q((size_t)G), Q0((size_t)G), dQ((size_t)G), qt;
    std::vector<int64_t> loops((size_t)G);
    nido_run_stats S{};
    int nphase = 0;
    auto any = [](const std::vector<unsigned char>& f) {
        for (unsigned char x : f)
            if (x) return true;
        return false;
    };
    while (any(alive) && nphase < NIDO_MAX_PHASES) {
        NIDO_TRY(do_singleton(g));
        NIDO_TRY(batch_modularity(g, Q));
        for (int i = 0; i < G; ++i) {
            Q0[(size_t)i] = Q[(size_t)i];
            dQ[(size_t)i] = 1.0e308;
            loops[(size_t)i] = 0;
            act[(size_t)i] = alive[(size_t)i] && g->twom_g[(size_t)i] != 0;
            if (alive[(size_t)i] && !act[(size_t)i]) {
                // no weight: passes with no moves and Q = 0, so dQ = 0 after
                // the first; the loop (louvain_gpu.cpp:53) goes on to
                // max_loops only when tol < 0
                loops[(size_t)i] = max_loops <= 0 ? 0 : tol < 0.0 ? max_loops : 1;
                Q[(size_t)i] = 0.0;
                dQ[(size_t)i] = 0.0;
            }
        }
        const double ts = now_s();
        int64_t iters = 0;
        while (any(act)) {
            NIDO_TRY(upload_flags(g, g->d_act, act));
            NIDO_TRY(run_iteration_launches(g));
            qt.resize((size_t)G + 1);
            HIP_TRY(hipMemcpyAsync(qt.data(), g->d_gq, sizeof(double) * ((size_t)G + 1), hipMemcpyDeviceToHost,
                                   g->stream));
            HIP_TRY(hipStreamSynchronize(g->stream));
            g->prof.harvest();
            std::memcpy(&g->last_moved, &qt[(size_t)G], sizeof(int64_t));
            bool rollback = false;
            for (int i = 0; i < G; ++i) {
                flag[(size_t)i] = 0;
                if (!act[(size_t)i]) continue;
                dQ[(size_t)i] = qt[(size_t)i] - Q[(size_t)i];
                loops[(size_t)i]++;
                if (dQ[(size_t)i] < 0) {
                    flag[(size_t)i] = 1;            // graph_gpu.cpp:822-842
                    rollback = true;
                } else {
                    Q[(size_t)i] = qt[(size_t)i];
                }
                act[(size_t)i] = tol < dQ[(size_t)i] && loops[(size_t)i] < max_loops;   // louvain_gpu.cpp:53
            }
            if (rollback) {
                NIDO_TRY(upload_flags(g, g->d_keep, flag));
                launch_commit_filtered(nullptr, 0, g->nv, g->d_C, g->d_Cn, g->d_C32, g->d_k, g->d_a, g->d_gid,
                                       g->d_keep, g->stream);
                HIP_TRY(hipGetLastError());
                // a graph that rolled back and goes on (tol < 0) needs its
                // terms and sums again: the next iteration sweeps
                for (int i = 0; i < G; ++i)
                    if (flag[(size_t)i] && act[(size_t)i]) g->inc_ok = false;
            }
            // Stopped graphs leave the work lists once the active ones hold
            // at most the set share of the listed vertices (their rows
            // would only be skipped; C' = C holds for them from here on).
            const int pct = g_list_rebuild_pct.load();
            if (G > 1 && pct > 0) {
                int64_t live = 0;
                for (int i = 0; i < G; ++i)
                    if (act[(size_t)i]) live += g->vb[(size_t)i + 1] - g->vb[(size_t)i];
                if (live > 0 && live < g->listed && live * 100 <= g->listed * pct) {
                    NIDO_TRY(build_lists(g, act.data()));
                    ++g->rebuilds;
                }
            }
            ++iters;
        }
        HIP_TRY(hipStreamSynchronize(g->stream));
        const double te = now_s();
        S.t_loop += te - ts;
        S.total_loops += iters;
        // phase end, per graph (louvain_gpu.cpp:107-118)
        std::vector<unsigned char> cand((size_t)G, 0);
        for (int i = 0; i < G; ++i) {
            if (!alive[(size_t)i]) continue;
            BatchResult& R = g->bres[(size_t)i];
            nido_phase_stats ps{};
            ps.nv = g->vb[(size_t)i + 1] - g->vb[(size_t)i];
            ps.ne = g->h_off[(size_t)g->vb[(size_t)i + 1]] - g->h_off[(size_t)g->vb[(size_t)i]];
            ps.q0 = Q0[(size_t)i];
            ps.q = Q[(size_t)i];
            ps.loops = loops[(size_t)i];
            ps.t_loop = te - ts;
            R.phases.push_back(ps);
            R.st.total_loops += loops[(size_t)i];
            R.st.q = Q[(size_t)i];
            R.st.phases = (int)R.phases.size();
            if (Q[(size_t)i] - Q0[(size_t)i] > tol_phase) {
                cand[(size_t)i] = 1;
            } else {
                alive[(size_t)i] = 0;                  // ends with the clustering of the phase before
                base[(size_t)i] = g->vb[(size_t)i];
                R.st.done = 1;
            }
        }
        ++nphase;
        if (!any(cand)) break;
        // aggregation of the candidates: renumber, update their clustering
        // (graph_gpu.cpp:1389), then stop those with K == N or K == 1
        // (graph_gpu.cpp:1391-1392) and coarsen the rest
        const double a0 = now_s();
        std::vector<int64_t> nvb;
        NIDO_TRY(batch_renumber(g, cand, nvb));
        std::vector<unsigned char> fin((size_t)G, 0), cont((size_t)G, 0);
        for (int i = 0; i < G; ++i) {
            if (!cand[(size_t)i]) continue;
            const int64_t K = nvb[(size_t)i + 1] - nvb[(size_t)i];
            const int64_t n = g->vb[(size_t)i + 1] - g->vb[(size_t)i];
            if (K == n || K == 1) {
                fin[(size_t)i] = 1;
                alive[(size_t)i] = 0;
                base[(size_t)i] = nvb[(size_t)i];
                g->bres[(size_t)i].st.done = 1;
            } else {
                cont[(size_t)i] = 1;
            }
        }
        if (any(fin)) {
            NIDO_TRY(upload_flags(g, g->d_keep, fin));
            launch_update_clustering(g->d_clus, g->n0, g->d_C, g->d_work, g->stream, g->d_gid0, g->d_keep);
            HIP_TRY(hipGetLastError());
        }
        if (any(cont)) {
            if (any(fin)) NIDO_TRY(batch_renumber(g, cont, nvb));
            NIDO_TRY(upload_flags(g, g->d_keep, cont));
            launch_update_clustering(g->d_clus, g->n0, g->d_C, g->d_work, g->stream, g->d_gid0, g->d_keep);
            HIP_TRY(hipGetLastError());
            NIDO_TRY(coarsen(g, g->d_work, nvb[(size_t)G], g->d_gid, g->d_keep));
            g->vb = nvb;
            HIP_TRY(hipMemcpyAsync(g->d_vb, g->vb.data(), sizeof(int64_t) * g->vb.size(), hipMemcpyHostToDevice,
                                   g->stream));
            launch_fill_gid(g->d_vb, G, g->d_gid, g->stream);
            HIP_TRY(hipGetLastError());
            NIDO_TRY(setup_phase(g));
        }
        const double at = now_s() - a0;
        S.t_agg += at;
        for (int i = 0; i < G; ++i)
            if (cand[(size_t)i]) g->bres[(size_t)i].phases.back().t_agg = at;
        if (!any(cont)) break;
    }
    // clusterings: graph-local ids on the device, then one copy to the host
    if (g->n0) {
        int64_t *d_loc = nullptr, *d_base = nullptr;
        int rc = dev_alloc(&d_loc, g->n0);
        if (rc == NIDO_OK) rc = dev_alloc(&d_base, G);
        if (rc == NIDO_OK && hipMemcpyAsync(d_base, base.data(), sizeof(int64_t) * (size_t)G,
                                            hipMemcpyHostToDevice, g->stream) != hipSuccess)
            rc = set_error(NIDO_ERR_HIP, "clustering bases upload failed");
        if (rc == NIDO_OK) {
            launch_local_ids(g->d_clus, g->n0, g->d_gid0, d_base, d_loc, g->stream);
            rc = host_clustering(g, d_loc);   // waits for the stream: base may go out of scope
        }
        dev_free(d_loc);
        dev_free(d_base);
        NIDO_TRY(rc);
    }
    S.t_total = now_s() - t0;
    S.phases = nphase;
    S.done = 1;
    for (int i = 0; i < G; ++i) {
        BatchResult& R = g->bres[(size_t)i];
        R.st.t_total = S.t_total;
        S.q += R.st.q;
    }
    g->last = S;
    if (st) *st = S;
    return NIDO_OK;
}

// ---- autotuner ----------------------------------------------------------
// Picks the launch knobs (degree-bin bounds, long-row partition size and
// table, stream overlap, graph replay) by timing `iters` iterations of the
// current phase after its first two from the singleton partition: one knob
// at a time, keeping a
// value only when it is at least 2% faster.  Results never depend on the
// knobs, so any choice is safe; the handle is left in the singleton
// partition of its current phase.
// Waits for the iteration and reads its moved-arc count (which picks the
// next iteration's form of Q).
bool sync_moved(nido_graph* g)
{
    if (g->G > 1 && hipMemcpyAsync(&g->last_moved, g->d_gq + g->G, sizeof(int64_t), hipMemcpyDeviceToHost,
                                   g->stream) != hipSuccess)
        return false;
    return hipStreamSynchronize(g->stream) == hipSuccess;
}

double tune_trial(nido_graph* g, const nido_tuning& t, int iters)
{
    g->tuning = t;
    if (setup_phase(g) != NIDO_OK || do_singleton(g) != NIDO_OK) return -1.0;
    if (g->G > 1) {
        std::vector<unsigned char> act((size_t)g->G);
        for (int i = 0; i < g->G; ++i) act[(size_t)i] = g->twom_g[(size_t)i] != 0;
        if (upload_flags(g, g->d_act, act) != NIDO_OK) return -1.0;
    }
    // Two untimed passes (the first iterations from the singleton partition,
    // which move most vertices and say little about the later ones; first
    // launches; the capture of both graph forms when replaying), then the
    // timed ones continue the same trajectory.
    for (int i = 0; i < 2; ++i)
        if (run_iteration_launches(g) != NIDO_OK || !sync_moved(g)) return -1.0;
    const double t0 = now_s();
    for (int i = 0; i < iters; ++i)
        if (run_iteration_launches(g) != NIDO_OK || !sync_moved(g)) return -1.0;
    return now_s() - t0;
}

int do_autotune(nido_graph* g, int iters, nido_tuning* out)
{
    nido_tuning best = g->tuning;
    const bool prof = g->prof.on;
    g->prof.on = false;
    double tb = tune_trial(g, best, iters);
    if (tb < 0) {
        g->prof.on = prof;
        return set_error(NIDO_ERR_HIP, "autotune trial failed: %s", nido_last_error());
    }
    auto try_value = [&](auto setter) {
        nido_tuning t = best;
        setter(t);
        if (std::memcmp(&t, &best, sizeof t) == 0) return;
        const double x = tune_trial(g, t, iters);
        if (x >= 0 && x < 0.98 * tb) {
            best = t;
            tb = x;
        }
    };
    for (int v : {0, 1, 2}) try_value([&](nido_tuning& t) { t.thread_max_deg = v; });
    for (int v : {128, 256, 512}) try_value([&](nido_tuning& t) { t.big_min_deg = v; });
    // table size and partition size move together: a partition of up to
    // big_part_arcs arcs fits a table of twice as many slots
    for (int v : {1024, 2048, 4096})
        for (int part : {v / 2, 3 * v / 8})
            try_value([&](nido_tuning& t) {
                t.big_slots = v;
                t.big_key_limit = v / 2;
                t.big_part_arcs = part;
            });
    for (int v : {0, 1}) try_value([&](nido_tuning& t) { t.overlap_long = v; });
    // small long-row tables beside the short rows on a second stream
    for (int v : {384, 512})
        try_value([&](nido_tuning& t) {
            t.big_slots = 1024;
            t.big_key_limit = 512;
            t.big_part_arcs = v;
            t.overlap_long = 1;
        });
    for (int v : {0, 2}) try_value([&](nido_tuning& t) { t.use_graph = v; });
    g->prof.on = prof;
    g->tuning = best;
    NIDO_TRY(setup_phase(g));
    NIDO_TRY(do_singleton(g));
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (out) *out = best;
    return NIDO_OK;
}

int create_impl(int64_t nv, const int64_t* off, const int64_t* col, const double* w, int nb,
                void* stream, nido_graph** out, int G = 1, const int64_t* vptr = nullptr, int64_t window = 0)
{
    if (!out) return set_error(NIDO_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (nv < 0 || !off) return set_error(NIDO_ERR_INVALID, "invalid vertex count or offsets");
    if (nv >= (1ll << 31) - 1) return set_error(NIDO_ERR_CAPACITY, "nv must be below 2^31-1");
    if (nb < 1) return set_error(NIDO_ERR_INVALID, "nbatches must be >= 1");
    const int64_t base = off[0];
    const int64_t ne = off[nv] - base;
    bool nonneg = true;
    const int64_t whole[2] = {0, nv};
    const int64_t* gv = G > 1 ? vptr : whole;
    // Inputs large enough for the pinned staging buffers are checked while
    // they pass through them (ids and weights are read there anyway); the
    // full host check runs first for smaller ones, and after the upload when
    // the quick checks saw anything it has to decide exactly (a bad arc, a
    // negative weight, a graph whose total of |w| nears the bound).
    const bool fused = !(window > 0 && ne > window) && ne > 0 && staged_path(sizeof(double) * (size_t)ne);
    if (fused) NIDO_TRY(check_offsets(nv, off, col, w));
    else NIDO_TRY(check_csr(nv, off, col, w, G, gv, &nonneg, nullptr));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_error(NIDO_ERR_HIP, "no HIP device available");
    }
    nido_graph* g = new nido_graph;
    g->nb = nb;
    default_tuning(g->tuning);
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
    if (hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_copied[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_copied[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_free[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->ev_free[1], hipEventDisableTiming) != hipSuccess)
        return fail(set_error(NIDO_ERR_HIP, "creating the side stream or its events failed"));
    AllocScope scope(g->stream);
    g->n0 = nv;
    g->nv = nv;
    g->ne = ne;
    g->h_off.resize((size_t)nv + 1);
    rebase_offsets(off, nv, g->h_off.data());
    // The device arc arrays hold the whole graph, or (window > 0 and more
    // arcs than that) one window; the arcs of a windowed graph stay in
    // pinned host memory.
    g->window = window;
    const bool windowed0 = window > 0 && ne > window;
    const int64_t cap = windowed0 ? window : ne;
    int rc;
    if (windowed0) {
        if (hipHostMalloc(reinterpret_cast<void**>(&g->h_col32), sizeof(int) * (size_t)ne) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void**>(&g->h_w), sizeof(double) * (size_t)ne) != hipSuccess)
            return fail(set_error(NIDO_ERR_NOMEM, "pinned host allocation of %lld arcs failed", (long long)ne));
        g->h_cap = ne;
#pragma omp parallel for schedule(static) if (ne >= (int64_t(1) << 22))
        for (int64_t e = 0; e < ne; ++e) {
            g->h_col32[e] = (int)col[base + e];
            g->h_w[e] = w[base + e];
        }
    }
    if ((rc = dev_alloc(&g->d_off, nv + 1)) || (rc = dev_alloc(&g->d_col, cap)) ||
        (rc = dev_alloc(&g->d_w, cap)) || (rc = dev_alloc(&g->d_C, nv)) ||
        (rc = dev_alloc(&g->d_Cn, nv)) || (rc = dev_alloc(&g->d_k, nv)) ||
        (rc = dev_alloc(&g->d_a, nv)) || (rc = dev_alloc(&g->d_kd, nv)) ||
        (rc = dev_alloc(&g->d_clus, nv)) || (rc = dev_alloc(&g->d_work, nv + 1)) ||
        (rc = dev_alloc(&g->d_acc, 4)) || (rc = dev_alloc(&g->d_q, 1)) ||
        (rc = dev_alloc(&g->d_tot, 2)) || (rc = dev_alloc(&g->d_src, cap)) ||
        (rc = dev_alloc(&g->d_col32, cap)) || (rc = dev_alloc(&g->d_C32, nv)) ||
        (rc = dev_alloc(&g->d_dcomms, nv)) || (rc = dev_alloc(&g->d_touched, nv)) ||
        (rc = dev_alloc(&g->d_term, nv)) || (rc = dev_alloc(&g->d_iacc, 2)) ||
        (rc = dev_alloc(&g->d_part, 2 * (int64_t)kPartBlocks)))
        return fail(rc);
    if ((rc = raw_alloc(&g->d_scan_tmp, scan_temp_bytes(nv + 1)))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void**>(&g->h_q), sizeof(double)) != hipSuccess)
        return fail(set_error(NIDO_ERR_NOMEM, "pinned allocation failed"));
    hipStream_t s = g->stream;
    launch_fill_fx(g->d_acc, 4, s);
    // Clustering::commIds_ starts as the identity (clustering.cpp:4-10)
    launch_iota_i64(g->d_clus, nv, s);
    ArcCheck chk;
    std::vector<int64_t> gstart((size_t)G + 1);
    for (int i = 0; i <= G; ++i) gstart[(size_t)i] = off[gv[i]] - base;
    chk.gstart = gstart.data();
    chk.vptr = gv;
    chk.G = G;
    // A handle whose graphs all have at most 2^16 vertices sends its ids as
    // 16-bit ids local to their graph, half the bytes of int32 (on the 4,096
    // SBMs 1.36 GB less over the host link), into the space of the int64 ids
    // (widened after the graph bases are on the device).
    int64_t maxg = 0;
    for (int i = 0; i < G; ++i) maxg = std::max(maxg, gv[i + 1] - gv[i]);
    const bool local16 = fused && G > 1 && maxg <= 65536;
    if (staged_h2d(g->d_off, g->h_off.data(), sizeof(int64_t) * (size_t)(nv + 1), s) != hipSuccess ||
        (local16 && staged_h2d_local_ids(reinterpret_cast<uint16_t*>(g->d_col), col + base, ne, s, chk) != hipSuccess) ||
        (ne > 0 && !windowed0 && !local16 &&
         staged_h2d_ids(g->d_col32, col + base, ne, s, fused ? &chk : nullptr) != hipSuccess) ||
        (ne > 0 && !windowed0 &&
         (fused ? staged_h2d_weights(g->d_w, w + base, ne, s, chk)
                : staged_h2d(g->d_w, w + base, sizeof(double) * (size_t)ne, s)) != hipSuccess))
        return fail(set_error(NIDO_ERR_HIP, "upload failed"));
    if (fused) {
        bool exact = chk.bad.load() != INT64_MAX || chk.neg.load();
        for (double x : chk.wabs) exact |= !(x <= 0x1p61);   // FP64 sums: far from 2^62 below this
        if (exact && (rc = check_csr(nv, off, col, w, G, gv, &nonneg, nullptr))) return fail(rc);
    }
    g->nonneg = nonneg;
    if (local16) {
        // graph-local ids back to handle ids, graph by graph in pieces of at
        // most 2^16 arcs
        std::vector<ArcPiece> pieces;
        for (int i = 0; i < G; ++i)
            for (int64_t e = gstart[(size_t)i]; e < gstart[(size_t)i + 1]; e += 65536)
                pieces.push_back(ArcPiece{e, std::min(gstart[(size_t)i + 1], e + 65536), gv[i]});
        ArcPiece* d_pieces = nullptr;
        if ((rc = dev_alloc(&d_pieces, std::max<int64_t>((int64_t)pieces.size(), 1)))) return fail(rc);
        const bool ok = pieces.empty() ||
                        (hipMemcpyAsync(d_pieces, pieces.data(), sizeof(ArcPiece) * pieces.size(),
                                        hipMemcpyHostToDevice, s) == hipSuccess &&
                         (launch_widen_local(d_pieces, (int64_t)pieces.size(),
                                             reinterpret_cast<const uint16_t*>(g->d_col), g->d_col32, s),
                          hipGetLastError() == hipSuccess) &&
                         hipStreamSynchronize(s) == hipSuccess);   // pieces go out of scope
        dev_free(d_pieces);
        if (!ok) return fail(set_error(NIDO_ERR_HIP, "widening the graph-local ids failed"));
    }
    // ids crossed the host link as int32 or 16-bit local ids (all below
    // nv < 2^31 - 1, checked); the int64 copy overwrites the local ids
    if (!windowed0) launch_widen_ids(g->d_col32, g->d_col, ne, s);
    if (G > 1) {
        g->G = G;
        g->vb.assign(vptr, vptr + G + 1);
        g->vb0 = g->vb;
        if ((rc = dev_alloc(&g->d_gid, nv)) || (rc = dev_alloc(&g->d_gid0, nv)) ||
            (rc = dev_alloc(&g->d_vb, G + 1)) || (rc = dev_alloc(&g->d_vbn, G + 1)) ||
            (rc = dev_alloc(&g->d_mg, 2 * (int64_t)G)) || (rc = dev_alloc(&g->d_twomg, G)) || (rc = dev_alloc(&g->d_act, G)) ||
            (rc = dev_alloc(&g->d_keep, G)) || (rc = dev_alloc(&g->d_gacc, 2 * (int64_t)G)) ||
            (rc = dev_alloc(&g->d_gq, (int64_t)G + 1)) ||
            (rc = dev_alloc(&g->d_moved, kMovedSlots)))
            return fail(rc);
        if (hipMemcpyAsync(g->d_vb, g->vb.data(), sizeof(int64_t) * (size_t)(G + 1), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return fail(set_error(NIDO_ERR_HIP, "upload failed"));
        launch_fill_gid(g->d_vb, G, g->d_gid, s);
        launch_fill_gid(g->d_vb, G, g->d_gid0, s);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            return fail(set_error(NIDO_ERR_HIP, "graph id set-up failed"));
    }
    if ((rc = setup_phase(g))) return fail(rc);
    if (G > 1) {
        // 2m per graph, exactly (graph_cuda.cu:1855): the sum of its k_v
        g->twom_g.assign((size_t)G, 0);
        launch_graph_sums(g->d_k, g->d_vb, G, g->d_gacc, s);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(g->twom_g.data(), g->d_gacc, sizeof(fx_t) * (size_t)G, hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(set_error(NIDO_ERR_HIP, "per-graph 2m failed"));
        g->twom_dg.assign((size_t)G, 0.0);
        std::vector<double> mg(2 * (size_t)G), tw((size_t)G);
        for (int i = 0; i < G; ++i) {
            tw[(size_t)i] = g->twom_dg[(size_t)i] = fx_to_double(g->twom_g[(size_t)i]);
            mg[2 * (size_t)i] = tw[(size_t)i] * 0.5;
            mg[2 * (size_t)i + 1] = 1.0 / mg[2 * (size_t)i];   // unused while the graph has no weight
        }
        if (hipMemcpyAsync(g->d_mg, mg.data(), sizeof(double) * 2 * (size_t)G, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(g->d_twomg, tw.data(), sizeof(double) * (size_t)G, hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(set_error(NIDO_ERR_HIP, "upload failed"));
    }
    if (!g->windowed && ne > 0) {
        fx_t* d_h = nullptr;   // four 64-bit sums
        unsigned long long h[4];
        if ((rc = dev_alloc(&d_h, 2))) return fail(rc);
        launch_fill_fx(d_h, 2, s);
        launch_sym_hash(g->d_src, g->d_col32, g->d_w, ne, reinterpret_cast<unsigned long long*>(d_h), s);
        const bool ok = hipGetLastError() == hipSuccess &&
                        hipMemcpyAsync(h, d_h, sizeof(h), hipMemcpyDeviceToHost, s) == hipSuccess &&
                        hipStreamSynchronize(s) == hipSuccess;
        dev_free(d_h);
        if (!ok) return fail(set_error(NIDO_ERR_HIP, "arc symmetry check failed"));
        g->symmetric = h[0] == h[1] && h[2] == h[3];
    } else if (g->windowed && ne > 0) {   // the arcs are on the host
        g->symmetric = host_arcs_symmetric(nv, g->h_off.data(), g->h_col32, g->h_w);
    }
    if ((rc = do_singleton(g))) return fail(rc);
    *out = g;
    return NIDO_OK;
}

// The n0 entries of src (device) into the handle's host clustering.
int host_clustering(nido_graph* g, const int64_t* src)
{
    if (g->bclus_n != g->n0) {
        g->bclus.reset(new int64_t[(size_t)std::max<int64_t>(g->n0, 1)]);   // no zero fill: copied over
        g->bclus_n = g->n0;
    }
    HIP_TRY(staged_d2h(g->bclus.get(), src, sizeof(int64_t) * (size_t)g->n0, g->stream));
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
    } catch (const std::exception& e_) {                                         \
        return set_error(NIDO_ERR_STATE, "internal error: %s", e_.what());       \
    }

extern "C" {

int nido_abi_version(void) { return NIDO_ABI_VERSION; }

int nido_set_device_lists(int64_t min_vertices)
{
    if (min_vertices < 0) return set_error(NIDO_ERR_INVALID, "min_vertices must be >= 0");
    g_device_lists_min.store(min_vertices ? min_vertices : kDeviceListsMin);
    return NIDO_OK;
}

int nido_set_list_rebuild(int percent)
{
    if (percent > 100) return set_error(NIDO_ERR_INVALID, "percent must be <= 100");
    g_list_rebuild_pct.store(percent < 0 ? kListRebuildPct : percent);
    return NIDO_OK;
}

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
    // determine_optimal_edges_per_partition (graph_gpu.cpp:362-391): arcs
    // that do not fit the free device memory stream through a window
    int64_t window = 0;
    if (off && nv >= 0) {
        const int64_t arcs = off[nv] - off[0];
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const double avail = 0.85 * (double)free_b - kDeviceBytesPerVertex * (double)nv;
            if (kDeviceBytesPerArc * (double)arcs > avail && avail > 0)
                window = std::max<int64_t>((int64_t)(avail / kWindowBytesPerArc), 1);
        } else {
            (void)hipGetLastError();
        }
    }
    API_GUARD(return create_impl(nv, off, col, w, nbatches, stream, out, 1, nullptr, window);)
}

int nido_graph_create_windowed(int64_t nv, int64_t ne, const int64_t* off, const int64_t* col, const double* w,
                               int nbatches, int64_t window_arcs, void* stream, nido_graph** out)
{
    (void)ne;
    if (window_arcs < 1) return set_error(NIDO_ERR_INVALID, "window_arcs must be >= 1");
    API_GUARD(return create_impl(nv, off, col, w, nbatches, stream, out, 1, nullptr, window_arcs);)
}

int nido_graph_destroy(nido_graph* g)
{
    if (!g) return NIDO_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    AllocScope scope(g->stream);
    drop_iteration_graph(g);
    free_graph_arrays(g);
    dev_free(g->d_C); dev_free(g->d_Cn); dev_free(g->d_k); dev_free(g->d_a); dev_free(g->d_kd);
    dev_free(g->d_clus); dev_free(g->d_work); dev_free(g->d_acc); dev_free(g->d_q); dev_free(g->d_tot);
    dev_free(g->d_src); dev_free(g->d_part); dev_free(g->d_col32); dev_free(g->d_C32); dev_free(g->d_wq);
    dev_free(g->d_gid); dev_free(g->d_gid0); dev_free(g->d_vb); dev_free(g->d_vbn); dev_free(g->d_mg);
    dev_free(g->d_twomg); dev_free(g->d_act); dev_free(g->d_keep); dev_free(g->d_gacc); dev_free(g->d_gq);
    dev_free(g->d_moved);
    dev_free(g->d_seg);
    dev_free(g->d_dcounts); dev_free(g->d_dchunks); dev_free(g->d_dcomms); dev_free(g->d_touched);
    dev_free(g->d_term); dev_free(g->d_iacc); dev_free(g->d_gp);
    dev_free(g->d_binv); dev_free(g->d_big_items); dev_free(g->d_big_poff); dev_free(g->d_big_toff);
    dev_free(g->d_key_chunks); dev_free(g->d_bkey); dev_free(g->d_bw); dev_free(g->d_cpo);
    dev_free(g->d_pg); dev_free(g->d_pc); dev_free(g->d_pself); dev_free(g->d_cnt);
    dev_free(g->d_tkey); dev_free(g->d_tval);
    dev_free(g->d_small); dev_free(g->d_items);
    raw_free(g->d_scan_tmp);
    if (g->h_q) (void)hipHostFree(g->h_q);
    if (g->h_col32) (void)hipHostFree(g->h_col32);
    if (g->h_w) (void)hipHostFree(g->h_w);
    if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
    if (g->ev_join) (void)hipEventDestroy(g->ev_join);
    for (int h = 0; h < 2; ++h) {
        if (g->ev_copied[h]) (void)hipEventDestroy(g->ev_copied[h]);
        if (g->ev_free[h]) (void)hipEventDestroy(g->ev_free[h]);
    }
    if (g->side) (void)hipStreamDestroy(g->side);
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
    if (!t) return set_error(NIDO_ERR_INVALID, "null argument");
    if (t->thread_max_deg < 0 || t->thread_max_deg > kThreadMaxDeg)
        return set_error(NIDO_ERR_INVALID, "thread_max_deg must be in [0, %d]", kThreadMaxDeg);
    if (t->big_min_deg < 0 || t->big_min_deg > kBinMaxDeg[BIN_B256])
        return set_error(NIDO_ERR_INVALID, "big_min_deg must be in [0, %d]", (int)kBinMaxDeg[BIN_B256]);
    if (t->big_part_arcs < 1)
        return set_error(NIDO_ERR_INVALID, "big_part_arcs must be >= 1");
    if (t->big_slots != 1024 && t->big_slots != 2048 && t->big_slots != 4096)
        return set_error(NIDO_ERR_INVALID, "big_slots must be 1024, 2048 or 4096");
    if (t->big_key_limit < 1 || t->big_key_limit > t->big_slots / 2)
        return set_error(NIDO_ERR_INVALID, "big_key_limit must be in [1, big_slots/2]");
    if (t->use_graph < 0 || t->use_graph > 2) return set_error(NIDO_ERR_INVALID, "use_graph must be 0, 1 or 2");
    if ((t->split_small | 1) != 1) return set_error(NIDO_ERR_INVALID, "split_small must be 0 or 1");
    if ((t->overlap_long | 1) != 1) return set_error(NIDO_ERR_INVALID, "overlap_long must be 0 or 1");
    if (t->full_modularity < 0 || t->full_modularity > 2)
        return set_error(NIDO_ERR_INVALID, "full_modularity must be 0, 1 or 2");
    if ((t->fixed_short | 1) != 1) return set_error(NIDO_ERR_INVALID, "fixed_short must be 0 or 1");
    if (!g) return set_error(NIDO_ERR_INVALID, "null handle");
    AllocScope scope(g->stream);
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
    NIDO_TRY(single_graph(g));
    return do_singleton(g);
}

int nido_compute_modularity(nido_graph* g, double* q)
{
    NIDO_TRY(single_graph(g));
    if (!q) return set_error(NIDO_ERR_INVALID, "null argument");
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
    NIDO_TRY(single_graph(g));
    return do_restore(g);
}

int nido_aggregation(nido_graph* g, int* done)
{
    NIDO_TRY(single_graph(g));
    if (!done) return set_error(NIDO_ERR_INVALID, "null argument");
    AllocScope scope(g->stream);
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
    if (g->windowed) {   // the arcs are on the host
        if (off) std::copy(g->h_off.begin(), g->h_off.end(), off);
        if (col) std::copy(g->h_col32, g->h_col32 + g->ne, col);
        if (w) std::copy(g->h_w, g->h_w + g->ne, w);
        return NIDO_OK;
    }
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
    return nido_dump_partition_orders(g, path, nullptr);
}

int nido_dump_partition_orders(nido_graph* g, const char* path, const int64_t* new_orders)
{
    if (!g || !path) return set_error(NIDO_ERR_INVALID, "null argument");
    API_GUARD(
        std::vector<int64_t> c((size_t)std::max<int64_t>(g->n0, 1));
        NIDO_TRY(d2h(g, c.data(), g->d_clus, g->n0));
        if (new_orders) {   // clustering.cpp:25-37: line new_orders[i] holds vertex i's community
            std::vector<int64_t> t((size_t)std::max<int64_t>(g->n0, 1), -1);
            for (int64_t i = 0; i < g->n0; ++i) {
                const int64_t id = new_orders[i];
                if (id < 0 || id >= g->n0 || t[(size_t)id] >= 0)
                    return set_error(NIDO_ERR_INVALID, "new_orders is not a permutation of [0, %lld)",
                                     (long long)g->n0);
                t[(size_t)id] = c[(size_t)i];
            }
            c.swap(t);
        }
        FILE* f = fopen(path, "wt");
        if (!f) return set_error(NIDO_ERR_IO, "cannot open %s", path);
        for (int64_t i = 0; i < g->n0; ++i) fprintf(f, "%lld\n", (long long)c[(size_t)i]);
        fclose(f);
        return NIDO_OK;)
}

int nido_louvain_iteration(nido_graph* g, double* q, double* dq, double* q_eval)
{
    NIDO_TRY(single_graph(g));
    if (!q || !dq) return set_error(NIDO_ERR_INVALID, "null argument");
    API_GUARD(return do_iteration(g, q, dq, q_eval);)
}

int nido_graph_set_profiling(nido_graph* g, int enable)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    HIP_TRY(hipStreamSynchronize(g->stream));
    API_GUARD(g->prof.harvest();)
    g->prof.reset();
    g->prof.on = enable != 0;
    g->prof.mask = ~0u;
    return NIDO_OK;
}

int nido_graph_set_profiling_classes(nido_graph* g, unsigned mask)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    HIP_TRY(hipStreamSynchronize(g->stream));
    API_GUARD(g->prof.harvest();)
    g->prof.reset();
    g->prof.on = mask != 0;
    g->prof.mask = mask;
    return NIDO_OK;
}

int nido_graph_set_profiling_stride(nido_graph* g, int stride)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (stride < 1) return set_error(NIDO_ERR_INVALID, "profiling stride must be >= 1");
    HIP_TRY(hipStreamSynchronize(g->stream));
    API_GUARD(g->prof.harvest();)
    g->prof.reset();
    g->prof.stride = stride;
    return NIDO_OK;
}

int nido_graph_kernel_stats(nido_graph* g, int cls, int64_t* launches, double* ms, double* bytes)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (cls < 0 || cls >= NIDO_K_CLASSES) return set_error(NIDO_ERR_INVALID, "kernel class %d", cls);
    HIP_TRY(hipStreamSynchronize(g->stream));
    API_GUARD(g->prof.harvest();)
    if (launches) *launches = g->prof.launches[cls];
    if (ms) *ms = g->prof.ms[cls];
    if (bytes) *bytes = g->prof.bytes[cls];
    return NIDO_OK;
}

int nido_louvain_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, int record,
                     int verbose, nido_run_stats* stats)
{
    NIDO_TRY(single_graph(g));
    AllocScope scope(g->stream);
    API_GUARD(return do_run(g, max_loops, tol, tol_phase, record, verbose, stats);)
}

int nido_batch_create(int ngraphs, const int64_t* vptr, const int64_t* off, const int64_t* col,
                      const double* w, int nbatches, void* stream, nido_graph** out)
{
    if (!out) return set_error(NIDO_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (ngraphs < 1 || !vptr || !off) return set_error(NIDO_ERR_INVALID, "need ngraphs >= 1, vptr and offsets");
    if (vptr[0] != 0) return set_error(NIDO_ERR_INVALID, "vptr[0] must be 0");
    for (int i = 0; i < ngraphs; ++i)
        if (vptr[i + 1] < vptr[i]) return set_error(NIDO_ERR_INVALID, "vptr decreases at graph %d", i);
    const int64_t nv = vptr[ngraphs];
    if (nv >= (1ll << 31) - 1) return set_error(NIDO_ERR_CAPACITY, "total vertices must be below 2^31-1");
    // graphs are disjoint (graph i's arcs stay inside [vptr[i], vptr[i+1])):
    // checked with the rest of the CSR when the handle is created
    API_GUARD(return create_impl(nv, off, col, w, nbatches, stream, out, ngraphs == 1 ? 1 : ngraphs, vptr);)
}

int nido_batch_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, nido_run_stats* stats)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    AllocScope scope(g->stream);
    if (g->G == 1) {
        // one graph: the plain path, with its result kept in the batched form
        nido_run_stats S{};
        API_GUARD(NIDO_TRY(do_run(g, max_loops, tol, tol_phase, 0, 0, &S));
                  g->bres.assign(1, BatchResult{});
                  g->bres[0].st = S;
                  for (const auto& h : g->hist) g->bres[0].phases.push_back(h.st);
                  NIDO_TRY(host_clustering(g, g->d_clus));
                  if (stats) *stats = S;
                  return NIDO_OK;)
    }
    API_GUARD(return do_batch_run(g, max_loops, tol, tol_phase, stats);)
}

int nido_autotune(nido_graph* g, int iterations, nido_tuning* chosen)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (iterations < 1) return set_error(NIDO_ERR_INVALID, "iterations must be >= 1");
    AllocScope scope(g->stream);
    API_GUARD(return do_autotune(g, iterations, chosen);)
}

int nido_batch_size(const nido_graph* g, int* ngraphs)
{
    if (!g || !ngraphs) return set_error(NIDO_ERR_INVALID, "null argument");
    *ngraphs = g->G;
    return NIDO_OK;
}

int nido_batch_list_rebuilds(const nido_graph* g, int64_t* n)
{
    if (!g || !n) return set_error(NIDO_ERR_INVALID, "null argument");
    *n = g->rebuilds;
    return NIDO_OK;
}

static int check_result(const nido_graph* g, int graph)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (graph < 0 || graph >= g->G) return set_error(NIDO_ERR_INVALID, "graph %d out of range", graph);
    if (g->bres.size() != (size_t)g->G) return set_error(NIDO_ERR_STATE, "no batched run yet");
    return NIDO_OK;
}

int nido_batch_graph_stats(const nido_graph* g, int graph, nido_run_stats* out)
{
    NIDO_TRY(check_result(g, graph));
    if (!out) return set_error(NIDO_ERR_INVALID, "null argument");
    *out = g->bres[(size_t)graph].st;
    return NIDO_OK;
}

int nido_batch_graph_phase(const nido_graph* g, int graph, int phase, nido_phase_stats* out)
{
    NIDO_TRY(check_result(g, graph));
    const auto& P = g->bres[(size_t)graph].phases;
    if (!out || phase < 0 || phase >= (int)P.size()) return set_error(NIDO_ERR_INVALID, "phase %d out of range", phase);
    *out = P[(size_t)phase];
    return NIDO_OK;
}

int nido_batch_clustering(const nido_graph* g, int graph, int64_t* comm)
{
    NIDO_TRY(check_result(g, graph));
    if (!comm) return set_error(NIDO_ERR_INVALID, "null argument");
    const int64_t o0 = g->G == 1 ? 0 : g->vb0[(size_t)graph];
    const int64_t o1 = g->G == 1 ? g->n0 : g->vb0[(size_t)graph + 1];
    if (o1 > o0) std::memcpy(comm, g->bclus.get() + o0, sizeof(int64_t) * (size_t)(o1 - o0));
    return NIDO_OK;
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
