// host_graph.cpp -- host CSR graph (the reference's Graph, graph.hpp:14-102):
// binary reader/writer and the seeded synthetic-input generators.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>
#include <omp.h>
#include <parallel/algorithm>
#include <random>

#include "../../include/nido_amd.h"
#include "errors.h"
#include "host_check.h"

// Vectors whose resize leaves the new elements uninitialised: every element
// of a CSR array is written right after the array is sized, and a zero fill
// of gigabytes on one thread (with its page faults) costs more than the
// writes that follow, which the threads that write take in parallel.
template <typename T>
struct NoInit : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <typename U>
    NoInit(const NoInit<U>&) noexcept {}
    template <typename U>
    void construct(U* p) noexcept
    {
        ::new (static_cast<void*>(p)) U;
    }
    template <typename U, typename... A>
    void construct(U* p, A&&... a)
    {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <typename T>
using hvec = std::vector<T, NoInit<T>>;

struct nido_hgraph {
    int64_t nv = 0, ne = 0;
    hvec<int64_t> off;
    hvec<int64_t> col;
    hvec<double> w;
};

namespace nido {

static thread_local char g_err[512];

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }

const char* last_error() { return g_err; }

namespace {

// ---- counter-based randomness ----------------------------------------------
inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() { return mix64(s++ * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull); }
    double uniform() { return (double)(next() >> 11) * 0x1p-53; }        // [0, 1)
    double uniform_pos() { return 1.0 - uniform(); }                     // (0, 1]
};

inline uint64_t stream_seed(uint64_t seed, uint64_t stream)
{
    return mix64(seed ^ mix64(stream + 0x5851F42D4C957F2Dull));
}

// One weight per undirected edge {u, v}, u < v: (0, 1] on a 2^-53 grid,
// a pure function of (seed, u, v) so duplicates and order never matter.
inline double edge_weight(uint64_t seed, int64_t u, int64_t v)
{
    const uint64_t r = mix64(seed ^ mix64((uint64_t)u * 0x9E3779B97F4A7C15ull ^ (uint64_t)v));
    return (double)((1ull << 53) - (r >> 11)) * 0x1p-53;
}

// Builds a symmetric CSR from unique undirected edges packed as (u<<32)|v
// with u < v, sorted ascending.  Rows come out sorted by neighbour id: row x
// receives its reverse arcs (from u < x) before its forward arcs (to v > x).
nido_hgraph* csr_from_sorted_pairs(int64_t nv, const std::vector<uint64_t>& e, int weighted,
                                   uint64_t wseed)
{
    nido_hgraph* g = new nido_hgraph;
    g->nv = nv;
    g->ne = 2 * (int64_t)e.size();
    g->off.assign((size_t)nv + 1, 0);
    for (uint64_t p : e) {
        g->off[(p >> 32) + 1]++;
        g->off[(p & 0xffffffffull) + 1]++;
    }
    for (int64_t i = 0; i < nv; ++i) g->off[i + 1] += g->off[i];
    g->col.resize((size_t)g->ne);
    g->w.resize((size_t)g->ne);
    std::vector<int64_t> cur(g->off.begin(), g->off.end() - 1);
    for (uint64_t p : e) {
        const int64_t u = (int64_t)(p >> 32), v = (int64_t)(p & 0xffffffffull);
        const double wt = weighted ? edge_weight(wseed, u, v) : 1.0;
        int64_t a = cur[u]++, b = cur[v]++;
        g->col[a] = v; g->w[a] = wt;
        g->col[b] = u; g->w[b] = wt;
    }
    return g;
}

void sort_unique(std::vector<uint64_t>& e)
{
    __gnu_parallel::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
}

int check_out(nido_hgraph** out)
{
    if (!out) return set_error(NIDO_ERR_INVALID, "null output handle");
    *out = nullptr;
    return NIDO_OK;
}

}  // namespace
}  // namespace nido

using namespace nido;

extern "C" {

const char* nido_last_error(void) { return nido::last_error(); }

int nido_hgraph_from_csr(int64_t nv, int64_t ne, const int64_t* off, const int64_t* col,
                         const double* w, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    if (nv < 0 || !off || (nv > 0 && off[nv] - off[0] > 0 && (!col || !w)))
        return set_error(NIDO_ERR_INVALID, "invalid CSR arguments");
    const int64_t base = off[0];
    const int64_t arcs = off[nv] - base;
    if (base < 0 || arcs < 0) return set_error(NIDO_ERR_INVALID, "offsets negative or decreasing");
    // cols/weights hold ne entries from offsets[0] on: offsets that claim
    // more arcs would read past them
    if (arcs > ne)
        return set_error(NIDO_ERR_INVALID, "offsets span %lld arcs, more than the %lld given", (long long)arcs,
                         (long long)ne);
    rc = nido::check_offsets(nv, off, col, w);
    if (rc) return rc;
    try {
        nido_hgraph* g = new nido_hgraph;
        g->nv = nv;
        g->ne = arcs;
        g->off.resize((size_t)nv + 1);
        for (int64_t i = 0; i <= nv; ++i) g->off[i] = off[i] - base;
        g->col.assign(col + base, col + base + arcs);
        g->w.assign(w + base, w + base + arcs);
        *out = g;
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

int nido_hgraph_read(const char* path, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    std::ifstream f(path ? path : "", std::ios::in | std::ios::binary);
    if (!f.is_open()) return set_error(NIDO_ERR_IO, "cannot open %s", path ? path : "(null)");
    int64_t hdr[2];
    if (!f.read(reinterpret_cast<char*>(hdr), sizeof(hdr)))
        return set_error(NIDO_ERR_IO, "%s: short header", path);
    const int64_t nv = hdr[0];
    if (nv < 0) return set_error(NIDO_ERR_INVALID, "%s: negative vertex count", path);
    try {
        nido_hgraph* g = new nido_hgraph;
        g->nv = nv;
        g->off.resize((size_t)nv + 1);
        if (!f.read(reinterpret_cast<char*>(g->off.data()), (std::streamsize)(sizeof(int64_t) * (nv + 1)))) {
            delete g;
            return set_error(NIDO_ERR_IO, "%s: short offsets", path);
        }
        // graph.cpp:376-382: the arc count comes from the offsets
        const int64_t base = g->off[0];
        const int64_t arcs = g->off[nv] - base;
        if (arcs < 0) {
            delete g;
            return set_error(NIDO_ERR_INVALID, "%s: offsets decrease", path);
        }
        for (int64_t i = 0; i <= nv; ++i) g->off[i] -= base;      // graph.cpp:470-472
        g->ne = arcs;
        g->col.resize((size_t)arcs);
        g->w.resize((size_t)arcs);
        // records {int64 tail, double weight} (types.hpp:48-55), read right
        // after the offsets (the GRAPH_FT_LOAD path, graph.cpp:397-435)
        const int64_t chunk = 1 << 20;
        std::vector<char> buf;
        buf.resize((size_t)std::min<int64_t>(arcs, chunk) * 16 + 16);
        for (int64_t e0 = 0; e0 < arcs; e0 += chunk) {
            const int64_t n = std::min(chunk, arcs - e0);
            if (!f.read(buf.data(), (std::streamsize)(n * 16))) {
                delete g;
                return set_error(NIDO_ERR_IO, "%s: short edge records", path);
            }
            for (int64_t i = 0; i < n; ++i) {
                std::memcpy(&g->col[e0 + i], buf.data() + 16 * i, 8);
                std::memcpy(&g->w[e0 + i], buf.data() + 16 * i + 8, 8);
            }
        }
        *out = g;
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

int nido_hgraph_write(const nido_hgraph* g, const char* path)
{
    if (!g || !path) return set_error(NIDO_ERR_INVALID, "null argument");
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return set_error(NIDO_ERR_IO, "cannot open %s for writing", path);
    const int64_t hdr[2] = {g->nv, g->ne};
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    f.write(reinterpret_cast<const char*>(g->off.data()), (std::streamsize)(sizeof(int64_t) * (g->nv + 1)));
    std::vector<char> buf;
    const int64_t chunk = 1 << 20;
    buf.resize((size_t)std::min<int64_t>(g->ne, chunk) * 16 + 16);
    for (int64_t e0 = 0; e0 < g->ne; e0 += chunk) {
        const int64_t n = std::min(chunk, g->ne - e0);
        for (int64_t i = 0; i < n; ++i) {
            std::memcpy(buf.data() + 16 * i, &g->col[e0 + i], 8);
            std::memcpy(buf.data() + 16 * i + 8, &g->w[e0 + i], 8);
        }
        f.write(buf.data(), (std::streamsize)(n * 16));
    }
    if (!f.good()) return set_error(NIDO_ERR_IO, "write to %s failed", path);
    return NIDO_OK;
}

int nido_hgraph_info(const nido_hgraph* g, int64_t* nv, int64_t* ne)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (nv) *nv = g->nv;
    if (ne) *ne = g->ne;
    return NIDO_OK;
}

int nido_hgraph_arrays(const nido_hgraph* g, const int64_t** off, const int64_t** col,
                       const double** w)
{
    if (!g) return set_error(NIDO_ERR_INVALID, "null graph");
    if (off) *off = g->off.data();
    if (col) *col = g->col.data();
    if (w) *w = g->w.data();
    return NIDO_OK;
}

void nido_hgraph_destroy(nido_hgraph* g) { delete g; }

int nido_gen_sbm(int64_t nv, int64_t block, double p_in, double p_out, int weighted,
                 uint64_t seed, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    if (nv < 0 || nv >= (1ll << 31) || block < 1 || !(p_in >= 0 && p_in <= 1) ||
        !(p_out >= 0 && p_out <= 1))
        return set_error(NIDO_ERR_INVALID, "invalid SBM parameters");
    try {
        // per source vertex u, successes among v in (u, nv) by geometric skips,
        // in lexicographic pair order; stream u of the seed drives vertex u
        std::vector<std::vector<uint64_t>> rows((size_t)nv);
        #pragma omp parallel for schedule(dynamic, 64)
        for (int64_t u = 0; u < nv; ++u) {
            Rng r(stream_seed(seed, (uint64_t)u));
            auto& row = rows[(size_t)u];
            const int64_t bend = std::min(nv, (u / block + 1) * block);
            auto sweep = [&](int64_t lo, int64_t hi, double p) {
                if (p <= 0 || lo >= hi) return;
                if (p >= 1) {
                    for (int64_t v = lo; v < hi; ++v) row.push_back(((uint64_t)u << 32) | (uint64_t)v);
                    return;
                }
                const double lq = std::log1p(-p);
                int64_t v = lo - 1;
                for (;;) {
                    const double skip = std::floor(std::log(r.uniform_pos()) / lq);
                    if (skip >= (double)(hi - v)) break;
                    v += 1 + (int64_t)skip;
                    if (v >= hi) break;
                    row.push_back(((uint64_t)u << 32) | (uint64_t)v);
                }
            };
            sweep(u + 1, bend, p_in);
            sweep(bend, nv, p_out);
        }
        std::vector<uint64_t> e;
        size_t total = 0;
        for (auto& r : rows) total += r.size();
        e.reserve(total);
        for (auto& r : rows) {
            e.insert(e.end(), r.begin(), r.end());
            std::vector<uint64_t>().swap(r);
        }
        *out = csr_from_sorted_pairs(nv, e, weighted, seed ^ 0xA5A5A5A5DEADBEEFull);
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

int nido_gen_rmat(int scale, int64_t edge_factor, double a, double b, double c, int weighted,
                  uint64_t seed, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1)
        return set_error(NIDO_ERR_INVALID, "invalid R-MAT parameters");
    const int64_t nv = 1ll << scale;
    const int64_t ns = edge_factor * nv;
    try {
        std::vector<uint64_t> e((size_t)ns);
        const double ab = a + b, abc = a + b + c;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < ns; ++i) {
            Rng r(stream_seed(seed, (uint64_t)i));
            uint64_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                const double x = r.uniform();
                const uint64_t bu = x >= ab;                        // c or d quadrant
                const uint64_t bv = (x >= a && x < ab) || x >= abc; // b or d quadrant
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            if (u == v) e[(size_t)i] = ~0ull;                       // self-loop, dropped
            else {
                const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
                e[(size_t)i] = (lo << 32) | hi;
            }
        }
        sort_unique(e);
        if (!e.empty() && e.back() == ~0ull) e.pop_back();
        *out = csr_from_sorted_pairs(nv, e, weighted, seed ^ 0xA5A5A5A5DEADBEEFull);
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

int nido_gen_chung_lu(int64_t nv, double alpha, double dmin, double dmax, double mean,
                      int weighted, uint64_t seed, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    if (nv < 2 || nv >= (1ll << 31) || !(alpha > 1) || !(dmin > 0) || !(dmax >= dmin) ||
        !(mean >= dmin && mean <= dmax))
        return set_error(NIDO_ERR_INVALID, "invalid Chung-Lu parameters");
    try {
        // expected degrees: Pareto(alpha) with x_min = 1, scaled so that the
        // mean after clipping to [dmin, dmax] is `mean`
        std::vector<double> base((size_t)nv), theta((size_t)nv);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nv; ++i) {
            Rng r(stream_seed(seed, (uint64_t)i));
            base[(size_t)i] = std::pow(r.uniform_pos(), -1.0 / (alpha - 1.0));
        }
        double s = 1.0;
        for (int it = 0; it < 40; ++it) {
            double tot = 0;
            #pragma omp parallel for reduction(+ : tot) schedule(static)
            for (int64_t i = 0; i < nv; ++i)
                tot += std::min(dmax, std::max(dmin, s * base[(size_t)i]));
            s *= mean / (tot / (double)nv);
        }
        double tot = 0;
        for (int64_t i = 0; i < nv; ++i) {
            theta[(size_t)i] = std::min(dmax, std::max(dmin, s * base[(size_t)i]));
            tot += theta[(size_t)i];
        }
        std::vector<double>().swap(base);
        // Vose alias table over the expected degrees
        std::vector<double> prob((size_t)nv);
        std::vector<int64_t> alias((size_t)nv);
        {
            std::vector<int64_t> small, large;
            std::vector<double> p((size_t)nv);
            for (int64_t i = 0; i < nv; ++i) {
                p[(size_t)i] = theta[(size_t)i] * (double)nv / tot;
                (p[(size_t)i] < 1.0 ? small : large).push_back(i);
            }
            while (!small.empty() && !large.empty()) {
                const int64_t l = small.back(); small.pop_back();
                const int64_t g = large.back(); large.pop_back();
                prob[(size_t)l] = p[(size_t)l];
                alias[(size_t)l] = g;
                p[(size_t)g] = (p[(size_t)g] + p[(size_t)l]) - 1.0;
                (p[(size_t)g] < 1.0 ? small : large).push_back(g);
            }
            for (int64_t i : large) { prob[(size_t)i] = 1.0; alias[(size_t)i] = i; }
            for (int64_t i : small) { prob[(size_t)i] = 1.0; alias[(size_t)i] = i; }
        }
        const int64_t ns = (int64_t)std::llround((double)nv * mean / 2.0);
        std::vector<uint64_t> e((size_t)ns);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < ns; ++i) {
            Rng r(stream_seed(seed ^ 0x3C6EF372FE94F82Bull, (uint64_t)i));
            int64_t end[2];
            for (int k = 0; k < 2; ++k) {
                const uint64_t x = r.next();
                const int64_t col = (int64_t)((x >> 32) * (uint64_t)nv >> 32);
                const double y = (double)(x & 0xffffffffull) * 0x1p-32;
                end[k] = y < prob[(size_t)col] ? col : alias[(size_t)col];
            }
            if (end[0] == end[1]) e[(size_t)i] = ~0ull;
            else {
                const uint64_t lo = (uint64_t)std::min(end[0], end[1]);
                const uint64_t hi = (uint64_t)std::max(end[0], end[1]);
                e[(size_t)i] = (lo << 32) | hi;
            }
        }
        sort_unique(e);
        if (!e.empty() && e.back() == ~0ull) e.pop_back();
        *out = csr_from_sorted_pairs(nv, e, weighted, seed ^ 0xA5A5A5A5DEADBEEFull);
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

// Graph::create_random_network_ba(m0), graph.cpp:76-110, the reference's -r
// generator, with the same random-number objects drawn in the same order:
// two default-constructed std::default_random_engine (one for the weights,
// one for the endpoints), one std::normal_distribution<double>(1, 1) whose
// absolute values are the weights, and a fresh std::uniform_int_distribution
// over the arcs listed so far for every new vertex.  Vertices 0..m0 start as
// a star around 0; each later vertex i attaches m0 times to the first
// endpoint of a uniformly chosen earlier arc (so by degree), duplicates
// kept.  The arcs are then ordered by (tail, head), as the reference's heap
// sort orders them (graph.cpp:62-72); the order among duplicate arcs, which
// its heap leaves unspecified, is the draw order here and changes no result
// (every sum is exact).  The draws are libstdc++'s (the C++ library this is
// built with); 2 * m0 * (nv - m0) arcs.
int nido_gen_ba(int64_t nv, int64_t m0, nido_hgraph** out)
{
    int rc = check_out(out);
    if (rc) return rc;
    if (m0 < 1 || nv < m0 + 1 || nv >= (1ll << 31) || m0 > (1ll << 20) ||
        (double)m0 * (double)(nv - m0) * 2.0 >= 0x1p40)
        return set_error(NIDO_ERR_INVALID, "invalid BA parameters (nv %lld, m0 %lld): need 1 <= m0 < nv < 2^31",
                         (long long)nv, (long long)m0);
    try {
        struct Tuple {
            int64_t x, y;
            double w;
        };
        const int64_t ne = 2 * m0 * (nv - m0);
        std::vector<Tuple> el((size_t)ne);
        std::default_random_engine rand_gen1, rand_gen2;
        std::normal_distribution<double> gaussian(1.0, 1.0);
        int64_t total = 0;
        for (int64_t i = 0; i < m0; ++i) {
            const double w = std::fabs(gaussian(rand_gen1));
            el[(size_t)total + 0] = {0, i + 1, w};
            el[(size_t)total + 1] = {i + 1, 0, w};
            total += 2;
        }
        for (int64_t i = m0 + 1; i < nv; ++i) {
            std::uniform_int_distribution<int64_t> uniform(0, total - 1);
            for (int64_t j = 0; j < m0; ++j) {
                const int64_t t = el[(size_t)uniform(rand_gen2)].x;
                const double w = std::fabs(gaussian(rand_gen1));
                el[(size_t)(total + 2 * j + 0)] = {i, t, w};
                el[(size_t)(total + 2 * j + 1)] = {t, i, w};
            }
            total += 2 * m0;
        }
        // (tail, head) order; counting sort by tail keeps the draw order
        // within a row, then each row sorted by head
        nido_hgraph* g = new nido_hgraph;
        g->nv = nv;
        g->ne = ne;
        g->off.assign((size_t)nv + 1, 0);
        for (const Tuple& t : el) g->off[(size_t)t.x + 1]++;
        for (int64_t i = 0; i < nv; ++i) g->off[(size_t)i + 1] += g->off[(size_t)i];
        g->col.resize((size_t)ne);
        g->w.resize((size_t)ne);
        {
            std::vector<int64_t> cur(g->off.begin(), g->off.end() - 1);
            for (const Tuple& t : el) {
                const int64_t p = cur[(size_t)t.x]++;
                g->col[(size_t)p] = t.y;
                g->w[(size_t)p] = t.w;
            }
        }
        std::vector<Tuple>().swap(el);
        #pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < nv; ++v) {
            const int64_t a = g->off[(size_t)v], b = g->off[(size_t)v + 1];
            std::vector<std::pair<int64_t, double>> row((size_t)(b - a));
            for (int64_t e = a; e < b; ++e) row[(size_t)(e - a)] = {g->col[(size_t)e], g->w[(size_t)e]};
            std::stable_sort(row.begin(), row.end(),
                             [](const std::pair<int64_t, double>& p, const std::pair<int64_t, double>& q) {
                                 return p.first < q.first;
                             });
            for (int64_t e = a; e < b; ++e) {
                g->col[(size_t)e] = row[(size_t)(e - a)].first;
                g->w[(size_t)e] = row[(size_t)(e - a)].second;
            }
        }
        *out = g;
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

// Graph::coloring, graph.cpp:547-694 (the reference's -c): random weights,
// the greedy colouring (on the device), then the vertices renumbered in
// (colour, id) order with the CSR permuted to match.
//
// Random weights (graph.cpp:558-571): `engines` std::mt19937_64 seeded
// 1 << t, each with a std::uniform_int_distribution<int64_t>(0, 4 nv);
// vertex i takes the next draw of engine i % engines.  With one engine this
// is the reference's sequence exactly; with several, the reference's threads
// share the engines without synchronisation (a data race), and this takes
// each engine's draws in vertex order, the race-free reading.  The colouring
// itself is the reference's loop run by one thread (graph.cpp:573-605):
// colour(i) = the smallest colour not held by a neighbour j < i with
// r_j >= r_i; with several threads the reference's result depends on their
// timing.
int nido_hgraph_coloring(nido_hgraph* g, int engines, int64_t* colors, int64_t* new_orders,
                         int64_t* num_colors)
{
    if (!g || !new_orders) return set_error(NIDO_ERR_INVALID, "null argument");
    if (engines < 1 || engines > 31) return set_error(NIDO_ERR_INVALID, "engines must be in [1, 31]");
    const int64_t nv = g->nv, ne = g->ne;
    if (nv >= (1ll << 31) - 1) return set_error(NIDO_ERR_CAPACITY, "coloring needs fewer than 2^31-1 vertices");
    if (nv == 0) {
        if (num_colors) *num_colors = 0;
        return NIDO_OK;
    }
    try {
        bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
        for (int64_t e = 0; e < ne; ++e) bad = bad || g->col[(size_t)e] < 0 || g->col[(size_t)e] >= nv;
        if (bad)
            for (int64_t e = 0; e < ne; ++e)
                if (g->col[(size_t)e] < 0 || g->col[(size_t)e] >= nv)
                    return set_error(NIDO_ERR_INVALID, "arc %lld: column outside [0, %lld)", (long long)e,
                                     (long long)nv);
        std::vector<int64_t> r((size_t)nv);
        {
            std::vector<std::mt19937_64> eng((size_t)engines);
            for (int t = 0; t < engines; ++t) eng[(size_t)t] = std::mt19937_64(1ull << t);
            std::uniform_int_distribution<int64_t> dist(0, nv * 4);
            for (int64_t i = 0; i < nv; ++i) {
                // stored through a double as the reference does (exact below 2^53)
                const double x = (double)dist(eng[(size_t)(i % engines)]);
                r[(size_t)i] = (int64_t)x;
            }
        }
        std::vector<int> color((size_t)nv);
        int64_t rounds = 0;
        int rc = device_coloring(nv, g->off.data(), g->col.data(), r.data(), color.data(), &rounds);
        if (rc != NIDO_OK) return rc;
        // (colour, id) order, graph.cpp:609-629: a counting sort by colour
        // keeps ids ascending within a colour
        int maxc = 0;
        for (int64_t i = 0; i < nv; ++i) maxc = std::max(maxc, color[(size_t)i]);
        const int64_t K = (int64_t)maxc + 1;
        std::vector<int64_t> start((size_t)K + 1, 0);
        for (int64_t i = 0; i < nv; ++i) start[(size_t)color[(size_t)i] + 1]++;
        for (int64_t c = 0; c < K; ++c) start[(size_t)c + 1] += start[(size_t)c];
        std::vector<int64_t> orders((size_t)nv);
        for (int64_t i = 0; i < nv; ++i) {
            const int64_t pos = start[(size_t)color[(size_t)i]]++;
            orders[(size_t)i] = pos;
            new_orders[pos] = i;
        }
        if (colors)
            for (int64_t i = 0; i < nv; ++i) colors[i] = color[(size_t)i];
        // renumbered CSR, graph.cpp:631-691: row pos is old row new_orders[pos]
        // with its arcs in their order, every column mapped through orders
        hvec<int64_t> noff((size_t)nv + 1, 0);
        for (int64_t p = 0; p < nv; ++p) {
            const int64_t v = new_orders[p];
            noff[(size_t)p + 1] = noff[(size_t)p] + (g->off[(size_t)v + 1] - g->off[(size_t)v]);
        }
        hvec<int64_t> ncol((size_t)ne);
        hvec<double> nw((size_t)ne);
        #pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t p = 0; p < nv; ++p) {
            const int64_t v = new_orders[p];
            const int64_t a = g->off[(size_t)v], d = g->off[(size_t)v + 1] - a, b = noff[(size_t)p];
            for (int64_t k = 0; k < d; ++k) {
                ncol[(size_t)(b + k)] = orders[(size_t)g->col[(size_t)(a + k)]];
                nw[(size_t)(b + k)] = g->w[(size_t)(a + k)];
            }
        }
        g->off.swap(noff);
        g->col.swap(ncol);
        g->w.swap(nw);
        if (num_colors) *num_colors = K;
    } catch (const std::exception& e) {   // bad_alloc, or a size no vector can hold (length_error)
        return set_error(NIDO_ERR_NOMEM, "host allocation failed: %s", e.what());
    }
    return NIDO_OK;
}

}  // extern "C"

// ---- checks and sums used by the device handle's constructors ------------
namespace nido {

int check_offsets(int64_t nv, const int64_t* off, const int64_t* col, const double* w)
{
    int64_t bad_off = INT64_MAX;
#pragma omp parallel for reduction(min : bad_off) schedule(static)
    for (int64_t i = 0; i < nv; ++i)
        if (off[i + 1] < off[i] && i < bad_off) bad_off = i;
    if (bad_off != INT64_MAX) return set_error(NIDO_ERR_INVALID, "offsets decrease at %lld", (long long)bad_off);
    if (off[nv] > off[0] && (!col || !w)) return set_error(NIDO_ERR_INVALID, "null cols/weights");
    return NIDO_OK;
}

int check_csr(int64_t nv, const int64_t* off, const int64_t* col, const double* w, int G,
              const int64_t* vptr, bool* nonneg, fx_t* twom)
{
    const int64_t base = off[0];
    {
        const int rc = check_offsets(nv, off, col, w);
        if (rc != NIDO_OK) return rc;
    }
    // first bad arc of graph g: column outside the graph, weight not finite
    // or too large; and the graph's exact total of |fx(w)| (saturating at
    // 2^126 units) and of fx(w).  One parallel loop over the graphs when
    // there are many, over each graph's arcs otherwise.
    int neg = 0;
    const ufx_t cap = (ufx_t)1 << 126;   // 2^62 in units of 2^-64
    auto sat_add = [&](ufx_t x, ufx_t y) { return x >= cap || y >= cap || x + y >= cap ? cap : x + y; };
    auto scan = [&](int g, int64_t e0, int64_t e1, int64_t& bad_col, int64_t& bad_w, int& ng, ufx_t& sabs,
                    fx_t& ssum, bool par) {
        const int64_t lo = vptr[g], hi = vptr[g + 1];
        int64_t bc = INT64_MAX, bw = INT64_MAX;
        int nn = 0;
        ufx_t sa = 0, ss = 0;   // ss: fx(w) summed modulo 2^128
#pragma omp parallel if (par)
        {
            int64_t tbc = INT64_MAX, tbw = INT64_MAX;
            int tn = 0;
            ufx_t ta = 0, ts = 0;
#pragma omp for schedule(static) nowait
            for (int64_t e = e0; e < e1; ++e) {
                const int64_t c = col[base + e];
                if ((c < lo || c >= hi) && e < tbc) tbc = e;
                const double x = w[base + e];
                if (!(x == x) || x >= 0x1p62 || x <= -0x1p62) {
                    if (e < tbw) tbw = e;
                    continue;
                }
                tn |= x < 0;
                const fx_t f = fx_from_double(x);
                ta = sat_add(ta, f < 0 ? (ufx_t)(-f) : (ufx_t)f);
                ts += (ufx_t)f;   // exact while the total of |w| stays below 2^62
            }
#pragma omp critical
            {
                bc = std::min(bc, tbc);
                bw = std::min(bw, tbw);
                nn |= tn;
                sa = sat_add(sa, ta);
                ss += ts;
            }
        }
        bad_col = bc;
        bad_w = bw;
        ng = nn;
        sabs = sa;
        ssum = (fx_t)ss;
    };
    std::vector<int64_t> gbc((size_t)G), gbw((size_t)G);
    std::vector<int> gneg((size_t)G);
    std::vector<ufx_t> gabs((size_t)G);
    std::vector<fx_t> gsum((size_t)G);
    const bool many = G >= 64;
#pragma omp parallel for schedule(dynamic, 16) if (many)
    for (int g = 0; g < G; ++g)
        scan(g, off[vptr[g]] - base, off[vptr[g + 1]] - base, gbc[(size_t)g], gbw[(size_t)g], gneg[(size_t)g],
             gabs[(size_t)g], gsum[(size_t)g], !many);
    for (int g = 0; g < G; ++g) {
        const int64_t bad_col = gbc[(size_t)g], bad_w = gbw[(size_t)g];
        neg |= gneg[(size_t)g];
        if (bad_col != INT64_MAX && G > 1)
            return set_error(NIDO_ERR_INVALID, "arc %lld (col %lld) leaves graph %d", (long long)bad_col,
                             (long long)col[base + bad_col], g);
        if (bad_col != INT64_MAX)
            return set_error(NIDO_ERR_INVALID, "col[%lld]=%lld out of range", (long long)bad_col,
                             (long long)col[base + bad_col]);
        if (bad_w != INT64_MAX)
            return set_error(NIDO_ERR_INVALID, "weight[%lld] is not finite or |w| >= 2^62", (long long)bad_w);
        // every exact sum (k_v, e_vc, a_c, 2m, coarse weights, the internal
        // weight) is bounded by the total of |w|: below 2^62 it fits Q64.64
        if (gabs[(size_t)g] >= cap)
            return set_error(NIDO_ERR_INVALID, "graph %d: the total of |w| over its arcs reaches 2^62", g);
        // with negative weights |a_c / 2m| <= sum|w| / |2m|: below 2^31 the
        // square terms of the modularity stay below 2^62
        const fx_t t = gsum[(size_t)g];
        const ufx_t tabs = t < 0 ? (ufx_t)(-t) : (ufx_t)t;
        if (gneg[(size_t)g] && tabs != 0 && (gabs[(size_t)g] >> 31) >= tabs)
            return set_error(NIDO_ERR_INVALID,
                             "graph %d: the total of |w| is 2^31 or more times |2m| (negative weights)", g);
    }
    if (nonneg) *nonneg = !neg;
    if (twom) std::copy(gsum.begin(), gsum.end(), twom);   // the exact sum of a graph's weights is its 2m
    return NIDO_OK;
}

}  // namespace nido

// ---- the Q64.64 conversions of fixed.h (host build) -----------------------
extern "C" {
void nido_fx_from_double(double w, uint64_t* lo, int64_t* hi)
{
    const fx_t x = fx_from_double(w);
    *lo = fx_lo(x);
    *hi = fx_hi(x);
}

double nido_fx_to_double(uint64_t lo, int64_t hi) { return fx_to_double(fx_make(lo, hi)); }

void nido_div_rcp(int64_t n, const double* a, const double* b, double* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = div_rcp(a[i], b[i], 1.0 / b[i]);
}
}
