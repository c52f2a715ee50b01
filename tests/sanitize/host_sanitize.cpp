// The library's host-side code and the CPU oracle, built with g++/gcc under
// AddressSanitizer and UndefinedBehaviorSanitizer (no GPU code, no HIP
// runtime): the graph generators, the file reader and writer, the CSR
// checks (valid and malformed input), batch bounds, degree bins, edge-window
// units, the arc-symmetry hash, exact vertex weights, host coarsening, the
// colouring's renumbering, and full oracle runs on the same graphs.
//
// device_coloring (coloring.hip) is the one GPU routine the host code calls;
// this build defines it as a sequential greedy colouring in id order, which
// is a valid colouring for the renumbering that follows it.
//
// Usage: host_sanitize <scratch-dir>; prints "sanitize ok" and exits 0 when
// every check holds (a sanitizer report aborts the run with a non-zero code).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/nido_amd.h"
#include "../../nido_amd/csrc/host_bins.h"
#include "../../nido_amd/csrc/host_check.h"

extern "C" {
typedef struct orc_ctx orc_ctx;
orc_ctx* orc_create(int64_t nv, int64_t ne, const int64_t* off, const int64_t* col, const double* w, int nbatches,
                    int nthreads);
void orc_destroy(orc_ctx* c);
void orc_run(orc_ctx* c, int64_t max_loops, double tol, double tol_phase, int record);
int orc_hist_phases(const orc_ctx* c);
int orc_hist_comm(const orc_ctx* c, int i, int64_t* out);
void orc_get_clustering(const orc_ctx* c, int64_t* out);
}

namespace nido {
int device_coloring(int64_t nv, const int64_t* off, const int64_t* col, const int64_t* r, int* colors,
                    int64_t* rounds)
{
    (void)r;
    std::vector<char> used;
    for (int64_t i = 0; i < nv; ++i) {
        used.assign((size_t)(off[i + 1] - off[i]) + 2, 0);
        for (int64_t e = off[i]; e < off[i + 1]; ++e)
            if (col[e] < i && (size_t)colors[col[e]] < used.size()) used[(size_t)colors[col[e]]] = 1;
        int c = 0;
        while (used[(size_t)c]) ++c;
        colors[i] = c;
    }
    *rounds = nv;
    return NIDO_OK;
}
}  // namespace nido

static int g_fails = 0;
#define CHECK(c)                                                                   \
    do {                                                                           \
        if (!(c)) {                                                                \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);           \
            ++g_fails;                                                             \
        }                                                                          \
    } while (0)

static nido_tuning default_knobs()
{
    nido_tuning t{};
    t.thread_max_deg = nido::kThreadMaxDeg;
    t.big_min_deg = (int32_t)nido::kBinMaxDeg[nido::BIN_B256];
    t.big_part_arcs = (int32_t)nido::kBigPartArcs;
    t.big_key_limit = nido::kBigLimit;
    t.use_graph = 1;
    t.big_slots = nido::kBigSlots;
    t.overlap_long = 1;
    t.fixed_short = 1;
    return t;
}

static void exercise(const char* name, nido_hgraph* h, const std::string& dir)
{
    int64_t nv = 0, ne = 0;
    CHECK(nido_hgraph_info(h, &nv, &ne) == NIDO_OK);
    const int64_t *off = nullptr, *col = nullptr;
    const double* w = nullptr;
    CHECK(nido_hgraph_arrays(h, &off, &col, &w) == NIDO_OK);
    printf("graph %s: %lld vertices, %lld arcs\n", name, (long long)nv, (long long)ne);

    // file round trip
    const std::string path = dir + "/" + name + ".bin";
    CHECK(nido_hgraph_write(h, path.c_str()) == NIDO_OK);
    nido_hgraph* back = nullptr;
    CHECK(nido_hgraph_read(path.c_str(), &back) == NIDO_OK);
    if (back) {
        int64_t nv2 = 0, ne2 = 0;
        const int64_t *off2 = nullptr, *col2 = nullptr;
        const double* w2 = nullptr;
        nido_hgraph_info(back, &nv2, &ne2);
        nido_hgraph_arrays(back, &off2, &col2, &w2);
        CHECK(nv2 == nv && ne2 == ne);
        CHECK(std::equal(off, off + nv + 1, off2) && std::equal(col, col + ne, col2) && std::equal(w, w + ne, w2));
        nido_hgraph_destroy(back);
    }

    // CSR checks: the graph passes, one malformed copy at a time fails
    bool nonneg = false;
    fx_t twom = 0;
    const int64_t whole[2] = {0, nv};
    CHECK(nido::check_csr(nv, off, col, w, 1, whole, &nonneg, &twom) == NIDO_OK);
    if (ne > 0) {
        std::vector<int64_t> bad(col, col + ne);
        bad[(size_t)ne / 2] = nv;
        CHECK(nido::check_csr(nv, off, bad.data(), w, 1, whole, &nonneg, &twom) == NIDO_ERR_INVALID);
        std::vector<double> wn(w, w + ne);
        wn[(size_t)ne / 3] = std::nan("");
        CHECK(nido::check_csr(nv, off, col, wn.data(), 1, whole, &nonneg, &twom) == NIDO_ERR_INVALID);
        std::vector<int64_t> offb(off, off + nv + 1);
        offb[(size_t)nv] = ne + 1;
        nido_hgraph* rej = nullptr;
        CHECK(nido_hgraph_from_csr(nv, ne, offb.data(), col, w, &rej) != NIDO_OK && rej == nullptr);
    }

    // batches, bins and window units
    const nido_tuning t = default_knobs();
    for (int nb : {1, 3, 8}) {
        std::vector<int64_t> parts;
        nido::batch_bounds(off, nv, nb, parts);
        CHECK((int)parts.size() == nb + 1 && parts.front() == 0 && parts.back() == nv);
        for (int b = 0; b < nb; ++b) CHECK(parts[(size_t)b] <= parts[(size_t)b + 1]);
        nido::PhaseBins pb;
        nido::build_phase_bins(off, nv, nb, 1, parts, std::vector<int64_t>(), t, pb);
        int64_t nonempty = 0;   // isolated vertices are left out of the bins
        for (int64_t v = 0; v < nv; ++v) nonempty += off[v + 1] > off[v];
        CHECK(pb.rows == nonempty && (int64_t)pb.binv.size() >= nonempty);
        const int64_t maxd = nido::max_degree(off, nv);
        for (int64_t window : {std::max<int64_t>(maxd, 1), std::max<int64_t>(ne / 3, maxd), ne}) {
            std::vector<int64_t> units;
            std::vector<int> first;
            if (window <= 0) continue;
            CHECK(nido::window_units(off, parts, nb, window, units, first) == NIDO_OK);
            CHECK((int)first.size() == nb + 1 && units.front() == 0 && units.back() == nv);
            for (size_t u = 0; u + 1 < units.size(); ++u) CHECK(off[units[u + 1]] - off[units[u]] <= window);
        }
        if (maxd > 1) {
            std::vector<int64_t> units;
            std::vector<int> first;
            CHECK(nido::window_units(off, parts, nb, maxd - 1, units, first) == NIDO_ERR_CAPACITY);
        }
    }

    // host-side arrays of a windowed phase
    std::vector<int> col32(col, col + ne);
    CHECK(nido::host_arcs_symmetric(nv, off, col32.data(), w));
    std::vector<fx_t> k((size_t)nv);
    std::vector<double> kd((size_t)nv);
    fx_t twom2 = 0;
    double maxk = 0.0;
    nido::host_vertex_weights(nv, off, w, k.data(), kd.data(), &twom2, &maxk);
    CHECK(twom2 == twom);

    // a full oracle run, then the coarse graph of its first phase
    orc_ctx* c = orc_create(nv, ne, off, col, w, 3, 2);
    CHECK(c != nullptr);
    if (!c) return;
    orc_run(c, 500, 1e-6, 1e-6, 1);
    std::vector<int64_t> clus((size_t)nv);
    orc_get_clustering(c, clus.data());
    for (int64_t v = 0; v < nv; ++v) CHECK(clus[(size_t)v] >= 0 && clus[(size_t)v] < std::max<int64_t>(nv, 1));
    if (orc_hist_phases(c) > 0 && nv > 0) {
        std::vector<int64_t> C((size_t)nv);
        CHECK(orc_hist_comm(c, 0, C.data()) >= 0);
        std::vector<int64_t> newid((size_t)nv, -1);
        int64_t K = 0;
        for (int64_t v = 0; v < nv; ++v) newid[(size_t)C[(size_t)v]] = 0;
        for (int64_t x = 0; x < nv; ++x)
            if (newid[(size_t)x] == 0) newid[(size_t)x] = K++;
        std::vector<int64_t> noff;
        std::vector<int> ncol;
        std::vector<double> nw;
        nido::host_coarsen(nv, off, col32.data(), w, C.data(), newid.data(), K, noff, ncol, nw);
        CHECK((int64_t)noff.size() == K + 1 && noff[(size_t)K] <= ne && (int64_t)ncol.size() == noff[(size_t)K]);
        double s0 = 0.0, s1 = 0.0;
        for (int64_t e = 0; e < ne; ++e) s0 += w[e];
        for (double x : nw) s1 += x;
        CHECK(std::fabs(s0 - s1) <= 1e-9 * std::max(1.0, std::fabs(s0)));
        for (int64_t r = 0; r < K; ++r)
            for (int64_t e = noff[(size_t)r]; e < noff[(size_t)r + 1]; ++e) {
                CHECK(ncol[(size_t)e] >= 0 && ncol[(size_t)e] < K);
                if (e > noff[(size_t)r]) CHECK(ncol[(size_t)e - 1] < ncol[(size_t)e]);
            }
    }
    orc_destroy(c);
}

static void colour(nido_hgraph* h)
{
    int64_t nv = 0, ne = 0;
    nido_hgraph_info(h, &nv, &ne);
    std::vector<int64_t> colors((size_t)nv), order((size_t)nv);
    int64_t K = 0;
    CHECK(nido_hgraph_coloring(h, 4, colors.data(), order.data(), &K) == NIDO_OK);
    std::vector<int64_t> sorted(order);
    std::sort(sorted.begin(), sorted.end());
    for (int64_t i = 0; i < nv; ++i) CHECK(sorted[(size_t)i] == i);
    int64_t nv2 = 0, ne2 = 0;
    nido_hgraph_info(h, &nv2, &ne2);
    CHECK(nv2 == nv && ne2 == ne);
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <scratch-dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    struct Gen {
        const char* name;
        int rc;
        nido_hgraph* h;
    } gens[5] = {};
    gens[0].name = "sbm";
    gens[0].rc = nido_gen_sbm(3000, 60, 0.2, 0.005, 1, 11, &gens[0].h);
    gens[1].name = "rmat";
    gens[1].rc = nido_gen_rmat(12, 16, 0.57, 0.19, 0.19, 1, 3, &gens[1].h);
    gens[2].name = "chung_lu";
    gens[2].rc = nido_gen_chung_lu(4000, 2.1, 2.0, 400.0, 8.0, 1, 5, &gens[2].h);
    gens[3].name = "ba";
    gens[3].rc = nido_gen_ba(3000, 4, &gens[3].h);
    gens[4].name = "empty";
    {
        const int64_t off0[4] = {0, 0, 0, 0};
        gens[4].rc = nido_hgraph_from_csr(3, 0, off0, nullptr, nullptr, &gens[4].h);
    }
    for (Gen& g : gens) {
        CHECK(g.rc == NIDO_OK && g.h != nullptr);
        if (g.rc != NIDO_OK || !g.h) continue;
        exercise(g.name, g.h, dir);
        colour(g.h);
        nido_hgraph_destroy(g.h);
    }

    // total weight beyond what Q64.64 sums hold: a path 0-1-2 with 3e18 weights
    {
        const int64_t off[4] = {0, 1, 3, 4}, col[4] = {1, 0, 2, 1};
        const double w[4] = {3e18, 3e18, 3e18, 3e18};
        const int64_t whole[2] = {0, 3};
        bool nonneg = false;
        fx_t twom = 0;
        CHECK(nido::check_csr(3, off, col, w, 1, whole, &nonneg, &twom) == NIDO_ERR_INVALID);
    }
    // a missing and a truncated file
    {
        nido_hgraph* h = nullptr;
        CHECK(nido_hgraph_read((dir + "/missing.bin").c_str(), &h) != NIDO_OK && h == nullptr);
        FILE* f = fopen((dir + "/short.bin").c_str(), "wb");
        const int64_t hdr[2] = {1000, 5000};
        if (f) {
            fwrite(hdr, sizeof hdr, 1, f);
            fclose(f);
        }
        CHECK(nido_hgraph_read((dir + "/short.bin").c_str(), &h) != NIDO_OK && h == nullptr);
    }
    if (g_fails) {
        fprintf(stderr, "%d check(s) failed\n", g_fails);
        return 1;
    }
    printf("sanitize ok\n");
    return 0;
}
