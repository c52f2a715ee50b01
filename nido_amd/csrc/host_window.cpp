// host_window.cpp -- host side of phases whose arcs stream through a device
// window (graphs larger than the device arc budget): the window units of
// each batch, the exact vertex weights, and the coarse graph of a windowed
// phase, all on the host CSR with OpenMP.
//
// The reference streams such graphs the same way: the device holds
// ne_per_partition arcs, each batch is cut into sub-partitions that fit
// (graph_gpu.cpp:341-359, 362-417) whose arcs are copied in before their
// kernels run (graph_gpu.cpp:967-1027), and the coarse graph is built on the
// host (graph_gpu.cpp:1088-1304, 1382-1545).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "errors.h"
#include "host_bins.h"

namespace nido {

int window_units(const int64_t* off, const std::vector<int64_t>& parts, int nb, int64_t window,
                 std::vector<int64_t>& units, std::vector<int>& first_unit)
{
    units.assign(1, parts[0]);
    first_unit.assign((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        first_unit[(size_t)b] = (int)units.size() - 1;
        const int64_t v0 = parts[(size_t)b], v1 = parts[(size_t)b + 1];
        int64_t start = v0;
        for (int64_t v = v0; v < v1; ++v) {
            const int64_t d = off[v + 1] - off[v];
            if (d > window)
                return set_error(NIDO_ERR_CAPACITY, "row %lld has %lld arcs, more than the %lld-arc device window",
                                 (long long)v, (long long)d, (long long)window);
            if (off[v + 1] - off[start] > window) {   // close the unit before v
                units.push_back(v);
                start = v;
            }
        }
        if (units.back() != v1 || v1 == v0) units.push_back(v1);
    }
    first_unit[(size_t)nb] = (int)units.size() - 1;
    return NIDO_OK;
}

namespace {
// splitmix64's output function (Vigna's public-domain splitmix64.c).
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint64_t arc_hash(uint32_t a, uint32_t b, uint64_t wb, uint64_t key)
{
    return mix64(mix64((((uint64_t)a << 32) | b) ^ key) ^ wb);
}
}  // namespace

bool host_arcs_symmetric(int64_t nv, const int64_t* off, const int* col, const double* w)
{
    constexpr uint64_t k1 = 0x6a09e667f3bcc909ull, k2 = 0xbb67ae8584caa73bull;   // sqrt(2), sqrt(3) fractions
    uint64_t h0 = 0, h1 = 0, h2 = 0, h3 = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : h0, h1, h2, h3)
    for (int64_t v = 0; v < nv; ++v)
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            const uint32_t a = (uint32_t)v, b = (uint32_t)col[e];
            uint64_t wb;
            std::memcpy(&wb, &w[e], sizeof(wb));
            h0 += arc_hash(a, b, wb, k1);
            h1 += arc_hash(b, a, wb, k1);
            h2 += arc_hash(a, b, wb, k2);
            h3 += arc_hash(b, a, wb, k2);
        }
    return h0 == h1 && h2 == h3;
}

void host_vertex_weights(int64_t nv, const int64_t* off, const double* w, fx_t* k, double* kd, fx_t* twom,
                         double* maxk)
{
    fx_t tot = 0;
    double mx = 0.0;
#pragma omp parallel
    {
        fx_t t = 0;
        double m = 0.0;
#pragma omp for schedule(dynamic, 4096) nowait
        for (int64_t v = 0; v < nv; ++v) {
            fx_t s = 0;
            for (int64_t e = off[v]; e < off[v + 1]; ++e) s += fx_from_double(w[e]);
            k[v] = s;
            kd[v] = fx_to_double(s);
            t += s;
            m = std::max(m, std::fabs(kd[v]));
        }
#pragma omp critical
        {
            tot += t;
            mx = std::max(mx, m);
        }
    }
    *twom = tot;
    *maxk = mx;
}

void host_coarsen(int64_t nv, const int64_t* off, const int* col, const double* w, const int64_t* C,
                  const int64_t* newid, int64_t K, std::vector<int64_t>& noff, std::vector<int>& ncol,
                  std::vector<double>& nw)
{
    // members of every coarse vertex, ascending
    std::vector<int64_t> mstart((size_t)K + 1, 0), members((size_t)nv);
    for (int64_t v = 0; v < nv; ++v) mstart[(size_t)newid[C[v]] + 1]++;
    for (int64_t i = 0; i < K; ++i) mstart[(size_t)i + 1] += mstart[(size_t)i];
    {
        std::vector<int64_t> cur(mstart.begin(), mstart.end() - 1);
        for (int64_t v = 0; v < nv; ++v) members[(size_t)cur[(size_t)newid[C[v]]]++] = v;
    }
    // per coarse row: (target, exact sum) merged by target, ascending
    std::vector<std::vector<std::pair<int, fx_t>>> rows((size_t)K);
#pragma omp parallel
    {
        std::vector<std::pair<int, fx_t>> buf;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < K; ++i) {
            buf.clear();
            for (int64_t m = mstart[(size_t)i]; m < mstart[(size_t)i + 1]; ++m) {
                const int64_t v = members[(size_t)m];
                for (int64_t e = off[v]; e < off[v + 1]; ++e)
                    buf.emplace_back((int)newid[C[col[e]]], fx_from_double(w[e]));
            }
            std::sort(buf.begin(), buf.end(),
                      [](const std::pair<int, fx_t>& a, const std::pair<int, fx_t>& b) { return a.first < b.first; });
            auto& r = rows[(size_t)i];
            for (size_t j = 0; j < buf.size();) {
                fx_t s = 0;
                size_t k = j;
                for (; k < buf.size() && buf[k].first == buf[j].first; ++k) s += buf[k].second;
                r.emplace_back(buf[j].first, s);
                j = k;
            }
        }
    }
    noff.assign((size_t)K + 1, 0);
    for (int64_t i = 0; i < K; ++i) noff[(size_t)i + 1] = noff[(size_t)i] + (int64_t)rows[(size_t)i].size();
    ncol.resize((size_t)noff[(size_t)K]);
    nw.resize((size_t)noff[(size_t)K]);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < K; ++i) {
        int64_t p = noff[(size_t)i];
        for (const auto& x : rows[(size_t)i]) {
            ncol[(size_t)p] = x.first;
            nw[(size_t)p] = fx_to_double(x.second);
            ++p;
        }
        std::vector<std::pair<int, fx_t>>().swap(rows[(size_t)i]);
    }
}

}  // namespace nido
