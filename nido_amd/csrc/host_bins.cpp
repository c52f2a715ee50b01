// host_bins.cpp -- per-phase work lists of the local-moving launches
// (GraphGPU ctor steps graph_gpu.cpp:94-139 and aggregation steps
// graph_gpu.cpp:1487-1516, in this library's layout), built on the host with
// OpenMP: one task per batch, merged in batch order, so the lists are the
// same whatever the thread count.
#include "host_bins.h"

#include <omp.h>

#include <algorithm>

namespace nido {
namespace {

int bin_of(int64_t d, const nido_tuning& t)
{
    if (d > t.big_min_deg) return BIN_BIG;
    if (d <= t.thread_max_deg) return BIN_THREAD;
    for (int b = BIN_G8; b < NUM_SMALL_BINS; ++b)
        if (d <= kBinMaxDeg[b]) return b;
    return BIN_BIG;
}

int64_t pow2_at_least(int64_t x)
{
    int64_t s = 2;
    while (s < x) s <<= 1;
    return s;
}

struct BatchWork {
    int64_t cnt[NUM_BINS] = {};
    std::vector<BinRow> rows;
    std::vector<BigItem> items;
    std::vector<KeyChunk> chunks;
    std::vector<int64_t> poff, toff;
    double sbytes = 0, lbytes = 0;
    int64_t ksum = 0, csum = 0, psum = 0, tsum = 0;
    int slot = 0;
};

}  // namespace

void batch_bounds(const int64_t* off, int64_t nv, int nb, std::vector<int64_t>& parts)
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

void batch_bounds_all(const int64_t* off, int G, const int64_t* vb, int nb, std::vector<int64_t>& gparts)
{
    const size_t nb1 = (size_t)nb + 1;
    gparts.assign((size_t)G * nb1, 0);
#pragma omp parallel
    {
        std::vector<int64_t> p;
#pragma omp for schedule(dynamic, 64)
        for (int g = 0; g < G; ++g) {
            const int64_t v0 = vb[g], n = vb[g + 1] - v0;
            batch_bounds(off + v0, n, nb, p);
            for (size_t i = 0; i < nb1; ++i) gparts[(size_t)g * nb1 + i] = v0 + p[i];
        }
    }
}

int64_t max_degree(const int64_t* off, int64_t nv)
{
    int64_t m = 0;
#pragma omp parallel for reduction(max : m) schedule(static) if (nv >= kParallelRows)
    for (int64_t v = 0; v < nv; ++v) m = std::max(m, off[v + 1] - off[v]);
    return m;
}

namespace {

// Long row v (d arcs) of a batch: P = 2^ceil(log2(d / part_arcs)) partition
// items, their partial slots, an HBM fallback table, and for P >
// kDirectParts the key-fill chunks and bucket space.
void add_long_row(BatchWork& w, int64_t v, int64_t d, const int64_t* off, const nido_tuning& t)
{
    int P = 1;
    while (P < kMaxParts && (int64_t)P * t.big_part_arcs < d) P <<= 1;
    const bool direct = P <= kDirectParts;
    for (int p = 0; p < P; ++p)
        w.items.push_back(BigItem{(int)v, p, P, w.slot, w.ksum, direct ? -1 : w.csum, off[v], d});
    if (!direct) {
        for (int64_t e = off[v]; e < off[v + 1]; e += kKeyChunkArcs) {
            w.chunks.push_back(KeyChunk{(int)v, P, e, std::min(off[v + 1], e + kKeyChunkArcs),
                                        w.ksum + (e - off[v]), w.csum});
            w.csum += P + 1;
        }
        w.ksum += d;
    }
    w.poff.push_back(w.psum);
    w.psum += P;
    w.tsum += pow2_at_least(2 * d);
    w.toff.push_back(w.tsum);
    ++w.slot;
}

// Longest rows first: their partition workgroups are the long poles.
void order_items(BatchWork& w, const int64_t* off)
{
    std::stable_sort(w.items.begin(), w.items.end(), [&](const BigItem& x, const BigItem& y) {
        return off[x.v + 1] - off[x.v] > off[y.v + 1] - off[y.v];
    });
}

// Batch offsets of every list, in batch order; binv is filled from the
// batches' rows when they were built on the host.
void merge_batches(const std::vector<BatchWork>& W, int nb, bool rows_on_host, PhaseBins& pb)
{
    const size_t nb1 = (size_t)nb + 1;
    pb.item_base.assign((size_t)nb, 0);
    pb.poff_base.assign((size_t)nb, 0);
    pb.toff_base.assign((size_t)nb, 0);
    pb.chunk_base.assign(nb1, 0);
    pb.small_start.assign((size_t)nb * NUM_SMALL_BINS, 0);
    pb.small_n.assign((size_t)nb * NUM_SMALL_BINS, 0);
    pb.blist0.assign((size_t)nb, 0);
    pb.blistn.assign((size_t)nb, 0);
    pb.short_bytes.assign((size_t)nb, 0.0);
    pb.long_bytes.assign((size_t)nb, 0.0);
    pb.tab_need = pb.part_need = pb.slot_need = pb.key_need = pb.cpo_need = 0;
    int64_t rows = 0, items = 0, chunks = 0, poffs = 0, toffs = 0;
    for (int b = 0; b < nb; ++b) {
        const BatchWork& w = W[(size_t)b];
        pb.blist0[(size_t)b] = rows;
        int64_t p = rows, n = 0;
        for (int k = 0; k < NUM_SMALL_BINS; ++k) {
            pb.small_start[(size_t)b * NUM_SMALL_BINS + k] = p;
            pb.small_n[(size_t)b * NUM_SMALL_BINS + k] = w.cnt[k];
            p += w.cnt[k];
        }
        for (int k = 0; k < NUM_BINS; ++k) n += w.cnt[k];
        pb.blistn[(size_t)b] = n;
        pb.item_base[(size_t)b] = items;
        pb.poff_base[(size_t)b] = poffs;
        pb.toff_base[(size_t)b] = toffs;
        pb.chunk_base[(size_t)b] = chunks;
        pb.short_bytes[(size_t)b] = w.sbytes;
        pb.long_bytes[(size_t)b] = w.lbytes;
        rows += n;
        items += (int64_t)w.items.size();
        chunks += (int64_t)w.chunks.size();
        poffs += (int64_t)w.poff.size();
        toffs += (int64_t)w.toff.size();
        pb.tab_need = std::max(pb.tab_need, w.tsum);
        pb.part_need = std::max(pb.part_need, w.psum);
        pb.slot_need = std::max<int64_t>(pb.slot_need, w.slot);
        pb.key_need = std::max(pb.key_need, w.ksum);
        pb.cpo_need = std::max(pb.cpo_need, w.csum);
    }
    pb.chunk_base[(size_t)nb] = chunks;
    pb.rows = rows;
    pb.binv.clear();
    if (rows_on_host) pb.binv.resize((size_t)std::max<int64_t>(rows, 1));
    pb.bitems.resize((size_t)items);
    pb.chunks.resize((size_t)chunks);
    pb.poff.resize((size_t)poffs);
    pb.toff.resize((size_t)toffs);
#pragma omp parallel for schedule(dynamic, 1) if (rows >= kParallelRows)
    for (int b = 0; b < nb; ++b) {
        const BatchWork& w = W[(size_t)b];
        if (rows_on_host) std::copy(w.rows.begin(), w.rows.end(), pb.binv.begin() + pb.blist0[(size_t)b]);
        std::copy(w.items.begin(), w.items.end(), pb.bitems.begin() + pb.item_base[(size_t)b]);
        std::copy(w.chunks.begin(), w.chunks.end(), pb.chunks.begin() + pb.chunk_base[(size_t)b]);
        std::copy(w.poff.begin(), w.poff.end(), pb.poff.begin() + pb.poff_base[(size_t)b]);
        std::copy(w.toff.begin(), w.toff.end(), pb.toff.begin() + pb.toff_base[(size_t)b]);
    }
}

}  // namespace

void build_phase_bins(const int64_t* off, int64_t nv, int nb, int G, const std::vector<int64_t>& parts,
                      const std::vector<int64_t>& gparts, const nido_tuning& t, PhaseBins& pb,
                      const unsigned char* keep)
{
    const size_t nb1 = (size_t)nb + 1;
    auto for_batch = [&](int b, auto&& fn) {
        if (G == 1) {
            fn(parts[(size_t)b], parts[(size_t)b + 1]);
            return;
        }
        for (int gg = 0; gg < G; ++gg)
            if (!keep || keep[gg]) fn(gparts[(size_t)gg * nb1 + b], gparts[(size_t)gg * nb1 + b + 1]);
    };
    std::vector<BatchWork> W((size_t)nb);
    // Short rows: vertices grouped by (batch, bin), ascending id inside a
    // group, then the batch's long rows; isolated vertices are left out
    // (they never move, and C' = C holds for them throughout), so a batch's
    // slice of binv lists every vertex its commit has to visit.
#pragma omp parallel for schedule(dynamic, 1) if (nv >= kParallelRows)
    for (int b = 0; b < nb; ++b) {
        BatchWork& w = W[(size_t)b];
        for_batch(b, [&](int64_t v0, int64_t v1) {
            for (int64_t v = v0; v < v1; ++v)
                if (off[v + 1] > off[v]) w.cnt[bin_of(off[v + 1] - off[v], t)]++;
        });
        int64_t cur[NUM_BINS];
        int64_t n = 0;
        for (int k = 0; k < NUM_BINS; ++k) {
            cur[k] = n;
            n += w.cnt[k];
        }
        w.rows.resize((size_t)n);
        w.toff.push_back(0);
        for_batch(b, [&](int64_t v0, int64_t v1) {
            for (int64_t v = v0; v < v1; ++v) {
                const int64_t d = off[v + 1] - off[v];
                const int k = bin_of(d, t);
                // algorithmic bytes of the sweep: 24 per arc (col, weight,
                // C[col]) + 40 per vertex (offset, C, k, a_C, C')
                (k != BIN_BIG ? w.sbytes : w.lbytes) += 24.0 * (double)d + 40.0;
                if (d == 0) continue;
                w.rows[(size_t)cur[k]++] = BinRow{(int)v, (int)std::min<int64_t>(d, INT32_MAX), off[v]};
                if (k == BIN_BIG) add_long_row(w, v, d, off, t);
            }
        });
        order_items(w, off);
    }
    merge_batches(W, nb, true, pb);
}

void bin_tiles(int nb, int G, const std::vector<int64_t>& parts, const std::vector<int64_t>& gparts,
               const unsigned char* keep, std::vector<BinTile>& tiles, std::vector<int64_t>& tb0)
{
    const size_t nb1 = (size_t)nb + 1;
    tiles.clear();
    tb0.assign(nb1, 0);
    for (int b = 0; b < nb; ++b) {
        tb0[(size_t)b] = (int64_t)tiles.size();
        auto add = [&](int64_t v0, int64_t v1) {
            for (int64_t v = v0; v < v1; v += kBinTileVerts)
                tiles.push_back(BinTile{v, std::min(v1, v + kBinTileVerts), 0, 0});
        };
        if (G == 1) add(parts[(size_t)b], parts[(size_t)b + 1]);
        else
            for (int gg = 0; gg < G; ++gg)
                if (!keep || keep[gg]) add(gparts[(size_t)gg * nb1 + b], gparts[(size_t)gg * nb1 + b + 1]);
        const int64_t t0 = tb0[(size_t)b], nt = (int64_t)tiles.size() - t0;
        for (int64_t i = 0; i < nt; ++i) {
            tiles[(size_t)(t0 + i)].at = NUM_BINS * t0 + i;
            tiles[(size_t)(t0 + i)].stride = nt;
        }
    }
    tb0[(size_t)nb] = (int64_t)tiles.size();
}

void build_phase_lists(const int64_t* off, int nb, const std::vector<int64_t>& cnt,
                       const std::vector<std::vector<BinRow>>& big, const std::vector<int64_t>& bverts,
                       const std::vector<int64_t>& barcs, const nido_tuning& t, PhaseBins& pb)
{
    std::vector<BatchWork> W((size_t)nb);
#pragma omp parallel for schedule(dynamic, 1) if (nb >= 8)
    for (int b = 0; b < nb; ++b) {
        BatchWork& w = W[(size_t)b];
        for (int k = 0; k < NUM_BINS; ++k) w.cnt[k] = cnt[(size_t)b * NUM_BINS + k];
        w.toff.push_back(0);
        int64_t la = 0;
        for (const BinRow& r : big[(size_t)b]) {
            const int64_t d = off[r.v + 1] - off[r.v];
            add_long_row(w, r.v, d, off, t);
            la += d;
        }
        const int64_t lv = (int64_t)big[(size_t)b].size();
        // the same integer-valued sums as the host builder's, exact in FP64
        w.lbytes = 24.0 * (double)la + 40.0 * (double)lv;
        w.sbytes = 24.0 * (double)(barcs[(size_t)b] - la) + 40.0 * (double)(bverts[(size_t)b] - lv);
        order_items(w, off);
    }
    merge_batches(W, nb, false, pb);
}

void build_row_lists(const int64_t* off, int64_t nv, std::vector<int>& small, std::vector<RowItem>& items)
{
    const int T = nv >= kParallelRows ? std::max(1, omp_get_max_threads()) : 1;
    std::vector<std::vector<int>> ls((size_t)T);
    std::vector<std::vector<RowItem>> li((size_t)T);
#pragma omp parallel num_threads(T)
    {
        const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
        const int64_t v0 = nv * tid / nt, v1 = nv * (tid + 1) / nt;
        auto& s = ls[(size_t)tid];
        auto& it = li[(size_t)tid];
        for (int64_t v = v0; v < v1; ++v) {
            const int64_t d = off[v + 1] - off[v];
            if (d <= kSmallRow) {
                s.push_back((int)v);
            } else {
                for (int64_t e = off[v]; e < off[v + 1]; e += kItemArcs)
                    it.push_back(RowItem{v, e, std::min(off[v + 1], e + kItemArcs)});
            }
        }
    }
    size_t ns = 0, ni = 0;
    for (int i = 0; i < T; ++i) {
        ns += ls[(size_t)i].size();
        ni += li[(size_t)i].size();
    }
    small.clear();
    items.clear();
    small.reserve(ns);
    items.reserve(ni);
    for (int i = 0; i < T; ++i) {
        small.insert(small.end(), ls[(size_t)i].begin(), ls[(size_t)i].end());
        items.insert(items.end(), li[(size_t)i].begin(), li[(size_t)i].end());
    }
}

}  // namespace nido
