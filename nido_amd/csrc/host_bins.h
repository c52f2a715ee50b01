// host_bins.h -- per-phase work lists of the local-moving launches, built on
// the host (OpenMP): batch boundaries, degree bins, long-row items and the
// row lists of the degree and modularity sweeps.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/nido_amd.h"
#include "internal.h"

namespace nido {

constexpr int64_t kSmallRow = kSmallRowArcs;
constexpr int64_t kItemArcs = kItemRowArcs;
// Host work lists are built with threads only from this many vertices up:
// smaller graphs are the serving loop's, whose requests already run side by
// side, one host thread each.
constexpr int64_t kParallelRows = int64_t(1) << 22;

// nb + 1 batch boundaries: partition_graph_edge_batch with part_on_batch=1
// on one device, graph_gpu.cpp:303-339.
void batch_bounds(const int64_t* off, int64_t nv, int nb, std::vector<int64_t>& parts);

// The same for each of G graphs (graph g owns vertices [vb[g], vb[g+1])),
// boundaries in handle-wide ids: gparts[g * (nb + 1) + i].
void batch_bounds_all(const int64_t* off, int G, const int64_t* vb, int nb, std::vector<int64_t>& gparts);

int64_t max_degree(const int64_t* off, int64_t nv);

// Degree bins and long-row work of every batch.  Batch b of the handle is
// [parts[b], parts[b+1]) for one graph, or the union over graphs g of
// [gparts[g (nb+1) + b], gparts[g (nb+1) + b + 1]) (over the graphs with
// keep[g] set when keep is given).
struct PhaseBins {
    std::vector<BinRow> binv;          // per batch: its rows by bin (ascending id inside a bin), long rows last
    std::vector<BigItem> bitems;       // per batch: partition items, longest row first
    std::vector<KeyChunk> chunks;      // per batch: key-fill chunks of rows with P > kDirectParts
    std::vector<int64_t> poff, toff;   // per batch: partial offsets per row slot; HBM table offsets (slots + 1)
    std::vector<int64_t> item_base, poff_base, toff_base, chunk_base;   // chunk_base has nb + 1 entries
    std::vector<int64_t> small_start, small_n;                          // [b * NUM_SMALL_BINS + k]
    std::vector<int64_t> blist0, blistn;                                // batch slices of binv
    std::vector<double> short_bytes, long_bytes;                        // algorithmic bytes per batch
    int64_t rows = 0;                                                   // entries of binv
    int64_t tab_need = 0, part_need = 0, slot_need = 0, key_need = 0, cpo_need = 0;
};
void build_phase_bins(const int64_t* off, int64_t nv, int nb, int G, const std::vector<int64_t>& parts,
                      const std::vector<int64_t>& gparts, const nido_tuning& t, PhaseBins& pb,
                      const unsigned char* keep = nullptr);

// Tiles of the device bin builder (BinTile): batch by batch, each batch's
// vertex ranges (graph by graph, the graphs with keep[g] set when keep is
// given) cut into tiles of kBinTileVerts; tb0 gets nb + 1 tile bounds.
void bin_tiles(int nb, int G, const std::vector<int64_t>& parts, const std::vector<int64_t>& gparts,
               const unsigned char* keep, std::vector<BinTile>& tiles, std::vector<int64_t>& tb0);
// The lists of a phase whose binv the device built: cnt[b * NUM_BINS + k]
// rows of batch b in bin k, big[b] the long rows of batch b (ascending id),
// bverts / barcs the vertices and arcs of batch b.  Leaves pb.binv empty.
void build_phase_lists(const int64_t* off, int nb, const std::vector<int64_t>& cnt,
                       const std::vector<std::vector<BinRow>>& big, const std::vector<int64_t>& bverts,
                       const std::vector<int64_t>& barcs, const nido_tuning& t, PhaseBins& pb);

// Arcs streamed through a device window of `window` arcs: every batch b
// is cut into units of consecutive rows holding at most `window` arcs
// (units[first_unit[b]] .. units[first_unit[b + 1]] are its bounds, nb + 1
// entries in first_unit).  A row longer than the window is refused.
int window_units(const int64_t* off, const std::vector<int64_t>& parts, int nb, int64_t window,
                 std::vector<int64_t>& units, std::vector<int>& first_unit);
// Whether every arc has a reverse arc of the same weight, on host arrays:
// the multisets {(v, col, w)} and {(col, v, w)} compared through two keyed
// hash sums of each (mod 2^64), as the device check for resident graphs
// does; a wrong "true" has probability about 2^-128.
bool host_arcs_symmetric(int64_t nv, const int64_t* off, const int* col, const double* w);
// Exact vertex weights on the host: k (Q64.64), kd = RN(k), 2m and max |kd|.
void host_vertex_weights(int64_t nv, const int64_t* off, const double* w, fx_t* k, double* kd, fx_t* twom,
                         double* maxk);
// The coarse graph of a phase on the host (graph_gpu.cpp:1088-1304): row
// newid[C[v]] gathers the arcs of v, merged by newid[C[target]] with exact
// sums, rows sorted by target.
void host_coarsen(int64_t nv, const int64_t* off, const int* col, const double* w, const int64_t* C,
                  const int64_t* newid, int64_t K, std::vector<int64_t>& noff, std::vector<int>& ncol,
                  std::vector<double>& nw);

// Rows of at most kSmallRow arcs (one lane each) and kItemArcs-arc slices
// of the longer rows, both in ascending vertex order.
void build_row_lists(const int64_t* off, int64_t nv, std::vector<int>& small, std::vector<RowItem>& items);

}  // namespace nido
