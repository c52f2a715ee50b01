// internal.h -- declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fixed.h"

namespace nido {

// Degree bins of the local-moving kernels.  Bin b holds the vertices whose
// row length d satisfies kBinMaxDeg[b-1] < d <= kBinMaxDeg[b] (defaults;
// the tuning struct can move the thread-bin bound).
enum LmBin : int {
    BIN_THREAD = 0,   // one lane per vertex, neighbours in registers
    BIN_G8,           // 8 lanes per vertex, 32-slot LDS table
    BIN_G16,          // 16 lanes, 64 slots
    BIN_G32,          // 32 lanes, 128 slots
    BIN_G64,          // one wavefront, 256 slots
    BIN_B256,         // one 256-lane workgroup, 1024 slots
    BIN_BIG,          // hash-partitioned rows, one workgroup per (row, partition)
    NUM_BINS
};
constexpr int NUM_SMALL_BINS = BIN_BIG;

// Registers hold up to this many arcs.  Rows longer than two arcs run
// faster with 8 lanes each (one arc per lane, LDS table) than with one lane
// holding the whole row: R-MAT s22 end to end, +5% over a bound of 8.
constexpr int kThreadMaxDeg = 2;
constexpr int64_t kBinMaxDeg[NUM_BINS] = {kThreadMaxDeg, 16, 32, 64, 128, 512, INT64_MAX};
constexpr int64_t kBigPartArcs = 1024;    // expected distinct keys per partition, at most
constexpr int kBigSlots = 2048;           // LDS table of a partition workgroup (40/48 KB: 4 or 3 per CU)
constexpr int kBigLimit = kBigSlots / 2;  // distinct keys before the exact HBM fallback

// Arcs [e0, e1) of a long row v split into P > 1 partitions.  The key fill
// writes the community keys and weights of these arcs grouped by partition
// into bkey/bw[k0, k0 + (e1 - e0)) (self arcs dropped) and the P + 1
// segment starts, relative to k0, into cpo[c0, c0 + P + 1).
struct KeyChunk {
    int v, P;
    int64_t e0, e1, k0, c0;
};
constexpr int64_t kKeyChunkArcs = 2048;
// Rows split into at most this many partitions skip the key fill: each
// partition workgroup reads the whole row and keeps the arcs of its
// partition (their items carry c0 = -1).  Measured on R-MAT s22: 2 is
// +0.5% over 1, 4 is -4%.
constexpr int kDirectParts = 2;
constexpr int kMaxParts = 1024;           // partitions per row, at most

// A slice of one row for the degree / modularity sweeps over long rows.
struct RowItem {
    int64_t v;
    int64_t e0, e1;
};

// Incremental modularity.  The commit of batch b lists the rows of its
// moved vertices (in chunks of kDeltaArcs arcs) and the communities whose
// total changed; the delta kernel then adds the exact change of the internal
// weight and of the sum of square terms to acc, so Q after the last batch
// equals a full sweep bit for bit.  A handle of several graphs (gid set)
// keeps one accumulator pair per graph, acc[2g] and acc[2g + 1].
constexpr int64_t kDeltaArcs = 1024;
constexpr int kMovedSlots = 256;           // counters of moved arcs (== kBlock: one thread each)
struct ModDelta {
    int* counts;             // per batch: [2b] row chunks, [2b+1] communities
    RowItem* chunks;         // row chunks of moved vertices (current batch)
    int* comms;              // communities touched (current batch)
    int* touched;            // per community: listed in the current batch
    fx_t* term;              // per community: fx((a_c / 2m)^2), 2m of its graph
    fx_t* acc;               // [0] internal weight, [1] sum of square terms (per graph with gid)
    const int* gid;          // several graphs: graph of every vertex; null for one graph
    const int64_t* gp;       // several graphs: batch b of graph g is [gp[g nb1 + b], gp[g nb1 + b + 1])
    const double* twomg;     // several graphs: 2m per graph
    int nb1;                 // batches + 1
    unsigned long long* moved;   // several graphs: kMovedSlots counters of the arcs of moved vertices
};
constexpr int kDeltaBlocks = 256;          // workgroups of a delta riding in a short-row launch

// A vertex of a degree bin with its row: arcs [e0, e0 + d) (saves the
// offset load on the local-moving kernels' dependent-load chain).
struct BinRow {
    int v, d;
    int64_t e0;
};

// Work lists built on the device (bins.hip) for phases of many vertices.
// A tile is up to kBinTileVerts consecutive vertices of one batch (of one
// graph); the tiles of batch b are [tb0[b], tb0[b + 1]), in vertex order,
// and tile t's count of bin k sits at cnt[at + k * stride] (at =
// NUM_BINS * tb0[b] + t - tb0[b], stride = the batch's tile count), so one
// exclusive scan of cnt orders the rows by (batch, bin, id).
constexpr int64_t kBinTileVerts = 2048;
constexpr int64_t kSmallRowArcs = 32;      // rows handled by one lane in the degree/modularity sweeps
constexpr int64_t kItemRowArcs = 1024;     // longer rows are cut into slices of this many arcs
struct BinTile {
    int64_t v0, v1, at, stride;
};
void launch_bin_count(const BinTile* tiles, int64_t ntiles, const int64_t* off, int thread_max, int64_t big_min,
                      int64_t* cnt, hipStream_t s);
void launch_bin_scatter(const BinTile* tiles, int64_t ntiles, const int64_t* off, int thread_max, int64_t big_min,
                        const int64_t* base, BinRow* binv, hipStream_t s);
// out[b * NUM_BINS + k]: the first position of group (b, k); -1 for a batch
// without tiles.
void launch_bin_starts(const int64_t* base, const int64_t* tb0, int nb, int64_t* out, hipStream_t s);
// Row lists over tiles of kBinTileVerts vertices: cs / ci receive each
// tile's rows of at most kSmallRowArcs arcs and slices of the longer rows;
// after their exclusive scans, the scatter writes small[] and items[] in
// ascending vertex order.
void launch_rowlist_count(const int64_t* off, int64_t nv, int64_t* cs, int64_t* ci, hipStream_t s);
void launch_rowlist_scatter(const int64_t* off, int64_t nv, const int64_t* cs, const int64_t* ci, int* small,
                            RowItem* items, hipStream_t s);

// The first launch of a batch: the small bins plus the key fill of the
// batch's partitioned long rows.
struct SmallBins {
    const BinRow* verts[NUM_SMALL_BINS];
    int64_t n[NUM_SMALL_BINS];
    const KeyChunk* chunks;
    int64_t nchunks;
    int* bkey;
    double* bw;
    int* cpo;
    // The incremental-modularity delta of the previous batch (db, vertices
    // [dv0, dv1)) rides in this launch: dblocks workgroups, 0 if none.
    int dblocks, db;
    int64_t dv0, dv1;
    ModDelta D;
    const int64_t* Cn64;                   // C' as int64 (previous communities after a commit)
    double twom_d;
    int blk0[NUM_SMALL_BINS + 3];          // first block of each bin, of the key fill, of the delta; end
};

// One workgroup's share of a long row: partition p of P (P a power of 2);
// `slot` indexes the row within its batch's list of long rows.  For P > 1
// the row's chunk j (of kKeyChunkArcs arcs) has its buckets at
// bkey/bw[k0 + j * kKeyChunkArcs ...] and its segment starts at
// cpo[c0 + j * (P + 1) ...]; a row with P == 1 reads its CSR row directly.
struct BigItem {
    int v, p, P, slot;
    int64_t k0, c0;
    int64_t e0, d;           // the row's first arc and length
};

struct BigArgs {
    const BigItem* items;
    int64_t nitems;
    const int* bkey;         // bucketed community keys of the batch's partitioned rows
    const double* bw;        // their weights
    const int* cpo;          // per chunk: P + 1 segment starts
    int key_limit;           // distinct keys a partition may hold in LDS (<= slots/2)
    int narrow;              // every row's weights are >= 0 and sum below 2^31
    int slots;               // LDS table size of a partition workgroup: 1024, 2048 or 4096
    const int64_t* poff;     // per slot: first partial
    double* pg;              // partial best gain
    int* pc;                 // partial best community
    fx_t* pself;             // partial e_{v,own}
    int* cnt;                // per slot: partitions finished (reset by the last one)
    const int64_t* toff;     // per slot: HBM fallback table [toff[s], toff[s+1])
    int* tkey;
    fx_t* tval;
};

// Everything one local-moving launch reads or writes.  Vertex and community
// ids are below 2^31 on the device, so the hot loops read 32-bit mirrors of
// the int64 column array (col) and community array (C).
struct LmArgs {
    const int64_t* off;
    const int* col;          // 32-bit mirror of the CSR columns
    const double* w;
    const int* C;            // 32-bit mirror of the communities (snapshot)
    int64_t* Cn;             // decided community of the batch's vertices
    const double* kd;        // k_v rounded to FP64
    const fx_t* k;           // k_v exact
    const fx_t* a;           // a_c exact
    double m;                // mass, FP64 (one graph)
    double ym;               // RN(1 / m)
    // Several graphs in one launch (gid null: one graph of mass m): vertex
    // v belongs to graph gid[v], of mass mg[2 gid[v]] with RN(1 / mass) at
    // mg[2 gid[v] + 1], and takes part only while act[gid[v]] is set.
    const int* gid;
    const double* mg;
    const unsigned char* act;
    // Short rows in 64-bit fixed point (wq null: Q64.64): every weight is
    // wq[e] * 2^-F exactly and every short row's sums stay below 2^63 in
    // those units, so the sums are the same exact values with one 64-bit
    // word instead of two; wscale = 2^-F.
    const int64_t* wq;
    double wscale;
};

// ---- launch helpers (kernels.hip) ----------------------------------------
void launch_fill_i64(int64_t* p, int64_t val, int64_t n, hipStream_t s);
void launch_iota_i64(int64_t* p, int64_t n, hipStream_t s);   // p[i] = i
void launch_fill_i32(int32_t* p, int32_t val, int64_t n, hipStream_t s);
void launch_fill_fx(fx_t* p, int64_t n, hipStream_t s);

// Blocks of the reduction kernels, each leaving one partial sum.
constexpr unsigned kPartBlocks = 1024;

// k_v and 2m, and src[e] = the row of arc e, in one pass over the arcs.
// `small` lists rows of at most kSmallRowArcs arcs (kRowLanes lanes each),
// `items` slices of longer rows (one wavefront each).  *acc receives sum(k)
// exactly, *maxabs the bits of max |FP64(k_v)|; part holds kPartBlocks
// partials.
constexpr int kRowLanes = 8;
void launch_vertex_weights(const int64_t* off, const double* w, const int* small, int64_t nsmall,
                           const RowItem* items, int64_t nitems, fx_t* k, double* kd,
                           int64_t nv, fx_t* acc, unsigned long long* maxabs, fx_t* part, int* src,
                           hipStream_t s);
void launch_singleton(int64_t* C, int64_t* Cn, int* C32, fx_t* a, const fx_t* k, int64_t nv,
                      hipStream_t s);
// *fneed = the fractional bits the finest weight needs (max over nonzero
// weights; 2000 for a subnormal or non-finite one, INT_MIN when every weight
// is 0), *wmax = the bits of max |w|.
void launch_weight_grid(const double* w, int64_t ne, int* fneed, unsigned long long* wmax, hipStream_t s);
// wq[e] = w[e] * scale, exact for weights on the 1/scale grid.
void launch_to_fixed(const double* w, int64_t* wq, int64_t ne, double scale, hipStream_t s);
// col32[e] = (int)col[e]
void launch_narrow_ids(const int64_t* col, int* col32, int64_t n, hipStream_t s);
void launch_widen_ids(const int* col32, int64_t* col, int64_t n, hipStream_t s);

// Fills L.blk0 from L.n and L.nchunks; returns the number of blocks of the
// fused launch.
int small_bins_blocks(SmallBins& L, unsigned cap_per_bin);
// bins / fill: launch the short-row bins / the long-row key fill (both: one
// fused launch); split: one launch per bin instead of one fused launch;
// delta: the previous batch's modularity delta rides with the bins.
void launch_lm_small(const LmArgs& A, const SmallBins& L, bool bins, bool fill, bool split, hipStream_t s,
                     bool delta = true);
void launch_lm_big(const LmArgs& A, const BigArgs& B, hipStream_t s);

// With moved set, the commits add the arcs of the vertices that moved to
// the kMovedSlots counters at moved (off: the CSR offsets, for a null list).
void launch_commit(int64_t v0, int64_t v1, int64_t* C, int64_t* Cn, int* C32, const fx_t* k, fx_t* a,
                   hipStream_t s, const int64_t* off = nullptr, unsigned long long* moved = nullptr);
// The commit over list[0, n) (or [v0, v0 + n) with a null list), restricted
// to the vertices v whose graph has flag[gid[v]] set.
void launch_commit_filtered(const BinRow* list, int64_t v0, int64_t n, int64_t* C, int64_t* Cn, int* C32,
                            const fx_t* k, fx_t* a, const int* gid, const unsigned char* flag, hipStream_t s,
                            unsigned long long* moved = nullptr);

// Modularity over the arc-parallel CSR (src = row ids): writes Q to *q;
// part holds 2 * kPartBlocks partials.  With term/acc set it also leaves
// every community's square term in term[] and the exact internal weight and
// sum of squares in acc[0], acc[1] (the state the incremental form updates).
// ident: the caller guarantees C[v] = v for every v (C is then not read).
void launch_modularity(const int* src, const int* col, const double* w, const int* C,
                       int64_t ne, const fx_t* a, int64_t nv, double twom_d, fx_t* part, double* q,
                       hipStream_t s, fx_t* term = nullptr, fx_t* acc = nullptr, bool ident = false);

// Modularity of a phase whose arcs stream through a window: per window,
// the internal weight of rows [v0, v1) added into acc[0] (col / w based so
// that CSR arc indices address the window); then the square terms into
// acc[1] and *q.  acc[0..1] must start at zero.  arcs = off[v1] - off[v0]
// (sizes the grid).
void launch_modularity_rows(const int64_t* off, const int* col, const double* w, const int* C, int64_t v0,
                            int64_t v1, int64_t arcs, fx_t* acc, hipStream_t s);
// term (if not null) receives every community's square term.
void launch_modularity_terms(const fx_t* a, int64_t nv, double twom_d, fx_t* acc, double* q, hipStream_t s,
                             fx_t* term = nullptr);

void launch_commit_tracked(int64_t v0, int64_t v1, int b, int64_t* C, int64_t* Cn, int* C32, const fx_t* k,
                           fx_t* a, const int64_t* off, const ModDelta& D, hipStream_t s);
// The same over the listed vertices list[0, n) of batch b of a handle of
// several graphs, those of graphs whose flag is set (gid from D).
void launch_commit_tracked_list(const BinRow* list, int64_t n, int b, int64_t* C, int64_t* Cn, int* C32,
                                const fx_t* k, fx_t* a, const int64_t* off, const ModDelta& D,
                                const unsigned char* flag, hipStream_t s);
// out[0..3] (zeroed by the caller) += hash sums of the arcs in both
// directions: out[0] == out[1] and out[2] == out[3] when every arc (v, u, w)
// has a reverse arc (u, v, w), the condition the incremental update needs.
void launch_sym_hash(const int* src, const int* col, const double* w, int64_t ne, unsigned long long* out,
                     hipStream_t s);
// Batch b's delta (v0, v1 are its bounds for one graph; unused with D.gid).
void launch_mod_delta(int b, int64_t v0, int64_t v1, const int64_t* Cn, const int* C32, const int* col,
                      const double* w, const fx_t* a, double twom_d, const ModDelta& D, hipStream_t s);
// q = I / 2m - T from acc; zeroes the per-batch counts for the next iteration.
void launch_mod_finish(const ModDelta& D, int nb, double twom_d, double* q, hipStream_t s);
// The same per graph (D.gid set): q[g] for the G graphs whose act flag is
// set, and the sum of the D.moved counters in the bits of q[G].
void launch_mod_finish_graphs(const ModDelta& D, int nb, int G, double* q, const unsigned char* act,
                              hipStream_t s);

// ---- several graphs in one CSR ----------------------------------------------
// Arcs [e0, e1) and vertices [v0, v1) of graph g.
struct SegItem {
    int g;
    int64_t e0, e1, v0, v1;
};
// Per-graph modularity: acc holds 2 accumulators per graph, q[g] receives Q_g;
// with act set, only for the graphs g with act[g] set (the others' q[g] is
// left as it was).
// With term set, also term[c] = fx((a_c / 2m_g)^2) for every community of
// the graphs it sums (the state the incremental update starts from); with
// moved set, the sum of its kMovedSlots counters in the bits of q[G].
void launch_modularity_seg(const SegItem* items, int64_t nitems, const int* src, const int* col,
                           const double* w, const int* C, const fx_t* a, const double* twom_g, int G,
                           fx_t* acc, double* q, hipStream_t s, const unsigned char* act = nullptr,
                           fx_t* term = nullptr, const unsigned long long* moved = nullptr);
// out[g] = newid[vb[g]] for g <= G (the total where vb[g] == nv).
void launch_gather_bases(const int64_t* newid, const int64_t* total, const int64_t* vb, int G, int64_t nv,
                         int64_t* out, hipStream_t s);
// gid[i] = g for i in [vb[g], vb[g+1]).
void launch_fill_gid(const int64_t* vb, int G, int* gid, hipStream_t s);
// out[g] = sum of k over [vb[g], vb[g+1]), exactly (one workgroup per graph).
void launch_graph_sums(const fx_t* k, const int64_t* vb, int G, fx_t* out, hipStream_t s);

// Exclusive scan of n int64 values (out may alias in); writes the total to
// *total (device) when total is non-null.
size_t scan_temp_bytes(int64_t n);
void launch_exclusive_scan(const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                           void* temp, hipStream_t s);

// ---- coarsening ------------------------------------------------------------
// With gid (per current vertex) / gid0 (per original vertex) set, only the
// vertices of graphs g with keep[g] set take part.
void launch_mark_communities(const int64_t* C, int64_t nv, int64_t* flag, hipStream_t s,
                             const int* gid = nullptr, const unsigned char* keep = nullptr);
void launch_update_clustering(int64_t* clus, int64_t n0, const int64_t* C, const int64_t* newid,
                              hipStream_t s, const int* gid0 = nullptr, const unsigned char* keep = nullptr);
// Arcs [e0, e1) belong to a graph whose first vertex is base.
struct ArcPiece {
    int64_t e0, e1;
    int64_t base;
};
// col32[e] = base + local[e] over the pieces: ids that crossed the host link
// as 16-bit graph-local ids.
void launch_widen_local(const ArcPiece* pieces, int64_t n, const uint16_t* local, int* col32, hipStream_t s);
// out[i] = clus[i] - base[gid0[i]] for the n0 original vertices.
void launch_local_ids(const int64_t* clus, int64_t n0, const int* gid0, const int64_t* base, int64_t* out,
                      hipStream_t s);
void launch_member_counts(const int64_t* C, const int64_t* newid, const int64_t* off, int64_t nv,
                          int64_t* cnt, int64_t* arcs, hipStream_t s, const int* gid = nullptr,
                          const unsigned char* keep = nullptr);
// Coarse rows of at most kLdsRowArcs arcs are aggregated in LDS, one
// workgroup each, longer ones in HBM hash tables: tsz[i] table slots (0 for
// an LDS row), csz[i] compact entries (0 for a table row).
constexpr int64_t kLdsRowArcs = 1024;
void launch_coarse_sizes(const int64_t* arcs, int64_t K, int64_t* tsz, int64_t* csz, hipStream_t s);
// members[mstart[i] ...]: the vertices of coarse row i (mfill zeroed, K);
// hlist[0, *nh): the vertices of rows with an HBM table (*nh zeroed).
void launch_member_scatter(const int64_t* C, const int64_t* newid, int64_t nv, const int64_t* mstart,
                           unsigned long long* mfill, int* members, const int64_t* toff, int* hlist,
                           unsigned long long* nh, hipStream_t s, const int* gid = nullptr,
                           const unsigned char* keep = nullptr);
// The LDS rows: distinct targets at ckey/cw[cpos[i] ...), their count in cnt[i].
void launch_coarse_lds(const int64_t* arcs, const int64_t* mstart, const int* members, const int64_t* off,
                       const int64_t* col, const double* w, const int64_t* C, const int64_t* newid, int64_t K,
                       const int64_t* cpos, int* ckey, double* cw, int64_t* cnt, hipStream_t s);
// The HBM-table rows: every listed vertex's arcs (hlist[0, *nh), nmax >= *nh
// sizes the grid).
void launch_coarse_insert(const int64_t* off, const int64_t* col, const double* w, const int64_t* C,
                          const int64_t* newid, const int* hlist, const unsigned long long* nh, int64_t nmax,
                          const int64_t* toff, int* tkey, fx_t* tval, hipStream_t s);
void launch_coarse_count(const int64_t* toff, const int* tkey, int64_t K, int64_t* cnt,
                         hipStream_t s);
void launch_coarse_emit(const int64_t* toff, const int* tkey, const fx_t* tval, int64_t K,
                        const int64_t* noff, int64_t* ncol, double* nw, const int64_t* cpos, const int* ckey,
                        const double* cw, hipStream_t s);

}  // namespace nido
