/*
 * nido_amd.h -- C ABI of the MI355X-native batched Louvain library
 * (libnido_amd.so).
 *
 * The reference (sg0/nido) is a monolithic executable with no FFI; the seam
 * this ABI replaces is the GraphGPU public method set that LouvainGPU::run
 * calls (graph_gpu.hpp:168-247, louvain_gpu.cpp:28-124) together with the
 * host Graph container (graph.hpp:14-102).  Each entry point below names the
 * reference method it stands in for.
 *
 * Conventions
 *   - Every call returns NIDO_OK (0) or a negative NIDO_ERR_* code; nothing
 *     calls exit().  nido_last_error() gives a message for the calling
 *     thread's last failure.
 *   - Ids are int64, weights FP64 (types.hpp:13-22).  CSR = offsets[nv+1],
 *     cols[ne], weights[ne] (graph.hpp:23-26).
 *   - The library owns every device buffer; callers own the host arrays
 *     they pass in or receive into.  One handle per host thread.
 *   - All sums of input weights are exact (Q64.64 fixed point) and gain
 *     ties go to the smallest community id, so results are bit-identical
 *     for every launch configuration and every run.
 */
#ifndef NIDO_AMD_H_
#define NIDO_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NIDO_OK            0
#define NIDO_ERR_INVALID  -1   /* bad argument or malformed graph */
#define NIDO_ERR_HIP      -2   /* HIP runtime error */
#define NIDO_ERR_NOMEM    -3   /* device or host allocation failed */
#define NIDO_ERR_CAPACITY -4   /* graph exceeds a supported size */
#define NIDO_ERR_STATE    -5   /* call out of order */
#define NIDO_ERR_IO       -6   /* file could not be read or written */

#define NIDO_ABI_VERSION   3   /* 2: nido_serve_config.group; 3: nido_tuning.fixed_short */

#define NIDO_MAX_PHASES  200   /* MAX_PHASES, types.hpp:24-26 */
#define NIDO_DEFAULT_ITERATIONS 500    /* types.hpp:28-30 */
#define NIDO_DEFAULT_BATCHES      2    /* types.hpp:32-34 */
#define NIDO_DEFAULT_THRESHOLD 1.0e-6  /* types.hpp:36-38 */

typedef struct nido_hgraph nido_hgraph;   /* host CSR graph (Graph) */
typedef struct nido_graph nido_graph;     /* device-resident state (GraphGPU) */

/* Per-phase record of a run (the PRINT output of louvain_gpu.cpp:43-118). */
typedef struct {
    int64_t nv;          /* vertices of the phase graph */
    int64_t ne;          /* arcs (CSR entries) of the phase graph */
    int64_t loops;       /* local-moving iterations */
    double q0;           /* modularity of the singleton partition */
    double q;            /* modularity at the end of the phase */
    double t_loop;       /* seconds in the iteration loop ("Time elapse") */
    double t_agg;        /* seconds in aggregation ("Aggregation time") */
} nido_phase_stats;

/* Whole-run summary (louvain_gpu.cpp:125-131). */
typedef struct {
    int32_t phases;
    int32_t done;        /* 1 when the run stopped by convergence */
    int64_t total_loops;
    double q;            /* final modularity */
    double t_total;      /* "Total time elapse" */
    double t_loop;       /* sum of the phases' iteration-loop time */
    double t_agg;        /* sum of aggregation time */
} nido_run_stats;

/* Knobs of the degree-binned local-moving kernels (results do not depend
 * on them; only speed does). */
typedef struct {
    int32_t thread_max_deg;   /* rows up to this length: one lane per vertex (0..2) */
    int32_t big_min_deg;      /* longer rows: hash-partitioned workgroups (0..512) */
    int32_t big_part_arcs;    /* distinct communities expected per partition, at most (>= 1) */
    int32_t big_key_limit;    /* LDS keys per partition before the exact HBM fallback (1..big_slots/2) */
    int32_t use_graph;        /* replay each phase's iterations from a hipGraph: 0 never, 1 for graphs of
                                 2^28+ arcs (default), 2 always; never while per-kernel events are on */
    int32_t big_slots;        /* LDS table of a long-row workgroup: 1024, 2048 or 4096 slots */
    int32_t split_small;      /* launch every short-row bin as a kernel of its own (0/1; profiling) */
    int32_t overlap_long;     /* run the long rows of a batch on a second stream, beside the short rows
                                 (0/1, default 1; only while iterations are replayed from a hipGraph) */
    int32_t full_modularity;  /* how Q is evaluated after an iteration (the same Q either way): 0 (default)
                                 the exact incremental update over the moved rows after each phase's first
                                 iteration, for a handle of several graphs only when the iteration before
                                 moved at most 1/8 of the arcs (else a sweep of every arc); 1 always the
                                 sweep; 2 the update after each phase's first iteration (0/1/2).  With 0
                                 or 2, a windowed phase whose batches each fit the window takes the
                                 update from its first iteration on */
    int32_t fixed_short;      /* 1 (default): rows up to big_min_deg arcs sum in 64-bit fixed point when every
                                 weight is a multiple of one power of two 2^-F and those rows' sums stay below
                                 2^63 in units of 2^-F (the same exact sums); 0: always Q64.64 (0/1) */
} nido_tuning;

const char* nido_last_error(void);
/* The exact Q64.64 <-> FP64 conversions every weight sum goes through
 * (round-to-nearest-even both ways), as compiled for the host: value =
 * hi + lo * 2^-64, two's complement.  No counterpart in the reference, whose
 * sums are FP64; exposed so that the conversions can be checked. */
void nido_fx_from_double(double w, uint64_t* lo, int64_t* hi);
double nido_fx_to_double(uint64_t lo, int64_t hi);
/* out[i] = a[i] / b[i] by the division of the gain formulas (the reciprocal
 * of b and two FMA corrections), as compiled for the host; exact for b in
 * [2^-399, 2^399] and a = 0 or |a| in [2^-400, 2^400].  Exposed so that its
 * rounding can be checked against IEEE division. */
void nido_div_rcp(int64_t n, const double* a, const double* b, double* out);
int nido_abi_version(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int nido_device_count(int* count);
/* Graph creation copies large inputs to the device, and a batched run its
 * clusterings back, through two pinned staging buffers (ids narrowed to
 * int32 on the way up; the buffers are pooled and reused): from min_bytes
 * crossing the host link, chunk_bytes per buffer (a multiple of 4096); 0
 * restores the defaults, 256 MiB and 64 MiB.  Process-wide.  No counterpart
 * in the reference (its uploads are plain pageable copies,
 * graph_gpu.cpp:967-1027); lowering it lets small inputs take the pinned
 * path. */
int nido_set_upload_staging(int64_t min_bytes, int64_t chunk_bytes);
/* Phases of at least min_vertices vertices whose arcs are resident build
 * their degree bins and sweep row lists on the device, smaller ones on the
 * host (the same lists); 0 restores the default, 2^17.  Process-wide.  No
 * counterpart in the reference (its bins are built by Thrust calls each
 * phase, graph_gpu.cpp:94-139); lowering it lets small inputs take the
 * device builder. */
int nido_set_device_lists(int64_t min_vertices);
/* A batched handle (several graphs) rebuilds its local-moving work lists
 * for the graphs still iterating in the phase once their vertices are at
 * most `percent` of the listed ones (0 never, 100 as soon as one graph
 * stops; results do not change); a negative value restores the default,
 * 50.  Process-wide.  No counterpart in the reference (it runs one graph
 * per handle, louvain_gpu.cpp:40-120). */
int nido_set_list_rebuild(int percent);

/* ---- host graph: Graph (graph.hpp:14-102) ------------------------------- */

/* Graph::Graph(binfile), graph.cpp:322-495: header {int64 nv, int64 ne},
 * (nv+1) int64 offsets, then ne records {int64 tail, double weight}
 * (types.hpp:48-55).  Offsets are rebased to start at 0 and the arc count is
 * taken from them (graph.cpp:376-382, 470-472). */
int nido_hgraph_read(const char* path, nido_hgraph** out);
/* Writes the same format (the converter in Vite that the README points at). */
int nido_hgraph_write(const nido_hgraph* g, const char* path);
/* Copies a caller CSR into a host graph: arcs [offsets[0], offsets[nv]) of
 * cols/weights, which hold ne entries from offsets[0] on (offsets spanning
 * more than ne arcs, or decreasing, are refused with NIDO_ERR_INVALID). */
int nido_hgraph_from_csr(int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* cols,
                         const double* weights, nido_hgraph** out);
/* get_num_vertices / get_num_edges (graph.hpp:72-73). */
int nido_hgraph_info(const nido_hgraph* g, int64_t* nv, int64_t* ne);
/* get_index_ranges / get_edges / get_edge_weights (graph.hpp:76-78): pointers
 * stay valid until nido_hgraph_destroy. */
int nido_hgraph_arrays(const nido_hgraph* g, const int64_t** offsets, const int64_t** cols,
                       const double** weights);
void nido_hgraph_destroy(nido_hgraph* g);

/* Seeded synthetic inputs (the workloads' graphs).  All are symmetric,
 * without self-loops or duplicate arcs, rows sorted by neighbour id; one
 * weight per undirected edge copied to both arcs; weighted=0 gives unit
 * weights, weighted=1 weights in (0,1] on a 2^-53 grid.  The reference's own
 * generator, the -r BA graph, is nido_gen_ba below. */
/* Planted partition: vertex v in block v/block_size, pairs u<v taken in
 * lexicographic order, Bernoulli(p_in) inside a block, Bernoulli(p_out)
 * across blocks. */
int nido_gen_sbm(int64_t nv, int64_t block_size, double p_in, double p_out, int weighted,
                 uint64_t seed, nido_hgraph** out);
/* R-MAT: edge_factor * 2^scale samples with quadrant probabilities
 * (a, b, c, 1-a-b-c); self-loops dropped, symmetrised, deduplicated. */
int nido_gen_rmat(int scale, int64_t edge_factor, double a, double b, double c, int weighted,
                  uint64_t seed, nido_hgraph** out);
/* Chung-Lu: expected degrees Pareto(alpha) clipped to [dmin, dmax] and
 * rescaled to the requested mean; nv*mean/2 endpoint pairs drawn in
 * proportion to the expected degrees; self-loops dropped, deduplicated. */
int nido_gen_chung_lu(int64_t nv, double alpha, double dmin, double dmax, double mean,
                      int weighted, uint64_t seed, nido_hgraph** out);
/* Graph::Graph(nv, m0) -> create_random_network_ba, graph.cpp:76-110 and
 * 127-148 (the reference's -r "<nv> <m0>"): a Barabasi-Albert graph drawn
 * with the reference's random-number objects in the reference's order (two
 * default std::default_random_engine, weights |N(1, 1)|, attachment to the
 * first endpoint of a uniformly drawn earlier arc), duplicates kept, arcs
 * ordered by (tail, head); 2*m0*(nv-m0) arcs.  Needs 1 <= m0 < nv < 2^31. */
int nido_gen_ba(int64_t nv, int64_t m0, nido_hgraph** out);
/* Graph::coloring, graph.cpp:547-694 (the reference's -c), in place: random
 * weights r (engine t = std::mt19937_64(1 << t), t < engines, each with a
 * std::uniform_int_distribution<int64_t>(0, 4 nv); vertex i draws from
 * engine i % engines, each engine's draws taken in vertex order), the
 * colouring of the reference's loop as one thread runs it -- colour(i) is the
 * smallest colour held by no neighbour j < i with r_j >= r_i -- computed on
 * the device, then the vertices renumbered in (colour, id) order and the CSR
 * permuted to match (rows keep their arc order).  new_orders[nv] receives the
 * old id of every new vertex, colors[nv] (if not NULL) the colour of every
 * old vertex, *num_colors (if not NULL) the colour count.  With engines > 1
 * the reference's own result depends on thread timing (its threads share
 * the engines and the colour array without synchronisation); engines = 1
 * is its reproducible case.  1 <= engines <= 31. */
int nido_hgraph_coloring(nido_hgraph* g, int engines, int64_t* colors, int64_t* new_orders,
                         int64_t* num_colors);

/* ---- device state: GraphGPU (graph_gpu.hpp:168-247) -------------------- */

/* GraphGPU::GraphGPU(graph, nbatches, ...), graph_gpu.cpp:19-185: uploads the
 * CSR, computes k_v (sum_vertex_weights), 2m (compute_mass), the maximum
 * order and the -b batch boundaries (partition_graph_edge_batch).  Offsets
 * must be non-decreasing, cols in [0, nv), weights finite with |w| < 2^62,
 * their |w| totalling less than 2^62 (each exact sum -- k_v, a_c, 2m, coarse
 * weights -- is then below 2^62), and with any negative weight the total of
 * |w| less than 2^31 |2m| (modularity square terms below 2^62); in a batched
 * handle these hold per graph.
 * stream: a hipStream_t to run on, or NULL for a stream owned by the
 * handle.  nv must be below 2^31. */
int nido_graph_create(int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* cols,
                      const double* weights, int nbatches, void* stream, nido_graph** out);
/* GraphGPU::~GraphGPU, graph_gpu.cpp:187-221. */
/* GraphGPU with out-of-core edge windows (graph_gpu.cpp:341-417, 967-1027):
 * the device holds at most window_arcs arcs of the graph at a time; a phase
 * with more arcs keeps them in pinned host memory, cuts every batch into
 * units of consecutive rows whose arcs fit the window, copies each unit's
 * arcs in before its local-moving launches and sweeps the modularity window
 * by window, and its coarse graph is built on the host; a phase whose arcs
 * fit runs resident.  When every batch fits half the window, a batch's
 * copy runs (on an internal stream, joined back by events) beside the
 * previous batch's launches.  Results are those of nido_graph_create.  A row
 * longer than the window is refused (NIDO_ERR_CAPACITY) when its phase
 * starts.
 * nido_graph_create picks such a window by itself, sized from the free
 * device memory, when a graph would not fit resident. */
int nido_graph_create_windowed(int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* cols,
                               const double* weights, int nbatches, int64_t window_arcs, void* stream,
                               nido_graph** out);
int nido_graph_destroy(nido_graph* g);
/* NV_, NE_, maxOrder_, mass_ of the current phase graph. */
int nido_graph_info(const nido_graph* g, int64_t* nv, int64_t* ne, int64_t* max_order,
                    double* mass);
/* vertex_per_batch_ (graph_gpu.cpp:338-339): nbatches+1 entries. */
int nido_graph_batch_bounds(const nido_graph* g, int64_t* parts);
int nido_graph_set_tuning(nido_graph* g, const nido_tuning* t);
int nido_graph_get_tuning(const nido_graph* g, nido_tuning* t);
/* Autotuner (no counterpart in the reference): times `iterations` local-
 * moving iterations of the current phase, after its first two from the
 * singleton partition, for candidate knob values, one knob at a time, keeps the fastest, applies it
 * and reports it in *chosen (may be NULL).  Results never depend on the
 * knobs.  Leaves the handle in the singleton partition of its phase. */
int nido_autotune(nido_graph* g, int iterations, nido_tuning* chosen);
/* Blocks until every call queued on the handle's stream has finished. */
int nido_graph_sync(nido_graph* g);

/* GraphGPU::singleton_partition, graph_gpu.cpp:519-536. */
int nido_singleton_partition(nido_graph* g);
/* GraphGPU::compute_modularity, graph_gpu.cpp:862-905 (blocks for the value). */
int nido_compute_modularity(nido_graph* g, double* q);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the modularity nido_compute_modularity gives, for the
 * partition comm of a CSR graph, all arrays device pointers (offsets nv+1,
 * cols/weights ne, comm nv with ids in [0, nv), *q one device double).  It
 * allocates nothing: its working memory is a fixed device array of the
 * library (16 B per vertex for NIDO_NOALLOC_MAX_VERTICES vertices, held on
 * every GPU the library runs on until the process ends); more vertices give
 * NIDO_ERR_CAPACITY.  Asynchronous on `stream` (a hipStream_t; NULL = the
 * null stream): *q is written when the stream reaches it, NaN if an id or
 * a row lies outside the graph.  Calls on one GPU share the working memory:
 * calls whose launches overlap (unordered streams) race. */
#define NIDO_NOALLOC_MAX_VERTICES (INT64_C(1) << 22)
int nido_modularity_noalloc(int64_t nv, int64_t ne, const int64_t* offsets, const int64_t* cols,
                            const double* weights, const int64_t* comm, double* q, void* stream);
/* GraphGPU::louvain_update(batch, g), graph_gpu.cpp:602-624: decides the
 * moves of one batch (writes newCommIds_). */
int nido_louvain_update(nido_graph* g, int batch);
/* GraphGPU::update_community_weights(batch, g), graph_gpu.cpp:779-791. */
int nido_update_community_weights(nido_graph* g, int batch);
/* GraphGPU::update_community_ids(batch, g), graph_gpu.cpp:808-820.  The
 * device applies the weight update and the swap in one kernel: after
 * nido_update_community_weights the swap is already done and this call only
 * checks the order of calls. */
int nido_update_community_ids(nido_graph* g, int batch);
/* GraphGPU::restore_community, graph_gpu.cpp:822-842. */
int nido_restore_community(nido_graph* g);
/* GraphGPU::aggregation, graph_gpu.cpp:1382-1545: renumbers, updates the
 * cumulative clustering, and (unless K==N or K==1, then *done=1) rebuilds the
 * coarse graph on the device and recomputes k, m and the batches. */
int nido_aggregation(nido_graph* g, int* done);

/* State readers (D2H copies; the arrays are the current phase's). */
int nido_get_communities(nido_graph* g, int64_t* comm);          /* commIds_ */
int nido_get_new_communities(nido_graph* g, int64_t* comm);      /* newCommIds_ */
int nido_get_vertex_weights(nido_graph* g, double* k);           /* vertexWeights_ */
int nido_get_community_weights(nido_graph* g, double* a);        /* commWeights_ */
/* Current phase CSR; rows of coarse graphs are in hash order (the results
 * do not depend on arc order). */
int nido_get_graph(nido_graph* g, int64_t* offsets, int64_t* cols, double* weights);
/* Clustering::commIds_ (clustering.cpp:12-21): original vertex -> community
 * of the last aggregated phase; n0 = original vertex count. */
int nido_get_clustering(nido_graph* g, int64_t* comm);
/* GraphGPU::dump_partition / Clustering::dump_partition, clustering.cpp:23-46:
 * one "%lld\n" line per original vertex. */
int nido_dump_partition(nido_graph* g, const char* path);
/* The same with the order from nido_hgraph_coloring (main.cpp:162-163,
 * clustering.cpp:25-37): the line of original vertex new_orders[i] holds
 * the community of renumbered vertex i; new_orders NULL = no renumbering. */
int nido_dump_partition_orders(nido_graph* g, const char* path, const int64_t* new_orders);

/* One pass of the LouvainGPU::run loop body (louvain_gpu.cpp:55-90): every
 * batch's louvain_update / update_community_weights / update_community_ids,
 * then compute_modularity; if the new modularity is below *q the iteration is
 * rolled back (restore_community), otherwise *q takes the new value.  *dq
 * receives Qnew - Q and *q_eval (if not NULL) Qnew itself. */
int nido_louvain_iteration(nido_graph* g, double* q, double* dq, double* q_eval);

/* Live per-kernel timing with HIP events on the handle's stream (off by
 * default).  Kernel classes: */
#define NIDO_K_LM_SHORT    0   /* short-row bins + long-row key fill, one launch per batch */
#define NIDO_K_LM_LONG     1   /* hash-partitioned long rows, one launch per batch */
#define NIDO_K_COMMIT      2   /* community-total update + swap, one launch per batch */
#define NIDO_K_MODULARITY  3   /* modularity sweep + finish, one pair per evaluation */
#define NIDO_K_CLASSES     4
int nido_graph_set_profiling(nido_graph* g, int enable);
/* Events only around the classes whose bit (1 << class) is set in mask. */
int nido_graph_set_profiling_classes(nido_graph* g, unsigned mask);
/* Events only around every stride-th launch of each profiled class (1, the
 * default, brackets every launch).  A stride coprime with the batch count
 * samples every batch equally often; each event pair costs a few
 * microseconds of GPU time. */
int nido_graph_set_profiling_stride(nido_graph* g, int stride);
/* Launch count, summed duration (ms) and algorithmic HBM bytes of one class
 * since profiling was enabled (24 B per arc + 40 B per vertex for local
 * moving, 24 B per arc + 24 B per vertex for modularity, 0 for commits). */
int nido_graph_kernel_stats(nido_graph* g, int kernel_class, int64_t* launches, double* ms,
                            double* bytes);

/* ---- driver: LouvainGPU (louvain_gpu.hpp:15-23) ------------------------- */

/* LouvainGPU::run, louvain_gpu.cpp:6-139: phases (<= 200) of iterations
 * (while tol < dQ and loops < max_loops) of -b batches, modularity
 * accept/rollback, aggregation while Q - Q0 > tol_phase.  record=1 keeps the
 * per-phase communities for nido_run_phase_communities.  verbose=1 prints the
 * reference's PRINT lines to stdout. */
int nido_louvain_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, int record,
                     int verbose, nido_run_stats* stats);
int nido_run_phase(const nido_graph* g, int phase, nido_phase_stats* out);
/* Modularity after every iteration of a phase (loops entries). */
int nido_run_iteration_q(const nido_graph* g, int phase, double* q);
/* Communities at the end of a phase, phase-local ids (record=1 runs). */
int nido_run_phase_communities(const nido_graph* g, int phase, int64_t* comm);

/* ---- many independent graphs in one handle ------------------------------
 * No counterpart in the reference, which runs one graph per process; this
 * runs LouvainGPU::run (louvain_gpu.cpp:6-139) on every graph of a batch at
 * once, sharing launches, and every graph's result is the one it gets alone
 * (same phases, iterations, modularity and communities).
 *
 * The graphs come as one CSR: graph i owns vertices [vptr[i], vptr[i+1]) and
 * the arcs of their rows; column ids are handle-wide (graph i's arcs stay in
 * graph i's vertex range).  Each graph is cut into nbatches batches by its
 * own rule (graph_gpu.cpp:303-339). */
int nido_batch_create(int ngraphs, const int64_t* vptr, const int64_t* offsets, const int64_t* cols,
                      const double* weights, int nbatches, void* stream, nido_graph** out);
/* Runs every graph to completion; stats receives the handle-wide totals
 * (phases = longest run, total_loops = shared iterations). */
int nido_batch_run(nido_graph* g, int64_t max_loops, double tol, double tol_phase, nido_run_stats* stats);
int nido_batch_size(const nido_graph* g, int* ngraphs);
/* Work-list rebuilds for the graphs still iterating (nido_set_list_rebuild)
 * in the last nido_batch_run.  No counterpart in the reference. */
int nido_batch_list_rebuilds(const nido_graph* g, int64_t* n);
/* One graph's run: q, phases, total_loops as LouvainGPU::run reports them. */
int nido_batch_graph_stats(const nido_graph* g, int graph, nido_run_stats* out);
int nido_batch_graph_phase(const nido_graph* g, int graph, int phase, nido_phase_stats* out);
/* Clustering::commIds_ of one graph (clustering.cpp:12-21): original vertex
 * (graph-local) -> final community (graph-local, dense from 0). */
int nido_batch_clustering(const nido_graph* g, int graph, int64_t* comm);

/* ---- request loop (continuous batching on one GPU) -----------------------
 * No counterpart in the reference.  Requests (one graph each, copied at
 * submission) are admitted in arrival order as soon as one of `slots`
 * workers is free and the request's estimated HBM footprint fits in the
 * unused part of hbm_budget; each worker runs LouvainGPU::run on its own
 * stream, so admitted requests overlap on the device.  A request's result
 * is the one it gets alone. */
typedef struct nido_server nido_server;
typedef struct {
    int32_t slots;         /* workers, each with its own stream (1..64) */
    int32_t nbatches;      /* -b */
    int64_t hbm_budget;    /* bytes of device memory the admitted requests may use together */
    int64_t max_loops;     /* -i */
    double tol;            /* -t */
    double tol_phase;
    int32_t group;         /* small requests (up to 2^23 arcs) a worker runs together in one batched
                              handle, at most (0 or 1: one request per worker; up to 4096) */
} nido_serve_config;
int nido_serve_create(const nido_serve_config* cfg, nido_server** out);
int64_t nido_serve_footprint(int64_t nv, int64_t ne);   /* admission estimate, bytes */
/* Copies the graph unless flags has NIDO_SERVE_BORROW, in which case the
 * caller keeps the arrays alive and unchanged until the request finishes. */
#define NIDO_SERVE_BORROW 1
int nido_serve_submit(nido_server* s, int64_t nv, const int64_t* offsets, const int64_t* cols,
                      const double* weights, int flags, int64_t* id);
/* Blocks until request `id` has finished; latency_s = finish - submission. */
int nido_serve_wait(nido_server* s, int64_t id, nido_run_stats* stats, double* latency_s);
int nido_serve_clustering(nido_server* s, int64_t id, int64_t* comm);   /* nv entries */
int nido_serve_release(nido_server* s, int64_t id);                      /* forget a finished request */
int nido_serve_status(nido_server* s, int64_t* queued, int64_t* running, int64_t* completed,
                      int64_t* peak_running);
/* Drains the queue, then stops the workers. */
int nido_serve_destroy(nido_server* s);

#ifdef __cplusplus
}
#endif
#endif /* NIDO_AMD_H_ */
