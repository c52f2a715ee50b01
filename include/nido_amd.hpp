// nido_amd.hpp -- C++ host API over the C ABI (nido_amd.h), with the
// reference's class and method names so that its driver code ports verbatim:
//   Graph       graph.hpp:14-102       (binary CSR reader, getters)
//   GraphGPU    graph_gpu.hpp:168-247  (device state, per-batch methods)
//   LouvainGPU  louvain_gpu.hpp:15-23  (the phase / iteration / batch driver)
// Errors raise nido::Error instead of exit() / silent failure.
#ifndef NIDO_AMD_HPP_
#define NIDO_AMD_HPP_

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "nido_amd.h"

namespace nido {

using GraphElem = int64_t;     // types.hpp:13-22
using GraphWeight = double;

class Error : public std::runtime_error {
  public:
    Error(const std::string& where, int code)
        : std::runtime_error(where + " failed (" + std::to_string(code) + "): " + nido_last_error()),
          code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc, const char* where)
{
    if (rc != NIDO_OK) throw Error(where, rc);
}

// Host CSR graph (graph.hpp:14-102).
class Graph {
  public:
    explicit Graph(const std::string& binfile) { check(nido_hgraph_read(binfile.c_str(), &h_), "Graph(binfile)"); }
    Graph(GraphElem nv, const GraphElem* offsets, const GraphElem* cols, const GraphWeight* weights)
    {
        check(nido_hgraph_from_csr(nv, offsets[nv] - offsets[0], offsets, cols, weights, &h_), "Graph(csr)");
    }
    // Graph(nv, m0): the -r Barabasi-Albert generator (graph.cpp:76-110, 127-148)
    Graph(GraphElem totalVertices, GraphElem m0) { check(nido_gen_ba(totalVertices, m0, &h_), "Graph(nv, m0)"); }
    explicit Graph(nido_hgraph* h) : h_(h) {}
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    ~Graph() { nido_hgraph_destroy(h_); }

    GraphElem get_num_vertices() const { int64_t n = 0; check(nido_hgraph_info(h_, &n, nullptr), "get_num_vertices"); return n; }
    GraphElem get_num_edges() const { int64_t e = 0; check(nido_hgraph_info(h_, nullptr, &e), "get_num_edges"); return e; }
    const GraphElem* get_index_ranges() const { const int64_t* p; check(nido_hgraph_arrays(h_, &p, nullptr, nullptr), "get_index_ranges"); return p; }
    const GraphElem* get_edges() const { const int64_t* p; check(nido_hgraph_arrays(h_, nullptr, &p, nullptr), "get_edges"); return p; }
    const GraphWeight* get_edge_weights() const { const double* p; check(nido_hgraph_arrays(h_, nullptr, nullptr, &p), "get_edge_weights"); return p; }
    void write(const std::string& binfile) const { check(nido_hgraph_write(h_, binfile.c_str()), "Graph::write"); }
    // Graph::coloring (graph.cpp:547-694): renumbers the graph in place and
    // returns new_orders (old id of every new vertex).
    std::vector<GraphElem> coloring(int engines = 1)
    {
        std::vector<GraphElem> no((size_t)get_num_vertices());
        check(nido_hgraph_coloring(h_, engines, nullptr, no.data(), &num_colors_), "coloring");
        return no;
    }
    GraphElem get_num_colors() const { return num_colors_; }
    nido_hgraph* handle() const { return h_; }

  private:
    nido_hgraph* h_ = nullptr;
    int64_t num_colors_ = 0;
};

// Device-resident state (graph_gpu.hpp:168-247).
class GraphGPU {
  public:
    GraphGPU(const Graph* graph, const int& nbatches, void* stream = nullptr) : nbatches_(nbatches)
    {
        check(nido_graph_create(graph->get_num_vertices(), graph->get_num_edges(), graph->get_index_ranges(),
                                graph->get_edges(), graph->get_edge_weights(), nbatches, stream, &g_),
              "GraphGPU()");
    }
    GraphGPU(const GraphGPU&) = delete;
    GraphGPU& operator=(const GraphGPU&) = delete;
    ~GraphGPU() { nido_graph_destroy(g_); }

    void singleton_partition() { check(nido_singleton_partition(g_), "singleton_partition"); }
    GraphWeight compute_modularity() { double q = 0; check(nido_compute_modularity(g_, &q), "compute_modularity"); return q; }
    void louvain_update(const int& batch) { check(nido_louvain_update(g_, batch), "louvain_update"); }
    void update_community_weights(const int& batch) { check(nido_update_community_weights(g_, batch), "update_community_weights"); }
    void update_community_ids(const int& batch) { check(nido_update_community_ids(g_, batch), "update_community_ids"); }
    void restore_community() { check(nido_restore_community(g_), "restore_community"); }
    bool aggregation() { int done = 0; check(nido_aggregation(g_, &done), "aggregation"); return done != 0; }
    void dump_partition(const std::string& filename, const GraphElem* new_orders = nullptr)
    {
        check(nido_dump_partition_orders(g_, filename.c_str(), new_orders), "dump_partition");
    }

    GraphElem get_num_vertices() const { int64_t n = 0; check(nido_graph_info(g_, &n, nullptr, nullptr, nullptr), "info"); return n; }
    GraphElem get_num_edges() const { int64_t e = 0; check(nido_graph_info(g_, nullptr, &e, nullptr, nullptr), "info"); return e; }
    GraphWeight get_mass() const { double m = 0; check(nido_graph_info(g_, nullptr, nullptr, nullptr, &m), "info"); return m; }
    std::vector<GraphElem> get_communities()
    {
        std::vector<GraphElem> c((size_t)get_num_vertices());
        check(nido_get_communities(g_, c.data()), "get_communities");
        return c;
    }
    std::vector<GraphElem> get_clustering(GraphElem n0)
    {
        std::vector<GraphElem> c((size_t)n0);
        check(nido_get_clustering(g_, c.data()), "get_clustering");
        return c;
    }
    int nbatches() const { return nbatches_; }
    nido_graph* handle() const { return g_; }

  private:
    nido_graph* g_ = nullptr;
    int nbatches_;
};

// The driver (louvain_gpu.hpp:15-23); run() executes LouvainGPU::run
// (louvain_gpu.cpp:6-139) inside the library and prints the reference's
// PRINT lines.
class LouvainGPU {
  public:
    LouvainGPU(const GraphElem& maxLoops, const GraphWeight& tol, const int& nbatches)
        : maxLoops_(maxLoops), tol_(tol), tol_phase_(tol), nbatches_(nbatches) {}
    LouvainGPU(const GraphElem& maxLoops, const GraphWeight& tol, const GraphWeight& tol_phase, const int& nbatches)
        : maxLoops_(maxLoops), tol_(tol), tol_phase_(tol_phase), nbatches_(nbatches) {}

    nido_run_stats run(GraphGPU* graph, bool verbose = true, bool record = false)
    {
        nido_run_stats st{};
        check(nido_louvain_run(graph->handle(), maxLoops_, tol_, tol_phase_, record ? 1 : 0, verbose ? 1 : 0, &st),
              "LouvainGPU::run");
        return st;
    }

  private:
    GraphElem maxLoops_;
    GraphWeight tol_, tol_phase_;
    int nbatches_;
};

}  // namespace nido

#endif  // NIDO_AMD_HPP_
