// main.cpp -- nido_louvain, the command line of the reference (main.cpp:15-177)
// on one MI355X: -f <bin-file> | -r "<nv> <m0>", -c, -b <batches>,
// -t <threshold>, -i <max iterations>, -o <output file>; same defaults
// (types.hpp:24-38) and output lines.  -p (the multi-GPU partition knobs) is
// rejected with a message: this driver runs one GPU with the graph in HBM.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/nido_amd.hpp"

int main(int argc, char** argv)
{
    std::string input, output, rand_args;
    int nbatches = NIDO_DEFAULT_BATCHES;
    int64_t max_loops = NIDO_DEFAULT_ITERATIONS;
    double tau = NIDO_DEFAULT_THRESHOLD;
    bool have_output = false, rand_graph = false, colored = false;
    int opt;
    while ((opt = getopt(argc, argv, "f:p:r:cb:i:t:ho:")) != -1) {
        switch (opt) {
        case 'f': input = optarg; break;
        case 'r': rand_graph = true; rand_args = optarg; break;
        case 'c': colored = true; break;
        case 'o': output = optarg; have_output = true; break;
        case 't': tau = atof(optarg); break;
        case 'i': max_loops = atoll(optarg); break;
        case 'b': nbatches = atoi(optarg); break;
        case 'p':
            std::cerr << "-p is not supported: this is a single-GPU Louvain driver with the graph resident in HBM"
                      << std::endl;
            return 2;
        case 'h':
            std::cout << "usage: nido_louvain (-f <graph.bin> | -r \"<#vertices> <edge-factor>\") [-c] [-b batches] "
                         "[-t threshold] [-i max-iterations] [-o communities.txt]" << std::endl;
            return 0;
        default:
            std::cout << "Reached default -- please check the passed options." << std::endl;
            break;
        }
    }
    try {
        nido::Graph* graph = nullptr;
        if (rand_graph) {   // main.cpp:84-104: "<#vertices> <edge-factor>"
            std::stringstream ss(rand_args);
            std::vector<std::string> args;
            std::string s;
            while (std::getline(ss, s, ' ')) args.push_back(s);
            if (args.size() != 2) {
                std::cerr << "For random graphs, expecting (in this order): <#vertices> <edge-factor>" << std::endl;
                return 1;
            }
            graph = new nido::Graph((nido::GraphElem)std::stoll(args[0]), (nido::GraphElem)std::stoll(args[1]));
        } else if (!input.empty()) {
            graph = new nido::Graph(input);
        } else {
            std::cerr << "Input graph arguments missing. Use -f or -r options." << std::endl;
            return 1;
        }
        std::cout << "#Batches: " << nbatches << "\n#Threshold: " << tau << "\nMax #iterations: " << max_loops
                  << "\n#NGPUs: 1\n--------------------------------------" << std::endl;
        std::vector<nido::GraphElem> new_orders;
        if (colored) {   // main.cpp:149-155
            const auto t0 = std::chrono::steady_clock::now();
            new_orders = graph->coloring();
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cout << "Coloring Time: " << dt << " s\n";
        }
        {
            nido::GraphGPU graph_gpu(graph, nbatches);
            nido::LouvainGPU louvain(max_loops, tau, nbatches);
            louvain.run(&graph_gpu, true);
            if (have_output) graph_gpu.dump_partition(output, colored ? new_orders.data() : nullptr);
        }
        delete graph;
    } catch (const nido::Error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
