// main.cpp -- nido_louvain, the command line of the reference (main.cpp:15-177)
// on one MI355X: -f <bin-file> -b <batches> -t <threshold> -i <max iterations>
// -o <output file>, same defaults (types.hpp:24-38) and output lines.
// Options of the reference outside this build: -p (multi-GPU partition
// knobs), -r (BA random graph), -c (Luby colouring) -- rejected with a message.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <unistd.h>

#include "../../include/nido_amd.hpp"

int main(int argc, char** argv)
{
    std::string input, output;
    int nbatches = NIDO_DEFAULT_BATCHES;
    int64_t max_loops = NIDO_DEFAULT_ITERATIONS;
    double tau = NIDO_DEFAULT_THRESHOLD;
    bool have_output = false;
    int opt;
    while ((opt = getopt(argc, argv, "f:p:r:cb:i:t:ho:")) != -1) {
        switch (opt) {
        case 'f': input = optarg; break;
        case 'o': output = optarg; have_output = true; break;
        case 't': tau = atof(optarg); break;
        case 'i': max_loops = atoll(optarg); break;
        case 'b': nbatches = atoi(optarg); break;
        case 'p':
        case 'r':
        case 'c':
            std::cerr << "-" << (char)opt << " is not supported: this is a single-GPU Louvain driver with the graph resident in HBM"
                      << std::endl;
            return 2;
        case 'h':
            std::cout << "usage: nido_louvain -f <graph.bin> [-b batches] [-t threshold] [-i max-iterations] "
                         "[-o communities.txt]" << std::endl;
            return 0;
        default:
            std::cout << "Reached default -- please check the passed options." << std::endl;
            break;
        }
    }
    if (input.empty()) {
        std::cerr << "Input graph arguments missing. Use -f." << std::endl;
        return 1;
    }
    try {
        nido::Graph graph(input);
        std::cout << "#Batches: " << nbatches << "\n#Threshold: " << tau << "\nMax #iterations: " << max_loops
                  << "\n#NGPUs: 1\n--------------------------------------" << std::endl;
        nido::GraphGPU graph_gpu(&graph, nbatches);
        nido::LouvainGPU louvain(max_loops, tau, nbatches);
        louvain.run(&graph_gpu, true);
        if (have_output) graph_gpu.dump_partition(output);
    } catch (const nido::Error& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
