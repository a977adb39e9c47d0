// Drop-in for GraphHET's GPU_CDLP_gpma.cuh: label propagation over in- and
// out-neighbours of the current edges.
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_CDLP_optimized (GPU_CDLP_gpma.cuh:361-371): one (vertex, label) pair
// per vertex id.  gpma_out's store holds both directions (see GPU_PR_gpma).
inline std::vector<std::pair<std::string, std::string>> Cuda_CDLP_optimized(graph_structure<double>& graph,
                                                                            GPMA& gpma_in, GPMA& gpma_out,
                                                                            int max_iterations) {
    (void)gpma_in;
    ghet_gpma::Snapshot s(gpma_out, GHET_GRAPH_NO_WEIGHTS);
    const std::size_t n = gpma_out.row_num;
    std::vector<int> labels(n);
    std::vector<std::pair<std::string, std::string>> res;
    if (!s.g || !ghet_shim::ok(ghet_cdlp(s.g, max_iterations, labels.data(), nullptr, nullptr), "Cuda_CDLP_optimized"))
        return res;
    for (std::size_t i = 0; i < n; i++)
        res.emplace_back(graph.vertex_id_to_str[i].first, graph.vertex_id_to_str[labels[i]].first);
    return res;
}
