// Drop-in for GraphHET's GPU_PR_csr.cuh: PageRank with the reference's
// update rule; FP32 compensated gather, FP64 dangling mass and rank.
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// GPU_PR (GPU_PR_csr.cuh:34): result resized to V; prints and leaves it empty
// on failure, as the reference does on allocation errors.
inline void GPU_PR(graph_structure<double>& graph, CSR_graph<double>& csr_graph, std::vector<double>& result,
                   int iterations, double damping) {
    result.assign(static_cast<std::size_t>(graph.V), 0.0);
    if (!csr_graph.handle ||
        !ghet_shim::ok(ghet_pr(csr_graph.handle, iterations, damping, result.data(), nullptr, nullptr), "GPU_PR"))
        result.clear();
}

// Cuda_PR (GPU_PR_csr.cuh:29,193-197).
inline std::vector<std::pair<std::string, double>> Cuda_PR(graph_structure<double>& graph,
                                                           CSR_graph<double>& csr_graph, int iterations,
                                                           double damping) {
    std::vector<double> result;
    GPU_PR(graph, csr_graph, result, iterations, damping);
    return graph.res_trans_id_val(result);
}
