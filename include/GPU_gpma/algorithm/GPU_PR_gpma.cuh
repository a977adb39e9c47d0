// Drop-in for GraphHET's GPU_PR_gpma.cuh: PageRank over the current edges.
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_PR_optimized (GPU_PR_gpma.cuh:241-245).  gpma_out's store holds both
// directions of the edge set, so it alone drives the pull over in-rows;
// gpma_in is accepted for the signature.
inline std::vector<std::pair<std::string, double>> Cuda_PR_optimized(graph_structure<double>& graph, GPMA& gpma_in,
                                                                     GPMA& gpma_out, int iterations, double damping) {
    (void)gpma_in;
    ghet_gpma::Snapshot s(gpma_out, GHET_GRAPH_NO_WEIGHTS | GHET_GRAPH_NO_ALL);
    std::vector<double> rank(static_cast<std::size_t>(gpma_out.row_num));
    if (!s.g || !ghet_shim::ok(ghet_pr(s.g, iterations, damping, rank.data(), nullptr, nullptr), "Cuda_PR_optimized"))
        return {};
    return graph.res_trans_id_val(rank);
}
