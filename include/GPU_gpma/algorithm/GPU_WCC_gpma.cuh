// Drop-in for GraphHET's GPU_WCC_gpma.cuh: weak components of the current
// edges of a GPMA store (either direction's store gives the same components).
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_WCC_optimized (GPU_WCC_gpma.cuh:89-93): each vertex named with its
// component's smallest-id vertex.
inline std::vector<std::pair<std::string, std::string>> Cuda_WCC_optimized(graph_structure<double>& graph,
                                                                           GPMA& gpma) {
    ghet_gpma::Snapshot s(gpma, GHET_GRAPH_NO_WEIGHTS | GHET_GRAPH_NO_ALL);
    std::vector<int> labels(static_cast<std::size_t>(gpma.row_num));
    if (!s.g || !ghet_shim::ok(ghet_wcc(s.g, labels.data(), nullptr, nullptr), "Cuda_WCC_optimized")) return {};
    return graph.res_trans_id_id(labels);
}
