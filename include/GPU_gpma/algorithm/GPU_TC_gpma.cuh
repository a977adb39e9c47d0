// Drop-in for GraphHET's GPU_TC_gpma.cuh: per-vertex triangle counts over the
// current edges, in the reference's convention.
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_TC_PerNode (GPU_TC_gpma.cuh:68-83) on a GPMA store.
inline std::vector<int> Cuda_TC_PerNode(GPMA& gpma_graph, int V) {
    std::vector<int> out(static_cast<std::size_t>(V), 0);
    std::vector<long long> cnt(static_cast<std::size_t>(gpma_graph.row_num), 0);
    ghet_gpma::Snapshot s(gpma_graph, GHET_GRAPH_NO_WEIGHTS | GHET_GRAPH_NO_IN);
    if (!s.g || !ghet_shim::ok(ghet_tc(s.g, reinterpret_cast<int64_t*>(cnt.data()), nullptr, nullptr), "Cuda_TC"))
        return out;
    for (std::size_t i = 0; i < out.size() && i < cnt.size(); ++i)
        out[i] = static_cast<int>(static_cast<unsigned int>(cnt[i]));
    return out;
}

// Cuda_TC (GPU_TC_gpma.cuh:86-90).
inline std::vector<std::pair<std::string, int>> Cuda_TC(graph_structure<double>& graph, GPMA& gpma_graph) {
    std::vector<int> counts = Cuda_TC_PerNode(gpma_graph, graph.size());
    return graph.res_trans_id_val(counts);
}
