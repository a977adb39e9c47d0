// Drop-in for GraphHET's GPU_WCC_csr.cuh: weakly connected components; the
// label of a vertex is the smallest vertex id of its component.
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// WCC_GPU (GPU_WCC_csr.cuh:83).
inline std::vector<int> WCC_GPU(graph_structure<double>& graph, CSR_graph<double>& input_graph) {
    std::vector<int> result(static_cast<std::size_t>(graph.size()));
    if (!input_graph.handle || !ghet_shim::ok(ghet_wcc(input_graph.handle, result.data(), nullptr, nullptr), "WCC_GPU"))
        return {};
    return result;
}

// Cuda_WCC (GPU_WCC_csr.cuh:126).
inline std::vector<std::pair<std::string, std::string>> Cuda_WCC(graph_structure<double>& graph,
                                                                 CSR_graph<double>& csr_graph) {
    std::vector<int> labels = WCC_GPU(graph, csr_graph);
    return graph.res_trans_id_id(labels);
}
