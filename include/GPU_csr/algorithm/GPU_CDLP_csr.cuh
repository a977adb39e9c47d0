// Drop-in for GraphHET's GPU_CDLP_csr.cuh: synchronous label propagation over
// in- and out-neighbours, ties to the smallest label; isolated vertices keep
// their label (the CPU/LDBC semantics).
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// CDLP_GPU (GPU_CDLP_csr.cuh:75): res[i] = name of vertex i's final label.
inline void CDLP_GPU(graph_structure<double>& graph, CSR_graph<double>& input_graph, std::vector<std::string>& res,
                     int max_iterations) {
    const int n = graph.size();
    std::vector<int> labels(static_cast<std::size_t>(n));
    res.clear();
    if (!input_graph.handle ||
        !ghet_shim::ok(ghet_cdlp(input_graph.handle, max_iterations, labels.data(), nullptr, nullptr), "CDLP_GPU"))
        return;
    res.resize(static_cast<std::size_t>(n));
    for (int i = 0; i < n; i++) res[i] = graph.vertex_id_to_str[labels[i]].first;
}

// Cuda_CDLP (GPU_CDLP_csr.cuh:25,200-210): one pair per vertex id.
inline std::vector<std::pair<std::string, std::string>> Cuda_CDLP(graph_structure<double>& graph,
                                                                  CSR_graph<double>& input_graph,
                                                                  int max_iterations) {
    std::vector<std::string> result;
    CDLP_GPU(graph, input_graph, result, max_iterations);
    std::vector<std::pair<std::string, std::string>> res;
    for (std::size_t i = 0; i < result.size(); i++) res.emplace_back(graph.vertex_id_to_str[i].first, result[i]);
    return res;
}
