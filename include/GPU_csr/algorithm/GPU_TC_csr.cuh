// Drop-in for GraphHET's GPU_TC_csr.cuh: per-vertex triangle counts in the
// reference's convention (6 per triangle on a symmetric loop-free graph).
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// Raw core (GPU_TC_csr.cuh:46): device out-CSR.  The library counts in int64;
// the int result wraps like the reference's int32 atomics would.
inline std::vector<int> Cuda_TC_PerNode(int* d_row_offsets, int* d_col_indices, int node_size) {
    std::vector<int> out(static_cast<std::size_t>(node_size), 0);
    std::vector<long long> cnt(static_cast<std::size_t>(node_size), 0);
    ghet_shim::Borrowed b(node_size, -1, d_row_offsets, d_col_indices, nullptr);
    if (!b.g || !ghet_shim::ok(ghet_tc(b.g, reinterpret_cast<int64_t*>(cnt.data()), nullptr, nullptr), "Cuda_TC"))
        return out;
    for (std::size_t i = 0; i < cnt.size(); ++i) out[i] = static_cast<int>(static_cast<unsigned int>(cnt[i]));
    return out;
}

// Cuda_TC (GPU_TC_csr.cuh:65).
inline std::vector<std::pair<std::string, int>> Cuda_TC(graph_structure<double>& graph, CSR_graph<double>& csr_graph) {
    std::vector<int> counts(static_cast<std::size_t>(graph.V), 0);
    std::vector<long long> cnt(static_cast<std::size_t>(graph.V), 0);
    if (!csr_graph.handle ||
        !ghet_shim::ok(ghet_tc(csr_graph.handle, reinterpret_cast<int64_t*>(cnt.data()), nullptr, nullptr), "Cuda_TC"))
        return graph.res_trans_id_val(counts);
    for (std::size_t i = 0; i < cnt.size(); ++i) counts[i] = static_cast<int>(static_cast<unsigned int>(cnt[i]));
    return graph.res_trans_id_val(counts);
}
