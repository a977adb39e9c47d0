// Drop-in for GraphHET's include/GPU_csr/algorithm/GPU_BFS_csr.cuh: same
// entry points, backed by libghet's direction-optimising HIP BFS.
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// Raw core (GPU_BFS_csr.cuh:258): device out-CSR (keys = column ids, values =
// weights, row_offsets), returns depths, INT_MAX when unreached.  Depths
// beyond max_depth are reported unreached, as CPU_BFS does (CPU_BFS.hpp:28).
inline std::vector<int> csr_bfs(int* keys, double* values, int* row_offsets, int node_size, int edge_size,
                                int start_node, int max_depth) {
    std::vector<int> result(static_cast<std::size_t>(node_size), INT_MAX);
    ghet_shim::Borrowed b(node_size, edge_size, row_offsets, keys, values);
    if (!b.g || !ghet_shim::ok(ghet_bfs(b.g, start_node, result.data(), nullptr, nullptr), "csr_bfs")) return result;
    if (max_depth != INT_MAX)
        for (int& d : result)
            if (d > max_depth) d = INT_MAX;
    return result;
}

// Cuda_BFS (GPU_BFS_csr.cuh:13,329-335); the source name is looked up with
// operator[] as the reference does (a missing name maps to id 0).
inline std::vector<std::pair<std::string, int>> Cuda_BFS(graph_structure<double>& graph, CSR_graph<double>& csr_graph,
                                                         std::string src_v, int min_depth = 0, int max_depth = INT_MAX) {
    (void)min_depth;
    const int src = graph.vertex_str_to_id[src_v];
    std::vector<int> depth(static_cast<std::size_t>(graph.V), INT_MAX);
    if (!csr_graph.handle || !ghet_shim::ok(ghet_bfs(csr_graph.handle, src, depth.data(), nullptr, nullptr), "Cuda_BFS"))
        return {};
    if (max_depth != INT_MAX)
        for (int& d : depth)
            if (d > max_depth) d = INT_MAX;
    return graph.res_trans_id_val(depth);
}
