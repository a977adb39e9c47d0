// Drop-in for GraphHET's GPU_BFS_gpma.cuh: BFS over the current edges of a
// GPMA store (a snapshot of the dynamic store, then libghet's BFS).
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_BFS_optimized (GPU_BFS_gpma.cuh:13,336-341): depths over out-edges,
// INT_MAX when unreached; the source name is looked up with operator[].
inline std::vector<std::pair<std::string, int>> Cuda_BFS_optimized(graph_structure<double>& graph, GPMA& gpma,
                                                                   std::string src_v, int min_depth = 0,
                                                                   int max_depth = INT_MAX) {
    (void)min_depth;
    const int src = graph.vertex_str_to_id[src_v];
    ghet_gpma::Snapshot s(gpma, GHET_GRAPH_NO_WEIGHTS | GHET_GRAPH_NO_ALL);
    std::vector<int> depth(static_cast<std::size_t>(gpma.row_num), INT_MAX);
    if (!s.g || !ghet_shim::ok(ghet_bfs(s.g, src, depth.data(), nullptr, nullptr), "Cuda_BFS_optimized")) return {};
    if (max_depth != INT_MAX)
        for (int& d : depth)
            if (d > max_depth) d = INT_MAX;
    return graph.res_trans_id_val(depth);
}
