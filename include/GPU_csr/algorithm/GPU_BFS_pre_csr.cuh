// Drop-in for GraphHET's GPU_BFS_pre_csr.cuh: BFS depths plus a parent per
// vertex.  The parent is deterministic -- the smallest in-neighbour one level
// up -- one of the valid parents the reference's racy kernels may return.
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// Raw core (GPU_BFS_pre_csr.cuh:62): device out- and in-CSR.
inline std::vector<int> bfs_pre(int* d_row_offsets_out, int* d_column_indices_out, int* d_row_offsets_in,
                                int* d_column_indices_in, int num_nodes, int start_node, std::vector<int>& pre_v,
                                int max_depth) {
    std::vector<int> result(static_cast<std::size_t>(num_nodes), INT_MAX);
    pre_v.assign(static_cast<std::size_t>(num_nodes), -1);
    ghet_shim::Borrowed b(num_nodes, -1, d_row_offsets_out, d_column_indices_out, nullptr, d_row_offsets_in,
                          d_column_indices_in, -1);
    if (b.g) ghet_shim::ok(ghet_bfs_pre(b.g, start_node, result.data(), pre_v.data(), nullptr, nullptr), "bfs_pre");
    if (max_depth != INT_MAX)
        for (std::size_t v = 0; v < result.size(); ++v)
            if (result[v] > max_depth) {
                result[v] = INT_MAX;
                pre_v[v] = -1;
            }
    return result;
}

// Cuda_BFS_pre (GPU_BFS_pre_csr.cuh:15,126-156), including the reference's
// host-side check (GPU_BFS_pre_csr.cuh:133-153): every parent chain is walked
// back towards the source, and a link pre -> now that is not an out-edge of
// pre in graph.OUTs prints "Not found!".  Parents are one level up, so a
// chain ends after at most depth[v] links; the walk is bounded by V anyway.
inline std::vector<std::pair<std::string, int>> Cuda_BFS_pre(graph_structure<double>& graph,
                                                             CSR_graph<double>& csr_graph, std::string src_v,
                                                             std::vector<int>& pre_v, int min_depth = 0,
                                                             int max_depth = INT_MAX) {
    (void)min_depth;
    const int src = graph.vertex_str_to_id[src_v];
    std::vector<int> depth(static_cast<std::size_t>(graph.V), INT_MAX);
    pre_v.assign(static_cast<std::size_t>(graph.V), -1);
    if (!csr_graph.handle ||
        !ghet_shim::ok(ghet_bfs_pre(csr_graph.handle, src, depth.data(), pre_v.data(), nullptr, nullptr),
                       "Cuda_BFS_pre"))
        return {};
    if (max_depth != INT_MAX)
        for (std::size_t v = 0; v < depth.size(); ++v)
            if (depth[v] > max_depth) {
                depth[v] = INT_MAX;
                pre_v[v] = -1;
            }
    for (int v = 0; v < graph.V; ++v) {
        int now = v, up = pre_v[v];
        for (int steps = 0; up != -1 && steps < graph.V; ++steps) {
            bool linked = false;
            for (const auto& e : graph.OUTs[up])
                if (e.first == now) {
                    linked = true;
                    break;
                }
            if (!linked) {
                std::printf("Not found!\n");
                break;
            }
            now = up;
            up = pre_v[up];
        }
    }
    return graph.res_trans_id_val(depth);
}
