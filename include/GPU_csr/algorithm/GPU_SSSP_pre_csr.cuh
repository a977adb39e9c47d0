// Drop-in for GraphHET's GPU_SSSP_pre_csr.cuh: distances plus a parent per
// vertex, with no per-vertex spin lock (GPU_SSSP_pre_csr.cuh:215 can livelock
// on wave64).  Parent = the smallest u with d[u] + w(u,v) == d[v], computed
// after the distances settle, so it is deterministic.
#pragma once

#include <cmath>

#include <GPU_csr/algorithm/ghet_shim.hpp>

// Raw core (GPU_SSSP_pre_csr.cuh:256): device out-CSR.
inline std::vector<double> csr_sssp_pre(int* keys, double* values, int* row_offsets, SIZE_TYPE node_size,
                                        SIZE_TYPE edge_size, SIZE_TYPE start_node, std::vector<int>& pre_v,
                                        double INF = 10000000000) {
    std::vector<double> result(node_size, INF);
    pre_v.assign(node_size, -1);
    ghet_shim::Borrowed b(static_cast<int>(node_size), edge_size, row_offsets, keys, values);
    if (b.g)
        ghet_shim::ok(ghet_sssp_pre(b.g, static_cast<int>(start_node), INF, result.data(), pre_v.data(), nullptr,
                                    nullptr),
                      "csr_sssp_pre");
    return result;
}

// Cuda_SSSP_pre (GPU_SSSP_pre_csr.cuh:332-368), including the reference's
// host-side check that every parent chain sums to the distance.
inline std::vector<std::pair<std::string, double>> Cuda_SSSP_pre(graph_structure<double>& graph,
                                                                 CSR_graph<double>& csr_graph, std::string src_v,
                                                                 std::vector<int>& pre_v, double max_dis) {
    const int src = graph.vertex_str_to_id[src_v];
    std::vector<double> dist(static_cast<std::size_t>(graph.V), max_dis);
    pre_v.assign(static_cast<std::size_t>(graph.V), -1);
    if (!csr_graph.handle ||
        !ghet_shim::ok(ghet_sssp_pre(csr_graph.handle, src, max_dis, dist.data(), pre_v.data(), nullptr, nullptr),
                       "Cuda_SSSP_pre"))
        return {};
    for (int i = 0; i < graph.V; i++) {
        int pre = pre_v[i], now = i;
        if (pre == -1) continue;
        double sum = 0;
        while (pre != -1) {
            bool found = false;
            for (const auto& e : graph.OUTs[pre])
                if (e.first == now) {
                    sum += e.second;
                    now = pre;
                    pre = pre_v[pre];
                    found = true;
                    break;
                }
            if (!found) {
                std::printf("Not found!\n");
                break;
            }
        }
        if (std::fabs(sum - dist[i]) > 1e-4)
            std::printf("Error: pre_v is wrong!\nvertex: %d, sum: %.5lf, dis: %.5lf\n", i, sum, dist[i]);
    }
    return graph.res_trans_id_val(dist);
}
