// Drop-in for GraphHET's GPU_SSSP_gpma.cuh: shortest paths over the current
// weighted edges of a GPMA store.
#pragma once

#include <GPU_gpma/GPU_gpma.hpp>

// Cuda_SSSP_optimized (GPU_SSSP_gpma.cuh:13,334-339): unreached vertices get
// max_dis, the reference's default 1e10.
inline std::vector<std::pair<std::string, double>> Cuda_SSSP_optimized(graph_structure<double>& graph, GPMA& gpma,
                                                                       std::string src_v, double max_dis = 10000000000) {
    const int src = graph.vertex_str_to_id[src_v];
    ghet_gpma::Snapshot s(gpma, GHET_GRAPH_NO_IN);
    std::vector<double> dist(static_cast<std::size_t>(gpma.row_num), max_dis);
    if (!s.g || !ghet_shim::ok(ghet_sssp(s.g, src, max_dis, dist.data(), nullptr, nullptr), "Cuda_SSSP_optimized"))
        return {};
    return graph.res_trans_id_val(dist);
}
