// Drop-in for GraphHET's GPU_SSSP_csr.cuh: FP64 single-source shortest paths
// (libghet near-far buckets; identical distances to the reference's Dijkstra).
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

// Raw core (GPU_SSSP_csr.cuh:253): device out-CSR; unreached vertices keep INF.
inline std::vector<double> csr_sssp(int* keys, double* values, int* row_offsets, SIZE_TYPE node_size,
                                    SIZE_TYPE edge_size, SIZE_TYPE start_node, double INF = 10000000000) {
    std::vector<double> result(node_size, INF);
    ghet_shim::Borrowed b(static_cast<int>(node_size), edge_size, row_offsets, keys, values);
    if (b.g) ghet_shim::ok(ghet_sssp(b.g, static_cast<int>(start_node), INF, result.data(), nullptr, nullptr), "csr_sssp");
    return result;
}

// Cuda_SSSP (GPU_SSSP_csr.cuh:321-327).  The reference's CSR overload has no
// default for max_dis; its drivers pass DBL_MAX.
inline std::vector<std::pair<std::string, double>> Cuda_SSSP(graph_structure<double>& graph,
                                                             CSR_graph<double>& csr_graph, std::string src_v,
                                                             double max_dis) {
    const int src = graph.vertex_str_to_id[src_v];
    std::vector<double> dist(static_cast<std::size_t>(graph.V), max_dis);
    if (!csr_graph.handle ||
        !ghet_shim::ok(ghet_sssp(csr_graph.handle, src, max_dis, dist.data(), nullptr, nullptr), "Cuda_SSSP"))
        return {};
    return graph.res_trans_id_val(dist);
}
