// Exercises every drop-in entry point of include/GPU_csr on one graph and
// prints the results as "<entry> <vertex-name> <value>" lines; the GPU tests
// compare them with the CPU oracle.
//
//   shim_check <edge-file> <source-name>
// edge file: "n" then one "u v w" line per edge (vertex names are u, v).
#include <climits>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include <GPU_csr/GPU_csr.hpp>
#include <GPU_csr/algorithm/GPU_BFS_csr.cuh>
#include <GPU_csr/algorithm/GPU_BFS_pre_csr.cuh>
#include <GPU_csr/algorithm/GPU_CDLP_csr.cuh>
#include <GPU_csr/algorithm/GPU_PR_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_pre_csr.cuh>
#include <GPU_csr/algorithm/GPU_TC_csr.cuh>
#include <GPU_csr/algorithm/GPU_WCC_csr.cuh>

template <typename T>
static void emit(const char* tag, const std::vector<std::pair<std::string, T>>& r) {
    for (const auto& x : r) std::cout << tag << " " << x.first << " " << x.second << "\n";
}
template <typename T>
static void emit_ids(const char* tag, const std::vector<T>& r) {
    for (size_t i = 0; i < r.size(); ++i) std::cout << tag << " " << i << " " << r[i] << "\n";
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: shim_check <edge-file> <source-name>\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    int n;
    in >> n;
    graph_structure<double> graph;
    for (int i = 0; i < n; ++i) graph.add_vertice(std::to_string(i));
    std::string u, v;
    double w;
    while (in >> u >> v >> w) graph.add_edge(u, v, w);
    const std::string src = argv[2];
    std::cout.precision(17);

    CSR_graph<double> csr = toCSR(graph);
    CSR_graph<double> copy = csr;             // deep copy
    CSR_graph<double> moved = std::move(copy);  // move
    emit("Cuda_BFS", Cuda_BFS(graph, moved, src));
    std::vector<int> pre;
    emit("Cuda_BFS_pre", Cuda_BFS_pre(graph, csr, src, pre));
    emit_ids("Cuda_BFS_pre.pre", pre);
    emit("Cuda_SSSP", Cuda_SSSP(graph, csr, src, std::numeric_limits<double>::max()));
    emit("Cuda_SSSP_pre", Cuda_SSSP_pre(graph, csr, src, pre, std::numeric_limits<double>::max()));
    emit_ids("Cuda_SSSP_pre.pre", pre);
    emit("Cuda_WCC", Cuda_WCC(graph, csr));
    emit("Cuda_PR", Cuda_PR(graph, csr, 10, 0.85));
    emit("Cuda_CDLP", Cuda_CDLP(graph, csr, 10));
    const int sid = graph.vertex_str_to_id[src];
    const int V = graph.V, E = static_cast<int>(csr.OUTs_Edges.size());
    emit_ids("csr_bfs", csr_bfs(csr.out_edge, csr.out_edge_weight, csr.out_pointer, V, E, sid, INT_MAX));
    std::vector<int> pre2;
    emit_ids("bfs_pre", bfs_pre(csr.out_pointer, csr.out_edge, csr.in_pointer, csr.in_edge, V, sid, pre2, INT_MAX));
    emit_ids("csr_sssp", csr_sssp(csr.out_edge, csr.out_edge_weight, csr.out_pointer, V, E, sid,
                                  std::numeric_limits<double>::max()));
    emit_ids("csr_sssp_pre", csr_sssp_pre(csr.out_edge, csr.out_edge_weight, csr.out_pointer, V, E, sid, pre2,
                                          std::numeric_limits<double>::max()));
    emit_ids("WCC_GPU", WCC_GPU(graph, csr));
    std::vector<double> pr;
    GPU_PR(graph, csr, pr, 10, 0.85);
    emit_ids("GPU_PR", pr);
    std::vector<std::string> lab;
    CDLP_GPU(graph, csr, lab, 10);
    emit_ids("CDLP_GPU", lab);
    // TC runs on the undirected symmetrisation, as the reference expects.
    graph_structure<double> und = graph;
    for (int a = 0; a < und.V; ++a)
        for (const auto& e : graph.OUTs[a])
            if (e.first != a) und.add_edge(e.first, a, e.second);
    for (int a = 0; a < und.V; ++a) und.remove_edge(a, a);
    CSR_graph<double> ucsr = toCSR(und, false);
    emit("Cuda_TC", Cuda_TC(und, ucsr));
    emit_ids("Cuda_TC_PerNode", Cuda_TC_PerNode(ucsr.out_pointer, ucsr.out_edge, und.V));
    return 0;
}
