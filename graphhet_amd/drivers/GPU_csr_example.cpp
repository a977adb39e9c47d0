// The reference's GPU CSR example (src/GPU_csr/GPU_csr_example.cu) rebuilt on
// the drop-in headers: the same 6-vertex graph and the same call shapes,
// compiled with a plain C++ compiler against libghet.
#include <chrono>
#include <cstdio>
#include <iostream>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include <GPU_csr/GPU_csr.hpp>
#include <GPU_csr/algorithm/GPU_BFS_csr.cuh>
#include <GPU_csr/algorithm/GPU_CDLP_csr.cuh>
#include <GPU_csr/algorithm/GPU_PR_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_pre_csr.cuh>
#include <GPU_csr/algorithm/GPU_TC_csr.cuh>
#include <GPU_csr/algorithm/GPU_WCC_csr.cuh>

static graph_structure<double> sample_graph() {
    graph_structure<double> g;
    for (const char* name : {"one", "two", "three", "four", "five", "R"}) g.add_vertice(name);
    g.add_edge("one", "two", 0.8);
    g.add_edge("two", "three", 1);
    g.add_edge("two", "R", 1);
    g.add_edge("two", "four", 0.1);
    g.add_edge("R", "three", 1);
    g.add_edge("one", "three", 1);
    g.add_edge("one", "four", 1);
    g.add_edge("four", "three", 1);
    g.add_edge("four", "five", 1);
    g.remove_vertice("R");
    g.add_vertice("six");
    g.remove_edge("two", "four");
    g.add_edge("one", "six", 1);
    return g;
}

template <typename A, typename B>
static void show(const char* title, const std::vector<std::pair<A, B>>& r) {
    std::cout << title << std::endl;
    for (const auto& x : r) std::cout << x.first << " " << x.second << std::endl;
}

int main() {
    graph_structure<double> graph = sample_graph();
    auto t0 = std::chrono::high_resolution_clock::now();
    CSR_graph<double> csr = toCSR(graph);
    auto t1 = std::chrono::high_resolution_clock::now();
    std::printf("LDBC to CSR cost time: %f s\n", std::chrono::duration<double>(t1 - t0).count());
    show("BFS result:", Cuda_BFS(graph, csr, "one"));
    show("Connected Components result:", Cuda_WCC(graph, csr));
    std::vector<int> pre_v;
    show("SSSP result:", Cuda_SSSP_pre(graph, csr, "one", pre_v, std::numeric_limits<double>::max()));
    show("PageRank result:", Cuda_PR(graph, csr, 10, 0.85));
    show("Community Detection result:", Cuda_CDLP(graph, csr, 10));
    return 0;
}
