// LDBC-style driver over the drop-in headers: the call shapes of GraphHET's
// src/LDBC/LDBC_GPU_csr.cu (reads the data directory and graph name from
// stdin, loads <name>.properties/.v/.e, builds the CSR with toCSR, runs every
// algorithm the properties list -- SSSP always, as the reference does
// (LDBC_GPU_csr.cu:137-139) -- checks each against <name>-<ALGO> baselines,
// prints the result row and appends it to ./result-gpu-csr.csv).
//
// Built with g++ against libghet.so; no HIP or CUDA headers.  Each result
// cell is the average seconds of the timed run, "wrong" when the checker
// rejects it, "failed!" when the call threw, "N/A" when not requested.
#include <GPU_csr/algorithm/GPU_BFS_csr.cuh>
#include <GPU_csr/algorithm/GPU_BFS_pre_csr.cuh>
#include <GPU_csr/algorithm/GPU_CDLP_csr.cuh>
#include <GPU_csr/algorithm/GPU_PR_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_pre_csr.cuh>
#include <GPU_csr/algorithm/GPU_WCC_csr.cuh>
#include <LDBC/checker.hpp>
#include <LDBC/ldbc.hpp>

#include <chrono>
#include <cstdio>
#include <iostream>
#include <limits>
#include <string>
#include <utility>
#include <vector>

namespace {

using Row = std::vector<std::pair<std::string, std::string>>;

template <typename F>
double seconds(int reps, F&& f) {
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int r = 0; r < reps; ++r) f();
    const auto t1 = std::chrono::high_resolution_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / reps;
}

void gpu_usage(size_t free_before) {
    size_t free_after = 0, total = 0;
    if (ghet_device_mem_info(&free_after, &total) != GHET_OK || total == 0) return;
    const double used = static_cast<double>(free_before) - static_cast<double>(free_after);
    std::printf("GPU usage: %.5f MB / %.5f MB (%.5f%% used)\n", used / (1024.0 * 1024.0), total / (1024.0 * 1024.0),
                used * 100.0 / static_cast<double>(total));
}

// One warm-up call, then `reps` timed calls; the checker validates the output.
template <typename Run, typename Check>
void checked(Row& row, const std::string& name, const char* label, int reps, Run&& run, Check&& check) {
    try {
        auto out = run();
        const double s = seconds(reps, [&] { out = run(); });
        std::printf("%s cost time: %f s\n", label, s);
        row.emplace_back(name, check(out) ? std::to_string(s) : "wrong");
    } catch (...) {
        row.emplace_back(name, "failed!");
    }
}

}  // namespace

int main() {
    std::ios::sync_with_stdio(false);
    std::string dir, name;
    std::cout << "Please input the data directory: " << std::endl;
    std::cin >> dir;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    std::cout << "Please input the graph name: " << std::endl;
    std::cin >> name;

    Row row;
    LDBC<double> graph(dir, name);
    graph.read_config(dir + name + ".properties");
    std::printf("load_ldbc_time cost time: %f s\n", seconds(1, [&] { graph.load_graph(); }));

    size_t free_before = 0, total_before = 0;
    ghet_device_mem_info(&free_before, &total_before);
    CSR_graph<double> csr;
    const double to_csr = seconds(1, [&] { csr = toCSR(graph, graph.is_directed); });
    std::cout << "Number of vertices: " << csr.OUTs_Neighbor_start_pointers.size() - 1 << std::endl;
    std::cout << "Number of edges: " << csr.OUTs_Edges.size() << std::endl;
    std::printf("graph_to_csr_time cost time: %f s\n", to_csr);
    gpu_usage(free_before);

    constexpr int kReps = 1;
    const double inf = std::numeric_limits<double>::max();
    if (graph.sup_bfs)
        checked(row, "BFS", "GPU BFS", kReps, [&] { return Cuda_BFS(graph, csr, graph.bfs_src_name); },
                [&](auto& o) { return Bfs_checker(graph, o, graph.base_path + "-BFS"); });
    else row.emplace_back("BFS", "N/A");

    graph.sup_sssp = true;
    checked(row, "SSSP", "GPU SSSP", kReps, [&] { return Cuda_SSSP(graph, csr, graph.sssp_src_name, inf); },
            [&](auto& o) { return SSSP_checker(graph, o, graph.base_path + "-SSSP"); });
    try {
        std::vector<int> pre;
        (void)Cuda_SSSP_pre(graph, csr, graph.sssp_src_name, pre, inf);
    } catch (...) {
    }

    if (graph.sup_wcc)
        checked(row, "WCC", "GPU WCC", kReps, [&] { return Cuda_WCC(graph, csr); },
                [&](auto& o) { return WCC_checker(graph, o, graph.base_path + "-WCC"); });
    else row.emplace_back("WCC", "N/A");

    if (graph.sup_pr)
        checked(row, "PR", "GPU PageRank", kReps, [&] { return Cuda_PR(graph, csr, graph.pr_its, graph.pr_damping); },
                [&](auto& o) { return PR_checker(graph, o, graph.base_path + "-PR"); });
    else row.emplace_back("PR", "N/A");

    if (graph.sup_cdlp)
        checked(row, "CDLP", "GPU Community Detection", kReps, [&] { return Cuda_CDLP(graph, csr, graph.cdlp_max_its); },
                [&](auto& o) { return CDLP_checker(graph, o, graph.base_path + "-CDLP"); });
    else row.emplace_back("CDLP", "N/A");

    std::cout << "Result: " << std::endl;
    for (size_t i = 0; i < row.size(); ++i) std::cout << row[i].second << (i + 1 < row.size() ? "," : "");
    std::cout << std::endl;
    graph.save_to_CSV(row, "./result-gpu-csr.csv");
    return 0;
}
