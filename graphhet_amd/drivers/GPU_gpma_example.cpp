// The reference's GPMA example (src/GPU_gpma/GPU_gpma_example.cu) rebuilt on
// the drop-in headers, plus a check: after batches of GPMA updates (inserts,
// weight changes, deletes), every GPMA analytics call must equal the CSR
// drop-in's result on a graph_structure given the same changes.
//   GPU_gpma_example           prints the example's results
//   GPU_gpma_example --check   prints "gpma selftest ok" or the first mismatch
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include <GPU_csr/GPU_csr.hpp>
#include <GPU_csr/algorithm/GPU_BFS_csr.cuh>
#include <GPU_csr/algorithm/GPU_CDLP_csr.cuh>
#include <GPU_csr/algorithm/GPU_PR_csr.cuh>
#include <GPU_csr/algorithm/GPU_SSSP_csr.cuh>
#include <GPU_csr/algorithm/GPU_TC_csr.cuh>
#include <GPU_csr/algorithm/GPU_WCC_csr.cuh>
#include <GPU_gpma/GPU_gpma.hpp>
#include <GPU_gpma/algorithm/GPU_BFS_gpma.cuh>
#include <GPU_gpma/algorithm/GPU_CDLP_gpma.cuh>
#include <GPU_gpma/algorithm/GPU_PR_gpma.cuh>
#include <GPU_gpma/algorithm/GPU_SSSP_gpma.cuh>
#include <GPU_gpma/algorithm/GPU_TC_gpma.cuh>
#include <GPU_gpma/algorithm/GPU_WCC_gpma.cuh>

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

static KEY_TYPE key(int u, int v) { return (static_cast<KEY_TYPE>(u) << 32) | static_cast<KEY_TYPE>(v); }

// gpma_out from OUTs, gpma_in from INs, as the reference example builds them.
static void build_gpma(graph_structure<double>& graph, GPMA& out, GPMA& in) {
    const int n = graph.size();
    DEV_VEC_KEY ko, ki;
    DEV_VEC_VALUE vo, vi;
    for (int u = 0; u < n; ++u) {
        for (const auto& e : graph.OUTs[u]) {
            ko.push_back(key(u, e.first));
            vo.push_back(e.second);
        }
        for (const auto& e : graph.INs[u]) {
            ki.push_back(key(u, e.first));
            vi.push_back(e.second);
        }
    }
    init_csr_gpma(out, static_cast<SIZE_TYPE>(n));
    init_csr_gpma(in, static_cast<SIZE_TYPE>(n));
    init_gpma_from_csr(out, ko, vo);
    init_gpma_from_csr(in, ki, vi);
}

template <typename A, typename B>
static void show(const char* title, const std::vector<std::pair<A, B>>& r) {
    std::cout << title << std::endl;
    for (const auto& x : r) std::cout << x.first << " " << x.second << std::endl;
}

template <typename A, typename B>
static bool same(const std::vector<std::pair<A, B>>& a, const std::vector<std::pair<A, B>>& b, double tol = 0) {
    if (a.size() != b.size()) return false;
    for (std::size_t i = 0; i < a.size(); ++i) {
        if (a[i].first != b[i].first) return false;
        if constexpr (std::is_floating_point<B>::value) {
            if (std::fabs(a[i].second - b[i].second) > tol) return false;
        } else if (a[i].second != b[i].second) {
            return false;
        }
    }
    return true;
}

// Compares every GPMA analytics call with the CSR drop-in on `graph`.
static bool check_all(graph_structure<double>& graph, GPMA& out, GPMA& in, const char* stage) {
    CSR_graph<double> csr = toCSR(graph);
    bool ok = true;
    auto fail = [&](const char* what) {
        std::printf("FAIL %s: %s\n", stage, what);
        ok = false;
    };
    if (!same(Cuda_BFS_optimized(graph, out, "one"), Cuda_BFS(graph, csr, "one"))) fail("BFS");
    if (!same(Cuda_SSSP_optimized(graph, out, "one"), Cuda_SSSP(graph, csr, "one", 10000000000.0))) fail("SSSP");
    if (!same(Cuda_WCC_optimized(graph, in), Cuda_WCC(graph, csr))) fail("WCC");
    if (!same(Cuda_PR_optimized(graph, in, out, 10, 0.85), Cuda_PR(graph, csr, 10, 0.85), 1e-12)) fail("PR");
    if (!same(Cuda_CDLP_optimized(graph, in, out, 10), Cuda_CDLP(graph, csr, 10))) fail("CDLP");
    if (!same(Cuda_TC(graph, out), Cuda_TC(graph, csr))) fail("TC");
    if (out.get_size() != graph.edge_number()) fail("edge count");
    return ok;
}

// One batch applied to both stores (the in-store takes reversed keys) and to
// the graph_structure.
static void batch(graph_structure<double>& graph, GPMA& out, GPMA& in, const std::vector<std::pair<int, int>>& e,
                  const std::vector<double>& w) {
    DEV_VEC_KEY ko, ki;
    DEV_VEC_VALUE vo, vi;
    for (std::size_t i = 0; i < e.size(); ++i) {
        ko.push_back(key(e[i].first, e[i].second));
        ki.push_back(key(e[i].second, e[i].first));
        vo.push_back(w[i]);
        vi.push_back(w[i]);
        if (w[i] == VALUE_NONE) graph.remove_edge(e[i].first, e[i].second);
        else graph.add_edge(e[i].first, e[i].second, w[i]);
    }
    update_gpma(out, ko, vo);
    update_gpma(in, ki, vi);
}

int main(int argc, char** argv) {
    const bool check = argc > 1 && std::strcmp(argv[1], "--check") == 0;
    graph_structure<double> graph = sample_graph();
    GPMA out, in;
    build_gpma(graph, out, in);
    if (!check) {
        show("BFS result:", Cuda_BFS_optimized(graph, out, "one"));
        show("SSSP result:", Cuda_SSSP_optimized(graph, out, "one"));
        show("Connected Components result:", Cuda_WCC_optimized(graph, in));
        show("PageRank result:", Cuda_PR_optimized(graph, in, out, 10, 0.85));
        show("Community Detection result:", Cuda_CDLP_optimized(graph, in, out, 10));
        return 0;
    }
    bool ok = check_all(graph, out, in, "initial");
    const int one = graph.vertex_str_to_id["one"], two = graph.vertex_str_to_id["two"];
    const int three = graph.vertex_str_to_id["three"], four = graph.vertex_str_to_id["four"];
    const int five = graph.vertex_str_to_id["five"], six = graph.vertex_str_to_id["six"];
    // inserts (one of them a loop), a weight change, deletes of a present and an absent edge
    batch(graph, out, in, {{five, one}, {six, two}, {three, three}, {one, three}, {four, five}, {two, one}},
          {0.5, 2.0, 1.0, 0.25, VALUE_NONE, VALUE_NONE});
    ok = check_all(graph, out, in, "after batch 1") && ok;
    // a larger batch: every ordered pair among the named vertices, alternating insert / delete
    std::vector<std::pair<int, int>> e;
    std::vector<double> w;
    const int ids[] = {one, two, three, four, five, six};
    int k = 0;
    for (int u : ids)
        for (int v : ids) {
            e.emplace_back(u, v);
            w.push_back((k++ % 3 == 0) ? VALUE_NONE : 0.125 * k);
        }
    batch(graph, out, in, e, w);
    ok = check_all(graph, out, in, "after batch 2") && ok;
    if (ok) std::printf("gpma selftest ok\n");
    return ok ? 0 : 1;
}
