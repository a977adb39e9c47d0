// ORACLE -- test infrastructure only (see ghet_oracle.cpp).
//
// C entry points over GraphHET's own CPU algorithms, compiled from the
// reference's headers where they lie (-I <reference>/include) into
// oracle/_ref/libghet_ref.so.  This file holds only the glue: it fills a
// graph_structure<double> from CSR arrays and calls the reference functions
// unchanged.  Used to write tests/golden/ fixtures and, when present, as the
// CPU baseline.
#include <CPU_adj_list/CPU_adj_list.hpp>
#include <CPU_adj_list/algorithm/CPU_BFS.hpp>
#include <CPU_adj_list/algorithm/CPU_Community_Detection.hpp>
#include <CPU_adj_list/algorithm/CPU_PageRank.hpp>
#include <CPU_adj_list/algorithm/CPU_TC.hpp>
#include <CPU_adj_list/algorithm/CPU_connected_components.hpp>
#include <CPU_adj_list/algorithm/CPU_shortest_paths.hpp>

#include <chrono>
#include <cstdint>
#include <string>

namespace {

// OUTs from the out-CSR and INs from the in-CSR; names are decimal ids.
graph_structure<double> make_graph(int32_t n, const int32_t* orp, const int32_t* ocol, const double* ow,
                                   const int32_t* irp, const int32_t* icol, const double* iw) {
    graph_structure<double> g(n);
    for (int32_t v = 0; v < n; ++v) {
        for (int32_t e = orp[v]; e < orp[v + 1]; ++e) g.OUTs[v].push_back({ocol[e], ow ? ow[e] : 1.0});
        if (irp)
            for (int32_t e = irp[v]; e < irp[v + 1]; ++e) g.INs[v].push_back({icol[e], iw ? iw[e] : 1.0});
        g.vertex_id_to_str.push_back({std::to_string(v), true});
        g.vertex_str_to_id[std::to_string(v)] = v;
    }
    return g;
}

}  // namespace

extern "C" {

void ref_bfs(int32_t n, const int32_t* rp, const int32_t* col, int32_t src, int32_t* depth) {
    auto g = make_graph(n, rp, col, nullptr, nullptr, nullptr, nullptr);
    std::vector<int> d = CPU_BFS(g.OUTs, src);
    for (int32_t i = 0; i < n; ++i) depth[i] = d[i];
}

void ref_sssp(int32_t n, const int32_t* rp, const int32_t* col, const double* w, int32_t src, double* dist) {
    auto g = make_graph(n, rp, col, w, nullptr, nullptr, nullptr);
    std::vector<double> d = CPU_shortest_paths(g.OUTs, src);
    for (int32_t i = 0; i < n; ++i) dist[i] = d[i];
}

void ref_wcc(int32_t n, const int32_t* orp, const int32_t* ocol, const int32_t* irp, const int32_t* icol,
             int32_t* comp) {
    auto g = make_graph(n, orp, ocol, nullptr, irp, icol, nullptr);
    std::vector<int> c = CPU_connected_components(g.OUTs, g.INs);
    for (int32_t i = 0; i < n; ++i) comp[i] = c[i];
}

void ref_pr(int32_t n, const int32_t* orp, const int32_t* ocol, const int32_t* irp, const int32_t* icol, double damp,
            int32_t iters, double* rank) {
    auto g = make_graph(n, orp, ocol, nullptr, irp, icol, nullptr);
    std::vector<double> r = PageRank(g.INs, g.OUTs, damp, iters);
    for (int32_t i = 0; i < n; ++i) rank[i] = r[i];
}

void ref_cdlp(int32_t n, const int32_t* orp, const int32_t* ocol, const int32_t* irp, const int32_t* icol,
              int32_t iters, int32_t* label) {
    auto g = make_graph(n, orp, ocol, nullptr, irp, icol, nullptr);
    std::vector<std::string> r = CDLP(g, iters);
    for (int32_t i = 0; i < n; ++i) label[i] = std::stoi(r[i]);
}

// Timed forms for the CPU baseline: the graph_structure is built once
// (construction is outside the reference's timed region too,
// LDBC_GPU_csr.cu:86-90) and each call returns the seconds spent in the
// reference function alone.
void* ref_graph_create(int32_t n, const int32_t* orp, const int32_t* ocol, const double* ow, const int32_t* irp,
                       const int32_t* icol, const double* iw) {
    return new graph_structure<double>(make_graph(n, orp, ocol, ow, irp, icol, iw));
}
void ref_graph_free(void* h) { delete static_cast<graph_structure<double>*>(h); }

double ref_graph_bfs(void* h, int32_t src, int32_t* depth) {
    auto& g = *static_cast<graph_structure<double>*>(h);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> d = CPU_BFS(g.OUTs, src);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (depth) for (size_t i = 0; i < d.size(); ++i) depth[i] = d[i];
    return s;
}

double ref_graph_sssp(void* h, int32_t src, double* dist) {
    auto& g = *static_cast<graph_structure<double>*>(h);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> d = CPU_shortest_paths(g.OUTs, src);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dist) for (size_t i = 0; i < d.size(); ++i) dist[i] = d[i];
    return s;
}

double ref_graph_pr(void* h, double damp, int32_t iters, double* rank) {
    auto& g = *static_cast<graph_structure<double>*>(h);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> r = PageRank(g.INs, g.OUTs, damp, iters);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rank) for (size_t i = 0; i < r.size(); ++i) rank[i] = r[i];
    return s;
}

double ref_graph_wcc(void* h, int32_t* comp) {
    auto& g = *static_cast<graph_structure<double>*>(h);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> c = CPU_connected_components(g.OUTs, g.INs);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (comp) for (size_t i = 0; i < c.size(); ++i) comp[i] = c[i];
    return s;
}

double ref_graph_cdlp(void* h, int32_t iters, int32_t* label) {
    auto& g = *static_cast<graph_structure<double>*>(h);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::string> r = CDLP(g, iters);
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (label) for (size_t i = 0; i < r.size(); ++i) label[i] = std::stoi(r[i]);
    return s;
}

void ref_tc(int32_t n, const int32_t* rp, const int32_t* col, int64_t* cnt) {
    auto g = make_graph(n, rp, col, nullptr, nullptr, nullptr, nullptr);
    std::vector<int> c = Tri_Counting(g.OUTs);
    for (int32_t i = 0; i < n; ++i) cnt[i] = c[i];
}

}  // extern "C"
