// CSR_graph<W> and toCSR: the GPU CSR container of GraphHET
// (include/GPU_csr/GPU_csr.hpp:62-118,282-339), backed by libghet.
//
// Public fields keep their names and meaning: host CSR vectors (out, in and
// "all" = in-list then out-list per vertex) and device pointers to the same
// arrays.  Undirected graphs alias in/all to out.  The device arrays live in
// one 256-byte-aligned arena owned by a ghet_graph handle; copying a CSR_graph
// uploads a new handle (deep copy), moving transfers it, the destructor frees
// it.  Plain C++: no CUDA or HIP headers are needed to include this file.
#pragma once

#include <cstdio>
#include <utility>
#include <vector>

#include <CPU_adj_list/CPU_adj_list.hpp>
#include <ghet.h>

namespace csr_detail {

inline void report_ghet_error(int status, const char* operation) {
    if (status != GHET_OK && status != GHET_WARN_FALLBACK)
        std::fprintf(stderr, "ghet error during %s: %s\n", operation, ghet_last_error());
}

}  // namespace csr_detail

template <typename weight_type>
class CSR_graph {
public:
    CSR_graph() = default;
    CSR_graph(const CSR_graph& other)
        : is_directed(other.is_directed),
          INs_Neighbor_start_pointers(other.INs_Neighbor_start_pointers),
          OUTs_Neighbor_start_pointers(other.OUTs_Neighbor_start_pointers),
          ALL_start_pointers(other.ALL_start_pointers),
          INs_Edges(other.INs_Edges),
          OUTs_Edges(other.OUTs_Edges),
          all_Edges(other.all_Edges),
          INs_Edge_weights(other.INs_Edge_weights),
          OUTs_Edge_weights(other.OUTs_Edge_weights),
          E_all(other.E_all) {
        if (!OUTs_Neighbor_start_pointers.empty()) upload();
    }
    CSR_graph(CSR_graph&& other) noexcept { swap(*this, other); }
    CSR_graph& operator=(CSR_graph other) noexcept {
        swap(*this, other);
        return *this;
    }
    ~CSR_graph() {
        if (handle) ghet_graph_destroy(handle);
    }

    friend void swap(CSR_graph& a, CSR_graph& b) noexcept {
        using std::swap;
        swap(a.is_directed, b.is_directed);
        swap(a.INs_Neighbor_start_pointers, b.INs_Neighbor_start_pointers);
        swap(a.OUTs_Neighbor_start_pointers, b.OUTs_Neighbor_start_pointers);
        swap(a.ALL_start_pointers, b.ALL_start_pointers);
        swap(a.INs_Edges, b.INs_Edges);
        swap(a.OUTs_Edges, b.OUTs_Edges);
        swap(a.all_Edges, b.all_Edges);
        swap(a.INs_Edge_weights, b.INs_Edge_weights);
        swap(a.OUTs_Edge_weights, b.OUTs_Edge_weights);
        swap(a.in_pointer, b.in_pointer);
        swap(a.out_pointer, b.out_pointer);
        swap(a.in_edge, b.in_edge);
        swap(a.out_edge, b.out_edge);
        swap(a.all_pointer, b.all_pointer);
        swap(a.all_edge, b.all_edge);
        swap(a.in_edge_weight, b.in_edge_weight);
        swap(a.out_edge_weight, b.out_edge_weight);
        swap(a.E_all, b.E_all);
        swap(a.handle, b.handle);
    }

    bool is_directed = true;
    // Neighbor_start_pointers[i] is where vertex i's neighbours start; [V] = edge count.
    std::vector<int> INs_Neighbor_start_pointers, OUTs_Neighbor_start_pointers, ALL_start_pointers;
    std::vector<int> INs_Edges, OUTs_Edges, all_Edges;
    std::vector<weight_type> INs_Edge_weights, OUTs_Edge_weights;
    int *in_pointer = nullptr, *out_pointer = nullptr, *in_edge = nullptr, *out_edge = nullptr;
    int *all_pointer = nullptr, *all_edge = nullptr;
    double *in_edge_weight = nullptr, *out_edge_weight = nullptr;
    std::size_t E_all = 0;

    // The library handle behind the device pointers.
    ghet_graph* handle = nullptr;

    // Creates the device copy from the host vectors (called by toCSR).
    void upload() {
        if (handle) {
            ghet_graph_destroy(handle);
            handle = nullptr;
        }
        const int n = static_cast<int>(OUTs_Neighbor_start_pointers.size()) - 1;
        std::vector<double> ow(OUTs_Edge_weights.begin(), OUTs_Edge_weights.end());
        std::vector<double> iw(INs_Edge_weights.begin(), INs_Edge_weights.end());
        const uint32_t flags = is_directed ? 0u : GHET_GRAPH_UNDIRECTED;
        const int st = ghet_graph_create(
            n, OUTs_Neighbor_start_pointers.data(), OUTs_Edges.data(), ow.data(),
            is_directed ? INs_Neighbor_start_pointers.data() : nullptr, is_directed ? INs_Edges.data() : nullptr,
            is_directed ? iw.data() : nullptr, is_directed ? ALL_start_pointers.data() : nullptr,
            is_directed ? all_Edges.data() : nullptr, flags, &handle);
        csr_detail::report_ghet_error(st, "toCSR upload");
        if (st != GHET_OK) {
            handle = nullptr;
            return;
        }
        ghet_device_ptrs p;
        ghet_graph_device_ptrs(handle, &p);
        out_pointer = p.out_rowptr;
        out_edge = p.out_col;
        out_edge_weight = p.out_w;
        in_pointer = p.in_rowptr;
        in_edge = p.in_col;
        in_edge_weight = p.in_w;
        all_pointer = p.all_rowptr;
        all_edge = p.all_col;
    }
};

// Flattens graph.OUTs / graph.INs (sorted neighbour order) into CSR vectors
// and uploads them.  For directed graphs the "all" list of a vertex is its
// in-list followed by its out-list (GPU_csr.hpp:315-330 of GraphHET).
template <typename weight_type>
CSR_graph<weight_type> toCSR(graph_structure<weight_type>& graph, bool is_directed = true) {
    CSR_graph<weight_type> csr;
    csr.is_directed = is_directed;
    const int n = graph.size();
    auto flatten = [n](const std::vector<std::vector<std::pair<int, weight_type>>>& lists, std::vector<int>& ptr,
                       std::vector<int>& idx, std::vector<weight_type>& w) {
        ptr.assign(static_cast<std::size_t>(n) + 1, 0);
        std::size_t total = 0;
        for (int v = 0; v < n; ++v) total += lists[v].size();
        idx.reserve(total);
        w.reserve(total);
        for (int v = 0; v < n; ++v) {
            ptr[v] = static_cast<int>(idx.size());
            for (const auto& e : lists[v]) {
                idx.push_back(e.first);
                w.push_back(e.second);
            }
        }
        ptr[n] = static_cast<int>(idx.size());
    };
    flatten(graph.OUTs, csr.OUTs_Neighbor_start_pointers, csr.OUTs_Edges, csr.OUTs_Edge_weights);
    if (is_directed) {
        flatten(graph.INs, csr.INs_Neighbor_start_pointers, csr.INs_Edges, csr.INs_Edge_weights);
        csr.ALL_start_pointers.assign(static_cast<std::size_t>(n) + 1, 0);
        csr.all_Edges.reserve(csr.INs_Edges.size() + csr.OUTs_Edges.size());
        for (int v = 0; v < n; ++v) {
            csr.ALL_start_pointers[v] = static_cast<int>(csr.all_Edges.size());
            for (const auto& e : graph.INs[v]) csr.all_Edges.push_back(e.first);
            for (const auto& e : graph.OUTs[v]) csr.all_Edges.push_back(e.first);
        }
        csr.ALL_start_pointers[n] = static_cast<int>(csr.all_Edges.size());
    }
    csr.E_all = is_directed ? csr.INs_Edges.size() + csr.OUTs_Edges.size() : csr.OUTs_Edges.size();
    csr.upload();
    return csr;
}
