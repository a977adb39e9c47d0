// graph_structure<W>: the host graph model every GraphHET entry point takes
// (interface of include/CPU_adj_list/CPU_adj_list.hpp:16-70 in GraphHET).
//
// Same public members and the same semantics: per-vertex OUTs/INs lists of
// (neighbour, weight) kept sorted by neighbour id, an insert of an existing
// edge overwrites its weight, vertex names map to dense ids, removed vertex
// ids are recycled in FIFO order, and results are mapped back to names
// skipping removed ids.  Written for this library; kept header-only so the
// drop-in headers need nothing else.
#pragma once

#include <algorithm>
#include <iostream>
#include <limits>
#include <queue>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace ghet_adj {

template <typename W>
using AdjList = std::vector<std::pair<int, W>>;

template <typename W>
typename AdjList<W>::iterator find_slot(AdjList<W>& list, int key) {
    return std::lower_bound(list.begin(), list.end(), key,
                            [](const std::pair<int, W>& e, int k) { return e.first < k; });
}

template <typename W>
typename AdjList<W>::const_iterator find_slot(const AdjList<W>& list, int key) {
    return std::lower_bound(list.begin(), list.end(), key,
                            [](const std::pair<int, W>& e, int k) { return e.first < k; });
}

// Insert keeping the list sorted; an existing key gets the new weight.
template <typename W>
int upsert(AdjList<W>& list, int key, const W& w) {
    auto it = find_slot(list, key);
    const int pos = static_cast<int>(it - list.begin());
    if (it != list.end() && it->first == key) it->second = w;
    else list.insert(it, std::make_pair(key, w));
    return pos;
}

template <typename W>
void erase_key(AdjList<W>& list, int key) {
    auto it = find_slot(list, key);
    if (it != list.end() && it->first == key) list.erase(it);
}

}  // namespace ghet_adj

template <typename weight_type>
class graph_structure {
public:
    int V = 0;       // number of vertex ids (including removed ones)
    long long E = 0; // number of add_edge(string, string, w) calls

    std::vector<std::vector<std::pair<int, weight_type>>> OUTs;  // OUTs[u]: (v, w) for u -> v
    std::vector<std::vector<std::pair<int, weight_type>>> INs;   // INs[v]: (u, w) for u -> v

    graph_structure() {}
    graph_structure(int n) : V(n), OUTs(n), INs(n) {}
    int size() { return V; }

    inline void add_edge(int e1, int e2, weight_type ec) {
        ghet_adj::upsert(OUTs[e1], e2, ec);
        ghet_adj::upsert(INs[e2], e1, ec);
    }
    inline void remove_edge(int e1, int e2) {
        ghet_adj::erase_key(OUTs[e1], e2);
        ghet_adj::erase_key(INs[e2], e1);
    }
    inline void remove_edge(std::string e1, std::string e2) {
        if (!vertex_str_to_id.count(e1)) {
            std::cerr << "vertex " << e1 << " not exist!" << std::endl;
            return;
        }
        if (!vertex_str_to_id.count(e2)) {
            std::cerr << "vertex " << e2 << " not exist!" << std::endl;
            return;
        }
        remove_edge(vertex_str_to_id[e1], vertex_str_to_id[e2]);
    }
    inline void remove_all_adjacent_edges(int v) {
        for (const auto& e : OUTs[v]) ghet_adj::erase_key(INs[e.first], v);
        for (const auto& e : INs[v]) ghet_adj::erase_key(OUTs[e.first], v);
        std::vector<std::pair<int, weight_type>>().swap(OUTs[v]);
        std::vector<std::pair<int, weight_type>>().swap(INs[v]);
    }
    inline bool contain_edge(int e1, int e2) {
        auto it = ghet_adj::find_slot(OUTs[e1], e2);
        return it != OUTs[e1].end() && it->first == e2;
    }
    inline weight_type edge_weight(int e1, int e2) {
        auto it = ghet_adj::find_slot(OUTs[e1], e2);
        if (it != OUTs[e1].end() && it->first == e2) return it->second;
        return std::numeric_limits<weight_type>::max();
    }
    inline long long int edge_number() {
        long long int m = 0;
        for (const auto& l : OUTs) m += static_cast<long long>(l.size());
        return m;
    }
    inline void print() {
        std::cout << "graph_structure_print:" << std::endl;
        for (int i = 0; i < V; i++) {
            std::cout << "Vertex " << i << " OUTs List: ";
            for (const auto& e : OUTs[i]) std::cout << "<" << e.first << "," << e.second << "> ";
            std::cout << std::endl;
        }
        std::cout << "graph_structure_print END" << std::endl;
    }
    inline void clear() {
        std::vector<std::vector<std::pair<int, weight_type>>>().swap(OUTs);
        std::vector<std::vector<std::pair<int, weight_type>>>().swap(INs);
    }
    inline int out_degree(int v) { return static_cast<int>(OUTs[v].size()); }
    inline int in_degree(int v) { return static_cast<int>(INs[v].size()); }

    std::unordered_map<std::string, int> vertex_str_to_id;      // name -> id
    std::vector<std::pair<std::string, bool>> vertex_id_to_str; // id -> (name, valid)
    std::queue<int> invalid_vertex_id;                         // recyclable ids

    int add_vertice(std::string vertex) {
        auto found = vertex_str_to_id.find(vertex);
        if (found != vertex_str_to_id.end()) return found->second;
        int id;
        if (invalid_vertex_id.empty()) {
            id = V++;
            vertex_id_to_str.emplace_back(vertex, true);
            OUTs.emplace_back();
            INs.emplace_back();
        } else {
            id = invalid_vertex_id.front();
            invalid_vertex_id.pop();
            vertex_id_to_str[id] = std::make_pair(vertex, true);
            std::cout << "Recover vertex id " << id << std::endl;
        }
        vertex_str_to_id[vertex] = id;
        return id;
    }
    void remove_vertice(std::string vertex) {
        auto found = vertex_str_to_id.find(vertex);
        if (found == vertex_str_to_id.end()) {
            std::cerr << "vertex " << vertex << " not exist!" << std::endl;
            return;
        }
        const int v = found->second;
        remove_all_adjacent_edges(v);
        vertex_str_to_id.erase(found);
        vertex_id_to_str[v].second = false;
        invalid_vertex_id.push(v);
    }
    void add_edge(std::string e1, std::string e2, weight_type ec) {
        E++;
        const int v1 = add_vertice(e1);
        const int v2 = add_vertice(e2);
        add_edge(v1, v2, ec);
    }

    // (name, value) for every valid id below res.size().
    template <typename T>
    std::vector<std::pair<std::string, T>> res_trans_id_val(std::vector<T>& res) {
        std::vector<std::pair<std::string, T>> out;
        const int k = static_cast<int>(res.size());
        for (int i = 0; i < k; i++)
            if (vertex_id_to_str[i].second) out.emplace_back(vertex_id_to_str[i].first, res[i]);
        return out;
    }
    // (name, name of the label id) for every valid id.
    std::vector<std::pair<std::string, std::string>> res_trans_id_id(std::vector<int>& wcc_res) {
        std::vector<std::pair<std::string, std::string>> out;
        const int k = static_cast<int>(wcc_res.size());
        for (int i = 0; i < k; i++) {
            if (!vertex_id_to_str[i].second) continue;
            const auto& lab = vertex_id_to_str[wcc_res[i]];
            if (lab.second) out.emplace_back(vertex_id_to_str[i].first, lab.first);
            else std::cerr << "vertex " << lab.first << " not exist!" << std::endl;
        }
        return out;
    }
};
