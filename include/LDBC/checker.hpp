// Result checkers against LDBC Graphalytics baseline files -- the interface
// and conventions of GraphHET's include/LDBC/checker.hpp, written for this
// library.  A baseline file holds one "vertex value" line per vertex, in any
// order, and must cover every vertex exactly once:
//   BFS   exact depth; the result's INT_MAX matches the baseline's
//         9223372036854775807 (unreached)                    (checker.hpp:57-58)
//   SSSP  |result - base| <= 1e-4 * base; "infinity"/"inf" matches DBL_MAX
//                                                             (checker.hpp:217-235)
//   PR    |result - base| <= 1e-4 * base                      (checker.hpp:295-296)
//   WCC   same partition into components (labels may differ) (checker.hpp:80-163)
//   CDLP  exact label names                                   (checker.hpp:361)
// Every checker prints why it fails and returns false; true when correct.
#pragma once

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include <CPU_adj_list/CPU_adj_list.hpp>
#include <CPU_adj_list/parse_string.hpp>

// Orders components by their smallest member (vectors are sorted first).
inline bool compare(std::vector<int>& a, std::vector<int>& b) { return a[0] < b[0]; }

namespace ldbc_check {

// Walks a baseline file: calls on_line(vertex_id, value_token, name) for each
// "name value" line; false (with a message) on a format error, an unknown
// vertex, more lines than `expected`, or a failing on_line.
inline bool walk_baseline(graph_structure<double>& graph, const std::string& path, int expected, const char* what,
                          const std::function<bool(int, const std::string&, const std::string&)>& on_line) {
    std::ifstream in(path);
    if (!in.is_open()) {
        std::cout << "Baseline file not found!" << std::endl;
        return false;
    }
    int seen = 0;
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> t = parse_string(line, " ");
        if (t.size() != 2) {
            std::cout << "Baseline file format error!" << std::endl;
            return false;
        }
        if (seen >= expected) {
            std::cout << "Size of baseline file is larger than the result!" << std::endl;
            return false;
        }
        auto it = graph.vertex_str_to_id.find(t[0]);
        if (it == graph.vertex_str_to_id.end()) {
            std::cout << "Baseline file contains a vertex that is not in the graph!" << std::endl;
            return false;
        }
        if (!on_line(it->second, t[1], t[0])) return false;
        ++seen;
    }
    if (seen != expected) {
        std::cout << "Size of baseline file is smaller than the result!" << std::endl;
        return false;
    }
    (void)what;
    return true;
}

template <typename T>
std::vector<T> by_id(graph_structure<double>& graph, std::vector<std::pair<std::string, T>>& res, T fill) {
    std::vector<T> v(static_cast<size_t>(graph.V), fill);
    for (auto& p : res) v[graph.vertex_str_to_id[p.first]] = p.second;
    return v;
}

inline bool within(double got, double base) { return std::fabs(got - base) <= 0.0001 * base; }

}  // namespace ldbc_check

inline bool Bfs_checker(graph_structure<double>& graph, std::vector<std::pair<std::string, int>>& res,
                        std::string base_line_file) {
    if (static_cast<int>(res.size()) != graph.V) {
        std::cout << "Size of BFS results is not equal to the number of vertices!" << std::endl;
        return false;
    }
    const std::vector<int> got = ldbc_check::by_id(graph, res, -1);
    const bool ok = ldbc_check::walk_baseline(
        graph, base_line_file, static_cast<int>(res.size()), "BFS",
        [&](int v, const std::string& val, const std::string& name) {
            const long long base = std::stoll(val);
            if (got[v] == base || (got[v] == INT_MAX && base == LLONG_MAX)) return true;
            std::cout << "Baseline file and GPU BFS results are not the same!" << std::endl
                      << "Baseline file: " << name << " " << val << std::endl
                      << "BFS result: " << graph.vertex_id_to_str[v].first << " " << got[v] << std::endl;
            return false;
        });
    if (ok) std::cout << "BFS results are correct!" << std::endl;
    return ok;
}

inline bool SSSP_checker(graph_structure<double>& graph, std::vector<std::pair<std::string, double>>& res,
                         std::string base_line_file) {
    if (static_cast<int>(res.size()) != graph.V) {
        std::cout << "Size of SSSP results is not equal to the number of vertices!" << std::endl;
        return false;
    }
    const std::vector<double> got = ldbc_check::by_id(graph, res, static_cast<double>(INT_MAX));
    const bool ok = ldbc_check::walk_baseline(
        graph, base_line_file, static_cast<int>(res.size()), "SSSP",
        [&](int v, const std::string& val, const std::string& name) {
            const bool same = (val == "infinity" || val == "inf") ? got[v] == std::numeric_limits<double>::max()
                                                                   : ldbc_check::within(got[v], std::stod(val));
            if (same) return true;
            std::cout << "Baseline file and SSSP results are not the same!" << std::endl
                      << "Baseline file: " << name << " " << val << std::endl
                      << "SSSP result: " << graph.vertex_id_to_str[v].first << " " << got[v] << std::endl;
            return false;
        });
    if (ok) std::cout << "SSSP results are correct!" << std::endl;
    return ok;
}

inline bool PR_checker(graph_structure<double>& graph, std::vector<std::pair<std::string, double>>& res,
                       std::string base_line_file) {
    const std::vector<double> got = ldbc_check::by_id(graph, res, 0.0);
    if (static_cast<int>(res.size()) != graph.V) {
        std::cout << "Size of PageRank results is not equal to the number of vertices!" << std::endl;
        return false;
    }
    const bool ok = ldbc_check::walk_baseline(
        graph, base_line_file, static_cast<int>(res.size()), "PR",
        [&](int v, const std::string& val, const std::string& name) {
            if (ldbc_check::within(got[v], std::stod(val))) return true;
            std::cout << "Baseline file and PageRank results are not the same!" << std::endl
                      << "Baseline file: " << name << " " << val << std::endl
                      << "PageRank result: " << graph.vertex_id_to_str[v].first << " " << got[v] << std::endl;
            return false;
        });
    if (ok) std::cout << "PageRank results are correct!" << std::endl;
    return ok;
}

// Labels are compared by position: result i belongs to vertex id i (results
// list the valid ids in order, as res_trans_id_id produces them).
inline bool CDLP_checker(graph_structure<double>& graph, std::vector<std::pair<std::string, std::string>>& res,
                         std::string base_line_file) {
    std::vector<std::string> got;
    got.reserve(res.size());
    for (auto& p : res) got.push_back(p.second);
    if (static_cast<int>(res.size()) != graph.V) {
        std::cout << "Size of CDLP results is not equal to the number of vertices!" << std::endl;
        return false;
    }
    const bool ok = ldbc_check::walk_baseline(
        graph, base_line_file, static_cast<int>(res.size()), "CDLP",
        [&](int v, const std::string& val, const std::string& name) {
            if (got[v] == val) return true;
            std::cout << "Baseline file and CDLP results are not the same!" << std::endl
                      << "Baseline file: " << name << " " << val << std::endl
                      << "CDLP result: " << graph.vertex_id_to_str[v].first << " " << got[v] << std::endl;
            return false;
        });
    if (ok) std::cout << "CDLP results are correct!" << std::endl;
    return ok;
}

// Same partition: components (sorted member lists, ordered by smallest
// member) of the result and of the baseline must be identical.
inline bool WCC_checker(graph_structure<double>& graph, std::vector<std::pair<std::string, std::string>>& res,
                        std::string base_line_file) {
    auto components = [](std::vector<std::vector<int>>& groups) {
        std::vector<std::vector<int>> out;
        for (auto& g : groups)
            if (!g.empty()) {
                std::sort(g.begin(), g.end());
                out.push_back(g);
            }
        std::sort(out.begin(), out.end(), compare);
        return out;
    };
    std::vector<std::vector<int>> mine(static_cast<size_t>(graph.V));
    for (auto& p : res) mine[graph.vertex_str_to_id[p.second]].push_back(graph.vertex_str_to_id[p.first]);
    std::vector<std::vector<int>> got = components(mine);

    std::ifstream in(base_line_file);
    if (!in.is_open()) {
        std::cout << "Baseline file not found!" << std::endl;
        return false;
    }
    std::unordered_map<std::string, int> label_group;
    std::vector<std::vector<int>> theirs;
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> t = parse_string(line, " ");
        if (t.size() != 2) {
            std::cout << "Baseline file format error!" << std::endl;
            return false;
        }
        auto it = label_group.find(t[1]);
        if (it == label_group.end()) {
            it = label_group.emplace(t[1], static_cast<int>(theirs.size())).first;
            theirs.emplace_back();
        }
        theirs[it->second].push_back(graph.vertex_str_to_id[t[0]]);
    }
    std::vector<std::vector<int>> base = components(theirs);
    if (got.size() != base.size()) {
        std::cout << "Baseline file and WCC results are not the same!" << std::endl
                  << "Baseline total component is " << base.size() << std::endl
                  << "WCC result total component is " << got.size() << std::endl;
        return false;
    }
    for (size_t i = 0; i < got.size(); ++i) {
        if (got[i] != base[i]) {
            std::cout << "Baseline file and WCC results are not the same!" << std::endl
                      << "Baseline component size is " << base[i].size() << std::endl
                      << "WCC result component size is " << got[i].size() << std::endl;
            return false;
        }
    }
    std::cout << "WCC results are correct!" << std::endl;
    return true;
}
