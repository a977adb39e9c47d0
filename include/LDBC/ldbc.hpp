// LDBC<W>: a graph_structure read from an LDBC Graphalytics dataset -- the
// interface of GraphHET's include/LDBC/ldbc.hpp:8-46, written for this
// library.
//
//   <dir><name>.properties   "key = value" lines; the last dot-separated part
//                            of the key selects the option (ldbc.hpp:57-171)
//   <dir><name>.v            one vertex name per line
//   <dir><name>.e            "src dst [weight]" per line; an undirected graph
//                            also gets the reverse edge (ldbc.hpp:176-254)
//
// Behaviour kept from the reference: unknown keys are ignored, a line without
// " = " is reported and skipped, a missing weight is 1, a BFS/SSSP source
// that is not a vertex name or a missing .v/.e file ends the program with
// exit status 1 (the reference also waits for a key press first; this one
// does not).  save_to_CSV appends one comma-separated row of result values.
#pragma once

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include <CPU_adj_list/CPU_adj_list.hpp>
#include <CPU_adj_list/parse_string.hpp>

template <typename weight_type>
class LDBC : public graph_structure<weight_type> {
public:
    LDBC() : graph_structure<weight_type>() {}
    LDBC(int n) : graph_structure<weight_type>(n) {}
    LDBC(std::string directory, std::string name) : graph_structure<weight_type>() { base_path = directory + name; }

    bool is_directed = true;
    bool is_weight = false;
    bool is_sssp_weight = true;

    bool sup_bfs = false;
    bool sup_cdlp = false;
    bool sup_pr = false;
    bool sup_wcc = false;
    bool sup_sssp = false;
    std::string bfs_src_name;
    std::string sssp_src_name;
    std::string base_path;
    int bfs_src = 0;
    int cdlp_max_its = 10;
    int pr_its = 10;
    int sssp_src = 0;
    double pr_damping = 0.85;

    void read_config(std::string config_path);
    void load_graph();
    void save_to_CSV(std::vector<std::pair<std::string, std::string>>& res, std::string file_path);

private:
    static bool flag(const std::string& v) { return v != "false"; }
    void set_option(const std::vector<std::string>& key, const std::string& value);
};

template <typename weight_type>
void LDBC<weight_type>::set_option(const std::vector<std::string>& key, const std::string& value) {
    const std::string& last = key.back();
    if (last == "vertex-file") {
        std::cout << "vertex_file: " << value << std::endl;
    } else if (last == "edge-file") {
        std::cout << "edge_file: " << value << std::endl;
    } else if (last == "vertices") {
        std::cout << "V: " << value << std::endl;
    } else if (last == "edges") {
        std::cout << "E: " << value << std::endl;
    } else if (last == "directed") {
        is_directed = flag(value);
        std::cout << "is_directed: " << is_directed << std::endl;
    } else if (last == "names") {
        is_weight = value == "weight";
        std::cout << "is_weight: " << is_weight << std::endl;
    } else if (last == "algorithms") {
        for (const std::string& a : parse_string(value, ", ")) {
            if (a == "bfs") sup_bfs = true;
            else if (a == "cdlp") sup_cdlp = true;
            else if (a == "pr") sup_pr = true;
            else if (a == "sssp") sup_sssp = true;
            else if (a == "wcc") sup_wcc = true;
        }
        std::cout << "bfs: " << sup_bfs << std::endl << "cdlp: " << sup_cdlp << std::endl
                  << "pr: " << sup_pr << std::endl << "sssp: " << sup_sssp << std::endl
                  << "wcc: " << sup_wcc << std::endl;
    } else if (last == "cdlp-max-iterations" || last == "max-iterations") {
        cdlp_max_its = std::stoi(value);
        std::cout << "cdlp_max_its: " << cdlp_max_its << std::endl;
    } else if (last == "pr-damping-factor" || last == "damping-factor") {
        pr_damping = std::stod(value);
        std::cout << "pr_damping: " << pr_damping << std::endl;
    } else if (last == "pr-num-iterations" || last == "num-iterations") {
        pr_its = std::stoi(value);
        std::cout << "pr_its: " << pr_its << std::endl;
    } else if (last == "sssp-weight-property" || last == "weight-property") {
        is_sssp_weight = value == "weight";
        std::cout << "is_sssp_weight: " << is_sssp_weight << std::endl;
    } else if (last == "source-vertex") {
        if (key[key.size() - 2] == "bfs") {
            bfs_src_name = value;
            std::cout << "bfs_source_vertex: " << value << std::endl;
        } else {
            sssp_src_name = value;
            std::cout << "sssp_source_vertex: " << value << std::endl;
        }
    }
}

template <typename weight_type>
void LDBC<weight_type>::read_config(std::string config_path) {
    std::ifstream in(config_path);
    if (!in.is_open()) {
        std::cerr << "Unable to open file: " << config_path << std::endl;
        return;
    }
    std::cout << "Reading config file..." << std::endl;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        const std::vector<std::string> kv = parse_string(line, " = ");
        if (kv.size() != 2) {
            std::cerr << "Invalid line: " << line << std::endl;
            continue;
        }
        const std::vector<std::string> key = parse_string(kv[0], ".");
        if (key.size() >= 2) set_option(key, kv[1]);
    }
    std::cout << "Done." << std::endl;
}

template <typename weight_type>
void LDBC<weight_type>::load_graph() {
    const std::string vpath = base_path + ".v";
    std::cout << "Loading vertices..." << std::endl;
    std::ifstream vin(vpath);
    if (!vin.is_open()) {
        std::cout << "Unable to open file " << vpath << std::endl << "Please check the file location or file name." << std::endl;
        std::exit(1);
    }
    std::string line;
    while (std::getline(vin, line)) this->add_vertice(line);
    vin.close();
    std::cout << "Done." << std::endl;

    auto resolve = [this](bool wanted, const std::string& name, int& id, const char* what) {
        if (!wanted) return;
        auto it = this->vertex_str_to_id.find(name);
        if (it == this->vertex_str_to_id.end()) {
            std::cout << "Invalid source vertex for " << what << std::endl;
            std::exit(1);
        }
        id = it->second;
    };
    resolve(sup_bfs, bfs_src_name, bfs_src, "BFS");
    resolve(sup_sssp, sssp_src_name, sssp_src, "SSSP");

    const std::string epath = base_path + ".e";
    std::cout << "Loading edges..." << std::endl;
    std::ifstream ein(epath);
    if (!ein.is_open()) {
        std::cout << "Unable to open file " << epath << std::endl << "Please check the file location or file name." << std::endl;
        std::exit(1);
    }
    while (std::getline(ein, line)) {
        const std::vector<std::string> f = parse_string(line, " ");
        if (f.size() < 2) {
            std::cerr << "Invalid edge input!" << std::endl;
            continue;
        }
        const weight_type w = f.size() > 2 ? static_cast<weight_type>(std::stod(f[2])) : weight_type(1);
        this->add_edge(f[0], f[1], w);
        if (!is_directed) this->add_edge(f[1], f[0], w);
    }
    std::cout << "Done." << std::endl;
}

template <typename weight_type>
void LDBC<weight_type>::save_to_CSV(std::vector<std::pair<std::string, std::string>>& res, std::string file_path) {
    std::ofstream out(file_path, std::ios::app);
    for (size_t i = 0; i < res.size(); ++i) out << res[i].second << (i + 1 < res.size() ? "," : "");
    out << std::endl;
}
