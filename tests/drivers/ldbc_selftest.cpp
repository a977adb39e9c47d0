// Host-only check of the LDBC loader and checkers (no GPU): loads
// <dir><name>.properties/.v/.e, prints the parsed configuration, then for
// every baseline <name>-<ALGO> present builds a result from the baseline's
// own values -- which the checker must accept -- and a copy with one vertex
// changed -- which it must reject.  WCC results use a different but
// equivalent labelling (each component named by its largest member).
//
//   ldbc_selftest <dir> <name>   prints "config ..." and "check ALGO accept=A reject=R"
#include <LDBC/checker.hpp>
#include <LDBC/ldbc.hpp>

#include <cfloat>
#include <climits>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::map<int, std::string> read_baseline(LDBC<double>& g, const std::string& path) {
    std::map<int, std::string> out;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> t = parse_string(line, " ");
        if (t.size() == 2) out[g.vertex_str_to_id[t[0]]] = t[1];
    }
    return out;
}

bool exists(const std::string& path) { return std::ifstream(path).good(); }

void report(const char* algo, bool accept, bool reject) {
    std::cout << "check " << algo << " accept=" << accept << " reject=" << reject << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: ldbc_selftest <dir> <name>\n";
        return 2;
    }
    std::string dir = argv[1];
    if (!dir.empty() && dir.back() != '/') dir += '/';
    const std::string name = argv[2];
    LDBC<double> g(dir, name);
    g.read_config(dir + name + ".properties");
    g.load_graph();
    std::cout << "config directed=" << g.is_directed << " weight=" << g.is_weight << " bfs=" << g.sup_bfs
              << " cdlp=" << g.sup_cdlp << " pr=" << g.sup_pr << " sssp=" << g.sup_sssp << " wcc=" << g.sup_wcc
              << " bfs_src=" << g.bfs_src << " sssp_src=" << g.sssp_src << " cdlp_its=" << g.cdlp_max_its
              << " pr_its=" << g.pr_its << " damping=" << g.pr_damping << " V=" << g.V
              << " E=" << g.edge_number() << std::endl;
    const std::string base = g.base_path;

    if (exists(base + "-BFS")) {
        auto b = read_baseline(g, base + "-BFS");
        std::vector<int> d(g.V);
        for (auto& kv : b) d[kv.first] = kv.second == "9223372036854775807" ? INT_MAX : std::stoi(kv.second);
        auto res = g.res_trans_id_val(d);
        const bool a = Bfs_checker(g, res, base + "-BFS");
        d[g.bfs_src] += 1;
        auto bad = g.res_trans_id_val(d);
        report("BFS", a, !Bfs_checker(g, bad, base + "-BFS"));
    }
    if (exists(base + "-SSSP")) {
        auto b = read_baseline(g, base + "-SSSP");
        std::vector<double> d(g.V);
        int reached = -1;
        for (auto& kv : b) {
            d[kv.first] = (kv.second == "infinity" || kv.second == "inf") ? DBL_MAX : std::stod(kv.second);
            if (d[kv.first] > 0 && d[kv.first] != DBL_MAX) reached = kv.first;
        }
        auto res = g.res_trans_id_val(d);
        const bool a = SSSP_checker(g, res, base + "-SSSP");
        if (reached >= 0) d[reached] *= 1.001;  // beyond the 1e-4 tolerance
        auto bad = g.res_trans_id_val(d);
        report("SSSP", a, reached >= 0 && !SSSP_checker(g, bad, base + "-SSSP"));
    }
    if (exists(base + "-PR")) {
        auto b = read_baseline(g, base + "-PR");
        std::vector<double> r(g.V);
        for (auto& kv : b) r[kv.first] = std::stod(kv.second);
        auto res = g.res_trans_id_val(r);
        const bool a = PR_checker(g, res, base + "-PR");
        r[0] *= 1.001;
        auto bad = g.res_trans_id_val(r);
        report("PR", a, !PR_checker(g, bad, base + "-PR"));
    }
    if (exists(base + "-CDLP")) {
        auto b = read_baseline(g, base + "-CDLP");
        std::vector<int> lab(g.V);
        for (auto& kv : b) lab[kv.first] = g.vertex_str_to_id[kv.second];
        auto res = g.res_trans_id_id(lab);
        const bool a = CDLP_checker(g, res, base + "-CDLP");
        lab[0] = lab[0] == 0 ? 1 : 0;
        auto bad = g.res_trans_id_id(lab);
        report("CDLP", a, !CDLP_checker(g, bad, base + "-CDLP"));
    }
    if (exists(base + "-WCC")) {
        auto b = read_baseline(g, base + "-WCC");
        std::map<std::string, int> largest;  // baseline label -> largest member id
        for (auto& kv : b) {
            auto it = largest.find(kv.second);
            if (it == largest.end() || kv.first > it->second) largest[kv.second] = kv.first;
        }
        std::vector<int> lab(g.V);
        for (auto& kv : b) lab[kv.first] = largest[kv.second];
        auto res = g.res_trans_id_id(lab);
        const bool a = WCC_checker(g, res, base + "-WCC");
        // move vertex 0 into a component it does not belong to (if there is one)
        int other = -1;
        for (auto& kv : b)
            if (lab[kv.first] != lab[0]) {
                other = lab[kv.first];
                break;
            }
        bool r = false;
        if (other >= 0) {
            lab[0] = other;
            auto bad = g.res_trans_id_id(lab);
            r = !WCC_checker(g, bad, base + "-WCC");
        }
        report("WCC", a, r);
    }
    return 0;
}
