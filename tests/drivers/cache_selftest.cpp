// Interoperability of the binary CSR cache with the vector-of-vectors header:
//   cache_selftest <dir>
// 1. rows written by binary_save_vector_of_vectors (unsorted, a repeated
//    target) load through ghet_host_graph_load canonicalised: ascending, the
//    later entry's weight kept;
// 2. a graph written by ghet_host_graph_save reads back through
//    binary_read_vector_of_vectors<int> / <double> equal to its out-view.
// Prints "cache selftest ok" and exits 0, or names the first mismatch.
#include <cstdio>
#include <string>
#include <vector>

#include "CPU_adj_list/binary_save_read_vector_of_vectors.hpp"
#include "ghet.h"

static int fail(const char* what) {
    std::printf("FAIL: %s (%s)\n", what, ghet_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: cache_selftest <dir>\n");
        return 2;
    }
    const std::string dir = argv[1];

    // 1. header writer -> library loader
    const std::vector<std::vector<int>> tg = {{2, 1}, {}, {0, 3, 0}, {1}};
    const std::vector<std::vector<double>> wt = {{0.5, 0.25}, {}, {1.0, 2.0, 3.0}, {4.0}};
    binary_save_vector_of_vectors(dir + "/vv.csr", tg);
    binary_save_vector_of_vectors(dir + "/vv.csr.w", wt);
    ghet_host_graph* g = nullptr;
    if (ghet_host_graph_load((dir + "/vv.csr").c_str(), 1, &g) != GHET_OK) return fail("load of header-written rows");
    ghet_csr_view v{};
    if (ghet_host_graph_view(g, GHET_VIEW_OUT, &v) != GHET_OK) return fail("view");
    const int want_rp[] = {0, 2, 2, 4, 5};
    const int want_col[] = {1, 2, 0, 3, 1};
    const double want_w[] = {0.25, 0.5, 3.0, 2.0, 4.0};  // row 2: target 0 twice, the later weight (3.0) wins
    if (v.n != 4 || v.m != 5) return fail("header-written rows: sizes");
    for (int i = 0; i <= 4; ++i)
        if (v.rowptr[i] != want_rp[i]) return fail("header-written rows: offsets");
    for (int i = 0; i < 5; ++i)
        if (v.col[i] != want_col[i] || v.w[i] != want_w[i]) return fail("header-written rows: entries");
    ghet_host_graph_free(g);

    // 2. library writer -> header reader
    if (ghet_gen_rmat(10, 8, 0.57, 0.19, 0.19, 3, 2, &g) != GHET_OK) return fail("rmat");
    if (ghet_host_graph_save(g, (dir + "/rmat.csr").c_str()) != GHET_OK) return fail("save");
    if (ghet_host_graph_view(g, GHET_VIEW_OUT, &v) != GHET_OK) return fail("view");
    std::vector<std::vector<int>> rt;
    std::vector<std::vector<double>> rw;
    binary_read_vector_of_vectors(dir + "/rmat.csr", rt);
    binary_read_vector_of_vectors(dir + "/rmat.csr.w", rw);
    if (static_cast<int>(rt.size()) != v.n || static_cast<int>(rw.size()) != v.n) return fail("row counts");
    for (int u = 0; u < v.n; ++u) {
        const int len = v.rowptr[u + 1] - v.rowptr[u];
        if (static_cast<int>(rt[u].size()) != len || static_cast<int>(rw[u].size()) != len) return fail("row lengths");
        for (int k = 0; k < len; ++k)
            if (rt[u][k] != v.col[v.rowptr[u] + k] || rw[u][k] != v.w[v.rowptr[u] + k]) return fail("row entries");
    }
    ghet_host_graph_free(g);
    std::printf("cache selftest ok\n");
    return 0;
}
