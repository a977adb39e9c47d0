// Shared plumbing of the drop-in algorithm headers: the typedefs and grid
// macros the reference's CSR algorithms borrow from GPU_gpma.hpp:17-40
// (values unchanged, they are part of the interface), and helpers that run a
// library call on a CSR_graph handle or on borrowed device arrays.
#pragma once

#include <climits>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include <CPU_adj_list/CPU_adj_list.hpp>
#include <GPU_csr/GPU_csr.hpp>
#include <ghet.h>

#ifndef GHET_GPMA_TYPES
#define GHET_GPMA_TYPES
typedef unsigned long long KEY_TYPE;
typedef double VALUE_TYPE;
typedef unsigned int SIZE_TYPE;
const SIZE_TYPE SIZE_NONE = 0xFFFFFFFF;
const KEY_TYPE COL_IDX_NONE = 0xFFFFFFFF;
const SIZE_TYPE MAX_BLOCKS_NUM = 96 * 8;
#ifndef CALC_BLOCKS_NUM
#define CALC_BLOCKS_NUM(ITEMS_PER_BLOCK, CALC_SIZE) \
    (((CALC_SIZE) - 1) / (ITEMS_PER_BLOCK) + 1 < MAX_BLOCKS_NUM ? ((CALC_SIZE) - 1) / (ITEMS_PER_BLOCK) + 1 : MAX_BLOCKS_NUM)
#endif
#endif

namespace ghet_shim {

inline bool ok(int st, const char* what) {
    if (st == GHET_OK || st == GHET_WARN_FALLBACK) return true;
    std::fprintf(stderr, "%s failed: %s\n", what, ghet_last_error());
    return false;
}

// Borrows device CSR arrays for one call (the raw cores take device pointers).
struct Borrowed {
    ghet_graph* g = nullptr;
    Borrowed(int n, long long m, const int* rowptr, const int* col, const double* w, const int* in_rowptr = nullptr,
             const int* in_col = nullptr, long long m_in = 0) {
        ok(ghet_graph_wrap_device(n, m, rowptr, col, w, m_in, in_rowptr, in_col, 0, nullptr, nullptr,
                                  in_rowptr ? 0u : GHET_GRAPH_NO_IN, &g),
           "ghet_graph_wrap_device");
    }
    ~Borrowed() {
        if (g) ghet_graph_destroy(g);
    }
    Borrowed(const Borrowed&) = delete;
    Borrowed& operator=(const Borrowed&) = delete;
};

}  // namespace ghet_shim
