// Drop-in for GraphHET's include/GPU_gpma/GPU_gpma.hpp: the GPMA handle and
// its build / update entry points (init_csr_gpma :928-941, init_gpma_from_csr
// :882-926, update_gpma :756-824), backed by libghet's dynamic graph store
// (ghet_dyn_*, include/ghet.h).  Plain C++: no CUDA, thrust or HIP headers.
//
// Differences a caller sees:
//  * the reference's thrust device vectors (DEV_VEC_KEY, DEV_VEC_VALUE,
//    DEV_VEC_SIZE) are host std::vectors here; callers fill them on the host,
//    as GPU_gpma_example.cu does, and the store itself lives on the GPU;
//  * one GPMA holds a whole directed edge set (keys (src << 32) | dst); the
//    analytics read its out-rows, and the in-rows are kept alongside;
//  * a key repeated within one batch: the last occurrence wins (GPMA leaves
//    this undefined); value VALUE_NONE (0) deletes, as in GPMA.
#pragma once

#include <GPU_csr/algorithm/ghet_shim.hpp>

const KEY_TYPE KEY_NONE = 0xFFFFFFFFFFFFFFFFULL;
const KEY_TYPE KEY_MAX = 0xFFFFFFFFFFFFFFFEULL;
const KEY_TYPE KEY_HIGH = 0xFFFFFFFF00000000ULL;
const VALUE_TYPE VALUE_NONE = 0;

typedef std::vector<KEY_TYPE> DEV_VEC_KEY;
typedef std::vector<VALUE_TYPE> DEV_VEC_VALUE;
typedef std::vector<SIZE_TYPE> DEV_VEC_SIZE;

class GPMA {
public:
    SIZE_TYPE row_num = 0;
    DEV_VEC_SIZE row_offset;  // kept for callers that set it; the store tracks its own rows
    ghet_dyn* handle = nullptr;

    GPMA() = default;
    GPMA(const GPMA&) = delete;
    GPMA& operator=(const GPMA&) = delete;
    GPMA(GPMA&& o) noexcept : row_num(o.row_num), row_offset(std::move(o.row_offset)), handle(o.handle) { o.handle = nullptr; }
    ~GPMA() {
        if (handle) ghet_dyn_destroy(handle);
    }
    // live edges (the reference's keys.size() counts slots, including gaps)
    int get_size() const { return handle ? static_cast<int>(ghet_dyn_edges(handle)) : 0; }
    int get_V() const { return static_cast<int>(row_num) + 1; }
};

namespace ghet_gpma {

// Splits (src << 32) | dst keys and applies them as one batch.
inline bool apply(GPMA& gpma, const DEV_VEC_KEY& keys, const DEV_VEC_VALUE& values, const char* what) {
    if (!gpma.handle) {
        std::fprintf(stderr, "%s: GPMA not initialised (call init_csr_gpma)\n", what);
        return false;
    }
    const std::size_t n = std::min(keys.size(), values.size());
    std::vector<int32_t> src(n), dst(n);
    std::vector<double> w(n);
    for (std::size_t i = 0; i < n; ++i) {
        src[i] = static_cast<int32_t>(keys[i] >> 32);
        dst[i] = static_cast<int32_t>(keys[i] & 0xFFFFFFFFULL);
        w[i] = values[i];
    }
    return ghet_shim::ok(ghet_dyn_update(gpma.handle, static_cast<int64_t>(n), src.data(), dst.data(), w.data(), nullptr),
                         what);
}

// A device graph of the current edges for one analytics call.
struct Snapshot {
    ghet_graph* g = nullptr;
    explicit Snapshot(GPMA& gpma, uint32_t flags = 0) {
        if (gpma.handle) ghet_shim::ok(ghet_dyn_snapshot(gpma.handle, flags, &g), "GPMA snapshot");
    }
    ~Snapshot() {
        if (g) ghet_graph_destroy(g);
    }
    Snapshot(const Snapshot&) = delete;
    Snapshot& operator=(const Snapshot&) = delete;
};

}  // namespace ghet_gpma

// An empty store of row_num vertices (init_csr_gpma, GPU_gpma.hpp:928-941).
inline void init_csr_gpma(GPMA& gpma, SIZE_TYPE row_num) {
    if (gpma.handle) ghet_dyn_destroy(gpma.handle);
    gpma.handle = nullptr;
    gpma.row_num = row_num;
    gpma.row_offset.assign(static_cast<std::size_t>(row_num) + 1, 0);
    std::vector<int32_t> rp(static_cast<std::size_t>(row_num) + 1, 0);
    ghet_shim::ok(ghet_dyn_create(static_cast<int32_t>(row_num), rp.data(), nullptr, nullptr, 0, &gpma.handle),
                  "init_csr_gpma");
}

// Loads the initial edges (init_gpma_from_csr, GPU_gpma.hpp:882-926); returns
// the number of keys given, as the reference does.
inline int init_gpma_from_csr(GPMA& gpma, DEV_VEC_KEY& init_keys, DEV_VEC_VALUE& init_values) {
    ghet_gpma::apply(gpma, init_keys, init_values, "init_gpma_from_csr");
    return static_cast<int>(init_keys.size());
}

// One batch of updates (update_gpma, GPU_gpma.hpp:756-824): value 0 deletes,
// a present key takes the new value, any other key is inserted.
inline void update_gpma(GPMA& gpma, DEV_VEC_KEY& update_keys, DEV_VEC_VALUE& update_values) {
    ghet_gpma::apply(gpma, update_keys, update_values, "update_gpma");
}
