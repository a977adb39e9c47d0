// Schedule network placeholder: no network installed yet.
#include "graph_internal.hpp"
#include "sched.hpp"

namespace ghet {
bool net_available(int) { return false; }
void net_forward_host(int, const float*, float* y) {
    for (int k = 0; k < GHET_MLP_OUT; ++k) y[k] = 0.f;
}
}  // namespace ghet

using namespace ghet;

extern "C" {
int ghet_sched_load(const char*) { return GHET_WARN_FALLBACK; }
int ghet_sched_save(const char*) { return set_error(GHET_ERR_UNSUPPORTED, "no network installed"); }
int ghet_sched_clear(void) { return GHET_OK; }
int ghet_sched_set_params(int, const float*) { return set_error(GHET_ERR_UNSUPPORTED, "not built yet"); }
int ghet_sched_get_params(int, float*) { return set_error(GHET_ERR_UNSUPPORTED, "not built yet"); }
int ghet_mlp_forward(int, int32_t, const float*, float*) { return set_error(GHET_ERR_UNSUPPORTED, "not built yet"); }
int ghet_mlp_train_step(int, int32_t, const float*, const int32_t*, float, float*, float*) {
    return set_error(GHET_ERR_UNSUPPORTED, "not built yet");
}
int ghet_bfs_multi(ghet_graph* g, int32_t k, const int32_t* srcs, int32_t* depth_out, const ghet_opts* o,
                   ghet_run_stats* st) {
    clear_error();
    if (!g) return set_error(GHET_ERR_INVALID, "graph handle is NULL");
    if (st) std::memset(st, 0, sizeof(*st));
    if (hipSetDevice(g->device) != hipSuccess) return set_error(GHET_ERR_HIP, "hipSetDevice failed");
    return run_bfs_multi(g, k, srcs, depth_out, o, st);
}
int ghet_sssp_multi(ghet_graph* g, int32_t k, const int32_t* srcs, double inf, double* dist_out, const ghet_opts* o,
                    ghet_run_stats* st) {
    clear_error();
    if (!g) return set_error(GHET_ERR_INVALID, "graph handle is NULL");
    if (st) std::memset(st, 0, sizeof(*st));
    if (hipSetDevice(g->device) != hipSuccess) return set_error(GHET_ERR_HIP, "hipSetDevice failed");
    return run_sssp_multi(g, k, srcs, inf, dist_out, o, st);
}
}
