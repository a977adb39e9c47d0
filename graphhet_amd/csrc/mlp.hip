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
int ghet_autotune(ghet_graph*, const char*, int32_t*) { return set_error(GHET_ERR_UNSUPPORTED, "not built yet"); }
int ghet_bfs_multi(ghet_graph*, int32_t, const int32_t*, int32_t*, const ghet_opts*, ghet_run_stats*) {
    return set_error(GHET_ERR_UNSUPPORTED, "not built yet");
}
int ghet_sssp_multi(ghet_graph*, int32_t, const int32_t*, double, double*, const ghet_opts*, ghet_run_stats*) {
    return set_error(GHET_ERR_UNSUPPORTED, "not built yet");
}
int ghet_wcc_many(int32_t, const int32_t*, const int32_t*, const int32_t*, int32_t*, const ghet_opts*,
                  ghet_run_stats*) {
    return set_error(GHET_ERR_UNSUPPORTED, "not built yet");
}
int ghet_tc_many(int32_t, const int32_t*, const int32_t*, const int32_t*, int64_t*, const ghet_opts*,
                 ghet_run_stats*) {
    return set_error(GHET_ERR_UNSUPPORTED, "not built yet");
}
}
