// Schedule selection: heuristic rules and the learned network's decisions.
#include "sched.hpp"

#include <algorithm>
#include <cmath>

namespace ghet {

// Installed networks (defined in mlp.hip); forward on the host copy of the
// FP16 parameters with FP32 accumulation -- the same arithmetic as the GPU
// inference kernel, used per level because the level loop is host-driven.
bool net_available(int kind);
void net_forward_host(int kind, const float* x, float* logits);

void sched_begin(SchedState* s, int schedule, int kind, int64_t n, int64_t m, int64_t m_in) {
    s->schedule = schedule;
    s->kind = kind;
    s->n = n;
    s->m = m;
    s->m_in = m_in;
    s->use_net = false;
    s->fallback = false;
    if (schedule == GHET_SCHED_LEARNED) {
        if (net_available(kind)) s->use_net = true;
        else s->fallback = true;
    }
}

namespace {

constexpr int kIn = GHET_MLP_IN, kH = GHET_MLP_H, kOut = GHET_MLP_OUT;
constexpr int kB1 = kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kOut * kH;

// Both variants compute every output as bias, then fused multiply-adds over
// the inputs in index order -- the GPU kernel's order -- so their results
// are bit-identical to it; the loops run over outputs innermost so the FMA
// build vectorises.  The bodies are spelled out twice because the FMA one
// must be compiled for the avx2+fma target as a whole.
#pragma GCC push_options
#pragma GCC target("avx2,fma")
void forward_fma(const float* wt, const float* x, float* y) {
    float h1[kH], h2[kH];
    for (int j = 0; j < kH; ++j) h1[j] = wt[kB1 + j];
    for (int k = 0; k < kIn; ++k)
        for (int j = 0; j < kH; ++j) h1[j] = __builtin_fmaf(wt[k * kH + j], x[k], h1[j]);
    for (int j = 0; j < kH; ++j) h1[j] = h1[j] > 0.f ? h1[j] : 0.f;
    for (int j = 0; j < kH; ++j) h2[j] = wt[kB2 + j];
    for (int i = 0; i < kH; ++i)
        for (int j = 0; j < kH; ++j) h2[j] = __builtin_fmaf(wt[kW2 + i * kH + j], h1[i], h2[j]);
    for (int j = 0; j < kH; ++j) h2[j] = h2[j] > 0.f ? h2[j] : 0.f;
    for (int o = 0; o < kOut; ++o) y[o] = wt[kB3 + o];
    for (int j = 0; j < kH; ++j)
        for (int o = 0; o < kOut; ++o) y[o] = __builtin_fmaf(wt[kW3 + j * kOut + o], h2[j], y[o]);
}
#pragma GCC pop_options

void forward_libm(const float* wt, const float* x, float* y) {
    float h1[kH], h2[kH];
    for (int j = 0; j < kH; ++j) h1[j] = wt[kB1 + j];
    for (int k = 0; k < kIn; ++k)
        for (int j = 0; j < kH; ++j) h1[j] = std::fma(wt[k * kH + j], x[k], h1[j]);
    for (int j = 0; j < kH; ++j) h1[j] = h1[j] > 0.f ? h1[j] : 0.f;
    for (int j = 0; j < kH; ++j) h2[j] = wt[kB2 + j];
    for (int i = 0; i < kH; ++i)
        for (int j = 0; j < kH; ++j) h2[j] = std::fma(wt[kW2 + i * kH + j], h1[i], h2[j]);
    for (int j = 0; j < kH; ++j) h2[j] = h2[j] > 0.f ? h2[j] : 0.f;
    for (int o = 0; o < kOut; ++o) y[o] = wt[kB3 + o];
    for (int j = 0; j < kH; ++j)
        for (int o = 0; o < kOut; ++o) y[o] = std::fma(wt[kW3 + j * kOut + o], h2[j], y[o]);
}

}  // namespace

void mlp_forward_host_t(const float* wt, const float* x, float* y) {
    static const bool has_fma = __builtin_cpu_supports("fma") && __builtin_cpu_supports("avx2");
    if (has_fma) forward_fma(wt, x, y);
    else forward_libm(wt, x, y);
}

void mlp_transpose(const float* w, float* wt) {
    for (int j = 0; j < kH; ++j)
        for (int k = 0; k < kIn; ++k) wt[k * kH + j] = w[j * kIn + k];
    for (int j = 0; j < kH; ++j) wt[kB1 + j] = w[kB1 + j];
    for (int j = 0; j < kH; ++j)
        for (int i = 0; i < kH; ++i) wt[kW2 + i * kH + j] = w[kW2 + j * kH + i];
    for (int j = 0; j < kH; ++j) wt[kB2 + j] = w[kB2 + j];
    for (int o = 0; o < kOut; ++o)
        for (int j = 0; j < kH; ++j) wt[kW3 + j * kOut + o] = w[kW3 + o * kH + j];
    for (int o = 0; o < kOut; ++o) wt[kB3 + o] = w[kB3 + o];
}

static inline float lg(double x) { return static_cast<float>(std::log2(1.0 + std::max(0.0, x))); }

void sched_features(const LevelInfo& li, float* f) {
    const double n = std::max<int64_t>(li.n, 1), m = std::max<int64_t>(li.m, 1);
    f[0] = lg(li.frontier) / 32.f;
    f[1] = lg(li.frontier_edges) / 32.f;
    f[2] = lg(li.unvisited) / 32.f;
    f[3] = lg(li.unvisited_in_edges) / 32.f;
    f[4] = static_cast<float>(li.frontier / n);
    f[5] = static_cast<float>(li.frontier_edges / m);
    f[6] = static_cast<float>(li.unvisited / n);
    f[7] = static_cast<float>(li.unvisited_in_edges / std::max<double>(li.m_in, 1));
    f[8] = static_cast<float>(std::min(li.level, 64)) / 64.f;
    f[9] = li.prev_mode == GHET_MODE_PUSH ? 1.f : 0.f;
    f[10] = li.prev_mode == GHET_MODE_PULL ? 1.f : 0.f;
    f[11] = li.prev_mode == GHET_MODE_BITMAP ? 1.f : 0.f;
    f[12] = static_cast<float>(std::tanh(std::log2((1.0 + li.frontier) / (1.0 + li.prev_frontier)) / 4.0));
    f[13] = lg(li.n) / 32.f;
    f[14] = static_cast<float>(std::log2(1.0 + m / n) / 8.0);
    f[15] = li.delta > 0 ? static_cast<float>(std::tanh(li.far_size / n * 8.0)) : 1.f;
}

int heuristic_bfs(const LevelInfo& li) {
    // Beamer-style direction optimisation: go bottom-up once the frontier's
    // out-edges exceed 1/alpha of the in-edges still unexplored, come back
    // top-down when the frontier falls below n/beta and is shrinking.
    constexpr double kAlpha = 14.0, kBeta = 24.0;
    const double n = static_cast<double>(li.n);
    if (!li.have_in) return li.frontier * 64 < li.n ? GHET_MODE_PUSH : GHET_MODE_BITMAP;
    if (li.prev_mode == GHET_MODE_PULL && li.level > 0) {
        if (li.frontier < n / kBeta && li.frontier < li.prev_frontier)
            return li.frontier * 64 < li.n ? GHET_MODE_PUSH : GHET_MODE_BITMAP;
        return GHET_MODE_PULL;
    }
    if (static_cast<double>(li.frontier_edges) > static_cast<double>(li.unvisited_in_edges) / kAlpha)
        return GHET_MODE_PULL;
    return li.frontier * 64 < li.n ? GHET_MODE_PUSH : GHET_MODE_BITMAP;
}

int sched_pick_bfs(SchedState* s, const LevelInfo& li) {
    switch (s->schedule) {
        case GHET_SCHED_PUSH: return GHET_MODE_PUSH;
        case GHET_SCHED_PULL: return li.have_in ? GHET_MODE_PULL : GHET_MODE_PUSH;
        case GHET_SCHED_BITMAP: return GHET_MODE_BITMAP;
        case GHET_SCHED_REFERENCE:
            // bfs_pre's rule (GPU_BFS_pre_csr.cuh:91): push while frontier < n/100.
            return (li.frontier < li.n / 100 || !li.have_in) ? GHET_MODE_PUSH : GHET_MODE_PULL;
        case GHET_SCHED_LEARNED:
            if (s->use_net) {
                float x[GHET_TRACE_FEATURES], y[GHET_MLP_OUT];
                sched_features(li, x);
                net_forward_host(GHET_NET_BFS, x, y);
                int best = 0;
                for (int k = 1; k < 3; ++k)
                    if (y[k] > y[best]) best = k;
                if (best == GHET_MODE_PULL && !li.have_in) best = GHET_MODE_PUSH;
                return best;
            }
            return heuristic_bfs(li);
        default: return heuristic_bfs(li);
    }
}

// The SSSP network sees only scale-free features: the absolute sizes (log
// frontier, frontier edges, unvisited, unvisited in-edges, frontier / n, log
// n) are zeroed, so that a network trained on small graphs transfers to large
// ones; the base width's size term carries the scale.  The trainer applies
// the same mask (graphhet_amd/learn.py, SSSP_MASKED).
void sssp_mask_features(float* f) {
    for (int k : {0, 1, 2, 3, 4, 13}) f[k] = 0.f;
}

int sched_pick_sssp(SchedState* s, const LevelInfo& li) {
    if (s->schedule == GHET_SCHED_LEARNED && s->use_net) {
        if (s->sssp_cls < 0) {
            float x[GHET_TRACE_FEATURES], y[GHET_MLP_OUT];
            sched_features(li, x);
            sssp_mask_features(x);
            net_forward_host(GHET_NET_SSSP, x, y);
            int best = 0;
            for (int k = 1; k < GHET_MLP_OUT; ++k)
                if (y[k] > y[best]) best = k;
            s->sssp_cls = best;
        }
        return s->sssp_cls;
    }
    return 1;  // the base width
}

}  // namespace ghet
