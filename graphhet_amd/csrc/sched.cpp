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

// log2(1 + x) / log2(1 + total): where x sits between 1 and the graph's own
// size, the same fraction at every scale.
static inline float rel_lg(double x, double total) {
    const double d = std::log2(1.0 + std::max(1.0, total));
    return static_cast<float>(std::log2(1.0 + std::max(0.0, x)) / d);
}

int heuristic_bfs(const LevelInfo& li);

// Sizes enter as fractions of n or m, or as logs relative to log n or log
// m, so that a network trained on small graphs sees the same ranges on large
// ones; slot 15 alone carries the scale.  The scale is needed: a bottom-up
// sweep of the tail costs about 2.3 ps per vertex on MI355X and a top-down
// tail level about 13 us, so which is cheaper flips near 2^22.4 vertices.
// Slot 15 is tanh((log2 n - 22.4) / 2): centred on that crossover and
// saturating, so graphs beyond the training scales fall on the side of the
// largest ones instead of extrapolating a linear input.  For BFS, slot 13
// carries the heuristic rule's own pick at this level (push 0, bitmap 0.5,
// pull 1).
void sched_features(const LevelInfo& li, float* f) {
    const double n = std::max<int64_t>(li.n, 1), m = std::max<int64_t>(li.m, 1);
    f[0] = rel_lg(li.frontier, n);
    f[1] = rel_lg(li.frontier_edges, m);
    f[2] = rel_lg(li.unvisited, n);
    f[3] = rel_lg(li.unvisited_in_edges, std::max<double>(li.m_in, 1));
    f[4] = static_cast<float>(li.frontier / n);
    f[5] = static_cast<float>(li.frontier_edges / m);
    f[6] = static_cast<float>(li.unvisited / n);
    f[7] = static_cast<float>(li.unvisited_in_edges / std::max<double>(li.m_in, 1));
    f[8] = static_cast<float>(std::min(li.level, 64)) / 64.f;
    f[9] = li.prev_mode == GHET_MODE_PUSH ? 1.f : 0.f;
    f[10] = li.prev_mode == GHET_MODE_PULL ? 1.f : 0.f;
    f[11] = li.prev_mode == GHET_MODE_BITMAP ? 1.f : 0.f;
    f[12] = static_cast<float>(std::tanh(std::log2((1.0 + li.frontier) / (1.0 + li.prev_frontier)) / 4.0));
    f[14] = static_cast<float>(std::log2(1.0 + m / n) / 8.0);
    if (li.delta > 0) {
        // SSSP decision: frontier = the bucket just finished (vertices and
        // relaxed edges), unvisited = vertices not yet final, far_size = the
        // far list; prev_mode = the previous width class
        f[3] = rel_lg(li.far_size, n);
        f[7] = static_cast<float>(li.far_size / n);
        f[9] = li.prev_mode == 0 ? 1.f : 0.f;
        f[10] = li.prev_mode == 1 ? 1.f : 0.f;
        f[11] = li.prev_mode == 2 ? 1.f : 0.f;
        f[12] = li.prev_mode == 3 ? 1.f : 0.f;
        f[13] = static_cast<float>(std::log2(1.0 + li.threshold / li.delta) / 16.0);
    } else {
        const int h = heuristic_bfs(li);
        f[13] = h == GHET_MODE_PULL ? 1.f : (h == GHET_MODE_BITMAP ? 0.5f : 0.f);
    }
    f[15] = static_cast<float>(std::tanh((std::log2(n) - 22.4) / 2.0));
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

// The SSSP network picks the width class of every bucket from the state at
// the threshold advance (all features scale-free).
int sched_pick_sssp(SchedState* s, const LevelInfo& li) {
    if (s->schedule == GHET_SCHED_LEARNED && s->use_net) {
        LevelInfo x = li;
        x.prev_mode = s->sssp_prev;
        float f[GHET_TRACE_FEATURES], y[GHET_MLP_OUT];
        sched_features(x, f);
        net_forward_host(GHET_NET_SSSP, f, y);
        int best = 0;
        for (int k = 1; k < GHET_MLP_OUT; ++k)
            if (y[k] > y[best]) best = k;
        s->sssp_prev = best;
        return best;
    }
    return 1;  // the base width
}

}  // namespace ghet
