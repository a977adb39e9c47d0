// Per-level schedule for BFS (push / pull / bitmap) and SSSP (bucket width):
// a hand-tuned heuristic, the reference's n/100 rule, and the learned MLP.
// The schedule only chooses how a level is computed; every choice yields the
// same result.
#pragma once

#include <cstdint>

#include "ghet.h"

namespace ghet {

struct LevelInfo {
    int64_t n = 0, m = 0, m_in = 0;
    int32_t level = 0;
    int64_t frontier = 0;           // vertices in the current frontier
    int64_t frontier_edges = 0;     // their out-edges
    int64_t unvisited = 0;          // vertices not yet reached
    int64_t unvisited_in_edges = 0; // in-edges of unreached vertices
    int64_t prev_frontier = 0;
    int32_t prev_mode = GHET_MODE_PUSH;
    bool have_in = true;
    // SSSP only
    double delta = 0, threshold = 0, avg_weight = 0;
    int64_t far_size = 0;
};

struct SchedState {
    int schedule = GHET_SCHED_HEURISTIC;
    int kind = GHET_NET_BFS;
    bool use_net = false;   // learned network installed and requested
    bool fallback = false;  // learned requested but unavailable
    int64_t n = 0, m = 0, m_in = 0;
    int sssp_prev = 1;      // SSSP class picked at the previous decision
};

void sched_begin(SchedState* s, int schedule, int kind, int64_t n, int64_t m, int64_t m_in);
// Picks the BFS mode of the next level.
int sched_pick_bfs(SchedState* s, const LevelInfo& li);
// Picks an SSSP bucket-width multiplier class in [0, 4): delta = base * 2^(class-1).
// The learned schedule picks it from the first phase's features and keeps it.
int sched_pick_sssp(SchedState* s, const LevelInfo& li);
// The 16 features the networks read (also written to traces).
void sched_features(const LevelInfo& li, float* f);
// The heuristic's own choice (used as fallback and as the training baseline).
int heuristic_bfs(const LevelInfo& li);
// Host forward pass of a schedule network over its parameters in transposed
// layout (mlp_transpose of the FP16-rounded parameters); bit-identical to the
// GPU forward kernel.
void mlp_transpose(const float* w, float* wt);
void mlp_forward_host_t(const float* wt, const float* x, float* y);

}  // namespace ghet
