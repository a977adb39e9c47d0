// ORACLE -- test infrastructure only.  Nothing in the product links, loads or
// calls this file: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker and as the timed CPU baseline.
//
// A CPU restatement of GraphHET's CPU adjacency-list algorithms
// (include/CPU_adj_list/algorithm/*.hpp), operating on CSR arrays whose rows
// are the reference's sorted OUTs/INs lists (toCSR order, GPU_csr.hpp:282-339).
// Each function keeps the reference's operation order where floating point
// depends on it, and its threading where the CPU baseline is timed.
//
// Pinning (see tests/test_oracle.py):
//   bfs   golden vectors from oracle/_ref; networkx shortest-path lengths;
//         closed forms (path, star, binary tree, grid).
//   bfs_parent, sssp_parent
//         validity (pre[v] -> v is an out-edge one level up, or with
//         d[pre] + w == d[v] in FP64) and minimality, both computed from the
//         out-CSR by a scatter-min over edges; planted mistakes (largest
//         parent, in/out lists swapped) must fail.
//   sssp  golden vectors; scipy.sparse.csgraph.dijkstra; edge invariants.
//   wcc   golden vectors; scipy connected_components(weak) + min-id closed form.
//   pr    golden vectors (bit-exact, same operation order); directed cycle
//         closed form 1/n; mass conservation.
//   cdlp  golden vectors; brute force on tiny graphs; K_n and isolated cases.
//   tc    golden vectors; networkx.triangles x 6; K_n closed form.
//   tc_forward
//         networkx.triangles; K_n closed form; equality with oracle_tc (6 t)
//         on R-MAT and Erdos-Renyi graphs.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <functional>
#include <limits>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// Runs fn(q) for q in [0, chunks) on `threads` std::threads (the reference's
// ThreadPool hands each chunk to its own pool thread).
void run_chunks(int chunks, int threads, const std::function<void(int)>& fn) {
    threads = std::max(1, std::min(threads, chunks));
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (int q; (q = next.fetch_add(1)) < chunks;) fn(q);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

// CPU_BFS.hpp:9-41 (min_depth = 0, max_depth = INT_MAX): FIFO queue, depth
// relaxed when depth[adj] > depth[v] + 1.
void oracle_bfs(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t src, int32_t* depth) {
    for (int32_t i = 0; i < n; ++i) depth[i] = INT_MAX;
    if (src < 0 || src >= n) return;
    depth[src] = 0;
    std::queue<int32_t> q;
    q.push(src);
    while (!q.empty()) {
        const int32_t v = q.front();
        q.pop();
        if (depth[v] == INT_MAX) continue;
        const int32_t nd = depth[v] + 1;
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            const int32_t a = col[e];
            if (depth[a] > nd) {
                depth[a] = nd;
                q.push(a);
            }
        }
    }
}

// The device path's parent rule, stated independently: smallest in-neighbour
// u (ascending in-list) with depth[u] == depth[v] - 1; -1 for src/unreached.
void oracle_bfs_parent(int32_t n, const int32_t* in_rowptr, const int32_t* in_col, const int32_t* depth,
                       int32_t* pre) {
    for (int32_t v = 0; v < n; ++v) {
        pre[v] = -1;
        if (depth[v] == INT_MAX || depth[v] == 0) continue;
        for (int32_t e = in_rowptr[v]; e < in_rowptr[v + 1]; ++e)
            if (depth[in_col[e]] == depth[v] - 1) {
                pre[v] = in_col[e];
                break;
            }
    }
}

// CPU_shortest_paths.hpp:19-64: binary-heap Dijkstra, inf = DBL_MAX, skip
// vertices already settled, relax when distances[v] > distances[u] + w.
void oracle_sssp(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w, int32_t src, double* dist) {
    const double inf = std::numeric_limits<double>::max();
    for (int32_t i = 0; i < n; ++i) dist[i] = inf;
    if (src < 0 || src >= n) return;
    dist[src] = 0;
    std::vector<char> vis(static_cast<size_t>(n), 0);
    using Item = std::pair<double, int32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    pq.push({0.0, src});
    while (!pq.empty()) {
        const int32_t u = pq.top().second;
        pq.pop();
        if (vis[u]) continue;
        vis[u] = 1;
        for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            const int32_t v = col[e];
            if (dist[v] > dist[u] + w[e]) {
                dist[v] = dist[u] + w[e];
                pq.push({dist[v], v});
            }
        }
    }
}

// Parent rule of the device path: smallest in-neighbour u != v with
// dist[u] + w(u,v) == dist[v]; -1 for src/unreached.
void oracle_sssp_parent(int32_t n, const int32_t* in_rowptr, const int32_t* in_col, const double* in_w,
                        const double* dist, double inf, int32_t src, int32_t* pre) {
    for (int32_t v = 0; v < n; ++v) {
        pre[v] = -1;
        if (dist[v] == inf || v == src) continue;
        for (int32_t e = in_rowptr[v]; e < in_rowptr[v + 1]; ++e) {
            const int32_t u = in_col[e];
            if (u != v && dist[u] != inf && dist[u] + in_w[e] == dist[v]) {
                pre[v] = u;
                break;
            }
        }
    }
}

// CPU_connected_components.hpp:11-66 over the out-lists: repeat {hook the
// higher component id under the lower one when it is still a root; pointer
// jumping} until nothing changes, `threads` workers over contiguous vertex
// ranges (relaxed atomics replace the reference's plain racy accesses).
void oracle_wcc(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t threads, int32_t* comp_out) {
    std::vector<std::atomic<int32_t>> comp(static_cast<size_t>(n));
    for (int32_t u = 0; u < n; ++u) comp[u].store(u, std::memory_order_relaxed);
    threads = std::max(1, threads);
    std::atomic<bool> change(true);
    while (change.load()) {
        change.store(false);
        run_chunks(threads, threads, [&](int q) {
            const int32_t start = static_cast<int32_t>(static_cast<int64_t>(q) * n / threads);
            const int32_t end = std::min<int64_t>(n - 1, static_cast<int64_t>(q + 1) * n / threads);
            for (int32_t u = start; u <= end; ++u)
                for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                    const int32_t v = col[e];
                    const int32_t cu = comp[u].load(std::memory_order_relaxed);
                    const int32_t cv = comp[v].load(std::memory_order_relaxed);
                    if (cu == cv) continue;
                    const int32_t hi = std::max(cu, cv), lo = std::min(cu, cv);
                    if (comp[hi].load(std::memory_order_relaxed) == hi) {
                        change.store(true, std::memory_order_relaxed);
                        comp[hi].store(lo, std::memory_order_relaxed);
                    }
                }
        });
        run_chunks(threads, threads, [&](int q) {
            const int32_t start = static_cast<int32_t>(static_cast<int64_t>(q) * n / threads);
            const int32_t end = std::min<int64_t>(n - 1, static_cast<int64_t>(q + 1) * n / threads);
            for (int32_t u = start; u <= end; ++u)
                for (;;) {
                    const int32_t c = comp[u].load(std::memory_order_relaxed);
                    const int32_t cc = comp[c].load(std::memory_order_relaxed);
                    if (c == cc) break;
                    comp[u].store(cc, std::memory_order_relaxed);
                }
        });
    }
    for (int32_t u = 0; u < n; ++u) comp_out[u] = comp[u].load(std::memory_order_relaxed);
}

// CPU_PageRank.hpp:12-56: rank0 = 1/N; per iteration sink_sum over sinks in
// ascending id order, common_add = sink_sum*damp/N + (1-damp)/N, and for every
// v, sum over its in-list (in order) of rank[u] / outdeg(u), new rank =
// common_add + damp*sum.  100 chunks on `threads` threads (the reference
// builds a 100-thread pool every iteration).
void oracle_pr(int32_t n, const int32_t* in_rowptr, const int32_t* in_col, const int32_t* out_rowptr, double damp,
               int32_t iters, int32_t threads, double* rank_out) {
    std::vector<double> rank(static_cast<size_t>(n), 1 / static_cast<double>(n)), next(static_cast<size_t>(n));
    const double teleport = (1 - damp) / n;
    std::vector<int32_t> sink;
    for (int32_t i = 0; i < n; ++i)
        if (out_rowptr[i + 1] == out_rowptr[i]) sink.push_back(i);
    for (int32_t it = 0; it < iters; ++it) {
        double sink_sum = 0;
        for (int32_t s : sink) sink_sum += rank[s];
        const double common_add = (sink_sum * damp / n) + teleport;
        run_chunks(100, threads, [&](int q) {
            const int64_t start = static_cast<int64_t>(q) * n / 100;
            const int64_t end = std::min<int64_t>(static_cast<int64_t>(n) - 1, static_cast<int64_t>(q + 1) * n / 100);
            for (int64_t v = start; v <= end; ++v) {
                double sum = 0;
                for (int32_t e = in_rowptr[v]; e < in_rowptr[v + 1]; ++e) {
                    const int32_t u = in_col[e];
                    sum += rank[u] / static_cast<double>(out_rowptr[u + 1] - out_rowptr[u]);
                }
                next[v] = common_add + damp * sum;
            }
        });
        rank.swap(next);
    }
    for (int32_t i = 0; i < n; ++i) rank_out[i] = rank[i];
}

// One iteration of oracle_pr (same operations, same order): next from rank.
// Lets a checker compare every iteration of a run.
void oracle_pr_step(int32_t n, const int32_t* in_rowptr, const int32_t* in_col, const int32_t* out_rowptr,
                    double damp, int32_t threads, const double* rank, double* next) {
    const double teleport = (1 - damp) / n;
    double sink_sum = 0;
    for (int32_t i = 0; i < n; ++i)
        if (out_rowptr[i + 1] == out_rowptr[i]) sink_sum += rank[i];
    const double common_add = (sink_sum * damp / n) + teleport;
    run_chunks(100, threads, [&](int q) {
        const int64_t start = static_cast<int64_t>(q) * n / 100;
        const int64_t end = std::min<int64_t>(static_cast<int64_t>(n) - 1, static_cast<int64_t>(q + 1) * n / 100);
        for (int64_t v = start; v <= end; ++v) {
            double sum = 0;
            for (int32_t e = in_rowptr[v]; e < in_rowptr[v + 1]; ++e) {
                const int32_t u = in_col[e];
                sum += rank[u] / static_cast<double>(out_rowptr[u + 1] - out_rowptr[u]);
            }
            next[v] = common_add + damp * sum;
        }
    });
}

// CPU_Community_Detection.hpp:12-64: synchronous; counts labels of in- then
// out-neighbours with multiplicity in a hash map; the result is the label with
// the largest count, ties to the smaller label, and a vertex without
// neighbours keeps its label.  100 chunks on `threads` threads.
void oracle_cdlp(int32_t n, const int32_t* in_rowptr, const int32_t* in_col, const int32_t* out_rowptr,
                 const int32_t* out_col, int32_t iters, int32_t threads, int32_t* label_out) {
    std::vector<int32_t> label(static_cast<size_t>(n)), next(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) label[i] = i;
    for (int32_t it = 0; it < iters; ++it) {
        run_chunks(100, threads, [&](int q) {
            const int64_t start = static_cast<int64_t>(q) * n / 100;
            const int64_t end = std::min<int64_t>(static_cast<int64_t>(n) - 1, static_cast<int64_t>(q + 1) * n / 100);
            for (int64_t i = start; i <= end; ++i) {
                std::unordered_map<int32_t, int32_t> count;
                for (int32_t e = in_rowptr[i]; e < in_rowptr[i + 1]; ++e) count[label[in_col[e]]]++;
                for (int32_t e = out_rowptr[i]; e < out_rowptr[i + 1]; ++e) count[label[out_col[e]]]++;
                int32_t maxcount = 0, maxlabel = label[i];
                for (const auto& p : count) {
                    if (p.second > maxcount) {
                        maxcount = p.second;
                        maxlabel = p.first;
                    } else if (p.second == maxcount) {
                        maxlabel = std::min(p.first, maxlabel);
                    }
                }
                next[i] = maxlabel;
            }
        });
        label.swap(next);
    }
    for (int32_t i = 0; i < n; ++i) label_out[i] = label[i];
}

// CPU_TC.hpp:31-60: for u, for v in N(u), two-pointer merge of N(u) and N(v);
// each common w adds 1 to u, v and w.  int64 counters (the reference's int
// wraps at large hubs).  `v_begin..v_end` restricts the outer loop (used to
// time a vertex sample); pass 0, n for the whole graph.
void oracle_tc(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t v_begin, int32_t v_end, int64_t* cnt) {
    for (int32_t i = 0; i < n; ++i) cnt[i] = 0;
    for (int32_t u = v_begin; u < v_end; ++u) {
        const int32_t ub = rowptr[u], ue = rowptr[u + 1];
        for (int32_t e = ub; e < ue; ++e) {
            const int32_t v = col[e];
            int32_t i = ub, j = rowptr[v];
            const int32_t je = rowptr[v + 1];
            while (i < ue && j < je) {
                const int32_t a = col[i], b = col[j];
                if (a == b) {
                    ++cnt[u];
                    ++cnt[v];
                    ++cnt[a];
                    ++i;
                    ++j;
                } else if (a < b) {
                    ++i;
                } else {
                    ++j;
                }
            }
        }
    }
}

// oracle_tc's outer loop over a list of vertices instead of a range (a
// seeded vertex sample for timing the CPU baseline); returns the number of
// merge steps taken, the unit the sample's rate is measured in.
int64_t oracle_tc_vertices(int32_t n, const int32_t* rowptr, const int32_t* col, const int32_t* verts, int32_t k,
                           int64_t* cnt) {
    for (int32_t i = 0; i < n; ++i) cnt[i] = 0;
    int64_t steps = 0;
    for (int32_t x = 0; x < k; ++x) {
        const int32_t u = verts[x];
        const int32_t ub = rowptr[u], ue = rowptr[u + 1];
        for (int32_t e = ub; e < ue; ++e) {
            const int32_t v = col[e];
            int32_t i = ub, j = rowptr[v];
            const int32_t je = rowptr[v + 1];
            while (i < ue && j < je) {
                ++steps;
                const int32_t a = col[i], b = col[j];
                if (a == b) {
                    ++cnt[u];
                    ++cnt[v];
                    ++cnt[a];
                    ++i;
                    ++j;
                } else if (a < b) {
                    ++i;
                } else {
                    ++j;
                }
            }
        }
    }
    return steps;
}

// An independent triangle counter for checking large symmetric loop-free
// graphs: the forward algorithm over a degree-ordered orientation (Schank and
// Wagner, "Finding, counting and listing all triangles in large graphs",
// WEA 2005; the orientation helper of CPU_TC.hpp:9-29 is the same idea).
// Vertices are ranked by (degree, id); u -> v when rank(u) < rank(v).  Every
// triangle a < b < c (by rank) is found once, at u = a, v = b, w = c, by
// probing N+(v) against a bitmap of N+(u), and adds 1 to t[a], t[b], t[c].
// tri[v] receives t(v), the number of triangles through v (6 * t(v) is the
// reference's convention, CPU_TC.hpp:31-60).  `threads` workers, each with a
// private bitmap and count array.
void oracle_tc_forward(int32_t n, const int32_t* rowptr, const int32_t* col, int32_t threads, int64_t* tri) {
    threads = std::max(1, threads);
    auto before = [&](int32_t a, int32_t b) {
        const int32_t da = rowptr[a + 1] - rowptr[a], db = rowptr[b + 1] - rowptr[b];
        return da < db || (da == db && a < b);
    };
    std::vector<int64_t> off(static_cast<size_t>(n) + 1, 0);
    for (int32_t u = 0; u < n; ++u) {
        int64_t k = 0;
        for (int32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) k += before(u, col[e]) ? 1 : 0;
        off[u + 1] = off[u] + k;
    }
    std::vector<int32_t> fwd(static_cast<size_t>(off[n]));
    run_chunks(threads, threads, [&](int q) {
        const int32_t b = static_cast<int32_t>(static_cast<int64_t>(q) * n / threads);
        const int32_t e = static_cast<int32_t>(static_cast<int64_t>(q + 1) * n / threads);
        for (int32_t u = b; u < e; ++u) {
            int64_t o = off[u];
            for (int32_t x = rowptr[u]; x < rowptr[u + 1]; ++x)
                if (before(u, col[x])) fwd[o++] = col[x];
        }
    });
    std::vector<std::vector<int64_t>> part(static_cast<size_t>(threads));
    std::atomic<int32_t> next(0);
    const int32_t grain = 1024;
    run_chunks(threads, threads, [&](int t) {
        std::vector<int64_t>& cnt = part[t];
        cnt.assign(static_cast<size_t>(n), 0);
        std::vector<uint64_t> mark((static_cast<size_t>(n) + 63) / 64, 0);
        for (int32_t b; (b = next.fetch_add(grain)) < n;) {
            const int32_t e = std::min(n, b + grain);
            for (int32_t u = b; u < e; ++u) {
                const int64_t ub = off[u], ue = off[u + 1];
                if (ue - ub < 2) continue;
                for (int64_t i = ub; i < ue; ++i) mark[fwd[i] >> 6] |= 1ull << (fwd[i] & 63);
                int64_t tu = 0;
                for (int64_t i = ub; i < ue; ++i) {
                    const int32_t v = fwd[i];
                    int64_t tv = 0;
                    for (int64_t j = off[v]; j < off[v + 1]; ++j) {
                        const int32_t w = fwd[j];
                        if (mark[w >> 6] >> (w & 63) & 1) {
                            ++tv;
                            ++cnt[w];
                        }
                    }
                    cnt[v] += tv;
                    tu += tv;
                }
                cnt[u] += tu;
                for (int64_t i = ub; i < ue; ++i) mark[fwd[i] >> 6] = 0;
            }
        }
    });
    run_chunks(threads, threads, [&](int q) {
        const int64_t b = static_cast<int64_t>(q) * n / threads, e = static_cast<int64_t>(q + 1) * n / threads;
        for (int64_t v = b; v < e; ++v) {
            int64_t s = 0;
            for (const auto& p : part) s += p[v];
            tri[v] = s;
        }
    });
}

}  // extern "C"
