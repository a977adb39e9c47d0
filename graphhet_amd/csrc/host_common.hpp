// Host-side helpers shared by the library's C++ translation units:
// thread-local error text, a fork-join parallel_for and the counter-based RNG.
#pragma once

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "ghet.h"

namespace ghet {

// Records the message returned by ghet_last_error() and passes `code` through.
int set_error(int code, const char* fmt, ...);
void clear_error();

inline int hw_threads(int requested) {
    if (requested > 0) return requested;
    unsigned h = std::thread::hardware_concurrency();
    return h == 0 ? 1 : static_cast<int>(std::min(h, 64u));
}

// Splits [0, n) into `threads` contiguous chunks and runs fn(begin, end, t).
template <typename F>
void parallel_for(int64_t n, int threads, F&& fn) {
    threads = std::max(1, threads);
    if (n <= 0) return;
    if (threads == 1 || n < 4096) {
        fn(int64_t(0), n, 0);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; ++t) {
        int64_t b = n * t / threads, e = n * (t + 1) / threads;
        pool.emplace_back([&fn, b, e, t] { fn(b, e, t); });
    }
    for (auto& th : pool) th.join();
}

// SplitMix64 output function (Steele, Lea, Flood 2014): a bijective 64-bit
// mixer.  Used as a counter-based generator: draw(key, j) = mix64(key + (j+1)*G).
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}
inline uint64_t stream_key(uint64_t seed, uint64_t k) { return mix64(seed + kGolden * (k + 1)); }
inline uint64_t draw(uint64_t key, uint64_t j) { return mix64(key + kGolden * (j + 1)); }
// Weight (k+1)*2^-53 with k = top 53 bits of a draw: uniform in (0,1], exact in FP64.
inline double weight_from(uint64_t r) { return static_cast<double>((r >> 11) + 1) * 0x1.0p-53; }

// Exclusive prefix sum in place over `v` (returns the total).  Parallel over chunks.
int64_t exclusive_scan_i64(std::vector<int64_t>& v, int threads);

}  // namespace ghet
