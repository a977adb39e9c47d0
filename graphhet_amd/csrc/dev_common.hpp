// Device-side building blocks for gfx950 (CDNA4, wave64): error macros,
// wave-level ballot/prefix helpers and relaxed agent-scope atomics.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "host_common.hpp"

#define GHET_HIP(call)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ::ghet::set_error(GHET_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                     __FILE__, __LINE__);                                          \
    } while (0)

#define GHET_CHECK(expr)            \
    do {                            \
        int st_ = (expr);           \
        if (st_ != GHET_OK) return st_; \
    } while (0)

namespace ghet {

constexpr int kWave = 64;
constexpr int32_t kIntInf = 0x7FFFFFFF;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set lanes in `mask` below this lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}

// Wave-aggregated append: every active lane calls it; lanes with pred=true get
// a unique slot in [*counter, *counter + count).  One atomic per wave.
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    return pred ? base + lanes_below(mask) : -1;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int32_t load_relaxed(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FP64 atomic min on non-negative doubles through the signed 64-bit bit
// pattern: for x, y >= 0 (finite or DBL_MAX), x < y  <=>  bits(x) < bits(y).
__device__ __forceinline__ double atomic_min_nonneg(double* addr, double v) {
    long long old = atomicMin(reinterpret_cast<long long*>(addr), __double_as_longlong(v));
    return __longlong_as_double(old);
}

// Compensated (Kahan-Babuska/Neumaier) FP32 accumulator.
struct Kahan32 {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void merge(const Kahan32& o) {
        add(o.s);
        c += o.c;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

__host__ __device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Device-scope atomics on one address serialise across the 8 XCDs (measured
// ~10 ns each on MI355X), so counters that many blocks bump are spread over
// kSlots cache-line-sized slots (block b uses slot b % kSlots) and summed on
// the host.
constexpr int kSlots = 64;
struct alignas(64) CtrSlot {
    unsigned long long v[8];  // v[7]: arrive_last's group count
};

__device__ __forceinline__ void slot_add(CtrSlot* slots, int field, unsigned long long x) {
    if (x) atomicAdd(&slots[blockIdx.x & (kSlots - 1)].v[field], x);
}
// slot_add that waits for the atomic and returns the slot's previous value
// (0 when x is 0): a token for arrive_last.
__device__ __forceinline__ unsigned long long slot_add_ret(CtrSlot* slots, int field, unsigned long long x) {
    return x ? atomicAdd(&slots[blockIdx.x & (kSlots - 1)].v[field], x) : 0ULL;
}

// Counter mailbox.  In place of a device-to-host copy and a stream
// synchronisation after each level or pass, the launch's last block writes
// the totals the host needs, then a sequence number, to fine-grained host
// memory that the host polls (about 6 us per round trip on MI355X against
// about 14 us for copy + synchronise).
struct alignas(128) HostMail {
    unsigned long long seq;
    unsigned long long v[15];
};

// Called by every block after its counter atomics (block-uniform).  True in
// the one block that arrives last; its counter reads then see every block's
// atomics.  No fences: a device-scope fence writes back the XCD's L2 (all of
// the level's dirty data) in every block.  Ordering comes from returned
// values instead: thread 0 passes `token`, a value computed from what its
// counter atomics returned, so the ticket atomic is issued only after they
// have completed (the test on it never fails in practice).  Arrival is
// counted in two levels so that no address takes more than about
// gridDim.x / kSlots atomics: block b first counts in slots[b % kSlots].v[7],
// and the last block of each group counts in *done.  The counters must be
// zero before the launch.
__device__ __forceinline__ bool arrive_last(CtrSlot* slots, unsigned int* done, unsigned long long token) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int G = gridDim.x, grp = blockIdx.x & (kSlots - 1);
        const unsigned int groups = G < kSlots ? G : kSlots;
        const unsigned long long in_grp = (G - grp + kSlots - 1) / kSlots;
        int last = 0;
        if (token != ~0ULL && atomicAdd(&slots[grp].v[7], 1ULL) == in_grp - 1)
            last = atomicAdd(done, 1u) == groups - 1;
        s_last = last;
    }
    __syncthreads();
    return s_last != 0;
}

__device__ __forceinline__ unsigned long long load_agent(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long slot_total(const CtrSlot* slots, int field) {
    unsigned long long t = 0;
    for (int i = 0; i < kSlots; ++i) t += load_agent(&slots[i].v[field]);
    return t;
}

// After the values: the sequence number, a system-scope release store (the
// values land first).
__device__ __forceinline__ void mail_post(HostMail* m, unsigned long long seq) {
    __hip_atomic_store(&m->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

inline unsigned long long slot_sum(const CtrSlot* host_slots, int field) {
    unsigned long long s = 0;
    for (int i = 0; i < kSlots; ++i) s += host_slots[i].v[field];
    return s;
}

// Block-wide sum of a 64-bit value (all threads call; result valid everywhere).
template <int BT>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long x, unsigned long long* s_red) {
    x = wave_sum(x);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned long long t = 0;
    for (int w = 0; w < BT / 64; ++w) t += s_red[w];
    __syncthreads();
    return t;
}

// Grid-wide barrier for cooperative launches, in place of the cooperative
// groups one (about 60 us per barrier with 768 blocks on MI355X, measured).
// Here thread 0 of each block fences once on each side and arrives in two
// levels: kBarGroups group counters, then the last
// arrival of each group bumps the top counter, which the waiters poll.
// Counters only grow (barrier k completes when the top counter reaches
// k * groups), so nothing is reset between barriers.  `bar` holds
// kBarWords zeroed words before the launch; `epoch` is the caller's count of
// barriers passed (thread 0's copy is the one used).
constexpr int kBarGroups = 16;
constexpr int kBarWords = (kBarGroups + 1) * 8;  // one 64-byte line per counter
__device__ __forceinline__ void grid_barrier(unsigned long long* bar, unsigned long long& epoch) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int G = gridDim.x;
        const unsigned int grp = blockIdx.x % kBarGroups;
        const unsigned long long groups = G < kBarGroups ? G : kBarGroups;
        const unsigned long long gsize = (G - grp + kBarGroups - 1) / kBarGroups;
        ++epoch;
        // The release fence writes back this XCD's L2 and waits for it before
        // the arrival atomic issues; the group's last arrival bumps the top
        // counter only after its own atomic has returned, so every block's
        // writes are in memory before a waiter sees the barrier complete.
        // One fence per side per block: about 13 us per barrier with 768
        // blocks, against about 60 us for the cooperative-groups barrier.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned long long old =
            __hip_atomic_fetch_add(bar + grp * 8, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == epoch * gsize)
            __hip_atomic_fetch_add(bar + kBarGroups * 8, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long target = epoch * groups;
        while (__hip_atomic_load(bar + kBarGroups * 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
            __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// LDS-staged output queue.  lq_push appends the lanes with pred (one LDS
// atomic per wave); lq_flush (block-uniform) moves the staged items to a
// global queue with ONE global atomic, calling on_item(v) for each item.
template <typename T>
__device__ __forceinline__ void lq_push(T* buf, int* cnt, bool pred, T v) {
    const uint64_t mask = __ballot(pred);
    if (!mask) return;
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, __popcll(mask));
    base = __shfl(base, leader);
    if (pred) buf[base + lanes_below(mask)] = v;
}

template <int BT, typename T, typename Count, typename OnItem>
__device__ __forceinline__ void lq_flush(T* buf, int* cnt, Count* s_base, Count* gcount, T* gqueue, int limit,
                                         OnItem&& on_item) {
    __syncthreads();
    const int c = *cnt;
    if (c == 0 || c <= limit) {
        __syncthreads();
        return;
    }
    if (threadIdx.x == 0) *s_base = atomicAdd(gcount, static_cast<Count>(c));
    __syncthreads();
    const Count base = *s_base;
    // four items per step, so on_item's loads for them overlap
    for (int j0 = threadIdx.x; j0 < c; j0 += 4 * BT) {
        T v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = j0 + k * BT;
            v[k] = buf[j < c ? j : j0];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (j0 + k * BT < c) gqueue[base + j0 + k * BT] = v[k];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (j0 + k * BT < c) on_item(v[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
}

}  // namespace ghet
