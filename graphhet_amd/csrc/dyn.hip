// Dynamic graph store: batched edge inserts, weight changes and deletes on
// the GPU, with compact CSR snapshots for the analytics.
//
// Reference: GPMA+ (GPU_gpma.hpp: class GPMA 42-73, update_gpma 756-824,
// init_csr_gpma 928-941, init_gpma_from_csr 882-926) keeps one packed-memory
// array of (src << 32 | dst) keys with a sentinel key per row, applies a
// batch by deleting (value 0), modifying (key present) or inserting, and
// rebalances segments within density bounds; the GPMA algorithms
// (GPU_*_gpma.cuh) then walk the slots and skip empty and deleted ones.
//
// Design here.  Every row owns a contiguous slot range [beg, beg + cap) of a
// pool and holds its `len` live edges there, sorted by target, with no holes
// and no sentinels, so a row is always a plain CSR row.  A batch, per
// direction (the out-store gets (src, dst), the in-store (dst, src)):
//   1. a stable radix sort of the (row, target) keys carrying the batch index;
//      the last occurrence of each key wins;
//   2. one thread per update binary-searches its row: weight 0 deletes a
//      present edge (an absent one is a no-op), a present edge takes the new
//      weight in place, an absent one becomes an insert;
//   3. rows with inserts or deletes get their new length; a row that
//      outgrows its capacity moves to fresh slots at the end of the pool
//      (capacity new length x 1.5); a full pool is compacted into a larger one
//      (every row at length x 1.25);
//   4. the new rows go to scratch -- old edge k to k - (deletes below it) +
//      (inserts below it), insert j to (old edges below it) - (deletes below
//      it) + j -- and are copied back.
// The work is proportional to the touched rows, not to m.  A snapshot scans
// the lengths and copies the rows into a ghet_graph (out-, in- and all-CSR)
// on which every analytics entry point runs unchanged.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "graph_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace ghet {

int graph_alloc_arena(int32_t n, int64_t m_out, bool want_in, int64_t m_in, bool want_all, int64_t m_all,
                      bool want_w, ghet_graph** out);
void graph_free(ghet_graph* g);
void launch_all_csr(const int32_t* orp, const int32_t* ocol, const int32_t* irp, const int32_t* icol, int32_t n,
                    int32_t* arp, int32_t* acol, hipStream_t s);

namespace {

constexpr int kBT = 256;
enum : int32_t { kNop = 0, kMod = 1, kDel = 2, kIns = 3 };

__host__ __device__ __forceinline__ int32_t slack_cap(int32_t len) { return len + (len / 4 > 4 ? len / 4 : 4); }
__host__ __device__ __forceinline__ int32_t grow_cap(int32_t len) { return len + (len / 2 > 4 ? len / 2 : 4); }

unsigned grid_for(int64_t items) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(ceil_div(items, kBT), 65536)));
}

template <typename T>
__device__ __forceinline__ int64_t lower_bound_d(const T* a, int64_t len, T key) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ int64_t upper_bound_d(const T* a, int64_t len, T key) {
    int64_t lo = 0, hi = len;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------- building ---
__global__ void __launch_bounds__(kBT) k_init_rows(const int32_t* __restrict__ rp, int32_t n, int32_t* __restrict__ len,
                                                   int32_t* __restrict__ cap, int64_t* __restrict__ cap64) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; r < n;
         r += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t d = rp[r + 1] - rp[r];
        len[r] = d;
        cap[r] = slack_cap(d);
        cap64[r] = slack_cap(d);
    }
}

// One wave per row: src row (CSR or pool at sbeg) -> pool at dbeg.
__global__ void __launch_bounds__(kBT) k_move_rows(int32_t n, const int32_t* __restrict__ srp,
                                                   const int64_t* __restrict__ sbeg, const int32_t* __restrict__ len,
                                                   const int32_t* __restrict__ scol, const double* __restrict__ sw,
                                                   const int64_t* __restrict__ dbeg, int32_t* __restrict__ dcol,
                                                   double* __restrict__ dw) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t r = wave; r < n; r += waves) {
        const int64_t s0 = srp ? srp[r] : sbeg[r];
        const int32_t l = len[r];
        const int64_t d0 = dbeg[r];
        for (int32_t k = lane; k < l; k += 64) {
            dcol[d0 + k] = scol[s0 + k];
            dw[d0 + k] = sw ? sw[s0 + k] : 1.0;
        }
    }
}

__global__ void __launch_bounds__(kBT) k_count_zero(const double* __restrict__ w, int64_t m, int32_t* __restrict__ flag) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * kBT)
        if (w[i] == 0.0) flag[0] = 1;
}

// -------------------------------------------------------------- updates ---
__global__ void __launch_bounds__(kBT) k_upd_keys(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int64_t B,
                                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < B;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        keys[i] = (static_cast<uint64_t>(static_cast<uint32_t>(a[i])) << 32) | static_cast<uint32_t>(b[i]);
        idx[i] = static_cast<uint32_t>(i);
    }
}

__global__ void __launch_bounds__(kBT) k_upd_last(const uint64_t* __restrict__ keys, int64_t B, int32_t* __restrict__ flag) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < B;
         i += static_cast<int64_t>(gridDim.x) * kBT)
        flag[i] = (i + 1 == B || keys[i + 1] != keys[i]) ? 1 : 0;
}

__global__ void __launch_bounds__(kBT) k_upd_compact(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                                     const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                                     int64_t B, const double* __restrict__ w,
                                                     uint64_t* __restrict__ ukey, double* __restrict__ uw) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < B;
         i += static_cast<int64_t>(gridDim.x) * kBT)
        if (flag[i]) {
            ukey[pos[i]] = keys[i];
            uw[pos[i]] = w ? w[idx[i]] : 1.0;
        }
}

// Classifies update i against its row; a present edge takes its new weight
// here.  Flags for the insert/delete/row-head scans.
__global__ void __launch_bounds__(kBT) k_upd_find(const uint64_t* __restrict__ ukey, const double* __restrict__ uw,
                                                  int64_t U, const int64_t* __restrict__ beg,
                                                  const int32_t* __restrict__ len, const int32_t* __restrict__ pcol,
                                                  double* __restrict__ pw, int32_t* __restrict__ fins,
                                                  int32_t* __restrict__ fdel, int32_t* __restrict__ fhead,
                                                  int32_t* __restrict__ fmod) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < U;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        const uint64_t k = ukey[i];
        const int32_t r = static_cast<int32_t>(k >> 32), c = static_cast<int32_t>(k & 0xFFFFFFFFu);
        const int64_t b = beg[r];
        const int32_t l = len[r];
        const int64_t p = lower_bound_d(pcol + b, l, c);
        const bool found = p < l && pcol[b + p] == c;
        const double w = uw[i];
        int32_t kind;
        if (w == 0.0) kind = found ? kDel : kNop;
        else if (found) {
            pw[b + p] = w;
            kind = kMod;
        } else kind = kIns;
        fins[i] = kind == kIns;
        fdel[i] = kind == kDel;
        fmod[i] = kind == kMod;
        fhead[i] = (i == 0 || (ukey[i - 1] >> 32) != (k >> 32)) ? 1 : 0;
    }
}

// Per distinct row rr (from the row heads): its row, and where its inserts
// and deletes start in the compacted lists; entry R closes the ranges.
__global__ void __launch_bounds__(kBT) k_upd_rows(const uint64_t* __restrict__ ukey, int64_t U,
                                                  const int32_t* __restrict__ fhead, const int32_t* __restrict__ hpos,
                                                  const int32_t* __restrict__ ipos, const int32_t* __restrict__ dpos,
                                                  int32_t* __restrict__ urow, int32_t* __restrict__ ifirst,
                                                  int32_t* __restrict__ dfirst) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < U;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        if (fhead[i]) {
            const int32_t rr = hpos[i];
            urow[rr] = static_cast<int32_t>(ukey[i] >> 32);
            ifirst[rr] = ipos[i];
            dfirst[rr] = dpos[i];
        }
        if (i == 0) {
            const int32_t R = hpos[U];
            ifirst[R] = ipos[U];
            dfirst[R] = dpos[U];
        }
    }
}

__global__ void __launch_bounds__(kBT) k_upd_lists(const uint64_t* __restrict__ ukey, const double* __restrict__ uw,
                                                   int64_t U, const int32_t* __restrict__ fins,
                                                   const int32_t* __restrict__ fdel, const int32_t* __restrict__ fhead,
                                                   const int32_t* __restrict__ hpos, const int32_t* __restrict__ ipos,
                                                   const int32_t* __restrict__ dpos, int32_t* __restrict__ icol,
                                                   double* __restrict__ iw, int32_t* __restrict__ irid,
                                                   int32_t* __restrict__ dcol) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; i < U;
         i += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t c = static_cast<int32_t>(ukey[i] & 0xFFFFFFFFu);
        if (fins[i]) {
            const int32_t j = ipos[i];
            icol[j] = c;
            iw[j] = uw[i];
            irid[j] = hpos[i] + fhead[i] - 1;
        }
        if (fdel[i]) dcol[dpos[i]] = c;
    }
}

// Per distinct row: new length, relocation and the sizes the scans place.
__global__ void __launch_bounds__(kBT) k_upd_plan(int32_t R, const int32_t* __restrict__ urow,
                                                  const int32_t* __restrict__ ifirst, const int32_t* __restrict__ dfirst,
                                                  const int32_t* __restrict__ len, const int32_t* __restrict__ cap,
                                                  int32_t* __restrict__ new_len, int32_t* __restrict__ new_cap,
                                                  int64_t* __restrict__ need, int32_t* __restrict__ slen,
                                                  int32_t* __restrict__ olen) {
    for (int64_t rr = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; rr < R;
         rr += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t r = urow[rr];
        const int32_t ni = ifirst[rr + 1] - ifirst[rr], nd = dfirst[rr + 1] - dfirst[rr];
        const int32_t l = len[r], nl = l + ni - nd;
        const bool structural = ni + nd > 0;
        const bool reloc = structural && nl > cap[r];
        new_len[rr] = nl;
        new_cap[rr] = reloc ? grow_cap(nl) : cap[r];
        need[rr] = reloc ? grow_cap(nl) : 0;
        slen[rr] = structural ? nl : 0;
        olen[rr] = structural ? l : 0;
    }
}

// Old edges of the restructured rows -> scratch (deleted ones dropped).
__global__ void __launch_bounds__(kBT) k_upd_merge_old(int64_t total, int32_t R, const int32_t* __restrict__ ooff,
                                                       const int32_t* __restrict__ urow, const int64_t* __restrict__ beg,
                                                       const int32_t* __restrict__ pcol, const double* __restrict__ pw,
                                                       const int32_t* __restrict__ ifirst, const int32_t* __restrict__ icol,
                                                       const int32_t* __restrict__ dfirst, const int32_t* __restrict__ dcol,
                                                       const int32_t* __restrict__ soff, int32_t* __restrict__ scol,
                                                       double* __restrict__ sw) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t rr = static_cast<int32_t>(upper_bound_d(ooff, int64_t(R) + 1, static_cast<int32_t>(e)) - 1);
        const int32_t k = static_cast<int32_t>(e - ooff[rr]);
        const int64_t b = beg[urow[rr]];
        const int32_t c = pcol[b + k];
        const int32_t* dl = dcol + dfirst[rr];
        const int32_t nd = dfirst[rr + 1] - dfirst[rr];
        const int32_t dn = static_cast<int32_t>(lower_bound_d(dl, nd, c));
        if (dn < nd && dl[dn] == c) continue;
        const int32_t ib = static_cast<int32_t>(lower_bound_d(icol + ifirst[rr], ifirst[rr + 1] - ifirst[rr], c));
        const int64_t o = int64_t(soff[rr]) + k - dn + ib;
        scol[o] = c;
        sw[o] = pw[b + k];
    }
}

// Inserted edges -> scratch.
__global__ void __launch_bounds__(kBT) k_upd_merge_ins(int64_t total, const int32_t* __restrict__ icol,
                                                       const double* __restrict__ iw, const int32_t* __restrict__ irid,
                                                       const int32_t* __restrict__ ifirst, const int32_t* __restrict__ urow,
                                                       const int64_t* __restrict__ beg, const int32_t* __restrict__ len,
                                                       const int32_t* __restrict__ pcol,
                                                       const int32_t* __restrict__ dfirst, const int32_t* __restrict__ dcol,
                                                       const int32_t* __restrict__ soff, int32_t* __restrict__ scol,
                                                       double* __restrict__ sw) {
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; j < total;
         j += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t rr = irid[j];
        const int32_t r = urow[rr];
        const int32_t c = icol[j];
        const int32_t ob = static_cast<int32_t>(lower_bound_d(pcol + beg[r], len[r], c));
        const int32_t dn = static_cast<int32_t>(lower_bound_d(dcol + dfirst[rr], dfirst[rr + 1] - dfirst[rr], c));
        const int64_t o = int64_t(soff[rr]) + ob - dn + (j - ifirst[rr]);
        scol[o] = c;
        sw[o] = iw[j];
    }
}

// Scratch rows -> pool (their old slots, or fresh slots at top + need_off).
__global__ void __launch_bounds__(kBT) k_upd_copy_back(int64_t total, int32_t R, const int32_t* __restrict__ soff,
                                                       const int32_t* __restrict__ urow, const int64_t* __restrict__ noff,
                                                       int64_t top, const int64_t* __restrict__ beg,
                                                       const int32_t* __restrict__ scol, const double* __restrict__ sw,
                                                       int32_t* __restrict__ pcol, double* __restrict__ pw) {
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; e < total;
         e += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t rr = static_cast<int32_t>(upper_bound_d(soff, int64_t(R) + 1, static_cast<int32_t>(e)) - 1);
        const int64_t k = e - soff[rr];
        const int64_t dst = (noff[rr + 1] > noff[rr]) ? top + noff[rr] : beg[urow[rr]];
        pcol[dst + k] = scol[e];
        pw[dst + k] = sw[e];
    }
}

__global__ void __launch_bounds__(kBT) k_upd_commit(int32_t R, const int32_t* __restrict__ urow,
                                                    const int64_t* __restrict__ noff, int64_t top,
                                                    const int32_t* __restrict__ new_len,
                                                    const int32_t* __restrict__ new_cap, int64_t* __restrict__ beg,
                                                    int32_t* __restrict__ len, int32_t* __restrict__ cap) {
    for (int64_t rr = static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x; rr < R;
         rr += static_cast<int64_t>(gridDim.x) * kBT) {
        const int32_t r = urow[rr];
        if (noff[rr + 1] > noff[rr]) beg[r] = top + noff[rr];
        len[r] = new_len[rr];
        cap[r] = new_cap[rr];
    }
}

// ------------------------------------------------------------- snapshot ---
__global__ void __launch_bounds__(kBT) k_snap_rows(int32_t n, const int64_t* __restrict__ beg,
                                                   const int32_t* __restrict__ len, const int32_t* __restrict__ pcol,
                                                   const double* __restrict__ pw, const int32_t* __restrict__ rp,
                                                   int32_t* __restrict__ col, double* __restrict__ w) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBT + threadIdx.x) >> 6;
    const int64_t waves = (static_cast<int64_t>(gridDim.x) * kBT) >> 6;
    for (int64_t r = wave; r < n; r += waves) {
        const int64_t b = beg[r];
        const int32_t l = len[r], d = rp[r];
        for (int32_t k = lane; k < l; k += 64) {
            col[d + k] = pcol[b + k];
            if (w) w[d + k] = pw[b + k];
        }
    }
}

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int need(size_t b) {
        b = std::max<size_t>(b, 256);
        if (b <= bytes) return GHET_OK;
        release();
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_error(GHET_ERR_NOMEM, "dynamic graph: hipMalloc(%zu) failed", b);
        }
        bytes = b;
        return GHET_OK;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// One direction: rows of the pool.
struct Store {
    int64_t pool = 0, top = 0, m = 0;
    Buf beg, len, cap, col, w;
};

}  // namespace
}  // namespace ghet

struct ghet_dyn {
    int32_t n = 0;
    int device = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ghet::Store dir[2];  // 0: out (src -> dst), 1: in (dst -> src)
    // per-batch scratch (grow-only)
    ghet::Buf keys[2], idx[2], flag, pos, ukey, uw, fins, fdel, fhead, fmod, ipos, dpos, hpos, mpos, urow, ifirst,
        dfirst, icol, iw, irid, dcol, new_len, new_cap, need, noff, slen, soff, olen, ooff, scol, sw, aux, aux64, tmp,
        d_src, d_dst, d_w;
    int64_t* pinned = nullptr;
    int64_t compactions = 0;
};

namespace ghet {
namespace {

int dyn_sync_read(ghet_dyn* d, const void* dev, size_t bytes, int slot) {
    GHET_HIP(hipMemcpyAsync(reinterpret_cast<char*>(d->pinned) + 8 * slot, dev, bytes, hipMemcpyDeviceToHost, d->s));
    return GHET_OK;
}

// Builds a store from a device CSR (rp, col, w may be null = 1.0).
int store_init(ghet_dyn* d, Store& st, const int32_t* rp, const int32_t* col, const double* w, int64_t m) {
    const int32_t n = d->n;
    GHET_CHECK(st.beg.need(8 * (size_t(n) + 1)));
    GHET_CHECK(st.len.need(4 * size_t(n)));
    GHET_CHECK(st.cap.need(4 * size_t(n)));
    GHET_CHECK(d->aux64.need(8 * size_t(scan_aux_ints(n) + 1)));
    Buf cap64;
    GHET_CHECK(cap64.need(8 * (size_t(n) + 1)));
    hipLaunchKernelGGL(k_init_rows, dim3(grid_for(n)), dim3(kBT), 0, d->s, rp, n, st.len.as<int32_t>(),
                       st.cap.as<int32_t>(), cap64.as<int64_t>());
    GHET_CHECK(exclusive_scan_i64(cap64.as<int64_t>(), st.beg.as<int64_t>(), n, d->aux64.as<int64_t>(), d->s));
    GHET_CHECK(dyn_sync_read(d, st.beg.as<int64_t>() + n, 8, 0));
    GHET_HIP(hipStreamSynchronize(d->s));
    const int64_t used = d->pinned[0];
    st.pool = used + std::max<int64_t>(used / 4, 1024);
    st.top = used;
    st.m = m;
    GHET_CHECK(st.col.need(4 * size_t(st.pool)));
    GHET_CHECK(st.w.need(8 * size_t(st.pool)));
    hipLaunchKernelGGL(k_move_rows, dim3(grid_for(int64_t(n) * 64)), dim3(kBT), 0, d->s, n, rp, nullptr,
                       st.len.as<int32_t>(), col, w, st.beg.as<int64_t>(), st.col.as<int32_t>(), st.w.as<double>());
    GHET_HIP(hipGetLastError());
    GHET_HIP(hipStreamSynchronize(d->s));
    return GHET_OK;
}

// Moves every row into a fresh pool at length x 1.25 with `extra` free slots
// beyond the rows.
int store_compact(ghet_dyn* d, Store& st, int64_t extra) {
    const int32_t n = d->n;
    Buf cap64, nbeg, ncol, nw;
    GHET_CHECK(cap64.need(8 * (size_t(n) + 1)));
    GHET_CHECK(nbeg.need(8 * (size_t(n) + 1)));
    // new capacities from the current lengths (rp = null: len drives the caps)
    Buf rp;
    GHET_CHECK(rp.need(4 * (size_t(n) + 1)));
    GHET_CHECK(d->aux.need(4 * size_t(scan_aux_ints(n) + 1)));
    GHET_CHECK(exclusive_scan_i32(st.len.as<int32_t>(), rp.as<int32_t>(), n, d->aux.as<int32_t>(), d->s));
    Buf len2;
    GHET_CHECK(len2.need(4 * size_t(n)));
    hipLaunchKernelGGL(k_init_rows, dim3(grid_for(n)), dim3(kBT), 0, d->s, rp.as<int32_t>(), n, len2.as<int32_t>(),
                       st.cap.as<int32_t>(), cap64.as<int64_t>());
    GHET_CHECK(exclusive_scan_i64(cap64.as<int64_t>(), nbeg.as<int64_t>(), n, d->aux64.as<int64_t>(), d->s));
    GHET_CHECK(dyn_sync_read(d, nbeg.as<int64_t>() + n, 8, 0));
    GHET_HIP(hipStreamSynchronize(d->s));
    const int64_t used = d->pinned[0];
    const int64_t pool = used + extra + std::max<int64_t>(used / 4, 1024);
    GHET_CHECK(ncol.need(4 * size_t(pool)));
    GHET_CHECK(nw.need(8 * size_t(pool)));
    hipLaunchKernelGGL(k_move_rows, dim3(grid_for(int64_t(n) * 64)), dim3(kBT), 0, d->s, n, nullptr,
                       st.beg.as<int64_t>(), st.len.as<int32_t>(), st.col.as<int32_t>(), st.w.as<double>(),
                       nbeg.as<int64_t>(), ncol.as<int32_t>(), nw.as<double>());
    GHET_HIP(hipGetLastError());
    GHET_HIP(hipStreamSynchronize(d->s));
    std::swap(st.beg.p, nbeg.p);
    std::swap(st.beg.bytes, nbeg.bytes);
    std::swap(st.col.p, ncol.p);
    std::swap(st.col.bytes, ncol.bytes);
    std::swap(st.w.p, nw.p);
    std::swap(st.w.bytes, nw.bytes);
    st.pool = pool;
    st.top = used;
    ++d->compactions;
    return GHET_OK;
}

struct BatchCounts {
    int64_t U = 0, inserted = 0, deleted = 0, modified = 0, relocated_slots = 0;
};

// Applies the batch (a[i] -> b[i], weight w[i] or 1.0) to one direction.
int store_update(ghet_dyn* d, Store& st, const int32_t* a, const int32_t* b, const double* w, int64_t B,
                 BatchCounts* bc) {
    const int32_t n = d->n;
    hipStream_t s = d->s;
    int bits = 1;
    while ((int64_t(1) << bits) < n) ++bits;
    // 1. sort (row, target) keys with the batch index, keep the last of each key
    GHET_CHECK(d->keys[0].need(8 * size_t(B)));
    GHET_CHECK(d->keys[1].need(8 * size_t(B)));
    GHET_CHECK(d->idx[0].need(4 * size_t(B)));
    GHET_CHECK(d->idx[1].need(4 * size_t(B)));
    hipLaunchKernelGGL(k_upd_keys, dim3(grid_for(B)), dim3(kBT), 0, s, a, b, B, d->keys[0].as<uint64_t>(),
                       d->idx[0].as<uint32_t>());
    rocprim::double_buffer<uint64_t> kd(d->keys[0].as<uint64_t>(), d->keys[1].as<uint64_t>());
    rocprim::double_buffer<uint32_t> vd(d->idx[0].as<uint32_t>(), d->idx[1].as<uint32_t>());
    size_t tb = 0;
    GHET_HIP(rocprim::radix_sort_pairs(nullptr, tb, kd, vd, B, 0, static_cast<unsigned>(32 + bits), s));
    GHET_CHECK(d->tmp.need(tb));
    GHET_HIP(rocprim::radix_sort_pairs(d->tmp.p, tb, kd, vd, B, 0, static_cast<unsigned>(32 + bits), s));
    GHET_CHECK(d->flag.need(4 * size_t(B)));
    GHET_CHECK(d->pos.need(4 * (size_t(B) + 1)));
    GHET_CHECK(d->aux.need(4 * size_t(scan_aux_ints(B) + scan_aux_ints(n) + 2)));
    hipLaunchKernelGGL(k_upd_last, dim3(grid_for(B)), dim3(kBT), 0, s, kd.current(), B, d->flag.as<int32_t>());
    GHET_CHECK(exclusive_scan_i32(d->flag.as<int32_t>(), d->pos.as<int32_t>(), B, d->aux.as<int32_t>(), s));
    GHET_CHECK(dyn_sync_read(d, d->pos.as<int32_t>() + B, 4, 0));
    GHET_HIP(hipStreamSynchronize(s));
    const int64_t U = *reinterpret_cast<int32_t*>(d->pinned);
    GHET_CHECK(d->ukey.need(8 * size_t(U)));
    GHET_CHECK(d->uw.need(8 * size_t(U)));
    hipLaunchKernelGGL(k_upd_compact, dim3(grid_for(B)), dim3(kBT), 0, s, kd.current(), vd.current(),
                       d->flag.as<int32_t>(), d->pos.as<int32_t>(), B, w, d->ukey.as<uint64_t>(), d->uw.as<double>());
    // 2. classify against the rows; weight changes land in place
    for (Buf* f : {&d->fins, &d->fdel, &d->fhead, &d->fmod}) GHET_CHECK(f->need(4 * size_t(U)));
    for (Buf* f : {&d->ipos, &d->dpos, &d->hpos, &d->mpos}) GHET_CHECK(f->need(4 * (size_t(U) + 1)));
    hipLaunchKernelGGL(k_upd_find, dim3(grid_for(U)), dim3(kBT), 0, s, d->ukey.as<uint64_t>(), d->uw.as<double>(), U,
                       st.beg.as<int64_t>(), st.len.as<int32_t>(), st.col.as<int32_t>(), st.w.as<double>(),
                       d->fins.as<int32_t>(), d->fdel.as<int32_t>(), d->fhead.as<int32_t>(), d->fmod.as<int32_t>());
    GHET_CHECK(exclusive_scan_i32(d->fins.as<int32_t>(), d->ipos.as<int32_t>(), U, d->aux.as<int32_t>(), s));
    GHET_CHECK(exclusive_scan_i32(d->fdel.as<int32_t>(), d->dpos.as<int32_t>(), U, d->aux.as<int32_t>(), s));
    GHET_CHECK(exclusive_scan_i32(d->fhead.as<int32_t>(), d->hpos.as<int32_t>(), U, d->aux.as<int32_t>(), s));
    GHET_CHECK(exclusive_scan_i32(d->fmod.as<int32_t>(), d->mpos.as<int32_t>(), U, d->aux.as<int32_t>(), s));
    GHET_CHECK(dyn_sync_read(d, d->ipos.as<int32_t>() + U, 4, 0));
    GHET_CHECK(dyn_sync_read(d, d->dpos.as<int32_t>() + U, 4, 1));
    GHET_CHECK(dyn_sync_read(d, d->hpos.as<int32_t>() + U, 4, 2));
    GHET_CHECK(dyn_sync_read(d, d->mpos.as<int32_t>() + U, 4, 3));
    GHET_HIP(hipStreamSynchronize(s));
    const int64_t NI = *reinterpret_cast<int32_t*>(d->pinned);
    const int64_t ND = *reinterpret_cast<int32_t*>(d->pinned + 1);
    const int32_t R = *reinterpret_cast<int32_t*>(d->pinned + 2);
    const int64_t NM = *reinterpret_cast<int32_t*>(d->pinned + 3);
    bc->U += U;
    bc->inserted += NI;
    bc->deleted += ND;
    bc->modified += NM;
    if (NI + ND == 0) return GHET_OK;
    for (Buf* f : {&d->urow, &d->ifirst, &d->dfirst}) GHET_CHECK(f->need(4 * (size_t(R) + 1)));
    GHET_CHECK(d->icol.need(4 * size_t(NI)));
    GHET_CHECK(d->iw.need(8 * size_t(NI)));
    GHET_CHECK(d->irid.need(4 * size_t(NI)));
    GHET_CHECK(d->dcol.need(4 * size_t(ND)));
    hipLaunchKernelGGL(k_upd_rows, dim3(grid_for(U)), dim3(kBT), 0, s, d->ukey.as<uint64_t>(), U,
                       d->fhead.as<int32_t>(), d->hpos.as<int32_t>(), d->ipos.as<int32_t>(), d->dpos.as<int32_t>(),
                       d->urow.as<int32_t>(), d->ifirst.as<int32_t>(), d->dfirst.as<int32_t>());
    hipLaunchKernelGGL(k_upd_lists, dim3(grid_for(U)), dim3(kBT), 0, s, d->ukey.as<uint64_t>(), d->uw.as<double>(), U,
                       d->fins.as<int32_t>(), d->fdel.as<int32_t>(), d->fhead.as<int32_t>(), d->hpos.as<int32_t>(),
                       d->ipos.as<int32_t>(), d->dpos.as<int32_t>(), d->icol.as<int32_t>(), d->iw.as<double>(),
                       d->irid.as<int32_t>(), d->dcol.as<int32_t>());
    // 3. new lengths, relocations; compact the pool when it cannot take them
    for (Buf* f : {&d->new_len, &d->new_cap, &d->slen, &d->olen}) GHET_CHECK(f->need(4 * size_t(R)));
    GHET_CHECK(d->need.need(8 * size_t(R)));
    GHET_CHECK(d->noff.need(8 * (size_t(R) + 1)));
    GHET_CHECK(d->soff.need(4 * (size_t(R) + 1)));
    GHET_CHECK(d->ooff.need(4 * (size_t(R) + 1)));
    GHET_CHECK(d->aux64.need(8 * size_t(scan_aux_ints(std::max<int64_t>(R, n)) + 1)));
    int64_t need_total = 0, s_total = 0, o_total = 0;
    for (int attempt = 0;; ++attempt) {
        hipLaunchKernelGGL(k_upd_plan, dim3(grid_for(R)), dim3(kBT), 0, s, R, d->urow.as<int32_t>(),
                           d->ifirst.as<int32_t>(), d->dfirst.as<int32_t>(), st.len.as<int32_t>(),
                           st.cap.as<int32_t>(), d->new_len.as<int32_t>(), d->new_cap.as<int32_t>(),
                           d->need.as<int64_t>(), d->slen.as<int32_t>(), d->olen.as<int32_t>());
        GHET_CHECK(exclusive_scan_i64(d->need.as<int64_t>(), d->noff.as<int64_t>(), R, d->aux64.as<int64_t>(), s));
        GHET_CHECK(exclusive_scan_i32(d->slen.as<int32_t>(), d->soff.as<int32_t>(), R, d->aux.as<int32_t>(), s));
        GHET_CHECK(exclusive_scan_i32(d->olen.as<int32_t>(), d->ooff.as<int32_t>(), R, d->aux.as<int32_t>(), s));
        GHET_CHECK(dyn_sync_read(d, d->noff.as<int64_t>() + R, 8, 0));
        GHET_CHECK(dyn_sync_read(d, d->soff.as<int32_t>() + R, 4, 1));
        GHET_CHECK(dyn_sync_read(d, d->ooff.as<int32_t>() + R, 4, 2));
        GHET_HIP(hipStreamSynchronize(s));
        need_total = d->pinned[0];
        s_total = *reinterpret_cast<int32_t*>(d->pinned + 1);
        o_total = *reinterpret_cast<int32_t*>(d->pinned + 2);
        if (st.top + need_total <= st.pool) break;
        if (attempt > 0) return set_error(GHET_ERR_NOMEM, "dynamic graph: pool still full after compaction");
        GHET_CHECK(store_compact(d, st, need_total));
    }
    // 4. restructured rows -> scratch -> pool
    GHET_CHECK(d->scol.need(4 * size_t(s_total)));
    GHET_CHECK(d->sw.need(8 * size_t(s_total)));
    if (o_total > 0)
        hipLaunchKernelGGL(k_upd_merge_old, dim3(grid_for(o_total)), dim3(kBT), 0, s, o_total, R,
                           d->ooff.as<int32_t>(), d->urow.as<int32_t>(), st.beg.as<int64_t>(), st.col.as<int32_t>(),
                           st.w.as<double>(), d->ifirst.as<int32_t>(), d->icol.as<int32_t>(), d->dfirst.as<int32_t>(),
                           d->dcol.as<int32_t>(), d->soff.as<int32_t>(), d->scol.as<int32_t>(), d->sw.as<double>());
    if (NI > 0)
        hipLaunchKernelGGL(k_upd_merge_ins, dim3(grid_for(NI)), dim3(kBT), 0, s, NI, d->icol.as<int32_t>(),
                           d->iw.as<double>(), d->irid.as<int32_t>(), d->ifirst.as<int32_t>(), d->urow.as<int32_t>(),
                           st.beg.as<int64_t>(), st.len.as<int32_t>(), st.col.as<int32_t>(), d->dfirst.as<int32_t>(),
                           d->dcol.as<int32_t>(), d->soff.as<int32_t>(), d->scol.as<int32_t>(), d->sw.as<double>());
    if (s_total > 0)
        hipLaunchKernelGGL(k_upd_copy_back, dim3(grid_for(s_total)), dim3(kBT), 0, s, s_total, R,
                           d->soff.as<int32_t>(), d->urow.as<int32_t>(), d->noff.as<int64_t>(), st.top,
                           st.beg.as<int64_t>(), d->scol.as<int32_t>(), d->sw.as<double>(), st.col.as<int32_t>(),
                           st.w.as<double>());
    hipLaunchKernelGGL(k_upd_commit, dim3(grid_for(R)), dim3(kBT), 0, s, R, d->urow.as<int32_t>(),
                       d->noff.as<int64_t>(), st.top, d->new_len.as<int32_t>(), d->new_cap.as<int32_t>(),
                       st.beg.as<int64_t>(), st.len.as<int32_t>(), st.cap.as<int32_t>());
    GHET_HIP(hipGetLastError());
    st.top += need_total;
    st.m += NI - ND;
    bc->relocated_slots += need_total;
    return GHET_OK;
}

// Frees a store made by dyn_begin; a pointer that is not live (already
// destroyed, or released by ghet_shutdown) is ignored.
void dyn_free(ghet_dyn* d) {
    if (!d || !live_remove(LIVE_DYN, d)) return;
    if (d->s) hipStreamSynchronize(d->s);
    for (Store& st : d->dir)
        for (Buf* b : {&st.beg, &st.len, &st.cap, &st.col, &st.w}) b->release();
    if (d->pinned) hipHostFree(d->pinned);
    if (d->ev0) hipEventDestroy(d->ev0);
    if (d->ev1) hipEventDestroy(d->ev1);
    if (d->s) hipStreamDestroy(d->s);
    delete d;
}

int dyn_begin(int32_t n, ghet_dyn** out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(GHET_ERR_NODEVICE, "no GPU visible to this process");
    }
    auto* d = new ghet_dyn();
    live_add(LIVE_DYN, d);
    d->n = n;
    hipError_t e = hipGetDevice(&d->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&d->ev0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev1);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&d->pinned), 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        dyn_free(d);
        return set_error(GHET_ERR_HIP, "dynamic graph setup: %s", hipGetErrorString(e));
    }
    *out = d;
    return GHET_OK;
}

}  // namespace
}  // namespace ghet

using namespace ghet;

extern "C" {

int ghet_dyn_create(int32_t n, const int32_t* rowptr, const int32_t* col, const double* w, int threads,
                    ghet_dyn** out) {
    clear_error();
    if (!out) return set_error(GHET_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (n < 0 || !rowptr) return set_error(GHET_ERR_INVALID, "bad CSR");
    // weight 0 means "no edge" (a delete in an update): drop such edges
    std::vector<int32_t> frp, fcol;
    std::vector<double> fw;
    const int32_t* rp = rowptr;
    const int32_t* cl = col;
    const double* wt = w;
    if (w) {
        bool any_zero = false;
        for (int64_t i = 0; i < rowptr[n] && !any_zero; ++i) any_zero = w[i] == 0.0;
        if (any_zero) {
            frp.assign(size_t(n) + 1, 0);
            for (int32_t u = 0; u < n; ++u) {
                for (int32_t i = rowptr[u]; i < rowptr[u + 1]; ++i)
                    if (w[i] != 0.0) {
                        fcol.push_back(col[i]);
                        fw.push_back(w[i]);
                    }
                frp[u + 1] = static_cast<int32_t>(fcol.size());
            }
            rp = frp.data();
            cl = fcol.data();
            wt = fw.data();
        }
    }
    ghet_host_graph* hg = nullptr;
    GHET_CHECK(ghet_host_graph_from_csr(n, rp, cl, wt, threads, &hg));
    std::unique_ptr<ghet_host_graph, void (*)(ghet_host_graph*)> hguard(hg, ghet_host_graph_free);
    ghet_csr_view vo{}, vi{};
    GHET_CHECK(ghet_host_graph_view(hg, GHET_VIEW_OUT, &vo));
    GHET_CHECK(ghet_host_graph_view(hg, GHET_VIEW_IN, &vi));
    ghet_dyn* d = nullptr;
    GHET_CHECK(dyn_begin(n, &d));
    std::unique_ptr<ghet_dyn, void (*)(ghet_dyn*)> dguard(d, dyn_free);
    for (int k = 0; k < 2; ++k) {
        const ghet_csr_view& v = k == 0 ? vo : vi;
        Buf drp, dcol, dw;
        GHET_CHECK(drp.need(4 * (size_t(n) + 1)));
        GHET_CHECK(dcol.need(4 * size_t(v.m)));
        GHET_CHECK(dw.need(8 * size_t(v.m)));
        GHET_HIP(hipMemcpyAsync(drp.p, v.rowptr, 4 * (size_t(n) + 1), hipMemcpyHostToDevice, d->s));
        if (v.m > 0) {
            GHET_HIP(hipMemcpyAsync(dcol.p, v.col, 4 * size_t(v.m), hipMemcpyHostToDevice, d->s));
            GHET_HIP(hipMemcpyAsync(dw.p, v.w, 8 * size_t(v.m), hipMemcpyHostToDevice, d->s));
        }
        GHET_CHECK(store_init(d, d->dir[k], drp.as<int32_t>(), dcol.as<int32_t>(), dw.as<double>(), v.m));
    }
    *out = dguard.release();
    return GHET_OK;
}

int ghet_dyn_from_graph(const ghet_graph* g, ghet_dyn** out) {
    clear_error();
    if (!g || !out) return set_error(GHET_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (!g->out.present() || !g->in.present() || g->in.rowptr == g->out.rowptr)
        return set_error(GHET_ERR_INVALID, "dynamic graph needs a directed graph with its in-CSR");
    GHET_HIP(hipSetDevice(g->device));
    ghet_dyn* d = nullptr;
    GHET_CHECK(dyn_begin(g->n, &d));
    std::unique_ptr<ghet_dyn, void (*)(ghet_dyn*)> dguard(d, dyn_free);
    if (g->out.w) {
        Buf zf;
        GHET_CHECK(zf.need(4));
        GHET_CHECK(fill<int32_t>(zf.as<int32_t>(), 0, 1, d->s));
        hipLaunchKernelGGL(k_count_zero, dim3(grid_for(g->out.m)), dim3(kBT), 0, d->s, g->out.w, g->out.m,
                           zf.as<int32_t>());
        GHET_CHECK(dyn_sync_read(d, zf.p, 4, 0));
        GHET_HIP(hipStreamSynchronize(d->s));
        if (*reinterpret_cast<int32_t*>(d->pinned) != 0)
            return set_error(GHET_ERR_INVALID, "graph has zero-weight edges (weight 0 means no edge here)");
    }
    GHET_CHECK(store_init(d, d->dir[0], g->out.rowptr, g->out.col, g->out.w, g->out.m));
    GHET_CHECK(store_init(d, d->dir[1], g->in.rowptr, g->in.col, g->in.w, g->in.m));
    *out = dguard.release();
    return GHET_OK;
}

int ghet_dyn_update(ghet_dyn* d, int64_t count, const int32_t* src, const int32_t* dst, const double* w,
                    ghet_run_stats* st) {
    clear_error();
    if (st) std::memset(st, 0, sizeof(*st));
    if (!d) return set_error(GHET_ERR_INVALID, "dynamic graph handle is NULL");
    if (count < 0 || (count > 0 && (!src || !dst))) return set_error(GHET_ERR_INVALID, "bad update batch");
    if (count >= INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "batch of %lld updates", (long long)count);
    for (int64_t i = 0; i < count; ++i) {
        if (src[i] < 0 || src[i] >= d->n || dst[i] < 0 || dst[i] >= d->n)
            return set_error(GHET_ERR_INVALID, "update %lld (%d,%d) out of range [0,%d)", (long long)i, src[i], dst[i],
                             d->n);
        if (w && !(w[i] >= 0.0)) return set_error(GHET_ERR_INVALID, "update %lld has weight %g", (long long)i, w[i]);
    }
    if (count == 0) return GHET_OK;
    GHET_HIP(hipSetDevice(d->device));
    const int64_t c0 = d->compactions;
    GHET_CHECK(d->d_src.need(4 * size_t(count)));
    GHET_CHECK(d->d_dst.need(4 * size_t(count)));
    if (w) GHET_CHECK(d->d_w.need(8 * size_t(count)));
    GHET_HIP(hipEventRecord(d->ev0, d->s));
    GHET_HIP(hipMemcpyAsync(d->d_src.p, src, 4 * size_t(count), hipMemcpyHostToDevice, d->s));
    GHET_HIP(hipMemcpyAsync(d->d_dst.p, dst, 4 * size_t(count), hipMemcpyHostToDevice, d->s));
    if (w) GHET_HIP(hipMemcpyAsync(d->d_w.p, w, 8 * size_t(count), hipMemcpyHostToDevice, d->s));
    const double* dw = w ? d->d_w.as<double>() : nullptr;
    BatchCounts out_c, in_c;
    GHET_CHECK(store_update(d, d->dir[0], d->d_src.as<int32_t>(), d->d_dst.as<int32_t>(), dw, count, &out_c));
    GHET_CHECK(store_update(d, d->dir[1], d->d_dst.as<int32_t>(), d->d_src.as<int32_t>(), dw, count, &in_c));
    GHET_HIP(hipEventRecord(d->ev1, d->s));
    GHET_HIP(hipStreamSynchronize(d->s));
    if (out_c.inserted != in_c.inserted || out_c.deleted != in_c.deleted || d->dir[0].m != d->dir[1].m)
        return set_error(GHET_ERR_HIP, "dynamic graph: out- and in-stores disagree after the batch");
    if (st) {
        float ms = 0;
        GHET_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
        st->device_ms = ms;
        st->edges_traversed = out_c.U;                    // distinct updates applied
        st->levels = static_cast<int32_t>(d->compactions - c0);  // pool compactions
        st->extra[0] = static_cast<double>(out_c.inserted);
        st->extra[1] = static_cast<double>(out_c.deleted);
        st->extra[2] = static_cast<double>(out_c.modified);
        st->extra[3] = static_cast<double>(out_c.relocated_slots + in_c.relocated_slots);
    }
    return GHET_OK;
}

int64_t ghet_dyn_edges(const ghet_dyn* d) { return d ? d->dir[0].m : 0; }

int ghet_dyn_snapshot(ghet_dyn* d, uint32_t flags, ghet_graph** out) {
    clear_error();
    if (!d || !out) return set_error(GHET_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (flags & GHET_GRAPH_UNDIRECTED) return set_error(GHET_ERR_INVALID, "dynamic graphs are directed");
    const int32_t n = d->n;
    const int64_t m = d->dir[0].m;
    if (m > INT32_MAX || 2 * m > INT32_MAX) return set_error(GHET_ERR_UNSUPPORTED, "snapshot of %lld edges", (long long)m);
    GHET_HIP(hipSetDevice(d->device));
    const bool want_w = !(flags & GHET_GRAPH_NO_WEIGHTS);
    const bool want_in = !(flags & GHET_GRAPH_NO_IN);
    const bool want_all = want_in && !(flags & GHET_GRAPH_NO_ALL);
    ghet_graph* g = nullptr;
    GHET_CHECK(graph_alloc_arena(n, m, want_in, m, want_all, want_all ? 2 * m : 0, want_w, &g));
    std::unique_ptr<ghet_graph, void (*)(ghet_graph*)> gguard(g, graph_free);
    GHET_CHECK(d->aux.need(4 * size_t(scan_aux_ints(n) + 1)));
    for (int k = 0; k < (want_in ? 2 : 1); ++k) {
        Store& st = d->dir[k];
        DevCSR& c = k == 0 ? g->out : g->in;
        GHET_CHECK(exclusive_scan_i32(st.len.as<int32_t>(), c.rowptr, n, d->aux.as<int32_t>(), d->s));
        hipLaunchKernelGGL(k_snap_rows, dim3(grid_for(int64_t(n) * 64)), dim3(kBT), 0, d->s, n, st.beg.as<int64_t>(),
                           st.len.as<int32_t>(), st.col.as<int32_t>(), st.w.as<double>(), c.rowptr, c.col, c.w);
        GHET_HIP(hipGetLastError());
    }
    if (want_all) launch_all_csr(g->out.rowptr, g->out.col, g->in.rowptr, g->in.col, n, g->all.rowptr, g->all.col, d->s);
    g->h_out_rowptr.resize(size_t(n) + 1);
    GHET_HIP(hipMemcpyAsync(g->h_out_rowptr.data(), g->out.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, d->s));
    if (want_in) {
        g->h_in_rowptr.resize(size_t(n) + 1);
        GHET_HIP(hipMemcpyAsync(g->h_in_rowptr.data(), g->in.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, d->s));
    }
    if (want_all) {
        g->h_all_rowptr.resize(size_t(n) + 1);
        GHET_HIP(hipMemcpyAsync(g->h_all_rowptr.data(), g->all.rowptr, 4 * (size_t(n) + 1), hipMemcpyDeviceToHost, d->s));
    }
    GHET_HIP(hipStreamSynchronize(d->s));
    *out = gguard.release();
    return GHET_OK;
}

void ghet_dyn_destroy(ghet_dyn* d) { dyn_free(d); }

}  // extern "C"

void ghet::dyn_release(void* d) { dyn_free(static_cast<ghet_dyn*>(d)); }
