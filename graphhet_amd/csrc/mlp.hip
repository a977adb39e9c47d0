// Learned schedule networks: 16 -> 64 -> 64 -> 4 MLPs (ReLU) that pick the
// BFS direction per level (GHET_NET_BFS: logits of push / pull / bitmap) and
// the SSSP bucket-width class per phase (GHET_NET_SSSP: classes 0..3).
//
// No reference counterpart: the reference schedules BFS with a fixed rule
// (GPU_BFS_pre_csr.cuh:91) and SSSP without buckets.
//
// Arithmetic: FP32 master parameters; every forward pass uses their FP16
// rounding (weights and biases), converted exactly to FP32, with FP32 fused
// multiply-adds accumulated in a fixed order (bias first, then inputs in
// index order).  The GPU kernels and the host mirror (sched.cpp, vectorised
// with x86 FMA) perform the same operations in the same order, so both give
// bit-identical logits; the host mirror makes the per-level decisions
// because the level loops are host-driven and a device round trip per level
// would cost more than the 5.4k multiply-adds (well under a microsecond).
//
// Training (ghet_mlp_train_step): one kernel runs forward + backward of
// softmax cross-entropy over 64-sample chunks (weights and activations in
// LDS) and writes per-chunk gradient partials; a second kernel sums the
// partials in chunk order (deterministic), applies Adam to the FP32 master
// and re-rounds the FP16 copy.  The gradient is that of the loss evaluated
// with the FP16 weights, taken as the gradient w.r.t. the master
// (straight-through rounding).
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "graph_internal.hpp"
#include "sched.hpp"

namespace ghet {
namespace {

constexpr int kIn = GHET_MLP_IN, kH = GHET_MLP_H, kOut = GHET_MLP_OUT, kP = GHET_MLP_PARAMS;
// Parameter layout (row-major, output-major): W1[kH][kIn], b1[kH], W2[kH][kH], b2[kH], W3[kOut][kH], b3[kOut].
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kEnd = kB3 + kOut;
static_assert(kEnd == kP, "parameter count");

constexpr int kChunk = 64;  // samples per block
constexpr int kMT = 256;    // threads per block
constexpr int kMaxBatch = 1 << 20;
constexpr float kBeta1 = 0.9f, kBeta2 = 0.999f, kEps = 1e-8f;

__device__ __forceinline__ float relu(float z) { return z > 0.f ? z : 0.f; }

// Forward (mode 0: logits) or forward + backward (mode 1: gradient and loss
// partials of this chunk) for samples [blockIdx.x*kChunk, +kChunk) of B.
__global__ void __launch_bounds__(kMT) k_mlp_chunk(const __half* __restrict__ w16, const float* __restrict__ x,
                                                   const int32_t* __restrict__ labels, int32_t B, int mode,
                                                   float* __restrict__ logits, float* __restrict__ grad_part,
                                                   float* __restrict__ loss_part) {
    __shared__ float W[kP];
    __shared__ float sx[kChunk][kIn];
    __shared__ float h1[kChunk][kH];
    __shared__ float h2[kChunk][kH];
    __shared__ float z3[kChunk][kOut];
    __shared__ float sloss[kChunk];
    const int tid = threadIdx.x;
    const int64_t s0 = static_cast<int64_t>(blockIdx.x) * kChunk;
    const int ns = static_cast<int>(min<int64_t>(kChunk, B - s0));
    for (int p = tid; p < kP; p += kMT) W[p] = __half2float(w16[p]);
    for (int e = tid; e < kChunk * kIn; e += kMT) {
        const int s = e / kIn, k = e % kIn;
        sx[s][k] = s < ns ? x[(s0 + s) * kIn + k] : 0.f;
    }
    __syncthreads();
    for (int e = tid; e < kChunk * kH; e += kMT) {
        const int s = e / kH, j = e % kH;
        float acc = W[kB1 + j];
        for (int k = 0; k < kIn; ++k) acc = fmaf(W[kW1 + j * kIn + k], sx[s][k], acc);
        h1[s][j] = relu(acc);
    }
    __syncthreads();
    for (int e = tid; e < kChunk * kH; e += kMT) {
        const int s = e / kH, j = e % kH;
        float acc = W[kB2 + j];
        for (int i = 0; i < kH; ++i) acc = fmaf(W[kW2 + j * kH + i], h1[s][i], acc);
        h2[s][j] = relu(acc);
    }
    __syncthreads();
    for (int e = tid; e < kChunk * kOut; e += kMT) {
        const int s = e / kOut, o = e % kOut;
        float acc = W[kB3 + o];
        for (int j = 0; j < kH; ++j) acc = fmaf(W[kW3 + o * kH + j], h2[s][j], acc);
        z3[s][o] = acc;
        if (mode == 0 && s < ns) logits[(s0 + s) * kOut + o] = acc;
    }
    if (mode == 0) return;
    __syncthreads();
    // softmax cross-entropy, mean over the whole batch: dz3 overwrites z3
    const float invB = 1.f / static_cast<float>(B);
    if (tid < kChunk) {
        const int s = tid;
        float loss = 0.f;
        if (s < ns) {
            const int y = labels[s0 + s];
            float mx = z3[s][0];
            for (int o = 1; o < kOut; ++o) mx = fmaxf(mx, z3[s][o]);
            float e[kOut], sum = 0.f;
            for (int o = 0; o < kOut; ++o) {
                e[o] = expf(z3[s][o] - mx);
                sum += e[o];
            }
            loss = logf(sum) + mx - z3[s][y];
            for (int o = 0; o < kOut; ++o) z3[s][o] = (e[o] / sum - (o == y ? 1.f : 0.f)) * invB;
        } else {
            for (int o = 0; o < kOut; ++o) z3[s][o] = 0.f;
        }
        sloss[s] = loss;
    }
    __syncthreads();
    float* gp = grad_part + static_cast<int64_t>(blockIdx.x) * kP;
    if (tid == 0) {
        float l = 0.f;
        for (int s = 0; s < kChunk; ++s) l += sloss[s];
        loss_part[blockIdx.x] = l;
    }
    // layer 3: dW3[o][j] = sum_s dz3[s][o] h2[s][j]; db3[o] = sum_s dz3[s][o]
    for (int e = tid; e < kOut * kH + kOut; e += kMT) {
        float acc = 0.f;
        if (e < kOut * kH) {
            const int o = e / kH, j = e % kH;
            for (int s = 0; s < kChunk; ++s) acc = fmaf(z3[s][o], h2[s][j], acc);
            gp[kW3 + e] = acc;
        } else {
            const int o = e - kOut * kH;
            for (int s = 0; s < kChunk; ++s) acc += z3[s][o];
            gp[kB3 + o] = acc;
        }
    }
    __syncthreads();
    // dz2 = (W3^T dz3) * relu'(z2), in place over h2
    for (int e = tid; e < kChunk * kH; e += kMT) {
        const int s = e / kH, j = e % kH;
        float acc = 0.f;
        for (int o = 0; o < kOut; ++o) acc = fmaf(W[kW3 + o * kH + j], z3[s][o], acc);
        h2[s][j] = h2[s][j] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // layer 2: dW2[j][i] = sum_s dz2[s][j] h1[s][i]; db2[j] = sum_s dz2[s][j]
    for (int e = tid; e < kH * kH + kH; e += kMT) {
        float acc = 0.f;
        if (e < kH * kH) {
            const int j = e / kH, i = e % kH;
            for (int s = 0; s < kChunk; ++s) acc = fmaf(h2[s][j], h1[s][i], acc);
            gp[kW2 + e] = acc;
        } else {
            const int j = e - kH * kH;
            for (int s = 0; s < kChunk; ++s) acc += h2[s][j];
            gp[kB2 + j] = acc;
        }
    }
    __syncthreads();
    // dz1 = (W2^T dz2) * relu'(z1), in place over h1
    for (int e = tid; e < kChunk * kH; e += kMT) {
        const int s = e / kH, i = e % kH;
        float acc = 0.f;
        for (int j = 0; j < kH; ++j) acc = fmaf(W[kW2 + j * kH + i], h2[s][j], acc);
        h1[s][i] = h1[s][i] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // layer 1: dW1[i][k] = sum_s dz1[s][i] x[s][k]; db1[i] = sum_s dz1[s][i]
    for (int e = tid; e < kH * kIn + kH; e += kMT) {
        float acc = 0.f;
        if (e < kH * kIn) {
            const int i = e / kIn, k = e % kIn;
            for (int s = 0; s < kChunk; ++s) acc = fmaf(h1[s][i], sx[s][k], acc);
            gp[kW1 + e] = acc;
        } else {
            const int i = e - kH * kIn;
            for (int s = 0; s < kChunk; ++s) acc += h1[s][i];
            gp[kB1 + i] = acc;
        }
    }
}

// Sums the chunk partials in chunk order, then one Adam step on the master
// and the FP16 re-rounding.  grad (optional) receives the summed gradient.
__global__ void __launch_bounds__(256) k_mlp_adam(const float* __restrict__ grad_part, int32_t chunks,
                                                  float* __restrict__ master, float* __restrict__ m1,
                                                  float* __restrict__ m2, __half* __restrict__ w16,
                                                  float* __restrict__ grad, float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kP) return;
    float g = 0.f;
    for (int c = 0; c < chunks; ++c) g += grad_part[static_cast<int64_t>(c) * kP + p];
    if (grad) grad[p] = g;
    const float a = kBeta1 * m1[p] + (1.f - kBeta1) * g;
    const float b = kBeta2 * m2[p] + (1.f - kBeta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    const float w = master[p] - lr * (a / bc1) / (sqrtf(b / bc2) + kEps);
    master[p] = w;
    w16[p] = __float2half_rn(w);
}

__global__ void k_mlp_quantize(const float* __restrict__ master, __half* __restrict__ w16) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kP) w16[p] = __float2half_rn(master[p]);
}

// Exact FP16 -> FP32 (the host mirror's weights).
float half_bits_to_float(uint16_t h) {
    const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t bits;
    if (ex == 0) {
        if (man == 0) bits = sign;
        else {  // subnormal: normalise
            int e = -1;
            uint32_t mm = man;
            do {
                ++e;
                mm <<= 1;
            } while ((mm & 0x400u) == 0);
            bits = sign | (static_cast<uint32_t>(127 - 15 - e) << 23) | ((mm & 0x3ffu) << 13);
        }
    } else if (ex == 0x1f) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((ex + 127 - 15) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct Net {
    bool installed = false;
    int device = -1;
    float *master = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    __half* w16 = nullptr;
    int64_t step = 0;
    // staging for batches
    float *x = nullptr, *out = nullptr, *gpart = nullptr, *lpart = nullptr;
    int32_t* lab = nullptr;
    int32_t cap = 0;  // samples the staging holds
    bool host_dirty = true;
    std::vector<float> hw;  // FP16 parameters as FP32, transposed layout (host mirror)
};

Net g_nets[2];
std::mutex g_mu;

void net_free_device(Net& n) {
    for (void* p : {static_cast<void*>(n.master), static_cast<void*>(n.m1), static_cast<void*>(n.m2),
                    static_cast<void*>(n.grad), static_cast<void*>(n.w16), static_cast<void*>(n.x),
                    static_cast<void*>(n.out), static_cast<void*>(n.gpart), static_cast<void*>(n.lpart),
                    static_cast<void*>(n.lab)})
        if (p) (void)hipFree(p);
    n = Net{};
}

int net_alloc(Net& n) {
    int dev = 0;
    GHET_HIP(hipGetDevice(&dev));
    if (n.master && n.device == dev) return GHET_OK;
    if (n.master) net_free_device(n);
    n.device = dev;
    GHET_HIP(hipMalloc(&n.master, 4 * kP));
    GHET_HIP(hipMalloc(&n.m1, 4 * kP));
    GHET_HIP(hipMalloc(&n.m2, 4 * kP));
    GHET_HIP(hipMalloc(&n.grad, 4 * kP));
    GHET_HIP(hipMalloc(&n.w16, 2 * kP));
    return GHET_OK;
}

int net_staging(Net& n, int32_t batch) {
    if (batch <= n.cap) return GHET_OK;
    for (void* p : {static_cast<void*>(n.x), static_cast<void*>(n.out), static_cast<void*>(n.gpart),
                    static_cast<void*>(n.lpart), static_cast<void*>(n.lab)})
        if (p) (void)hipFree(p);
    n.x = n.out = n.gpart = n.lpart = nullptr;
    n.lab = nullptr;
    n.cap = 0;
    const int64_t chunks = ceil_div(batch, kChunk);
    GHET_HIP(hipMalloc(&n.x, 4 * size_t(batch) * kIn));
    GHET_HIP(hipMalloc(&n.out, 4 * size_t(batch) * kOut));
    GHET_HIP(hipMalloc(&n.lab, 4 * size_t(batch)));
    GHET_HIP(hipMalloc(&n.gpart, 4 * size_t(chunks) * kP));
    GHET_HIP(hipMalloc(&n.lpart, 4 * size_t(chunks)));
    n.cap = batch;
    return GHET_OK;
}

int net_refresh_host(Net& n) {
    if (!n.host_dirty) return GHET_OK;
    std::vector<uint16_t> h(kP);
    GHET_HIP(hipMemcpy(h.data(), n.w16, 2 * kP, hipMemcpyDeviceToHost));
    std::vector<float> w(kP);
    for (int p = 0; p < kP; ++p) w[p] = half_bits_to_float(h[p]);
    n.hw.resize(kP);
    mlp_transpose(w.data(), n.hw.data());
    n.host_dirty = false;
    return GHET_OK;
}

int net_install(Net& n, const float* params) {
    for (int p = 0; p < kP; ++p)
        if (!std::isfinite(params[p])) return set_error(GHET_ERR_INVALID, "parameter %d is not finite", p);
    GHET_CHECK(net_alloc(n));
    GHET_HIP(hipMemcpy(n.master, params, 4 * kP, hipMemcpyHostToDevice));
    GHET_HIP(hipMemset(n.m1, 0, 4 * kP));
    GHET_HIP(hipMemset(n.m2, 0, 4 * kP));
    hipLaunchKernelGGL(k_mlp_quantize, dim3(ceil_div(kP, 256)), dim3(256), 0, nullptr, n.master, n.w16);
    GHET_HIP(hipGetLastError());
    GHET_HIP(hipDeviceSynchronize());
    n.step = 0;
    n.installed = true;
    n.host_dirty = true;
    return net_refresh_host(n);
}

bool valid_kind(int kind) { return kind == GHET_NET_BFS || kind == GHET_NET_SSSP; }

}  // namespace

bool net_available(int kind) {
    if (!valid_kind(kind)) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    Net& n = g_nets[kind];
    if (!n.installed) return false;
    return net_refresh_host(n) == GHET_OK;
}

void net_forward_host(int kind, const float* x, float* y) {
    std::lock_guard<std::mutex> lk(g_mu);
    mlp_forward_host_t(g_nets[kind].hw.data(), x, y);
}

void mlp_release() {
    std::lock_guard<std::mutex> lk(g_mu);
    for (Net& n : g_nets) net_free_device(n);
}

}  // namespace ghet

using namespace ghet;

extern "C" {

int ghet_sched_set_params(int kind, const float* params) {
    clear_error();
    if (!valid_kind(kind)) return set_error(GHET_ERR_INVALID, "unknown network kind %d", kind);
    if (!params) return set_error(GHET_ERR_INVALID, "params is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    return net_install(g_nets[kind], params);
}

int ghet_sched_get_params(int kind, float* params) {
    clear_error();
    if (!valid_kind(kind)) return set_error(GHET_ERR_INVALID, "unknown network kind %d", kind);
    if (!params) return set_error(GHET_ERR_INVALID, "params is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    Net& n = g_nets[kind];
    if (!n.installed) return set_error(GHET_ERR_INVALID, "network %d is not installed", kind);
    GHET_HIP(hipMemcpy(params, n.master, 4 * kP, hipMemcpyDeviceToHost));
    return GHET_OK;
}

int ghet_sched_clear(void) {
    clear_error();
    std::lock_guard<std::mutex> lk(g_mu);
    for (Net& n : g_nets) {
        if (n.master) (void)hipSetDevice(n.device);
        net_free_device(n);
    }
    return GHET_OK;
}

int ghet_sched_save(const char* path) {
    clear_error();
    if (!path) return set_error(GHET_ERR_INVALID, "path is NULL");
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<float> buf(kP);
    std::string text = "ghet-sched 1\n";
    int saved = 0;
    for (int kind = 0; kind < 2; ++kind) {
        Net& n = g_nets[kind];
        if (!n.installed) continue;
        GHET_HIP(hipMemcpy(buf.data(), n.master, 4 * kP, hipMemcpyDeviceToHost));
        char line[64];
        std::snprintf(line, sizeof line, "net %d %d\n", kind, kP);
        text += line;
        for (int p = 0; p < kP; ++p) {
            std::snprintf(line, sizeof line, "%.9g\n", static_cast<double>(buf[p]));
            text += line;
        }
        ++saved;
    }
    if (!saved) return set_error(GHET_ERR_INVALID, "no network installed");
    FILE* f = std::fopen(path, "w");
    if (!f) return set_error(GHET_ERR_INVALID, "cannot write %s", path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) return set_error(GHET_ERR_INVALID, "write to %s failed", path);
    return GHET_OK;
}

// Missing or malformed file: nothing changes, the heuristic stays in charge
// and GHET_WARN_FALLBACK is returned (results never depend on the network).
int ghet_sched_load(const char* path) {
    clear_error();
    if (!path) return set_error(GHET_ERR_INVALID, "path is NULL");
    FILE* f = std::fopen(path, "r");
    if (!f) {
        set_error(GHET_WARN_FALLBACK, "cannot open %s: heuristic schedule in use", path);
        return GHET_WARN_FALLBACK;
    }
    std::vector<std::vector<float>> nets(2);
    char magic[32] = {};
    int version = 0;
    bool ok = std::fscanf(f, "%31s %d", magic, &version) == 2 && std::strcmp(magic, "ghet-sched") == 0 && version == 1;
    while (ok) {
        char tag[8] = {};
        int kind = -1, count = 0;
        const int r = std::fscanf(f, "%7s %d %d", tag, &kind, &count);
        if (r == EOF) break;
        if (r != 3 || std::strcmp(tag, "net") != 0 || !valid_kind(kind) || count != kP || !nets[kind].empty()) {
            ok = false;
            break;
        }
        nets[kind].resize(kP);
        for (int p = 0; p < kP && ok; ++p) ok = std::fscanf(f, "%f", &nets[kind][p]) == 1 && std::isfinite(nets[kind][p]);
    }
    std::fclose(f);
    if (!ok || (nets[0].empty() && nets[1].empty())) {
        set_error(GHET_WARN_FALLBACK, "%s is not a schedule file: heuristic schedule in use", path);
        return GHET_WARN_FALLBACK;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    for (int kind = 0; kind < 2; ++kind)
        if (!nets[kind].empty()) GHET_CHECK(net_install(g_nets[kind], nets[kind].data()));
    return GHET_OK;
}

int ghet_mlp_forward(int kind, int32_t batch, const float* x, float* logits) {
    clear_error();
    if (!valid_kind(kind)) return set_error(GHET_ERR_INVALID, "unknown network kind %d", kind);
    if (batch < 0 || batch > kMaxBatch) return set_error(GHET_ERR_INVALID, "batch %d outside [0,%d]", batch, kMaxBatch);
    if (batch > 0 && (!x || !logits)) return set_error(GHET_ERR_INVALID, "NULL buffer");
    std::lock_guard<std::mutex> lk(g_mu);
    Net& n = g_nets[kind];
    if (!n.installed) return set_error(GHET_ERR_INVALID, "network %d is not installed", kind);
    if (batch == 0) return GHET_OK;
    GHET_HIP(hipSetDevice(n.device));
    GHET_CHECK(net_staging(n, batch));
    GHET_HIP(hipMemcpy(n.x, x, 4 * size_t(batch) * kIn, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_chunk, dim3(static_cast<unsigned>(ceil_div(batch, kChunk))), dim3(kMT), 0, nullptr, n.w16,
                       n.x, nullptr, batch, 0, n.out, nullptr, nullptr);
    GHET_HIP(hipGetLastError());
    GHET_HIP(hipMemcpy(logits, n.out, 4 * size_t(batch) * kOut, hipMemcpyDeviceToHost));
    return GHET_OK;
}

int ghet_mlp_forward_host(int kind, int32_t batch, const float* x, float* logits) {
    clear_error();
    if (!valid_kind(kind)) return set_error(GHET_ERR_INVALID, "unknown network kind %d", kind);
    if (batch < 0) return set_error(GHET_ERR_INVALID, "batch %d < 0", batch);
    if (batch > 0 && (!x || !logits)) return set_error(GHET_ERR_INVALID, "NULL buffer");
    std::lock_guard<std::mutex> lk(g_mu);
    Net& n = g_nets[kind];
    if (!n.installed) return set_error(GHET_ERR_INVALID, "network %d is not installed", kind);
    GHET_CHECK(net_refresh_host(n));
    for (int32_t b = 0; b < batch; ++b) mlp_forward_host_t(n.hw.data(), x + int64_t(b) * kIn, logits + int64_t(b) * kOut);
    return GHET_OK;
}

int ghet_mlp_train_step(int kind, int32_t batch, const float* x, const int32_t* labels, float lr, float* loss,
                        float* grad_out) {
    clear_error();
    if (!valid_kind(kind)) return set_error(GHET_ERR_INVALID, "unknown network kind %d", kind);
    if (batch <= 0 || batch > kMaxBatch) return set_error(GHET_ERR_INVALID, "batch %d outside [1,%d]", batch, kMaxBatch);
    if (!x || !labels) return set_error(GHET_ERR_INVALID, "NULL buffer");
    if (!(lr >= 0.f) || !std::isfinite(lr)) return set_error(GHET_ERR_INVALID, "bad learning rate");
    for (int32_t b = 0; b < batch; ++b)
        if (labels[b] < 0 || labels[b] >= kOut) return set_error(GHET_ERR_INVALID, "label %d outside [0,%d)", labels[b], kOut);
    std::lock_guard<std::mutex> lk(g_mu);
    Net& n = g_nets[kind];
    if (!n.installed) return set_error(GHET_ERR_INVALID, "network %d is not installed", kind);
    GHET_HIP(hipSetDevice(n.device));
    GHET_CHECK(net_staging(n, batch));
    const int32_t chunks = static_cast<int32_t>(ceil_div(batch, kChunk));
    GHET_HIP(hipMemcpy(n.x, x, 4 * size_t(batch) * kIn, hipMemcpyHostToDevice));
    GHET_HIP(hipMemcpy(n.lab, labels, 4 * size_t(batch), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_chunk, dim3(chunks), dim3(kMT), 0, nullptr, n.w16, n.x, n.lab, batch, 1, nullptr, n.gpart,
                       n.lpart);
    ++n.step;
    const float bc1 = 1.f - std::pow(kBeta1, static_cast<float>(n.step));
    const float bc2 = 1.f - std::pow(kBeta2, static_cast<float>(n.step));
    hipLaunchKernelGGL(k_mlp_adam, dim3(ceil_div(kP, 256)), dim3(256), 0, nullptr, n.gpart, chunks, n.master, n.m1,
                       n.m2, n.w16, n.grad, lr, bc1, bc2);
    GHET_HIP(hipGetLastError());
    std::vector<float> lp(static_cast<size_t>(chunks));
    GHET_HIP(hipMemcpy(lp.data(), n.lpart, 4 * size_t(chunks), hipMemcpyDeviceToHost));
    if (grad_out) GHET_HIP(hipMemcpy(grad_out, n.grad, 4 * kP, hipMemcpyDeviceToHost));
    if (loss) {
        double t = 0;
        for (float v : lp) t += v;
        *loss = static_cast<float>(t / batch);
    }
    n.host_dirty = true;
    return GHET_OK;
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

}  // extern "C"
