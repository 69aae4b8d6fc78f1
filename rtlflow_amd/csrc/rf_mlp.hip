// rf_mlp.hip -- training and inference kernels of the learned mtask cost model.
//
// F -> 64 -> 64 -> 1, ReLU, FP32, mean squared error, Adam. The network is
// tiny (4,673 parameters for F = 6), so the kernels are organised for
// determinism and simplicity rather than matrix cores:
//   * forward / backward: one thread per sample, parameters staged in LDS;
//   * gradient reduction: one thread per parameter, fixed-order sum over the
//     batch (no atomics: two runs give bit-identical gradients, which the CPU
//     FP64 gradient check relies on);
//   * loss: one block, fixed-order tree reduction;
//   * Adam: one thread per parameter.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rf_mlp.h"

namespace {

constexpr int H = RF_MLP_HIDDEN;
constexpr int kMaxF = 16;

thread_local std::string g_mlp_err;

int mfail(const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_mlp_err = buf;
    return -2;
}

#define MLP_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) return mfail("%s: %s", #call, hipGetErrorString(e_));        \
    } while (0)

struct Offsets {
    int w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline Offsets offsets(int F) {
    Offsets o;
    o.w1 = 0;
    o.b1 = o.w1 + H * F;
    o.w2 = o.b1 + H;
    o.b2 = o.w2 + H * H;
    o.w3 = o.b2 + H;
    o.b3 = o.w3 + H;
    o.total = o.b3 + 1;
    return o;
}

// forward pass; h1 / h2 may be null (inference only)
__global__ void __launch_bounds__(256) mlp_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                               int n, int F, float* __restrict__ h1o,
                                               float* __restrict__ h2o, float* __restrict__ y) {
    __shared__ float sp[H * kMaxF + H + H * H + H + H + 1];
    const Offsets o = offsets(F);
    for (int i = threadIdx.x; i < o.total; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[kMaxF];
    for (int k = 0; k < F; ++k) x[k] = X[(size_t)s * F + k];
    float h1[H];
    for (int i = 0; i < H; ++i) {
        float a = sp[o.b1 + i];
        for (int k = 0; k < F; ++k) a = fmaf(sp[o.w1 + i * F + k], x[k], a);
        h1[i] = a > 0.f ? a : 0.f;
        if (h1o) h1o[(size_t)s * H + i] = h1[i];
    }
    float out = sp[o.b3];
    for (int j = 0; j < H; ++j) {
        float a = sp[o.b2 + j];
        for (int i = 0; i < H; ++i) a = fmaf(sp[o.w2 + j * H + i], h1[i], a);
        const float r = a > 0.f ? a : 0.f;
        if (h2o) h2o[(size_t)s * H + j] = r;
        out = fmaf(sp[o.w3 + j], r, out);
    }
    y[s] = out;
}

// backward to the hidden layers: dy = 2 (y - t) / n, dh2, dh1 (pre-activation)
__global__ void __launch_bounds__(256) mlp_bwd(const float* __restrict__ P, const float* __restrict__ T,
                                               int n, int F, const float* __restrict__ h1,
                                               const float* __restrict__ h2, const float* __restrict__ y,
                                               float* __restrict__ dy, float* __restrict__ dh1,
                                               float* __restrict__ dh2) {
    __shared__ float sp[H * H + H];
    const Offsets o = offsets(F);
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) sp[i] = P[o.w2 + i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) sp[H * H + i] = P[o.w3 + i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float g = 2.f * (y[s] - T[s]) / (float)n;
    dy[s] = g;
    float d2[H];
    for (int j = 0; j < H; ++j) {
        d2[j] = h2[(size_t)s * H + j] > 0.f ? sp[H * H + j] * g : 0.f;
        dh2[(size_t)s * H + j] = d2[j];
    }
    for (int i = 0; i < H; ++i) {
        float a = 0.f;
        for (int j = 0; j < H; ++j) a = fmaf(sp[j * H + i], d2[j], a);
        dh1[(size_t)s * H + i] = h1[(size_t)s * H + i] > 0.f ? a : 0.f;
    }
}

// one thread per parameter: fixed-order sum over the batch
__global__ void __launch_bounds__(256) mlp_grad(const float* __restrict__ X, int n, int F,
                                                const float* __restrict__ h1, const float* __restrict__ h2,
                                                const float* __restrict__ dy, const float* __restrict__ dh1,
                                                const float* __restrict__ dh2, float* __restrict__ G) {
    const Offsets o = offsets(F);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= o.total) return;
    float acc = 0.f;
    if (p < o.b1) {
        const int i = (p - o.w1) / F, k = (p - o.w1) % F;
        for (int s = 0; s < n; ++s) acc = fmaf(dh1[(size_t)s * H + i], X[(size_t)s * F + k], acc);
    } else if (p < o.w2) {
        const int i = p - o.b1;
        for (int s = 0; s < n; ++s) acc += dh1[(size_t)s * H + i];
    } else if (p < o.b2) {
        const int j = (p - o.w2) / H, i = (p - o.w2) % H;
        for (int s = 0; s < n; ++s) acc = fmaf(dh2[(size_t)s * H + j], h1[(size_t)s * H + i], acc);
    } else if (p < o.w3) {
        const int j = p - o.b2;
        for (int s = 0; s < n; ++s) acc += dh2[(size_t)s * H + j];
    } else if (p < o.b3) {
        const int j = p - o.w3;
        for (int s = 0; s < n; ++s) acc = fmaf(dy[s], h2[(size_t)s * H + j], acc);
    } else {
        for (int s = 0; s < n; ++s) acc += dy[s];
    }
    G[p] = acc;
}

// mean squared error, one block, fixed-order tree reduction
__global__ void __launch_bounds__(256) mlp_loss(const float* __restrict__ y, const float* __restrict__ T,
                                                int n, float* __restrict__ loss) {
    __shared__ float red[256];
    float a = 0.f;
    for (int s = threadIdx.x; s < n; s += 256) {
        const float d = y[s] - T[s];
        a = fmaf(d, d, a);
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

__global__ void __launch_bounds__(256) mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                                                float* __restrict__ M, float* __restrict__ V, int total,
                                                float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const float g = G[p];
    const float m = 0.9f * M[p] + 0.1f * g;
    const float v = 0.999f * V[p] + 0.001f * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

}  // namespace

struct rf_mlp {
    int F = 0;
    Offsets o{};
    hipStream_t stream = nullptr;
    float* P = nullptr;   // parameters
    float* G = nullptr;   // gradients
    float* M = nullptr;   // Adam first moment
    float* V = nullptr;   // Adam second moment
    float* loss = nullptr;
    uint64_t step = 0;
    int cap = 0;          // batch capacity of the work buffers
    float *X = nullptr, *T = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr;
    float *DY = nullptr, *DH1 = nullptr, *DH2 = nullptr;
};

namespace {

int ensure_cap(rf_mlp* m, int n) {
    if (n <= m->cap) return 0;
    float** bufs[] = {&m->X, &m->T, &m->Y, &m->H1, &m->H2, &m->DY, &m->DH1, &m->DH2};
    for (float** b : bufs)
        if (*b) (void)hipFree(*b);
    const size_t sizes[] = {(size_t)n * m->F, (size_t)n, (size_t)n, (size_t)n * H, (size_t)n * H,
                            (size_t)n, (size_t)n * H, (size_t)n * H};
    for (int i = 0; i < 8; ++i) MLP_HIP(hipMalloc(bufs[i], sizes[i] * 4));
    m->cap = n;
    return 0;
}

int run_fwd(rf_mlp* m, const float* x, int n, bool keep) {
    int rc = ensure_cap(m, n);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(m->X, x, (size_t)n * m->F * 4, hipMemcpyHostToDevice, m->stream));
    const int grid = (n + 255) / 256;
    hipLaunchKernelGGL(mlp_fwd, dim3(grid), dim3(256), 0, m->stream, m->P, m->X, n, m->F,
                       keep ? m->H1 : nullptr, keep ? m->H2 : nullptr, m->Y);
    MLP_HIP(hipGetLastError());
    return 0;
}

int run_grad(rf_mlp* m, const float* x, const float* t, int n, float* loss_out) {
    int rc = run_fwd(m, x, n, true);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(m->T, t, (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
    const int grid = (n + 255) / 256;
    hipLaunchKernelGGL(mlp_bwd, dim3(grid), dim3(256), 0, m->stream, m->P, m->T, n, m->F, m->H1, m->H2,
                       m->Y, m->DY, m->DH1, m->DH2);
    hipLaunchKernelGGL(mlp_grad, dim3((m->o.total + 255) / 256), dim3(256), 0, m->stream, m->X, n, m->F,
                       m->H1, m->H2, m->DY, m->DH1, m->DH2, m->G);
    hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(256), 0, m->stream, m->Y, m->T, n, m->loss);
    MLP_HIP(hipGetLastError());
    if (loss_out) {
        MLP_HIP(hipMemcpyAsync(loss_out, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
    }
    return 0;
}

uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

uint32_t rf_mlp_param_count(uint32_t F) { return (uint32_t)offsets((int)F).total; }

const char* rf_mlp_last_error(void) { return g_mlp_err.c_str(); }

int rf_mlp_destroy(rf_mlp* m) {
    if (!m) return 0;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    float* bufs[] = {m->P, m->G, m->M, m->V, m->loss, m->X, m->T, m->Y, m->H1, m->H2, m->DY, m->DH1, m->DH2};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

int rf_mlp_create(uint32_t F, const float* params, uint64_t seed, rf_mlp** out) {
    if (!out || F == 0 || F > (uint32_t)kMaxF) return mfail("rf_mlp_create: bad arguments (F=%u)", F);
    *out = nullptr;
    rf_mlp* m = new rf_mlp();
    m->F = (int)F;
    m->o = offsets((int)F);
    auto bail = [&](int rc) {
        rf_mlp_destroy(m);
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(mfail("hipStreamCreate failed"));
    const size_t bytes = (size_t)m->o.total * 4;
    float** bufs[] = {&m->P, &m->G, &m->M, &m->V};
    for (float** b : bufs) {
        if (hipMalloc(b, bytes) != hipSuccess) return bail(mfail("hipMalloc failed"));
        (void)hipMemsetAsync(*b, 0, bytes, m->stream);
    }
    if (hipMalloc(&m->loss, 4) != hipSuccess) return bail(mfail("hipMalloc failed"));
    std::vector<float> p(m->o.total);
    if (params) {
        memcpy(p.data(), params, bytes);
    } else {
        // He-uniform init from a counter-based generator; biases zero
        auto uni = [&](uint64_t k) { return (float)((mix(seed ^ (k * 0x9E3779B97F4A7C15ull)) >> 11) * 0x1.0p-53) * 2.f - 1.f; };
        const float a1 = std::sqrt(6.f / F), a2 = std::sqrt(6.f / H);
        for (int i = 0; i < m->o.total; ++i) p[i] = 0.f;
        for (int i = 0; i < H * (int)F; ++i) p[m->o.w1 + i] = a1 * uni(i);
        for (int i = 0; i < H * H; ++i) p[m->o.w2 + i] = a2 * uni(100000 + i);
        for (int i = 0; i < H; ++i) p[m->o.w3 + i] = a2 * uni(200000 + i);
    }
    if (hipMemcpyAsync(m->P, p.data(), bytes, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return bail(mfail("parameter upload failed"));
    *out = m;
    return 0;
}

int rf_mlp_get_params(rf_mlp* m, float* out) {
    if (!m || !out) return mfail("null argument");
    MLP_HIP(hipMemcpyAsync(out, m->P, (size_t)m->o.total * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int rf_mlp_set_params(rf_mlp* m, const float* params) {
    if (!m || !params) return mfail("null argument");
    MLP_HIP(hipMemcpyAsync(m->P, params, (size_t)m->o.total * 4, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int rf_mlp_forward(rf_mlp* m, const float* x, uint32_t n, float* y) {
    if (!m || !x || !y) return mfail("null argument");
    if (!n) return 0;
    int rc = run_fwd(m, x, (int)n, false);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(y, m->Y, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int rf_mlp_gradients(rf_mlp* m, const float* x, const float* t, uint32_t n, float* loss, float* grads) {
    if (!m || !x || !t || !grads || !n) return mfail("bad argument");
    int rc = run_grad(m, x, t, (int)n, loss);
    if (rc) return rc;
    MLP_HIP(hipMemcpyAsync(grads, m->G, (size_t)m->o.total * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int rf_mlp_train_step(rf_mlp* m, const float* x, const float* t, uint32_t n, float lr, float* loss) {
    if (!m || !x || !t || !n) return mfail("bad argument");
    int rc = run_grad(m, x, t, (int)n, loss);
    if (rc) return rc;
    m->step++;
    const float bc1 = 1.f - std::pow(0.9f, (float)m->step);
    const float bc2 = 1.f - std::pow(0.999f, (float)m->step);
    hipLaunchKernelGGL(mlp_adam, dim3((m->o.total + 255) / 256), dim3(256), 0, m->stream, m->P, m->G,
                       m->M, m->V, m->o.total, lr, bc1, bc2);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

}  // extern "C"
