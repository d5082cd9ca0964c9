// mlp.hip -- the small MLP that predicts the Delta-stepping bucket width and the
// mapping regime of a graph, with hand-written HIP inference, backward and Adam
// kernels.  No reference counterpart (the reference only has the Davidson
// default, src/gpu_sssp.cu:202-206).
//
// Network: in -> H -> H -> out, ReLU hidden layers, linear output; FP32.
// Parameter layout (one flat float array):
//   W1[H][in], b1[H], W2[H][H], b2[H], W3[out][H], b3[out]
// Loss: masked mean squared error  L = sum_ij m_ij (yhat_ij - y_ij)^2 / sum_ij m_ij.
//
// Training runs three kernels per step: a per-sample forward+backward kernel
// (one thread per sample, writes that sample's gradient), a deterministic
// per-parameter reduction over samples (fixed summation order, so gradients do
// not depend on scheduling), and an Adam update.  The layers are a few
// thousand multiply-adds per sample: scalar FP32 on vector ALUs, no MFMA.
#include "../../include/pf_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int MAXW = 64;   // widest layer supported by the per-thread kernels

thread_local std::string g_mlpError;

int mfail(int code, const std::string& m) {
    g_mlpError = m;
    return code;
}

#define MLP_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return mfail(PF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Dims {
    int in, h, out;
    __host__ __device__ int w1() const { return 0; }
    __host__ __device__ int b1() const { return h * in; }
    __host__ __device__ int w2() const { return b1() + h; }
    __host__ __device__ int b2() const { return w2() + h * h; }
    __host__ __device__ int w3() const { return b2() + h; }
    __host__ __device__ int b3() const { return w3() + out * h; }
    __host__ __device__ int count() const { return b3() + out; }
};

// Forward of one sample into registers; returns activations for backward.
__device__ __forceinline__ void forward1(const Dims d, const float* p, const float* x, float* h1, float* h2, float* y) {
#pragma clang fp contract(fast)
    for (int j = 0; j < d.h; ++j) {
        float a = p[d.b1() + j];
        for (int i = 0; i < d.in; ++i) a += p[d.w1() + j * d.in + i] * x[i];
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < d.h; ++j) {
        float a = p[d.b2() + j];
        for (int i = 0; i < d.h; ++i) a += p[d.w2() + j * d.h + i] * h1[i];
        h2[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < d.out; ++j) {
        float a = p[d.b3() + j];
        for (int i = 0; i < d.h; ++i) a += p[d.w3() + j * d.h + i] * h2[i];
        y[j] = a;
    }
}

__global__ void forwardKernel(Dims d, const float* p, const float* X, int n, float* Y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float h1[MAXW], h2[MAXW], y[MAXW];
    forward1(d, p, X + (size_t)s * d.in, h1, h2, y);
    for (int j = 0; j < d.out; ++j) Y[(size_t)s * d.out + j] = y[j];
}

// Per-sample gradient of sum_j m_j (y_j - t_j)^2 (unnormalised); also the sample's loss terms.
__global__ void fwdBwdKernel(Dims d, const float* p, const float* X, const float* T, const float* M, int n, float* G,
                             float* lossTerms) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* x = X + (size_t)s * d.in;
    float h1[MAXW], h2[MAXW], y[MAXW], dy[MAXW], d2[MAXW], d1[MAXW];
    forward1(d, p, x, h1, h2, y);
    float* g = G + (size_t)s * d.count();
    float l = 0.f;
    for (int j = 0; j < d.out; ++j) {
        const float m = M[(size_t)s * d.out + j];
        const float e = y[j] - T[(size_t)s * d.out + j];
        l += m * e * e;
        dy[j] = 2.f * m * e;
    }
    lossTerms[s] = l;
    // layer 3
    for (int j = 0; j < d.out; ++j) {
        g[d.b3() + j] = dy[j];
        for (int i = 0; i < d.h; ++i) g[d.w3() + j * d.h + i] = dy[j] * h2[i];
    }
    for (int i = 0; i < d.h; ++i) {
        float a = 0.f;
        for (int j = 0; j < d.out; ++j) a += p[d.w3() + j * d.h + i] * dy[j];
        d2[i] = h2[i] > 0.f ? a : 0.f;
    }
    // layer 2
    for (int j = 0; j < d.h; ++j) {
        g[d.b2() + j] = d2[j];
        for (int i = 0; i < d.h; ++i) g[d.w2() + j * d.h + i] = d2[j] * h1[i];
    }
    for (int i = 0; i < d.h; ++i) {
        float a = 0.f;
        for (int j = 0; j < d.h; ++j) a += p[d.w2() + j * d.h + i] * d2[j];
        d1[i] = h1[i] > 0.f ? a : 0.f;
    }
    // layer 1
    for (int j = 0; j < d.h; ++j) {
        g[d.b1() + j] = d1[j];
        for (int i = 0; i < d.in; ++i) g[d.w1() + j * d.in + i] = d1[j] * x[i];
    }
}

// grad[q] = scale * sum_s G[s][q], samples summed in index order.
__global__ void reduceKernel(const float* G, int n, int P, float scale, float* grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    float a = 0.f;
    for (int s = 0; s < n; ++s) a += G[(size_t)s * P + q];
    grad[q] = a * scale;
}

__global__ void sumKernel(const float* v, int n, float* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float a = 0.f;
        for (int i = 0; i < n; ++i) a += v[i];
        *out = a;
    }
}

// Adam (Kingma & Ba 2015), bias-corrected.
__global__ void adamKernel(float* p, const float* g, float* m1, float* m2, int P, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const float gq = g[q];
    const float a = b1 * m1[q] + (1.f - b1) * gq;
    const float b = b2 * m2[q] + (1.f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = b;
    p[q] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace

struct pf_mlp {
    Dims d{};
    float* d_p = nullptr;
    float* d_m1 = nullptr;
    float* d_m2 = nullptr;
    float* d_g = nullptr;
    long long step = 0;
};

extern "C" {

const char* pf_mlp_last_error(void) { return g_mlpError.c_str(); }

int pf_mlp_create(uint32_t n_in, uint32_t n_hidden, uint32_t n_out, uint64_t seed, pf_mlp** out) {
    *out = nullptr;
    if (!n_in || !n_hidden || !n_out || n_in > MAXW || n_hidden > MAXW || n_out > MAXW)
        return mfail(PF_ERR_INVALID, "layer widths must be in [1, 64]");
    pf_mlp* m = new pf_mlp();
    m->d = Dims{(int)n_in, (int)n_hidden, (int)n_out};
    const int P = m->d.count();
    // He-uniform initialisation from a SplitMix64 stream, biases zero
    std::vector<float> p(P, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int cnt, int fanin) {
        const double lim = std::sqrt(6.0 / fanin);
        for (int i = 0; i < cnt; ++i) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(m->d.w1(), m->d.h * m->d.in, m->d.in);
    fill(m->d.w2(), m->d.h * m->d.h, m->d.h);
    fill(m->d.w3(), m->d.out * m->d.h, m->d.h);
    if (hipMalloc((void**)&m->d_p, P * 4) != hipSuccess || hipMalloc((void**)&m->d_m1, P * 4) != hipSuccess ||
        hipMalloc((void**)&m->d_m2, P * 4) != hipSuccess || hipMalloc((void**)&m->d_g, P * 4) != hipSuccess ||
        hipMemcpy(m->d_p, p.data(), P * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->d_m1, 0, P * 4) != hipSuccess || hipMemset(m->d_m2, 0, P * 4) != hipSuccess) {
        (void)hipGetLastError();
        hipFree(m->d_p); hipFree(m->d_m1); hipFree(m->d_m2); hipFree(m->d_g);
        delete m;
        return mfail(PF_ERR_NO_DEVICE, "MLP device allocation failed (no GPU?)");
    }
    *out = m;
    return PF_OK;
}

void pf_mlp_destroy(pf_mlp* m) {
    if (!m) return;
    hipFree(m->d_p); hipFree(m->d_m1); hipFree(m->d_m2); hipFree(m->d_g);
    delete m;
}

uint32_t pf_mlp_num_params(const pf_mlp* m) { return (uint32_t)m->d.count(); }

int pf_mlp_get_params(pf_mlp* m, float* params) {
    MLP_HIP(hipMemcpy(params, m->d_p, m->d.count() * 4, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_mlp_set_params(pf_mlp* m, const float* params) {
    MLP_HIP(hipMemcpy(m->d_p, params, m->d.count() * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemset(m->d_m1, 0, m->d.count() * 4));
    MLP_HIP(hipMemset(m->d_m2, 0, m->d.count() * 4));
    m->step = 0;
    return PF_OK;
}

int pf_mlp_forward(pf_mlp* m, const float* x, uint32_t n, float* y) {
    if (!n) return PF_OK;
    float *dx = nullptr, *dy = nullptr;
    MLP_HIP(hipMalloc((void**)&dx, (size_t)n * m->d.in * 4));
    MLP_HIP(hipMalloc((void**)&dy, (size_t)n * m->d.out * 4));
    hipError_t e = hipMemcpy(dx, x, (size_t)n * m->d.in * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(forwardKernel, dim3((n + 63) / 64), dim3(64), 0, 0, m->d, m->d_p, dx, (int)n, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * m->d.out * 4, hipMemcpyDeviceToHost);
    hipFree(dx);
    hipFree(dy);
    if (e != hipSuccess) return mfail(PF_ERR_HIP, hipGetErrorString(e));
    return PF_OK;
}

namespace {
struct Batch {
    float *X = nullptr, *T = nullptr, *M = nullptr, *G = nullptr, *L = nullptr, *loss = nullptr;
    float msum = 0.f;
    ~Batch() { hipFree(X); hipFree(T); hipFree(M); hipFree(G); hipFree(L); hipFree(loss); }
};

int upload(pf_mlp* m, const float* x, const float* t, const float* mask, uint32_t n, Batch& b) {
    const size_t ni = (size_t)n * m->d.in, no = (size_t)n * m->d.out;
    MLP_HIP(hipMalloc((void**)&b.X, ni * 4));
    MLP_HIP(hipMalloc((void**)&b.T, no * 4));
    MLP_HIP(hipMalloc((void**)&b.M, no * 4));
    MLP_HIP(hipMalloc((void**)&b.G, (size_t)n * m->d.count() * 4));
    MLP_HIP(hipMalloc((void**)&b.L, (size_t)n * 4));
    MLP_HIP(hipMalloc((void**)&b.loss, 4));
    MLP_HIP(hipMemcpy(b.X, x, ni * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(b.T, t, no * 4, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(b.M, mask, no * 4, hipMemcpyHostToDevice));
    double s = 0.0;
    for (size_t i = 0; i < no; ++i) s += mask[i];
    b.msum = (float)std::max(s, 1e-12);
    return PF_OK;
}

int lossGrad(pf_mlp* m, Batch& b, uint32_t n) {
    const int P = m->d.count();
    hipLaunchKernelGGL(fwdBwdKernel, dim3((n + 63) / 64), dim3(64), 0, 0, m->d, m->d_p, b.X, b.T, b.M, (int)n, b.G, b.L);
    hipLaunchKernelGGL(reduceKernel, dim3((P + 255) / 256), dim3(256), 0, 0, b.G, (int)n, P, 1.f / b.msum, m->d_g);
    hipLaunchKernelGGL(sumKernel, dim3(1), dim3(64), 0, 0, b.L, (int)n, b.loss);
    MLP_HIP(hipGetLastError());
    return PF_OK;
}
}  // namespace

int pf_mlp_loss_grad(pf_mlp* m, const float* x, const float* t, const float* mask, uint32_t n, float* loss,
                     float* grad) {
    if (!n) return mfail(PF_ERR_INVALID, "empty batch");
    Batch b;
    int rc = upload(m, x, t, mask, n, b);
    if (rc) return rc;
    rc = lossGrad(m, b, n);
    if (rc) return rc;
    float l = 0.f;
    MLP_HIP(hipMemcpy(&l, b.loss, 4, hipMemcpyDeviceToHost));
    *loss = l / b.msum;
    if (grad) MLP_HIP(hipMemcpy(grad, m->d_g, m->d.count() * 4, hipMemcpyDeviceToHost));
    return PF_OK;
}

int pf_mlp_train(pf_mlp* m, const float* x, const float* t, const float* mask, uint32_t n, uint32_t epochs, float lr,
                 float* final_loss) {
    if (!n) return mfail(PF_ERR_INVALID, "empty batch");
    Batch b;
    int rc = upload(m, x, t, mask, n, b);
    if (rc) return rc;
    const int P = m->d.count();
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t e = 0; e < epochs; ++e) {
        rc = lossGrad(m, b, n);
        if (rc) return rc;
        m->step++;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(adamKernel, dim3((P + 255) / 256), dim3(256), 0, 0, m->d_p, m->d_g, m->d_m1, m->d_m2, P, lr,
                           b1, b2, eps, c1, c2);
        MLP_HIP(hipGetLastError());
    }
    rc = lossGrad(m, b, n);
    if (rc) return rc;
    float l = 0.f;
    MLP_HIP(hipMemcpy(&l, b.loss, 4, hipMemcpyDeviceToHost));
    if (final_loss) *final_loss = l / b.msum;
    return PF_OK;
}

}  // extern "C"
