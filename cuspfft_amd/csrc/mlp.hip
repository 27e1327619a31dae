// mlp.hip -- the learned plan selector's network: a small MLP (D -> H -> H -> 1,
// ReLU, FP32) with hand-written gfx950 forward, backward and Adam kernels.
//
// It predicts log(GPU time) of a (matrix, plan) pair from features; the
// selector evaluates every candidate plan of a matrix and picks the smallest
// prediction.  Every reduction runs in a fixed order (no atomics), so training
// is deterministic for a given seed and batch order.
//
// Parameters live in one device buffer, in this order:
//   W1 [H x D], b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1]
// Loss: L = 1/(2B) sum_b (yhat_b - y_b)^2 over a batch of B samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <random>
#include <vector>

#include "../../include/spfft_c.h"
#include "spfft_plan_internal.h"

using namespace spfft_internal;

namespace {

constexpr int kT = 256;

// Z[b][n] = sum_k X[b][k] W[n][k] + bias[n]; A = relu(Z) (or Z when !relu)
__global__ void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W,
                            const float* __restrict__ bias, float* __restrict__ Z, float* __restrict__ A,
                            int B, int K, int N, int relu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * N) return;
    const int b = i / N, n = i - b * N;
    const float* x = X + (size_t)b * K;
    const float* w = W + (size_t)n * K;
    float acc = bias[n];
    for (int k = 0; k < K; k++) acc = fmaf(x[k], w[k], acc);
    Z[i] = acc;
    if (A) A[i] = relu ? fmaxf(acc, 0.f) : acc;
}

// dY[b] = (yhat[b] - y[b]) / B; loss[0] = sum (yhat - y)^2 / (2B)  (one block)
__global__ void k_out_grad(const float* __restrict__ yhat, const float* __restrict__ y, float* __restrict__ dy,
                           float* __restrict__ loss, int B) {
    __shared__ float s[kT];
    float acc = 0.f;
    for (int b = threadIdx.x; b < B; b += kT) {
        const float d = yhat[b] - y[b];
        dy[b] = d / (float)B;
        acc += d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = s[0] / (2.f * (float)B);
}

// dW[n][k] = sum_b dZ[b][n] Aprev[b][k]; db[n] = sum_b dZ[b][n]
__global__ void k_dense_bwd_w(const float* __restrict__ dZ, const float* __restrict__ Ap, float* __restrict__ dW,
                              float* __restrict__ db, int B, int K, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * (K + 1)) return;
    const int n = i / (K + 1), k = i - n * (K + 1);
    float acc = 0.f;
    if (k < K) {
        for (int b = 0; b < B; b++) acc = fmaf(dZ[(size_t)b * N + n], Ap[(size_t)b * K + k], acc);
        dW[(size_t)n * K + k] = acc;
    } else {
        for (int b = 0; b < B; b++) acc += dZ[(size_t)b * N + n];
        db[n] = acc;
    }
}

// dZp[b][k] = (sum_n dZ[b][n] W[n][k]) * (Zp[b][k] > 0)
__global__ void k_dense_bwd_x(const float* __restrict__ dZ, const float* __restrict__ W, const float* __restrict__ Zp,
                              float* __restrict__ dZp, int B, int K, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * K) return;
    const int b = i / K, k = i - b * K;
    float acc = 0.f;
    for (int n = 0; n < N; n++) acc = fmaf(dZ[(size_t)b * N + n], W[(size_t)n * K + k], acc);
    dZp[i] = Zp[i] > 0.f ? acc : 0.f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows of X / y gathered by an index list (one batch of a shuffled epoch)
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ y, const int* __restrict__ idx,
                         float* __restrict__ Xb, float* __restrict__ yb, int B, int D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * (D + 1)) return;
    const int b = i / (D + 1), k = i - b * (D + 1);
    const int r = idx[b];
    if (k < D) Xb[(size_t)b * D + k] = X[(size_t)r * D + k];
    else yb[b] = y[r];
}

inline int nblk(long n) { return (int)((n + kT - 1) / kT); }

}  // namespace

struct spfft_mlp_s {
    int D = 0, H = 0;
    int maxB = 0;
    long nparam = 0;
    DevBuf params, grads, m, v;
    DevBuf Z1, A1, Z2, A2, Z3, dZ3, dZ2, dZ1, loss, Xb, yb, idx;
    int step = 0;
    float* W1() const { return params.as<float>(); }
    float* b1() const { return W1() + (size_t)H * D; }
    float* W2() const { return b1() + H; }
    float* b2() const { return W2() + (size_t)H * H; }
    float* W3() const { return b2() + H; }
    float* b3() const { return W3() + H; }
    float* gW1() const { return grads.as<float>(); }
    float* gb1() const { return gW1() + (size_t)H * D; }
    float* gW2() const { return gb1() + H; }
    float* gb2() const { return gW2() + (size_t)H * H; }
    float* gW3() const { return gb2() + H; }
    float* gb3() const { return gW3() + H; }
};

namespace {

spfft_status ensure_batch(spfft_mlp_s& n, int B) {
    if (B <= n.maxB) return SPFFT_OK;
    const size_t H = n.H;
    SPFFT_HIP(n.Z1.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.A1.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.Z2.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.A2.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.Z3.alloc((size_t)B * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.dZ3.alloc((size_t)B * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.dZ2.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.dZ1.alloc((size_t)B * H * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.Xb.alloc((size_t)B * n.D * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.yb.alloc((size_t)B * 4), "hipMalloc(mlp)");
    SPFFT_HIP(n.idx.alloc((size_t)B * 4), "hipMalloc(mlp)");
    n.maxB = B;
    return SPFFT_OK;
}

void forward(spfft_mlp_s& n, const float* X, int B, hipStream_t s) {
    const int D = n.D, H = n.H;
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long)B * H)), dim3(kT), 0, s, X, n.W1(), n.b1(), n.Z1.as<float>(),
                       n.A1.as<float>(), B, D, H, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk((long)B * H)), dim3(kT), 0, s, n.A1.as<float>(), n.W2(), n.b2(),
                       n.Z2.as<float>(), n.A2.as<float>(), B, H, H, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblk(B)), dim3(kT), 0, s, n.A2.as<float>(), n.W3(), n.b3(),
                       n.Z3.as<float>(), (float*)nullptr, B, H, 1, 0);
}

void backward(spfft_mlp_s& n, const float* X, const float* y, int B, hipStream_t s) {
    const int D = n.D, H = n.H;
    hipLaunchKernelGGL(k_out_grad, dim3(1), dim3(kT), 0, s, n.Z3.as<float>(), y, n.dZ3.as<float>(),
                       n.loss.as<float>(), B);
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((long)1 * (H + 1))), dim3(kT), 0, s, n.dZ3.as<float>(),
                       n.A2.as<float>(), n.gW3(), n.gb3(), B, H, 1);
    hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk((long)B * H)), dim3(kT), 0, s, n.dZ3.as<float>(), n.W3(),
                       n.Z2.as<float>(), n.dZ2.as<float>(), B, H, 1);
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((long)H * (H + 1))), dim3(kT), 0, s, n.dZ2.as<float>(),
                       n.A1.as<float>(), n.gW2(), n.gb2(), B, H, H);
    hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblk((long)B * H)), dim3(kT), 0, s, n.dZ2.as<float>(), n.W2(),
                       n.Z1.as<float>(), n.dZ1.as<float>(), B, H, H);
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblk((long)H * (D + 1))), dim3(kT), 0, s, n.dZ1.as<float>(), X,
                       n.gW1(), n.gb1(), B, D, H);
}

}  // namespace

extern "C" {

spfft_status spfft_mlp_create(int in_dim, int hidden, uint64_t seed, spfft_mlp* out) {
    if (!out || in_dim <= 0 || hidden <= 0) return fail(SPFFT_ERR_INVALID_ARG, "bad MLP shape");
    *out = nullptr;
    std::unique_ptr<spfft_mlp_s> n(new (std::nothrow) spfft_mlp_s());
    if (!n) return fail(SPFFT_ERR_ALLOC, "host allocation failed");
    n->D = in_dim;
    n->H = hidden;
    n->nparam = (long)hidden * in_dim + hidden + (long)hidden * hidden + hidden + hidden + 1;
    // He-uniform initialisation, biases zero
    std::vector<float> p(n->nparam, 0.f);
    std::mt19937_64 g(seed);
    auto fill = [&](float* w, int fan_in, long count) {
        const float lim = std::sqrt(6.f / (float)fan_in);
        std::uniform_real_distribution<float> u(-lim, lim);
        for (long i = 0; i < count; i++) w[i] = u(g);
    };
    fill(p.data(), in_dim, (long)hidden * in_dim);
    fill(p.data() + (long)hidden * in_dim + hidden, hidden, (long)hidden * hidden);
    fill(p.data() + (long)hidden * in_dim + hidden + (long)hidden * hidden + hidden, hidden, hidden);
    SPFFT_HIP(upload(n->params, p), "upload MLP");
    std::vector<float> z(n->nparam, 0.f);
    SPFFT_HIP(upload(n->grads, z), "upload MLP");
    SPFFT_HIP(upload(n->m, z), "upload MLP");
    SPFFT_HIP(upload(n->v, z), "upload MLP");
    SPFFT_HIP(n->loss.alloc(4), "hipMalloc(mlp)");
    *out = n.release();
    return SPFFT_OK;
}

spfft_status spfft_mlp_destroy(spfft_mlp n) {
    delete n;
    return SPFFT_OK;
}

int64_t spfft_mlp_num_params(spfft_mlp n) { return n ? n->nparam : 0; }

spfft_status spfft_mlp_get_params(spfft_mlp n, float* h_params) {
    if (!n || !h_params) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    SPFFT_HIP(hipMemcpy(h_params, n->params.p, n->nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return SPFFT_OK;
}

spfft_status spfft_mlp_set_params(spfft_mlp n, const float* h_params) {
    if (!n || !h_params) return fail(SPFFT_ERR_INVALID_ARG, "NULL argument");
    SPFFT_HIP(hipMemcpy(n->params.p, h_params, n->nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
    n->step = 0;
    std::vector<float> z(n->nparam, 0.f);
    SPFFT_HIP(hipMemcpy(n->m.p, z.data(), n->nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
    SPFFT_HIP(hipMemcpy(n->v.p, z.data(), n->nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
    return SPFFT_OK;
}

// Predictions for n_rows feature rows (device in, device out).  Synchronous.
spfft_status spfft_mlp_predict(spfft_mlp n, const float* d_X, int n_rows, float* d_y, void* stream) {
    if (!n || !d_X || !d_y || n_rows <= 0) return fail(SPFFT_ERR_INVALID_ARG, "bad MLP arguments");
    hipStream_t s = (hipStream_t)stream;
    for (int r0 = 0; r0 < n_rows; r0 += 4096) {
        const int B = std::min(4096, n_rows - r0);
        spfft_status st = ensure_batch(*n, B);
        if (st != SPFFT_OK) return st;
        forward(*n, d_X + (size_t)r0 * n->D, B, s);
        SPFFT_HIP(hipMemcpyAsync(d_y + r0, n->Z3.p, (size_t)B * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
    }
    SPFFT_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return SPFFT_OK;
}

// Loss and gradient of one batch (device X, y) into h_grads (host, num_params
// floats, same layout as the parameters).  Synchronous.
spfft_status spfft_mlp_loss_grad(spfft_mlp n, const float* d_X, const float* d_y, int B, float* loss,
                                 float* h_grads, void* stream) {
    if (!n || !d_X || !d_y || B <= 0) return fail(SPFFT_ERR_INVALID_ARG, "bad MLP arguments");
    hipStream_t s = (hipStream_t)stream;
    spfft_status st = ensure_batch(*n, B);
    if (st != SPFFT_OK) return st;
    forward(*n, d_X, B, s);
    backward(*n, d_X, d_y, B, s);
    SPFFT_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (loss) SPFFT_HIP(hipMemcpy(loss, n->loss.p, 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (h_grads) SPFFT_HIP(hipMemcpy(h_grads, n->grads.p, n->nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return SPFFT_OK;
}

// Adam over `epochs` passes of mini-batches of size `batch`; the sample order of
// each epoch is a permutation from `seed`.  Returns the mean loss of the last epoch.
spfft_status spfft_mlp_train(spfft_mlp n, const float* d_X, const float* d_y, int n_rows, int epochs, int batch,
                             float lr, float weight_decay, uint64_t seed, float* last_loss, void* stream) {
    if (!n || !d_X || !d_y || n_rows <= 0 || batch <= 0 || epochs <= 0)
        return fail(SPFFT_ERR_INVALID_ARG, "bad MLP arguments");
    hipStream_t s = (hipStream_t)stream;
    const int B = std::min(batch, n_rows);
    spfft_status st = ensure_batch(*n, B);
    if (st != SPFFT_OK) return st;
    std::vector<int> perm(n_rows);
    for (int i = 0; i < n_rows; i++) perm[i] = i;
    std::mt19937_64 g(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double epoch_loss = 0.0;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), g);
        epoch_loss = 0.0;
        int nb = 0;
        for (int r0 = 0; r0 + B <= n_rows; r0 += B) {
            SPFFT_HIP(hipMemcpyAsync(n->idx.p, perm.data() + r0, (size_t)B * 4, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
            hipLaunchKernelGGL(k_gather, dim3(nblk((long)B * (n->D + 1))), dim3(kT), 0, s, d_X, d_y,
                               n->idx.as<int>(), n->Xb.as<float>(), n->yb.as<float>(), B, n->D);
            forward(*n, n->Xb.as<float>(), B, s);
            backward(*n, n->Xb.as<float>(), n->yb.as<float>(), B, s);
            n->step++;
            const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
            hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparam)), dim3(kT), 0, s, n->params.as<float>(),
                               n->grads.as<float>(), n->m.as<float>(), n->v.as<float>(), (int)n->nparam, lr, b1,
                               b2, eps, c1, c2, weight_decay);
            float l = 0.f;
            SPFFT_HIP(hipMemcpyAsync(&l, n->loss.p, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
            SPFFT_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
            epoch_loss += l;
            nb++;
        }
        epoch_loss /= std::max(nb, 1);
    }
    if (last_loss) *last_loss = (float)epoch_loss;
    return SPFFT_OK;
}

}  // extern "C"
