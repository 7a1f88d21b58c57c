// Learned solver selection: feature extraction and a small MLP (gfx950).
//
// The selector maps cheap per-system features to a choice of factorization
// precision / refinement scheme.  Features (features_* kernels) come from one
// or two passes over A: per-tile max |a_ij| (tiles of 256 x 256) and a
// 16-bin histogram of the binary exponent of |a_ij| relative to the largest
// one.  The network is a 3-layer perceptron in FP32,
//   h1 = relu(W1^T x + b1), h2 = relu(W2^T h1 + b2), z = W3^T h2 + b3,
// with softmax cross-entropy; parameters are one contiguous float array in
// the order W1 (F x H, row-major), b1, W2 (H x H), b2, W3 (H x K), b3.
// Training runs on the GPU in one workgroup per step (mlp_train_kernel):
// forward and backward per sample, gradient sums over the minibatch, and an
// Adam update.  No counterpart in the reference: SLATE chooses precisions
// through template arguments and options only.
#include "common.hh"
#include "kernels.hh"

#include <cmath>

namespace sa {

static constexpr int MLP_F = 24, MLP_H = 64, MLP_K = 8;
static constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_F * MLP_H, MLP_W2 = MLP_B1 + MLP_H,
                     MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H,
                     MLP_B3 = MLP_W3 + MLP_H * MLP_K, MLP_NPARAM = MLP_B3 + MLP_K;

int mlp_num_params() { return MLP_NPARAM; }
int mlp_scratch_per_sample() { return 2 * MLP_H + MLP_K; }   // floats of mlp_train scratch per sample

//------------------------------------------------------------------------------
// Features.
__global__ __launch_bounds__(256)
void tile_maxabs_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int64_t tiles,
                        double* __restrict__ tmax)
{
    const int64_t ti = blockIdx.x, tj = blockIdx.y;
    const int64_t i0 = ti * 256, j0 = tj * 256;
    const int64_t i = i0 + threadIdx.x;
    double m = 0.0;
    if (i < n) {
        const int64_t j1 = std::min<int64_t>(j0 + 256, n);
        for (int64_t j = j0; j < j1; ++j) m = fmax(m, fabs(A[i + j * lda]));
    }
    __shared__ double red[4];
    m = fmax(m, __shfl_xor(m, 32, 64));
    for (int o = 16; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) tmax[ti * tiles + tj] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// hist[b] += number of entries with emax - exponent(|a|) in bin b (b = 15
// collects d >= 15 and exact zeros); exponent as frexp's.
__global__ __launch_bounds__(256)
void exp_hist_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int64_t cstep, int emax,
                     unsigned long long* __restrict__ hist)
{
    // per-thread counters in registers (compile-time indices), one LDS
    // reduction per workgroup, one global atomic per bin and workgroup
    unsigned int c[16];
    #pragma unroll
    for (int k = 0; k < 16; ++k) c[k] = 0;
    const int64_t total = n * ceildiv(n, cstep);            // columns 0, cstep, 2 cstep, ...
    for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total; t += int64_t(gridDim.x) * 256) {
        const int64_t i = t % n, j = (t / n) * cstep;
        const double a = fabs(A[i + j * lda]);
        int b = 15;
        if (a != 0.0) {
            int e;
            frexp(a, &e);
            b = min(15, max(0, emax - e));
        }
        #pragma unroll
        for (int k = 0; k < 16; ++k) c[k] += (b == k);
    }
    __shared__ unsigned int h[16][257];
    #pragma unroll
    for (int k = 0; k < 16; ++k) h[k][threadIdx.x] = c[k];
    __syncthreads();
    if (threadIdx.x < 16) {
        unsigned long long s = 0;
        for (int i = 0; i < 256; ++i) s += h[threadIdx.x][i];
        atomicAdd(&hist[threadIdx.x], s);
    }
}

// The same tile maxima, and each tile's row sums of |a_ij| (part[tj * n + i]
// for tile column tj) for ||A||_inf, in the same single pass over A; the
// partial sums are added in tile order by rowsum_reduce_kernel, so the norm
// does not depend on scheduling.
__global__ __launch_bounds__(256)
void tile_maxabs_rows_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int64_t tiles,
                             double* __restrict__ tmax, double* __restrict__ part)
{
    const int64_t ti = blockIdx.x, tj = blockIdx.y;
    const int64_t i0 = ti * 256, j0 = tj * 256;
    const int64_t i = i0 + threadIdx.x;
    double m = 0.0, s = 0.0;
    if (i < n) {
        const int64_t j1 = std::min<int64_t>(j0 + 256, n);
        for (int64_t j = j0; j < j1; ++j) {
            const double a = fabs(__builtin_nontemporal_load(&A[i + j * lda]));
            m = fmax(m, a);
            s += a;
        }
        part[tj * n + i] = s;
    }
    __shared__ double red[4];
    m = fmax(m, __shfl_xor(m, 32, 64));
    for (int o = 16; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) tmax[ti * tiles + tj] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ __launch_bounds__(256)
void rowsum_reduce_kernel(const double* __restrict__ part, int64_t n, int64_t tiles, double* __restrict__ rowsum)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t tj = 0; tj < tiles; ++tj) s += part[tj * n + i];
    rowsum[i] = s;
}

void feature_tiles_rows(const double* A, int64_t lda, int64_t n, double* tmax, double* part, double* rowsum,
                        hipStream_t s)
{
    const int64_t t = ceildiv(n, 256);
    SA_REQUIRE(t < 65536, -2, "feature_tiles: matrix too large");
    hipLaunchKernelGGL(tile_maxabs_rows_kernel, dim3(unsigned(t), unsigned(t)), dim3(256), 0, s, A, lda, n, t,
                       tmax, part);
    SA_CHECK_LAUNCH();
    hipLaunchKernelGGL(rowsum_reduce_kernel, dim3(unsigned(ceildiv(n, 256))), dim3(256), 0, s, part, n, t, rowsum);
    SA_CHECK_LAUNCH();
}

void feature_tiles(const double* A, int64_t lda, int64_t n, double* tmax, hipStream_t s)
{
    const int64_t t = ceildiv(n, 256);
    SA_REQUIRE(t < 65536, -2, "feature_tiles: matrix too large");
    hipLaunchKernelGGL(tile_maxabs_kernel, dim3(unsigned(t), unsigned(t)), dim3(256), 0, s, A, lda, n, t, tmax);
    SA_CHECK_LAUNCH();
}

void feature_hist(const double* A, int64_t lda, int64_t n, int64_t cstep, int emax, unsigned long long* hist,
                  hipStream_t s)
{
    SA_REQUIRE(cstep >= 1, -2, "feature_hist: column step < 1");
    const unsigned g = unsigned(std::min<int64_t>(ceildiv(n * ceildiv(n, cstep), 256), 4096));
    hipLaunchKernelGGL(exp_hist_kernel, dim3(g), dim3(256), 0, s, A, lda, n, cstep, emax, hist);
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// Network.
__device__ inline void mlp_fwd(const float* __restrict__ P, const float* x, float* h1, float* h2, float* z)
{
    for (int j = 0; j < MLP_H; ++j) {
        float s = P[MLP_B1 + j];
        for (int f = 0; f < MLP_F; ++f) s += x[f] * P[MLP_W1 + f * MLP_H + j];
        h1[j] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < MLP_H; ++j) {
        float s = P[MLP_B2 + j];
        for (int i = 0; i < MLP_H; ++i) s += h1[i] * P[MLP_W2 + i * MLP_H + j];
        h2[j] = s > 0.f ? s : 0.f;
    }
    for (int k = 0; k < MLP_K; ++k) {
        float s = P[MLP_B3 + k];
        for (int i = 0; i < MLP_H; ++i) s += h2[i] * P[MLP_W3 + i * MLP_K + k];
        z[k] = s;
    }
}

__device__ inline void softmax(float* z)
{
    float m = z[0];
    for (int k = 1; k < MLP_K; ++k) m = fmaxf(m, z[k]);
    float s = 0.f;
    for (int k = 0; k < MLP_K; ++k) { z[k] = expf(z[k] - m); s += z[k]; }
    for (int k = 0; k < MLP_K; ++k) z[k] /= s;
}

// probs[b][k] for B samples (one thread per sample).
__global__ __launch_bounds__(256)
void mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X, int B, float* __restrict__ probs)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    float x[MLP_F], h1[MLP_H], h2[MLP_H], z[MLP_K];
    for (int f = 0; f < MLP_F; ++f) x[f] = X[size_t(b) * MLP_F + f];
    mlp_fwd(P, x, h1, h2, z);
    softmax(z);
    for (int k = 0; k < MLP_K; ++k) probs[size_t(b) * MLP_K + k] = z[k];
}

// One training step on a minibatch of B <= 4096 samples.  Single workgroup.
// grad (out): d(mean loss)/dP; scratch: B x (2H + K) floats; with lr > 0
// the Adam update (beta1 0.9, beta2 0.999, eps 1e-8, step t) is applied to P
// using the moment arrays m, v; loss (out): mean cross-entropy.
__global__ __launch_bounds__(256)
void mlp_train_kernel(float* __restrict__ P, float* __restrict__ grad, float* __restrict__ m,
                      float* __restrict__ v, const float* __restrict__ X, const int* __restrict__ y, int B,
                      float* __restrict__ scratch, float lr, int t, float* __restrict__ loss)
{
    const int tid = threadIdx.x;
    float* H1 = scratch;                         // B x H (post-relu)
    float* H2 = H1 + size_t(B) * MLP_H;          // B x H (post-relu), later delta2
    float* D3 = H2 + size_t(B) * MLP_H;          // B x K
    __shared__ float s_loss[256];
    float lsum = 0.f;
    // (1) per sample: forward, softmax, output delta (p - onehot) / B
    for (int b = tid; b < B; b += 256) {
        float x[MLP_F], h1[MLP_H], h2[MLP_H], z[MLP_K];
        for (int f = 0; f < MLP_F; ++f) x[f] = X[size_t(b) * MLP_F + f];
        mlp_fwd(P, x, h1, h2, z);
        softmax(z);
        const int yb = y[b];
        lsum += -logf(fmaxf(z[yb], 1e-30f));
        for (int j = 0; j < MLP_H; ++j) { H1[size_t(b) * MLP_H + j] = h1[j]; H2[size_t(b) * MLP_H + j] = h2[j]; }
        for (int k = 0; k < MLP_K; ++k) D3[size_t(b) * MLP_K + k] = (z[k] - (k == yb ? 1.f : 0.f)) / float(B);
    }
    s_loss[tid] = lsum;
    __syncthreads();
    // (2) gradients of layer 3: dW3 = H2^T D3, db3 = sum D3
    for (int e = tid; e < MLP_H * MLP_K + MLP_K; e += 256) {
        float s = 0.f;
        if (e < MLP_H * MLP_K) {
            const int i = e / MLP_K, k = e % MLP_K;
            for (int b = 0; b < B; ++b) s += H2[size_t(b) * MLP_H + i] * D3[size_t(b) * MLP_K + k];
            grad[MLP_W3 + e] = s;
        }
        else {
            const int k = e - MLP_H * MLP_K;
            for (int b = 0; b < B; ++b) s += D3[size_t(b) * MLP_K + k];
            grad[MLP_B3 + k] = s;
        }
    }
    __syncthreads();
    // (3) delta2 = (W3 D3) * 1[h2 > 0], overwriting H2 (dW3 already used it)
    for (int e = tid; e < B * MLP_H; e += 256) {
        const int b = e / MLP_H, i = e % MLP_H;
        float s = 0.f;
        for (int k = 0; k < MLP_K; ++k) s += P[MLP_W3 + i * MLP_K + k] * D3[size_t(b) * MLP_K + k];
        H2[e] = H2[e] > 0.f ? s : 0.f;
    }
    __syncthreads();
    // (4) layer 2: dW2 = H1^T delta2, db2 = sum delta2
    for (int e = tid; e < MLP_H * MLP_H + MLP_H; e += 256) {
        float s = 0.f;
        if (e < MLP_H * MLP_H) {
            const int i = e / MLP_H, j = e % MLP_H;
            for (int b = 0; b < B; ++b) s += H1[size_t(b) * MLP_H + i] * H2[size_t(b) * MLP_H + j];
            grad[MLP_W2 + e] = s;
        }
        else {
            const int j = e - MLP_H * MLP_H;
            for (int b = 0; b < B; ++b) s += H2[size_t(b) * MLP_H + j];
            grad[MLP_B2 + j] = s;
        }
    }
    __syncthreads();
    // (5) delta1 = (W2 delta2) * 1[h1 > 0], overwriting H1 (dW2 already used it)
    for (int b = tid; b < B; b += 256) {
        float d1[MLP_H];
        for (int i = 0; i < MLP_H; ++i) {
            float s = 0.f;
            for (int j = 0; j < MLP_H; ++j) s += P[MLP_W2 + i * MLP_H + j] * H2[size_t(b) * MLP_H + j];
            d1[i] = H1[size_t(b) * MLP_H + i] > 0.f ? s : 0.f;
        }
        for (int i = 0; i < MLP_H; ++i) H1[size_t(b) * MLP_H + i] = d1[i];
    }
    __syncthreads();
    // (6) layer 1: dW1 = X^T delta1, db1 = sum delta1
    for (int e = tid; e < MLP_F * MLP_H + MLP_H; e += 256) {
        float s = 0.f;
        if (e < MLP_F * MLP_H) {
            const int f = e / MLP_H, j = e % MLP_H;
            for (int b = 0; b < B; ++b) s += X[size_t(b) * MLP_F + f] * H1[size_t(b) * MLP_H + j];
            grad[MLP_W1 + e] = s;
        }
        else {
            const int j = e - MLP_F * MLP_H;
            for (int b = 0; b < B; ++b) s += H1[size_t(b) * MLP_H + j];
            grad[MLP_B1 + j] = s;
        }
    }
    if (tid == 0) {
        float s = 0.f;
        for (int i = 0; i < 256; ++i) s += s_loss[i];
        *loss = s / float(B);
    }
    __syncthreads();
    // (7) Adam
    if (lr > 0.f) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - powf(b1, float(t)), c2 = 1.f - powf(b2, float(t));
        for (int e = tid; e < MLP_NPARAM; e += 256) {
            const float g = grad[e];
            const float mm = b1 * m[e] + (1.f - b1) * g;
            const float vv = b2 * v[e] + (1.f - b2) * g * g;
            m[e] = mm;
            v[e] = vv;
            P[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
        }
    }
}

void mlp_forward(const float* P, const float* X, int B, float* probs, hipStream_t s)
{
    if (B <= 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(unsigned(ceildiv(B, 256))), dim3(256), 0, s, P, X, B, probs);
    SA_CHECK_LAUNCH();
}

void mlp_train(float* P, float* grad, float* m, float* v, const float* X, const int* y, int B, float* scratch,
               float lr, int t, float* loss, hipStream_t s)
{
    SA_REQUIRE(B >= 1 && B <= 4096, -2, "mlp_train: batch must be 1..4096");
    hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(256), 0, s, P, grad, m, v, X, y, B, scratch, lr, t, loss);
    SA_CHECK_LAUNCH();
}

} // namespace sa
