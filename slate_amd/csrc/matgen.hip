// Test-matrix generation with a prescribed spectrum (gfx950).
//
// A = U diag(sigma) V^T with U, V products of n Householder reflectors
// H_i = I - tau_i v_i v_i^T, tau_i = 2 / (v_i^T v_i), v_i the i-th column of
// a matgen "randn" matrix (Philox normal, matgen/random.cc:120-127).  The
// reflectors are applied in blocks as compact-WY products
// H_1 ... H_b = I - Y T Y^T (T upper triangular, the forward column-wise
// recurrence of LAPACK dlarft) so the work is GEMM-shaped.  sigma follows
// the distributions of matgen/generate_sigma.hh:61-111 (arith, geo, cluster0,
// cluster1).  SLATE's own "svd" kind builds Haar-distributed U, V from the
// QR of a Gaussian matrix (matgen/generate_type_svd.hh:90-157); a product of
// n random reflectors is the generator the refinement configurations name.
#include "common.hh"
#include "kernels.hh"

#include <cmath>

namespace sa {

// sigma_i for i in [0, n): 1 = arith, 2 = geo, 3 = cluster0, 4 = cluster1.
__global__ __launch_bounds__(256)
void sigma_diag_kernel(int dist, double cond, int64_t n, double* __restrict__ A, int64_t lda)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double t = n > 1 ? double(i) / double(n - 1) : 0.0;
        double s;
        switch (dist) {
            case 1: s = 1.0 - t * (1.0 - 1.0 / cond); break;
            case 2: s = pow(cond, -t); break;
            case 3: s = (i == 0) ? 1.0 : 1.0 / cond; break;
            default: s = (i == n - 1) ? 1.0 / cond : 1.0; break;
        }
        A[i * lda + i] = s;
    }
}

void sigma_diag(int dist, double cond, int64_t n, double* A, int64_t lda, hipStream_t s)
{
    if (n <= 0) return;
    SA_REQUIRE(dist >= 1 && dist <= 4, -2, "sigma_diag: unknown distribution");
    SA_REQUIRE(cond >= 1.0, -2, "sigma_diag: cond must be >= 1");
    hipLaunchKernelGGL(sigma_diag_kernel, dim3(unsigned(std::min<int64_t>(ceildiv(n, 256), 1024))), dim3(256),
                       0, s, dist, cond, n, A, lda);
    SA_CHECK_LAUNCH();
}

// T (b x b, row-major, upper) of H_1 ... H_b = I - Y T Y^T from the Gram
// matrix G = Y^T Y (b x b, row-major):
//   T(i, i) = tau_i = 2 / G(i, i),
//   T(r, i) = -tau_i * sum_{j = r}^{i-1} T(r, j) G(j, i),  r < i.
// Row r only needs row r of T, so each thread owns one row; no barriers.
__global__ __launch_bounds__(256)
void larft_kernel(const double* __restrict__ G, int64_t b, double* __restrict__ T)
{
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= b) return;
    double* Tr = T + r * b;
    for (int64_t j = 0; j < r; ++j) Tr[j] = 0.0;
    Tr[r] = 2.0 / G[r * b + r];
    for (int64_t i = r + 1; i < b; ++i) {
        const double tau = 2.0 / G[i * b + i];
        double s = 0.0;
        for (int64_t j = r; j < i; ++j) s += Tr[j] * G[j * b + i];
        Tr[i] = -tau * s;
    }
}

void larft_from_gram(const double* G, int64_t b, double* T, hipStream_t s)
{
    if (b <= 0) return;
    hipLaunchKernelGGL(larft_kernel, dim3(unsigned(ceildiv(b, 256))), dim3(256), 0, s, G, b, T);
    SA_CHECK_LAUNCH();
}

} // namespace sa
