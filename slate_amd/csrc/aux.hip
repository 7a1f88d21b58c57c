// Memory-bound FP64 kernels of the refinement loops (gfx950): the infinity
// norm of A, per-column max norms, the FP64 residual R = B - A X, vector
// updates, the GMRES dot products, and the matgen Philox generator.
//
// Reductions are two-pass and deterministic: partial results per column chunk
// go to a workspace and a second kernel folds them in a fixed order, so a
// residual is bit-reproducible run to run.  These replace the per-tile partial
// norms that SLATE finishes on the host (src/cuda/device_genorm.cu:178-197,
// :285-316) and the FP64 gemm/hemm residuals of src/gesv_mixed.cc:198-204 and
// src/posv_mixed.cc:200-205.
#include "common.hh"
#include "kernels.hh"

namespace sa {

__device__ inline double max_nan(double a, double b) { return (b > a || isnan(b)) ? b : a; }

static constexpr int RCH = 1024;   // columns per residual/norm chunk

//------------------------------------------------------------------------------
// Row partial sums of |A| over a column chunk (sym: only j <= i, the stored
// lower triangle).
__global__ __launch_bounds__(256)
void rowabs_partial_kernel(const double* __restrict__ A, int64_t m, int64_t n, int64_t lda,
                           int sym, double* __restrict__ part)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t j0 = c * RCH, j1 = (j0 + RCH < n) ? j0 + RCH : n;
    if (i >= m) return;
    double s = 0;
    int64_t je = sym ? ((i + 1 < j1) ? i + 1 : j1) : j1;
    for (int64_t j = j0; j < je; ++j) s += fabs(A[i + j * lda]);
    part[c * m + i] = s;
}

// Column sums of |A| strictly below the diagonal (the mirrored upper part of
// a lower-stored symmetric matrix): one wave per column.
__global__ __launch_bounds__(256)
void colabs_below_kernel(const double* __restrict__ A, int64_t n, int64_t lda,
                         double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    double s = 0;
    for (int64_t i = j + 1 + lane; i < n; i += 64) s += fabs(A[i + j * lda]);
    s = wave_sum(s);
    if (lane == 0) out[j] = s;
}

__global__ __launch_bounds__(1024)
void norm_final_kernel(const double* __restrict__ part, int64_t m, int64_t nchunk,
                       const double* __restrict__ extra, double* out)
{
    __shared__ double red[1024];
    double mx = 0;
    for (int64_t i = threadIdx.x; i < m; i += 1024) {
        double s = 0;
        for (int64_t c = 0; c < nchunk; ++c) s += part[c * m + i];
        if (extra) s += extra[i];
        mx = max_nan(mx, s);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

void norm_inf(const double* A, int64_t m, int64_t n, int64_t lda, int sym, double* work,
              int64_t work_elems, double* out, hipStream_t s)
{
    int64_t nchunk = ceildiv(n, RCH);
    SA_REQUIRE(work_elems >= nchunk * m + (sym ? n : 0), -2, "norm_inf: workspace too small");
    hipLaunchKernelGGL(rowabs_partial_kernel, dim3(unsigned(ceildiv(m, 256)), unsigned(nchunk)),
                       dim3(256), 0, s, A, m, n, lda, sym, work);
    SA_CHECK_LAUNCH();
    double* extra = nullptr;
    if (sym) {
        extra = work + nchunk * m;
        hipLaunchKernelGGL(colabs_below_kernel, dim3(unsigned(ceildiv(n, 4))), dim3(256), 0, s,
                           A, n, lda, extra);
        SA_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(1024), 0, s, work, m, nchunk, extra, out);
    SA_CHECK_LAUNCH();
}

__global__ __launch_bounds__(1024)
void colmax_kernel(const double* __restrict__ X, int64_t m, int64_t ldx, double* out)
{
    __shared__ double red[1024];
    const int64_t j = blockIdx.x;
    double mx = 0;
    for (int64_t i = threadIdx.x; i < m; i += 1024) mx = max_nan(mx, fabs(X[i + j * ldx]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

void col_norms_max(const double* X, int64_t m, int64_t n, int64_t ldx, double* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(colmax_kernel, dim3(unsigned(n)), dim3(1024), 0, s, X, m, ldx, out);
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// Residual.  Pass 1, rows: part[c][r][i] = sum_{j in chunk c (j <= i if sym)}
// A[i,j] X[j,r].  Pass 1', sym only: colpart[b][r][j] = sum_{i > j, i in row
// chunk b} A[i,j] X[i,r] (the mirrored upper triangle, column dots, one wave
// per (column, row chunk)).  Pass 2 folds the partials in a fixed order.
// A is streamed once per call, so its loads are nontemporal (measured at
// n = 65536: 6.15 -> 6.45 TB/s for one row per thread).
template <int NR>
__global__ __launch_bounds__(256)
void gemv_rows_kernel(const double* __restrict__ A, int64_t n, int64_t lda, int sym,
                      const double* __restrict__ X, int64_t ldx, double* __restrict__ part)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t j0 = c * RCH, j1 = (j0 + RCH < n) ? j0 + RCH : n;
    if (i >= n) return;
    double acc[NR];
    #pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    int64_t je = sym ? ((i + 1 < j1) ? i + 1 : j1) : j1;
    for (int64_t j = j0; j < je; ++j) {
        double a = __builtin_nontemporal_load(&A[i + j * lda]);
        #pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += a * X[j + r * ldx];
    }
    #pragma unroll
    for (int r = 0; r < NR; ++r) part[(c * NR + r) * n + i] = acc[r];
}

// The same with two rows (i, i + 1) per thread and one 16-byte nontemporal
// load per column (A 16-byte aligned, lda and n even): half the load
// instructions, each row summed in the same order as above (bitwise equal);
// n = 65536: 5.59 -> 5.11 ms per pass over the 34 GB matrix (6.72 TB/s).
typedef double gv_d2 __attribute__((ext_vector_type(2)));
template <int NR>
__global__ __launch_bounds__(256)
void gemv_rows2_kernel(const double* __restrict__ A, int64_t n, int64_t lda, int sym,
                       const double* __restrict__ X, int64_t ldx, double* __restrict__ part)
{
    const int64_t i = (int64_t(blockIdx.x) * 256 + threadIdx.x) * 2;
    const int64_t c = blockIdx.y;
    const int64_t j0 = c * RCH, j1 = (j0 + RCH < n) ? j0 + RCH : n;
    if (i >= n) return;
    double a0[NR], a1[NR];
    #pragma unroll
    for (int r = 0; r < NR; ++r) a0[r] = a1[r] = 0;
    // sym: row i runs over j <= i, row i + 1 over j <= i + 1
    const int64_t je0 = sym ? ((i + 1 < j1) ? i + 1 : j1) : j1;
    const int64_t je1 = sym ? ((i + 2 < j1) ? i + 2 : j1) : j1;
    for (int64_t j = j0; j < je0; ++j) {
        const gv_d2 a = __builtin_nontemporal_load(reinterpret_cast<const gv_d2*>(&A[i + j * lda]));
        #pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double x = X[j + r * ldx];
            a0[r] += a[0] * x;
            a1[r] += a[1] * x;
        }
    }
    for (int64_t j = (je0 > j0 ? je0 : j0); j < je1; ++j) {
        const double a = A[i + 1 + j * lda];
        #pragma unroll
        for (int r = 0; r < NR; ++r) a1[r] += a * X[j + r * ldx];
    }
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
        part[(c * NR + r) * n + i] = a0[r];
        part[(c * NR + r) * n + i + 1] = a1[r];
    }
}

static constexpr int CROWS = 4096;   // rows per column-dot chunk (sym residual)

template <int NR>
__global__ __launch_bounds__(256)
void gemv_cols_below_kernel(const double* __restrict__ A, int64_t n, int64_t lda,
                            const double* __restrict__ X, int64_t ldx, double* __restrict__ part)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int64_t b = blockIdx.y;
    if (j >= n) return;
    int64_t i0 = b * CROWS, i1 = (i0 + CROWS < n) ? i0 + CROWS : n;
    if (i0 < j + 1) i0 = j + 1;
    double acc[NR];
    #pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int64_t i = i0 + lane; i < i1; i += 64) {
        double a = A[i + j * lda];
        #pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += a * X[i + r * ldx];
    }
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
        double v = wave_sum(acc[r]);
        if (lane == 0) part[(b * NR + r) * n + j] = v;
    }
}

template <int NR>
__global__ __launch_bounds__(256)
void residual_final_kernel(int64_t n, int64_t nchunk, int64_t nbchunk, const double* __restrict__ part,
                           const double* __restrict__ cpart, const double* __restrict__ B,
                           int64_t ldb, double* __restrict__ R, int64_t ldr)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = 0;
        for (int64_t c = 0; c < nchunk; ++c) s += part[(c * NR + r) * n + i];
        if (cpart)
            for (int64_t b = 0; b < nbchunk; ++b) s += cpart[(b * NR + r) * n + i];
        R[i + r * ldr] = B[i + r * ldb] - s;
    }
}

int64_t residual_work_elems(int64_t n, int64_t nrhs)
{
    return (ceildiv(n, RCH) + ceildiv(n, CROWS)) * nrhs * n + 64;
}

template <int NR>
static void residual_nr(const double* A, int64_t n, int64_t lda, int sym, const double* X,
                        int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                        double* work, hipStream_t s)
{
    int64_t nchunk = ceildiv(n, RCH);
    int64_t nb = ceildiv(n, CROWS);
    double* part = work;
    double* cpart = sym ? work + nchunk * NR * n : nullptr;
    if (n % 2 == 0 && lda % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0)
        hipLaunchKernelGGL(gemv_rows2_kernel<NR>, dim3(unsigned(ceildiv(n, 512)), unsigned(nchunk)),
                           dim3(256), 0, s, A, n, lda, sym, X, ldx, part);
    else
        hipLaunchKernelGGL(gemv_rows_kernel<NR>, dim3(unsigned(ceildiv(n, 256)), unsigned(nchunk)),
                           dim3(256), 0, s, A, n, lda, sym, X, ldx, part);
    SA_CHECK_LAUNCH();
    if (sym) {
        hipLaunchKernelGGL(gemv_cols_below_kernel<NR>, dim3(unsigned(ceildiv(n, 4)), unsigned(nb)),
                           dim3(256), 0, s, A, n, lda, X, ldx, cpart);
        SA_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(residual_final_kernel<NR>, dim3(unsigned(ceildiv(n, 256))), dim3(256), 0, s,
                       n, nchunk, nb, part, cpart, B, ldb, R, ldr);
    SA_CHECK_LAUNCH();
}

void residual(const double* A, int64_t n, int64_t lda, int sym, const double* X, int64_t ldx,
              const double* B, int64_t ldb, double* R, int64_t ldr, int64_t nrhs,
              double* work, int64_t work_elems, hipStream_t s)
{
    SA_REQUIRE(work_elems >= residual_work_elems(n, std::min<int64_t>(nrhs, 4)), -2,
               "residual: workspace too small");
    for (int64_t c = 0; c < nrhs; c += 4) {
        int64_t k = std::min<int64_t>(4, nrhs - c);
        const double* Xc = X + c * ldx;
        const double* Bc = B + c * ldb;
        double* Rc = R + c * ldr;
        switch (k) {
            case 1: residual_nr<1>(A, n, lda, sym, Xc, ldx, Bc, ldb, Rc, ldr, work, s); break;
            case 2: residual_nr<2>(A, n, lda, sym, Xc, ldx, Bc, ldb, Rc, ldr, work, s); break;
            case 3: residual_nr<3>(A, n, lda, sym, Xc, ldx, Bc, ldb, Rc, ldr, work, s); break;
            default: residual_nr<4>(A, n, lda, sym, Xc, ldx, Bc, ldb, Rc, ldr, work, s); break;
        }
    }
}

__global__ void axpy_kernel(int64_t m, double alpha, const double* __restrict__ X, int64_t ldx,
                            double* __restrict__ Y, int64_t ldy)
{
    const int64_t j = blockIdx.y;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < m; i += int64_t(gridDim.x) * 256)
        Y[i + j * ldy] += alpha * X[i + j * ldx];
}

void axpy(int64_t m, int64_t n, double alpha, const double* X, int64_t ldx, double* Y,
          int64_t ldy, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(m, 256), 1024));
    hipLaunchKernelGGL(axpy_kernel, dim3(gx, unsigned(n)), dim3(256), 0, s, m, alpha, X, ldx, Y, ldy);
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// GMRES helpers (k <= 64 basis vectors).
static constexpr int DOTCH = 8192;

__global__ __launch_bounds__(256)
void multi_dot_partial_kernel(const double* __restrict__ V, int64_t ldv, int64_t n, int k,
                              const double* __restrict__ w, double* __restrict__ part)
{
    __shared__ double red[4];
    const int64_t c = blockIdx.x;
    const int p = blockIdx.y;
    const int64_t i0 = c * DOTCH, i1 = (i0 + DOTCH < n) ? i0 + DOTCH : n;
    double s = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += V[i + p * ldv] * w[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[c * k + p] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void multi_dot_final_kernel(const double* __restrict__ part, int64_t nchunk, int k, double* h)
{
    int p = threadIdx.x;
    if (p >= k) return;
    double s = 0;
    for (int64_t c = 0; c < nchunk; ++c) s += part[c * k + p];
    h[p] = s;
}

void multi_dot(const double* V, int64_t ldv, int64_t n, int64_t k, const double* w,
               double* h, double* work, hipStream_t s)
{
    if (k <= 0) return;
    SA_REQUIRE(k <= 64, -2, "multi_dot: k > 64");
    int64_t nchunk = ceildiv(n, DOTCH);
    hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(unsigned(nchunk), unsigned(k)), dim3(256), 0, s,
                       V, ldv, n, int(k), w, work);
    SA_CHECK_LAUNCH();
    hipLaunchKernelGGL(multi_dot_final_kernel, dim3(1), dim3(64), 0, s, work, nchunk, int(k), h);
    SA_CHECK_LAUNCH();
}

__global__ __launch_bounds__(256)
void multi_axpy_kernel(const double* __restrict__ V, int64_t ldv, int64_t n, int k,
                       const double* __restrict__ h, double* __restrict__ w)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        double s = w[i];
        for (int p = 0; p < k; ++p) s -= V[i + p * ldv] * h[p];
        w[i] = s;
    }
}

void multi_axpy(const double* V, int64_t ldv, int64_t n, int64_t k, const double* h,
                double* w, hipStream_t s)
{
    if (k <= 0 || n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(n, 256), 2048));
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(gx), dim3(256), 0, s, V, ldv, n, int(k), h, w);
    SA_CHECK_LAUNCH();
}

__global__ void sqrt_kernel(double* x) { x[0] = sqrt(x[0]); }

void nrm2(const double* w, int64_t n, double* work, double* out, hipStream_t s)
{
    // sum of squares via multi_dot(w, w), then sqrt
    int64_t nchunk = ceildiv(n, DOTCH);
    hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(unsigned(nchunk), 1), dim3(256), 0, s,
                       w, n, n, 1, w, work);
    SA_CHECK_LAUNCH();
    hipLaunchKernelGGL(multi_dot_final_kernel, dim3(1), dim3(64), 0, s, work, nchunk, 1, out);
    SA_CHECK_LAUNCH();
    hipLaunchKernelGGL(sqrt_kernel, dim3(1), dim3(1), 0, s, out);
    SA_CHECK_LAUNCH();
}

__global__ void scale_kernel(double* __restrict__ w, int64_t n, const double* __restrict__ denom)
{
    const double sc = 1.0 / denom[0];     // slate::scale(1, denom): x *= numer/denom
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        w[i] *= sc;
}

void scale(double* w, int64_t n, const double* denom_dev, hipStream_t s)
{
    if (n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(n, 256), 2048));
    hipLaunchKernelGGL(scale_kernel, dim3(gx), dim3(256), 0, s, w, n, denom_dev);
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// matgen Philox (matgen/random.cc:53-127): entry (i, j) depends only on
// (seed, i, j), so the device copy is bit-identical to the host generator for
// the uniform distributions.
__device__ inline void philox_dev(uint64_t s0, uint64_t s1, uint64_t seed, uint64_t& o0, uint64_t& o1)
{
    const uint64_t bump = 0xD2B74407B1CE6E93ull;
    const uint64_t mult = 0x9E3779B97F4A7C15ull;
    #pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r != 0) seed += bump;
        uint64_t lo = s1 * mult;
        uint64_t hi = __umul64hi(s1, mult);
        uint64_t n1 = hi ^ seed ^ s0;
        s0 = lo; s1 = n1;
    }
    o0 = s0; o1 = s1;
}

__global__ __launch_bounds__(256)
void generate_kernel(int dist, uint64_t seed, int64_t m, int64_t n, int64_t ioff, int64_t joff,
                     double* __restrict__ A, int64_t lda)
{
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y)
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < m; i += int64_t(gridDim.x) * 256) {
        uint64_t b0, b1;
        philox_dev(uint64_t(i + ioff), uint64_t(j + joff), seed, b0, b1);
        double u1 = double(b0 >> 11) * (1.0 / 9007199254740992.0);
        double u2 = double(b1 >> 11) * (1.0 / 9007199254740992.0);
        double v;
        if (dist == 1) v = u1;
        else if (dist == 2) v = 2 * u1 - 1;
        else v = sqrt(-2 * log(1 - u1)) * cos(2 * 3.1415926535897932385 * u2);
        A[i + j * lda] = v;
    }
}

void generate(int dist, int64_t seed, int64_t m, int64_t n, int64_t ioff, int64_t joff,
              double* A, int64_t lda, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    // a bounded grid (<= 256 x 2048 workgroups) that strides over the matrix
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(m, 256), 256));
    unsigned gy = unsigned(std::min<int64_t>(n, 2048));
    hipLaunchKernelGGL(generate_kernel, dim3(gx, gy), dim3(256), 0, s, dist,
                       uint64_t(seed), m, n, ioff, joff, A, lda);
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// Row permutation of a vector fused with a precision cast:
// gather: dst[i] = src[perm[i]]  (applies P: the interchanges of getrf, forward)
// scatter: dst[perm[i]] = src[i] (applies P^T)
template <class S, class D, int SCATTER>
__global__ void permute_convert_kernel(const S* __restrict__ src, D* __restrict__ dst,
                                       const int64_t* __restrict__ perm, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        if (SCATTER) dst[perm[i]] = D(src[i]);
        else dst[i] = D(src[perm[i]]);
    }
}

template <class S, class D>
void permute_convert(const S* src, D* dst, const int64_t* perm, int64_t n, int scatter, hipStream_t s)
{
    if (n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(n, 256), 2048));
    if (scatter) hipLaunchKernelGGL((permute_convert_kernel<S, D, 1>), dim3(gx), dim3(256), 0, s, src, dst, perm, n);
    else hipLaunchKernelGGL((permute_convert_kernel<S, D, 0>), dim3(gx), dim3(256), 0, s, src, dst, perm, n);
    SA_CHECK_LAUNCH();
}
template void permute_convert<double, float>(const double*, float*, const int64_t*, int64_t, int, hipStream_t);
template void permute_convert<double, double>(const double*, double*, const int64_t*, int64_t, int, hipStream_t);
template void permute_convert<float, float>(const float*, float*, const int64_t*, int64_t, int, hipStream_t);

// Y += convert(X) for vectors (the "copy X_lo to R; add R to X" step of
// src/gesv_mixed.cc:228-232 in one pass).
template <class S>
__global__ void add_convert_kernel(const S* __restrict__ x, double* __restrict__ y, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] += double(x[i]);
}

template <class S>
void add_convert(const S* x, double* y, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(n, 256), 2048));
    hipLaunchKernelGGL(add_convert_kernel<S>, dim3(gx), dim3(256), 0, s, x, y, n);
    SA_CHECK_LAUNCH();
}
template void add_convert<float>(const float*, double*, int64_t, hipStream_t);
template void add_convert<double>(const double*, double*, int64_t, hipStream_t);

__global__ void zero_u32_kernel(unsigned* p, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) p[i] = 0;
}
__global__ void zero_f64_kernel(double* p, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) p[i] = 0;
}

void zero_u32(unsigned* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(unsigned(std::min<int64_t>(ceildiv(n, 256), 1024))), dim3(256),
                       0, s, p, n);
    SA_CHECK_LAUNCH();
}
void zero_f64(double* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_f64_kernel, dim3(unsigned(std::min<int64_t>(ceildiv(n, 256), 1024))), dim3(256),
                       0, s, p, n);
    SA_CHECK_LAUNCH();
}

} // namespace sa
