// Random butterfly transforms (gfx950): the two-sided A := U^T A V and the
// one-sided B := U^T B, B := V B of SLATE's RBT solver (src/gerbt.cc,
// src/internal/Tile_gerbt.hh, src/internal/internal_rbt_generate.cc).
//
// A depth-d transform has d levels.  Level k (0 <= k < d) cuts [0, n) into
// blocks of 2h indices, h = 2^(d-k-1) * ceil(nt / 2^d) * nb with nt = ceil(n/nb)
// (the reference's half length in tiles times the tile size, gerbt.cc:60-100);
// index i1 = b*2h + t (t < h) pairs with i2 = i1 + h when i2 < n.  With one
// weight vector per side (the reference reads column 0 of its n x d weight
// matrix at every level: U1(ii, 0) in internal_gerbt.cc, U1(i, 0) in
// Tile_gerbt.hh):
//   U^T (left, transposed):      y1 = u1 (x1 + x2) / sqrt2,   y2 = u2 (x1 - x2) / sqrt2
//   V   (left, not transposed):  y1 = (u1 x1 + u2 x2) / sqrt2, y2 = (u1 x1 - u2 x2) / sqrt2
//   A V (right):                 c1 = (a1 + a2) v1 / sqrt2,   c2 = (a1 - a2) v2 / sqrt2
//   an index without a partner:  y1 = u1 x1
// U^T levels run smallest first and V levels largest first (gerbt.cc
// gerbt_iterate_1d); the two-sided form runs smallest first and applies both
// sides of a level in one pass (Tile_gerbt.hh gerbt).
//
// Every kernel is a streaming pass: one thread owns one index pair (i1, i2)
// of a column (or a 2 x 2 group of a matrix), reads it once and writes it
// once, so a level moves 2 n^2 (two-sided) or 2 n nrhs (one-sided) elements
// through HBM.  Consecutive threads own consecutive i1, so the loads and
// stores of both rows of a pair are coalesced.
#include "common.hh"
#include "kernels.hh"

#include <cmath>

namespace sa {

int64_t rbt_half(int64_t n, int64_t nb, int depth, int k)
{
    const int64_t nt = ceildiv(n, nb);
    const int64_t inner = (nt + (int64_t(1) << depth) - 1) >> depth;
    return (int64_t(1) << (depth - k - 1)) * inner * nb;
}

// Number of first-of-pair indices in [0, n) for half length h.
static inline int64_t rbt_pairs(int64_t n, int64_t h)
{
    return (n / (2 * h)) * h + std::min(n % (2 * h), h);
}

__device__ __forceinline__ uint32_t rbt_first(uint32_t p, uint32_t h)
{
    return (p / h) * 2u * h + p % h;
}

template <class T>
__global__ __launch_bounds__(256)
void rbt2_kernel(T* __restrict__ A, int64_t lda, uint32_t n, uint32_t h, uint32_t P,
                 const T* __restrict__ u, const T* __restrict__ v)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= P) return;
    const T half = T(0.5);
    const T s2 = T(1.0 / 1.4142135623730950488);
    const uint32_t i1 = rbt_first(p, h), i2 = i1 + h;
    const bool hr = i2 < n;
    const T u1 = u[i1];
    const T u2 = hr ? u[i2] : T(0);
    for (uint32_t q = blockIdx.y; q < P; q += gridDim.y) {
        const uint32_t j1 = rbt_first(q, h), j2 = j1 + h;
        const bool hc = j2 < n;
        const T v1 = v[j1];
        T* c1 = A + int64_t(j1) * lda;
        if (hc) {
            T* c2 = A + int64_t(j2) * lda;
            const T v2 = v[j2];
            if (hr) {
                const T a11 = c1[i1], a21 = c1[i2], a12 = c2[i1], a22 = c2[i2];
                const T sum1 = a11 + a12, sum2 = a21 + a22;
                const T dif1 = a11 - a12, dif2 = a21 - a22;
                c1[i1] = half * u1 * (sum1 + sum2) * v1;
                c2[i1] = half * u1 * (dif1 + dif2) * v2;
                c1[i2] = half * u2 * (sum1 - sum2) * v1;
                c2[i2] = half * u2 * (dif1 - dif2) * v2;
            }
            else {
                const T a11 = c1[i1], a12 = c2[i1];
                c1[i1] = s2 * u1 * (a11 + a12) * v1;
                c2[i1] = s2 * u1 * (a11 - a12) * v2;
            }
        }
        else if (hr) {
            const T a11 = c1[i1], a21 = c1[i2];
            c1[i1] = s2 * u1 * (a11 + a21) * v1;
            c1[i2] = s2 * u2 * (a11 - a21) * v1;
        }
        else {
            c1[i1] = u1 * c1[i1] * v1;
        }
    }
}

template <class T, int TRANS>
__global__ __launch_bounds__(256)
void rbt1_kernel(T* __restrict__ B, int64_t ldb, uint32_t n, uint32_t nrhs, uint32_t h, uint32_t P,
                 const T* __restrict__ u)
{
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= P) return;
    const T s2 = T(1.0 / 1.4142135623730950488);
    const uint32_t i1 = rbt_first(p, h), i2 = i1 + h;
    const bool hr = i2 < n;
    const T u1 = u[i1];
    const T u2 = hr ? u[i2] : T(0);
    for (uint32_t j = blockIdx.y; j < nrhs; j += gridDim.y) {
        T* b = B + int64_t(j) * ldb;
        if (hr) {
            const T b1 = b[i1], b2 = b[i2];
            if (TRANS) {
                b[i1] = s2 * u1 * (b1 + b2);
                b[i2] = s2 * u2 * (b1 - b2);
            }
            else {
                b[i1] = s2 * (u1 * b1 + u2 * b2);
                b[i2] = s2 * (u1 * b1 - u2 * b2);
            }
        }
        else {
            b[i1] = u1 * b[i1];
        }
    }
}

// u[i] = exp(r / 20) for the matgen rands entry r (a uniform on (-1, 1)).
__global__ __launch_bounds__(256)
void rbt_exp_kernel(double* __restrict__ u, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        u[i] = exp(u[i] / 20.0);
}

static void rbt_check(int64_t n, int64_t nb, int depth)
{
    SA_REQUIRE(depth >= 0 && depth <= 20, -2, "rbt: depth must be in [0, 20]");
    SA_REQUIRE(nb >= 1, -2, "rbt: block size must be positive");
    SA_REQUIRE(n >= 0 && n <= (int64_t(1) << 30), -3, "rbt: n out of range");
}

template <class T>
void gerbt2(T* A, int64_t lda, int64_t n, int64_t nb, int depth, const T* u, const T* v, hipStream_t s)
{
    rbt_check(n, nb, depth);
    if (n == 0 || depth == 0) return;
    SA_REQUIRE(lda >= n, -3, "gerbt: lda < n");
    for (int k = depth - 1; k >= 0; --k) {
        const int64_t h = rbt_half(n, nb, depth, k);
        const int64_t P = rbt_pairs(n, h);
        // h can exceed n (few tiles, large depth): then no index has a partner
        // and the level only scales by the weights.
        const unsigned gx = unsigned(ceildiv(P, 256));
        const unsigned gy = unsigned(std::min<int64_t>(P, 2048));
        hipLaunchKernelGGL(rbt2_kernel<T>, dim3(gx, gy), dim3(256), 0, s, A, lda, uint32_t(n),
                           uint32_t(std::min<int64_t>(h, n)), uint32_t(P), u, v);
        SA_CHECK_LAUNCH();
    }
}

template <class T>
void gerbt1(int trans, T* B, int64_t ldb, int64_t n, int64_t nrhs, int64_t nb, int depth, const T* u,
            hipStream_t s)
{
    rbt_check(n, nb, depth);
    if (n == 0 || nrhs == 0 || depth == 0) return;
    SA_REQUIRE(ldb >= n && nrhs > 0 && nrhs <= (int64_t(1) << 30), -4, "gerbt: bad nrhs/ldb");
    for (int it = 0; it < depth; ++it) {
        const int k = trans ? depth - 1 - it : it;
        const int64_t h = rbt_half(n, nb, depth, k);
        const int64_t P = rbt_pairs(n, h);
        const unsigned gx = unsigned(ceildiv(P, 256));
        const unsigned gy = unsigned(std::min<int64_t>(nrhs, 1024));
        if (trans)
            hipLaunchKernelGGL((rbt1_kernel<T, 1>), dim3(gx, gy), dim3(256), 0, s, B, ldb, uint32_t(n),
                               uint32_t(nrhs), uint32_t(std::min<int64_t>(h, n)), uint32_t(P), u);
        else
            hipLaunchKernelGGL((rbt1_kernel<T, 0>), dim3(gx, gy), dim3(256), 0, s, B, ldb, uint32_t(n),
                               uint32_t(nrhs), uint32_t(std::min<int64_t>(h, n)), uint32_t(P), u);
        SA_CHECK_LAUNCH();
    }
}

void rbt_weights(int64_t seed, int64_t n, double* u, hipStream_t s)
{
    if (n <= 0) return;
    generate(2, seed, n, 1, 0, 0, u, n, s);
    hipLaunchKernelGGL(rbt_exp_kernel, dim3(unsigned(std::min<int64_t>(ceildiv(n, 256), 1024))), dim3(256),
                       0, s, u, n);
    SA_CHECK_LAUNCH();
}

template void gerbt2<double>(double*, int64_t, int64_t, int64_t, int, const double*, const double*, hipStream_t);
template void gerbt1<double>(int, double*, int64_t, int64_t, int64_t, int64_t, int, const double*, hipStream_t);

} // namespace sa
