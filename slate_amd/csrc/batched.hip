// Batched factor-and-solve of many small systems of uneven size (gfx950).
//
// One persistent launch: each workgroup takes the next system from a work
// queue (an atomic counter over a largest-first order), copies it into its
// own workspace in the factor precision T (row-major, so interchanges move
// contiguous rows), factors it with partial pivoting, solves, and - when
// T = float - refines in FP64 against the original matrix to SLATE's
// stopping rule (src/gesv_mixed.cc:190-259, internal_util.hh:121-138), with
// the FP64 factorization as the fallback when refinement misses the
// tolerance.  The per-system steps mirror slate::gesv / gesv_mixed on a
// single tile column; the batching (one launch for thousands of systems of
// different n, no per-system launches or host round trips) is this
// library's own design.
//
// Factorization: blocked right-looking LU with NB-column panels held in LDS
// (up to 1024 rows x NB), pivot search by wave shuffles, the panel's
// interchanges applied to the rest of the rows, U12 = L11^{-1} A12, and the
// trailing update A22 -= L21 U12 register-blocked over the workgroup.
// Triangular solves: 64-row blocks, each diagonal block solved by one wave
// (no barriers), the rows below updated by the whole workgroup.
#include "common.hh"
#include "kernels.hh"

#include <limits>

namespace sa {

static constexpr int BT = 256;              // threads per workgroup
static constexpr int BMAXN = 1024;          // largest system

template <class T> struct BNB;
template <> struct BNB<float> { static constexpr int v = 32; };
template <> struct BNB<double> { static constexpr int v = 16; };

__device__ inline void block_sync() { __syncthreads(); }

template <class T>
__device__ inline bool bt_better(T a, int ia, T b, int ib) { return a > b || (a == b && ia < ib); }

// In-workgroup LU with partial pivoting of the row-major n x n matrix W
// (leading dimension ldw).  piv[j] = 0-based row swapped with row j.
// Returns info (first exactly-zero pivot, 1-based) in *info (thread 0 writes).
template <class T>
__device__ __noinline__ void wg_getrf(T* __restrict__ W, int n, int ldw, int* __restrict__ piv, int* info,
                                      T* s_p, T* s_red, int* s_ired)
{
    constexpr int NB = BNB<T>::v;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T safe_min = std::numeric_limits<T>::min();
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int mr = n - k0;                           // panel rows
        // panel -> LDS (row-major mr x NB)
        for (int e = tid; e < mr * NB; e += BT) {
            const int i = e / NB, c = e % NB;
            s_p[e] = c < kb ? W[(k0 + i) * ldw + k0 + c] : T(0);
        }
        block_sync();
        for (int j = 0; j < kb; ++j) {
            // argmax |s_p[i][j]| for i >= j
            T ba = T(-1);
            int bi = 0x7fffffff;
            for (int i = j + tid; i < mr; i += BT) {
                const T a = fabs(s_p[i * NB + j]);
                if (bt_better(a, i, ba, bi)) { ba = a; bi = i; }
            }
            #pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const T oa = __shfl_xor(ba, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (bt_better(oa, oi, ba, bi)) { ba = oa; bi = oi; }
            }
            if (lane == 0) { s_red[wave] = ba; s_ired[wave] = bi; }
            block_sync();
            ba = s_red[0]; bi = s_ired[0];
            #pragma unroll
            for (int w = 1; w < BT / 64; ++w)
                if (bt_better(s_red[w], s_ired[w], ba, bi)) { ba = s_red[w]; bi = s_ired[w]; }
            const int p = (bi == 0x7fffffff) ? j : bi;
            if (tid == 0) {
                piv[k0 + j] = k0 + p;
                if (s_p[p * NB + j] == T(0) && *info == 0) *info = k0 + j + 1;
            }
            // swap rows j and p of the panel
            if (p != j && tid < NB) {
                const T t = s_p[j * NB + tid];
                s_p[j * NB + tid] = s_p[p * NB + tid];
                s_p[p * NB + tid] = t;
            }
            block_sync();
            const T pv = s_p[j * NB + j];
            const bool tiny = fabs(pv) < safe_min;
            const T inv = T(1) / pv;
            for (int i = j + 1 + tid; i < mr; i += BT) {
                T l = s_p[i * NB + j];
                if (pv != T(0)) l = tiny ? l / pv : l * inv;
                s_p[i * NB + j] = l;
                for (int c = j + 1; c < kb; ++c) s_p[i * NB + c] -= l * s_p[j * NB + c];
            }
            block_sync();
        }
        // panel back to W
        for (int e = tid; e < mr * NB; e += BT) {
            const int i = e / NB, c = e % NB;
            if (c < kb) W[(k0 + i) * ldw + k0 + c] = s_p[e];
        }
        // interchanges of this panel on the columns left and right of it
        for (int j = 0; j < kb; ++j) {
            const int p = piv[k0 + j];
            block_sync();                                  // piv visible, previous swap done
            if (p != k0 + j) {
                for (int c = tid; c < n; c += BT) {
                    if (c >= k0 && c < k0 + kb) continue;
                    const T t = W[(k0 + j) * ldw + c];
                    W[(k0 + j) * ldw + c] = W[p * ldw + c];
                    W[p * ldw + c] = t;
                }
            }
        }
        block_sync();
        const int c1 = k0 + kb, nr = n - c1;
        if (nr <= 0) break;
        // U12 = L11^{-1} A12: forward substitution down the kb rows, one
        // column per thread (the column stays in the L1 cache)
        for (int c = c1 + tid; c < n; c += BT) {
            for (int r = 1; r < kb; ++r) {
                T s = W[(k0 + r) * ldw + c];
                for (int q = 0; q < r; ++q) s -= s_p[r * NB + q] * W[(k0 + q) * ldw + c];
                W[(k0 + r) * ldw + c] = s;
            }
        }
        block_sync();
        // A22 -= L21 U12: each thread a 4 x 4 block per pass over the trailing matrix
        const int tr = (nr + 3) / 4;                       // 4-row groups
        const int tc = (nr + 3) / 4;
        for (int t = tid; t < tr * tc; t += BT) {
            const int bi = (t / tc) * 4, bj = (t % tc) * 4;
            T acc[4][4];
            #pragma unroll
            for (int a = 0; a < 4; ++a)
                #pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
            for (int q = 0; q < kb; ++q) {
                T l[4], u[4];
                #pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = bi + a;
                    l[a] = i < nr ? s_p[(kb + i) * NB + q] : T(0);
                }
                #pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int jj = bj + b;
                    u[b] = jj < nr ? W[(k0 + q) * ldw + c1 + jj] : T(0);
                }
                #pragma unroll
                for (int a = 0; a < 4; ++a)
                    #pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
            }
            #pragma unroll
            for (int a = 0; a < 4; ++a)
                #pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int i = bi + a, jj = bj + b;
                    if (i < nr && jj < nr) W[(c1 + i) * ldw + c1 + jj] -= acc[a][b];
                }
        }
        block_sync();
    }
}

// x := op(LU)^{-1} x for the factors in W (row-major), x of length n in
// global memory.  Lower (unit) then upper, 64-row blocks: one wave solves the
// diagonal block, the workgroup updates the rows below (above).
template <class T, class V>
__device__ __noinline__ void wg_trsv(const T* __restrict__ W, int n, int ldw, V* __restrict__ x, bool lower)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (n + 63) / 64;
    for (int bb = 0; bb < nblk; ++bb) {
        const int blk = lower ? bb : nblk - 1 - bb;
        const int b0 = blk * 64, bn = min(64, n - b0);
        if (wave == 0) {
            V y = lane < bn ? x[b0 + lane] : V(0);
            if (lower) {
                for (int j = 0; j < bn; ++j) {
                    const V yj = __shfl(y, j, 64);
                    if (lane > j && lane < bn) y -= V(W[(b0 + lane) * ldw + b0 + j]) * yj;
                }
            }
            else {
                for (int j = bn - 1; j >= 0; --j) {
                    if (lane == j) y /= V(W[(b0 + j) * ldw + b0 + j]);
                    const V yj = __shfl(y, j, 64);
                    if (lane < j) y -= V(W[(b0 + lane) * ldw + b0 + j]) * yj;
                }
            }
            if (lane < bn) x[b0 + lane] = y;
        }
        block_sync();
        // rows outside the block: x_i -= sum_j W[i][b0 + j] y_j
        const int r0 = lower ? b0 + bn : 0, r1 = lower ? n : b0;
        for (int i = r0 + tid; i < r1; i += BT) {
            V s = V(0);
            for (int j = 0; j < bn; ++j) s += V(W[i * ldw + b0 + j]) * x[b0 + j];
            x[i] -= s;
        }
        block_sync();
    }
}

struct BatchArgs {
    int count;
    const int* order;               // systems, largest first
    const int* n;                   // sizes
    const int* mode;                // per system: 0 float factor + FP64 refinement, 1 FP64 factor
    const int64_t* offA;            // A_i at A + offA[i], column-major, ld n_i (FP64)
    const int64_t* offB;            // b_i / x_i at B + offB[i]
    const double* A;
    const double* B;
    double* X;
    int* iters;                     // refinement steps (>= 0), -(itermax+1) fallback, -3 singular factor
    int* info;                      // factorization info (FP64 path or fallback)
    double* berr;                   // ||b - A x||_inf / (||A||_inf ||x||_inf)
    int* queue;                     // work counter (zeroed)
    char* work;                     // per-workgroup workspace
    int64_t work_stride;            // bytes per workgroup
    int itermax;
    double tol;                     // <= 0: eps * sqrt(n)
};

// r = b - A x (FP64, A column-major) and the infinity norms; returns ||r||_inf
// in s_dred[0], ||x||_inf in s_dred[1].
__device__ void wg_residual(const double* __restrict__ A, int n, const double* __restrict__ b,
                            const double* __restrict__ x, double* __restrict__ r, double* s_dred)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double rmax = 0, xmax = 0;
    for (int i = tid; i < n; i += BT) {
        double s = b[i];
        for (int j = 0; j < n; ++j) s -= A[i + int64_t(j) * n] * x[j];
        r[i] = s;
        rmax = fmax(rmax, fabs(s));
        xmax = fmax(xmax, fabs(x[i]));
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        rmax = fmax(rmax, __shfl_xor(rmax, o, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, o, 64));
    }
    block_sync();
    if (lane == 0) { s_dred[2 + wave] = rmax; s_dred[6 + wave] = xmax; }
    block_sync();
    if (tid == 0) {
        double a = 0, c = 0;
        for (int w = 0; w < BT / 64; ++w) { a = fmax(a, s_dred[2 + w]); c = fmax(c, s_dred[6 + w]); }
        s_dred[0] = a;
        s_dred[1] = c;
    }
    block_sync();
}

// ||A||_inf (FP64 column-major) into s_dred[0].
__device__ void wg_norm_inf(const double* __restrict__ A, int n, double* s_dred)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double m = 0;
    for (int i = tid; i < n; i += BT) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += fabs(A[i + int64_t(j) * n]);
        m = fmax(m, s);
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    block_sync();
    if (lane == 0) s_dred[2 + wave] = m;
    block_sync();
    if (tid == 0) {
        double a = 0;
        for (int w = 0; w < BT / 64; ++w) a = fmax(a, s_dred[2 + w]);
        s_dred[0] = a;
    }
    block_sync();
}

// Factor (precision T) and solve every system of the batch.
template <class T>
__global__ __launch_bounds__(BT)
void batched_solve_kernel(BatchArgs a)
{
    constexpr int NB = BNB<T>::v;
    __shared__ T s_p[BMAXN * NB];
    __shared__ T s_red[BT / 64];
    __shared__ int s_ired[BT / 64];
    __shared__ double s_dred[16];
    __shared__ int s_sys, s_info;
    const int tid = threadIdx.x;
    char* ws = a.work + int64_t(blockIdx.x) * a.work_stride;
    // workspace: Wlo (n*n T) | W64 (n*n double, fallback) | piv (n int) | r, y (n double)
    T* W = reinterpret_cast<T*>(ws);
    double* W64 = reinterpret_cast<double*>(ws + int64_t(BMAXN) * BMAXN * sizeof(T));
    int* piv = reinterpret_cast<int*>(W64 + int64_t(BMAXN) * BMAXN);
    double* r = reinterpret_cast<double*>(piv + BMAXN);
    double* y = r + BMAXN;
    const double eps = std::numeric_limits<double>::epsilon();
    while (true) {
        if (tid == 0) {
            const int q = atomicAdd(a.queue, 1);
            s_sys = q < a.count ? a.order[q] : -1;
            s_info = 0;
        }
        block_sync();
        const int sys = s_sys;
        if (sys < 0) break;                               // every wave leaves here
        const int n = a.n[sys];
        const double* A = a.A + a.offA[sys];
        const double* b = a.B + a.offB[sys];
        double* x = a.X + a.offB[sys];
        // copy / cast into the row-major working matrix
        for (int e = tid; e < n * n && !(sizeof(T) != 8 && a.mode && a.mode[sys] == 1); e += BT) {
            const int i = e / n, j = e % n;
            W[i * n + j] = T(A[i + int64_t(j) * n]);
        }
        const bool direct64 = sizeof(T) != 8 && a.mode && a.mode[sys] == 1;
        block_sync();
        if (!direct64) wg_getrf<T>(W, n, n, piv, &s_info, s_p, s_red, s_ired);
        block_sync();
        int info = s_info;
        int iters = 0;
        bool need64 = direct64;
        // x = (LU)^{-1} P b, in FP64 storage
        auto lo_solve = [&](const double* rhs, double* out) {
            for (int i = tid; i < n; i += BT) out[i] = rhs[i];
            block_sync();
            if (tid == 0)
                for (int j = 0; j < n; ++j) {
                    const int p = piv[j];
                    if (p != j) { const double t = out[j]; out[j] = out[p]; out[p] = t; }
                }
            block_sync();
            if constexpr (sizeof(T) == 8) {
                wg_trsv<T, double>(W, n, n, out, true);
                wg_trsv<T, double>(W, n, n, out, false);
            }
            else {
                wg_trsv<T, double>(W, n, n, out, true);
                wg_trsv<T, double>(W, n, n, out, false);
            }
        };
        if (direct64) {
            // the FP64 factorization below is the chosen path, not a fallback
        }
        else if (info != 0) {
            need64 = sizeof(T) != 8;
            iters = -3;
        }
        else {
            lo_solve(b, x);
            if constexpr (sizeof(T) != 8) {
                wg_norm_inf(A, n, s_dred);
                const double anorm = s_dred[0];
                const double tol = a.tol > 0 ? a.tol : eps * sqrt(double(n));
                bool conv = false;
                for (int it = 0; it <= a.itermax; ++it) {
                    wg_residual(A, n, b, x, r, s_dred);
                    if (s_dred[0] <= s_dred[1] * anorm * tol) { iters = it; conv = true; break; }
                    if (it == a.itermax) break;
                    lo_solve(r, y);
                    for (int i = tid; i < n; i += BT) x[i] += y[i];
                    block_sync();
                }
                if (!conv) { need64 = true; iters = -a.itermax - 1; }
            }
        }
        if (need64) {
            // FP64 fallback: the same steps in double
            for (int e = tid; e < n * n; e += BT) {
                const int i = e / n, j = e % n;
                W64[i * n + j] = A[i + int64_t(j) * n];
            }
            if (tid == 0) s_info = 0;
            block_sync();
            wg_getrf<double>(W64, n, n, piv, &s_info, reinterpret_cast<double*>(s_p),
                             reinterpret_cast<double*>(s_dred + 8), s_ired);
            block_sync();
            info = s_info;
            for (int i = tid; i < n; i += BT) x[i] = b[i];
            block_sync();
            if (tid == 0)
                for (int j = 0; j < n; ++j) {
                    const int p = piv[j];
                    if (p != j) { const double t = x[j]; x[j] = x[p]; x[p] = t; }
                }
            block_sync();
            wg_trsv<double, double>(W64, n, n, x, true);
            wg_trsv<double, double>(W64, n, n, x, false);
        }
        // backward error
        wg_norm_inf(A, n, s_dred);
        const double anorm = s_dred[0];
        wg_residual(A, n, b, x, r, s_dred);
        if (tid == 0) {
            a.iters[sys] = iters;
            a.info[sys] = info;
            const double den = anorm * s_dred[1];
            a.berr[sys] = den > 0 ? s_dred[0] / den : 0.0;
        }
        block_sync();
    }
}

int64_t batched_work_stride(int elem_bytes)
{
    int64_t b = int64_t(BMAXN) * BMAXN * (elem_bytes + 8) + int64_t(BMAXN) * (4 + 16) + 256;
    return (b + 255) & ~int64_t(255);
}

void batched_solve(int elem_bytes, int count, const int* order, const int* n, const int* mode,
                   const int64_t* offA, const int64_t* offB, const double* A, const double* B, double* X,
                   int* iters, int* info, double* berr, int* queue, char* work, int grid, int itermax,
                   double tol, hipStream_t s)
{
    if (count <= 0) return;
    BatchArgs a{count, order, n, mode, offA, offB, A, B, X, iters, info, berr, queue, work,
                batched_work_stride(elem_bytes), itermax, tol};
    zero_u32(reinterpret_cast<unsigned*>(queue), 1, s);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(batched_solve_kernel<double>, dim3(unsigned(grid)), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL(batched_solve_kernel<float>, dim3(unsigned(grid)), dim3(BT), 0, s, a);
    SA_CHECK_LAUNCH();
}

int batched_max_n() { return BMAXN; }

} // namespace sa
