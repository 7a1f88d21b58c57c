// Single-launch triangular solve with one right-hand side (gfx950).
//
// The refinement loops of gesv_mixed / posv_mixed / the GMRES preconditioner
// solve with the low-precision factors once or twice per iteration and one
// vector at a time; SLATE does this with work::trsmA's block sweep, one
// batched trsm + gemm launch pair per block row (src/work/work_trsmA.cc,
// reached from src/getrs.cc:94-97).  Here the whole sweep is one launch:
// block q (BS rows) is handled by workgroup q, which accumulates the products
// with every already-solved block as soon as a completion counter says it is
// ready (a "look-back" chain: blocks finish in order, so one monotonic counter
// is enough), then applies the precomputed inverse of its diagonal block,
// publishes its piece of the solution and bumps the counter.
//
// The diagonal blocks are inverted once per factor by trsv_diag_inverse (the
// inverted-diagonal-block form of trsm that GPU BLAS libraries use), so the
// chain's critical path per block is: look-back wait, one BS x BS tile
// product, an LDS row reduction and a BS x BS matrix-vector product whose
// matrix was prefetched into registers before the wait.
//
// Workgroups are dispatched in index order and only wait on lower indices,
// so the chain always progresses; every wait is bounded and reports a timeout
// through an error flag instead of hanging.
#include "common.hh"
#include "kernels.hh"

namespace sa {

template <class T> struct TvBS { static constexpr int v = 128; };
template <> struct TvBS<double> { static constexpr int v = 64; };

int64_t trsv_block(int elem_bytes) { return elem_bytes == 8 ? TvBS<double>::v : TvBS<float>::v; }

//------------------------------------------------------------------------------
// Dinv[b] = inverse of diagonal block b (padded with identity past n), stored
// so that element (t, k) of block b is at Dinv[b*BS*BS + k*BS + t].
// tri(i, k) = M[i*ld + k] (CM == 0) or M[i + k*ld] (CM == 1).
template <class T, int BS, int LOWER, int UNIT, int CM>
__global__ __launch_bounds__(256)
void diag_inverse_kernel(const T* __restrict__ M, int64_t ld, int64_t n, T* __restrict__ Dinv)
{
    __shared__ T Tl[BS][BS + 1];
    __shared__ T Xs[BS][BS + 1];          // Xs[i][t]: row i of inverse column t
    const int b = blockIdx.x;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x;
    for (int e = tid; e < BS * BS; e += 256) {
        int i = e / BS, k = e % BS;
        T v;
        if (i < rows && k < rows) v = CM ? M[(r0 + i) + (r0 + k) * ld] : M[(r0 + i) * ld + r0 + k];
        else v = (i == k) ? T(1) : T(0);
        Tl[i][k] = v;
    }
    __syncthreads();
    // thread t < BS solves tri * x = e_t by substitution, x kept in Xs[:, t]
    if (tid < BS) {
        const int t = tid;
        if (LOWER) {
            for (int i = 0; i < BS; ++i) {
                T s = (i == t) ? T(1) : T(0);
                for (int k = t; k < i; ++k) s -= Tl[i][k] * Xs[k][t];
                if (i < t) s = T(0);
                Xs[i][t] = UNIT ? s : s / Tl[i][i];
            }
        }
        else {
            for (int i = BS - 1; i >= 0; --i) {
                T s = (i == t) ? T(1) : T(0);
                for (int k = i + 1; k <= t; ++k) s -= Tl[i][k] * Xs[k][t];
                if (i > t) s = T(0);
                Xs[i][t] = UNIT ? s : s / Tl[i][i];
            }
        }
    }
    __syncthreads();
    T* out = Dinv + int64_t(b) * BS * BS;
    for (int e = tid; e < BS * BS; e += 256) {
        int k = e / BS, t = e % BS;          // out[k*BS + t] = inv(t, k)
        out[e] = Xs[t][k];
    }
}

template <class T>
void trsv_diag_inverse(const T* M, int64_t ld, int cm, int lower, int unit, int64_t n, T* Dinv,
                       hipStream_t s)
{
    if (n <= 0) return;
    constexpr int BS = TvBS<T>::v;
    unsigned nblk = unsigned(ceildiv(n, BS));
    int key = (lower ? 4 : 0) | (unit ? 2 : 0) | (cm ? 1 : 0);
#define SA_DI_CASE(K, LO, UN, CMV)                                                                  \
    case K:                                                                                         \
        hipLaunchKernelGGL((diag_inverse_kernel<T, BS, LO, UN, CMV>), dim3(nblk), dim3(256), 0, s, \
                           M, ld, n, Dinv);                                                         \
        break;
    switch (key) {
        SA_DI_CASE(0, 0, 0, 0)
        SA_DI_CASE(1, 0, 0, 1)
        SA_DI_CASE(2, 0, 1, 0)
        SA_DI_CASE(3, 0, 1, 1)
        SA_DI_CASE(4, 1, 0, 0)
        SA_DI_CASE(5, 1, 0, 1)
        SA_DI_CASE(6, 1, 1, 0)
        default:
        SA_DI_CASE(7, 1, 1, 1)
    }
#undef SA_DI_CASE
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
template <class T, int BS, int LOWER, int CM>
__global__ __launch_bounds__(256)
void trsv_lookback_kernel(const T* __restrict__ M, int64_t ld, int64_t n, const T* __restrict__ Dinv,
                          T* __restrict__ X, unsigned* ndone, int* error)
{
    constexpr int RPW = BS / 4;                 // rows per wave (row-major product)
    constexpr int NPART = 256 / BS;             // threads per row in the column-major product
    constexpr int KP = BS / NPART;              // columns per thread
    const int nblk = int(ceildiv(n, BS));
    const int q = blockIdx.x;
    const int b = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    __shared__ T sr[BS];
    __shared__ T spart[NPART][BS];
    __shared__ T sred[4][RPW][65];
    __shared__ unsigned s_ready;

    // Prefetch this block's inverse: thread (it, ih) holds inv(it, ih*KP .. ih*KP+KP-1).
    const int it = tid % BS, ih = tid / BS;
    T dinv[KP];
    {
        const T* dp = Dinv + int64_t(b) * BS * BS;
        #pragma unroll
        for (int k = 0; k < KP; ++k) dinv[k] = dp[int64_t(ih * KP + k) * BS + it];
    }

    // ---- products with the already-solved blocks
    const int ndep = LOWER ? b : nblk - 1 - b;
    T acc_rm[RPW];
    T acc_cm = T(0);
    #pragma unroll
    for (int r = 0; r < RPW; ++r) acc_rm[r] = T(0);

    unsigned ready = 0;                         // uniform private copy of s_ready
    for (int t = 0; t < ndep; ++t) {
        if (t >= int(ready)) {
            if (tid == 0) {
                uint64_t t0 = wall_clock64();
                unsigned v;
                while ((v = __hip_atomic_load(ndone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) < unsigned(t + 1)) {
                    // blocks far from their turn poll rarely, so the few that
                    // are next see the counter move without contention
                    if (unsigned(t + 1) - v > 2) __builtin_amdgcn_s_sleep(16);
                    else __builtin_amdgcn_s_sleep(1);
                    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { v = unsigned(ndep); break; }
                    if (wall_clock64() - t0 > 400000000ull) {
                        __hip_atomic_store(error, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = unsigned(ndep);
                        break;
                    }
                }
                sa_publish_fence();
                s_ready = v;
            }
            __syncthreads();
            ready = s_ready;
            __syncthreads();                    // s_ready is rewritten only after everyone read it
        }
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        // All loads below are unconditional with clamped addresses (edge
        // values are zeroed afterwards), so the compiler issues a whole tile's
        // loads back to back instead of one guarded round trip per element.
        if (CM) {
            // each thread reads the x entries of its own columns (the same
            // addresses across the BS threads of a column group: one
            // broadcast request), so a tile needs no LDS staging or barrier
            const int64_t row = r0 + (it < rows ? it : rows - 1);
            const T* mp = M + row + c0 * ld;
            #pragma unroll 16
            for (int c = 0; c < KP; ++c) {
                const int cc = ih * KP + c;
                const int ccl = cc < cols ? cc : cols - 1;
                const T a = mp[int64_t(ccl) * ld];
                const T xv = X[c0 + ccl];
                acc_cm += (cc < cols) ? a * xv : T(0);
            }
        }
        else {
            // lanes run along the columns (BS/64 per lane), the wave over RPW rows
            constexpr int CPL = BS / 64;
            T y[CPL];
            int cl[CPL];
            #pragma unroll
            for (int u = 0; u < CPL; ++u) {
                const int c = CPL * lane + u;
                cl[u] = c < cols ? c : cols - 1;
                T v = X[c0 + cl[u]];
                y[u] = (c < cols) ? v : T(0);
            }
            #pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int i = wave * RPW + r;
                const T* mp = M + (r0 + (i < rows ? i : rows - 1)) * ld + c0;
                #pragma unroll
                for (int u = 0; u < CPL; ++u) acc_rm[r] += mp[cl[u]] * y[u];
            }
        }
    }
    // ---- residual of this block's rows
    if (CM) {
        spart[ih][it] = acc_cm;
        __syncthreads();
        if (tid < BS) {
            T sum = T(0);
            #pragma unroll
            for (int p = 0; p < NPART; ++p) sum += spart[p][tid];
            T xv = X[r0 + (tid < rows ? tid : rows - 1)];
            sr[tid] = (tid < rows) ? xv - sum : T(0);
        }
    }
    else {
        // transpose-reduce the per-lane partials through LDS
        #pragma unroll
        for (int r = 0; r < RPW; ++r) sred[wave][r][lane] = acc_rm[r];
        __syncthreads();
        if (tid < BS) {
            const int w = tid / RPW, r = tid % RPW;
            T sum = T(0);
            #pragma unroll 16
            for (int l = 0; l < 64; ++l) sum += sred[w][r][l];
            T xv = X[r0 + (tid < rows ? tid : rows - 1)];
            sr[tid] = (tid < rows) ? xv - sum : T(0);
        }
    }
    __syncthreads();
    // ---- y = Dinv_b * r (partials over this thread's k range, folded in LDS)
    {
        T acc = T(0);
        #pragma unroll
        for (int k = 0; k < KP; ++k) acc += dinv[k] * sr[ih * KP + k];
        spart[ih][it] = acc;
    }
    __syncthreads();
    if (tid < BS) {
        T yv = T(0);
        #pragma unroll
        for (int p = 0; p < NPART; ++p) yv += spart[p][tid];
        if (tid < rows) X[r0 + tid] = yv;
    }
    __syncthreads();
    if (tid == 0) {
        sa_publish_fence();
        __hip_atomic_store(ndone, unsigned(q + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class T>
void trsv_lookback(const T* M, int64_t ld, int cm, int lower, int64_t n, const T* Dinv, T* X,
                   unsigned* ndone, int* error, hipStream_t s)
{
    if (n <= 0) return;
    constexpr int BS = TvBS<T>::v;
    zero_u32(ndone, 1, s);
    unsigned nblk = unsigned(ceildiv(n, BS));
    int key = (lower ? 2 : 0) | (cm ? 1 : 0);
#define SA_TV_CASE(K, LO, CMV)                                                                      \
    case K:                                                                                         \
        hipLaunchKernelGGL((trsv_lookback_kernel<T, BS, LO, CMV>), dim3(nblk), dim3(256), 0, s,    \
                           M, ld, n, Dinv, X, ndone, error);                                        \
        break;
    switch (key) {
        SA_TV_CASE(0, 0, 0)
        SA_TV_CASE(1, 0, 1)
        SA_TV_CASE(2, 1, 0)
        default:
        SA_TV_CASE(3, 1, 1)
    }
#undef SA_TV_CASE
    SA_CHECK_LAUNCH();
}

template void trsv_diag_inverse<double>(const double*, int64_t, int, int, int, int64_t, double*, hipStream_t);
template void trsv_diag_inverse<float>(const float*, int64_t, int, int, int, int64_t, float*, hipStream_t);
template void trsv_lookback<double>(const double*, int64_t, int, int, int64_t, const double*, double*,
                                    unsigned*, int*, hipStream_t);
template void trsv_lookback<float>(const float*, int64_t, int, int, int64_t, const float*, float*,
                                   unsigned*, int*, hipStream_t);

} // namespace sa
