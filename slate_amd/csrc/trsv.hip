// Single-launch triangular solve with one right-hand side (gfx950).
//
// The refinement loops of gesv_mixed / posv_mixed / the GMRES preconditioner
// solve with the low-precision factors once or twice per iteration and one
// vector at a time; SLATE does this with work::trsmA's block sweep, one
// batched trsm + gemm launch pair per block row (src/work/work_trsmA.cc,
// reached from src/getrs.cc:94-97).  Here the whole sweep is one launch:
// block q (BS rows) is handled by workgroup q, which accumulates the products
// with every already-solved block as soon as that block's solved piece has
// been published (a "look-back" chain: blocks finish in order), then applies
// the precomputed inverse of its diagonal block and publishes its own piece.
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
// Look-back sweep.  Block q publishes its solved piece twice: plainly into X
// (the result) and as tagged words (32-bit payload + the launch's tag) into
// xw, which is what later blocks read: a reader that sees the tag has the
// value, so there is no completion counter on the critical path and no
// cache-wide fence.  A separate progress word (tag, blocks done) only paces
// the waiting: blocks far from their turn sleep in proportion to the
// distance instead of polling.
//
// Each dependency's BS x BS tile of the factor is loaded into registers
// before its x arrives (it does not depend on x), so after the wait only the
// x words, one LDS pass and the FMAs remain.
template <class T>
__device__ inline void tv_put(uint64_t* p, T v, uint32_t tag) { tw_store_val<T>(p, v, tag); }

template <class T, int BS, int LOWER, int CM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void trsv_lookback_kernel(const T* __restrict__ M, int64_t ld, int64_t n, const T* __restrict__ Dinv,
                          T* __restrict__ X, uint64_t* __restrict__ xw, uint32_t tag,
                          uint64_t* __restrict__ progress, int* error)
{
    constexpr int NWV = int(sizeof(T)) / 4;
    constexpr int RPW = BS / 4;                 // rows per wave (row-major product)
    constexpr int CPL = BS / 64;                // columns per lane (row-major product)
    constexpr int NPART = 256 / BS;             // threads per row in the column-major product
    constexpr int KP = BS / NPART;              // columns per thread
    const int nblk = int(ceildiv(n, BS));
    const int q = blockIdx.x;
    const int b = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x, lane = tid & 63;
    // Wave-uniform, so a tile's row (row-major) or column (column-major)
    // addresses are scalar; that keeps the kernel at two waves per SIMD, so
    // every block of an n = 65536 solve is resident from the start.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    __shared__ uint32_t sxw[2][BS * NWV];       // x words of the dependency being applied
    __shared__ T sr[BS];
    __shared__ T spart[NPART][BS];
    __shared__ T sred[4][RPW][65];

    // Prefetch this block's inverse: thread (it, ih) holds inv(it, ih*KP .. ih*KP+KP-1).
    const int it = tid % BS, ih = (wave * 64) / BS;   // BS is a multiple of 64
    T dinv[KP];
    {
        const T* dp = Dinv + int64_t(b) * BS * BS;
        #pragma unroll
        for (int k = 0; k < KP; ++k) dinv[k] = dp[int64_t(ih * KP + k) * BS + it];
    }

    const int ndep = LOWER ? b : nblk - 1 - b;
    T acc_rm[RPW];
    T acc_cm = T(0);
    #pragma unroll
    for (int r = 0; r < RPW; ++r) acc_rm[r] = T(0);

    // Tile registers.  Loads use clamped addresses (edge values are zeroed
    // when applied), so a tile's loads issue back to back.
    T mrm[CM ? 1 : RPW][CM ? 1 : CPL];
    T mcm[CM ? KP : 1];
    auto load_tile = [&](int t) {
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        if constexpr (CM) {
            const int64_t row = r0 + (it < rows ? it : rows - 1);
            const T* mp = M + row + c0 * ld;
            #pragma unroll
            for (int c = 0; c < KP; ++c) {
                const int cc = ih * KP + c;
                mcm[c] = mp[int64_t(cc < cols ? cc : cols - 1) * ld];
            }
        }
        else {
            #pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int i = wave * RPW + r;
                const T* mp = M + (r0 + (i < rows ? i : rows - 1)) * ld + c0;
                #pragma unroll
                for (int u = 0; u < CPL; ++u) {
                    const int c = CPL * lane + u;
                    mrm[r][u] = mp[c < cols ? c : cols - 1];
                }
            }
        }
    };

    const uint64_t t_start = wall_clock64();
    bool failed = false;
    unsigned known = 0;                         // thread 0: chain progress last seen
    if (ndep > 0) load_tile(0);
    #pragma unroll 1
    for (int t = 0; t < ndep; ++t) {
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        const int buf = t & 1;
        // pace: sleep while the chain is more than two blocks away from t;
        // progress is re-read only while the last value seen is that far
        if (tid == 0 && !failed) {
            while (known + 2 < unsigned(t + 1)) {
                const uint64_t pw = tw_load(progress);
                const unsigned done = (uint32_t(pw >> 32) == tag) ? uint32_t(pw) : 0u;
                known = done > known ? done : known;
                if (known + 2 >= unsigned(t + 1)) break;
                const unsigned dist = unsigned(t + 1) - known;
                const unsigned naps = dist > 16 ? 8 : dist / 2;
                for (unsigned z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(32);
                if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0
                    || wall_clock64() - t_start > 400000000ull) {
                    __hip_atomic_store(error, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __syncthreads();
        // the dependency's x words (bounded spin inside tw_wait)
        if (tid < BS * NWV) {
            uint32_t wv = 0;
            if (tid < cols * NWV) wv = tw_wait(xw + c0 * NWV + tid, tag, error, 2);
            sxw[buf][tid] = wv;
        }
        __syncthreads();
        if (CM) {
            #pragma unroll
            for (int c = 0; c < KP; ++c) {
                const int cc = ih * KP + c;
                const T xv = lds_word_val<T>(&sxw[buf][cc * NWV]);
                acc_cm += (cc < cols) ? mcm[c] * xv : T(0);
            }
        }
        else {
            #pragma unroll
            for (int u = 0; u < CPL; ++u) {
                const int c = CPL * lane + u;
                const T xv = lds_word_val<T>(&sxw[buf][c * NWV]);
                const T y = (c < cols) ? xv : T(0);
                #pragma unroll
                for (int r = 0; r < RPW; ++r) acc_rm[r] += mrm[r][u] * y;
            }
        }
        if (t + 1 < ndep) load_tile(t + 1);
        failed = failed || (tid == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0);
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
    if (tid < rows) {
        T yv = T(0);
        #pragma unroll
        for (int p = 0; p < NPART; ++p) yv += spart[p][tid];
        tv_put<T>(xw + (r0 + tid) * NWV, yv, tag);
        X[r0 + tid] = yv;
    }
    if (tid == 0)
        __hip_atomic_fetch_max(progress, (uint64_t(tag) << 32) | uint32_t(q + 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
void trsv_lookback(const T* M, int64_t ld, int cm, int lower, int64_t n, const T* Dinv, T* X,
                   const TvSync& sync, hipStream_t s)
{
    if (n <= 0) return;
    constexpr int BS = TvBS<T>::v;
    SA_REQUIRE(sync.tag != 0, -2, "trsv_lookback: tag 0 is reserved");
    unsigned nblk = unsigned(ceildiv(n, BS));
    int key = (lower ? 2 : 0) | (cm ? 1 : 0);
#define SA_TV_CASE(K, LO, CMV)                                                                      \
    case K:                                                                                         \
        hipLaunchKernelGGL((trsv_lookback_kernel<T, BS, LO, CMV>), dim3(nblk), dim3(256), 0, s,    \
                           M, ld, n, Dinv, X, sync.words, sync.tag, sync.progress, sync.error);    \
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
                                    const TvSync&, hipStream_t);
template void trsv_lookback<float>(const float*, int64_t, int, int, int64_t, const float*, float*,
                                   const TvSync&, hipStream_t);

} // namespace sa
