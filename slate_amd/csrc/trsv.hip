// Single-launch triangular solve with one right-hand side (gfx950).
//
// The refinement loops of gesv_mixed / posv_mixed / the GMRES preconditioner
// solve with the low-precision factors once or twice per iteration and one
// vector at a time; SLATE does this with work::trsmA's block sweep, one
// batched trsm + gemm launch pair per block row (src/work/work_trsmA.cc,
// reached from src/getrs.cc:94-97).  Here the whole sweep is one launch: block
// q (128 rows) is handled by workgroup q, which accumulates the products with
// every already-solved block as soon as a completion counter says it is ready
// (a "look-back" chain: blocks finish in order, so one monotonic counter is
// enough), then solves its diagonal block in one wavefront with lane
// broadcasts, publishes its piece of the solution and bumps the counter.
// Workgroups are dispatched in index order and only wait on lower indices,
// so the chain always progresses; every wait is bounded and reports a timeout
// through an error flag instead of hanging.
#include "common.hh"
#include "kernels.hh"

namespace sa {

static constexpr int TV_BS = 128;

// tri(i, k) = M[i*ld + k] (CM == 0) or M[i + k*ld] (CM == 1), global indices.
template <class T, int LOWER, int UNIT, int CM>
__global__ __launch_bounds__(256)
void trsv_lookback_kernel(const T* __restrict__ M, int64_t ld, int64_t n, T* __restrict__ X,
                          unsigned* ndone, int* error)
{
    constexpr int BS = TV_BS;
    const int nblk = int(ceildiv(n, BS));
    const int q = blockIdx.x;
    const int b = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    __shared__ T sdiag[BS * (BS + 1) / 2];
    __shared__ T sy[BS];
    __shared__ T sr[BS];
    __shared__ T spart[2][BS];
    __shared__ unsigned s_ready;

    auto tri = [&](int64_t i, int64_t k) -> T {
        return CM ? M[i + k * ld] : M[i * ld + k];
    };
    // packed diagonal triangle: lower (k <= i) at i(i+1)/2 + k;
    // upper (k >= i) at (BS-1-i)(BS-i)/2 + (BS-1-k)
    auto pidx = [&](int i, int k) -> int {
        return LOWER ? i * (i + 1) / 2 + k : (BS - 1 - i) * (BS - i) / 2 + (BS - 1 - k);
    };
    for (int e = tid; e < BS * BS; e += 256) {
        int i = e / BS, k = e % BS;
        bool in = LOWER ? (k <= i) : (k >= i);
        if (in) sdiag[pidx(i, k)] = (i < rows && k < rows) ? tri(r0 + i, r0 + k) : T(i == k ? 1 : 0);
    }
    __syncthreads();

    // ---- products with the already-solved blocks
    const int ndep = LOWER ? b : nblk - 1 - b;
    T acc_rm[32];                       // CM == 0: wave owns 32 rows, lanes run along columns
    T acc_cm = T(0);                    // CM == 1: thread owns one row, half the columns
    #pragma unroll
    for (int r = 0; r < 32; ++r) acc_rm[r] = T(0);

    unsigned ready = 0;                 // uniform private copy of s_ready
    for (int t = 0; t < ndep; ++t) {
        if (t >= int(ready)) {
            if (tid == 0) {
                uint64_t t0 = wall_clock64();
                unsigned v;
                while ((v = __hip_atomic_load(ndone, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) < unsigned(t + 1)) {
                    __builtin_amdgcn_s_sleep(1);
                    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { v = unsigned(ndep); break; }
                    if (wall_clock64() - t0 > 400000000ull) {
                        __hip_atomic_store(error, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = unsigned(ndep);
                        break;
                    }
                }
                __threadfence();
                s_ready = v;
            }
            __syncthreads();
            ready = s_ready;
            __syncthreads();            // s_ready is rewritten only after everyone read it
        }
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        if (CM) {
            // y_j into LDS, then thread (row, half) walks its 64 columns
            if (tid < BS) sy[tid] = (tid < cols) ? X[c0 + tid] : T(0);
            __syncthreads();
            const int row = tid & (BS - 1), half = tid >> 7;
            if (row < rows) {
                const T* mp = M + (r0 + row) + (c0 + half * 64) * ld;
                #pragma unroll 8
                for (int c = 0; c < 64; ++c)
                    if (half * 64 + c < cols) acc_cm += mp[int64_t(c) * ld] * sy[half * 64 + c];
            }
            __syncthreads();
        }
        else {
            const int c = 2 * lane;
            const T y0 = (c < cols) ? X[c0 + c] : T(0);
            const T y1 = (c + 1 < cols) ? X[c0 + c + 1] : T(0);
            #pragma unroll
            for (int r = 0; r < 32; ++r) {
                const int i = wave * 32 + r;
                if (i < rows) {
                    const T* mp = M + (r0 + i) * ld + c0 + c;
                    T l0 = (c < cols) ? mp[0] : T(0);
                    T l1 = (c + 1 < cols) ? mp[1] : T(0);
                    acc_rm[r] += l0 * y0 + l1 * y1;
                }
            }
        }
    }
    // ---- residual of this block's rows
    if (CM) {
        const int row = tid & (BS - 1), half = tid >> 7;
        spart[half][row] = acc_cm;
        __syncthreads();
        if (tid < BS) sr[tid] = (tid < rows) ? X[r0 + tid] - (spart[0][tid] + spart[1][tid]) : T(0);
    }
    else {
        #pragma unroll
        for (int r = 0; r < 32; ++r) {
            T v = wave_sum(acc_rm[r]);
            const int i = wave * 32 + r;
            if (lane == 0) sr[i] = (i < rows) ? X[r0 + i] - v : T(0);
        }
    }
    __syncthreads();

    // ---- diagonal block in one wavefront: lane owns rows lane and lane + 64
    if (wave == 0) {
        T ra = sr[lane], rb = sr[lane + 64];
        if (LOWER) {
            for (int k = 0; k < BS; ++k) {
                T own = (k < 64) ? ra : rb;
                T yk = __shfl(own, k & 63, 64);
                if (!UNIT) yk = yk / sdiag[pidx(k, k)];
                if (lane == (k & 63)) { if (k < 64) ra = yk; else rb = yk; }
                if (lane > k) ra -= sdiag[pidx(lane, k)] * yk;
                if (lane + 64 > k) rb -= sdiag[pidx(lane + 64, k)] * yk;
            }
        }
        else {
            for (int k = BS - 1; k >= 0; --k) {
                T own = (k < 64) ? ra : rb;
                T yk = __shfl(own, k & 63, 64);
                if (!UNIT) yk = yk / sdiag[pidx(k, k)];
                if (lane == (k & 63)) { if (k < 64) ra = yk; else rb = yk; }
                if (lane < k) ra -= sdiag[pidx(lane, k)] * yk;
                if (lane + 64 < k) rb -= sdiag[pidx(lane + 64, k)] * yk;
            }
        }
        if (lane < rows) X[r0 + lane] = ra;
        if (lane + 64 < rows) X[r0 + lane + 64] = rb;
        __threadfence();
        if (lane == 0)
            __hip_atomic_store(ndone, unsigned(q + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class T>
void trsv_lookback(const T* M, int64_t ld, int cm, int lower, int unit, int64_t n, T* X,
                   unsigned* ndone, int* error, hipStream_t s)
{
    if (n <= 0) return;
    zero_u32(ndone, 1, s);
    unsigned nblk = unsigned(ceildiv(n, TV_BS));
    int key = (lower ? 4 : 0) | (unit ? 2 : 0) | (cm ? 1 : 0);
#define SA_TV_CASE(K, LO, UN, CMV)                                                                  \
    case K:                                                                                         \
        hipLaunchKernelGGL((trsv_lookback_kernel<T, LO, UN, CMV>), dim3(nblk), dim3(256), 0, s,    \
                           M, ld, n, X, ndone, error);                                              \
        break;
    switch (key) {
        SA_TV_CASE(0, 0, 0, 0)
        SA_TV_CASE(1, 0, 0, 1)
        SA_TV_CASE(2, 0, 1, 0)
        SA_TV_CASE(3, 0, 1, 1)
        SA_TV_CASE(4, 1, 0, 0)
        SA_TV_CASE(5, 1, 0, 1)
        SA_TV_CASE(6, 1, 1, 0)
        default:
        SA_TV_CASE(7, 1, 1, 1)
    }
#undef SA_TV_CASE
    SA_CHECK_LAUNCH();
}

template void trsv_lookback<double>(const double*, int64_t, int, int, int, int64_t, double*, unsigned*,
                                    int*, hipStream_t);
template void trsv_lookback<float>(const float*, int64_t, int, int, int, int64_t, float*, unsigned*,
                                   int*, hipStream_t);

} // namespace sa
