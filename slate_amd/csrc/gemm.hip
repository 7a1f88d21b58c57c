// MFMA tile GEMM for gfx950: C = alpha * A * B + beta * C, C row-major.
//
// This is the kernel behind every level-3 update of the factorizations (the
// trailing-matrix update of getrf, the herk/gemm update of potrf, the inner
// updates of the recursive panel and of the blocked triangular solves).  It
// replaces the vendor blas::batch::gemm / batch::herk calls of
// src/internal/internal_gemm.cc:498-504 and src/internal/internal_herk.cc:491-516:
// one launch covers the whole trailing matrix instead of a batch of nb x nb
// tiles, so there are no host-built pointer arrays and no per-tile launches.
//
// Shape: WM x WN wave64s (8 waves = 2 per SIMD for the main update kernels);
// each wave owns a (BM/WM) x (BN/WN) block of C held as 16x16 MFMA
// accumulators.  The K loop stages BK-deep slices of A and B through
// double-buffered LDS; the next slice is fetched into registers while the
// current one feeds the matrix cores and is written to LDS half-way through.
#include "common.hh"
#include "kernels.hh"
#include <type_traits>

namespace sa {

template <class T> struct Mf;

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator register r of lane l holds D[4r + l/16][l%16].
template <> struct Mf<double> {
    using acc = v4d;
    static constexpr int KM = 4;
    __device__ static acc op(double a, double b, acc c)
    { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int l, int r) { return 4 * r + (l >> 4); }
};

// v_mfma_f32_16x16x4f32: same operand layout; register r holds D[4(l/16) + r][l%16].
template <> struct Mf<float> {
    using acc = v4f;
    static constexpr int KM = 4;
    __device__ static acc op(float a, float b, acc c)
    { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// A layout: A_KC = K-contiguous (A[m*lda + k]), A_MC = M-contiguous (A[m + k*lda]).
// B layout: B_NC = N-contiguous (B[k*ldb + n]), B_KC = K-contiguous (B[k + n*ldb]).
// ROLE only names the instantiation: 1 = the trailing-matrix update of the
// blocked LU (profiled on its own), 0 = every other product.
// PF: prefetch depth in K slices.  PF = 2 keeps two register sets, so a
// slice's global loads are issued 1.5 slices before its LDS store (more
// latency cover at the cost of one more set of staging registers).
template <class T, int ALAY, int BLAY, int BM, int BN, int BK, int WM = 2, int WN = 2, int MINB = 1,
          int ROLE = 0, int PF = 1>
__global__ __launch_bounds__(WM * WN * 64, MINB)
void gemm_kernel(GemmArgs p)
{
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM, TN = BN / WN;       // per-wave C block
    constexpr int FM = TM / 16, FN = TN / 16;       // MFMA tiles per wave
    constexpr int KM = Mf<T>::KM;
    // LDS tiles mirror the global layout so the copy is a straight store.
    constexpr int A_CD = (ALAY == A_KC) ? BK : BM;  // contiguous extent
    constexpr int A_OD = (ALAY == A_KC) ? BM : BK;
    constexpr int B_CD = (BLAY == B_NC) ? BN : BK;
    constexpr int B_OD = (BLAY == B_NC) ? BK : BN;
    // Padding so fragment reads are bank-conflict free: 64-bit reads are served
    // 16 lanes at a time (one odd 8-byte pad suffices); 32-bit reads 32 lanes at
    // a time, where lanes 16..31 read the next k, so that k-row must start 2
    // banks (M/N-fastest index along lanes: 16 banks) further.
    constexpr int A_LD = (sizeof(T) == 8) ? A_CD + 1 : ((ALAY == A_KC) ? A_CD + 2 : A_CD + 16);
    constexpr int B_LD = (sizeof(T) == 8) ? B_CD + 1 : ((BLAY == B_NC) ? B_CD + 16 : B_CD + 2);
    constexpr int A_PER = A_CD * A_OD / NT;         // elements per thread
    constexpr int B_PER = B_CD * B_OD / NT;
    static_assert(NT % A_CD == 0 && NT % B_CD == 0, "tile shape");

    static_assert((A_CD * A_OD) % NT == 0 && (B_CD * B_OD) % NT == 0, "tile/threads");
    __shared__ T As[2][A_OD * A_LD];
    __shared__ T Bs[2][B_OD * B_LD];

    const int64_t M = p.M, N = p.N, K = p.K;
    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    const int64_t ntiles = tiles_m * tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    // Group GM tile-rows together so the A rows of a group stay in L2.
    constexpr int64_t GM = 8;
    int64_t group = t / (GM * tiles_n);
    int64_t first_m = group * GM;
    int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
    int64_t within = t - group * GM * tiles_n;
    int64_t tm = first_m + within % gm;
    int64_t tn = within / gm;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    if (p.uplo == 1 && n0 + BN <= m0) return;       // strictly below the diagonal

    const T* __restrict__ A = static_cast<const T*>(p.A);
    const T* __restrict__ B = static_cast<const T*>(p.B);
    T* __restrict__ C = static_cast<T*>(p.C);
    const int64_t lda = p.lda, ldb = p.ldb, ldc = p.ldc;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;

    T ra[PF][A_PER], rb[PF][B_PER];

    auto load_tiles = [&](int64_t k0, int set) {
        #pragma unroll
        for (int i = 0; i < A_PER; ++i) {
            int e = tid + i * NT;
            int cd = e % A_CD, od = e / A_CD;
            int64_t gm_, gk_;
            if (ALAY == A_KC) { gk_ = k0 + cd; gm_ = m0 + od; }
            else              { gm_ = m0 + cd; gk_ = k0 + od; }
            // branch-free edges: clamp the address, zero the value
            bool ok = gm_ < M && gk_ < K;
            int64_t cm_ = gm_ < M ? gm_ : M - 1, ck_ = gk_ < K ? gk_ : K - 1;
            int64_t off = (ALAY == A_KC) ? cm_ * lda + ck_ : cm_ + ck_ * lda;
            T v = A[off];
            #pragma unroll
            for (int q = 0; q < PF; ++q) if (q == set) ra[q][i] = ok ? v : T(0);
        }
        #pragma unroll
        for (int i = 0; i < B_PER; ++i) {
            int e = tid + i * NT;
            int cd = e % B_CD, od = e / B_CD;
            int64_t gk_, gn_;
            if (BLAY == B_NC) { gn_ = n0 + cd; gk_ = k0 + od; }
            else              { gk_ = k0 + cd; gn_ = n0 + od; }
            bool ok = gn_ < N && gk_ < K;
            int64_t cn_ = gn_ < N ? gn_ : N - 1, ck_ = gk_ < K ? gk_ : K - 1;
            int64_t off = (BLAY == B_NC) ? ck_ * ldb + cn_ : ck_ + cn_ * ldb;
            T v = B[off];
            #pragma unroll
            for (int q = 0; q < PF; ++q) if (q == set) rb[q][i] = ok ? v : T(0);
        }
    };
    auto store_tiles = [&](int buf, int set) {
        #pragma unroll
        for (int q = 0; q < PF; ++q) {
            if (q != set) continue;
            #pragma unroll
            for (int i = 0; i < A_PER; ++i) {
                int e = tid + i * NT;
                As[buf][(e / A_CD) * A_LD + (e % A_CD)] = ra[q][i];
            }
            #pragma unroll
            for (int i = 0; i < B_PER; ++i) {
                int e = tid + i * NT;
                Bs[buf][(e / B_CD) * B_LD + (e % B_CD)] = rb[q][i];
            }
        }
    };
    auto a_at = [&](int buf, int m, int k) -> T {
        return (ALAY == A_KC) ? As[buf][m * A_LD + k] : As[buf][k * A_LD + m];
    };
    auto b_at = [&](int buf, int k, int n) -> T {
        return (BLAY == B_NC) ? Bs[buf][k * B_LD + n] : Bs[buf][n * B_LD + k];
    };

    typename Mf<T>::acc acc[FM][FN];
    #pragma unroll
    for (int i = 0; i < FM; ++i)
        #pragma unroll
        for (int j = 0; j < FN; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);

    const int64_t ktiles = ceildiv(K, BK);
    const int fr = lane & 15, fk = lane >> 4;
    // One K slice: MFMAs on LDS buffer kt&1; the slice kt+1 (register set
    // `sn`) is written to the other buffer half-way through, so the LDS writes
    // overlap the remaining MFMAs (that buffer was last read before the
    // previous barrier); with PF = 2 the loads of slice kt+2 go out first.
    auto slice = [&](int64_t kt, int sn) {
        const int cur = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (PF == 1) { if (more) load_tiles((kt + 1) * BK, 0); }
        else if (kt + 2 < ktiles) load_tiles((kt + 2) * BK, 1 - sn);
        #pragma unroll
        for (int kk = 0; kk < BK; kk += KM) {
            if (kk == BK / 2 && more) store_tiles(cur ^ 1, PF == 1 ? 0 : sn);
            T af[FM], bf[FN];
            #pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = a_at(cur, wm * TM + i * 16 + fr, kk + fk);
            #pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = b_at(cur, kk + fk, wn * TN + j * 16 + fr);
            #pragma unroll
            for (int i = 0; i < FM; ++i)
                #pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mf<T>::op(af[i], bf[j], acc[i][j]);
        }
        if (more) __syncthreads();
    };
    load_tiles(0, 0);
    store_tiles(0, 0);
    if (PF == 2 && ktiles > 1) load_tiles(BK, 1);
    __syncthreads();
    if (PF == 1) {
        for (int64_t kt = 0; kt < ktiles; ++kt) slice(kt, 0);
    }
    else {
        for (int64_t kt = 0; kt < ktiles; kt += 2) {
            slice(kt, 1);
            if (kt + 1 < ktiles) slice(kt + 1, 0);
        }
    }

    const T alpha = T(p.alpha), beta = T(p.beta);
    // Interior tiles: C read per 16-row fragment row with no per-element
    // branches, so the loads are in flight together (the guarded form below
    // waits one memory round trip per element).
    if (m0 + BM <= M && n0 + BN <= N && (p.uplo == 0 || n0 >= m0 + BM)) {
        #pragma unroll
        for (int i = 0; i < FM; ++i) {
            T cv[FN][4];
            #pragma unroll
            for (int j = 0; j < FN; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r)
                    cv[j][r] = beta != T(0)
                        ? C[(m0 + wm * TM + i * 16 + Mf<T>::row(lane, r)) * ldc + n0 + wn * TN + j * 16 + fr] : T(0);
            #pragma unroll
            for (int j = 0; j < FN; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) {
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * cv[j][r];
                    C[(m0 + wm * TM + i * 16 + Mf<T>::row(lane, r)) * ldc + n0 + wn * TN + j * 16 + fr] = v;
                }
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < FM; ++i)
        #pragma unroll
        for (int j = 0; j < FN; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gm_ = m0 + wm * TM + i * 16 + Mf<T>::row(lane, r);
                int64_t gn_ = n0 + wn * TN + j * 16 + fr;
                if (gm_ < M && gn_ < N && (p.uplo == 0 || gn_ >= gm_)) {
                    T* c = C + gm_ * ldc + gn_;
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * *c;
                    *c = v;
                }
            }
}

//------------------------------------------------------------------------------
// FP64 product with A K-contiguous and B N-contiguous (the LU trailing update,
// the panel and lookahead updates): the generic kernel above moves one
// element per load and per LDS access and recomputes clamped addresses for
// each, which measured 47 TF (60% of the FP64 MFMA peak) with about three
// VALU instructions per MFMA.  Here:
//  * global loads are 16-byte vectors (8 doubles per thread per operand and
//    slice), unguarded for interior tiles;
//  * the K order inside a 16-deep slice is permuted so that lane l's four
//    k-steps use k = 4 (l/16) + s: its A and B fragments for the whole slice
//    are 4 consecutive doubles (two ds_read_b128 each) — B is stored
//    transposed in LDS ([n][k]) for that;
//  * LDS rows are 18 doubles (144 bytes): the 16-byte fragment reads of 8
//    consecutive rows fall on 8 distinct bank groups.
template <int ROLE>
__global__ __launch_bounds__(256, 2)
void gemm_f64_kernel(GemmArgs p)
{
    constexpr int BM = 128, BN = 128, BK = 16, LDP = 18;
    __shared__ double As[2][BM * LDP];
    __shared__ double Bt[2][BN * LDP];

    const int64_t M = p.M, N = p.N, K = p.K;
    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    // uplo with a square tile grid: only the upper-triangle tiles are launched
    const bool tri = p.uplo == 1 && tiles_m == tiles_n;
    const int64_t ntiles = tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t m0, n0;
    if (tri) {
        int64_t tm, tn;
        upper_tile(t, tiles_m, tm, tn);
        m0 = tm * BM;
        n0 = tn * BN;
    }
    else {
        constexpr int64_t GM = 8;
        int64_t group = t / (GM * tiles_n);
        int64_t first_m = group * GM;
        int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
        int64_t within = t - group * GM * tiles_n;
        m0 = (first_m + within % gm) * BM;
        n0 = (within / gm) * BN;
    }
    if (p.uplo == 1 && n0 + BN <= m0) return;

    const double* __restrict__ A = static_cast<const double*>(p.A);
    const double* __restrict__ B = static_cast<const double*>(p.B);
    double* __restrict__ C = static_cast<double*>(p.C);
    const int64_t lda = p.lda, ldb = p.ldb, ldc = p.ldc;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int fr = lane & 15, fk = lane >> 4;
    // staging: A row ar, k ak .. ak+7; B k row bk, columns bc .. bc+7
    const int ar = tid >> 1, ak = (tid & 1) * 8;
    const int bk = tid & 15, bc = (tid >> 4) * 8;
    const bool full_mn = m0 + BM <= M && n0 + BN <= N;
    const bool vec_ok = (lda % 2 == 0) && (ldb % 2 == 0) && (reinterpret_cast<uintptr_t>(A) % 16 == 0)
                        && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
    double ra[8], rb[8];
    auto load = [&](int64_t k0) {
        if (full_mn && vec_ok && k0 + BK <= K) {
            const double* ap = A + (m0 + ar) * lda + k0 + ak;
            const double* bp = B + (k0 + bk) * ldb + n0 + bc;
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                const v2d x = *reinterpret_cast<const v2d*>(ap + 2 * q);
                const v2d y = *reinterpret_cast<const v2d*>(bp + 2 * q);
                ra[2 * q] = x[0]; ra[2 * q + 1] = x[1];
                rb[2 * q] = y[0]; rb[2 * q + 1] = y[1];
            }
        }
        else {
            const int64_t gmr = m0 + ar, gkb = k0 + bk;
            const int64_t cm = gmr < M ? gmr : M - 1, ckb = gkb < K ? gkb : K - 1;
            #pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t gk = k0 + ak + e, gn = n0 + bc + e;
                const double x = A[cm * lda + (gk < K ? gk : K - 1)];
                const double y = B[ckb * ldb + (gn < N ? gn : N - 1)];
                ra[e] = (gmr < M && gk < K) ? x : 0.0;
                rb[e] = (gkb < K && gn < N) ? y : 0.0;
            }
        }
    };
    auto store = [&](int buf) {
        #pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<v2d*>(&As[buf][ar * LDP + ak + 2 * q]) = v2d{ra[2 * q], ra[2 * q + 1]};
        #pragma unroll
        for (int e = 0; e < 8; ++e) Bt[buf][(bc + e) * LDP + bk] = rb[e];
    };

    v4d acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    const int64_t ktiles = ceildiv(K, BK);
    load(0);
    store(0);
    __syncthreads();
    for (int64_t kt = 0; kt < ktiles; ++kt) {
        const int cur = int(kt & 1);
        const bool more = kt + 1 < ktiles;
        if (more) load((kt + 1) * BK);
        // two k-step pairs: each lane's fragments for a pair are one 16-byte read
        #pragma unroll
        for (int h = 0; h < 2; ++h) {
            v2d af[4], bf[4];
            #pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const v2d*>(&As[cur][(wm * 64 + i * 16 + fr) * LDP + 4 * fk + 2 * h]);
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                bf[j] = *reinterpret_cast<const v2d*>(&Bt[cur][(wn * 64 + j * 16 + fr) * LDP + 4 * fk + 2 * h]);
            #pragma unroll
            for (int s_ = 0; s_ < 2; ++s_)
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    #pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][s_], bf[j][s_], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store(cur ^ 1);
            __syncthreads();
        }
    }

    const double alpha = p.alpha, beta = p.beta;
    // Interior tiles: each 16-row block reads its 16 C values per lane with
    // no per-element branches, so they are in flight together (the guarded
    // element-by-element form below waits one memory round trip per element).
    const bool full = m0 + BM <= M && n0 + BN <= N && (p.uplo == 0 || n0 >= m0 + BM);
    if (full) {
        double* cb = C + (m0 + wm * 64 + fk) * ldc + n0 + wn * 64 + fr;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            double cv[4][4];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cv[j][r] = beta != 0.0 ? cb[(i * 16 + 4 * r) * ldc + j * 16] : 0.0;
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) {
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v += beta * cv[j][r];
                    cb[(i * 16 + 4 * r) * ldc + j * 16] = v;
                }
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm_ = m0 + wm * 64 + i * 16 + 4 * r + fk;
                const int64_t gn_ = n0 + wn * 64 + j * 16 + fr;
                if (gm_ < M && gn_ < N && (p.uplo == 0 || gn_ >= gm_)) {
                    double* c = C + gm_ * ldc + gn_;
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v += beta * *c;
                    *c = v;
                }
            }
}

//------------------------------------------------------------------------------
// FP32 product with A K-contiguous and B N-contiguous (the float LU's
// L11^{-1} applications, the panel's node updates, the solves' off-diagonal
// blocks, and the trailing update of the FP32-MFMA factor): the FP64 kernel
// above with float operands and v_mfma_f32_16x16x4f32.  Slices are 32 deep
// (the same 128-byte rows), lane l's eight k-steps use k = 8 (l/16) + s, so
// its A and B fragments for a slice are 8 consecutive floats (two
// ds_read_b128 each); LDS rows are 36 floats (144 bytes) as in the FP64
// kernel.  The accumulator layout differs from FP64's: register r of lane l
// holds D[4 (l/16) + r][l%16].
template <int ROLE>
__global__ __launch_bounds__(256, 2)
void gemm_f32_kernel(GemmArgs p)
{
    constexpr int BM = 128, BN = 128, BK = 32, LDP = 36;
    __shared__ float As[2][BM * LDP];
    __shared__ float Bt[2][BN * LDP];

    const int64_t M = p.M, N = p.N, K = p.K;
    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    // uplo with a square tile grid: only the upper-triangle tiles are launched
    const bool tri = p.uplo == 1 && tiles_m == tiles_n;
    const int64_t ntiles = tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t m0, n0;
    if (tri) {
        int64_t tm, tn;
        upper_tile(t, tiles_m, tm, tn);
        m0 = tm * BM;
        n0 = tn * BN;
    }
    else {
        constexpr int64_t GM = 8;
        int64_t group = t / (GM * tiles_n);
        int64_t first_m = group * GM;
        int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
        int64_t within = t - group * GM * tiles_n;
        m0 = (first_m + within % gm) * BM;
        n0 = (within / gm) * BN;
    }
    if (p.uplo == 1 && n0 + BN <= m0) return;

    const float* __restrict__ A = static_cast<const float*>(p.A);
    const float* __restrict__ B = static_cast<const float*>(p.B);
    float* __restrict__ C = static_cast<float*>(p.C);
    const int64_t lda = p.lda, ldb = p.ldb, ldc = p.ldc;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int fr = lane & 15, fk = lane >> 4;
    // staging: A row ar, k ak .. ak+15; B k row bk, columns bc .. bc+15
    const int ar = tid >> 1, ak = (tid & 1) * 16;
    const int bk = tid & 31, bc = (tid >> 5) * 16;
    const bool full_mn = m0 + BM <= M && n0 + BN <= N;
    const bool vec_ok = (lda % 4 == 0) && (ldb % 4 == 0) && (reinterpret_cast<uintptr_t>(A) % 16 == 0)
                        && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
    v4f ra[4], rb[4];
    auto load = [&](int64_t k0) {
        if (full_mn && vec_ok && k0 + BK <= K) {
            const float* ap = A + (m0 + ar) * lda + k0 + ak;
            const float* bp = B + (k0 + bk) * ldb + n0 + bc;
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                ra[q] = *reinterpret_cast<const v4f*>(ap + 4 * q);
                rb[q] = *reinterpret_cast<const v4f*>(bp + 4 * q);
            }
        }
        else {
            const int64_t gmr = m0 + ar, gkb = k0 + bk;
            const int64_t cm = gmr < M ? gmr : M - 1, ckb = gkb < K ? gkb : K - 1;
            #pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int64_t gk = k0 + ak + e, gn = n0 + bc + e;
                const float x = A[cm * lda + (gk < K ? gk : K - 1)];
                const float y = B[ckb * ldb + (gn < N ? gn : N - 1)];
                ra[e >> 2][e & 3] = (gmr < M && gk < K) ? x : 0.0f;
                rb[e >> 2][e & 3] = (gkb < K && gn < N) ? y : 0.0f;
            }
        }
    };
    auto store = [&](int buf) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) *reinterpret_cast<v4f*>(&As[buf][ar * LDP + ak + 4 * q]) = ra[q];
        #pragma unroll
        for (int e = 0; e < 16; ++e) Bt[buf][(bc + e) * LDP + bk] = rb[e >> 2][e & 3];
    };

    v4f acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    const int64_t ktiles = ceildiv(K, BK);
    load(0);
    store(0);
    __syncthreads();
    for (int64_t kt = 0; kt < ktiles; ++kt) {
        const int cur = int(kt & 1);
        const bool more = kt + 1 < ktiles;
        if (more) load((kt + 1) * BK);
        #pragma unroll
        for (int h = 0; h < 2; ++h) {
            v4f af[4], bf[4];
            #pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const v4f*>(&As[cur][(wm * 64 + i * 16 + fr) * LDP + 8 * fk + 4 * h]);
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                bf[j] = *reinterpret_cast<const v4f*>(&Bt[cur][(wn * 64 + j * 16 + fr) * LDP + 8 * fk + 4 * h]);
            #pragma unroll
            for (int s_ = 0; s_ < 4; ++s_)
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    #pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][s_], bf[j][s_], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store(cur ^ 1);
            __syncthreads();
        }
    }

    const float alpha = float(p.alpha), beta = float(p.beta);
    // Interior tiles: C read per 16-row block with no per-element branches
    // (see gemm_f64_kernel).
    const bool full = m0 + BM <= M && n0 + BN <= N && (p.uplo == 0 || n0 >= m0 + BM);
    if (full) {
        float* cb = C + (m0 + wm * 64 + 4 * fk) * ldc + n0 + wn * 64 + fr;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            float cv[4][4];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cv[j][r] = beta != 0.0f ? cb[(i * 16 + r) * ldc + j * 16] : 0.0f;
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * cv[j][r];
                    cb[(i * 16 + r) * ldc + j * 16] = v;
                }
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm_ = m0 + wm * 64 + i * 16 + 4 * fk + r;
                const int64_t gn_ = n0 + wn * 64 + j * 16 + fr;
                if (gm_ < M && gn_ < N && (p.uplo == 0 || gn_ >= gm_)) {
                    float* c = C + gm_ * ldc + gn_;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *c;
                    *c = v;
                }
            }
}

//------------------------------------------------------------------------------
// Narrow products (A_KC / B_NC, N <= 64, K <= 64, all of C): the recursive
// panel's lowest splits (a tall M x 32 or M x 64 block of L times the square
// block of U beside it, ~1500 launches per factorization) and the small
// off-diagonal blocks of the recursive triangular solve.  On the 128 x 128
// MFMA tile each workgroup stages one slice, waits and writes, and such a
// launch moves ~0.14 TB/s (M = 40000, N = K = 32: ~110 us on 32 CUs).  Here B
// sits in LDS, each thread reads its own row of A (the lanes of one row share
// its addresses) and owns 16 bytes of C, and each element of C sums its K
// products in k order on the VALU.
template <class T>
__global__ __launch_bounds__(256)
void gemm_narrow_kernel(GemmArgs p)
{
    constexpr int VN = 16 / int(sizeof(T));
    constexpr int BP = 64 + VN;                        // padded row pitch of B in LDS
    typedef T vec_t __attribute__((ext_vector_type(VN)));
    __shared__ __align__(16) T Bs[64 * BP];
    const int64_t M = p.M;
    const int N = int(p.N), K = int(p.K);
    const T* __restrict__ A = static_cast<const T*>(p.A);
    const T* __restrict__ B = static_cast<const T*>(p.B);
    T* __restrict__ C = static_cast<T*>(p.C);
    const int64_t lda = p.lda, ldb = p.ldb, ldc = p.ldc;
    const int tid = threadIdx.x;
    for (int e = tid; e < K * 64; e += 256) {
        const int k = e >> 6, n = e & 63;
        Bs[k * BP + n] = n < N ? B[int64_t(k) * ldb + n] : T(0);
    }
    __syncthreads();
    const int tpr = (N + VN - 1) / VN;                 // threads per row of C
    const int rw = 256 / tpr;                          // rows per workgroup
    const int lr = tid / tpr, g = tid - lr * tpr;
    const int64_t i = int64_t(blockIdx.x) * rw + lr;
    if (lr >= rw || i >= M) return;                    // no barrier below
    const T* Ai = A + i * lda;
    const T* Bg = Bs + g * VN;
    const T alpha = T(p.alpha), beta = T(p.beta);
    const int c0 = g * VN;
    T* Ci = C + i * ldc + c0;
    const bool cvec = c0 + VN <= N && reinterpret_cast<uintptr_t>(C) % 16 == 0 && ldc % VN == 0;
    vec_t cv;                                          // C read before the products: one round trip less
    if (cvec && beta != T(0)) cv = *reinterpret_cast<const vec_t*>(Ci);
    T acc[VN];
    #pragma unroll
    for (int e = 0; e < VN; ++e) acc[e] = T(0);
    int k = 0;
    if (reinterpret_cast<uintptr_t>(A) % 16 == 0 && lda % VN == 0) {
        for (; k + 8 * VN <= K; k += 8 * VN) {
            vec_t av[8];
            #pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = *reinterpret_cast<const vec_t*>(Ai + k + u * VN);
            #pragma unroll
            for (int u = 0; u < 8; ++u)
                #pragma unroll
                for (int t = 0; t < VN; ++t) {
                    const vec_t b = *reinterpret_cast<const vec_t*>(Bg + (k + u * VN + t) * BP);
                    #pragma unroll
                    for (int e = 0; e < VN; ++e) acc[e] += av[u][t] * b[e];
                }
        }
    }
    for (; k < K; ++k) {
        const T a = Ai[k];
        const vec_t b = *reinterpret_cast<const vec_t*>(Bg + k * BP);
        #pragma unroll
        for (int e = 0; e < VN; ++e) acc[e] += a * b[e];
    }
    if (cvec) {
        vec_t v;
        if (beta != T(0)) {
            #pragma unroll
            for (int e = 0; e < VN; ++e) v[e] = alpha * acc[e] + beta * cv[e];
        }
        else {
            #pragma unroll
            for (int e = 0; e < VN; ++e) v[e] = alpha * acc[e];
        }
        *reinterpret_cast<vec_t*>(Ci) = v;
    }
    else {
        #pragma unroll
        for (int e = 0; e < VN; ++e)
            if (c0 + e < N) Ci[e] = beta != T(0) ? alpha * acc[e] + beta * Ci[e] : alpha * acc[e];
    }
}

template <class T, int ALAY, int BLAY, int ROLE = 0>
static void launch(const GemmArgs& a, hipStream_t s)
{
    if (a.M <= 0 || a.N <= 0) return;
    if constexpr (ALAY == A_KC && BLAY == B_NC && ROLE == 0) {
        if (a.uplo == 0 && a.N <= 64 && a.K <= 64) {
            const int64_t rw = 256 / ((a.N + 16 / int64_t(sizeof(T)) - 1) / (16 / int64_t(sizeof(T))));
            const int64_t nwg = ceildiv(a.M, rw);
            SA_REQUIRE(nwg < (int64_t(1) << 31), -2, "gemm: grid too large");
            hipLaunchKernelGGL((gemm_narrow_kernel<T>), dim3(unsigned(nwg)), dim3(256), 0, s, a);
            SA_CHECK_LAUNCH();
            return;
        }
    }
    // 128 x 128 tiles, 4 waves, two workgroups per CU (2 waves per SIMD) —
    // the fastest of the tilings measured for both FP64 and FP32.
    constexpr int BM = 128, BN = 128, BK = (sizeof(T) == 8) ? 16 : 32;
    int64_t nt = ceildiv(a.M, BM) * ceildiv(a.N, BN);
    // the 16-byte-staged kernels launch only the upper-triangle tiles of a
    // square triangular update (same tile size in both directions)
    if constexpr (ALAY == A_KC && BLAY == B_NC)
        if (a.uplo == 1 && ceildiv(a.M, BM) == ceildiv(a.N, BN))
            nt = ceildiv(a.M, BM) * (ceildiv(a.M, BM) + 1) / 2;
    SA_REQUIRE(nt < (int64_t(1) << 31), -2, "gemm: grid too large");
    if constexpr (std::is_same<T, double>::value && ALAY == A_KC && BLAY == B_NC) {
        hipLaunchKernelGGL((gemm_f64_kernel<ROLE>), dim3(unsigned(nt)), dim3(256), 0, s, a);
    }
    else if constexpr (std::is_same<T, float>::value && ALAY == A_KC && BLAY == B_NC) {
        hipLaunchKernelGGL((gemm_f32_kernel<ROLE>), dim3(unsigned(nt)), dim3(256), 0, s, a);
    }
    else {
        hipLaunchKernelGGL((gemm_kernel<T, ALAY, BLAY, BM, BN, BK, 2, 2, 2, ROLE>), dim3(unsigned(nt)), dim3(256),
                           0, s, a);
    }
    SA_CHECK_LAUNCH();
}

template <class T>
void gemm(const GemmArgs& a, int alay, int blay, hipStream_t s, int role)
{
    if (a.K <= 0) {
        // C = beta * C only; the factorizations never request this.
        SA_REQUIRE(a.beta == 1.0, -2, "gemm: K == 0 with beta != 1 unsupported");
        return;
    }
    if (role == 1) {
        SA_REQUIRE(alay == A_KC && blay == B_NC, -2, "gemm: trailing-update role needs A_KC/B_NC");
        launch<T, A_KC, B_NC, 1>(a, s);
    }
    else if (alay == A_KC && blay == B_NC) launch<T, A_KC, B_NC>(a, s);
    else if (alay == A_MC && blay == B_NC) launch<T, A_MC, B_NC>(a, s);
    else if (alay == A_KC && blay == B_KC) launch<T, A_KC, B_KC>(a, s);
    else launch<T, A_MC, B_KC>(a, s);
}

template void gemm<double>(const GemmArgs&, int, int, hipStream_t, int);
template void gemm<float>(const GemmArgs&, int, int, hipStream_t, int);

} // namespace sa
