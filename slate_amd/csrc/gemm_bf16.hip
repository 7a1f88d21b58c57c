// Split-BF16 MFMA update for the low-precision LU (gfx950).
//
// The mixed drivers factor a float copy of A (src/gesv_mixed.cc:179 keeps the
// low-precision factor in scalar_lo).  Here the bulk of that factorization's
// flops, the trailing update C -= L U (src/internal/internal_gemm.cc:355-518
// for the low-precision matrix), runs on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation: each FP32 operand is split into NS bfloat16 terms
// (x = x0 + x1 + ... , each the round-to-nearest bf16 of the remainder), and
// the product keeps the terms x_i y_j with i + j < NS:
//   NS = 1: x0 y0                            (8-bit significands, bf16 LU)
//   NS = 2: x0 y0 + x0 y1 + x1 y0            (~16-bit)
//   NS = 3: the six terms with i + j <= 2    (~24-bit: FP32-equivalent)
// so one kernel gives a BF16, a "two-term" and an FP32-accurate update at
// 1, 3 and 6 bf16 products, against FP32 MFMA at 1/16 of the bf16 rate.
//
// Operands are prepared once per step by the split kernels below: the L panel
// (row-major, K contiguous) as NS planes of M x Kp bf16 and the U row block
// (row-major K x N) transposed into NS planes of N x Kp, so both MFMA
// operands read 8 consecutive k (16 bytes) per lane.  Kp is K rounded up to
// 32 with zero padding.
#include "common.hh"
#include "kernels.hh"

namespace sa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

//------------------------------------------------------------------------------
// Splitting.  Each thread handles 8 consecutive k of one row (one 16-byte
// store per plane).
template <int NS>
__device__ inline void split8(const float (&x)[8], bf16x8 (&h)[NS > 0 ? NS : 1])
{
    #pragma unroll
    for (int e = 0; e < 8; ++e) {
        float r = x[e];
        #pragma unroll
        for (int s = 0; s < NS; ++s) {
            __bf16 b = __bf16(r);
            h[s][e] = b;
            r -= float(b);
        }
    }
}

// src: M x K row-major fp32 (ld); dst plane s: dst + s*plane, row m at m*Kp.
template <int NS>
__global__ __launch_bounds__(256)
void split_rows_kernel(const float* __restrict__ src, int64_t ld, int64_t M, int64_t K,
                       __bf16* __restrict__ dst, int64_t Kp, int64_t plane)
{
    const int64_t kg = Kp / 8;
    const int64_t total = M * kg;
    for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total; t += int64_t(gridDim.x) * 256) {
        const int64_t m = t / kg, k0 = (t % kg) * 8;
        float x[8];
        #pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = (k0 + e < K) ? src[m * ld + k0 + e] : 0.0f;
        bf16x8 h[NS > 0 ? NS : 1];
        split8<NS>(x, h);
        #pragma unroll
        for (int s = 0; s < NS; ++s)
            *reinterpret_cast<bf16x8*>(dst + s * plane + m * Kp + k0) = h[s];
    }
}

// src: K x N row-major fp32 (ld); dst plane s: N x Kp (k contiguous).  A 32 x 64
// (k x n) tile is staged through LDS so both the reads and the writes are
// contiguous.
template <int NS>
__global__ __launch_bounds__(256)
void split_cols_t_kernel(const float* __restrict__ src, int64_t ld, int64_t K, int64_t N,
                         __bf16* __restrict__ dst, int64_t Kp, int64_t plane)
{
    __shared__ float t[32][65];
    const int64_t n0 = int64_t(blockIdx.x) * 64, k0 = int64_t(blockIdx.y) * 32;
    const int tid = threadIdx.x;
    #pragma unroll
    for (int u = 0; u < 8; ++u) {
        int kk = (tid >> 6) + 4 * u, nn = tid & 63;
        int64_t k = k0 + kk, n = n0 + nn;
        t[kk][nn] = (k < K && n < N) ? src[k * ld + n] : 0.0f;
    }
    __syncthreads();
    // 64 columns x 4 groups of 8 k = 256 threads
    const int nn = tid >> 2, kg = (tid & 3) * 8;
    const int64_t n = n0 + nn;
    if (n < N) {
        float x[8];
        #pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = t[kg + e][nn];
        if constexpr (NS == 0) {
            // plain transposed float copy (dst is float, plane unused)
            float* d = reinterpret_cast<float*>(dst) + n * Kp + k0 + kg;
            *reinterpret_cast<v4f*>(d) = v4f{x[0], x[1], x[2], x[3]};
            *reinterpret_cast<v4f*>(d + 4) = v4f{x[4], x[5], x[6], x[7]};
        }
        else {
            bf16x8 h[NS > 0 ? NS : 1];
            split8<NS>(x, h);
            #pragma unroll
            for (int s = 0; s < NS; ++s)
                *reinterpret_cast<bf16x8*>(dst + s * plane + n * Kp + k0 + kg) = h[s];
        }
    }
}

//------------------------------------------------------------------------------
// C (fp32 row-major, ldc) += alpha * sum_{i+j<NS} A_i B_j^T, A_i: M x Kp,
// B_j: N x Kp (bf16, k contiguous); uplo = 1 updates only entries with
// column >= row (the herk of the Cholesky update).  128 x 128 C tile per workgroup, 4 wave64s
// in 2 x 2, each a 64 x 64 block of 16 x 16 accumulators; BK = 32 (one MFMA
// deep) slices of every plane double-buffered in LDS.  LDS rows are padded to
// 40 bf16 (80 bytes) so the 16-byte fragment reads of 16 rows hit 16
// different 4-bank groups.
template <int NS, int ROLE>
__global__ __launch_bounds__(256, 1)
void gemm_bf16s_kernel(int64_t M, int64_t N, int64_t Kp, const __bf16* __restrict__ A, int64_t lda,
                       int64_t planeA, const __bf16* __restrict__ B, int64_t ldb, int64_t planeB,
                       float* __restrict__ C, int64_t ldc, float alpha, int uplo)
{
    constexpr int BM = 128, BN = 128, BK = 32, LDK = 40;
    constexpr int TILE = BM * LDK;                   // bf16 per plane per buffer
    __shared__ __bf16 As[2][NS][TILE];
    __shared__ __bf16 Bs[2][NS][TILE];

    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    const int64_t ntiles = tiles_m * tiles_n;
    int64_t t = xcd_remap(blockIdx.x, ntiles);
    constexpr int64_t GM = 8;
    int64_t group = t / (GM * tiles_n);
    int64_t first_m = group * GM;
    int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
    int64_t within = t - group * GM * tiles_n;
    const int64_t m0 = (first_m + within % gm) * BM;
    const int64_t n0 = (within / gm) * BN;
    if (uplo && n0 + BN <= m0) return;               // strictly below the diagonal

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    // Global -> register staging: per plane, 128 rows x 4 chunks of 8 bf16;
    // thread tid takes chunks tid and tid + 256 (rows tid/4 and tid/4 + 64).
    // Rows beyond M (N) are clamped: they only feed C rows (columns) that are
    // never stored.
    u32x4 ra[NS][2], rb[NS][2];
    const int crow = tid >> 2, cchunk = (tid & 3) * 8;
    int64_t arow[2], brow[2];
    #pragma unroll
    for (int h = 0; h < 2; ++h) {
        int64_t r = m0 + crow + 64 * h;
        arow[h] = (r < M ? r : M - 1) * lda;
        int64_t c = n0 + crow + 64 * h;
        brow[h] = (c < N ? c : N - 1) * ldb;
    }
    auto load = [&](int64_t k0) {
        #pragma unroll
        for (int s = 0; s < NS; ++s)
            #pragma unroll
            for (int h = 0; h < 2; ++h) {
                ra[s][h] = *reinterpret_cast<const u32x4*>(A + s * planeA + arow[h] + k0 + cchunk);
                rb[s][h] = *reinterpret_cast<const u32x4*>(B + s * planeB + brow[h] + k0 + cchunk);
            }
    };
    auto store = [&](int buf) {
        #pragma unroll
        for (int s = 0; s < NS; ++s)
            #pragma unroll
            for (int h = 0; h < 2; ++h) {
                *reinterpret_cast<u32x4*>(&As[buf][s][(crow + 64 * h) * LDK + cchunk]) = ra[s][h];
                *reinterpret_cast<u32x4*>(&Bs[buf][s][(crow + 64 * h) * LDK + cchunk]) = rb[s][h];
            }
    };

    v4f acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    const int64_t ktiles = Kp / BK;
    load(0);
    store(0);
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int64_t kt = 0; kt < ktiles; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < ktiles;
        if (more) load((kt + 1) * BK);
        bf16x8 af[NS][4], bfr[NS][4];
        #pragma unroll
        for (int s = 0; s < NS; ++s) {
            #pragma unroll
            for (int i = 0; i < 4; ++i)
                af[s][i] = *reinterpret_cast<const bf16x8*>(&As[cur][s][(wm * 64 + i * 16 + fr) * LDK + fk]);
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                bfr[s][j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][s][(wn * 64 + j * 16 + fr) * LDK + fk]);
        }
        // small terms first, the leading x0 y0 last
        #pragma unroll
        for (int sa_ = NS - 1; sa_ >= 0; --sa_)
            #pragma unroll
            for (int sb = NS - 1 - sa_; sb >= 0; --sb)
                #pragma unroll
                for (int i = 0; i < 4; ++i)
                    #pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[sa_][i], bfr[sb][j], acc[i][j],
                                                                           0, 0, 0);
        if (more) {
            store(cur ^ 1);
            __syncthreads();
        }
    }

    // Interior tiles: C read per 16-row block with no per-element branches,
    // so the loads are in flight together (the guarded element-by-element
    // update below waits one memory round trip per element).
    if (m0 + BM <= M && n0 + BN <= N && (uplo == 0 || n0 >= m0 + BM)) {
        float* cb = C + (m0 + wm * 64 + 4 * (lane >> 4)) * ldc + n0 + wn * 64 + fr;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            float cv[4][4];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cv[j][r] = cb[(i * 16 + r) * ldc + j * 16];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cb[(i * 16 + r) * ldc + j * 16] = cv[j][r] + alpha * acc[i][j][r];
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gm_ = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
                int64_t gn_ = n0 + wn * 64 + j * 16 + fr;
                if (gm_ < M && gn_ < N && (uplo == 0 || gn_ >= gm_)) {
                    float* c = C + gm_ * ldc + gn_;
                    *c += alpha * acc[i][j][r];
                }
            }
}

//------------------------------------------------------------------------------
// The same update from float operands split inside the kernel: C (row-major)
// += alpha * L R^T, L: M x K and R: N x K, both float with k contiguous, with
// no pre-split planes (half the operand traffic of three bf16 planes and no
// separate split pass over the L panel).  Each float element is split once per
// workgroup, by the thread that stages it into LDS, rather than by every wave
// that reads it (each A row is read by two waves, each B row by two).  LDS
// holds NS bf16 planes per operand, a 128 x 32 tile each as 128 rows of four
// 16-byte chunks (8 consecutive k); chunk c of row r sits in slot
// c ^ ((r >> 2) & 3), so the 16 rows of a fragment read and the 4 rows x 4
// chunks of a staging write both cover all 64 banks.  One buffer (16 KB per
// term), refilled between two barriers while the next slice's global loads are
// in flight; MINB workgroups share a CU to cover the barrier gaps.

typedef float v2f_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// 8 floats -> NS packed bf16x8 terms (round to nearest, residual carried).
template <int NS>
__device__ inline void split8_packed(float (&r)[8], u32x4 (&h)[NS])
{
    #pragma unroll
    for (int s = 0; s < NS; ++s) {
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
            const bf16x2_t b = __builtin_convertvector((v2f_t{r[2 * e], r[2 * e + 1]}), bf16x2_t);
            const unsigned u = __builtin_bit_cast(unsigned, b);
            h[s][e] = u;
            if (s + 1 < NS) {
                r[2 * e] -= __uint_as_float(u << 16);
                r[2 * e + 1] -= __uint_as_float(u & 0xffff0000u);
            }
        }
    }
}

template <int NS, int ROLE, int MINB>
__global__ __launch_bounds__(256, MINB)
void gemm_sp_kernel(int64_t M, int64_t N, int64_t K, const float* __restrict__ L, int64_t ldl,
                    const float* __restrict__ R, int64_t ldr, float* __restrict__ C, int64_t ldc, float alpha,
                    int uplo)
{
    constexpr int BM = 128, BN = 128, BK = 32;
    __shared__ u32x4 As[NS][BM * 4];
    __shared__ u32x4 Bs[NS][BN * 4];

    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    // uplo with a square tile grid: only the upper-triangle tiles are launched
    const bool tri = uplo && tiles_m == tiles_n;
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
    if (uplo && n0 + BN <= m0) return;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // staging: rows lr and lr + 64 of each operand tile, k chunk lc (8 floats
    // as two float4); rows past M (N) clamped (not stored), k past K zeroed
    const int lr = tid >> 2, lc = tid & 3;
    int64_t aoff[2], boff[2];
    #pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t r = m0 + lr + 64 * h, c = n0 + lr + 64 * h;
        aoff[h] = (r < M ? r : M - 1) * ldl;
        boff[h] = (c < N ? c : N - 1) * ldr;
    }
    v4f ra[2][2], rb[2][2];
    auto load = [&](int64_t k0) {
        if (k0 + BK <= K) {
            // whole slice inside K (uniform): plain 16-byte loads, no edge selects
            #pragma unroll
            for (int q = 0; q < 2; ++q)
                #pragma unroll
                for (int h = 0; h < 2; ++h) {
                    ra[h][q] = *reinterpret_cast<const v4f*>(L + aoff[h] + k0 + lc * 8 + q * 4);
                    rb[h][q] = *reinterpret_cast<const v4f*>(R + boff[h] + k0 + lc * 8 + q * 4);
                }
            return;
        }
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t k = k0 + lc * 8 + q * 4;
            const bool full = k + 4 <= K;
            const int64_t kc = full ? k : (K >= 4 ? K - 4 : 0);
            #pragma unroll
            for (int h = 0; h < 2; ++h) {
                v4f a = *reinterpret_cast<const v4f*>(L + aoff[h] + kc);
                v4f b = *reinterpret_cast<const v4f*>(R + boff[h] + kc);
                if (!full) {
                    #pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int64_t ke = k + e;
                        const int src = int(ke - kc);
                        float av = 0.f, bv = 0.f;
                        #pragma unroll
                        for (int w = 0; w < 4; ++w) if (w == src) { av = a[w]; bv = b[w]; }
                        a[e] = (ke < K) ? av : 0.f;
                        b[e] = (ke < K) ? bv : 0.f;
                    }
                }
                ra[h][q] = a;
                rb[h][q] = b;
            }
        }
    };
    auto stage = [&]() {
        #pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int row = lr + 64 * h;
            const int slot = row * 4 + (lc ^ ((row >> 2) & 3));
            float x[8] = {ra[h][0][0], ra[h][0][1], ra[h][0][2], ra[h][0][3],
                          ra[h][1][0], ra[h][1][1], ra[h][1][2], ra[h][1][3]};
            u32x4 p[NS];
            split8_packed<NS>(x, p);
            #pragma unroll
            for (int s = 0; s < NS; ++s) As[s][slot] = p[s];
            float y[8] = {rb[h][0][0], rb[h][0][1], rb[h][0][2], rb[h][0][3],
                          rb[h][1][0], rb[h][1][1], rb[h][1][2], rb[h][1][3]};
            split8_packed<NS>(y, p);
            #pragma unroll
            for (int s = 0; s < NS; ++s) Bs[s][slot] = p[s];
        }
    };

    v4f acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    const int fr = lane & 15, fslot = (lane >> 4) ^ ((fr >> 2) & 3);
    auto compute = [&]() {
        bf16x8 bs[4][NS];
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int s = 0; s < NS; ++s)
                bs[j][s] = __builtin_bit_cast(bf16x8, Bs[s][(wn * 64 + j * 16 + fr) * 4 + fslot]);
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            bf16x8 as[NS];
            #pragma unroll
            for (int s = 0; s < NS; ++s)
                as[s] = __builtin_bit_cast(bf16x8, As[s][(wm * 64 + i * 16 + fr) * 4 + fslot]);
            #pragma unroll
            for (int sa_ = NS - 1; sa_ >= 0; --sa_)
                #pragma unroll
                for (int sb = NS - 1 - sa_; sb >= 0; --sb)
                    #pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as[sa_], bs[j][sb], acc[i][j], 0, 0, 0);
        }
    };

    const int64_t ktiles = ceildiv(K, BK);
    load(0);
    stage();
    __syncthreads();
    for (int64_t kt = 0; kt < ktiles; ++kt) {
        const bool more = kt + 1 < ktiles;
        if (more) load((kt + 1) * BK);
        compute();
        if (more) {
            __syncthreads();
            stage();
            __syncthreads();
        }
    }

    // Interior tiles: C read per 16-row block with no per-element branches,
    // so the loads are in flight together (the guarded element-by-element
    // update below waits one memory round trip per element).
    if (m0 + BM <= M && n0 + BN <= N && (uplo == 0 || n0 >= m0 + BM)) {
        float* cb = C + (m0 + wm * 64 + 4 * (lane >> 4)) * ldc + n0 + wn * 64 + fr;
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            float cv[4][4];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cv[j][r] = cb[(i * 16 + r) * ldc + j * 16];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cb[(i * 16 + r) * ldc + j * 16] = cv[j][r] + alpha * acc[i][j][r];
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm_ = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + r;
                const int64_t gn_ = n0 + wn * 64 + j * 16 + fr;
                if (gm_ < M && gn_ < N && (uplo == 0 || gn_ >= gm_)) {
                    float* c = C + gm_ * ldc + gn_;
                    *c += alpha * acc[i][j][r];
                }
            }
}

void transpose_f32(const float* src, int64_t ld, int64_t K, int64_t N, float* dst, int64_t Kp, hipStream_t s)
{
    if (N <= 0) return;
    SA_REQUIRE(Kp % 32 == 0 && Kp >= K, -2, "transpose_f32: bad Kp");
    dim3 grid(unsigned(ceildiv(N, 64)), unsigned(Kp / 32));
    hipLaunchKernelGGL(split_cols_t_kernel<0>, grid, dim3(256), 0, s, src, ld, K, N,
                       reinterpret_cast<__bf16*>(dst), Kp, int64_t(0));
    SA_CHECK_LAUNCH();
}

void gemm_s32(int ns, int64_t M, int64_t N, int64_t K, const float* L, int64_t ldl, const float* R, int64_t ldr,
              float* C, int64_t ldc, float alpha, hipStream_t s, int role, int uplo)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    SA_REQUIRE(ldl % 4 == 0 && ldr % 4 == 0 && reinterpret_cast<uintptr_t>(L) % 16 == 0
               && reinterpret_cast<uintptr_t>(R) % 16 == 0, -2, "gemm_s32: operands must be 16-byte aligned");
    SA_REQUIRE(K >= 4, -2, "gemm_s32: K must be at least 4");
    int64_t nt = ceildiv(M, 128) * ceildiv(N, 128);
    if (uplo && ceildiv(M, 128) == ceildiv(N, 128)) nt = ceildiv(M, 128) * (ceildiv(M, 128) + 1) / 2;
    SA_REQUIRE(nt < (int64_t(1) << 31), -2, "gemm_s32: grid too large");
    // split at staging, three workgroups per CU (48 KB of LDS each at NS = 3)
#define SA_S32(NSV, RV) hipLaunchKernelGGL((gemm_sp_kernel<NSV, RV, 3>), dim3(unsigned(nt)), dim3(256), 0, s, \
                                           M, N, K, L, ldl, R, ldr, C, ldc, alpha, uplo)
    if (role == 1) {
        switch (ns) { case 1: SA_S32(1, 1); break; case 2: SA_S32(2, 1); break; case 3: SA_S32(3, 1); break;
                      default: SA_REQUIRE(false, -2, "gemm_s32: ns must be 1..3"); }
    }
    else {
        switch (ns) { case 1: SA_S32(1, 0); break; case 2: SA_S32(2, 0); break; case 3: SA_S32(3, 0); break;
                      default: SA_REQUIRE(false, -2, "gemm_s32: ns must be 1..3"); }
    }
#undef SA_S32
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
void split_rows(int ns, const float* src, int64_t ld, int64_t M, int64_t K, __bf16* dst, int64_t Kp,
                int64_t plane, hipStream_t s)
{
    if (M <= 0) return;
    SA_REQUIRE(Kp % 32 == 0 && Kp >= K, -2, "split_rows: bad Kp");
    int64_t total = M * (Kp / 8);
    unsigned g = unsigned(std::min<int64_t>(ceildiv(total, 256), 8192));
    switch (ns) {
        case 1: hipLaunchKernelGGL(split_rows_kernel<1>, dim3(g), dim3(256), 0, s, src, ld, M, K, dst, Kp, plane); break;
        case 2: hipLaunchKernelGGL(split_rows_kernel<2>, dim3(g), dim3(256), 0, s, src, ld, M, K, dst, Kp, plane); break;
        case 3: hipLaunchKernelGGL(split_rows_kernel<3>, dim3(g), dim3(256), 0, s, src, ld, M, K, dst, Kp, plane); break;
        default: SA_REQUIRE(false, -2, "split_rows: ns must be 1..3");
    }
    SA_CHECK_LAUNCH();
}

void split_cols_t(int ns, const float* src, int64_t ld, int64_t K, int64_t N, __bf16* dst, int64_t Kp,
                  int64_t plane, hipStream_t s)
{
    if (N <= 0) return;
    SA_REQUIRE(Kp % 32 == 0 && Kp >= K, -2, "split_cols_t: bad Kp");
    SA_REQUIRE(ceildiv(N, 64) < (int64_t(1) << 31) && Kp / 32 < 65536, -2, "split_cols_t: too large");
    dim3 grid(unsigned(ceildiv(N, 64)), unsigned(Kp / 32));
    switch (ns) {
        case 1: hipLaunchKernelGGL(split_cols_t_kernel<1>, grid, dim3(256), 0, s, src, ld, K, N, dst, Kp, plane); break;
        case 2: hipLaunchKernelGGL(split_cols_t_kernel<2>, grid, dim3(256), 0, s, src, ld, K, N, dst, Kp, plane); break;
        case 3: hipLaunchKernelGGL(split_cols_t_kernel<3>, grid, dim3(256), 0, s, src, ld, K, N, dst, Kp, plane); break;
        default: SA_REQUIRE(false, -2, "split_cols_t: ns must be 1..3");
    }
    SA_CHECK_LAUNCH();
}

void gemm_bf16s(int ns, int64_t M, int64_t N, int64_t Kp, const __bf16* A, int64_t lda, int64_t planeA,
                const __bf16* B, int64_t ldb, int64_t planeB, float* C, int64_t ldc, float alpha,
                hipStream_t s, int role, int uplo)
{
    if (M <= 0 || N <= 0 || Kp <= 0) return;
    SA_REQUIRE(Kp % 32 == 0 && lda % 8 == 0 && ldb % 8 == 0 && planeA % 8 == 0 && planeB % 8 == 0, -2,
               "gemm_bf16s: operands must be 16-byte aligned in k");
    SA_REQUIRE(lda >= Kp && ldb >= Kp, -2, "gemm_bf16s: leading dimension below Kp");
    SA_REQUIRE(reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0, -2,
               "gemm_bf16s: operand base not 16-byte aligned");
    int64_t nt = ceildiv(M, 128) * ceildiv(N, 128);
    SA_REQUIRE(nt < (int64_t(1) << 31), -2, "gemm_bf16s: grid too large");
#define SA_BF(NSV, R) hipLaunchKernelGGL((gemm_bf16s_kernel<NSV, R>), dim3(unsigned(nt)), dim3(256), 0, s, \
                                         M, N, Kp, A, lda, planeA, B, ldb, planeB, C, ldc, alpha, uplo)
    if (role == 1) {
        switch (ns) { case 1: SA_BF(1, 1); break; case 2: SA_BF(2, 1); break; case 3: SA_BF(3, 1); break;
                      default: SA_REQUIRE(false, -2, "gemm_bf16s: ns must be 1..3"); }
    }
    else {
        switch (ns) { case 1: SA_BF(1, 0); break; case 2: SA_BF(2, 0); break; case 3: SA_BF(3, 0); break;
                      default: SA_REQUIRE(false, -2, "gemm_bf16s: ns must be 1..3"); }
    }
#undef SA_BF
    SA_CHECK_LAUNCH();
}

} // namespace sa
