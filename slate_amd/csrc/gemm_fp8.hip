// FP8 e4m3 MFMA update for the low-precision LU (gfx950).
//
// factor_precision FP8 / FP8X2: the float factor's trailing update
// C -= L U (src/internal/internal_gemm.cc:355-518 for the low-precision
// matrix) multiplies OCP E4M3FN operands on v_mfma_scale_f32_16x16x128_f8f6f4
// with FP32 accumulation.  Each operand is quantized once per update, in tiles
// of 128 rows of the L panel (128 columns of the U row block) over the whole
// inner dimension k, with a power-of-two scale per tile: the largest 2^s with
// max |x| 2^s <= 448 (the e4m3 maximum), q0 = e4m3(x 2^s), round to nearest
// even.  FP8X2 adds the remainder plane q1 = e4m3((x 2^s - q0) 2^4) and the
// products q0 q0' + 2^-4 (q0 q1' + q1 q0'); the MFMA's E8M0 operand scales
// fold the 2^-4 in, so one accumulator holds all three.  The tile scales are
// applied to the accumulator at the end: C += alpha 2^-(sa+sb) acc.
//
// Operand planes: rows padded to a multiple of 128 (zero rows), k padded to Kp
// (a multiple of 128, zero columns), k contiguous, one byte per element, so
// the GEMM needs no edge handling on loads.
#include "common.hh"
#include "kernels.hh"

namespace sa {

typedef int v8i_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int FP8_TILE = 128;

__device__ inline int fp8_scale_exp(float amax)
{
    // largest s with amax 2^s <= 448 = 0.875 2^9 (0 for an all-zero tile)
    if (!(amax > 0.f)) return 0;
    int e;
    const float f = frexpf(amax, &e);
    return f <= 0.875f ? 9 - e : 8 - e;
}

// Four floats (already scaled) -> four e4m3 bytes (round to nearest even).
__device__ inline unsigned fp8x4(float a, float b, float c, float d)
{
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return unsigned(w);
}

__device__ inline float fp8_value(unsigned w, int byte)
{
    switch (byte) {
        case 0: return __builtin_amdgcn_cvt_f32_fp8(int(w), 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8(int(w), 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8(int(w), 2);
        default: return __builtin_amdgcn_cvt_f32_fp8(int(w), 3);
    }
}

// Quantize 4 values (x already multiplied by 2^s): plane-0 word, and with
// NT = 2 the remainder plane word.
template <int NT>
__device__ inline void quant4(const float (&x)[4], unsigned (&w)[2])
{
    w[0] = fp8x4(x[0], x[1], x[2], x[3]);
    if constexpr (NT > 1) {
        float r[4];
        #pragma unroll
        for (int e = 0; e < 4; ++e) r[e] = (x[e] - fp8_value(w[0], e)) * 16.0f;
        w[1] = fp8x4(r[0], r[1], r[2], r[3]);
    }
}

__device__ inline float block_max(float v, float* red)
{
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float m = red[0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
    __syncthreads();
    return m;
}

//------------------------------------------------------------------------------
// src: M x K row-major float (ld), k contiguous -> NT planes of Mp x Kp bytes
// (plane stride `plane`), sexp[tile] = s.  One workgroup per 128-row tile.
template <int NT>
__global__ __launch_bounds__(256)
void quant_rows_kernel(const float* __restrict__ src, int64_t ld, int64_t M, int64_t K, uint8_t* __restrict__ dst,
                       int64_t Kp, int64_t plane, int* __restrict__ sexp)
{
    __shared__ float red[4];
    const int64_t r0 = int64_t(blockIdx.x) * FP8_TILE;
    const int64_t rows = M - r0 < FP8_TILE ? M - r0 : FP8_TILE;
    float amax = 0.f;
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t k = threadIdx.x; k < K; k += 256) amax = fmaxf(amax, fabsf(src[(r0 + r) * ld + k]));
    amax = block_max(amax, red);
    const int s = fp8_scale_exp(amax);
    if (threadIdx.x == 0) sexp[blockIdx.x] = s;
    for (int64_t r = 0; r < FP8_TILE; ++r)
        for (int64_t k = 4 * int64_t(threadIdx.x); k < Kp; k += 4 * 256) {
            float x[4];
            #pragma unroll
            for (int e = 0; e < 4; ++e)
                x[e] = (r < rows && k + e < K) ? ldexpf(src[(r0 + r) * ld + k + e], s) : 0.f;
            unsigned w[2];
            quant4<NT>(x, w);
            #pragma unroll
            for (int p = 0; p < NT; ++p)
                *reinterpret_cast<unsigned*>(dst + p * plane + (r0 + r) * Kp + k) = w[p];
        }
}

// src: K x N row-major float (ld) -> NT planes of Np x Kp bytes holding the
// transpose (k contiguous), sexp[tile] = s.  One workgroup per 128-column tile;
// 32-row chunks of the tile pass through LDS so reads and writes are both
// contiguous.
template <int NT>
__global__ __launch_bounds__(256)
void quant_cols_t_kernel(const float* __restrict__ src, int64_t ld, int64_t K, int64_t N,
                         uint8_t* __restrict__ dst, int64_t Kp, int64_t plane, int* __restrict__ sexp)
{
    __shared__ float red[4];
    __shared__ float t[32][FP8_TILE + 1];
    const int64_t c0 = int64_t(blockIdx.x) * FP8_TILE;
    const int64_t cols = N - c0 < FP8_TILE ? N - c0 : FP8_TILE;
    const int tid = threadIdx.x;
    float amax = 0.f;
    if ((tid & (FP8_TILE - 1)) < cols)
        for (int64_t k = tid >> 7; k < K; k += 2) amax = fmaxf(amax, fabsf(src[k * ld + c0 + (tid & (FP8_TILE - 1))]));
    amax = block_max(amax, red);
    const int s = fp8_scale_exp(amax);
    if (tid == 0) sexp[blockIdx.x] = s;
    const int cc = tid & (FP8_TILE - 1), kh = (tid >> 7) * 16;     // column, 16-k half of the chunk
    for (int64_t k0 = 0; k0 < Kp; k0 += 32) {
        #pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int kk = (tid >> 7) + 2 * u;
            const int64_t k = k0 + kk;
            t[kk][cc] = (k < K && cc < cols) ? ldexpf(src[k * ld + c0 + cc], s) : 0.f;
        }
        __syncthreads();
        unsigned w[NT][4];
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float x[4] = {t[kh + 4 * q][cc], t[kh + 4 * q + 1][cc], t[kh + 4 * q + 2][cc], t[kh + 4 * q + 3][cc]};
            unsigned ww[2];
            quant4<NT>(x, ww);
            #pragma unroll
            for (int p = 0; p < NT; ++p) w[p][q] = ww[p];
        }
        #pragma unroll
        for (int p = 0; p < NT; ++p)
            *reinterpret_cast<u32x4_t*>(dst + p * plane + (c0 + cc) * Kp + k0 + kh) =
                u32x4_t{w[p][0], w[p][1], w[p][2], w[p][3]};
        __syncthreads();
    }
}

//------------------------------------------------------------------------------
// C (fp32 row-major, ldc) += alpha 2^-(sa[tm] + sb[tn]) sum_k A(m, k) B(n, k)
// over the e4m3 planes (NT = 2: plus the 2^-4 cross products).  128 x 128 C
// tile per workgroup, 4 wave64s in 2 x 2, each a 64 x 64 block of 16 x 16
// accumulators; one 128-deep k slice per MFMA.  LDS holds NT planes per
// operand, 128 rows x 8 chunks of 16 bytes; chunk c of row r in slot
// c ^ (r & 7).  The next slice is loaded into registers while the current one
// is multiplied; uplo = 1 skips tiles below the diagonal and updates only
// column >= row (the herk of the Cholesky update).
template <int NT, int ROLE>
__global__ __launch_bounds__(256, NT == 1 ? 2 : 1)
void gemm_f8_kernel(int64_t M, int64_t N, int64_t Kp, const uint8_t* __restrict__ A, int64_t planeA,
                    const int* __restrict__ sa, const uint8_t* __restrict__ B, int64_t planeB,
                    const int* __restrict__ sb, float* __restrict__ C, int64_t ldc, float alpha, int uplo)
{
    constexpr int BM = 128, BN = 128, BK = 128;
    __shared__ u32x4_t As[NT][BM * 8];
    __shared__ u32x4_t Bs[NT][BN * 8];

    const int64_t tiles_m = ceildiv(M, BM), tiles_n = ceildiv(N, BN);
    const bool tri = uplo && tiles_m == tiles_n;
    const int64_t ntiles = tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    const int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    if (tri) {
        upper_tile(t, tiles_m, tm, tn);
    }
    else {
        constexpr int64_t GM = 8;
        const int64_t group = t / (GM * tiles_n);
        const int64_t first_m = group * GM;
        const int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
        const int64_t within = t - group * GM * tiles_n;
        tm = first_m + within % gm;
        tn = within / gm;
    }
    const int64_t m0 = tm * BM, n0 = tn * BN;
    if (uplo && n0 + BN <= m0) return;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // staging: chunk lc of rows lr + 32 u (u = 0..3) of each tile
    const int lr = tid >> 3, lc = tid & 7;
    u32x4_t ra[NT][4], rb[NT][4];
    auto load = [&](int64_t k0) {
        #pragma unroll
        for (int p = 0; p < NT; ++p)
            #pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t r = lr + 32 * u;
                ra[p][u] = *reinterpret_cast<const u32x4_t*>(A + p * planeA + (m0 + r) * Kp + k0 + lc * 16);
                rb[p][u] = *reinterpret_cast<const u32x4_t*>(B + p * planeB + (n0 + r) * Kp + k0 + lc * 16);
            }
    };
    auto stage = [&]() {
        #pragma unroll
        for (int p = 0; p < NT; ++p)
            #pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = lr + 32 * u;
                const int slot = r * 8 + (lc ^ (r & 7));
                As[p][slot] = ra[p][u];
                Bs[p][slot] = rb[p][u];
            }
    };

    v4f acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // fragment of a 16 x 128 block: lane l holds row l & 15, k bytes
    // 32 (l >> 4) .. + 32 (chunks 2 (l >> 4) and 2 (l >> 4) + 1); the same map
    // for both operands, so each k meets its partner whatever the instruction's
    // internal k order
    const int fr = lane & 15, fc = 2 * (lane >> 4);
    auto frag = [&](const u32x4_t* S, int row) {
        const u32x4_t lo = S[row * 8 + (fc ^ (row & 7))];
        const u32x4_t hi = S[row * 8 + ((fc + 1) ^ (row & 7))];
        return v8i_t{int(lo[0]), int(lo[1]), int(lo[2]), int(lo[3]), int(hi[0]), int(hi[1]), int(hi[2]), int(hi[3])};
    };
    auto compute = [&]() {
        v8i_t bq[4][NT];
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int p = 0; p < NT; ++p) bq[j][p] = frag(Bs[p], wn * 64 + j * 16 + fr);
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            v8i_t aq[NT];
            #pragma unroll
            for (int p = 0; p < NT; ++p) aq[p] = frag(As[p], wm * 64 + i * 16 + fr);
            #pragma unroll
            for (int j = 0; j < 4; ++j) {
                // E8M0 operand scales: 127 = 2^0, 123 = 2^-4 (remainder planes)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aq[0], bq[j][0], acc[i][j],
                                                                             0, 0, 0, 127, 0, 127);
                if constexpr (NT > 1) {
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aq[0], bq[j][1], acc[i][j],
                                                                                 0, 0, 0, 127, 0, 123);
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aq[1], bq[j][0], acc[i][j],
                                                                                 0, 0, 0, 123, 0, 127);
                }
            }
        }
    };

    const int64_t ktiles = Kp / BK;
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

    const float scale = ldexpf(alpha, -(sa[tm] + sb[tn]));
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
                    *c += scale * acc[i][j][r];
                }
            }
}

//------------------------------------------------------------------------------
int64_t fp8_kp(int64_t K) { return ceildiv(K, FP8_TILE) * FP8_TILE; }
int64_t fp8_rows(int64_t M) { return ceildiv(M, FP8_TILE) * FP8_TILE; }
int64_t fp8_tiles(int64_t M) { return ceildiv(M, FP8_TILE); }

void quant_fp8_rows(int nt, const float* src, int64_t ld, int64_t M, int64_t K, uint8_t* dst, int64_t Kp,
                    int64_t plane, int* sexp, hipStream_t s)
{
    if (M <= 0) return;
    SA_REQUIRE(Kp % FP8_TILE == 0 && Kp >= K && K > 0, -2, "quant_fp8_rows: bad Kp");
    SA_REQUIRE(plane >= fp8_rows(M) * Kp && Kp % 16 == 0, -2, "quant_fp8_rows: plane too small");
    SA_REQUIRE(fp8_tiles(M) < (int64_t(1) << 31), -2, "quant_fp8_rows: too large");
    const dim3 grid(unsigned(fp8_tiles(M)));
    if (nt == 1) hipLaunchKernelGGL(quant_rows_kernel<1>, grid, dim3(256), 0, s, src, ld, M, K, dst, Kp, plane, sexp);
    else if (nt == 2) hipLaunchKernelGGL(quant_rows_kernel<2>, grid, dim3(256), 0, s, src, ld, M, K, dst, Kp, plane, sexp);
    else SA_REQUIRE(false, -2, "quant_fp8_rows: terms must be 1 or 2");
    SA_CHECK_LAUNCH();
}

void quant_fp8_cols_t(int nt, const float* src, int64_t ld, int64_t K, int64_t N, uint8_t* dst, int64_t Kp,
                      int64_t plane, int* sexp, hipStream_t s)
{
    if (N <= 0) return;
    SA_REQUIRE(Kp % FP8_TILE == 0 && Kp >= K && K > 0, -2, "quant_fp8_cols_t: bad Kp");
    SA_REQUIRE(plane >= fp8_rows(N) * Kp, -2, "quant_fp8_cols_t: plane too small");
    SA_REQUIRE(fp8_tiles(N) < (int64_t(1) << 31), -2, "quant_fp8_cols_t: too large");
    const dim3 grid(unsigned(fp8_tiles(N)));
    if (nt == 1) hipLaunchKernelGGL(quant_cols_t_kernel<1>, grid, dim3(256), 0, s, src, ld, K, N, dst, Kp, plane, sexp);
    else if (nt == 2) hipLaunchKernelGGL(quant_cols_t_kernel<2>, grid, dim3(256), 0, s, src, ld, K, N, dst, Kp, plane, sexp);
    else SA_REQUIRE(false, -2, "quant_fp8_cols_t: terms must be 1 or 2");
    SA_CHECK_LAUNCH();
}

void gemm_fp8(int nt, int64_t M, int64_t N, int64_t Kp, const uint8_t* A, int64_t planeA, const int* sa,
              const uint8_t* B, int64_t planeB, const int* sb, float* C, int64_t ldc, float alpha, hipStream_t s,
              int role, int uplo)
{
    if (M <= 0 || N <= 0 || Kp <= 0) return;
    SA_REQUIRE(Kp % FP8_TILE == 0, -2, "gemm_fp8: Kp must be a multiple of 128");
    SA_REQUIRE(planeA >= fp8_rows(M) * Kp && planeB >= fp8_rows(N) * Kp, -2, "gemm_fp8: planes too small");
    SA_REQUIRE(reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0
               && planeA % 16 == 0 && planeB % 16 == 0, -2, "gemm_fp8: planes must be 16-byte aligned");
    const int64_t tm = fp8_tiles(M), tn = fp8_tiles(N);
    int64_t ntl = tm * tn;
    if (uplo && tm == tn) ntl = tm * (tm + 1) / 2;
    SA_REQUIRE(ntl < (int64_t(1) << 31), -2, "gemm_fp8: grid too large");
#define SA_F8(NTV, RV) hipLaunchKernelGGL((gemm_f8_kernel<NTV, RV>), dim3(unsigned(ntl)), dim3(256), 0, s, \
                                          M, N, Kp, A, planeA, sa, B, planeB, sb, C, ldc, alpha, uplo)
    if (role == 1) {
        if (nt == 1) SA_F8(1, 1); else if (nt == 2) SA_F8(2, 1); else SA_REQUIRE(false, -2, "gemm_fp8: terms 1 or 2");
    }
    else {
        if (nt == 1) SA_F8(1, 0); else if (nt == 2) SA_F8(2, 0); else SA_REQUIRE(false, -2, "gemm_fp8: terms 1 or 2");
    }
#undef SA_F8
    SA_CHECK_LAUNCH();
}

} // namespace sa
