// Split-BF16 trailing update on 256 x 256 tiles with direct global->LDS
// staging (gfx950).
//
// C (fp32 row-major, ldc) += alpha * sum_{i+j<NS} A_i B_j^T over pre-split
// bf16 planes (A_i: M x Kp, B_j: N x Kp, k contiguous, row stride Kp, zero
// padded past K; the planes split_rows / split_cols_t write).  The float LU's
// trailing update C -= L21 U12 (src/internal/internal_gemm.cc:355-518 for the
// low-precision matrix) at the large early steps, where the 128 x 128
// split-at-staging kernel spends most of its time staging operands.
//
// One workgroup per 256 x 256 C tile, eight wave64s in 2 (M) x 4 (N), each a
// 128 x 64 block of 8 x 4 v_mfma_f32_16x16x32_bf16 accumulators.  The
// operands of a 32-deep k slice (NS planes x 256 rows x 64 bytes for each of
// A and B) go from global memory straight into LDS with
// global_load_lds_dwordx4 (no staging registers, no VALU work), in two
// alternating buffers: the loads of slice s+1 are in flight while slice s is
// multiplied, one barrier per slice.  Each wave's load writes 1 KB of LDS
// contiguously (lane l at 16 l: row l/4, 16-byte slot l%4 of a 16-row
// group), so the swizzle is chosen on the global side: slot c of row r holds
// k-chunk c ^ ((r >> 2) & 3), which spreads both the loads' writes and the
// 16-row fragment reads over all LDS banks.
//
// Per C element the accumulation is the same as the 128 x 128 kernels':
// slices in k order, within a slice the term pairs smallest first and x0 y0
// last, one MFMA each, then C += alpha * acc.  So this kernel and gemm_s32
// give bitwise equal results (the LU's lookahead block and its trailing
// update may use either).
#include "common.hh"
#include "kernels.hh"

namespace sa {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void* lds_ptr_t;

constexpr int LT_BM = 256, LT_BN = 256, LT_BK = 32;
constexpr int LT_ROWB = LT_BK * 2;                     // bytes per row per plane and slice
constexpr int LT_OPB = LT_BM * LT_ROWB;                // 16 KB: one operand plane of a slice

__device__ inline void lds_load16(const void* g, lds_ptr_t l)
{
    __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

} // namespace

template <int NS, int ROLE>
__global__ __launch_bounds__(512, 1)
void gemm_lt_kernel(int64_t M, int64_t N, int64_t Kp, const __bf16* __restrict__ A, int64_t lda,
                    int64_t planeA, const __bf16* __restrict__ B, int64_t ldb, int64_t planeB,
                    float* __restrict__ C, int64_t ldc, float alpha, int uplo)
{
    // [buffer][operand A=0 / B=1][plane][256 rows x 64 bytes]
    __shared__ __attribute__((aligned(1024))) char lds[2][2][NS][LT_OPB];

    const int64_t tiles_m = ceildiv(M, LT_BM), tiles_n = ceildiv(N, LT_BN);
    const bool tri = uplo && tiles_m == tiles_n;
    const int64_t ntiles = tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    const int64_t t = xcd_remap(blockIdx.x, ntiles);
    int64_t tm, tn;
    if (tri) {
        upper_tile(t, tiles_m, tm, tn);
    }
    else {
        // each XCD's contiguous run of tiles goes down 8 tile rows before
        // moving right, so the workgroups resident on one XCD share a few
        // L and U stripes in its L2
        constexpr int64_t GM = 8;
        const int64_t group = t / (GM * tiles_n);
        const int64_t first_m = group * GM;
        const int64_t gm = (tiles_m - first_m < GM) ? (tiles_m - first_m) : GM;
        const int64_t within = t - group * GM * tiles_n;
        tm = first_m + within % gm;
        tn = within / gm;
    }
    const int64_t m0 = tm * LT_BM, n0 = tn * LT_BN;
    if (uplo && n0 + LT_BN <= m0) return;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 2, wn = wave & 3;

    // Loads: wave w fills 16-row groups 2w and 2w + 1 of every operand plane.
    // Lane l: row 16 g + l/4, LDS slot l%4, global k-chunk slot ^ ((row >> 2) & 3).
    int64_t aoff[2], boff[2];
    int lofs[2];
    #pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int grp = 2 * wave + h;
        const int row = 16 * grp + (lane >> 2);
        const int chunk = (lane & 3) ^ ((row >> 2) & 3);
        const int64_t ra = m0 + row, rb = n0 + row;
        aoff[h] = (ra < M ? ra : M - 1) * lda + chunk * 8;
        boff[h] = (rb < N ? rb : N - 1) * ldb + chunk * 8;
        lofs[h] = grp * 1024;                          // wave-uniform
    }
    auto issue = [&](int buf, int64_t k0) {
        #pragma unroll
        for (int p = 0; p < NS; ++p)
            #pragma unroll
            for (int h = 0; h < 2; ++h) {
                lds_load16(A + p * planeA + aoff[h] + k0, (lds_ptr_t)(&lds[buf][0][p][lofs[h]]));
                lds_load16(B + p * planeB + boff[h] + k0, (lds_ptr_t)(&lds[buf][1][p][lofs[h]]));
            }
    };

    v4f acc[8][4];
    #pragma unroll
    for (int i = 0; i < 8; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // fragment of rows r0 .. r0 + 15: lane l reads row r0 + (l & 15), k-chunk l >> 4
    const int fr = lane & 15, fg = lane >> 4;
    auto frag = [&](int buf, int op, int p, int r0) {
        const int r = r0 + fr;
        const int slot = fg ^ ((r >> 2) & 3);
        return *reinterpret_cast<const bf16x8_t*>(&lds[buf][op][p][r * LT_ROWB + slot * 16]);
    };

    const int64_t ktiles = Kp / LT_BK;
    issue(0, 0);
    for (int64_t kt = 0; kt < ktiles; ++kt) {
        const int buf = int(kt & 1);
        __builtin_amdgcn_s_waitcnt(0x0f70);            // vmcnt(0): this wave's loads of slice kt
        __syncthreads();                               // every wave's; buffer buf^1 no longer read
        if (kt + 1 < ktiles) issue(buf ^ 1, (kt + 1) * LT_BK);
        bf16x8_t bq[4][NS];
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int p = 0; p < NS; ++p) bq[j][p] = frag(buf, 1, p, wn * 64 + j * 16);
        #pragma unroll
        for (int i = 0; i < 8; ++i) {
            bf16x8_t aq[NS];
            #pragma unroll
            for (int p = 0; p < NS; ++p) aq[p] = frag(buf, 0, p, wm * 128 + i * 16);
            #pragma unroll
            for (int sa_ = NS - 1; sa_ >= 0; --sa_)
                #pragma unroll
                for (int sb = NS - 1 - sa_; sb >= 0; --sb)
                    #pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aq[sa_], bq[j][sb], acc[i][j], 0, 0, 0);
        }
    }

    // C += alpha acc.  Interior tiles read 16 C values per lane with no
    // per-element branches, so the loads of a row block are in flight
    // together (a guarded element-by-element update waits one memory round
    // trip per element).
    const bool full = m0 + LT_BM <= M && n0 + LT_BN <= N && (uplo == 0 || n0 >= m0 + LT_BM);
    if (full) {
        float* cb = C + (m0 + wm * 128 + 4 * (lane >> 4)) * ldc + n0 + wn * 64 + fr;
        #pragma unroll
        for (int i = 0; i < 8; ++i) {
            float cv[4][4];
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r) cv[j][r] = __builtin_nontemporal_load(&cb[(i * 16 + r) * ldc + j * 16]);
            // C is read and written once per launch: nontemporal, so it does
            // not displace the operand planes from L2 (replay of the n = 65536
            // updates: 8.98 -> 8.93 ms per launch)
            #pragma unroll
            for (int j = 0; j < 4; ++j)
                #pragma unroll
                for (int r = 0; r < 4; ++r)
                    __builtin_nontemporal_store(cv[j][r] + alpha * acc[i][j][r], &cb[(i * 16 + r) * ldc + j * 16]);
        }
        return;
    }
    #pragma unroll
    for (int i = 0; i < 8; ++i)
        #pragma unroll
        for (int j = 0; j < 4; ++j)
            #pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm_ = m0 + wm * 128 + i * 16 + 4 * (lane >> 4) + r;
                const int64_t gn_ = n0 + wn * 64 + j * 16 + fr;
                if (gm_ < M && gn_ < N && (uplo == 0 || gn_ >= gm_)) {
                    float* c = C + gm_ * ldc + gn_;
                    *c += alpha * acc[i][j][r];
                }
            }
}

int64_t gemm_lt_tiles(int64_t M, int64_t N) { return ceildiv(M, LT_BM) * ceildiv(N, LT_BN); }

void gemm_bf16_lt(int ns, int64_t M, int64_t N, int64_t Kp, const __bf16* A, int64_t lda, int64_t planeA,
                  const __bf16* B, int64_t ldb, int64_t planeB, float* C, int64_t ldc, float alpha,
                  hipStream_t s, int role, int uplo)
{
    if (M <= 0 || N <= 0 || Kp <= 0) return;
    SA_REQUIRE(ns == 1 || ns == 2, -2, "gemm_bf16_lt: ns must be 1 or 2");
    SA_REQUIRE(Kp % LT_BK == 0 && lda >= Kp && ldb >= Kp && lda % 8 == 0 && ldb % 8 == 0
               && planeA % 8 == 0 && planeB % 8 == 0, -2, "gemm_bf16_lt: planes must be 16-byte aligned in k");
    SA_REQUIRE(reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0, -2,
               "gemm_bf16_lt: operand base not 16-byte aligned");
    SA_REQUIRE(ns == 1 || (planeA >= (M - 1) * lda + Kp && planeB >= (N - 1) * ldb + Kp), -2,
               "gemm_bf16_lt: planes overlap");
    const int64_t tm = ceildiv(M, LT_BM), tn = ceildiv(N, LT_BN);
    int64_t nt = tm * tn;
    if (uplo && tm == tn) nt = tm * (tm + 1) / 2;
    SA_REQUIRE(nt < (int64_t(1) << 31), -2, "gemm_bf16_lt: grid too large");
#define SA_LT(NSV, RV) hipLaunchKernelGGL((gemm_lt_kernel<NSV, RV>), dim3(unsigned(nt)), dim3(512), 0, s, \
                                          M, N, Kp, A, lda, planeA, B, ldb, planeB, C, ldc, alpha, uplo)
    if (role == 1) { if (ns == 1) SA_LT(1, 1); else SA_LT(2, 1); }
    else { if (ns == 1) SA_LT(1, 0); else SA_LT(2, 0); }
#undef SA_LT
    SA_CHECK_LAUNCH();
}

} // namespace sa
