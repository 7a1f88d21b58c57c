// Host-side launchers of the slate_amd device kernels.  All matrices handed to
// the factorization kernels are row-major unless stated: element (i, j) at
// base[i * ld + j].  (A column-major n x n matrix is turned into this layout by
// an in-place transpose, the same conversion SLATE performs with its RowMajor
// device tiles, src/getrf.cc:54-55.)
#pragma once
#include "common.hh"

namespace sa {

enum { A_KC = 0, A_MC = 1 };     // A operand: K-contiguous / M-contiguous
enum { B_NC = 0, B_KC = 1 };     // B operand: N-contiguous / K-contiguous

struct GemmArgs {                // C[M x N] = alpha * A * B + beta * C, C row-major
    int64_t M, N, K;
    const void* A; int64_t lda;
    const void* B; int64_t ldb;
    void* C; int64_t ldc;
    double alpha, beta;
    int uplo;                    // 0: all of C; 1: only entries with col >= row
};

// role 1 marks the LU trailing-matrix update (a separately named kernel).
template <class T> void gemm(const GemmArgs& a, int alay, int blay, hipStream_t s, int role = 0);

// Split-BF16 update (gemm_bf16.hip): split_rows turns an M x K row-major
// fp32 block into ns bf16 planes of M x Kp (plane stride `plane`),
// split_cols_t a K x N row-major block into ns planes of N x Kp (transposed);
// gemm_bf16s: C (fp32 row-major) += alpha * sum_{i+j<ns} A_i B_j^T (uplo = 1:
// only entries with column >= row).
void split_rows(int ns, const float* src, int64_t ld, int64_t M, int64_t K, __bf16* dst, int64_t Kp,
                int64_t plane, hipStream_t s);
void split_cols_t(int ns, const float* src, int64_t ld, int64_t K, int64_t N, __bf16* dst, int64_t Kp,
                  int64_t plane, hipStream_t s);
// 256 x 256-tile split-BF16 update over pre-split planes (direct global->LDS
// staging), ns = 1 or 2; gemm_lt_tiles = its workgroup count without uplo.
void gemm_bf16_lt(int ns, int64_t M, int64_t N, int64_t Kp, const __bf16* A, int64_t lda, int64_t planeA,
                  const __bf16* B, int64_t ldb, int64_t planeB, float* C, int64_t ldc, float alpha,
                  hipStream_t s, int role, int uplo);
int64_t gemm_lt_tiles(int64_t M, int64_t N);
void gemm_bf16s(int ns, int64_t M, int64_t N, int64_t Kp, const __bf16* A, int64_t lda, int64_t planeA,
                const __bf16* B, int64_t ldb, int64_t planeB, float* C, int64_t ldc, float alpha,
                hipStream_t s, int role = 0, int uplo = 0);

// Spectrum-controlled test matrices (matgen.hip): sigma_diag writes sigma_i
// of distribution dist (1 arith, 2 geo, 3 cluster0, 4 cluster1) on the
// diagonal; larft_from_gram builds the upper-triangular T of
// H_1 ... H_b = I - Y T Y^T from G = Y^T Y (both b x b row-major).
void sigma_diag(int dist, double cond, int64_t n, double* A, int64_t lda, hipStream_t s);
void larft_from_gram(const double* G, int64_t b, double* T, hipStream_t s);

// Selector (mlp.hip): per-tile max |a| (tiles of 256, t x t row-major into
// tmax), 16-bin exponent histogram (hist zeroed by the caller), and the MLP
// forward pass / one training step (parameters in one float array).
int mlp_num_params();
int mlp_scratch_per_sample();
void feature_tiles(const double* A, int64_t lda, int64_t n, double* tmax, hipStream_t s);
void feature_tiles_rows(const double* A, int64_t lda, int64_t n, double* tmax, double* part, double* rowsum,
                        hipStream_t s);
void feature_hist(const double* A, int64_t lda, int64_t n, int64_t cstep, int emax, unsigned long long* hist,
                  hipStream_t s);
void mlp_forward(const float* P, const float* X, int B, float* probs, hipStream_t s);
void mlp_train(float* P, float* grad, float* m, float* v, const float* X, const int* y, int B, float* scratch,
               float lr, int t, float* loss, hipStream_t s);

// Batched uneven solves (batched.hip): one persistent launch of `grid`
// workgroups of a tile configuration (float panel width nbf 16 / 32,
// workgroup size bt 256 / 512).  Per-system mode: 0 FP32 factor + FP64 IR,
// 1 FP64, 2 BF16 factor + FP64 IR, 3 learned choice, 4..6 learned path with
// the choice forced to 0..2.  Device arrays throughout; work: grid *
// batched_work_stride() bytes; queue: two ints, zero before the first launch
// (each launch leaves them zero).  batched_fill_args writes the kernel's
// argument block (at most BATCH_ARGS_BYTES) into out.
constexpr int BATCH_ARGS_BYTES = 256;
int64_t batched_work_stride(int elem_bytes = 4);
int batched_max_n();
bool batched_config_valid(int nbf, int bt);
int batched_max_wgs_per_cu(int nbf, int bt);
void batched_fill_args(void* out, int count, const int* order, const int* n, const int* mode, const int64_t* offA,
                       const int64_t* offB, const double* A, const double* B, double* X, int* iters, int* info,
                       double* berr, int* choice, float* feat, unsigned long long* ticks, const float* sel,
                       int* queue, char* work, int itermax, double tol, int probe_bf16);
void batched_launch(const void* args, int nbf, int bt, int grid, hipStream_t s);
void batched_graph(void* args, int nbf, int bt, int grid, hipGraph_t* g, hipGraphExec_t* ge);

// gemm_s32: C (fp32 row-major) += alpha * L R^T with float L (M x K) and
// R (N x K), k contiguous, rows 16-byte aligned; each operand split into ns
// bf16 terms inside the kernel.  transpose_f32: K x N row-major float ->
// N x Kp (zero padded past K).
void gemm_s32(int ns, int64_t M, int64_t N, int64_t K, const float* L, int64_t ldl, const float* R, int64_t ldr,
              float* C, int64_t ldc, float alpha, hipStream_t s, int role = 0, int uplo = 0);
void transpose_f32(const float* src, int64_t ld, int64_t K, int64_t N, float* dst, int64_t Kp, hipStream_t s);

// FP8 e4m3 update (gemm_fp8.hip): operands quantized in tiles of 128 rows
// (128 columns of a K x N block) over the whole k with a power-of-two scale
// per tile (sexp[t] = s, largest with max|x| 2^s <= 448), nt = 1 or 2 planes
// (the second the 2^4-scaled remainder) of fp8_rows(M) x Kp bytes each,
// Kp = fp8_kp(K); gemm_fp8: C (fp32 row-major) += alpha * the scaled product.
int64_t fp8_kp(int64_t K);
int64_t fp8_rows(int64_t M);
int64_t fp8_tiles(int64_t M);
void quant_fp8_rows(int nt, const float* src, int64_t ld, int64_t M, int64_t K, uint8_t* dst, int64_t Kp,
                    int64_t plane, int* sexp, hipStream_t s);
void quant_fp8_cols_t(int nt, const float* src, int64_t ld, int64_t K, int64_t N, uint8_t* dst, int64_t Kp,
                      int64_t plane, int* sexp, hipStream_t s);
void gemm_fp8(int nt, int64_t M, int64_t N, int64_t Kp, const uint8_t* A, int64_t planeA, const int* sa,
              const uint8_t* B, int64_t planeB, const int* sb, float* C, int64_t ldc, float alpha, hipStream_t s,
              int role = 0, int uplo = 0);

// SLATE tile storage (tiles.hip): copy between mt x nt independent column-major
// tiles (device pointer table dtiles, per-tile ld dld, both device arrays,
// tile (i, j) at index i + j*mt, NULL = not stored) and a dense column-major
// matrix D; to_tiles = 1 scatters D into the tiles, 0 gathers.
template <class T>
void tiles_copy(T* const* dtiles, const int64_t* dld, int64_t m, int64_t n, int64_t mb, int64_t nb, T* D,
                int64_t ldd, int to_tiles, hipStream_t s);

// Workspace shared by the cross-workgroup panel kernel.
struct PanelWork {
    void* slots;                 // tagged candidate words exchanged between workgroups
    int64_t slots_bytes;
    int* error;                  // set to 1 if a wait for another workgroup times out
};

// Partial-pivot LU of a leaf panel: rows [r0, m_total) x columns [c0, c0+w) of
// a row-major matrix.  Writes ipiv[c0 + j] = global 0-based pivot row and
// sets *info (1-based global column) at the first exactly-zero pivot.
// Column j exchanges words tagged tag0 + j: the caller hands every launch on
// the same slot area fresh tags (it advances tag0 by w), and slot words start
// at tag 0, which is never used.
// With xpw > 0 the leaf also applies its interchanges to the other columns
// of the enclosing panel [xp0, xp0 + xpw) (rows from r0 = c0 down).
template <class T>
void panel_leaf(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w,
                int64_t* ipiv, int64_t* info, double threshold, const PanelWork& pw,
                uint32_t tag0, hipStream_t s, int64_t xp0 = 0, int64_t xpw = 0);
int panel_leaf_width(int elem_bytes);   // largest leaf width the kernel supports

// Tournament-pivoting leaf (MethodLU::CALU): same contract as panel_leaf
// with the pivots chosen by a butterfly tournament over the workgroups; uses
// tags tag0 .. tag0 + 8.  calu_slot_bytes: exchange area it needs for m rows.
template <class T>
void calu_leaf(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w, int64_t* ipiv, int64_t* info,
               const PanelWork& pw, uint32_t tag0, hipStream_t s);
int calu_leaf_width(int elem_bytes);
int64_t calu_slot_bytes(int elem_bytes, int64_t m);

// Row interchanges: apply swaps (i, ipiv[i]) for i in [k1, k2), in order, to
// columns [c0, c1) of a row-major matrix with nrows rows.  Two launches: a
// one-block plan that turns the sequence into a (dst <- src) row list, then a
// gather/scatter over column chunks.  plan must hold 4 * (k2 - k1) + 1 int64.
// Pivots outside [i, nrows) are treated as "no interchange".
// With a workspace (tmp, tmp_elems >= 2 (k2 - k1) (c1 - c0)) a ready plan is
// applied in two passes instead — gather every source row into tmp, then
// scatter — each a full-width, 16-byte-vector row copy: the one-pass form
// stages whole column chunks of all moved rows in LDS, which limits a chunk
// to a few columns (32-byte row segments) when 2048 rows move.
template <class T>
void laswp(T* A, int64_t ld, int64_t nrows, int64_t c0, int64_t c1, int64_t k1, int64_t k2,
           const int64_t* ipiv, int64_t* plan, hipStream_t s, bool plan_ready = false,
           T* tmp = nullptr, int64_t tmp_elems = 0);
void laswp_plan(int64_t k1, int64_t k2, int64_t nrows, const int64_t* ipiv, int64_t* plan,
                hipStream_t s);
// Turn a (dst <- src) plan into its inverse permutation (dst/src swapped).
void laswp_invert_plan(int64_t* plan, int64_t npiv, hipStream_t s);

// Triangular solve leaf: B (nb x N, row-major, ldb) := op(T)^{-1} B for an
// nb x nb triangle (nb <= TRSM_LEAF).  tri(i,k) = M[i*ld + k] (cm == 0) or
// M[i + k*ld] (cm == 1).  lower/unit select the triangle and diagonal.
constexpr int TRSM_LEAF = 32;
template <class T>
void trsm_leaf(const T* M, int64_t ld, int cm, int lower, int unit, int64_t nb,
               T* B, int64_t ldb, int64_t N, hipStream_t s);
// Unit-lower row-major triangle (m <= trsm_fused_max) solved against a
// row-major B (m x N) in one launch: L X = B in place.
int64_t trsm_fused_max(int elem_bytes);
template <class T>
void trsm_lunit_fused(const T* L, int64_t ldl, int64_t m, T* B, int64_t ldb, int64_t N, hipStream_t s);
// The same for a lower triangle with its own diagonal when unit == 0.
template <class T>
void trsm_lower_fused(const T* L, int64_t ldl, int64_t m, T* B, int64_t ldb, int64_t N, int unit, hipStream_t s);

// One-launch triangular solve of tri * x = b (one right-hand side, x overwrites
// b, contiguous) with a look-back chain between blocks of trsv_block() rows.
// Aux (trsv_aux_elems(n) elements) holds what trsv_diag_inverse prepares per
// factor: the inverted diagonal blocks (the unit/non-unit choice is folded
// into them) and each block's inverse times its last off-diagonal tile.
// TvSync: words = 2n tagged exchange words, progress = one tagged word, both
// zeroed when allocated; tag = nonzero and unused by earlier launches on those
// words; error set to 2 if a wait times out.
struct TvSync {
    uint64_t* words;
    uint64_t* progress;
    int* error;
    uint32_t tag;
};
int64_t trsv_block(int elem_bytes);
int64_t trsv_aux_elems(int64_t n, int elem_bytes);
template <class T>
void trsv_diag_inverse(const T* M, int64_t ld, int cm, int lower, int unit, int64_t n, T* Aux,
                       hipStream_t s);
template <class T>
void trsv_lookback(const T* M, int64_t ld, int cm, int lower, int64_t n, const T* Aux, T* X,
                   const TvSync& sync, hipStream_t s);

// Cholesky leaf on an nb x nb row-major block (nb <= 64), upper factor U with
// A = U^T U stored in the upper triangle.  Sets *info = col0 + j + 1 at the
// first non-positive pivot (if *info == 0).
template <class T>
void potrf_leaf(T* A, int64_t ld, int64_t nb, int64_t col0, int64_t* info, hipStream_t s);

// A := I (n x n, row-major, ld).
template <class T> void set_identity(T* A, int64_t n, int64_t ld, hipStream_t s);
template <class T> void unit_lower_copy(const T* L, int64_t ld, int64_t n, T* out, hipStream_t s);

// In-place transpose of an n x n matrix (column-major <-> row-major).
template <class T> void transpose_inplace(T* A, int64_t n, int64_t ld, hipStream_t s);
// Out-of-place transpose with conversion: dst(j, i) = convert(src(i, j)),
// src m x n with leading dim lds (element (i,j) at src[i + j*lds]).
template <class S, class D>
void transpose_convert(const S* src, int64_t lds, D* dst, int64_t ldd, int64_t m, int64_t n,
                       hipStream_t s);
// Elementwise copy/convert of an m x n column-major block.
template <class S, class D>
void copy_convert(const S* src, int64_t lds, D* dst, int64_t ldd, int64_t m, int64_t n,
                  hipStream_t s);

// ||A||_inf of a column-major matrix (general, or symmetric stored lower).
// Result written to out[0] on the device; work holds ceil(n/256)*m doubles.
void norm_inf(const double* A, int64_t m, int64_t n, int64_t lda, int sym, double* work,
              int64_t work_elems, double* out, hipStream_t s);
// Per-column max |X(:, j)| into out[j].
void col_norms_max(const double* X, int64_t m, int64_t n, int64_t ldx, double* out,
                   hipStream_t s);
// R = B - op(A) X for column-major A (general or symmetric-lower), nrhs <= 4,
// deterministic two-pass reduction through work.
void residual(const double* A, int64_t n, int64_t lda, int sym, const double* X, int64_t ldx,
              const double* B, int64_t ldb, double* R, int64_t ldr, int64_t nrhs,
              double* work, int64_t work_elems, hipStream_t s);
int64_t residual_work_elems(int64_t n, int64_t nrhs);
// Y += alpha * X (column-major m x n).
void axpy(int64_t m, int64_t n, double alpha, const double* X, int64_t ldx, double* Y,
          int64_t ldy, hipStream_t s);

// Dense dot-products used by GMRES: h[p] = sum_i V[i + p*ldv] * w[i], p < k.
void multi_dot(const double* V, int64_t ldv, int64_t n, int64_t k, const double* w,
               double* h, double* work, hipStream_t s);
// w -= V(:, 0:k) h
void multi_axpy(const double* V, int64_t ldv, int64_t n, int64_t k, const double* h,
                double* w, hipStream_t s);
// out[0] = ||w||_2
void nrm2(const double* w, int64_t n, double* work, double* out, hipStream_t s);
void scale(double* w, int64_t n, const double* denom_dev, hipStream_t s);

// dst[i] = src[perm[i]] (scatter = 0) or dst[perm[i]] = src[i] (scatter = 1),
// with conversion; perm is the row permutation equivalent to getrf's ipiv.
template <class S, class D>
void permute_convert(const S* src, D* dst, const int64_t* perm, int64_t n, int scatter, hipStream_t s);
// y += convert(x) for vectors.
template <class S> void add_convert(const S* x, double* y, int64_t n, hipStream_t s);

// Matrix generation (SLATE matgen Philox): A[i + j*lda] = f(seed, i+ioff, j+joff).
void generate(int dist, int64_t seed, int64_t m, int64_t n, int64_t ioff, int64_t joff,
              double* A, int64_t lda, hipStream_t s);

// Random butterfly transforms (rbt.hip): half length of level k, the
// two-sided A := U^T A V, the one-sided B := U^T B (trans = 1) or B := V B
// (trans = 0), and the weights u[i] = exp(rands(seed, i, 0) / 20).
int64_t rbt_half(int64_t n, int64_t nb, int depth, int k);
template <class T>
void gerbt2(T* A, int64_t lda, int64_t n, int64_t nb, int depth, const T* u, const T* v, hipStream_t s);
template <class T>
void gerbt1(int trans, T* B, int64_t ldb, int64_t n, int64_t nrhs, int64_t nb, int depth, const T* u,
            hipStream_t s);
void rbt_weights(int64_t seed, int64_t n, double* u, hipStream_t s);

// Fill an int/unsigned buffer with zeros (graph-safe replacement for memset).
void zero_u32(unsigned* p, int64_t n, hipStream_t s);
void zero_f64(double* p, int64_t n, hipStream_t s);

//------------------------------------------------------------------------------
// No-allocation scratch (noalloc.hip): one device array per GPU, NOALLOC_BYTES,
// holding the workspace of slate_amd_dgetrs_noalloc for n <= NOALLOC_NMAX.
constexpr int64_t NOALLOC_NMAX = 16384;
constexpr size_t NOALLOC_BYTES = size_t(36) << 20;
constexpr size_t noalloc_align(size_t b) { return (b + 255) & ~size_t(255); }
struct NoallocLayout {
    static constexpr size_t ipiv1 = 0;
    static constexpr size_t ipiv0 = ipiv1 + noalloc_align(8 * size_t(NOALLOC_NMAX));
    static constexpr size_t plan = ipiv0 + noalloc_align(8 * size_t(NOALLOC_NMAX));
    static constexpr size_t plan_words = 4 * 2048 + 8;
    static constexpr size_t err = plan + noalloc_align(8 * plan_words);
    static constexpr size_t tvw = err + 256;
    static constexpr size_t tv_words = 2 * size_t(NOALLOC_NMAX) + 32;
    static constexpr size_t dinv = tvw + noalloc_align(8 * tv_words);
    static constexpr size_t dinv_per = 2 * size_t(NOALLOC_NMAX) * 64; // doubles per triangle (64-row blocks)
    static constexpr size_t end = dinv + 2 * 8 * dinv_per;
};
static_assert(NoallocLayout::end <= NOALLOC_BYTES, "no-allocation scratch too small");
void* noalloc_scratch_setup(int device);
void ipiv_zero_based(const int64_t* ipiv1, int64_t* ipiv0, int64_t n, hipStream_t s);

} // namespace sa
