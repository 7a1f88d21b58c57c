/*
 * slate_amd — C ABI of the MI355X-native factor-and-solve library.
 *
 * All matrix arguments are DEVICE pointers to column-major storage with a
 * leading dimension, i.e. the memory that slate::Matrix<T>::fromLAPACK /
 * fromDevices describes: tile (i, j) of an nb-blocked matrix is the nb x nb
 * view at A + i*nb + j*nb*lda.  Pivots are HOST arrays of int64 in LAPACK
 * convention (1-based global row indices), the form lapack_api/lapack_gesv.cc
 * :90-100 converts slate::Pivots into.
 *
 * Return value of every solver entry point: >= 0 is SLATE's `info`
 * (0 = success, i > 0 = U(i,i) exactly zero / leading minor i not positive
 * definite); < 0 is an error code, with a message in slate_amd_last_error().
 * No C++ exception crosses this boundary.  Every entry point is synchronous:
 * it waits for the device work it issued before returning.
 */
#ifndef SLATE_AMD_H
#define SLATE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slate_amd_context_t* slate_amd_context;

/* Factorization precision of the mixed-precision drivers.  The BF16 forms keep
 * a float factor whose trailing updates run on bf16 MFMA with fp32
 * accumulation, each float operand split into 1, 2 or 3 bf16 terms (about 8,
 * 16 or 24 significant bits per product; BF16X3 is float-accurate).  The FP8
 * forms run them on fp8 MFMA: operands quantized to OCP e4m3 with a
 * power-of-two scale per tile of 128 rows (the largest 2^s with
 * max|x| 2^s <= 448), one plane (about 4 significant bits) or two (the
 * remainder in a second plane, about 8 bits, 3 products); block size a
 * multiple of 128.  Only SLATE's double -> float exists in the reference
 * (src/gesv_mixed.cc:293-316); the others are this library's. */
enum {
    SLATE_AMD_PREC_FP32 = 0,   /* float factor (SLATE's scalar_lo = float)       */
    SLATE_AMD_PREC_BF16 = 1,   /* float storage, bf16 MFMA updates, fp32 accum   */
    SLATE_AMD_PREC_BF16X2 = 2, /* two-term bf16 split: 3 products per update     */
    SLATE_AMD_PREC_FP64 = 3,   /* no refinement: plain FP64 factor and solve     */
    SLATE_AMD_PREC_BF16X3 = 4, /* three-term bf16 split: 6 products per update   */
    SLATE_AMD_PREC_FP8 = 5,    /* float storage, e4m3 MFMA updates, tile scales  */
    SLATE_AMD_PREC_FP8X2 = 6   /* e4m3 with a remainder plane: 3 products        */
};

enum { SLATE_AMD_LOWER = 0, SLATE_AMD_UPPER = 1 };   /* slate::Uplo                */
enum { SLATE_AMD_NOTRANS = 0, SLATE_AMD_TRANS = 1 }; /* slate::Op                  */

/* Mirrors the slate::Option values used on this path (include/slate/enums.hh
 * :453-490, include/slate/types.hh:241-265).  A field left at its
 * slate_amd_options_init() value takes SLATE's default. */
typedef struct {
    int64_t block_size;        /* nb; 0 = choose automatically                   */
    int64_t inner_blocking;    /* Option::InnerBlocking: panel leaf width, 0=auto */
    int64_t lookahead;         /* Option::Lookahead (default 1)                  */
    double  pivot_threshold;   /* Option::PivotThreshold (default 1.0)           */
    int64_t max_iterations;    /* Option::MaxIterations (default 30)             */
    double  tolerance;         /* Option::Tolerance; <= 0 -> eps * sqrt(n)       */
    int32_t use_fallback;      /* Option::UseFallbackSolver (default 1)          */
    int32_t factor_precision;  /* SLATE_AMD_PREC_* for the mixed drivers         */
    int32_t method_lu;         /* Option::MethodLU: SLATE_AMD_LU_* (default PP)  */
    int32_t depth;             /* Option::Depth: butterfly depth of gesv_rbt (2) */
    int32_t use_graph;         /* replay the LU step sequence from a HIP graph   */
                               /* recorded on the first call with the same       */
                               /* matrix, pivots, sizes and options (default 0;  */
                               /* no counterpart in the reference)               */
} slate_amd_options;

/* slate::MethodLU (include/slate/enums.hh:294-301) for getrf/gesv and the
 * float factor of the mixed drivers. */
enum { SLATE_AMD_LU_PARTIALPIV = 0, SLATE_AMD_LU_NOPIV = 1, SLATE_AMD_LU_CALU = 2 };

/* ---- context ------------------------------------------------------------ */
/* Creates the per-device state (two HIP streams, workspaces).  No counterpart
 * in the reference: SLATE keeps this state in MatrixStorage (queues,
 * include/slate/internal/MatrixStorage.hh:574-590) and the Memory pool
 * (src/core/Memory.cc:17-140). */
int  slate_amd_init(slate_amd_context* ctx, int device);
int  slate_amd_finalize(slate_amd_context ctx);
const char* slate_amd_last_error(void);
void slate_amd_options_init(slate_amd_options* opts);
/* Device milliseconds of the last solver call (for benchmarking). */
double slate_amd_last_time_ms(slate_amd_context ctx);

/* ---- LU ----------------------------------------------------------------- */
/* slate::getrf(Matrix<T>& A, Pivots&, Options) — include/slate/slate.hh:574-577,
 * src/getrf.cc:319-356.  Square n x n only. */
int64_t slate_amd_dgetrf(slate_amd_context ctx, int64_t n, double* dA, int64_t lda,
                         int64_t* ipiv, const slate_amd_options* opts);
int64_t slate_amd_sgetrf(slate_amd_context ctx, int64_t n, float* dA, int64_t lda,
                         int64_t* ipiv, const slate_amd_options* opts);

/* slate::getrs(A, Pivots, B, Options) — slate.hh:603-607, src/getrs.cc:66-117.
 * trans = SLATE_AMD_TRANS solves A^T X = B (A.op() == Op::Trans). */
int64_t slate_amd_dgetrs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs,
                         const double* dA, int64_t lda, const int64_t* ipiv,
                         double* dB, int64_t ldb, const slate_amd_options* opts);
int64_t slate_amd_sgetrs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs,
                         const float* dA, int64_t lda, const int64_t* ipiv,
                         float* dB, int64_t ldb, const slate_amd_options* opts);

/* getrs for one right-hand side (B := A^{-1} B, NoTrans) that allocates
 * nothing: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under graph capture, serving loops, harnesses
 * that forbid hipMalloc).  Its workspace (device pivots, interchange plans,
 * inverted diagonal blocks, look-back exchange words) is a static device
 * array of 36 MiB that every GPU the library runs on holds until the process
 * ends; n <= 16384 (larger n: -3).  Calls on one device share that array:
 * calls that overlap on one device race.  ipiv: host, 1-based (getrf's).
 * Same result, bit for bit, as slate_amd_dgetrs with nrhs = 1. */
int64_t slate_amd_dgetrs_noalloc(slate_amd_context ctx, int64_t n, const double* dA, int64_t lda,
                                 const int64_t* ipiv, double* dB);

/* slate::gesv(A, Pivots, B, Options) — slate.hh:504-508, src/gesv.cc:85-109. */
int64_t slate_amd_dgesv(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                        int64_t* ipiv, double* dB, int64_t ldb, const slate_amd_options* opts);

/* ---- Cholesky ----------------------------------------------------------- */
/* slate::potrf(HermitianMatrix<T>&, Options) — slate.hh:714-717, src/potrf.cc:23-283. */
int64_t slate_amd_dpotrf(slate_amd_context ctx, int uplo, int64_t n, double* dA, int64_t lda,
                         const slate_amd_options* opts);
int64_t slate_amd_spotrf(slate_amd_context ctx, int uplo, int64_t n, float* dA, int64_t lda,
                         const slate_amd_options* opts);
/* slate::potrs — slate.hh:741-745, src/potrs.cc:55-77. */
int64_t slate_amd_dpotrs(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs,
                         const double* dA, int64_t lda, double* dB, int64_t ldb,
                         const slate_amd_options* opts);
/* slate::posv — slate.hh:646-650, src/posv.cc:71-94. */
int64_t slate_amd_dposv(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, double* dA,
                        int64_t lda, double* dB, int64_t ldb, const slate_amd_options* opts);

/* ---- mixed precision + iterative refinement ------------------------------ */
/* slate::gesv_mixed<double,float> — slate.hh:521-535, src/gesv_mixed.cc:106-290.
 * A is left unchanged unless the FP64 fallback runs.  *iter as in SLATE:
 * >= 0 iterations, -3 low-precision factor singular, -(itermax+1) no
 * convergence. */
int64_t slate_amd_dsgesv(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                         int64_t* ipiv, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                         int* iter, const slate_amd_options* opts);
/* slate::gesv_mixed_gmres — slate.hh:539-553, src/gesv_mixed_gmres.cc:111-410 (nrhs = 1). */
int64_t slate_amd_dsgesv_gmres(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA,
                               int64_t lda, int64_t* ipiv, const double* dB, int64_t ldb,
                               double* dX, int64_t ldx, int* iter, const slate_amd_options* opts);
/* slate::posv_mixed — slate.hh:667-681, src/posv_mixed.cc:105-292. */
int64_t slate_amd_dsposv(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, double* dA,
                         int64_t lda, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                         int* iter, const slate_amd_options* opts);
/* slate::posv_mixed_gmres — slate.hh:687-701, src/posv_mixed_gmres.cc:110-410 (nrhs = 1). */
int64_t slate_amd_dsposv_gmres(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs,
                               double* dA, int64_t lda, const double* dB, int64_t ldb,
                               double* dX, int64_t ldx, int* iter, const slate_amd_options* opts);

/* ---- random butterfly transform solver ---------------------------------- */
/* slate::gesv_rbt(A, B, X, iter, Options) — slate.hh:558-563,
 * src/gesv_rbt.cc:94-236.  A := U^T A V (depth opts->depth, seed 42), LU
 * without pivoting, X = V (LU)^{-1} U^T B, FP64 refinement against a copy of
 * A with the PP gesv as fallback.  On return A holds the factors of U^T A V,
 * or the original A when the fallback ran.  *iter: >= 0 refinement steps,
 * -(itermax+1) no convergence.  The weights come from matgen rands (seeds 42
 * and 42 + ceil(n/nb)), not LAPACK larnv. */
int64_t slate_amd_dgesv_rbt(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                            const double* dB, int64_t ldb, double* dX, int64_t ldx, int* iter,
                            const slate_amd_options* opts);
/* internal::rbt_generate — src/internal/internal_rbt_generate.cc:26-140:
 * du[i] = exp(r_i / 20), r_i the matgen rands entry (i, 0) of `seed`.  The
 * vector is column 0 of the reference's n x depth weight matrix, the only
 * column its butterflies read. */
int slate_amd_drbt_generate(slate_amd_context ctx, int64_t n, int64_t seed, double* du);
/* slate::gerbt(U, A, V) — slate.hh:166-169, src/gerbt.cc:124-217: two-sided
 * transform A := U^T A V of depth `depth` over nb-row tiles. */
int slate_amd_dgerbt(slate_amd_context ctx, int64_t n, int depth, int64_t nb, const double* du,
                     const double* dv, double* dA, int64_t lda);
/* slate::gerbt(U, B) — slate.hh:171-173, src/gerbt.cc:251-323: B := U^T B
 * (trans = SLATE_AMD_TRANS) or B := U B (SLATE_AMD_NOTRANS, used with V). */
int slate_amd_dgerbt_rhs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs, int depth, int64_t nb,
                         const double* du, double* dB, int64_t ldb);

/* ---- instrumentation ----------------------------------------------------- */
/* Times every trailing-update GEMM launch of the factorizations with HIP
 * events on the stream it runs on; read returns the summed device time, the
 * algorithmic flops (2*m*n*k per launch) and the launch count since enable.
 * Closest reference counterpart: the slate::timers map (include/slate/slate.hh
 * :35-38) filled by the drivers' Timer blocks. */
int slate_amd_profile_enable(slate_amd_context ctx, int on);
/* Graphs recorded by use_graph (captured), keys whose capture failed and so
 * run eagerly (failed), and replays so far.  No counterpart in the
 * reference. */
int slate_amd_graph_stats(slate_amd_context ctx, int64_t* captured, int64_t* failed, int64_t* replays);
int slate_amd_profile_read(slate_amd_context ctx, double* gemm_ms, double* gemm_flops,
                           int64_t* launches);
/* The same accumulated per trailing-update kernel: SLATE_AMD_PROF_LT (the
 * 256 x 256-tile kernel over pre-split bf16 planes), SLATE_AMD_PROF_SP (the
 * 128 x 128 kernel splitting float operands at LDS staging), or
 * SLATE_AMD_PROF_OTHER (FP64 / FP32 / FP8 / unaligned forms). */
enum { SLATE_AMD_PROF_OTHER = 0, SLATE_AMD_PROF_LT = 1, SLATE_AMD_PROF_SP = 2 };
int slate_amd_profile_read_kernel(slate_amd_context ctx, int kernel, double* gemm_ms, double* gemm_flops,
                                  int64_t* launches);

/* ---- utilities on the path ---------------------------------------------- */
/* slate::gemm(alpha, A, B, beta, C) for column-major device matrices,
 * C (m x n) = alpha * op(A) * op(B) + beta * C, op = NoTrans only —
 * include/slate/slate.hh (gemm), src/gemm.cc:11-105.  The MFMA kernel behind
 * every trailing-matrix update of the factorizations. */
int slate_amd_dgemm(slate_amd_context ctx, int64_t m, int64_t n, int64_t k, double alpha,
                    const double* dA, int64_t lda, const double* dB, int64_t ldb, double beta,
                    double* dC, int64_t ldc);
int slate_amd_sgemm(slate_amd_context ctx, int64_t m, int64_t n, int64_t k, float alpha,
                    const float* dA, int64_t lda, const float* dB, int64_t ldb, float beta,
                    float* dC, int64_t ldc);

/* C (m x n, column-major) += alpha * A * B for float A (m x k), B (k x n),
 * computed by the split-BF16 MFMA kernel behind the BF16 factor precisions:
 * each operand split into `terms` = 1, 2 or 3 bf16 terms, fp32 accumulation.
 * No counterpart in the reference: SLATE's low-precision update is the
 * vendor sgemm (src/internal/internal_gemm.cc:498-504); exposed for the
 * accuracy tests of that kernel. */
int slate_amd_sgemm_bf16(slate_amd_context ctx, int terms, int64_t m, int64_t n, int64_t k, float alpha,
                         const float* dA, int64_t lda, const float* dB, int64_t ldb, float* dC, int64_t ldc);
/* The same on the fp8 MFMA kernel behind the FP8 factor precisions: A
 * quantized to e4m3 per tile of 128 rows, B per tile of 128 columns (each
 * over all of k) with power-of-two scales, terms = 1 or 2 planes, fp32
 * accumulation.  No counterpart in the reference; for the accuracy tests. */
int slate_amd_sgemm_fp8(slate_amd_context ctx, int terms, int64_t m, int64_t n, int64_t k, float alpha,
                        const float* dA, int64_t lda, const float* dB, int64_t ldb, float* dC, int64_t ldc);

/* slate::norm(Norm::Inf, A) — src/norm.cc:24-240; sym != 0: symmetric, lower stored. */
int slate_amd_dnorm_inf(slate_amd_context ctx, int sym, int64_t n, const double* dA, int64_t lda,
                        double* result);
/* matgen random entries (matgen/random.cc:202-231): dist 1 uniform [0,1),
 * 2 uniform [-1,1) ("rands"), 3 normal ("randn"). */
int slate_amd_generate(slate_amd_context ctx, int dist, int64_t seed, int64_t m, int64_t n,
                       int64_t ioffset, int64_t joffset, double* dA, int64_t lda);

/* slate::gecondest(Norm, A, Anorm, Options) — src/gecondest.cc:15-118:
 * reciprocal condition number estimate from the LU factors of getrf
 * (Hager/Higham estimate of ||A^{-1}||, src/internal/internal_norm1est.cc).
 * norm = '1' or 'O' (one norm) or 'I' (infinity norm); anorm = ||A|| in that
 * norm (slate_amd_dnorm_inf for 'I').  Returns 0 or a negative error. */
int64_t slate_amd_dgecondest(slate_amd_context ctx, int norm, int64_t n, const double* dLU, int64_t lda,
                             const int64_t* ipiv, double anorm, double* rcond);
/* slate::pocondest — src/pocondest.cc: the same from the Cholesky factor. */
int64_t slate_amd_dpocondest(slate_amd_context ctx, int uplo, int64_t n, const double* dL, int64_t lda,
                             double anorm, double* rcond);

/* ---- batched uneven solves (no counterpart in the reference) ------------
 * count systems A_i x_i = b_i, n[i] in [1, 1024] (host array), packed on the
 * device back to back: A_i column-major n_i x n_i at dA + sum_{k<i} n_k^2,
 * b_i and x_i at dB / dX + sum_{k<i} n_k.  One persistent launch (a work
 * queue over the systems, largest first; options.use_graph = 1 runs it from
 * a one-node HIP graph, rebuilt only when an argument changes).
 * factor_precision FP32: float factor + FP64 refinement to SLATE's stopping
 * rule with the FP64 fallback; BF16: the same with bf16-operand MFMA
 * trailing updates (FP32 accumulation); FP64: plain FP64 LU.  mode (host,
 * may be NULL): mode[i] = 1 sends system i straight to FP64.  Outputs
 * (host): iters[i] as in dsgesv (0 for FP64), info[i], berr[i] =
 * ||b - A x||_inf / (||A||_inf ||x||_inf).  Per-system steps mirror
 * src/gesv.cc:85-109 and src/gesv_mixed.cc:106-290. */
int64_t slate_amd_dgesv_batched(slate_amd_context ctx, int count, const int* n, const double* dA,
                                const double* dB, double* dX, const int* mode, int factor_precision,
                                int* iters, int* info, double* berr, const slate_amd_options* opts);
/* The same with the factorization precision chosen per system by the
 * selector network (SLATE_AMD_BATCHED_* choices): each system is factored
 * first in the probe precision (options.factor_precision: FP32, the default,
 * or BF16), solved once, and the probe's normwise quantities (n,
 * ||A|| ||x|| / ||b||, ||b - A x|| / (||A|| ||x||), ||b - A x|| / ||b||) go
 * through the network, which picks BF16 + IR (0), FP32 + IR (1) or FP64 (2);
 * the probe factor is kept when it has the chosen precision (with an FP32
 * probe, 0 and 1 both keep it), and a refinement that misses the tolerance
 * falls back to FP64.  params: host array of slate_amd_mlp_num_params()
 * floats (the layout of the selector network); force in 0..2 runs that
 * choice for every system instead (params may then be NULL).  choice (may
 * be NULL): what ran per system; features (may be NULL): count x 24 floats;
 * ms (may be NULL): each system's time on its workgroup. */
enum { SLATE_AMD_BATCHED_BF16 = 0, SLATE_AMD_BATCHED_FP32 = 1, SLATE_AMD_BATCHED_FP64 = 2 };
int64_t slate_amd_dgesv_batched_auto(slate_amd_context ctx, int count, const int* n, const double* dA,
                                     const double* dB, double* dX, const float* params, int force, int* iters,
                                     int* info, double* berr, int* choice, float* features, double* ms,
                                     const slate_amd_options* opts);
/* Graphs built and replayed by the batched entries under use_graph. */
int slate_amd_batched_graph_stats(slate_amd_context ctx, int64_t* builds, int64_t* replays);

/* ---- tuning ---------------------------------------------------------------
 * Block size nb per kind of factorization, used when options.block_size is 0
 * (SLATE takes nb from the matrix tiling; the autotuner of
 * slate_amd/autotune.py measures and installs this table).  Entry i applies
 * to n <= n_upper[i] (ascending); count = 0 restores the built-in choice. */
enum { SLATE_AMD_TUNE_LU_FP64 = 0, SLATE_AMD_TUNE_LU_FP32 = 1, SLATE_AMD_TUNE_LU_BF16 = 2,
       SLATE_AMD_TUNE_CHOL = 3 };
int slate_amd_tune_set_nb(slate_amd_context ctx, int kind, int count, const int64_t* n_upper,
                          const int64_t* nb);
/* The block size a factorization of this kind and size uses when
 * Options::block_size is 0 (the installed table, else the built-in one). */
int64_t slate_amd_tune_get_nb(slate_amd_context ctx, int kind, int64_t n);
/* Tile configuration of the batched solver: float panel width nbf (16 or
 * 32; the FP64 panel is half as wide), workgroup size bt (256, or 512 with
 * nbf 16) and workgroups per CU (at most what LDS and registers allow, reported in
 * max_wgs_per_cu).  Default 32 / 256 / 1 until the autotuner installs its
 * measured choice. */
int slate_amd_tune_set_batched(slate_amd_context ctx, int nbf, int bt, int wgs_per_cu);
int slate_amd_tune_get_batched(slate_amd_context ctx, int* nbf, int* bt, int* wgs_per_cu, int* max_wgs_per_cu);

/* ---- learned solver selection (no counterpart in the reference) ---------
 * For callers that want the factorization precision and refinement scheme
 * picked per system.  A probe factor in BF16X2 (the float LU with two-term
 * split-BF16 updates) and its first solve give the features: 24 floats
 * (log2 n; log10 ||A||_inf ||x0||_inf / ||b||_inf, a condition estimate from
 * x0 = (LU)^{-1} b; per-tile max |a| spread; a 16-bin exponent histogram; a
 * constant; log10 of the probe solve's backward error and relative residual;
 * whether x0 already meets the stopping rule).  x0 and its residual are
 * iteration 0 of the refinement that follows when the probe factor is kept.
 * The MLP (3 layers, 64
 * hidden, FP32; slate_amd_mlp_num_params() floats) maps them to one of
 * SLATE_AMD_SELECTOR_CHOICES choices:
 *   0 BF16X2 + classic IR   1 BF16X2 + GMRES-IR   (keep the probe factor)
 *   2 BF16X3 + classic IR   3 BF16X3 + GMRES-IR   4 FP32 + classic IR
 *   5 FP32 + GMRES-IR       6 FP8X2 + GMRES-IR    7 plain FP64 getrf + getrs
 * Training runs on the GPU with slate_amd_mlp_train_step (softmax
 * cross-entropy, Adam).  Device pointers except where noted. */
enum { SLATE_AMD_SELECTOR_CHOICES = 8 };
int slate_amd_mlp_num_params(void);
int64_t slate_amd_selector_features(slate_amd_context ctx, int64_t n, const double* dA, int64_t lda,
                                    const double* dB, double* dX /* n: receives x0 */,
                                    const slate_amd_options* opts, float* features /* host, 24 */);
int slate_amd_mlp_forward(slate_amd_context ctx, const float* dparams, const float* dX, int batch,
                          float* dprobs);
/* One step on a minibatch (batch <= 4096): dgrad <- d(mean loss)/dparams,
 * then Adam (step >= 1) when lr > 0; *loss (host) <- mean loss. */
int slate_amd_mlp_train_step(slate_amd_context ctx, float* dparams, float* dgrad, float* dm, float* dv,
                             const float* dX, const int* dy, int batch, float lr, int step, float* loss);
/* gesv_mixed with the selector (nrhs = 1).  params (host): the trained MLP;
 * force >= 0: run that choice instead of the network (how the training labels
 * are measured); params NULL and force < 0: SLATE's default, float factor +
 * classic IR with no probe (*choice = 4).  *choice the choice made, *iter as
 * in dsgesv, features (host, 24, may be NULL) the inputs it saw (zeros
 * without a probe).  The call's device time (slate_amd_last_time_ms)
 * includes the probe. */
int64_t slate_amd_dsgesv_auto(slate_amd_context ctx, int64_t n, double* dA, int64_t lda, int64_t* ipiv,
                              const double* dB, double* dX, const float* params, int force, int* iter,
                              int* choice, float* features, const slate_amd_options* opts);

/* Test matrix with a prescribed spectrum: A = U diag(sigma) V^T, U and V
 * products of n Householder reflectors whose vectors are the columns of
 * matgen randn matrices (seed + 1 for U, seed for V); sym != 0 gives
 * A = V diag(sigma) V^T.  sigma_dist 1 arith, 2 geo, 3 cluster0, 4 cluster1
 * with condition number cond (matgen/generate_sigma.hh:61-111).  The
 * reference's "svd" kind (matgen/generate_type_svd.hh:34-200) builds U, V
 * from the QR of a Gaussian matrix instead. */
int slate_amd_generate_svd(slate_amd_context ctx, int sigma_dist, double cond, int64_t seed, int sym,
                           int64_t n, double* dA, int64_t lda);

/* ---- SLATE tile storage -------------------------------------------------
 * A matrix as SLATE's MatrixStorage holds it after insertLocalTiles(
 * Target::Devices) on one device (include/slate/internal/MatrixStorage.hh:151
 * tile map, include/slate/Matrix.hh:164 insertLocalTiles,
 * include/slate/BaseMatrix.hh:2566-2646 tileGet): mt x nt tiles of mb x nb,
 * mt = ceil(m/mb), nt = ceil(n/nb), the last row / column of tiles ragged;
 * tile (i, j) is an independent column-major DEVICE buffer tiles[i + j*mt]
 * with leading dimension lds[i + j*mt] >= its row count.  For the Hermitian
 * (Cholesky) entries only tiles with i >= j are read (lower storage, as
 * SLATE's HermitianMatrix keeps); the others may be NULL.  The two arrays are
 * host memory.  Each entry gathers the tiles into one dense workspace of the
 * context, runs the dense entry of the same name, and writes the outputs back
 * into the tiles (A: factors; B: solution; X: solution; for the mixed drivers
 * A only when the FP64 fallback ran).  Cost: one read and one write of each
 * tile on the way in and out (about 2 x 34 GB of HBM traffic for A at
 * n = 65536), and an n x n FP64 workspace. */
typedef struct {
    int64_t m, n;              /* matrix size                                  */
    int64_t mb, nb;            /* tile size (uniform; last tiles ragged)      */
    void* const* tiles;        /* mt * nt device pointers, tile (i,j) at i+j*mt */
    const int64_t* lds;        /* mt * nt leading dimensions                  */
} slate_amd_tiles;

/* slate::getrf / getrs / gesv on tiles (slate.hh:574-577, :603-607, :504-508). */
int64_t slate_amd_dgetrf_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                               const slate_amd_options* opts);
int64_t slate_amd_dgetrs_tiles(slate_amd_context ctx, int trans, const slate_amd_tiles* A, const int64_t* ipiv,
                               const slate_amd_tiles* B, const slate_amd_options* opts);
int64_t slate_amd_dgesv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                              const slate_amd_tiles* B, const slate_amd_options* opts);
/* slate::potrf / potrs / posv on lower tiles (slate.hh:714-717, :741-745, :646-650). */
int64_t slate_amd_dpotrf_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_options* opts);
int64_t slate_amd_dpotrs_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                               const slate_amd_options* opts);
int64_t slate_amd_dposv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                              const slate_amd_options* opts);
/* slate::gesv_mixed / gesv_mixed_gmres / posv_mixed / posv_mixed_gmres on
 * tiles (slate.hh:521-553, :667-701). */
int64_t slate_amd_dsgesv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                               const slate_amd_tiles* B, const slate_amd_tiles* X, int* iter,
                               const slate_amd_options* opts);
int64_t slate_amd_dsgesv_gmres_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                                     const slate_amd_tiles* B, const slate_amd_tiles* X, int* iter,
                                     const slate_amd_options* opts);
int64_t slate_amd_dsposv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                               const slate_amd_tiles* X, int* iter, const slate_amd_options* opts);
int64_t slate_amd_dsposv_gmres_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                                     const slate_amd_tiles* X, int* iter, const slate_amd_options* opts);

#ifdef __cplusplus
}
#endif

#endif /* SLATE_AMD_H */
