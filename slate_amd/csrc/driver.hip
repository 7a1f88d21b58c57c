// Host drivers of the factor-and-solve path and the C ABI (include/slate_amd.h).
//
// The factorizations run on a row-major view of the matrix (element (i, j) at
// A[i*ld + j]), which keeps the pivot row interchanges coalesced; the caller's
// column-major matrix is flipped in place on entry and flipped back on exit,
// as SLATE does tile by tile for its GPU LU (src/getrf.cc:54-55, :238).
// For a symmetric matrix the row-major view of the stored lower triangle is
// an upper triangle, so potrf computes A = U^T U with U = L^T in place.
#include "common.hh"
#include "kernels.hh"
#include "../../include/slate_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace sa {

//------------------------------------------------------------------------------
// Workspace generation: bumped whenever any workspace is (re)allocated, so a
// captured graph (which holds workspace addresses) knows it went stale.
static uint64_t g_buf_generation = 1;
// LU graph replay (Options use_graph) is switched off: the first capture of
// an FP64 LU sequence (n = 2048, nb = 256, streams s0/s1 only) ended the
// process with a host segmentation fault inside run_graph, and no cause in
// this code has been shown.  Until one is, read_opts refuses use_graph
// before any work is issued.
static constexpr bool g_graph_enabled = false;

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) SA_HIP(hipFree(p));
        p = nullptr; bytes = 0;
        ++g_buf_generation;
        SA_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    void release() { if (p) { (void)hipFree(p); ++g_buf_generation; } p = nullptr; bytes = 0; }
    template <class T> T* as() { return static_cast<T*>(p); }
};

// Identity of one captured factorization: the kernels' arguments are fixed
// by these (matrix and pivot addresses, sizes, options, workspace
// generation).
struct GraphKey {
    int kind = 0;                // 0 LU float, 1 LU double
    const void* A = nullptr;
    const void* ipiv = nullptr;
    int64_t n = 0, ld = 0, nb = 0, ib = 0, lookahead = 0;
    double threshold = 0;
    int prec = 0, method = 0;
    uint64_t generation = 0;
    bool operator==(const GraphKey& o) const {
        return kind == o.kind && A == o.A && ipiv == o.ipiv && n == o.n && ld == o.ld && nb == o.nb && ib == o.ib
               && lookahead == o.lookahead && threshold == o.threshold && prec == o.prec && method == o.method
               && generation == o.generation;
    }
};

struct Context {
    int device = 0;
    hipStream_t s0 = nullptr;          // trailing updates, solves
    hipStream_t s1 = nullptr;          // panel + lookahead column
    hipStream_t s2 = nullptr;          // trailing updates on a CU subset (when a panel reserve is set)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int64_t panel_switch[2] = {0, 0};  // trailing size below which s2 takes the updates (float, double)
    hipEvent_t t_start = nullptr, t_stop = nullptr;
    std::vector<hipEvent_t> evs;       // step events for the two-stream schedule
    Buf bfl, bfu0, bfu1, bfg;    // split-BF16 operand planes (LU update, gemm entry)
    Buf f8l, f8u0, f8u1, f8s;    // FP8 operand planes (L panel slots, U blocks) and tile scales
    Buf tlA, tlB, tlX, tlP;      // tile-storage entries: dense copies of A, B, X and the tile tables
    Buf npu, npt;                // no-pivoting panel: U11^{-1} and staging
    Buf bmeta, bwork;            // batched solves: per-system metadata, per-workgroup workspace
    Buf bqueue;                  // batched ticket counters (zeroed when allocated; launches leave them zero)
    int b_nbf = 32, b_bt = 256, b_wgs = 1;   // batched tile configuration (slate_amd_tune_set_batched)
    hipGraph_t bgraph = nullptr;             // one-node graph of the last batched launch (use_graph)
    hipGraphExec_t bgraph_exec = nullptr;
    alignas(16) char bgraph_args[BATCH_ARGS_BYTES] = {};
    int bgraph_cfg[3] = {0, 0, 0};
    int64_t bgraph_builds = 0, bgraph_replays = 0;
    Buf slots, err, dinfo, plan0, plan1, plans, ipiv, work, lo, vec, basis, rhs, perm, flags, tinv, ttmp, sinv, stmp1, dinv;
    double last_ms = 0;
    uint32_t leaf_tag = 0;       // next free tag of the panel exchange words
    bool capturing = false;      // a factorization is being recorded into a graph
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;   // captured factorizations
    std::vector<GraphKey> graph_failed;                         // keys whose capture failed
    int64_t graph_replays = 0;
    Buf tvw;                     // trsv look-back exchange words (+ progress word)
    unsigned char* scratch = nullptr;   // this device's no-allocation scratch (noalloc.hip)
    Buf swaptmp;                 // two-pass row interchange workspace (trailing-update stream)
    Buf swaptmp1;                // the same for the lookahead block (tall blocks only)
    Buf ttri;                    // row-major copy of a column-major-view triangle (Cholesky solves)
    Buf cut0, cut1;              // Cholesky: K-contiguous U12^T for the trailing (s0) and block-row (s1) updates
    uint32_t tv_tag = 0;         // next free tag of tvw
    bool calu = false;           // leaves use tournament pivoting (MethodLU::CALU)
    // tuned block sizes per kind (slate_amd_tune_set_nb): (largest n, nb), ascending
    std::vector<std::pair<int64_t, int64_t>> nb_table[4];
    // trailing-update GEMM timing (enabled by slate_amd_profile_enable)
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;
    std::vector<double> prof_flops_pending;
    std::vector<int> prof_kind_pending;
    size_t prof_used = 0;
    double prof_ms = 0, prof_flops = 0;
    int64_t prof_launches = 0;
    // the same split by kernel (SLATE_AMD_PROF_*: other, 256 x 256 planes, 128 x 128 split at staging)
    double prof_ms_k[3] = {0, 0, 0}, prof_flops_k[3] = {0, 0, 0};
    int64_t prof_launches_k[3] = {0, 0, 0};
};

// Event pair around one trailing-update GEMM on stream s (profiling only).
static void prof_begin(Context& c, hipStream_t s)
{
    if (!c.prof) return;
    while (c.prof_ev.size() < c.prof_used + 2) {
        hipEvent_t e;
        SA_HIP(hipEventCreate(&e));
        c.prof_ev.push_back(e);
    }
    SA_HIP(hipEventRecord(c.prof_ev[c.prof_used], s));
}
static void prof_end(Context& c, hipStream_t s, double flops, int kind)
{
    if (!c.prof) return;
    SA_HIP(hipEventRecord(c.prof_ev[c.prof_used + 1], s));
    c.prof_flops_pending.push_back(flops);
    c.prof_kind_pending.push_back(kind);
    c.prof_used += 2;
}
// Fold the recorded pairs into the totals (after the streams are synchronised).
static void prof_collect(Context& c)
{
    if (!c.prof) return;
    for (size_t i = 0; i + 1 < c.prof_used; i += 2) {
        float ms = 0;
        SA_HIP(hipEventElapsedTime(&ms, c.prof_ev[i], c.prof_ev[i + 1]));
        c.prof_ms += ms;
        c.prof_flops += c.prof_flops_pending[i / 2];
        c.prof_launches += 1;
        const int k = c.prof_kind_pending[i / 2];
        c.prof_ms_k[k] += ms;
        c.prof_flops_k[k] += c.prof_flops_pending[i / 2];
        c.prof_launches_k[k] += 1;
    }
    c.prof_used = 0;
    c.prof_flops_pending.clear();
    c.prof_kind_pending.clear();
}

struct Opts {
    int64_t nb = 0, ib = 0, lookahead = 1, itermax = 30;
    double threshold = 1.0, tol = 0;
    bool fallback = true;
    int prec = SLATE_AMD_PREC_FP32;
    int method = SLATE_AMD_LU_PARTIALPIV;
    int depth = 2;
    bool graph = false;
};

// Number of bf16 terms per operand of the float LU's trailing update
// (0 = FP32 MFMA).
static int split_terms(int prec)
{
    switch (prec) {
        case SLATE_AMD_PREC_BF16: return 1;
        case SLATE_AMD_PREC_BF16X2: return 2;
        case SLATE_AMD_PREC_BF16X3: return 3;
        default: return 0;
    }
}

// Number of e4m3 planes per operand of the float factor's trailing update
// (0 = not an FP8 form).
static int fp8_terms(int prec)
{
    switch (prec) {
        case SLATE_AMD_PREC_FP8: return 1;
        case SLATE_AMD_PREC_FP8X2: return 2;
        default: return 0;
    }
}

static Opts read_opts(const slate_amd_options* o)
{
    Opts r;
    if (!o) return r;
    r.nb = o->block_size;
    r.ib = o->inner_blocking;
    r.lookahead = o->lookahead;
    r.threshold = o->pivot_threshold;
    r.itermax = o->max_iterations;
    r.tol = o->tolerance;
    r.fallback = o->use_fallback != 0;
    r.prec = o->factor_precision;
    r.method = o->method_lu;
    r.depth = o->depth;
    r.graph = o->use_graph != 0;
    SA_REQUIRE(!r.graph || g_graph_enabled, -2, "use_graph: LU graph replay is disabled (its capture crashed)");
    SA_REQUIRE(r.method == SLATE_AMD_LU_PARTIALPIV || r.method == SLATE_AMD_LU_NOPIV
               || r.method == SLATE_AMD_LU_CALU, -2, "unknown method_lu");
    SA_REQUIRE(r.prec >= SLATE_AMD_PREC_FP32 && r.prec <= SLATE_AMD_PREC_FP8X2, -2, "unknown factor_precision");
    return r;
}

// kind: SLATE_AMD_TUNE_* (LU FP64, LU float, LU split-BF16, Cholesky).
static int64_t choose_nb(const Context& c, int kind, int64_t n, int64_t requested)
{
    if (requested > 0) return requested;
    for (const auto& e : c.nb_table[kind])
        if (n <= e.first) return e.second;
    // Measured at n = 65536 (float LU with BF16X3 updates / FP64 LU):
    // nb 512 1736 / 4400 ms, 768 1604 / 4244, 896 1580 / 4195, 1024 1569 /
    // 4146 ms (larger K per trailing update, fewer steps).
    if (n <= 2048) return 128;
    if (n <= 16384) return 256;
    if (n <= 32768) return 512;
    return 1024;
}

// Tags of the panel exchange words advance monotonically over the life of the
// context; the slot area is cleared when it is first allocated and whenever
// the 32-bit tag space would wrap, so a stale word never carries a live tag.
static void prepare_panel_work(Context& c, int64_t n, int64_t W)
{
    (void)W;
    const size_t bytes = std::max<size_t>(size_t(1) << 22, size_t(calu_slot_bytes(8, n)) + 4096);
    const bool fresh = c.slots.bytes < bytes;
    if (fresh) c.slots.release();
    c.slots.ensure(bytes);
    // a captured factorization clears the words and restarts the tags itself,
    // so every replay finds no stale tag
    if (c.capturing || fresh || c.leaf_tag > 0xF0000000u - uint32_t(std::min<int64_t>(n, 1 << 26))) {
        zero_u32(static_cast<unsigned*>(c.slots.p), int64_t(c.slots.bytes / 4), c.s0);
        c.leaf_tag = 1;
    }
    c.err.ensure(sizeof(int) * 4);
    zero_u32(c.err.as<unsigned>(), 4, c.s0);
}

static PanelWork panel_work(Context& c)
{
    return PanelWork{c.slots.p, int64_t(c.slots.bytes), c.err.as<int>()};
}

static void check_kernel_error(Context& c)
{
    int e = 0;
    SA_HIP(hipMemcpy(&e, c.err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (e != 0) SA_HIP(hipMemset(c.err.p, 0, sizeof(int) * 4));
    SA_REQUIRE(e != 1, -101, "panel kernel: cross-workgroup barrier timed out");
    SA_REQUIRE(e != 2, -102, "triangular solve: completion chain timed out");
}

//------------------------------------------------------------------------------
// Recursive triangular solve with an nb x nb triangle and an nb x N row-major
// right-hand side; off-diagonal blocks go through the MFMA GEMM.
// tri(i,k) = M[i*ld + k] (cm = 0) or M[i + k*ld] (cm = 1).
template <class T>
static void trsm_rec(const T* M, int64_t ld, int cm, int lower, int unit, int64_t nb,
                     T* B, int64_t ldb, int64_t N, hipStream_t s)
{
    if (nb <= 0 || N <= 0) return;
    constexpr int64_t L = TRSM_LEAF;
    if (nb <= L) { trsm_leaf<T>(M, ld, cm, lower, unit, nb, B, ldb, N, s); return; }
    // the lower row-major solves (the panel's, L11^{-1}, getrs' L sweep, the
    // Cholesky blocks after their transposed copy) run as one launch up to
    // trsm_fused_max rows
    if (cm == 0 && lower && nb <= trsm_fused_max(int(sizeof(T)))) {
        trsm_lower_fused<T>(M, ld, nb, B, ldb, N, unit, s);
        return;
    }
    int64_t n1 = ((nb / 2 + L - 1) / L) * L;
    if (n1 >= nb) n1 = nb - L;
    int64_t n2 = nb - n1;
    const T* M11 = M;
    const T* M22 = M + n1 * ld + n1;
    T* B1 = B;
    T* B2 = B + n1 * ldb;
    GemmArgs g{};
    g.alpha = -1.0; g.beta = 1.0; g.uplo = 0;
    if (lower) {
        trsm_rec<T>(M11, ld, cm, 1, unit, n1, B1, ldb, N, s);
        // B2 -= tri(n1:, 0:n1) * B1
        g.M = n2; g.N = N; g.K = n1;
        g.A = cm ? static_cast<const void*>(M + n1) : static_cast<const void*>(M + n1 * ld);
        g.lda = ld; g.B = B1; g.ldb = ldb; g.C = B2; g.ldc = ldb;
        gemm<T>(g, cm ? A_MC : A_KC, B_NC, s);
        trsm_rec<T>(M22, ld, cm, 1, unit, n2, B2, ldb, N, s);
    }
    else {
        trsm_rec<T>(M22, ld, cm, 0, unit, n2, B2, ldb, N, s);
        // B1 -= tri(0:n1, n1:) * B2
        g.M = n1; g.N = N; g.K = n2;
        g.A = cm ? static_cast<const void*>(M + n1 * ld) : static_cast<const void*>(M + n1);
        g.lda = ld; g.B = B2; g.ldb = ldb; g.C = B1; g.ldc = ldb;
        gemm<T>(g, cm ? A_MC : A_KC, B_NC, s);
        trsm_rec<T>(M11, ld, cm, 0, unit, n1, B1, ldb, N, s);
    }
}

// Lower triangle of a column-major view (cm = 1: tri(i, k) = M[i + k*ld], the
// transposed row-major upper factor of the Cholesky): copied row-major into
// c.ttri (nb^2 elements) so that the one-launch lower solve applies; false
// when the triangle is outside that solve's range.
template <class T>
static bool trsm_lower_cm_fused(Context& c, const T* M, int64_t ld, int unit, int64_t nb, T* B, int64_t ldb,
                                int64_t N, hipStream_t s)
{
    if (nb <= TRSM_LEAF || nb > trsm_fused_max(int(sizeof(T))) || N <= 0) return false;
    c.ttri.ensure(sizeof(T) * size_t(nb) * size_t(nb));
    T* Lr = c.ttri.as<T>();
    transpose_convert<T, T>(M, ld, Lr, nb, nb, nb, s);               // Lr[i*nb + k] = M[i + k*ld]
    trsm_lower_fused<T>(Lr, nb, nb, B, ldb, N, unit, s);
    return true;
}

// Wide triangular solve B := tri^{-1} B for an nb x nb triangle and an
// nb x N row-major B with N >> nb: the triangle is inverted once (the same
// recursive solve applied to an identity, nb x nb) and the product with B is
// one full-width MFMA GEMM into a workspace that is copied back.  This is the
// inverted-diagonal-block form of trsm that GPU BLAS libraries use; it trades
// nb^2 N extra flops at MFMA rate for 2 nb/32 narrow launches.  Runs on s0
// only (one workspace).
template <class T>
static void trsm_wide(Context& c, const T* M, int64_t ld, int cm, int lower, int unit, int64_t nb,
                      T* B, int64_t ldb, int64_t N, hipStream_t s)
{
    if (N < 4 * nb || nb <= TRSM_LEAF) {
        if (!(cm == 1 && lower && trsm_lower_cm_fused<T>(c, M, ld, unit, nb, B, ldb, N, s)))
            trsm_rec<T>(M, ld, cm, lower, unit, nb, B, ldb, N, s);
        return;
    }
    c.tinv.ensure(sizeof(T) * size_t(nb) * size_t(nb));
    c.ttmp.ensure(sizeof(T) * size_t(nb) * size_t(N));
    T* inv = c.tinv.as<T>();
    T* tmp = c.ttmp.as<T>();
    set_identity<T>(inv, nb, nb, s);
    if (!(cm == 1 && lower && trsm_lower_cm_fused<T>(c, M, ld, unit, nb, inv, nb, nb, s)))
        trsm_rec<T>(M, ld, cm, lower, unit, nb, inv, nb, nb, s);
    GemmArgs g{};
    g.M = nb; g.N = N; g.K = nb;
    g.A = inv; g.lda = nb; g.B = B; g.ldb = ldb; g.C = tmp; g.ldc = N;
    g.alpha = 1.0; g.beta = 0.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);
    SA_HIP(hipMemcpy2DAsync(B, sizeof(T) * ldb, tmp, sizeof(T) * N, sizeof(T) * N, nb,
                            hipMemcpyDeviceToDevice, s));
}

// B := inv * B for an nb x nb (full, row-major, ld nb) inverse and an nb x N
// row-major B, through the workspace tmp (nb x N).
template <class T>
static void apply_inv(const T* inv, int64_t nb, T* B, int64_t ldb, int64_t N, T* tmp, hipStream_t s)
{
    if (N <= 0) return;
    GemmArgs g{};
    g.M = nb; g.N = N; g.K = nb;
    g.A = inv; g.lda = nb; g.B = B; g.ldb = ldb; g.C = tmp; g.ldc = N;
    g.alpha = 1.0; g.beta = 0.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);
    SA_HIP(hipMemcpy2DAsync(B, sizeof(T) * ldb, tmp, sizeof(T) * N, sizeof(T) * N, nb,
                            hipMemcpyDeviceToDevice, s));
}

// The same product followed by one refinement step against the unit-lower
// triangle itself (lfull: its explicit nb x nb copy): X = inv B,
// R = B - L X, X += inv R.  The explicit inverse alone leaves a residual
// B - L X of order kappa(L) eps |B|, which at nb = 1024 measured ~10x the
// backward error of a triangular solve in the float LU; after the step the
// residual is at the solve's level (eps |L| |X|), at two more GEMMs of the
// same shape.
template <class T>
static void apply_inv_refined(const T* inv, const T* lfull, int64_t nb, T* B, int64_t ldb, int64_t N, T* tmp,
                              hipStream_t s)
{
    if (N <= 0) return;
    GemmArgs g{};
    g.M = nb; g.N = N; g.K = nb;
    g.A = inv; g.lda = nb; g.B = B; g.ldb = ldb; g.C = tmp; g.ldc = N;
    g.alpha = 1.0; g.beta = 0.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);                                          // X = inv B
    g.A = lfull; g.B = tmp; g.ldb = N; g.C = B; g.ldc = ldb;
    g.alpha = -1.0; g.beta = 1.0;
    gemm<T>(g, A_KC, B_NC, s);                                          // B := B - L X (R)
    g.A = inv; g.B = B; g.ldb = ldb; g.C = tmp; g.ldc = N;
    g.alpha = 1.0; g.beta = 1.0;
    gemm<T>(g, A_KC, B_NC, s);                                          // X += inv R
    SA_HIP(hipMemcpy2DAsync(B, sizeof(T) * ldb, tmp, sizeof(T) * N, sizeof(T) * N, nb,
                            hipMemcpyDeviceToDevice, s));
}

// inv := L^{-1} for the unit-lower nb x nb diagonal block of a row-major matrix.
// From nb = 256 on, in two half-height sweeps instead of one full one: X11 =
// L11^{-1} (top-left block), then the bottom rows [X21 X22] solve
// L22 [X21 X22] = [-L21 X11  I] with the same substitution kernel.  Every
// block comes from substitution, as in the recursive solve (multiplying by
// an explicit X22 instead loses accuracy when L22 is ill-conditioned).
template <class T>
static void unit_lower_inverse(const T* L, int64_t ld, int64_t nb, T* inv, hipStream_t s)
{
    set_identity<T>(inv, nb, nb, s);
    const int64_t h = ceildiv(nb / 2, 32) * 32, h2 = nb - h;
    if (nb < 256 || h > trsm_fused_max(int(sizeof(T)))) {
        trsm_rec<T>(L, ld, 0, 1, 1, nb, inv, nb, nb, s);
        return;
    }
    trsm_lunit_fused<T>(L, ld, h, inv, nb, h, s);                       // X11
    GemmArgs g{};
    g.M = h2; g.N = h; g.K = h;
    g.A = L + h * ld; g.lda = ld; g.B = inv; g.ldb = nb; g.C = inv + h * nb; g.ldc = nb;
    g.alpha = -1.0; g.beta = 0.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);                                          // -L21 X11
    trsm_lunit_fused<T>(L + h * ld + h, ld, h2, inv + h * nb, nb, nb, s);  // [X21 X22]
}

//------------------------------------------------------------------------------
// Recursive panel LU of rows [c0, n) x columns [c0, c0+w) of a row-major n x n
// matrix (the panel's top row is the diagonal), inside the enclosing panel
// columns [p0, p0 + pw).  Leaves of width <= W use the cross-workgroup
// kernel, which also applies its interchanges to the rest of the enclosing
// panel; the splits solve with the left half's unit-lower block, update with
// the GEMM and recurse.  (The CALU leaf only swaps its own columns: there the
// splits apply the left half's interchanges to the right half first and bring
// the right half's back to the left half at the end.)
template <class T>
static void panel_rec(Context& c, T* A, int64_t ld, int64_t n, int64_t c0, int64_t w,
                      int64_t* ipiv, int64_t* info, double thr, int64_t leaf_w,
                      int64_t* plan, hipStream_t s, int64_t p0, int64_t pw)
{
    if (w <= leaf_w) {
        if (c.calu) {
            calu_leaf<T>(A, ld, c0, n, c0, int(w), ipiv, info, panel_work(c), c.leaf_tag, s);
            c.leaf_tag += 16;
        }
        else {
            panel_leaf<T>(A, ld, c0, n, c0, int(w), ipiv, info, thr, panel_work(c), c.leaf_tag, s, p0, pw);
            c.leaf_tag += uint32_t(w);
        }
        return;
    }
    int64_t w1 = ((w / 2 + leaf_w - 1) / leaf_w) * leaf_w;
    if (w1 >= w) w1 = w - leaf_w;
    int64_t w2 = w - w1;
    panel_rec<T>(c, A, ld, n, c0, w1, ipiv, info, thr, leaf_w, plan, s, p0, pw);
    if (c.calu) laswp<T>(A, ld, n, c0 + w1, c0 + w, c0, c0 + w1, ipiv, plan, s);
    trsm_rec<T>(A + c0 * ld + c0, ld, 0, 1, 1, w1, A + c0 * ld + c0 + w1, ld, w2, s);
    GemmArgs g{};
    g.M = n - c0 - w1; g.N = w2; g.K = w1;
    g.A = A + (c0 + w1) * ld + c0; g.lda = ld;
    g.B = A + c0 * ld + c0 + w1; g.ldb = ld;
    g.C = A + (c0 + w1) * ld + c0 + w1; g.ldc = ld;
    g.alpha = -1.0; g.beta = 1.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);
    panel_rec<T>(c, A, ld, n, c0 + w1, w2, ipiv, info, thr, leaf_w, plan, s, p0, pw);
    if (c.calu) laswp<T>(A, ld, n, c0, c0 + w1, c0 + w1, c0 + w, ipiv, plan, s);
}

// Panel of the no-pivoting LU (src/getrf_nopiv.cc): the kb x kb diagonal
// block is factored without interchanges (the leaf kernel with threshold 0
// keeps every diagonal pivot), then L21 = A21 U11^{-1} with the inverse of the
// upper triangle applied by one GEMM.  No cross-workgroup pivot search over
// the tall panel.  ipiv is left as the identity.
template <class T>
static void panel_nopiv(Context& c, T* A, int64_t ld, int64_t n, int64_t k0, int64_t kb, int64_t* ipiv,
                        int64_t* info, int64_t leaf_w, int64_t* plan, T* uinv, T* tmp, hipStream_t s)
{
    panel_rec<T>(c, A, ld, k0 + kb, k0, kb, ipiv, info, 0.0, leaf_w, plan, s, k0, kb);
    const int64_t m = n - k0 - kb;
    if (m <= 0) return;
    T* Akk = A + k0 * ld + k0;
    set_identity<T>(uinv, kb, kb, s);
    trsm_rec<T>(Akk, ld, 0, 0, 0, kb, uinv, kb, kb, s);          // U11^{-1}
    GemmArgs g{};
    g.M = m; g.N = kb; g.K = kb;
    g.A = Akk + kb * ld; g.lda = ld; g.B = uinv; g.ldb = kb; g.C = tmp; g.ldc = kb;
    g.alpha = 1.0; g.beta = 0.0; g.uplo = 0;
    gemm<T>(g, A_KC, B_NC, s);
    SA_HIP(hipMemcpy2DAsync(Akk + kb * ld, sizeof(T) * ld, tmp, sizeof(T) * kb, sizeof(T) * kb, m,
                            hipMemcpyDeviceToDevice, s));
}

// Right-looking blocked LU of a row-major n x n matrix with lookahead 1 on a
// second stream: the panel of step k+1 is factored on s1 while s0 applies
// step k to the columns beyond k+1 (src/getrf.cc:85-235, with the OpenMP
// task DAG turned into two streams and events).
// ipiv: device, global 0-based rows; info: device int64 (zeroed by caller).
template <class T>
static void getrf_rm(Context& c, T* A, int64_t n, int64_t ld, const Opts& o, int64_t* ipiv,
                     int64_t* info)
{
    const int kind = sizeof(T) == 8 ? SLATE_AMD_TUNE_LU_FP64
                   : (split_terms(o.prec) > 0 || fp8_terms(o.prec) > 0 ? SLATE_AMD_TUNE_LU_BF16
                                                                      : SLATE_AMD_TUNE_LU_FP32);
    const int64_t nb = choose_nb(c, kind, n, o.nb);
    const int f8 = std::is_same<T, float>::value ? fp8_terms(o.prec) : 0;
    SA_REQUIRE(f8 == 0 || nb % 128 == 0, -2, "FP8 factor precision: block size must be a multiple of 128");
    c.calu = o.method == SLATE_AMD_LU_CALU;
    int64_t leaf_w = c.calu ? calu_leaf_width(int(sizeof(T))) : panel_leaf_width(int(sizeof(T)));
    if (o.ib > 0 && o.ib < leaf_w) leaf_w = o.ib;
    prepare_panel_work(c, n, leaf_w);
    c.plan1.ensure(sizeof(int64_t) * (4 * 2048 + 8));
    int64_t* plan1 = c.plan1.as<int64_t>();
    const int64_t nsteps = ceildiv(n, nb);
    const bool la = o.lookahead > 0 && nsteps > 2;
    if (int64_t(c.evs.size()) < 2 * nsteps + 4) {
        for (int64_t i = int64_t(c.evs.size()); i < 2 * nsteps + 4; ++i) {
            hipEvent_t e;
            SA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c.evs.push_back(e);
        }
    }
    // With a panel reserve the trailing updates run on s2, whose CU mask
    // leaves the reserved CUs to the panel stream: a trailing GEMM of
    // thousands of workgroups otherwise holds every CU slot and the panel's
    // small launches (and the leaf, whose workgroups must be co-resident)
    // wait for slots at every step.
    // The switch happens once, when the trailing size drops below
    // c.panel_switch (the steps where the panel, not the update, is the
    // critical path).
    hipStream_t s0 = c.s0, s1 = la ? c.s1 : c.s0;
    // s1 must see the counter reset done on s0
    if (la) {
        SA_HIP(hipEventRecord(c.evs[0], s0));
        SA_HIP(hipStreamWaitEvent(s1, c.evs[0], 0));
    }
    // ev_panel[k]: panel k factored and its column block complete (on s1)
    // ev_trail[k]: step k-1 applied to every column block beyond k (on s0)
    auto ev_panel = [&](int64_t k) { return c.evs[2 + 2 * k]; };
    auto ev_trail = [&](int64_t k) { return c.evs[3 + 2 * k]; };

    // One swap plan per step, built on s1 right after the step's panel and
    // read by both streams (plan slots are never reused within a call).
    // L11^{-1} of each step, built on s1 after the panel (two alternating
    // slots: s0 is done with slot k%2 before s1 starts step k+2), applied by
    // GEMM to the lookahead block (s1) and to the rest (s0), so the result
    // is independent of where the columns are split.
    // The float-accurate factor precisions (FP32, BF16X3) refine each
    // application once against an explicit unit-lower copy of L11 (same
    // slots), so U12 carries a triangular solve's error, not the explicit
    // inverse's.
    // SLATE_AMD_REFINE_INV=1 / 0 turns it on / off for every float factor.
    static const int refine_env = [] {
        const char* e = std::getenv("SLATE_AMD_REFINE_INV");
        return e ? std::atoi(e) : -1;
    }();
    const bool refine = std::is_same<T, float>::value
                        && (refine_env >= 0 ? refine_env != 0
                                            : (o.prec == SLATE_AMD_PREC_FP32 || o.prec == SLATE_AMD_PREC_BF16X3));
    c.sinv.ensure(sizeof(T) * size_t((refine ? 4 : 2) * nb * nb));
    c.stmp1.ensure(sizeof(T) * size_t(nb) * size_t(nb));
    c.ttmp.ensure(sizeof(T) * size_t(nb) * size_t(std::max<int64_t>(n, 1)));
    auto step_inv = [&](int64_t k) { return c.sinv.as<T>() + (k % 2) * nb * nb; };
    auto step_lfull = [&](int64_t k) { return c.sinv.as<T>() + (2 + k % 2) * nb * nb; };
    T* tmp0 = c.ttmp.as<T>();
    T* tmp1 = c.stmp1.as<T>();
    // L11^{-1} (and with refine its unit-lower copy) of the panel at [p0, p0 + pw)
    auto make_inv = [&](int64_t k, int64_t p0, int64_t pw) {
        unit_lower_inverse<T>(A + p0 * ld + p0, ld, pw, step_inv(k), s1);
        if (refine) unit_lower_copy<T>(A + p0 * ld + p0, ld, pw, step_lfull(k), s1);
    };
    // U12 := L11^{-1} U12 for the nb_k x N block at B (step k)
    auto use_inv = [&](int64_t k, int64_t kb_, T* B, int64_t N, T* tmp, hipStream_t st) {
        if (refine) apply_inv_refined<T>(step_inv(k), step_lfull(k), kb_, B, ld, N, tmp, st);
        else apply_inv<T>(step_inv(k), kb_, B, ld, N, tmp, st);
    };

    // Float factor with split-BF16 updates (factor_precision BF16, BF16X2,
    // BF16X3).  With 16-byte aligned rows the update reads the float L panel
    // directly and a transposed float copy of the U row block (k contiguous),
    // splitting both into bf16 terms inside the GEMM (gemm_s32).  Otherwise
    // the L panel of each step is split once on s1 into one of two
    // alternating bf16-plane slots (s0 is done with slot k%2 before s1
    // refills it at step k+2), and each U row block right before its update.
    const int ns = std::is_same<T, float>::value ? split_terms(o.prec) : 0;
    const int64_t Kp = ceildiv(nb, 32) * 32;
    const bool direct = ns > 0 && ld % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0;
    const int64_t lplane = n * Kp, uplane0 = n * Kp, uplane1 = nb * Kp;
    if (ns > 0) {
        const size_t ub = direct ? sizeof(float) : sizeof(__bf16) * size_t(ns);
        if (!direct) c.bfl.ensure(sizeof(__bf16) * size_t(2 * ns) * size_t(lplane));
        c.bfu0.ensure(ub * size_t(uplane0));
        c.bfu1.ensure(ub * size_t(uplane1));
    }
    // One- and two-term splits: trailing updates of at least LT_MIN 256 x 256
    // tiles read pre-split planes instead (gemm_bf16_lt: direct global->LDS
    // staging, no split work in the GEMM); the step's L panel is split into
    // its slot on s1 and the U row block on s0 before the update.  Both
    // kernels accumulate each element identically, so results do not depend
    // on which one a step uses.
    constexpr int64_t LT_MIN = 512;
    auto use_lt = [&](int64_t k) -> bool {
        if (!direct || (ns != 1 && ns != 2)) return false;
        const int64_t k0 = k * nb, kb = std::min(nb, n - k0), r = n - k0 - kb;
        const int64_t nb1 = std::min(nb, r);
        const int64_t ncols = (la && r > nb1) ? r - nb1 : r;
        return r > 0 && ncols > 0 && gemm_lt_tiles(r, ncols) >= LT_MIN;
    };
    if (ns > 0 && direct && use_lt(0)) c.bfl.ensure(sizeof(__bf16) * size_t(2 * ns) * size_t(lplane));
    auto lslot = [&](int64_t k) { return c.bfl.as<__bf16>() + (k % 2) * ns * lplane; };
    auto split_panel = [&](int64_t k, int64_t k0, int64_t kb, hipStream_t st) {
        if constexpr (std::is_same<T, float>::value) {
            if (ns > 0 && (!direct || use_lt(k)))
                split_rows(ns, A + (k0 + kb) * ld + k0, ld, n - k0 - kb, kb, lslot(k), Kp, lplane, st);
        }
    };
    // FP8 updates (factor_precision FP8, FP8X2): the L panel of each step is
    // quantized once on s1 (two alternating slots, as the bf16 planes), each
    // U row block right before its update; scales per 128-row tile.
    const int64_t Kp8 = f8 ? fp8_kp(nb) : 0;
    const int64_t p8l = f8 ? fp8_rows(n) * Kp8 : 0, p8u1 = f8 ? fp8_rows(nb) * Kp8 : 0;
    const int64_t t8 = f8 ? fp8_tiles(n) : 0;
    if (f8) {
        c.f8l.ensure(size_t(2 * f8) * size_t(p8l));
        c.f8u0.ensure(size_t(f8) * size_t(p8l));
        c.f8u1.ensure(size_t(f8) * size_t(p8u1));
        c.f8s.ensure(sizeof(int) * size_t(4 * t8 + 4));
    }
    auto l8slot = [&](int64_t k) { return c.f8l.as<uint8_t>() + (k % 2) * f8 * p8l; };
    auto l8scale = [&](int64_t k) { return c.f8s.as<int>() + (k % 2) * t8; };
    auto quant_panel = [&](int64_t k, int64_t k0, int64_t kb, hipStream_t st) {
        if constexpr (std::is_same<T, float>::value) {
            if (f8 && n - k0 - kb > 0)
                quant_fp8_rows(f8, A + (k0 + kb) * ld + k0, ld, n - k0 - kb, kb, l8slot(k), Kp8, p8l, l8scale(k), st);
        }
    };
    // Update of columns [cb, cb + ncols) by step k (rows below the panel):
    // C -= L21 U12, U12 already multiplied by L11^{-1}.
    // role 1 (the trailing update on s0) is the one timed by the profiling
    // events, around the GEMM launch alone.
    auto update = [&](int64_t k, int64_t k0, int64_t kb, int64_t cb, int64_t ncols, bool lookahead_block,
                      hipStream_t st, int role) {
        const int64_t c1 = k0 + kb;
        if (ncols <= 0 || n - c1 <= 0) return;
        const double flops = 2.0 * double(n - c1) * double(ncols) * double(kb);
        auto timed = [&](int kind, auto&& launch) {
            if (role == 1) prof_begin(c, st);
            launch();
            if (role == 1) prof_end(c, st, flops, kind);
        };
        if constexpr (std::is_same<T, float>::value) {
            if (f8) {
                uint8_t* u8 = lookahead_block ? c.f8u1.as<uint8_t>() : c.f8u0.as<uint8_t>();
                const int64_t pu = lookahead_block ? p8u1 : p8l;
                int* su = c.f8s.as<int>() + (lookahead_block ? 3 : 2) * t8;
                quant_fp8_cols_t(f8, A + k0 * ld + cb, ld, kb, ncols, u8, Kp8, pu, su, st);
                timed(SLATE_AMD_PROF_OTHER, [&] { gemm_fp8(f8, n - c1, ncols, Kp8, l8slot(k), p8l, l8scale(k), u8, pu, su,
                                     A + c1 * ld + cb, ld, -1.0f, st, role, 0); });
                return;
            }
            if (ns > 0 && direct && !lookahead_block && use_lt(k)) {
                __bf16* uh = c.bfu0.as<__bf16>();
                split_cols_t(ns, A + k0 * ld + cb, ld, kb, ncols, uh, Kp, uplane0, st);
                timed(SLATE_AMD_PROF_LT, [&] { gemm_bf16_lt(ns, n - c1, ncols, Kp, lslot(k), Kp, lplane, uh, Kp, uplane0,
                                         A + c1 * ld + cb, ld, -1.0f, st, role, 0); });
                return;
            }
            if (ns > 0 && direct) {
                float* ut = lookahead_block ? c.bfu1.as<float>() : c.bfu0.as<float>();
                transpose_f32(A + k0 * ld + cb, ld, kb, ncols, ut, Kp, st);
                timed(SLATE_AMD_PROF_SP, [&] { gemm_s32(ns, n - c1, ncols, kb, A + c1 * ld + k0, ld, ut, Kp, A + c1 * ld + cb, ld,
                                     -1.0f, st, role, 0); });
                return;
            }
            if (ns > 0) {
                __bf16* uh = lookahead_block ? c.bfu1.as<__bf16>() : c.bfu0.as<__bf16>();
                const int64_t up = lookahead_block ? uplane1 : uplane0;
                split_cols_t(ns, A + k0 * ld + cb, ld, kb, ncols, uh, Kp, up, st);
                timed(SLATE_AMD_PROF_OTHER, [&] { gemm_bf16s(ns, n - c1, ncols, Kp, lslot(k), Kp, lplane, uh, Kp, up,
                                       A + c1 * ld + cb, ld, -1.0f, st, role); });
                return;
            }
        }
        GemmArgs g{};
        g.M = n - c1; g.N = ncols; g.K = kb;
        g.A = A + c1 * ld + k0; g.lda = ld;
        g.B = A + k0 * ld + cb; g.ldb = ld;
        g.C = A + c1 * ld + cb; g.ldc = ld;
        g.alpha = -1.0; g.beta = 1.0; g.uplo = 0;
        timed(SLATE_AMD_PROF_OTHER, [&] { gemm<T>(g, A_KC, B_NC, st, role); });
    };
    const int64_t plan_stride = 4 * nb + 8;
    c.plans.ensure(sizeof(int64_t) * size_t(plan_stride) * size_t(nsteps + 1));
    auto step_plan = [&](int64_t k) { return c.plans.as<int64_t>() + k * plan_stride; };
    // MethodLU::NoPiv (src/getrf_nopiv.cc): panels without interchanges and
    // no row swaps; the pivot vector stays the identity.
    const bool nopiv = o.method == SLATE_AMD_LU_NOPIV;
    T* npu = nullptr;
    T* npt = nullptr;
    if (nopiv) {
        c.npu.ensure(sizeof(T) * size_t(nb) * size_t(nb));
        c.npt.ensure(sizeof(T) * size_t(nb) * size_t(std::max<int64_t>(n, 1)));
        npu = c.npu.as<T>();
        npt = c.npt.as<T>();
    }
    auto factor_panel = [&](int64_t k, int64_t p0, int64_t pw) {       // on s1
        if (nopiv) {
            panel_nopiv<T>(c, A, ld, n, p0, pw, ipiv, info, leaf_w, plan1, npu, npt, s1);
        }
        else {
            // FP64 panels taller than 32768 rows keep 16-column leaves: the
            // 32-column register form needs 2 rows per thread there and drops
            // to one workgroup per CU (measured 3564 -> 3980 ms at n = 65536),
            // while below it the wider leaves save launches (643 -> 625 ms at
            // n = 32768)
            const int64_t lw = (sizeof(T) == 8 && !c.calu && n - p0 > 32768 && leaf_w > 16 && o.ib <= 0) ? 16 : leaf_w;
            panel_rec<T>(c, A, ld, n, p0, pw, ipiv, info, o.threshold, lw, plan1, s1, p0, pw);
            laswp_plan(p0, p0 + pw, n, ipiv, step_plan(k), s1);
        }
    };
    // Wide interchanges on the trailing-update stream go through a two-pass
    // gather/scatter workspace; the lookahead block's, on s1, stay one-pass
    // unless the block is too tall for the one-pass LDS staging (nb > 1024
    // in FP64 / 2048 in float), when they use a workspace of their own.
    const int64_t swap_elems = 2 * nb * n, swap1_elems = 2 * nb * nb;
    const bool s1_two_pass = int64_t(4) * 2 * nb * int64_t(sizeof(T)) > 65536;
    if (!nopiv) {
        c.swaptmp.ensure(sizeof(T) * size_t(swap_elems));
        if (s1_two_pass) c.swaptmp1.ensure(sizeof(T) * size_t(swap1_elems));
    }
    auto swap = [&](int64_t a0, int64_t a1, int64_t p0, int64_t p1, int64_t* pk, hipStream_t st) {
        if (nopiv) return;
        if (st != s1 || !la)
            laswp<T>(A, ld, n, a0, a1, p0, p1, ipiv, pk, st, true, c.swaptmp.as<T>(), swap_elems);
        else if (s1_two_pass)
            laswp<T>(A, ld, n, a0, a1, p0, p1, ipiv, pk, st, true, c.swaptmp1.as<T>(), swap1_elems);
        else
            laswp<T>(A, ld, n, a0, a1, p0, p1, ipiv, pk, st, true);
    };
    // panel 0
    factor_panel(0, 0, std::min(nb, n));
    make_inv(0, 0, std::min(nb, n));
    split_panel(0, 0, std::min(nb, n), s1);
    quant_panel(0, 0, std::min(nb, n), s1);
    SA_HIP(hipEventRecord(ev_panel(0), s1));
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t k0 = k * nb, kb = std::min(nb, n - k0);
        const int64_t c1 = k0 + kb;                   // first column after the panel
        const int64_t r = n - c1;                     // trailing size
        {
            const int pi = sizeof(T) == 8 ? 1 : 0;
            hipStream_t want = c.s0;
            if (la && c.s2 && r < c.panel_switch[pi]) want = c.s2;
            if (want != s0) {
                SA_HIP(hipEventRecord(c.ev_in, s0));
                SA_HIP(hipStreamWaitEvent(want, c.ev_in, 0));
                s0 = want;
            }
        }
        SA_HIP(hipStreamWaitEvent(s0, ev_panel(k), 0));
        // pivots of step k applied to the columns on the left (finished L) on s0
        int64_t* pk = step_plan(k);
        swap(0, k0, k0, c1, pk, s0);
        if (r <= 0) break;
        const int64_t nb1 = std::min(nb, r);          // lookahead column block width
        if (la && r > nb1) {
            // s1: step k on column block k+1, then panel k+1
            SA_HIP(hipStreamWaitEvent(s1, k > 0 ? ev_trail(k) : ev_panel(0), 0));
            swap(c1, c1 + nb1, k0, c1, pk, s1);
            use_inv(k, kb, A + k0 * ld + c1, nb1, tmp1, s1);
            update(k, k0, kb, c1, nb1, true, s1, 0);
            factor_panel(k + 1, c1, nb1);
            make_inv(k + 1, c1, nb1);
            split_panel(k + 1, c1, nb1, s1);
            quant_panel(k + 1, c1, nb1, s1);
            SA_HIP(hipEventRecord(ev_panel(k + 1), s1));
            // s0: step k on the columns beyond k+1
            const int64_t c2 = c1 + nb1;
            swap(c2, n, k0, c1, pk, s0);
            use_inv(k, kb, A + k0 * ld + c2, n - c2, tmp0, s0);
            update(k, k0, kb, c2, n - c2, false, s0, 1);
            SA_HIP(hipEventRecord(ev_trail(k + 1), s0));
        }
        else {
            // no lookahead: whole trailing matrix on s0, then the next panel
            swap(c1, n, k0, c1, pk, s0);
            use_inv(k, kb, A + k0 * ld + c1, r, tmp0, s0);
            update(k, k0, kb, c1, r, false, s0, 1);
            if (la) {
                SA_HIP(hipEventRecord(ev_trail(k + 1), s0));
                SA_HIP(hipStreamWaitEvent(s1, ev_trail(k + 1), 0));
            }
            factor_panel(k + 1, c1, nb1);
            make_inv(k + 1, c1, nb1);
            split_panel(k + 1, c1, nb1, s1);
            quant_panel(k + 1, c1, nb1, s1);
            SA_HIP(hipEventRecord(ev_panel(k + 1), s1));
        }
    }
    if (la) {
        SA_HIP(hipEventRecord(c.evs[1], s1));
        SA_HIP(hipStreamWaitEvent(s0, c.evs[1], 0));
    }
    if (s0 != c.s0) {
        SA_HIP(hipEventRecord(c.ev_out, s0));
        SA_HIP(hipStreamWaitEvent(c.s0, c.ev_out, 0));
    }
    if (c.prof) {
        SA_HIP(hipStreamSynchronize(s0));
        prof_collect(c);
    }
}

// Row-major upper Cholesky A = U^T U (the column-major lower factor).
template <class T>
static void potrf_diag_rec(Context& c, T* M, int64_t ld, int64_t nb, int64_t col0,
                           int64_t* info, hipStream_t s)
{
    if (nb <= 64) { potrf_leaf<T>(M, ld, nb, col0, info, s); return; }
    int64_t n1 = ((nb / 2 + 63) / 64) * 64;
    if (n1 >= nb) n1 = nb - 64;
    int64_t n2 = nb - n1;
    potrf_diag_rec<T>(c, M, ld, n1, col0, info, s);
    if (!trsm_lower_cm_fused<T>(c, M, ld, 0, n1, M + n1, ld, n2, s))
        trsm_rec<T>(M, ld, 1, 1, 0, n1, M + n1, ld, n2, s);
    GemmArgs g{};
    g.M = n2; g.N = n2; g.K = n1;
    g.A = M + n1; g.lda = ld; g.B = M + n1; g.ldb = ld;
    g.C = M + n1 * ld + n1; g.ldc = ld;
    g.alpha = -1.0; g.beta = 1.0; g.uplo = 1;
    gemm<T>(g, A_MC, B_NC, s);
    potrf_diag_rec<T>(c, M + n1 * ld + n1, ld, n2, col0 + n1, info, s);
}

// Right-looking tile Cholesky (src/potrf.cc:84-194): diagonal block, the
// block row solve (the column-major trsm of A(k+1:, k) by L(k,k)^{-H}), and
// the herk/gemm update of the trailing triangle as one triangular GEMM.
template <class T>
static void potrf_rm(Context& c, T* A, int64_t n, int64_t ld, const Opts& o, int64_t* info)
{
    const int64_t nb = choose_nb(c, SLATE_AMD_TUNE_CHOL, n, o.nb);
    // Float factor with split-BF16 updates: U12 (kb x r) is split and
    // transposed once per step into r x Kp planes, which are both operands
    // of the triangular update C -= U12^T U12.
    const int ns = std::is_same<T, float>::value ? split_terms(o.prec) : 0;
    const int64_t Kp = ceildiv(nb, 32) * 32, plane = n * Kp;
    if (ns > 0) c.bfu0.ensure(std::max(sizeof(__bf16) * size_t(ns), sizeof(float)) * size_t(plane));
    // FP8 forms: U12^T quantized once per step (scales per 128-row tile) is
    // both operands of the triangular update.
    const int f8 = std::is_same<T, float>::value ? fp8_terms(o.prec) : 0;
    SA_REQUIRE(f8 == 0 || nb % 128 == 0, -2, "FP8 factor precision: block size must be a multiple of 128");
    const int64_t Kp8 = f8 ? fp8_kp(nb) : 0, p8 = f8 ? fp8_rows(n) * Kp8 : 0;
    if (f8) {
        c.f8u0.ensure(size_t(f8) * size_t(p8));
        c.f8u1.ensure(size_t(f8) * size_t(p8));
        c.f8s.ensure(sizeof(int) * size_t(2 * fp8_tiles(n) + 4));
    }
    if (ns > 0) c.bfu1.ensure(sizeof(float) * size_t(plane));
    const int64_t nsteps = ceildiv(n, nb);
    // Lookahead as in the LU (src/potrf.cc:84-194 with Option::Lookahead): s1
    // applies step k to the next block row and factors it (diagonal block and
    // row-panel solve), while s0 applies step k to the rows beyond it.  The
    // panel's workspaces (tinv, ttmp, ttri) belong to s1; s0 uses cut0.
    const bool la = o.lookahead > 0 && nsteps > 2;
    hipStream_t s0 = c.s0, s1 = la ? c.s1 : c.s0;
    if (int64_t(c.evs.size()) < 2 * nsteps + 4) {
        for (int64_t i = int64_t(c.evs.size()); i < 2 * nsteps + 4; ++i) {
            hipEvent_t e;
            SA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c.evs.push_back(e);
        }
    }
    auto ev_panel = [&](int64_t k) { return c.evs[2 + 2 * k]; };
    auto ev_trail = [&](int64_t k) { return c.evs[3 + 2 * k]; };
    if (la) {
        SA_HIP(hipEventRecord(c.evs[0], s0));
        SA_HIP(hipStreamWaitEvent(s1, c.evs[0], 0));
    }
    // diagonal block k0 (kb) and its row-panel solve U12 := U11^{-T} A12
    auto panel = [&](int64_t k0, int64_t kb, hipStream_t st) {
        T* Akk = A + k0 * ld + k0;
        potrf_diag_rec<T>(c, Akk, ld, kb, k0, info, st);
        const int64_t r = n - k0 - kb;
        if (r > 0) trsm_wide<T>(c, Akk, ld, 1, 1, 0, kb, Akk + kb, ld, r, st);
    };
    // rows [r0, n) x columns [r0 + ..., n) of the trailing upper triangle
    // -= U12(:, r0:)^T U12(:, r0:) for the step whose row panel is at k0
    auto herk = [&](int64_t k0, int64_t kb, int64_t r0, Buf& wk, hipStream_t st) {
        const int64_t r = n - r0;
        if (r <= 0) return;
        const T* U = A + k0 * ld + r0;                 // kb x r, row-major
        T* C = A + r0 * ld + r0;
        if constexpr (std::is_same<T, float>::value) {
            if (f8 && &wk == &c.bfu0) {
                uint8_t* u8 = c.f8u0.as<uint8_t>();
                int* su = c.f8s.as<int>();
                quant_fp8_cols_t(f8, U, ld, kb, r, u8, Kp8, p8, su, st);
                gemm_fp8(f8, r, r, Kp8, u8, p8, su, u8, p8, su, C, ld, -1.0f, st, 0, 1);
                return;
            }
            if (ns > 0 && &wk == &c.bfu0) {
                float* ut = c.bfu0.as<float>();
                transpose_f32(U, ld, kb, r, ut, Kp, st);
                gemm_s32(ns, r, r, kb, ut, Kp, ut, Kp, C, ld, -1.0f, st, 0, 1);
                return;
            }
        }
        // U^T copied K-contiguous so the update runs on the 16-byte-staged kernel
        wk.ensure(sizeof(T) * size_t(r) * size_t(kb));
        T* ut = wk.as<T>();
        transpose_convert<T, T>(U, ld, ut, kb, r, kb, st);               // ut[i*kb + k] = U[k][i]
        GemmArgs g{};
        g.M = r; g.N = r; g.K = kb;
        g.A = ut; g.lda = kb; g.B = U; g.ldb = ld; g.C = C; g.ldc = ld;
        g.alpha = -1.0; g.beta = 1.0; g.uplo = 1;
        gemm<T>(g, A_KC, B_NC, st);
    };
    // block row [c1, c2) -= U12(:, c1:c2)^T U12(:, c1:n), upper part only.
    // The low-precision forms use the same kernel and the same operand tiles
    // (counted from c1) as the herk, so each entry is computed exactly as the
    // update without lookahead computes it; one transposed / quantized copy of
    // U12(:, c1:n) serves as both operands (its first w rows are the left one).
    auto block_row = [&](int64_t k0, int64_t kb, int64_t c1, int64_t c2, hipStream_t st) {
        const int64_t w = c2 - c1, r = n - c1;
        const T* U = A + k0 * ld + c1;
        if constexpr (std::is_same<T, float>::value) {
            if (f8) {
                uint8_t* u8 = c.f8u1.as<uint8_t>();
                int* su = c.f8s.as<int>() + fp8_tiles(n);
                quant_fp8_cols_t(f8, U, ld, kb, r, u8, Kp8, p8, su, st);
                gemm_fp8(f8, w, r, Kp8, u8, p8, su, u8, p8, su, A + c1 * ld + c1, ld, -1.0f, st, 0, 1);
                return;
            }
            if (ns > 0) {
                float* ut = c.bfu1.as<float>();
                transpose_f32(U, ld, kb, r, ut, Kp, st);
                gemm_s32(ns, w, r, kb, ut, Kp, ut, Kp, A + c1 * ld + c1, ld, -1.0f, st, 0, 1);
                return;
            }
        }
        c.cut1.ensure(sizeof(T) * size_t(w) * size_t(kb));
        T* ut = c.cut1.as<T>();
        transpose_convert<T, T>(U, ld, ut, kb, w, kb, st);
        GemmArgs g{};
        g.M = w; g.N = r; g.K = kb;
        g.A = ut; g.lda = kb; g.B = U; g.ldb = ld; g.C = A + c1 * ld + c1; g.ldc = ld;
        g.alpha = -1.0; g.beta = 1.0; g.uplo = 1;
        gemm<T>(g, A_KC, B_NC, st);
    };
    Buf& wk0 = ns > 0 || f8 > 0 ? c.bfu0 : c.cut0;
    if (!la) {
        for (int64_t k0 = 0; k0 < n; k0 += nb) {
            const int64_t kb = std::min(nb, n - k0);
            panel(k0, kb, s0);
            herk(k0, kb, k0 + kb, wk0, s0);
        }
        return;
    }
    panel(0, std::min(nb, n), s1);
    SA_HIP(hipEventRecord(ev_panel(0), s1));
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t k0 = k * nb, kb = std::min(nb, n - k0), c1 = k0 + kb;
        const int64_t r = n - c1;
        if (r <= 0) break;
        const int64_t nb1 = std::min(nb, r), c2 = c1 + nb1;
        // s1: step k on block row k+1, then panel k+1
        SA_HIP(hipStreamWaitEvent(s1, k > 0 ? ev_trail(k) : ev_panel(0), 0));
        block_row(k0, kb, c1, c2, s1);
        panel(c1, nb1, s1);
        SA_HIP(hipEventRecord(ev_panel(k + 1), s1));
        // s0: step k on the rows beyond block row k+1
        SA_HIP(hipStreamWaitEvent(s0, ev_panel(k), 0));
        herk(k0, kb, c2, wk0, s0);
        SA_HIP(hipEventRecord(ev_trail(k + 1), s0));
    }
    SA_HIP(hipEventRecord(c.evs[1], s1));
    SA_HIP(hipStreamWaitEvent(s0, c.evs[1], 0));
}

// Solve with LU factors in row-major (cm = 0) or column-major (cm = 1) layout;
// B row-major n x nrhs (ldb).  src/getrs.cc:82-115.
// Row permutation equivalent to the interchange sequence: perm[i] = original
// row that getrf moves to position i (swap perm[i] <-> perm[ipiv[i]] in order).
static void perm_from_ipiv(const int64_t* ipiv0, int64_t n, std::vector<int64_t>& perm)
{
    perm.resize(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[ipiv0[i]]);
}

static int64_t* upload_perm(Context& c, const int64_t* ipiv0_host, int64_t n)
{
    std::vector<int64_t> perm;
    perm_from_ipiv(ipiv0_host, n, perm);
    c.perm.ensure(sizeof(int64_t) * std::max<int64_t>(n, 1));
    SA_HIP(hipMemcpyAsync(c.perm.p, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    return c.perm.as<int64_t>();
}

static unsigned* tv_counter(Context& c)
{
    c.flags.ensure(sizeof(unsigned) * 16);
    return c.flags.as<unsigned>();
}

static int* err_flag(Context& c)
{
    if (c.err.bytes == 0) {
        c.err.ensure(sizeof(int) * 4);
        SA_HIP(hipMemsetAsync(c.err.p, 0, sizeof(int) * 4, c.s0));
    }
    return c.err.as<int>();
}

// Exchange words for one trsv_lookback launch over n rows: a fresh tag per
// launch; the words are zeroed when (re)allocated and when the tags wrap.
static TvSync tv_next(Context& c, int64_t n)
{
    const size_t words = size_t(2 * std::max<int64_t>(n, 1)) + 32;
    const bool fresh = c.tvw.bytes < words * sizeof(uint64_t);
    if (fresh || c.tv_tag == 0 || c.tv_tag > 0xFFFFFF00u) {
        c.tvw.ensure(words * sizeof(uint64_t));
        zero_u32(c.tvw.as<unsigned>(), int64_t(c.tvw.bytes / sizeof(unsigned)), c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        c.tv_tag = 1;
    }
    TvSync t;
    t.words = c.tvw.as<uint64_t>();
    t.progress = t.words + (c.tvw.bytes / sizeof(uint64_t) - 16);
    t.error = err_flag(c);
    t.tag = c.tv_tag++;
    return t;
}

// Workspace for the look-back solve data (trsv_diag_inverse) of a
// triangular pair.
template <class T>
static T* dinv_pair(Context& c, int64_t n)
{
    c.dinv.ensure(sizeof(T) * size_t(2 * trsv_aux_elems(n, int(sizeof(T)))));
    return c.dinv.as<T>();
}

// Single right-hand side: x := op(LU)^{-1} x with x contiguous, using the
// one-launch look-back triangular solves.  perm is the getrf permutation.
template <class T>
static void getrs_vec(Context& c, const T* LU, int64_t ld, int cm, int64_t n, const int64_t* perm,
                      T* x, T* tmp, int trans, hipStream_t s)
{
    T* d0 = dinv_pair<T>(c, n);
    T* d1 = d0 + trsv_aux_elems(n, int(sizeof(T)));
    if (!trans) {
        trsv_diag_inverse<T>(LU, ld, cm, 1, 1, n, d0, s);
        trsv_diag_inverse<T>(LU, ld, cm, 0, 0, n, d1, s);
        permute_convert<T, T>(x, tmp, perm, n, 0, s);
        trsv_lookback<T>(LU, ld, cm, 1, n, d0, tmp, tv_next(c, n), s);
        trsv_lookback<T>(LU, ld, cm, 0, n, d1, tmp, tv_next(c, n), s);
        copy_convert<T, T>(tmp, n, x, n, n, 1, s);
    }
    else {
        trsv_diag_inverse<T>(LU, ld, 1 - cm, 1, 0, n, d0, s);
        trsv_diag_inverse<T>(LU, ld, 1 - cm, 0, 1, n, d1, s);
        trsv_lookback<T>(LU, ld, 1 - cm, 1, n, d0, x, tv_next(c, n), s);
        trsv_lookback<T>(LU, ld, 1 - cm, 0, n, d1, x, tv_next(c, n), s);
        permute_convert<T, T>(x, tmp, perm, n, 1, s);
        copy_convert<T, T>(tmp, n, x, n, n, 1, s);
    }
}

template <class T>
static void getrs_any(Context& c, const T* LU, int64_t ld, int cm, int64_t n, const int64_t* ipiv,
                      T* B, int64_t ldb, int64_t nrhs, int trans, hipStream_t s)
{
    c.plan0.ensure(sizeof(int64_t) * (4 * 2048 + 8));
    int64_t* plan = c.plan0.as<int64_t>();
    if (!trans) {
        for (int64_t k1 = 0; k1 < n; k1 += 1024)
            laswp<T>(B, ldb, n, 0, nrhs, k1, std::min(n, k1 + 1024), ipiv, plan, s);
        trsm_rec<T>(LU, ld, cm, 1, 1, n, B, ldb, nrhs, s);
        trsm_rec<T>(LU, ld, cm, 0, 0, n, B, ldb, nrhs, s);
    }
    else {
        // A^T = U^T L^T P^T: solve U^T then L^T (transposed access), then P.
        trsm_rec<T>(LU, ld, 1 - cm, 1, 0, n, B, ldb, nrhs, s);
        trsm_rec<T>(LU, ld, 1 - cm, 0, 1, n, B, ldb, nrhs, s);
        // interchanges applied backward (src/getrs.cc:106-114): chunks from the
        // end, each as the inverse of its forward (dst <- src) plan
        for (int64_t k2 = n; k2 > 0; k2 -= 1024) {
            int64_t k1 = std::max<int64_t>(0, k2 - 1024);
            laswp_plan(k1, k2, n, ipiv, plan, s);
            laswp_invert_plan(plan, k2 - k1, s);
            laswp<T>(B, ldb, n, 0, nrhs, k1, k2, ipiv, plan, s, true);
        }
    }
}

// Solve with the row-major upper Cholesky factor: U^T y = b, U x = y.
template <class T>
static void potrs_rm(const T* U, int64_t ld, int64_t n, T* B, int64_t ldb, int64_t nrhs, hipStream_t s)
{
    trsm_rec<T>(U, ld, 1, 1, 0, n, B, ldb, nrhs, s);
    trsm_rec<T>(U, ld, 0, 0, 0, n, B, ldb, nrhs, s);
}

//------------------------------------------------------------------------------
struct Timer {
    Context& c;
    explicit Timer(Context& cc) : c(cc) { SA_HIP(hipEventRecord(c.t_start, c.s0)); }
    void stop() {
        SA_HIP(hipEventRecord(c.t_stop, c.s0));
        SA_HIP(hipEventSynchronize(c.t_stop));
        float ms = 0;
        SA_HIP(hipEventElapsedTime(&ms, c.t_start, c.t_stop));
        c.last_ms = ms;
    }
};

static void ipiv_to_host(Context& c, const int64_t* d, int64_t n, int64_t* host)
{
    if (!host || n <= 0) return;
    SA_HIP(hipMemcpyAsync(host, d, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    for (int64_t i = 0; i < n; ++i) host[i] += 1;     // LAPACK 1-based
}

static int64_t* ipiv_from_host(Context& c, const int64_t* host, int64_t n)
{
    c.ipiv.ensure(sizeof(int64_t) * std::max<int64_t>(n, 1));
    std::vector<int64_t> tmp(host, host + n);
    for (int64_t i = 0; i < n; ++i) {
        // getrf output satisfies i+1 <= ipiv[i] <= n (LAPACK convention)
        SA_REQUIRE(tmp[i] >= i + 1 && tmp[i] <= n, -5, "ipiv entry out of range");
        tmp[i] -= 1;
    }
    SA_HIP(hipMemcpy(c.ipiv.p, tmp.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    return c.ipiv.as<int64_t>();
}

static int64_t read_info(Context& c)
{
    int64_t info = 0;
    SA_HIP(hipMemcpyAsync(&info, c.dinfo.p, sizeof(int64_t), hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    return info;
}

static void reset_info(Context& c)
{
    c.dinfo.ensure(sizeof(int64_t) * 8);
    zero_u32(c.dinfo.as<unsigned>(), 16, c.s0);
}

// Run body() (a factorization issued on s0, forking to s1/s2 through events
// and joining back) through a HIP graph when o.graph is set: the first call
// with a given key runs eagerly (sizing every workspace) and then records the
// same launches into a graph; later calls with the same key replay it.  No
// memset, allocation or host synchronisation happens inside body() once the
// workspaces are sized; the panel exchange words are cleared inside the
// recorded sequence and its tags restart at 1 (the tag counter is restored
// afterwards, above every tag the graph writes).  A capture that fails leaves
// that key to the eager path.
template <class F>
static void run_graph(Context& c, const Opts& o, const GraphKey& key0, F&& body)
{
    if (!o.graph || c.prof) { body(); return; }
    GraphKey key = key0;
    key.generation = g_buf_generation;
    for (auto& g : c.graphs)
        if (g.first == key) { SA_HIP(hipGraphLaunch(g.second, c.s0)); ++c.graph_replays; return; }
    for (auto& k : c.graph_failed) if (k == key) { body(); return; }
    body();                                   // eager: the result, and every workspace sized
    SA_HIP(hipStreamSynchronize(c.s0));
    key.generation = g_buf_generation;
    const uint32_t saved_tag = c.leaf_tag;
    hipGraph_t graph = nullptr;
    c.capturing = true;
    try {
        SA_HIP(hipStreamBeginCapture(c.s0, hipStreamCaptureModeThreadLocal));
        try { body(); }
        catch (...) { (void)hipStreamEndCapture(c.s0, &graph); throw; }
        SA_HIP(hipStreamEndCapture(c.s0, &graph));
        c.capturing = false;
        SA_REQUIRE(key.generation == g_buf_generation, -2, "workspace changed during capture");
        hipGraphExec_t exec = nullptr;
        SA_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        c.graphs.push_back({key, exec});
    }
    catch (...) {
        c.graph_failed.push_back(key);
        (void)hipGetLastError();
    }
    c.capturing = false;
    if (graph) (void)hipGraphDestroy(graph);
    c.leaf_tag = std::max(saved_tag, c.leaf_tag);
}

template <class T>
static GraphKey lu_key(const T* A, const int64_t* ipiv, int64_t n, int64_t ld, const Opts& o)
{
    GraphKey k;
    k.kind = sizeof(T) == 8 ? 1 : 0;
    k.A = A; k.ipiv = ipiv; k.n = n; k.ld = ld; k.nb = o.nb; k.ib = o.ib; k.lookahead = o.lookahead;
    k.threshold = o.threshold; k.prec = o.prec; k.method = o.method;
    return k;
}

// In-place getrf of a column-major n x n matrix.
template <class T>
static int64_t getrf_cm(Context& c, T* A, int64_t n, int64_t lda, const Opts& o, int64_t* d_ipiv)
{
    reset_info(c);
    transpose_inplace<T>(A, n, lda, c.s0);
    run_graph(c, o, lu_key<T>(A, d_ipiv, n, lda, o), [&] { getrf_rm<T>(c, A, n, lda, o, d_ipiv, c.dinfo.as<int64_t>()); });
    transpose_inplace<T>(A, n, lda, c.s0);
    SA_HIP(hipStreamSynchronize(c.s0));
    check_kernel_error(c);
    return read_info(c);
}

// Row-major copy (n x nrhs) of a column-major B into a workspace and back.
template <class T>
static T* rhs_rm(Context& c, Buf& buf, const T* B, int64_t ldb, int64_t n, int64_t nrhs)
{
    buf.ensure(sizeof(T) * size_t(n) * size_t(nrhs));
    T* W = buf.as<T>();
    if (nrhs == 1) copy_convert<T, T>(B, ldb, W, n, n, 1, c.s0);
    else transpose_convert<T, T>(B, ldb, W, nrhs, n, nrhs, c.s0);
    return W;
}

template <class T>
static void rhs_back(Context& c, const T* W, T* B, int64_t ldb, int64_t n, int64_t nrhs)
{
    if (nrhs == 1) copy_convert<T, T>(W, n, B, ldb, n, 1, c.s0);
    else transpose_convert<T, T>(W, nrhs, B, ldb, nrhs, n, c.s0);
}

template <class T>
static void getrs_cm(Context& c, int trans, const T* A, int64_t lda, int64_t n, const int64_t* d_ipiv,
                     T* B, int64_t ldb, int64_t nrhs)
{
    if (nrhs == 1) {
        std::vector<int64_t> ip(n);
        SA_HIP(hipMemcpyAsync(ip.data(), d_ipiv, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c.s0));
        SA_HIP(hipStreamSynchronize(c.s0));
        const int64_t* perm = upload_perm(c, ip.data(), n);
        c.rhs.ensure(sizeof(T) * size_t(n));
        getrs_vec<T>(c, A, lda, 1, n, perm, B, c.rhs.as<T>(), trans, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        check_kernel_error(c);
        return;
    }
    T* W = rhs_rm<T>(c, c.rhs, B, ldb, n, nrhs);
    getrs_any<T>(c, A, lda, 1, n, d_ipiv, W, nrhs, nrhs, trans, c.s0);
    rhs_back<T>(c, W, B, ldb, n, nrhs);
}

template <class T>
static int64_t potrf_cm(Context& c, int uplo, T* A, int64_t n, int64_t lda, const Opts& o)
{
    reset_info(c);
    // Lower column-major == upper row-major: factor in place.  Upper storage is
    // flipped to lower first and flipped back (a permutation, so the
    // unreferenced triangle is restored exactly).
    if (uplo == SLATE_AMD_UPPER) transpose_inplace<T>(A, n, lda, c.s0);
    potrf_rm<T>(c, A, n, lda, o, c.dinfo.as<int64_t>());
    if (uplo == SLATE_AMD_UPPER) transpose_inplace<T>(A, n, lda, c.s0);
    return read_info(c);
}

template <class T>
static void potrs_cm(Context& c, int uplo, T* A, int64_t lda, int64_t n, T* B, int64_t ldb, int64_t nrhs)
{
    if (nrhs == 1) {
        // Lower column-major L (A = L L^T): L y = b, L^T x = y.
        // Upper column-major U (A = U^T U): U^T y = b, U x = y.
        const int cmL = (uplo == SLATE_AMD_UPPER) ? 0 : 1;
        T* d0 = dinv_pair<T>(c, n);
        T* d1 = d0 + trsv_aux_elems(n, int(sizeof(T)));
        trsv_diag_inverse<T>(A, lda, cmL, 1, 0, n, d0, c.s0);
        trsv_diag_inverse<T>(A, lda, 1 - cmL, 0, 0, n, d1, c.s0);
        trsv_lookback<T>(A, lda, cmL, 1, n, d0, B, tv_next(c, n), c.s0);
        trsv_lookback<T>(A, lda, 1 - cmL, 0, n, d1, B, tv_next(c, n), c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        check_kernel_error(c);
        return;
    }
    T* W = rhs_rm<T>(c, c.rhs, B, ldb, n, nrhs);
    if (uplo == SLATE_AMD_UPPER) {
        // column-major upper U (A = U^T U) is row-major lower U^T
        trsm_rec<T>(A, lda, 0, 1, 0, n, W, nrhs, nrhs, c.s0);
        trsm_rec<T>(A, lda, 1, 0, 0, n, W, nrhs, nrhs, c.s0);
    }
    else {
        potrs_rm<T>(A, lda, n, W, nrhs, nrhs, c.s0);
    }
    rhs_back<T>(c, W, B, ldb, n, nrhs);
}

//------------------------------------------------------------------------------
// Mixed precision.  The low-precision factor lives in a row-major float copy
// of A (one fused transpose + cast, src/gesv_mixed.cc:179) and never goes back
// to column-major; the FP64 residual reads the caller's A.
struct MixedState {
    float* Alo;           // row-major n x n
    int64_t* ipiv;        // device
    int64_t* perm;        // device, getrf permutation of the float factor
    float* xlo;           // row-major n x nrhs
    float* xtmp;          // scratch vector
    float* dinvL;         // look-back solve data (trsv_diag_inverse) of the first solve
    float* dinvU;         // ... and of the second
    double* R;            // column-major n x nrhs
    double* work;
    int64_t work_elems;
    double* norms;        // [0]: ||A||, [1..]: col norms
};

// Iteration 0 of a refinement already done by the selector's probe (one
// right-hand side): X holds x0 = (LU)^{-1} b from the probe factor, the
// state's R holds r0 = b - A x0, with ||A||_inf, ||x0||_inf and ||r0||_inf.
struct PreStart {
    bool valid = false;
    double anorm = 0, nx = 0, nr = 0;
};

// internal_util.hh:121-138 (converged unless ||r_i|| > ||x_i|| * cte), with
// one difference: a non-finite norm is "not converged".  The reference's
// comparison is false for NaN and for inf > inf, so a refinement that diverges
// to overflow would stop as converged with a non-finite solution; here it runs
// on to itermax and the FP64 fallback.
static bool converged_host(const std::vector<double>& nr, const std::vector<double>& nx, double cte)
{
    for (size_t i = 0; i < nr.size(); ++i) {
        if (!std::isfinite(nr[i]) || !std::isfinite(nx[i])) return false;
        if (nr[i] > nx[i] * cte) return false;
    }
    return true;
}

static void colnorms_host(Context& c, const double* X, int64_t n, int64_t ld, int64_t nrhs,
                          double* dev, std::vector<double>& out)
{
    col_norms_max(X, n, nrhs, ld, dev, c.s0);
    out.resize(nrhs);
    SA_HIP(hipMemcpyAsync(out.data(), dev, sizeof(double) * nrhs, hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
}

static double norm_host(Context& c, const double* A, int64_t n, int64_t lda, int sym, double* work,
                        int64_t work_elems, double* dev)
{
    norm_inf(A, n, n, lda, sym, work, work_elems, dev, c.s0);
    double v = 0;
    SA_HIP(hipMemcpyAsync(&v, dev, sizeof(double), hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    return v;
}

static void lo_solve(Context& c, const MixedState& st, int sym, int64_t n, int64_t nrhs)
{
    if (sym) potrs_rm<float>(st.Alo, n, n, st.xlo, nrhs, nrhs, c.s0);
    else getrs_any<float>(c, st.Alo, n, 0, n, st.ipiv, st.xlo, nrhs, nrhs, 0, c.s0);
}

// Single vector: dst (+)= A_lo^{-1} src with the float factor, the permutation
// fused into the cast to float and the cast back fused into the update
// (copy / getrs / copy / add of src/gesv_mixed.cc:220-232).
static void lo_solve_vec(Context& c, const MixedState& st, int sym, int64_t n, const double* src,
                         double* dst, bool accumulate)
{
    if (sym) {
        copy_convert<double, float>(src, n, st.xlo, n, n, 1, c.s0);
        trsv_lookback<float>(st.Alo, n, 1, 1, n, st.dinvL, st.xlo, tv_next(c, n), c.s0);   // U^T y = b
        trsv_lookback<float>(st.Alo, n, 0, 0, n, st.dinvU, st.xlo, tv_next(c, n), c.s0);   // U x = y
    }
    else {
        permute_convert<double, float>(src, st.xlo, st.perm, n, 0, c.s0);
        trsv_lookback<float>(st.Alo, n, 0, 1, n, st.dinvL, st.xlo, tv_next(c, n), c.s0);   // L y = P b
        trsv_lookback<float>(st.Alo, n, 0, 0, n, st.dinvU, st.xlo, tv_next(c, n), c.s0);   // U x = y
    }
    if (accumulate) add_convert<float>(st.xlo, dst, n, c.s0);
    else copy_convert<float, double>(st.xlo, n, dst, n, n, 1, c.s0);
}

// x_lo (row-major float) <- V (column-major double)
static void to_lo(Context& c, const double* V, int64_t ldv, float* xlo, int64_t n, int64_t nrhs)
{
    if (nrhs == 1) copy_convert<double, float>(V, ldv, xlo, n, n, 1, c.s0);
    else transpose_convert<double, float>(V, ldv, xlo, nrhs, n, nrhs, c.s0);
}
// V (column-major double) <- x_lo
static void from_lo(Context& c, const float* xlo, double* V, int64_t ldv, int64_t n, int64_t nrhs)
{
    if (nrhs == 1) copy_convert<float, double>(xlo, n, V, ldv, n, 1, c.s0);
    else transpose_convert<float, double>(xlo, nrhs, V, ldv, nrhs, n, c.s0);
}

static MixedState mixed_setup(Context& c, int64_t n, int64_t nrhs)
{
    MixedState st{};
    const int64_t dper = trsv_aux_elems(n, 4);
    c.lo.ensure(sizeof(float) * (size_t(n) * size_t(n) + size_t(2 * dper)) + 1024);
    st.Alo = c.lo.as<float>();
    st.dinvL = st.Alo + size_t(n) * size_t(n);
    st.dinvU = st.dinvL + dper;
    int64_t we = std::max<int64_t>(residual_work_elems(n, 4), ceildiv(n, 1024) * n + n + 64);
    // seven pieces, each rounded up to 256 bytes by take()
    size_t vec_bytes = 2 * sizeof(int64_t) * n + 2 * sizeof(float) * n * nrhs + sizeof(double) * n * nrhs
                       + sizeof(double) * we + sizeof(double) * (nrhs + 8) + 7 * 256;
    c.vec.ensure(vec_bytes);
    char* p = static_cast<char*>(c.vec.p);
    auto take = [&](size_t b) { char* q = p; p += (b + 255) & ~size_t(255); return q; };
    st.ipiv = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * n));
    st.perm = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * n));
    st.xlo = reinterpret_cast<float*>(take(sizeof(float) * n * nrhs));
    st.xtmp = reinterpret_cast<float*>(take(sizeof(float) * n * nrhs));
    st.R = reinterpret_cast<double*>(take(sizeof(double) * n * nrhs));
    st.work = reinterpret_cast<double*>(take(sizeof(double) * we));
    st.work_elems = we;
    st.norms = reinterpret_cast<double*>(take(sizeof(double) * (nrhs + 8)));
    return st;
}

static int64_t factor_lo(Context& c, MixedState& st, const double* A, int64_t lda, int64_t n, int sym,
                         const Opts& o)
{
    reset_info(c);
    // Fused transpose + cast into the row-major float copy.  For the symmetric
    // case the column-major lower triangle already is the row-major upper
    // triangle, so the cast is a plain copy.
    if (sym) copy_convert<double, float>(A, lda, st.Alo, n, n, n, c.s0);
    else transpose_convert<double, float>(A, lda, st.Alo, n, n, n, c.s0);
    if (sym) {
        // transposing lower-stored A gives the upper triangle in row-major,
        // which is what potrf_rm factors
        potrf_rm<float>(c, st.Alo, n, n, o, c.dinfo.as<int64_t>());
    }
    else {
        run_graph(c, o, lu_key<float>(st.Alo, st.ipiv, n, n, o),
                  [&] { getrf_rm<float>(c, st.Alo, n, n, o, st.ipiv, c.dinfo.as<int64_t>()); });
    }
    SA_HIP(hipStreamSynchronize(c.s0));
    if (read_info(c) == 0) {
        // inverted diagonal blocks of both triangles, reused by every solve
        if (sym) {
            trsv_diag_inverse<float>(st.Alo, n, 1, 1, 0, n, st.dinvL, c.s0);
            trsv_diag_inverse<float>(st.Alo, n, 0, 0, 0, n, st.dinvU, c.s0);
        }
        else {
            trsv_diag_inverse<float>(st.Alo, n, 0, 1, 1, n, st.dinvL, c.s0);
            trsv_diag_inverse<float>(st.Alo, n, 0, 0, 0, n, st.dinvU, c.s0);
        }
    }
    if (!sym) {
        check_kernel_error(c);
        std::vector<int64_t> ip(n), perm;
        SA_HIP(hipMemcpy(ip.data(), st.ipiv, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
        perm_from_ipiv(ip.data(), n, perm);
        SA_HIP(hipMemcpy(st.perm, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    }
    return read_info(c);
}

template <class T> static T* as_mut(const T* p) { return const_cast<T*>(p); }

// gesv_mixed / posv_mixed: src/gesv_mixed.cc:106-290, src/posv_mixed.cc:105-292.
// pre: a state whose float factor is already computed (info in pre_info),
// or null to set up and factor here.
static int64_t mixed_ir(Context& c, int sym, int uplo, int64_t n, int64_t nrhs, double* A, int64_t lda,
                        int64_t* ipiv_host, const double* B, int64_t ldb, double* X, int64_t ldx,
                        int* iter, const Opts& o, const MixedState* pre = nullptr, int64_t pre_info = 0,
                        const PreStart* ps = nullptr)
{
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = o.tol > 0 ? o.tol : eps * std::sqrt(double(n));
    const int64_t itermax = o.itermax;
    *iter = 0;
    SA_REQUIRE(!(sym && uplo == SLATE_AMD_UPPER), -2, "posv_mixed: Upper storage not supported");
    MixedState st = pre ? *pre : mixed_setup(c, n, nrhs);
    const bool started = ps && ps->valid && pre && nrhs == 1;
    const double Anorm = started ? ps->anorm : norm_host(c, A, n, lda, sym, st.work, st.work_elems, st.norms);
    const double cte = Anorm * tol;
    std::vector<double> nx, nr;
    bool converged = false;
    int64_t info = pre ? pre_info : factor_lo(c, st, A, lda, n, sym, o);
    if (info != 0) {
        *iter = -3;
    }
    else if (started) {
        nx.assign(1, ps->nx);                  // x0, r0 from the probe
        nr.assign(1, ps->nr);
    }
    else {
        if (nrhs == 1) {
            lo_solve_vec(c, st, sym, n, B, X, false);
        }
        else {
            to_lo(c, B, ldb, st.xlo, n, nrhs);
            lo_solve(c, st, sym, n, nrhs);
            from_lo(c, st.xlo, X, ldx, n, nrhs);
        }
        residual(A, n, lda, sym, X, ldx, B, ldb, st.R, n, nrhs, st.work, st.work_elems, c.s0);
        colnorms_host(c, X, n, ldx, nrhs, st.norms + 1, nx);
        colnorms_host(c, st.R, n, n, nrhs, st.norms + 1, nr);
    }
    if (info == 0) {
        if (converged_host(nr, nx, cte)) { *iter = 0; converged = true; }
        for (int64_t it = 0; it < itermax && !converged; ++it) {
            if (nrhs == 1) {
                lo_solve_vec(c, st, sym, n, st.R, X, true);
            }
            else {
                to_lo(c, st.R, n, st.xlo, n, nrhs);
                lo_solve(c, st, sym, n, nrhs);
                from_lo(c, st.xlo, st.R, n, n, nrhs);
                axpy(n, nrhs, 1.0, st.R, n, X, ldx, c.s0);
            }
            residual(A, n, lda, sym, X, ldx, B, ldb, st.R, n, nrhs, st.work, st.work_elems, c.s0);
            colnorms_host(c, X, n, ldx, nrhs, st.norms + 1, nx);
            colnorms_host(c, st.R, n, n, nrhs, st.norms + 1, nr);
            if (converged_host(nr, nx, cte)) { *iter = int(it + 1); converged = true; }
        }
    }
    if (!converged) {
        if (info == 0) *iter = int(-itermax - 1);
        if (o.fallback) {
            c.ipiv.ensure(sizeof(int64_t) * n);
            int64_t* dip = c.ipiv.as<int64_t>();
            if (sym) info = potrf_cm<double>(c, uplo, A, n, lda, o);
            else info = getrf_cm<double>(c, A, n, lda, o, dip);
            if (info == 0) {
                copy_convert<double, double>(B, ldb, X, ldx, n, nrhs, c.s0);
                if (sym) potrs_cm<double>(c, uplo, A, lda, n, X, ldx, nrhs);
                else getrs_cm<double>(c, 0, A, lda, n, dip, X, ldx, nrhs);
            }
            if (!sym) ipiv_to_host(c, dip, n, ipiv_host);
        }
    }
    else if (!sym) {
        ipiv_to_host(c, st.ipiv, n, ipiv_host);
    }
    SA_HIP(hipStreamSynchronize(c.s0));
    check_kernel_error(c);
    return info;
}

//------------------------------------------------------------------------------
// gesv_rbt: src/gesv_rbt.cc:94-236.  A := U^T A V with depth-d random
// butterflies, LU without pivoting, X = V (LU)^{-1} U^T B, then classic
// refinement in FP64 against a copy of the original A; the PP gesv of the copy
// is the fallback.  The butterfly weights are exp(r/20) with r the matgen
// rands stream of seed 42 (U) and 42 + nt (V), nt = ceil(n/nb)
// (internal_rbt_generate.cc:45-51, :133-134; the reference draws r with
// LAPACK larnv instead).  A holds the factors of the transformed matrix on
// return, or the original A when the fallback ran (gesv_rbt.cc:220-225).
static int64_t gesv_rbt(Context& c, int64_t n, int64_t nrhs, double* A, int64_t lda, const double* B,
                        int64_t ldb, double* X, int64_t ldx, int* iter, const Opts& o)
{
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = o.tol > 0 ? o.tol : eps * std::sqrt(double(n));
    const int64_t itermax = o.itermax;
    const int depth = o.depth;
    const int64_t nb = choose_nb(c, SLATE_AMD_TUNE_LU_FP64, n, o.nb);
    *iter = 0;

    const int64_t we = residual_work_elems(n, std::min<int64_t>(nrhs, 4));
    c.vec.ensure(sizeof(double) * (2 * size_t(n) + size_t(n) * size_t(nrhs) + size_t(we) + size_t(nrhs) + 8)
                 + 5 * 256);
    char* p = static_cast<char*>(c.vec.p);
    auto take = [&](size_t b) { char* q = p; p += (b + 255) & ~size_t(255); return q; };
    double* u = reinterpret_cast<double*>(take(sizeof(double) * n));
    double* v = reinterpret_cast<double*>(take(sizeof(double) * n));
    double* R = reinterpret_cast<double*>(take(sizeof(double) * n * nrhs));
    double* work = reinterpret_cast<double*>(take(sizeof(double) * we));
    double* norms = reinterpret_cast<double*>(take(sizeof(double) * (nrhs + 8)));
    rbt_weights(42, n, u, c.s0);
    rbt_weights(42 + ceildiv(n, nb), n, v, c.s0);

    double* Acopy = nullptr;
    double Anorm = 0;
    if (itermax > 0) {
        c.lo.ensure(sizeof(double) * size_t(n) * size_t(n));
        Acopy = c.lo.as<double>();
        copy_convert<double, double>(A, lda, Acopy, n, n, n, c.s0);
        Anorm = norm_host(c, A, n, lda, 0, work, we, norms);
    }
    copy_convert<double, double>(B, ldb, X, ldx, n, nrhs, c.s0);

    // factor and solve the transformed system
    gerbt2<double>(A, lda, n, nb, depth, u, v, c.s0);
    Opts on = o;
    on.method = SLATE_AMD_LU_NOPIV;
    c.ipiv.ensure(sizeof(int64_t) * n);
    int64_t* dip = c.ipiv.as<int64_t>();
    int64_t info = getrf_cm<double>(c, A, n, lda, on, dip);
    if (info == 0) {
        gerbt1<double>(1, X, ldx, n, nrhs, nb, depth, u, c.s0);
        getrs_cm<double>(c, 0, A, lda, n, dip, X, ldx, nrhs);
        gerbt1<double>(0, X, ldx, n, nrhs, nb, depth, v, c.s0);
    }
    if (itermax == 0) {
        SA_HIP(hipStreamSynchronize(c.s0));
        return info;
    }

    const double cte = Anorm * tol;
    std::vector<double> nx, nr;
    bool converged = false;
    if (info == 0) {
        residual(Acopy, n, n, 0, X, ldx, B, ldb, R, n, nrhs, work, we, c.s0);
        colnorms_host(c, X, n, ldx, nrhs, norms, nx);
        colnorms_host(c, R, n, n, nrhs, norms, nr);
        if (converged_host(nr, nx, cte)) converged = true;
        for (int64_t it = 0; it < itermax && !converged; ++it) {
            gerbt1<double>(1, R, n, n, nrhs, nb, depth, u, c.s0);
            getrs_cm<double>(c, 0, A, lda, n, dip, R, n, nrhs);
            gerbt1<double>(0, R, n, n, nrhs, nb, depth, v, c.s0);
            axpy(n, nrhs, 1.0, R, n, X, ldx, c.s0);
            residual(Acopy, n, n, 0, X, ldx, B, ldb, R, n, nrhs, work, we, c.s0);
            colnorms_host(c, X, n, ldx, nrhs, norms, nx);
            colnorms_host(c, R, n, n, nrhs, norms, nr);
            if (converged_host(nr, nx, cte)) { *iter = int(it + 1); converged = true; }
        }
    }
    // A zero pivot of the unpivoted factor (info > 0) leaves X undefined; the
    // reference would refine a non-finite X, which is "not converged" here.
    if (!converged) {
        *iter = int(-itermax - 1);
        if (o.fallback) {
            copy_convert<double, double>(Acopy, n, A, lda, n, n, c.s0);
            info = getrf_cm<double>(c, Acopy, n, n, o, dip);
            copy_convert<double, double>(B, ldb, X, ldx, n, nrhs, c.s0);
            if (info == 0) getrs_cm<double>(c, 0, Acopy, n, n, dip, X, ldx, nrhs);
        }
    }
    SA_HIP(hipStreamSynchronize(c.s0));
    check_kernel_error(c);
    return info;
}

// Givens rotation with the reference-BLAS drotg conventions.
static void rotg(double* a, double* b, double* cc, double* ss)
{
    double aa = *a, bb = *b;
    double roe = std::fabs(aa) > std::fabs(bb) ? aa : bb;
    double scale = std::fabs(aa) + std::fabs(bb);
    double r, z;
    if (scale == 0) { *cc = 1; *ss = 0; r = 0; z = 0; }
    else {
        r = scale * std::sqrt((aa / scale) * (aa / scale) + (bb / scale) * (bb / scale));
        r = std::copysign(1.0, roe) * r;
        *cc = aa / r; *ss = bb / r;
        z = 1;
        if (std::fabs(aa) > std::fabs(bb)) z = *ss;
        if (std::fabs(bb) >= std::fabs(aa) && *cc != 0) z = 1 / *cc;
    }
    *a = r; *b = z;
}

static inline void rot(double* x, double* y, double cc, double ss)
{
    double t = cc * *x + ss * *y;
    *y = cc * *y - ss * *x;
    *x = t;
}

// gesv_mixed_gmres / posv_mixed_gmres: src/gesv_mixed_gmres.cc:111-410.  The
// Krylov bases V, W are FP64 on the device; the (restart+1)^2 Hessenberg
// matrix and Givens rotations are on the host like the reference (:168-177).
static int64_t mixed_gmres(Context& c, int sym, int uplo, int64_t n, double* A, int64_t lda,
                           int64_t* ipiv_host, const double* B, double* X, int* iter, const Opts& o,
                           int64_t nb_hint, const MixedState* pre = nullptr, int64_t pre_info = 0,
                           const PreStart* ps = nullptr)
{
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = o.tol > 0 ? o.tol : eps * std::sqrt(double(n));
    const int64_t itermax = o.itermax;
    const int64_t nb = choose_nb(c, sym ? SLATE_AMD_TUNE_CHOL : SLATE_AMD_TUNE_LU_FP32, n, nb_hint);
    const int64_t restart = std::min<int64_t>(std::min<int64_t>(30, itermax), std::min(nb, n) - 1);
    SA_REQUIRE(restart >= 1, -2, "gmres: restart length < 1");
    SA_REQUIRE(!(sym && uplo == SLATE_AMD_UPPER), -2, "posv_mixed_gmres: Upper storage not supported");
    *iter = 0;
    MixedState st = pre ? *pre : mixed_setup(c, n, 1);
    const int64_t R1 = restart + 1;
    c.basis.ensure(sizeof(double) * size_t(n) * size_t(2 * R1) + sizeof(double) * 256);
    double* V = c.basis.as<double>();
    double* W = V + n * R1;
    double* dsc = W + n * R1;          // device scalars: [0..R1) h, [64..] z, [128] norm
    bool started = ps && ps->valid && pre;
    const double Anorm = started ? ps->anorm : norm_host(c, A, n, lda, sym, st.work, st.work_elems, st.norms);
    const double cte = Anorm * tol;
    std::vector<double> H(R1 * R1, 0.0), S(R1, 0.0), ga(restart), gb(restart), hz(R1), zz(R1);
    std::vector<double> nx(1), nr(1);
    bool converged = false;
    int64_t info = pre ? pre_info : factor_lo(c, st, A, lda, n, sym, o);
    if (info != 0) {
        *iter = -3;
    }
    else {
        const double negone = -1.0;
        SA_HIP(hipMemcpyAsync(dsc + 129, &negone, sizeof(double), hipMemcpyHostToDevice, c.s0));
        if (!started) lo_solve_vec(c, st, sym, n, B, X, false);
        int64_t iiter = 0;
        while (iiter < itermax) {
            if (started) {                     // x0, r0 and their norms from the probe
                nx.assign(1, ps->nx);
                nr.assign(1, ps->nr);
                started = false;
            }
            else {
                residual(A, n, lda, sym, X, n, B, n, st.R, n, 1, st.work, st.work_elems, c.s0);
                colnorms_host(c, X, n, n, 1, st.norms + 1, nx);
                colnorms_host(c, st.R, n, n, 1, st.norms + 1, nr);
            }
            if (converged_host(nr, nx, cte)) { *iter = int(iiter); converged = true; break; }
            // v0 = r / ||r||
            copy_convert<double, double>(st.R, n, V, n, n, 1, c.s0);
            nrm2(V, n, st.work, dsc + 128, c.s0);
            double beta = 0;
            SA_HIP(hipMemcpyAsync(&beta, dsc + 128, sizeof(double), hipMemcpyDeviceToHost, c.s0));
            SA_HIP(hipStreamSynchronize(c.s0));
            std::vector<double> arn = {beta};
            if (beta == 0) { *iter = int(iiter); converged = false; break; }
            scale(V, n, dsc + 128, c.s0);
            S.assign(R1, 0.0);
            S[0] = beta;
            int64_t j = 0;
            for (; j < restart && iiter < itermax && !converged_host(arn, nx, cte); ++j, ++iiter) {
                double* Vj = V + j * n;
                double* Vj1 = V + (j + 1) * n;
                double* Wj1 = W + (j + 1) * n;
                // W(:, j+1) = M^{-1} V(:, j) in low precision (:281-286)
                lo_solve_vec(c, st, sym, n, Vj, Wj1, false);
                // V(:, j+1) = A W(:, j+1) in FP64 (:288-293): R = 0 - A w, then negate
                zero_f64(st.R, n, c.s0);
                residual(A, n, lda, sym, Wj1, n, st.R, n, Vj1, n, 1, st.work, st.work_elems, c.s0);
                scale(Vj1, n, dsc + 129, c.s0);                 // dsc[129] = -1
                // CGS2 (:296-327): h = V^T v; v -= V h; z = V^T v; v -= V z; h += z
                multi_dot(V, n, n, j + 1, Vj1, dsc, st.work, c.s0);
                multi_axpy(V, n, n, j + 1, dsc, Vj1, c.s0);
                multi_dot(V, n, n, j + 1, Vj1, dsc + 64, st.work, c.s0);
                multi_axpy(V, n, n, j + 1, dsc + 64, Vj1, c.s0);
                nrm2(Vj1, n, st.work, dsc + 128, c.s0);
                scale(Vj1, n, dsc + 128, c.s0);
                double nrm = 0;
                SA_HIP(hipMemcpyAsync(hz.data(), dsc, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c.s0));
                SA_HIP(hipMemcpyAsync(zz.data(), dsc + 64, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c.s0));
                SA_HIP(hipMemcpyAsync(&nrm, dsc + 128, sizeof(double), hipMemcpyDeviceToHost, c.s0));
                SA_HIP(hipStreamSynchronize(c.s0));
                double* Hj = H.data() + j * R1;
                for (int64_t p = 0; p <= j; ++p) Hj[p] = hz[p] + zz[p];
                Hj[j + 1] = nrm;
                for (int64_t i = 0; i < j; ++i) rot(&Hj[i], &Hj[i + 1], ga[i], gb[i]);
                double hjj = Hj[j], hj1 = Hj[j + 1];
                rotg(&hjj, &hj1, &ga[j], &gb[j]);
                rot(&Hj[j], &Hj[j + 1], ga[j], gb[j]);
                rot(&S[j], &S[j + 1], ga[j], gb[j]);
                arn[0] = std::fabs(S[j + 1]);
            }
            // X += W(:, 1:j) * (H(0:j, 0:j)^{-1} S(0:j)) (:360-374)
            std::vector<double> y(S.begin(), S.begin() + j);
            for (int64_t p = j - 1; p >= 0; --p) {
                y[p] /= H[p + p * R1];
                for (int64_t q = 0; q < p; ++q) y[q] -= H[q + p * R1] * y[p];
            }
            if (j > 0) {
                for (int64_t p = 0; p < j; ++p) y[p] = -y[p];       // multi_axpy subtracts
                SA_HIP(hipMemcpyAsync(dsc + 160, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, c.s0));
                multi_axpy(W + n, n, n, j, dsc + 160, X, c.s0);
            }
        }
    }
    if (!converged) {
        if (info == 0) *iter = int(-itermax - 1);
        if (o.fallback) {
            c.ipiv.ensure(sizeof(int64_t) * n);
            int64_t* dip = c.ipiv.as<int64_t>();
            if (sym) info = potrf_cm<double>(c, uplo, A, n, lda, o);
            else info = getrf_cm<double>(c, A, n, lda, o, dip);
            if (info == 0) {
                copy_convert<double, double>(B, n, X, n, n, 1, c.s0);
                if (sym) potrs_cm<double>(c, uplo, A, lda, n, X, n, 1);
                else getrs_cm<double>(c, 0, A, lda, n, dip, X, n, 1);
            }
            if (!sym) ipiv_to_host(c, dip, n, ipiv_host);
        }
    }
    else if (!sym) {
        ipiv_to_host(c, st.ipiv, n, ipiv_host);
    }
    SA_HIP(hipStreamSynchronize(c.s0));
    check_kernel_error(c);
    return info;
}

// Estimate of ||A^{-1}||_1 by Hager's method with Higham's refinements, the
// steps of LAPACK dlacn2 that SLATE's internal::norm1est restates
// (src/internal/internal_norm1est.cc): x = 1/n, then alternately
// y = A^{-1} x and z = A^{-T} sign(y), moving x to the unit vector of the
// largest |z_j| until the sign vector repeats, the estimate stops growing or
// five iterations pass; finally the alternating vector
// x_i = (-1)^i (1 + i/(n-1)) guards against unlucky starts.  The vector logic
// runs on the host (n doubles per step); solve(x, t) overwrites the device
// vector x with A^{-1} x (t = 0) or A^{-T} x (t = 1).
template <class Solve>
static double norm1est_inv(int64_t n, double* xd, Solve solve)
{
    std::vector<double> x(size_t(n), 1.0 / double(n));
    std::vector<int> isgn(size_t(n), 0);
    auto run = [&](int t) {
        SA_HIP(hipMemcpy(xd, x.data(), sizeof(double) * size_t(n), hipMemcpyHostToDevice));
        solve(xd, t);
        SA_HIP(hipMemcpy(x.data(), xd, sizeof(double) * size_t(n), hipMemcpyDeviceToHost));
    };
    auto asum = [&]() { double s = 0; for (double e : x) s += std::fabs(e); return s; };
    auto imax = [&]() {
        int64_t j = 0;
        for (int64_t i = 1; i < n; ++i) if (std::fabs(x[size_t(i)]) > std::fabs(x[size_t(j)])) j = i;
        return j;
    };
    run(0);
    if (n == 1) return std::fabs(x[0]);
    double est = asum();
    for (int64_t i = 0; i < n; ++i) { isgn[size_t(i)] = x[size_t(i)] >= 0 ? 1 : -1; x[size_t(i)] = isgn[size_t(i)]; }
    run(1);
    int64_t j = imax();
    for (int iter = 2;; ++iter) {
        std::fill(x.begin(), x.end(), 0.0);
        x[size_t(j)] = 1.0;
        run(0);
        const double estold = est;
        est = asum();
        bool repeated = true;
        for (int64_t i = 0; i < n && repeated; ++i) repeated = (x[size_t(i)] >= 0 ? 1 : -1) == isgn[size_t(i)];
        if (repeated || est <= estold) break;
        for (int64_t i = 0; i < n; ++i) { isgn[size_t(i)] = x[size_t(i)] >= 0 ? 1 : -1; x[size_t(i)] = isgn[size_t(i)]; }
        run(1);
        const int64_t jlast = j;
        j = imax();
        if (!(x[size_t(jlast)] != std::fabs(x[size_t(j)]) && iter < 5)) break;
    }
    double alt = 1.0;
    for (int64_t i = 0; i < n; ++i) { x[size_t(i)] = alt * (1.0 + double(i) / double(n - 1)); alt = -alt; }
    run(0);
    const double temp = 2.0 * asum() / double(3 * n);
    return std::max(est, temp);
}

//------------------------------------------------------------------------------
// Learned selection (no counterpart in the reference).  Features of a general
// FP64 system A x = b, computed on the device:
//   f0  log2 n / 16
//   f1  log10 of ||A||_inf ||x0||_inf / ||b||_inf / 16, with x0 = (LU)^{-1} b
//       from the probe factor: a lower bound on kappa_inf(A) that a random
//       right-hand side makes tight within about sqrt(n) (it reads ~20 when
//       the probe factor is singular)
//   f2  log2(max / min) of the per-256x256-tile max |a_ij| / 64
//   f3  standard deviation of log2 of those tile maxima / 16
//   f4..f19  fractions of entries whose binary exponent lies 0..14 (and
//       >= 15 or zero) below the largest one, over every column when
//       n <= 4096 and about 4096 evenly spaced columns above
//   f20 1
//   f21 log10 of the probe solve's backward error ||r0|| / (||A|| ||x0||) / 16,
//       r0 = b - A x0 in FP64
//   f22 log10 ||r0|| / ||b|| / 16
//   f23 1 if x0 already meets the refinement's stopping rule
// One full pass over A gives the tile maxima and the row sums (so ||A||_inf),
// one pass over the sampled columns the histogram; x0, r0 and the norms are
// iteration 0 of the refinement, handed to the chosen driver in *ps when the
// probe factor is kept.  The float factor is left in st.
static constexpr int SEL_F = 24;

static void selector_features(Context& c, MixedState& st, int64_t n, const double* A, int64_t lda,
                              const double* B, double* X, const Opts& o, int64_t* info, float* f, PreStart* ps)
{
    const int64_t t = ceildiv(n, 256);
    c.flags.ensure(sizeof(double) * size_t(t * t + n + t * n) + sizeof(unsigned long long) * 16 + 512);
    double* tmax = c.flags.as<double>();
    double* rowsum = tmax + t * t;
    unsigned long long* hist = reinterpret_cast<unsigned long long*>(rowsum + n);
    double* part = reinterpret_cast<double*>(hist + 16);
    feature_tiles_rows(A, lda, n, tmax, part, rowsum, c.s0);
    std::vector<double> tm(static_cast<size_t>(t * t));
    std::vector<double> rs(static_cast<size_t>(n));
    SA_HIP(hipMemcpyAsync(tm.data(), tmax, sizeof(double) * tm.size(), hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipMemcpyAsync(rs.data(), rowsum, sizeof(double) * rs.size(), hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    double mx = 0, mn = std::numeric_limits<double>::infinity(), s1 = 0, s2 = 0;
    int64_t cnt = 0;
    for (double v : tm) {
        mx = std::max(mx, v);
        if (v > 0) { mn = std::min(mn, v); double l = std::log2(v); s1 += l; s2 += l * l; ++cnt; }
    }
    double anorm = 0;
    for (double v : rs) anorm = std::max(anorm, v);
    int emax = 0;
    std::frexp(mx > 0 ? mx : 1.0, &emax);
    const int64_t cstep = std::max<int64_t>(1, n / 4096);
    zero_u32(reinterpret_cast<unsigned*>(hist), 32, c.s0);
    feature_hist(A, lda, n, cstep, emax, hist, c.s0);
    unsigned long long h[16];
    SA_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, c.s0));
    *info = factor_lo(c, st, A, lda, n, 0, o);
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = o.tol > 0 ? o.tol : eps * std::sqrt(double(n));
    double nx = 0, nr = 0, nb = 0;
    if (*info == 0) {
        lo_solve_vec(c, st, 0, n, B, X, false);
        residual(A, n, lda, 0, X, n, B, n, st.R, n, 1, st.work, st.work_elems, c.s0);
        std::vector<double> v;
        colnorms_host(c, X, n, n, 1, st.norms + 1, v);
        nx = v[0];
        colnorms_host(c, st.R, n, n, 1, st.norms + 1, v);
        nr = v[0];
    }
    {
        std::vector<double> v;
        colnorms_host(c, B, n, n, 1, st.norms + 1, v);
        nb = v[0];
    }
    SA_HIP(hipStreamSynchronize(c.s0));
    const double ncols = double(ceildiv(n, cstep));
    const double total = double(n) * ncols;
    const double tiny = 1e-300;
    const bool ok = *info == 0 && std::isfinite(nx) && std::isfinite(nr) && nx > 0 && nb > 0 && anorm > 0;
    for (int k = 0; k < SEL_F; ++k) f[k] = 0.f;
    f[0] = float(std::log2(double(n)) / 16.0);
    f[1] = float((ok ? std::log10(std::max(anorm * nx / nb, 1.0)) : 20.0) / 16.0);
    f[2] = float((cnt > 0 && mn < std::numeric_limits<double>::infinity()) ? std::log2(mx / mn) / 64.0 : 1.0);
    const double mean = cnt ? s1 / double(cnt) : 0.0;
    f[3] = float(cnt ? std::sqrt(std::max(0.0, s2 / double(cnt) - mean * mean)) / 16.0 : 0.0);
    for (int k = 0; k < 16; ++k) f[4 + k] = float(double(h[k]) / total);
    f[20] = 1.0f;                                   // bias-like constant input
    if (ok) {
        f[21] = float(std::log10(std::max(nr / (anorm * nx), tiny)) / 16.0);
        f[22] = float(std::log10(std::max(nr / nb, tiny)) / 16.0);
        f[23] = (nr <= nx * anorm * tol) ? 1.f : 0.f;
    }
    if (ps) {
        ps->valid = ok;
        ps->anorm = anorm;
        ps->nx = nx;
        ps->nr = nr;
    }
}

// The selector's choices: (factor precision, scheme).  The features come from
// a probe factor in SEL_PROBE precision, which choices 0 and 1 keep; the
// others factor again in their own precision (FP8X2 is cheaper but only pays
// when the probe is not kept, which the labels decide).  Choice 7 is the plain
// FP64 getrf + getrs.
struct SelChoice { int prec; int gmres; };
static constexpr int SEL_K = 8;
static constexpr int SEL_PROBE = SLATE_AMD_PREC_BF16X2;
static constexpr SelChoice SEL_CHOICES[SEL_K] = {
    {SLATE_AMD_PREC_BF16X2, 0}, {SLATE_AMD_PREC_BF16X2, 1}, {SLATE_AMD_PREC_BF16X3, 0},
    {SLATE_AMD_PREC_BF16X3, 1}, {SLATE_AMD_PREC_FP32, 0},   {SLATE_AMD_PREC_FP32, 1},
    {SLATE_AMD_PREC_FP8X2, 1},  {SLATE_AMD_PREC_FP64, 0}};

// gesv_mixed with the learned selector.  params (host, mlp_num_params()
// floats): the trained MLP; force >= 0 runs that choice instead of the
// network (the labels are the measured cost of each forced choice); with
// neither, SLATE's default runs (float factor + classic IR,
// src/gesv_mixed.cc) without a probe, *choice = 4.  A refinement that misses
// the tolerance falls back to FP64 as in the reference.
static int64_t mixed_select(Context& c, int64_t n, double* A, int64_t lda, int64_t* ipiv_host, const double* B,
                            double* X, int* iter, int* choice, const float* params, int force, float* feat_out,
                            const Opts& o)
{
    SA_REQUIRE(force < SEL_K, -2, "selector: choice out of range");
    if (!params && force < 0) {
        Opts od = o;
        od.prec = SLATE_AMD_PREC_FP32;
        *choice = 4;
        if (feat_out) std::fill(feat_out, feat_out + SEL_F, 0.f);
        return mixed_ir(c, 0, SLATE_AMD_LOWER, n, 1, A, lda, ipiv_host, B, n, X, n, iter, od);
    }
    Opts op = o;
    op.prec = SEL_PROBE;
    MixedState st = mixed_setup(c, n, 1);
    int64_t finfo = 0;
    float f[SEL_F];
    PreStart ps;
    selector_features(c, st, n, A, lda, B, X, op, &finfo, f, &ps);
    if (feat_out) std::copy(f, f + SEL_F, feat_out);
    int ch = force;
    if (ch < 0) {
        ch = SEL_K - 1;                 // a singular probe factor: straight to FP64
        if (finfo == 0) {
            const int np = mlp_num_params();
            c.tinv.ensure(sizeof(float) * size_t(np + SEL_F + 16));
            float* dp = c.tinv.as<float>();
            float* dx = dp + np;
            float* dprob = dx + SEL_F;
            SA_HIP(hipMemcpyAsync(dp, params, sizeof(float) * np, hipMemcpyHostToDevice, c.s0));
            SA_HIP(hipMemcpyAsync(dx, f, sizeof(float) * SEL_F, hipMemcpyHostToDevice, c.s0));
            mlp_forward(dp, dx, 1, dprob, c.s0);
            float pr[SEL_K];
            SA_HIP(hipMemcpyAsync(pr, dprob, sizeof(pr), hipMemcpyDeviceToHost, c.s0));
            SA_HIP(hipStreamSynchronize(c.s0));
            ch = 0;
            for (int k = 1; k < SEL_K; ++k) if (pr[k] > pr[ch]) ch = k;
        }
    }
    *choice = ch;
    const SelChoice sc = SEL_CHOICES[ch];
    if (sc.prec == SLATE_AMD_PREC_FP64) {
        *iter = 0;
        c.ipiv.ensure(sizeof(int64_t) * n);
        int64_t* dip = c.ipiv.as<int64_t>();
        int64_t info = getrf_cm<double>(c, A, n, lda, o, dip);
        if (info == 0) {
            copy_convert<double, double>(B, n, X, n, n, 1, c.s0);
            getrs_cm<double>(c, 0, A, lda, n, dip, X, n, 1);
        }
        ipiv_to_host(c, dip, n, ipiv_host);
        SA_HIP(hipStreamSynchronize(c.s0));
        return info;
    }
    Opts oc = o;
    oc.prec = sc.prec;
    const bool keep = sc.prec == SEL_PROBE;       // the probe factor is this choice's factor
    if (sc.gmres)
        return mixed_gmres(c, 0, SLATE_AMD_LOWER, n, A, lda, ipiv_host, B, X, iter, oc, o.nb,
                           keep ? &st : nullptr, keep ? finfo : 0, keep ? &ps : nullptr);
    return mixed_ir(c, 0, SLATE_AMD_LOWER, n, 1, A, lda, ipiv_host, B, n, X, n, iter, oc,
                    keep ? &st : nullptr, keep ? finfo : 0, keep ? &ps : nullptr);
}

} // namespace sa

//==============================================================================
// C ABI
//==============================================================================
struct slate_amd_context_t { sa::Context c; };

static thread_local std::string g_last_error;

// Inputs may have been written on another stream (e.g. torch's): settle them.
#define SA_ENTRY_SYNC() SA_HIP(hipDeviceSynchronize())

template <class F>
static int64_t guarded(slate_amd_context ctx, F&& f)
{
    try {
        if (!ctx) throw sa::Error(-1, "null context");
        SA_HIP(hipSetDevice(ctx->c.device));
        return f(ctx->c);
    }
    catch (const sa::Error& e) { g_last_error = e.what(); return e.code < 0 ? e.code : -1; }
    catch (const std::exception& e) { g_last_error = e.what(); return -99; }
    catch (...) { g_last_error = "unknown error"; return -99; }
}

template <class T>
static int gemm_entry(slate_amd_context ctx, int64_t m, int64_t n, int64_t k, double alpha,
                      const T* A, int64_t lda, const T* B, int64_t ldb, double beta, T* C, int64_t ldc)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(m >= 0 && n >= 0 && k >= 0, -3, "bad m/n/k");
        SA_REQUIRE(lda >= std::max<int64_t>(1, m) && ldb >= std::max<int64_t>(1, k)
                   && ldc >= std::max<int64_t>(1, m), -3, "bad leading dimension");
        if (m == 0 || n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Timer t(c);
        // column-major C = row-major C^T = alpha * B^T A^T + beta * C^T
        sa::GemmArgs g{};
        g.M = n; g.N = m; g.K = k;
        g.A = B; g.lda = ldb; g.B = A; g.ldb = lda; g.C = C; g.ldc = ldc;
        g.alpha = alpha; g.beta = beta; g.uplo = 0;
        if (k == 0) {
            SA_REQUIRE(beta == 1.0, -2, "gemm: k == 0 requires beta == 1");
        }
        else {
            sa::gemm<T>(g, sa::A_KC, sa::B_NC, c.s0);
        }
        t.stop();
        return 0;
    }));
}

// C (column-major m x n) += alpha * A B through the split-BF16 kernel:
// column-major C is row-major C^T = B^T A^T, so B^T is the K-contiguous left
// operand (split_rows) and A^T the right one (split_cols_t).
static int gemm_bf16_entry(slate_amd_context ctx, int terms, int64_t m, int64_t n, int64_t k, float alpha,
                           const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(terms >= 1 && terms <= 3, -2, "terms must be 1, 2 or 3");
        SA_REQUIRE(m >= 0 && n >= 0 && k >= 0, -3, "bad m/n/k");
        SA_REQUIRE(lda >= std::max<int64_t>(1, m) && ldb >= std::max<int64_t>(1, k)
                   && ldc >= std::max<int64_t>(1, m), -3, "bad leading dimension");
        if (m == 0 || n == 0 || k == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Timer t(c);
        const int64_t Kp = sa::ceildiv(k, 32) * 32;
        const int64_t pa = n * Kp, pb = m * Kp;
        if (k >= 4 && ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0) {
            // B^T (n x k) is k-contiguous in place; A^T needs a transposed copy
            c.bfg.ensure(sizeof(float) * size_t(pb));
            float* at = c.bfg.as<float>();
            sa::transpose_f32(A, lda, k, m, at, Kp, c.s0);
            sa::gemm_s32(terms, n, m, k, B, ldb, at, Kp, C, ldc, alpha, c.s0);
        }
        else {
            c.bfg.ensure(sizeof(__bf16) * size_t(terms) * size_t(pa + pb));
            __bf16* a = c.bfg.as<__bf16>();
            __bf16* b = a + terms * pa;
            sa::split_rows(terms, B, ldb, n, k, a, Kp, pa, c.s0);
            sa::split_cols_t(terms, A, lda, k, m, b, Kp, pb, c.s0);
            sa::gemm_bf16s(terms, n, m, Kp, a, Kp, pa, b, Kp, pb, C, ldc, alpha, c.s0, 0, 0);
        }
        t.stop();
        return 0;
    }));
}

int slate_amd_sgemm_fp8(slate_amd_context ctx, int terms, int64_t m, int64_t n, int64_t k, float alpha,
                        const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(terms == 1 || terms == 2, -2, "terms must be 1 or 2");
        SA_REQUIRE(m >= 0 && n >= 0 && k >= 0, -3, "bad m/n/k");
        SA_REQUIRE(lda >= std::max<int64_t>(1, m) && ldb >= std::max<int64_t>(1, k)
                   && ldc >= std::max<int64_t>(1, m), -3, "bad leading dimension");
        if (m == 0 || n == 0 || k == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Timer t(c);
        // column-major C (m x n) is row-major C^T (n x m): C^T += alpha B^T A^T
        // with B^T (n x k) k-contiguous in place and A^T from A (k x m row-major view)
        const int64_t Kp = sa::fp8_kp(k);
        const int64_t pb = sa::fp8_rows(n) * Kp, pa = sa::fp8_rows(m) * Kp;
        c.bfg.ensure(size_t(terms) * size_t(pa + pb) + 256);
        uint8_t* bq = c.bfg.as<uint8_t>();
        uint8_t* aq = bq + terms * pb;
        c.f8s.ensure(sizeof(int) * size_t(sa::fp8_tiles(n) + sa::fp8_tiles(m) + 4));
        int* sb = c.f8s.as<int>();
        int* sa_ = sb + sa::fp8_tiles(n);
        sa::quant_fp8_rows(terms, B, ldb, n, k, bq, Kp, pb, sb, c.s0);
        sa::quant_fp8_cols_t(terms, A, lda, k, m, aq, Kp, pa, sa_, c.s0);
        sa::gemm_fp8(terms, n, m, Kp, bq, pb, sb, aq, pa, sa_, C, ldc, alpha, c.s0, 0, 0);
        t.stop();
        return 0;
    }));
}


namespace sa {

// Spectrum-controlled test matrix, column-major n x n:
//   A = U diag(sigma) V^T, U's reflectors the columns of randn(seed + 1) and
//   V's those of randn(seed);  sym: A = V diag(sigma) V^T (SPD for cond >= 1).
// Works on the row-major view R = A^T = V diag(sigma) U^T, applying the
// compact-WY blocks Q_k = I - Y_k T_k Y_k^T of V = Q_1 ... Q_m from the left
// and of U^T from the right, last block first; each block is 7 MFMA GEMMs.
static void generate_svd(Context& c, int sdist, double cond, int64_t seed, int sym, int64_t n, double* R,
                         int64_t ld)
{
    hipStream_t s = c.s0;
    SA_HIP(hipMemset2DAsync(R, sizeof(double) * ld, 0, sizeof(double) * n, n, s));
    sigma_diag(sdist, cond, n, R, ld, s);
    const int64_t b = std::min<int64_t>(256, n);
    c.work.ensure(sizeof(double) * size_t(3 * n * b + 2 * b * b + 64));
    double* Y = c.work.as<double>();
    double* W1 = Y + n * b;
    double* W2 = W1 + n * b;
    double* G = W2 + n * b;
    double* T = G + b * b;
    auto gm = [&](int64_t M, int64_t N, int64_t K, const double* A_, int64_t lda_, int alay, const double* B_,
                  int64_t ldb_, int blay, double alpha, double beta, double* C_, int64_t ldc_) {
        GemmArgs g{};
        g.M = M; g.N = N; g.K = K; g.A = A_; g.lda = lda_; g.B = B_; g.ldb = ldb_;
        g.C = C_; g.ldc = ldc_; g.alpha = alpha; g.beta = beta; g.uplo = 0;
        gemm<double>(g, alay, blay, s);
    };
    // Y (n x bk, column-major), its T; Y(k, j) = Y[k + j n].
    auto block = [&](int64_t sd, int64_t j0, int64_t bk) {
        generate(3, sd, n, bk, 0, j0, Y, n, s);
        gm(bk, bk, n, Y, n, A_KC, Y, n, B_KC, 1.0, 0.0, G, bk);          // G = Y^T Y
        larft_from_gram(G, bk, T, s);
    };
    auto left = [&](int64_t bk) {                                        // R := Q R
        gm(bk, n, n, Y, n, A_KC, R, ld, B_NC, 1.0, 0.0, W1, n);           // W1 = Y^T R
        gm(bk, n, bk, T, bk, A_KC, W1, n, B_NC, 1.0, 0.0, W2, n);         // W2 = T W1
        gm(n, n, bk, Y, n, A_MC, W2, n, B_NC, -1.0, 1.0, R, ld);          // R -= Y W2
    };
    auto right = [&](int64_t bk) {                                       // R := R Q^T
        gm(n, bk, n, R, ld, A_KC, Y, n, B_KC, 1.0, 0.0, W1, bk);          // W1 = R Y
        gm(n, bk, bk, W1, bk, A_KC, T, bk, B_KC, 1.0, 0.0, W2, bk);       // W2 = W1 T^T
        gm(n, n, bk, W2, bk, A_KC, Y, n, B_NC, -1.0, 1.0, R, ld);         // R -= W2 Y^T
    };
    const int64_t nblk = ceildiv(n, b);
    for (int64_t k = nblk - 1; k >= 0; --k) {
        const int64_t j0 = k * b, bk = std::min(b, n - j0);
        block(seed, j0, bk);
        left(bk);
        if (sym) right(bk);
    }
    if (!sym) {
        for (int64_t k = nblk - 1; k >= 0; --k) {
            const int64_t j0 = k * b, bk = std::min(b, n - j0);
            block(seed + 1, j0, bk);
            right(bk);
        }
    }
}

} // namespace sa

namespace {
void check_dims(int64_t n, int64_t lda) { SA_REQUIRE(n >= 0 && lda >= std::max<int64_t>(1, n), -3, "bad n/lda"); }
void check_rhs(int64_t n, int64_t nrhs, int64_t ldb) { SA_REQUIRE(nrhs >= 0 && ldb >= std::max<int64_t>(1, n), -4, "bad nrhs/ldb"); }
}

extern "C" {

int slate_amd_init(slate_amd_context* out, int device)
{
    try {
        SA_REQUIRE(out != nullptr, -1, "null output pointer");
        SA_HIP(hipSetDevice(device));
        auto* h = new slate_amd_context_t;
        h->c.device = device;
        SA_HIP(hipStreamCreateWithFlags(&h->c.s0, hipStreamNonBlocking));
        SA_HIP(hipStreamCreateWithFlags(&h->c.s1, hipStreamNonBlocking));
        SA_HIP(hipEventCreateWithFlags(&h->c.ev_in, hipEventDisableTiming));
        SA_HIP(hipEventCreateWithFlags(&h->c.ev_out, hipEventDisableTiming));
        // Panel reserve: the float LU's trailing updates run on s2, whose CU
        // mask leaves `reserve` CUs (32 of 256 by default) to the panel
        // stream.  With split-BF16 updates the panel is the critical path and
        // the reserve shortens the n = 65536 factorization by about a quarter;
        // the FP64 LU is update-bound and keeps every CU (switch 0).
        // SLATE_AMD_PANEL_CUS overrides the reserve (0 = none) and
        // SLATE_AMD_PANEL_SWITCH the trailing size below which both
        // precisions use it.  Measured at n = 65536, nb = 1024, same box:
        // reserve 32 / 48 / 64 / 80 CUs: 1485 / 1560 / 1540 / 1611 ms (with
        // the current leaf; before it was shortened, 64 was best), a smaller
        // reserve for the early, update-bound steps only: slower.  The
        // non-monotonic response suggests the reserve works best in whole
        // XCDs (32 CUs each).
        {
            hipDeviceProp_t prop;
            SA_HIP(hipGetDeviceProperties(&prop, device));
            const int ncu = prop.multiProcessorCount;
            int reserve = ncu / 8;
            if (const char* e = std::getenv("SLATE_AMD_PANEL_CUS")) reserve = std::atoi(e);
            if (reserve > 0 && reserve < ncu) {
                std::vector<uint32_t> mask(size_t((ncu + 31) / 32), 0u);
                for (int cu = reserve; cu < ncu; ++cu) mask[size_t(cu / 32)] |= 1u << (cu % 32);
                SA_HIP(hipExtStreamCreateWithCUMask(&h->c.s2, uint32_t(mask.size()), mask.data()));
                h->c.panel_switch[0] = std::numeric_limits<int64_t>::max();
                h->c.panel_switch[1] = 0;
                if (const char* sw = std::getenv("SLATE_AMD_PANEL_SWITCH"))
                    h->c.panel_switch[0] = h->c.panel_switch[1] = std::atoll(sw);
            }
        }
        SA_HIP(hipEventCreate(&h->c.t_start));
        SA_HIP(hipEventCreate(&h->c.t_stop));
        h->c.scratch = static_cast<unsigned char*>(sa::noalloc_scratch_setup(device));
        *out = h;
        return 0;
    }
    catch (const sa::Error& e) { g_last_error = e.what(); return e.code; }
    catch (const std::exception& e) { g_last_error = e.what(); return -99; }
}

int slate_amd_finalize(slate_amd_context ctx)
{
    if (!ctx) return 0;
    sa::Context& c = ctx->c;
    (void)hipSetDevice(c.device);
    (void)hipDeviceSynchronize();
    for (sa::Buf* b : {&c.bfl, &c.bfu0, &c.bfu1, &c.bfg, &c.f8l, &c.f8u0, &c.f8u1, &c.f8s, &c.tlA, &c.tlB, &c.tlX, &c.tlP, &c.npu, &c.npt, &c.bmeta, &c.bwork, &c.slots, &c.err, &c.dinfo, &c.plan0, &c.plan1, &c.ipiv, &c.work,
                       &c.lo, &c.vec, &c.basis, &c.rhs, &c.plans, &c.perm, &c.flags, &c.tinv, &c.ttmp, &c.sinv, &c.stmp1, &c.dinv,
                       &c.tvw, &c.swaptmp, &c.swaptmp1, &c.ttri, &c.cut0, &c.cut1})
        b->release();
    for (auto e : c.evs) (void)hipEventDestroy(e);
    for (auto e : c.prof_ev) (void)hipEventDestroy(e);
    for (auto& g : c.graphs) (void)hipGraphExecDestroy(g.second);
    if (c.bgraph_exec) (void)hipGraphExecDestroy(c.bgraph_exec);
    if (c.bgraph) (void)hipGraphDestroy(c.bgraph);
    c.bqueue.release();
    (void)hipEventDestroy(c.t_start);
    (void)hipEventDestroy(c.t_stop);
    (void)hipEventDestroy(c.ev_in);
    (void)hipEventDestroy(c.ev_out);
    (void)hipStreamDestroy(c.s0);
    (void)hipStreamDestroy(c.s1);
    if (c.s2) (void)hipStreamDestroy(c.s2);
    delete ctx;
    return 0;
}

const char* slate_amd_last_error(void) { return g_last_error.c_str(); }

void slate_amd_options_init(slate_amd_options* o)
{
    if (!o) return;
    o->block_size = 0;
    o->inner_blocking = 0;
    o->lookahead = 1;
    o->pivot_threshold = 1.0;
    o->max_iterations = 30;
    o->tolerance = 0.0;
    o->use_fallback = 1;
    o->factor_precision = SLATE_AMD_PREC_FP32;
    o->method_lu = SLATE_AMD_LU_PARTIALPIV;
    o->depth = 2;
    o->use_graph = 0;
}

double slate_amd_last_time_ms(slate_amd_context ctx) { return ctx ? ctx->c.last_ms : 0.0; }


int64_t slate_amd_dgetrf(slate_amd_context ctx, int64_t n, double* dA, int64_t lda, int64_t* ipiv,
                         const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        c.ipiv.ensure(sizeof(int64_t) * n);
        sa::Timer t(c);
        int64_t info = sa::getrf_cm<double>(c, dA, n, lda, o, c.ipiv.as<int64_t>());
        t.stop();
        sa::ipiv_to_host(c, c.ipiv.as<int64_t>(), n, ipiv);
        return info;
    });
}

int64_t slate_amd_sgetrf(slate_amd_context ctx, int64_t n, float* dA, int64_t lda, int64_t* ipiv,
                         const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        c.ipiv.ensure(sizeof(int64_t) * n);
        sa::Timer t(c);
        int64_t info = sa::getrf_cm<float>(c, dA, n, lda, o, c.ipiv.as<int64_t>());
        t.stop();
        sa::ipiv_to_host(c, c.ipiv.as<int64_t>(), n, ipiv);
        return info;
    });
}

int64_t slate_amd_dgetrs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs, const double* dA,
                         int64_t lda, const int64_t* ipiv, double* dB, int64_t ldb,
                         const slate_amd_options*)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb);
        if (n == 0 || nrhs == 0) return 0;
        SA_REQUIRE(ipiv != nullptr, -5, "null ipiv");
        SA_ENTRY_SYNC();
        int64_t* dip = sa::ipiv_from_host(c, ipiv, n);
        sa::Timer t(c);
        sa::getrs_cm<double>(c, trans, dA, lda, n, dip, dB, ldb, nrhs);
        t.stop();
        return 0;
    });
}

int64_t slate_amd_dgetrs_noalloc(slate_amd_context ctx, int64_t n, const double* dA, int64_t lda,
                                 const int64_t* ipiv, double* dB)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        if (n == 0) return 0;
        SA_REQUIRE(n <= sa::NOALLOC_NMAX, -3, "dgetrs_noalloc: n above the scratch limit (16384)");
        SA_REQUIRE(ipiv != nullptr && dB != nullptr, -5, "null ipiv or B");
        for (int64_t i = 0; i < n; ++i)
            SA_REQUIRE(ipiv[i] >= i + 1 && ipiv[i] <= n, -5, "ipiv entry out of range");
        SA_REQUIRE(c.scratch != nullptr, -100, "dgetrs_noalloc: no scratch on this device");
        SA_ENTRY_SYNC();
        using L = sa::NoallocLayout;
        unsigned char* sc = c.scratch;
        int64_t* ip1 = reinterpret_cast<int64_t*>(sc + L::ipiv1);
        int64_t* ip0 = reinterpret_cast<int64_t*>(sc + L::ipiv0);
        int64_t* plan = reinterpret_cast<int64_t*>(sc + L::plan);
        int* err = reinterpret_cast<int*>(sc + L::err);
        uint64_t* words = reinterpret_cast<uint64_t*>(sc + L::tvw);
        double* d0 = reinterpret_cast<double*>(sc + L::dinv);
        double* d1 = d0 + L::dinv_per;
        hipStream_t s = c.s0;
        sa::Timer t(c);
        // the previous call's exchange words and error flag are cleared here
        sa::zero_u32(reinterpret_cast<unsigned*>(err), 64, s);
        sa::zero_u32(reinterpret_cast<unsigned*>(words), int64_t(2 * L::tv_words), s);
        SA_HIP(hipMemcpyAsync(ip1, ipiv, sizeof(int64_t) * size_t(n), hipMemcpyHostToDevice, s));
        sa::ipiv_zero_based(ip1, ip0, n, s);
        // B := P B (the LAPACK interchange sequence, 1024 pivots per plan), then
        // the unit-lower and upper look-back solves in place
        for (int64_t k1 = 0; k1 < n; k1 += 1024)
            sa::laswp<double>(dB, 1, n, 0, 1, k1, std::min(n, k1 + 1024), ip0, plan, s);
        sa::trsv_diag_inverse<double>(dA, lda, 1, 1, 1, n, d0, s);
        sa::trsv_diag_inverse<double>(dA, lda, 1, 0, 0, n, d1, s);
        sa::TvSync tv{words, words + (L::tv_words - 16), err, 1u};
        sa::trsv_lookback<double>(dA, lda, 1, 1, n, d0, dB, tv, s);
        tv.tag = 2u;
        sa::trsv_lookback<double>(dA, lda, 1, 0, n, d1, dB, tv, s);
        int e = 0;
        SA_HIP(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
        SA_HIP(hipStreamSynchronize(s));
        t.stop();
        SA_REQUIRE(e == 0, -102, "triangular solve: completion chain timed out");
        return 0;
    });
}

int64_t slate_amd_sgetrs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs, const float* dA,
                         int64_t lda, const int64_t* ipiv, float* dB, int64_t ldb,
                         const slate_amd_options*)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb);
        if (n == 0 || nrhs == 0) return 0;
        SA_REQUIRE(ipiv != nullptr, -5, "null ipiv");
        SA_ENTRY_SYNC();
        int64_t* dip = sa::ipiv_from_host(c, ipiv, n);
        sa::Timer t(c);
        sa::getrs_cm<float>(c, trans, dA, lda, n, dip, dB, ldb, nrhs);
        t.stop();
        return 0;
    });
}

int64_t slate_amd_dgesv(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                        int64_t* ipiv, double* dB, int64_t ldb, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        c.ipiv.ensure(sizeof(int64_t) * n);
        sa::Timer t(c);
        int64_t info = sa::getrf_cm<double>(c, dA, n, lda, o, c.ipiv.as<int64_t>());
        if (info == 0 && nrhs > 0)
            sa::getrs_cm<double>(c, 0, dA, lda, n, c.ipiv.as<int64_t>(), dB, ldb, nrhs);
        t.stop();
        sa::ipiv_to_host(c, c.ipiv.as<int64_t>(), n, ipiv);
        return info;
    });
}

int64_t slate_amd_dpotrf(slate_amd_context ctx, int uplo, int64_t n, double* dA, int64_t lda,
                         const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::potrf_cm<double>(c, uplo, dA, n, lda, o);
        t.stop();
        return info;
    });
}

int64_t slate_amd_spotrf(slate_amd_context ctx, int uplo, int64_t n, float* dA, int64_t lda,
                         const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::potrf_cm<float>(c, uplo, dA, n, lda, o);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dpotrs(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, const double* dA,
                         int64_t lda, double* dB, int64_t ldb, const slate_amd_options*)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb);
        if (n == 0 || nrhs == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Timer t(c);
        sa::potrs_cm<double>(c, uplo, sa::as_mut(dA), lda, n, dB, ldb, nrhs);
        t.stop();
        return 0;
    });
}

int64_t slate_amd_dposv(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, double* dA,
                        int64_t lda, double* dB, int64_t ldb, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb);
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::potrf_cm<double>(c, uplo, dA, n, lda, o);
        if (info == 0 && nrhs > 0) sa::potrs_cm<double>(c, uplo, dA, lda, n, dB, ldb, nrhs);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dsgesv(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                         int64_t* ipiv, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                         int* iter, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb); check_rhs(n, nrhs, ldx);
        SA_REQUIRE(iter != nullptr, -6, "null iter");
        if (n == 0 || nrhs == 0) { *iter = 0; return 0; }
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::mixed_ir(c, 0, SLATE_AMD_LOWER, n, nrhs, dA, lda, ipiv, dB, ldb, dX, ldx, iter, o);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dsgesv_gmres(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                               int64_t* ipiv, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                               int* iter, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb); check_rhs(n, nrhs, ldx);
        SA_REQUIRE(iter != nullptr, -6, "null iter");
        // src/gesv_mixed_gmres.cc:144-146
        SA_REQUIRE(nrhs == 1, -7, "block-GMRES for multiple RHS is not yet supported");
        if (n == 0) { *iter = 0; return 0; }
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        (void)ldb; (void)ldx;
        int64_t info = sa::mixed_gmres(c, 0, SLATE_AMD_LOWER, n, dA, lda, ipiv, dB, dX, iter, o, o.nb);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dsposv(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, double* dA,
                         int64_t lda, const double* dB, int64_t ldb, double* dX, int64_t ldx, int* iter,
                         const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb); check_rhs(n, nrhs, ldx);
        SA_REQUIRE(iter != nullptr, -6, "null iter");
        if (n == 0 || nrhs == 0) { *iter = 0; return 0; }
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::mixed_ir(c, 1, uplo, n, nrhs, dA, lda, nullptr, dB, ldb, dX, ldx, iter, o);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dsposv_gmres(slate_amd_context ctx, int uplo, int64_t n, int64_t nrhs, double* dA,
                               int64_t lda, const double* dB, int64_t ldb, double* dX, int64_t ldx,
                               int* iter, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb); check_rhs(n, nrhs, ldx);
        SA_REQUIRE(iter != nullptr, -6, "null iter");
        SA_REQUIRE(nrhs == 1, -7, "block-GMRES for multiple RHS is not yet supported");
        if (n == 0) { *iter = 0; return 0; }
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        (void)ldb; (void)ldx;
        int64_t info = sa::mixed_gmres(c, 1, uplo, n, dA, lda, nullptr, dB, dX, iter, o, o.nb);
        t.stop();
        return info;
    });
}

int64_t slate_amd_dgesv_rbt(slate_amd_context ctx, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                            const double* dB, int64_t ldb, double* dX, int64_t ldx, int* iter,
                            const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda); check_rhs(n, nrhs, ldb); check_rhs(n, nrhs, ldx);
        SA_REQUIRE(iter != nullptr, -6, "null iter");
        if (n == 0 || nrhs == 0) { *iter = 0; return 0; }
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        sa::Timer t(c);
        int64_t info = sa::gesv_rbt(c, n, nrhs, dA, lda, dB, ldb, dX, ldx, iter, o);
        t.stop();
        return info;
    });
}

int slate_amd_drbt_generate(slate_amd_context ctx, int64_t n, int64_t seed, double* du)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(n >= 0, -2, "bad n");
        SA_ENTRY_SYNC();
        sa::rbt_weights(seed, n, du, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int slate_amd_dgerbt(slate_amd_context ctx, int64_t n, int depth, int64_t nb, const double* du,
                     const double* dv, double* dA, int64_t lda)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_ENTRY_SYNC();
        sa::gerbt2<double>(dA, lda, n, nb, depth, du, dv, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int slate_amd_dgerbt_rhs(slate_amd_context ctx, int trans, int64_t n, int64_t nrhs, int depth, int64_t nb,
                         const double* du, double* dB, int64_t ldb)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_rhs(n, nrhs, ldb);
        SA_REQUIRE(trans == SLATE_AMD_NOTRANS || trans == SLATE_AMD_TRANS, -2, "bad trans");
        SA_ENTRY_SYNC();
        sa::gerbt1<double>(trans, dB, ldb, n, nrhs, nb, depth, du, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int slate_amd_dgemm(slate_amd_context ctx, int64_t m, int64_t n, int64_t k, double alpha,
                    const double* dA, int64_t lda, const double* dB, int64_t ldb, double beta,
                    double* dC, int64_t ldc)
{ return gemm_entry<double>(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc); }

int slate_amd_sgemm(slate_amd_context ctx, int64_t m, int64_t n, int64_t k, float alpha,
                    const float* dA, int64_t lda, const float* dB, int64_t ldb, float beta,
                    float* dC, int64_t ldc)
{ return gemm_entry<float>(ctx, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc); }

int slate_amd_sgemm_bf16(slate_amd_context ctx, int terms, int64_t m, int64_t n, int64_t k, float alpha,
                         const float* dA, int64_t lda, const float* dB, int64_t ldb, float* dC, int64_t ldc)
{ return gemm_bf16_entry(ctx, terms, m, n, k, alpha, dA, lda, dB, ldb, dC, ldc); }

int slate_amd_graph_stats(slate_amd_context ctx, int64_t* captured, int64_t* failed, int64_t* replays)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        if (captured) *captured = int64_t(c.graphs.size());
        if (failed) *failed = int64_t(c.graph_failed.size());
        if (replays) *replays = c.graph_replays;
        return 0;
    }));
}

int slate_amd_profile_enable(slate_amd_context ctx, int on)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        c.prof = on != 0;
        c.prof_ms = 0; c.prof_flops = 0; c.prof_launches = 0;
        c.prof_used = 0; c.prof_flops_pending.clear(); c.prof_kind_pending.clear();
        for (int k = 0; k < 3; ++k) { c.prof_ms_k[k] = 0; c.prof_flops_k[k] = 0; c.prof_launches_k[k] = 0; }
        return 0;
    }));
}

int slate_amd_profile_read(slate_amd_context ctx, double* gemm_ms, double* gemm_flops, int64_t* launches)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        if (gemm_ms) *gemm_ms = c.prof_ms;
        if (gemm_flops) *gemm_flops = c.prof_flops;
        if (launches) *launches = c.prof_launches;
        return 0;
    }));
}

int slate_amd_profile_read_kernel(slate_amd_context ctx, int kernel, double* gemm_ms, double* gemm_flops,
                                  int64_t* launches)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(kernel >= 0 && kernel < 3, -2, "profile_read_kernel: unknown kernel");
        if (gemm_ms) *gemm_ms = c.prof_ms_k[kernel];
        if (gemm_flops) *gemm_flops = c.prof_flops_k[kernel];
        if (launches) *launches = c.prof_launches_k[kernel];
        return 0;
    }));
}

int slate_amd_dnorm_inf(slate_amd_context ctx, int sym, int64_t n, const double* dA, int64_t lda,
                        double* result)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_REQUIRE(result != nullptr, -5, "null result");
        if (n == 0) { *result = 0; return 0; }
        SA_ENTRY_SYNC();
        int64_t we = sa::ceildiv(n, 1024) * n + n + 64;
        c.work.ensure(sizeof(double) * (we + 8));
        double* w = c.work.as<double>();
        *result = sa::norm_host(c, dA, n, lda, sym, w, we, w + we);
        return 0;
    }));
}

int slate_amd_tune_set_nb(slate_amd_context ctx, int kind, int count, const int64_t* n_upper,
                          const int64_t* nb)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(kind >= 0 && kind < 4, -2, "unknown tuning kind");
        SA_REQUIRE(count >= 0 && (count == 0 || (n_upper && nb)), -3, "bad table");
        std::vector<std::pair<int64_t, int64_t>> t;
        for (int i = 0; i < count; ++i) {
            SA_REQUIRE(nb[i] >= 16 && nb[i] <= 4096, -5, "nb must be in [16, 4096]");
            SA_REQUIRE(i == 0 || n_upper[i] > n_upper[i - 1], -4, "n_upper must increase");
            t.emplace_back(n_upper[i], nb[i]);
        }
        c.nb_table[kind] = t;
        return 0;
    }));
}

int64_t slate_amd_tune_get_nb(slate_amd_context ctx, int kind, int64_t n)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(kind >= 0 && kind < 4, -2, "unknown tuning kind");
        SA_REQUIRE(n >= 0, -3, "bad n");
        return sa::choose_nb(c, kind, n, 0);
    });
}

// Learned selection: features, MLP forward / training step, and the solve
// driven by it (no counterpart in the reference).
int slate_amd_mlp_num_params(void) { return sa::mlp_num_params(); }

int64_t slate_amd_selector_features(slate_amd_context ctx, int64_t n, const double* dA, int64_t lda,
                                    const double* dB, double* dX, const slate_amd_options* opts, float* features)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_REQUIRE(features != nullptr, -8, "null features");
        SA_REQUIRE(dB != nullptr && dX != nullptr, -5, "null right-hand side or solution vector");
        SA_REQUIRE(n >= 2, -3, "selector features need n >= 2");
        SA_ENTRY_SYNC();
        sa::Opts o = sa::read_opts(opts);
        o.prec = sa::SEL_PROBE;                       // the features the selector sees
        sa::Timer t(c);
        sa::MixedState st = sa::mixed_setup(c, n, 1);
        int64_t info = 0;
        sa::selector_features(c, st, n, dA, lda, dB, dX, o, &info, features, nullptr);
        t.stop();
        sa::check_kernel_error(c);
        return info;
    });
}

int slate_amd_mlp_forward(slate_amd_context ctx, const float* dparams, const float* dX, int batch, float* dprobs)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(batch >= 0, -4, "bad batch");
        SA_ENTRY_SYNC();
        sa::mlp_forward(dparams, dX, batch, dprobs, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int slate_amd_mlp_train_step(slate_amd_context ctx, float* dparams, float* dgrad, float* dm, float* dv,
                             const float* dX, const int* dy, int batch, float lr, int step, float* loss)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(batch >= 1 && batch <= 4096, -8, "batch must be 1..4096");
        SA_REQUIRE(step >= 1, -10, "step counts from 1");
        SA_ENTRY_SYNC();
        const size_t per = size_t(sa::mlp_scratch_per_sample());
        c.work.ensure(sizeof(float) * size_t(batch) * per + sizeof(float) * 64);
        float* scratch = c.work.as<float>();
        float* dloss = scratch + size_t(batch) * per;
        sa::mlp_train(dparams, dgrad, dm, dv, dX, dy, batch, scratch, lr, step, dloss, c.s0);
        if (loss) SA_HIP(hipMemcpyAsync(loss, dloss, sizeof(float), hipMemcpyDeviceToHost, c.s0));
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int64_t slate_amd_dsgesv_auto(slate_amd_context ctx, int64_t n, double* dA, int64_t lda, int64_t* ipiv,
                              const double* dB, double* dX, const float* params, int force, int* iter,
                              int* choice, float* features, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_REQUIRE(n >= 2, -2, "n >= 2 required");
        SA_REQUIRE(ipiv && iter && choice, -5, "null output");
        SA_ENTRY_SYNC();
        sa::Timer t(c);
        sa::Opts o = sa::read_opts(opts);
        int64_t info = sa::mixed_select(c, n, dA, lda, ipiv, dB, dX, iter, choice, params, force, features, o);
        t.stop();
        return info;
    });
}

// Batched uneven solves (no counterpart in the reference; SLATE solves one
// distributed system per call).  Shared by the fixed-precision and the
// learned entries: per-system modes as in batched.hip.
static int64_t batched_run(sa::Context& c, int count, const int* n, const double* dA, const double* dB, double* dX,
                           const int* modes, const float* sel, int* iters, int* info, double* berr, int* choice,
                           float* features, double* ms, const slate_amd_options* opts)
{
    const int maxn = sa::batched_max_n();
    std::vector<int64_t> offA(count), offB(count);
    std::vector<int> order(count);
    int64_t a = 0, b = 0;
    for (int i = 0; i < count; ++i) {
        SA_REQUIRE(n[i] >= 1 && n[i] <= maxn, -3, "system sizes must be in [1, 1024]");
        offA[i] = a; offB[i] = b;
        a += int64_t(n[i]) * n[i];
        b += n[i];
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
    // use_graph means the batched launch's own one-node graph here (the LU
    // step-sequence graph that read_opts guards is a different path)
    slate_amd_options oc;
    if (opts) oc = *opts;
    else slate_amd_options_init(&oc);
    const bool use_graph = oc.use_graph != 0;
    oc.use_graph = 0;
    sa::Opts o = sa::read_opts(&oc);
    SA_ENTRY_SYNC();
    sa::Timer t(c);
    // metadata: offA, offB, ticks (8 bytes) | berr (double) | n, mode, order, iters, info, choice (int) |
    //           features (float) | selector parameters (float)
    const int nsel = sel ? sa::mlp_num_params() : 0;
    const size_t meta = (sizeof(int64_t) * 3 + sizeof(double)) * size_t(count) + sizeof(int) * 6 * size_t(count)
                        + sizeof(float) * (24 * size_t(count) + nsel) + 1024;
    c.bmeta.ensure(meta);
    char* p = static_cast<char*>(c.bmeta.p);
    int64_t* d_offA = reinterpret_cast<int64_t*>(p);
    int64_t* d_offB = d_offA + count;
    unsigned long long* d_ticks = reinterpret_cast<unsigned long long*>(d_offB + count);
    double* d_berr = reinterpret_cast<double*>(d_ticks + count);
    int* d_n = reinterpret_cast<int*>(d_berr + count);
    int* d_mode = d_n + count;
    int* d_order = d_mode + count;
    int* d_iters = d_order + count;
    int* d_info = d_iters + count;
    int* d_choice = d_info + count;
    float* d_feat = reinterpret_cast<float*>(d_choice + count);
    float* d_sel = d_feat + 24 * size_t(count);
    if (!c.bqueue.p) {
        c.bqueue.ensure(256);
        SA_HIP(hipMemsetAsync(c.bqueue.p, 0, 256, c.s0));
    }
    SA_HIP(hipMemcpyAsync(d_offA, offA.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipMemcpyAsync(d_offB, offB.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipMemcpyAsync(d_n, n, sizeof(int) * count, hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipMemcpyAsync(d_mode, modes, sizeof(int) * count, hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * count, hipMemcpyHostToDevice, c.s0));
    if (nsel) SA_HIP(hipMemcpyAsync(d_sel, sel, sizeof(float) * nsel, hipMemcpyHostToDevice, c.s0));
    int cus = 0;
    SA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    const int grid = std::max(1, std::min(count, cus * c.b_wgs));
    c.bwork.ensure(size_t(grid) * size_t(sa::batched_work_stride()));
    alignas(16) char args[sa::BATCH_ARGS_BYTES] = {};
    sa::batched_fill_args(args, count, d_order, d_n, d_mode, d_offA, d_offB, dA, dB, dX, d_iters, d_info, d_berr,
                          d_choice, features ? d_feat : nullptr, ms ? d_ticks : nullptr, nsel ? d_sel : nullptr,
                          c.bqueue.as<int>(), static_cast<char*>(c.bwork.p), int(o.itermax), o.tol,
                          oc.factor_precision == SLATE_AMD_PREC_BF16 ? 1 : 0);
    if (use_graph) {
        const int cfg[3] = {c.b_nbf, c.b_bt, grid};
        if (!c.bgraph_exec || memcmp(args, c.bgraph_args, sizeof(args)) != 0
            || memcmp(cfg, c.bgraph_cfg, sizeof(cfg)) != 0) {
            if (c.bgraph_exec) { SA_HIP(hipGraphExecDestroy(c.bgraph_exec)); c.bgraph_exec = nullptr; }
            if (c.bgraph) { SA_HIP(hipGraphDestroy(c.bgraph)); c.bgraph = nullptr; }
            memcpy(c.bgraph_args, args, sizeof(args));
            memcpy(c.bgraph_cfg, cfg, sizeof(cfg));
            sa::batched_graph(c.bgraph_args, c.b_nbf, c.b_bt, grid, &c.bgraph, &c.bgraph_exec);
            ++c.bgraph_builds;
        }
        else {
            ++c.bgraph_replays;
        }
        SA_HIP(hipGraphLaunch(c.bgraph_exec, c.s0));
    }
    else {
        sa::batched_launch(args, c.b_nbf, c.b_bt, grid, c.s0);
    }
    SA_HIP(hipMemcpyAsync(iters, d_iters, sizeof(int) * count, hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipMemcpyAsync(info, d_info, sizeof(int) * count, hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipMemcpyAsync(berr, d_berr, sizeof(double) * count, hipMemcpyDeviceToHost, c.s0));
    if (choice) SA_HIP(hipMemcpyAsync(choice, d_choice, sizeof(int) * count, hipMemcpyDeviceToHost, c.s0));
    if (features)
        SA_HIP(hipMemcpyAsync(features, d_feat, sizeof(float) * 24 * size_t(count), hipMemcpyDeviceToHost, c.s0));
    std::vector<unsigned long long> ticks(ms ? count : 0);
    if (ms) SA_HIP(hipMemcpyAsync(ticks.data(), d_ticks, sizeof(unsigned long long) * count, hipMemcpyDeviceToHost, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
    if (ms) {
        int khz = 0;
        SA_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device));
        SA_REQUIRE(khz > 0, -100, "wall clock rate unavailable");
        for (int i = 0; i < count; ++i) ms[i] = double(ticks[i]) / double(khz);
    }
    t.stop();
    return 0;
}

int64_t slate_amd_dgesv_batched(slate_amd_context ctx, int count, const int* n, const double* dA,
                                const double* dB, double* dX, const int* mode, int factor_precision,
                                int* iters, int* info, double* berr, const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(count >= 0, -2, "bad count");
        if (count == 0) return 0;
        SA_REQUIRE(n && dA && dB && dX && iters && info && berr, -3, "null argument");
        SA_REQUIRE(factor_precision == SLATE_AMD_PREC_FP32 || factor_precision == SLATE_AMD_PREC_FP64
                   || factor_precision == SLATE_AMD_PREC_BF16, -8, "factor_precision must be FP32, BF16 or FP64");
        const int base = factor_precision == SLATE_AMD_PREC_FP64 ? 1 : factor_precision == SLATE_AMD_PREC_BF16 ? 2 : 0;
        std::vector<int> md(count, base);
        for (int i = 0; i < count && mode; ++i) {
            SA_REQUIRE(mode[i] == 0 || mode[i] == 1, -7, "mode must be 0 or 1");
            if (mode[i] == 1) md[i] = 1;
        }
        return batched_run(c, count, n, dA, dB, dX, md.data(), nullptr, iters, info, berr, nullptr, nullptr,
                           nullptr, opts);
    });
}

int64_t slate_amd_dgesv_batched_auto(slate_amd_context ctx, int count, const int* n, const double* dA,
                                     const double* dB, double* dX, const float* params, int force, int* iters,
                                     int* info, double* berr, int* choice, float* features, double* ms,
                                     const slate_amd_options* opts)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(count >= 0, -2, "bad count");
        if (count == 0) return 0;
        SA_REQUIRE(n && dA && dB && dX && iters && info && berr, -3, "null argument");
        SA_REQUIRE(force >= -1 && force <= 2, -7, "force must be -1 (network) or a choice 0..2");
        SA_REQUIRE(force >= 0 || params, -7, "selector parameters required unless a choice is forced");
        SA_REQUIRE(!opts || opts->factor_precision == SLATE_AMD_PREC_FP32 || opts->factor_precision == SLATE_AMD_PREC_BF16,
                   -8, "batched learned path: the probe factor_precision must be FP32 or BF16");
        std::vector<int> md(count, force < 0 ? 3 : 4 + force);
        return batched_run(c, count, n, dA, dB, dX, md.data(), force < 0 ? params : nullptr, iters, info, berr,
                           choice, features, ms, opts);
    });
}

int slate_amd_tune_set_batched(slate_amd_context ctx, int nbf, int bt, int wgs_per_cu)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(sa::batched_config_valid(nbf, bt), -2, "batched tile configuration: nbf 16/32, bt 256/512");
        SA_REQUIRE(wgs_per_cu >= 1 && wgs_per_cu <= sa::batched_max_wgs_per_cu(nbf, bt), -4,
                   "batched: workgroups per CU above what LDS and registers allow");
        c.b_nbf = nbf;
        c.b_bt = bt;
        c.b_wgs = wgs_per_cu;
        return 0;
    }));
}

int slate_amd_tune_get_batched(slate_amd_context ctx, int* nbf, int* bt, int* wgs_per_cu, int* max_wgs_per_cu)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        if (nbf) *nbf = c.b_nbf;
        if (bt) *bt = c.b_bt;
        if (wgs_per_cu) *wgs_per_cu = c.b_wgs;
        if (max_wgs_per_cu) *max_wgs_per_cu = sa::batched_max_wgs_per_cu(c.b_nbf, c.b_bt);
        return 0;
    }));
}

int slate_amd_batched_graph_stats(slate_amd_context ctx, int64_t* builds, int64_t* replays)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        if (builds) *builds = c.bgraph_builds;
        if (replays) *replays = c.bgraph_replays;
        return 0;
    }));
}

// slate::gecondest (src/gecondest.cc): reciprocal condition number in the 1
// (norm = 'O' or '1') or infinity (norm = 'I') norm from the LU factors.
int64_t slate_amd_dgecondest(slate_amd_context ctx, int norm, int64_t n, const double* dLU, int64_t lda,
                             const int64_t* ipiv, double anorm, double* rcond)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_REQUIRE(norm == '1' || norm == 'O' || norm == 'o' || norm == 'I' || norm == 'i', -2, "invalid norm");
        SA_REQUIRE(anorm >= 0 && !std::isnan(anorm), -7, "invalid Anorm");
        SA_REQUIRE(rcond != nullptr, -8, "null rcond");
        if (n <= 1) { *rcond = 1.0; return 0; }
        if (anorm == 0.0) { *rcond = 0.0; return 0; }
        SA_REQUIRE(ipiv != nullptr, -6, "null ipiv");
        SA_ENTRY_SYNC();
        int64_t* dip = sa::ipiv_from_host(c, ipiv, n);
        c.basis.ensure(sizeof(double) * size_t(n));
        const bool inf = (norm == 'I' || norm == 'i');
        // ||A^{-1}||_inf = ||A^{-T}||_1: swap the two solves
        double ainv = sa::norm1est_inv(n, c.basis.as<double>(), [&](double* x, int t) {
            const int tr = (t == 1) != inf ? SLATE_AMD_TRANS : SLATE_AMD_NOTRANS;
            sa::getrs_cm<double>(c, tr, dLU, lda, n, dip, x, n, 1);
        });
        *rcond = ainv != 0.0 ? (1.0 / ainv) / anorm : 0.0;
        return 0;
    });
}

// slate::pocondest (src/pocondest.cc): reciprocal condition number from the
// Cholesky factor (A symmetric, so both norms coincide).
int64_t slate_amd_dpocondest(slate_amd_context ctx, int uplo, int64_t n, const double* dL, int64_t lda,
                             double anorm, double* rcond)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        check_dims(n, lda);
        SA_REQUIRE(uplo == SLATE_AMD_LOWER || uplo == SLATE_AMD_UPPER, -2, "invalid uplo");
        SA_REQUIRE(anorm >= 0 && !std::isnan(anorm), -6, "invalid Anorm");
        SA_REQUIRE(rcond != nullptr, -7, "null rcond");
        if (n <= 1) { *rcond = 1.0; return 0; }
        if (anorm == 0.0) { *rcond = 0.0; return 0; }
        SA_ENTRY_SYNC();
        c.basis.ensure(sizeof(double) * size_t(n));
        double ainv = sa::norm1est_inv(n, c.basis.as<double>(), [&](double* x, int) {
            sa::potrs_cm<double>(c, uplo, const_cast<double*>(dL), lda, n, x, n, 1);
        });
        *rcond = ainv != 0.0 ? (1.0 / ainv) / anorm : 0.0;
        return 0;
    });
}

int slate_amd_generate_svd(slate_amd_context ctx, int sigma_dist, double cond, int64_t seed, int sym,
                           int64_t n, double* dA, int64_t lda)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(n >= 0 && lda >= std::max<int64_t>(1, n), -3, "bad n/lda");
        SA_REQUIRE(sigma_dist >= 1 && sigma_dist <= 4, -2, "unknown sigma distribution");
        SA_REQUIRE(cond >= 1.0, -2, "cond must be >= 1");
        if (n == 0) return 0;
        SA_ENTRY_SYNC();
        sa::generate_svd(c, sigma_dist, cond, seed, sym != 0, n, dA, lda);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

int slate_amd_generate(slate_amd_context ctx, int dist, int64_t seed, int64_t m, int64_t n,
                       int64_t ioffset, int64_t joffset, double* dA, int64_t lda)
{
    return int(guarded(ctx, [&](sa::Context& c) -> int64_t {
        SA_REQUIRE(m >= 0 && n >= 0 && lda >= std::max<int64_t>(1, m), -3, "bad m/n/lda");
        SA_REQUIRE(dist >= 1 && dist <= 3, -2, "unknown distribution");
        SA_ENTRY_SYNC();
        sa::generate(dist, seed, m, n, ioffset, joffset, dA, lda, c.s0);
        SA_HIP(hipStreamSynchronize(c.s0));
        return 0;
    }));
}

} // extern "C"

//==============================================================================
// SLATE tile storage (slate_amd_tiles): the dense entries on gathered copies.
//==============================================================================
namespace {

struct TileView {
    const slate_amd_tiles* d;
    int64_t mt, nt;
};

TileView check_tiles(const slate_amd_tiles* t, bool lower_only, const char* what)
{
    SA_REQUIRE(t && t->tiles && t->lds, -2, std::string(what) + ": null tile descriptor");
    SA_REQUIRE(t->m >= 0 && t->n >= 0 && t->mb >= 1 && t->nb >= 1, -2, std::string(what) + ": bad tile sizes");
    TileView v{t, sa::ceildiv(t->m, t->mb), sa::ceildiv(t->n, t->nb)};
    for (int64_t j = 0; j < v.nt; ++j)
        for (int64_t i = 0; i < v.mt; ++i) {
            const int64_t k = i + j * v.mt;
            const bool needed = !lower_only || i >= j;
            if (!needed) continue;
            SA_REQUIRE(t->tiles[k] != nullptr, -2, std::string(what) + ": missing tile");
            const int64_t rows = std::min(t->mb, t->m - i * t->mb);
            SA_REQUIRE(t->lds[k] >= std::max<int64_t>(1, rows), -2, std::string(what) + ": tile ld below its rows");
        }
    return v;
}

// Device copies of the pointer table (NULL for unread tiles) and the lds,
// `slot` selects one of three table areas (A, B, X).
void upload_tables(sa::Context& c, const TileView& v, bool lower_only, int slot, double*** dptr, int64_t** dld)
{
    const int64_t cnt = v.mt * v.nt;
    static thread_local std::vector<void*> hp;
    hp.assign(size_t(cnt), nullptr);
    for (int64_t j = 0; j < v.nt; ++j)
        for (int64_t i = 0; i < v.mt; ++i)
            if (!lower_only || i >= j) hp[size_t(i + j * v.mt)] = v.d->tiles[i + j * v.mt];
    // three equal areas sized for the largest table seen
    const size_t area = (sizeof(void*) + sizeof(int64_t)) * size_t(cnt) + 256;
    if (c.tlP.bytes < 3 * area) { c.tlP.release(); c.tlP.ensure(3 * area); }
    const size_t a = c.tlP.bytes / 3 / 256 * 256;
    char* base = static_cast<char*>(c.tlP.p) + size_t(slot) * a;
    *dptr = reinterpret_cast<double**>(base);
    *dld = reinterpret_cast<int64_t*>(base + sizeof(void*) * size_t(cnt));
    SA_HIP(hipMemcpyAsync(*dptr, hp.data(), sizeof(void*) * size_t(cnt), hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipMemcpyAsync(*dld, v.d->lds, sizeof(int64_t) * size_t(cnt), hipMemcpyHostToDevice, c.s0));
    SA_HIP(hipStreamSynchronize(c.s0));
}

// Gather (to_tiles = 0) or scatter (1) between a tiled matrix and a dense
// workspace buffer of ld = m.
double* tiles_dense(sa::Context& c, const TileView& v, sa::Buf& buf, bool lower_only, int slot, int to_tiles)
{
    double** dp = nullptr;
    int64_t* dl = nullptr;
    upload_tables(c, v, lower_only, slot, &dp, &dl);
    const int64_t m = v.d->m, n = v.d->n;
    if (!to_tiles) buf.ensure(sizeof(double) * size_t(std::max<int64_t>(m, 1)) * size_t(std::max<int64_t>(n, 1)));
    sa::tiles_copy<double>(dp, dl, m, n, v.d->mb, v.d->nb, buf.as<double>(), std::max<int64_t>(m, 1), to_tiles, c.s0);
    SA_HIP(hipStreamSynchronize(c.s0));
    return buf.as<double>();
}

// Run `dense(A, lda, B, ldb, X, ldx)` on gathered copies; scatter A back when
// scatter_a(rc), B (when scatter_b) and X when given.  slate_amd_last_time_ms
// reports the dense call alone (without the gather and scatter).
template <class F, class G>
int64_t with_tiles(slate_amd_context ctx, const slate_amd_tiles* A, bool lower, const slate_amd_tiles* B,
                   const slate_amd_tiles* X, bool scatter_b, F&& dense, G&& scatter_a, const char* what)
{
    return guarded(ctx, [&](sa::Context& c) -> int64_t {
        const TileView va = check_tiles(A, lower, what);
        SA_REQUIRE(A->m == A->n, -2, std::string(what) + ": A must be square");
        TileView vb{}, vx{};
        if (B) { vb = check_tiles(B, false, what); SA_REQUIRE(B->m == A->n, -2, std::string(what) + ": B rows != n"); }
        if (X) {
            vx = check_tiles(X, false, what);
            SA_REQUIRE(X->m == A->n && B && X->n == B->n, -2, std::string(what) + ": X shape != B shape");
        }
        SA_ENTRY_SYNC();
        const int64_t n = A->n, ld = std::max<int64_t>(n, 1);
        double* dA = tiles_dense(c, va, c.tlA, lower, 0, 0);
        double* dB = B ? tiles_dense(c, vb, c.tlB, false, 1, 0) : nullptr;
        double* dX = nullptr;
        if (X) { c.tlX.ensure(sizeof(double) * size_t(ld) * size_t(std::max<int64_t>(X->n, 1))); dX = c.tlX.as<double>(); }
        const int64_t rc = dense(dA, ld, dB, ld, dX, ld);
        if (rc < 0) return rc;
        if (scatter_a(rc)) tiles_dense(c, va, c.tlA, lower, 0, 1);
        if (B && scatter_b) tiles_dense(c, vb, c.tlB, false, 1, 1);
        if (X) tiles_dense(c, vx, c.tlX, false, 2, 1);
        return rc;
    });
}

} // namespace

int64_t slate_amd_dgetrf_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                               const slate_amd_options* opts)
{
    return with_tiles(ctx, A, false, nullptr, nullptr, false,
        [&](double* a, int64_t lda, double*, int64_t, double*, int64_t) {
            return slate_amd_dgetrf(ctx, A->n, a, lda, ipiv, opts); },
        [](int64_t) { return true; }, "dgetrf_tiles");
}

int64_t slate_amd_dgetrs_tiles(slate_amd_context ctx, int trans, const slate_amd_tiles* A, const int64_t* ipiv,
                               const slate_amd_tiles* B, const slate_amd_options* opts)
{
    return with_tiles(ctx, A, false, B, nullptr, true,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double*, int64_t) {
            return slate_amd_dgetrs(ctx, trans, A->n, B->n, a, lda, ipiv, b, ldb, opts); },
        [](int64_t) { return false; }, "dgetrs_tiles");
}

int64_t slate_amd_dgesv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                              const slate_amd_tiles* B, const slate_amd_options* opts)
{
    return with_tiles(ctx, A, false, B, nullptr, true,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double*, int64_t) {
            return slate_amd_dgesv(ctx, A->n, B->n, a, lda, ipiv, b, ldb, opts); },
        [](int64_t) { return true; }, "dgesv_tiles");
}

int64_t slate_amd_dpotrf_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_options* opts)
{
    return with_tiles(ctx, A, true, nullptr, nullptr, false,
        [&](double* a, int64_t lda, double*, int64_t, double*, int64_t) {
            return slate_amd_dpotrf(ctx, SLATE_AMD_LOWER, A->n, a, lda, opts); },
        [](int64_t) { return true; }, "dpotrf_tiles");
}

int64_t slate_amd_dpotrs_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                               const slate_amd_options* opts)
{
    return with_tiles(ctx, A, true, B, nullptr, true,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double*, int64_t) {
            return slate_amd_dpotrs(ctx, SLATE_AMD_LOWER, A->n, B->n, a, lda, b, ldb, opts); },
        [](int64_t) { return false; }, "dpotrs_tiles");
}

int64_t slate_amd_dposv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                              const slate_amd_options* opts)
{
    return with_tiles(ctx, A, true, B, nullptr, true,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double*, int64_t) {
            return slate_amd_dposv(ctx, SLATE_AMD_LOWER, A->n, B->n, a, lda, b, ldb, opts); },
        [](int64_t) { return true; }, "dposv_tiles");
}

// Mixed drivers: A is written back only when the FP64 fallback factored it
// (*iter < 0), as the reference leaves A untouched on success.
int64_t slate_amd_dsgesv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                               const slate_amd_tiles* B, const slate_amd_tiles* X, int* iter,
                               const slate_amd_options* opts)
{
    return with_tiles(ctx, A, false, B, X, false,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double* x, int64_t ldx) {
            return slate_amd_dsgesv(ctx, A->n, B->n, a, lda, ipiv, b, ldb, x, ldx, iter, opts); },
        [&](int64_t) { return iter && *iter < 0; }, "dsgesv_tiles");
}

int64_t slate_amd_dsgesv_gmres_tiles(slate_amd_context ctx, const slate_amd_tiles* A, int64_t* ipiv,
                                     const slate_amd_tiles* B, const slate_amd_tiles* X, int* iter,
                                     const slate_amd_options* opts)
{
    return with_tiles(ctx, A, false, B, X, false,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double* x, int64_t ldx) {
            return slate_amd_dsgesv_gmres(ctx, A->n, B->n, a, lda, ipiv, b, ldb, x, ldx, iter, opts); },
        [&](int64_t) { return iter && *iter < 0; }, "dsgesv_gmres_tiles");
}

int64_t slate_amd_dsposv_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                               const slate_amd_tiles* X, int* iter, const slate_amd_options* opts)
{
    return with_tiles(ctx, A, true, B, X, false,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double* x, int64_t ldx) {
            return slate_amd_dsposv(ctx, SLATE_AMD_LOWER, A->n, B->n, a, lda, b, ldb, x, ldx, iter, opts); },
        [&](int64_t) { return iter && *iter < 0; }, "dsposv_tiles");
}

int64_t slate_amd_dsposv_gmres_tiles(slate_amd_context ctx, const slate_amd_tiles* A, const slate_amd_tiles* B,
                                     const slate_amd_tiles* X, int* iter, const slate_amd_options* opts)
{
    return with_tiles(ctx, A, true, B, X, false,
        [&](double* a, int64_t lda, double* b, int64_t ldb, double* x, int64_t ldx) {
            return slate_amd_dsposv_gmres(ctx, SLATE_AMD_LOWER, A->n, B->n, a, lda, b, ldb, x, ldx, iter, opts); },
        [&](int64_t) { return iter && *iter < 0; }, "dsposv_gmres_tiles");
}
