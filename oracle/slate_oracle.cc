// ============================================================================
// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain restatement, on the CPU, of SLATE's Target::HostTask single-rank
// algorithms for the factor-and-solve path.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library, and only as the checker
// or as the timed CPU baseline.  The product (slate_amd/) never calls it.
//
// Matrices are column-major (element (i,j) at A[i + j*lda]) and are viewed as a
// grid of nb x nb tiles exactly as slate::Matrix::fromLAPACK does: tile (i,j)
// covers rows [i*nb, min(m,(i+1)*nb)) and columns [j*nb, min(n,(j+1)*nb)).
//
// Parity status of each function (what pins it) is stated at the function and
// collected in the header comment of oracle/__init__.py.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <dlfcn.h>
#include <omp.h>

namespace {

//------------------------------------------------------------------------------
// Philox-2x64 variant used by SLATE's matgen.  Follows matgen/random.cc:53-75:
// ten rounds; per round the key is bumped by 0xD2B74407B1CE6E93 (except round 0)
// and the state (L, R) becomes (lo(R*M), hi(R*M) ^ key ^ L) with
// M = 0x9E3779B97F4A7C15.
inline void philox(uint64_t s0, uint64_t s1, uint64_t seed, uint64_t* o0, uint64_t* o1)
{
    const uint64_t bump = 0xD2B74407B1CE6E93ull;
    const uint64_t mult = 0x9E3779B97F4A7C15ull;
    for (int r = 0; r < 10; ++r) {
        if (r != 0) seed += bump;
        __uint128_t p = (__uint128_t)s1 * mult;
        uint64_t lo = (uint64_t)p, hi = (uint64_t)(p >> 64);
        uint64_t n0 = lo, n1 = hi ^ seed ^ s0;
        s0 = n0; s1 = n1;
    }
    *o0 = s0; *o1 = s1;
}

// matgen/random.cc:80-88 — top `digits` bits scaled into [0,1).
inline double bits_to_double(uint64_t b) { return double(b >> 11) / double(uint64_t(1) << 53); }

// matgen/random.cc:92-127 for the real case: dist 1 = Uniform [0,1),
// 2 = UniformSigned [-1,1), 3 = Normal (Box-Muller, cosine branch).
inline double gen_one(int dist, int64_t seed, int64_t i, int64_t j)
{
    uint64_t b0, b1;
    philox(uint64_t(i), uint64_t(j), uint64_t(seed), &b0, &b1);
    double u1 = bits_to_double(b0), u2 = bits_to_double(b1);
    const double pi = 3.1415926535897932385;
    switch (dist) {
        case 1: return u1;
        case 2: return 2*u1 - 1;
        case 3: return std::sqrt(-2*std::log(1 - u1)) * std::cos(2*pi*u2);
        default: return 0;
    }
}

template <class T> inline T cabs1(T x) { return std::fabs(x); }   // util.hh:125-140 (real)

//------------------------------------------------------------------------------
// Optional tile BLAS.  SLATE's HostTask tiles call a vendor BLAS
// (src/internal/internal_gemm.cc, Tile_blas.hh); for the timed CPU baseline the
// oracle can route its tile gemm/trsm through an OpenBLAS loaded at run time
// (oracle_set_blas; Fortran LP64 interface, symbols scipy_<name>_ as exported
// by scipy's bundled OpenBLAS), one BLAS thread per OpenMP task as in SLATE.
// Without it the plain loops below run.  Results agree to rounding only.
typedef void (*gemm_d_t)(const char*, const char*, const int*, const int*, const int*, const double*,
                         const double*, const int*, const double*, const int*, const double*, double*,
                         const int*);
typedef void (*gemm_s_t)(const char*, const char*, const int*, const int*, const int*, const float*,
                         const float*, const int*, const float*, const int*, const float*, float*,
                         const int*);
typedef void (*trsm_d_t)(const char*, const char*, const char*, const char*, const int*, const int*,
                         const double*, const double*, const int*, double*, const int*);
typedef void (*trsm_s_t)(const char*, const char*, const char*, const char*, const int*, const int*,
                         const float*, const float*, const int*, float*, const int*);
struct Blas {
    gemm_d_t dgemm = nullptr;
    gemm_s_t sgemm = nullptr;
    trsm_d_t dtrsm = nullptr;
    trsm_s_t strsm = nullptr;
};
Blas g_blas;

template <class T> inline bool blas_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, T alpha,
                                         const T* A, int64_t lda, const T* B, int64_t ldb, T beta,
                                         T* C, int64_t ldc)
{
    const int M = int(m), N = int(n), K = int(k), la = int(lda), lb = int(ldb), lc = int(ldc);
    if constexpr (sizeof(T) == 8) {
        if (!g_blas.dgemm) return false;
        g_blas.dgemm(&ta, &tb, &M, &N, &K, &alpha, A, &la, B, &lb, &beta, C, &lc);
    }
    else {
        if (!g_blas.sgemm) return false;
        g_blas.sgemm(&ta, &tb, &M, &N, &K, &alpha, A, &la, B, &lb, &beta, C, &lc);
    }
    return true;
}

//------------------------------------------------------------------------------
// Low-precision trailing updates of the float factor (this library's factor
// precisions beyond SLATE's float; the reference has double -> float only,
// src/gesv_mixed.cc:293-316).  With g_q.kind != 0 the float factor's trailing
// update C -= A B (LU gemm, Cholesky herk/gemm) multiplies rounded operands,
// products exact, accumulated in FP32 one MFMA-depth chunk at a time and
// subtracted in float (gemm_quantized):
//  * kind 1, terms t = 1..3: each float x split into bf16 terms
//    x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1) (round to
//    nearest even); the products x_i y_j with i + j < t are kept.
//  * kind 2, terms t = 1..2: FP8 e4m3 (OCP E4M3FN) with a power-of-two scale
//    per tile.  A tile is a block of `qrows` rows (of A) or columns (of B)
//    over the whole inner dimension k of the update, counted from the first
//    row/column of the update; its scale 2^s is the largest with
//    max|x| 2^s <= 448.  q0 = e4m3(x 2^s); with t = 2 the remainder plane
//    q1 = e4m3((x 2^s - q0) 2^4); the product is
//    2^-(sa+sb) (qa0 qb0 + 2^-4 (qa0 qb1 + qa1 qb0)).
struct UpdQ { int kind = 0; int terms = 0; int64_t qrows = 128; };
UpdQ g_q;

inline float bf16_rne(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// OCP FP8 E4M3FN: bias 7, no infinities, S.1111.111 = NaN, max 448 =
// 1.75 * 2^8, min normal 2^-6, subnormals m * 2^-9.
inline double e4m3_decode(uint8_t b)
{
    const int sgn = b >> 7, e = (b >> 3) & 15, m = b & 7;
    if (e == 15 && m == 7) return std::numeric_limits<double>::quiet_NaN();
    const double v = e == 0 ? std::ldexp(double(m), -9) : std::ldexp(1.0 + m / 8.0, e - 7);
    return sgn ? -v : v;
}

// Round to the nearest e4m3 value, ties to even, saturating at +-448 (the
// scaled operands never exceed it).  The quantum of the binade of |x| is
// 2^(e-3), 2^-9 below the normal range; rint rounds half to even.
inline double e4m3_round(double x)
{
    if (!(std::fabs(x) > 0)) return x == x ? 0.0 * x : x;
    int e;
    std::frexp(std::fabs(x), &e);          // |x| in [2^(e-1), 2^e)
    const int q = std::max(e - 1, -6) - 3;
    double r = std::ldexp(std::nearbyint(std::ldexp(x, -q)), q);
    if (r > 448) r = 448;
    if (r < -448) r = -448;
    return r;
}

// Encoded byte of e4m3_round(x) (x finite).
inline uint8_t e4m3_encode(double x)
{
    const double r = e4m3_round(x);
    const uint8_t sgn = std::signbit(r) ? 0x80 : 0;
    const double a = std::fabs(r);
    if (a == 0) return sgn;
    int e;
    const double f = std::frexp(a, &e);    // a = f 2^e, f in [0.5, 1)
    int be = e - 1 + 7;
    if (be <= 0) return uint8_t(sgn | int(std::ldexp(a, 9)));          // subnormal: m 2^-9
    const int m = int(std::ldexp(f, 4)) - 8;                             // (1 + m/8) 2^(e-1)
    return uint8_t(sgn | (be << 3) | m);
}

// Largest s with amax 2^s <= 448 (0 for an all-zero tile).
inline int e4m3_scale_exp(double amax)
{
    if (!(amax > 0)) return 0;
    int e;
    const double f = std::frexp(amax, &e); // amax = f 2^e, 448 = 0.875 2^9
    return f <= 0.875 ? 9 - e : 8 - e;
}

// Quantized planes of an m x k operand (element (i, p) at X[i + p*ldx] when
// rows_are_m, else X[p + i*ldx]); tiles of g_q.qrows along m starting at row
// `tile0` of the operand.  out[t][i + p*m] = term t in the unscaled units,
// scale[i] = 2^-s of row i's tile (1 for the bf16 split).
template <class T>
void quantize_operand(int64_t m, int64_t k, const T* X, int64_t ldx, bool rows_are_m,
                      std::vector<double>* out, std::vector<double>& scale)
{
    auto at = [&](int64_t i, int64_t p) -> double { return rows_are_m ? X[i + p*ldx] : X[p + i*ldx]; };
    const int t = g_q.terms;
    for (int q = 0; q < t; ++q) out[q].assign(m * k, 0.0);
    scale.assign(m, 1.0);
    if (g_q.kind == 1) {
        for (int64_t p = 0; p < k; ++p)
            for (int64_t i = 0; i < m; ++i) {
                float r = float(at(i, p));
                for (int q = 0; q < t; ++q) {
                    const float h = bf16_rne(r);
                    out[q][i + p*m] = h;
                    r = r - h;
                }
            }
        return;
    }
    const int64_t qr = g_q.qrows;
    for (int64_t i0 = 0; i0 < m; i0 += qr) {
        const int64_t i1 = std::min(m, i0 + qr);
        double amax = 0;
        for (int64_t p = 0; p < k; ++p)
            for (int64_t i = i0; i < i1; ++i) amax = std::max(amax, std::fabs(at(i, p)));
        const int s = e4m3_scale_exp(amax);
        for (int64_t i = i0; i < i1; ++i) scale[i] = std::ldexp(1.0, -s);
        for (int64_t p = 0; p < k; ++p)
            for (int64_t i = i0; i < i1; ++i) {
                const double xs = std::ldexp(double(float(at(i, p))), s);
                const double q0 = e4m3_round(xs);
                out[0][i + p*m] = q0;
                if (t > 1) out[1][i + p*m] = e4m3_round((xs - q0) * 16.0);
            }
    }
}

// C[m x n] -= A B on quantized operands; A m x k (column-major, lda), B given
// as op(B) = k x n from Bsrc: b_is_kn ? Bsrc[p + j*ldb] : Bsrc[j + p*ldb].
// Only C(i, j) with (!lower_only || i >= j) are updated.
// Accumulation follows the GPU's matrix-core update: k in chunks of one MFMA
// depth (32 for bf16, 128 for e4m3), each chunk's product sum formed exactly
// (in double: the products of two bf16 or e4m3 values are exact) and added to
// an FP32 accumulator with one rounding, term pairs in the kernels' order
// (bf16: smaller terms first, x0 y0 last; e4m3: q0 q0', then the two
// 2^-4 cross terms), and C := C - scale * acc rounded once at the end.
template <class T>
void gemm_quantized(int64_t m, int64_t n, int64_t k, const T* A, int64_t lda, const T* Bsrc, int64_t ldb,
                    bool b_is_kn, T* C, int64_t ldc, bool lower_only = false)
{
    std::vector<double> qa[3], qb[3], sa, sb;
    quantize_operand<T>(m, k, A, lda, true, qa, sa);
    // B as an n x k operand: element (j, p)
    quantize_operand<T>(n, k, Bsrc, ldb, !b_is_kn, qb, sb);
    const int t = g_q.terms;
    const int64_t ck = g_q.kind == 1 ? 32 : 128;
    // (a, b, weight) term pairs in accumulation order
    int pa[6], pb[6], np = 0;
    double pw[6];
    if (g_q.kind == 1) {
        for (int a = t - 1; a >= 0; --a)
            for (int b = t - 1 - a; b >= 0; --b) { pa[np] = a; pb[np] = b; pw[np] = 1.0; ++np; }
    }
    else {
        pa[np] = 0; pb[np] = 0; pw[np] = 1.0; ++np;
        if (t > 1) {
            pa[np] = 0; pb[np] = 1; pw[np] = 1.0 / 16.0; ++np;
            pa[np] = 1; pb[np] = 0; pw[np] = 1.0 / 16.0; ++np;
        }
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = lower_only ? j : 0; i < m; ++i) {
            float acc = 0.0f;
            for (int64_t p0 = 0; p0 < k; p0 += ck) {
                const int64_t p1 = std::min(k, p0 + ck);
                for (int u = 0; u < np; ++u) {
                    double d = 0;
                    for (int64_t p = p0; p < p1; ++p) d += qa[pa[u]][i + p*m] * qb[pb[u]][j + p*n];
                    acc = float(double(acc) + pw[u] * d);
                }
            }
            const double scale = sa[i] * sb[j];
            C[i + j*ldc] = T(double(C[i + j*ldc]) - double(float(scale * double(acc))));
        }
}

//------------------------------------------------------------------------------
// Plain loop kernels on column-major blocks.
// C[m x n] += alpha * A[m x k] * B[k x n]   (op = NoTrans/NoTrans)
template <class T>
void gemm_nn(int64_t m, int64_t n, int64_t k, T alpha,
             const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (blas_gemm<T>('N', 'N', m, n, k, alpha, A, lda, B, ldb, T(1), C, ldc)) return;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = 0; p < k; ++p) {
            T b = alpha * B[p + j*ldb];
            if (b == T(0)) continue;
            const T* a = A + p*lda;
            T* c = C + j*ldc;
            for (int64_t i = 0; i < m; ++i) c[i] += a[i] * b;
        }
}

// C[m x n] += alpha * A[m x k] * B[n x k]^T
template <class T>
void gemm_nt(int64_t m, int64_t n, int64_t k, T alpha,
             const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (blas_gemm<T>('N', 'T', m, n, k, alpha, A, lda, B, ldb, T(1), C, ldc)) return;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = 0; p < k; ++p) {
            T b = alpha * B[j + p*ldb];
            if (b == T(0)) continue;
            const T* a = A + p*lda;
            T* c = C + j*ldc;
            for (int64_t i = 0; i < m; ++i) c[i] += a[i] * b;
        }
}

// Solve op(L) X = B in place, L m x m lower triangular (unit or not), B m x n.
template <class T>
void trsm_left_lower(int64_t m, int64_t n, bool unit, const T* L, int64_t ldl, T* B, int64_t ldb)
{
    if (m > 0 && n > 0) {
        const char sd = 'L', ul = 'L', tr = 'N', dg = unit ? 'U' : 'N';
        const int M = int(m), N = int(n), l1 = int(ldl), l2 = int(ldb);
        const T one = 1;
        if constexpr (sizeof(T) == 8) {
            if (g_blas.dtrsm) { g_blas.dtrsm(&sd, &ul, &tr, &dg, &M, &N, &one, L, &l1, B, &l2); return; }
        }
        else {
            if (g_blas.strsm) { g_blas.strsm(&sd, &ul, &tr, &dg, &M, &N, &one, L, &l1, B, &l2); return; }
        }
    }
    for (int64_t j = 0; j < n; ++j) {
        T* b = B + j*ldb;
        for (int64_t p = 0; p < m; ++p) {
            if (!unit) b[p] /= L[p + p*ldl];
            T x = b[p];
            if (x == T(0)) continue;
            for (int64_t i = p+1; i < m; ++i) b[i] -= x * L[i + p*ldl];
        }
    }
}

// Solve U X = B in place, U m x m upper triangular non-unit.
template <class T>
void trsm_left_upper(int64_t m, int64_t n, const T* U, int64_t ldu, T* B, int64_t ldb)
{
    for (int64_t j = 0; j < n; ++j) {
        T* b = B + j*ldb;
        for (int64_t p = m-1; p >= 0; --p) {
            b[p] /= U[p + p*ldu];
            T x = b[p];
            if (x == T(0)) continue;
            for (int64_t i = 0; i < p; ++i) b[i] -= x * U[i + p*ldu];
        }
    }
}

// Solve L^T X = B in place (L lower non-unit): back substitution.
template <class T>
void trsm_left_lower_trans(int64_t m, int64_t n, bool unit, const T* L, int64_t ldl, T* B, int64_t ldb)
{
    for (int64_t j = 0; j < n; ++j) {
        T* b = B + j*ldb;
        for (int64_t p = m-1; p >= 0; --p) {
            T s = b[p];
            for (int64_t i = p+1; i < m; ++i) s -= L[i + p*ldl] * b[i];
            b[p] = unit ? s : s / L[p + p*ldl];
        }
    }
}

// Solve U^T X = B in place (U upper non-unit): forward substitution.
template <class T>
void trsm_left_upper_trans(int64_t m, int64_t n, const T* U, int64_t ldu, T* B, int64_t ldb)
{
    for (int64_t j = 0; j < n; ++j) {
        T* b = B + j*ldb;
        for (int64_t p = 0; p < m; ++p) {
            T s = b[p];
            for (int64_t i = 0; i < p; ++i) s -= U[i + p*ldu] * b[i];
            b[p] = s / U[p + p*ldu];
        }
    }
}

// X L^T = B in place, i.e. B := B * L^{-T}, L n x n lower non-unit, B m x n.
template <class T>
void trsm_right_lower_trans(int64_t m, int64_t n, const T* L, int64_t ldl, T* B, int64_t ldb)
{
    for (int64_t j = 0; j < n; ++j) {
        T* bj = B + j*ldb;
        for (int64_t p = 0; p < j; ++p) {
            T l = L[j + p*ldl];
            if (l == T(0)) continue;
            const T* bp = B + p*ldb;
            for (int64_t i = 0; i < m; ++i) bj[i] -= bp[i] * l;
        }
        T d = L[j + j*ldl];
        for (int64_t i = 0; i < m; ++i) bj[i] /= d;
    }
}

struct Grid {
    int64_t m, n, nb;
    int64_t mt() const { return (m + nb - 1) / nb; }
    int64_t nt() const { return (n + nb - 1) / nb; }
    int64_t tmb(int64_t i) const { return std::min(nb, m - i*nb); }
    int64_t tnb(int64_t j) const { return std::min(nb, n - j*nb); }
};

//------------------------------------------------------------------------------
// Panel LU of tile column k: follows src/internal/Tile_getrf.hh:159-447 on one
// rank (MPI reductions are identities), with `nthreads` emulated panel threads
// so that exact-tie behaviour matches the reference thread partition
// (thread r scans tiles r, r+T, ...; strict '>' within and across threads,
// Tile_getrf.hh:210-250).  Writes pivot (tile index relative to k, offset).
template <class T>
int64_t panel_getrf(T* A, int64_t lda, const Grid& g, int64_t k, int64_t ib,
                    int64_t nthreads, double threshold,
                    int64_t* piv_tile, int64_t* piv_off)
{
    int64_t mt = g.mt();
    int64_t ntiles = mt - k;
    int64_t nbk = g.tnb(k);
    int64_t diag_len = std::min(g.tmb(k), nbk);
    int64_t T_ = std::max<int64_t>(1, std::min<int64_t>(nthreads, ntiles));
    int64_t info = 0;
    const T one = 1, zero = 0;
    const T safe_min = std::numeric_limits<T>::min();
    auto tile = [&](int64_t idx) { return A + (k + idx)*g.nb + k*g.nb*lda; };  // tile (k+idx, k)
    auto tmb = [&](int64_t idx) { return g.tmb(k + idx); };
    std::vector<T> maxv(T_);
    std::vector<int64_t> maxi(T_), maxo(T_);
    std::vector<T> top(ib * nbk);
    std::vector<T> pivval(diag_len);

    for (int64_t k0 = 0; k0 < diag_len; k0 += ib) {
        int64_t kb = std::min(diag_len - k0, ib);
        for (int64_t j = k0; j < k0 + kb; ++j) {
            // thread max search (Tile_getrf.hh:197-237)
            for (int64_t r = 0; r < T_; ++r) {
                if (r == 0) { maxv[r] = tile(0)[j + j*lda]; maxi[r] = 0; maxo[r] = j; }
                else        { maxv[r] = tile(r)[0 + j*lda]; maxi[r] = r; maxo[r] = 0; }
                for (int64_t idx = r; idx < ntiles; idx += T_) {
                    const T* t = tile(idx);
                    int64_t i0 = (idx == 0) ? j+1 : 0;
                    for (int64_t i = i0; i < tmb(idx); ++i)
                        if (cabs1(t[i + j*lda]) > cabs1(maxv[r])) {
                            maxv[r] = t[i + j*lda]; maxi[r] = idx; maxo[r] = i;
                        }
                }
            }
            // thread reduction (Tile_getrf.hh:244-250)
            for (int64_t r = 1; r < T_; ++r)
                if (cabs1(maxv[r]) > cabs1(maxv[0])) {
                    maxv[0] = maxv[r]; maxi[0] = maxi[r]; maxo[0] = maxo[r];
                }
            // diagonal kept if |diag| >= threshold*|max| (Tile_getrf.hh:256-289)
            T diag = tile(0)[j + j*lda];
            if (cabs1(diag) >= cabs1(maxv[0]) * T(threshold)) {
                maxo[0] = j; maxi[0] = 0; maxv[0] = diag;
            }
            piv_tile[j] = maxi[0];
            piv_off[j] = maxo[0];
            pivval[j] = maxv[0];
            // swap across the full panel width (Tile_getrf.hh:52-95, 304-306)
            if (maxi[0] > 0 || maxo[0] > j) {
                T* t0 = tile(0);
                T* tp = tile(maxi[0]);
                for (int64_t c = 0; c < nbk; ++c)
                    std::swap(t0[j + c*lda], tp[maxo[0] + c*lda]);
            }
            // top row for geru (Tile_getrf.hh:309-316)
            for (int64_t c = j+1; c < k0 + kb; ++c) top[c - (j+1)] = tile(0)[j + c*lda];
            // column scaling and rank-1 update (Tile_getrf.hh:326-386)
            for (int64_t idx = 0; idx < ntiles; ++idx) {
                T* t = tile(idx);
                int64_t i0 = (idx == 0) ? j+1 : 0;
                T pv = pivval[j];
                if (cabs1(pv) >= safe_min) {
                    T alpha = (idx == 0) ? one / t[j + j*lda] : one / pv;
                    for (int64_t i = i0; i < tmb(idx); ++i) t[i + j*lda] *= alpha;
                }
                else if (pv != zero) {
                    T d = (idx == 0) ? t[j + j*lda] : pv;
                    for (int64_t i = i0; i < tmb(idx); ++i) t[i + j*lda] /= d;
                }
                else if (info == 0 && idx == 0) {
                    info = j + 1;
                }
                for (int64_t c = j+1; c < k0 + kb; ++c) {
                    T y = top[c - (j+1)];
                    for (int64_t i = i0; i < tmb(idx); ++i) t[i + c*lda] -= t[i + j*lda] * y;
                }
            }
        }
        // trsm + rank-kb update to the right within the panel (Tile_getrf.hh:392-445)
        if (k0 + kb < nbk) {
            T* t0 = tile(0);
            trsm_left_lower<T>(kb, nbk - k0 - kb, true, t0 + k0 + k0*lda, lda,
                               t0 + k0 + (k0+kb)*lda, lda);
            for (int64_t idx = 0; idx < ntiles; ++idx) {
                T* t = tile(idx);
                if (idx == 0) {
                    if (k0 + kb < tmb(0))
                        gemm_nn<T>(tmb(0) - k0 - kb, nbk - k0 - kb, kb, T(-1),
                                   t + (k0+kb) + k0*lda, lda, t0 + k0 + (k0+kb)*lda, lda,
                                   t + (k0+kb) + (k0+kb)*lda, lda);
                }
                else {
                    gemm_nn<T>(tmb(idx), nbk - k0 - kb, kb, T(-1),
                               t + k0*lda, lda, t0 + k0 + (k0+kb)*lda, lda,
                               t + (k0+kb)*lda, lda);
                }
            }
        }
    }
    return info;
}

// Apply panel-k pivots (forward) to tile columns [j0, j1) of rows k*nb..
// (src/internal/internal_swap.cc:198-435, sequential LAPACK-style semantics).
template <class T>
void permute_rows_fwd(T* A, int64_t lda, const Grid& g, int64_t k, int64_t col0, int64_t col1,
                      const int64_t* pt, const int64_t* po, int64_t len)
{
    for (int64_t i = 0; i < len; ++i) {
        int64_t r1 = k*g.nb + i;
        int64_t r2 = (k + pt[i])*g.nb + po[i];
        if (r1 == r2) continue;
        for (int64_t c = col0; c < col1; ++c) std::swap(A[r1 + c*lda], A[r2 + c*lda]);
    }
}

//------------------------------------------------------------------------------
// Right-looking tile LU, src/getrf.cc:23-245 (lookahead only reorders tasks and
// does not change results; it is not emulated).  ipiv is LAPACK-style, 1-based
// global rows (lapack_api/lapack_gesv.cc:90-100).
template <class T>
int64_t getrf_tiles(int64_t m, int64_t n, T* A, int64_t lda, int64_t nb, int64_t ib,
                    int64_t nthreads, double threshold, int64_t* ipiv)
{
    Grid g{m, n, nb};
    int64_t mt = g.mt(), nt = g.nt();
    int64_t kmax = std::min(mt, nt);
    int64_t info = 0, kk = 0;
    std::vector<int64_t> pt(nb), po(nb);
    for (int64_t k = 0; k < kmax; ++k) {
        int64_t diag_len = std::min(g.tmb(k), g.tnb(k));
        int64_t iinfo = panel_getrf<T>(A, lda, g, k, ib, nthreads, threshold, pt.data(), po.data());
        if (info == 0 && iinfo > 0) info = kk + iinfo;
        for (int64_t i = 0; i < diag_len; ++i) ipiv[k*nb + i] = (k + pt[i])*nb + po[i] + 1;
        // pivot to the left and to the right (getrf.cc:121-215)
        permute_rows_fwd<T>(A, lda, g, k, 0, k*nb, pt.data(), po.data(), diag_len);
        permute_rows_fwd<T>(A, lda, g, k, k*nb + g.tnb(k), n, pt.data(), po.data(), diag_len);
        // A(k, j) = L(k,k)^{-1} A(k, j), j > k (internal_trsm)
        T* Akk = A + k*nb + k*nb*lda;
        int64_t mbk = g.tmb(k);
        #pragma omp parallel for schedule(dynamic)
        for (int64_t j = k+1; j < nt; ++j)
            trsm_left_lower<T>(mbk, g.tnb(j), true, Akk, lda, A + k*nb + j*nb*lda, lda);
        // A(i, j) -= A(i, k) A(k, j), i, j > k (internal_gemm)
        int64_t cnt = (mt - k - 1) * (nt - k - 1);
        #pragma omp parallel for schedule(dynamic)
        for (int64_t t = 0; t < cnt; ++t) {
            int64_t i = k + 1 + t % (mt - k - 1);
            int64_t j = k + 1 + t / (mt - k - 1);
            if (sizeof(T) == 4 && g_q.kind != 0)
                gemm_quantized<T>(g.tmb(i), g.tnb(j), g.tnb(k), A + i*nb + k*nb*lda, lda,
                                  A + k*nb + j*nb*lda, lda, true, A + i*nb + j*nb*lda, lda);
            else
                gemm_nn<T>(g.tmb(i), g.tnb(j), g.tnb(k), T(-1),
                           A + i*nb + k*nb*lda, lda, A + k*nb + j*nb*lda, lda,
                           A + i*nb + j*nb*lda, lda);
        }
        kk += g.tnb(k);
    }
    return info;
}

// Forward/backward tile sweeps of src/work/work_trsm.cc:52-266 (Side::Left).
template <class T>
void trsm_tiles_lower(int64_t n, int64_t nrhs, bool unit, const T* L, int64_t ldl,
                      T* B, int64_t ldb, int64_t nb)
{
    Grid g{n, n, nb};
    for (int64_t k = 0; k < g.mt(); ++k) {
        trsm_left_lower<T>(g.tmb(k), nrhs, unit, L + k*nb + k*nb*ldl, ldl, B + k*nb, ldb);
        for (int64_t i = k+1; i < g.mt(); ++i)
            gemm_nn<T>(g.tmb(i), nrhs, g.tmb(k), T(-1), L + i*nb + k*nb*ldl, ldl,
                       B + k*nb, ldb, B + i*nb, ldb);
    }
}

template <class T>
void trsm_tiles_upper(int64_t n, int64_t nrhs, const T* U, int64_t ldu,
                      T* B, int64_t ldb, int64_t nb)
{
    Grid g{n, n, nb};
    for (int64_t k = g.mt() - 1; k >= 0; --k) {
        trsm_left_upper<T>(g.tmb(k), nrhs, U + k*nb + k*nb*ldu, ldu, B + k*nb, ldb);
        for (int64_t i = 0; i < k; ++i)
            gemm_nn<T>(g.tmb(i), nrhs, g.tmb(k), T(-1), U + i*nb + k*nb*ldu, ldu,
                       B + k*nb, ldb, B + i*nb, ldb);
    }
}

// src/getrs.cc:66-117 (NoTrans: pivot B, L solve, U solve; Trans: U^T, L^T, pivot back).
template <class T>
void getrs_tiles(int64_t n, int64_t nrhs, const T* A, int64_t lda, const int64_t* ipiv,
                 T* B, int64_t ldb, int64_t nb, int trans)
{
    if (!trans) {
        for (int64_t i = 0; i < n; ++i) {
            int64_t p = ipiv[i] - 1;
            if (p != i) for (int64_t c = 0; c < nrhs; ++c) std::swap(B[i + c*ldb], B[p + c*ldb]);
        }
        trsm_tiles_lower<T>(n, nrhs, true, A, lda, B, ldb, nb);
        trsm_tiles_upper<T>(n, nrhs, A, lda, B, ldb, nb);
    }
    else {
        trsm_left_upper_trans<T>(n, nrhs, A, lda, B, ldb);
        trsm_left_lower_trans<T>(n, nrhs, true, A, lda, B, ldb);
        for (int64_t i = n-1; i >= 0; --i) {
            int64_t p = ipiv[i] - 1;
            if (p != i) for (int64_t c = 0; c < nrhs; ++c) std::swap(B[i + c*ldb], B[p + c*ldb]);
        }
    }
}

// Unblocked lower Cholesky of one tile (LAPACK dpotrf semantics, the callee of
// src/internal/Tile_lapack.hh:268-275).  Returns the 1-based column of the first
// non-positive pivot, 0 on success.
template <class T>
int64_t potrf_tile_lower(int64_t n, T* A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j) {
        T d = A[j + j*lda];
        for (int64_t p = 0; p < j; ++p) d -= A[j + p*lda] * A[j + p*lda];
        if (!(d > T(0))) { A[j + j*lda] = d; return j + 1; }
        d = std::sqrt(d);
        A[j + j*lda] = d;
        for (int64_t i = j+1; i < n; ++i) {
            T s = A[i + j*lda];
            for (int64_t p = 0; p < j; ++p) s -= A[i + p*lda] * A[j + p*lda];
            A[i + j*lda] = s / d;
        }
    }
    return 0;
}

// Right-looking lower tile Cholesky, src/potrf.cc:23-211 (herk on diagonal
// tiles, gemm below, src/internal/internal_herk.cc:47-110).  Stops at the
// first failing tile and reports the global 1-based column, like the
// reference's info = kk + iinfo (potrf.cc:104-105).
template <class T>
int64_t potrf_tiles(int64_t n, T* A, int64_t lda, int64_t nb)
{
    Grid g{n, n, nb};
    int64_t nt = g.nt();
    for (int64_t k = 0; k < nt; ++k) {
        T* Akk = A + k*nb + k*nb*lda;
        int64_t iinfo = potrf_tile_lower<T>(g.tnb(k), Akk, lda);
        if (iinfo != 0) return k*nb + iinfo;
        #pragma omp parallel for schedule(dynamic)
        for (int64_t i = k+1; i < nt; ++i)
            trsm_right_lower_trans<T>(g.tmb(i), g.tnb(k), Akk, lda, A + i*nb + k*nb*lda, lda);
        int64_t r = nt - k - 1;
        #pragma omp parallel for schedule(dynamic)
        for (int64_t t = 0; t < r*r; ++t) {
            int64_t i = k + 1 + t % r, j = k + 1 + t / r;
            if (i < j) continue;
            T* Cij = A + i*nb + j*nb*lda;
            if (sizeof(T) == 4 && g_q.kind != 0) {
                // herk (lower triangle of the diagonal tile) or gemm on quantized operands
                gemm_quantized<T>(g.tmb(i), g.tnb(j), g.tnb(k), A + i*nb + k*nb*lda, lda,
                                  A + j*nb + k*nb*lda, lda, false, Cij, lda, i == j);
            }
            else if (i == j) {
                // herk, lower triangle only
                std::vector<T> tmp(g.tmb(i) * g.tmb(i), T(0));
                gemm_nt<T>(g.tmb(i), g.tmb(i), g.tnb(k), T(1), A + i*nb + k*nb*lda, lda,
                           A + i*nb + k*nb*lda, lda, tmp.data(), g.tmb(i));
                for (int64_t c = 0; c < g.tmb(i); ++c)
                    for (int64_t rr = c; rr < g.tmb(i); ++rr) Cij[rr + c*lda] -= tmp[rr + c*g.tmb(i)];
            }
            else {
                gemm_nt<T>(g.tmb(i), g.tnb(j), g.tnb(k), T(-1), A + i*nb + k*nb*lda, lda,
                           A + j*nb + k*nb*lda, lda, Cij, lda);
            }
        }
    }
    return 0;
}

// src/potrs.cc:55-77: L solve then L^T solve.
template <class T>
void potrs_tiles(int64_t n, int64_t nrhs, const T* A, int64_t lda, T* B, int64_t ldb, int64_t nb)
{
    trsm_tiles_lower<T>(n, nrhs, false, A, lda, B, ldb, nb);
    trsm_left_lower_trans<T>(n, nrhs, false, A, lda, B, ldb);
}

//------------------------------------------------------------------------------
// Norms and residuals (closed forms of src/norm.cc, src/colNorms.cc).
double norm_inf_ge(int64_t m, int64_t n, const double* A, int64_t lda)
{
    std::vector<double> rs(m, 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) rs[i] += std::fabs(A[i + j*lda]);
    double mx = 0;
    for (double v : rs) mx = (v > mx || std::isnan(v)) ? v : mx;
    return mx;
}

// Inf (= One) norm of a symmetric matrix stored in its lower triangle.
double norm_inf_sy_lower(int64_t n, const double* A, int64_t lda)
{
    std::vector<double> rs(n, 0.0);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = j; i < n; ++i) {
            double a = std::fabs(A[i + j*lda]);
            rs[i] += a;
            if (i != j) rs[j] += a;
        }
    double mx = 0;
    for (double v : rs) mx = (v > mx || std::isnan(v)) ? v : mx;
    return mx;
}

void col_norms_max(int64_t m, int64_t n, const double* X, int64_t ldx, double* out)
{
    for (int64_t j = 0; j < n; ++j) {
        double mx = 0;
        for (int64_t i = 0; i < m; ++i) {
            double a = std::fabs(X[i + j*ldx]);
            mx = (a > mx || std::isnan(a)) ? a : mx;
        }
        out[j] = mx;
    }
}

// R = B - op(A) X with A general (sym=0) or symmetric lower-stored (sym=1).
void residual(int64_t n, int64_t nrhs, const double* A, int64_t lda, int sym,
              const double* X, int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr)
{
    for (int64_t c = 0; c < nrhs; ++c) {
        double* r = R + c*ldr;
        const double* x = X + c*ldx;
        for (int64_t i = 0; i < n; ++i) r[i] = B[i + c*ldb];
        for (int64_t j = 0; j < n; ++j) {
            double xj = x[j];
            if (!sym) {
                for (int64_t i = 0; i < n; ++i) r[i] -= A[i + j*lda] * xj;
            }
            else {
                r[j] -= A[j + j*lda] * xj;
                double acc = 0;
                for (int64_t i = j+1; i < n; ++i) {
                    r[i] -= A[i + j*lda] * xj;
                    acc += A[i + j*lda] * x[i];
                }
                r[j] -= acc;
            }
        }
    }
}

// internal_util.hh:121-138
bool ir_converged(const std::vector<double>& nr, const std::vector<double>& nx, double cte)
{
    // a non-finite norm is "not converged" (the comparison alone would accept
    // NaN and inf > inf); the library does the same
    for (size_t i = 0; i < nr.size(); ++i) {
        if (!std::isfinite(nr[i]) || !std::isfinite(nx[i])) return false;
        if (nr[i] > nx[i] * cte) return false;
    }
    return true;
}

//------------------------------------------------------------------------------
// Low-precision factor used by the refinement drivers: a float copy of A
// factored by the same tile algorithm (src/gesv_mixed.cc:176-183).
struct LoFactor {
    int64_t n, nb, ib, nthreads;
    double threshold;
    bool chol;
    std::vector<float> A;
    std::vector<int64_t> ipiv;
    int64_t factor(const double* Ahi, int64_t lda) {
        A.assign(n*n, 0.f);
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) A[i + j*n] = float(Ahi[i + j*lda]);
        if (chol) return potrf_tiles<float>(n, A.data(), n, nb);
        ipiv.assign(n, 0);
        return getrf_tiles<float>(n, n, A.data(), n, nb, ib, nthreads, threshold, ipiv.data());
    }
    // x := A_lo^{-1} x, vector rounded to float and back (copy/getrs/copy of
    // src/gesv_mixed.cc:220-228).
    void solve(int64_t nrhs, double* X, int64_t ldx) {
        std::vector<float> xl(n * nrhs);
        for (int64_t c = 0; c < nrhs; ++c)
            for (int64_t i = 0; i < n; ++i) xl[i + c*n] = float(X[i + c*ldx]);
        if (chol) potrs_tiles<float>(n, nrhs, A.data(), n, xl.data(), n, nb);
        else getrs_tiles<float>(n, nrhs, A.data(), n, ipiv.data(), xl.data(), n, nb, 0);
        for (int64_t c = 0; c < nrhs; ++c)
            for (int64_t i = 0; i < n; ++i) X[i + c*ldx] = double(xl[i + c*n]);
    }
};

// Classic IR, src/gesv_mixed.cc:106-290 and src/posv_mixed.cc:105-292.
int64_t mixed_ir(int64_t n, int64_t nrhs, double* A, int64_t lda, int sym,
                 const double* B, int64_t ldb, double* X, int64_t ldx, int64_t* ipiv,
                 int64_t nb, int64_t ib, int64_t nthreads, double threshold,
                 int64_t itermax, double tol, int use_fallback, int* iter)
{
    const double eps = std::numeric_limits<double>::epsilon();
    if (tol <= 0) tol = eps * std::sqrt(double(n));
    double Anorm = sym ? norm_inf_sy_lower(n, A, lda) : norm_inf_ge(n, n, A, lda);
    double cte = Anorm * tol;
    LoFactor lo{n, nb, ib, nthreads, threshold, sym != 0, {}, {}};
    int64_t info = lo.factor(A, lda);
    bool converged = false;
    *iter = 0;
    std::vector<double> R(n * nrhs), nx(nrhs), nr(nrhs);
    if (info != 0) {
        *iter = -3;
    }
    else {
        for (int64_t c = 0; c < nrhs; ++c)
            for (int64_t i = 0; i < n; ++i) X[i + c*ldx] = B[i + c*ldb];
        lo.solve(nrhs, X, ldx);
        residual(n, nrhs, A, lda, sym, X, ldx, B, ldb, R.data(), n);
        col_norms_max(n, nrhs, X, ldx, nx.data());
        col_norms_max(n, nrhs, R.data(), n, nr.data());
        if (ir_converged(nr, nx, cte)) { *iter = 0; converged = true; }
        for (int64_t it = 0; it < itermax && !converged; ++it) {
            lo.solve(nrhs, R.data(), n);
            for (int64_t c = 0; c < nrhs; ++c)
                for (int64_t i = 0; i < n; ++i) X[i + c*ldx] += R[i + c*n];
            residual(n, nrhs, A, lda, sym, X, ldx, B, ldb, R.data(), n);
            col_norms_max(n, nrhs, X, ldx, nx.data());
            col_norms_max(n, nrhs, R.data(), n, nr.data());
            if (ir_converged(nr, nx, cte)) { *iter = int(it + 1); converged = true; }
        }
    }
    if (!converged) {
        if (info == 0) *iter = int(-itermax - 1);
        if (use_fallback) {
            if (sym) info = potrf_tiles<double>(n, A, lda, nb);
            else info = getrf_tiles<double>(n, n, A, lda, nb, ib, nthreads, threshold, ipiv);
            if (info == 0) {
                for (int64_t c = 0; c < nrhs; ++c)
                    for (int64_t i = 0; i < n; ++i) X[i + c*ldx] = B[i + c*ldb];
                if (sym) potrs_tiles<double>(n, nrhs, A, lda, X, ldx, nb);
                else getrs_tiles<double>(n, nrhs, A, lda, ipiv, X, ldx, nb, 0);
            }
        }
    }
    else if (!sym) {
        for (int64_t i = 0; i < n; ++i) ipiv[i] = lo.ipiv[i];
    }
    return info;
}

// Givens generation with the reference-BLAS drotg conventions (the callee of
// blas::rotg in src/gesv_mixed_gmres.cc:345): c*a + s*b = r, -s*a + c*b = 0.
void rotg(double* a, double* b, double* c, double* s)
{
    double aa = *a, bb = *b;
    double roe = std::fabs(aa) > std::fabs(bb) ? aa : bb;
    double scale = std::fabs(aa) + std::fabs(bb);
    double r, z;
    if (scale == 0) { *c = 1; *s = 0; r = 0; z = 0; }
    else {
        r = scale * std::sqrt((aa/scale)*(aa/scale) + (bb/scale)*(bb/scale));
        r = std::copysign(1.0, roe) * r;
        *c = aa / r; *s = bb / r;
        z = 1;
        if (std::fabs(aa) > std::fabs(bb)) z = *s;
        if (std::fabs(bb) >= std::fabs(aa) && *c != 0) z = 1 / *c;
    }
    *a = r; *b = z;
}

inline void rot(double* x, double* y, double c, double s)
{
    double t = c * *x + s * *y;
    *y = c * *y - s * *x;
    *x = t;
}

// Inspection of the first restart cycle of mixed_gmres (tests only): with
// g_gd.active set, the cycle's basis V (n x (restart+1)), the Hessenberg
// matrix before the Givens rotations (Hbar, (restart+1) x restart, column j =
// the CGS2 coefficients of step j and the norm of the new vector), |S(j+1)|
// after each step and the basis W = M^{-1} V are copied out, and the driver
// stops after that cycle.  cgs_passes = 1 plants a one-pass classical
// Gram-Schmidt (the reference reorthogonalises, gesv_mixed_gmres.cc:310-327),
// so that a test can show the orthogonality check detects it.
struct GmresDebug {
    bool active = false;
    int cgs_passes = 2;
    double* V = nullptr;
    double* W = nullptr;
    double* Hbar = nullptr;
    double* Sabs = nullptr;
    int64_t steps = 0;
};
GmresDebug g_gd;

// GMRES-IR, src/gesv_mixed_gmres.cc:111-410 (single RHS), and the same with a
// Cholesky preconditioner for posv_mixed_gmres.  The preconditioner applies
// the low-precision factor to a float-rounded vector (:282-286), the Arnoldi
// product and CGS2 (:289-327) run in double, Givens rotations on the host (:336-352).
int64_t mixed_gmres(int64_t n, double* A, int64_t lda, int sym, const double* B,
                    double* X, int64_t* ipiv, int64_t nb, int64_t ib, int64_t nthreads,
                    double threshold, int64_t itermax, double tol, int use_fallback, int* iter)
{
    const double eps = std::numeric_limits<double>::epsilon();
    if (tol <= 0) tol = eps * std::sqrt(double(n));
    int64_t restart = std::min<int64_t>(std::min<int64_t>(30, itermax), std::min(nb, n) - 1);
    double Anorm = sym ? norm_inf_sy_lower(n, A, lda) : norm_inf_ge(n, n, A, lda);
    double cte = Anorm * tol;
    LoFactor lo{n, nb, ib, nthreads, threshold, sym != 0, {}, {}};
    int64_t info = lo.factor(A, lda);
    bool converged = false;
    *iter = 0;
    int64_t R1 = restart + 1;
    std::vector<double> R(n), V(n * R1), W(n * R1), z(R1), H(R1 * R1, 0.0), S(R1, 0.0);
    std::vector<double> ga(restart), gb(restart);
    std::vector<double> nx(1), nr(1);
    if (info != 0) {
        *iter = -3;
    }
    else {
        for (int64_t i = 0; i < n; ++i) X[i] = B[i];
        lo.solve(1, X, n);
        int64_t iiter = 0;
        while (iiter < itermax) {
            residual(n, 1, A, lda, sym, X, n, B, n, R.data(), n);
            col_norms_max(n, 1, X, n, nx.data());
            col_norms_max(n, 1, R.data(), n, nr.data());
            if (ir_converged(nr, nx, cte)) { *iter = int(iiter); converged = true; break; }
            double beta = 0;
            for (int64_t i = 0; i < n; ++i) beta += R[i] * R[i];
            beta = std::sqrt(beta);
            std::vector<double> arn = {beta};
            if (beta == 0) { *iter = int(iiter); converged = false; break; }
            const double sc = 1.0 / beta;   // scale(1, beta): x *= numer/denom
            for (int64_t i = 0; i < n; ++i) V[i] = R[i] * sc;
            S[0] = beta;
            for (int64_t i = 1; i < R1; ++i) S[i] = 0;
            int64_t j = 0;
            for (; j < restart && iiter < itermax && !ir_converged(arn, nx, cte); ++j, ++iiter) {
                double* Vj = V.data() + j*n;
                double* Vj1 = V.data() + (j+1)*n;
                double* Wj1 = W.data() + (j+1)*n;
                for (int64_t i = 0; i < n; ++i) Wj1[i] = Vj[i];
                lo.solve(1, Wj1, n);
                residual(n, 1, A, lda, sym, Wj1, n, std::vector<double>(n, 0.0).data(), n, Vj1, n);
                for (int64_t i = 0; i < n; ++i) Vj1[i] = -Vj1[i];   // Vj1 = A * Wj1
                // CGS2 (gesv_mixed_gmres.cc:296-327)
                double* Hj = H.data() + j*R1;
                for (int64_t p = 0; p <= j; ++p) {
                    double s = 0; const double* Vp = V.data() + p*n;
                    for (int64_t i = 0; i < n; ++i) s += Vp[i] * Vj1[i];
                    Hj[p] = s;
                }
                for (int64_t p = 0; p <= j; ++p) {
                    const double* Vp = V.data() + p*n;
                    for (int64_t i = 0; i < n; ++i) Vj1[i] -= Vp[i] * Hj[p];
                }
                if (!(g_gd.active && g_gd.cgs_passes == 1)) {
                    for (int64_t p = 0; p <= j; ++p) {
                        double s = 0; const double* Vp = V.data() + p*n;
                        for (int64_t i = 0; i < n; ++i) s += Vp[i] * Vj1[i];
                        z[p] = s;
                    }
                    for (int64_t p = 0; p <= j; ++p) {
                        const double* Vp = V.data() + p*n;
                        for (int64_t i = 0; i < n; ++i) Vj1[i] -= Vp[i] * z[p];
                    }
                    for (int64_t p = 0; p <= j; ++p) Hj[p] += z[p];
                }
                double nrm = 0;
                for (int64_t i = 0; i < n; ++i) nrm += Vj1[i] * Vj1[i];
                nrm = std::sqrt(nrm);
                const double sc1 = 1.0 / nrm;
                for (int64_t i = 0; i < n; ++i) Vj1[i] *= sc1;
                Hj[j+1] = nrm;
                if (g_gd.active && g_gd.Hbar)
                    for (int64_t p = 0; p <= j + 1; ++p) g_gd.Hbar[p + j*R1] = Hj[p];
                for (int64_t i = 0; i < j; ++i) rot(&Hj[i], &Hj[i+1], ga[i], gb[i]);
                double hjj = Hj[j], hj1 = Hj[j+1];
                rotg(&hjj, &hj1, &ga[j], &gb[j]);
                rot(&Hj[j], &Hj[j+1], ga[j], gb[j]);
                rot(&S[j], &S[j+1], ga[j], gb[j]);
                arn[0] = std::fabs(S[j+1]);
                if (g_gd.active && g_gd.Sabs) g_gd.Sabs[j] = arn[0];
            }
            if (g_gd.active) {
                g_gd.steps = j;
                if (g_gd.V) std::copy(V.begin(), V.begin() + n * (j + 1), g_gd.V);
                if (g_gd.W) std::copy(W.begin() + n, W.begin() + n * (j + 1), g_gd.W);
                converged = true;
                break;
            }
            // X += W(:,1:j) * (H(0:j-1,0:j-1)^{-1} S(0:j-1)) (:360-374)
            std::vector<double> y(S.begin(), S.begin() + j);
            for (int64_t p = j-1; p >= 0; --p) {
                y[p] /= H[p + p*R1];
                for (int64_t q = 0; q < p; ++q) y[q] -= H[q + p*R1] * y[p];
            }
            for (int64_t p = 0; p < j; ++p) {
                const double* Wp = W.data() + (p+1)*n;
                for (int64_t i = 0; i < n; ++i) X[i] += Wp[i] * y[p];
            }
        }
    }
    if (!converged) {
        if (info == 0) *iter = int(-itermax - 1);
        if (use_fallback) {
            if (sym) info = potrf_tiles<double>(n, A, lda, nb);
            else info = getrf_tiles<double>(n, n, A, lda, nb, ib, nthreads, threshold, ipiv);
            if (info == 0) {
                for (int64_t i = 0; i < n; ++i) X[i] = B[i];
                if (sym) potrs_tiles<double>(n, 1, A, lda, X, n, nb);
                else getrs_tiles<double>(n, 1, A, lda, ipiv, X, n, nb, 0);
            }
        }
    }
    else if (!sym) {
        for (int64_t i = 0; i < n; ++i) ipiv[i] = lo.ipiv[i];
    }
    return info;
}

//------------------------------------------------------------------------------
// Random butterfly transforms and the RBT solver.
//
// Level k of a depth-d transform pairs index i with i + h for (i mod 2h) < h,
// h = 2^(d-k-1) * ceil(nt / 2^d) * nb, nt = ceil(n/nb): the tile ranges of
// src/gerbt.cc:60-100 (gerbt_iterate_2d / _1d) in rows.  The element formulas
// are those of src/internal/Tile_gerbt.hh:18-96 (two-sided) and :98-230
// (one-sided), with the same single weight vector per side at every level
// (column 0 of the reference's n x d weight matrix).  An index whose partner
// lies beyond n is scaled by its weight only, in both the one- and two-sided
// forms (the reference's one-sided pass skips whole partnerless tiles).
int64_t rbt_half(int64_t n, int64_t nb, int64_t d, int64_t k)
{
    const int64_t nt = (n + nb - 1) / nb;
    const int64_t inner = (nt + (int64_t(1) << d) - 1) >> d;
    return (int64_t(1) << (d - k - 1)) * inner * nb;
}

// u[i] = exp(r_i / 20) (internal_rbt_generate.cc:45-51), r_i = matgen rands
// entry (seed, i, 0) in place of LAPACK larnv.
void rbt_weights(int64_t seed, int64_t n, double* u)
{
    for (int64_t i = 0; i < n; ++i) u[i] = std::exp(gen_one(2, seed, i, 0) / 20.0);
}

// A := U^T A V, levels smallest first (gerbt.cc:60-80, Tile_gerbt.hh:18-96).
void gerbt2(int64_t n, int64_t nb, int64_t d, const double* u, const double* v, double* A, int64_t lda)
{
    const double half = 0.5, s2 = 1.0 / std::sqrt(2.0);
    for (int64_t k = d - 1; k >= 0; --k) {
        const int64_t h = rbt_half(n, nb, d, k);
        for (int64_t j1 = 0; j1 < n; ++j1) {
            if (j1 % (2 * h) >= h) continue;
            const int64_t j2 = j1 + h;
            const bool hc = j2 < n;
            for (int64_t i1 = 0; i1 < n; ++i1) {
                if (i1 % (2 * h) >= h) continue;
                const int64_t i2 = i1 + h;
                const bool hr = i2 < n;
                double& a11 = A[i1 + j1*lda];
                const double u1 = u[i1], v1 = v[j1];
                if (hr && hc) {
                    double& a12 = A[i1 + j2*lda];
                    double& a21 = A[i2 + j1*lda];
                    double& a22 = A[i2 + j2*lda];
                    const double s1 = a11 + a12, t1 = a21 + a22, d1 = a11 - a12, e1 = a21 - a22;
                    a11 = half * u1 * (s1 + t1) * v1;
                    a12 = half * u1 * (d1 + e1) * v[j2];
                    a21 = half * u[i2] * (s1 - t1) * v1;
                    a22 = half * u[i2] * (d1 - e1) * v[j2];
                }
                else if (hc) {
                    double& a12 = A[i1 + j2*lda];
                    const double x = a11, y = a12;
                    a11 = s2 * u1 * (x + y) * v1;
                    a12 = s2 * u1 * (x - y) * v[j2];
                }
                else if (hr) {
                    double& a21 = A[i2 + j1*lda];
                    const double x = a11, y = a21;
                    a11 = s2 * u1 * (x + y) * v1;
                    a21 = s2 * u[i2] * (x - y) * v1;
                }
                else {
                    a11 = u1 * a11 * v1;
                }
            }
        }
    }
}

// B := U^T B (trans, levels smallest first) or B := V B (largest first):
// gerbt.cc:85-100, Tile_gerbt.hh gerbt_left_trans / gerbt_left_notrans.
void gerbt1(int trans, int64_t n, int64_t nrhs, int64_t nb, int64_t d, const double* u,
            double* B, int64_t ldb)
{
    const double s2 = 1.0 / std::sqrt(2.0);
    for (int64_t it = 0; it < d; ++it) {
        const int64_t k = trans ? d - 1 - it : it;
        const int64_t h = rbt_half(n, nb, d, k);
        for (int64_t c = 0; c < nrhs; ++c) {
            double* b = B + c*ldb;
            for (int64_t i1 = 0; i1 < n; ++i1) {
                if (i1 % (2 * h) >= h) continue;
                const int64_t i2 = i1 + h;
                if (i2 >= n) { b[i1] = u[i1] * b[i1]; continue; }
                const double x = b[i1], y = b[i2];
                if (trans) {
                    b[i1] = s2 * u[i1] * (x + y);
                    b[i2] = s2 * u[i2] * (x - y);
                }
                else {
                    b[i1] = s2 * (u[i1] * x + u[i2] * y);
                    b[i2] = s2 * (u[i1] * x - u[i2] * y);
                }
            }
        }
    }
}

// src/gesv_rbt.cc:94-236: transform, LU without pivoting (pivot threshold 0
// keeps every diagonal, Tile_getrf.hh:256-289), solve, refine in FP64 against
// the original A, PP gesv of the copy as fallback.
int64_t gesv_rbt(int64_t n, int64_t nrhs, double* A, int64_t lda, const double* B, int64_t ldb,
                 double* X, int64_t ldx, int64_t nb, int64_t ib, int64_t depth, int64_t itermax,
                 double tol, int use_fallback, int* iter)
{
    const double eps = std::numeric_limits<double>::epsilon();
    if (tol <= 0) tol = eps * std::sqrt(double(n));
    std::vector<double> u(n), v(n), Acopy, R(n * nrhs), nx(nrhs), nr(nrhs);
    std::vector<int64_t> ipiv(n);
    rbt_weights(42, n, u.data());
    rbt_weights(42 + (n + nb - 1) / nb, n, v.data());
    double Anorm = 0;
    if (itermax > 0) {
        Acopy.resize(n * n);
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) Acopy[i + j*n] = A[i + j*lda];
        Anorm = norm_inf_ge(n, n, A, lda);
    }
    for (int64_t c = 0; c < nrhs; ++c)
        for (int64_t i = 0; i < n; ++i) X[i + c*ldx] = B[i + c*ldb];
    *iter = 0;
    gerbt2(n, nb, depth, u.data(), v.data(), A, lda);
    int64_t info = getrf_tiles<double>(n, n, A, lda, nb, ib, 1, 0.0, ipiv.data());
    if (info == 0) {
        gerbt1(1, n, nrhs, nb, depth, u.data(), X, ldx);
        getrs_tiles<double>(n, nrhs, A, lda, ipiv.data(), X, ldx, nb, 0);
        gerbt1(0, n, nrhs, nb, depth, v.data(), X, ldx);
    }
    if (itermax == 0) return info;
    const double cte = Anorm * tol;
    bool converged = false;
    if (info == 0) {
        residual(n, nrhs, Acopy.data(), n, 0, X, ldx, B, ldb, R.data(), n);
        col_norms_max(n, nrhs, X, ldx, nx.data());
        col_norms_max(n, nrhs, R.data(), n, nr.data());
        if (ir_converged(nr, nx, cte)) converged = true;
        for (int64_t it = 0; it < itermax && !converged; ++it) {
            gerbt1(1, n, nrhs, nb, depth, u.data(), R.data(), n);
            getrs_tiles<double>(n, nrhs, A, lda, ipiv.data(), R.data(), n, nb, 0);
            gerbt1(0, n, nrhs, nb, depth, v.data(), R.data(), n);
            for (int64_t c = 0; c < nrhs; ++c)
                for (int64_t i = 0; i < n; ++i) X[i + c*ldx] += R[i + c*n];
            residual(n, nrhs, Acopy.data(), n, 0, X, ldx, B, ldb, R.data(), n);
            col_norms_max(n, nrhs, X, ldx, nx.data());
            col_norms_max(n, nrhs, R.data(), n, nr.data());
            if (ir_converged(nr, nx, cte)) { *iter = int(it + 1); converged = true; }
        }
    }
    if (!converged) {
        *iter = int(-itermax - 1);
        if (use_fallback) {
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = 0; i < n; ++i) A[i + j*lda] = Acopy[i + j*n];
            info = getrf_tiles<double>(n, n, Acopy.data(), n, nb, ib, 1, 1.0, ipiv.data());
            for (int64_t c = 0; c < nrhs; ++c)
                for (int64_t i = 0; i < n; ++i) X[i + c*ldx] = B[i + c*ldb];
            if (info == 0) getrs_tiles<double>(n, nrhs, Acopy.data(), n, ipiv.data(), X, ldx, nb, 0);
        }
    }
    return info;
}

} // namespace

//==============================================================================
extern "C" {

// Route the tile gemm/trsm through the OpenBLAS at `path` (scipy's bundled
// LP64 build, symbols scipy_dgemm_ ...), one BLAS thread per OpenMP task.
// path = NULL unloads.  Returns 0, or -1 if a symbol is missing.
int oracle_set_blas(const char* path)
{
    g_blas = Blas{};
    if (!path) return 0;
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    g_blas.dgemm = reinterpret_cast<gemm_d_t>(dlsym(h, "scipy_dgemm_"));
    g_blas.sgemm = reinterpret_cast<gemm_s_t>(dlsym(h, "scipy_sgemm_"));
    g_blas.dtrsm = reinterpret_cast<trsm_d_t>(dlsym(h, "scipy_dtrsm_"));
    g_blas.strsm = reinterpret_cast<trsm_s_t>(dlsym(h, "scipy_strsm_"));
    auto set_threads = reinterpret_cast<void (*)(int)>(dlsym(h, "scipy_openblas_set_num_threads"));
    if (set_threads) set_threads(1);
    if (!g_blas.dgemm || !g_blas.sgemm || !g_blas.dtrsm || !g_blas.strsm || !set_threads) {
        g_blas = Blas{};
        return -1;
    }
    return 0;
}

// Low-precision trailing updates of the float factor: kind 0 float, 1 split
// bf16 (terms 1..3), 2 FP8 e4m3 with per-tile power-of-two scales (terms
// 1..2, tiles of qrows rows/columns).
int oracle_set_update_precision(int kind, int terms, int64_t qrows)
{
    if (kind < 0 || kind > 2) return -1;
    if (kind == 1 && (terms < 1 || terms > 3)) return -1;
    if (kind == 2 && (terms < 1 || terms > 2)) return -1;
    if (qrows < 1) return -1;
    g_q.kind = kind; g_q.terms = kind ? terms : 0; g_q.qrows = qrows;
    return 0;
}

// C (m x n, float, column-major) -= A B with the current update precision
// (A m x k, B k x n, both float column-major); kind 0 = plain float gemm.
void oracle_sgemm_update(int64_t m, int64_t n, int64_t k, const float* A, int64_t lda, const float* B,
                         int64_t ldb, float* C, int64_t ldc)
{
    if (g_q.kind != 0) gemm_quantized<float>(m, n, k, A, lda, B, ldb, true, C, ldc);
    else gemm_nn<float>(m, n, k, -1.0f, A, lda, B, ldb, C, ldc);
}

double oracle_e4m3_decode(int b) { return e4m3_decode(uint8_t(b)); }
double oracle_e4m3_round(double x) { return e4m3_round(x); }
int oracle_e4m3_encode(double x) { return e4m3_encode(x); }
int oracle_e4m3_scale_exp(double amax) { return e4m3_scale_exp(amax); }
float oracle_bf16_rne(float x) { return bf16_rne(x); }

// The oracle's gesv_mixed_gmres (sym 0) / posv_mixed_gmres (sym 1, lower)
// stopped after its first restart cycle; see GmresDebug.  Returns the number
// of Arnoldi steps j; V gets n x (j+1), W n x j, Hbar (restart+1) x restart,
// Sabs j values.
int64_t oracle_gmres_first_cycle(int64_t n, double* A, int64_t lda, int sym, const double* B, double* X,
                                 int64_t nb, int64_t itermax, int cgs_passes, double* V, double* W,
                                 double* Hbar, double* Sabs)
{
    g_gd = GmresDebug{};
    g_gd.active = true; g_gd.cgs_passes = cgs_passes;
    g_gd.V = V; g_gd.W = W; g_gd.Hbar = Hbar; g_gd.Sabs = Sabs;
    std::vector<int64_t> ipiv(n);
    int it = 0;
    mixed_gmres(n, A, lda, sym, B, X, ipiv.data(), nb, 16, 1, 1.0, itermax, 0.0, 0, &it);
    const int64_t steps = g_gd.steps;
    g_gd = GmresDebug{};
    return steps;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_num_threads(int t) { omp_set_num_threads(t); }

void oracle_generate(int dist, int64_t seed, int64_t m, int64_t n,
                     int64_t ioff, int64_t joff, double* A, int64_t lda)
{
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) A[i + j*lda] = gen_one(dist, seed, i + ioff, j + joff);
}

void oracle_philox(uint64_t s0, uint64_t s1, uint64_t seed, uint64_t* out)
{
    philox(s0, s1, seed, &out[0], &out[1]);
}

int64_t oracle_dgetrf(int64_t m, int64_t n, double* A, int64_t lda, int64_t nb, int64_t ib,
                      int64_t nthreads, double threshold, int64_t* ipiv)
{ return getrf_tiles<double>(m, n, A, lda, nb, ib, nthreads, threshold, ipiv); }

int64_t oracle_sgetrf(int64_t m, int64_t n, float* A, int64_t lda, int64_t nb, int64_t ib,
                      int64_t nthreads, double threshold, int64_t* ipiv)
{ return getrf_tiles<float>(m, n, A, lda, nb, ib, nthreads, threshold, ipiv); }

void oracle_dgetrs(int64_t n, int64_t nrhs, const double* A, int64_t lda, const int64_t* ipiv,
                   double* B, int64_t ldb, int64_t nb, int trans)
{ getrs_tiles<double>(n, nrhs, A, lda, ipiv, B, ldb, nb, trans); }

void oracle_sgetrs(int64_t n, int64_t nrhs, const float* A, int64_t lda, const int64_t* ipiv,
                   float* B, int64_t ldb, int64_t nb, int trans)
{ getrs_tiles<float>(n, nrhs, A, lda, ipiv, B, ldb, nb, trans); }

int64_t oracle_dgesv(int64_t n, int64_t nrhs, double* A, int64_t lda, int64_t* ipiv,
                     double* B, int64_t ldb, int64_t nb, int64_t ib, int64_t nthreads)
{
    // src/gesv.cc:85-109
    int64_t info = getrf_tiles<double>(n, n, A, lda, nb, ib, nthreads, 1.0, ipiv);
    if (info == 0) getrs_tiles<double>(n, nrhs, A, lda, ipiv, B, ldb, nb, 0);
    return info;
}

int64_t oracle_dpotrf(int64_t n, double* A, int64_t lda, int64_t nb)
{ return potrf_tiles<double>(n, A, lda, nb); }

int64_t oracle_spotrf(int64_t n, float* A, int64_t lda, int64_t nb)
{ return potrf_tiles<float>(n, A, lda, nb); }

void oracle_dpotrs(int64_t n, int64_t nrhs, const double* A, int64_t lda,
                   double* B, int64_t ldb, int64_t nb)
{ potrs_tiles<double>(n, nrhs, A, lda, B, ldb, nb); }

int64_t oracle_dposv(int64_t n, int64_t nrhs, double* A, int64_t lda,
                     double* B, int64_t ldb, int64_t nb)
{
    // src/posv.cc:71-94
    int64_t info = potrf_tiles<double>(n, A, lda, nb);
    if (info == 0) potrs_tiles<double>(n, nrhs, A, lda, B, ldb, nb);
    return info;
}

int64_t oracle_dsgesv(int64_t n, int64_t nrhs, double* A, int64_t lda, int64_t* ipiv,
                      const double* B, int64_t ldb, double* X, int64_t ldx,
                      int64_t nb, int64_t ib, int64_t nthreads, int64_t itermax,
                      double tol, int use_fallback, int* iter)
{ return mixed_ir(n, nrhs, A, lda, 0, B, ldb, X, ldx, ipiv, nb, ib, nthreads, 1.0,
                  itermax, tol, use_fallback, iter); }

int64_t oracle_dsposv(int64_t n, int64_t nrhs, double* A, int64_t lda,
                      const double* B, int64_t ldb, double* X, int64_t ldx,
                      int64_t nb, int64_t itermax, double tol, int use_fallback, int* iter)
{ return mixed_ir(n, nrhs, A, lda, 1, B, ldb, X, ldx, nullptr, nb, 16, 1, 1.0,
                  itermax, tol, use_fallback, iter); }

int64_t oracle_dsgesv_gmres(int64_t n, double* A, int64_t lda, int64_t* ipiv,
                            const double* B, double* X, int64_t nb, int64_t ib,
                            int64_t nthreads, int64_t itermax, double tol,
                            int use_fallback, int* iter)
{ return mixed_gmres(n, A, lda, 0, B, X, ipiv, nb, ib, nthreads, 1.0, itermax, tol,
                     use_fallback, iter); }

int64_t oracle_dsposv_gmres(int64_t n, double* A, int64_t lda, const double* B, double* X,
                            int64_t nb, int64_t itermax, double tol, int use_fallback,
                            int* iter)
{ return mixed_gmres(n, A, lda, 1, B, X, nullptr, nb, 16, 1, 1.0, itermax, tol,
                     use_fallback, iter); }

double oracle_norm_inf(int64_t m, int64_t n, const double* A, int64_t lda)
{ return norm_inf_ge(m, n, A, lda); }

double oracle_norm_inf_sy(int64_t n, const double* A, int64_t lda)
{ return norm_inf_sy_lower(n, A, lda); }

void oracle_residual(int64_t n, int64_t nrhs, const double* A, int64_t lda, int sym,
                     const double* X, int64_t ldx, const double* B, int64_t ldb,
                     double* R, int64_t ldr)
{ residual(n, nrhs, A, lda, sym, X, ldx, B, ldb, R, ldr); }

void oracle_rotg(double* a, double* b, double* c, double* s) { rotg(a, b, c, s); }

void oracle_rbt_weights(int64_t seed, int64_t n, double* u) { rbt_weights(seed, n, u); }

int64_t oracle_rbt_half(int64_t n, int64_t nb, int64_t d, int64_t k) { return rbt_half(n, nb, d, k); }

void oracle_gerbt(int64_t n, int64_t nb, int64_t d, const double* u, const double* v,
                  double* A, int64_t lda)
{ gerbt2(n, nb, d, u, v, A, lda); }

void oracle_gerbt_rhs(int trans, int64_t n, int64_t nrhs, int64_t nb, int64_t d, const double* u,
                      double* B, int64_t ldb)
{ gerbt1(trans, n, nrhs, nb, d, u, B, ldb); }

int64_t oracle_dgesv_rbt(int64_t n, int64_t nrhs, double* A, int64_t lda, const double* B, int64_t ldb,
                         double* X, int64_t ldx, int64_t nb, int64_t ib, int64_t depth, int64_t itermax,
                         double tol, int use_fallback, int* iter)
{ return gesv_rbt(n, nrhs, A, lda, B, ldb, X, ldx, nb, ib, depth, itermax, tol, use_fallback, iter); }

} // extern "C"
