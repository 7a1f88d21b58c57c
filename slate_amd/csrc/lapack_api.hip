// LAPACK-compatible entry points (Fortran calling convention: every argument
// by pointer, host column-major arrays, 32-bit integers, 1-based ipiv).
//
// Mirrors SLATE's compatibility API (lapack_api/lapack_gesv.cc:32-100,
// lapack_gesv_mixed.cc:24-67, lapack_getrf.cc, lapack_getrs.cc,
// lapack_posv.cc, lapack_potrf.cc): the same slate_<x><routine>_ names, the
// LAPACK argument order, and the SLATE_LAPACK_NB / SLATE_LAPACK_IB
// environment variables of lapack_api/lapack_slate.hh.  The host arrays are
// staged through device memory around one call of the C ABI.  Illegal
// arguments return info = -i (i = position of the argument), as LAPACK does;
// a factorization failure returns the positive info of the driver.
#include "common.hh"
#include "../../include/slate_amd.h"
#include "../../include/slate_amd_lapack.h"

#include <cctype>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace {

slate_amd_context lapack_context()
{
    static std::mutex mu;
    static slate_amd_context ctx[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx[dev] && slate_amd_init(&ctx[dev], dev) < 0) return nullptr;
    return ctx[dev];
}

int64_t env_int(const char* name)
{
    const char* s = std::getenv(name);
    return s ? std::strtol(s, nullptr, 0) : 0;
}

slate_amd_options lapack_options()
{
    slate_amd_options o;
    slate_amd_options_init(&o);
    o.block_size = env_int("SLATE_LAPACK_NB");     // 0: the library's own choice
    o.inner_blocking = env_int("SLATE_LAPACK_IB");
    return o;
}

// Device copy of a host column-major m x n array (leading dimension ld).
template <class T>
struct Staged {
    T* d = nullptr;
    int64_t m = 0, n = 0, ld = 0;
    T* h = nullptr;
    int64_t hld = 0;
    bool ok = false;
    Staged(T* host, int64_t m_, int64_t n_, int64_t hld_, bool copy_in) : m(m_), n(n_), h(host), hld(hld_)
    {
        ld = std::max<int64_t>(1, m);
        size_t bytes = sizeof(T) * size_t(ld) * size_t(std::max<int64_t>(1, n));
        if (hipMalloc(reinterpret_cast<void**>(&d), bytes) != hipSuccess) { d = nullptr; return; }
        ok = true;
        if (copy_in && m > 0 && n > 0)
            ok = hipMemcpy2D(d, sizeof(T) * ld, h, sizeof(T) * hld, sizeof(T) * m, n,
                             hipMemcpyHostToDevice) == hipSuccess;
    }
    bool back()
    {
        if (m <= 0 || n <= 0) return true;
        return hipMemcpy2D(h, sizeof(T) * hld, d, sizeof(T) * ld, sizeof(T) * m, n,
                           hipMemcpyDeviceToHost) == hipSuccess;
    }
    ~Staged() { if (d) (void)hipFree(d); }
};

// Negative C-ABI codes that are not argument errors are reported as
// info = -1000 + code (never a LAPACK argument position).
int map_error(int64_t rc) { return rc >= 0 ? int(rc) : int(-1000 + rc); }

bool is_uplo(char c) { c = char(std::toupper(c)); return c == 'L' || c == 'U'; }
int uplo_of(char c) { return std::toupper(c) == 'L' ? SLATE_AMD_LOWER : SLATE_AMD_UPPER; }

// ---- getrf ----------------------------------------------------------------
template <class T, class F>
void getrf_t(F fn, int m, int n, T* a, int lda, int* ipiv, int* info)
{
    *info = 0;
    if (m < 0) { *info = -1; return; }
    if (n < 0) { *info = -2; return; }
    if (lda < std::max(1, m)) { *info = -4; return; }
    if (m != n) { *info = -1; return; }          // square factorizations only
    if (n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<T> A(a, n, n, lda, true);
    if (!ctx || !A.ok) { *info = -1000 - 100; return; }
    std::vector<int64_t> piv(static_cast<size_t>(n), 0);
    auto o = lapack_options();
    int64_t rc = fn(ctx, n, A.d, A.ld, piv.data(), &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back()) { *info = -1000 - 100; return; }
    for (int i = 0; i < n; ++i) ipiv[i] = int(piv[size_t(i)]);
    *info = int(rc);
}

// ---- getrs ----------------------------------------------------------------
template <class T, class F>
void getrs_t(F fn, char trans, int n, int nrhs, const T* a, int lda, const int* ipiv, T* b, int ldb, int* info)
{
    *info = 0;
    char t = char(std::toupper(trans));
    if (t != 'N' && t != 'T' && t != 'C') { *info = -1; return; }
    if (n < 0) { *info = -2; return; }
    if (nrhs < 0) { *info = -3; return; }
    if (lda < std::max(1, n)) { *info = -5; return; }
    if (ldb < std::max(1, n)) { *info = -8; return; }
    if (n == 0 || nrhs == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<T> A(const_cast<T*>(a), n, n, lda, true);
    Staged<T> B(b, n, nrhs, ldb, true);
    if (!ctx || !A.ok || !B.ok) { *info = -1000 - 100; return; }
    std::vector<int64_t> piv(ipiv, ipiv + n);
    auto o = lapack_options();
    int64_t rc = fn(ctx, t == 'N' ? SLATE_AMD_NOTRANS : SLATE_AMD_TRANS, n, nrhs, A.d, A.ld, piv.data(), B.d,
                    B.ld, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!B.back()) *info = -1000 - 100;
}

// ---- potrf ----------------------------------------------------------------
template <class T, class F>
void potrf_t(F fn, char uplo, int n, T* a, int lda, int* info)
{
    *info = 0;
    if (!is_uplo(uplo)) { *info = -1; return; }
    if (n < 0) { *info = -2; return; }
    if (lda < std::max(1, n)) { *info = -4; return; }
    if (n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<T> A(a, n, n, lda, true);
    if (!ctx || !A.ok) { *info = -1000 - 100; return; }
    auto o = lapack_options();
    int64_t rc = fn(ctx, uplo_of(uplo), n, A.d, A.ld, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back()) { *info = -1000 - 100; return; }
    *info = int(rc);
}

} // namespace

extern "C" {

void slate_dgetrf_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info)
{ getrf_t<double>(slate_amd_dgetrf, *m, *n, a, *lda, ipiv, info); }

void slate_sgetrf_(const int* m, const int* n, float* a, const int* lda, int* ipiv, int* info)
{ getrf_t<float>(slate_amd_sgetrf, *m, *n, a, *lda, ipiv, info); }

void slate_dgetrs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda,
                   const int* ipiv, double* b, const int* ldb, int* info)
{ getrs_t<double>(slate_amd_dgetrs, *trans, *n, *nrhs, a, *lda, ipiv, b, *ldb, info); }

void slate_sgetrs_(const char* trans, const int* n, const int* nrhs, const float* a, const int* lda,
                   const int* ipiv, float* b, const int* ldb, int* info)
{ getrs_t<float>(slate_amd_sgetrs, *trans, *n, *nrhs, a, *lda, ipiv, b, *ldb, info); }

// lapack_api/lapack_gesv.cc:32-35
void slate_dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
                  const int* ldb, int* info)
{
    *info = 0;
    if (*n < 0) { *info = -1; return; }
    if (*nrhs < 0) { *info = -2; return; }
    if (*lda < std::max(1, *n)) { *info = -4; return; }
    if (*ldb < std::max(1, *n)) { *info = -7; return; }
    if (*n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<double> A(a, *n, *n, *lda, true);
    Staged<double> B(b, *n, *nrhs, *ldb, true);
    if (!ctx || !A.ok || !B.ok) { *info = -1000 - 100; return; }
    std::vector<int64_t> piv(static_cast<size_t>(*n), 0);
    auto o = lapack_options();
    int64_t rc = slate_amd_dgesv(ctx, *n, *nrhs, A.d, A.ld, piv.data(), B.d, B.ld, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back() || !B.back()) { *info = -1000 - 100; return; }
    for (int i = 0; i < *n; ++i) ipiv[i] = int(piv[size_t(i)]);
    *info = int(rc);
}

// LAPACK dsgesv argument list (lapack_api/lapack_gesv_mixed.cc:24); work and
// swork are accepted for compatibility and not used (the device holds the
// workspace).  iter as in LAPACK: >= 0 refinement steps, < 0 the FP64
// fallback ran.
void slate_dsgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
                   const int* ldb, double* x, const int* ldx, double* work, float* swork, int* iter, int* info)
{
    (void)work; (void)swork;
    *info = 0;
    *iter = 0;
    if (*n < 0) { *info = -1; return; }
    if (*nrhs < 0) { *info = -2; return; }
    if (*lda < std::max(1, *n)) { *info = -4; return; }
    if (*ldb < std::max(1, *n)) { *info = -7; return; }
    if (*ldx < std::max(1, *n)) { *info = -9; return; }
    if (*n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<double> A(a, *n, *n, *lda, true);
    Staged<double> B(b, *n, *nrhs, *ldb, true);
    Staged<double> X(x, *n, *nrhs, *ldx, false);
    if (!ctx || !A.ok || !B.ok || !X.ok) { *info = -1000 - 100; return; }
    std::vector<int64_t> piv(static_cast<size_t>(*n), 0);
    auto o = lapack_options();
    int it = 0;
    int64_t rc = slate_amd_dsgesv(ctx, *n, *nrhs, A.d, A.ld, piv.data(), B.d, B.ld, X.d, X.ld, &it, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back() || !X.back()) { *info = -1000 - 100; return; }
    for (int i = 0; i < *n; ++i) ipiv[i] = int(piv[size_t(i)]);
    *iter = it;
    *info = int(rc);
}

void slate_dpotrf_(const char* uplo, const int* n, double* a, const int* lda, int* info)
{ potrf_t<double>(slate_amd_dpotrf, *uplo, *n, a, *lda, info); }

void slate_spotrf_(const char* uplo, const int* n, float* a, const int* lda, int* info)
{ potrf_t<float>(slate_amd_spotrf, *uplo, *n, a, *lda, info); }

void slate_dpotrs_(const char* uplo, const int* n, const int* nrhs, const double* a, const int* lda, double* b,
                   const int* ldb, int* info)
{
    *info = 0;
    if (!is_uplo(*uplo)) { *info = -1; return; }
    if (*n < 0) { *info = -2; return; }
    if (*nrhs < 0) { *info = -3; return; }
    if (*lda < std::max(1, *n)) { *info = -5; return; }
    if (*ldb < std::max(1, *n)) { *info = -7; return; }
    if (*n == 0 || *nrhs == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<double> A(const_cast<double*>(a), *n, *n, *lda, true);
    Staged<double> B(b, *n, *nrhs, *ldb, true);
    if (!ctx || !A.ok || !B.ok) { *info = -1000 - 100; return; }
    auto o = lapack_options();
    int64_t rc = slate_amd_dpotrs(ctx, uplo_of(*uplo), *n, *nrhs, A.d, A.ld, B.d, B.ld, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!B.back()) *info = -1000 - 100;
}

void slate_dposv_(const char* uplo, const int* n, const int* nrhs, double* a, const int* lda, double* b,
                  const int* ldb, int* info)
{
    *info = 0;
    if (!is_uplo(*uplo)) { *info = -1; return; }
    if (*n < 0) { *info = -2; return; }
    if (*nrhs < 0) { *info = -3; return; }
    if (*lda < std::max(1, *n)) { *info = -5; return; }
    if (*ldb < std::max(1, *n)) { *info = -7; return; }
    if (*n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<double> A(a, *n, *n, *lda, true);
    Staged<double> B(b, *n, *nrhs, *ldb, true);
    if (!ctx || !A.ok || !B.ok) { *info = -1000 - 100; return; }
    auto o = lapack_options();
    int64_t rc = slate_amd_dposv(ctx, uplo_of(*uplo), *n, *nrhs, A.d, A.ld, B.d, B.ld, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back() || !B.back()) { *info = -1000 - 100; return; }
    *info = int(rc);
}

// LAPACK dsposv argument list.
void slate_dsposv_(const char* uplo, const int* n, const int* nrhs, double* a, const int* lda, double* b,
                   const int* ldb, double* x, const int* ldx, double* work, float* swork, int* iter, int* info)
{
    (void)work; (void)swork;
    *info = 0;
    *iter = 0;
    if (!is_uplo(*uplo)) { *info = -1; return; }
    if (*n < 0) { *info = -2; return; }
    if (*nrhs < 0) { *info = -3; return; }
    if (*lda < std::max(1, *n)) { *info = -5; return; }
    if (*ldb < std::max(1, *n)) { *info = -7; return; }
    if (*ldx < std::max(1, *n)) { *info = -9; return; }
    if (*n == 0) return;
    slate_amd_context ctx = lapack_context();
    Staged<double> A(a, *n, *n, *lda, true);
    Staged<double> B(b, *n, *nrhs, *ldb, true);
    Staged<double> X(x, *n, *nrhs, *ldx, false);
    if (!ctx || !A.ok || !B.ok || !X.ok) { *info = -1000 - 100; return; }
    auto o = lapack_options();
    int it = 0;
    int64_t rc = slate_amd_dsposv(ctx, uplo_of(*uplo), *n, *nrhs, A.d, A.ld, B.d, B.ld, X.d, X.ld, &it, &o);
    if (rc < 0) { *info = map_error(rc); return; }
    if (!A.back() || !X.back()) { *info = -1000 - 100; return; }
    *iter = it;
    *info = int(rc);
}

} // extern "C"
