// slate_amd.hh — C++ mirror of the slate:: factor-and-solve interface over the
// C ABI of include/slate_amd.h.
//
// The types and drivers keep the names, argument meaning, return values and
// error behaviour of include/slate/slate.hh so that code written against
//   slate::gesv(A, pivots, B, opts)  (include/slate/slate.hh:504-508)
// compiles against slate_amd:: with the namespace changed.  Matrices are
// single-device, 1 x 1 grid: fromDevices() wraps a dense column-major device
// array (tile (i, j) at A + i*nb + j*nb*lda, exactly the tiles
// slate::Matrix<T>::fromDevices builds there, include/slate/Matrix.hh:397),
// fromLAPACK() wraps host memory, which the drivers stage through device
// memory (Target::Devices semantics of src/getrf.cc:54-55 for host-origin
// tiles).  Errors surface as slate_amd::Exception, as slate:: raises
// slate::Exception (include/slate/Exception.hh:16-120); `info` is returned.
#ifndef SLATE_AMD_HH
#define SLATE_AMD_HH

#include "slate_amd.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace slate_amd {

//------------------------------------------------------------------------------
// Exceptions (include/slate/Exception.hh:16-120).
class Exception : public std::exception {
public:
    explicit Exception(std::string msg, int64_t code = -1) : msg_(std::move(msg)), code_(code) {}
    const char* what() const noexcept override { return msg_.c_str(); }
    int64_t code() const { return code_; }
private:
    std::string msg_;
    int64_t code_;
};
class NotImplemented : public Exception {
public:
    explicit NotImplemented(const std::string& what) : Exception("not implemented: " + what, -2) {}
};

//------------------------------------------------------------------------------
// Enums and options (include/slate/enums.hh, include/slate/types.hh:32-80,
// :241-265).  Only the values used on the factor-and-solve path.
enum class Target : char { Host = 'H', HostTask = 'T', HostNest = 'N', HostBatch = 'B', Devices = 'D' };
enum class Uplo : char { Upper = 'U', Lower = 'L', General = 'G' };
enum class Op : char { NoTrans = 'N', Trans = 'T', ConjTrans = 'C' };

// Factorization precision of the mixed drivers; SLATE picks it with the
// scalar_lo template argument, the BF16 choice has no SLATE counterpart.
enum class FactorPrecision : int {
    FP32 = SLATE_AMD_PREC_FP32, BF16 = SLATE_AMD_PREC_BF16, BF16X2 = SLATE_AMD_PREC_BF16X2,
    BF16X3 = SLATE_AMD_PREC_BF16X3, FP8 = SLATE_AMD_PREC_FP8, FP8X2 = SLATE_AMD_PREC_FP8X2,
    FP64 = SLATE_AMD_PREC_FP64
};

enum class MethodLU : char { PartialPiv = 'P', CALU = 'C', NoPiv = 'N' };   // include/slate/enums.hh:294-301

enum class Option : char {
    Lookahead, InnerBlocking, MaxPanelThreads, PivotThreshold, Tolerance, MaxIterations,
    UseFallbackSolver, Target, HoldLocalWorkspace, FactorPrecision, MethodLU, Depth,
    UseGraph,   // replay the LU step sequence from a recorded HIP graph (no counterpart in slate::Option)
};

class OptionValue {
public:
    OptionValue() : i_(0) {}
    OptionValue(int i) : i_(i) {}
    OptionValue(int64_t i) : i_(i) {}
    OptionValue(double d) : d_(d) {}
    OptionValue(bool b) : i_(b) {}
    OptionValue(slate_amd::Target t) : i_(int64_t(t)) {}
    OptionValue(slate_amd::FactorPrecision p) : i_(int64_t(p)) {}
    OptionValue(slate_amd::MethodLU m) : i_(int64_t(m)) {}
    int64_t i_;
    double d_;
};
using Options = std::map<Option, OptionValue>;

//------------------------------------------------------------------------------
// Pivots (include/slate/types.hh:84-117): per block column k, one entry per
// row with the tile index relative to k and the offset inside the tile.
class Pivot {
public:
    Pivot() : tile_index_(0), element_offset_(0) {}
    Pivot(int64_t tile_index, int64_t element_offset)
        : tile_index_(tile_index), element_offset_(element_offset) {}
    int64_t tileIndex() const { return tile_index_; }
    int64_t elementOffset() const { return element_offset_; }
    bool operator==(const Pivot& o) const
    { return tile_index_ == o.tile_index_ && element_offset_ == o.element_offset_; }
private:
    int64_t tile_index_, element_offset_;
};
using Pivots = std::vector<std::vector<Pivot>>;

namespace detail {

inline void check(int64_t rc)
{
    if (rc < 0) throw Exception(slate_amd_last_error(), rc);
}

// One context per device, created on first use and kept for the process.
inline slate_amd_context context(int device)
{
    static std::mutex mu;
    static std::map<int, slate_amd_context> ctx;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ctx.find(device);
    if (it != ctx.end()) return it->second;
    slate_amd_context c = nullptr;
    check(slate_amd_init(&c, device));
    ctx[device] = c;
    return c;
}

inline int current_device()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) throw Exception("hipGetDevice failed");
    return d;
}

// Pivots <-> LAPACK ipiv (1-based global rows), the conversion of
// lapack_api/lapack_gesv.cc:90-100.
inline Pivots pivots_from_ipiv(const std::vector<int64_t>& ipiv, int64_t nb)
{
    const int64_t n = int64_t(ipiv.size());
    Pivots piv;
    for (int64_t k0 = 0, k = 0; k0 < n; k0 += nb, ++k) {
        std::vector<Pivot> col;
        for (int64_t i = k0; i < std::min(n, k0 + nb); ++i) {
            int64_t row = ipiv[size_t(i)] - 1;
            col.emplace_back(row / nb - k, row % nb);
        }
        piv.push_back(std::move(col));
    }
    return piv;
}

inline std::vector<int64_t> ipiv_from_pivots(const Pivots& piv, int64_t nb, int64_t n)
{
    std::vector<int64_t> ipiv;
    ipiv.reserve(size_t(n));
    int64_t k = 0;
    for (const auto& col : piv) {
        for (const auto& p : col) ipiv.push_back((k + p.tileIndex()) * nb + p.elementOffset() + 1);
        ++k;
    }
    if (int64_t(ipiv.size()) != n) throw Exception("pivots do not match the matrix size", -2);
    return ipiv;
}

inline int64_t opt_i(const Options& o, Option k, int64_t def)
{
    auto it = o.find(k);
    return it == o.end() ? def : it->second.i_;
}
inline double opt_d(const Options& o, Option k, double def)
{
    auto it = o.find(k);
    return it == o.end() ? def : it->second.d_;
}

// Device-resident storage of a matrix: the caller's device array, or a
// staging copy of a host (fromLAPACK) array.
template <class T>
struct Storage {
    int64_t m = 0, n = 0, lda = 0, nb = 0;
    T* data = nullptr;
    bool host = false;
};

} // namespace detail

//------------------------------------------------------------------------------
// Matrix<T> (include/slate/Matrix.hh:41-54, :293-300, :397): a shallow handle
// on an m x n column-major matrix in nb x nb tiles, one device, 1 x 1 grid.
template <typename scalar_t>
class Matrix {
public:
    Matrix() : st_(std::make_shared<detail::Storage<scalar_t>>()) {}

    static Matrix fromDevices(int64_t m, int64_t n, scalar_t** Aarray, int num_devices,
                              int64_t lda, int64_t nb, int p, int q)
    {
        if (num_devices != 1 || p != 1 || q != 1)
            throw NotImplemented("fromDevices: one device on a 1 x 1 grid only");
        return wrap(m, n, Aarray[0], lda, nb, false);
    }
    static Matrix fromLAPACK(int64_t m, int64_t n, scalar_t* A, int64_t lda, int64_t nb,
                             int p, int q)
    {
        if (p != 1 || q != 1) throw NotImplemented("fromLAPACK: 1 x 1 grid only");
        return wrap(m, n, A, lda, nb, true);
    }

    int64_t m() const { return op_ == Op::NoTrans ? st_->m : st_->n; }
    int64_t n() const { return op_ == Op::NoTrans ? st_->n : st_->m; }
    int64_t mt() const { return (m() + st_->nb - 1) / st_->nb; }
    int64_t nt() const { return (n() + st_->nb - 1) / st_->nb; }
    int64_t tileMb(int64_t i) const { return std::min(st_->nb, m() - i * st_->nb); }
    int64_t tileNb(int64_t j) const { return std::min(st_->nb, n() - j * st_->nb); }
    Op op() const { return op_; }
    // Column-major element access in the storage's own orientation.
    scalar_t* data() const { return st_->data; }
    int64_t stride() const { return st_->lda; }
    bool hostOrigin() const { return st_->host; }
    int64_t blockSize() const { return st_->nb; }

    template <typename T> friend Matrix<T> transpose(const Matrix<T>& A);

protected:
    static Matrix wrap(int64_t m, int64_t n, scalar_t* A, int64_t lda, int64_t nb, bool host)
    {
        if (m < 0 || n < 0 || nb <= 0 || lda < std::max<int64_t>(1, m))
            throw Exception("Matrix: invalid dimensions", -2);
        Matrix M;
        M.st_->m = m; M.st_->n = n; M.st_->lda = lda; M.st_->nb = nb;
        M.st_->data = A; M.st_->host = host;
        return M;
    }
    std::shared_ptr<detail::Storage<scalar_t>> st_;
    Op op_ = Op::NoTrans;
};

template <typename T>
Matrix<T> transpose(const Matrix<T>& A)
{
    Matrix<T> AT = A;
    AT.op_ = (A.op_ == Op::NoTrans) ? Op::Trans : Op::NoTrans;
    return AT;
}

// HermitianMatrix<T> (include/slate/HermitianMatrix.hh): the uplo triangle of
// an n x n symmetric matrix is referenced.
template <typename scalar_t>
class HermitianMatrix : public Matrix<scalar_t> {
public:
    HermitianMatrix() = default;
    static HermitianMatrix fromDevices(Uplo uplo, int64_t n, scalar_t** Aarray, int num_devices,
                                       int64_t lda, int64_t nb, int p, int q)
    {
        HermitianMatrix H;
        static_cast<Matrix<scalar_t>&>(H) = Matrix<scalar_t>::fromDevices(n, n, Aarray, num_devices,
                                                                           lda, nb, p, q);
        H.uplo_ = uplo;
        return H;
    }
    static HermitianMatrix fromLAPACK(Uplo uplo, int64_t n, scalar_t* A, int64_t lda, int64_t nb,
                                      int p, int q)
    {
        HermitianMatrix H;
        static_cast<Matrix<scalar_t>&>(H) = Matrix<scalar_t>::fromLAPACK(n, n, A, lda, nb, p, q);
        H.uplo_ = uplo;
        return H;
    }
    Uplo uplo() const { return uplo_; }
private:
    Uplo uplo_ = Uplo::Lower;
};

namespace detail {

// Device view of a matrix for one driver call: host-origin matrices are
// copied in on entry and (when written) back on exit.
template <class T>
class DeviceView {
public:
    DeviceView(const Matrix<T>& A, bool copy_in, bool copy_out)
        : A_(A), out_(copy_out)
    {
        if (!A.hostOrigin()) { ptr_ = A.data(); ld_ = A.stride(); return; }
        const int64_t m = A.op() == Op::NoTrans ? A.m() : A.n();
        const int64_t n = A.op() == Op::NoTrans ? A.n() : A.m();
        ld_ = std::max<int64_t>(1, m);
        if (hipMalloc(reinterpret_cast<void**>(&ptr_), sizeof(T) * size_t(ld_) * size_t(std::max<int64_t>(n, 1)))
            != hipSuccess)
            throw Exception("hipMalloc failed for a host-origin matrix", -100);
        owned_ = true;
        if (copy_in && m > 0 && n > 0 &&
            hipMemcpy2D(ptr_, sizeof(T) * ld_, A.data(), sizeof(T) * A.stride(), sizeof(T) * m, n,
                        hipMemcpyHostToDevice) != hipSuccess)
            throw Exception("hipMemcpy2D to device failed", -100);
        rows_ = m; cols_ = n;
    }
    void finish()
    {
        if (owned_ && out_ && rows_ > 0 && cols_ > 0 &&
            hipMemcpy2D(A_.data(), sizeof(T) * A_.stride(), ptr_, sizeof(T) * ld_, sizeof(T) * rows_, cols_,
                        hipMemcpyDeviceToHost) != hipSuccess)
            throw Exception("hipMemcpy2D to host failed", -100);
    }
    ~DeviceView() { if (owned_) (void)hipFree(ptr_); }
    T* ptr() const { return ptr_; }
    int64_t ld() const { return ld_; }
private:
    Matrix<T> A_;
    bool out_;
    bool owned_ = false;
    T* ptr_ = nullptr;
    int64_t ld_ = 0, rows_ = 0, cols_ = 0;
};

inline slate_amd_options to_c(const Options& o, int64_t nb)
{
    slate_amd_options c;
    slate_amd_options_init(&c);
    auto t = Target(opt_i(o, Option::Target, int64_t(Target::Devices)));
    if (t != Target::Devices) throw NotImplemented("only Target::Devices runs on this library");
    c.block_size = nb;
    c.inner_blocking = opt_i(o, Option::InnerBlocking, 0);
    c.lookahead = opt_i(o, Option::Lookahead, 1);
    c.pivot_threshold = opt_d(o, Option::PivotThreshold, 1.0);
    c.max_iterations = opt_i(o, Option::MaxIterations, 30);
    c.tolerance = opt_d(o, Option::Tolerance, 0.0);
    c.use_fallback = int32_t(opt_i(o, Option::UseFallbackSolver, 1));
    c.factor_precision = int32_t(opt_i(o, Option::FactorPrecision, int64_t(FactorPrecision::FP32)));
    auto m = MethodLU(opt_i(o, Option::MethodLU, int64_t(MethodLU::PartialPiv)));
    if (m != MethodLU::PartialPiv && m != MethodLU::NoPiv && m != MethodLU::CALU)
        throw NotImplemented("MethodLU");
    c.method_lu = m == MethodLU::NoPiv ? SLATE_AMD_LU_NOPIV
                : m == MethodLU::CALU ? SLATE_AMD_LU_CALU : SLATE_AMD_LU_PARTIALPIV;
    c.depth = int32_t(opt_i(o, Option::Depth, 2));
    c.use_graph = int32_t(opt_i(o, Option::UseGraph, 0));
    return c;
}

template <class T>
inline void require_square(const Matrix<T>& A, const char* who)
{
    if (A.m() != A.n()) throw Exception(std::string(who) + ": A must be square", -2);
}

template <class T>
inline void require_rhs(const Matrix<T>& A, const Matrix<T>& B, const char* who)
{
    if (B.m() != A.n()) throw Exception(std::string(who) + ": B rows must equal A columns", -2);
    if (B.op() != Op::NoTrans) throw NotImplemented(std::string(who) + ": transposed B");
}

inline int c_uplo(Uplo u)
{
    if (u == Uplo::General) throw Exception("HermitianMatrix with Uplo::General", -2);
    return u == Uplo::Lower ? SLATE_AMD_LOWER : SLATE_AMD_UPPER;
}

} // namespace detail

//------------------------------------------------------------------------------
// LU: getrf / getrs / gesv (include/slate/slate.hh:504-508, :574-577, :603-607).
inline int64_t getrf(Matrix<double>& A, Pivots& pivots, const Options& opts = Options())
{
    detail::require_square(A, "getrf");
    if (A.op() != Op::NoTrans) throw NotImplemented("getrf of a transposed matrix");
    const int64_t n = A.n();
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<double> dA(A, true, true);
    std::vector<int64_t> ipiv{};
    ipiv.resize(size_t(n));
    int64_t info = slate_amd_dgetrf(detail::context(detail::current_device()), n, dA.ptr(), dA.ld(),
                                    ipiv.data(), &c);
    detail::check(info);
    dA.finish();
    pivots = detail::pivots_from_ipiv(ipiv, A.blockSize());
    return info;
}

inline void getrs(Matrix<double>& A, Pivots& pivots, Matrix<double>& B, const Options& opts = Options())
{
    detail::require_square(A, "getrs");
    detail::require_rhs(A, B, "getrs");
    const int64_t n = A.n();
    auto c = detail::to_c(opts, A.blockSize());
    std::vector<int64_t> ipiv = detail::ipiv_from_pivots(pivots, A.blockSize(), n);
    // A holds the factors in its storage orientation; op(A) selects the solve.
    detail::DeviceView<double> dA(A, true, false);
    detail::DeviceView<double> dB(B, true, true);
    int trans = A.op() == Op::NoTrans ? SLATE_AMD_NOTRANS : SLATE_AMD_TRANS;
    detail::check(slate_amd_dgetrs(detail::context(detail::current_device()), trans, n, B.n(), dA.ptr(),
                                   dA.ld(), ipiv.data(), dB.ptr(), dB.ld(), &c));
    dB.finish();
}

inline int64_t gesv(Matrix<double>& A, Pivots& pivots, Matrix<double>& B, const Options& opts = Options())
{
    detail::require_square(A, "gesv");
    detail::require_rhs(A, B, "gesv");
    if (A.op() != Op::NoTrans) throw NotImplemented("gesv of a transposed matrix");
    const int64_t n = A.n();
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<double> dA(A, true, true);
    detail::DeviceView<double> dB(B, true, true);
    std::vector<int64_t> ipiv{};
    ipiv.resize(size_t(n));
    int64_t info = slate_amd_dgesv(detail::context(detail::current_device()), n, B.n(), dA.ptr(), dA.ld(),
                                   ipiv.data(), dB.ptr(), dB.ld(), &c);
    detail::check(info);
    dA.finish();
    dB.finish();
    pivots = detail::pivots_from_ipiv(ipiv, A.blockSize());
    return info;
}

// Single-precision LU (slate::getrf / getrs instantiated for float).
inline int64_t getrf(Matrix<float>& A, Pivots& pivots, const Options& opts = Options())
{
    detail::require_square(A, "getrf");
    if (A.op() != Op::NoTrans) throw NotImplemented("getrf of a transposed matrix");
    const int64_t n = A.n();
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<float> dA(A, true, true);
    std::vector<int64_t> ipiv{};
    ipiv.resize(size_t(n));
    int64_t info = slate_amd_sgetrf(detail::context(detail::current_device()), n, dA.ptr(), dA.ld(),
                                    ipiv.data(), &c);
    detail::check(info);
    dA.finish();
    pivots = detail::pivots_from_ipiv(ipiv, A.blockSize());
    return info;
}

inline void getrs(Matrix<float>& A, Pivots& pivots, Matrix<float>& B, const Options& opts = Options())
{
    detail::require_square(A, "getrs");
    detail::require_rhs(A, B, "getrs");
    const int64_t n = A.n();
    auto c = detail::to_c(opts, A.blockSize());
    std::vector<int64_t> ipiv = detail::ipiv_from_pivots(pivots, A.blockSize(), n);
    detail::DeviceView<float> dA(A, true, false);
    detail::DeviceView<float> dB(B, true, true);
    int trans = A.op() == Op::NoTrans ? SLATE_AMD_NOTRANS : SLATE_AMD_TRANS;
    detail::check(slate_amd_sgetrs(detail::context(detail::current_device()), trans, n, B.n(), dA.ptr(),
                                   dA.ld(), ipiv.data(), dB.ptr(), dB.ld(), &c));
    dB.finish();
}

//------------------------------------------------------------------------------
// Cholesky: potrf / potrs / posv (slate.hh:646-650, :714-717, :741-745).
inline int64_t potrf(HermitianMatrix<double>& A, const Options& opts = Options())
{
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<double> dA(A, true, true);
    int64_t info = slate_amd_dpotrf(detail::context(detail::current_device()), detail::c_uplo(A.uplo()),
                                    A.n(), dA.ptr(), dA.ld(), &c);
    detail::check(info);
    dA.finish();
    return info;
}

inline void potrs(HermitianMatrix<double>& A, Matrix<double>& B, const Options& opts = Options())
{
    detail::require_rhs(A, B, "potrs");
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<double> dA(A, true, false);
    detail::DeviceView<double> dB(B, true, true);
    detail::check(slate_amd_dpotrs(detail::context(detail::current_device()), detail::c_uplo(A.uplo()),
                                   A.n(), B.n(), dA.ptr(), dA.ld(), dB.ptr(), dB.ld(), &c));
    dB.finish();
}

inline int64_t posv(HermitianMatrix<double>& A, Matrix<double>& B, const Options& opts = Options())
{
    detail::require_rhs(A, B, "posv");
    auto c = detail::to_c(opts, A.blockSize());
    detail::DeviceView<double> dA(A, true, true);
    detail::DeviceView<double> dB(B, true, true);
    int64_t info = slate_amd_dposv(detail::context(detail::current_device()), detail::c_uplo(A.uplo()),
                                   A.n(), B.n(), dA.ptr(), dA.ld(), dB.ptr(), dB.ld(), &c);
    detail::check(info);
    dA.finish();
    dB.finish();
    return info;
}

//------------------------------------------------------------------------------
// Mixed precision + refinement (slate.hh:521-553, :667-701).  A is written
// only when the FP64 fallback runs (then it holds the FP64 factors).
namespace detail {

template <class Fn>
inline int64_t mixed_call(Matrix<double>& A, Matrix<double>& B, Matrix<double>& X, const Options& opts,
                          Fn fn)
{
    require_square(A, "mixed solver");
    require_rhs(A, B, "mixed solver");
    if (X.m() != B.m() || X.n() != B.n()) throw Exception("X must have the shape of B", -2);
    auto c = to_c(opts, A.blockSize());
    DeviceView<double> dA(A, true, true);
    DeviceView<double> dB(B, true, false);
    DeviceView<double> dX(X, false, true);
    int64_t info = fn(c, dA, dB, dX);
    check(info);
    dA.finish();
    dX.finish();
    return info;
}

} // namespace detail

template <typename scalar_hi = double, typename scalar_lo = float>
int64_t gesv_mixed(Matrix<double>& A, Pivots& pivots, Matrix<double>& B, Matrix<double>& X, int& iter,
                   const Options& opts = Options())
{
    static_assert(std::is_same<scalar_hi, double>::value && std::is_same<scalar_lo, float>::value,
                  "gesv_mixed: double with float factors (BF16 via Option::FactorPrecision)");
    std::vector<int64_t> ipiv{};
    ipiv.resize(size_t(A.n()));
    int64_t info = detail::mixed_call(A, B, X, opts, [&](slate_amd_options& c, auto& dA, auto& dB, auto& dX) {
        return slate_amd_dsgesv(detail::context(detail::current_device()), A.n(), B.n(), dA.ptr(), dA.ld(),
                                ipiv.data(), dB.ptr(), dB.ld(), dX.ptr(), dX.ld(), &iter, &c);
    });
    pivots = detail::pivots_from_ipiv(ipiv, A.blockSize());
    return info;
}

template <typename scalar_hi = double, typename scalar_lo = float>
int64_t gesv_mixed_gmres(Matrix<double>& A, Pivots& pivots, Matrix<double>& B, Matrix<double>& X, int& iter,
                         const Options& opts = Options())
{
    static_assert(std::is_same<scalar_hi, double>::value && std::is_same<scalar_lo, float>::value,
                  "gesv_mixed_gmres: double with float factors (BF16 via Option::FactorPrecision)");
    std::vector<int64_t> ipiv{};
    ipiv.resize(size_t(A.n()));
    int64_t info = detail::mixed_call(A, B, X, opts, [&](slate_amd_options& c, auto& dA, auto& dB, auto& dX) {
        return slate_amd_dsgesv_gmres(detail::context(detail::current_device()), A.n(), B.n(), dA.ptr(),
                                      dA.ld(), ipiv.data(), dB.ptr(), dB.ld(), dX.ptr(), dX.ld(), &iter, &c);
    });
    pivots = detail::pivots_from_ipiv(ipiv, A.blockSize());
    return info;
}

template <typename scalar_hi = double, typename scalar_lo = float>
int64_t posv_mixed(HermitianMatrix<double>& A, Matrix<double>& B, Matrix<double>& X, int& iter,
                   const Options& opts = Options())
{
    static_assert(std::is_same<scalar_hi, double>::value && std::is_same<scalar_lo, float>::value,
                  "posv_mixed: double with float factors");
    const int uplo = detail::c_uplo(A.uplo());
    return detail::mixed_call(A, B, X, opts, [&](slate_amd_options& c, auto& dA, auto& dB, auto& dX) {
        return slate_amd_dsposv(detail::context(detail::current_device()), uplo, A.n(), B.n(), dA.ptr(),
                                dA.ld(), dB.ptr(), dB.ld(), dX.ptr(), dX.ld(), &iter, &c);
    });
}

template <typename scalar_hi = double, typename scalar_lo = float>
int64_t posv_mixed_gmres(HermitianMatrix<double>& A, Matrix<double>& B, Matrix<double>& X, int& iter,
                         const Options& opts = Options())
{
    static_assert(std::is_same<scalar_hi, double>::value && std::is_same<scalar_lo, float>::value,
                  "posv_mixed_gmres: double with float factors");
    const int uplo = detail::c_uplo(A.uplo());
    return detail::mixed_call(A, B, X, opts, [&](slate_amd_options& c, auto& dA, auto& dB, auto& dX) {
        return slate_amd_dsposv_gmres(detail::context(detail::current_device()), uplo, A.n(), B.n(),
                                      dA.ptr(), dA.ld(), dB.ptr(), dB.ld(), dX.ptr(), dX.ld(), &iter, &c);
    });
}

// slate::gesv_rbt (include/slate/slate.hh:558-563, src/gesv_rbt.cc:94-236):
// random butterfly transform, LU without pivoting, FP64 refinement.
template <typename scalar_t>
void gesv_rbt(Matrix<scalar_t>& A, Matrix<scalar_t>& B, Matrix<scalar_t>& X, int& iter,
              const Options& opts = Options())
{
    static_assert(std::is_same<scalar_t, double>::value, "gesv_rbt: double only");
    detail::mixed_call(A, B, X, opts, [&](slate_amd_options& c, auto& dA, auto& dB, auto& dX) {
        return slate_amd_dgesv_rbt(detail::context(detail::current_device()), A.n(), B.n(), dA.ptr(),
                                   dA.ld(), dB.ptr(), dB.ld(), dX.ptr(), dX.ld(), &iter, &c);
    });
}

} // namespace slate_amd

#endif // SLATE_AMD_HH
