// Drives the slate_amd:: C++ interface the way SLATE's tester drives slate::
// (test/test_gesv.cc:232-380, test/test_posv.cc:209-345): seeded matgen
// inputs, the driver call, and the scaled residual
//   ||B - A X||_1 / (n ||A||_1 ||X||_1 eps) <= 3.
// Writes the LAPACK-style pivots of the n = 2000 gesv to <outdir>/ipiv_gesv.txt
// for comparison with LAPACK.  Exit status 0 when every check passes.
#include "slate_amd.hh"

#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

using slate_amd::Matrix;
using slate_amd::HermitianMatrix;

static int failures = 0;
#define CHECK(cond, ...)                                        \
    do {                                                        \
        if (!(cond)) { std::printf("FAIL: " __VA_ARGS__); std::printf("\n"); ++failures; } \
    } while (0)

static const double eps = std::numeric_limits<double>::epsilon();

// Column-major host helpers.
static double norm1(const std::vector<double>& A, int64_t m, int64_t n, int64_t ld)
{
    double r = 0;
    for (int64_t j = 0; j < n; ++j) {
        double s = 0;
        for (int64_t i = 0; i < m; ++i) s += std::fabs(A[size_t(i + j * ld)]);
        r = std::max(r, s);
    }
    return r;
}

// ||B - op(A) X||_1 / (n ||A||_1 ||X||_1 eps), A n x n, X and B n x k.
static double scaled_residual(const std::vector<double>& A, const std::vector<double>& X,
                              const std::vector<double>& B, int64_t n, int64_t k, bool trans)
{
    std::vector<double> R(B);
    for (int64_t c = 0; c < k; ++c)
        for (int64_t j = 0; j < n; ++j) {
            double x = X[size_t(j + c * n)];
            for (int64_t i = 0; i < n; ++i)
                R[size_t(i + c * n)] -= (trans ? A[size_t(j + i * n)] : A[size_t(i + j * n)]) * x;
        }
    return norm1(R, n, k, n) / (double(n) * norm1(A, n, n, n) * norm1(X, n, k, n) * eps);
}

static std::vector<double> generate(int dist, int64_t seed, int64_t m, int64_t n)
{
    double* d = nullptr;
    (void)hipMalloc(reinterpret_cast<void**>(&d), sizeof(double) * size_t(m * n));
    auto ctx = slate_amd::detail::context(0);
    slate_amd::detail::check(slate_amd_generate(ctx, dist, seed, m, n, 0, 0, d, m));
    std::vector<double> h(size_t(m * n));
    (void)hipMemcpy(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return h;
}

static double* to_device(const std::vector<double>& h)
{
    double* d = nullptr;
    (void)hipMalloc(reinterpret_cast<void**>(&d), sizeof(double) * h.size());
    (void)hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice);
    return d;
}

static std::vector<double> to_host(const double* d, size_t count)
{
    std::vector<double> h(count);
    (void)hipMemcpy(h.data(), d, sizeof(double) * count, hipMemcpyDeviceToHost);
    return h;
}

int main(int argc, char** argv)
{
    const char* outdir = argc > 1 ? argv[1] : ".";
    const int64_t n = 2000, nb = 256, nrhs = 1;
    const std::vector<double> A0 = generate(2, 0, n, n);     // rands, seed 0
    const std::vector<double> B0 = generate(2, 1, n, nrhs);  // rands, seed 1

    // 1. gesv on device matrices (Matrix::fromDevices, Target::Devices).
    {
        double* dA = to_device(A0);
        double* dB = to_device(B0);
        auto A = Matrix<double>::fromDevices(n, n, &dA, 1, n, nb, 1, 1);
        auto B = Matrix<double>::fromDevices(n, nrhs, &dB, 1, n, nb, 1, 1);
        slate_amd::Pivots pivots;
        int64_t info = slate_amd::gesv(A, pivots, B, {{slate_amd::Option::Lookahead, 1},
                                                        {slate_amd::Option::Target, slate_amd::Target::Devices}});
        auto X = to_host(dB, size_t(n * nrhs));
        double r = scaled_residual(A0, X, B0, n, nrhs, false);
        std::printf("gesv n=%ld nb=%ld info=%ld scaled_residual=%.3f\n", long(n), long(nb), long(info), r);
        CHECK(info == 0, "gesv info %ld", long(info));
        CHECK(r <= 3.0, "gesv residual %g", r);
        CHECK(int64_t(pivots.size()) == A.nt(), "pivots has %zu block columns", pivots.size());
        auto ipiv = slate_amd::detail::ipiv_from_pivots(pivots, nb, n);
        std::string path = std::string(outdir) + "/ipiv_gesv.txt";
        FILE* f = std::fopen(path.c_str(), "w");
        if (f) {
            for (int64_t v : ipiv) std::fprintf(f, "%ld\n", long(v));
            std::fclose(f);
        }
        // getrs with the returned Pivots reuses the factors: solve again.
        std::vector<double> B1 = generate(2, 5, n, 2);
        double* dB1 = to_device(B1);
        auto Bm = Matrix<double>::fromDevices(n, 2, &dB1, 1, n, nb, 1, 1);
        slate_amd::getrs(A, pivots, Bm);
        auto X1 = to_host(dB1, size_t(n * 2));
        double r1 = scaled_residual(A0, X1, B1, n, 2, false);
        std::printf("getrs (factors from gesv) nrhs=2 scaled_residual=%.3f\n", r1);
        CHECK(r1 <= 3.0, "getrs residual %g", r1);
        (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dB1);
    }

    // 2. getrf + transposed getrs on host matrices (Matrix::fromLAPACK).
    {
        std::vector<double> Ah(A0), Bh(B0);
        auto A = Matrix<double>::fromLAPACK(n, n, Ah.data(), n, nb, 1, 1);
        auto B = Matrix<double>::fromLAPACK(n, nrhs, Bh.data(), n, nb, 1, 1);
        slate_amd::Pivots pivots;
        int64_t info = slate_amd::getrf(A, pivots);
        auto AT = slate_amd::transpose(A);
        slate_amd::getrs(AT, pivots, B);
        double r = scaled_residual(A0, Bh, B0, n, nrhs, true);
        std::printf("getrf + getrs(transpose(A)) host-origin info=%ld scaled_residual=%.3f\n", long(info), r);
        CHECK(info == 0, "getrf info %ld", long(info));
        CHECK(r <= 3.0, "trans getrs residual %g", r);
    }

    // 3. posv on an SPD matrix G G^T / n + I (host origin, lower).
    const int64_t ns = 1000;
    std::vector<double> S(size_t(ns * ns));
    {
        std::vector<double> G = generate(3, 2, ns, ns);   // randn, seed 2
        for (int64_t j = 0; j < ns; ++j)
            for (int64_t i = 0; i < ns; ++i) {
                double s = 0;
                for (int64_t k = 0; k < ns; ++k) s += G[size_t(i + k * ns)] * G[size_t(j + k * ns)];
                S[size_t(i + j * ns)] = s / double(ns) + (i == j ? 1.0 : 0.0);
            }
        std::vector<double> Sh(S), Bh = generate(2, 1, ns, 1), B0s(Bh);
        auto A = HermitianMatrix<double>::fromLAPACK(slate_amd::Uplo::Lower, ns, Sh.data(), ns, 128, 1, 1);
        auto B = Matrix<double>::fromLAPACK(ns, 1, Bh.data(), ns, 128, 1, 1);
        int64_t info = slate_amd::posv(A, B);
        double r = scaled_residual(S, Bh, B0s, ns, 1, false);
        std::printf("posv n=%ld info=%ld scaled_residual=%.3f\n", long(ns), long(info), r);
        CHECK(info == 0, "posv info %ld", long(info));
        CHECK(r <= 3.0, "posv residual %g", r);
    }

    // 4. mixed-precision drivers reach the FP64 bound; A stays unchanged.
    {
        double* dA = to_device(A0);
        double* dB = to_device(B0);
        double* dX = nullptr;
        (void)hipMalloc(reinterpret_cast<void**>(&dX), sizeof(double) * size_t(n));
        auto A = Matrix<double>::fromDevices(n, n, &dA, 1, n, nb, 1, 1);
        auto B = Matrix<double>::fromDevices(n, 1, &dB, 1, n, nb, 1, 1);
        auto X = Matrix<double>::fromDevices(n, 1, &dX, 1, n, nb, 1, 1);
        slate_amd::Pivots pivots;
        int iter = -99;
        int64_t info = slate_amd::gesv_mixed(A, pivots, B, X, iter);
        double r = scaled_residual(A0, to_host(dX, size_t(n)), B0, n, 1, false);
        std::printf("gesv_mixed info=%ld iter=%d scaled_residual=%.3f\n", long(info), iter, r);
        CHECK(info == 0 && iter >= 0, "gesv_mixed info %ld iter %d", long(info), iter);
        CHECK(r <= 3.0, "gesv_mixed residual %g", r);
        CHECK(to_host(dA, size_t(n * n)) == A0, "gesv_mixed modified A after converging");
        iter = -99;
        info = slate_amd::gesv_mixed_gmres(A, pivots, B, X, iter);
        r = scaled_residual(A0, to_host(dX, size_t(n)), B0, n, 1, false);
        std::printf("gesv_mixed_gmres info=%ld iter=%d scaled_residual=%.3f\n", long(info), iter, r);
        CHECK(info == 0 && iter >= 0, "gesv_mixed_gmres info %ld iter %d", long(info), iter);
        CHECK(r <= 3.0, "gesv_mixed_gmres residual %g", r);
        iter = -99;
        slate_amd::gesv_rbt(A, B, X, iter);
        r = scaled_residual(A0, to_host(dX, size_t(n)), B0, n, 1, false);
        std::printf("gesv_rbt iter=%d scaled_residual=%.3f\n", iter, r);
        CHECK(iter >= 0, "gesv_rbt iter %d", iter);
        CHECK(r <= 3.0, "gesv_rbt residual %g", r);
        (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dX);

        std::vector<double> Sh(S), Bh = generate(2, 1, ns, 1), Xh(static_cast<size_t>(ns), 0.0);
        auto H = HermitianMatrix<double>::fromLAPACK(slate_amd::Uplo::Lower, ns, Sh.data(), ns, 128, 1, 1);
        auto Bs = Matrix<double>::fromLAPACK(ns, 1, Bh.data(), ns, 128, 1, 1);
        auto Xs = Matrix<double>::fromLAPACK(ns, 1, Xh.data(), ns, 128, 1, 1);
        iter = -99;
        info = slate_amd::posv_mixed(H, Bs, Xs, iter);
        r = scaled_residual(S, Xh, Bh, ns, 1, false);
        std::printf("posv_mixed info=%ld iter=%d scaled_residual=%.3f\n", long(info), iter, r);
        CHECK(info == 0 && iter >= 0, "posv_mixed info %ld iter %d", long(info), iter);
        CHECK(r <= 3.0, "posv_mixed residual %g", r);
        iter = -99;
        info = slate_amd::posv_mixed_gmres(H, Bs, Xs, iter);
        r = scaled_residual(S, Xh, Bh, ns, 1, false);
        std::printf("posv_mixed_gmres info=%ld iter=%d scaled_residual=%.3f\n", long(info), iter, r);
        CHECK(info == 0 && iter >= 0, "posv_mixed_gmres info %ld iter %d", long(info), iter);
        CHECK(r <= 3.0, "posv_mixed_gmres residual %g", r);
    }

    // 5. misuse raises exceptions instead of returning garbage.
    {
        std::vector<double> Ah(16, 1.0), Bh(3, 1.0);
        auto A = Matrix<double>::fromLAPACK(4, 4, Ah.data(), 4, 2, 1, 1);
        auto B = Matrix<double>::fromLAPACK(3, 1, Bh.data(), 3, 2, 1, 1);
        slate_amd::Pivots pivots;
        bool threw = false;
        try { slate_amd::gesv(A, pivots, B); } catch (const slate_amd::Exception&) { threw = true; }
        CHECK(threw, "gesv with mismatched B did not throw");
        threw = false;
        try {
            slate_amd::getrf(A, pivots, {{slate_amd::Option::Target, slate_amd::Target::HostTask}});
        } catch (const slate_amd::NotImplemented&) { threw = true; }
        CHECK(threw, "Target::HostTask did not throw NotImplemented");
        std::vector<double> Z(16, 0.0);
        auto Zm = Matrix<double>::fromLAPACK(4, 4, Z.data(), 4, 2, 1, 1);
        int64_t info = slate_amd::getrf(Zm, pivots);
        CHECK(info == 1, "getrf of the zero matrix: info %ld, expected 1", long(info));
        std::printf("exceptions and info checked\n");
    }

    std::printf("%s (%d failures)\n", failures == 0 ? "ALL PASSED" : "FAILED", failures);
    return failures == 0 ? 0 : 1;
}
