// Host-only sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer) over
// the CPU oracle: it is compiled into this executable with the sanitizers on
// (tests/cpp/Makefile target oracle_sanitize, g++ only, no GPU code) and every
// solver family runs once on small ragged inputs: tile LU / Cholesky and their
// solves, the refinement drivers (classic and GMRES, general and SPD, with
// the split-BF16 and e4m3 update emulations), the random butterfly solver,
// matgen and the GMRES first-cycle inspection.  A sanitizer report makes the
// process fail; the checks below only make sure each call did its work.
#include "../../oracle/slate_oracle.cc"

#include <cstdio>
#include <vector>

static int failures = 0;
#define EXPECT(cond, msg) do { if (!(cond)) { std::printf("FAIL: %s\n", msg); ++failures; } } while (0)

static std::vector<double> rands(int64_t seed, int64_t m, int64_t n)
{
    std::vector<double> A(static_cast<size_t>(m * n));
    oracle_generate(2, seed, m, n, 0, 0, A.data(), m);
    return A;
}

static double resid(int64_t n, const std::vector<double>& A, const std::vector<double>& x,
                    const std::vector<double>& b)
{
    double r = 0, a = 0, xn = 0;
    for (int64_t i = 0; i < n; ++i) {
        double s = b[size_t(i)];
        for (int64_t j = 0; j < n; ++j) s -= A[size_t(i + j * n)] * x[size_t(j)];
        r = std::max(r, std::fabs(s));
        xn = std::max(xn, std::fabs(x[size_t(i)]));
    }
    for (int64_t i = 0; i < n; ++i) {
        double row = 0;
        for (int64_t j = 0; j < n; ++j) row += std::fabs(A[size_t(i + j * n)]);
        a = std::max(a, row);
    }
    return r / (a * xn * double(n) * 2.220446049250313e-16);
}

int main()
{
    oracle_set_num_threads(2);
    const int64_t n = 150, nb = 32;
    const auto A0 = rands(0, n, n);
    const auto b0 = rands(1, n, 1);
    // LU, solve
    {
        auto A = A0;
        auto b = b0;
        std::vector<int64_t> ipiv(static_cast<size_t>(n));
        EXPECT(oracle_dgesv(n, 1, A.data(), n, ipiv.data(), b.data(), n, nb, 8, 2) == 0, "dgesv info");
        EXPECT(resid(n, A0, b, b0) < 10, "dgesv residual");
        std::vector<float> S(static_cast<size_t>(n * n));
        for (size_t i = 0; i < S.size(); ++i) S[i] = float(A0[i]);
        EXPECT(oracle_sgetrf(n, n, S.data(), n, nb, 8, 2, 1.0, ipiv.data()) == 0, "sgetrf info");
    }
    // SPD: G G^T / n + I
    std::vector<double> P(static_cast<size_t>(n * n));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int64_t k = 0; k < n; ++k) s += A0[size_t(i + k * n)] * A0[size_t(j + k * n)] / double(n);
            P[size_t(i + j * n)] = s;
        }
    {
        auto A = P;
        auto b = b0;
        EXPECT(oracle_dposv(n, 1, A.data(), n, b.data(), n, nb) == 0, "dposv info");
        EXPECT(resid(n, P, b, b0) < 10, "dposv residual");
    }
    // refinement drivers, float factor and the emulated update precisions
    const int kinds[][2] = {{0, 0}, {1, 1}, {1, 2}, {1, 3}, {2, 1}, {2, 2}};
    for (const auto& kt : kinds) {
        EXPECT(oracle_set_update_precision(kt[0], kt[1], 128) == 0, "update precision");
        for (int gm = 0; gm < 2; ++gm) {
            auto A = A0;
            std::vector<double> x(static_cast<size_t>(n));
            std::vector<int64_t> ipiv(static_cast<size_t>(n));
            int it = 0;
            int64_t info = gm ? oracle_dsgesv_gmres(n, A.data(), n, ipiv.data(), b0.data(), x.data(), nb, 8, 2,
                                                    30, 0.0, 1, &it)
                              : oracle_dsgesv(n, 1, A.data(), n, ipiv.data(), b0.data(), n, x.data(), n, nb, 8,
                                              2, 30, 0.0, 1, &it);
            EXPECT(info == 0 && resid(n, A0, x, b0) < 10, "dsgesv");
            auto S = P;
            info = gm ? oracle_dsposv_gmres(n, S.data(), n, b0.data(), x.data(), nb, 30, 0.0, 1, &it)
                      : oracle_dsposv(n, 1, S.data(), n, b0.data(), n, x.data(), n, nb, 30, 0.0, 1, &it);
            EXPECT(info == 0 && resid(n, P, x, b0) < 10, "dsposv");
        }
    }
    oracle_set_update_precision(0, 0, 128);
    // butterfly solver
    {
        auto A = A0;
        std::vector<double> x(static_cast<size_t>(n));
        int it = 0;
        const int64_t info = oracle_dgesv_rbt(n, 1, A.data(), n, b0.data(), n, x.data(), n, nb, 8, 2, 30, 0.0, 1,
                                              &it);
        EXPECT(info == 0 && resid(n, A0, x, b0) < 10, "dgesv_rbt");
    }
    // GMRES first cycle
    {
        auto A = A0;
        std::vector<double> x(static_cast<size_t>(n)), V(static_cast<size_t>(n * 31)), W(static_cast<size_t>(n * 31)),
            H(31 * 31), Sa(31);
        const int64_t j = oracle_gmres_first_cycle(n, A.data(), n, 0, b0.data(), x.data(), nb, 30, 2, V.data(),
                                                   W.data(), H.data(), Sa.data());
        EXPECT(j >= 0 && j <= 30, "gmres first cycle");
    }
    std::printf("oracle sanitizer run: %d failures\n", failures);
    return failures ? 1 : 0;
}
