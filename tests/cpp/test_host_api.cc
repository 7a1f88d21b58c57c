// Host-only checks of the slate_amd:: C++ interface: the Pivots <-> LAPACK
// ipiv conversion of lapack_api/lapack_gesv.cc:90-100, Options translation,
// and argument errors raised before any device work.  Needs no GPU.
#include "slate_amd.hh"

#include <cstdio>

static int failures = 0;
#define CHECK(cond, msg) do { if (!(cond)) { std::printf("FAIL: %s\n", msg); ++failures; } } while (0)

int main()
{
    using namespace slate_amd;
    // ipiv (1-based) of a 7 x 7 matrix with nb = 3: panels {0,1,2}, {3,4,5}, {6}.
    std::vector<int64_t> ipiv = {3, 2, 7, 4, 6, 6, 7};
    Pivots piv = detail::pivots_from_ipiv(ipiv, 3);
    CHECK(piv.size() == 3, "three block columns");
    CHECK(piv[0].size() == 3 && piv[1].size() == 3 && piv[2].size() == 1, "panel sizes");
    // row 2 (0-based) in panel 0: tile 0, offset 2; row 6 in panel 0: tile 2, offset 0
    CHECK(piv[0][0] == Pivot(0, 2), "piv[0][0]");
    CHECK(piv[0][2] == Pivot(2, 0), "piv[0][2]");
    // panel 1 starts at tile 1: row 3 -> tile index 0 relative, offset 0
    CHECK(piv[1][0] == Pivot(0, 0), "piv[1][0]");
    CHECK(piv[1][1] == Pivot(0, 2), "piv[1][1]");
    CHECK(piv[2][0] == Pivot(0, 0), "piv[2][0]");
    CHECK(detail::ipiv_from_pivots(piv, 3, 7) == ipiv, "round trip");
    bool threw = false;
    try { detail::ipiv_from_pivots(piv, 3, 8); } catch (const Exception&) { threw = true; }
    CHECK(threw, "size mismatch throws");

    // Options -> C struct: SLATE defaults and overrides.
    slate_amd_options c = detail::to_c(Options(), 256);
    CHECK(c.block_size == 256 && c.lookahead == 1 && c.max_iterations == 30, "defaults");
    CHECK(c.pivot_threshold == 1.0 && c.use_fallback == 1 && c.tolerance <= 0.0, "defaults 2");
    c = detail::to_c({{Option::Lookahead, 0}, {Option::MaxIterations, 7}, {Option::Tolerance, 1e-9},
                      {Option::PivotThreshold, 0.5}, {Option::UseFallbackSolver, false},
                      {Option::FactorPrecision, FactorPrecision::BF16}}, 128);
    CHECK(c.lookahead == 0 && c.max_iterations == 7 && c.tolerance == 1e-9, "overrides");
    CHECK(c.pivot_threshold == 0.5 && c.use_fallback == 0 && c.factor_precision == SLATE_AMD_PREC_BF16,
          "overrides 2");
    threw = false;
    try { detail::to_c({{Option::Target, Target::HostTask}}, 64); } catch (const NotImplemented&) { threw = true; }
    CHECK(threw, "HostTask rejected");

    // Shape errors surface before any device call.
    std::vector<double> A(12, 1.0), B(3, 1.0);
    threw = false;
    try { Matrix<double>::fromLAPACK(4, 3, A.data(), 2, 2, 1, 1); } catch (const Exception&) { threw = true; }
    CHECK(threw, "lda < m rejected");
    threw = false;
    try { Matrix<double>::fromLAPACK(4, 3, A.data(), 4, 2, 2, 1); } catch (const NotImplemented&) { threw = true; }
    CHECK(threw, "2 x 1 grid rejected");
    auto R = Matrix<double>::fromLAPACK(4, 3, A.data(), 4, 2, 1, 1);
    auto Bm = Matrix<double>::fromLAPACK(3, 1, B.data(), 3, 2, 1, 1);
    Pivots p2;
    threw = false;
    try { getrf(R, p2); } catch (const Exception&) { threw = true; }
    CHECK(threw, "non-square getrf rejected");
    auto RT = transpose(R);
    CHECK(RT.m() == 3 && RT.n() == 4 && RT.op() == Op::Trans, "transpose view");
    CHECK(R.mt() == 2 && R.nt() == 2 && R.tileNb(1) == 1, "tiles");
    (void)Bm;

    std::printf("%s (%d failures)\n", failures == 0 ? "ALL PASSED" : "FAILED", failures);
    return failures == 0 ? 0 : 1;
}
