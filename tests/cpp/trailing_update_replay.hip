// Replays the trailing-update launches of the n = 65536 float LU (nb = 1024,
// lookahead 1) standalone: for each step k, the s0 update of the columns
// beyond the lookahead block, C (r x (r - nb)) -= L (r x nb) U (nb x (r - nb)),
// with the same shapes, strides and kernels as the factorization, on a
// resident 65536 x 65536 float matrix: kernel "lt" (the factorization's
// choice for one- and two-term splits) splits the L panel and the U block
// into bf16 planes and runs gemm_lt_kernel<terms, 1> where the update has at
// least 512 of its 256 x 256 tiles and gemm_sp_kernel<terms, 1, 3> below;
// kernel "sp" runs gemm_sp_kernel<terms, 1, 3> for every step.  A plain copy kernel of known size runs first:
// the calibration point for HBM byte counters.  Kernel-level counters
// (FETCH_SIZE, WRITE_SIZE) of a launch depend on its shape, not on the
// values, so this gives the per-launch HBM traffic of the benchmark's
// dominant kernel under counter collection.
//   trailing_update_replay [n=65536] [nb=1024] [terms=2] [copy_mib=1024] [kernel=lt]
#include "../../slate_amd/csrc/gemm_bf16.hip"
#include "../../slate_amd/csrc/gemm_bf16_lds.hip"
#include <chrono>
#include <string>
#include <cstdio>
#include <cstdlib>

#define CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); return 2; } } while (0)

__global__ void copy_probe(const float* __restrict__ x, float* __restrict__ y, int64_t count)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < count; i += int64_t(gridDim.x) * 256)
        y[i] = 2.0f * x[i];
}

__global__ void fill_uniform(float* __restrict__ p, int64_t count, uint64_t seed)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < count; i += int64_t(gridDim.x) * 256) {
        uint64_t x = (uint64_t(i) + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        x ^= x >> 29; x *= 0xBF58476D1CE4E5B9ull; x ^= x >> 32;
        p[i] = float(double(x >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    }
}

int main(int argc, char** argv)
{
    const int64_t n = argc > 1 ? std::atoll(argv[1]) : 65536;
    const int64_t nb = argc > 2 ? std::atoll(argv[2]) : 1024;
    const int terms = argc > 3 ? std::atoi(argv[3]) : 2;
    const int64_t copy_mib = argc > 4 ? std::atoll(argv[4]) : 1024;
    const bool lt = !(argc > 5 && std::string(argv[5]) == "sp");
    if (n <= 0 || nb <= 0 || nb % 32 != 0 || n % 4 != 0 || terms < 1 || terms > 3 || (lt && terms > 2)) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    if (copy_mib > 0) {
        const int64_t count = copy_mib * (int64_t(1) << 20) / 4;
        float *x = nullptr, *y = nullptr;
        CHECK(hipMalloc(&x, sizeof(float) * count));
        CHECK(hipMalloc(&y, sizeof(float) * count));
        hipLaunchKernelGGL(fill_uniform, dim3(4096), dim3(256), 0, 0, x, count, uint64_t(7));
        CHECK(hipDeviceSynchronize());
        hipLaunchKernelGGL(copy_probe, dim3(4096), dim3(256), 0, 0, x, y, count);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        std::printf("copy probe: %lld MiB read + %lld MiB written\n", (long long)copy_mib, (long long)copy_mib);
        std::fflush(stdout);
        CHECK(hipFree(x));
        CHECK(hipFree(y));
    }
    float *A = nullptr, *ut = nullptr;
    __bf16 *lp = nullptr, *up = nullptr;
    CHECK(hipMalloc(&A, sizeof(float) * size_t(n) * size_t(n)));
    CHECK(hipMalloc(&ut, sizeof(float) * size_t(n) * size_t(nb)));
    if (lt) {
        CHECK(hipMalloc(&lp, sizeof(__bf16) * size_t(terms) * size_t(n) * size_t(nb)));
        CHECK(hipMalloc(&up, sizeof(__bf16) * size_t(terms) * size_t(n) * size_t(nb)));
    }
    hipLaunchKernelGGL(fill_uniform, dim3(8192), dim3(256), 0, 0, A, n * n, uint64_t(1));
    CHECK(hipDeviceSynchronize());
    const int64_t ld = n, nsteps = (n + nb - 1) / nb;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double gemm_ms = 0, flops = 0;
    int launches = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < nsteps; ++k) {
        const int64_t k0 = k * nb, kb = std::min(nb, n - k0), c1 = k0 + kb, r = n - c1;
        const int64_t nb1 = std::min(nb, r), c2 = c1 + nb1;
        if (r <= nb1) break;                      // the last step has no columns beyond the lookahead block
        const bool use_lt = lt && sa::gemm_lt_tiles(r, n - c2) >= 512;
        if (use_lt) {
            sa::split_rows(terms, A + c1 * ld + k0, ld, r, kb, lp, nb, n * nb, 0);
            sa::split_cols_t(terms, A + k0 * ld + c2, ld, kb, n - c2, up, nb, n * nb, 0);
        }
        else {
            sa::transpose_f32(A + k0 * ld + c2, ld, kb, n - c2, ut, nb, 0);
        }
        CHECK(hipEventRecord(e0, 0));
        if (use_lt)
            sa::gemm_bf16_lt(terms, r, n - c2, nb, lp, nb, n * nb, up, nb, n * nb, A + c1 * ld + c2, ld, -1.0f, 0, 1, 0);
        else
            sa::gemm_s32(terms, r, n - c2, kb, A + c1 * ld + k0, ld, ut, nb, A + c1 * ld + c2, ld, -1.0f, 0, 1, 0);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        gemm_ms += ms;
        flops += 2.0 * double(r) * double(n - c2) * double(kb);
        ++launches;
    }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("replay n=%lld nb=%lld terms=%d: %d trailing updates, %.1f ms in the GEMM (avg %.3f ms, %.1f TFLOP/s "
                "FP32-equivalent), %.1f ms wall\n", (long long)n, (long long)nb, terms, launches, gemm_ms,
                gemm_ms / launches, flops / (gemm_ms * 1e-3) / 1e12, wall);
    CHECK(hipFree(A));
    CHECK(hipFree(ut));
    if (lt) { CHECK(hipFree(lp)); CHECK(hipFree(up)); }
    return 0;
}
