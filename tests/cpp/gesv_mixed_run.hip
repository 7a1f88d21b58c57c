// The benchmark workload through the C ABI alone (no Python, no torch): one
// plain copy kernel of known size (the calibration point for HBM byte
// counters), then matgen rands A (seed 0) and b (seed 1) generated on the GPU
// and one slate_amd_dsgesv (gesv_mixed) call, timed.  Used to profile the
// solve with rocprofv3 counter collection:
//   gesv_mixed_run [n=65536] [factor_precision=4] [copy_mib=1024]
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "slate_amd.h"

#define CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_)); return 2; } } while (0)

// y = 2 x over `count` floats: reads and writes count * 4 bytes once each.
__global__ void copy_probe(const float* __restrict__ x, float* __restrict__ y, int64_t count)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < count; i += int64_t(gridDim.x) * 256)
        y[i] = 2.0f * x[i];
}

int main(int argc, char** argv)
{
    const int64_t n = argc > 1 ? std::atoll(argv[1]) : 65536;
    const int prec = argc > 2 ? std::atoi(argv[2]) : SLATE_AMD_PREC_BF16X3;
    const int64_t copy_mib = argc > 3 ? std::atoll(argv[3]) : 1024;
    if (n <= 0 || copy_mib < 0) { std::fprintf(stderr, "bad arguments\n"); return 2; }

    if (copy_mib > 0) {
        const int64_t count = copy_mib * (int64_t(1) << 20) / 4;
        float *x = nullptr, *y = nullptr;
        CHECK(hipMalloc(&x, sizeof(float) * count));
        CHECK(hipMalloc(&y, sizeof(float) * count));
        CHECK(hipMemset(x, 0, sizeof(float) * count));
        CHECK(hipDeviceSynchronize());
        hipLaunchKernelGGL(copy_probe, dim3(4096), dim3(256), 0, 0, x, y, count);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        std::printf("copy probe: %lld MiB read + %lld MiB written\n", (long long)copy_mib, (long long)copy_mib);
        CHECK(hipFree(x));
        CHECK(hipFree(y));
    }

    slate_amd_context ctx = nullptr;
    if (slate_amd_init(&ctx, 0) != 0) { std::fprintf(stderr, "init: %s\n", slate_amd_last_error()); return 2; }
    double *A = nullptr, *B = nullptr, *X = nullptr;
    CHECK(hipMalloc(&A, sizeof(double) * size_t(n) * size_t(n)));
    CHECK(hipMalloc(&B, sizeof(double) * size_t(n)));
    CHECK(hipMalloc(&X, sizeof(double) * size_t(n)));
    if (slate_amd_generate(ctx, 2, 0, n, n, 0, 0, A, n) != 0 || slate_amd_generate(ctx, 2, 1, n, 1, 0, 0, B, n) != 0) {
        std::fprintf(stderr, "generate: %s\n", slate_amd_last_error());
        return 2;
    }
    std::vector<int64_t> ipiv(static_cast<size_t>(n));
    slate_amd_options opts;
    slate_amd_options_init(&opts);
    opts.factor_precision = prec;
    int iter = 0;
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t info = slate_amd_dsgesv(ctx, n, 1, A, n, ipiv.data(), B, n, X, n, &iter, &opts);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("gesv_mixed n=%lld factor_precision=%d: info %lld iter %d  %.1f ms  %.2f TFLOP/s\n",
                (long long)n, prec, (long long)info, iter, ms, 2.0 / 3.0 * double(n) * n * n / (ms * 1e-3) / 1e12);
    CHECK(hipFree(A));
    CHECK(hipFree(B));
    CHECK(hipFree(X));
    slate_amd_finalize(ctx);
    return info == 0 && iter >= 0 ? 0 : 1;
}
