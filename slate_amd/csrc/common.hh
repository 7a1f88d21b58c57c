// Shared device/host helpers for the slate_amd HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sa {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Host-side error with a message that the C ABI turns into a negative return code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SA_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            throw ::sa::Error(-100, std::string(#expr) + ": " + hipGetErrorString(e_) \
                              + " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define SA_CHECK_LAUNCH() SA_HIP(hipGetLastError())

#define SA_REQUIRE(cond, code, msg)                                               \
    do { if (!(cond)) throw ::sa::Error((code), (msg)); } while (0)

__host__ __device__ inline int64_t ceildiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Wavefront (64-lane) reductions via cross-lane shuffles.
template <class T>
__device__ inline T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Map a linear block id onto tiles so that the round-robin XCD dispatch
// (block b runs on XCD b % 8) hands each XCD a contiguous chunk of tiles.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblocks)
{
    int64_t q = nblocks / 8;
    if (bid < 8 * q) return (bid % 8) * q + bid / 8;
    return bid;
}

} // namespace sa
