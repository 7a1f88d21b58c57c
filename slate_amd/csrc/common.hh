// Shared device/host helpers for the slate_amd HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sa {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
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

// Cross-workgroup publication fences.  Release/acquire atomics already order
// the publishing thread's own accesses; SA_CUMULATIVE_FENCE additionally
// issues a full agent-scope fence (L2 writeback + invalidate) where data
// written by other lanes/waves of the group is published by one thread.
#ifndef SA_FENCE_MODE
#define SA_FENCE_MODE 1
#endif
__device__ inline void sa_publish_fence()
{
#if SA_FENCE_MODE
    __threadfence();
#endif
}

// Tagged-word exchange between workgroups.  Each 64-bit word carries 32 bits
// of payload and a 32-bit tag naming the step that wrote it; words are written
// and read with relaxed agent-scope atomics, so a reader that sees the tag it
// waits for also sees that word's payload (single-location atomicity and
// coherence), with no fences and no separate ready flag.  A value wider than
// 32 bits spans several words, each tagged.
__device__ inline void tw_store(uint64_t* p, uint32_t payload, uint32_t tag)
{
    __hip_atomic_store(p, (uint64_t(tag) << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t tw_load(const uint64_t* p)
{
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Spin (bounded, ~4 s) until the word carries `tag`; returns the payload.
// On timeout sets *error = code and returns 0.
__device__ inline uint32_t tw_wait(const uint64_t* p, uint32_t tag, int* error, int code)
{
    uint64_t w = tw_load(p);
    if (uint32_t(w >> 32) == tag) return uint32_t(w);
    uint64_t t0 = wall_clock64();
    while (true) {
        __builtin_amdgcn_s_sleep(1);
        w = tw_load(p);
        if (uint32_t(w >> 32) == tag) return uint32_t(w);
        if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return 0;
        if (wall_clock64() - t0 > 400000000ull) {
            __hip_atomic_store(error, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 0;
        }
    }
}
// Batched wait: the N words p + off[u] (u with valid[u]) are loaded together
// and reloaded together until every one carries `tag`, so one iteration costs
// one memory round trip however many words are still missing (a word-by-word
// wait would serialise the round trips).  Bounded like tw_wait.
template <int N>
__device__ inline void tw_gather(const uint64_t* p, const int64_t (&off)[N], const bool (&valid)[N],
                                 uint32_t tag, uint32_t (&out)[N], int* error, int code)
{
    uint64_t w[N];
    #pragma unroll
    for (int u = 0; u < N; ++u) w[u] = tw_load(p + off[u]);
    uint64_t t0 = 0;
    bool started = false;
    while (true) {
        bool ok = true;
        #pragma unroll
        for (int u = 0; u < N; ++u) ok = ok && (!valid[u] || uint32_t(w[u] >> 32) == tag);
        if (ok) break;
        if (!started) { t0 = wall_clock64(); started = true; }
        else if (wall_clock64() - t0 > 400000000ull) {
            __hip_atomic_store(error, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        __builtin_amdgcn_s_sleep(1);
        #pragma unroll
        for (int u = 0; u < N; ++u) w[u] = tw_load(p + off[u]);
    }
    #pragma unroll
    for (int u = 0; u < N; ++u) out[u] = uint32_t(w[u]);
}

template <class T> struct TwWords { static constexpr int n = sizeof(T) / 4; };
// A value from consecutive 32-bit payload words (as gathered into LDS).
template <class T>
__device__ inline T lds_word_val(const uint32_t* w)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, (uint64_t(w[1]) << 32) | w[0]);
    else return __builtin_bit_cast(T, w[0]);
}
template <class T>
__device__ inline void tw_store_val(uint64_t* p, T v, uint32_t tag)
{
    if constexpr (sizeof(T) == 8) {
        uint64_t u = __builtin_bit_cast(uint64_t, v);
        tw_store(p, uint32_t(u), tag);
        tw_store(p + 1, uint32_t(u >> 32), tag);
    }
    else {
        tw_store(p, __builtin_bit_cast(uint32_t, v), tag);
    }
}
template <class T>
__device__ inline T tw_wait_val(const uint64_t* p, uint32_t tag, int* error, int code)
{
    if constexpr (sizeof(T) == 8) {
        uint64_t lo = tw_wait(p, tag, error, code);
        uint64_t hi = tw_wait(p + 1, tag, error, code);
        return __builtin_bit_cast(T, (hi << 32) | lo);
    }
    else {
        return __builtin_bit_cast(T, tw_wait(p, tag, error, code));
    }
}

// Wavefront (64-lane) reductions via cross-lane shuffles.
template <class T>
__device__ inline T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__device__ inline T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
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

// Tile t of the upper triangle (tn >= tm) of a T x T tile grid, enumerated
// row by row: row tm starts at off(tm) = tm (2T - tm + 1) / 2.  Launching only
// these tiles keeps every workgroup (and every XCD's band) busy in the
// triangular update, where a full grid would idle below the diagonal.
__device__ inline void upper_tile(int64_t t, int64_t T, int64_t& tm, int64_t& tn)
{
    const double b = double(2 * T + 1);
    int64_t r = int64_t((b - sqrt(b * b - 8.0 * double(t))) * 0.5);
    auto off = [T](int64_t x) { return x * (2 * T - x + 1) / 2; };
    if (r < 0) r = 0;
    if (r > T - 1) r = T - 1;
    while (r > 0 && off(r) > t) --r;
    while (r + 1 < T && off(r + 1) <= t) ++r;
    tm = r;
    tn = r + (t - off(r));
}

} // namespace sa
