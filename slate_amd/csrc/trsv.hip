// Single-launch triangular solve with one right-hand side (gfx950).
//
// The refinement loops of gesv_mixed / posv_mixed / the GMRES preconditioner
// solve with the low-precision factors once or twice per iteration and one
// vector at a time; SLATE does this with work::trsmA's block sweep, one
// batched trsm + gemm launch pair per block row (src/work/work_trsmA.cc,
// reached from src/getrs.cc:94-97).  Here the whole sweep is one launch:
// block q (BS rows) is handled by workgroup q, which accumulates the products
// with every already-solved block as soon as that block's solved piece has
// been published (a "look-back" chain: blocks finish in order) and then
// publishes its own piece.
//
// Per factor, trsv_diag_inverse prepares two BS x BS matrices per block b:
// the inverse D_b^{-1} of its diagonal block (the inverted-diagonal-block
// form of trsm that GPU BLAS libraries use) and P_b = D_b^{-1} M_{b,l}, l the
// block solved just before b.  With r' = x_b - sum over the other solved
// blocks j of M_{b,j} x_j,
//     x_b = D_b^{-1} r' - P_b x_l,
// so everything but the last term is done while block l is still being
// solved, and the chain's critical path per block is: look-back wait for
// x_l, one BS x BS matrix-vector product with P_b held in registers, a
// lane-shuffle reduction, the publication.
//
// Workgroups are dispatched in index order and only wait on lower indices,
// so the chain always progresses; every wait is bounded and reports a timeout
// through an error flag instead of hanging.
#include "common.hh"
#include "kernels.hh"

namespace sa {

template <class T> struct TvBS { static constexpr int v = 128; };
template <> struct TvBS<double> { static constexpr int v = 64; };

int64_t trsv_block(int elem_bytes) { return elem_bytes == 8 ? TvBS<double>::v : TvBS<float>::v; }

int64_t trsv_aux_elems(int64_t n, int elem_bytes)
{
    const int64_t bs = trsv_block(elem_bytes);
    return 2 * ceildiv(std::max<int64_t>(n, 1), bs) * bs * bs;
}

//------------------------------------------------------------------------------
// Aux[b*BS*BS + t*BS + k] = D_b^{-1}(t, k) (padded with identity past n) and
// Aux[(nblk + b)*BS*BS + t*BS + k] = P_b(t, k) = (D_b^{-1} M_{b,l})(t, k),
// l = b - 1 (LOWER) or b + 1 (upper), zero where l does not exist and in the
// rows and columns past n.
// tri(i, k) = M[i*ld + k] (CM == 0) or M[i + k*ld] (CM == 1).
template <class T, int BS, int LOWER, int UNIT, int CM>
__global__ __launch_bounds__(256)
void diag_inverse_kernel(const T* __restrict__ M, int64_t ld, int64_t n, T* __restrict__ Aux)
{
    __shared__ T Tl[BS][BS + 1];
    __shared__ T Xs[BS][BS + 1];          // Xs[i][t]: row i of inverse column t
    const int nblk = int(ceildiv(n, BS));
    const int b = blockIdx.x;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x;
    auto elem = [&](int64_t i, int64_t k) { return CM ? M[i + k * ld] : M[i * ld + k]; };
    for (int e = tid; e < BS * BS; e += 256) {
        int i = e / BS, k = e % BS;
        T v;
        if (i < rows && k < rows) v = elem(r0 + i, r0 + k);
        else v = (i == k) ? T(1) : T(0);
        Tl[i][k] = v;
    }
    __syncthreads();
    // thread t < BS solves tri * x = e_t by substitution, x kept in Xs[:, t]
    if (tid < BS) {
        const int t = tid;
        if (LOWER) {
            for (int i = 0; i < BS; ++i) {
                T s = (i == t) ? T(1) : T(0);
                for (int k = t; k < i; ++k) s -= Tl[i][k] * Xs[k][t];
                if (i < t) s = T(0);
                Xs[i][t] = UNIT ? s : s / Tl[i][i];
            }
        }
        else {
            for (int i = BS - 1; i >= 0; --i) {
                T s = (i == t) ? T(1) : T(0);
                for (int k = i + 1; k <= t; ++k) s -= Tl[i][k] * Xs[k][t];
                if (i > t) s = T(0);
                Xs[i][t] = UNIT ? s : s / Tl[i][i];
            }
        }
    }
    __syncthreads();
    T* out = Aux + int64_t(b) * BS * BS;
    for (int e = tid; e < BS * BS; e += 256) out[e] = Xs[e / BS][e % BS];
    // the off-diagonal tile of the block solved just before this one
    const int l = LOWER ? b - 1 : b + 1;
    const bool has = l >= 0 && l < nblk;
    const int64_t c0 = int64_t(has ? l : 0) * BS;
    const int cols = has ? int((n - c0 < BS) ? n - c0 : BS) : 0;
    for (int e = tid; e < BS * BS; e += 256) {
        int i = e / BS, k = e % BS;
        Tl[i][k] = (i < rows && k < cols) ? elem(r0 + i, c0 + k) : T(0);
    }
    __syncthreads();
    T* pout = Aux + int64_t(nblk + b) * BS * BS;
    for (int e = tid; e < BS * BS; e += 256) {
        const int t = e / BS, k = e % BS;
        T s = T(0);
        for (int m = 0; m < BS; ++m) s += Xs[t][m] * Tl[m][k];
        pout[e] = s;
    }
}

template <class T>
void trsv_diag_inverse(const T* M, int64_t ld, int cm, int lower, int unit, int64_t n, T* Aux,
                       hipStream_t s)
{
    if (n <= 0) return;
    constexpr int BS = TvBS<T>::v;
    unsigned nblk = unsigned(ceildiv(n, BS));
    int key = (lower ? 4 : 0) | (unit ? 2 : 0) | (cm ? 1 : 0);
#define SA_DI_CASE(K, LO, UN, CMV)                                                                  \
    case K:                                                                                         \
        hipLaunchKernelGGL((diag_inverse_kernel<T, BS, LO, UN, CMV>), dim3(nblk), dim3(256), 0, s, \
                           M, ld, n, Aux);                                                          \
        break;
    switch (key) {
        SA_DI_CASE(0, 0, 0, 0)
        SA_DI_CASE(1, 0, 0, 1)
        SA_DI_CASE(2, 0, 1, 0)
        SA_DI_CASE(3, 0, 1, 1)
        SA_DI_CASE(4, 1, 0, 0)
        SA_DI_CASE(5, 1, 0, 1)
        SA_DI_CASE(6, 1, 1, 0)
        default:
        SA_DI_CASE(7, 1, 1, 1)
    }
#undef SA_DI_CASE
    SA_CHECK_LAUNCH();
}

//------------------------------------------------------------------------------
// Sum of v over the lanes whose index differs only in the bits of MASK
// (a butterfly: every lane of the group ends with the same sum).
template <int MASK, class T>
__device__ inline T lanes_sum(T v)
{
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (MASK & o) v += __shfl_xor(v, o, 64);
    return v;
}

// Row sums of N rows of per-lane partials spread over the 32 lanes of a
// half-wave: a register butterfly in which each exchange halves the rows a
// lane keeps (the lanes whose index bit o is set keep the upper half and add
// their partner's partials of it), o = 16, 8, ...  Afterwards a[0] holds the
// sum over the lanes that agree in the exchanged bits, for row
// (lane >> (5 - log2 N)) & (N - 1).  H is a template argument so every
// register index is a constant (a run-time index into a register array
// becomes a compare-and-select chain over the whole array).
template <int H, int N, class T>
__device__ __forceinline__ void row_butterfly(T (&a)[N], int lane)
{
    constexpr int o = 32 * H / N;
    const bool up = (lane & o) != 0;
    #pragma unroll
    for (int i = 0; i < H; ++i) {
        const T send = up ? a[i] : a[i + H];
        const T keep = up ? a[i + H] : a[i];
        a[i] = keep + __shfl_xor(send, o, 64);
    }
    if constexpr (H > 1) row_butterfly<H / 2, N>(a, lane);
}

// Look-back sweep.  Block q publishes its solved piece twice: plainly into X
// (the result) and as tagged words (32-bit payload + the launch's tag) into
// xw, which is what later blocks read: a reader that sees the tag has the
// value, so there is no completion counter on the critical path and no
// cache-wide fence.  A separate progress word (tag, blocks done) only paces
// the waiting: blocks far from their turn sleep in proportion to the
// distance instead of polling.
//
// Each earlier block's BS x BS tile of the factor is loaded into registers
// before its x arrives (it does not depend on x), so after the wait only the
// x words, one LDS pass and the FMAs remain.  Row-major tiles are read with
// 16-byte loads: 32 lanes cover one row of the tile (VEC elements each), a
// wave two rows per load, so every cache line is fetched by one instruction;
// lane l keeps NI rows x VEC columns and the per-lane partial row sums are
// folded by a register butterfly.  Column-major tiles are spread as one row
// per lane and a wave-uniform column range per wave (scalar column
// addresses), folded through LDS.
template <class T>
__device__ inline void tv_put(uint64_t* p, T v, uint32_t tag) { tw_store_val<T>(p, v, tag); }

template <class T, int BS, int LOWER, int CM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void trsv_lookback_kernel(const T* __restrict__ M, int64_t ld, int64_t n, const T* __restrict__ Aux,
                          T* __restrict__ X, uint64_t* __restrict__ xw, uint32_t tag,
                          uint64_t* __restrict__ progress, int* error)
{
    constexpr int NWV = int(sizeof(T)) / 4;
    constexpr int VEC = 16 / int(sizeof(T));    // elements per 16-byte load (row-major tile)
    constexpr int NI = BS * BS / (256 * VEC);   // row-major tile rows per lane
    constexpr int LG = (NI == 16) ? 4 : (NI == 8) ? 3 : 2;         // log2 NI
    constexpr int NPART = 256 / BS;             // lanes per row (column-major and diagonal products)
    constexpr int KP = BS / NPART;              // columns per lane
    constexpr int DLD = BS + 16 / int(sizeof(T));                  // padded row of D^{-1} in LDS
    static_assert(BS == 32 * VEC && NI == (1 << LG) && NPART * KP == BS, "trsv tiling");
    const int nblk = int(ceildiv(n, BS));
    const int q = blockIdx.x;
    const int b = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = int64_t(b) * BS;
    const int rows = int((n - r0 < BS) ? n - r0 : BS);
    const int tid = threadIdx.x, lane = tid & 63;
    // Wave-uniform, so a tile's row (row-major) or column (column-major)
    // bases are scalar; that keeps the kernel at two waves per SIMD, so every
    // block of an n = 65536 solve is resident from the start.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int vc = VEC * (lane & 31), vr = lane >> 5;   // row-major tile: column group, row parity
    const int it = tid / NPART, ih = tid % NPART;       // row and column range of the diagonal products
    const int ct = tid % BS, cw = (wave * 64) / BS;     // row and column range of a column-major tile

    __shared__ __attribute__((aligned(16))) uint32_t sxw[2][BS * NWV];   // x words of a dependency
    __shared__ __attribute__((aligned(16))) T sr[BS];                    // r'
    __shared__ __attribute__((aligned(16))) T sdinv[BS * DLD];           // D_b^{-1}
    __shared__ T spart[CM ? NPART : 1][BS];                              // column-major partial sums

    // P_b in registers (lane (it, ih) holds P_b(it, ih*KP .. ih*KP+KP-1)),
    // D_b^{-1} in LDS, and this block's right-hand side entries.
    T pm[KP];
    {
        const T* pp = Aux + int64_t(nblk + b) * BS * BS + it * BS + ih * KP;
        #pragma unroll
        for (int k = 0; k < KP; ++k) pm[k] = pp[k];
        const T* dp = Aux + int64_t(b) * BS * BS;
        for (int e = tid; e < BS * BS; e += 256) sdinv[(e / BS) * DLD + e % BS] = dp[e];
    }
    // the row whose r' this lane finishes (row-major: lane l keeps rows
    // wave*2*NI + 2i + (l >> 5), and the butterfly leaves i = (l >> (5 - LG)) & (NI - 1))
    const int xrow = CM ? ct : wave * 2 * NI + 2 * ((lane >> (5 - LG)) & (NI - 1)) + vr;
    const T xrhs = X[r0 + (xrow < rows ? xrow : rows - 1)];

    const int ndep = LOWER ? b : nblk - 1 - b;
    T acc_rm[NI];
    T acc_cm = T(0);
    #pragma unroll
    for (int r = 0; r < NI; ++r) acc_rm[r] = T(0);

    // Tile registers.  Loads use clamped addresses (edge values are zeroed
    // when applied), so a tile's loads issue back to back.  A tile narrower
    // than BS (the last block column) is only ever dependency 0 of an upper
    // solve, so it takes element loads before the loop; inside the loop every
    // row-major tile is BS wide and is read with 16-byte loads, the rows
    // clamped through a 32-bit lane offset from the block's base.
    T mrm[CM ? 1 : NI][CM ? 1 : VEC];
    T mcm[CM ? KP : 1];
    auto tile_cols = [&](int t) {
        const int64_t c0 = int64_t(LOWER ? t : nblk - 1 - t) * BS;
        return int((n - c0 < BS) ? n - c0 : BS);
    };
    auto load_tile = [&](int t) {
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        if constexpr (CM) {
            const int cols = tile_cols(t);
            const T* mp = M + r0 + (ct < rows ? ct : rows - 1) + c0 * ld;
            #pragma unroll
            for (int c = 0; c < KP; ++c) {
                const int cc = cw * KP + c;
                mcm[c] = mp[int64_t(cc < cols ? cc : cols - 1) * ld];
            }
        }
        else {
            // row 2r + vr of the wave's rows: a wave-uniform (clamped) base
            // row and one of two lane offsets (the odd row, or the base row
            // again where the odd row lies past the block's last row)
            const T* sb = M + r0 * ld + c0;
            const uint32_t off0 = uint32_t(vc * int(sizeof(T)));
            const uint32_t off1 = uint32_t((int64_t(vr) * ld + vc) * int64_t(sizeof(T)));
            // the factor is streamed once per solve: nontemporal loads
            typedef T tv_vec __attribute__((ext_vector_type(VEC), aligned(sizeof(T))));
            #pragma unroll
            for (int r = 0; r < NI; ++r) {
                const int rb = wave * 2 * NI + 2 * r;
                const char* base = reinterpret_cast<const char*>(sb + int64_t(rb < rows ? rb : rows - 1) * ld);
                const tv_vec v = __builtin_nontemporal_load(
                    reinterpret_cast<const tv_vec*>(base + (rb + 1 < rows ? off1 : off0)));
                #pragma unroll
                for (int u = 0; u < VEC; ++u) mrm[r][u] = v[u];
            }
        }
    };
    auto load_tile_edge = [&](int t) {            // row-major, narrower than BS
        const int64_t c0 = int64_t(LOWER ? t : nblk - 1 - t) * BS;
        const int cols = tile_cols(t);
        #pragma unroll
        for (int r = 0; r < NI; ++r) {
            const int i = wave * 2 * NI + 2 * r + vr;
            const T* mp = M + (r0 + (i < rows ? i : rows - 1)) * ld + c0;
            #pragma unroll
            for (int u = 0; u < VEC; ++u) {
                const int c = vc + u;
                mrm[r][u] = mp[c < cols ? c : cols - 1];
            }
        }
    };

    const uint64_t t_start = wall_clock64();
    bool failed = false;
    unsigned known = 0;                         // thread 0: chain progress last seen
    // Wait until dependency t's x words are in sxw[t & 1].
    auto wait_x = [&](int t) {
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        // pace: sleep while the chain is more than two blocks away from t;
        // progress is re-read only while the last value seen is that far
        if (tid == 0 && !failed) {
            while (known + 2 < unsigned(t + 1)) {
                const uint64_t pw = tw_load(progress);
                const unsigned done = (uint32_t(pw >> 32) == tag) ? uint32_t(pw) : 0u;
                known = done > known ? done : known;
                if (known + 2 >= unsigned(t + 1)) break;
                const unsigned dist = unsigned(t + 1) - known;
                const unsigned naps = dist > 16 ? 8 : dist / 2;
                for (unsigned z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(32);
                if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0
                    || wall_clock64() - t_start > 400000000ull) {
                    __hip_atomic_store(error, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    failed = true;
                    break;
                }
            }
        }
        __syncthreads();
        // the dependency's x words (bounded spin inside tw_wait)
        if (tid < BS * NWV) {
            uint32_t wv = 0;
            if (tid < cols * NWV) wv = tw_wait(xw + c0 * NWV + tid, tag, error, 2);
            sxw[t & 1][tid] = wv;
        }
        __syncthreads();
    };

    // ---- every dependency but the last: r' accumulation
    if (ndep > 1) {
        if (!CM && tile_cols(0) < BS) load_tile_edge(0);
        else load_tile(0);
    }
    #pragma unroll 1
    for (int t = 0; t + 1 < ndep; ++t) {
        const int j = LOWER ? t : nblk - 1 - t;
        const int64_t c0 = int64_t(j) * BS;
        const int cols = int((n - c0 < BS) ? n - c0 : BS);
        const int buf = t & 1;
        wait_x(t);
        if (CM) {
            #pragma unroll
            for (int c = 0; c < KP; ++c) {
                const int cc = cw * KP + c;
                const T xv = lds_word_val<T>(&sxw[buf][cc * NWV]);
                acc_cm += (cc < cols) ? mcm[c] * xv : T(0);
            }
        }
        else {
            #pragma unroll
            for (int u = 0; u < VEC; ++u) {
                const int c = vc + u;
                const T xv = lds_word_val<T>(&sxw[buf][c * NWV]);
                const T y = (c < cols) ? xv : T(0);
                #pragma unroll
                for (int r = 0; r < NI; ++r) acc_rm[r] += mrm[r][u] * y;
            }
        }
        // No error-flag read here: a load issued after the next tile's loads
        // completes after them, and thread 0 would wait for the whole tile
        // before its next poll (two memory round trips per dependency
        // instead of one).  The waits read the flag while they spin.
        if (t + 2 < ndep) load_tile(t + 1);
    }
    // ---- r' of this block's rows into LDS
    if (CM) {
        spart[cw][ct] = acc_cm;
        __syncthreads();
        if (tid < BS) {
            T sum = T(0);
            #pragma unroll
            for (int p = 0; p < NPART; ++p) sum += spart[p][tid];
            sr[tid] = (tid < rows) ? xrhs - sum : T(0);
        }
    }
    else {
        row_butterfly<NI / 2, NI>(acc_rm, lane);
        const T sum = lanes_sum<(1 << (5 - LG)) - 1>(acc_rm[0]);
        if ((lane & ((1 << (5 - LG)) - 1)) == 0) sr[xrow] = (xrow < rows) ? xrhs - sum : T(0);
    }
    __syncthreads();
    // ---- y' = D_b^{-1} r' (row it, folded over the NPART lanes of the row)
    T yv;
    {
        T a0 = T(0), a1 = T(0);
        const T* dr = sdinv + it * DLD + ih * KP;
        const T* rr = sr + ih * KP;
        #pragma unroll
        for (int k = 0; k < KP; k += 2) {
            a0 += dr[k] * rr[k];
            a1 += dr[k + 1] * rr[k + 1];
        }
        yv = lanes_sum<NPART - 1>(a0 + a1);
    }
    // ---- the last dependency: x_b = y' - P_b x_l
    if (ndep > 0) {
        const int t = ndep - 1;
        wait_x(t);
        const uint32_t* xs = &sxw[t & 1][ih * KP * NWV];
        T z0 = T(0), z1 = T(0);
        #pragma unroll
        for (int k = 0; k < KP; k += 2) {
            z0 += pm[k] * lds_word_val<T>(xs + k * NWV);
            z1 += pm[k + 1] * lds_word_val<T>(xs + (k + 1) * NWV);
        }
        yv -= lanes_sum<NPART - 1>(z0 + z1);
    }
    if (ih == 0 && it < rows) {
        tv_put<T>(xw + (r0 + it) * NWV, yv, tag);
        X[r0 + it] = yv;
    }
    if (tid == 0)
        __hip_atomic_fetch_max(progress, (uint64_t(tag) << 32) | uint32_t(q + 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
void trsv_lookback(const T* M, int64_t ld, int cm, int lower, int64_t n, const T* Aux, T* X,
                   const TvSync& sync, hipStream_t s)
{
    if (n <= 0) return;
    constexpr int BS = TvBS<T>::v;
    SA_REQUIRE(sync.tag != 0, -2, "trsv_lookback: tag 0 is reserved");
    unsigned nblk = unsigned(ceildiv(n, BS));
    int key = (lower ? 2 : 0) | (cm ? 1 : 0);
#define SA_TV_CASE(K, LO, CMV)                                                                      \
    case K:                                                                                         \
        hipLaunchKernelGGL((trsv_lookback_kernel<T, BS, LO, CMV>), dim3(nblk), dim3(256), 0, s,    \
                           M, ld, n, Aux, X, sync.words, sync.tag, sync.progress, sync.error);     \
        break;
    switch (key) {
        SA_TV_CASE(0, 0, 0)
        SA_TV_CASE(1, 0, 1)
        SA_TV_CASE(2, 1, 0)
        default:
        SA_TV_CASE(3, 1, 1)
    }
#undef SA_TV_CASE
    SA_CHECK_LAUNCH();
}

template void trsv_diag_inverse<double>(const double*, int64_t, int, int, int, int64_t, double*, hipStream_t);
template void trsv_diag_inverse<float>(const float*, int64_t, int, int, int, int64_t, float*, hipStream_t);
template void trsv_lookback<double>(const double*, int64_t, int, int, int64_t, const double*, double*,
                                    const TvSync&, hipStream_t);
template void trsv_lookback<float>(const float*, int64_t, int, int, int64_t, const float*, float*,
                                   const TvSync&, hipStream_t);

} // namespace sa
