// Panel LU leaf and row interchanges (gfx950).
//
// The leaf kernel factors a tall, narrow block (up to 65536+ rows x W columns)
// with partial pivoting.  It replaces SLATE's host panel (internal::getrf_panel
// <HostTask>, src/internal/internal_getrf.cc:20-121, which pulls the panel to
// the CPU every step) by a GPU kernel in which G workgroups each keep a slice
// of rows in registers.  Per column, each workgroup finds its local pivot
// candidate with wavefront shuffles and publishes it (value, row, and the row's
// W entries) as tagged words to a double-buffered exchange area; every
// workgroup reads the G (value, row) heads (one memory round trip once they
// have landed, with no grid barrier and no fences), reduces them itself,
// fetches the winning row, applies the interchange to the rows it owns, scales
// and performs the rank-1 update.
// Pivot rules follow src/internal/Tile_getrf.hh:197-366: largest |a_ij| with
// the smallest row winning exact ties, the diagonal kept when
// |a_jj| >= threshold * max, scaling by the reciprocal unless the pivot is
// below the smallest normal number, and info = first exactly-zero pivot.
#include "common.hh"
#include "kernels.hh"
#include <limits>

namespace sa {

static constexpr int LEAF_GMAX = 128;     // workgroups per leaf launch

template <class T> struct LeafW;
template <> struct LeafW<double> { static constexpr int W = 32; };
template <> struct LeafW<float> { static constexpr int W = 32; };

int panel_leaf_width(int elem_bytes) { return elem_bytes == 8 ? LeafW<double>::W : LeafW<float>::W; }

// Exchange area (64-bit tagged words, see tw_store in common.hh), per column
// parity: G candidate entries of CW words [|a| (NWV words), row index (1
// word), the candidate row (W values)], then row j (W values).  NWV = words
// per value.  Column j of a launch carries tag tag0 + j; alternating parity
// areas keep a workgroup that runs one column ahead from overwriting words
// that a slower one has not read yet.
template <class T, int W>
struct LeafSlots {
    static constexpr int NWV = int(sizeof(T)) / 4;
    static constexpr int CW = NWV * (W + 1) + 1;
    static constexpr int MAXWORDS = LEAF_GMAX * CW + W * NWV;
    __host__ __device__ static int64_t per(int G) { return (int64_t(G) * CW + W * NWV + 63) & ~int64_t(63); }
};

template <class T>
__device__ inline T lds_val(const uint32_t* w)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, (uint64_t(w[1]) << 32) | w[0]);
    else return __builtin_bit_cast(T, w[0]);
}

template <class T>
__device__ inline void lds_put_val(uint32_t* w, T v)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        w[0] = uint32_t(u);
        w[1] = uint32_t(u >> 32);
    }
    else {
        w[0] = __builtin_bit_cast(uint32_t, v);
    }
}

template <class T>
__device__ inline bool better(T a, int64_t ia, T b, int64_t ib)
{
    return a > b || (a == b && ia < ib);
}

//------------------------------------------------------------------------------
// The leaf's w interchanges (row j <-> row piv[j], sequentially, rows local to
// r0) applied to the other columns of the enclosing panel, [xp0, c0) and
// [c0 + w, xp0 + xpw), so that the recursive panel needs no row-swap launches
// of its own.  Both ranges are untouched while the leaf runs: the right part
// is only read by the solves that follow the leaf, the left part (finished
// columns of L) only needs the interchanges brought back.  Every workgroup
// builds the same list of at most 2w (dst <- src) row moves: a touched row
// r receives the original row t_0(t_1(...t_{w-1}(r))), t_j the j-th
// transposition; then it moves its share of the columns through LDS
// (gather, barrier, scatter), so sources are read before any is overwritten.
template <class T, int W, int NT>
__device__ void leaf_swap_others(T* __restrict__ A, int64_t ld, int64_t r0, int64_t c0, int w, int64_t xp0,
                                 int64_t xpw, const int* s_pivs, int G, int g)
{
    constexpr int CH = 32;                      // columns per chunk
    __shared__ int s_dst[2 * W], s_src[2 * W];
    __shared__ int s_nmv;
    __shared__ T s_stage[2 * W * CH];
    const int tid = threadIdx.x;
    if (tid == 0) s_nmv = 0;
    __syncthreads();                            // s_pivs complete, counter cleared
    if (tid < 2 * W) {
        const int t = tid < W ? tid : tid - W;
        int dst = -1;
        if (t < w) {
            if (tid < W) {
                dst = t;
            }
            else {
                const int pt = s_pivs[t];
                bool first = pt >= w;
                for (int u = 0; u < t; ++u) first = first && s_pivs[u] != pt;
                if (first) dst = pt;
            }
        }
        if (dst >= 0) {
            int x = dst;
            for (int jj = w - 1; jj >= 0; --jj) {
                const int pj = s_pivs[jj];
                x = (x == jj) ? pj : (x == pj ? jj : x);
            }
            if (x != dst) {
                const int slot = atomicAdd(&s_nmv, 1);
                s_dst[slot] = dst;
                s_src[slot] = x;
            }
        }
    }
    __syncthreads();
    const int nmv = s_nmv;
    if (nmv == 0) return;
    const int64_t nother = xpw - w, left = c0 - xp0;
    const int64_t share = (nother + G - 1) / G;
    const int64_t q0 = int64_t(g) * share;
    const int64_t q1 = q0 + share < nother ? q0 + share : nother;
    for (int64_t qc = q0; qc < q1; qc += CH) {
        const int nc = int(q1 - qc < CH ? q1 - qc : CH);
        for (int e = tid; e < nmv * CH; e += NT) {
            const int mv = e / CH, cc = e % CH;
            if (cc < nc) {
                const int64_t q = qc + cc;
                const int64_t col = q < left ? xp0 + q : c0 + w + (q - left);
                s_stage[e] = A[(r0 + s_src[mv]) * ld + col];
            }
        }
        __syncthreads();
        for (int e = tid; e < nmv * CH; e += NT) {
            const int mv = e / CH, cc = e % CH;
            if (cc < nc) {
                const int64_t q = qc + cc;
                const int64_t col = q < left ? xp0 + q : c0 + w + (q - left);
                A[(r0 + s_dst[mv]) * ld + col] = s_stage[e];
            }
        }
        __syncthreads();
    }
}

//------------------------------------------------------------------------------
// Leaf with rotated register columns.  v[q][c] holds column (j + c) mod W at
// step j, so the active column is always v[q][0] and every other access uses
// compile-time register indices (no per-element column selects); one
// circular shift per step keeps the invariant.  Rows are published and kept
// in LDS in the same rotated order.  The exchange is in two stages: every
// workgroup gathers only the G (|a|, row) heads and row j (G*(NWV+1) + W*NWV
// words, about one per thread), decides the pivot, and then wave 0 reads the
// winner's row, which the winner published in the same round (no second read
// when the diagonal is kept).
template <class T, int W, int RPT, int VEC, int NT>
__global__ __launch_bounds__(NT)
void panel_leaf_rot_kernel(T* __restrict__ A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0,
                           int w, int64_t* __restrict__ ipiv, int64_t* __restrict__ info, T threshold,
                           uint32_t tag0, uint64_t* slots, int* error, int64_t xp0, int64_t xpw)
{
    using SL = LeafSlots<T, W>;
    constexpr int NWV = SL::NWV, CW = SL::CW;
    constexpr int HW = NWV + 1;                 // head words of an entry: |a|, row
    constexpr int AMAX = LEAF_GMAX * HW + W * NWV;
    constexpr int NCR = (AMAX + NT - 1) / NT;
    static_assert((W & (W - 1)) == 0 && W % 4 == 0, "W must be a power of two");
    typedef T vec4_t __attribute__((ext_vector_type(4)));
    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m = m_total - r0;
    constexpr int64_t R = int64_t(NT) * RPT;
    const int64_t gbase = int64_t(g) * R;
    const int totalA = G * HW + W * NWV;
    const int64_t per = SL::per(G);

    constexpr int NW = NT / 64;                 // waves per workgroup
    __shared__ T s_wabs[NW];
    __shared__ uint32_t s_widx[NW];
    __shared__ uint32_t s_pub[2][W * NWV];      // rows being published (candidate, row j)
    __shared__ uint32_t s_words[AMAX];
    __shared__ __attribute__((aligned(16))) T s_rowj[W];   // row j, rotated
    __shared__ __attribute__((aligned(16))) T s_piv[W];    // pivot row, rotated
    __shared__ __attribute__((aligned(16))) T s_prot[W];   // pivot row, rotated, 0 outside columns j+1 .. w-1
    __shared__ int64_t s_p;
    __shared__ int s_pivs[W];                   // local pivot row of each column

    constexpr int VE = 16 / int(sizeof(T));
    T v[RPT][W];
    #pragma unroll
    for (int q = 0; q < RPT; ++q) {
        int64_t i = gbase + tid + int64_t(NT) * q;
        const T* rowp = A + (r0 + (i < m ? i : 0)) * ld + c0;
        if (VEC && w == W) {
            #pragma unroll
            for (int c = 0; c < W; c += VE) {
                typedef T vec_t __attribute__((ext_vector_type(VE)));
                vec_t x = *reinterpret_cast<const vec_t*>(rowp + c);
                #pragma unroll
                for (int e = 0; e < VE; ++e) v[q][c + e] = (i < m) ? x[e] : T(0);
            }
        }
        else {
            #pragma unroll
            for (int c = 0; c < W; ++c)
                v[q][c] = (i < m && c < w) ? rowp[c] : T(0);
        }
    }
    const T safe_min = std::numeric_limits<T>::min();
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    auto better = [](T a, uint32_t ia, T b, uint32_t ib) { return a > b || (a == b && ia < ib); };

    #pragma unroll 1
    for (int j = 0; j < w; ++j) {
        uint64_t* buf = slots + (j & 1) * per;
        const uint32_t tag = tag0 + uint32_t(j);

        // (1) local candidate: first row i >= j of maximal |a_ij|
        T babs = T(-1);
        uint32_t bidx = NONE;
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = gbase + tid + int64_t(NT) * q;
            const T a = fabs(v[q][0]);
            const bool take = i >= j && i < m && a > babs;
            babs = take ? a : babs;
            bidx = take ? uint32_t(i) : bidx;
        }
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T oa = __shfl_xor(babs, o, 64);
            const uint32_t oi = __shfl_xor(bidx, o, 64);
            const bool b = better(oa, oi, babs, bidx);
            babs = b ? oa : babs;
            bidx = b ? oi : bidx;
        }
        if (lane == 0) { s_wabs[wave] = babs; s_widx[wave] = bidx; }
        __syncthreads();
        babs = s_wabs[0]; bidx = s_widx[0];
        #pragma unroll
        for (int k = 1; k < NW; ++k) {
            const bool b = better(s_wabs[k], s_widx[k], babs, bidx);
            babs = b ? s_wabs[k] : babs;
            bidx = b ? s_widx[k] : bidx;
        }

        // (2) publish the head (|a|, row) first — it is what every workgroup
        // polls for — then the candidate row and, from the owner of row j,
        // row j (rotated order).  A row is copied by its owning thread into
        // LDS and stored by the lanes of that thread's wave, one word each
        // (one vector store instead of W single-lane stores).
        uint64_t* ent = buf + int64_t(g) * CW;
        if (tid == 0) {
            tw_store_val<T>(ent, babs, tag);
            tw_store(ent + NWV, bidx, tag);
        }
        const int64_t cl = (bidx != NONE) ? int64_t(bidx) - gbase : -1;
        const int64_t jl = int64_t(j) - gbase;
        const int cand_t = (cl >= 0 && cl < R) ? int(cl % NT) : -1;
        const int rowj_t = (jl >= 0 && jl < R) ? int(jl % NT) : -1;
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (tid == cand_t && q == int(cl / NT)) {
                #pragma unroll
                for (int c = 0; c < W; ++c) lds_put_val<T>(&s_pub[0][c * NWV], v[q][c]);
            }
            if (tid == rowj_t && q == int(jl / NT)) {
                #pragma unroll
                for (int c = 0; c < W; ++c) lds_put_val<T>(&s_pub[1][c * NWV], v[q][c]);
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (cand_t >= 0 && wave == cand_t / 64 && lane < W * NWV) tw_store(ent + HW + lane, s_pub[0][lane], tag);
        if (rowj_t >= 0 && wave == rowj_t / 64 && lane < W * NWV)
            tw_store(buf + int64_t(G) * CW + lane, s_pub[1][lane], tag);

        // (3a) gather the heads of all workgroups and row j
        {
            int64_t off[NCR];
            bool val[NCR];
            uint32_t got[NCR];
            #pragma unroll
            for (int u = 0; u < NCR; ++u) {
                const int e = tid + u * NT;
                const int ee = e < totalA ? e : totalA - 1;
                val[u] = e < totalA;
                off[u] = ee < G * HW ? int64_t(ee / HW) * CW + ee % HW : int64_t(G) * CW + (ee - G * HW);
            }
            tw_gather<NCR>(buf, off, val, tag, got, error, 1);
            #pragma unroll
            for (int u = 0; u < NCR; ++u) {
                const int e = tid + u * NT;
                if (e < totalA) s_words[e] = got[u];
            }
        }
        __syncthreads();
        // (3b) decide; wave 0 fetches the winner's row unless the diagonal stays
        if (wave == 0) {
            T ba = T(-1);
            uint32_t bi = NONE;
            int bg = -1;
            for (int e = lane; e < G; e += 64) {
                const T a = lds_word_val<T>(&s_words[e * HW]);
                const uint32_t x = s_words[e * HW + NWV];
                const bool b = better(a, x, ba, bi);
                ba = b ? a : ba;
                bi = b ? x : bi;
                bg = b ? e : bg;
            }
            #pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const T oa = __shfl_xor(ba, o, 64);
                const uint32_t oi = __shfl_xor(bi, o, 64);
                const int og = __shfl_xor(bg, o, 64);
                const bool b = better(oa, oi, ba, bi);
                ba = b ? oa : ba;
                bi = b ? oi : bi;
                bg = b ? og : bg;
            }
            const int lc = lane < W ? lane : 0;
            const T rj = lds_word_val<T>(&s_words[G * HW + lc * NWV]);
            const T ajj = lds_word_val<T>(&s_words[G * HW]);
            // Tile_getrf.hh:283-289: keep the diagonal if it is large enough.
            const bool diag = bi == NONE || bg < 0 || fabs(ajj) >= threshold * ba;
            int64_t p = diag ? int64_t(j) : int64_t(bi);
            if (p < j || p >= m) p = j;            // never trust a row outside the leaf
            T pr = rj;
            if (p != j) {
                // the winner's row (published with its head)
                int64_t off[1] = {int64_t(bg) * CW + HW + (lane < W * NWV ? lane : 0)};
                bool val[1] = {lane < W * NWV};
                uint32_t got[1];
                tw_gather<1>(buf, off, val, tag, got, error, 1);
                if (lane < W * NWV) s_words[lane] = got[0];   // the heads are consumed
                __builtin_amdgcn_wave_barrier();
                pr = lds_word_val<T>(&s_words[lc * NWV]);
            }
            if (lane < W) {
                s_rowj[lane] = rj;
                s_piv[lane] = pr;
                s_prot[lane] = (lane >= 1 && lane < w - j) ? pr : T(0);
            }
            if (lane == 0) s_p = p;
        }
        __syncthreads();
        const int64_t p = s_p;
        if (tid == 0) s_pivs[j] = int(p);

        // (4) interchange rows j and p among the rows this thread owns
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = gbase + tid + int64_t(NT) * q;
            if (i == j) {
                #pragma unroll
                for (int c = 0; c < W; c += 4) {
                    const vec4_t x = *reinterpret_cast<const vec4_t*>(&s_piv[c]);
                    v[q][c] = x[0]; v[q][c + 1] = x[1]; v[q][c + 2] = x[2]; v[q][c + 3] = x[3];
                }
            }
            else if (i == p) {
                #pragma unroll
                for (int c = 0; c < W; c += 4) {
                    const vec4_t x = *reinterpret_cast<const vec4_t*>(&s_rowj[c]);
                    v[q][c] = x[0]; v[q][c + 1] = x[1]; v[q][c + 2] = x[2]; v[q][c + 3] = x[3];
                }
            }
        }
        const T pv = s_piv[0];
        if (g == 0 && tid == 0) {
            ipiv[c0 + j] = r0 + p;
            if (pv == T(0) && *info == 0) *info = c0 + j + 1;
        }
        // (5) scale the column below the pivot and update columns j+1 .. w-1
        // (rotated 1 .. w-j-1; the others see a zero in s_prot and are kept
        // by the select, so finished columns stay bit-identical)
        const bool tiny = fabs(pv) < safe_min;
        const T inv = T(1) / pv;
        const int lim = w - j;
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = gbase + tid + int64_t(NT) * q;
            if (i > j && i < m) {
                T l = v[q][0];
                if (pv != T(0)) l = tiny ? l / pv : l * inv;
                v[q][0] = l;
                #pragma unroll
                for (int c = 0; c < W; c += 4) {
                    const vec4_t x = *reinterpret_cast<const vec4_t*>(&s_prot[c]);
                    #pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (c + e >= 1) v[q][c + e] = (c + e < lim) ? v[q][c + e] - l * x[e] : v[q][c + e];
                }
            }
        }
        // (6) rotate: column j + 1 becomes v[q][0]
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const T t0 = v[q][0];
            #pragma unroll
            for (int c = 0; c < W - 1; ++c) v[q][c] = v[q][c + 1];
            v[q][W - 1] = t0;
        }
    }

    // after w steps v[q][c] holds column (w + c) mod W
    #pragma unroll
    for (int q = 0; q < RPT; ++q) {
        int64_t i = gbase + tid + int64_t(NT) * q;
        if (i < m) {
            T* rowp = A + (r0 + i) * ld + c0;
            if (VEC && w == W) {
                #pragma unroll
                for (int c = 0; c < W; c += VE) {
                    typedef T vec_t __attribute__((ext_vector_type(VE)));
                    vec_t x;
                    #pragma unroll
                    for (int e = 0; e < VE; ++e) x[e] = v[q][c + e];
                    *reinterpret_cast<vec_t*>(rowp + c) = x;
                }
            }
            else {
                #pragma unroll
                for (int c = 0; c < W; ++c) {
                    const int col = (w + c) & (W - 1);
                    if (col < w) rowp[col] = v[q][c];
                }
            }
        }
    }
    if (xpw > w) leaf_swap_others<T, W, NT>(A, ld, r0, c0, w, xp0, xpw, s_pivs, G, g);
}

// Leaf shape (NT threads, RPT rows each).  Measured per column: FP32 W = 32
// at m = 65536: 256 x 4 7.8 us, 512 x 2 7.1 us, 1024 x 1 7.5 us, 1024 x 2
// 22.7 us; at m = 16384: 256 x 1 (G = 64) 5.8 us, 512 x 1 (G = 32) 5.1 us;
// at m = 8192: 256 x 1 (G = 32) 4.8 us.  Fewer, fatter workgroups lose
// more in their own barriers and updates than they save in gathered heads,
// and a one-round gather of every candidate row (no second read of the
// winner's row) measured no faster at G <= 16 (m = 8192, 512 x 1: 5.1 us).
static void leaf_shape(int64_t m, int& nt, int& rpt)
{
    if (m <= 32 * 256) { nt = 256; rpt = 1; }
    else if (m <= 64 * 512) { nt = 512; rpt = 1; }
    else { nt = 512; rpt = 2; }
}

template <class T, int W>
static void panel_leaf_w(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w,
                         int64_t* ipiv, int64_t* info, double threshold, const PanelWork& pw,
                         uint32_t tag0, hipStream_t s, int64_t xp0, int64_t xpw);

template <class T>
void panel_leaf(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w,
                int64_t* ipiv, int64_t* info, double threshold, const PanelWork& pw,
                uint32_t tag0, hipStream_t s, int64_t xp0, int64_t xpw)
{
    // FP64 leaves of at most 16 columns run the 16-wide register form (its
    // 2-row shape keeps two workgroups per CU on tall panels)
    if constexpr (sizeof(T) == 8) {
        if (w <= 16) {
            panel_leaf_w<T, 16>(A, ld, r0, m_total, c0, w, ipiv, info, threshold, pw, tag0, s, xp0, xpw);
            return;
        }
    }
    panel_leaf_w<T, LeafW<T>::W>(A, ld, r0, m_total, c0, w, ipiv, info, threshold, pw, tag0, s, xp0, xpw);
}

template <class T, int W>
static void panel_leaf_w(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w,
                         int64_t* ipiv, int64_t* info, double threshold, const PanelWork& pw,
                         uint32_t tag0, hipStream_t s, int64_t xp0, int64_t xpw)
{
    const int64_t m = m_total - r0;
    SA_REQUIRE(w >= 1 && w <= W, -2, "panel_leaf: bad width");
    SA_REQUIRE(m >= w, -2, "panel_leaf: fewer rows than columns");
    SA_REQUIRE(m < int64_t(0xFFFFFFFFu), -2, "panel_leaf: panel too tall");
    int nt = 256, rpt = 1;
    leaf_shape(m, nt, rpt);
    int64_t G = ceildiv(m, int64_t(nt) * rpt);
    SA_REQUIRE(G <= LEAF_GMAX, -2, "panel_leaf: panel too tall");
    SA_REQUIRE(xpw == 0 || (r0 == c0 && xp0 <= c0 && c0 + w <= xp0 + xpw), -2,
               "panel_leaf: leaf outside the enclosing panel");
    const int64_t slot_bytes = int64_t(2 * sizeof(uint64_t)) * LeafSlots<T, W>::per(int(G));
    SA_REQUIRE(slot_bytes <= pw.slots_bytes, -2, "panel_leaf: slot area too small");
    T thr = T(threshold);
    constexpr int VE = 16 / int(sizeof(T));
    const bool vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (ld % VE == 0) && (c0 % VE == 0);
    uint64_t* slots = static_cast<uint64_t*>(pw.slots);
#define SA_LEAF(NTV, R, V) hipLaunchKernelGGL((panel_leaf_rot_kernel<T, W, R, V, NTV>), dim3(G), dim3(NTV), 0, s, \
                                               A, ld, r0, m_total, c0, w, ipiv, info, thr, tag0, slots, pw.error, \
                                               xp0, xpw)
#define SA_LEAF_V(NTV, R) do { if (vec) SA_LEAF(NTV, R, 1); else SA_LEAF(NTV, R, 0); } while (0)
    if (nt == 256) SA_LEAF_V(256, 1);
    else if (rpt == 1) SA_LEAF_V(512, 1);
    else SA_LEAF_V(512, 2);
#undef SA_LEAF_V
#undef SA_LEAF
    SA_CHECK_LAUNCH();
}

template void panel_leaf<double>(double*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t*,
                                 double, const PanelWork&, uint32_t, hipStream_t, int64_t, int64_t);
template void panel_leaf<float>(float*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t*,
                                double, const PanelWork&, uint32_t, hipStream_t, int64_t, int64_t);

//------------------------------------------------------------------------------
// Row interchanges.  The sequential LAPACK-style swap list (i, ipiv[i]),
// i = k1..k2-1, with ipiv[i] >= i (src/internal/internal_swap.cc:198-435 and
// the parallel-pivot map of :33-77), is turned into an explicit list of
// (dst row <- src row) moves: every touched row r receives the original row
// src(r) = t_k1(t_k1+1(...t_k2-1(r))), t_i the i-th transposition.  The apply
// kernel then gathers the sources of a column chunk into LDS and scatters
// them, which keeps every row access coalesced in the row-major layout.
static constexpr int PLAN_MAX = 2048;

// Rows are kept relative to k1 as 32-bit ints in LDS, padded to a multiple
// of 16 with identity transpositions, so both per-thread sweeps (the
// duplicate test and the backward walk of src_of) read sixteen entries per
// group of four broadcast LDS loads and do only register compares in between.
__global__ __launch_bounds__(1024)
void laswp_plan_kernel(int64_t k1, int64_t k2, int64_t nrows, const int64_t* __restrict__ ipiv,
                       int64_t* __restrict__ plan)
{
    __shared__ __attribute__((aligned(16))) int piv[PLAN_MAX];
    __shared__ int keep[PLAN_MAX];
    __shared__ int scan[PLAN_MAX];
    typedef int int4_t __attribute__((ext_vector_type(4)));
    const int tid = threadIdx.x, nt = blockDim.x;
    const int np = int(k2 - k1);
    const int np4 = (np + 15) & ~15;                     // padded to whole 16-entry groups
    for (int i = tid; i < np4; i += nt) {
        int v = i;                                          // padding / out-of-range: no swap
        if (i < np) {
            const int64_t p = ipiv[k1 + i];
            if (p >= k1 + i && p < nrows) v = int(p - k1);
        }
        piv[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < np; i += nt) {
        const int p = piv[i];
        int k = p >= np;                                    // a row below the swapped block
        for (int i2 = 0; i2 < (i & ~15); i2 += 16) {
            int4_t q[4];
            #pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = *reinterpret_cast<const int4_t*>(&piv[i2 + 4 * u]);
            #pragma unroll
            for (int u = 0; u < 4; ++u) k &= (q[u][0] != p) & (q[u][1] != p) & (q[u][2] != p) & (q[u][3] != p);
        }
        for (int i2 = i & ~15; i2 < i; ++i2) k &= (piv[i2] != p);
        keep[i] = k;
        scan[i] = k;
    }
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int off = 1; off < np; off <<= 1) {
        int vals[2];
        int cnt = 0;
        for (int i = tid; i < np; i += nt) vals[cnt++] = (i >= off) ? scan[i - off] : 0;
        __syncthreads();
        cnt = 0;
        for (int i = tid; i < np; i += nt) scan[i] += vals[cnt++];
        __syncthreads();
    }
    const int nkeep = np > 0 ? scan[np - 1] : 0;
    const int count = np + nkeep;
    // source of relative row x: t_0(t_1(...t_{np-1}(x))), t_i = (i piv[i])
    auto src_of = [&](int x) {
        for (int i = np4 - 16; i >= 0; i -= 16) {
            int4_t q[4];
            #pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = *reinterpret_cast<const int4_t*>(&piv[i + 4 * u]);
            #pragma unroll
            for (int u = 15; u >= 0; --u) {
                const int a = i + u, b = q[u >> 2][u & 3];
                x = (x == a) ? b : ((x == b) ? a : x);
            }
        }
        return k1 + int64_t(x);
    };
    for (int e = tid; e < np; e += nt) {
        plan[1 + 2 * e] = k1 + e;
        plan[2 + 2 * e] = src_of(e);
    }
    for (int i = tid; i < np; i += nt)
        if (keep[i]) {
            const int e = np + scan[i] - 1;
            plan[1 + 2 * e] = k1 + int64_t(piv[i]);
            plan[2 + 2 * e] = src_of(piv[i]);
        }
    if (tid == 0) plan[0] = count;
}

template <class T>
__global__ __launch_bounds__(256)
void laswp_apply_kernel(T* __restrict__ A, int64_t ld, int64_t c0, int64_t c1, int ch,
                        const int64_t* __restrict__ plan)
{
    extern __shared__ unsigned char smem_raw[];
    T* buf = reinterpret_cast<T*>(smem_raw);
    const int count = int(plan[0]);
    const int64_t cb = c0 + int64_t(blockIdx.x) * ch;
    const int tid = threadIdx.x;
    const int c = tid % ch, e0 = tid / ch, es = 256 / ch;
    const bool ok = cb + c < c1;
    const int64_t cc = ok ? cb + c : c1 - 1;
    // Batches of 8 independent loads (clamped), so row fetches overlap.
    for (int e = e0; e < count; e += 8 * es) {
        T v[8];
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            int ee = e + u * es;
            int64_t src = plan[2 + 2 * (ee < count ? ee : count - 1)];
            v[u] = A[src * ld + cc];
        }
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            int ee = e + u * es;
            if (ee < count) buf[ee * ch + c] = ok ? v[u] : T(0);
        }
    }
    __syncthreads();
    for (int e = e0; e < count; e += es) {
        int64_t dst = plan[1 + 2 * e];
        if (ok) A[dst * ld + cb + c] = buf[e * ch + c];
    }
}

// Small interchange sequences (<= 64 pivots, the ones inside the recursive
// panel): every workgroup derives the (dst <- src) list itself in LDS and
// applies it to its column chunk, so no separate plan launch is needed.
// Entry e < np is row k1+e; entry np+i is pivot row ipiv[k1+i] when it lies
// at or beyond k2 and is its first occurrence, otherwise unused.
template <class T>
__global__ __launch_bounds__(256)
void laswp_small_kernel(T* __restrict__ A, int64_t ld, int64_t nrows, int64_t c0, int64_t c1, int ch,
                        int64_t k1, int np, const int64_t* __restrict__ ipiv)
{
    __shared__ int64_t piv[64];
    __shared__ int64_t dst[128], src[128];
    extern __shared__ unsigned char smem_raw[];
    T* buf = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    const int64_t k2 = k1 + np;
    if (tid < np) {
        int64_t p = ipiv[k1 + tid];
        piv[tid] = (p >= k1 + tid && p < nrows) ? p : k1 + tid;
    }
    __syncthreads();
    if (tid < 2 * np) {
        int64_t r = -1;
        if (tid < np) r = k1 + tid;
        else {
            int i = tid - np;
            int64_t p = piv[i];
            bool first = p >= k2;
            for (int i2 = 0; i2 < i; ++i2) first = first && (piv[i2] != p);
            if (first) r = p;
        }
        int64_t x = r;
        if (r >= 0)
            for (int i = np - 1; i >= 0; --i) {
                int64_t a = k1 + i, b = piv[i];
                if (x == a) x = b;
                else if (x == b) x = a;
            }
        dst[tid] = r;
        src[tid] = x;
    }
    __syncthreads();
    const int64_t cb = c0 + int64_t(blockIdx.x) * ch;
    const int c = tid % ch, e0 = tid / ch, es = 256 / ch;
    const bool ok = cb + c < c1;
    const int64_t cc = ok ? cb + c : c1 - 1;
    for (int e = e0; e < 2 * np; e += 8 * es) {
        T v[8];
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            int ee = e + u * es;
            int64_t sr = (ee < 2 * np && dst[ee] >= 0) ? src[ee] : k1;
            v[u] = A[sr * ld + cc];
        }
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            int ee = e + u * es;
            if (ee < 2 * np && dst[ee] >= 0) buf[ee * ch + c] = ok ? v[u] : T(0);
        }
    }
    __syncthreads();
    for (int e = e0; e < 2 * np; e += es)
        if (dst[e] >= 0 && ok) A[dst[e] * ld + cb + c] = buf[e * ch + c];
}

void laswp_plan(int64_t k1, int64_t k2, int64_t nrows, const int64_t* ipiv, int64_t* plan,
                hipStream_t s)
{
    SA_REQUIRE(k2 - k1 <= PLAN_MAX, -2, "laswp: too many pivots in one call");
    SA_REQUIRE(nrows - k1 < (int64_t(1) << 31), -2, "laswp: rows beyond 32-bit plan range");
    if (k2 <= k1) return;
    hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(1024), 0, s, k1, k2, nrows, ipiv, plan);
    SA_CHECK_LAUNCH();
}

__global__ void laswp_invert_kernel(int64_t* plan, int64_t maxcount)
{
    const int64_t count = plan[0];
    for (int64_t e = threadIdx.x; e < count && e < maxcount; e += blockDim.x) {
        int64_t d = plan[1 + 2 * e];
        plan[1 + 2 * e] = plan[2 + 2 * e];
        plan[2 + 2 * e] = d;
    }
}

void laswp_invert_plan(int64_t* plan, int64_t npiv, hipStream_t s)
{
    if (npiv <= 0) return;
    hipLaunchKernelGGL(laswp_invert_kernel, dim3(1), dim3(1024), 0, s, plan, 2 * npiv);
    SA_CHECK_LAUNCH();
}

// Two-pass interchange: pass 0 copies source row plan[2+2e] (columns
// [c0, c1)) into tmp row e, pass 1 copies tmp row e into destination row
// plan[1+2e].  grid.y strides over entries, grid.x over 16-byte column groups.
template <class T, int PASS, int VEC>
__global__ __launch_bounds__(256)
void laswp_pass_kernel(T* __restrict__ A, int64_t ld, int64_t c0, int64_t ncol, T* __restrict__ tmp,
                       const int64_t* __restrict__ plan)
{
    constexpr int VE = VEC ? 16 / int(sizeof(T)) : 1;
    typedef T vec_t __attribute__((ext_vector_type(VE)));
    const int count = int(plan[0]);
    const int64_t col = (int64_t(blockIdx.x) * 256 + threadIdx.x) * VE;
    if (col >= ncol) return;
    for (int e = blockIdx.y; e < count; e += gridDim.y) {
        const int64_t row = plan[(PASS == 0 ? 2 : 1) + 2 * e];
        T* a = A + row * ld + c0 + col;
        T* t = tmp + int64_t(e) * ncol + col;
        if (VEC && col + VE <= ncol) {
            // each source element is read once: nontemporal loads
            if (PASS == 0) *reinterpret_cast<vec_t*>(t) = __builtin_nontemporal_load(reinterpret_cast<const vec_t*>(a));
            else *reinterpret_cast<vec_t*>(a) = __builtin_nontemporal_load(reinterpret_cast<const vec_t*>(t));
        }
        else {
            for (int q = 0; q < VE && col + q < ncol; ++q) {
                if (PASS == 0) t[q] = a[q];
                else a[q] = t[q];
            }
        }
    }
}

template <class T>
void laswp(T* A, int64_t ld, int64_t nrows, int64_t c0, int64_t c1, int64_t k1, int64_t k2,
           const int64_t* ipiv, int64_t* plan, hipStream_t s, bool plan_ready, T* tmp, int64_t tmp_elems)
{
    if (k2 <= k1) return;
    if (plan_ready && tmp && c1 > c0 && tmp_elems >= 2 * (k2 - k1) * (c1 - c0)) {
        constexpr int VE = 16 / int(sizeof(T));
        const int64_t ncol = c1 - c0;
        const bool vec = ld % VE == 0 && c0 % VE == 0 && ncol % VE == 0
                         && reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(tmp) % 16 == 0;
        const int64_t per = vec ? VE : 1;
        dim3 grid(unsigned(ceildiv(ncol, 256 * per)), unsigned(std::min<int64_t>(2 * (k2 - k1), 2048)));
        if (vec) {
            hipLaunchKernelGGL((laswp_pass_kernel<T, 0, 1>), grid, dim3(256), 0, s, A, ld, c0, ncol, tmp, plan);
            hipLaunchKernelGGL((laswp_pass_kernel<T, 1, 1>), grid, dim3(256), 0, s, A, ld, c0, ncol, tmp, plan);
        }
        else {
            hipLaunchKernelGGL((laswp_pass_kernel<T, 0, 0>), grid, dim3(256), 0, s, A, ld, c0, ncol, tmp, plan);
            hipLaunchKernelGGL((laswp_pass_kernel<T, 1, 0>), grid, dim3(256), 0, s, A, ld, c0, ncol, tmp, plan);
        }
        SA_CHECK_LAUNCH();
        return;
    }
    if (!plan_ready && k2 - k1 <= 64) {
        if (c1 <= c0) return;
        const int np = int(k2 - k1);
        int ch = 256;
        while (ch > 4 && int64_t(ch) * 2 * np * int64_t(sizeof(T)) > 65536) ch /= 2;
        while (ch > 4 && ch / 2 >= (c1 - c0)) ch /= 2;
        size_t lds = size_t(ch) * size_t(2 * np) * sizeof(T);
        int64_t nblk = ceildiv(c1 - c0, ch);
        hipLaunchKernelGGL(laswp_small_kernel<T>, dim3(unsigned(nblk)), dim3(256), lds, s,
                           A, ld, nrows, c0, c1, ch, k1, np, ipiv);
        SA_CHECK_LAUNCH();
        return;
    }
    if (!plan_ready) laswp_plan(k1, k2, nrows, ipiv, plan, s);
    if (c1 <= c0) return;
    const int64_t maxcount = 2 * (k2 - k1);
    // column chunk: power of two, LDS staging <= 64 KiB, 4..256 columns
    int ch = 256;
    while (ch > 4 && int64_t(ch) * maxcount * int64_t(sizeof(T)) > 65536) ch /= 2;
    while (ch > 4 && ch / 2 >= (c1 - c0)) ch /= 2;
    size_t lds = size_t(ch) * size_t(maxcount) * sizeof(T);
    SA_REQUIRE(lds <= 65536, -2, "laswp: LDS staging too large");
    int64_t nblk = ceildiv(c1 - c0, ch);
    hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3(unsigned(nblk)), dim3(256), lds, s,
                       A, ld, c0, c1, ch, plan);
    SA_CHECK_LAUNCH();
}

template void laswp<double>(double*, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t,
                            const int64_t*, int64_t*, hipStream_t, bool, double*, int64_t);
template void laswp<float>(float*, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t,
                           const int64_t*, int64_t*, hipStream_t, bool, float*, int64_t);

} // namespace sa
