// Tournament-pivoting panel leaf (MethodLU::CALU, src/getrf_tntpiv.cc) for
// gfx950.
//
// SLATE's CALU panel (internal::getrf_tntpiv_panel) picks the nb pivot rows of
// a tall panel by a reduction tree: each tile is factored with partial
// pivoting to nominate candidate rows, candidates are merged pairwise by
// partial pivoting on the stacked originals, and the winners are the pivots
// of the whole panel; the rest of the panel is then L = A U^{-1}.  Here the
// tree is over the G workgroups of the leaf (each holds ~1024 rows in
// registers): a local partial-pivot elimination nominates W candidates per
// workgroup, then ceil(log2 G) butterfly rounds merge the candidate sets
// through tagged words (each workgroup reads the set of the partner subtree
// and both sides compute the same merge), so a leaf needs ceil(log2 G)
// cross-workgroup exchanges instead of one per column.  Every workgroup ends
// with the same W winners, factors them (the top W x W block: L11, U11),
// and writes its rows of L21 = A21 U11^{-1} in the interchanged order.
//
// The pivot sequence differs from partial pivoting (and from SLATE's tree,
// whose leaves are nb-row tiles); the factorization is P A = L U with the
// growth of tournament pivoting.
#include "common.hh"
#include "kernels.hh"
#include <limits>

namespace sa {

static constexpr int CALU_THREADS = 256;
static constexpr int CALU_GMAX = 128;
static constexpr int CALU_RPT_MAX = 4;
static constexpr int CALU_MAX_ROUNDS = 7;    // log2(CALU_GMAX)

template <class T> struct CaluW;
template <> struct CaluW<double> { static constexpr int W = 16; };
template <> struct CaluW<float> { static constexpr int W = 32; };

// Tagged-word layout: [top rows: W x W values] then, per round r, per
// workgroup g, a candidate set [W row indices (1 word each) | W x W values].
template <class T, int W>
struct CaluSlots {
    static constexpr int NWV = int(sizeof(T)) / 4;
    static constexpr int64_t TOP = int64_t(W) * W * NWV;
    static constexpr int64_t SET = W + int64_t(W) * W * NWV;
    __host__ __device__ static int64_t round_base(int r, int G) { return TOP + int64_t(r) * G * SET; }
    __host__ __device__ static int64_t words(int rounds, int G) { return TOP + int64_t(rounds) * G * SET; }
};

template <class T>
__device__ inline T raw_val(const uint32_t* w)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, (uint64_t(w[1]) << 32) | w[0]);
    else return __builtin_bit_cast(T, w[0]);
}

template <class T>
__device__ inline bool calu_better(T a, int ia, T b, int ib)
{
    return a > b || (a == b && ia < ib);
}

// Partial-pivot elimination of `rows` stacked rows (values in work[rows][W],
// row ids in id[]) for w columns, run by wave 0 (one row per lane, rows <= 64).
// Returns the chosen ids in pivot order in out_id[0..w) (-1 when fewer valid
// rows), and leaves the factors of the chosen rows (L strictly below the
// diagonal, U on and above, in pivot order) in lu[w][W].
template <class T, int W>
__device__ void calu_gepp(T (*work)[W], const int* id, int rows, int w, int* out_id, T (*lu)[W])
{
    const int lane = threadIdx.x & 63;
    const bool have = lane < rows && id[lane] >= 0;
    T r[W];
    #pragma unroll
    for (int c = 0; c < W; ++c) r[c] = have ? work[lane][c] : T(0);
    bool done = !have;
    const T safe_min = std::numeric_limits<T>::min();
    for (int j = 0; j < w; ++j) {
        T x = T(0);
        #pragma unroll
        for (int c = 0; c < W; ++c) if (c == j) x = r[c];
        T a = done ? T(-1) : fabs(x);
        int who = done ? 0x7fffffff : lane;
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            T oa = __shfl_xor(a, o, 64);
            int ow = __shfl_xor(who, o, 64);
            if (calu_better(oa, ow, a, who)) { a = oa; who = ow; }
        }
        if (who == 0x7fffffff) {                     // no rows left
            if (lane == 0) out_id[j] = -1;
            if (lane < W) lu[j][lane] = T(0);
            continue;
        }
        // broadcast the pivot row from lane `who`
        T pr[W];
        #pragma unroll
        for (int c = 0; c < W; ++c) pr[c] = __shfl(r[c], who, 64);
        if (lane == who) {
            done = true;
            out_id[j] = id[who];
            #pragma unroll
            for (int c = 0; c < W; ++c) lu[j][c] = r[c];     // U row (L part set below)
        }
        const T pv = pr[j];
        const bool tiny = fabs(pv) < safe_min;
        const T inv = T(1) / pv;
        if (!done) {
            T l = T(0);
            #pragma unroll
            for (int c = 0; c < W; ++c) if (c == j) l = r[c];
            if (pv != T(0)) l = tiny ? l / pv : l * inv;
            else l = T(0);
            #pragma unroll
            for (int c = 0; c < W; ++c) {
                if (c == j) r[c] = l;
                else if (c > j) r[c] -= l * pr[c];
            }
        }
    }
    // lu[j] holds the pivot row as it was when chosen: its entries left of j
    // are its L multipliers (set when earlier pivots eliminated it), the rest U.
}

template <class T, int W, int RPT>
__global__ __launch_bounds__(CALU_THREADS)
void calu_leaf_kernel(T* __restrict__ A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w,
                      int64_t* __restrict__ ipiv, int64_t* __restrict__ info, uint32_t tag0,
                      uint64_t* __restrict__ slots, int* error)
{
    using SL = CaluSlots<T, W>;
    constexpr int NWV = SL::NWV;
    const int G = gridDim.x, g = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m = m_total - r0;
    constexpr int64_t R = int64_t(CALU_THREADS) * RPT;
    const int64_t gbase = int64_t(g) * R;
    int rounds = 0;
    while ((1 << rounds) < G) ++rounds;

    __shared__ T s_wabs[4];
    __shared__ int s_widx[4];
    __shared__ T s_piv[W];
    __shared__ int s_cand[W];
    __shared__ int s_mid[2 * W];
    __shared__ T s_m[2 * W][W];          // stacked candidate originals
    __shared__ T s_set[W][W];            // own candidate set (originals)
    __shared__ int s_sid[W];
    __shared__ T s_lu[W][W];             // factors of the final winners
    __shared__ int s_win[W];
    __shared__ int s_dn_pos[W], s_dn_row[W];
    __shared__ int s_ndn;
    __shared__ T s_top[W][W];
    __shared__ uint32_t s_raw[SL::SET];   // payloads of a partner's set
    auto rowp = [&](int64_t i) { return A + (r0 + i) * ld + c0; };
    T v[RPT][W];
    #pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t i = gbase + tid + int64_t(CALU_THREADS) * q;
        const T* p = rowp(i < m ? i : 0);
        #pragma unroll
        for (int c = 0; c < W; ++c) v[q][c] = (i < m && c < w) ? p[c] : T(0);
    }
    // (0) workgroup 0 publishes the original top w rows (needed by the owners
    // of the positions they will be swapped to)
    if (g == 0 && tid < w) {
        #pragma unroll
        for (int c = 0; c < W; ++c) tw_store_val<T>(slots + (int64_t(tid) * W + c) * NWV, v[0][c], tag0);
    }

    // (1) local partial-pivot elimination nominates w candidate rows
    unsigned chosen = 0;
    for (int j = 0; j < w; ++j) {
        T babs = T(-1);
        int bidx = 0x7fffffff;
        #pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = gbase + tid + int64_t(CALU_THREADS) * q;
            T x = T(0);
            #pragma unroll
            for (int c = 0; c < W; ++c) if (c == j) x = v[q][c];
            const T a = fabs(x);
            if (i < m && !((chosen >> q) & 1u) && calu_better(a, int(i), babs, bidx)) { babs = a; bidx = int(i); }
        }
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            T oa = __shfl_xor(babs, o, 64);
            int oi = __shfl_xor(bidx, o, 64);
            if (calu_better(oa, oi, babs, bidx)) { babs = oa; bidx = oi; }
        }
        if (lane == 0) { s_wabs[wave] = babs; s_widx[wave] = bidx; }
        __syncthreads();
        babs = s_wabs[0]; bidx = s_widx[0];
        #pragma unroll
        for (int k = 1; k < 4; ++k)
            if (calu_better(s_wabs[k], s_widx[k], babs, bidx)) { babs = s_wabs[k]; bidx = s_widx[k]; }
        const bool any = bidx != 0x7fffffff;
        if (any && int64_t(bidx) >= gbase && (int64_t(bidx) - gbase) % CALU_THREADS == tid) {
            const int qq = int((int64_t(bidx) - gbase) / CALU_THREADS);
            #pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (q == qq) {
                    chosen |= 1u << q;
                    #pragma unroll
                    for (int c = 0; c < W; ++c) s_piv[c] = v[q][c];
                }
        }
        if (tid == 0) s_cand[j] = any ? bidx : -1;
        __syncthreads();
        if (any) {
            const T pv = s_piv[j];
            #pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int64_t i = gbase + tid + int64_t(CALU_THREADS) * q;
                if (i < m && !((chosen >> q) & 1u) && pv != T(0)) {
                    T x = T(0);
                    #pragma unroll
                    for (int c = 0; c < W; ++c) if (c == j) x = v[q][c];
                    const T l = x / pv;
                    #pragma unroll
                    for (int c = 0; c < W; ++c) if (c > j) v[q][c] -= l * s_piv[c];
                }
            }
        }
        __syncthreads();
    }
    // originals of the candidates (rows of this workgroup, unchanged in memory)
    for (int e = tid; e < W * W; e += CALU_THREADS) {
        const int j = e / W, c = e % W;
        const int id = j < w ? s_cand[j] : -1;
        s_set[j][c] = (id >= 0 && c < w) ? rowp(id)[c] : T(0);
        if (c == 0) s_sid[j] = id;
    }
    __syncthreads();

    // (2) butterfly tournament over the workgroups
    for (int r = 0; r < rounds; ++r) {
        const uint32_t tag = tag0 + 1 + uint32_t(r);
        uint64_t* mine = slots + SL::round_base(r, G) + int64_t(g) * SL::SET;
        for (int e = tid; e < W + W * W; e += CALU_THREADS) {
            if (e < W) tw_store(mine + e, uint32_t(s_sid[e]), tag);
            else tw_store_val<T>(mine + W + int64_t(e - W) * NWV, s_set[(e - W) / W][(e - W) % W], tag);
        }
        const int bit = 1 << r;
        const int partner = (g ^ bit) & ~(bit - 1);
        if (partner < G) {                            // uniform over the workgroup
            const uint64_t* theirs = slots + SL::round_base(r, G) + int64_t(partner) * SL::SET;
            const bool lo_first = partner < g;        // lower subtree's rows first
            {
                // all words of the partner's set in one batched wait
                constexpr int NSW = int((SL::SET + CALU_THREADS - 1) / CALU_THREADS);
                int64_t off[NSW];
                bool val[NSW];
                uint32_t got[NSW];
                #pragma unroll
                for (int u = 0; u < NSW; ++u) {
                    const int64_t e = tid + int64_t(u) * CALU_THREADS;
                    val[u] = e < SL::SET;
                    off[u] = e < SL::SET ? e : SL::SET - 1;
                }
                tw_gather<NSW>(theirs, off, val, tag, got, error, 1);
                #pragma unroll
                for (int u = 0; u < NSW; ++u) {
                    const int64_t e = tid + int64_t(u) * CALU_THREADS;
                    if (e < SL::SET) s_raw[e] = got[u];
                }
            }
            __syncthreads();
            for (int e = tid; e < W + W * W; e += CALU_THREADS) {
                if (e < W) {
                    s_mid[lo_first ? e : W + e] = int(s_raw[e]);
                    s_mid[lo_first ? W + e : e] = s_sid[e];
                }
                else {
                    const int j = (e - W) / W, c = (e - W) % W;
                    const T x = raw_val<T>(&s_raw[W + (e - W) * NWV]);
                    s_m[lo_first ? j : W + j][c] = x;
                    s_m[lo_first ? W + j : j][c] = s_set[j][c];
                }
            }
            __syncthreads();
            if (wave == 0) {
                // GEPP on a working copy of the stacked originals
                __shared__ T s_work[2 * W][W];
                __shared__ int s_out[W];
                __shared__ T s_tmp_lu[W][W];
                for (int e = lane; e < 2 * W * W; e += 64) s_work[e / W][e % W] = s_m[e / W][e % W];
                __builtin_amdgcn_wave_barrier();
                calu_gepp<T, W>(s_work, s_mid, 2 * W, w, s_out, s_tmp_lu);
                // new set = winners with their original values
                for (int e = lane; e < W * W; e += 64) {
                    const int j = e / W, c = e % W;
                    const int id = j < w ? s_out[j] : -1;
                    int src = -1;
                    for (int k = 0; k < 2 * W; ++k) if (id >= 0 && s_mid[k] == id) src = k;
                    s_set[j][c] = src >= 0 ? s_m[src][c] : T(0);
                    if (c == 0) s_sid[j] = id;
                }
            }
            __syncthreads();
        }
        else {
            __syncthreads();
        }
    }

    // (3) factor the winners (their pivot order is the final one)
    if (wave == 0) {
        __shared__ T s_work2[W][W];
        for (int e = lane; e < W * W; e += 64) s_work2[e / W][e % W] = s_set[e / W][e % W];
        __builtin_amdgcn_wave_barrier();
        calu_gepp<T, W>(s_work2, s_sid, W, w, s_win, s_lu);
    }
    __syncthreads();
    // interchange sequence: position j <- winner j; track the top rows that
    // move down (only positions < w ever move)
    if (tid == 0) {
        int at[W];                                    // original row now at position j < w
        for (int j = 0; j < W; ++j) at[j] = j;
        int ndn = 0;
        for (int j = 0; j < w; ++j) {
            const int wr = s_win[j] >= 0 ? s_win[j] : j;
            int pos = -1;
            for (int k = j; k < w; ++k) if (at[k] == wr) pos = k;          // still in the top block
            if (pos < 0) {
                for (int k = 0; k < ndn; ++k) if (s_dn_row[k] == wr) pos = s_dn_pos[k];
                if (pos < 0) pos = wr;                                    // never moved
            }
            if (g == 0) ipiv[c0 + j] = r0 + pos;
            const int disp = at[j];                   // row displaced from position j
            if (pos < w) {
                at[pos] = disp;
            }
            else {
                int k = 0;
                for (; k < ndn; ++k) if (s_dn_pos[k] == pos) break;
                s_dn_pos[k] = pos;
                s_dn_row[k] = disp;
                if (k == ndn) ++ndn;
            }
            at[j] = wr;
        }
        s_ndn = ndn;
        if (g == 0) {
            for (int j = 0; j < w; ++j)
                if (s_lu[j][j] == T(0)) { if (*info == 0) *info = c0 + j + 1; break; }
        }
    }
    __syncthreads();
    // original top rows for the displaced positions this workgroup owns, in
    // one batched wait
    {
        constexpr int TW = W * NWV;                   // words per top row
        constexpr int NTW = (W * TW + CALU_THREADS - 1) / CALU_THREADS;
        int64_t off[NTW];
        bool val[NTW];
        uint32_t got[NTW];
        const int ndn = s_ndn;
        #pragma unroll
        for (int u = 0; u < NTW; ++u) {
            const int e = tid + u * CALU_THREADS;
            const int k = e / TW;
            const bool mine = k < ndn && int64_t(s_dn_pos[k < W ? k : 0]) >= gbase
                              && int64_t(s_dn_pos[k < W ? k : 0]) < gbase + R;
            val[u] = e < W * TW && mine;
            off[u] = val[u] ? int64_t(s_dn_row[k]) * TW + e % TW : 0;
        }
        tw_gather<NTW>(slots, off, val, tag0, got, error, 1);
        #pragma unroll
        for (int u = 0; u < NTW; ++u) {
            const int e = tid + u * CALU_THREADS;
            if (val[u]) reinterpret_cast<uint32_t*>(&s_top[0][0])[(e / TW) * TW + e % TW] = got[u];
        }
    }
    __syncthreads();

    // (4) new rows: positions < w get [L11 | U11] of the winners, the others
    // l = a U11^{-1} from their (interchanged) original row
    const T safe_min = std::numeric_limits<T>::min();
    #pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t i = gbase + tid + int64_t(CALU_THREADS) * q;
        if (i >= m) continue;
        T row[W];
        if (i < w) {
            #pragma unroll
            for (int c = 0; c < W; ++c) row[c] = s_lu[i][c];
        }
        else {
            int k = -1;
            for (int t = 0; t < s_ndn; ++t) if (s_dn_pos[t] == i) k = t;
            if (k >= 0) {
                #pragma unroll
                for (int c = 0; c < W; ++c) row[c] = s_top[k][c];
            }
            else {
                const T* p = rowp(i);
                #pragma unroll
                for (int c = 0; c < W; ++c) row[c] = c < w ? p[c] : T(0);
            }
            #pragma unroll
            for (int c = 0; c < W; ++c) {
                if (c < w) {
                    T s = row[c];
                    #pragma unroll
                    for (int kk = 0; kk < W; ++kk) if (kk < c) s -= row[kk] * s_lu[kk][c];
                    const T u = s_lu[c][c];
                    row[c] = (u == T(0)) ? s : (fabs(u) < safe_min ? s / u : s * (T(1) / u));
                }
            }
        }
        T* p = rowp(i);
        #pragma unroll
        for (int c = 0; c < W; ++c) if (c < w) p[c] = row[c];
    }
}

int calu_leaf_width(int elem_bytes) { return elem_bytes == 8 ? CaluW<double>::W : CaluW<float>::W; }

int64_t calu_slot_bytes(int elem_bytes, int64_t m)
{
    int rpt = 1;
    while (rpt < CALU_RPT_MAX && ceildiv(m, int64_t(CALU_THREADS) * rpt) > 32) rpt *= 2;
    const int G = int(ceildiv(m, int64_t(CALU_THREADS) * rpt));
    int rounds = 0;
    while ((1 << rounds) < G) ++rounds;
    const int64_t words = elem_bytes == 8 ? CaluSlots<double, 16>::words(rounds, G)
                                          : CaluSlots<float, 32>::words(rounds, G);
    return words * int64_t(sizeof(uint64_t));
}

template <class T>
void calu_leaf(T* A, int64_t ld, int64_t r0, int64_t m_total, int64_t c0, int w, int64_t* ipiv, int64_t* info,
               const PanelWork& pw, uint32_t tag0, hipStream_t s)
{
    constexpr int W = CaluW<T>::W;
    const int64_t m = m_total - r0;
    SA_REQUIRE(w >= 1 && w <= W, -2, "calu_leaf: bad width");
    SA_REQUIRE(m >= w, -2, "calu_leaf: fewer rows than columns");
    SA_REQUIRE(m < int64_t(1) << 30, -2, "calu_leaf: panel too tall");
    int rpt = 1;
    while (rpt < CALU_RPT_MAX && ceildiv(m, int64_t(CALU_THREADS) * rpt) > 32) rpt *= 2;
    const int64_t G = ceildiv(m, int64_t(CALU_THREADS) * rpt);
    SA_REQUIRE(G <= CALU_GMAX, -2, "calu_leaf: panel too tall");
    SA_REQUIRE(calu_slot_bytes(int(sizeof(T)), m) <= pw.slots_bytes, -2, "calu_leaf: slot area too small");
    uint64_t* slots = static_cast<uint64_t*>(pw.slots);
#define SA_CALU(R) hipLaunchKernelGGL((calu_leaf_kernel<T, W, R>), dim3(unsigned(G)), dim3(CALU_THREADS), 0, s, \
                                      A, ld, r0, m_total, c0, w, ipiv, info, tag0, slots, pw.error)
    switch (rpt) { case 1: SA_CALU(1); break; case 2: SA_CALU(2); break; default: SA_CALU(4); break; }
#undef SA_CALU
    SA_CHECK_LAUNCH();
}

template void calu_leaf<double>(double*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t*,
                                const PanelWork&, uint32_t, hipStream_t);
template void calu_leaf<float>(float*, int64_t, int64_t, int64_t, int64_t, int, int64_t*, int64_t*,
                               const PanelWork&, uint32_t, hipStream_t);

} // namespace sa
