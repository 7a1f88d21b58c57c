// Batched factor-and-solve of many small systems of uneven size (gfx950).
//
// One persistent launch: each workgroup takes the next system from a work
// queue (an atomic ticket over a largest-first order), copies it into its
// own workspace in row-major order (interchanges then move contiguous rows),
// factors it with partial pivoting, solves, and refines in FP64 against the
// original matrix to SLATE's stopping rule (src/gesv_mixed.cc:190-259,
// internal_util.hh:121-138), with the FP64 factorization as the fallback when
// the refinement misses the tolerance.  The per-system steps mirror
// slate::gesv / gesv_mixed on a single tile column; the batching (one launch
// for thousands of systems of different n, no per-system launches or host
// round trips) is this library's own design.
//
// Per-system modes (BatchArgs::mode):
//   0 FP32 factor + FP64 IR, 1 FP64 factor, 2 BF16 factor (FP32 storage,
//   bf16-operand MFMA trailing updates with FP32 accumulation) + FP64 IR,
//   3 learned choice, 4/5/6 the learned path with its choice forced to
//   0/1/2 (labelling).
// The learned path factors in BF16 or FP32 first (the probe), solves once and
// forms the FP64 residual (iteration 0 of the refinement), and feeds the
// probe's normwise quantities to the selector network (the MLP of mlp.hip,
// run by one wave), which picks what runs: BF16 + IR (0), FP32 + IR (1) or
// FP64 (2) — keeping the probe factor when it has the chosen precision.
//
// Factorization: blocked right-looking LU with NB-column panels held in LDS
// (up to 1024 rows, odd row stride so column reads are conflict free).  A
// panel column costs two barriers: the pivot candidates of column j + 1 are
// found while column j's rank-1 update is applied, and each thread reads the
// pivot row and row j into registers before either is rewritten.  The
// panel's interchanges are composed by one wave into at most 2 NB row moves
// (destination, source), which every column left and right of the panel and
// the permutation vector apply with all loads before any store (one memory
// round trip per panel instead of one per swap); U12 = L11^{-1} A12 runs one
// column per thread in registers, and A22 -= L21 U12 as 32 x 32 wave tiles of MFMA
// (v_mfma_f32_16x16x4f32, v_mfma_f32_16x16x32_bf16 or v_mfma_f64_16x16x4f64)
// with L21 read from the LDS panel.
// Triangular solves: 64-row blocks, each diagonal block solved by one wave
// (no barriers), the rows below updated by the whole workgroup.
//
// The ticket counter and a count of finished workgroups live in `queue`; the
// last workgroup to finish zeroes both, so consecutive launches (and replays
// of a graph holding the launch) need no separate reset.
#include "common.hh"
#include "kernels.hh"

#include <cstring>
#include <limits>

namespace sa {

static constexpr int BMAXN = 1024;          // largest system
static constexpr int SEL_F = 24, SEL_H = 64, SEL_K = 8;   // selector network (mlp.hip layout)
static constexpr int SEL_CHOICES = 3;

// Panel of storage type T for a float panel width NBF: the FP64 panel is half
// as wide, so both fit the same LDS allocation.
template <class T, int NBF> struct Pan {
    static constexpr int NB = sizeof(T) == 8 ? NBF / 2 : NBF;
    static constexpr int LD = NB + 1;
};
template <int NBF> constexpr int pan_lds_bytes()
{
    return BMAXN * Pan<float, NBF>::LD * 4 > BMAXN * Pan<double, NBF>::LD * 8
         ? BMAXN * Pan<float, NBF>::LD * 4 : BMAXN * Pan<double, NBF>::LD * 8;
}

__host__ __device__ inline int b_ldw(int n) { return (n + 3) & ~3; }

template <class T>
__device__ inline bool bt_better(T a, int ia, T b, int ib) { return a > b || (a == b && ia < ib); }

// W (row-major, ldw) = A (column-major n x n, FP64), cast to T.  Wave tiles of
// 64 rows x 32 columns: the column reads are coalesced, each lane then writes
// 32 consecutive elements of its row.
template <class T, int BT>
__device__ void wg_load_rowmajor(const double* __restrict__ A, int n, T* __restrict__ W, int ldw)
{
    constexpr int NW = BT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = (n + 63) >> 6, tj = (n + 31) >> 5;
    for (int t = wave; t < ti * tj; t += NW) {
        const int i = (t % ti) * 64 + lane, j0 = (t / ti) * 32;
        T v[32];
        #pragma unroll
        for (int c = 0; c < 32; ++c)
            v[c] = (i < n && j0 + c < n) ? T(A[i + int64_t(j0 + c) * n]) : T(0);
        if (i < n) {
            T* w = W + int64_t(i) * ldw + j0;
            if (j0 + 32 <= n) {
                #pragma unroll
                for (int c = 0; c < 32; ++c) w[c] = v[c];
            }
            else {
                #pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (j0 + c < n) w[c] = v[c];
            }
        }
    }
}

// Factor the mr x kb panel in LDS (row stride LD) with partial pivoting
// (Tile_getrf.hh:197-366: largest magnitude, ties to the lower index, scaling
// by the reciprocal unless the pivot is below the safe minimum).  s_piv[j]
// receives the chosen local row of column j; *s_info the first exactly-zero
// pivot (1-based global column k0 + j + 1) if none was recorded yet.
template <class T, int NB, int LD, int BT>
__device__ void wg_panel(T* __restrict__ s_p, int mr, int kb, int k0, int* s_piv, int* s_info, T* s_red,
                         int* s_ired)
{
    constexpr int NW = BT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T safe_min = std::numeric_limits<T>::min();
    T ba = T(-1);
    int bi = 0x7fffffff;
    for (int i = tid; i < mr; i += BT) {
        const T a = fabs(s_p[i * LD]);
        if (bt_better(a, i, ba, bi)) { ba = a; bi = i; }
    }
    for (int j = 0; j < kb; ++j) {
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const T oa = __shfl_xor(ba, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (bt_better(oa, oi, ba, bi)) { ba = oa; bi = oi; }
        }
        if (lane == 0) { s_red[wave] = ba; s_ired[wave] = bi; }
        __syncthreads();
        ba = s_red[0];
        bi = s_ired[0];
        #pragma unroll
        for (int w = 1; w < NW; ++w)
            if (bt_better(s_red[w], s_ired[w], ba, bi)) { ba = s_red[w]; bi = s_ired[w]; }
        const int p = (bi == 0x7fffffff) ? j : bi;
        T prow[NB], rj[NB];
        #pragma unroll
        for (int c = 0; c < NB; ++c) {
            prow[c] = s_p[p * LD + c];
            rj[c] = s_p[j * LD + c];
        }
        __syncthreads();                                   // rows p and j read by every thread
        T pv = T(0);
        #pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c == j) pv = prow[c];
        if (tid == 0) {
            s_piv[j] = p;
            if (pv == T(0) && *s_info == 0) *s_info = k0 + j + 1;
            if (p != j) {
                #pragma unroll
                for (int c = 0; c < NB; ++c) s_p[j * LD + c] = prow[c];
            }
        }
        const bool tiny = fabs(pv) < safe_min;
        const T inv = T(1) / pv;
        ba = T(-1);
        bi = 0x7fffffff;
        for (int i = j + 1 + tid; i < mr; i += BT) {
            const bool isp = (i == p);
            T* row = s_p + i * LD;
            // the row's values (old row j's for the pivot row's position),
            // selected as values so no load goes through a mixed pointer
            T cur[NB];
            #pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c >= j) {
                    const T rv = row[c];
                    cur[c] = isp ? rj[c] : rv;
                }
                else {
                    cur[c] = rj[c];                 // read only for the pivot row's position
                }
            }
            T l = T(0);
            #pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c == j) l = cur[c];
            if (pv != T(0)) l = tiny ? l / pv : l * inv;
            T nxt = T(0);
            #pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c < j) {
                    if (isp) row[c] = cur[c];               // L part of old row j moves to row p
                }
                else if (c == j) {
                    row[c] = l;
                }
                else if (c < kb) {
                    const T v = cur[c] - l * prow[c];
                    row[c] = v;
                    if (c == j + 1) nxt = v;
                }
            }
            if (j + 1 < kb) {
                const T a = fabs(nxt);
                if (bt_better(a, i, ba, bi)) { ba = a; bi = i; }
            }
        }
    }
    __syncthreads();
}

// C (nr x nr, row stride ldw) -= L (nr x NB, LDS, row stride LD) * U (NB x nr,
// row stride ldw), 32 x 32 wave tiles.  A tile inside the matrix reads its 16
// C values per lane before the products, unguarded, so they are in flight
// together with the U operand (one memory round trip per tile; guarded
// element-by-element updates after the products wait once per element).
// Row of accumulator register r: 4 r + l/16 (FP64), 4 (l/16) + r (FP32).
template <class T, int NB, int LD, int BT, bool BFU>
__device__ void wg_update(T* __restrict__ C, const T* __restrict__ U, int ldw, const T* __restrict__ L, int nr)
{
    constexpr int NW = BT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int fr = lane & 15, fg = lane >> 4;
    const int tt = (nr + 31) >> 5;
    auto crow = [&](int i0, int a, int r) {
        return sizeof(T) == 8 ? i0 + 16 * a + 4 * r + fg : i0 + 16 * a + 4 * fg + r;
    };
    for (int t = wave; t < tt * tt; t += NW) {
        const int i0 = (t / tt) << 5, j0 = (t % tt) << 5;
        const bool full = i0 + 32 <= nr && j0 + 32 <= nr;
        T cv[2][2][4];
        if (full) {
            #pragma unroll
            for (int a = 0; a < 2; ++a)
                #pragma unroll
                for (int b = 0; b < 2; ++b)
                    #pragma unroll
                    for (int r = 0; r < 4; ++r)
                        cv[a][b][r] = C[int64_t(crow(i0, a, r)) * ldw + j0 + 16 * b + fr];
        }
        if constexpr (sizeof(T) == 8) {
            v4d acc[2][2];
            #pragma unroll
            for (int a = 0; a < 2; ++a)
                #pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
            #pragma unroll
            for (int s = 0; s < NB / 4; ++s) {
                const int k = 4 * s + fg;
                double av[2], bv[2];
                #pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int i = i0 + 16 * a + fr;
                    av[a] = i < nr ? L[i * LD + k] : 0.0;
                }
                #pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int jj = j0 + 16 * b + fr;
                    bv[b] = jj < nr ? U[int64_t(k) * ldw + jj] : 0.0;
                }
                #pragma unroll
                for (int a = 0; a < 2; ++a)
                    #pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
            }
            // register r holds D[4 r + l/16][l % 16]
            #pragma unroll
            for (int a = 0; a < 2; ++a)
                #pragma unroll
                for (int b = 0; b < 2; ++b)
                    #pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = crow(i0, a, r), jj = j0 + 16 * b + fr;
                        if (full) C[int64_t(i) * ldw + jj] = cv[a][b][r] - acc[a][b][r];
                        else if (i < nr && jj < nr) C[int64_t(i) * ldw + jj] -= acc[a][b][r];
                    }
        }
        else {
            v4f acc[2][2];
            #pragma unroll
            for (int a = 0; a < 2; ++a)
                #pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
            if constexpr (BFU) {
                // lane supplies 8 consecutive k (8 * (l/16) + e); k >= NB is zero padding
                v8bf av[2], bv[2];
                #pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int i = i0 + 16 * a + fr;
                    #pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = 8 * fg + e;
                        av[a][e] = __bf16((i < nr && k < NB) ? L[i * LD + k] : 0.f);
                    }
                }
                #pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int jj = j0 + 16 * b + fr;
                    #pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = 8 * fg + e;
                        bv[b][e] = __bf16((jj < nr && k < NB) ? U[int64_t(k) * ldw + jj] : 0.f);
                    }
                }
                #pragma unroll
                for (int a = 0; a < 2; ++a)
                    #pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
            }
            else {
                #pragma unroll
                for (int s = 0; s < NB / 4; ++s) {
                    const int k = 4 * s + fg;
                    float av[2], bv[2];
                    #pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        const int i = i0 + 16 * a + fr;
                        av[a] = i < nr ? L[i * LD + k] : 0.f;
                    }
                    #pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const int jj = j0 + 16 * b + fr;
                        bv[b] = jj < nr ? U[int64_t(k) * ldw + jj] : 0.f;
                    }
                    #pragma unroll
                    for (int a = 0; a < 2; ++a)
                        #pragma unroll
                        for (int b = 0; b < 2; ++b)
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
                }
            }
            // register r holds D[4 (l/16) + r][l % 16]
            #pragma unroll
            for (int a = 0; a < 2; ++a)
                #pragma unroll
                for (int b = 0; b < 2; ++b)
                    #pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = crow(i0, a, r), jj = j0 + 16 * b + fr;
                        if (full) C[int64_t(i) * ldw + jj] = cv[a][b][r] - acc[a][b][r];
                        else if (i < nr && jj < nr) C[int64_t(i) * ldw + jj] -= acc[a][b][r];
                    }
        }
    }
}

// In-workgroup LU with partial pivoting of the row-major n x n matrix W.
// perm[i] = original row now at position i (the interchanges composed).
// *s_info: first exactly-zero pivot.
template <class T, int NBF, int BT, bool BFU>
__device__ __forceinline__ void wg_getrf(T* __restrict__ W, int n, int ldw, int* __restrict__ perm, int* s_info,
                                      T* s_p, T* s_red, int* s_ired, int* s_piv, int* s_dst, int* s_src)
{
    constexpr int NB = Pan<T, NBF>::NB, LD = Pan<T, NBF>::LD;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < n; i += BT) perm[i] = i;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        const int mr = n - k0;
        for (int e = tid; e < mr * NB; e += BT) {
            const int i = e / NB, c = e % NB;
            s_p[i * LD + c] = c < kb ? W[int64_t(k0 + i) * ldw + k0 + c] : T(0);
        }
        __syncthreads();
        wg_panel<T, NB, LD, BT>(s_p, mr, kb, k0, s_piv, s_info, s_red, s_ired);
        // Net row moves of the panel's interchanges: lane t < kb follows row
        // k0 + t, lane kb + t the pivot row of column t; each lane tracks
        // which original row ends up in its row (lanes naming the same row
        // stay equal).  A row whose source is itself does not move.
        if (wave == 0) {
            const int cand = lane < kb ? k0 + lane : (lane < 2 * kb ? k0 + s_piv[lane - kb] : -1);
            int src = cand;
            for (int j = 0; j < kb; ++j) {
                const int ra = k0 + j, rb = k0 + s_piv[j];
                const int va = __shfl(src, j, 64), vb = __shfl(src, kb + j, 64);
                if (ra != rb) {
                    if (cand == ra) src = vb;
                    else if (cand == rb) src = va;
                }
            }
            s_dst[lane] = cand;
            s_src[lane] = src;
        }
        // panel back to W
        for (int e = tid; e < mr * NB; e += BT) {
            const int i = e / NB, c = e % NB;
            if (c < kb) W[int64_t(k0 + i) * ldw + k0 + c] = s_p[i * LD + c];
        }
        __syncthreads();
        // the moves on the columns left and right of the panel and on perm
        // (column n): one column per thread, every source loaded before any
        // row is written
        for (int c = tid; c <= n; c += BT) {
            if (c >= k0 && c < k0 + kb) continue;
            asm volatile("" ::: "memory");          // reread the move list per column (keeps registers free)
            if (c < n) {
                T v[2 * NB];
                #pragma unroll
                for (int t = 0; t < 2 * NB; ++t) {
                    const int d = s_dst[t], sr = s_src[t];
                    v[t] = (d >= 0 && sr != d) ? W[sr * ldw + c] : T(0);
                }
                asm volatile("" ::: "memory");      // reread the list below instead of keeping it live
                #pragma unroll
                for (int t = 0; t < 2 * NB; ++t) {
                    const int d = s_dst[t], sr = s_src[t];
                    if (d >= 0 && sr != d) W[d * ldw + c] = v[t];
                }
            }
            else {
                int v[2 * NB];
                #pragma unroll
                for (int t = 0; t < 2 * NB; ++t) {
                    const int d = s_dst[t], sr = s_src[t];
                    v[t] = (d >= 0 && sr != d) ? perm[sr] : 0;
                }
                asm volatile("" ::: "memory");
                #pragma unroll
                for (int t = 0; t < 2 * NB; ++t) {
                    const int d = s_dst[t], sr = s_src[t];
                    if (d >= 0 && sr != d) perm[d] = v[t];
                }
            }
        }
        __syncthreads();
        const int c1 = k0 + kb, nr = n - c1;
        if (nr <= 0) break;
        // U12 = L11^{-1} A12, one column per thread in registers
        for (int c = c1 + tid; c < n; c += BT) {
            asm volatile("" ::: "memory");          // read L11 from LDS per column, not hoisted into registers
            T x[NB];
            #pragma unroll
            for (int r = 0; r < NB; ++r) x[r] = W[int64_t(k0 + r) * ldw + c];
            #pragma unroll
            for (int r = 1; r < NB; ++r) {
                asm volatile("" ::: "memory");      // row r's L11 reads stay inside its step
                T s = x[r];
                #pragma unroll
                for (int q = 0; q < r; ++q) s -= s_p[r * LD + q] * x[q];
                x[r] = s;
            }
            #pragma unroll
            for (int r = 1; r < NB; ++r) W[int64_t(k0 + r) * ldw + c] = x[r];
        }
        __syncthreads();
        wg_update<T, NB, LD, BT, BFU>(W + int64_t(c1) * ldw + c1, W + int64_t(k0) * ldw + c1, ldw,
                                      s_p + kb * LD, nr);
        __syncthreads();
    }
}

// x := op(LU)^{-1} x for the factors in W (row-major), x of length n in
// global memory (FP64).  Lower (unit) or upper, 64-row blocks: the diagonal
// block is staged in LDS (s_blk, 64 x 65 T) and solved by wave 0 with
// shuffles, the solved piece is kept in LDS (s_x) and the rows below (above)
// are updated by the whole workgroup, one row per thread with eight
// independent partial sums.
template <class T, int BT>
__device__ __forceinline__ void wg_trsv(const T* __restrict__ W, int n, int ldw, double* __restrict__ x, bool lower,
                                        T* s_blk, double* s_x)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (n + 63) / 64;
    for (int bb = 0; bb < nblk; ++bb) {
        const int blk = lower ? bb : nblk - 1 - bb;
        const int b0 = blk * 64, bn = min(64, n - b0);
        for (int e = tid; e < 64 * 64; e += BT) {
            const int r = e >> 6, c = e & 63;
            s_blk[r * 65 + c] = (r < bn && c < bn) ? W[int64_t(b0 + r) * ldw + b0 + c] : T(0);
        }
        __syncthreads();
        if (wave == 0) {
            double y = lane < bn ? x[b0 + lane] : 0.0;
            if (lower) {
                for (int j = 0; j < bn; ++j) {
                    const double yj = __shfl(y, j, 64);
                    if (lane > j) y -= double(s_blk[lane * 65 + j]) * yj;
                }
            }
            else {
                for (int j = bn - 1; j >= 0; --j) {
                    if (lane == j) y /= double(s_blk[j * 65 + j]);
                    const double yj = __shfl(y, j, 64);
                    if (lane < j) y -= double(s_blk[lane * 65 + j]) * yj;
                }
            }
            if (lane < bn) x[b0 + lane] = y;
            s_x[lane] = lane < bn ? y : 0.0;
        }
        __syncthreads();
        const int r0 = lower ? b0 + bn : 0, r1 = lower ? n : b0;
        for (int i = r0 + tid; i < r1; i += BT) {
            const T* wr = W + int64_t(i) * ldw + b0;
            double ps[8];
            #pragma unroll
            for (int u = 0; u < 8; ++u) ps[u] = 0.0;
            if (bn == 64) {
                #pragma unroll
                for (int j = 0; j < 64; j += 8)
                    #pragma unroll
                    for (int u = 0; u < 8; ++u) ps[u] += double(wr[j + u]) * s_x[j + u];
            }
            else {
                for (int j = 0; j < bn; ++j) ps[j & 7] += double(wr[j]) * s_x[j];
            }
            x[i] -= ((ps[0] + ps[1]) + (ps[2] + ps[3])) + ((ps[4] + ps[5]) + (ps[6] + ps[7]));
        }
        __syncthreads();
    }
}

struct BatchArgs {
    int count;
    const int* order;               // systems, largest first
    const int* n;                   // sizes
    const int* mode;                // per system, see the file comment
    const int64_t* offA;            // A_i at A + offA[i], column-major, ld n_i (FP64)
    const int64_t* offB;            // b_i / x_i at B + offB[i]
    const double* A;
    const double* B;
    double* X;
    int* iters;                     // refinement steps (>= 0), -(itermax+1) fallback, -3 singular factor
    int* info;                      // factorization info (FP64 path or fallback)
    double* berr;                   // ||b - A x||_inf / (||A||_inf ||x||_inf)
    int* choice;                    // what ran: 0 BF16 + IR, 1 FP32 + IR, 2 FP64
    float* feat;                    // learned path: SEL_F features per system (may be null)
    unsigned long long* ticks;      // wall-clock ticks per system (100 MHz; may be null)
    const float* sel;               // selector network parameters (learned path)
    int* queue;                     // [0] ticket, [1] finished workgroups; zero between launches
    char* work;                     // per-workgroup workspace
    int64_t work_stride;            // bytes per workgroup
    int itermax;
    double tol;                     // <= 0: eps * sqrt(n)
    int probe_bf16;                 // learned path: probe factor in BF16 (1) or FP32 (0)
};

// sum_j |A(i, j)| (ABS) or sum_j A(i, j) v_j over the n columns of the
// column-major A, with v read from LDS; eight independent partial sums keep
// eight loads in flight per thread.
template <bool ABS>
__device__ __forceinline__ double row_dot(const double* __restrict__ A, int n, int i, const double* s_v)
{
    double ps[8];
    #pragma unroll
    for (int u = 0; u < 8; ++u) ps[u] = 0.0;
    int j = 0;
    // 32 loads in flight per thread; column j + u still goes to partial sum
    // u mod 8 in column order, as in the 8-column loop below
    for (; j + 32 <= n; j += 32) {
        double av[32];
        #pragma unroll
        for (int u = 0; u < 32; ++u) av[u] = A[i + int64_t(j + u) * n];
        #pragma unroll
        for (int u = 0; u < 32; ++u) ps[u & 7] += ABS ? fabs(av[u]) : av[u] * s_v[j + u];
    }
    for (; j + 8 <= n; j += 8)
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double a = A[i + int64_t(j + u) * n];
            ps[u] += ABS ? fabs(a) : a * s_v[j + u];
        }
    for (; j < n; ++j) {
        const double a = A[i + int64_t(j) * n];
        ps[j & 7] += ABS ? fabs(a) : a * s_v[j];
    }
    return ((ps[0] + ps[1]) + (ps[2] + ps[3])) + ((ps[4] + ps[5]) + (ps[6] + ps[7]));
}

// r = b - A x (FP64, A column-major) and infinity norms: ||r|| in s_dred[0],
// ||x|| in s_dred[1].  s_v: n doubles of LDS for x.
template <int BT>
__device__ void wg_residual(const double* __restrict__ A, int n, const double* __restrict__ b,
                            const double* __restrict__ x, double* __restrict__ r, double* s_dred, double* s_v)
{
    constexpr int NW = BT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < n; i += BT) s_v[i] = x[i];
    __syncthreads();
    double rmax = 0, xmax = 0;
    for (int i = tid; i < n; i += BT) {
        const double s = b[i] - row_dot<false>(A, n, i, s_v);
        r[i] = s;
        rmax = fmax(rmax, fabs(s));
        xmax = fmax(xmax, fabs(s_v[i]));
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        rmax = fmax(rmax, __shfl_xor(rmax, o, 64));
        xmax = fmax(xmax, __shfl_xor(xmax, o, 64));
    }
    __syncthreads();
    if (lane == 0) { s_dred[2 + wave] = rmax; s_dred[2 + NW + wave] = xmax; }
    __syncthreads();
    if (tid == 0) {
        double a = 0, c = 0;
        for (int w = 0; w < NW; ++w) { a = fmax(a, s_dred[2 + w]); c = fmax(c, s_dred[2 + NW + w]); }
        s_dred[0] = a;
        s_dred[1] = c;
    }
    __syncthreads();
}

// max over rows of the row sums of |A| (column-major) -> s_dred[0];
// max |v| of v (length n) -> s_dred[1].
template <int BT>
__device__ void wg_norms(const double* __restrict__ A, int n, const double* __restrict__ v, double* s_dred)
{
    constexpr int NW = BT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double m = 0, vm = 0;
    for (int i = tid; i < n; i += BT) {
        m = fmax(m, row_dot<true>(A, n, i, nullptr));
        vm = fmax(vm, fabs(v[i]));
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_xor(m, o, 64));
        vm = fmax(vm, __shfl_xor(vm, o, 64));
    }
    __syncthreads();
    if (lane == 0) { s_dred[2 + wave] = m; s_dred[2 + NW + wave] = vm; }
    __syncthreads();
    if (tid == 0) {
        double a = 0, c = 0;
        for (int w = 0; w < NW; ++w) { a = fmax(a, s_dred[2 + w]); c = fmax(c, s_dred[2 + NW + w]); }
        s_dred[0] = a;
        s_dred[1] = c;
    }
    __syncthreads();
}

// Selector network forward pass by wave 0 (h1 = relu(W1^T f + b1),
// h2 = relu(W2^T h1 + b2), z = W3^T h2 + b3; lane h holds hidden unit h);
// returns argmax over the first SEL_CHOICES outputs (lane 0 valid).
__device__ int sel_choose(const float* __restrict__ P, const float* f)
{
    const int lane = threadIdx.x & 63;
    const float* W1 = P;
    const float* b1 = W1 + SEL_F * SEL_H;
    const float* W2 = b1 + SEL_H;
    const float* b2 = W2 + SEL_H * SEL_H;
    const float* W3 = b2 + SEL_H;
    const float* b3 = W3 + SEL_H * SEL_K;
    float h = b1[lane];
    for (int q = 0; q < SEL_F; ++q) h += f[q] * W1[q * SEL_H + lane];
    h = fmaxf(h, 0.f);
    float h2 = b2[lane];
    for (int g = 0; g < SEL_H; ++g) h2 += __shfl(h, g, 64) * W2[g * SEL_H + lane];
    h2 = fmaxf(h2, 0.f);
    const int k = lane & (SEL_K - 1);
    float z = b3[k];
    for (int g = 0; g < SEL_H; ++g) z += __shfl(h2, g, 64) * W3[g * SEL_K + k];
    int best = 0;
    float bz = __shfl(z, 0, 64);
    for (int c = 1; c < SEL_CHOICES; ++c) {
        const float zc = __shfl(z, c, 64);
        if (zc > bz) { bz = zc; best = c; }
    }
    return best;
}

template <int NBF, int BT>
__global__ __launch_bounds__(BT)
void batched_solve_kernel(BatchArgs a)
{
    constexpr int NW = BT / 64;
    __shared__ __attribute__((aligned(16))) char s_pan[pan_lds_bytes<NBF>()];
    __shared__ double s_red[NW];
    __shared__ int s_ired[NW];
    __shared__ int s_piv[NBF];
    __shared__ int s_dst[64], s_src[64];
    __shared__ double s_dred[2 + 2 * NW];
    __shared__ float s_feat[SEL_F];
    __shared__ int s_sys, s_info, s_choice;
    const int tid = threadIdx.x;
    char* ws = a.work + int64_t(blockIdx.x) * a.work_stride;
    // workspace: W32 (1024^2 float) | W64 (1024^2 double) | perm (1024 int) | r, y (1024 double)
    float* W = reinterpret_cast<float*>(ws);
    double* W64 = reinterpret_cast<double*>(ws + int64_t(BMAXN) * BMAXN * sizeof(float));
    int* perm = reinterpret_cast<int*>(W64 + int64_t(BMAXN) * BMAXN);
    double* r = reinterpret_cast<double*>(perm + BMAXN);
    double* y = r + BMAXN;
    float* s_pf = reinterpret_cast<float*>(s_pan);
    double* s_pd = reinterpret_cast<double*>(s_pan);
    // outside the factorizations the panel LDS holds the solves' diagonal
    // block (64 x 65, up to FP64) and vectors: x for residuals, a solved piece
    static_assert(pan_lds_bytes<NBF>() >= 64 * 65 * 8 + BMAXN * 8 + 64 * 8, "panel LDS too small for the solves");
    double* s_v = reinterpret_cast<double*>(s_pan + 64 * 65 * 8);
    double* s_x = s_v + BMAXN;
    const double eps = std::numeric_limits<double>::epsilon();
    while (true) {
        if (tid == 0) {
            const int q = atomicAdd(&a.queue[0], 1);
            s_sys = q < a.count ? a.order[q] : -1;
            s_info = 0;
        }
        __syncthreads();
        const int sys = s_sys;
        if (sys < 0) break;                               // every wave leaves here
        const unsigned long long t0 = wall_clock64();
        const int n = a.n[sys], ldw = b_ldw(n);
        const int mode = a.mode[sys];
        const double* A = a.A + a.offA[sys];
        const double* b = a.B + a.offB[sys];
        double* x = a.X + a.offB[sys];
        const double tol = a.tol > 0 ? a.tol : eps * sqrt(double(n));

        // ||A||_inf and ||b||_inf
        wg_norms<BT>(A, n, b, s_dred);
        const double anorm = s_dred[0], bnorm = s_dred[1];
        __syncthreads();

        // out = (LU)^{-1} P rhs in FP64 storage, LU in W (float) or W64
        auto lu_solve = [&](bool f64, const double* rhs, double* out) {
            for (int i = tid; i < n; i += BT) out[i] = rhs[perm[i]];
            __syncthreads();
            if (f64) {
                wg_trsv<double, BT>(W64, n, ldw, out, true, s_pd, s_x);
                wg_trsv<double, BT>(W64, n, ldw, out, false, s_pd, s_x);
            }
            else {
                wg_trsv<float, BT>(W, n, ldw, out, true, s_pf, s_x);
                wg_trsv<float, BT>(W, n, ldw, out, false, s_pf, s_x);
            }
        };
        auto factor32 = [&](bool bf) {
            wg_load_rowmajor<float, BT>(A, n, W, ldw);
            if (tid == 0) s_info = 0;
            __syncthreads();
            if (bf)
                wg_getrf<float, NBF, BT, true>(W, n, ldw, perm, &s_info, s_pf, reinterpret_cast<float*>(s_red),
                                               s_ired, s_piv, s_dst, s_src);
            else
                wg_getrf<float, NBF, BT, false>(W, n, ldw, perm, &s_info, s_pf, reinterpret_cast<float*>(s_red),
                                                s_ired, s_piv, s_dst, s_src);
            __syncthreads();
            return s_info;
        };
        // Refinement of x with the float factor in W; have_r: the residual of
        // x is already in r / s_dred.  Returns the iterations, or -1 on a miss.
        auto refine = [&](bool have_r) {
            for (int it = 0; it <= a.itermax; ++it) {
                if (!(have_r && it == 0)) wg_residual<BT>(A, n, b, x, r, s_dred, s_v);
                const bool conv = s_dred[0] <= s_dred[1] * anorm * tol;
                __syncthreads();
                if (conv) return it;
                if (it == a.itermax) break;
                lu_solve(false, r, y);
                for (int i = tid; i < n; i += BT) x[i] += y[i];
                __syncthreads();
            }
            return -1;
        };

        int info = 0, iters = 0, ch;
        bool need64 = false;
        if (mode == 1) {
            ch = 2;
            need64 = true;
        }
        else if (mode == 0 || mode == 2) {
            ch = mode == 0 ? 1 : 0;
            if (factor32(mode == 2) != 0) { need64 = true; iters = -3; }
            else {
                lu_solve(false, b, x);
                const int it = refine(false);
                if (it < 0) { need64 = true; iters = -a.itermax - 1; }
                else iters = it;
            }
        }
        else {
            // learned path: probe factor (BF16 or FP32), one solve, iteration-0 residual
            const int pinfo = factor32(a.probe_bf16 != 0);
            double xnorm = 0, rnorm = 0;
            if (pinfo == 0) {
                lu_solve(false, b, x);
                wg_residual<BT>(A, n, b, x, r, s_dred, s_v);
                rnorm = s_dred[0];
                xnorm = s_dred[1];
            }
            if (tid == 0) {
                for (int q = 0; q < SEL_F; ++q) s_feat[q] = 0.f;
                const double tiny = 1e-300;
                s_feat[0] = 1.f;
                s_feat[1] = float(log2(double(n)) / 10.0);
                const bool ok = pinfo == 0 && isfinite(rnorm) && isfinite(xnorm) && xnorm > 0 && bnorm > 0
                                && anorm > 0;
                if (ok) {
                    s_feat[2] = float(log10(fmax(anorm * xnorm / bnorm, tiny)) / 8.0);
                    s_feat[3] = float(log10(fmax(rnorm / (anorm * xnorm), tiny)) / 8.0);
                    s_feat[5] = float(log10(fmax(rnorm / bnorm, tiny)) / 8.0);
                    s_feat[6] = (rnorm <= xnorm * anorm * tol) ? 1.f : 0.f;
                }
                else {
                    s_feat[4] = 1.f;
                }
                s_feat[7] = a.probe_bf16 ? 1.f : 0.f;
            }
            __syncthreads();
            if (a.feat && tid < SEL_F) a.feat[int64_t(sys) * SEL_F + tid] = s_feat[tid];
            int c = mode - 4;
            if (mode == 3) {
                if (tid < 64) {
                    const int cc = sel_choose(a.sel, s_feat);
                    if (tid == 0) s_choice = cc;
                }
                __syncthreads();
                c = s_choice;
            }
            // what runs: 0 BF16 + IR, 1 FP32 + IR, 2 FP64.  A singular or
            // non-finite probe cannot be kept (BF16 probe: FP32 instead; FP32
            // probe: FP64); with an FP32 probe choices 0 and 1 both keep it.
            const bool bad = s_feat[4] != 0.f;
            bool keep;
            if (a.probe_bf16) {
                if (c == 0 && bad) c = 1;
                keep = c == 0;
            }
            else {
                if (c <= 1) c = bad ? 2 : 1;
                keep = c == 1;
            }
            ch = c;
            if (keep) {
                const int it = refine(true);
                if (it < 0) { need64 = true; iters = -a.itermax - 1; }
                else iters = it;
            }
            else if (c == 1) {
                if (factor32(false) != 0) { need64 = true; iters = -3; }
                else {
                    lu_solve(false, b, x);
                    const int it = refine(false);
                    if (it < 0) { need64 = true; iters = -a.itermax - 1; }
                    else iters = it;
                }
            }
            else {
                need64 = true;
                iters = 0;
            }
        }
        if (need64) {
            wg_load_rowmajor<double, BT>(A, n, W64, ldw);
            if (tid == 0) s_info = 0;
            __syncthreads();
            wg_getrf<double, NBF, BT, false>(W64, n, ldw, perm, &s_info, s_pd, s_red, s_ired, s_piv, s_dst, s_src);
            __syncthreads();
            info = s_info;
            lu_solve(true, b, x);
        }
        // backward error
        wg_residual<BT>(A, n, b, x, r, s_dred, s_v);
        if (tid == 0) {
            a.iters[sys] = iters;
            a.info[sys] = info;
            const double den = anorm * s_dred[1];
            a.berr[sys] = den > 0 ? s_dred[0] / den : 0.0;
            if (a.choice) a.choice[sys] = ch;
            if (a.ticks) a.ticks[sys] = wall_clock64() - t0;
        }
        __syncthreads();
    }
    // the last workgroup to finish leaves the queue zeroed for the next launch
    if (tid == 0) {
        __threadfence();
        const int done = atomicAdd(&a.queue[1], 1);
        if (done == int(gridDim.x) - 1) {
            atomicExch(&a.queue[0], 0);
            atomicExch(&a.queue[1], 0);
        }
    }
}

int64_t batched_work_stride(int)
{
    int64_t b = int64_t(BMAXN) * BMAXN * (4 + 8) + int64_t(BMAXN) * (4 + 16) + 256;
    return (b + 255) & ~int64_t(255);
}

// Kernel of a tile configuration: float panel width nbf (16 or 32) and
// workgroup size bt (256; 512 with nbf 16 only: the 32-wide panel at 512
// threads spills most of its working registers).
static const void* batched_kernel_ptr(int nbf, int bt)
{
    if (nbf == 16) return bt == 512 ? (const void*)batched_solve_kernel<16, 512> : (const void*)batched_solve_kernel<16, 256>;
    return (const void*)batched_solve_kernel<32, 256>;
}

bool batched_config_valid(int nbf, int bt) { return (nbf == 16 && (bt == 256 || bt == 512)) || (nbf == 32 && bt == 256); }

int batched_max_wgs_per_cu(int nbf, int bt)
{
    SA_REQUIRE(batched_config_valid(nbf, bt), -2, "batched: bad tile configuration");
    int nb = 0;
    SA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batched_kernel_ptr(nbf, bt), bt, 0));
    return nb;
}

void batched_fill_args(void* out, int count, const int* order, const int* n, const int* mode, const int64_t* offA,
                       const int64_t* offB, const double* A, const double* B, double* X, int* iters, int* info,
                       double* berr, int* choice, float* feat, unsigned long long* ticks, const float* sel,
                       int* queue, char* work, int itermax, double tol, int probe_bf16)
{
    BatchArgs a{count, order, n, mode, offA, offB, A, B, X, iters, info, berr, choice, feat, ticks, sel, queue,
                work, batched_work_stride(4), itermax, tol, probe_bf16};
    static_assert(sizeof(BatchArgs) <= BATCH_ARGS_BYTES, "BatchArgs size");
    std::memcpy(out, &a, sizeof(a));
}

void batched_launch(const void* args, int nbf, int bt, int grid, hipStream_t s)
{
    SA_REQUIRE(batched_config_valid(nbf, bt), -2, "batched: bad tile configuration");
    const BatchArgs& a = *static_cast<const BatchArgs*>(args);
    if (a.count <= 0) return;
    if (nbf == 16) {
        if (bt == 512) hipLaunchKernelGGL((batched_solve_kernel<16, 512>), dim3(unsigned(grid)), dim3(512), 0, s, a);
        else hipLaunchKernelGGL((batched_solve_kernel<16, 256>), dim3(unsigned(grid)), dim3(256), 0, s, a);
    }
    else {
        hipLaunchKernelGGL((batched_solve_kernel<32, 256>), dim3(unsigned(grid)), dim3(256), 0, s, a);
    }
    SA_CHECK_LAUNCH();
}

// A one-node graph holding the launch (the kernel resets its own queue, so
// replays need nothing else).  args must stay valid while the graph exists.
void batched_graph(void* args, int nbf, int bt, int grid, hipGraph_t* g, hipGraphExec_t* ge)
{
    SA_REQUIRE(batched_config_valid(nbf, bt), -2, "batched: bad tile configuration");
    SA_HIP(hipGraphCreate(g, 0));
    hipKernelNodeParams kp{};
    void* kargs[1] = {args};
    kp.func = const_cast<void*>(batched_kernel_ptr(nbf, bt));
    kp.gridDim = dim3(unsigned(grid));
    kp.blockDim = dim3(unsigned(bt));
    kp.sharedMemBytes = 0;
    kp.kernelParams = kargs;
    kp.extra = nullptr;
    hipGraphNode_t node;
    SA_HIP(hipGraphAddKernelNode(&node, *g, nullptr, 0, &kp));
    SA_HIP(hipGraphInstantiate(ge, *g, nullptr, nullptr, 0));
}

int batched_max_n() { return BMAXN; }

} // namespace sa
