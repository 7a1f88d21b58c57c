// Small dense tile kernels: triangular-solve leaf, Cholesky leaf, layout
// transposes and precision conversions (gfx950).
//
// trsm_leaf and potrf_leaf are the diagonal-block work of the blocked
// triangular solve and Cholesky; the blocks above them are recursive and run
// their off-diagonal updates through the MFMA GEMM.  They replace the vendor
// batch::trsm (src/internal/internal_trsm.cc:244-249) and device lapack::potrf
// (src/internal/internal_potrf.cc:72-78, which also synchronised the host per
// step to read info; here info stays on the device).
// transpose_inplace is the square in-place layout flip that SLATE performs per
// tile with device::transpose (src/cuda/device_transpose.cu:282-316) when it
// moves tiles to RowMajor for the GPU LU (src/getrf.cc:54-55).
// copy_convert / transpose_convert are the precision casts of slate::copy
// (src/internal/internal_gecopy.cc:121 -> device_gecopy.cu:40-115).
#include "common.hh"
#include "kernels.hh"

namespace sa {

//------------------------------------------------------------------------------
// One thread per right-hand-side column; the triangle lives in LDS and is
// padded to NB with an identity so the fully unrolled substitution has no
// data-dependent bounds.
// FULL (nb == TRSM_LEAF) has no row predicates: predicated stores of the
// unrolled register array make the compiler spill it.
template <class T, int LOWER, int UNIT, int CM, int FULL, int NB>
__global__ __launch_bounds__(256)
void trsm_leaf_kernel(const T* __restrict__ M, int64_t ld, int nb, T* __restrict__ B,
                      int64_t ldb, int64_t N)
{
    __shared__ T L[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += 256) {
        int i = e / NB, k = e % NB;
        T val;
        if (i < nb && k < nb) val = CM ? M[i + int64_t(k) * ld] : M[int64_t(i) * ld + k];
        else val = (i == k) ? T(1) : T(0);
        L[i][k] = val;
    }
    __syncthreads();
    const int64_t j = int64_t(blockIdx.x) * 256 + tid;
    if (j >= N) return;
    T x[NB];
    #pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = (FULL || i < nb) ? B[int64_t(i) * ldb + j] : T(0);
    // Row-oriented substitution; the scheduling barrier per row keeps the
    // compiler from hoisting every LDS read of the triangle (which spills).
    if (LOWER) {
        #pragma unroll
        for (int i = 0; i < NB; ++i) {
            T s = x[i];
            #pragma unroll
            for (int k = 0; k < i; ++k) s -= L[i][k] * x[k];
            x[i] = UNIT ? s : s / L[i][i];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    else {
        #pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            T s = x[i];
            #pragma unroll
            for (int k = i + 1; k < NB; ++k) s -= L[i][k] * x[k];
            x[i] = UNIT ? s : s / L[i][i];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    #pragma unroll
    for (int i = 0; i < NB; ++i)
        if (FULL || i < nb) B[int64_t(i) * ldb + j] = x[i];
}

template <class T>
void trsm_leaf(const T* M, int64_t ld, int cm, int lower, int unit, int64_t nb,
               T* B, int64_t ldb, int64_t N, hipStream_t s)
{
    if (N <= 0 || nb <= 0) return;
    SA_REQUIRE(nb <= TRSM_LEAF, -2, "trsm_leaf: block larger than the leaf");
    dim3 grid(unsigned(ceildiv(N, 256))), blk(256);
    int key = (lower ? 4 : 0) | (unit ? 2 : 0) | (cm ? 1 : 0);
    int inb = int(nb);
#define SA_TRSM_CASE(K, LO, UN, CMV)                                                              \
    case K:                                                                                       \
        if (nb == TRSM_LEAF)                                                                      \
            hipLaunchKernelGGL((trsm_leaf_kernel<T, LO, UN, CMV, 1, TRSM_LEAF>), grid, blk, 0, s, M, ld, inb, B, ldb, N); \
        else if (nb == TRSM_LEAF / 2)                                                             \
            hipLaunchKernelGGL((trsm_leaf_kernel<T, LO, UN, CMV, 1, TRSM_LEAF / 2>), grid, blk, 0, s, M, ld, inb, B, ldb, N); \
        else                                                                                      \
            hipLaunchKernelGGL((trsm_leaf_kernel<T, LO, UN, CMV, 0, TRSM_LEAF>), grid, blk, 0, s, M, ld, inb, B, ldb, N); \
        break;
    switch (key) {
        SA_TRSM_CASE(0, 0, 0, 0)
        SA_TRSM_CASE(1, 0, 0, 1)
        SA_TRSM_CASE(2, 0, 1, 0)
        SA_TRSM_CASE(3, 0, 1, 1)
        SA_TRSM_CASE(4, 1, 0, 0)
        SA_TRSM_CASE(5, 1, 0, 1)
        SA_TRSM_CASE(6, 1, 1, 0)
        default:
        SA_TRSM_CASE(7, 1, 1, 1)
    }
#undef SA_TRSM_CASE
    SA_CHECK_LAUNCH();
}

template void trsm_leaf<double>(const double*, int64_t, int, int, int, int64_t, double*, int64_t,
                                int64_t, hipStream_t);
template void trsm_leaf<float>(const float*, int64_t, int, int, int, int64_t, float*, int64_t,
                               int64_t, hipStream_t);

//------------------------------------------------------------------------------
// L X = B in place for a lower triangle (tri(i, k) = L[i*ldl + k], m x m;
// unit diagonal or not) and a row-major B (m x N, ldb), in one launch,
// m <= trsm_fused_max.
// Workgroup x owns the columns [8x, 8x + 8), kept in LDS k-contiguous.  The
// sweep is left-looking over 32-row blocks: every thread owns one (row,
// column) of the block and forms y = b - L(i, 0:p0) X(0:p0, c) from LDS; then
// the block's forward substitution runs in registers, each wave holding two
// columns (one per 32-lane half) and broadcasting each solved entry with
// v_readlane.  The block row L(p0:p0+32, 0:p0+32) reaches LDS in chunks of
// 32 rows x 512 bytes, each loaded once per workgroup with coalesced 16-byte
// loads (consecutive threads read consecutive pieces of a row) and the next
// chunk's loads in flight while the current one is multiplied.  Reading the
// row straight from memory instead, one row per lane, makes every load touch
// 32 rows and leaves the kernel bound by address processing (1.7-2x slower,
// the same sums).  Two barriers per chunk, no launches inside the sweep: the
// recursive form needs 2 m/32 - 1 launches.
template <class T> __device__ __forceinline__ T bcast_half(T v, int k, bool upper);
template <> __device__ __forceinline__ float bcast_half<float>(float v, int k, bool upper)
{
    const int iv = __builtin_bit_cast(int, v);
    const int a = __builtin_amdgcn_readlane(iv, k);
    const int b = __builtin_amdgcn_readlane(iv, k + 32);
    return __builtin_bit_cast(float, upper ? b : a);
}
template <> __device__ __forceinline__ double bcast_half<double>(double v, int k, bool upper)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = int(uint32_t(u)), hi = int(uint32_t(u >> 32));
    const uint32_t alo = uint32_t(__builtin_amdgcn_readlane(lo, k));
    const uint32_t ahi = uint32_t(__builtin_amdgcn_readlane(hi, k));
    const uint32_t blo = uint32_t(__builtin_amdgcn_readlane(lo, k + 32));
    const uint32_t bhi = uint32_t(__builtin_amdgcn_readlane(hi, k + 32));
    const uint64_t r = upper ? (uint64_t(bhi) << 32 | blo) : (uint64_t(ahi) << 32 | alo);
    return __builtin_bit_cast(double, r);
}

template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float __attribute__((ext_vector_type(4))) type; static constexpr int n = 4; };
template <> struct Vec16<double> { typedef double __attribute__((ext_vector_type(2))) type; static constexpr int n = 2; };

template <class T>
struct TrsmProb {
    const T* L; int64_t ldl; T* B; int64_t ldb; int64_t N; int m; int ldx; int vec_ok;
};

// 16-byte vectors per row of a staged chunk of the block row
static constexpr int TRSM_NV = 32;

template <class T, int UNIT>
__global__ __launch_bounds__(256)
void trsm_lunit_kernel(TrsmProb<T> P)
{
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::n;
    constexpr int NV = TRSM_NV;
    constexpr int KC = NV * VN;                         // columns per chunk
    constexpr int LP = KC + VN;                         // padded chunk row pitch
    constexpr int NU = 32 * NV / 256;                   // chunk vectors per thread
    const int64_t j0 = int64_t(blockIdx.x) * 8;
    if (j0 >= P.N) return;                              // uniform per workgroup, before any barrier
    const T* __restrict__ L = P.L;
    T* __restrict__ B = P.B;
    const int64_t ldl = P.ldl, ldb = P.ldb, N = P.N;
    const int m = P.m, ldx = P.ldx;
    extern __shared__ __align__(16) unsigned char trsm_lds[];
    T* Xs = reinterpret_cast<T*>(trsm_lds);            // column c at Xs + c*ldx
    T* Lc = Xs + 8 * ldx;                               // 32 x KC chunk of the block row, pitch LP
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = lane & 31;
    const bool upper = lane >= 32;
    const int cl = 2 * (tid >> 6) + (upper ? 1 : 0);
    {
        const int c = tid & 7;
        const bool cv = j0 + c < N;
        for (int i = tid >> 3; i < m; i += 32)
            Xs[c * ldx + i] = cv ? B[int64_t(i) * ldb + j0 + c] : T(0);
    }
    T* Xc = Xs + cl * ldx;
    const V* Lr = reinterpret_cast<const V*>(Lc + r * LP);
    // chunk (q0, kc): rows [q0, q0 + 32), columns [kc, kc + KC), strictly-lower
    // entries only (zeros elsewhere), into registers
    V pv[NU];
    auto fetch = [&](int q0, int kc) {
        #pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int e = tid + 256 * u;
            const int col = kc + (e % NV) * VN, row = q0 + e / NV;
            #pragma unroll
            for (int t = 0; t < VN; ++t) pv[u][t] = T(0);
            if (row < m && col < row) {
                const T* src = L + int64_t(row) * ldl + col;
                if (P.vec_ok && col + VN <= row) pv[u] = *reinterpret_cast<const V*>(src);
                else {
                    #pragma unroll
                    for (int t = 0; t < VN; ++t) if (col + t < row) pv[u][t] = src[t];
                }
            }
        }
    };
    fetch(0, 0);
    for (int p0 = 0; p0 < m; p0 += 32) {
        const int i = p0 + r;
        T acc[VN];
        #pragma unroll
        for (int e = 0; e < VN; ++e) acc[e] = T(0);
        T lr[32];
        // chunks cover columns [0, p0 + 32); the last one holds the diagonal block
        for (int kc = 0; kc < p0 + 32; kc += KC) {
            __syncthreads();                            // previous chunk consumed, X rows written
            #pragma unroll
            for (int u = 0; u < NU; ++u) {
                const int e = tid + 256 * u;
                *reinterpret_cast<V*>(Lc + (e / NV) * LP + (e % NV) * VN) = pv[u];
            }
            __syncthreads();
            {
                int np0 = p0, nkc = kc + KC;
                if (nkc >= p0 + 32) { np0 = p0 + 32; nkc = 0; }
                if (np0 < m) fetch(np0, nkc);           // in flight during the products below
            }
            const int kdot = (p0 - kc < KC ? p0 - kc : KC) / VN;   // vectors left of the diagonal block
            #pragma unroll 8
            for (int v = 0; v < kdot; ++v) {
                const V l = Lr[v];
                const V x = *reinterpret_cast<const V*>(Xc + kc + v * VN);
                #pragma unroll
                for (int e = 0; e < VN; ++e) acc[e] += l[e] * x[e];
            }
            if (p0 + 32 <= kc + KC) {
                const T* d = Lc + r * LP + (p0 - kc);
                #pragma unroll
                for (int k = 0; k < 32; ++k) lr[k] = d[k];
            }
        }
        T y = T(0);
        if (i < m) {
            T sum = acc[0];
            #pragma unroll
            for (int e = 1; e < VN; ++e) sum += acc[e];
            y = Xc[i] - sum;
        }
        const T diag = (!UNIT && i < m) ? L[int64_t(i) * ldl + i] : T(1);
        #pragma unroll
        for (int k = 0; k < 31; ++k) {
            if (!UNIT && r == k) y = y / diag;
            const T xk = bcast_half<T>(y, k, upper);
            if (r > k) y -= lr[k] * xk;
        }
        if (!UNIT && r == 31) y = y / diag;
        if (i < m) Xc[i] = y;
    }
    __syncthreads();
    {
        const int c = tid & 7;
        if (j0 + c < N)
            for (int i = tid >> 3; i < m; i += 32) B[int64_t(i) * ldb + j0 + c] = Xs[c * ldx + i];
    }
}

int64_t trsm_fused_max(int elem_bytes) { return elem_bytes == 8 ? 768 : 1024; }

template <class T>
static TrsmProb<T> trsm_prob(const T* L, int64_t ldl, int64_t m, T* B, int64_t ldb, int64_t N, size_t& lds)
{
    SA_REQUIRE(m <= trsm_fused_max(int(sizeof(T))), -2, "trsm_lunit_fused: triangle too large");
    SA_REQUIRE(ceildiv(N, 8) < (int64_t(1) << 31), -2, "trsm_lunit_fused: too many columns");
    constexpr int VN = Vec16<T>::n;
    TrsmProb<T> q;
    q.L = L; q.ldl = ldl; q.B = B; q.ldb = ldb; q.N = N; q.m = int(m);
    q.ldx = int(ceildiv(m, 32) * 32) + VN;
    q.vec_ok = (reinterpret_cast<uintptr_t>(L) % 16 == 0 && ldl % VN == 0) ? 1 : 0;
    lds = std::max(lds, sizeof(T) * (size_t(8) * size_t(q.ldx) + 32 * size_t(TRSM_NV * VN + VN)));
    return q;
}

template <class T>
void trsm_lower_fused(const T* L, int64_t ldl, int64_t m, T* B, int64_t ldb, int64_t N, int unit, hipStream_t s)
{
    if (m <= 0 || N <= 0) return;
    size_t lds = 0;
    const TrsmProb<T> q = trsm_prob(L, ldl, m, B, ldb, N, lds);
    if (unit) hipLaunchKernelGGL((trsm_lunit_kernel<T, 1>), dim3(unsigned(ceildiv(N, 8))), dim3(256), lds, s, q);
    else hipLaunchKernelGGL((trsm_lunit_kernel<T, 0>), dim3(unsigned(ceildiv(N, 8))), dim3(256), lds, s, q);
    SA_CHECK_LAUNCH();
}

template <class T>
void trsm_lunit_fused(const T* L, int64_t ldl, int64_t m, T* B, int64_t ldb, int64_t N, hipStream_t s)
{
    trsm_lower_fused<T>(L, ldl, m, B, ldb, N, 1, s);
}

template void trsm_lower_fused<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, int, hipStream_t);
template void trsm_lower_fused<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, int, hipStream_t);
template void trsm_lunit_fused<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template void trsm_lunit_fused<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, hipStream_t);
//------------------------------------------------------------------------------
// Right-looking upper Cholesky (A = U^T U) of one nb x nb block in LDS.
// Row-major upper == column-major lower, so this is LAPACK's lower dpotrf on
// the caller's column-major tile.  Stops at the first non-positive pivot.
template <class T>
__global__ __launch_bounds__(256)
void potrf_leaf_kernel(T* __restrict__ A, int64_t ld, int nb, int64_t col0, int64_t* info)
{
    constexpr int NB = 64;
    __shared__ T S[NB][NB + 1];
    __shared__ int fail;
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e / nb, j = e % nb;
        if (j >= i) S[i][j] = A[int64_t(i) * ld + j];
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    int nbad = nb;
    for (int k = 0; k < nb; ++k) {
        if (tid == 0) {
            T d = S[k][k];
            if (!(d > T(0))) {
                fail = 1;
                if (*info == 0) *info = col0 + k + 1;
            }
            else S[k][k] = sqrt(d);
        }
        __syncthreads();
        if (fail) { nbad = k; break; }
        const T dk = S[k][k];
        if (tid > k && tid < nb) S[k][tid] /= dk;
        __syncthreads();
        const int r = nb - k - 1;
        for (int e = tid; e < r * r; e += 256) {
            int i = k + 1 + e / r, j = k + 1 + e % r;
            if (j >= i) S[i][j] -= S[k][i] * S[k][j];
        }
        __syncthreads();
    }
    (void)nbad;
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e / nb, j = e % nb;
        if (j >= i) A[int64_t(i) * ld + j] = S[i][j];
    }
}

template <class T>
void potrf_leaf(T* A, int64_t ld, int64_t nb, int64_t col0, int64_t* info, hipStream_t s)
{
    SA_REQUIRE(nb >= 1 && nb <= 64, -2, "potrf_leaf: bad block size");
    hipLaunchKernelGGL(potrf_leaf_kernel<T>, dim3(1), dim3(256), 0, s, A, ld, int(nb), col0, info);
    SA_CHECK_LAUNCH();
}

template void potrf_leaf<double>(double*, int64_t, int64_t, int64_t, int64_t*, hipStream_t);
template void potrf_leaf<float>(float*, int64_t, int64_t, int64_t, int64_t*, hipStream_t);

//------------------------------------------------------------------------------
// In-place square transpose by 32 x 32 tile pairs (ti <= tj), one pair per block.
template <class T>
__global__ __launch_bounds__(256)
void transpose_inplace_kernel(T* __restrict__ A, int64_t n, int64_t ld, int64_t ntile)
{
    __shared__ T X[32][33];
    __shared__ T Y[32][33];
    // decode blockIdx.x -> (ti, tj), ti <= tj, enumerated column by column
    int64_t b = blockIdx.x;
    int64_t tj = int64_t((sqrt(8.0 * double(b) + 1.0) - 1.0) * 0.5);
    while ((tj + 1) * (tj + 2) / 2 <= b) ++tj;
    while (tj * (tj + 1) / 2 > b) --tj;
    int64_t ti = b - tj * (tj + 1) / 2;
    const int64_t I0 = ti * 32, J0 = tj * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    (void)ntile;
    // X[a][b] = A(I0+a, J0+b) where A(i,j) = A[i + j*ld]
    for (int bb = ty; bb < 32; bb += 8) {
        int64_t i = I0 + tx, j = J0 + bb;
        X[tx][bb] = (i < n && j < n) ? A[i + j * ld] : T(0);
    }
    if (ti != tj) {
        for (int aa = ty; aa < 32; aa += 8) {
            int64_t i = J0 + tx, j = I0 + aa;   // Y[b][a] = A(J0+b, I0+a)
            Y[tx][aa] = (i < n && j < n) ? A[i + j * ld] : T(0);
        }
    }
    __syncthreads();
    if (ti != tj) {
        // A(I0+a, J0+b) <- Y[b][a];  A(J0+b, I0+a) <- X[a][b]
        for (int bb = ty; bb < 32; bb += 8) {
            int64_t i = I0 + tx, j = J0 + bb;
            if (i < n && j < n) A[i + j * ld] = Y[bb][tx];
        }
        for (int aa = ty; aa < 32; aa += 8) {
            int64_t i = J0 + tx, j = I0 + aa;
            if (i < n && j < n) A[i + j * ld] = X[aa][tx];
        }
    }
    else {
        for (int bb = ty; bb < 32; bb += 8) {
            int64_t i = I0 + tx, j = J0 + bb;
            if (i < n && j < n) A[i + j * ld] = X[bb][tx];
        }
    }
}

template <class T>
void transpose_inplace(T* A, int64_t n, int64_t ld, hipStream_t s)
{
    if (n <= 1) return;
    int64_t nt = ceildiv(n, 32);
    int64_t nblk = nt * (nt + 1) / 2;
    SA_REQUIRE(nblk < (int64_t(1) << 31), -2, "transpose: too large");
    hipLaunchKernelGGL(transpose_inplace_kernel<T>, dim3(unsigned(nblk)), dim3(256), 0, s, A, n, ld, nt);
    SA_CHECK_LAUNCH();
}

template void transpose_inplace<double>(double*, int64_t, int64_t, hipStream_t);
template void transpose_inplace<float>(float*, int64_t, int64_t, hipStream_t);

//------------------------------------------------------------------------------
template <class T>
__global__ void set_identity_kernel(T* A, int64_t n, int64_t ld)
{
    const int64_t i = blockIdx.y;
    for (int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x; j < n; j += int64_t(gridDim.x) * 256)
        A[i * ld + j] = (i == j) ? T(1) : T(0);
}

template <class T>
void set_identity(T* A, int64_t n, int64_t ld, hipStream_t s)
{
    if (n <= 0) return;
    SA_REQUIRE(n < 65536, -2, "set_identity: too large");
    hipLaunchKernelGGL(set_identity_kernel<T>, dim3(unsigned(ceildiv(n, 256)), unsigned(n)), dim3(256), 0, s,
                       A, n, ld);
    SA_CHECK_LAUNCH();
}
template void set_identity<double>(double*, int64_t, int64_t, hipStream_t);
template void set_identity<float>(float*, int64_t, int64_t, hipStream_t);

// out (n x n, row-major, ld n) = the unit-lower triangle of the row-major
// block L (ld): strictly lower entries copied, ones on the diagonal, zeros above.
template <class T>
__global__ void unit_lower_copy_kernel(const T* __restrict__ L, int64_t ld, int64_t n, T* __restrict__ out)
{
    const int64_t i = blockIdx.y;
    for (int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x; j < n; j += int64_t(gridDim.x) * 256)
        out[i * n + j] = (j < i) ? L[i * ld + j] : (i == j ? T(1) : T(0));
}

template <class T>
void unit_lower_copy(const T* L, int64_t ld, int64_t n, T* out, hipStream_t s)
{
    if (n <= 0) return;
    SA_REQUIRE(n < 65536, -2, "unit_lower_copy: too large");
    hipLaunchKernelGGL(unit_lower_copy_kernel<T>, dim3(unsigned(ceildiv(n, 256)), unsigned(n)), dim3(256), 0, s,
                       L, ld, n, out);
    SA_CHECK_LAUNCH();
}
template void unit_lower_copy<double>(const double*, int64_t, int64_t, double*, hipStream_t);
template void unit_lower_copy<float>(const float*, int64_t, int64_t, float*, hipStream_t);

//------------------------------------------------------------------------------
template <class D, class S> __device__ inline D cvt(S x) { return D(x); }

// dst[j*ldd + i] = src[i + j*lds]: column-major src -> row-major dst.
template <class S, class D>
__global__ __launch_bounds__(256)
void transpose_convert_kernel(const S* __restrict__ src, int64_t lds, D* __restrict__ dst,
                              int64_t ldd, int64_t m, int64_t n)
{
    __shared__ D X[32][33];
    const int64_t I0 = int64_t(blockIdx.x) * 32, J0 = int64_t(blockIdx.y) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // the source is read once: nontemporal loads
    for (int bb = ty; bb < 32; bb += 8) {
        int64_t i = I0 + tx, j = J0 + bb;
        X[tx][bb] = (i < m && j < n) ? cvt<D>(__builtin_nontemporal_load(&src[i + j * lds])) : D(0);
    }
    __syncthreads();
    for (int aa = ty; aa < 32; aa += 8) {
        int64_t i = I0 + aa, j = J0 + tx;
        if (i < m && j < n) dst[i * ldd + j] = X[aa][tx];
    }
}

template <class S, class D>
void transpose_convert(const S* src, int64_t lds, D* dst, int64_t ldd, int64_t m, int64_t n,
                       hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    SA_REQUIRE(ceildiv(n, 32) < 65536, -2, "transpose_convert: too many columns");
    dim3 grid(unsigned(ceildiv(m, 32)), unsigned(ceildiv(n, 32)));
    hipLaunchKernelGGL((transpose_convert_kernel<S, D>), grid, dim3(256), 0, s, src, lds, dst, ldd, m, n);
    SA_CHECK_LAUNCH();
}

template <class S, class D>
__global__ __launch_bounds__(256)
void copy_convert_kernel(const S* __restrict__ src, int64_t lds, D* __restrict__ dst, int64_t ldd,
                         int64_t m, int64_t n)
{
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y)
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < m; i += int64_t(gridDim.x) * 256)
            dst[i + j * ldd] = cvt<D>(src[i + j * lds]);
}

template <class S, class D>
void copy_convert(const S* src, int64_t lds, D* dst, int64_t ldd, int64_t m, int64_t n,
                  hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    unsigned gx = unsigned(std::min<int64_t>(ceildiv(m, 256), 1024));
    unsigned gy = unsigned(std::min<int64_t>(n, 32768));
    hipLaunchKernelGGL((copy_convert_kernel<S, D>), dim3(gx, gy), dim3(256), 0, s,
                       src, lds, dst, ldd, m, n);
    SA_CHECK_LAUNCH();
}

#define SA_INST_CONV(S, D)                                                                       \
    template void transpose_convert<S, D>(const S*, int64_t, D*, int64_t, int64_t, int64_t,      \
                                          hipStream_t);                                         \
    template void copy_convert<S, D>(const S*, int64_t, D*, int64_t, int64_t, int64_t, hipStream_t);
SA_INST_CONV(double, double)
SA_INST_CONV(double, float)
SA_INST_CONV(float, double)
SA_INST_CONV(float, float)

} // namespace sa
