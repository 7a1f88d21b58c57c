// SLATE tile storage at the boundary (gfx950).
//
// A slate::Matrix built with insertLocalTiles(Target::Devices) keeps every
// tile as an independent column-major device buffer with its own leading
// dimension (include/slate/internal/MatrixStorage.hh:151, Tile.hh:106-...);
// a HermitianMatrix stores the lower tiles only.  The drivers of this library
// factor one dense column-major array, so the tile entries gather the tiles
// into a context workspace, run the dense path, and scatter the outputs back.
// One kernel copies both ways: a workgroup per (tile, 16-column group),
// threads striding down the rows of each column (contiguous on both sides).
#include "common.hh"
#include "kernels.hh"

namespace sa {

template <class T>
__global__ __launch_bounds__(256)
void tiles_copy_kernel(T* const* __restrict__ tp, const int64_t* __restrict__ tld, int64_t m, int64_t n,
                       int64_t mb, int64_t nb, int64_t mt, T* __restrict__ D, int64_t ldd, int to_tiles)
{
    const int64_t t = blockIdx.x;
    const int64_t i = t % mt, j = t / mt;
    T* tile = tp[t];
    if (tile == nullptr) return;                    // tile not stored (upper part of a Hermitian matrix)
    const int64_t r0 = i * mb, c0 = j * nb;
    const int64_t rows = m - r0 < mb ? m - r0 : mb;
    const int64_t cols = n - c0 < nb ? n - c0 : nb;
    const int64_t ld = tld[t];
    for (int64_t cc = int64_t(blockIdx.y) * 16; cc < cols && cc < int64_t(blockIdx.y) * 16 + 16; ++cc) {
        T* d = D + r0 + (c0 + cc) * ldd;
        T* s = tile + cc * ld;
        if (to_tiles)
            for (int64_t r = threadIdx.x; r < rows; r += 256) s[r] = d[r];
        else
            for (int64_t r = threadIdx.x; r < rows; r += 256) d[r] = s[r];
    }
}

template <class T>
void tiles_copy(T* const* dtiles, const int64_t* dld, int64_t m, int64_t n, int64_t mb, int64_t nb, T* D,
                int64_t ldd, int to_tiles, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    const int64_t mt = ceildiv(m, mb), nt = ceildiv(n, nb);
    SA_REQUIRE(mt * nt < (int64_t(1) << 31) && ceildiv(nb, 16) < 65536, -2, "tiles_copy: grid too large");
    dim3 grid(unsigned(mt * nt), unsigned(ceildiv(nb, 16)));
    hipLaunchKernelGGL(tiles_copy_kernel<T>, grid, dim3(256), 0, s, dtiles, dld, m, n, mb, nb, mt, D, ldd,
                       to_tiles);
    SA_CHECK_LAUNCH();
}

template void tiles_copy<double>(double* const*, const int64_t*, int64_t, int64_t, int64_t, int64_t, double*,
                                 int64_t, int, hipStream_t);

} // namespace sa
