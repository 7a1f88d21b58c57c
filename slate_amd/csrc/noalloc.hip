// Device scratch for the entry points that allocate nothing
// (slate_amd_dgetrs_noalloc).
//
// Callers that cannot call hipMalloc (code under graph capture, serving loops,
// harnesses that forbid allocation) still need workspace: the pivot vector on
// the device, the interchange plan, the inverted diagonal blocks and the
// look-back exchange words of the triangular solves.  They live in one
// file-scope device array of NOALLOC_BYTES, which every GPU the library runs
// on holds until the process ends.  Host code never uses the array's name as
// an address: at set-up (slate_amd_init, with the device current) its address
// on that device is looked up once with hipGetSymbolAddress and kept per
// device; kernels receive that pointer.
//
// Every call on a device shares that device's array and finds the previous
// call's contents: whatever must start at zero is cleared by the call itself
// with a fill kernel.  Calls that overlap on one device race on the array.
#include "common.hh"
#include "kernels.hh"

#include <mutex>
#include <vector>

namespace sa {

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_noalloc_mu;
static std::vector<void*> g_noalloc_ptr;        // per device, set once

void* noalloc_scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    SA_REQUIRE(device >= 0, -1, "noalloc scratch: bad device");
    if (int(g_noalloc_ptr.size()) <= device) g_noalloc_ptr.resize(size_t(device) + 1, nullptr);
    if (!g_noalloc_ptr[size_t(device)]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(-100, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[size_t(device)] = p;
    }
    return g_noalloc_ptr[size_t(device)];
}

// ipiv0[i] = ipiv1[i] - 1 (LAPACK 1-based pivots, already range-checked on the host)
__global__ void ipiv_zero_based_kernel(const int64_t* __restrict__ ipiv1, int64_t* __restrict__ ipiv0, int64_t n)
{
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        ipiv0[i] = ipiv1[i] - 1;
}

void ipiv_zero_based(const int64_t* ipiv1, int64_t* ipiv0, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(ipiv_zero_based_kernel, dim3(unsigned(std::min<int64_t>(ceildiv(n, 256), 1024))), dim3(256),
                       0, s, ipiv1, ipiv0, n);
    SA_CHECK_LAUNCH();
}

} // namespace sa
