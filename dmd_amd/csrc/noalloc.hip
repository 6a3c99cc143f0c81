// noalloc.hip -- device memory of the allocation-free component compute
// (dmdx_component_compute_noalloc in include/dmdx.h).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) compute a stateless force component with every array
// the kernels need -- system constants, term lists and parameters, term-force
// buffers, per-atom contribution lists, energy partials, PME grids,
// twiddles -- carved from one file-scope device array, `scratch`.  Each GPU
// the library runs on holds its own copy of the array for the life of the
// process; calls on one GPU share it, so two calls that overlap on one GPU
// race.  The host refers to the array only through the address that
// hipGetSymbolAddress returns for the current device.
#include <mutex>

#include "dmdx_internal.h"

namespace dmdx {

static __device__ __align__(256) unsigned char scratch[DMDX_NOALLOC_SCRATCH_BYTES];

__global__ void k_zero_words(uint32_t* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_add_energy(double* __restrict__ energy, const SysState* __restrict__ st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) energy[0] += st[0].pe;
}

hipError_t launch_zero_bytes(void* p, size_t bytes, hipStream_t s)
{
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>(1024, (n + 255) / 256);
    k_zero_words<<<g, 256, 0, s>>>((uint32_t*)p, n);
    return hipGetLastError();
}

hipError_t launch_add_energy(double* energy, const SysState* st, hipStream_t s)
{
    k_add_energy<<<1, 64, 0, s>>>(energy, st);
    return hipGetLastError();
}

namespace {
constexpr int MAX_DEVICES = 64;
NoallocDevice g_dev[MAX_DEVICES];
bool g_ready[MAX_DEVICES] = {};
std::mutex g_mu;
}  // namespace

int noalloc_device(NoallocDevice** out)
{
    int dev = 0;
    DMDX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(DMDX_E_GPU, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ready[dev]) {
        NoallocDevice& D = g_dev[dev];
        void* p = nullptr;
        // the array's name is a host-side stand-in: only this address may be
        // handed to kernels and copies
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(DMDX_E_GPU, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
        D.scratch = p;
        D.bytes = DMDX_NOALLOC_SCRATCH_BYTES;
        DMDX_HIP(hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking));
        DMDX_HIP(hipStreamCreateWithFlags(&D.side, hipStreamNonBlocking));
        for (auto& ev : D.events) DMDX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        g_ready[dev] = true;
    }
    *out = &g_dev[dev];
    return DMDX_OK;
}

}  // namespace dmdx
