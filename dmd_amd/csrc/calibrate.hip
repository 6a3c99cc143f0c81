// calibrate.hip -- a plain device-to-device copy kernel of known size.
//
// Used to calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE read by
// rocprofv3 for this kernel against its known traffic) and to report the
// HBM bandwidth reachable on the device.
#include "engine.h"

namespace dmdx {

__global__ void k_copy_calibration(const double4* __restrict__ src, double4* __restrict__ dst,
                                   int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace dmdx

extern "C" int dmdx_copy_bandwidth(int64_t nbytes, int32_t repeats, double* gbps)
{
    using namespace dmdx;
    if (nbytes < 4096 || repeats < 1 || !gbps) return fail(DMDX_E_INVALID, "bad arguments");
    int64_t n = nbytes / (int64_t)sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    DMDX_HIP(hipMalloc((void**)&a, n * sizeof(double4)));
    hipError_t e = hipMalloc((void**)&b, n * sizeof(double4));
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(DMDX_E_OOM, "copy buffers");
    }
    (void)hipMemset(a, 0, n * sizeof(double4));
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    unsigned blocks = 256 * 8 * 4;
    k_copy_calibration<<<blocks, 256>>>(a, b, n);
    (void)hipEventRecord(t0, nullptr);
    for (int r = 0; r < repeats; ++r) k_copy_calibration<<<blocks, 256>>>(a, b, n);
    (void)hipEventRecord(t1, nullptr);
    e = hipEventSynchronize(t1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(DMDX_E_GPU, hipGetErrorString(e));
    *gbps = 2.0 * (double)n * sizeof(double4) * repeats / (ms * 1e-3) / 1e9;
    return DMDX_OK;
}
