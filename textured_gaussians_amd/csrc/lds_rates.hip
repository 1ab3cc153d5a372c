// Measurement hook: throughput of LDS atomic adds on this GPU (no reference
// counterpart).  The texture-gradient gather keeps its per-row sums in LDS as
// float64 because, on gfx950, a float32 LDS atomic add over distinct addresses is
// far slower than a float64 one; bench.py reports these rates beside the gather.
#include "common.h"

namespace tgs {

__device__ __forceinline__ uint32_t lds_rate_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Every lane adds into `iters` pseudo-random addresses of a 40 KB LDS buffer (the size of
// a 50x50 float32 texture row).  MODE 0: float32 atomic add; 1: float64 atomic add;
// 2: uint32 atomic add; 3: float32 read, add and write (not atomic: the plain-access
// rate, for comparison).
template <int MODE>
__global__ __launch_bounds__(256) void lds_rate_kernel(int32_t iters, float *__restrict__ sink) {
    constexpr int kWords = 10000;  // 40 KB
    __shared__ __align__(16) uint32_t buf[kWords];
    for (int i = threadIdx.x; i < kWords; i += blockDim.x) buf[i] = 0u;
    __syncthreads();
    uint32_t s = lds_rate_hash(threadIdx.x * 7919u + blockIdx.x * 104729u);
    for (int32_t it = 0; it < iters; ++it) {
        s = lds_rate_hash(s + (uint32_t)it);
        if constexpr (MODE == 0) {
            atomicAdd((float *)buf + s % kWords, 1.0f);
        } else if constexpr (MODE == 1) {
            atomicAdd((double *)buf + s % (kWords / 2), 1.0);
        } else if constexpr (MODE == 2) {
            atomicAdd(buf + s % kWords, 1u);
        } else {
            float *f = (float *)buf + s % kWords;
            *f += 1.0f;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = __uint_as_float(buf[0]);
}

}  // namespace tgs

using namespace tgs;

extern "C" int tgs_lds_atomic_rates(double *lanes_per_clock_per_cu, int32_t n, tgs_stream_t stream) {
    TGS_CHECK_ARG(lanes_per_clock_per_cu && n >= 1 && n <= 4, "lds_atomic_rates: n must be 1..4");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev) != hipSuccess || cus <= 0 || khz <= 0) {
        set_error("lds_atomic_rates: cannot query the device");
        return TGS_ERR_LAUNCH;
    }
    const int blocks = cus * 4, threads = 256, iters = 2048;
    float *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TGS_OK;
    if (hipMalloc(&sink, (size_t)blocks * sizeof(float)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        set_error("lds_atomic_rates: allocation failed");
        rc = TGS_ERR_LAUNCH;
    }
    for (int m = 0; m < n && rc == TGS_OK; ++m) {
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
            if (rep == 1) (void)hipEventRecord(e0, st);
            if (m == 0) lds_rate_kernel<0><<<blocks, threads, 0, st>>>(iters, sink);
            if (m == 1) lds_rate_kernel<1><<<blocks, threads, 0, st>>>(iters, sink);
            if (m == 2) lds_rate_kernel<2><<<blocks, threads, 0, st>>>(iters, sink);
            if (m == 3) lds_rate_kernel<3><<<blocks, threads, 0, st>>>(iters, sink);
        }
        (void)hipEventRecord(e1, st);
        float ms = 0.0f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f) {
            set_error("lds_atomic_rates: kernel failed");
            rc = TGS_ERR_LAUNCH;
            break;
        }
        const double lanes = (double)blocks * threads * iters;
        lanes_per_clock_per_cu[m] = lanes / cus / ((double)ms * 1e-3 * (double)khz * 1e3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    return rc;
}
