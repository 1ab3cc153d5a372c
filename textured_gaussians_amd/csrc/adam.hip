// Fused selective Adam step (textured_gaussians/cuda/csrc/adam.cu:16-47 semantics:
// element-wise Adam without bias correction on rows whose visibility flag is
// set).  Each thread updates four consecutive elements of one row with 16-byte
// loads and stores when the row length allows it.
#include "common.h"

namespace tgs {

__device__ __forceinline__ void adam_elem(float &p, float g, float &m, float &v, float lr, float b1, float b2,
                                          float eps) {
    m = b1 * m + (1.0f - b1) * g;
    v = b2 * v + (1.0f - b2) * g * g;
    p += -lr * m / (sqrtf(v) + eps);
}

__global__ __launch_bounds__(256) void adam_vec4_kernel(int64_t n4, int64_t M4, float4 *__restrict__ param,
                                                        const float4 *__restrict__ grad, float4 *__restrict__ m,
                                                        float4 *__restrict__ v, const uint8_t *__restrict__ vis,
                                                        float lr, float b1, float b2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4 || !vis[i / M4]) return;
    float4 p = param[i], g = grad[i], mm = m[i], vv = v[i];
    adam_elem(p.x, g.x, mm.x, vv.x, lr, b1, b2, eps);
    adam_elem(p.y, g.y, mm.y, vv.y, lr, b1, b2, eps);
    adam_elem(p.z, g.z, mm.z, vv.z, lr, b1, b2, eps);
    adam_elem(p.w, g.w, mm.w, vv.w, lr, b1, b2, eps);
    param[i] = p;
    m[i] = mm;
    v[i] = vv;
}

__global__ __launch_bounds__(256) void adam_kernel(int64_t n, int64_t M, float *__restrict__ param,
                                                   const float *__restrict__ grad, float *__restrict__ m,
                                                   float *__restrict__ v, const uint8_t *__restrict__ vis, float lr,
                                                   float b1, float b2, float eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !vis[i / M]) return;
    float p = param[i], mm = m[i], vv = v[i];
    adam_elem(p, grad[i], mm, vv, lr, b1, b2, eps);
    param[i] = p;
    m[i] = mm;
    v[i] = vv;
}

}  // namespace tgs

using namespace tgs;

extern "C" int tgs_selective_adam_update(int64_t N, int64_t M, float *param, const float *grad, float *exp_avg,
                                         float *exp_avg_sq, const uint8_t *visibility, double lr, double b1,
                                         double b2, double eps, tgs_stream_t stream) {
    TGS_CHECK_ARG(N >= 0 && M >= 0, "selective_adam_update: bad sizes");
    const int64_t n = N * M;
    if (n == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    const bool aligned = ((uintptr_t)param | (uintptr_t)grad | (uintptr_t)exp_avg | (uintptr_t)exp_avg_sq) % 16 == 0;
    if (M % 4 == 0 && aligned) {
        const int64_t n4 = n / 4;
        adam_vec4_kernel<<<(unsigned)ceil_div(n4, 256), 256, 0, st>>>(
            n4, M / 4, (float4 *)param, (const float4 *)grad, (float4 *)exp_avg, (float4 *)exp_avg_sq, visibility,
            (float)lr, (float)b1, (float)b2, (float)eps);
    } else {
        adam_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, M, param, grad, exp_avg, exp_avg_sq, visibility,
                                                                 (float)lr, (float)b1, (float)b2, (float)eps);
    }
    TGS_CHECK_LAUNCH("selective_adam_update");
    return TGS_OK;
}
