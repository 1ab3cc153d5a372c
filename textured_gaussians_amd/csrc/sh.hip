// Real spherical-harmonics colour evaluation (degree <= 4) and its VJP.
//
// Same basis, band order and sign convention as the reference
// (textured_gaussians/cuda/csrc/spherical_harmonics.cuh:17-111): the
// orthonormal real SH with Condon-Shortley phase, evaluated at dir/|dir|.
// One thread per direction computes the basis once for all three channels.
#include "common.h"

namespace tgs {

// Basis constants: normalisation factors of the real SH, sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) times
// the integer factors of the polynomial forms below.
#define C0 0.28209479177387814
#define C1 0.4886025119029199
#define C2a 1.0925484305920792
#define C2b 0.31539156525252005
#define C2c 0.5462742152960396
#define C3a 0.5900435899266435
#define C3b 2.890611442640554
#define C3c 0.4570457994644658
#define C3d 0.3731763325901154
#define C3e 1.445305721320277
#define C4a 2.5033429417967046
#define C4b 1.7701307697799304
#define C4c 0.9461746957575601
#define C4d 0.6690465435572892
#define C4e 0.10578554691520431
#define C4f 0.47308734787878004
#define C4g 0.6258357354491761

// Calls f(k, Y_k, dY_k/dx, dY_k/dy, dY_k/dz) for k < nb at the unit vector (x,y,z),
// with k a literal at each call, so callers index registers, not arrays.
template <typename T, typename F> __device__ __forceinline__ void sh_each(int nb, T x, T y, T z, F &&f) {
    f(0, T(C0), T(0), T(0), T(0));
    if (nb <= 1) return;
    f(1, -T(C1) * y, T(0), -T(C1), T(0));
    f(2, T(C1) * z, T(0), T(0), T(C1));
    f(3, -T(C1) * x, -T(C1), T(0), T(0));
    if (nb <= 4) return;
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    f(4, T(C2a) * x * y, T(C2a) * y, T(C2a) * x, T(0));
    f(5, -T(C2a) * y * z, T(0), -T(C2a) * z, -T(C2a) * y);
    f(6, T(C2b) * (T(3) * z2 - T(1)), T(0), T(0), T(6 * C2b) * z);
    f(7, -T(C2a) * x * z, -T(C2a) * z, T(0), -T(C2a) * x);
    f(8, T(C2c) * (x2 - y2), T(2 * C2c) * x, -T(2 * C2c) * y, T(0));
    if (nb <= 9) return;
    f(9, -T(C3a) * y * (T(3) * x2 - y2), -T(6 * C3a) * x * y, -T(3 * C3a) * (x2 - y2), T(0));
    f(10, T(C3b) * x * y * z, T(C3b) * y * z, T(C3b) * x * z, T(C3b) * x * y);
    f(11, -T(C3c) * y * (T(5) * z2 - T(1)), T(0), -T(C3c) * (T(5) * z2 - T(1)), -T(10 * C3c) * y * z);
    f(12, T(C3d) * z * (T(5) * z2 - T(3)), T(0), T(0), T(C3d) * (T(15) * z2 - T(3)));
    f(13, -T(C3c) * x * (T(5) * z2 - T(1)), -T(C3c) * (T(5) * z2 - T(1)), T(0), -T(10 * C3c) * x * z);
    f(14, T(C3e) * z * (x2 - y2), T(2 * C3e) * x * z, -T(2 * C3e) * y * z, T(C3e) * (x2 - y2));
    f(15, -T(C3a) * x * (x2 - T(3) * y2), -T(3 * C3a) * (x2 - y2), T(6 * C3a) * x * y, T(0));
    if (nb <= 16) return;
    const T q = T(7) * z2 - T(1), p = T(7) * z2 - T(3);
    f(16, T(C4a) * x * y * (x2 - y2), T(C4a) * (T(3) * x2 * y - y2 * y), T(C4a) * (x2 * x - T(3) * x * y2), T(0));
    f(17, -T(C4b) * y * z * (T(3) * x2 - y2),
      -T(6 * C4b) * x * y * z, -T(3 * C4b) * (x2 - y2) * z, -T(C4b) * (T(3) * x2 * y - y2 * y));
    f(18, T(C4c) * x * y * q, T(C4c) * y * q, T(C4c) * x * q, T(14 * C4c) * x * y * z);
    f(19, -T(C4d) * y * z * p, T(0), -T(C4d) * z * p, -T(C4d) * y * (T(21) * z2 - T(3)));
    f(20, T(C4e) * (T(35) * z2 * z2 - T(30) * z2 + T(3)), T(0), T(0), T(C4e) * (T(140) * z2 * z - T(60) * z));
    f(21, -T(C4d) * x * z * p, -T(C4d) * z * p, T(0), -T(C4d) * x * (T(21) * z2 - T(3)));
    f(22, T(C4f) * (x2 - y2) * q, T(2 * C4f) * x * q, -T(2 * C4f) * y * q, T(14 * C4f) * z * (x2 - y2));
    f(23, -T(C4b) * x * z * (x2 - T(3) * y2),
      -T(3 * C4b) * (x2 - y2) * z, T(6 * C4b) * x * y * z, -T(C4b) * (x2 * x - T(3) * x * y2));
    f(24, T(C4g) * (x2 * x2 - T(6) * x2 * y2 + y2 * y2),
      T(4 * C4g) * (x2 * x - T(3) * x * y2), T(4 * C4g) * (y2 * y - T(3) * x2 * y), T(0));
}

template <typename T>
__global__ __launch_bounds__(256) void sh_fwd_kernel(int64_t n, int K, int degree, const T *__restrict__ dirs,
                                                     const T *__restrict__ coeffs, const uint8_t *__restrict__ masks,
                                                     T *__restrict__ colors) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (masks && !masks[i]) {
        colors[i * 3] = 0; colors[i * 3 + 1] = 0; colors[i * 3 + 2] = 0;
        return;
    }
    const int nb = (degree + 1) * (degree + 1);
    T x = 0, y = 0, z = 0;
    if (nb > 1) {
        const T dx = dirs[i * 3], dy = dirs[i * 3 + 1], dz = dirs[i * 3 + 2];
        const T inv = tgs_rsqrt<T>(dx * dx + dy * dy + dz * dz);
        x = dx * inv; y = dy * inv; z = dz * inv;
    }
    const T *c = coeffs + i * (int64_t)K * 3;
    T r = 0, g = 0, b = 0;
    sh_each<T>(nb, x, y, z, [&](int k, T v, T, T, T) {
        r += v * c[k * 3];
        g += v * c[k * 3 + 1];
        b += v * c[k * 3 + 2];
    });
    colors[i * 3] = r; colors[i * 3 + 1] = g; colors[i * 3 + 2] = b;
}

template <typename T>
__global__ __launch_bounds__(256) void sh_bwd_kernel(int64_t n, int K, int degree, const T *__restrict__ dirs,
                                                     const T *__restrict__ coeffs, const uint8_t *__restrict__ masks,
                                                     const T *__restrict__ v_colors, T *__restrict__ v_coeffs,
                                                     T *__restrict__ v_dirs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T *vc = v_coeffs + i * (int64_t)K * 3;
    if (masks && !masks[i]) {
        for (int k = 0; k < K * 3; ++k) vc[k] = 0;
        return;
    }
    const int nb = (degree + 1) * (degree + 1);
    T x = 0, y = 0, z = 0, inv = 0;
    if (nb > 1) {
        const T dx = dirs[i * 3], dy = dirs[i * 3 + 1], dz = dirs[i * 3 + 2];
        inv = tgs_rsqrt<T>(dx * dx + dy * dy + dz * dz);
        x = dx * inv; y = dy * inv; z = dz * inv;
    }
    const bool want_dir = v_dirs != nullptr && nb > 1;
    const T vr = v_colors[i * 3], vg = v_colors[i * 3 + 1], vb = v_colors[i * 3 + 2];
    const T *c = coeffs + i * (int64_t)K * 3;
    T gu = 0, gv = 0, gw = 0;
    sh_each<T>(nb, x, y, z, [&](int k, T v, T gx, T gy, T gz) {  // nb <= K (checked by the entry point)
        vc[k * 3] = v * vr;
        vc[k * 3 + 1] = v * vg;
        vc[k * 3 + 2] = v * vb;
        if (want_dir) {
            const T s = c[k * 3] * vr + c[k * 3 + 1] * vg + c[k * 3 + 2] * vb;
            gu += s * gx;
            gv += s * gy;
            gw += s * gz;
        }
    });
    for (int k = nb; k < K; ++k) {
        vc[k * 3] = 0; vc[k * 3 + 1] = 0; vc[k * 3 + 2] = 0;
    }
    if (want_dir) {
        // through u = d / |d|: dL/dd = (g - (g.u) u) / |d|
        const T gd = gu * x + gv * y + gw * z;
        // one lane per direction: stored (the caller zero-fills v_dirs)
        v_dirs[i * 3] = (gu - gd * x) * inv;
        v_dirs[i * 3 + 1] = (gv - gd * y) * inv;
        v_dirs[i * 3 + 2] = (gw - gd * z) * inv;
    }
}

// float32 with K * 3 <= 48 (SH degree 3): the same per-Gaussian arithmetic with the
// wave's coefficient block (64 Gaussians x K * 3 values, contiguous in memory)
// loaded and its v_coeffs block stored through LDS, so global accesses are
// contiguous 256-byte runs instead of 64 lines 192 bytes apart.  Rows are padded
// to an odd stride (49) against bank conflicts.  The basis size NB is a template
// argument, so the basis and its gradient stay in registers (indexed by unrolled
// loops) instead of private memory.
constexpr int kShStage = 49;

// A wave's coefficient block (the tot = nw * K3 values of nw consecutive Gaussians,
// contiguous at src) into st[t * kShStage + k] and back: element e is (e / K3, e % K3).
// With vec (K3 a multiple of 4, the block 16-byte aligned) it moves as 16-byte vectors,
// whose four values then share a Gaussian.
__device__ __forceinline__ void sh_stage_in(float *st, const float *__restrict__ src, int64_t tot, int K3, bool vec,
                                            int lane) {
    const int64_t tot4 = vec ? tot / 4 : 0;
    for (int64_t e4 = lane; e4 < tot4; e4 += 64) {
        const float4 v = ((const float4 *)src)[e4];
        const int t = (int)(e4 * 4 / K3), k = (int)(e4 * 4 - (int64_t)t * K3);
        float *d = st + t * kShStage + k;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    for (int64_t e = tot4 * 4 + lane; e < tot; e += 64) {
        const int t = (int)(e / K3), k = (int)(e - (int64_t)t * K3);
        st[t * kShStage + k] = src[e];
    }
}

__device__ __forceinline__ void sh_stage_out(float *__restrict__ dst, const float *st, int64_t tot, int K3, bool vec,
                                             int lane) {
    const int64_t tot4 = vec ? tot / 4 : 0;
    for (int64_t e4 = lane; e4 < tot4; e4 += 64) {
        const int t = (int)(e4 * 4 / K3), k = (int)(e4 * 4 - (int64_t)t * K3);
        const float *q = st + t * kShStage + k;
        ((float4 *)dst)[e4] = float4{q[0], q[1], q[2], q[3]};
    }
    for (int64_t e = tot4 * 4 + lane; e < tot; e += 64) {
        const int t = (int)(e / K3), k = (int)(e - (int64_t)t * K3);
        dst[e] = st[t * kShStage + k];
    }
}

template <int NB>
__global__ __launch_bounds__(256) void sh_bwd_staged_kernel(int64_t n, int K, const float *__restrict__ dirs,
                                                            const float *__restrict__ coeffs,
                                                            const uint8_t *__restrict__ masks,
                                                            const float *__restrict__ v_colors,
                                                            float *__restrict__ v_coeffs, float *__restrict__ v_dirs) {
    __shared__ float sm[4][64 * kShStage];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + wv * 64;
    if (base >= n) return;  // uniform over the wave
    const int K3 = K * 3;
    const int nw = (int)min((int64_t)64, n - base);
    const int64_t tot = (int64_t)nw * K3;
    float *st = sm[wv];
    // 16-byte moves when both buffers are 16-byte aligned (then so is every block: base % 64 == 0)
    const bool vec = K3 % 4 == 0 && (((uintptr_t)coeffs | (uintptr_t)v_coeffs) & 15) == 0;
    sh_stage_in(st, coeffs + base * K3, tot, K3, vec, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int64_t i = base + lane;
    if (i < n) {
        float *c = st + lane * kShStage;  // coefficients in, v_coeffs out (same slots)
        if (masks && !masks[i]) {
            for (int k = 0; k < K3; ++k) c[k] = 0.f;
        } else {
            float x = 0, y = 0, z = 0, inv = 0;
            if (NB > 1) {
                const float dx = dirs[i * 3], dy = dirs[i * 3 + 1], dz = dirs[i * 3 + 2];
                inv = tgs_rsqrt<float>(dx * dx + dy * dy + dz * dz);
                x = dx * inv; y = dy * inv; z = dz * inv;
            }
            const bool want_dir = v_dirs != nullptr && NB > 1;
            const float vr = v_colors[i * 3], vg = v_colors[i * 3 + 1], vb = v_colors[i * 3 + 2];
            float gu = 0, gv = 0, gw = 0;
            // NB <= K (checked by the entry point); c[k] is read before it is overwritten
            if (want_dir) {
                sh_each<float>(NB, x, y, z, [&](int k, float v, float gx, float gy, float gz) {
                    const float s = c[k * 3] * vr + c[k * 3 + 1] * vg + c[k * 3 + 2] * vb;
                    gu += s * gx;
                    gv += s * gy;
                    gw += s * gz;
                    c[k * 3] = v * vr;
                    c[k * 3 + 1] = v * vg;
                    c[k * 3 + 2] = v * vb;
                });
            } else {
                sh_each<float>(NB, x, y, z, [&](int k, float v, float, float, float) {
                    c[k * 3] = v * vr;
                    c[k * 3 + 1] = v * vg;
                    c[k * 3 + 2] = v * vb;
                });
            }
            for (int k = NB; k < K; ++k) {
                c[k * 3] = 0.f; c[k * 3 + 1] = 0.f; c[k * 3 + 2] = 0.f;
            }
            if (want_dir) {
                const float gd = gu * x + gv * y + gw * z;
                v_dirs[i * 3] = (gu - gd * x) * inv;  // one lane per direction
                v_dirs[i * 3 + 1] = (gv - gd * y) * inv;
                v_dirs[i * 3 + 2] = (gw - gd * z) * inv;
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    sh_stage_out(v_coeffs + base * K3, st, tot, K3, vec, lane);
}


// Several cameras over one set of coefficients, float32, K * 3 <= 48: the grid layout
// with C > 1 views.  The reference expands the coefficients to [C, N, K, 3] (a copy)
// and evaluates C * N rows; here a lane owns one Gaussian, stages its coefficients
// once (as the staged backward does) and loops over the cameras.  dirs, masks,
// colors, v_colors and v_dirs are [C, N, ...]; coeffs and v_coeffs [N, K, 3].
// Per (camera, Gaussian) the arithmetic is that of sh_fwd_kernel / the staged
// backward; v_coeffs sums the cameras in order 0..C-1.
template <int NB>
__global__ __launch_bounds__(256) void sh_fwd_cams_kernel(int64_t C, int64_t n, int K, const float *__restrict__ dirs,
                                                          const float *__restrict__ coeffs,
                                                          const uint8_t *__restrict__ masks,
                                                          float *__restrict__ colors) {
    __shared__ float sm[4][64 * kShStage];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + wv * 64;
    if (base >= n) return;  // uniform over the wave
    const int K3 = K * 3;
    const int nw = (int)min((int64_t)64, n - base);
    float *st = sm[wv];
    sh_stage_in(st, coeffs + base * K3, (int64_t)nw * K3, K3, K3 % 4 == 0 && ((uintptr_t)coeffs & 15) == 0, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int64_t i = base + lane;
    if (i >= n) return;
    const float *c = st + lane * kShStage;
    for (int64_t cam = 0; cam < C; ++cam) {
        const int64_t ci = cam * n + i;
        float r = 0, g = 0, b = 0;
        if (!masks || masks[ci]) {
            float x = 0, y = 0, z = 0;
            if (NB > 1) {
                const float dx = dirs[ci * 3], dy = dirs[ci * 3 + 1], dz = dirs[ci * 3 + 2];
                const float inv = tgs_rsqrt<float>(dx * dx + dy * dy + dz * dz);
                x = dx * inv; y = dy * inv; z = dz * inv;
            }
            sh_each<float>(NB, x, y, z, [&](int k, float v, float, float, float) {
                r += v * c[k * 3];
                g += v * c[k * 3 + 1];
                b += v * c[k * 3 + 2];
            });
        }
        colors[ci * 3] = r; colors[ci * 3 + 1] = g; colors[ci * 3 + 2] = b;
    }
}

// v_dirs (optional) is written in full, zero where masked: one lane per (camera, Gaussian).
template <int NB>
__global__ __launch_bounds__(256) void sh_bwd_cams_kernel(int64_t C, int64_t n, int K, const float *__restrict__ dirs,
                                                          const float *__restrict__ coeffs,
                                                          const uint8_t *__restrict__ masks,
                                                          const float *__restrict__ v_colors,
                                                          float *__restrict__ v_coeffs, float *__restrict__ v_dirs) {
    __shared__ float sm[4][64 * kShStage];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + wv * 64;
    if (base >= n) return;  // uniform over the wave
    const int K3 = K * 3;
    const int nw = (int)min((int64_t)64, n - base);
    const int64_t tot = (int64_t)nw * K3;
    float *st = sm[wv];
    const bool vec = K3 % 4 == 0 && (((uintptr_t)coeffs | (uintptr_t)v_coeffs) & 15) == 0;
    sh_stage_in(st, coeffs + base * K3, tot, K3, vec, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int64_t i = base + lane;
    if (i < n) {
        float *c = st + lane * kShStage;  // coefficients in, v_coeffs out (same slots)
        float acc[NB * 3];
#pragma unroll
        for (int q = 0; q < NB * 3; ++q) acc[q] = 0.f;
        const bool want_dir = v_dirs != nullptr && NB > 1;
        for (int64_t cam = 0; cam < C; ++cam) {
            const int64_t ci = cam * n + i;
            if (masks && !masks[ci]) {
                if (v_dirs) { v_dirs[ci * 3] = 0.f; v_dirs[ci * 3 + 1] = 0.f; v_dirs[ci * 3 + 2] = 0.f; }
                continue;
            }
            float x = 0, y = 0, z = 0, inv = 0;
            if (NB > 1) {
                const float dx = dirs[ci * 3], dy = dirs[ci * 3 + 1], dz = dirs[ci * 3 + 2];
                inv = tgs_rsqrt<float>(dx * dx + dy * dy + dz * dz);
                x = dx * inv; y = dy * inv; z = dz * inv;
            }
            const float vr = v_colors[ci * 3], vg = v_colors[ci * 3 + 1], vb = v_colors[ci * 3 + 2];
            float gu = 0, gv = 0, gw = 0;
            sh_each<float>(NB, x, y, z, [&](int k, float v, float gx, float gy, float gz) {
                if (want_dir) {
                    const float s = c[k * 3] * vr + c[k * 3 + 1] * vg + c[k * 3 + 2] * vb;
                    gu += s * gx;
                    gv += s * gy;
                    gw += s * gz;
                }
                acc[k * 3] += v * vr;
                acc[k * 3 + 1] += v * vg;
                acc[k * 3 + 2] += v * vb;
            });
            if (want_dir) {
                const float gd = gu * x + gv * y + gw * z;
                v_dirs[ci * 3] = (gu - gd * x) * inv;
                v_dirs[ci * 3 + 1] = (gv - gd * y) * inv;
                v_dirs[ci * 3 + 2] = (gw - gd * z) * inv;
            } else if (v_dirs) {
                v_dirs[ci * 3] = 0.f; v_dirs[ci * 3 + 1] = 0.f; v_dirs[ci * 3 + 2] = 0.f;
            }
        }
#pragma unroll
        for (int q = 0; q < NB * 3; ++q) c[q] = acc[q];  // NB <= K (checked by the entry point)
        for (int q = NB * 3; q < K3; ++q) c[q] = 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    sh_stage_out(v_coeffs + base * K3, st, tot, K3, vec, lane);
}

}  // namespace tgs

using namespace tgs;

extern "C" int tgs_sh_fwd(int dtype, int64_t n, int32_t K, int32_t degree, const void *dirs, const void *coeffs,
                          const uint8_t *masks, void *colors, tgs_stream_t stream) {
    TGS_CHECK_ARG(degree >= 0 && degree <= 4, "sh_fwd: degree %d not in [0, 4]", degree);
    TGS_CHECK_ARG(K >= (degree + 1) * (degree + 1), "sh_fwd: K=%d < (degree+1)^2", K);
    if (n <= 0) return TGS_OK;
    const unsigned nb = (unsigned)ceil_div(n, 256);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32)
        sh_fwd_kernel<float><<<nb, 256, 0, st>>>(n, K, degree, (const float *)dirs, (const float *)coeffs, masks,
                                                 (float *)colors);
    else if (dtype == TGS_F64)
        sh_fwd_kernel<double><<<nb, 256, 0, st>>>(n, K, degree, (const double *)dirs, (const double *)coeffs, masks,
                                                  (double *)colors);
    else {
        set_error("sh_fwd: unsupported dtype %d", dtype);
        return TGS_ERR_UNSUPPORTED;
    }
    TGS_CHECK_LAUNCH("sh_fwd");
    return TGS_OK;
}

extern "C" int tgs_sh_bwd(int dtype, int64_t n, int32_t K, int32_t degree, const void *dirs, const void *coeffs,
                          const uint8_t *masks, const void *v_colors, void *v_coeffs, void *v_dirs,
                          tgs_stream_t stream) {
    TGS_CHECK_ARG(degree >= 0 && degree <= 4, "sh_bwd: degree %d not in [0, 4]", degree);
    TGS_CHECK_ARG(K >= (degree + 1) * (degree + 1), "sh_bwd: K=%d < (degree+1)^2", K);
    if (n <= 0) return TGS_OK;
    const unsigned nb = (unsigned)ceil_div(n, 256);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32 && K * 3 <= kShStage - 1) {
#define TGS_SHB(NBV)                                                                                              \
    sh_bwd_staged_kernel<NBV><<<nb, 256, 0, st>>>(n, K, (const float *)dirs, (const float *)coeffs, masks,        \
                                                  (const float *)v_colors, (float *)v_coeffs, (float *)v_dirs)
        switch (degree) {
            case 0: TGS_SHB(1); break;
            case 1: TGS_SHB(4); break;
            case 2: TGS_SHB(9); break;
            case 3: TGS_SHB(16); break;
            default: TGS_SHB(25); break;
        }
#undef TGS_SHB
    }
    else if (dtype == TGS_F32)
        sh_bwd_kernel<float><<<nb, 256, 0, st>>>(n, K, degree, (const float *)dirs, (const float *)coeffs, masks,
                                                 (const float *)v_colors, (float *)v_coeffs, (float *)v_dirs);
    else if (dtype == TGS_F64)
        sh_bwd_kernel<double><<<nb, 256, 0, st>>>(n, K, degree, (const double *)dirs, (const double *)coeffs, masks,
                                                  (const double *)v_colors, (double *)v_coeffs, (double *)v_dirs);
    else {
        set_error("sh_bwd: unsupported dtype %d", dtype);
        return TGS_ERR_UNSUPPORTED;
    }
    TGS_CHECK_LAUNCH("sh_bwd");
    return TGS_OK;
}

// degree -> basis size template argument
#define TGS_SH_NB_SWITCH(degree, M)    \
    switch (degree) {                  \
        case 0: M(1); break;           \
        case 1: M(4); break;           \
        case 2: M(9); break;           \
        case 3: M(16); break;          \
        default: M(25); break;         \
    }

extern "C" int tgs_sh_fwd_cams(int dtype, int64_t C, int64_t n, int32_t K, int32_t degree, const void *dirs,
                               const void *coeffs, const uint8_t *masks, void *colors, tgs_stream_t stream) {
    TGS_CHECK_ARG(degree >= 0 && degree <= 4, "sh_fwd_cams: degree %d not in [0, 4]", degree);
    TGS_CHECK_ARG(K >= (degree + 1) * (degree + 1), "sh_fwd_cams: K=%d < (degree+1)^2", K);
    TGS_CHECK_ARG(C >= 0 && n >= 0, "sh_fwd_cams: bad sizes C=%lld n=%lld", (long long)C, (long long)n);
    if (dtype != TGS_F32 || K * 3 > kShStage - 1) {
        set_error("sh_fwd_cams: float32 with K * 3 <= %d only (got dtype %d, K=%d)", kShStage - 1, dtype, K);
        return TGS_ERR_UNSUPPORTED;
    }
    if (n == 0 || C == 0) return TGS_OK;
    TGS_CHECK_ARG(dirs && coeffs && colors, "sh_fwd_cams: null buffer");
    const unsigned nb = (unsigned)ceil_div(n, 256);
    hipStream_t st = (hipStream_t)stream;
#define TGS_SHF(NBV)                                                                                     \
    sh_fwd_cams_kernel<NBV><<<nb, 256, 0, st>>>(C, n, K, (const float *)dirs, (const float *)coeffs, masks, \
                                                (float *)colors)
    TGS_SH_NB_SWITCH(degree, TGS_SHF)
#undef TGS_SHF
    TGS_CHECK_LAUNCH("sh_fwd_cams");
    return TGS_OK;
}

extern "C" int tgs_sh_bwd_cams(int dtype, int64_t C, int64_t n, int32_t K, int32_t degree, const void *dirs,
                               const void *coeffs, const uint8_t *masks, const void *v_colors, void *v_coeffs,
                               void *v_dirs, tgs_stream_t stream) {
    TGS_CHECK_ARG(degree >= 0 && degree <= 4, "sh_bwd_cams: degree %d not in [0, 4]", degree);
    TGS_CHECK_ARG(K >= (degree + 1) * (degree + 1), "sh_bwd_cams: K=%d < (degree+1)^2", K);
    TGS_CHECK_ARG(C >= 0 && n >= 0, "sh_bwd_cams: bad sizes C=%lld n=%lld", (long long)C, (long long)n);
    if (dtype != TGS_F32 || K * 3 > kShStage - 1) {
        set_error("sh_bwd_cams: float32 with K * 3 <= %d only (got dtype %d, K=%d)", kShStage - 1, dtype, K);
        return TGS_ERR_UNSUPPORTED;
    }
    if (n == 0) return TGS_OK;
    TGS_CHECK_ARG(dirs && coeffs && v_colors && v_coeffs, "sh_bwd_cams: null buffer");
    const unsigned nb = (unsigned)ceil_div(n, 256);
    hipStream_t st = (hipStream_t)stream;
#define TGS_SHB(NBV)                                                                                      \
    sh_bwd_cams_kernel<NBV><<<nb, 256, 0, st>>>(C, n, K, (const float *)dirs, (const float *)coeffs, masks,  \
                                                (const float *)v_colors, (float *)v_coeffs, (float *)v_dirs)
    TGS_SH_NB_SWITCH(degree, TGS_SHB)
#undef TGS_SHB
    TGS_CHECK_LAUNCH("sh_bwd_cams");
    return TGS_OK;
}
