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

// Y[k] and, if G != nullptr, dY[k]/d(x,y,z) in G[k][0..2], at the unit vector (x,y,z).
template <typename T> __device__ __forceinline__ void sh_eval(int nb, T x, T y, T z, T *Y, T (*G)[3]) {
    Y[0] = T(C0);
    if (G) { G[0][0] = 0; G[0][1] = 0; G[0][2] = 0; }
    if (nb <= 1) return;
    Y[1] = -T(C1) * y; Y[2] = T(C1) * z; Y[3] = -T(C1) * x;
    if (G) {
        G[1][0] = 0; G[1][1] = -T(C1); G[1][2] = 0;
        G[2][0] = 0; G[2][1] = 0; G[2][2] = T(C1);
        G[3][0] = -T(C1); G[3][1] = 0; G[3][2] = 0;
    }
    if (nb <= 4) return;
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    Y[4] = T(C2a) * x * y;
    Y[5] = -T(C2a) * y * z;
    Y[6] = T(C2b) * (T(3) * z2 - T(1));
    Y[7] = -T(C2a) * x * z;
    Y[8] = T(C2c) * (x2 - y2);
    if (G) {
        G[4][0] = T(C2a) * y; G[4][1] = T(C2a) * x; G[4][2] = 0;
        G[5][0] = 0; G[5][1] = -T(C2a) * z; G[5][2] = -T(C2a) * y;
        G[6][0] = 0; G[6][1] = 0; G[6][2] = T(6 * C2b) * z;
        G[7][0] = -T(C2a) * z; G[7][1] = 0; G[7][2] = -T(C2a) * x;
        G[8][0] = T(2 * C2c) * x; G[8][1] = -T(2 * C2c) * y; G[8][2] = 0;
    }
    if (nb <= 9) return;
    Y[9] = -T(C3a) * y * (T(3) * x2 - y2);
    Y[10] = T(C3b) * x * y * z;
    Y[11] = -T(C3c) * y * (T(5) * z2 - T(1));
    Y[12] = T(C3d) * z * (T(5) * z2 - T(3));
    Y[13] = -T(C3c) * x * (T(5) * z2 - T(1));
    Y[14] = T(C3e) * z * (x2 - y2);
    Y[15] = -T(C3a) * x * (x2 - T(3) * y2);
    if (G) {
        G[9][0] = -T(6 * C3a) * x * y; G[9][1] = -T(3 * C3a) * (x2 - y2); G[9][2] = 0;
        G[10][0] = T(C3b) * y * z; G[10][1] = T(C3b) * x * z; G[10][2] = T(C3b) * x * y;
        G[11][0] = 0; G[11][1] = -T(C3c) * (T(5) * z2 - T(1)); G[11][2] = -T(10 * C3c) * y * z;
        G[12][0] = 0; G[12][1] = 0; G[12][2] = T(C3d) * (T(15) * z2 - T(3));
        G[13][0] = -T(C3c) * (T(5) * z2 - T(1)); G[13][1] = 0; G[13][2] = -T(10 * C3c) * x * z;
        G[14][0] = T(2 * C3e) * x * z; G[14][1] = -T(2 * C3e) * y * z; G[14][2] = T(C3e) * (x2 - y2);
        G[15][0] = -T(3 * C3a) * (x2 - y2); G[15][1] = T(6 * C3a) * x * y; G[15][2] = 0;
    }
    if (nb <= 16) return;
    const T q = T(7) * z2 - T(1), p = T(7) * z2 - T(3);
    Y[16] = T(C4a) * x * y * (x2 - y2);
    Y[17] = -T(C4b) * y * z * (T(3) * x2 - y2);
    Y[18] = T(C4c) * x * y * q;
    Y[19] = -T(C4d) * y * z * p;
    Y[20] = T(C4e) * (T(35) * z2 * z2 - T(30) * z2 + T(3));
    Y[21] = -T(C4d) * x * z * p;
    Y[22] = T(C4f) * (x2 - y2) * q;
    Y[23] = -T(C4b) * x * z * (x2 - T(3) * y2);
    Y[24] = T(C4g) * (x2 * x2 - T(6) * x2 * y2 + y2 * y2);
    if (G) {
        G[16][0] = T(C4a) * (T(3) * x2 * y - y2 * y); G[16][1] = T(C4a) * (x2 * x - T(3) * x * y2); G[16][2] = 0;
        G[17][0] = -T(6 * C4b) * x * y * z; G[17][1] = -T(3 * C4b) * (x2 - y2) * z;
        G[17][2] = -T(C4b) * (T(3) * x2 * y - y2 * y);
        G[18][0] = T(C4c) * y * q; G[18][1] = T(C4c) * x * q; G[18][2] = T(14 * C4c) * x * y * z;
        G[19][0] = 0; G[19][1] = -T(C4d) * z * p; G[19][2] = -T(C4d) * y * (T(21) * z2 - T(3));
        G[20][0] = 0; G[20][1] = 0; G[20][2] = T(C4e) * (T(140) * z2 * z - T(60) * z);
        G[21][0] = -T(C4d) * z * p; G[21][1] = 0; G[21][2] = -T(C4d) * x * (T(21) * z2 - T(3));
        G[22][0] = T(2 * C4f) * x * q; G[22][1] = -T(2 * C4f) * y * q; G[22][2] = T(14 * C4f) * z * (x2 - y2);
        G[23][0] = -T(3 * C4b) * (x2 - y2) * z; G[23][1] = T(6 * C4b) * x * y * z;
        G[23][2] = -T(C4b) * (x2 * x - T(3) * x * y2);
        G[24][0] = T(4 * C4g) * (x2 * x - T(3) * x * y2); G[24][1] = T(4 * C4g) * (y2 * y - T(3) * x2 * y);
        G[24][2] = 0;
    }
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
    T Y[25];
    sh_eval<T>(nb, x, y, z, Y, nullptr);
    const T *c = coeffs + i * (int64_t)K * 3;
    T r = 0, g = 0, b = 0;
    for (int k = 0; k < nb; ++k) {
        r += Y[k] * c[k * 3];
        g += Y[k] * c[k * 3 + 1];
        b += Y[k] * c[k * 3 + 2];
    }
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
    T Y[25], G[25][3];
    const bool want_dir = v_dirs != nullptr && nb > 1;
    sh_eval<T>(nb, x, y, z, Y, want_dir ? G : nullptr);
    const T vr = v_colors[i * 3], vg = v_colors[i * 3 + 1], vb = v_colors[i * 3 + 2];
    const T *c = coeffs + i * (int64_t)K * 3;
    T gu = 0, gv = 0, gw = 0;
    for (int k = 0; k < K; ++k) {
        if (k < nb) {
            vc[k * 3] = Y[k] * vr;
            vc[k * 3 + 1] = Y[k] * vg;
            vc[k * 3 + 2] = Y[k] * vb;
            if (want_dir) {
                const T s = c[k * 3] * vr + c[k * 3 + 1] * vg + c[k * 3 + 2] * vb;
                gu += s * G[k][0];
                gv += s * G[k][1];
                gw += s * G[k][2];
            }
        } else {
            vc[k * 3] = 0; vc[k * 3 + 1] = 0; vc[k * 3 + 2] = 0;
        }
    }
    if (want_dir) {
        // through u = d / |d|: dL/dd = (g - (g.u) u) / |d|
        const T gd = gu * x + gv * y + gw * z;
        atomic_add(v_dirs + i * 3, (gu - gd * x) * inv);
        atomic_add(v_dirs + i * 3 + 1, (gv - gd * y) * inv);
        atomic_add(v_dirs + i * 3 + 2, (gw - gd * z) * inv);
    }
}

// float32 with K * 3 <= 48 (SH degree 3): the same per-Gaussian arithmetic with the
// wave's coefficient block (64 Gaussians x K * 3 values, contiguous in memory)
// loaded and its v_coeffs block stored through LDS, so global accesses are
// contiguous 256-byte runs instead of 64 lines 192 bytes apart.  Rows are padded
// to an odd stride (49) against bank conflicts.
constexpr int kShStage = 49;

__global__ __launch_bounds__(256) void sh_bwd_staged_kernel(int64_t n, int K, int degree, const float *__restrict__ dirs,
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
    {
        int t = lane / K3, k = lane - (lane / K3) * K3;  // element lane + 64 j of the block = (t, k)
        for (int64_t e = lane; e < tot; e += 64) {
            st[t * kShStage + k] = coeffs[base * K3 + e];
            k += 64;
            while (k >= K3) {
                k -= K3;
                ++t;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int64_t i = base + lane;
    if (i < n) {
        float *c = st + lane * kShStage;  // coefficients in, v_coeffs out (same slots)
        if (masks && !masks[i]) {
            for (int k = 0; k < K3; ++k) c[k] = 0.f;
        } else {
            const int nb = (degree + 1) * (degree + 1);
            float x = 0, y = 0, z = 0, inv = 0;
            if (nb > 1) {
                const float dx = dirs[i * 3], dy = dirs[i * 3 + 1], dz = dirs[i * 3 + 2];
                inv = tgs_rsqrt<float>(dx * dx + dy * dy + dz * dz);
                x = dx * inv; y = dy * inv; z = dz * inv;
            }
            float Y[25], G[25][3];
            const bool want_dir = v_dirs != nullptr && nb > 1;
            sh_eval<float>(nb, x, y, z, Y, want_dir ? G : nullptr);
            const float vr = v_colors[i * 3], vg = v_colors[i * 3 + 1], vb = v_colors[i * 3 + 2];
            float gu = 0, gv = 0, gw = 0;
            for (int k = 0; k < K; ++k) {
                if (k < nb) {
                    if (want_dir) {
                        const float s = c[k * 3] * vr + c[k * 3 + 1] * vg + c[k * 3 + 2] * vb;
                        gu += s * G[k][0];
                        gv += s * G[k][1];
                        gw += s * G[k][2];
                    }
                    c[k * 3] = Y[k] * vr;
                    c[k * 3 + 1] = Y[k] * vg;
                    c[k * 3 + 2] = Y[k] * vb;
                } else {
                    c[k * 3] = 0.f; c[k * 3 + 1] = 0.f; c[k * 3 + 2] = 0.f;
                }
            }
            if (want_dir) {
                const float gd = gu * x + gv * y + gw * z;
                atomic_add(v_dirs + i * 3, (gu - gd * x) * inv);
                atomic_add(v_dirs + i * 3 + 1, (gv - gd * y) * inv);
                atomic_add(v_dirs + i * 3 + 2, (gw - gd * z) * inv);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    {
        int t = lane / K3, k = lane - (lane / K3) * K3;
        for (int64_t e = lane; e < tot; e += 64) {
            v_coeffs[base * K3 + e] = st[t * kShStage + k];
            k += 64;
            while (k >= K3) {
                k -= K3;
                ++t;
            }
        }
    }
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
    if (dtype == TGS_F32 && K * 3 <= kShStage - 1)
        sh_bwd_staged_kernel<<<nb, 256, 0, st>>>(n, K, degree, (const float *)dirs, (const float *)coeffs, masks,
                                                 (const float *)v_colors, (float *)v_coeffs, (float *)v_dirs);
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
