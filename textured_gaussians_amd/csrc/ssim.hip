// SSIM of the trainer's loss (torchmetrics StructuralSimilarityIndexMeasure with
// data_range 1: 11-tap Gaussian window, sigma 1.5, c1 = 0.01^2, c2 = 0.03^2, the map
// averaged over the windows inside the image), forward and gradient, on
// channels-last images [B, H, W, C] as the rasterizer writes them.
//
// Forward: pass 1 blurs x, y, x^2, y^2, xy along rows (valid columns only), pass 2
// along columns; pass 2 also writes per-block partial sums of the SSIM map and, for
// the gradient with respect to x, three maps per window: dS/dmu_x, dS/dE[x^2],
// dS/dE[xy] (variances clamped at 0 contribute nothing where clamped).
// Gradient: the window is symmetric, so dL/dx(p) = sum over windows c containing p
// of w(p - c) (A(c) + 2 x(p) B(c) + y(p) C(c)): the transposed column pass (pass 3)
// and row pass (pass 4, which also combines).  Everything is memory-bound and
// touches a few hundred MB at 1920x1080.
#include "common.h"

namespace tgs {

constexpr int kWin = 11, kRad = 5;

template <typename S> struct SsimWin {
    S g[kWin];
};

template <typename S> static SsimWin<S> ssim_window() {
    SsimWin<S> w;
    double sum = 0, t[kWin];
    for (int j = 0; j < kWin; ++j) {
        const double d = (double)(j - kRad) / 1.5;
        t[j] = exp(-0.5 * d * d);
        sum += t[j];
    }
    for (int j = 0; j < kWin; ++j) w.g[j] = (S)(t[j] / sum);
    return w;
}

// rows: T[q][b][y][xc][c] = sum_j g[j] f_q(b, y, xc + j, c), xc < Wv = W - 10
template <typename S>
__global__ __launch_bounds__(256) void ssim_rows_kernel(int32_t H, int32_t W, int32_t C, const S *__restrict__ x,
                                                        const S *__restrict__ y, SsimWin<S> w, S *__restrict__ T) {
    const int32_t Wv = W - 2 * kRad;
    const int64_t row = blockIdx.y;  // b * H + yy
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // xc * C + c
    if (i >= Wv * C) return;
    const S *xr = x + row * W * C + i, *yr = y + row * W * C + i;
    S s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
    for (int j = 0; j < kWin; ++j) {
        const S a = xr[j * C], b = yr[j * C], g = w.g[j];
        s0 += g * a;
        s1 += g * b;
        s2 += g * (a * a);
        s3 += g * (b * b);
        s4 += g * (a * b);
    }
    const int64_t plane = (int64_t)gridDim.y * Wv * C, o = row * Wv * C + i;
    T[o] = s0;
    T[plane + o] = s1;
    T[2 * plane + o] = s2;
    T[3 * plane + o] = s3;
    T[4 * plane + o] = s4;
}

// columns + SSIM: one thread per valid window (b, yc, xc, c); writes A, B, Cm and a
// partial sum of S per block
template <typename S>
__global__ __launch_bounds__(256) void ssim_cols_kernel(int32_t B, int32_t H, int32_t W, int32_t C,
                                                        const S *__restrict__ T, SsimWin<S> w, S c1, S c2,
                                                        S *__restrict__ A, S *__restrict__ Bm, S *__restrict__ Cm,
                                                        S *__restrict__ partial) {
    const int32_t Wv = W - 2 * kRad, Hv = H - 2 * kRad;
    const int64_t rc = blockIdx.y;  // b * Hv + yc
    const int32_t b = (int32_t)(rc / Hv), yc = (int32_t)(rc - (int64_t)b * Hv);
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    S ssim = 0;
    if (i < Wv * C) {
        const int64_t plane = (int64_t)B * H * Wv * C;
        const S *t = T + ((int64_t)b * H + yc) * Wv * C + i;
        S mx = 0, my = 0, exx = 0, eyy = 0, exy = 0;
#pragma unroll
        for (int j = 0; j < kWin; ++j) {
            const int64_t o = (int64_t)j * Wv * C;
            const S g = w.g[j];
            mx += g * t[o];
            my += g * t[plane + o];
            exx += g * t[2 * plane + o];
            eyy += g * t[3 * plane + o];
            exy += g * t[4 * plane + o];
        }
        const S vxr = exx - mx * mx, vyr = eyy - my * my;
        const bool ux = vxr >= S(0);
        const S vx = ux ? vxr : S(0), vy = vyr >= S(0) ? vyr : S(0);
        const S sxy = exy - mx * my;
        const S n1 = S(2) * mx * my + c1, n2 = S(2) * sxy + c2;
        const S d1 = mx * mx + my * my + c1, d2 = vx + vy + c2;
        const S inv = S(1) / (d1 * d2);
        ssim = n1 * n2 * inv;
        // derivatives with respect to mu_x, E[x^2], E[xy] (without dividing by n1, n2)
        const S a = (S(2) * my * n2 - S(2) * my * n1) * inv - ssim * (S(2) * mx / d1 - (ux ? S(2) * mx / d2 : S(0)));
        const S bb = ux ? -ssim / d2 : S(0);
        const S cc = S(2) * n1 * inv;
        const int64_t o = rc * Wv * C + i;
        A[o] = a;
        Bm[o] = bb;
        Cm[o] = cc;
    }
    // block sum (deterministic per block; the host sums the partials)
    __shared__ S red[4];
    S v = ssim;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[rc * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// transposed columns: U[q][b][y][xc][c] = sum_{yc = y - j valid} g[j] map_q(b, yc, xc, c)
template <typename S>
__global__ __launch_bounds__(256) void ssim_cols_t_kernel(int32_t B, int32_t H, int32_t W, int32_t C,
                                                          const S *__restrict__ A, const S *__restrict__ Bm,
                                                          const S *__restrict__ Cm, SsimWin<S> w, S *__restrict__ U) {
    const int32_t Wv = W - 2 * kRad, Hv = H - 2 * kRad;
    const int64_t row = blockIdx.y;  // b * H + y
    const int32_t b = (int32_t)(row / H), yy = (int32_t)(row - (int64_t)b * H);
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Wv * C) return;
    S sa = 0, sb = 0, sc = 0;
#pragma unroll
    for (int j = 0; j < kWin; ++j) {
        const int32_t yc = yy - j;  // window row whose tap j lands on row yy
        if (yc < 0 || yc >= Hv) continue;
        const int64_t o = ((int64_t)b * Hv + yc) * Wv * C + i;
        const S g = w.g[j];
        sa += g * A[o];
        sb += g * Bm[o];
        sc += g * Cm[o];
    }
    const int64_t plane = (int64_t)B * H * Wv * C, o = row * Wv * C + i;
    U[o] = sa;
    U[plane + o] = sb;
    U[2 * plane + o] = sc;
}

// transposed rows + combine: dx(b, y, xx, c) = scale (sA + 2 x sB + y sC)
template <typename S>
__global__ __launch_bounds__(256) void ssim_rows_t_kernel(int32_t B, int32_t H, int32_t W, int32_t C,
                                                          const S *__restrict__ U, const S *__restrict__ x,
                                                          const S *__restrict__ y, SsimWin<S> w, S scale,
                                                          S *__restrict__ dx) {
    const int32_t Wv = W - 2 * kRad;
    const int64_t row = blockIdx.y;  // b * H + y
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // xx * C + c
    if (i >= W * C) return;
    const int32_t xx = i / C, c = i - xx * C;
    const int64_t plane = (int64_t)B * H * Wv * C;
    S sa = 0, sb = 0, sc = 0;
#pragma unroll
    for (int j = 0; j < kWin; ++j) {
        const int32_t xc = xx - j;
        if (xc < 0 || xc >= Wv) continue;
        const int64_t o = row * Wv * C + (int64_t)xc * C + c;
        const S g = w.g[j];
        sa += g * U[o];
        sb += g * U[plane + o];
        sc += g * U[2 * plane + o];
    }
    const int64_t p = row * W * C + i;
    dx[p] = scale * (sa + S(2) * x[p] * sb + y[p] * sc);
}

template <typename S>
static int ssim_fwd(int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y, void *work,
                    void *partial, hipStream_t st) {
    const int32_t Wv = W - 2 * kRad, Hv = H - 2 * kRad;
    const SsimWin<S> w = ssim_window<S>();
    S *T = (S *)work;
    const int64_t plane = (int64_t)B * H * Wv * C;
    S *A = T + 5 * plane, *Bm = A + (int64_t)B * Hv * Wv * C, *Cm = Bm + (int64_t)B * Hv * Wv * C;
    const unsigned gx = (unsigned)ceil_div((int64_t)Wv * C, 256);
    ssim_rows_kernel<S><<<dim3(gx, (unsigned)(B * H)), 256, 0, st>>>(H, W, C, (const S *)x, (const S *)y, w, T);
    ssim_cols_kernel<S><<<dim3(gx, (unsigned)(B * Hv)), 256, 0, st>>>(B, H, W, C, T, w, (S)1e-4, (S)9e-4, A, Bm, Cm,
                                                                      (S *)partial);
    TGS_CHECK_LAUNCH("ssim_fwd");
    return TGS_OK;
}

template <typename S>
static int ssim_bwd(int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y, void *work,
                    double scale, void *dx, hipStream_t st) {
    const int32_t Wv = W - 2 * kRad, Hv = H - 2 * kRad;
    const SsimWin<S> w = ssim_window<S>();
    S *T = (S *)work;  // the forward's window sums, reused for U
    const int64_t plane = (int64_t)B * H * Wv * C;
    const S *A = T + 5 * plane, *Bm = A + (int64_t)B * Hv * Wv * C, *Cm = Bm + (int64_t)B * Hv * Wv * C;
    const unsigned gx = (unsigned)ceil_div((int64_t)Wv * C, 256);
    ssim_cols_t_kernel<S><<<dim3(gx, (unsigned)(B * H)), 256, 0, st>>>(B, H, W, C, A, Bm, Cm, w, T);
    const unsigned gx2 = (unsigned)ceil_div((int64_t)W * C, 256);
    ssim_rows_t_kernel<S><<<dim3(gx2, (unsigned)(B * H)), 256, 0, st>>>(B, H, W, C, T, (const S *)x, (const S *)y, w,
                                                                        (S)scale, (S *)dx);
    TGS_CHECK_LAUNCH("ssim_bwd");
    return TGS_OK;
}

static int ssim_check(int32_t B, int32_t H, int32_t W, int32_t C) {
    TGS_CHECK_ARG(B >= 1 && C >= 1 && C <= 4 && H >= kWin && W >= kWin, "ssim: bad sizes B=%d H=%d W=%d C=%d", B, H,
                  W, C);
    TGS_CHECK_ARG((int64_t)B * H <= 65535 && (int64_t)W * C < (int64_t(1) << 30), "ssim: image too large");
    return TGS_OK;
}

}  // namespace tgs

using namespace tgs;

extern "C" size_t tgs_ssim_work_bytes(int dtype, int32_t B, int32_t H, int32_t W, int32_t C) {
    if (H < kWin || W < kWin) return 0;
    const int64_t Wv = W - 2 * kRad, Hv = H - 2 * kRad;
    const int64_t n = 5 * (int64_t)B * H * Wv * C + 3 * (int64_t)B * Hv * Wv * C;
    return (size_t)n * (dtype == TGS_F64 ? sizeof(double) : sizeof(float));
}

extern "C" int64_t tgs_ssim_partials(int32_t B, int32_t H, int32_t W, int32_t C) {
    if (H < kWin || W < kWin) return 0;
    return (int64_t)B * (H - 2 * kRad) * ceil_div((int64_t)(W - 2 * kRad) * C, 256);
}

extern "C" int tgs_ssim_fwd(int dtype, int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y,
                            void *work, void *partials, tgs_stream_t stream) {
    int rc = ssim_check(B, H, W, C);
    if (rc) return rc;
    TGS_CHECK_ARG(x && y && work && partials, "ssim: null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32) return ssim_fwd<float>(B, H, W, C, x, y, work, partials, st);
    if (dtype == TGS_F64) return ssim_fwd<double>(B, H, W, C, x, y, work, partials, st);
    set_error("ssim: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}

extern "C" int tgs_ssim_bwd(int dtype, int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y,
                            void *work, double scale, void *dx, tgs_stream_t stream) {
    int rc = ssim_check(B, H, W, C);
    if (rc) return rc;
    TGS_CHECK_ARG(x && y && work && dx, "ssim: null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32) return ssim_bwd<float>(B, H, W, C, x, y, work, scale, dx, st);
    if (dtype == TGS_F64) return ssim_bwd<double>(B, H, W, C, x, y, work, scale, dx, st);
    set_error("ssim: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}
