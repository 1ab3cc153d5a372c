// Mipmap pyramid of square power-of-two textures for the `mipmapped2` filter
// (generate_mipmap_fwd.cu / generate_mipmap_bwd.cu semantics).  Layout per
// Gaussian: all levels flat, coarsest first, level l (resolution R >> l) at
// texel (4^(L-l) - 1) / 3, channels innermost: [N, (4^(L+1) - 1) / 3, C].
//
// Forward: the finest level is copied, then each coarser level is the 2x2 box
// average of the one below, summed ((t00 + t01) + t10) + t11 and divided by 4 as
// the reference does, one launch per level (a level is a few hundred KB at 1M
// Gaussians, so the launches are memory-bound and short).
// Backward: the adjoint.  Each finest texel walks its ancestors from the
// coarsest down, acc = v[level][ancestor] + acc / 4, which is the reference's
// in-place coarse-to-fine accumulation in the same operation order, and writes
// v_textures directly (no in-place update of v_mip).
#include "common.h"

namespace tgs {

__host__ __device__ __forceinline__ int64_t mip_base(int32_t log_res, int32_t level) {
    return ((int64_t(1) << (2 * (log_res - level))) - 1) / 3;
}

template <typename S>
__global__ __launch_bounds__(256) void mip_copy_kernel(int64_t n, int64_t per_g, int64_t total_c, int64_t fine_off,
                                                       const S *__restrict__ tex, S *__restrict__ mip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t g = i / per_g, r = i - g * per_g;
    mip[g * total_c + fine_off + r] = tex[i];
}

template <typename S>
__global__ __launch_bounds__(256) void mip_down_kernel(int64_t n, int32_t C, int64_t total_c, int32_t log_res,
                                                       int32_t level, S *__restrict__ mip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one output texel, all channels
    if (i >= n) return;
    const int64_t r_out = (int64_t(1) << log_res) >> level, r_in = 2 * r_out;
    const int64_t per_g = r_out * r_out;
    const int64_t g = i / per_g, t = i - g * per_g;
    const int64_t v = t / r_out, u = t - v * r_out;
    const S *in = mip + g * total_c + mip_base(log_res, level - 1) * C;
    S *out = mip + g * total_c + mip_base(log_res, level) * C;
    const int64_t a = (2 * v * r_in + 2 * u) * C, b = a + C, c = a + r_in * C, d = c + C;
    for (int32_t k = 0; k < C; ++k) out[t * C + k] = (in[a + k] + in[b + k] + in[c + k] + in[d + k]) / S(4);
}

template <typename S>
__global__ __launch_bounds__(256) void mip_bwd_kernel(int64_t n, int32_t C, int64_t total_c, int32_t log_res,
                                                      const S *__restrict__ v_mip, S *__restrict__ v_tex) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one finest texel channel
    if (i >= n) return;
    const int64_t R = int64_t(1) << log_res;
    const int64_t per_g = R * R * C;
    const int64_t g = i / per_g, rem = i - g * per_g;
    const int64_t texel = rem / C, k = rem - texel * C;
    const int64_t v = texel / R, u = texel - v * R;
    const S *vm = v_mip + g * total_c;
    S acc = S(0);
    for (int32_t level = log_res; level >= 0; --level) {
        const int64_t r = R >> level;
        const S x = vm[(mip_base(log_res, level) + (v >> level) * r + (u >> level)) * C + k];
        acc = level == log_res ? x : x + acc / S(4);
    }
    v_tex[i] = acc;
}

template <typename S>
static int mip_fwd(int64_t N, int32_t C, int32_t L, const void *tex, void *mip, hipStream_t st) {
    const int64_t R = int64_t(1) << L, total_c = mip_base(L, -1) * C;
    const int64_t n = N * R * R * C;
    mip_copy_kernel<S><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, R * R * C, total_c, mip_base(L, 0) * C,
                                                                    (const S *)tex, (S *)mip);
    for (int32_t level = 1; level <= L; ++level) {
        const int64_t r = R >> level, m = N * r * r;
        mip_down_kernel<S><<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, C, total_c, L, level, (S *)mip);
    }
    TGS_CHECK_LAUNCH("generate_mipmap_fwd");
    return TGS_OK;
}

}  // namespace tgs

using namespace tgs;

extern "C" int64_t tgs_mipmap_texels(int32_t log_texture_res) {
    return (log_texture_res < 0 || log_texture_res > 15) ? -1 : mip_base(log_texture_res, -1);
}

extern "C" int tgs_generate_mipmap_fwd(int dtype, int64_t N, int32_t C, int32_t log_texture_res,
                                       const void *textures, void *mip_textures, tgs_stream_t stream) {
    TGS_CHECK_ARG(N >= 0 && C >= 1 && C <= 5 && log_texture_res >= 0 && log_texture_res <= 15,
                  "generate_mipmap: bad sizes (N=%lld, C=%d, log_res=%d)", (long long)N, C, log_texture_res);
    TGS_CHECK_ARG(N == 0 || (textures && mip_textures), "generate_mipmap: null buffer");
    TGS_CHECK_ARG(N * mip_base(log_texture_res, -1) * C < (int64_t(1) << 46), "generate_mipmap: too large");
    if (N == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == TGS_F32) return mip_fwd<float>(N, C, log_texture_res, textures, mip_textures, st);
    if (dtype == TGS_F64) return mip_fwd<double>(N, C, log_texture_res, textures, mip_textures, st);
    set_error("generate_mipmap: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}

extern "C" int tgs_generate_mipmap_bwd(int dtype, int64_t N, int32_t C, int32_t log_texture_res,
                                       const void *v_mip_textures, void *v_textures, tgs_stream_t stream) {
    TGS_CHECK_ARG(N >= 0 && C >= 1 && C <= 5 && log_texture_res >= 0 && log_texture_res <= 15,
                  "generate_mipmap_bwd: bad sizes");
    TGS_CHECK_ARG(N == 0 || (v_mip_textures && v_textures), "generate_mipmap_bwd: null buffer");
    if (N == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t R = int64_t(1) << log_texture_res, total_c = mip_base(log_texture_res, -1) * C;
    const int64_t n = N * R * R * C;
    if (dtype == TGS_F32)
        mip_bwd_kernel<float><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(
            n, C, total_c, log_texture_res, (const float *)v_mip_textures, (float *)v_textures);
    else if (dtype == TGS_F64)
        mip_bwd_kernel<double><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(
            n, C, total_c, log_texture_res, (const double *)v_mip_textures, (double *)v_textures);
    else {
        set_error("generate_mipmap_bwd: unsupported dtype %d", dtype);
        return TGS_ERR_UNSUPPORTED;
    }
    TGS_CHECK_LAUNCH("generate_mipmap_bwd");
    return TGS_OK;
}
