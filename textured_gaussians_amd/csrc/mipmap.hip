// Mipmap pyramid of square power-of-two textures for the `mipmapped2` filter
// (generate_mipmap_fwd.cu / generate_mipmap_bwd.cu semantics).  Layout per
// Gaussian: all levels flat, coarsest first, level l (resolution R >> l) at
// texel (4^(L-l) - 1) / 3, channels innermost: [N, (4^(L+1) - 1) / 3, C].
//
// Forward: one pass copies the finest level and writes level 1 (each thread reads
// its 2x2 block once), then each coarser level is the 2x2 box average of the one
// below, summed ((t00 + t01) + t10) + t11 and divided by 4 as
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

// Every extent is a power of two but the channel count, so indices split with
// shifts.  RGBA (C = 4) moves as 4-vectors; other channel counts element-wise.
template <typename S> using vec4_t = S __attribute__((ext_vector_type(4)));

// Finest level and level 1 in one pass: a thread reads a 2x2 block of the input
// texture once, copies it into the pyramid and writes its average.
template <typename S, bool C4>
__global__ __launch_bounds__(256) void mip_top_kernel(int64_t n, int32_t C, int64_t total_c, int32_t log_res,
                                                      const S *__restrict__ tex, S *__restrict__ mip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one level-1 texel
    if (i >= n) return;
    const int32_t lo = log_res - 1;
    const int64_t g = i >> (2 * lo), t = i & ((int64_t(1) << (2 * lo)) - 1);
    const int64_t r1 = int64_t(1) << lo, r0 = 2 * r1;
    const int64_t v = t >> lo, u = t & (r1 - 1);
    const int64_t a = 2 * v * r0 + 2 * u, b = a + 1, c = a + r0, d = c + 1;  // texels of the block
    const S *src = tex + g * r0 * r0 * C;
    S *fine = mip + g * total_c + mip_base(log_res, 0) * C;
    S *out = mip + g * total_c + mip_base(log_res, 1) * C;
    if constexpr (C4) {
        const vec4_t<S> *s4 = (const vec4_t<S> *)src;
        const vec4_t<S> ta = __builtin_nontemporal_load(s4 + a), tb = __builtin_nontemporal_load(s4 + b);
        const vec4_t<S> tc = __builtin_nontemporal_load(s4 + c), td = __builtin_nontemporal_load(s4 + d);
        vec4_t<S> *f4 = (vec4_t<S> *)fine;
        f4[a] = ta;
        f4[b] = tb;
        f4[c] = tc;
        f4[d] = td;
        ((vec4_t<S> *)out)[t] = (ta + tb + tc + td) / S(4);
    } else {
        for (int32_t k = 0; k < C; ++k) {
            const S xa = src[a * C + k], xb = src[b * C + k], xc = src[c * C + k], xd = src[d * C + k];
            fine[a * C + k] = xa;
            fine[b * C + k] = xb;
            fine[c * C + k] = xc;
            fine[d * C + k] = xd;
            out[t * C + k] = (xa + xb + xc + xd) / S(4);
        }
    }
}

template <typename S, bool C4>
__global__ __launch_bounds__(256) void mip_down_kernel(int64_t n, int32_t C, int64_t total_c, int32_t log_res,
                                                       int32_t level, S *__restrict__ mip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one output texel, all channels
    if (i >= n) return;
    const int32_t lo = log_res - level;  // log2 of the output resolution
    const int64_t g = i >> (2 * lo), t = i & ((int64_t(1) << (2 * lo)) - 1);
    const int64_t r_out = int64_t(1) << lo, r_in = 2 * r_out;
    const int64_t v = t >> lo, u = t & (r_out - 1);
    const S *in = mip + g * total_c + mip_base(log_res, level - 1) * C;
    S *out = mip + g * total_c + mip_base(log_res, level) * C;
    const int64_t a = 2 * v * r_in + 2 * u, b = a + 1, c = a + r_in, d = c + 1;
    if constexpr (C4) {
        const vec4_t<S> *i4 = (const vec4_t<S> *)in;
        ((vec4_t<S> *)out)[t] = (i4[a] + i4[b] + i4[c] + i4[d]) / S(4);
    } else {
        for (int32_t k = 0; k < C; ++k)
            out[t * C + k] = (in[a * C + k] + in[b * C + k] + in[c * C + k] + in[d * C + k]) / S(4);
    }
}

template <typename S, bool C4>
__global__ __launch_bounds__(256) void mip_bwd_kernel(int64_t n, int32_t C, int64_t total_c, int32_t log_res,
                                                      const S *__restrict__ v_mip, S *__restrict__ v_tex) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one finest texel, all channels
    if (i >= n) return;
    const int64_t g = i >> (2 * log_res), t = i & ((int64_t(1) << (2 * log_res)) - 1);
    const int64_t v = t >> log_res, u = t & ((int64_t(1) << log_res) - 1);
    const S *vm = v_mip + g * total_c;
    if constexpr (C4) {
        vec4_t<S> acc;
        for (int32_t level = log_res; level >= 0; --level) {
            const int32_t lr = log_res - level;  // log2 of this level's resolution
            const vec4_t<S> x = ((const vec4_t<S> *)vm)[mip_base(log_res, level) + ((v >> level) << lr) + (u >> level)];
            acc = level == log_res ? x : x + acc / S(4);
        }
        __builtin_nontemporal_store(acc, (vec4_t<S> *)v_tex + i);
    } else {
        S acc[5];
        for (int32_t level = log_res; level >= 0; --level) {
            const int32_t lr = log_res - level;
            const S *x = vm + (mip_base(log_res, level) + ((v >> level) << lr) + (u >> level)) * C;
#pragma unroll
            for (int32_t k = 0; k < 5; ++k)
                if (k < C) acc[k] = level == log_res ? x[k] : x[k] + acc[k] / S(4);
        }
#pragma unroll
        for (int32_t k = 0; k < 5; ++k)
            if (k < C) v_tex[i * C + k] = acc[k];
    }
}

template <typename S, bool C4>
static int mip_fwd(int64_t N, int32_t C, int32_t L, const void *tex, void *mip, hipStream_t st) {
    const int64_t R = int64_t(1) << L, total_c = mip_base(L, -1) * C;
    if (L == 0) {  // 1x1 textures: the pyramid is the texture
        if (hipMemcpyAsync(mip, tex, (size_t)(N * C) * sizeof(S), hipMemcpyDeviceToDevice, st) != hipSuccess) {
            set_error("generate_mipmap: copy failed");
            return TGS_ERR_LAUNCH;
        }
        return TGS_OK;
    }
    const int64_t m1 = N * (R / 2) * (R / 2);
    mip_top_kernel<S, C4><<<(unsigned)ceil_div(m1, 256), 256, 0, st>>>(m1, C, total_c, L, (const S *)tex, (S *)mip);
    for (int32_t level = 2; level <= L; ++level) {
        const int64_t r = R >> level, m = N * r * r;
        mip_down_kernel<S, C4><<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(m, C, total_c, L, level, (S *)mip);
    }
    TGS_CHECK_LAUNCH("generate_mipmap_fwd");
    return TGS_OK;
}

template <typename S>
static void mip_bwd(int64_t n, int32_t C, int64_t total_c, int32_t L, const void *v_mip, void *v_tex,
                    hipStream_t st) {
    if (C == 4)
        mip_bwd_kernel<S, true><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, C, total_c, L, (const S *)v_mip,
                                                                            (S *)v_tex);
    else
        mip_bwd_kernel<S, false><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, C, total_c, L, (const S *)v_mip,
                                                                             (S *)v_tex);
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
    // RGBA moves as 4-vectors (the buffers must then be 16-byte aligned, as torch's are)
    const bool c4 = C == 4 && ((uintptr_t)textures % 16) == 0 && ((uintptr_t)mip_textures % 16) == 0;
    if (dtype == TGS_F32)
        return c4 ? mip_fwd<float, true>(N, C, log_texture_res, textures, mip_textures, st)
                  : mip_fwd<float, false>(N, C, log_texture_res, textures, mip_textures, st);
    if (dtype == TGS_F64)
        return c4 ? mip_fwd<double, true>(N, C, log_texture_res, textures, mip_textures, st)
                  : mip_fwd<double, false>(N, C, log_texture_res, textures, mip_textures, st);
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
    const int64_t n = N * R * R;
    TGS_CHECK_ARG(C != 4 || (((uintptr_t)v_mip_textures | (uintptr_t)v_textures) % 16) == 0,
                  "generate_mipmap_bwd: RGBA buffers must be 16-byte aligned");
    if (dtype == TGS_F32)
        mip_bwd<float>(n, C, total_c, log_texture_res, v_mip_textures, v_textures, st);
    else if (dtype == TGS_F64)
        mip_bwd<double>(n, C, total_c, log_texture_res, v_mip_textures, v_textures, st);
    else {
        set_error("generate_mipmap_bwd: unsupported dtype %d", dtype);
        return TGS_ERR_UNSUPPORTED;
    }
    TGS_CHECK_LAUNCH("generate_mipmap_bwd");
    return TGS_OK;
}
