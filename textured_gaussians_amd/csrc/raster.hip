// Tile rasterizer for (textured) 2D Gaussian splats: front-to-back alpha
// compositing with a per-Gaussian bilinear RGBA texture, and its backward.
//
// Forward semantics: textured_gaussians/cuda/csrc/rasterize_to_pixels_textured_gaussians_2dgs_fwd.cu:105-445
// (TEX=true) and rasterize_to_pixels_2dgs_fwd.cu (TEX=false); bilinear lookup:
// filters/bilinear.cuh:15-64.  Backward: ..._textured_gaussians_2dgs_bwd.cu:118-702,
// ..._bwd2.cu:611-676 (uv rule), rasterize_to_pixels_2dgs_bwd.cu.
//
// Work decomposition (CDNA4): one workgroup of tile_size^2 threads per screen
// tile (256 for 16x16 = four wave64s, each owning 4 pixel rows).  Every lane of
// a wavefront processes the same list entry at the same time, so the entry's
// Gaussian is wave-uniform data: it is read with scalar loads from an 80-byte
// packed record, built by a small packing kernel, straight into SGPRs that the
// vector ALUs use as operands.  There is no LDS staging of Gaussian batches and
// no workgroup barrier: each wave walks the tile's depth-sorted list on its own
// and stops as soon as a ballot shows all its pixels finished.  Finished lanes
// stay in the loop predicated off so wave reductions (per-Gaussian gradient
// sums, contribution counters) run with a full exec mask.  Workgroups are
// remapped so each XCD renders one contiguous band of tiles (neighbouring
// tiles share Gaussians, hence texels, in that XCD's L2).
//
// The loop is bound by vector-ALU issue, and most (pixel, entry) pairs it
// visits contribute nothing (alpha < 1/255), so the per-pair geometry is kept
// to a handful of instructions and everything texture-related runs only on
// lanes inside the texture range:
//  * the splat-plane cross product of the reference, cr = hu x hv with
//    hu = px T2 - T0 and hv = py T2 - T1, is affine in the pixel position:
//    cr = px (T1 x T2) + py (T2 x T0) + (T0 x T1).  Written around the
//    Gaussian's 2D mean m, cr = C0 - dx A0 - dy A1 with dx = mx - px,
//    A0 = T1 x T2, A1 = T2 x T0 and C0 = cr(m), which the packing kernel
//    computes in float64 (its terms are pixel-sized multiples of A that cancel).
//    cr then costs six FMAs per pixel with the conditioning of the reference's
//    hu x hv; the backward accumulates the matching centred sums and a float64
//    post-pass turns them into ray-transform gradients (cross-product VJP).
//  * s = cr.xy / cr.z uses one reciprocal; the texel coordinate
//    u = (s.x + range) / (2 range) (res - 1) is one FMA with host constants.
//  * texels are addressed as the Gaussian's texture base (SGPRs) plus a 32-bit
//    per-lane byte offset; one RGBA texel is one 16-byte load.
// The forward and the backward evaluate a pair through the same helpers with FP
// contraction disabled and explicit FMAs, so both take identical keep/skip
// decisions.
#include <mutex>
#include <type_traits>

#include "common.h"

namespace tgs {

// Read-only global memory viewed through the constant address space: loads at
// wave-uniform addresses become scalar (s_load) loads into SGPRs.
template <typename T> using cptr = const __attribute__((address_space(4))) T *;
template <typename T> __device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)(uintptr_t)p; }

template <typename S> __device__ __forceinline__ S tfma(S a, S b, S c);
template <> __device__ __forceinline__ float tfma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double tfma<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename S> __device__ __forceinline__ S trcp(S x);
template <> __device__ __forceinline__ float trcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double trcp<double>(double x) { return 1.0 / x; }

template <typename S> struct Texel {
    S c[4];
};
// RGBA texel at byte offset `off` from the texture base `tb`
template <typename S> __device__ __forceinline__ Texel<S> load_texel(const S *tb, uint32_t off);
template <> __device__ __forceinline__ Texel<float> load_texel<float>(const float *tb, uint32_t off) {
    const float4 v = *(const float4 *)((const char *)tb + off);
    return {{v.x, v.y, v.z, v.w}};
}
template <> __device__ __forceinline__ Texel<double> load_texel<double>(const double *tb, uint32_t off) {
    const double2 a = *(const double2 *)((const char *)tb + off);
    const double2 b = *(const double2 *)((const char *)tb + off + 16);
    return {{a.x, a.y, b.x, b.y}};
}

// Packed per-(camera, Gaussian) record, geometry part then shading part:
// mean2d(2) A(9) texture row(1) | opacity(1) colour(4, K used) normal(3), with
// A = (T1 x T2, T2 x T0, C0 = cr at the mean) for the ray transform rows T0, T1, T2.  The
// texture row is the Gaussian id modulo N (int32 bits in float records, an
// exact double otherwise): textures and contribution counters are indexed per
// Gaussian, not per flatten id (the reference's flatten-id indexing runs out of
// bounds for cameras >= 1).
constexpr int kRec = 20;
constexpr int kRecA = 2, kRecRow = 11, kRecOp = 12, kRecCol = 13, kRecNrm = 17;
constexpr int kAbc = 9;  // backward scratch: centred sums of dL/dcr per (camera, Gaussian)

template <typename S> struct RasterArgs {
    int64_t C, N, n_isects;
    int64_t E;            // per-entry inputs: C*N (grid layout) or nnz (packed layout)
    const int64_t *gids;  // packed layout: texture row (Gaussian) of each entry; NULL: entry % N
    const S *means2d, *rt, *colors, *opac, *normals, *textures, *bg;
    const S *rec;  // packed records [C*N][kRec]
    const uint8_t *masks;
    int32_t tex_h, tex_w;
    int64_t tex_stride;                   // elements per texture (tex_h * tex_w * 4)
    S trx, try_, su, sv, uh, vh, umax, vmax;  // u = s.x su + uh, v = s.y sv + vh, clamped to [0, umax/vmax]
    int32_t mip_log;                          // mip2 filter: log2 of the finest resolution
    int32_t cull;                             // forward: skip list entries no pixel of a wave can keep
    // optional: the tiles in launch order, longest list first (tile_order_kernel); order_ws is
    // the workspace region it is written to, when the caller's workspace has room for it
    const int32_t *tile_order;
    int32_t *order_ws;
    S g_weight, vis_bias, thr;
    int32_t W, H, ts, tw, th;
    int32_t bpt;  // workgroups per tile: a tile's ts * ts pixels, row-major, over bpt workgroups
    const int32_t *offsets, *flatten_ids;
    // forward outputs
    S *out_c, *out_a, *out_n, *out_d, *out_m, *csum, *ccnt;
    S *out_accd;  // optional: sum_i w_i depth_i per pixel (the forward's distortion accumulator)
    S *out_T;     // optional: the final transmittance per pixel (the backward's starting T)
    int32_t *last_ids, *median_ids;
    // optional: per wave of a tile and list entry, the 64-bit mask of the wave's lanes
    // that composited the entry ([waves per tile][n_isects]; forward writes, backward reads)
    uint64_t *live;
    // optional (float32, native or bf16 textures, K <= 3, with live): the texture sample
    // of every composited pair inside the texture range, RGB (K channels) and alpha,
    // [waves per tile][n_isects][64][4], packed per (wave, entry) in lane order over the
    // lanes that have one; the backward reads them instead of the texels
    S *samples;
    // measurement only (tgs_set_texel_marks): one byte per texel of the textures,
    // set to 1 by the forward for every texel it reads (native float textures)
    uint8_t *tex_seen;
    // optional texture-gradient records: the forward counts per texture row the
    // (pixel, entry) pairs that will add texture gradient (tex_cnt); the backward
    // claims slots of each row's segment (rec_incl: inclusive prefix sum of the
    // counts, rec_cur: per-row cursor) and writes one record per pair into recs
    int32_t *tex_cnt;
    const int64_t *rec_incl;
    const int32_t *rec_cnt;
    int32_t rec_rows;  // 1: only rows with rec_cnt > 0 take records, the others float atomics
    int32_t *rec_cur;
    void *recs;
    // backward inputs
    const S *r_c, *r_a, *v_c, *v_a, *v_n, *v_d, *v_m;
    const S *r_accd;  // optional: see out_accd; NULL = the reference's seed, render_colors[..., K-1]
    const S *r_T;     // optional: see out_T; NULL = 1 - render_alphas (cancels in float32 where T is small)
    const int32_t *r_last, *r_med;
    int uv_mode;
    // backward outputs (g_abc: scratch, [C*N][kAbc], turned into g_rt by the post-pass)
    S *g_xy_abs, *g_xy, *g_rt, *g_col, *g_op, *g_tex, *g_nrm, *g_abc;
    // optional fused zero-fill of rows of a [N][zrow16 x 16 bytes] buffer with zmask[row] == zsel
    // and bit (row & 7) of zres set
    void *zdst;
    const uint8_t *zmask;
    int64_t zrow16;
    int32_t zsel, zres;
    int32_t zsteps;  // fill stores (1 KB each) per visited list entry and wave
};

template <typename S> __device__ __forceinline__ S rec_int(int32_t v);
template <> __device__ __forceinline__ float rec_int<float>(int32_t v) { return __int_as_float(v); }
template <> __device__ __forceinline__ double rec_int<double>(int32_t v) { return (double)v; }
__device__ __forceinline__ int32_t rec_int_get(float v) { return __float_as_int(v); }
__device__ __forceinline__ int32_t rec_int_get(double v) { return (int32_t)v; }

template <typename S, int K>
__global__ __launch_bounds__(256) void pack_kernel(int64_t n, int64_t N, const int64_t *__restrict__ gids,
                                                   const S *__restrict__ means2d,
                                                   const S *__restrict__ rt, const S *__restrict__ opac,
                                                   const S *__restrict__ colors, const S *__restrict__ normals,
                                                   S *__restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // float64 whatever S is: C0 sums terms of size |pixel| |A| that cancel to a small value
    const v3<double> T0 = mk3<double>(rt[i * 9 + 0], rt[i * 9 + 1], rt[i * 9 + 2]);
    const v3<double> T1 = mk3<double>(rt[i * 9 + 3], rt[i * 9 + 4], rt[i * 9 + 5]);
    const v3<double> T2 = mk3<double>(rt[i * 9 + 6], rt[i * 9 + 7], rt[i * 9 + 8]);
    const double mx = means2d[i * 2], my = means2d[i * 2 + 1];
    const v3<double> A0 = cross(T1, T2), A1 = cross(T2, T0);
    const v3<double> C0 = mx * A0 + my * A1 + cross(T0, T1);
    const v3<double> A[3] = {A0, A1, C0};
    S r[kRec];
    r[0] = (S)mx;
    r[1] = (S)my;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r[kRecA + 3 * k] = (S)A[k].x;
        r[kRecA + 3 * k + 1] = (S)A[k].y;
        r[kRecA + 3 * k + 2] = (S)A[k].z;
    }
    r[kRecRow] = rec_int<S>((int32_t)(gids ? gids[i] : i % N));
    r[kRecOp] = opac[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) r[kRecCol + k] = k < K ? colors[i * K + k] : S(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) r[kRecNrm + k] = normals[i * 3 + k];
    S *o = rec + i * kRec;
#pragma unroll
    for (int k = 0; k < kRec; ++k) o[k] = r[k];
}

// A list entry's Gaussian, in SGPRs: the part the pair geometry and the texel
// fetch need, and the part the blend needs.
template <typename S> struct GaussGeom {
    S mx, my, A[9];
    int32_t row;
};
template <typename S, int K> struct GaussShade {
    S op, col[K], nrm[3];
};
template <typename S> __device__ __forceinline__ GaussGeom<S> load_geom(const RasterArgs<S> &a, int32_t g) {
    cptr<S> R = as_const(a.rec) + (int64_t)g * kRec;
    GaussGeom<S> G;
    G.mx = R[0];
    G.my = R[1];
#pragma unroll
    for (int k = 0; k < 9; ++k) G.A[k] = R[kRecA + k];
    G.row = rec_int_get(R[kRecRow]);
    return G;
}
template <typename S, int K> __device__ __forceinline__ GaussShade<S, K> load_shade(const RasterArgs<S> &a, int32_t g) {
    cptr<S> R = as_const(a.rec) + (int64_t)g * kRec;
    GaussShade<S, K> G;
    G.op = R[kRecOp];
#pragma unroll
    for (int k = 0; k < K; ++k) G.col[k] = R[kRecCol + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) G.nrm[k] = R[kRecNrm + k];
    return G;
}

// Per-(pixel, Gaussian) geometry, shared by the forward and the backward so that
// both take identical keep/skip decisions.
template <typename S> struct PairGeom {
    S cz, rz, sx, sy, w3, w2, dx, dy, sigma, ker, vis;
    bool okz, use3;
};

template <typename S>
__device__ __forceinline__ PairGeom<S> pair_geom(S px, S py, const GaussGeom<S> &G, S vis_bias, S g_weight) {
#pragma clang fp contract(off)
    PairGeom<S> p;
    // cr = C0 - dx A0 - dy A1, centred on the Gaussian's mean (dx = mx - px)
    p.dx = G.mx - px;
    p.dy = G.my - py;
    const S crx = tfma(-p.dx, G.A[0], tfma(-p.dy, G.A[3], G.A[6]));
    const S cry = tfma(-p.dx, G.A[1], tfma(-p.dy, G.A[4], G.A[7]));
    p.cz = tfma(-p.dx, G.A[2], tfma(-p.dy, G.A[5], G.A[8]));
    p.okz = p.cz != S(0);
    p.rz = trcp(p.okz ? p.cz : S(1));
    p.sx = crx * p.rz;
    p.sy = cry * p.rz;
    p.w3 = tfma(p.sx, p.sx, p.sy * p.sy);
    p.w2 = S(2) * tfma(p.dx, p.dx, p.dy * p.dy);
    p.use3 = p.w3 <= p.w2;
    p.sigma = S(0.5) * (p.use3 ? p.w3 : p.w2);
    p.ker = tgs_exp<S>(-p.sigma);
    p.vis = tfma(g_weight, p.ker, vis_bias);
    return p;
}

// inside [-range, range]^2; NaN coordinates count as outside
template <typename S> __device__ __forceinline__ bool in_tex_range(const PairGeom<S> &p, const RasterArgs<S> &a) {
    return p.okz && (p.sx >= -a.trx) && (p.sx <= a.trx) && (p.sy >= -a.try_) && (p.sy <= a.try_);
}

template <typename S> __device__ __forceinline__ S pair_alpha(S op, S vis, S a_s) {
#pragma clang fp contract(off)
    const S al = op * vis * a_s;
    return al < S(0.999) ? al : S(0.999);
}

template <typename S> __device__ __forceinline__ bool pair_keep(const PairGeom<S> &p, S alpha) {
    return p.okz && !(p.sigma < S(0)) && !(alpha < S(1) / S(255));
}

// Bilinear lookup state for one in-range (pixel, Gaussian) pair: weights and
// byte offsets of the four texels from the Gaussian's texture base.
template <typename S> struct Bilerp {
    S wu, wv;
    uint32_t o00, du, dv;  // texel (u0,v0); +du -> (u1,v0); +dv -> (u0,v1)
};

template <typename S> __device__ __forceinline__ Bilerp<S> bilerp_setup(S sx, S sy, const RasterArgs<S> &a) {
#pragma clang fp contract(off)
    Bilerp<S> b;
    S u = tfma(sx, a.su, a.uh), v = tfma(sy, a.sv, a.vh);
    u = u < S(0) ? S(0) : (u > a.umax ? a.umax : u);
    v = v < S(0) ? S(0) : (v > a.vmax ? a.vmax : v);
    const S fu = floor(u), fv = floor(v);
    b.wu = u - fu;
    b.wv = v - fv;
    const uint32_t tb = 4 * sizeof(S);  // bytes per texel
    b.o00 = ((uint32_t)fv * (uint32_t)a.tex_w + (uint32_t)fu) * tb;
    // the reference's second texel is at ceil(u): the next one unless u is integral
    b.du = b.wu > S(0) ? tb : 0u;
    b.dv = b.wv > S(0) ? (uint32_t)a.tex_w * tb : 0u;
    return b;
}

template <typename S> struct Texels {
    Texel<S> t00, t10, t01, t11;
};

template <typename S> __device__ __forceinline__ Texels<S> load_texels(const S *tb, const Bilerp<S> &b) {
    Texels<S> t;
    t.t00 = load_texel(tb, b.o00);
    t.t10 = load_texel(tb, b.o00 + b.du);
    t.t01 = load_texel(tb, b.o00 + b.dv);
    t.t11 = load_texel(tb, b.o00 + b.du + b.dv);
    return t;
}

// Texture storage: none (plain 2DGS), the compute type S, or bfloat16 RGBA texels
// (8 bytes, float compute only).  Texel offsets are kept in units of S texels
// (16 bytes) whatever the storage, so the gradient side (records, v_textures in
// float) is the same for both; a bf16 texel sits at half that byte offset.
constexpr int kTexNone = 0, kTexNative = 1, kTexBF16 = 2, kTexMip2 = 3;

__device__ __forceinline__ Texel<float> load_texel_bf16(const uint16_t *tb, uint32_t off_f32) {
    const uint2 w = *(const uint2 *)((const char *)tb + (off_f32 >> 1));
    return {{__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xffff0000u), __uint_as_float(w.y << 16),
             __uint_as_float(w.y & 0xffff0000u)}};
}

template <typename S, int TEX>
__device__ __forceinline__ Texels<S> fetch_texels(const S *textures, int64_t row_off, const Bilerp<S> &b) {
    if constexpr (TEX == kTexBF16) {
        const uint16_t *tb = (const uint16_t *)textures + row_off;
        Texels<S> t;
        t.t00 = load_texel_bf16(tb, b.o00);
        t.t10 = load_texel_bf16(tb, b.o00 + b.du);
        t.t01 = load_texel_bf16(tb, b.o00 + b.dv);
        t.t11 = load_texel_bf16(tb, b.o00 + b.du + b.dv);
        return t;
    } else {
        return load_texels(textures + row_off, b);
    }
}

template <typename S> struct BWeights {
    S w00, w10, w01, w11;
};
template <typename S> __device__ __forceinline__ BWeights<S> bweights(const Bilerp<S> &b) {
#pragma clang fp contract(off)
    const S ou = S(1) - b.wu, ov = S(1) - b.wv;
    return {ou * ov, b.wu * ov, ou * b.wv, b.wu * b.wv};
}

// One texture-gradient contribution of a (pixel, Gaussian) pair: the bilinear
// lookup (texel (u0,v0) byte offset, whether the +u / +v neighbours are distinct,
// the fractional weights) and the cotangent of the RGBA sample.  32 bytes (f32).
template <typename S> struct __align__(16) TexRecord {
    uint32_t o00, flags;  // flags: bit 0 -> du != 0, bit 1 -> dv != 0
    S wu, wv;
    S q[4];
};

// Trilinear (`mipmapped2`) lookup of one pair (filters/mip2.cuh precompute): the
// level of detail comes from the splat-coordinate change across the pixel.  With
// cr affine in the pixel position, the reference's crosses at px -/+ 1/2 and
// py -/+ 1/2 are cr -/+ A0/2 and cr -/+ A1/2; a zero z in any of them drops the
// pair (zok).  valid: every texel of both levels lies inside its level.  Texel
// offsets are element offsets (texel * 4) in the Gaussian's pyramid.
template <typename S> struct Mip2 {
    S w[8];
    uint32_t off[8];
    bool zok, valid;
};

template <typename S>
__device__ __forceinline__ Mip2<S> mip2_setup(const PairGeom<S> &p, const GaussGeom<S> &G, const RasterArgs<S> &a) {
    Mip2<S> m;
    const S crx = tfma(-p.dx, G.A[0], tfma(-p.dy, G.A[3], G.A[6]));
    const S cry = tfma(-p.dx, G.A[1], tfma(-p.dy, G.A[4], G.A[7]));
    const S h = S(0.5);
    const S x0z = p.cz - h * G.A[2], x1z = p.cz + h * G.A[2];
    const S y0z = p.cz - h * G.A[5], y1z = p.cz + h * G.A[5];
    m.zok = x0z != S(0) && x1z != S(0) && y0z != S(0) && y1z != S(0);
    m.valid = false;
    if (!m.zok || !p.okz) return m;
    const S dsdx_x = (crx + h * G.A[0]) / x1z - (crx - h * G.A[0]) / x0z;
    const S dsdx_y = (cry + h * G.A[1]) / x1z - (cry - h * G.A[1]) / x0z;
    const S dsdy_x = (crx + h * G.A[3]) / y1z - (crx - h * G.A[3]) / y0z;
    const S dsdy_y = (cry + h * G.A[4]) / y1z - (cry - h * G.A[4]) / y0z;
    const int32_t res = 1 << a.mip_log;
    // duv/dx scales both components by the x range, duv/dy by the y range (as the reference)
    const S ex = dsdx_x * a.su, ey = dsdx_y * a.su, fx = dsdy_x * a.sv, fy = dsdy_y * a.sv;
    const S duv = max(sqrt(ex * ex + ey * ey), sqrt(fx * fx + fy * fy));
    S t = log2(duv);
    if (t != t) return m;  // NaN footprint: no texture (every comparison below would pass)
    t = t < S(0) ? S(0) : (t > S(31) ? S(31) : t);
    const S u = tfma(p.sx, a.su, a.uh), v = tfma(p.sy, a.sv, a.vh);  // unclamped
    const int32_t tl = (int32_t)floor(t), th = (int32_t)ceil(t);
    const S wt = t - (S)tl;
    int32_t lvl[2] = {tl, th};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (lvl[j] > a.mip_log) return m;  // beyond the coarsest level: always invalid
        const S inv = S(1) / (S)(1 << lvl[j]);
        const S ul = (u + S(0.5)) * inv - S(0.5), vl = (v + S(0.5)) * inv - S(0.5);
        const S fu = floor(ul), fv = floor(vl), cu = ceil(ul), cv = ceil(vl);
        const int32_t hi = (res - 1) >> lvl[j];
        if (!(fu >= S(0) && fv >= S(0) && cu <= (S)hi && cv <= (S)hi)) return m;  // NaN fails too
        const int32_t lres = res >> lvl[j];
        const uint32_t base = (uint32_t)(((1u << (2 * (a.mip_log - lvl[j]))) - 1) / 3);
        const S wu = ul - fu, wv = vl - fv, wl = j == 0 ? S(1) - wt : wt;
        const uint32_t r0 = base + (uint32_t)fv * lres, r1 = base + (uint32_t)cv * lres;
        m.off[4 * j + 0] = 4 * (r0 + (uint32_t)fu);
        m.off[4 * j + 1] = 4 * (r0 + (uint32_t)cu);
        m.off[4 * j + 2] = 4 * (r1 + (uint32_t)fu);
        m.off[4 * j + 3] = 4 * (r1 + (uint32_t)cu);
        m.w[4 * j + 0] = (S(1) - wu) * (S(1) - wv) * wl;
        m.w[4 * j + 1] = wu * (S(1) - wv) * wl;
        m.w[4 * j + 2] = (S(1) - wu) * wv * wl;
        m.w[4 * j + 3] = wu * wv * wl;
    }
    m.valid = true;
    return m;
}

// a_s and the K colour channels of the trilinear sample, summed over the 8 texels in order
template <typename S, int K>
__device__ __forceinline__ S mip2_blend(const Mip2<S> &m, const S *tb, S tex[K]) {
    S as = S(0);
#pragma unroll
    for (int k = 0; k < K; ++k) tex[k] = S(0);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const Texel<S> t = load_texel(tb, m.off[i] * (uint32_t)sizeof(S));
#pragma unroll
        for (int k = 0; k < K; ++k) tex[k] += m.w[i] * t.c[k];
        as += m.w[i] * t.c[3];
    }
    return as;
}

// Bilinear blend of the four texels: a_s (alpha channel) and the K colour channels.
template <typename S, int K>
__device__ __forceinline__ S bilerp_blend(const BWeights<S> &w, const Texels<S> &t, S tex[K]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < K; ++k)
        tex[k] = tfma(w.w11, t.t11.c[k], tfma(w.w01, t.t01.c[k], tfma(w.w10, t.t10.c[k], w.w00 * t.t00.c[k])));
    return tfma(w.w11, t.t11.c[3], tfma(w.w01, t.t01.c[3], tfma(w.w10, t.t10.c[3], w.w00 * t.t00.c[3])));
}

// One cached texture sample (RGB, alpha) of a pair: 16 bytes in float32.
template <typename S> __device__ __forceinline__ void store_sample(S *q, const S v[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = v[k];
}
template <> __device__ __forceinline__ void store_sample<float>(float *q, const float v[4]) {
    *(float4 *)q = make_float4(v[0], v[1], v[2], v[3]);
}
template <typename S> __device__ __forceinline__ void load_sample(const S *q, S v[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = q[k];
}
template <> __device__ __forceinline__ void load_sample<float>(const float *q, float v[4]) {
    const float4 t = *(const float4 *)q;
    v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w;
}

// Texture-gradient update of one pixel for one texel row: the two horizontally
// adjacent texels (u0, v) and (u1, v), RGBA each, 8 values; off = byte offset of
// the first, du = byte distance to the second (0 when u0 == u1).
template <typename S> struct TexRec {
    uint32_t off, du;
    S v[8];
};

// Fused zero-fill of buffer rows, spread over a kernel's main loop: each wave owns
// rows w, w + W, w + 2W, ... (w = global wave index, W = waves in the grid) and,
// per step(), issues one 1 KB store instruction (16 bytes per lane,
// non-temporal) into the current selected row.  Which of its rows are selected
// is read 64 at a time (one byte load per lane, a ballot) into a uniform bit
// mask, so finding the next row is scalar bit arithmetic.  drain() finishes
// the wave's rows; every wave calls it before exiting.
typedef uint32_t z32x4 __attribute__((ext_vector_type(4)));
struct RowZeroer {
    z32x4 *dst;
    const uint8_t *mask;
    uint64_t bits;
    int32_t N, row16, stride, rb;  // rb: row of bit 0 of `bits`
    int32_t o, sel, res;
    __device__ __forceinline__ void reload(int32_t lane) {
        // next chunk of 64 rows with at least one selected row (bounded by N / stride chunks)
        while (bits == 0 && rb < N) {
            const int64_t r = (int64_t)rb + (int64_t)lane * stride;
            const bool s = r < N && (int32_t)mask[r] == sel && ((res >> (r & 7)) & 1);
            bits = __ballot(s);
            if (bits == 0) rb = (int32_t)min((int64_t)N, (int64_t)rb + 64 * (int64_t)stride);
        }
    }
    __device__ __forceinline__ void init(void *d, const uint8_t *m, int64_t n, int64_t r16, int32_t selv,
                                         int32_t resv, int32_t wave, int32_t waves, int32_t lane) {
        dst = (z32x4 *)d;
        mask = m;
        N = d ? (int32_t)n : 0;
        row16 = (int32_t)r16;
        stride = waves;
        rb = wave;
        bits = 0;
        o = 0;
        sel = selv;
        res = resv;
        reload(lane);
    }
    __device__ __forceinline__ void step(int32_t lane) {
        if (bits == 0) return;
        const int64_t r = (int64_t)rb + (int64_t)__builtin_ctzll(bits) * stride;
        const int32_t k = o + lane;
        if (k < row16) __builtin_nontemporal_store(z32x4{0, 0, 0, 0}, dst + r * row16 + k);
        o += kWave;
        if (o >= row16) {
            o = 0;
            bits &= bits - 1;
            if (bits == 0) {
                rb = (int32_t)min((int64_t)N, (int64_t)rb + 64 * (int64_t)stride);
                reload(lane);
            }
        }
    }
    __device__ __forceinline__ void drain(int32_t lane) {
        while (bits != 0) step(lane);
    }
};

template <typename S> __device__ __forceinline__ void zeroer_init(RowZeroer &z, const RasterArgs<S> &a, int lane) {
    const int32_t wpb = blockDim.x / kWave;
    z.init(a.zdst, a.zmask, a.N, a.zrow16, a.zsel, a.zres, (int32_t)(blockIdx.x * wpb + threadIdx.x / kWave),
           (int32_t)(gridDim.x * wpb), lane);
}

// ------------------------------------------------------------------ forward
// The first 13 values of an entry's packed record (mean2d, A, texture row, opacity), as
// vector loads from the 16-byte aligned record.
template <typename S> struct RecHead {
    S v[16];
};
template <typename S> __device__ __forceinline__ RecHead<S> load_rec_head(const S *R);
template <> __device__ __forceinline__ RecHead<float> load_rec_head<float>(const float *R) {
    RecHead<float> h;
    const float4 *q = (const float4 *)R;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float4 t = q[k];
        h.v[4 * k] = t.x, h.v[4 * k + 1] = t.y, h.v[4 * k + 2] = t.z, h.v[4 * k + 3] = t.w;
    }
    return h;
}
template <> __device__ __forceinline__ RecHead<double> load_rec_head<double>(const double *R) {
    RecHead<double> h;
    const double2 *q = (const double2 *)R;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const double2 t = q[k];
        h.v[2 * k] = t.x, h.v[2 * k + 1] = t.y;
    }
    h.v[14] = h.v[15] = 0;
    return h;
}

// Candidate test of one list entry for a wavefront's pixels, whose centres lie in the
// box [x0, x1] x [y0, y1]: false only if no pixel of the box can keep the pair in the
// kernel's own arithmetic, so skipping the entry changes no output.  With cr affine in
// the pixel (cr = crc + u A0 + v A1 around the box centre, |u| <= hx, |v| <= hy), each
// |cr_k| lies within |crc_k| -/+ (hx |A0_k| + hy |A1_k|), widened by err_k, a bound on
// the rounding of the kernel's two FMAs and on the rounding of this test itself; |s| is
// then at least |cr.xy|_min / |cr.z|_max.  The pair's sigma = 1/2 min(|s|^2, 2 |d|^2) is
// bounded below by that and by the distance from the mean to the box.  A texture can hold
// any alpha, so a textured entry is skipped only when the whole box lies outside the
// texture range (a_s = 1 there) and op (vis_bias + g e^-sigma) stays below 1/255; the plain
// kernel needs the second part only.  Non-positive or non-finite opacities and NaN bounds
// keep the entry.
//
// Float32 entries are tested in float32 (a float64 test held the forward at 119 VGPRs,
// four waves per SIMD; this one at 75, six).  Rounding, with u = 2^-24 and
// M = (|dxc| + hx) |A0_k| + (|dyc| + hy) |A1_k| + |C0_k|:
//  * the kernel's cr_k: dx = mx - px rounded, then two FMAs: <= 3 u M;
//  * this test's crc_k (mx - cx rounded, two FMAs) <= 3 u M, the spread's FMA and product
//    <= 2 u M, the two subtractions of lo / additions of hi <= 4 u M: <= 9 u M;
//  * err_k = 1.2e-6 M >= (3 + 9) u M = 7.2e-7 M, with room for err_k's own rounding;
//  * the quotient, squares and sums of sigma's bound: a few u relative, inside the 1e-5
//    relative and 1e-3 absolute margins (which also cover the kernel's rcp and squares);
//    an |dxc| - hx that cancels leaves an absolute error <= 2 u (|dxc| + hx) ddx in ddx^2,
//    <= 2e-6 for ddx < 1 (tile boxes have hx, hy <= 16) and <= 34 u relative above;
//  * __expf and the products against 1/255: relative errors ~1e-6, inside the 1e-3 alpha
//    margin.
// Float64 entries keep the float64 test with err_k = 1e-15 M.
template <typename S, int TEX>
__device__ __forceinline__ bool fwd_candidate(const RasterArgs<S> &a, int32_t g, float x0, float x1, float y0,
                                              float y1) {
    using B = typename std::conditional<sizeof(S) == 4, float, double>::type;
    const RecHead<S> R = load_rec_head<S>(a.rec + (int64_t)g * kRec);
    const B mx = (B)R.v[0], my = (B)R.v[1], op = (B)R.v[kRecOp];
    const B cx = B(0.5) * (B(x0) + B(x1)), cy = B(0.5) * (B(y0) + B(y1));
    const B hx = B(0.5) * (B(x1) - B(x0)), hy = B(0.5) * (B(y1) - B(y0));
    const B dxc = mx - cx, dyc = my - cy;
    const B dxm = fabs(dxc) + hx, dym = fabs(dyc) + hy;
    const B eps = sizeof(S) == 4 ? B(1.2e-6) : B(1e-15);
    B lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const B a0 = (B)R.v[kRecA + k], a1 = (B)R.v[kRecA + 3 + k], c0 = (B)R.v[kRecA + 6 + k];
        const B crc = c0 - dxc * a0 - dyc * a1;
        const B spread = hx * fabs(a0) + hy * fabs(a1);
        const B err = eps * (dxm * fabs(a0) + dym * fabs(a1) + fabs(c0));
        lo[k] = fabs(crc) - spread - err;
        hi[k] = fabs(crc) + spread + err;
    }
    const B ddx = fmax(B(0), fabs(dxc) - hx), ddy = fmax(B(0), fabs(dyc) - hy);
    B sig = ddx * ddx + ddy * ddy;  // 1/2 of 2 |d|^2
    bool out_of_range = false;
    {
        // hi[2] == 0 means cr.z == 0 over the box (no pixel keeps the pair): the bounds go
        // to +inf, or NaN where lo is 0 too, which keeps the entry
        const B shrink = B(1) - B(1e-5);
        const B lx = fmax(B(0), lo[0]) / hi[2] * shrink, ly = fmax(B(0), lo[1]) / hi[2] * shrink;
        sig = fmin(sig, B(0.5) * (lx * lx + ly * ly));
        out_of_range = lx > (B)a.trx * (B(1) + B(1e-5)) || ly > (B)a.try_ * (B(1) + B(1e-5));
    }
    sig = sig * (B(1) - B(1e-5)) - B(1e-3);
    if (!(op > B(0)) || !(op < B(1e30)) || !(sig == sig)) return true;
    const B vis_ub = (B)a.vis_bias + fmax(B(0), (B)a.g_weight * tgs_exp<B>(-fmax(sig, B(0))));
    const bool faint = op * vis_ub < (B(1) / B(255)) * (B(1) - B(1e-3));
    if constexpr (TEX == kTexNone) return !faint;
    return !(faint && out_of_range);
}

// ------------------------------------------------------------ launch order
// Tiles go to workgroups longest list first (tile_order_kernel): the dispatcher hands
// workgroup b to XCD b mod 8, so consecutive entries of the order spread the heavy tiles
// (the middle of a frame holds ~1.3x the mean list length) over all XCDs, and the short
// tiles come last, where they even out the finishing times.  Measured at 1M Gaussians,
// 1297x840: forward 3.67 -> 2.93 ms, backward 8.07 -> 7.17 ms against contiguous bands
// of tiles per XCD.  Without an order (no workspace room), runs of 16 tiles take turns
// over the XCDs.  A tile of bpt workgroups has them consecutive in b.
template <typename S> __device__ __forceinline__ int64_t tile_block(const RasterArgs<S> &a) {
    if (a.tile_order)
        return (int64_t)as_const(a.tile_order)[blockIdx.x / a.bpt] * a.bpt + blockIdx.x % a.bpt;
    return xcd_runs(blockIdx.x, gridDim.x, 16u);
}

// The longest-first order of the n_tiles tiles (list length from the offsets, 0 for
// tiles the mask switches off), by a counting sort in the LDS of one workgroup: bins of
// exact lengths (kOrderBins - 1 and longer share the first bin), counted, scanned in
// descending length, then each tile claims its place with an LDS atomic (tiles of equal
// length in any order).  ~10x faster than the bitonic sort of (length, tile) keys it
// replaced (60 us for 4346 tiles).
constexpr int kOrderThreads = 1024;
constexpr int32_t kOrderBins = 8192;                    // 8 per thread
constexpr int32_t kOrderMaxTiles = 1 << 18;

__device__ __forceinline__ int32_t order_bin(int32_t t, int32_t n_tiles, const int32_t *offsets, int64_t n_isects,
                                             const uint8_t *masks) {
    const int64_t e = t + 1 < n_tiles ? (int64_t)offsets[t + 1] : n_isects;
    int64_t len = (masks && !masks[t]) ? 0 : e - (int64_t)offsets[t];
    len = len < 0 ? 0 : (len > kOrderBins - 1 ? kOrderBins - 1 : len);
    return (int32_t)(kOrderBins - 1 - len);  // ascending bin = descending length
}

__global__ __launch_bounds__(kOrderThreads) void tile_order_kernel(int32_t n_tiles, const int32_t *__restrict__ offsets,
                                                                   int64_t n_isects, const uint8_t *__restrict__ masks,
                                                                   int32_t *__restrict__ order) {
    __shared__ uint32_t hist[kOrderBins];
    __shared__ uint32_t wsum[kOrderThreads / kWave];
    constexpr int kPer = kOrderBins / kOrderThreads;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    for (int32_t i = tid; i < kOrderBins; i += kOrderThreads) hist[i] = 0;
    __syncthreads();
    for (int32_t t = tid; t < n_tiles; t += kOrderThreads) atomicAdd(hist + order_bin(t, n_tiles, offsets, n_isects, masks), 1u);
    __syncthreads();
    // exclusive scan: thread tid owns bins [kPer tid, kPer tid + kPer)
    uint32_t v[kPer], run = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = hist[kPer * tid + k];
        run += v[k];
    }
    uint32_t inc = run;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t before = inc - run;
    for (int i = 0; i < w; ++i) before += wsum[i];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        hist[kPer * tid + k] = before;
        before += v[k];
    }
    __syncthreads();
    for (int32_t t = tid; t < n_tiles; t += kOrderThreads)
        order[atomicAdd(hist + order_bin(t, n_tiles, offsets, n_isects, masks), 1u)] = t;
}

template <typename S, int K, int TEX>
__global__ __launch_bounds__(256) void raster_fwd_kernel(const RasterArgs<S> a) {
    const int64_t tiles_per_cam = (int64_t)a.tw * a.th;
    const int64_t blk = tile_block(a);
    const int64_t tile_lin = blk / a.bpt;
    const int64_t c = tile_lin / tiles_per_cam;
    const int64_t tid = tile_lin % tiles_per_cam;
    const int ty = (int)(tid / a.tw), tx = (int)(tid % a.tw);
    const int r = (int)(blk % a.bpt) * (int)blockDim.x + (int)threadIdx.x;  // pixel of the tile
    const int lane = r & (kWave - 1);
    const int i = ty * a.ts + r / a.ts, j = tx * a.ts + r % a.ts;
    const bool inside = r < a.ts * a.ts && i < a.H && j < a.W;
    const int64_t pix = (c * a.H + i) * a.W + j;
    const S px = S(j) + S(0.5), py = S(i) + S(0.5);
    RowZeroer zr;
    zeroer_init(zr, a, lane);

    if (a.masks && !as_const(a.masks)[tile_lin]) {
        if (inside) {
#pragma unroll
            for (int k = 0; k < K; ++k) a.out_c[pix * K + k] = a.bg ? a.bg[c * K + k] : S(0);
            a.out_a[pix] = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) a.out_n[pix * 3 + k] = 0;
            a.out_d[pix] = 0;
            a.out_m[pix] = 0;
            a.last_ids[pix] = 0;
            a.median_ids[pix] = 0;
            if (a.out_accd) a.out_accd[pix] = 0;
            if (a.out_T) a.out_T[pix] = 1;
        }
        zr.drain(lane);
        return;
    }

    cptr<int32_t> ids = as_const(a.flatten_ids);
    const int32_t start = as_const(a.offsets)[tile_lin];
    const int32_t end = (tile_lin == a.C * tiles_per_cam - 1) ? (int32_t)a.n_isects : as_const(a.offsets)[tile_lin + 1];

    S T = 1, distort = 0, acc_vd = 0, median = 0;
    int32_t cur = 0, median_id = 0;
    S pc[K], pn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < K; ++k) pc[k] = 0;
    bool done = !inside;
    // composited-lane masks of this wave, one per list entry, written per chunk of 64 entries
    uint64_t *live = a.live ? a.live + (int64_t)(r / kWave) * a.n_isects : nullptr;
    S *smp = (a.samples && live) ? a.samples + (int64_t)(r / kWave) * a.n_isects * (kWave * 4) : nullptr;
    // the box of this wavefront's pixel centres (full tile width: a wave may span partial rows)
    const int wv = r / kWave;
    const int rr0 = (wv * kWave) / a.ts, rr1 = min(a.ts - 1, (wv * kWave + kWave - 1) / a.ts);
    const float bx0 = tx * a.ts + 0.5f, bx1 = tx * a.ts + a.ts - 0.5f;
    const float by0 = ty * a.ts + rr0 + 0.5f, by1 = ty * a.ts + rr1 + 0.5f;
    constexpr bool kCullable = TEX != kTexMip2;  // the mip2 sample's validity depends on the footprint

    // chunks of 64 entries: each lane tests one entry of the chunk, then the wave walks the
    // candidates front to back (all entries when culling is off)
    for (int32_t lo = start; lo < end; lo += kWave) {
        if (__ballot(!done) == 0) break;  // every pixel of this wave is finished
        const int32_t e = lo + lane;
        bool cand = e < end;
        if (kCullable && a.cull && cand) cand = fwd_candidate<S, TEX>(a, ids[e], bx0, bx1, by0, by1);
        uint64_t todo = __ballot(cand);
        uint64_t my_mask = 0;  // lanes of this wave that composited entry lo + lane
        while (todo) {
            if (__ballot(!done) == 0) break;
            const int bit = __builtin_ctzll(todo);
            todo &= todo - 1;
            const int32_t idx = __builtin_amdgcn_readfirstlane(lo + bit);
#pragma unroll
            for (int zs = 0; zs < a.zsteps; ++zs) zr.step(lane);
            const int32_t g = ids[idx];
            const GaussGeom<S> G = load_geom<S>(a, g);
            const GaussShade<S, K> H = load_shade<S, K>(a, g);
            PairGeom<S> p = pair_geom<S>(px, py, G, a.vis_bias, a.g_weight);
            S a_s = 1, tex[K];
#pragma unroll
            for (int k = 0; k < K; ++k) tex[k] = 0;
            bool inr = false;
            if constexpr (TEX == kTexMip2) {
                const Mip2<S> mm = mip2_setup(p, G, a);
                p.okz = p.okz && mm.zok;
                inr = !done && p.okz && mm.valid;
                if (inr) a_s = mip2_blend<S, K>(mm, a.textures + (int64_t)G.row * a.tex_stride, tex);
            } else if constexpr (TEX) {
                inr = !done && in_tex_range(p, a);
                if (inr) {
                    const Bilerp<S> bl = bilerp_setup<S>(p.sx, p.sy, a);
                    a_s = bilerp_blend<S, K>(bweights(bl),
                                             fetch_texels<S, TEX>(a.textures, (int64_t)G.row * a.tex_stride, bl), tex);
                    if (TEX == kTexNative && a.tex_seen) {
                        uint8_t *m = a.tex_seen + (int64_t)G.row * (a.tex_stride / 4);
                        constexpr uint32_t tb = 4 * sizeof(S);
                        m[bl.o00 / tb] = 1;
                        m[(bl.o00 + bl.du) / tb] = 1;
                        m[(bl.o00 + bl.dv) / tb] = 1;
                        m[(bl.o00 + bl.du + bl.dv) / tb] = 1;
                    }
                }
            }
            const S alpha = pair_alpha(H.op, p.vis, a_s);
            S contrib = 0;
            bool counted = false, used = false;
            if (!done && pair_keep(p, alpha)) {
                const S nT = T * (S(1) - alpha);
                if (nT <= S(1e-4)) {
                    done = true;  // this entry is excluded
                } else {
                    used = true;
                    const S w = alpha * T;
#pragma unroll
                    for (int k = 0; k < K; ++k) pc[k] += (H.col[k] + tex[k]) * w;
#pragma unroll
                    for (int k = 0; k < 3; ++k) pn[k] += H.nrm[k] * w;
                    const S depth = H.col[K - 1];
                    distort += S(2) * (w * depth * (S(1) - T) - w * acc_vd);
                    acc_vd += w * depth;
                    if (T > S(0.5)) {
                        median = depth;
                        median_id = idx;
                    }
                    cur = idx;
                    T = nT;
                    if (alpha > a.thr) {
                        contrib = alpha;
                        counted = true;
                    }
                }
            }
            // per-Gaussian contribution counters: one wave sum, one atomic pair per wave
            const uint64_t cm = __ballot(counted);
            if (cm) {
                const S sum = wave_sum(contrib);
                if (lane == 0) {
                    atomic_add(a.csum + G.row, sum);
                    atomic_add(a.ccnt + G.row, S(__popcll(cm)));
                }
            }
            if constexpr (TEX) {
                if (a.tex_cnt) {
                    // pairs whose backward adds texture gradient: composited and in range
                    const uint64_t tu = __ballot(used && inr);
                    if (tu && lane == 0) atomicAdd(a.tex_cnt + G.row, (int32_t)__popcll(tu));
                }
            }
            const uint64_t um = __ballot(used);
            if (lane == bit) my_mask = um;
            if constexpr ((TEX == kTexNative || TEX == kTexBF16) && K <= 3) {
                if (smp && um) {
                    // this pair's sample, at its rank among the wave's in-range composited lanes
                    const uint64_t ut = __ballot(used && inr);
                    if (used && inr) {
                        const int pos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ut >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)ut, 0u));
                        S v4[4] = {0, 0, 0, a_s};
#pragma unroll
                        for (int k = 0; k < K; ++k) v4[k] = tex[k];
                        store_sample(smp + ((int64_t)idx * kWave + pos) * 4, v4);
                    }
                }
            }
        }
        // the backward reads these masks only up to the wave's last composited entry
        if (live && e < end) live[e] = my_mask;
    }
    zr.drain(lane);
    if (inside) {
        a.out_a[pix] = S(1) - T;
#pragma unroll
        for (int k = 0; k < K; ++k) a.out_c[pix * K + k] = a.bg ? pc[k] + T * a.bg[c * K + k] : pc[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) a.out_n[pix * 3 + k] = pn[k];
        a.last_ids[pix] = cur;
        a.out_d[pix] = distort;
        a.out_m[pix] = median;
        a.median_ids[pix] = median_id;
        if (a.out_accd) a.out_accd[pix] = acc_vd;
        if (a.out_T) a.out_T[pix] = T;
    }
}

// ----------------------------------------------------------------- backward
// LEAN: the instantiations for the common float32 case, with the texture gradient through
// per-pair records only and the reference's uv rule (no d(sample)/ds term); LEAN = 1 reads
// the texels, LEAN = 2 the forward's pair samples (K <= 3).  Dropping the texel values
// kept for the uv term and the float-atomic staging takes LEAN = 1 from 125 to 96 VGPRs at
// K = 3: five waves per SIMD instead of four (K = 4 stays at four: at five it would
// spill).  LEAN = 2 needs 100 and runs at four: at five it spilled 8 VGPRs and measured
// 0.1 ms slower per backward at 1M Gaussians, 1297x840 (6.86 against 6.77 ms).
template <typename S, int K, int TEX, int LEAN>
__global__ __launch_bounds__(256, sizeof(S) == 4 ? (LEAN == 1 && K <= 3 ? 5 : 4) : 1) void raster_bwd_kernel(const RasterArgs<S> a) {
    extern __shared__ __align__(16) char smem[];
    // per-wave staging of texture-gradient updates (texel rows x 64 pixels: 2 rows for
    // bilinear, 4 for mip2, two per level)
    constexpr int kTexRows = TEX == kTexMip2 ? 4 : 2;
    TexRec<S> *stage = (TexRec<S> *)smem + (threadIdx.x / kWave) * (kTexRows * kWave);

    const int64_t tiles_per_cam = (int64_t)a.tw * a.th;
    const int64_t blk = tile_block(a);
    const int64_t tile_lin = blk / a.bpt;
    const int64_t c = tile_lin / tiles_per_cam;
    const int64_t tid = tile_lin % tiles_per_cam;
    const int r = (int)(blk % a.bpt) * (int)blockDim.x + (int)threadIdx.x;  // pixel of the tile
    const int lane = r & (kWave - 1);
    RowZeroer zr;
    zeroer_init(zr, a, lane);
    if (a.masks && !as_const(a.masks)[tile_lin]) {
        zr.drain(lane);
        return;
    }
    const int ty = (int)(tid / a.tw), tx = (int)(tid % a.tw);
    const int i = ty * a.ts + r / a.ts, j = tx * a.ts + r % a.ts;
    const bool inside = r < a.ts * a.ts && i < a.H && j < a.W;
    const int64_t pix = inside ? (c * a.H + i) * a.W + j : (c * a.H + a.H - 1) * a.W + a.W - 1;
    const S px = S(j) + S(0.5), py = S(i) + S(0.5);

    cptr<int32_t> ids = as_const(a.flatten_ids);
    const int32_t start = as_const(a.offsets)[tile_lin];
    const int32_t end = (tile_lin == a.C * tiles_per_cam - 1) ? (int32_t)a.n_isects : as_const(a.offsets)[tile_lin + 1];

    // transmittance after the last contributing entry, as the forward left it (1 - alpha
    // loses the small T of saturated pixels to cancellation); walked back by division
    const S T_final = a.r_T ? a.r_T[pix] : S(1) - a.r_a[pix];
    S T = T_final;
    S buf[K], bufn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < K; ++k) buf[k] = 0;
    const int32_t bin_final = inside ? a.r_last[pix] : 0;
    const int32_t median_idx = inside ? a.r_med[pix] : 0;
    S vc[K], vn[3];
#pragma unroll
    for (int k = 0; k < K; ++k) vc[k] = a.v_c[pix * K + k];
    const S va = a.v_a[pix];
#pragma unroll
    for (int k = 0; k < 3; ++k) vn[k] = a.v_n[pix * 3 + k];
    S vdist = 0, acc_d = 0, acc_w = 0, acc_d_buf = 0, acc_w_buf = 0, dist_buf = 0;
    if (a.v_d) {
        vdist = a.v_d[pix];
        // The reference seeds the depth accumulator with render_colors[..., K-1], which also holds
        // the texture colour and T*background of that channel; with r_accd the forward's own
        // accumulator is used instead and the distortion gradient is the true one.
        acc_d_buf = a.r_accd ? a.r_accd[pix] : a.r_c[pix * K + K - 1];
        acc_d = acc_d_buf;
        acc_w_buf = a.r_a[pix];
        acc_w = acc_w_buf;
    }
    const bool has_d = vdist != S(0);  // the distortion terms all carry a factor v_distort
    const S vmed = a.v_m[pix];
    S bg_dot = 0;
    if (a.bg) {
#pragma unroll
        for (int k = 0; k < K; ++k) bg_dot += a.bg[c * K + k] * vc[k];
    }
    // the wave starts at the last entry any of its pixels used
    const int32_t wave_last = __builtin_amdgcn_readfirstlane(wave_max(inside ? bin_final : int32_t(-1)));
    S ds_u = 0, ds_v = 0;  // d(u)/d(s.x), d(v)/d(s.y) under the selected rule
    if (TEX && a.uv_mode == TGS_UV_EXACT) {
        ds_u = a.su;
        ds_v = a.sv;
    } else if (TEX && a.uv_mode == TGS_UV_DOUBLE) {
        ds_u = S(2) * a.su;
        ds_v = S(2) * a.sv;
    }
    // mip2: the reference's rule only
    const bool uv_term = !LEAN && TEX && TEX != kTexMip2 && a.uv_mode != TGS_UV_NONE;
    const int32_t first = min(end - 1, wave_last);
    const uint64_t *live = a.live ? a.live + (int64_t)(r / kWave) * a.n_isects : nullptr;
    // the forward's samples: only for the reference uv rule (the others need the texels)
    const S *smp = (LEAN == 2 || (!LEAN && a.samples && live && !uv_term))
                       ? a.samples + (int64_t)(r / kWave) * a.n_isects * (kWave * 4)
                       : nullptr;
    // per-Gaussian gradient fields in reduce-scatter order: dL/dA (9), colour (K),
    // normal (3), mean2d (2), opacity, |mean2d| (2); this lane adds field `slot`
    constexpr int kGradFields = 17 + K;
    S *gdst = nullptr;
    int32_t gstride = 0;
    {
        const int slot = wave_rs_slot<kGradFields>(lane);
        if (slot >= 0 && slot < 9) { gdst = a.g_abc + slot; gstride = kAbc; }
        else if (slot >= 9 && slot < 9 + K) { gdst = a.g_col + (slot - 9); gstride = K; }
        else if (slot >= 9 + K && slot < 12 + K) { gdst = a.g_nrm + (slot - 9 - K); gstride = 3; }
        else if (slot >= 12 + K && slot < 14 + K) { gdst = a.g_xy + (slot - 12 - K); gstride = 2; }
        else if (slot == 14 + K) { gdst = a.g_op; gstride = 1; }
        else if (slot >= 15 + K && slot < 17 + K && a.g_xy_abs) { gdst = a.g_xy_abs + (slot - 15 - K); gstride = 2; }
    }

    // back to front in chunks of 64 entries; with the forward's live bytes only the
    // entries some pixel of this wave composited are visited
    for (int32_t hi = first; hi >= start; hi -= kWave) {
        const int32_t lo = max(start, hi - (kWave - 1));
        uint64_t todo;
        if (live) {
            const int32_t e = lo + lane;
            todo = __ballot(e <= hi && live[e] != 0);
        } else {
            const int32_t cnt = hi - lo + 1;
            todo = cnt == kWave ? ~0ull : ((1ull << cnt) - 1);
        }
    while (todo) {
        const int bit = 63 - __builtin_clzll(todo);
        todo ^= 1ull << bit;
        // wave-uniform (todo is a ballot); stated so the list and record reads stay scalar
        const int32_t idx = __builtin_amdgcn_readfirstlane(lo + bit);
        // with the forward's samples, its composited-lane mask decides which lanes keep the pair
        const uint64_t kept = smp ? as_const(live)[idx] : 0;
#pragma unroll
        for (int zs = 0; zs < a.zsteps; ++zs) zr.step(lane);
        const int32_t g = ids[idx];
        const GaussGeom<S> G = load_geom<S>(a, g);
        const GaussShade<S, K> H = load_shade<S, K>(a, g);
        PairGeom<S> p = pair_geom<S>(px, py, G, a.vis_bias, a.g_weight);
        bool valid = inside && idx <= bin_final && p.okz;
        if (smp) valid = valid && ((kept >> lane) & 1);
        S a_s = 1, tex[K];
#pragma unroll
        for (int k = 0; k < K; ++k) tex[k] = 0;
        Bilerp<S> bl;
        BWeights<S> bw;
        Texels<S> tt;
        bool inr = false;
        Mip2<S> mm;
        if constexpr (TEX == kTexMip2) {
            mm = mip2_setup(p, G, a);
            p.okz = p.okz && mm.zok;
            inr = valid && p.okz && mm.valid;
            if (inr) a_s = mip2_blend<S, K>(mm, a.textures + (int64_t)G.row * a.tex_stride, tex);
        } else if constexpr (TEX) {
            inr = valid && in_tex_range(p, a);
            bool cached = false;
            if constexpr (K <= 3) {
                if (LEAN == 2 || smp) {
                    // the sample the forward stored, at the lane's rank among the in-range kept lanes
                    cached = true;
                    const uint64_t ut = __ballot(inr);
                    if (inr) {
                        bl = bilerp_setup<S>(p.sx, p.sy, a);
                        bw = bweights(bl);
                        const int pos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(ut >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)ut, 0u));
                        S v4[4];
                        // entry base uniform (scalar), lane offset 32-bit: a scalar-base load
                        const S *se = smp + (int64_t)idx * (kWave * 4);
                        load_sample(se + (uint32_t)pos * 4u, v4);
#pragma unroll
                        for (int k = 0; k < K; ++k) tex[k] = v4[k];
                        a_s = v4[3];
                    }
                }
            }
            if constexpr (LEAN != 2) {
                if (inr && !cached) {
                    bl = bilerp_setup<S>(p.sx, p.sy, a);
                    bw = bweights(bl);
                    tt = fetch_texels<S, TEX>(a.textures, (int64_t)G.row * a.tex_stride, bl);
                    a_s = bilerp_blend<S, K>(bw, tt, tex);
                }
            }
        }
        const S alpha = pair_alpha(H.op, p.vis, a_s);
        valid = valid && pair_keep(p, alpha);
        if (__ballot(valid) == 0) continue;  // wave-uniform

        // lanes with texture-gradient updates are staged compactly: this lane at tpos
        uint64_t tl = 0;
        int tpos = 0;
        int32_t rbase = 0;
        bool use_rec = false;
        if constexpr (TEX) {
            tl = __ballot(valid && inr);
            tpos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(tl >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)tl, 0u));
            // records for this entry's row, or float atomics (uniform: one row per entry)
            use_rec = LEAN || (a.recs && (!a.rec_rows || as_const(a.rec_cnt)[G.row] > 0));
            if (use_rec && tl) {
                // claim this entry's slots in the row's record segment now; used at the end
                if (lane == 0) rbase = atomicAdd(a.rec_cur + G.row, (int32_t)__popcll(tl));
            }
        }
        S g_col[K], g_n[3] = {0, 0, 0}, g_A[9], g_xy[2] = {0, 0}, g_abs[2] = {0, 0}, g_op = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) g_col[k] = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) g_A[k] = 0;
        if (valid) {
            if (idx == median_idx) g_col[K - 1] += vmed;
            const S ra = S(1) / (S(1) - alpha);
            T *= ra;
            const S fac = alpha * T;
            S v_alpha = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                g_col[k] += fac * vc[k];
                v_alpha += ((H.col[k] + tex[k]) * T - buf[k] * ra) * vc[k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                g_n[k] = fac * vn[k];
                v_alpha += (H.nrm[k] * T - bufn[k] * ra) * vn[k];
            }
            v_alpha += T_final * ra * va;
            v_alpha += -T_final * ra * bg_dot;
            if (has_d) {
                const S depth = H.col[K - 1];
                const S dl_dw = S(2) * (S(2) * (depth * acc_w_buf - acc_d_buf) + (acc_d - depth * acc_w));
                v_alpha += (dl_dw * T - dist_buf * ra) * vdist;
                acc_d_buf -= fac * depth;
                acc_w_buf -= fac;
                dist_buf += dl_dw * fac;
                g_col[K - 1] += S(2) * fac * (S(2) - S(2) * T - acc_w + fac) * vdist;
            }
            const bool unclamped = H.op * p.vis * a_s <= S(0.999);
            S vsx = 0, vsy = 0;
            if (unclamped) {
                const S v_G = H.op * v_alpha * a_s;
                const S dk = -a.g_weight * p.ker;  // d(vis)/d(sigma)
                if (p.use3) {
                    vsx = v_G * dk * p.sx;
                    vsy = v_G * dk * p.sy;
                } else {
                    g_xy[0] = v_G * dk * S(2) * p.dx;
                    g_xy[1] = v_G * dk * S(2) * p.dy;
                    g_abs[0] = fabs(g_xy[0]);
                    g_abs[1] = fabs(g_xy[1]);
                }
                g_op = p.vis * v_alpha * a_s;
            }
            if constexpr (TEX) {
                if (inr) {
                    // d(loss)/d(sample) per channel: colour channels k < K, alpha channel 3
                    // (for K = 4 channel 3 carries both, as in the forward's sum)
                    const S v_as = unclamped ? p.vis * H.op * v_alpha : S(0);
                    S q[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int k = 0; k < K; ++k) q[k] = fac * vc[k];
                    q[3] += v_as;
                    if constexpr (TEX == kTexMip2) {
                        // four texel rows (two per level) of two adjacent texels each
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            TexRec<S> &r = stage[j * kWave + tpos];
                            r.off = mm.off[2 * j] * (uint32_t)sizeof(S);
                            r.du = (mm.off[2 * j + 1] - mm.off[2 * j]) * (uint32_t)sizeof(S);
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                r.v[k] = mm.w[2 * j] * q[k];
                                r.v[4 + k] = mm.w[2 * j + 1] * q[k];
                            }
                        }
                    } else if (use_rec) {
                        // lane 0 claimed the slots; read its value whatever lanes are active here
                        const int32_t pos = __builtin_amdgcn_readlane(rbase, 0) + tpos;
                        S *gtb = a.g_tex + (int64_t)G.row * a.tex_stride;
                        if (pos < as_const(a.rec_cnt)[G.row]) {
                            const int64_t slot = (G.row ? as_const(a.rec_incl)[G.row - 1] : 0) + pos;
                            TexRecord<S> rec;
                            rec.o00 = bl.o00;
                            rec.flags = (bl.du ? 1u : 0u) | (bl.dv ? 2u : 0u);
                            rec.wu = bl.wu;
                            rec.wv = bl.wv;
#pragma unroll
                            for (int k = 0; k < 4; ++k) rec.q[k] = q[k];
                            ((TexRecord<S> *)a.recs)[slot] = rec;
                        } else if (a.g_tex) {
                            // more pairs than the forward counted (not expected): add directly.
                            // Without v_textures (the records feed the row-wise Adam step) the
                            // pair is dropped; the cursor exceeding the count shows it.
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                if (q[k] == S(0)) continue;
                                atomic_add((S *)((char *)gtb + bl.o00) + k, bw.w00 * q[k]);
                                atomic_add((S *)((char *)gtb + bl.o00 + bl.du) + k, bw.w10 * q[k]);
                                atomic_add((S *)((char *)gtb + bl.o00 + bl.dv) + k, bw.w01 * q[k]);
                                atomic_add((S *)((char *)gtb + bl.o00 + bl.du + bl.dv) + k, bw.w11 * q[k]);
                            }
                        }
                    } else {
                        TexRec<S> &r0 = stage[tpos];
                        TexRec<S> &r1 = stage[kWave + tpos];
                        r0.off = bl.o00;
                        r0.du = bl.du;
                        r1.off = bl.o00 + bl.dv;
                        r1.du = bl.du;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            r0.v[k] = bw.w00 * q[k];
                            r0.v[4 + k] = bw.w10 * q[k];
                            r1.v[k] = bw.w01 * q[k];
                            r1.v[4 + k] = bw.w11 * q[k];
                        }
                    }
                    if (uv_term) {
                        // d(sample)/du and d(sample)/dv of the bilinear blend, against the
                        // sample's cotangent (colour channels, plus alpha when unclamped)
                        S gu = 0, gv = 0;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const S vk = fac * vc[k];
                            gu += vk * ((S(1) - bl.wv) * (tt.t10.c[k] - tt.t00.c[k]) + bl.wv * (tt.t11.c[k] - tt.t01.c[k]));
                            gv += vk * ((S(1) - bl.wu) * (tt.t01.c[k] - tt.t00.c[k]) + bl.wu * (tt.t11.c[k] - tt.t10.c[k]));
                        }
                        gu += v_as * ((S(1) - bl.wv) * (tt.t10.c[3] - tt.t00.c[3]) + bl.wv * (tt.t11.c[3] - tt.t01.c[3]));
                        gv += v_as * ((S(1) - bl.wu) * (tt.t01.c[3] - tt.t00.c[3]) + bl.wu * (tt.t11.c[3] - tt.t10.c[3]));
                        vsx += gu * ds_u;
                        vsy += gv * ds_v;
                    }
                }
            }
            if (vsx != S(0) || vsy != S(0)) {
                // s = cr.xy / cr.z with cr = C0 - dx A0 - dy A1
                const S ax = vsx * p.rz, ay = vsy * p.rz;
                const S az = -(ax * p.sx + ay * p.sy);
                // centred sums: (-dx, -dy, 1) x dL/dcr; the post-pass restores dL/dA
                g_A[0] = -p.dx * ax; g_A[1] = -p.dx * ay; g_A[2] = -p.dx * az;
                g_A[3] = -p.dy * ax; g_A[4] = -p.dy * ay; g_A[5] = -p.dy * az;
                g_A[6] = ax; g_A[7] = ay; g_A[8] = az;
            }
#pragma unroll
            for (int k = 0; k < K; ++k) buf[k] += (tex[k] + H.col[k]) * fac;
#pragma unroll
            for (int k = 0; k < 3; ++k) bufn[k] += H.nrm[k] * fac;
        }
        if constexpr (TEX) {
            // Texture-gradient flush, transposed through LDS: in each round 8 consecutive
            // lanes add the 8 values of one pixel's texel row (two adjacent RGBA texels,
            // contiguous), so one atomic instruction touches 8 such runs instead of 64
            // scattered values; the n staged pixels take ceil(n / 8) rounds.
            const int n_tl = use_rec ? 0 : __popcll(tl);
            if (n_tl) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                S *gtb = a.g_tex + (int64_t)G.row * a.tex_stride;
                const int sub = lane >> 3, e = lane & 7;
#pragma unroll 1
                for (int grp = 0; grp * 8 < n_tl; ++grp) {
                    const int src = grp * 8 + sub;
                    if (src < n_tl) {
#pragma unroll
                        for (int row = 0; row < kTexRows; ++row) {
                            const TexRec<S> &rr = stage[row * kWave + src];
                            const S val = rr.v[e];
                            const uint32_t off = rr.off + (e >> 2) * rr.du + (e & 3) * (uint32_t)sizeof(S);
                            if (val != S(0)) atomic_add((S *)((char *)gtb + off), val);
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        // per-Gaussian sums over the wavefront, reduce-scattered: the lane holding a
        // field's total adds it (one atomic instruction for all fields)
        {
            S f[kGradFields];
#pragma unroll
            for (int k = 0; k < 9; ++k) f[k] = g_A[k];
#pragma unroll
            for (int k = 0; k < K; ++k) f[9 + k] = g_col[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) f[9 + K + k] = g_n[k];
            f[12 + K] = g_xy[0];
            f[13 + K] = g_xy[1];
            f[14 + K] = g_op;
            f[15 + K] = g_abs[0];
            f[16 + K] = g_abs[1];
            const S v = wave_reduce_scatter(f, lane);
            if (gdst && v != S(0)) atomic_add(gdst + (int64_t)g * gstride, v);
        }
    }
    }
    zr.drain(lane);
}

template <typename S>
__global__ __launch_bounds__(256) void zero_kernel(int64_t n, S *__restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Post-pass of the backward, in float64: the kernel accumulated the centred sums
// S_x = sum (px - mx) dL/dcr, S_y = sum (py - my) dL/dcr, S_0 = sum dL/dcr, so the
// gradient with respect to A = (T1 x T2, T2 x T0, T0 x T1) is (S_x + mx S_0,
// S_y + my S_0, S_0) (the mean only recentres cr: its terms cancel).  The
// cross-product VJP (for z = x X y: dL/dx = y X dL/dz, dL/dy = dL/dz X x) turns
// that into the ray-transform gradient, added to v_ray_transforms; then the
// densification gradient v_densify = (dL/dT[0][2], dL/dT[1][2]) * T[2][2] from
// the settled value.
template <typename S>
__global__ __launch_bounds__(256) void rt_grad_kernel(int64_t n, const S *__restrict__ rt, const S *__restrict__ means2d,
                                                      const S *__restrict__ g_abc, S *__restrict__ g_rt,
                                                      S *__restrict__ v_densify) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const S *T = rt + i * 9;
    const S *q = g_abc + i * kAbc;
    const v3<double> T0 = mk3<double>(T[0], T[1], T[2]), T1 = mk3<double>(T[3], T[4], T[5]);
    const v3<double> T2 = mk3<double>(T[6], T[7], T[8]);
    const double mx = means2d[i * 2], my = means2d[i * 2 + 1];
    const v3<double> s0 = mk3<double>(q[6], q[7], q[8]);
    const v3<double> va = mk3<double>(q[0], q[1], q[2]) + mx * s0;
    const v3<double> vb = mk3<double>(q[3], q[4], q[5]) + my * s0;
    const v3<double> vc = s0;
    const v3<double> v0 = cross(vb, T2) + cross(T1, vc);
    const v3<double> v1 = cross(T2, va) + cross(vc, T0);
    const v3<double> v2 = cross(va, T1) + cross(T0, vb);
    S *o = g_rt + i * 9;
    const double d[9] = {v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z};
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] += (S)d[k];
    if (v_densify) {
        v_densify[i * 2] = o[2] * T[8];
        v_densify[i * 2 + 1] = o[5] * T[8];
    }
}

// Texture-gradient gather: per texture row, the sum of its records (one per
// (pixel, Gaussian) pair the backward found in the texture range), written once.
// Float atomics into L2 would cost one request per 32 bytes of texel row per
// pair; here the additions happen in LDS and HBM sees each row once, so the
// kernel streams at write bandwidth.
//
// gather_list_kernel first compacts the rows to write into a list of
// (row, records, first record) entries.  The gather then takes list entries in
// a stride over its workgroups (balanced whatever the rows' order) and reads
// them with scalar loads.
//
// A wave's vector loads and stores share one completion counter on this
// architecture: a wave that waits for a vector load also waits for every store
// it issued before it, and a __syncthreads() (a release fence) waits for all of
// its stores.  So the 512-thread workgroup splits by role, with two LDS row
// buffers and an LDS-only barrier:
//   waves 0-3 (loaders) add the records of entry k into buffer k & 1 with LDS
//   float atomics, after requesting the records of entry k + 1 (every lane
//   loads, so the number of loads in flight is fixed and the wait for entry k's
//   records does not wait for them); they store nothing to global memory;
//   waves 4-7 (writers) meanwhile read buffer (k - 1) & 1, zero it and store it
//   non-temporally; their only global loads are scalar, so nothing waits for
//   their stores.
// overwrite: the listed rows (mask[row] == 1) are written in full, so they need
// no zero-fill beforehand; other rows are not touched.  accumulate: the listed
// rows (those with records) get their sums added.
struct __align__(16) GatherEntry {
    int32_t row, n;  // texture row, records actually written for it
    int64_t start;   // its first record
};

// Rows are taken kGatherListRows per workgroup (4 per thread): the workgroup sums its
// selected rows in LDS and claims its place in the list with one atomic (per-wave
// atomics on the one counter serialize at the memory side: 0.18 ms at N = 1M).
constexpr int kGatherListThreads = 1024, kGatherListPer = 4;
constexpr int64_t kGatherListRows = (int64_t)kGatherListThreads * kGatherListPer;

__global__ __launch_bounds__(kGatherListThreads) void gather_list_kernel(int64_t N, const uint8_t *__restrict__ mask,
                                                                         const int32_t *__restrict__ cnt,
                                                                         const int32_t *__restrict__ cur,
                                                                         const int64_t *__restrict__ incl,
                                                                         GatherEntry *__restrict__ list, int64_t cap,
                                                                         uint32_t *__restrict__ count) {
    constexpr int kWaves = kGatherListThreads / kWave;
    __shared__ uint32_t wave_base[kWaves];
    __shared__ uint32_t block_base;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    bool sel[kGatherListPer];
    int32_t n[kGatherListPer];
    uint32_t before[kGatherListPer];  // selected rows of this wave ahead of this lane's row k
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kGatherListPer; ++k) {
        const int64_t r = (int64_t)blockIdx.x * kGatherListRows + (int64_t)k * kGatherListThreads + threadIdx.x;
        sel[k] = false;
        n[k] = 0;
        if (r < N) {
            n[k] = min(cnt[r], cur[r]);  // records actually written
            sel[k] = mask ? mask[r] == 1 : n[k] > 0;
        }
        const uint64_t bits = __ballot(sel[k]);
        before[k] = total + __builtin_amdgcn_mbcnt_hi((uint32_t)(bits >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bits, 0u));
        total += (uint32_t)__popcll(bits);
    }
    if (lane == 0) wave_base[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t t = wave_base[w];
            wave_base[w] = sum;
            sum += t;
        }
        block_base = sum ? atomicAdd(count, sum) : 0u;
    }
    __syncthreads();
    const uint32_t base = block_base + wave_base[wave];
#pragma unroll
    for (int k = 0; k < kGatherListPer; ++k) {
        const int64_t r = (int64_t)blockIdx.x * kGatherListRows + (int64_t)k * kGatherListThreads + threadIdx.x;
        // the list has room for every row a mask of tgs_mark_rows can select; rows past
        // `cap` (another mask) are not listed, and the gather reads at most `cap` entries
        const int64_t pos = (int64_t)base + before[k];
        if (sel[k] && pos < cap) list[pos] = GatherEntry{(int32_t)r, n[k], r ? incl[r - 1] : 0};
    }
}

__device__ __forceinline__ GatherEntry load_entry(cptr<GatherEntry> L, int32_t k) {
    GatherEntry g;
    g.row = L[k].row;
    g.n = L[k].n;
    g.start = L[k].start;
    return g;
}

// Loader lanes take runs of kGatherRun consecutive records of a row.  Consecutive
// records are neighbouring pixels of one Gaussian (the backward writes them in pixel
// order), and where the splat is magnified they hit the same four texels: the run's
// weighted cotangents are summed in registers while the texel quad stays the same, and
// LDS float atomics (which serialize on equal addresses) are issued only when it changes.
template <typename S> constexpr int gather_run() { return sizeof(S) == 4 ? 4 : 2; }  // registers: 32 / 32 dwords
template <typename S> struct RecRun {
    TexRecord<S> r[gather_run<S>()];
};

template <typename S>
__device__ __forceinline__ void load_run(RecRun<S> &run, const TexRecord<S> *recs, const GatherEntry &g,
                                         int32_t first) {
    // every lane loads kGatherRun records (record 0 where it has none), so the number of
    // loads in flight is fixed for the waits that follow
#pragma unroll
    for (int j = 0; j < gather_run<S>(); ++j) {
        const int32_t q = first + j;
        run.r[j] = recs[q < g.n ? g.start + q : 0];
    }
}

template <typename S>
__device__ __forceinline__ void flush_quad(S *acc, const uint32_t e[4], const S sum[16]) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (sum[4 * k + c] != S(0)) atomicAdd(acc + e[k] + c, sum[4 * k + c]);
}

// records first .. first + gather_run<S>() - 1 (those below n) of a row into the LDS row acc
template <typename S>
__device__ __forceinline__ void add_run(S *acc, const RecRun<S> &run, int32_t first, int32_t n, int32_t tex_w) {
    const uint32_t tb = 4 * sizeof(S);
    S sum[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) sum[k] = S(0);
    uint32_t e[4] = {0, 0, 0, 0};
    uint32_t key = 0xffffffffu;  // texel (u0, v0) and the two neighbour flags of the summed quad
#pragma unroll
    for (int j = 0; j < gather_run<S>(); ++j) {
        if (first + j >= n) break;
        const TexRecord<S> &rec = run.r[j];
        Bilerp<S> bl;
        bl.wu = rec.wu;
        bl.wv = rec.wv;
        bl.o00 = rec.o00;
        bl.du = (rec.flags & 1u) ? tb : 0u;
        bl.dv = (rec.flags & 2u) ? (uint32_t)tex_w * tb : 0u;
        const uint32_t k2 = (rec.o00 << 2) | (rec.flags & 3u);
        if (k2 != key) {
            if (key != 0xffffffffu) {
                flush_quad(acc, e, sum);
#pragma unroll
                for (int k = 0; k < 16; ++k) sum[k] = S(0);
            }
            key = k2;
            e[0] = bl.o00 / sizeof(S);
            e[1] = (bl.o00 + bl.du) / sizeof(S);
            e[2] = (bl.o00 + bl.dv) / sizeof(S);
            e[3] = (bl.o00 + bl.du + bl.dv) / sizeof(S);
        }
        const BWeights<S> bw = bweights(bl);
        const S w[4] = {bw.w00, bw.w10, bw.w01, bw.w11};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c) sum[4 * k + c] += w[k] * rec.q[c];
    }
    if (key != 0xffffffffu) flush_quad(acc, e, sum);
}

// Workgroup barrier for LDS traffic only: this wave's LDS operations complete, then
// s_barrier.  Unlike __syncthreads() it does not wait for the wave's global stores.
// The empty asm statements keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    asm volatile("" ::: "memory");
}

template <typename S>
__global__ __launch_bounds__(512, 4) void tex_gather_kernel(int32_t tex_w, int64_t row_vec,
                                                         const GatherEntry *__restrict__ list, int64_t cap,
                                                         const uint32_t *__restrict__ count,
                                                         const TexRecord<S> *__restrict__ recs, int accumulate,
                                                         S *__restrict__ g_tex) {
    typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
    extern __shared__ __align__(16) char smem[];
    // two row buffers, addressed from the LDS base every time so accesses stay ds_*
    S *const acc = (S *)smem;
    u32x4_t *const acc4 = (u32x4_t *)smem;
    constexpr int kT = 256;                     // threads per role
    const int64_t e = 16 / (int64_t)sizeof(S);  // elements per 16-byte piece
    const int t = threadIdx.x, tt = t & (kT - 1);
    const bool writer = t >= kT;  // uniform per wave
    cptr<GatherEntry> L = as_const(list);
    const int64_t total = min((int64_t)as_const(count)[0], cap);
    const int64_t G = gridDim.x;
    for (int64_t i = t; i < 2 * row_vec; i += 2 * kT) acc4[i] = u32x4_t{0, 0, 0, 0};

    // loaders: entry k of this workgroup is list[blockIdx.x + k G]; each loader lane takes
    // the run of R = gather_run<S>() records at R tt of each chunk of kT R records.  The
    // first chunk of the current and the next entry swap roles between the two halves of
    // the unrolled loop (copying a register whose load is in flight would wait for it).
    constexpr int32_t kR = gather_run<S>(), kChunk = kT * kR;
    RecRun<S> ra, rb;
    int32_t ek = blockIdx.x;  // wave-uniform; kept in an SGPR so the list reads are scalar
    GatherEntry cur{-1, 0, 0};
    if (!writer) {
        if (ek < total) cur = load_entry(L, ek);
        load_run(ra, recs, cur, tt * kR);
    }
    lds_barrier();
    // step s: loaders add entry s into buffer s & 1; writers drain entry s - 1 from the
    // other buffer.  The loop ends after the step in which the loaders had no entry.
    auto step = [&](int64_t s, RecRun<S> &rc, RecRun<S> &rn) -> bool {
        const int b = (int)(s & 1);
        if (!writer) {
            const int32_t e2 = __builtin_amdgcn_readfirstlane(ek + (int32_t)G);
            GatherEntry nxt{-1, 0, 0};
            if (e2 < total) nxt = load_entry(L, e2);
            load_run(rn, recs, nxt, tt * kR);
            if (ek < total) {
                S *ab = acc + b * row_vec * e;
                add_run(ab, rc, tt * kR, cur.n, tex_w);
                // rows of more than kChunk records: further chunks in rc's registers (its
                // first chunk is added; the next step loads into it afresh)
                for (int32_t q = kChunk + tt * kR; q < cur.n; q += kChunk) {
                    load_run(rc, recs, cur, q);
                    add_run(ab, rc, q, cur.n, tex_w);
                }
            }
            ek = e2;
            cur = nxt;
        } else if (s >= 1) {
            const int32_t ew = __builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x + (s - 1) * G));
            if (ew < total) {
                const int64_t r = L[ew].row;
                const int64_t o = (int64_t)(b ^ 1) * row_vec;
                if (!accumulate) {
                    u32x4_t *dst = (u32x4_t *)(g_tex + r * row_vec * e);
#pragma unroll 4
                    for (int64_t i = tt; i < row_vec; i += kT) {
                        const u32x4_t v = acc4[o + i];
                        acc4[o + i] = u32x4_t{0, 0, 0, 0};
                        __builtin_nontemporal_store(v, dst + i);
                    }
                } else {
                    // accumulate mode reads the destination: these writers do wait for loads
                    S *d = g_tex + r * row_vec * e;
                    for (int64_t i = tt; i < row_vec * e; i += kT) {
                        d[i] += acc[o * e + i];
                        acc[o * e + i] = S(0);
                    }
                }
            }
        }
        lds_barrier();
        return blockIdx.x + s * G >= total;  // the loaders had no entry in this step
    };
    for (int64_t s = 0;; s += 2) {
        if (step(s, ra, rb) || step(s + 1, rb, ra)) break;
    }
}

// ------------------------------------- texture-gradient gather, adders and writers
// One workgroup is AD adder waves and WR writer waves sharing two LDS row buffers.
// Step s: the adders sum the records of list entry b + s G into buffer s & 1 with LDS
// atomics, while the writers stream the row of entry b + (s - 1) G out of the other
// buffer, clearing it behind them; one barrier per step.
//  * The buffers hold float64 sums: on this hardware an LDS float32 atomic add costs
//    about 3 clocks per active lane (~190 clocks per wave instruction, whatever the
//    addresses), a float64 one ~28 clocks per wave instruction.  Rows are rounded to
//    float32 once, as they are written.
//  * Each buffer has a bitmask of the texels its row's records touch (set by the
//    adders with LDS integer atomics).  A writer reads and clears only marked texels
//    and stores zeros elsewhere: a typical row touches a sixth of its texels, and
//    reading and clearing whole 80 KB float64 rows made LDS traffic the writers' bound.
//  * Streaming the rows out is the bound (every listed row is written once, 40 KB at
//    50x50 float32).  A record's 16 contributions take ~50 vector instructions, so the
//    adders split each row: records come in chunks of 64, one per lane, and adder a
//    takes chunks a, a + AD, a + 2 AD, ...  Each adder loads its first kG2Pre chunks of
//    row s + 2 in step s, into the third of three register sets that rotate over the
//    steps, so those loads have two steps to arrive; further chunks (rows of more than
//    64 AD kG2Pre records) load during the step, four at a time.  List entries are
//    read 16 steps ahead in register blocks (4 dwords per entry, one per lane).
//  * The writers issue no global loads, so nothing they do waits for their stores
//    (vmcnt counts loads and stores together); adder 0 hands them the row through LDS.
//    All loads are vector loads: the step barrier waits on lgkmcnt, which scalar loads
//    would share with the LDS traffic.
//  * LDS layout: texel t's four sums at doubles 4t .. 4t + 3 with the channel index
//    XOR-ed by g2_swz(t), so the atomics of one channel over random texels reach every
//    bank (unswizzled, channel c of every texel falls in the same few banks).  A writer
//    lane reads a texel as two 16-byte pieces and undoes the permutation in registers.
// 6 + 10 waves (16: the largest workgroup; its lanes fit in 120 registers), measured in one
// process against other splits: 1M Gaussians, 50x50 rows 5.08-5.10 ms against 5.35-5.50
// for 6 + 6, 5.08-5.12 for 5 + 11, 5.18-5.20 for 4 + 12, 5.21-5.28 for 7 + 9 and 5 + 9,
// 5.75-5.88 for 6 + 8 and 8 + 8; 8 cameras x 500k, 32x32 rows 6.9-7.2 ms, as for 6 + 6.
// Of 12 waves, 6 + 6 had measured best (4 + 8 5.56 ms, 8 + 4 6.7 ms).
constexpr int kG2Adders = 6;
constexpr int kG2Writers = 10;
constexpr int kG2Pre = 2;     // record chunks per adder and row loaded two steps ahead
constexpr int kG2Block = 16;  // list entries per register block

__device__ __forceinline__ uint32_t g2_swz(uint32_t t) { return ((t >> 2) ^ (t >> 4)) & 3u; }

// LDS bytes of the gather for rows of `texels` texels: two float64 row buffers, their
// texel bitmasks, the row of each buffer, and one zero texel (what writer lanes whose
// texel is untouched read: one broadcast read instead of a branch)
__host__ __device__ constexpr size_t g2_mask_words(int64_t texels) { return (size_t)(texels + 31) / 32; }
__host__ __device__ constexpr size_t g2_pad_offset(int64_t texels) {
    return (2 * (size_t)texels * 32 + 2 * 4 * g2_mask_words(texels) + 16 + 15) / 16 * 16;
}
__host__ __device__ constexpr size_t g2_lds_bytes(int64_t texels) { return g2_pad_offset(texels) + 32; }

struct G2Set {
    TexRecord<float> r[kG2Pre];
};

struct G2Entry {
    int32_t row, n;
    int64_t start;
};

// Block m of list entries: entry j is list[b + (16 m + j) G]; lane l holds dword l % 4
// (row, n, start lo, start hi) of entry l / 4.  Past the end: row -1, no records.
__device__ __forceinline__ int32_t g2_block_load(const GatherEntry *L, int64_t total, int64_t b, int64_t G,
                                                 int64_t m, int lane) {
    const int64_t k = b + (m * kG2Block + lane / 4) * G;
    const int32_t none = (lane & 3) == 0 ? -1 : 0;
    return k < total ? ((const int32_t *)(L + k))[lane & 3] : none;
}

__device__ __forceinline__ G2Entry g2_entry(int32_t blk, int j) {
    G2Entry e;
    e.row = __builtin_amdgcn_readlane(blk, 4 * j);
    e.n = __builtin_amdgcn_readlane(blk, 4 * j + 1);
    e.start = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane(blk, 4 * j + 3) << 32) |
                        (uint64_t)(uint32_t)__builtin_amdgcn_readlane(blk, 4 * j + 2));
    return e;
}

// this lane's record of chunk c of entry e, or the row's first (record 0 for an empty
// row) where it has none: every lane loads, so the loads in flight are counted exactly
__device__ __forceinline__ TexRecord<float> g2_rec(const TexRecord<float> *recs, const G2Entry &e, int32_t c,
                                                   int lane) {
    const int32_t q = c * kWave + lane;
    return recs[q < e.n ? e.start + q : (e.n > 0 ? e.start : 0)];
}

// one record's 16 contributions (4 texels x RGBA) into the float64 row, and its texels'
// bits in the row's mask
__device__ __forceinline__ void g2_add1(double *acc, uint32_t *mask, const TexRecord<float> &rec, bool ok,
                                        int32_t tex_w) {
    if (!ok) return;
    uint32_t t[4];
    t[0] = rec.o00 / 16u;
    t[1] = t[0] + (rec.flags & 1u);
    t[2] = t[0] + ((rec.flags & 2u) ? (uint32_t)tex_w : 0u);
    t[3] = t[2] + (rec.flags & 1u);
    Bilerp<float> bl;
    bl.wu = rec.wu;
    bl.wv = rec.wv;
    const BWeights<float> bw = bweights(bl);
    const float w[4] = {bw.w00, bw.w10, bw.w01, bw.w11};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t sw = g2_swz(t[k]);
        bool any = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float v = w[k] * rec.q[c];  // the float32 product of the float atomic path
            if (v != 0.0f) {
                atomicAdd(acc + 4 * t[k] + (c ^ sw), (double)v);
                any = true;
            }
        }
        if (any) atomicOr(mask + (t[k] >> 5), 1u << (t[k] & 31u));
    }
}

template <int AD, int WR>
__global__ __launch_bounds__((AD + WR) * kWave) void tex_gather2_kernel(int32_t tex_w, int32_t texels,
                                                                        const GatherEntry *__restrict__ list,
                                                                        int64_t cap, const uint32_t *__restrict__ count,
                                                                        const TexRecord<float> *__restrict__ recs,
                                                                        float *__restrict__ g_tex) {
    typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
    typedef double f64x2_t __attribute__((ext_vector_type(2)));
    constexpr int kThreads = (AD + WR) * kWave;
    extern __shared__ __align__(16) char smem[];
    double *const acc = (double *)smem;
    f64x2_t *const acc2 = (f64x2_t *)smem;
    const int32_t mw = (int32_t)g2_mask_words(texels);
    uint32_t *const masks = (uint32_t *)(smem + 2 * (size_t)texels * 32);
    int32_t *const slot = (int32_t *)(masks + 2 * mw);
    const int32_t pad = (int32_t)(g2_pad_offset(texels) / 16);  // f64x2 index of the zero texel
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;  // adders 0 .. AD - 1, then the writers
    const bool writer = wave >= AD;
    const int64_t total = min((int64_t)count[0], cap);
    const int64_t G = gridDim.x, b = blockIdx.x;
    // entries b + s G, s < n_add, are added in steps s; the writers drain entry s - 1 in step s
    const int64_t n_add = b < total ? (total - b + G - 1) / G : 0;
    if (n_add == 0) return;  // uniform over the workgroup
    const int64_t n_steps = n_add + 1;
    for (int32_t i = threadIdx.x; i < 4 * texels; i += kThreads) acc2[i] = f64x2_t{0.0, 0.0};
    for (int32_t i = threadIdx.x; i < 2 * mw; i += kThreads) masks[i] = 0u;
    if (threadIdx.x < 2) acc2[pad + threadIdx.x] = f64x2_t{0.0, 0.0};

    // adders: entry blocks of steps [16 m, 16 m + 16) and the next; three record sets
    // (rows s, s + 1, s + 2)
    G2Set x0, x1, x2;
    int32_t blk = 0, blk_next = 0;
    if (!writer) {
        blk = g2_block_load(list, total, b, G, 0, lane);
        blk_next = g2_block_load(list, total, b, G, 1, lane);
        const G2Entry e0 = g2_entry(blk, 0), e1 = g2_entry(blk, 1);
#pragma unroll
        for (int i = 0; i < kG2Pre; ++i) {
            x0.r[i] = g2_rec(recs, e0, wave + AD * i, lane);
            x1.r[i] = g2_rec(recs, e1, wave + AD * i, lane);
        }
    }
    lds_barrier();
    // step s: add row s from `cur`, load row s + 2 into `nxt2` (the set row s - 1 used)
    auto step = [&](int64_t s, G2Set &cur, G2Set &nxt2) -> bool {
        const int buf = (int)(s & 1);
        if (!writer) {
            const int j = (int)(s % kG2Block);
            if (j == 0 && s > 0) {
                blk = blk_next;
                blk_next = g2_block_load(list, total, b, G, s / kG2Block + 1, lane);
            }
            const G2Entry e = g2_entry(blk, j);
            const int j2 = j + 2;
            const G2Entry e2 = j2 < kG2Block ? g2_entry(blk, j2) : g2_entry(blk_next, j2 - kG2Block);
#pragma unroll
            for (int i = 0; i < kG2Pre; ++i) nxt2.r[i] = g2_rec(recs, e2, wave + AD * i, lane);
            double *ab = acc + (size_t)buf * texels * 4;
            uint32_t *mb = masks + buf * mw;
            if (e.n > 0) {
#pragma unroll
                for (int i = 0; i < kG2Pre; ++i)
                    g2_add1(ab, mb, cur.r[i], (wave + AD * i) * kWave + lane < e.n, tex_w);
                // chunks past the prefetched ones, four per round of loads
                for (int32_t c = wave + AD * kG2Pre; c * kWave < e.n; c += 4 * AD) {
                    TexRecord<float> r4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) r4[u] = g2_rec(recs, e, c + u * AD, lane);
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        g2_add1(ab, mb, r4[u], (c + u * AD) * kWave + lane < e.n, tex_w);
                }
            }
            if (wave == 0 && lane == 0) slot[buf] = e.row;
        } else if (s >= 1) {
            const int32_t row = slot[buf ^ 1];
            if (row >= 0) {
                const int32_t src = (buf ^ 1) * texels * 2;  // f64x2 index of the buffer
                uint32_t *mb = masks + (buf ^ 1) * mw;
                u32x4_t *dst = (u32x4_t *)(g_tex + (int64_t)row * texels * 4);
                // kB = 2 texels per lane and batch, measured in one process: with 6 + 10 waves,
                // 50x50 rows (1M) 5.04-5.09 ms against 5.12-5.13 for 4 and 5.19-5.23 with one
                // prefetched record chunk instead of two; 32x32 rows (8 cameras x 500k) 6.6-7.1
                // against 7.05-7.13.  With 6 + 6 waves, 4 had beaten 8 (32x32 rows 6.9-7.2 ms against
                // 7.4-7.7, and 122 registers against 162, which let the workgroup grow to 16 waves)
                constexpr int kStride = WR * kWave, kB = 2;
                const int wl = (wave - AD) * kWave + lane;  // lane among the writers
                for (int32_t base = wl; base < texels; base += kB * kStride) {
                    // a batch of kB texels per lane: mask words, then sums (a marked texel's
                    // pair of 16-byte pieces, else the zero texel), then stores, so each phase
                    // issues all its LDS accesses before waiting on any
                    uint32_t word[kB];
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        const int32_t t = min(base + i * kStride, texels - 1);
                        word[i] = mb[t >> 5];
                    }
                    // the 32 texels of a mask word lie in one 64-texel block, which one writer
                    // wave reads in one batch: the word is cleared once that wave has read it
                    // (the LDS accesses of a wave complete in order)
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        const int32_t t = base + i * kStride;
                        if (t < texels && (t & 31) == 0) mb[t >> 5] = 0u;
                    }
                    int32_t at[kB];
                    f64x2_t lo[kB], hi[kB];
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        const int32_t t = base + i * kStride;
                        const bool hit = t < texels && ((word[i] >> (t & 31)) & 1u);
                        at[i] = hit ? src + 2 * t : pad;
                        lo[i] = acc2[at[i]];
                        hi[i] = acc2[at[i] + 1];
                    }
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        if (at[i] != pad) {
                            acc2[at[i]] = f64x2_t{0.0, 0.0};
                            acc2[at[i] + 1] = f64x2_t{0.0, 0.0};
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kB; ++i) {
                        const int32_t t = base + i * kStride;
                        u32x4_t v = u32x4_t{__float_as_uint((float)lo[i].x), __float_as_uint((float)lo[i].y),
                                            __float_as_uint((float)hi[i].x), __float_as_uint((float)hi[i].y)};
                        const uint32_t sw = g2_swz((uint32_t)t);
                        // position p holds channel p ^ sw
                        if (sw & 1u) v = u32x4_t{v.y, v.x, v.w, v.z};
                        if (sw & 2u) v = u32x4_t{v.z, v.w, v.x, v.y};
                        if (t < texels) __builtin_nontemporal_store(v, dst + t);
                    }
                }
            }
        }
        lds_barrier();
        return s + 1 >= n_steps;
    };
    for (int64_t s = 0;; s += 3) {
        if (step(s, x0, x2) || step(s + 1, x1, x0) || step(s + 2, x2, x1)) break;
    }
}

// ------------------------------------------------------- row-wise texture Adam
// Selective Adam (adam.cu:16-47 semantics: no bias correction, rows whose
// visibility flag is clear are left alone) applied to bfloat16 textures straight
// from the backward's texture-gradient records, so no dense gradient exists.
// State per texel channel, 8 bytes:
//   hi  bf16   the render copy, = the float32 master rounded to nearest (ties away)
//   lo  int16  master bits - (hi << 16): hi and lo together are the float32 master
//   m,v bf16   Adam moments, stochastically rounded (unbiased; round-to-nearest
//              would freeze v, whose per-step change (1 - b2) is below half a bf16 ulp)
// One workgroup per row at a time: zero an LDS float row, add the row's records,
// then step the row with all waves.  The LDS row is single-buffered (64 KB at
// 64x64) so two workgroups share a CU and cover each other's barriers.

__device__ __forceinline__ uint32_t mix32(uint32_t x) {  // murmur3 finaliser
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// float32 -> bf16 bits, stochastic: add 16 random bits below the kept ones, truncate
__device__ __forceinline__ uint32_t bf16_stochastic(float x, uint32_t r16) {
    return (__float_as_uint(x) + (r16 & 0xffffu)) >> 16;
}

__device__ __forceinline__ void adam_texel(uint32_t &hi, uint32_t &lo, uint32_t &m, uint32_t &v, float g, float lr,
                                           float b1, float b2, float eps, uint32_t rnd) {
    float p = __uint_as_float((hi << 16) + (uint32_t)(int32_t)(int16_t)lo);
    float mm = __uint_as_float(m << 16), vv = __uint_as_float(v << 16);
    mm = b1 * mm + (1.0f - b1) * g;
    vv = b2 * vv + (1.0f - b2) * g * g;
    // hardware square root and reciprocal (about 1 ulp each) in place of the
    // correctly rounded forms: the step is a learning-rate-sized quantity
    p -= lr * mm * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(vv) + eps);
    const uint32_t pb = __float_as_uint(p);
    hi = (pb + 0x8000u) >> 16;
    lo = (pb - (hi << 16)) & 0xffffu;
    m = bf16_stochastic(mm, rnd);
    v = bf16_stochastic(vv, rnd >> 16);
}

__global__ __launch_bounds__(512) void tex_adam_kernel(int64_t N, int64_t rows_per_block, int32_t tex_w,
                                                       int64_t row_elems, const int64_t *__restrict__ rec_incl,
                                                       const int32_t *__restrict__ rec_cnt,
                                                       const int32_t *__restrict__ rec_cur,
                                                       const TexRecord<float> *__restrict__ recs,
                                                       const uint8_t *__restrict__ visible, uint16_t *__restrict__ hi,
                                                       uint16_t *__restrict__ lo, uint16_t *__restrict__ m,
                                                       uint16_t *__restrict__ v, float lr, float b1, float b2,
                                                       float eps, uint32_t seed) {
    typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
    extern __shared__ __align__(16) char smem[];
    float *const acc = (float *)smem;  // one float row, addressed from the LDS base (ds_* accesses)
    u32x4_t *const acc4 = (u32x4_t *)smem;
    const int64_t row_vec = row_elems / 4;  // 16-byte units of the float row
    const int64_t row_oct = row_elems / 8;  // 8-channel groups (16 bytes of each bf16 plane)
    const int t = threadIdx.x;
    constexpr int kT = 512;
    const uint32_t tb = 16;  // float texel stride of the record offsets
    cptr<int32_t> cnt = as_const(rec_cnt), cur = as_const(rec_cur);
    cptr<int64_t> incl = as_const(rec_incl);
    constexpr int kU = 4;  // groups per thread whose loads are issued together

    const int64_t r_begin = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r_end = min(N, r_begin + rows_per_block);
    for (int64_t row = r_begin; row < r_end; ++row) {
        // the visibility byte through a scalar load of its (4-byte aligned) word, but the
        // last partial word, which may end the buffer
        const uint32_t word = (row | 3) < N ? as_const((const uint32_t *)visible)[row >> 2]
                                            : (uint32_t)visible[row] << (8 * (row & 3));
        if (((word >> (8 * (row & 3))) & 0xffu) == 0) continue;  // uniform over the workgroup
        const int32_t nrec = min(cnt[row], cur[row]);  // records actually written
        if (nrec > 0) {
            for (int64_t i = t; i < row_vec; i += kT) acc4[i] = u32x4_t{0, 0, 0, 0};
            __syncthreads();
            const TexRecord<float> *rr = recs + (row ? incl[row - 1] : 0);
            for (int32_t r = t; r < nrec; r += kT) {
                const TexRecord<float> rec = rr[r];
                Bilerp<float> bl;
                bl.wu = rec.wu;
                bl.wv = rec.wv;
                bl.o00 = rec.o00;
                bl.du = (rec.flags & 1u) ? tb : 0u;
                bl.dv = (rec.flags & 2u) ? (uint32_t)tex_w * tb : 0u;
                const BWeights<float> bw = bweights(bl);
                const uint32_t e00 = bl.o00 / 4, e10 = (bl.o00 + bl.du) / 4;
                const uint32_t e01 = (bl.o00 + bl.dv) / 4, e11 = (bl.o00 + bl.du + bl.dv) / 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float qc = rec.q[c];
                    if (qc == 0.0f) continue;
                    atomicAdd(acc + c + e00, bw.w00 * qc);
                    atomicAdd(acc + c + e10, bw.w10 * qc);
                    atomicAdd(acc + c + e01, bw.w01 * qc);
                    atomicAdd(acc + c + e11, bw.w11 * qc);
                }
            }
            __syncthreads();
        }
        u32x4_t *H = (u32x4_t *)(hi + row * row_elems), *L = (u32x4_t *)(lo + row * row_elems);
        u32x4_t *Mv = (u32x4_t *)(m + row * row_elems), *Vv = (u32x4_t *)(v + row * row_elems);
        for (int64_t i0 = t; i0 < row_oct; i0 += kU * kT) {
            u32x4_t hv[kU], lv[kU], mv[kU], vv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = i0 + u * kT;
                if (i < row_oct) {
                    hv[u] = __builtin_nontemporal_load(H + i);
                    lv[u] = __builtin_nontemporal_load(L + i);
                    mv[u] = __builtin_nontemporal_load(Mv + i);
                    vv[u] = __builtin_nontemporal_load(Vv + i);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = i0 + u * kT;
                if (i >= row_oct) continue;
                float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                if (nrec > 0) {
                    const u32x4_t g0 = acc4[2 * i], g1 = acc4[2 * i + 1];
                    g[0] = __uint_as_float(g0.x); g[1] = __uint_as_float(g0.y);
                    g[2] = __uint_as_float(g0.z); g[3] = __uint_as_float(g0.w);
                    g[4] = __uint_as_float(g1.x); g[5] = __uint_as_float(g1.y);
                    g[6] = __uint_as_float(g1.z); g[7] = __uint_as_float(g1.w);
                }
                // one hash per 8 channels, spread to each by an odd multiplier and a shift
                const uint32_t base = mix32((uint32_t)(row * row_oct + i) ^ seed);
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    uint32_t r0 = base * (4u * w + 1u) + 0x9e3779b9u * w, r1 = base * (4u * w + 3u) + 0x7f4a7c15u * w;
                    r0 ^= r0 >> 15;
                    r1 ^= r1 >> 15;
                    uint32_t h0 = hv[u][w] & 0xffffu, h1 = hv[u][w] >> 16, l0 = lv[u][w] & 0xffffu,
                             l1 = lv[u][w] >> 16;
                    uint32_t m0 = mv[u][w] & 0xffffu, m1 = mv[u][w] >> 16, v0 = vv[u][w] & 0xffffu,
                             v1 = vv[u][w] >> 16;
                    adam_texel(h0, l0, m0, v0, g[2 * w], lr, b1, b2, eps, r0);
                    adam_texel(h1, l1, m1, v1, g[2 * w + 1], lr, b1, b2, eps, r1);
                    hv[u][w] = h0 | (h1 << 16);
                    lv[u][w] = l0 | (l1 << 16);
                    mv[u][w] = m0 | (m1 << 16);
                    vv[u][w] = v0 | (v1 << 16);
                }
                __builtin_nontemporal_store(hv[u], H + i);
                __builtin_nontemporal_store(lv[u], L + i);
                __builtin_nontemporal_store(mv[u], Mv + i);
                __builtin_nontemporal_store(vv[u], Vv + i);
            }
        }
        if (nrec > 0) __syncthreads();  // the LDS row is read before the next row zeroes it
    }
}

template <typename S> static size_t rec_bytes(int64_t C, int64_t N) {
    return (((size_t)(C * N) * kRec * sizeof(S)) + 255) / 256 * 256;
}
template <typename S> static size_t abc_bytes(int64_t C, int64_t N) {
    return (((size_t)(C * N) * kAbc * sizeof(S)) + 255) / 256 * 256;
}

template <typename S, int K>
static void launch_pack(const RasterArgs<S> &a, hipStream_t st) {
    const int64_t n = a.E;
    if (n == 0) return;  // packed layout without entries
    pack_kernel<S, K><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, a.N, a.gids, a.means2d, a.rt, a.opac, a.colors,
                                                                    a.normals, (S *)a.rec);
}
// Measurement hook (tgs_set_texel_marks): the byte map the forward marks texel reads in.
static uint8_t *g_texel_marks;
static int64_t g_texel_marks_n;

// Measurement hook (tgs_set_raster_events): events recorded on the launch stream right
// before and after the rasterizer kernel itself, [0] forward, [1] backward, and around
// the texture-gradient gather of the backward's records, [2].
static hipEvent_t g_raster_ev[3][2];

static void record_raster_event(int which, int end, hipStream_t st) {
    if (g_raster_ev[which][end]) (void)hipEventRecord(g_raster_ev[which][end], st);
}

template <typename S, int K, int TEX> static void launch_fwd(const RasterArgs<S> &a, unsigned grid, int bs, hipStream_t st) {
    launch_pack<S, K>(a, st);
    record_raster_event(0, 0, st);
    raster_fwd_kernel<S, K, TEX><<<grid, bs, 0, st>>>(a);
    record_raster_event(0, 1, st);
}
template <typename S, int K, int TEX> static void launch_bwd(const RasterArgs<S> &a, unsigned grid, int bs, hipStream_t st) {
    launch_pack<S, K>(a, st);
    const size_t rows = TEX == kTexMip2 ? 4 : 2;
    const size_t stage = TEX ? (size_t)(bs / kWave) * rows * kWave * sizeof(TexRec<S>) : 0;
    if (stage > 64 * 1024) {  // mip2 in float64: 72 KB of staging per 256-thread block
        static bool raised = false;
        if (!raised)
            raised = hipFuncSetAttribute((const void *)raster_bwd_kernel<S, K, TEX, 0>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)stage) == hipSuccess;
    }
    // the lean instantiations: float32, texture gradient through records for every row,
    // reference uv rule (they need no LDS staging); 1 reads the texels, 2 the forward's
    // pair samples (K <= 3)
    constexpr bool kLeanOk = sizeof(S) == 4 && (TEX == kTexNative || TEX == kTexBF16);
    const bool lean = kLeanOk && a.recs && !a.rec_rows && a.uv_mode == TGS_UV_NONE;
    record_raster_event(1, 0, st);
    if constexpr (kLeanOk) {
        if (lean && !a.samples) raster_bwd_kernel<S, K, TEX, 1><<<grid, bs, 0, st>>>(a);
        else if (K <= 3 && lean && a.samples && a.live) raster_bwd_kernel<S, K, TEX, (K <= 3 ? 2 : 1)><<<grid, bs, 0, st>>>(a);
        else raster_bwd_kernel<S, K, TEX, 0><<<grid, bs, stage, st>>>(a);
    } else {
        raster_bwd_kernel<S, K, TEX, 0><<<grid, bs, stage, st>>>(a);
    }
    record_raster_event(1, 1, st);
}

// Writes the longest-first tile order into a.order_ws (room for C * tw * th tiles, at most
// kOrderMaxTiles; the caller checks both).
template <typename S> static int launch_tile_order(const RasterArgs<S> &a, hipStream_t st) {
    const int32_t n_tiles = (int32_t)(a.C * a.tw * a.th);
    tile_order_kernel<<<1, kOrderThreads, 0, st>>>(n_tiles, a.offsets, a.n_isects, a.masks, a.order_ws);
    TGS_CHECK_LAUNCH("rasterize tile order");
    return TGS_OK;
}

template <typename S, bool BWD>
static int dispatch(const RasterArgs<S> &a0, int K, int tex, hipStream_t st) {
    // a tile's ts * ts pixels over workgroups of at most 256 threads (whole wavefronts);
    // every workgroup of a tile walks the tile's whole list for its own pixels
    RasterArgs<S> a = a0;
    const int bs = std::min(256, kWave * (int)ceil_div(a.ts * a.ts, kWave));
    a.bpt = (int32_t)ceil_div(a.ts * a.ts, bs);
    a.tile_order = nullptr;
    if (a.order_ws) {
        const int rc = launch_tile_order(a, st);
        if (rc) return rc;
        a.tile_order = a.order_ws;
    }
    const unsigned grid = (unsigned)(a.C * a.tw * a.th * a.bpt);
    if (grid == 0) {  // no kernel to fuse the row fill into
        return a.zdst ? tgs_zero_fill_rows(a.zdst, (size_t)a.zrow16 * 16, a.N, a.zmask, a.zsel, st) : TGS_OK;
    }
    // fill stores per visited entry: each wave drains what is left of its rows at its end.
    // Pacing measured at 1M x 50x50 (backward, 18.4 GB of rows): 2 to 12 stores per entry
    // give the same kernel time; the walk and the fill share the memory bandwidth.
    if (a.zdst && a.zsteps == 0) a.zsteps = BWD ? 2 : 1;
// bf16 texture storage is instantiated for float compute only
#define TGS_CASE(KK)                                                                      \
    case KK:                                                                              \
        if (BWD) {                                                                        \
            if (tex == kTexNative) launch_bwd<S, KK, kTexNative>(a, grid, bs, st);        \
            else if (tex == kTexMip2) launch_bwd<S, KK, kTexMip2>(a, grid, bs, st);       \
            else if constexpr (sizeof(S) == 4) {                                          \
                if (tex == kTexBF16) launch_bwd<S, KK, kTexBF16>(a, grid, bs, st);        \
                else launch_bwd<S, KK, kTexNone>(a, grid, bs, st);                        \
            } else launch_bwd<S, KK, kTexNone>(a, grid, bs, st);                          \
        } else {                                                                          \
            if (tex == kTexNative) launch_fwd<S, KK, kTexNative>(a, grid, bs, st);        \
            else if (tex == kTexMip2) launch_fwd<S, KK, kTexMip2>(a, grid, bs, st);       \
            else if constexpr (sizeof(S) == 4) {                                          \
                if (tex == kTexBF16) launch_fwd<S, KK, kTexBF16>(a, grid, bs, st);        \
                else launch_fwd<S, KK, kTexNone>(a, grid, bs, st);                        \
            } else launch_fwd<S, KK, kTexNone>(a, grid, bs, st);                          \
        }                                                                                 \
        break;
    switch (K) {
        TGS_CASE(1)
        TGS_CASE(2)
        TGS_CASE(3)
        TGS_CASE(4)
        default:
            set_error("rasterize: Unsupported number of channels: %d", K);
            return TGS_ERR_UNSUPPORTED;
    }
#undef TGS_CASE
    TGS_CHECK_LAUNCH(BWD ? "rasterize_bwd" : "rasterize_fwd");
    return TGS_OK;
}

template <typename S>
static void fill_common(RasterArgs<S> &a, int64_t C, int64_t N, const void *means2d, const void *rt,
                        const void *colors, const void *opac, const void *normals, const void *textures,
                        int64_t tex_h, int64_t tex_w, double trx, double try_, const void *bg, const uint8_t *masks,
                        int32_t W, int32_t H, int32_t ts, int32_t tw, int32_t th, const int32_t *offsets,
                        const int32_t *flatten_ids, int64_t n_isects, double g_weight, void *workspace,
                        void *zdst, const uint8_t *zmask, size_t zrow_bytes, int zsel, const int64_t *gids,
                        int64_t nnz) {
    // zsel: low byte the mask value to fill, bits 8-15 the rows (by row & 7) to take (0 there =
    // all), bits 16-23 the fill stores per visited list entry (0 there = fill_steps' choice)
    memset(&a, 0, sizeof(a));
    a.gids = gids;
    a.E = gids ? nnz : C * N;
    a.zdst = zdst;
    a.zmask = zmask;
    a.zrow16 = (int64_t)(zrow_bytes / 16);
    a.zsel = zsel & 0xff;
    a.zres = (zsel >> 8) & 0xff ? (zsel >> 8) & 0xff : 0xff;
    a.zsteps = (zsel >> 16) & 0xff;
    a.C = C; a.N = N; a.n_isects = n_isects;
    a.means2d = (const S *)means2d; a.rt = (const S *)rt; a.colors = (const S *)colors;
    a.opac = (const S *)opac; a.normals = (const S *)normals; a.textures = (const S *)textures;
    a.bg = (const S *)bg; a.masks = masks;
    a.rec = (const S *)workspace;
    a.g_abc = (S *)((char *)workspace + rec_bytes<S>(1, a.E));
    if (textures) {
        a.tex_h = (int32_t)tex_h; a.tex_w = (int32_t)tex_w;
        a.tex_stride = tex_h * tex_w * 4;
        a.trx = (S)trx; a.try_ = (S)try_;
        a.umax = (S)(tex_w - 1); a.vmax = (S)(tex_h - 1);
        a.su = (S)((double)(tex_w - 1) / (2.0 * trx));
        a.sv = (S)((double)(tex_h - 1) / (2.0 * try_));
        a.uh = (S)(0.5 * (double)(tex_w - 1));
        a.vh = (S)(0.5 * (double)(tex_h - 1));
        a.g_weight = (S)g_weight;
        a.vis_bias = S(0.998) - (S)g_weight;
    } else {
        a.g_weight = S(1);
        a.vis_bias = S(0);
    }
    a.W = W; a.H = H; a.ts = ts; a.tw = tw; a.th = th;
    a.offsets = offsets; a.flatten_ids = flatten_ids;
}

static int check_common(int64_t C, int64_t N, int32_t K, const void *textures, int64_t tex_h, int64_t tex_w,
                        double trx, double try_, int32_t W, int32_t H, int32_t ts, int32_t tw, int32_t th,
                        int64_t n_isects) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "rasterize: bad sizes C=%lld N=%lld",
                  (long long)C, (long long)N);
    TGS_CHECK_ARG(K >= 1 && K <= 4, "rasterize: Unsupported number of channels: %d", K);
    // the reference launches tile_size^2 threads per tile
    // (rasterize_to_pixels_textured_gaussians_2dgs_fwd.cu:510), so at most 32
    TGS_CHECK_ARG(ts >= 1 && ts <= 32, "rasterize: tile_size %d unsupported (tile_size^2 must be 1..1024)", ts);
    TGS_CHECK_ARG((int64_t)tw * ts >= W && (int64_t)th * ts >= H, "rasterize: tile grid does not cover the image");
    TGS_CHECK_ARG(n_isects >= 0 && n_isects < (int64_t(1) << 31), "rasterize: bad n_isects");
    TGS_CHECK_ARG(!textures || (tex_h >= 1 && tex_w >= 1 && tex_h * tex_w * 4 * 8 < (int64_t(1) << 31)),
                  "rasterize: bad texture size");
    TGS_CHECK_ARG(!textures || (trx > 0 && try_ > 0), "rasterize: texture ranges must be positive");
    return TGS_OK;
}

static int tex_mode_of(const void *textures, int dtype) {
    if (!textures) return kTexNone;
    if (dtype & TGS_TEX_BF16) return kTexBF16;
    return (dtype & TGS_TEX_MIP2) ? kTexMip2 : kTexNative;
}

// log2 of the resolution of square power-of-two textures (<= 4096), else -1
static int mip_log_of(int64_t tex_h, int64_t tex_w) {
    if (tex_h != tex_w || tex_h < 1 || tex_h > 4096 || (tex_h & (tex_h - 1))) return -1;
    int l = 0;
    while ((int64_t(1) << l) < tex_h) ++l;
    return l;
}

// mip2: the textures are the flat pyramid [N, (4^(L+1) - 1) / 3, 4]
template <typename S> static void set_mip(RasterArgs<S> &a, int tex_mode, int64_t tex_h, int64_t tex_w) {
    if (tex_mode != kTexMip2) return;
    a.mip_log = mip_log_of(tex_h, tex_w);
    a.tex_stride = ((int64_t(1) << (2 * (a.mip_log + 1))) - 1) / 3 * 4;
}

// pair samples: float32 compute, native or bfloat16 textures, at most 3 colour channels
// (the fourth value is the alpha sample), with the composited-lane masks
static int check_pair_samples(const char *who, const void *samples, const void *live, int dtype, int tex_mode,
                              int32_t K) {
    TGS_CHECK_ARG(!samples || (live && dtype == TGS_F32 && (tex_mode == kTexNative || tex_mode == kTexBF16) && K <= 3 &&
                               ((uintptr_t)samples % 16) == 0),
                  "%s: pair_samples need live_entries, float32 compute, bilinear textures (float32 or bfloat16), "
                  "at most 3 channels and 16-byte alignment", who);
    return TGS_OK;
}

static int check_zero_rows(const void *dst, const uint8_t *mask, size_t row_bytes) {
    TGS_CHECK_ARG(!dst || (mask && ((uintptr_t)dst % 16) == 0 && row_bytes > 0 && row_bytes % 16 == 0),
                  "rasterize: zero_rows_dst needs a mask, 16-byte alignment and rows of a multiple of 16 bytes");
    return TGS_OK;
}

constexpr size_t kGatherMaxLds = 160 * 1024;  // LDS per workgroup: two texture rows in the gather

// two texture rows fit in one CU's LDS
template <typename S> static bool gather_fits(int64_t tex_h, int64_t tex_w) {
    return 2 * (size_t)(tex_h * tex_w * 4) * sizeof(S) <= kGatherMaxLds;
}

template <typename S>
static int launch_gather(int64_t N, int64_t tex_h, int64_t tex_w, const RasterArgs<S> &a, const uint8_t *mask,
                         int overwrite, hipStream_t st) {
    const int64_t row_bytes = tex_h * tex_w * 4 * (int64_t)sizeof(S);
    const size_t lds = 2 * (size_t)row_bytes;  // two row buffers
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)tex_gather_kernel<S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess) {
        set_error("rasterize_bwd: cannot reserve %zu bytes of LDS for the texture gather", lds);
        return TGS_ERR_LAUNCH;
    }
    // the row list goes where the per-entry records were: the backward kernel, their only
    // reader, has run (16 bytes per row; with the mask of tgs_mark_rows the rows to write
    // are at most the entries, 80 bytes each or more)
    if (a.E == 0) return TGS_OK;  // no entries: no listed rows, no records
    record_raster_event(2, 0, st);
    uint32_t *count = (uint32_t *)a.rec;
    GatherEntry *list = (GatherEntry *)((char *)a.rec + 16);
    const int64_t cap = ((int64_t)rec_bytes<S>(1, a.E) - 16) / (int64_t)sizeof(GatherEntry);
    zero_kernel<uint32_t><<<1, 64, 0, st>>>(1, count);
    if (N > 0)
        gather_list_kernel<<<(unsigned)ceil_div(N, kGatherListRows), kGatherListThreads, 0, st>>>(
            N, mask, a.rec_cnt, a.rec_cur, a.rec_incl, list, cap, count);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t lds2 = g2_lds_bytes(tex_h * tex_w);
    if (std::is_same<S, float>::value && overwrite && lds2 <= kGatherMaxLds) {
        // adder/writer workgroups, one or more per CU
        if (lds2 > 64 * 1024 && hipFuncSetAttribute((const void *)tex_gather2_kernel<kG2Adders, kG2Writers>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2) != hipSuccess) {
            set_error("rasterize_bwd: cannot reserve %zu bytes of LDS for the texture gather", lds2);
            return TGS_ERR_LAUNCH;
        }
        const int per_cu = (int)std::min<size_t>(8, kGatherMaxLds / lds2);
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(N, (int64_t)cus * per_cu));
        tex_gather2_kernel<kG2Adders, kG2Writers><<<(unsigned)blocks, (kG2Adders + kG2Writers) * kWave, lds2, st>>>(
            (int32_t)tex_w, (int32_t)(tex_h * tex_w), list, cap, count, (const TexRecord<float> *)a.recs,
            (float *)a.g_tex);
        TGS_CHECK_LAUNCH("rasterize_bwd texture gather");
        record_raster_event(2, 1, st);
        return TGS_OK;
    }
    // as many 512-thread workgroups per CU as their buffers allow (at most 2: 16 waves)
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, kGatherMaxLds / std::max<size_t>(lds, 1)));
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(N, (int64_t)cus * per_cu));
    tex_gather_kernel<S><<<(unsigned)blocks, 512, lds, st>>>((int32_t)tex_w, row_bytes / 16, list, cap, count,
                                                             (const TexRecord<S> *)a.recs, overwrite ? 0 : 1, a.g_tex);
    TGS_CHECK_LAUNCH("rasterize_bwd texture gather");
    record_raster_event(2, 1, st);
    return TGS_OK;
}

static size_t workspace_bytes(int dtype, int64_t C, int64_t N) {
    return dtype == TGS_F64 ? rec_bytes<double>(C, N) + abc_bytes<double>(C, N)
                            : rec_bytes<float>(C, N) + abc_bytes<float>(C, N);
}

// Bytes of the optional tile-order region that follows the workspace proper (0 when the
// tiles exceed what tile_order_kernel sorts).
static size_t tile_order_bytes(int64_t C, int64_t tw, int64_t th) {
    const int64_t n = C * tw * th;
    return n > 0 && n <= kOrderMaxTiles ? ((size_t)n * sizeof(int32_t) + 255) / 256 * 256 : 0;
}

// The tile-order region of a workspace of `size` bytes, or null when it has no room for it.
static int32_t *order_region(void *workspace, size_t size, int dtype, int64_t E, int64_t C, int64_t tw, int64_t th) {
    const size_t base = workspace_bytes(dtype, 1, E), ob = tile_order_bytes(C, tw, th);
    return workspace && ob > 0 && size >= base + ob ? (int32_t *)((char *)workspace + base) : nullptr;
}

}  // namespace tgs

using namespace tgs;

extern "C" size_t tgs_rasterize_workspace_bytes(int dtype, int64_t C, int64_t N) { return workspace_bytes(dtype, C, N); }

extern "C" size_t tgs_rasterize_tile_order_bytes(int64_t C, int32_t tile_width, int32_t tile_height) {
    return C > 0 && tile_width > 0 && tile_height > 0 ? tile_order_bytes(C, tile_width, tile_height) : 0;
}

extern "C" int tgs_set_raster_events(int which, void *begin, void *end) {
    TGS_CHECK_ARG(which >= 0 && which <= 2,
                  "set_raster_events: which must be 0 (forward), 1 (backward) or 2 (texture gather)");
    TGS_CHECK_ARG((begin == nullptr) == (end == nullptr), "set_raster_events: give both events or neither");
    g_raster_ev[which][0] = (hipEvent_t)begin;
    g_raster_ev[which][1] = (hipEvent_t)end;
    return TGS_OK;
}

extern "C" int tgs_set_texel_marks(uint8_t *marks, int64_t n_texels) {
    TGS_CHECK_ARG((marks == nullptr) == (n_texels == 0) && n_texels >= 0,
                  "set_texel_marks: give the map and its size, or NULL and 0");
    g_texel_marks = marks;
    g_texel_marks_n = n_texels;
    return TGS_OK;
}

extern "C" size_t tgs_tex_record_bytes(int dtype) {
    return dtype == TGS_F64 ? sizeof(TexRecord<double>) : sizeof(TexRecord<float>);
}

extern "C" int tgs_texture_gather_fast(int dtype, int64_t tex_h, int64_t tex_w) {
    return (dtype & ~TGS_TEX_BF16) == TGS_F32 && tex_h > 0 && tex_w > 0 && tex_h * tex_w <= (int64_t)1 << 20 &&
                   g2_lds_bytes(tex_h * tex_w) <= kGatherMaxLds
               ? 1
               : 0;
}

extern "C" int tgs_texture_adam_rows(int64_t N, int32_t tex_h, int32_t tex_w, const int32_t *tex_counts,
                                     const int64_t *tex_rec_incl, const int32_t *tex_rec_cursor,
                                     const void *tex_records, const uint8_t *visibility, uint16_t *tex_hi,
                                     uint16_t *tex_lo, uint16_t *exp_avg, uint16_t *exp_avg_sq, double lr, double b1,
                                     double b2, double eps, uint32_t seed, tgs_stream_t stream) {
    TGS_CHECK_ARG(N >= 0 && tex_h >= 1 && tex_w >= 1, "texture_adam_rows: bad sizes");
    const int64_t row_elems = (int64_t)tex_h * tex_w * 4;
    TGS_CHECK_ARG(gather_fits<float>(tex_h, tex_w) && row_elems % 8 == 0,
                  "texture_adam_rows: a texture must hold an even number of texels and two rows must fit in LDS");
    TGS_CHECK_ARG(((uintptr_t)visibility % 4) == 0, "texture_adam_rows: visibility must be 4-byte aligned");
    TGS_CHECK_ARG(N == 0 || (tex_counts && tex_rec_incl && tex_rec_cursor && tex_records && visibility && tex_hi &&
                             tex_lo && exp_avg && exp_avg_sq),
                  "texture_adam_rows: null buffer");
    TGS_CHECK_ARG(((uintptr_t)tex_hi | (uintptr_t)tex_lo | (uintptr_t)exp_avg | (uintptr_t)exp_avg_sq) % 16 == 0,
                  "texture_adam_rows: state planes must be 16-byte aligned");
    if (N == 0) return TGS_OK;
    const size_t lds = (size_t)row_elems * sizeof(float);  // one float row
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)tex_adam_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess) {
        set_error("texture_adam_rows: cannot reserve %zu bytes of LDS", lds);
        return TGS_ERR_LAUNCH;
    }
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, kGatherMaxLds / std::max<size_t>(lds, 1)));
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(N, (int64_t)cus * per_cu));
    const int64_t rows_per_block = (N + blocks - 1) / blocks;
    tex_adam_kernel<<<(unsigned)blocks, 512, lds, (hipStream_t)stream>>>(
        N, rows_per_block, tex_w, row_elems, tex_rec_incl, tex_counts, tex_rec_cursor,
        (const TexRecord<float> *)tex_records, visibility, tex_hi, tex_lo, exp_avg, exp_avg_sq, (float)lr, (float)b1,
        (float)b2, (float)eps, seed);
    TGS_CHECK_LAUNCH("texture_adam_rows");
    return TGS_OK;
}

extern "C" int tgs_rasterize_fwd(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                                 const void *ray_transforms, const void *colors, const void *opacities,
                                 const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                                 double texture_range_x, double texture_range_y, const void *backgrounds,
                                 const uint8_t *masks, const int64_t *gaussian_ids, int64_t nnz, int32_t width,
                                 int32_t height, int32_t tile_size,
                                 int32_t tile_width, int32_t tile_height, const int32_t *tile_offsets,
                                 const int32_t *flatten_ids, int64_t n_isects, double gs_contrib_threshold,
                                 double g_weight, void *render_colors, void *render_alphas, void *render_normals,
                                 void *render_distort, void *render_median, int32_t *last_ids, int32_t *median_ids,
                                 void *contrib_sum, void *contrib_count, void *render_accum_depth,
                                 void *render_trans, uint64_t *live_entries, void *pair_samples,
                                 int32_t *tex_counts, void *zero_rows_dst,
                                 const uint8_t *zero_rows_mask,
                                 size_t zero_row_bytes, int zero_rows_select, void *workspace, size_t workspace_size,
                                 tgs_stream_t stream) {
    const int tex_mode = tex_mode_of(textures, dtype);
    const int cull = (dtype & TGS_FWD_NO_CULL) ? 0 : 1;
    dtype &= ~(TGS_TEX_BF16 | TGS_TEX_MIP2 | TGS_FWD_NO_CULL);
    TGS_CHECK_ARG(tex_mode != kTexBF16 || dtype == TGS_F32, "rasterize_fwd: bfloat16 textures need float32 compute");
    TGS_CHECK_ARG(tex_mode != kTexMip2 || mip_log_of(tex_h, tex_w) >= 0,
                  "rasterize_fwd: the mip2 filter needs square power-of-two textures up to 4096");
    int rc = check_common(C, N, K, textures, tex_h, tex_w, texture_range_x, texture_range_y, width, height,
                          tile_size, tile_width, tile_height, n_isects);
    if (rc) return rc;
    const int64_t E = gaussian_ids ? nnz : C * N;  // per-entry inputs
    TGS_CHECK_ARG(!gaussian_ids || (nnz >= 0 && nnz < (int64_t(1) << 31)), "rasterize_fwd: bad nnz");
    TGS_CHECK_ARG((workspace || workspace_bytes(dtype, 1, E) == 0) && workspace_size >= workspace_bytes(dtype, 1, E),
                  "rasterize_fwd: workspace of %zu bytes required", workspace_bytes(dtype, 1, E));
    rc = check_zero_rows(zero_rows_dst, zero_rows_mask, zero_row_bytes);
    if (rc) return rc;
    rc = check_pair_samples("rasterize_fwd", pair_samples, live_entries, dtype, tex_mode, K);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
#define TGS_FWD(S)                                                                                              \
    {                                                                                                           \
        RasterArgs<S> a;                                                                                        \
        fill_common<S>(a, C, N, means2d, ray_transforms, colors, opacities, normals, textures, tex_h, tex_w,     \
                       texture_range_x, texture_range_y, backgrounds, masks, width, height, tile_size,          \
                       tile_width, tile_height, tile_offsets, flatten_ids, n_isects, g_weight, workspace,        \
                       zero_rows_dst, zero_rows_mask, zero_row_bytes, zero_rows_select, gaussian_ids, nnz);     \
        a.order_ws = order_region(workspace, workspace_size, dtype, E, C, tile_width, tile_height);             \
        set_mip<S>(a, tex_mode, tex_h, tex_w);                                                                 \
        a.thr = (S)gs_contrib_threshold;                                                                        \
        a.out_c = (S *)render_colors; a.out_a = (S *)render_alphas; a.out_n = (S *)render_normals;              \
        a.out_d = (S *)render_distort; a.out_m = (S *)render_median; a.last_ids = last_ids;                     \
        a.median_ids = median_ids; a.csum = (S *)contrib_sum; a.ccnt = (S *)contrib_count;                      \
        a.out_accd = (S *)render_accum_depth;                                                                   \
        a.out_T = (S *)render_trans;                                                                            \
        a.cull = cull;                                                                                          \
        a.live = live_entries;                                                                                  \
        a.samples = (S *)pair_samples;                                                                          \
        a.tex_cnt = textures ? tex_counts : nullptr;                                                            \
        a.tex_seen = (tex_mode == kTexNative && g_texel_marks && g_texel_marks_n >= N * tex_h * tex_w)          \
                         ? g_texel_marks : nullptr;                                                             \
        return dispatch<S, false>(a, K, tex_mode, st);                                                          \
    }
    if (dtype == TGS_F32) TGS_FWD(float)
    if (dtype == TGS_F64) TGS_FWD(double)
#undef TGS_FWD
    set_error("rasterize_fwd: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}

namespace tgs {

// Workspace of tgs_rasterize_fwd_noalloc: a file-scope device array, one copy per
// GPU, resident from the library's first use on that GPU until the process ends.
// In host code `scratch` names a host-side stand-in; kernels get the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[TGS_RASTERIZE_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device

// The current device's address of `scratch`, looked up on the first call per device.
static int scratch_address(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("rasterize_fwd_noalloc: no usable current device (%d)", dev);
        return TGS_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("rasterize_fwd_noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return TGS_ERR_LAUNCH;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return TGS_OK;
}

}  // namespace tgs

extern "C" int tgs_rasterize_fwd_noalloc(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                                         const void *ray_transforms, const void *colors, const void *opacities,
                                         const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                                         double texture_range_x, double texture_range_y, const void *backgrounds,
                                         const uint8_t *masks, const int64_t *gaussian_ids, int64_t nnz,
                                         int32_t width, int32_t height, int32_t tile_size, int32_t tile_width,
                                         int32_t tile_height, const int32_t *tile_offsets, const int32_t *flatten_ids,
                                         int64_t n_isects, double gs_contrib_threshold, double g_weight,
                                         void *render_colors, void *render_alphas, void *render_normals,
                                         void *render_distort, void *render_median, int32_t *last_ids,
                                         int32_t *median_ids, void *contrib_sum, void *contrib_count,
                                         void *render_accum_depth, void *render_trans, uint64_t *live_entries,
                                         void *pair_samples,
                                         int32_t *tex_counts, void *zero_rows_dst, const uint8_t *zero_rows_mask,
                                         size_t zero_row_bytes, int zero_rows_select, tgs_stream_t stream) {
    constexpr size_t cap = TGS_RASTERIZE_NOALLOC_SCRATCH_BYTES;
    const int base = dtype & ~(TGS_TEX_BF16 | TGS_TEX_MIP2 | TGS_FWD_NO_CULL);
    // sizes tgs_rasterize_fwd rejects are left to its own checks
    if (C >= 0 && N >= 0 && C < (int64_t(1) << 31) && N < (int64_t(1) << 31) && nnz < (int64_t(1) << 31) &&
        (base == TGS_F32 || base == TGS_F64)) {
        const int64_t E = gaussian_ids ? nnz : C * N;
        const size_t need = E > 0 ? workspace_bytes(base, 1, E) : 0;
        TGS_CHECK_ARG(need <= cap,
                      "rasterize_fwd_noalloc: %lld entries need %zu bytes of workspace; the static scratch holds "
                      "%zu bytes per GPU (%zu bytes per entry of this dtype)",
                      (long long)E, need, cap, (size_t)(kRec + kAbc) * (base == TGS_F64 ? 8 : 4));
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    return tgs_rasterize_fwd(dtype, C, N, K, means2d, ray_transforms, colors, opacities, normals, textures, tex_h,
                             tex_w, texture_range_x, texture_range_y, backgrounds, masks, gaussian_ids, nnz, width,
                             height, tile_size, tile_width, tile_height, tile_offsets, flatten_ids, n_isects,
                             gs_contrib_threshold, g_weight, render_colors, render_alphas, render_normals,
                             render_distort, render_median, last_ids, median_ids, contrib_sum, contrib_count,
                             render_accum_depth, render_trans, live_entries, pair_samples, tex_counts, zero_rows_dst,
                             zero_rows_mask, zero_row_bytes, zero_rows_select, ws, cap, stream);
}

extern "C" int tgs_rasterize_bwd(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                                 const void *ray_transforms, const void *colors, const void *opacities,
                                 const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                                 double texture_range_x, double texture_range_y, const void *backgrounds,
                                 const uint8_t *masks, const int64_t *gaussian_ids, int64_t nnz, int32_t width,
                                 int32_t height, int32_t tile_size,
                                 int32_t tile_width, int32_t tile_height, const int32_t *tile_offsets,
                                 const int32_t *flatten_ids, int64_t n_isects, double g_weight,
                                 const void *render_colors, const void *render_alphas, const int32_t *last_ids,
                                 const int32_t *median_ids, const void *render_accum_depth,
                                 const void *render_trans, const uint64_t *live_entries, const void *pair_samples,
                                 const int32_t *tex_counts,
                                 const int64_t *tex_rec_incl, int32_t *tex_rec_cursor, void *tex_records,
                                 const void *v_render_colors, const void *v_render_alphas,
                                 const void *v_render_normals, const void *v_render_distort,
                                 const void *v_render_median, int uv_grad_mode, void *v_means2d_abs, void *v_means2d,
                                 void *v_ray_transforms, void *v_colors, void *v_opacities, void *v_textures,
                                 void *v_normals, void *v_densify, void *zero_rows_dst,
                                 const uint8_t *zero_rows_mask, size_t zero_row_bytes, int zero_rows_select,
                                 void *workspace, size_t workspace_size, tgs_stream_t stream) {
    const int tex_mode = tex_mode_of(textures, dtype);
    const int rec_rows = (dtype & TGS_BWD_RECORD_ROWS) ? 1 : 0;
    dtype &= ~(TGS_TEX_BF16 | TGS_TEX_MIP2 | TGS_BWD_RECORD_ROWS);
    TGS_CHECK_ARG(tex_mode != kTexBF16 || dtype == TGS_F32, "rasterize_bwd: bfloat16 textures need float32 compute");
    TGS_CHECK_ARG(tex_mode != kTexMip2 || (mip_log_of(tex_h, tex_w) >= 0 && uv_grad_mode == TGS_UV_NONE &&
                                          v_textures && !tex_records),
                  "rasterize_bwd: the mip2 filter needs square power-of-two textures up to 4096, the reference "
                  "gradient rule and v_textures");
    int rc = check_common(C, N, K, textures, tex_h, tex_w, texture_range_x, texture_range_y, width, height,
                          tile_size, tile_width, tile_height, n_isects);
    if (rc) return rc;
    TGS_CHECK_ARG(uv_grad_mode >= 0 && uv_grad_mode <= 2, "rasterize_bwd: bad uv_grad_mode %d", uv_grad_mode);
    // without v_textures the texture gradient leaves only as records (for tgs_texture_adam_rows)
    const bool records_only = textures && !v_textures;
    TGS_CHECK_ARG(!records_only || (tex_records && tex_counts && tex_rec_incl && tex_rec_cursor && !zero_rows_dst &&
                                    (dtype == TGS_F64 ? gather_fits<double>(tex_h, tex_w)
                                                      : gather_fits<float>(tex_h, tex_w))),
                  "rasterize_bwd: v_textures required with textures (or the record buffers, for textures whose "
                  "two rows fit in LDS)");
    const int64_t E = gaussian_ids ? nnz : C * N;  // per-entry inputs
    TGS_CHECK_ARG(!gaussian_ids || (nnz >= 0 && nnz < (int64_t(1) << 31)), "rasterize_bwd: bad nnz");
    TGS_CHECK_ARG((workspace || workspace_bytes(dtype, 1, E) == 0) && workspace_size >= workspace_bytes(dtype, 1, E),
                  "rasterize_bwd: workspace of %zu bytes required", workspace_bytes(dtype, 1, E));
    rc = check_zero_rows(zero_rows_dst, zero_rows_mask, zero_row_bytes);
    if (rc) return rc;
    rc = check_pair_samples("rasterize_bwd", pair_samples, live_entries, dtype, tex_mode, K);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
#define TGS_BWD(S)                                                                                              \
    {                                                                                                           \
        RasterArgs<S> a;                                                                                        \
        fill_common<S>(a, C, N, means2d, ray_transforms, colors, opacities, normals, textures, tex_h, tex_w,     \
                       texture_range_x, texture_range_y, backgrounds, masks, width, height, tile_size,          \
                       tile_width, tile_height, tile_offsets, flatten_ids, n_isects, g_weight, workspace,        \
                       zero_rows_dst, zero_rows_mask, zero_row_bytes, zero_rows_select, gaussian_ids, nnz);     \
        a.order_ws = order_region(workspace, workspace_size, dtype, E, C, tile_width, tile_height);             \
        set_mip<S>(a, tex_mode, tex_h, tex_w);                                                                 \
        a.r_c = (const S *)render_colors; a.r_a = (const S *)render_alphas; a.r_last = last_ids;                \
        a.r_med = median_ids; a.v_c = (const S *)v_render_colors; a.v_a = (const S *)v_render_alphas;           \
        a.r_accd = (const S *)render_accum_depth;                                                               \
        a.r_T = (const S *)render_trans;                                                                        \
        a.live = (uint64_t *)live_entries;                                                                      \
        a.samples = (S *)pair_samples;                                                                          \
        const bool recs = textures && tex_records && tex_counts && tex_rec_incl && tex_rec_cursor &&           \
                          gather_fits<S>(tex_h, tex_w) &&                                                        \
                          (!zero_rows_dst || ((uintptr_t)zero_rows_mask % 4) == 0);                              \
        if (recs) {                                                                                             \
            a.rec_cnt = tex_counts; a.rec_incl = tex_rec_incl; a.rec_cur = tex_rec_cursor;                      \
            a.recs = tex_records; a.rec_rows = rec_rows;                                                        \
        } else if (textures && tex_records && zero_rows_dst && !rec_rows) {                                     \
            /* records requested but the row does not fit the gather's LDS: the rows the */                   \
            /* forward left for the gather are zero-filled here and added to directly    */                   \
            rc = tgs_zero_fill_rows(zero_rows_dst, zero_row_bytes, N, zero_rows_mask, 1, st);                   \
            if (rc) return rc;                                                                                  \
        }                                                                                                       \
        a.v_n = (const S *)v_render_normals; a.v_d = (const S *)v_render_distort;                               \
        a.v_m = (const S *)v_render_median; a.uv_mode = uv_grad_mode;                                           \
        a.g_xy_abs = (S *)v_means2d_abs; a.g_xy = (S *)v_means2d; a.g_rt = (S *)v_ray_transforms;              \
        a.g_col = (S *)v_colors; a.g_op = (S *)v_opacities; a.g_tex = (S *)v_textures; a.g_nrm = (S *)v_normals; \
        const int64_t n = E;                                                                                    \
        const unsigned nb = (unsigned)ceil_div(n, 256);                                                         \
        if (n > 0) zero_kernel<S><<<(unsigned)ceil_div(n * kAbc, 256), 256, 0, st>>>(n * kAbc, a.g_abc);        \
        TGS_CHECK_LAUNCH("rasterize_bwd zero");                                                                 \
        if (n_isects > 0) {                                                                                     \
            rc = dispatch<S, true>(a, K, tex_mode, st);                                                         \
        } else if (zero_rows_dst) {                                                                             \
            rc = tgs_zero_fill_rows(zero_rows_dst, zero_row_bytes, N, zero_rows_mask, zero_rows_select & 0xff, st);\
        }                                                                                                       \
        if (rc) return rc;                                                                                      \
        if (recs && !records_only) {                                                                            \
            /* the rows with records (given all rows: every row with mask 1) are written in full */           \
            rc = rec_rows ? launch_gather<S>(N, tex_h, tex_w, a, nullptr, 1, st)                                \
                          : launch_gather<S>(N, tex_h, tex_w, a, zero_rows_dst ? zero_rows_mask : nullptr,      \
                                             zero_rows_dst ? 1 : 0, st);                                          \
            if (rc) return rc;                                                                                  \
        }                                                                                                       \
        if (n > 0)                                                                                              \
            rt_grad_kernel<S><<<nb, 256, 0, st>>>(n, (const S *)ray_transforms, (const S *)means2d, a.g_abc,    \
                                                  (S *)v_ray_transforms, (S *)v_densify);                       \
        TGS_CHECK_LAUNCH("rasterize_bwd ray-transform gradient");                                               \
        return TGS_OK;                                                                                          \
    }
    if (dtype == TGS_F32) TGS_BWD(float)
    if (dtype == TGS_F64) TGS_BWD(double)
#undef TGS_BWD
    set_error("rasterize_bwd: unsupported dtype %d", dtype);
    return TGS_ERR_UNSUPPORTED;
}
