/*
 * tgs_hip.h - C ABI of libtgs_hip.so, the MI355X (gfx950) kernels of the
 * textured 2D-Gaussian-splatting render path.
 *
 * Conventions
 *  - Every pointer is a device pointer unless its name ends in _host.
 *  - Sizes are int64_t; floating-point scalars are passed as double and
 *    converted to the working precision selected by `dtype`.
 *  - `dtype` selects the arithmetic of the floating-point tensors:
 *    TGS_F32 (float) or TGS_F64 (double).  Integer tensors are fixed-width.
 *  - The library never allocates or frees device memory.  Outputs and
 *    scratch ("temp") buffers are owned by the caller; outputs documented as
 *    "accumulated" must be zero-filled by the caller before the call.
 *  - All work is enqueued on `stream` (a hipStream_t; NULL = legacy default
 *    stream).  No entry point synchronises the device.
 *  - Return value: TGS_OK (0) or a nonzero error code; tgs_last_error()
 *    then returns a message (thread-local, valid until the next call).
 *
 * Each entry point names the reference operator it replaces: the pybind op
 * in textured_gaussians/cuda/csrc/ext.cpp and its declaration in
 * textured_gaussians/cuda/csrc/bindings.h.
 */
#ifndef TGS_HIP_H
#define TGS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *tgs_stream_t;

enum { TGS_F32 = 0, TGS_F64 = 1 };
/* OR-ed into the dtype of the rasterizers: textures stored as bfloat16 RGBA (float32
 * compute only; the texture gradient stays float32). */
enum { TGS_TEX_BF16 = 0x100 };
/* OR-ed into the dtype of the rasterizers: the `mipmapped2` filter.  textures is then
 * the flat pyramid [N, (4^(L+1) - 1) / 3, 4] of tgs_generate_mipmap_fwd over
 * tex_h = tex_w = 2^L (L <= 12) textures; the backward takes uv_grad_mode
 * TGS_UV_NONE (the reference's rule) and v_textures of the pyramid's shape. */
enum { TGS_TEX_MIP2 = 0x200 };
/* OR-ed into the dtype of tgs_rasterize_fwd (no reference counterpart): visit every
 * list entry.  By default each wavefront first tests 64 entries at a time against its
 * pixels with a conservative bound (outside the texture range and op * vis below
 * 1/255 everywhere) and walks only the candidates; the outputs are the same. */
enum { TGS_FWD_NO_CULL = 0x400 };
/* OR-ed into the dtype of tgs_rasterize_bwd with the texture-gradient records (no
 * reference counterpart): only the rows whose tex_counts entry is nonzero take records
 * (and the gather writes exactly those rows in full); the other rows take float atomics
 * into v_textures, which must be zero there.  tex_counts may then hold zeros for rows
 * that do have pairs: the caller picks the rows (e.g. those with many pairs). */
enum { TGS_BWD_RECORD_ROWS = 0x800 };
enum { TGS_OK = 0, TGS_ERR_ARG = 1, TGS_ERR_LAUNCH = 2, TGS_ERR_UNSUPPORTED = 3 };

/* Gradient rule for the bilinear texture lookup w.r.t. the splat coordinate.
 * TGS_UV_NONE   - no d(texture sample)/ds term: the reference "bilinear"
 *                 backward (rasterize_to_pixels_textured_gaussians_2dgs_bwd.cu).
 * TGS_UV_DOUBLE - du/ds = (res-1)/range: the reference "bilinear_bwd2"
 *                 backward (rasterize_to_pixels_textured_gaussians_2dgs_bwd2.cu:611-676).
 * TGS_UV_EXACT  - du/ds = (res-1)/(2 range): the true derivative.          */
enum { TGS_UV_NONE = 0, TGS_UV_EXACT = 1, TGS_UV_DOUBLE = 2 };

/* Projection VJP rule.
 * TGS_VJP_REFERENCE      - fully_fused_projection_bwd_2dgs (utils.cuh:659-724)
 * TGS_VJP_REFERENCE_NORM - fully_fused_projection_bwd2_2dgs (utils.cuh:726-791)
 * TGS_VJP_EXACT          - the true derivative of the forward.             */
enum { TGS_VJP_REFERENCE = 0, TGS_VJP_REFERENCE_NORM = 1, TGS_VJP_EXACT = 2 };

const char *tgs_last_error(void);
int tgs_abi_version(void);

/* Zero-fills nbytes bytes at dst (16-byte aligned, nbytes a multiple of 4) in
 * one launch of non-temporal stores.  No reference counterpart: the reference
 * allocates the texture gradient with torch::zeros inside its backward
 * (rasterize_to_pixels_textured_gaussians_2dgs_bwd.cu).  The autograd layer
 * normally fuses the fill into the rasterizer kernels (zero_rows_* below);
 * this dense fill is its alternative (same result, one launch).           */
int tgs_zero_fill(void *dst, size_t nbytes, tgs_stream_t stream);

/* Row-selective variant for the texture gradient [N, row_bytes]: mask[N] (uint8)
 * is built by tgs_mark_rows (1 for every Gaussian id % N that occurs in
 * flatten_ids, i.e. rows the backward can touch, else 0); tgs_zero_fill_rows
 * zeroes the rows whose mask byte equals `select` (0 or 1).  Splitting the fill
 * lets the rows the backward never touches be zeroed while it runs.  No
 * reference counterpart (see tgs_zero_fill).                              */
int tgs_mark_rows(int64_t n_isects, const int32_t *flatten_ids, int64_t N, uint8_t *mask,
                  tgs_stream_t stream);
int tgs_zero_fill_rows(void *dst, size_t row_bytes, int64_t N, const uint8_t *mask, int select,
                       tgs_stream_t stream);

/* ---------------------------------------------------------------- projection
 * Replaces `fully_fused_projection_fwd_2dgs` (ext.cpp:53-55, bindings.h:282-293).
 * means[N,3] quats[N,4] scales[N,3] viewmats[C,4,4] Ks[C,3,3] ->
 * radii int32[C,N], means2d[C,N,2], depths[C,N], ray_transforms[C,N,3,3],
 * normals[C,N,3].  Culled entries get radii 0 and zeros elsewhere.        */
int tgs_projection_2dgs_fwd(int dtype, int64_t C, int64_t N, const void *means,
                            const void *quats, const void *scales, const void *viewmats,
                            const void *Ks, int32_t width, int32_t height,
                            double near_plane, double far_plane, double radius_clip,
                            int32_t *radii, void *means2d, void *depths,
                            void *ray_transforms, void *normals, tgs_stream_t stream);

/* Replaces `fully_fused_projection_bwd_2dgs` and `fully_fused_projection_bwd2_2dgs`
 * (ext.cpp:56-61, bindings.h:295-330).  v_means[N,3], v_quats[N,4],
 * v_scales[N,3] are accumulated.  v_viewmats[C,4,4] (may be NULL) is
 * accumulated in TGS_VJP_EXACT mode and left untouched otherwise (the
 * reference never writes it).                                              */
int tgs_projection_2dgs_bwd(int dtype, int64_t C, int64_t N, const void *means,
                            const void *quats, const void *scales, const void *viewmats,
                            const void *Ks, const int32_t *radii,
                            const void *ray_transforms, const void *v_means2d,
                            const void *v_depths, const void *v_normals,
                            const void *v_ray_transforms, int vjp_mode, void *v_means,
                            void *v_quats, void *v_scales, void *v_viewmats,
                            tgs_stream_t stream);

/* Packed layout (replaces `fully_fused_projection_packed_fwd_2dgs` /
 * `..._packed_bwd_2dgs` / `..._packed_bwd2_2dgs`, ext.cpp:63-71, bindings.h:430-486): only the visible
 * (camera, Gaussian) entries, in (camera, Gaussian) order.
 * tgs_projection_2dgs_packed_count: flags int32[C*N] (1 = visible) and their
 * inclusive prefix sum int64[C*N] (temp: tgs_isect_count_temp_bytes(C*N)); the
 * caller reads nnz = inclusive[C*N-1].
 * tgs_projection_2dgs_packed_fill: indptr int64[C+1] (entries of camera c are
 * [indptr[c], indptr[c+1])), camera_ids, gaussian_ids int64[nnz], radii
 * int32[nnz], means2d[nnz,2], depths[nnz], ray_transforms[nnz,3,3],
 * normals[nnz,3]; the values are those of the grid layout's visible entries.
 * tgs_projection_2dgs_packed_bwd: the backward over the nnz entries; sparse
 * = 0 accumulates v_means[N,3], v_quats[N,4], v_scales[N,3] (caller
 * zero-fills), sparse = 1 writes them per entry ([nnz,...], the values of the
 * reference's sparse gradients).  v_viewmats as in tgs_projection_2dgs_bwd. */
int tgs_projection_2dgs_packed_count(int dtype, int64_t C, int64_t N, const void *means,
                                     const void *quats, const void *scales, const void *viewmats,
                                     const void *Ks, int32_t width, int32_t height, double near_plane,
                                     double far_plane, double radius_clip, int32_t *flags,
                                     int64_t *inclusive, void *temp, size_t temp_bytes,
                                     tgs_stream_t stream);
int tgs_projection_2dgs_packed_fill(int dtype, int64_t C, int64_t N, const void *means,
                                    const void *quats, const void *scales, const void *viewmats,
                                    const void *Ks, int32_t width, int32_t height, double near_plane,
                                    double far_plane, double radius_clip, const int64_t *inclusive,
                                    int64_t *indptr, int64_t *camera_ids, int64_t *gaussian_ids,
                                    int32_t *radii, void *means2d, void *depths, void *ray_transforms,
                                    void *normals, tgs_stream_t stream);
int tgs_projection_2dgs_packed_bwd(int dtype, int64_t C, int64_t N, int64_t nnz,
                                   const int64_t *camera_ids, const int64_t *gaussian_ids, int sparse,
                                   const void *means, const void *quats, const void *scales,
                                   const void *viewmats, const void *Ks, const int32_t *radii,
                                   const void *ray_transforms, const void *v_means2d,
                                   const void *v_depths, const void *v_normals,
                                   const void *v_ray_transforms, int vjp_mode, void *v_means,
                                   void *v_quats, void *v_scales, void *v_viewmats, tgs_stream_t stream);

/* ------------------------------------------------------- tile intersection
 * Replaces `isect_tiles` (ext.cpp:32, bindings.h:110-121) in three phases:
 * count -> (caller reads cum_tiles_per_gauss[C*N-1] = n_isects) -> fill -> sort.
 * means2d f32[C,N,2], radii i32[C,N], depths f32[C,N] (unpacked layout).    */
size_t tgs_isect_count_temp_bytes(int64_t n_elems);
int tgs_isect_tiles_count(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                          int32_t tile_size, int32_t tile_width, int32_t tile_height,
                          int32_t *tiles_per_gauss, int64_t *cum_tiles_per_gauss,
                          void *temp, size_t temp_bytes, tgs_stream_t stream);
/* keys: cam << (32+tile_bits) | tile << 32 | int32 bits of depth (sign-extended) */
int tgs_isect_tiles_fill(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                         const float *depths, const int64_t *cum_tiles_per_gauss,
                         int32_t tile_size, int32_t tile_width, int32_t tile_height,
                         int64_t *isect_ids, int32_t *flatten_ids, tgs_stream_t stream);
/* Packed entries (the reference's isect_tiles with packed=True, [IS]:16-105):
 * entry i of camera camera_ids[i]; tiles_per_gauss / cum from
 * tgs_isect_tiles_count(1, nnz, ...); values are entry indices. */
int tgs_isect_tiles_fill_packed(int64_t C, int64_t nnz, const int64_t *camera_ids, const float *means2d,
                                const int32_t *radii, const float *depths,
                                const int64_t *cum_tiles_per_gauss, int32_t tile_size, int32_t tile_width,
                                int32_t tile_height, int64_t *isect_ids, int32_t *flatten_ids,
                                tgs_stream_t stream);
/* Stable LSD radix sort of (key, value) pairs on key bits [0, end_bit)
 * (CUB DeviceRadixSort::SortPairs with DoubleBuffer, isect_tiles.cu:252-287).
 * The result is in (keys, values) if *selector_host == 0 after the call,
 * else in (keys_alt, values_alt).  The selector is known on the host.     */
size_t tgs_sort_pairs_temp_bytes(int64_t n);
int tgs_sort_pairs(int64_t n, int32_t end_bit, int64_t *keys, int32_t *values,
                   int64_t *keys_alt, int32_t *values_alt, void *temp, size_t temp_bytes,
                   int32_t *selector_host, tgs_stream_t stream);
/* Replaces `isect_offset_encode` (ext.cpp:33, bindings.h:123-127).
 * offsets i32[C, tile_height, tile_width]; zero-filled when n_isects == 0. */
int tgs_isect_offset_encode(int64_t n_isects, const int64_t *isect_ids, int64_t C,
                            int32_t tile_width, int32_t tile_height, int32_t *offsets,
                            tgs_stream_t stream);

/* ------------------------------------------------------ spherical harmonics
 * Replaces `compute_sh_fwd` / `compute_sh_bwd` (ext.cpp:11-12, bindings.h:194-206).
 * n = number of directions; coeffs[n,K,3]; masks uint8[n] or NULL.
 * Masked-out colours are written as 0.  v_coeffs[n,K,3] is written (not
 * accumulated; zero for masked rows and bands above `degree`); v_dirs (may
 * be NULL) is accumulated.                                                 */
int tgs_sh_fwd(int dtype, int64_t n, int32_t K, int32_t degree, const void *dirs,
               const void *coeffs, const uint8_t *masks, void *colors, tgs_stream_t stream);
int tgs_sh_bwd(int dtype, int64_t n, int32_t K, int32_t degree, const void *dirs,
               const void *coeffs, const uint8_t *masks, const void *v_colors,
               void *v_coeffs, void *v_dirs, tgs_stream_t stream);

/* Several cameras over one coefficient set (no reference counterpart: the
 * reference expands coeffs to [C, N, K, 3] and calls compute_sh_fwd/bwd on
 * C*N rows, rendering.py:1573-1583).  dirs [C,n,3], masks uint8[C,n] or
 * NULL, colors / v_colors [C,n,3]; coeffs and v_coeffs [n,K,3]; v_coeffs is
 * written with the sum over the cameras; v_dirs [C,n,3] (may be NULL) is
 * written in full (zero where masked).  float32 with K*3 <= 48 only
 * (TGS_ERR_UNSUPPORTED otherwise).                                          */
int tgs_sh_fwd_cams(int dtype, int64_t C, int64_t n, int32_t K, int32_t degree, const void *dirs,
                    const void *coeffs, const uint8_t *masks, void *colors, tgs_stream_t stream);
int tgs_sh_bwd_cams(int dtype, int64_t C, int64_t n, int32_t K, int32_t degree, const void *dirs,
                    const void *coeffs, const uint8_t *masks, const void *v_colors, void *v_coeffs,
                    void *v_dirs, tgs_stream_t stream);

/* --------------------------------------------------------------- rasterizer
 * Replaces `rasterize_to_pixels_fwd_textured_gaussians` (ext.cpp:86-88,
 * bindings.h:509-538) when textures != NULL, and `rasterize_to_pixels_fwd_2dgs`
 * (ext.cpp:76-77) when textures == NULL.  K = colour channels, 1..4.
 * means2d[C,N,2] ray_transforms[C,N,3,3] colors[C,N,K] opacities[C,N]
 * normals[C,N,3] textures[N,tex_h,tex_w,4] backgrounds[C,K] (may be NULL)
 * masks uint8[C,tile_h,tile_w] (may be NULL).
 * Packed layout (gaussian_ids != NULL; the reference's `packed=True`): the
 * per-Gaussian inputs are nnz entries ([nnz,2], [nnz,3,3], [nnz,K], [nnz],
 * [nnz,3]) from tgs_projection_2dgs_packed_fill, flatten_ids index them, and
 * entry i samples texture row gaussian_ids[i] (the reference indexes the
 * textures with the entry index there, Q3).  N is then the number of
 * Gaussians (texture rows, contribution counters), and the workspace is
 * tgs_rasterize_workspace_bytes(dtype, 1, nnz) (0 bytes, may be NULL, when
 * nnz = 0; gaussian_ids must still be non-NULL then, it is not read).
 * Outputs (written): render_colors[C,H,W,K] render_alphas[C,H,W]
 * render_normals[C,H,W,3] render_distort[C,H,W] render_median[C,H,W]
 * last_ids i32[C,H,W] median_ids i32[C,H,W].
 * Outputs (accumulated): contrib_sum[N], contrib_count[N].
 * render_accum_depth[C,H,W] (may be NULL; no reference counterpart): the
 * forward's distortion accumulator sum_i w_i colors[i, K-1], for the exact
 * distortion gradient.
 * render_trans[C,H,W] (may be NULL; no reference counterpart): the final
 * transmittance T of each pixel.  The reference's backward starts from
 * 1 - render_alphas ([TB]:456-457), which in float32 keeps T only to an
 * absolute 6e-8 (6e-4 relative at the 1e-4 stop); given to
 * tgs_rasterize_bwd, the backward starts from T itself.
 * live_entries (may be NULL; no reference counterpart): uint64[waves][n_isects]
 * with waves = ceil(tile_size^2 / 64), the wavefronts of a tile; the forward
 * sets [w][i] to the mask of the lanes of wave w of entry i's tile that
 * composited entry i (bit l: pixel 64 w + l of the tile; 0 where none did), for
 * every i up to the last entry wave w reached.  Given to tgs_rasterize_bwd, the
 * backward visits only entries with a nonzero mask (the others contribute
 * nothing to any gradient).
 * pair_samples (may be NULL; no reference counterpart; needs live_entries,
 * float32 compute, float32 or bfloat16 textures without mip2, K <= 3, 16-byte
 * alignment): float32[waves][n_isects][64][4].  For each composited pair inside
 * the texture range the forward stores its bilinear sample (K colour channels,
 * zeros up to 3, then alpha) at [w][i][r], r = the pair's rank among the lanes
 * of mask [w][i] that are inside the range.  Given to tgs_rasterize_bwd with the
 * same live_entries (uv_grad_mode TGS_UV_NONE), the backward reads the samples
 * instead of the texels and takes the masks' keep decisions.
 * tex_counts int32[N] (may be NULL; no reference counterpart; accumulated,
 * caller zero-fills): per texture row, the number of (pixel, entry) pairs
 * whose backward adds texture gradient (composited and inside the texture
 * range).  Given to tgs_rasterize_bwd, it sizes the texture-gradient records.
 * workspace: device scratch of at least tgs_rasterize_workspace_bytes(dtype,
 * C, N) bytes (no reference counterpart: the reference reads Gaussian
 * attributes in place); it holds packed per-Gaussian records for the call.
 * When it holds tgs_rasterize_tile_order_bytes(C, tile_width, tile_height)
 * more bytes, the call also writes there the order in which it hands the tiles
 * to workgroups (longest list first, which balances the XCDs and shortens the
 * tail); without that room the tiles go in runs of 16 that take turns over the
 * XCDs.  The outputs are the same either way.
 * Calls that share one workspace must be ordered on one stream.
 * zero_rows_dst (may be NULL; no reference counterpart): a buffer of N rows of
 * zero_row_bytes bytes (16-byte aligned, a multiple of 16) whose rows with
 * zero_rows_mask[n] == 1 this call zero-fills along the way, using memory
 * bandwidth the rasterizer leaves idle.  zero_rows_select picks the rows: its
 * low byte is the mask value to fill (0 or 1), bits 8-15 a set of residues of
 * row mod 8 (0 = every row), bits 16-23 the fill stores (1 KB each) per visited
 * list entry and wavefront (0 = 1 in the forward, 2 in the backward).  With the mask of tgs_mark_rows this prepares the
 * texture gradient: the forward fills the rows with mask 1 (those the backward
 * accumulates into) and the backward those with mask 0, or, with the records
 * of tgs_rasterize_bwd, the rows with mask 0 are split between forward and
 * backward (disjoint residues) and the gather writes the rows with mask 1.  */
size_t tgs_rasterize_workspace_bytes(int dtype, int64_t C, int64_t N);
/* Bytes of the optional tile-order region after the rasterizer workspace (no
 * reference counterpart): ceil(C * tile_width * tile_height * 4 / 256) * 256, or 0
 * when the tiles exceed 262144 (no order is sorted then). */
size_t tgs_rasterize_tile_order_bytes(int64_t C, int32_t tile_width, int32_t tile_height);
/* Measurement hook (no reference counterpart): while set, tgs_rasterize_fwd
 * (which = 0) or tgs_rasterize_bwd (which = 1) records `begin` and `end`
 * (hipEvent_t) on its stream right before and right after the rasterizer
 * kernel, so that the pair times that kernel alone; which = 2: around the
 * texture-gradient gather that tgs_rasterize_bwd runs after its kernel when it
 * writes records (row list and gather kernels).  NULL, NULL clears it.  One
 * setting per process, not per thread. */
int tgs_set_raster_events(int which, void *begin, void *end);
/* Measurement hook (no reference counterpart): while set, tgs_rasterize_fwd with
 * float textures (no bf16, no mip2) stores 1 into marks[row * tex_h * tex_w + texel]
 * for every texel it reads (uint8[n_texels], n_texels >= N * tex_h * tex_w; caller
 * zero-fills); the number of nonzero bytes is the count of distinct texels read.
 * NULL, 0 clears it.  One setting per process. */
int tgs_set_texel_marks(uint8_t *marks, int64_t n_texels);
/* Measurement hook (no reference counterpart): LDS atomic-add throughput of the
 * current GPU, in lanes per clock per CU, for n (1..4) access kinds over random
 * addresses of a 40 KB LDS buffer: [0] float32 atomic add, [1] float64 atomic
 * add, [2] uint32 atomic add, [3] float32 read, add and write (plain accesses).
 * The texture-gradient gather sums in float64 because of [0] against [1] on
 * gfx950.  Synchronizes the stream; allocates a few KB of device memory. */
int tgs_lds_atomic_rates(double *lanes_per_clock_per_cu, int32_t n, tgs_stream_t stream);
/* Bytes of one texture-gradient record (see tgs_rasterize_bwd). */
size_t tgs_tex_record_bytes(int dtype);
/* 1 if tgs_rasterize_bwd sums texture-gradient records of this compute dtype and
 * texture size with the streaming row gather (float32 compute, rows whose float64
 * sums fit twice in one CU's LDS: up to about 50x50 texels), 0 if the records take
 * the slower fallback gather.  No reference counterpart: callers choose between
 * records and float atomics with it. */
int tgs_texture_gather_fast(int dtype, int64_t tex_h, int64_t tex_w);
/* inclusive[i] = sum_{j <= i} counts[j] (int64); temp of tgs_isect_count_temp_bytes(n)
 * bytes.  No reference counterpart (sizes the texture-gradient records). */
int tgs_prefix_sum_i32(int64_t n, const int32_t *counts, int64_t *inclusive, void *temp,
                       size_t temp_bytes, tgs_stream_t stream);

int tgs_rasterize_fwd(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                      const void *ray_transforms, const void *colors, const void *opacities,
                      const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                      double texture_range_x, double texture_range_y,
                      const void *backgrounds, const uint8_t *masks,
                      const int64_t *gaussian_ids, int64_t nnz, int32_t width,
                      int32_t height, int32_t tile_size, int32_t tile_width,
                      int32_t tile_height, const int32_t *tile_offsets,
                      const int32_t *flatten_ids, int64_t n_isects,
                      double gs_contrib_threshold, double g_weight, void *render_colors,
                      void *render_alphas, void *render_normals, void *render_distort,
                      void *render_median, int32_t *last_ids, int32_t *median_ids,
                      void *contrib_sum, void *contrib_count, void *render_accum_depth,
                      void *render_trans, uint64_t *live_entries, void *pair_samples,
                      int32_t *tex_counts, void *zero_rows_dst, const uint8_t *zero_rows_mask,
                      size_t zero_row_bytes, int zero_rows_select, void *workspace, size_t workspace_size,
                      tgs_stream_t stream);

/* Allocation-free variant of tgs_rasterize_fwd (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and outputs as tgs_rasterize_fwd minus the workspace, which is a
 * static device array of TGS_RASTERIZE_NOALLOC_SCRATCH_BYTES bytes inside the
 * library: one copy per GPU, held from the first call on that GPU until the
 * process ends.  Calls needing more than that
 * (tgs_rasterize_workspace_bytes(dtype, 1, entries) > the limit; entries = C*N,
 * or nnz when packed: about 578k float32 or 289k float64 entries) fail with
 * TGS_ERR_ARG.  Nothing in the scratch needs to start at zero.  Every call on
 * one GPU uses that GPU's one array: calls that overlap on one GPU (other
 * streams, other threads) race; order them on one stream.  This entry point
 * calls no hipMalloc/hipFree; its first call on a GPU looks the array up with
 * hipGetSymbolAddress. */
#define TGS_RASTERIZE_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tgs_rasterize_fwd_noalloc(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                              const void *ray_transforms, const void *colors, const void *opacities,
                              const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                              double texture_range_x, double texture_range_y,
                              const void *backgrounds, const uint8_t *masks,
                              const int64_t *gaussian_ids, int64_t nnz, int32_t width,
                              int32_t height, int32_t tile_size, int32_t tile_width,
                              int32_t tile_height, const int32_t *tile_offsets,
                              const int32_t *flatten_ids, int64_t n_isects,
                              double gs_contrib_threshold, double g_weight, void *render_colors,
                              void *render_alphas, void *render_normals, void *render_distort,
                              void *render_median, int32_t *last_ids, int32_t *median_ids,
                              void *contrib_sum, void *contrib_count, void *render_accum_depth,
                              void *render_trans, uint64_t *live_entries, void *pair_samples,
                              int32_t *tex_counts,
                              void *zero_rows_dst, const uint8_t *zero_rows_mask, size_t zero_row_bytes,
                              int zero_rows_select, tgs_stream_t stream);

/* Replaces `rasterize_to_pixels_bwd_textured_gaussians` /
 * `rasterize_to_pixels_bwd2_textured_gaussians` (ext.cpp:89-94,
 * bindings.h:540-629; uv_grad_mode TGS_UV_NONE / TGS_UV_DOUBLE) and
 * `rasterize_to_pixels_bwd_2dgs` (ext.cpp:78-80) when textures == NULL.
 * v_render_distort may be NULL (no distortion term).  render_accum_depth
 * NULL selects the reference's distortion rule, which seeds the depth
 * accumulator with render_colors[..., K-1] (that includes the texture colour
 * and T*background of the channel); non-NULL gives the true gradient.
 * live_entries: NULL (visit every entry) or the array the forward of the same
 * inputs wrote; pair_samples: NULL or the forward's, with its live_entries.  render_trans: NULL (start from 1 - render_alphas, as the
 * reference) or the forward's final transmittance.
 * Texture-gradient records (all NULL, or all given; no reference counterpart):
 * tex_counts from the forward of the same inputs, tex_rec_incl int64[N] its
 * inclusive prefix sum (tgs_prefix_sum_i32), tex_rec_cursor int32[N] zeroed,
 * tex_records of max(1, tex_rec_incl[N-1]) * tgs_tex_record_bytes(dtype) bytes.  The
 * backward then writes one record per pair instead of adding to v_textures with
 * float atomics, and a gather kernel sums each texture row's records in LDS (in
 * float64 where tgs_texture_gather_fast says so) and writes the row once: rows with zero_rows_mask == 1 are written in full (so they need no
 * zero-fill beforehand; the rows with mask 0 that zero_rows_select names are
 * zero-filled by this call), or, with zero_rows_dst NULL, the sums are added
 * to v_textures.  Rows larger than 64 KB fall back to direct accumulation.  With v_textures NULL (and
 * zero_rows_dst NULL) only the records are written, for tgs_texture_adam_rows.
 * dtype | TGS_TEX_BF16: textures are bfloat16 (float32 compute); the gradient
 * outputs stay float32.
 * Accumulated outputs (caller zero-fills): v_means2d_abs[C,N,2] (may be
 * NULL), v_means2d[C,N,2], v_ray_transforms[C,N,3,3], v_colors[C,N,K],
 * v_opacities[C,N], v_textures[N,tex_h,tex_w,4] (NULL if textures NULL),
 * v_normals[C,N,3].  v_densify[C,N,2] (may be NULL) is written after the
 * accumulation: v_ray_transforms[.,0,2] * T[2,2], v_ray_transforms[.,1,2] * T[2,2]
 * (the reference reads it while other waves still accumulate; this is the
 * settled value).  workspace: as for tgs_rasterize_fwd.                  */
int tgs_rasterize_bwd(int dtype, int64_t C, int64_t N, int32_t K, const void *means2d,
                      const void *ray_transforms, const void *colors, const void *opacities,
                      const void *normals, const void *textures, int64_t tex_h, int64_t tex_w,
                      double texture_range_x, double texture_range_y,
                      const void *backgrounds, const uint8_t *masks,
                      const int64_t *gaussian_ids, int64_t nnz, int32_t width,
                      int32_t height, int32_t tile_size, int32_t tile_width,
                      int32_t tile_height, const int32_t *tile_offsets,
                      const int32_t *flatten_ids, int64_t n_isects, double g_weight,
                      const void *render_colors, const void *render_alphas,
                      const int32_t *last_ids, const int32_t *median_ids,
                      const void *render_accum_depth, const void *render_trans,
                      const uint64_t *live_entries, const void *pair_samples, const int32_t *tex_counts,
                      const int64_t *tex_rec_incl,
                      int32_t *tex_rec_cursor, void *tex_records, const void *v_render_colors,
                      const void *v_render_alphas,
                      const void *v_render_normals, const void *v_render_distort,
                      const void *v_render_median, int uv_grad_mode, void *v_means2d_abs,
                      void *v_means2d, void *v_ray_transforms, void *v_colors,
                      void *v_opacities, void *v_textures, void *v_normals, void *v_densify,
                      void *zero_rows_dst, const uint8_t *zero_rows_mask, size_t zero_row_bytes,
                      int zero_rows_select, void *workspace, size_t workspace_size,
                      tgs_stream_t stream);

/* Mipmap pyramid for the `mipmapped2` filter (generate_mipmap_fwd.cu,
 * generate_mipmap_bwd.cu; ext.cpp:188-193): textures [N, R, R, C] (R = 2^L,
 * L <= 15, C <= 5) -> mip_textures [N, tgs_mipmap_texels(L), C], levels coarsest
 * first (level l at texel (4^(L-l) - 1) / 3), each the 2x2 box average of the
 * finer one.  The backward is its adjoint: v_mip_textures -> v_textures
 * [N, R, R, C] (written, not accumulated). */
int64_t tgs_mipmap_texels(int32_t log_texture_res);
int tgs_generate_mipmap_fwd(int dtype, int64_t N, int32_t C, int32_t log_texture_res,
                            const void *textures, void *mip_textures, tgs_stream_t stream);
int tgs_generate_mipmap_bwd(int dtype, int64_t N, int32_t C, int32_t log_texture_res,
                            const void *v_mip_textures, void *v_textures, tgs_stream_t stream);

/* SSIM of the trainer's loss (no C++ counterpart in the reference: it calls
 * torchmetrics' StructuralSimilarityIndexMeasure(data_range=1.0),
 * examples/simple_trainer_textured_gaussians.py:430, 1426-1429): 11-tap Gaussian
 * window (sigma 1.5), c1 = 1e-4, c2 = 9e-4, mean over the windows inside the image.
 * x, y: [B, H, W, C] channels-last (C <= 4, H, W >= 11, B * H <= 65535).
 * tgs_ssim_fwd writes tgs_ssim_partials(...) per-block partial sums of the SSIM map
 * (their sum / (B C (H-10) (W-10)) is the SSIM) and keeps in `work`
 * (tgs_ssim_work_bytes) what tgs_ssim_bwd needs: dx = scale * dSSIM_sum/dx, the
 * gradient of the sum of the map with respect to x (y held constant).  The two
 * calls on one work buffer must be ordered; the backward overwrites the window sums. */
size_t tgs_ssim_work_bytes(int dtype, int32_t B, int32_t H, int32_t W, int32_t C);
int64_t tgs_ssim_partials(int32_t B, int32_t H, int32_t W, int32_t C);
int tgs_ssim_fwd(int dtype, int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y,
                 void *work, void *partials, tgs_stream_t stream);
int tgs_ssim_bwd(int dtype, int32_t B, int32_t H, int32_t W, int32_t C, const void *x, const void *y,
                 void *work, double scale, void *dx, tgs_stream_t stream);

/* Selective Adam step of bfloat16 textures from texture-gradient records (no
 * reference counterpart; the arithmetic is the reference's `selective_adam_update`,
 * adam.cu:16-47: no bias correction, rows with visibility[n] == 0 untouched).
 * It replaces building a dense float32 texture gradient and calling
 * selective_adam_update on it, which at 2M x 64x64 textures does not fit in HBM.
 * The records come from tgs_rasterize_bwd with v_textures NULL (tex_counts,
 * tex_rec_incl, tex_rec_cursor, tex_records as there, float32).  State, each
 * [N, tex_h, tex_w, 4] 16-bit, 16-byte aligned:
 *   tex_hi  bfloat16 render copy = the float32 master rounded to nearest (ties away)
 *   tex_lo  int16, master bits - (hi << 16): hi and lo hold the master exactly
 *   exp_avg, exp_avg_sq  bfloat16 moments, stochastically rounded (hash of the
 *           element index and seed; pass a new seed each step).
 * visibility uint8[N], 4-byte aligned.  Needs tex_h * tex_w even and two float
 * rows in LDS (tex_h * tex_w <= 5120).                                    */
int tgs_texture_adam_rows(int64_t N, int32_t tex_h, int32_t tex_w, const int32_t *tex_counts,
                          const int64_t *tex_rec_incl, const int32_t *tex_rec_cursor,
                          const void *tex_records, const uint8_t *visibility, uint16_t *tex_hi,
                          uint16_t *tex_lo, uint16_t *exp_avg, uint16_t *exp_avg_sq, double lr,
                          double b1, double b2, double eps, uint32_t seed, tgs_stream_t stream);

/* ------------------------------------------------ neural-texture decoder
 * No reference counterpart (its closest relative is the coordinate-MLP
 * implicit texture, examples/texture_models/mlp.py:6-18).  The decoder maps a
 * per-Gaussian latent through 32 -> 128 -> 128 -> 1024 (ReLU, ReLU, sigmoid) to
 * a 16x16 RGBA texture; these are its building blocks.
 *
 * tgs_gemm_bf16: C[M,N] = epilogue(A[M,K] B[K,N] + bias[N]) on the matrix cores
 * (v_mfma_f32_16x16x32_bf16): operands are float32 in memory, rounded to
 * bfloat16 (nearest even) when staged, accumulated in float32; an operand
 * flagged TGS_GEMM_A_BF16 / TGS_GEMM_B_BF16 (OR-ed into `epilogue`) is
 * bfloat16 bits in memory instead and is staged as is.  A element
 * (m,k) is A[m*lda+k] (trans_a 0) or A[k*lda+m] (trans_a 1); B element (k,n)
 * is B[k*ldb+n] (trans_b 0) or B[n*ldb+k] (trans_b 1).  Epilogues: NONE, RELU,
 * SIGMOID, RELU_MASK (keep where mask[m*ldm+n] > 0, for ReLU backward), ATOMIC
 * (C += A B with float atomics; no bias; k_splits > 1 splits K over
 * workgroups, C zeroed by the caller).
 * tgs_sigmoid_bwd: g = g_out * out * (1 - out).
 * tgs_colsum: db[n] += sum_m G[m*ldg+n] (float32).                         */
enum { TGS_EPI_NONE = 0, TGS_EPI_RELU = 1, TGS_EPI_SIGMOID = 2, TGS_EPI_RELU_MASK = 3, TGS_EPI_ATOMIC = 4 };
enum { TGS_GEMM_A_BF16 = 0x10, TGS_GEMM_B_BF16 = 0x20 };
int tgs_gemm_bf16(int trans_a, int trans_b, int epilogue, int64_t M, int64_t N, int64_t K,
                  const void *A, int64_t lda, const void *B, int64_t ldb, const float *bias,
                  float *C, int64_t ldc, const float *mask, int64_t ldm, int32_t k_splits,
                  tgs_stream_t stream);
int tgs_sigmoid_bwd(int64_t n, const float *g_out, const float *out, float *g, tgs_stream_t stream);
/* Sigmoid backward fused with the column sums of its result (no reference
 * counterpart; the decoder's output layer): g = g_out * o * (1 - o) for [M, N]
 * float32 (N a multiple of 4, 16-byte aligned), and db[N] += g.sum(0) of the
 * float32 values.  g is stored as float32 (g_bf16 0) or as bfloat16 bits
 * rounded to nearest even (g_bf16 1: what tgs_gemm_bf16 would stage). */
int tgs_sigmoid_bwd_colsum(int64_t M, int64_t N, const float *g_out, const float *out, void *g,
                           int g_bf16, float *db, tgs_stream_t stream);
int tgs_colsum(int64_t M, int64_t N, const float *G, int64_t ldg, float *db, tgs_stream_t stream);

/* Fused decoder inference: all three layers in one kernel, activations kept in
 * registers, with the GEMM path's numerics (bfloat16 operands, float32
 * accumulation, bias and activations in float32).
 * tgs_decoder_pack: W1 [32,128], W2 [128,128], W3 [128,1024] (row-major
 * [in, out], float32) -> tgs_decoder_pack_bytes() bytes of bfloat16 MFMA
 * fragments at `packed` (16-byte aligned).  Repack after every weight change.
 * tgs_decode_textures: out[i, :] = sigmoid(W3^T relu(W2^T relu(W1^T z[i] + b1)
 * + b2) + b3) for i < n; z [n,32], out [n,1024] (the [n,16,16,4] RGBA textures);
 * z, packed, biases and out 16-byte aligned.                                */
size_t tgs_decoder_pack_bytes(void);
int tgs_decoder_pack(const float *W1, const float *W2, const float *W3, void *packed, tgs_stream_t stream);
int tgs_decode_textures(int64_t n, const float *z, const void *packed, const float *b1, const float *b2,
                        const float *b3, float *out, tgs_stream_t stream);

/* ------------------------------------------------------------ fused Adam
 * Replaces `selective_adam_update` (ext.cpp:73, bindings.h:486-497, adam.cu):
 * rows g < N with visibility[g] != 0 are updated element-wise,
 * m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; p -= lr m / (sqrt(v) + eps)
 * (no bias correction).  param/grad/exp_avg/exp_avg_sq are f32[N*M].     */
int tgs_selective_adam_update(int64_t N, int64_t M, float *param, const float *grad,
                              float *exp_avg, float *exp_avg_sq, const uint8_t *visibility,
                              double lr, double b1, double b2, double eps,
                              tgs_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TGS_HIP_H */
