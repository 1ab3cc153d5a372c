// Neural-texture decoder kernels: a BF16 matrix-core GEMM with fused epilogues,
// and the element-wise pieces of the decoder MLP's forward and backward.
//
// The decoder maps a per-Gaussian latent z (32) through 32 -> 128 -> 128 -> 1024
// (ReLU, ReLU, sigmoid) to a 16x16 RGBA texture.  No reference counterpart (the
// reference's closest relative is its coordinate-MLP implicit texture,
// examples/texture_models/mlp.py:6-18).
//
// Numerics (mixed precision, as the decoder is specified): every GEMM operand is
// rounded to bfloat16 (round to nearest even) as it is staged, products are
// accumulated in float32 by v_mfma_f32_16x16x32_bf16, and bias, activations and
// their derivatives are float32.  Activations are stored in float32 and rounded
// again when the next GEMM stages them, so forward and backward see the same
// bfloat16 values.
//
// GEMM: C[M,N] = epi(A[M,K] B[K,N] + bias[N]).  A element (m,k) is A[m*lda+k]
// (TA = false) or A[k*lda+m] (TA = true); B element (k,n) is B[k*ldb+n] (TB =
// false) or B[n*ldb+k] (TB = true).  A 256-thread workgroup computes a 64x64
// tile: the four waves own 32x32 quadrants, each four 16x16 MFMA tiles, over K
// steps of 32 staged through LDS in the operand layout the MFMA reads (16 bytes
// per lane: lane l takes row l%16 of A and column l%16 of B, k = 8(l/16)..+8).
// Split-K accumulates partial tiles into C with atomics (C zeroed by the
// caller); it is used for weight gradients, whose K is the batch.  Either
// operand may also be given as bfloat16 bits (staged as is; the rounding then
// happened where it was written, to the same values).
// Tile order: consecutive workgroup ids go to the 8 XCDs in turn, so workgroup w
// runs tile (w % 8) * ceil(tiles / 8) + w / 8 and each XCD walks a contiguous
// range of tiles.  Tiles are numbered with the N tile fastest when K is not split
// (the N tiles of one row block share its A rows, read once into that XCD's L2)
// and with the M tile fastest under split-K (the tiles of one K slice share it).
#include "common.h"

namespace tgs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t bf16_bits(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// logistic function; the reciprocal is the hardware's (1 ulp)
__device__ __forceinline__ float sigmoidf(float v) { return __builtin_amdgcn_rcpf(1.f + __expf(-v)); }

enum { EPI_NONE = 0, EPI_RELU = 1, EPI_SIGMOID = 2, EPI_RELU_MASK = 3, EPI_ATOMIC = 4 };

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdsK = kBK + 8;  // row pitch in bf16 (80 bytes): spreads the 16 rows a wave reads over banks

// Column of the 8-element (16-byte) chunk holding k in row r of a staged tile:
// chunks are XOR-swizzled by bits 3-4 of the row, so the transposing stores
// (8 rows 8 apart, one k) fall into four different chunk columns instead of one
// (rows 8 apart are 160 dwords apart, the same bank).  Reads of 16-byte chunks
// stay aligned.
__device__ __forceinline__ int lds_col(int r, int k) { return (((k >> 3) ^ ((r >> 3) & 3)) << 3) | (k & 7); }

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// 8 consecutive operand elements as loaded: 8 float32, or 8 bfloat16 bits in `a`.
// They stay in registers while the previous K step's products run.
struct Raw8 {
    f32x4 a, b;
};

// Loads the 8 elements at P + off: P holds float32 or, with bf, bfloat16 bits.
// full: all 8 exist and 16-byte loads may be used; else element i is loaded when
// i < lim and is zero otherwise.
__device__ __forceinline__ Raw8 load8(const void *__restrict__ P, bool bf, int64_t off, bool full, int lim) {
    Raw8 r;
    if (bf) {
        const uint16_t *src = (const uint16_t *)P + off;
        u16x8 v;
        if (full) {
            v = *(const u16x8 *)src;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = i < lim ? src[i] : (uint16_t)0;
        }
        r.a = __builtin_bit_cast(f32x4, v);
        r.b = f32x4{0, 0, 0, 0};
    } else {
        const float *src = (const float *)P + off;
        if (full) {
            r.a = *(const f32x4 *)src;
            r.b = *(const f32x4 *)(src + 4);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                r.a[i] = i < lim ? src[i] : 0.f;
                r.b[i] = 4 + i < lim ? src[4 + i] : 0.f;
            }
        }
    }
    return r;
}

// the 8 elements as bfloat16 bits (float32 rounded to nearest even)
__device__ __forceinline__ u16x8 to_bf16x8(const Raw8 &r, bool bf) {
    if (bf) return __builtin_bit_cast(u16x8, r.a);
    u16x8 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = bf16_bits(r.a[i]);
        v[4 + i] = bf16_bits(r.b[i]);
    }
    return v;
}

// A 64 (rows r) x 32 (k) operand tile, staged into S[r][k] as bfloat16, zero
// outside the matrix.  KCONTIG: element (r, k) is P[r*ld + k]; else P[k*ld + r].
// Each thread moves 8 elements, along k or along r: stage_load issues its loads,
// stage_store converts and writes them to LDS.
template <bool KCONTIG>
__device__ __forceinline__ Raw8 stage_load(const void *__restrict__ P, bool bf, int64_t ld, int r0, int R, int kb,
                                           int k_end, bool vec, int t) {
    if (KCONTIG) {
        const int gr = r0 + (t >> 2), gk = kb + (t & 3) * 8;
        return load8(P, bf, (int64_t)gr * ld + gk, vec && gr < R && gk + 8 <= k_end, gr < R ? k_end - gk : 0);
    } else {
        const int gk = kb + (t >> 3), gr = r0 + (t & 7) * 8;
        return load8(P, bf, (int64_t)gk * ld + gr, vec && gk < k_end && gr + 8 <= R, gk < k_end ? R - gr : 0);
    }
}

template <bool KCONTIG>
__device__ __forceinline__ void stage_store(uint16_t (*S)[kLdsK], const Raw8 &raw, bool bf, int t) {
    const u16x8 v = to_bf16x8(raw, bf);
    if (KCONTIG) {
        const int rr = t >> 2, kk = (t & 3) * 8;
        *(u16x8 *)&S[rr][lds_col(rr, kk)] = v;
    } else {
        const int kk = t >> 3, rr = (t & 7) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) S[rr + i][lds_col(rr + i, kk)] = v[i];
    }
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const void *__restrict__ A, int64_t lda,
                                                        const void *__restrict__ B, int64_t ldb, int a_bf, int b_bf,
                                                        const float *__restrict__ bias, float *__restrict__ C,
                                                        int64_t ldc, const float *__restrict__ mask, int64_t ldm,
                                                        int k_chunk, int tiles_m, int tiles_n, int64_t tiles) {
    __shared__ __align__(16) uint16_t As[kBM][kLdsK];  // [m][k]
    __shared__ __align__(16) uint16_t Bs[kBN][kLdsK];  // [n][k]
    const int64_t w = blockIdx.x, T = (w % 8) * ((tiles + 7) / 8) + w / 8;
    if (T >= tiles) return;  // uniform over the workgroup
    int tm, tn, tz;
    if (k_chunk >= K) {  // K not split: N tile fastest
        tn = (int)(T % tiles_n);
        tm = (int)(T / tiles_n);
        tz = 0;
    } else {
        tm = (int)(T % tiles_m);
        tn = (int)((T / tiles_m) % tiles_n);
        tz = (int)(T / ((int64_t)tiles_m * tiles_n));
    }
    const int t = threadIdx.x, wave = t / kWave, lane = t % kWave;
    const int m0 = tm * kBM, n0 = tn * kBN;
    const int k_begin = tz * k_chunk;
    const int k_end = min(K, k_begin + k_chunk);
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;  // wave quadrant

    // 16-byte loads when the operand allows them
    const bool a_vec = (lda % (a_bf ? 8 : 4)) == 0 && ((uintptr_t)A % 16) == 0;
    const bool b_vec = (ldb % (b_bf ? 8 : 4)) == 0 && ((uintptr_t)B % 16) == 0;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

    // A tile (64 x 32) and B tile (32 x 64) of the first K step in registers
    Raw8 ra, rb;
    if (k_begin < k_end) {
        ra = stage_load<!TA>(A, a_bf != 0, lda, m0, M, k_begin, k_end, a_vec, t);
        rb = stage_load<TB>(B, b_bf != 0, ldb, n0, N, k_begin, k_end, b_vec, t);
    }
    for (int kb = k_begin; kb < k_end; kb += kBK) {
        // ---- stage this step's tiles as bf16, then issue the next step's loads, which
        // are in flight during this step's products
        stage_store<!TA>(As, ra, a_bf != 0, t);
        stage_store<TB>(Bs, rb, b_bf != 0, t);
        __syncthreads();
        if (kb + kBK < k_end) {
            ra = stage_load<!TA>(A, a_bf != 0, lda, m0, M, kb + kBK, k_end, a_vec, t);
            rb = stage_load<TB>(B, b_bf != 0, ldb, n0, N, kb + kBK, k_end, b_vec, t);
        }
        // ---- 2 x 2 MFMA tiles per wave
        bf16x8 af[2], bfr[2];
        const int r = lane % 16, kq = (lane / 16) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = *(const bf16x8 *)&As[wm + i * 16 + r][lds_col(wm + i * 16 + r, kq)];
#pragma unroll
        for (int j = 0; j < 2; ++j) bfr[j] = *(const bf16x8 *)&Bs[wn + j * 16 + r][lds_col(wn + j * 16 + r, kq)];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // ---- epilogue: lane holds rows 4(l/16)..+3 of column l%16 of each 16x16 tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int gn = n0 + wn + j * 16 + lane % 16;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gm = m0 + wm + i * 16 + (lane / 16) * 4 + q;
                if (gm >= M || gn >= N) continue;
                float v = acc[i][j][q];
                float *dst = C + (int64_t)gm * ldc + gn;
                if (EPI == EPI_ATOMIC) {
                    atomic_add(dst, v);
                    continue;
                }
                if (bias) v += bias[gn];
                if (EPI == EPI_RELU) v = v > 0.f ? v : 0.f;
                if (EPI == EPI_SIGMOID) v = sigmoidf(v);
                if (EPI == EPI_RELU_MASK) v = mask[(int64_t)gm * ldm + gn] > 0.f ? v : 0.f;
                *dst = v;
            }
        }
}

// g = g_out * out * (1 - out)  (sigmoid derivative, output already sigmoid)
__global__ __launch_bounds__(256) void sigmoid_bwd_kernel(int64_t n, const float *__restrict__ g_out,
                                                          const float *__restrict__ out, float *__restrict__ g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float o = out[i];
        g[i] = g_out[i] * o * (1.f - o);
    }
}

// sigmoid backward and the column sums of its result in one pass (the output
// layer's bias gradient): g = g_out * o * (1 - o) with 4-vectors, each thread
// keeps running sums for its 4 columns; one workgroup per (256-column strip,
// chunk of rows), one atomic per column per workgroup.  Saves re-reading g.
// BF: g is stored as bfloat16 (nearest even: the values the GEMMs that consume
// it would stage anyway), the sums are of the float32 values.
template <bool BF>
__global__ __launch_bounds__(256) void sigmoid_bwd_colsum_kernel(int M, int N4, const float4 *__restrict__ g_out,
                                                                 const float4 *__restrict__ out,
                                                                 void *__restrict__ g, int rows_per_block,
                                                                 float *__restrict__ db) {
    __shared__ float4 part[4][64];
    const int t = threadIdx.x, c = t % 64, rq = t / 64;
    const int n4 = blockIdx.x * 64 + c;
    const int r0 = blockIdx.y * rows_per_block, r1 = min(M, r0 + rows_per_block);
    float4 s = {0.f, 0.f, 0.f, 0.f};
    if (n4 < N4)
        for (int r = r0 + rq; r < r1; r += 4) {
            const int64_t i = (int64_t)r * N4 + n4;
            const float4 o = out[i], go = g_out[i];
            float4 gg;
            gg.x = go.x * o.x * (1.f - o.x);
            gg.y = go.y * o.y * (1.f - o.y);
            gg.z = go.z * o.z * (1.f - o.z);
            gg.w = go.w * o.w * (1.f - o.w);
            if (BF) {
                const uint32_t lo = bf16_bits(gg.x) | ((uint32_t)bf16_bits(gg.y) << 16);
                const uint32_t hi = bf16_bits(gg.z) | ((uint32_t)bf16_bits(gg.w) << 16);
                ((uint2 *)g)[i] = uint2{lo, hi};
            } else {
                ((float4 *)g)[i] = gg;
            }
            s.x += gg.x; s.y += gg.y; s.z += gg.z; s.w += gg.w;
        }
    part[rq][c] = s;
    __syncthreads();
    if (rq == 0 && n4 < N4) {
        const float4 a = part[0][c], b = part[1][c], cc = part[2][c], d = part[3][c];
        atomic_add(db + 4 * n4, a.x + b.x + cc.x + d.x);
        atomic_add(db + 4 * n4 + 1, a.y + b.y + cc.y + d.y);
        atomic_add(db + 4 * n4 + 2, a.z + b.z + cc.z + d.z);
        atomic_add(db + 4 * n4 + 3, a.w + b.w + cc.w + d.w);
    }
}

// column sums of G[M,N] added into db[N] (float32, unrounded); one workgroup per
// (64-column strip, chunk of rows), one atomic per column per workgroup
__global__ __launch_bounds__(256) void colsum_kernel(int M, int N, const float *__restrict__ G, int64_t ldg,
                                                     int rows_per_block, float *__restrict__ db) {
    __shared__ float part[4][64];
    const int t = threadIdx.x, c = t % 64, rq = t / 64;
    const int n = blockIdx.x * 64 + c;
    const int r0 = blockIdx.y * rows_per_block, r1 = min(M, r0 + rows_per_block);
    float s = 0.f;
    if (n < N)
        for (int r = r0 + rq; r < r1; r += 4) s += G[(int64_t)r * ldg + n];
    part[rq][c] = s;
    __syncthreads();
    if (rq == 0 && n < N) atomic_add(db + n, part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

// The same with 4-vectors: 32 column quads by 8 row lanes per workgroup, eight
// rows in flight per lane, and row chunks long enough that few workgroups add
// into each column (the final atomics to one address serialise).
__global__ __launch_bounds__(256) void colsum4_kernel(int M, int N4, const float4 *__restrict__ G, int64_t ldg4,
                                                      int rows_per_block, float *__restrict__ db) {
    __shared__ float4 part[8][32];
    const int t = threadIdx.x, c = t % 32, rq = t / 32;
    const int n4 = blockIdx.x * 32 + c;
    const int r0 = blockIdx.y * rows_per_block, r1 = min(M, r0 + rows_per_block);
    float4 s = {0.f, 0.f, 0.f, 0.f};
    if (n4 < N4) {
        int r = r0 + rq;
        for (; r + 56 < r1; r += 64) {
            float4 x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = G[(int64_t)(r + 8 * u) * ldg4 + n4];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                s.x += x[u].x;
                s.y += x[u].y;
                s.z += x[u].z;
                s.w += x[u].w;
            }
        }
        for (; r < r1; r += 8) {
            const float4 x = G[(int64_t)r * ldg4 + n4];
            s.x += x.x;
            s.y += x.y;
            s.z += x.z;
            s.w += x.w;
        }
    }
    part[rq][c] = s;
    __syncthreads();
    if (rq == 0 && n4 < N4) {
        float4 a = part[0][c];
#pragma unroll
        for (int q = 1; q < 8; ++q) {
            a.x += part[q][c].x;
            a.y += part[q][c].y;
            a.z += part[q][c].z;
            a.w += part[q][c].w;
        }
        atomic_add(db + 4 * n4, a.x);
        atomic_add(db + 4 * n4 + 1, a.y);
        atomic_add(db + 4 * n4 + 2, a.z);
        atomic_add(db + 4 * n4 + 3, a.w);
    }
}

// ------------------------------------------------------------ fused decode
// Inference of the whole decoder in one kernel: latents in, float32 textures out,
// activations never leave registers.  Each layer is computed transposed,
// H_out^T = W^T H_in^T, so a 16x16 MFMA tile holds 16 output features (rows) x
// 16 latents (columns) and lane l holds features 4(l/16)..+3 of latent l%16.
// Rounded to bfloat16, those 4 values of two consecutive feature tiles are exactly
// the lane's 8 B-operand values of the next layer's K step, provided that layer
// visits K in the order the lanes hold it (kperm); the weights are packed once per
// call into 1 KB MFMA A-fragments in that order, so no activation is exchanged
// between lanes.  A wave owns 4 column tiles (64 latents) and reuses each weight
// fragment (read from L2, 296 KB in all) across them.  The output row of a latent
// is its 1024 features in order, i.e. the [16, 16, 4] RGBA texture, written through
// an LDS slab in 512-byte pieces.  The kernel is bound by those 4 KB-per-latent
// writes; the matrix-core work is a small fraction of its time.
constexpr int kDecIn = 32, kDecHid = 128, kDecOut = 1024;
constexpr int kDecNT = 4;                                      // 16-latent column tiles per wave
constexpr int kDecRows = 4 * kDecNT * 16;                      // latents per 256-thread workgroup
constexpr int kF1 = 0, kF2 = 8, kF3 = 8 + 32, kFN = 8 + 32 + 256;  // fragment offsets per layer

// input feature held in k-slot s (0..31) of K step kt, for layers fed by a D-tile pair
__device__ __forceinline__ int kperm(int kt, int s) {
    const int g = s / 8, j = s % 8;
    return 32 * kt + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

// fragment f, lane l, slot j -> bf16(W^T[row][k]) with row = 16 mt + l%16, k per layer
__global__ __launch_bounds__(256) void decoder_pack_kernel(const float *__restrict__ W1, const float *__restrict__ W2,
                                                           const float *__restrict__ W3, uint16_t *__restrict__ frag) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= kFN * 512) return;
    const int f = e / 512, lane = (e / 8) % 64, j = e % 8;
    const int g = lane / 16, c = lane % 16;
    float v;
    if (f < kF2) {
        const int mt = f;
        v = W1[(8 * g + j) * kDecHid + 16 * mt + c];
    } else if (f < kF3) {
        const int mt = (f - kF2) / 4, kt = (f - kF2) % 4;
        v = W2[kperm(kt, 8 * g + j) * kDecHid + 16 * mt + c];
    } else {
        const int mt = (f - kF3) / 4, kt = (f - kF3) % 4;
        v = W3[kperm(kt, 8 * g + j) * kDecOut + 16 * mt + c];
    }
    frag[e] = bf16_bits(v);
}

__device__ __forceinline__ bf16x8 ld_frag(const u32x4v *__restrict__ frag, int f, int lane) {
    return __builtin_bit_cast(bf16x8, frag[f * 64 + lane]);
}

// two D tiles (features 4g..+3 of tiles 2kt and 2kt+1) -> one bfloat16 B fragment
__device__ __forceinline__ bf16x8 to_bfrag(const f32x4 &lo, const f32x4 &hi) {
    return bf16x8{(__bf16)lo[0], (__bf16)lo[1], (__bf16)lo[2], (__bf16)lo[3],
                  (__bf16)hi[0], (__bf16)hi[1], (__bf16)hi[2], (__bf16)hi[3]};
}

__device__ __forceinline__ f32x4 ld4(const float *p) { return *(const f32x4 *)p; }

// Layer 3 of the fused decode: 64 tiles of 16 features, each fragment group and
// bias prefetched one tile ahead.  Finished tiles are staged in the wave's LDS
// slab (64 latents x 128 features); every 8 tiles the wave writes them out with
// non-temporal stores, two latents x 512 contiguous bytes per instruction (direct
// 64-byte pieces per tile reach about half the write bandwidth).  FULL (every
// latent of the wave exists) keeps the stores unconditional, so the wait for the
// prefetch never waits for them.
constexpr int kDecGroup = 8;                         // tiles per staged group
constexpr int kDecPitch = 16 * kDecGroup + 4;        // floats per staged latent row
constexpr int kDecSlab = 16 * kDecNT * kDecPitch;    // floats per wave

template <bool FULL>
__device__ __forceinline__ void decode_out(const u32x4v *__restrict__ frag, const bf16x8 (&h2)[4][kDecNT],
                                           const float *__restrict__ b3, float *__restrict__ out, int64_t base,
                                           int64_t n, int lane, float *__restrict__ slab) {
    const int g = lane / 16, c = lane % 16;
    const f32x4 zero = {0, 0, 0, 0};
    bf16x8 A[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) A[kt] = ld_frag(frag, kF3 + kt, lane);
    f32x4 bias = ld4(b3 + 4 * g);
    auto tile = [&](int mt) {
        f32x4 d[kDecNT];
#pragma unroll
        for (int t = 0; t < kDecNT; ++t) d[t] = zero;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int t = 0; t < kDecNT; ++t) d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kt], h2[kt][t], d[t], 0, 0, 0);
        const f32x4 bc = bias;
        if (mt + 1 < kDecOut / 16) {
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) A[kt] = ld_frag(frag, kF3 + (mt + 1) * 4 + kt, lane);
            bias = ld4(b3 + 16 * (mt + 1) + 4 * g);
        }
#pragma unroll
        for (int t = 0; t < kDecNT; ++t) {
            f32x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = sigmoidf(d[t][r] + bc[r]);
            *(f32x4 *)(slab + (16 * t + c) * kDecPitch + 16 * (mt % kDecGroup) + 4 * g) = o;
        }
        if (mt % kDecGroup == kDecGroup - 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            constexpr int kLanesPer = 4 * kDecGroup;  // lanes per latent row piece
            constexpr int kLpi = kWave / kLanesPer;   // latents per store instruction
            const int col = 4 * (lane % kLanesPer);
#pragma unroll
            for (int i = 0; i < 16 * kDecNT / kLpi; ++i) {
                const int L = kLpi * i + lane / kLanesPer;
                const f32x4 v = *(const f32x4 *)(slab + L * kDecPitch + col);
                if (FULL || base + L < n)
                    __builtin_nontemporal_store(
                        v, (f32x4 *)(out + (base + L) * kDecOut + 16 * kDecGroup * (mt / kDecGroup) + col));
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    };
    // tile 0 peeled: every entry into the loop then has the prefetch followed by a
    // tile's work in flight, and the wait for the prefetch leaves the stores alone
    tile(0);
#pragma unroll 1
    for (int mt = 1; mt < kDecOut / 16; ++mt) tile(mt);
}

__global__ __launch_bounds__(256) void decode_kernel(int64_t n, const float *__restrict__ z,
                                                     const u32x4v *__restrict__ frag, const float *__restrict__ b1,
                                                     const float *__restrict__ b2, const float *__restrict__ b3,
                                                     float *__restrict__ out) {
    __shared__ __align__(16) float slabs[4 * kDecSlab];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t base = (int64_t)blockIdx.x * kDecRows + (int64_t)wave * (16 * kDecNT);
    if (base >= n) return;  // wave-uniform; the waves share no LDS and no barrier
    const int g = lane / 16, c = lane % 16;
    // layer-1 B operand: latent base + 16 t + c, input features 8g..8g+7
    bf16x8 bz[kDecNT];
#pragma unroll
    for (int t = 0; t < kDecNT; ++t) {
        const int64_t row = base + 16 * t + c;
        f32x4 lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
        if (row < n) {
            lo = ld4(z + row * kDecIn + 8 * g);
            hi = ld4(z + row * kDecIn + 8 * g + 4);
        }
        bz[t] = to_bfrag(lo, hi);
    }
    const f32x4 zero = {0, 0, 0, 0};
    // ---- layer 1: 32 -> 128, ReLU
    bf16x8 h1[4][kDecNT];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
        f32x4 d[2][kDecNT];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int mt = 2 * kt + h;
            const bf16x8 A = ld_frag(frag, kF1 + mt, lane);
            const f32x4 bias = ld4(b1 + 16 * mt + 4 * g);
#pragma unroll
            for (int t = 0; t < kDecNT; ++t) {
                d[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, bz[t], zero, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) d[h][t][r] = fmaxf(d[h][t][r] + bias[r], 0.f);
            }
        }
#pragma unroll
        for (int t = 0; t < kDecNT; ++t) h1[kt][t] = to_bfrag(d[0][t], d[1][t]);
    }
    // ---- layer 2: 128 -> 128, ReLU
    bf16x8 h2[4][kDecNT];
#pragma unroll
    for (int kt2 = 0; kt2 < 4; ++kt2) {
        f32x4 d[2][kDecNT];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int mt = 2 * kt2 + h;
#pragma unroll
            for (int t = 0; t < kDecNT; ++t) d[h][t] = zero;
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const bf16x8 A = ld_frag(frag, kF2 + mt * 4 + kt, lane);
#pragma unroll
                for (int t = 0; t < kDecNT; ++t)
                    d[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, h1[kt][t], d[h][t], 0, 0, 0);
            }
            const f32x4 bias = ld4(b2 + 16 * mt + 4 * g);
#pragma unroll
            for (int t = 0; t < kDecNT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) d[h][t][r] = fmaxf(d[h][t][r] + bias[r], 0.f);
        }
#pragma unroll
        for (int t = 0; t < kDecNT; ++t) h2[kt2][t] = to_bfrag(d[0][t], d[1][t]);
    }
    // ---- layer 3: 128 -> 1024, sigmoid, streamed to the output tile by tile
    float *slab = slabs + wave * kDecSlab;
    if (base + 16 * kDecNT <= n) decode_out<true>(frag, h2, b3, out, base, n, lane, slab);
    else decode_out<false>(frag, h2, b3, out, base, n, lane, slab);
}

template <bool TA, bool TB, int EPI>
static void launch_gemm(int M, int N, int K, const void *A, int64_t lda, const void *B, int64_t ldb, int a_bf,
                        int b_bf, const float *bias, float *C, int64_t ldc, const float *mask, int64_t ldm,
                        int splits, hipStream_t st) {
    const int chunk = K > 0 ? ((K + splits - 1) / splits + kBK - 1) / kBK * kBK : kBK;
    const int nz = K > 0 ? (K + chunk - 1) / chunk : 1;
    const int tm = (int)ceil_div(M, kBM), tn = (int)ceil_div(N, kBN);
    const int64_t tiles = (int64_t)tm * tn * nz;
    const int64_t grid = (tiles + 7) / 8 * 8;
    gemm_bf16_kernel<TA, TB, EPI><<<(unsigned)grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, a_bf, b_bf, bias, C, ldc,
                                                                   mask, ldm, chunk, tm, tn, tiles);
}

}  // namespace tgs

using namespace tgs;

extern "C" int tgs_gemm_bf16(int trans_a, int trans_b, int epilogue, int64_t M, int64_t N, int64_t K, const void *A,
                             int64_t lda, const void *B, int64_t ldb, const float *bias, float *C, int64_t ldc,
                             const float *mask, int64_t ldm, int32_t k_splits, tgs_stream_t stream) {
    TGS_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && M < (int64_t(1) << 31) && N < (int64_t(1) << 31) &&
                      K < (int64_t(1) << 31),
                  "gemm_bf16: bad sizes M=%lld N=%lld K=%lld", (long long)M, (long long)N, (long long)K);
    const int a_bf = (epilogue & TGS_GEMM_A_BF16) != 0, b_bf = (epilogue & TGS_GEMM_B_BF16) != 0;
    epilogue &= ~(TGS_GEMM_A_BF16 | TGS_GEMM_B_BF16);
    TGS_CHECK_ARG(epilogue >= EPI_NONE && epilogue <= EPI_ATOMIC, "gemm_bf16: bad epilogue %d", epilogue);
    TGS_CHECK_ARG((K == 0 || (A && B)) && (M == 0 || N == 0 || C), "gemm_bf16: null operand");
    TGS_CHECK_ARG(epilogue != EPI_RELU_MASK || mask, "gemm_bf16: the ReLU-mask epilogue needs a mask");
    TGS_CHECK_ARG(k_splits >= 1 && (k_splits == 1 || epilogue == EPI_ATOMIC),
                  "gemm_bf16: split-K needs the atomic epilogue");
    TGS_CHECK_ARG(epilogue != EPI_ATOMIC || !bias, "gemm_bf16: the atomic epilogue takes no bias");
    if (M == 0 || N == 0) return TGS_OK;
    const int64_t tiles_mn = ceil_div(M, kBM) * ceil_div(N, kBN);  // < 2^50
    TGS_CHECK_ARG(tiles_mn < (int64_t(1) << 31) && tiles_mn * (K / kBK + 1) < (int64_t(1) << 31),
                  "gemm_bf16: too many tiles (M=%lld N=%lld K=%lld)", (long long)M, (long long)N, (long long)K);
    hipStream_t st = (hipStream_t)stream;
    const int m = (int)M, n = (int)N, k = (int)K;
#define TGS_G(TA, TB, E) \
    launch_gemm<TA, TB, E>(m, n, k, A, lda, B, ldb, a_bf, b_bf, bias, C, ldc, mask, ldm, k_splits, st)
#define TGS_GE(TA, TB)                                  \
    switch (epilogue) {                                 \
        case EPI_NONE: TGS_G(TA, TB, EPI_NONE); break;  \
        case EPI_RELU: TGS_G(TA, TB, EPI_RELU); break;  \
        case EPI_SIGMOID: TGS_G(TA, TB, EPI_SIGMOID); break; \
        case EPI_RELU_MASK: TGS_G(TA, TB, EPI_RELU_MASK); break; \
        default: TGS_G(TA, TB, EPI_ATOMIC); break;      \
    }
    if (!trans_a && !trans_b) { TGS_GE(false, false) }
    else if (!trans_a && trans_b) { TGS_GE(false, true) }
    else if (trans_a && !trans_b) { TGS_GE(true, false) }
    else { TGS_GE(true, true) }
#undef TGS_GE
#undef TGS_G
    TGS_CHECK_LAUNCH("gemm_bf16");
    return TGS_OK;
}

extern "C" size_t tgs_decoder_pack_bytes(void) { return (size_t)kFN * 512 * sizeof(uint16_t); }

extern "C" int tgs_decoder_pack(const float *W1, const float *W2, const float *W3, void *packed,
                                tgs_stream_t stream) {
    TGS_CHECK_ARG(W1 && W2 && W3 && packed, "decoder_pack: null argument");
    TGS_CHECK_ARG(((uintptr_t)packed & 15) == 0, "decoder_pack: packed weights must be 16-byte aligned");
    decoder_pack_kernel<<<(unsigned)ceil_div(kFN * 512, 256), 256, 0, (hipStream_t)stream>>>(W1, W2, W3,
                                                                                           (uint16_t *)packed);
    TGS_CHECK_LAUNCH("decoder_pack");
    return TGS_OK;
}

extern "C" int tgs_decode_textures(int64_t n, const float *z, const void *packed, const float *b1, const float *b2,
                                   const float *b3, float *out, tgs_stream_t stream) {
    TGS_CHECK_ARG(n >= 0 && n < (int64_t(1) << 40), "decode_textures: bad n=%lld", (long long)n);
    if (n == 0) return TGS_OK;
    TGS_CHECK_ARG(z && packed && b1 && b2 && b3 && out, "decode_textures: null argument");
    TGS_CHECK_ARG(((uintptr_t)z & 15) == 0 && ((uintptr_t)packed & 15) == 0 && ((uintptr_t)out & 15) == 0 &&
                      ((uintptr_t)b1 & 15) == 0 && ((uintptr_t)b2 & 15) == 0 && ((uintptr_t)b3 & 15) == 0,
                  "decode_textures: z, packed weights, biases and out must be 16-byte aligned");
    decode_kernel<<<(unsigned)ceil_div(n, kDecRows), 256, 0, (hipStream_t)stream>>>(
        n, z, (const u32x4v *)packed, b1, b2, b3, out);
    TGS_CHECK_LAUNCH("decode_textures");
    return TGS_OK;
}

extern "C" int tgs_sigmoid_bwd(int64_t n, const float *g_out, const float *out, float *g, tgs_stream_t stream) {
    TGS_CHECK_ARG(n >= 0 && (n == 0 || (g_out && out && g)), "sigmoid_bwd: bad arguments");
    if (n == 0) return TGS_OK;
    sigmoid_bwd_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(n, g_out, out, g);
    TGS_CHECK_LAUNCH("sigmoid_bwd");
    return TGS_OK;
}

extern "C" int tgs_sigmoid_bwd_colsum(int64_t M, int64_t N, const float *g_out, const float *out, void *g,
                                      int g_bf16, float *db, tgs_stream_t stream) {
    TGS_CHECK_ARG(M >= 0 && N >= 0 && M < (int64_t(1) << 31) && N % 4 == 0 && (M == 0 || (g_out && out && g && db)),
                  "sigmoid_bwd_colsum: bad arguments (N must be a multiple of 4)");
    TGS_CHECK_ARG((((uintptr_t)g_out | (uintptr_t)out | (uintptr_t)g) % 16) == 0,
                  "sigmoid_bwd_colsum: buffers must be 16-byte aligned");
    if (M == 0 || N == 0) return TGS_OK;
    const int rows = 4096;
    const int64_t N4 = N / 4;
    dim3 grid((unsigned)ceil_div(N4, 64), (unsigned)ceil_div(M, rows));
    if (g_bf16)
        sigmoid_bwd_colsum_kernel<true><<<grid, 256, 0, (hipStream_t)stream>>>((int)M, (int)N4, (const float4 *)g_out,
                                                                               (const float4 *)out, g, rows, db);
    else
        sigmoid_bwd_colsum_kernel<false><<<grid, 256, 0, (hipStream_t)stream>>>(
            (int)M, (int)N4, (const float4 *)g_out, (const float4 *)out, g, rows, db);
    TGS_CHECK_LAUNCH("sigmoid_bwd_colsum");
    return TGS_OK;
}

extern "C" int tgs_colsum(int64_t M, int64_t N, const float *G, int64_t ldg, float *db, tgs_stream_t stream) {
    TGS_CHECK_ARG(M >= 0 && N >= 0 && M < (int64_t(1) << 31) && N < (int64_t(1) << 31) && (M == 0 || (G && db)),
                  "colsum: bad arguments");
    if (M == 0 || N == 0) return TGS_OK;
    if (N % 4 == 0 && ldg % 4 == 0 && ((uintptr_t)G % 16) == 0) {
        const int rows = 2048;
        dim3 grid((unsigned)ceil_div(N / 4, 32), (unsigned)ceil_div(M, rows));
        colsum4_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((int)M, (int)(N / 4), (const float4 *)G, ldg / 4, rows,
                                                              db);
    } else {
        const int rows = 4096;
        dim3 grid((unsigned)ceil_div(N, 64), (unsigned)ceil_div(M, rows));
        colsum_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((int)M, (int)N, G, ldg, rows, db);
    }
    TGS_CHECK_LAUNCH("colsum");
    return TGS_OK;
}
