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
// Split-K (grid.z > 1) accumulates partial tiles into C with atomics (C zeroed
// by the caller); it is used for weight gradients, whose K is the batch.
#include "common.h"

namespace tgs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

enum { EPI_NONE = 0, EPI_RELU = 1, EPI_SIGMOID = 2, EPI_RELU_MASK = 3, EPI_ATOMIC = 4 };

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdsK = kBK + 8;  // row pitch in bf16 (80 bytes): spreads the 16 rows a wave reads over banks

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                                                        const float *__restrict__ B, int64_t ldb,
                                                        const float *__restrict__ bias, float *__restrict__ C,
                                                        int64_t ldc, const float *__restrict__ mask, int64_t ldm,
                                                        int k_chunk) {
    __shared__ __align__(16) uint16_t As[kBM][kLdsK];  // [m][k]
    __shared__ __align__(16) uint16_t Bs[kBN][kLdsK];  // [n][k]
    const int t = threadIdx.x, wave = t / kWave, lane = t % kWave;
    const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
    const int k_begin = blockIdx.z * k_chunk;
    const int k_end = min(K, k_begin + k_chunk);
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;  // wave quadrant

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

    for (int kb = k_begin; kb < k_end; kb += kBK) {
        // ---- stage A tile (64 x 32) and B tile (32 x 64) as bf16, zero outside the matrices
#pragma unroll
        for (int it = 0; it < (kBM * kBK) / 256; ++it) {
            const int e = it * 256 + t;
            int mm, kk;
            if (!TA) { mm = e / kBK; kk = e % kBK; }  // k fastest: coalesced along k
            else { kk = e / kBM; mm = e % kBM; }      // m fastest: coalesced along m
            const int gm = m0 + mm, gk = kb + kk;
            float v = 0.f;
            if (gm < M && gk < k_end) v = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
            As[mm][kk] = bf16_bits(v);
        }
#pragma unroll
        for (int it = 0; it < (kBN * kBK) / 256; ++it) {
            const int e = it * 256 + t;
            int nn, kk;
            if (!TB) { kk = e / kBN; nn = e % kBN; }  // n fastest
            else { nn = e / kBK; kk = e % kBK; }      // k fastest
            const int gn = n0 + nn, gk = kb + kk;
            float v = 0.f;
            if (gn < N && gk < k_end) v = TB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
            Bs[nn][kk] = bf16_bits(v);
        }
        __syncthreads();
        // ---- 2 x 2 MFMA tiles per wave
        bf16x8 af[2], bfr[2];
        const int r = lane % 16, kq = (lane / 16) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) af[i] = *(const bf16x8 *)&As[wm + i * 16 + r][kq];
#pragma unroll
        for (int j = 0; j < 2; ++j) bfr[j] = *(const bf16x8 *)&Bs[wn + j * 16 + r][kq];
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
                if (EPI == EPI_SIGMOID) v = 1.f / (1.f + __expf(-v));
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

template <bool TA, bool TB, int EPI>
static void launch_gemm(int M, int N, int K, const float *A, int64_t lda, const float *B, int64_t ldb,
                        const float *bias, float *C, int64_t ldc, const float *mask, int64_t ldm, int splits,
                        hipStream_t st) {
    const int chunk = ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;
    const int nz = (K + chunk - 1) / chunk;
    dim3 grid((unsigned)ceil_div(M, kBM), (unsigned)ceil_div(N, kBN), (unsigned)max(nz, 1));
    gemm_bf16_kernel<TA, TB, EPI><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, bias, C, ldc, mask, ldm, chunk);
}

}  // namespace tgs

using namespace tgs;

extern "C" int tgs_gemm_bf16(int trans_a, int trans_b, int epilogue, int64_t M, int64_t N, int64_t K, const float *A,
                             int64_t lda, const float *B, int64_t ldb, const float *bias, float *C, int64_t ldc,
                             const float *mask, int64_t ldm, int32_t k_splits, tgs_stream_t stream) {
    TGS_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && M < (int64_t(1) << 31) && N < (int64_t(1) << 31) &&
                      K < (int64_t(1) << 31),
                  "gemm_bf16: bad sizes M=%lld N=%lld K=%lld", (long long)M, (long long)N, (long long)K);
    TGS_CHECK_ARG(epilogue >= EPI_NONE && epilogue <= EPI_ATOMIC, "gemm_bf16: bad epilogue %d", epilogue);
    TGS_CHECK_ARG(A && B && C, "gemm_bf16: null operand");
    TGS_CHECK_ARG(epilogue != EPI_RELU_MASK || mask, "gemm_bf16: the ReLU-mask epilogue needs a mask");
    TGS_CHECK_ARG(k_splits >= 1 && (k_splits == 1 || epilogue == EPI_ATOMIC),
                  "gemm_bf16: split-K needs the atomic epilogue");
    TGS_CHECK_ARG(epilogue != EPI_ATOMIC || !bias, "gemm_bf16: the atomic epilogue takes no bias");
    if (M == 0 || N == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    const int m = (int)M, n = (int)N, k = (int)K;
#define TGS_G(TA, TB, E) launch_gemm<TA, TB, E>(m, n, k, A, lda, B, ldb, bias, C, ldc, mask, ldm, k_splits, st)
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

extern "C" int tgs_sigmoid_bwd(int64_t n, const float *g_out, const float *out, float *g, tgs_stream_t stream) {
    TGS_CHECK_ARG(n >= 0 && (n == 0 || (g_out && out && g)), "sigmoid_bwd: bad arguments");
    if (n == 0) return TGS_OK;
    sigmoid_bwd_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(n, g_out, out, g);
    TGS_CHECK_LAUNCH("sigmoid_bwd");
    return TGS_OK;
}

extern "C" int tgs_colsum(int64_t M, int64_t N, const float *G, int64_t ldg, float *db, tgs_stream_t stream) {
    TGS_CHECK_ARG(M >= 0 && N >= 0 && M < (int64_t(1) << 31) && N < (int64_t(1) << 31) && (M == 0 || (G && db)),
                  "colsum: bad arguments");
    if (M == 0 || N == 0) return TGS_OK;
    const int rows = 4096;
    dim3 grid((unsigned)ceil_div(N, 64), (unsigned)ceil_div(M, rows));
    colsum_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((int)M, (int)N, G, ldg, rows, db);
    TGS_CHECK_LAUNCH("colsum");
    return TGS_OK;
}
