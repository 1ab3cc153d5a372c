// Shared device helpers and C-ABI error plumbing for libtgs_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "tgs_hip.h"

namespace tgs {

// ------------------------------------------------------------ error plumbing
void set_error(const char *fmt, ...);

#define TGS_CHECK_ARG(cond, ...)              \
    do {                                      \
        if (!(cond)) {                        \
            ::tgs::set_error(__VA_ARGS__);    \
            return TGS_ERR_ARG;               \
        }                                     \
    } while (0)

#define TGS_CHECK_LAUNCH(what)                                                  \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) {                                                 \
            ::tgs::set_error("%s: launch failed: %s", what, hipGetErrorString(e_)); \
            return TGS_ERR_LAUNCH;                                              \
        }                                                                       \
    } while (0)

constexpr int kWave = 64;  // CDNA wavefront width

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------ small vectors
template <typename T> struct v3 {
    T x, y, z;
};
template <typename T> __device__ __forceinline__ v3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <typename T> __device__ __forceinline__ v3<T> operator+(v3<T> a, v3<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T> __device__ __forceinline__ v3<T> operator-(v3<T> a, v3<T> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T> __device__ __forceinline__ v3<T> operator*(T s, v3<T> a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <typename T> __device__ __forceinline__ T dot(v3<T> a, v3<T> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T> __device__ __forceinline__ v3<T> cross(v3<T> a, v3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 3x3 matrix, row-major: m[r][c]
template <typename T> struct m33 {
    T m[3][3];
};
template <typename T> __device__ __forceinline__ m33<T> matmul(const m33<T> &a, const m33<T> &b) {
    m33<T> o;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) o.m[r][c] = a.m[r][0] * b.m[0][c] + a.m[r][1] * b.m[1][c] + a.m[r][2] * b.m[2][c];
    return o;
}
template <typename T> __device__ __forceinline__ m33<T> transpose(const m33<T> &a) {
    m33<T> o;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) o.m[r][c] = a.m[c][r];
    return o;
}
template <typename T> __device__ __forceinline__ v3<T> matvec(const m33<T> &a, v3<T> v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}

template <typename T> __device__ __forceinline__ T tgs_exp(T x);
template <> __device__ __forceinline__ float tgs_exp<float>(float x) { return __expf(x); }
template <> __device__ __forceinline__ double tgs_exp<double>(double x) { return exp(x); }

template <typename T> __device__ __forceinline__ T tgs_rsqrt(T x);
template <> __device__ __forceinline__ float tgs_rsqrt<float>(float x) { return rsqrtf(x); }
template <> __device__ __forceinline__ double tgs_rsqrt<double>(double x) { return 1.0 / sqrt(x); }

// Atomic add without return value (lowers to global_atomic_add_f32/f64).
template <typename T> __device__ __forceinline__ void atomic_add(T *p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// DPP lane permutes (gfx9 encodings): quad_perm, row_ror:n, row_bcast:15/31.
constexpr int kDppQuad1032 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuad2301 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppRowRor4 = 0x124;
constexpr int kDppRowRor8 = 0x128;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF> __device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK = 0xF> __device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROW_MASK = 0xF> __device__ __forceinline__ float dpp(float v) { return dpp_f32<CTRL, ROW_MASK>(v); }
template <int CTRL, int ROW_MASK = 0xF> __device__ __forceinline__ double dpp(double v) { return dpp_f64<CTRL, ROW_MASK>(v); }

__device__ __forceinline__ float readlane63(float v) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double readlane63(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of the wavefront, returned (uniform) in every lane;
// every lane must be active.  Quad and row rotations leave each 16-lane row
// holding its row sum; row_bcast:15 then row_bcast:31 fold the rows so lane 63
// holds the total.
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    v += dpp<kDppQuad1032>(v);
    v += dpp<kDppQuad2301>(v);
    v += dpp<kDppRowRor4>(v);
    v += dpp<kDppRowRor8>(v);
    v += dpp<kDppRowBcast15, 0xA>(v);
    v += dpp<kDppRowBcast31, 0xC>(v);
    return readlane63(v);
}
// Reduce-scatter of N per-lane values over the wavefront.  Six halving steps, one
// per lane bit from 5 down to 0: the lanes whose bit is clear keep the first half of
// the values, the others the second half, and each adds its partner's copy of the
// half it keeps.  Partners differ in that bit: v_permlane32_swap / v_permlane16_swap
// exchange wave halves / 16-lane rows (bits 5, 4), DPP row_mirror, row_half_mirror
// and two quad permutations pair lanes across bits 3..0.  Afterwards lane l holds
// the wave total of value wave_rs_slot<N>(l) (-1: none); each value is held by
// exactly one lane.  About 2.6 instructions per value instead of 7 for a wave_sum.
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppQuad3210 = 0x1B;

__device__ __forceinline__ void permlane_swap32(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void permlane_swap16(float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void permlane_swap32(double &a, double &b) {
    uint2 x = __builtin_bit_cast(uint2, a), y = __builtin_bit_cast(uint2, b);
    const auto lo = __builtin_amdgcn_permlane32_swap(x.x, y.x, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(x.y, y.y, false, false);
    a = __builtin_bit_cast(double, make_uint2(lo[0], hi[0]));
    b = __builtin_bit_cast(double, make_uint2(lo[1], hi[1]));
}
__device__ __forceinline__ void permlane_swap16(double &a, double &b) {
    uint2 x = __builtin_bit_cast(uint2, a), y = __builtin_bit_cast(uint2, b);
    const auto lo = __builtin_amdgcn_permlane16_swap(x.x, y.x, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(x.y, y.y, false, false);
    a = __builtin_bit_cast(double, make_uint2(lo[0], hi[0]));
    b = __builtin_bit_cast(double, make_uint2(lo[1], hi[1]));
}

// a, b: this lane's values of the two halves; returns the kept half plus the partner's
template <int B, typename T> __device__ __forceinline__ T rs_pair(T a, T b, bool bit) {
    if constexpr (B == 5) {
        permlane_swap32(a, b);
        return a + b;
    } else if constexpr (B == 4) {
        permlane_swap16(a, b);
        return a + b;
    } else {
        constexpr int ctrl = B == 3 ? kDppRowMirror : B == 2 ? kDppRowHalfMirror : B == 1 ? kDppQuad3210 : kDppQuad1032;
        const T keep = bit ? b : a, send = bit ? a : b;
        return keep + dpp<ctrl>(send);
    }
}

template <int n, int B, typename T, int CAP>
__device__ __forceinline__ void rs_step(T (&v)[CAP], int lane) {
    constexpr int m = (n + 1) / 2;
    const bool bit = (lane >> B) & 1;
#pragma unroll
    for (int i = 0; i < m; ++i) v[i] = rs_pair<B>(v[i], m + i < n ? v[m + i] : T(0), bit);
    if constexpr (B > 0) rs_step<m, B - 1>(v, lane);
}

template <int N, typename T> __device__ __forceinline__ T wave_reduce_scatter(T (&v)[N], int lane) {
    rs_step<N, 5>(v, lane);
    return v[0];
}

template <int N> __device__ __forceinline__ int wave_rs_slot(int lane) {
    int n = N, off = 0, cnt = N;
#pragma unroll
    for (int b = 5; b >= 0; --b) {
        const int m = (n + 1) / 2;
        if ((lane >> b) & 1) {
            off += m;
            cnt = cnt > m ? cnt - m : 0;
        } else {
            cnt = cnt < m ? cnt : m;
        }
        n = m;
    }
    return cnt >= 1 ? off : -1;
}

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}

// Block-index remap in runs: run j of `run` consecutive work items goes to XCD j mod 8,
// so every XCD takes runs from the whole range (its share of the work follows the
// average when the work per item varies smoothly over the range, e.g. over an image)
// while neighbouring items of a run share its L2.  Items past the last whole group of
// 8 runs keep the dispatcher's order.
__device__ __forceinline__ uint32_t xcd_runs(uint32_t b, uint32_t nb, uint32_t run) {
    constexpr uint32_t X = 8;
    const uint32_t full = nb / (X * run) * (X * run);
    if (b >= full) return b;
    const uint32_t x = b % X, k = b / X;
    return ((k / run) * X + x) * run + k % run;
}

}  // namespace tgs
