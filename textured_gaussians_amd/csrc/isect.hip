// Tile intersection: per-Gaussian tile rectangles, prefix sum, key emission,
// a stable LDS-ranked LSD radix sort on the (camera | tile | depth) keys, and
// the per-tile offset table.
//
// Integer semantics (bit-exact contract) follow textured_gaussians/cuda/csrc/
// isect_tiles.cu: tile rectangle :60-70 (float32 arithmetic, clamp of
// floor/ceil to [0, tile_width]), key layout :90-101, stable sort over the low
// 32 + tile_bits + cam_bits bits :262-299, offsets :309-355.
#include "common.h"

namespace tgs {

// ---------------------------------------------------------------- scan
// Three-kernel reduce-then-scan over n items of TIn producing TOut.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

template <typename T> __device__ __forceinline__ T wave_incl_scan(T x, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over the block; returns the block total.
template <typename T> __device__ __forceinline__ T block_excl_scan(T v, T *lds_wave, T &total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    T inc = wave_incl_scan(v, lane);
    if (lane == kWave - 1) lds_wave[w] = inc;
    __syncthreads();
    T before = 0, tot = 0;
    for (int i = 0; i < nw; ++i) {
        const T s = lds_wave[i];
        if (i < w) before += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

template <typename TIn, typename TOut>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const TIn *__restrict__ in, int64_t n,
                                                                    TOut *__restrict__ partial) {
    __shared__ TOut lds[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    TOut s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += (TOut)in[base + i];
    TOut total;
    block_excl_scan<TOut>(s, lds, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

template <typename TIn, typename TOut>
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const TIn *__restrict__ in, int64_t n,
                                                                   TOut *__restrict__ out,
                                                                   const TOut *__restrict__ block_prefix,
                                                                   int inclusive) {
    __shared__ TOut lds[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    TOut v[kScanItems];
    TOut s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < n ? (TOut)in[base + i] : TOut(0);
        s += v[i];
    }
    TOut total;
    TOut run = block_excl_scan<TOut>(s, lds, total) + (block_prefix ? block_prefix[blockIdx.x] : TOut(0));
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = inclusive ? run + v[i] : run;
        run += v[i];
    }
}

template <typename TOut> static size_t scan_temp_bytes(int64_t n) {
    size_t b = 0;
    while (n > kScanTile) {
        n = ceil_div(n, kScanTile);
        b += (size_t)n * sizeof(TOut) * 2 + 256;
    }
    return b + 256;
}

// out[i] = sum_{j <= i} in[j] (inclusive) or sum_{j < i} (exclusive); temp from scan_temp_bytes.
template <typename TIn, typename TOut>
static int device_scan(const TIn *in, TOut *out, int64_t n, bool inclusive, char *temp, size_t temp_bytes,
                       hipStream_t st) {
    if (n == 0) return TGS_OK;
    const int64_t nb = ceil_div(n, kScanTile);
    if (nb == 1) {
        scan_apply_kernel<TIn, TOut><<<1, kScanThreads, 0, st>>>(in, n, out, nullptr, inclusive);
        TGS_CHECK_LAUNCH("scan");
        return TGS_OK;
    }
    const size_t need = (size_t)nb * sizeof(TOut) * 2 + 256;
    TGS_CHECK_ARG(temp_bytes >= need, "scan: temp too small (%zu < %zu)", temp_bytes, need);
    TOut *partial = (TOut *)temp;
    TOut *prefix = partial + nb;
    char *rest = temp + (((size_t)nb * sizeof(TOut) * 2 + 255) / 256) * 256;
    scan_reduce_kernel<TIn, TOut><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, partial);
    TGS_CHECK_LAUNCH("scan_reduce");
    int rc = device_scan<TOut, TOut>(partial, prefix, nb, false, rest, temp_bytes - (rest - temp), st);
    if (rc) return rc;
    scan_apply_kernel<TIn, TOut><<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, out, prefix, inclusive);
    TGS_CHECK_LAUNCH("scan_apply");
    return TGS_OK;
}

// ------------------------------------------------------- tile rectangles
struct Rect {
    int32_t x0, y0, x1, y1;
};

// clamp(floor/ceil(v), 0, hi) as the reference's saturating uint32 cast then min.
__device__ __forceinline__ int32_t sat_tile(float v, int32_t hi) {
    if (!(v > 0.f)) return 0;  // negatives and NaN saturate to 0
    if (v >= (float)hi) return hi;
    return (int32_t)v;
}

__device__ __forceinline__ Rect tile_rect(const float *means2d, int32_t radius, int64_t idx, int32_t ts,
                                          int32_t tw, int32_t th) {
    const float r = (float)radius / (float)ts;
    const float tx = means2d[idx * 2] / (float)ts;
    const float ty = means2d[idx * 2 + 1] / (float)ts;
    Rect R;
    R.x0 = sat_tile(floorf(tx - r), tw);
    R.y0 = sat_tile(floorf(ty - r), th);
    R.x1 = sat_tile(ceilf(tx + r), tw);
    R.y1 = sat_tile(ceilf(ty + r), th);
    return R;
}

__global__ __launch_bounds__(256) void isect_count_kernel(int64_t n, const float *__restrict__ means2d,
                                                          const int32_t *__restrict__ radii, int32_t ts,
                                                          int32_t tw, int32_t th, int32_t *__restrict__ tpg) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int32_t r = radii[idx];
    if (r <= 0) {
        tpg[idx] = 0;
        return;
    }
    const Rect R = tile_rect(means2d, r, idx, ts, tw, th);
    tpg[idx] = (R.y1 - R.y0) * (R.x1 - R.x0);
}

// One wavefront per 64 consecutive Gaussians, emitting their intersections
// together: the wave's output range is contiguous (cumulative counts), each lane
// takes every 64th position of it and finds the Gaussian that owns the position
// by a binary search over the lanes' start offsets (6 shuffles), so the writes
// are coalesced and no lane waits for a large Gaussian's loop.  The keys and
// their positions are those of a per-Gaussian loop over its tile rectangle in
// row-major order.
__global__ __launch_bounds__(256) void isect_fill_kernel(int64_t n, int64_t N, const int64_t *__restrict__ camera_ids,
                                                         const float *__restrict__ means2d,
                                                         const int32_t *__restrict__ radii,
                                                         const float *__restrict__ depths,
                                                         const int64_t *__restrict__ cum, int32_t ts, int32_t tw,
                                                         int32_t th, int32_t tile_bits,
                                                         int64_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);  // the wave's first Gaussian
    if (base >= n) return;                                                           // uniform over the wave
    const int64_t idx = base + lane;
    long long start = LLONG_MAX;  // lanes past n own nothing
    int32_t x0 = 0, y0 = 0, w = 1, dbits = 0;
    long long hi = 0;
    if (idx < n) {
        start = idx == 0 ? 0 : cum[idx - 1];
        const int32_t r = radii[idx];
        if (r > 0) {
            const Rect R = tile_rect(means2d, r, idx, ts, tw, th);
            x0 = R.x0;
            y0 = R.y0;
            w = max(R.x1 - R.x0, 1);
            const int64_t cam = camera_ids ? camera_ids[idx] : idx / N;  // packed entries carry their camera
            hi = (long long)(cam << (32 + tile_bits));
            dbits = __float_as_int(depths[idx]);
        }
    }
    const int64_t last = min(n, base + 64) - 1;
    const long long first = __shfl(start, 0);
    const long long end = cum[last];
    for (long long pos = first + lane; __ballot(pos < end) != 0; pos += 64) {
        // owner: the last lane whose start <= pos (lanes of empty Gaussians share the
        // next lane's start, so the last of them is the one that owns positions)
        int owner = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const long long s = __shfl(start, owner + step);
            if (s <= pos) owner += step;
        }
        const long long so = __shfl(start, owner);
        const int32_t ox0 = __shfl(x0, owner), oy0 = __shfl(y0, owner), ow = __shfl(w, owner);
        const long long ohi = __shfl(hi, owner);
        const int32_t od = __shfl(dbits, owner);
        if (pos < end) {
            const int32_t k = (int32_t)(pos - so);
            const int32_t dy = k / ow, dx = k - dy * ow;
            keys[pos] = ohi | ((int64_t)((oy0 + dy) * tw + ox0 + dx) << 32) | (int64_t)od;  // depth bits sign-extended
            vals[pos] = (int32_t)(base + owner);
        }
    }
}

// ------------------------------------------------------------ radix sort
// LSD radix sort, 8-bit digits.  Per pass: (1) per-block digit histograms,
// written digit-major so that (2) one exclusive scan yields every block's
// scatter base per digit, (3) a stable scatter: each round the 256 threads of
// a block take 256 consecutive keys; a key's rank is its digit's running
// block count + the count in earlier wavefronts of the round (LDS) + the
// count in earlier lanes of its wavefront (wave64 ballot match).
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 keys per block
constexpr int kRadix = 256;

__device__ __forceinline__ uint32_t digit_of(int64_t key, int shift, uint64_t mask) {
    return (uint32_t)((((uint64_t)key) & mask) >> shift) & 0xffu;
}

__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(const int64_t *__restrict__ keys, int64_t n,
                                                                 int shift, uint64_t mask, int64_t nb,
                                                                 uint32_t *__restrict__ counts) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of(keys[i], shift, mask)], 1u);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const int64_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in, int64_t n, int shift,
    uint64_t mask, int64_t nb, const uint32_t *__restrict__ offsets, int64_t *__restrict__ keys_out,
    int32_t *__restrict__ vals_out) {
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t wcount[kSortThreads / kWave][kRadix];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    base[threadIdx.x] = offsets[(int64_t)threadIdx.x * nb + blockIdx.x];
#pragma unroll
    for (int i = 0; i < kSortThreads / kWave; ++i) wcount[i][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const int64_t i = tile0 + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        int64_t key = 0;
        int32_t val = 0;
        uint32_t d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in[i];
            d = digit_of(key, shift, mask);
        }
        // lanes of this wavefront holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = __popcll(peers & lt);
        const bool leader = valid && (peers & lt) == 0;
        if (leader) wcount[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (int ww = 0; ww < w; ++ww) pos += wcount[ww][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        {
            const int dd = threadIdx.x;
            uint32_t add = 0;
#pragma unroll
            for (int ww = 0; ww < kSortThreads / kWave; ++ww) {
                add += wcount[ww][dd];
                wcount[ww][dd] = 0;
            }
            base[dd] += add;
        }
        __syncthreads();
    }
}

static size_t sort_counts_bytes(int64_t n) {
    const int64_t nb = ceil_div(n, kSortTile);
    return (((size_t)nb * kRadix * sizeof(uint32_t) + 255) / 256) * 256;
}

// ------------------------------------------------------------ offsets
// offsets[j] = lower_bound over the sorted (camera, tile) ids of value j:
// the first intersection of tile j or of the next occupied tile.
__global__ __launch_bounds__(256) void offset_encode_kernel(int64_t n_isects, const int64_t *__restrict__ keys,
                                                            int64_t n_bins, int32_t n_tiles, int32_t tile_bits,
                                                            int32_t *__restrict__ offsets) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_bins) return;
    int64_t lo = 0, hi = n_isects;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int64_t k = keys[mid] >> 32;
        const int64_t id = (k >> tile_bits) * n_tiles + (k & ((int64_t(1) << tile_bits) - 1));
        if (id < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    offsets[j] = (int32_t)lo;
}

static int bit_length(int64_t x) {
    int b = 0;
    while (x > 0) {
        ++b;
        x >>= 1;
    }
    return b;
}

}  // namespace tgs

using namespace tgs;

extern "C" size_t tgs_isect_count_temp_bytes(int64_t n_elems) { return scan_temp_bytes<int64_t>(n_elems); }

extern "C" int tgs_isect_tiles_count(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                                     int32_t tile_size, int32_t tile_width, int32_t tile_height,
                                     int32_t *tiles_per_gauss, int64_t *cum_tiles_per_gauss, void *temp,
                                     size_t temp_bytes, tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "isect_tiles: bad sizes");
    TGS_CHECK_ARG(tile_size > 0 && tile_width > 0 && tile_height > 0, "isect_tiles: bad tile grid");
    TGS_CHECK_ARG(bit_length((int64_t)tile_width * tile_height) + bit_length(C) <= 32,
                  "isect_tiles: camera and tile ids need more than 32 bits");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = C * N;
    if (n == 0) return TGS_OK;
    isect_count_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(n, means2d, radii, tile_size, tile_width,
                                                                    tile_height, tiles_per_gauss);
    TGS_CHECK_LAUNCH("isect_count");
    return device_scan<int32_t, int64_t>(tiles_per_gauss, cum_tiles_per_gauss, n, true, (char *)temp, temp_bytes,
                                         st);
}

extern "C" int tgs_prefix_sum_i32(int64_t n, const int32_t *counts, int64_t *inclusive, void *temp,
                                  size_t temp_bytes, tgs_stream_t stream) {
    TGS_CHECK_ARG(n >= 0 && (n == 0 || (counts && inclusive)), "prefix_sum: bad arguments");
    if (n == 0) return TGS_OK;
    return device_scan<int32_t, int64_t>(counts, inclusive, n, true, (char *)temp, temp_bytes, (hipStream_t)stream);
}

extern "C" int tgs_isect_tiles_fill(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                                    const float *depths, const int64_t *cum_tiles_per_gauss, int32_t tile_size,
                                    int32_t tile_width, int32_t tile_height, int64_t *isect_ids,
                                    int32_t *flatten_ids, tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && N >= 0 && C * N < (int64_t(1) << 31), "isect_tiles: bad sizes");
    const int64_t n = C * N;
    if (n == 0) return TGS_OK;
    const int tb = bit_length((int64_t)tile_width * tile_height);
    isect_fill_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(
        n, N, nullptr, means2d, radii, depths, cum_tiles_per_gauss, tile_size, tile_width, tile_height, tb,
        isect_ids, flatten_ids);
    TGS_CHECK_LAUNCH("isect_fill");
    return TGS_OK;
}

extern "C" int tgs_isect_tiles_fill_packed(int64_t C, int64_t nnz, const int64_t *camera_ids, const float *means2d,
                                           const int32_t *radii, const float *depths,
                                           const int64_t *cum_tiles_per_gauss, int32_t tile_size, int32_t tile_width,
                                           int32_t tile_height, int64_t *isect_ids, int32_t *flatten_ids,
                                           tgs_stream_t stream) {
    TGS_CHECK_ARG(C >= 0 && nnz >= 0 && nnz < (int64_t(1) << 31), "isect_tiles: bad sizes");
    TGS_CHECK_ARG(bit_length((int64_t)tile_width * tile_height) + bit_length(C) <= 32,
                  "isect_tiles: camera and tile ids need more than 32 bits");
    TGS_CHECK_ARG(nnz == 0 || camera_ids, "isect_tiles: packed entries need camera_ids");
    if (nnz == 0) return TGS_OK;
    const int tb = bit_length((int64_t)tile_width * tile_height);
    isect_fill_kernel<<<(unsigned)ceil_div(nnz, 256), 256, 0, (hipStream_t)stream>>>(
        nnz, 1, camera_ids, means2d, radii, depths, cum_tiles_per_gauss, tile_size, tile_width, tile_height, tb,
        isect_ids, flatten_ids);
    TGS_CHECK_LAUNCH("isect_fill");
    return TGS_OK;
}

extern "C" size_t tgs_sort_pairs_temp_bytes(int64_t n) {
    const int64_t nb = ceil_div(n > 0 ? n : 1, kSortTile);
    return 2 * sort_counts_bytes(n > 0 ? n : 1) + scan_temp_bytes<uint32_t>(nb * kRadix);
}

extern "C" int tgs_sort_pairs(int64_t n, int32_t end_bit, int64_t *keys, int32_t *values, int64_t *keys_alt,
                              int32_t *values_alt, void *temp, size_t temp_bytes, int32_t *selector_host,
                              tgs_stream_t stream) {
    TGS_CHECK_ARG(n >= 0 && n < (int64_t(1) << 32), "sort_pairs: bad n");
    TGS_CHECK_ARG(end_bit > 0 && end_bit <= 64, "sort_pairs: bad end_bit %d", end_bit);
    TGS_CHECK_ARG(temp_bytes >= tgs_sort_pairs_temp_bytes(n), "sort_pairs: temp too small");
    *selector_host = 0;
    if (n == 0) return TGS_OK;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nb = ceil_div(n, kSortTile);
    const size_t cb = sort_counts_bytes(n);
    uint32_t *counts = (uint32_t *)temp;
    uint32_t *offs = (uint32_t *)((char *)temp + cb);
    char *scan_tmp = (char *)temp + 2 * cb;
    const size_t scan_bytes = temp_bytes - 2 * cb;
    const uint64_t mask = end_bit >= 64 ? ~0ull : ((1ull << end_bit) - 1);
    const int passes = (end_bit + 7) / 8;
    int64_t *kin = keys, *kout = keys_alt;
    int32_t *vin = values, *vout = values_alt;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        sort_hist_kernel<<<(unsigned)nb, kSortThreads, 0, st>>>(kin, n, shift, mask, nb, counts);
        TGS_CHECK_LAUNCH("sort_hist");
        int rc = device_scan<uint32_t, uint32_t>(counts, offs, nb * kRadix, false, scan_tmp, scan_bytes, st);
        if (rc) return rc;
        sort_scatter_kernel<<<(unsigned)nb, kSortThreads, 0, st>>>(kin, vin, n, shift, mask, nb, offs, kout, vout);
        TGS_CHECK_LAUNCH("sort_scatter");
        int64_t *tk = kin; kin = kout; kout = tk;
        int32_t *tv = vin; vin = vout; vout = tv;
    }
    *selector_host = passes & 1;
    return TGS_OK;
}

extern "C" int tgs_isect_offset_encode(int64_t n_isects, const int64_t *isect_ids, int64_t C, int32_t tile_width,
                                       int32_t tile_height, int32_t *offsets, tgs_stream_t stream) {
    TGS_CHECK_ARG(n_isects >= 0 && C >= 0 && tile_width > 0 && tile_height > 0, "isect_offset_encode: bad sizes");
    const int64_t n_tiles = (int64_t)tile_width * tile_height;
    const int64_t bins = C * n_tiles;
    if (bins == 0) return TGS_OK;
    offset_encode_kernel<<<(unsigned)ceil_div(bins, 256), 256, 0, (hipStream_t)stream>>>(
        n_isects, isect_ids, bins, (int32_t)n_tiles, bit_length(n_tiles), offsets);
    TGS_CHECK_LAUNCH("isect_offset_encode");
    return TGS_OK;
}
