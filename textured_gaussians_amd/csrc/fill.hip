// Bulk zero-fill for the texture-gradient buffer ([N, tex_h, tex_w, 4], tens of
// GB at the headline size): one launch, 16-byte non-temporal stores (the
// zeros are not re-read soon, so they should not displace L2 lines of the
// kernels running beside the fill), grid-stride over the buffer.
#include "common.h"

namespace tgs {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void zero_fill_kernel(u32x4 *__restrict__ dst, uint64_t n16, uint32_t *__restrict__ tail,
                                                        uint32_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const u32x4 z = {0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(z, dst + i);
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) tail[t] = 0;
}

}  // namespace tgs

extern "C" int tgs_zero_fill(void *dst, size_t nbytes, tgs_stream_t stream) {
    using namespace tgs;
    if (nbytes == 0) return TGS_OK;
    TGS_CHECK_ARG(dst && ((uintptr_t)dst % 16) == 0 && nbytes % 4 == 0,
                  "zero_fill: dst must be 16-byte aligned and nbytes a multiple of 4");
    const uint64_t n16 = nbytes / 16;
    const uint32_t n4 = (uint32_t)((nbytes % 16) / 4);
    // Two workgroups per CU at most: the fill runs beside the rasterizer on another stream,
    // and a grid larger than what fits at once would hold every CU slot until it drained,
    // starving the other stream.  Each thread strides; stores are fire-and-forget, so a few
    // waves per CU keep the write path busy.
    const uint64_t want = (n16 + 255) / 256;
    const unsigned grid = (unsigned)(want < 512 ? (want > 0 ? want : 1) : 512);
    zero_fill_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((u32x4 *)dst, n16, (uint32_t *)((char *)dst + n16 * 16),
                                                            n4);
    TGS_CHECK_LAUNCH("zero_fill");
    return TGS_OK;
}

namespace tgs {

__global__ __launch_bounds__(256) void zero_bytes_kernel(uint8_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(256) void mark_rows_kernel(int64_t n, const int32_t *__restrict__ flatten_ids, int64_t N,
                                                        uint8_t *__restrict__ mask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[flatten_ids[i] % N] = 1;  // every writer stores the same value
}

// One row per workgroup iteration: rows whose mask byte equals `select` are zeroed
// with 16-byte non-temporal stores, the others skipped.
__global__ __launch_bounds__(256) void zero_rows_kernel(u32x4 *__restrict__ dst, int64_t row16, int64_t N,
                                                        const uint8_t *__restrict__ mask, uint8_t select) {
    const u32x4 z = {0, 0, 0, 0};
    for (int64_t r = blockIdx.x; r < N; r += gridDim.x) {
        if (mask[r] != select) continue;  // uniform across the workgroup
        u32x4 *row = dst + r * row16;
        for (int64_t k = threadIdx.x; k < row16; k += blockDim.x) __builtin_nontemporal_store(z, row + k);
    }
}

}  // namespace tgs

extern "C" int tgs_mark_rows(int64_t n_isects, const int32_t *flatten_ids, int64_t N, uint8_t *mask,
                             tgs_stream_t stream) {
    using namespace tgs;
    TGS_CHECK_ARG(N > 0 && n_isects >= 0 && mask && (n_isects == 0 || flatten_ids), "mark_rows: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    zero_bytes_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, st>>>(mask, N);
    if (n_isects > 0)
        mark_rows_kernel<<<(unsigned)ceil_div(n_isects, 256), 256, 0, st>>>(n_isects, flatten_ids, N, mask);
    TGS_CHECK_LAUNCH("mark_rows");
    return TGS_OK;
}

extern "C" int tgs_zero_fill_rows(void *dst, size_t row_bytes, int64_t N, const uint8_t *mask, int select,
                                  tgs_stream_t stream) {
    using namespace tgs;
    TGS_CHECK_ARG(dst && mask && N > 0 && ((uintptr_t)dst % 16) == 0 && row_bytes % 16 == 0 && row_bytes > 0,
                  "zero_fill_rows: dst must be 16-byte aligned and row_bytes a positive multiple of 16");
    const unsigned grid = (unsigned)(N < 512 ? N : 512);  // two workgroups per CU at most (see tgs_zero_fill)
    zero_rows_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((u32x4 *)dst, (int64_t)(row_bytes / 16), N, mask,
                                                            (uint8_t)(select != 0));
    TGS_CHECK_LAUNCH("zero_fill_rows");
    return TGS_OK;
}
