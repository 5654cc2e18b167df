// Device-wide exclusive prefix sum of uint32 counts (two launches: block
// reduce, then scan with the block offsets).  Used to turn per-bucket and
// per-particle counts into offsets.
#pragma once

#include "lm_common.h"

namespace lm {

constexpr int kScanItems = 8;                       // items per thread
constexpr int kScanTile = kBlock * kScanItems;      // 2048 items per workgroup

inline int64_t scan_num_blocks(int64_t n) {
    int64_t b = (n + kScanTile - 1) / kScanTile;
    return b < 1 ? 1 : b;
}

// out[0..n) = exclusive prefix of in[0..n), out[n] = total.  If total64 is not
// NULL the total is also stored there as int64.  block_sums must hold
// scan_num_blocks(n) entries.  in and out may not alias.
int exclusive_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* block_sums,
                       int64_t* total64, hipStream_t stream);

// nseg independent scans: segment g reads in + g*in_stride and writes
// out + g*out_stride (n + 1 entries); totals (if not NULL) to total64[g].
// block_sums must hold nseg * scan_num_blocks(n) entries; strides are
// multiples of 4 and out_stride >= n + 1.
int exclusive_scan_u32_batched(const uint32_t* in, int64_t n, int64_t in_stride, uint32_t* out, int64_t out_stride,
                               int32_t nseg, uint32_t* block_sums, int64_t* total64, hipStream_t stream);

// Row stride of per-particle arrays that are scanned (n + 1 entries, 16-byte rows).
inline int64_t scan_row_stride(int64_t n) { return (n + 4) & ~(int64_t)3; }

// ---- wave/block building blocks (wave64) ---------------------------------
__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Block-wide sum for blockDim.x == kBlock; every thread gets the result.
__device__ inline uint32_t block_sum(uint32_t x, uint32_t* lds /* >= 4 */) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    x = wave_sum(x);
    __syncthreads();
    if (lane == 0) lds[wave] = x;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    return t;
}

// Block-wide exclusive scan for blockDim.x == kBlock; *total gets the sum.
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* lds /* >= 4 */,
                                                uint32_t* total) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t inc = wave_inclusive_scan(x);
    __syncthreads();
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        uint32_t v = lds[w];
        before += (w < wave) ? v : 0u;
        all += v;
    }
    *total = all;
    return before + inc - x;
}

}  // namespace lm
