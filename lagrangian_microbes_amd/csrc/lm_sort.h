// Stable LSD radix sort of (uint32 key, uint32 value) pairs, segmented: nseg
// independent segments of a common layout (segment g at base + g*stride), each
// with its own length (device array, so no host round trip).  No global
// atomics: digit counts are LDS histograms, the in-tile rank of an element is
// computed from wavefront ballots, and tile offsets come from the device-wide
// scan.
#pragma once

#include "lm_common.h"

namespace lm {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 elements per tile
constexpr int kMaxDigitBits = 10;  // digit width of a pass (2 passes sort 20-bit keys)
constexpr int kSortGroup = 4;      // consecutive tiles per block

// Tiles of max_len elements, rounded up to whole groups.
inline int64_t sort_padded_tiles(int64_t max_len) {
    const int64_t t = (max_len + kSortTile - 1) / kSortTile;
    return (t + kSortGroup - 1) / kSortGroup * kSortGroup;
}

struct SortScratch {
    uint32_t* hist = nullptr;     // [nseg][row]  digit-major tile counts
    uint32_t* offs = nullptr;     // [nseg][row]  their exclusive scan
    uint32_t* bsums = nullptr;    // scan block sums
    int64_t row = 0;              // per-segment stride of hist / offs
    int64_t max_len = 0;
    int32_t nseg_max = 0;
    int64_t bytes = 0;
};

int sort_scratch_alloc(SortScratch* s, int64_t max_len, int32_t nseg_max);
void sort_scratch_free(SortScratch* s);

// Sort keys[0]/vals[0] (bits [0, key_bits) of the keys) into either buffer
// pair; *result = 0 or 1 tells which of keys[]/vals[] holds the result.
// Segment g has min(seg_len[g], max_len) elements if seg_len is not NULL,
// else uniform_len.  stride >= max_len.
int radix_sort_u32(uint32_t* keys[2], uint32_t* vals[2], int64_t stride, int32_t nseg, const int64_t* seg_len,
                   int64_t uniform_len, int64_t max_len, int key_bits, SortScratch* scratch, hipStream_t stream,
                   int* result);

// One pass over the digit (key >> shift) & (2^bits - 1) of nseg segments of
// len elements each (bits <= kMaxDigitBits), from keys/vals[0] into
// keys/vals[1]: the elements grouped by digit, in their input order inside a
// digit when stable, else in any order.  Digit d of segment g starts at
// digit_start[g * scratch->row + d * digit_stride] (the offsets the pass
// scattered with, valid until the scratch is used again).  With index_vals,
// vals[0] is not read: each element's value is its index in the segment.
int radix_partition_u32(uint32_t* keys[2], uint32_t* vals[2], int64_t stride, int32_t nseg, int64_t len,
                        int shift, int bits, SortScratch* scratch, hipStream_t stream, const uint32_t** digit_start,
                        int64_t* digit_stride, bool index_vals = false, bool stable = true);

}  // namespace lm
