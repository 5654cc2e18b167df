// Segmented stable LSD radix sort (see lm_sort.h).
//
// Per pass over a digit of up to 10 bits (R = 2^bits buckets), with blocks of
// kSortGroup consecutive tiles of 2048 elements:
//   hist    : per (segment, tile) the digit counts (LDS atomics: counting
//             needs no order, and the keys here give few conflicts), stored
//             digit-major, hist[g][d * ntp + tile] (ntp = tiles rounded up to
//             the group), one 16-byte store per digit and block
//   scan    : exclusive scan of each segment's histogram -> the output offset
//             of (digit, tile); digit-major order puts every tile's run of a
//             digit after the earlier tiles' runs of that digit
//   scatter : the block walks its tiles in order.  Each wavefront ranks its
//             contiguous quarter of the tile, 64 elements at a time: the lanes
//             holding the same digit are found with one ballot per digit bit,
//             a lane's rank is the number of such lanes below it, and a
//             per-wavefront running count per digit (LDS, touched only by that
//             wavefront, so in program order) carries the rank across rounds.
//             One scan over the digits then gives each element its place in the
//             tile sorted by digit; the tile is staged there in LDS and written
//             out in that order, so consecutive lanes store to consecutive
//             addresses of a run.  The next tile's keys load meanwhile.
// Equal digits keep their input order, which makes every pass stable.  A
// caller that needs no order inside a digit (a partition whose buckets are
// placed in any order later) asks for an unstable pass: the rank is then an
// LDS atomic's return value, with no ballots.
#include "lm_sort.h"
#include "lm_scan.h"

namespace lm {

constexpr int kSortWaves = kSortThreads / 64;
constexpr int kWaveChunk = kSortTile / kSortWaves;  // 512 contiguous elements per wavefront

__device__ __forceinline__ int64_t seg_len_of(const int64_t* seg_len, int64_t uniform_len, int64_t max_len, int g) {
    int64_t L = seg_len ? seg_len[g] : uniform_len;
    if (L < 0) L = 0;
    return L > max_len ? max_len : L;
}

// Lanes of the wavefront whose digit equals this lane's (valid lanes only).
template <int BITS>
__device__ __forceinline__ unsigned long long match_digit(bool valid, uint32_t d) {
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bm = __ballot(valid && bit);
        m &= bit ? bm : ~bm;
    }
    return m;
}

template <int BITS>
__global__ void __launch_bounds__(kSortThreads)
radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t stride, const int64_t* __restrict__ seg_len,
                  int64_t uniform_len, int64_t max_len, int shift, uint32_t dmask, uint32_t* __restrict__ hist,
                  int64_t row, int64_t ntp) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    constexpr int R = 1 << BITS;
    static_assert(kSortGroup == 4, "one uint4 of counts per digit");
    __shared__ uint32_t h[kSortGroup][R];
    const int g = by;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t len = seg_len_of(seg_len, uniform_len, max_len, g);
    for (int q = t; q < kSortGroup * R; q += kSortThreads) (&h[0][0])[q] = 0;
    __syncthreads();
    const uint32_t* k = keys + (int64_t)g * stride;
    const int64_t first = (int64_t)bx * kSortGroup * kSortTile + wave * kWaveChunk;
    uint32_t kk[kSortGroup][kSortItems];
#pragma unroll
    for (int j = 0; j < kSortGroup; ++j)
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t e = first + (int64_t)j * kSortTile + r * 64 + lane;
            kk[j][r] = e < len ? k[e] : 0u;
        }
#pragma unroll
    for (int j = 0; j < kSortGroup; ++j)
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const bool valid = first + (int64_t)j * kSortTile + r * 64 + lane < len;
            if (valid) atomicAdd(&h[j][(kk[j][r] >> shift) & dmask], 1u);
        }
    __syncthreads();
    for (int q = t; q < R; q += kSortThreads)
        *reinterpret_cast<uint4*>(hist + (int64_t)g * row + (int64_t)q * ntp + (int64_t)bx * kSortGroup) =
            make_uint4(h[0][q], h[1][q], h[2][q], h[3][q]);
}

template <int BITS, bool STABLE>
__global__ void __launch_bounds__(kSortThreads)
radix_scatter_kernel(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                     uint32_t* __restrict__ vout, int64_t stride, const int64_t* __restrict__ seg_len,
                     int64_t uniform_len, int64_t max_len, int shift, uint32_t dmask,
                     const uint32_t* __restrict__ offs, int64_t row, int64_t ntp) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    constexpr int R = 1 << BITS;
    constexpr int PER = R / kSortThreads;  // digits per thread in the digit scan
    static_assert(PER >= 1 && R % kSortThreads == 0, "digit count must be a multiple of the block");
    __shared__ uint32_t wcnt[kSortWaves][R];  // running counts, then per-wavefront offsets
    __shared__ uint32_t lstart[R];            // tile-local start of each digit
    __shared__ uint32_t gdelta[R];            // global start - local start (mod 2^32)
    __shared__ uint32_t skey[kSortTile], sval[kSortTile];
    __shared__ uint32_t scan_lds[kSortWaves];
    const int g = by;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t len = seg_len_of(seg_len, uniform_len, max_len, g);
    const int64_t group0 = (int64_t)bx * kSortGroup * kSortTile;
    if (group0 >= len) return;  // uniform over the block
    const uint32_t* ki = kin + (int64_t)g * stride;
    // vin == NULL: each element's value is its index in the segment
    const uint32_t* vi = vin ? vin + (int64_t)g * stride : nullptr;
    uint32_t* ko = kout + (int64_t)g * stride;
    uint32_t* vo = vout + (int64_t)g * stride;
    uint4 gb[PER];  // output offsets of my digits in the group's tiles
#pragma unroll
    for (int q = 0; q < PER; ++q)
        gb[q] = *reinterpret_cast<const uint4*>(offs + (int64_t)g * row + (int64_t)(t * PER + q) * ntp +
                                                (int64_t)bx * kSortGroup);
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t key[kSortItems], val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t e = group0 + wave * kWaveChunk + r * 64 + lane;
        key[r] = e < len ? ki[e] : 0u;
        val[r] = e < len ? (vi ? vi[e] : (uint32_t)e) : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSortGroup; ++j) {
        const int64_t tile0 = group0 + (int64_t)j * kSortTile;
        if (tile0 >= len) break;  // uniform over the block
        const int64_t base = tile0 + wave * kWaveChunk;
        for (int q = t; q < R; q += kSortThreads)
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) wcnt[w][q] = 0;
        __syncthreads();
        // rank within the wavefront's chunk (this wavefront alone uses wcnt[wave])
        uint32_t mine[kSortItems];
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const bool valid = base + r * 64 + lane < len;
            const uint32_t d = (key[r] >> shift) & dmask;
            if constexpr (STABLE) {
                const unsigned long long m = match_digit<BITS>(valid, d);
                const uint32_t rank = (uint32_t)__popcll(m & below);
                const uint32_t run = valid ? wcnt[wave][d] : 0u;
                mine[r] = run + rank;
                if (valid && rank == 0) wcnt[wave][d] = run + (uint32_t)__popcll(m);
            } else {
                mine[r] = valid ? atomicAdd(&wcnt[wave][d], 1u) : 0u;
            }
        }
        // next tile's elements, in flight while this one is placed
        uint32_t nkey[kSortItems], nval[kSortItems];
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            const int64_t e = base + kSortTile + r * 64 + lane;
            const bool ok = j + 1 < kSortGroup && e < len;
            nkey[r] = ok ? ki[e] : 0u;
            nval[r] = ok ? (vi ? vi[e] : (uint32_t)e) : 0u;
        }
        __syncthreads();
        // per digit: offsets of the wavefronts' runs, tile-local start, global start
        uint32_t tot[PER], sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int d = t * PER + q;
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const uint32_t c = wcnt[w][d];
                wcnt[w][d] = acc;
                acc += c;
            }
            tot[q] = acc;
            sum += acc;
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan(sum, scan_lds, &all);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int d = t * PER + q;
            const uint32_t gbase = j == 0 ? gb[q].x : j == 1 ? gb[q].y : j == 2 ? gb[q].z : gb[q].w;
            lstart[d] = run;
            gdelta[d] = gbase - run;
            run += tot[q];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            if (base + r * 64 + lane < len) {
                const uint32_t d = (key[r] >> shift) & dmask;
                const uint32_t lpos = lstart[d] + wcnt[wave][d] + mine[r];
                if (lpos < (uint32_t)kSortTile) {
                    skey[lpos] = key[r];
                    sval[lpos] = val[r];
                }
            }
        }
        __syncthreads();
        const int64_t left = len - tile0;
        const uint32_t cnt = left < kSortTile ? (uint32_t)left : (uint32_t)kSortTile;
#pragma unroll
        for (int q = 0; q < kSortItems; ++q) {
            const uint32_t slot = q * kSortThreads + t;
            if (slot < cnt) {
                const uint32_t k = skey[slot];
                const uint32_t pos = gdelta[(k >> shift) & dmask] + slot;
                if (pos < (uint64_t)len) {
                    ko[pos] = k;
                    vo[pos] = sval[slot];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kSortItems; ++r) {
            key[r] = nkey[r];
            val[r] = nval[r];
        }
        // the next tile's first barrier orders these LDS reads before its writes
    }
}

__global__ void __launch_bounds__(kSortThreads)
seg_copy_kernel(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ vs, uint32_t* __restrict__ kd,
                uint32_t* __restrict__ vd, int64_t stride, const int64_t* __restrict__ seg_len, int64_t uniform_len,
                int64_t max_len) {
    const int g = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * kSortThreads + threadIdx.x;
    if (e < seg_len_of(seg_len, uniform_len, max_len, g)) {
        kd[(int64_t)g * stride + e] = ks[(int64_t)g * stride + e];
        vd[(int64_t)g * stride + e] = vs[(int64_t)g * stride + e];
    }
}

int sort_scratch_alloc(SortScratch* s, int64_t max_len, int32_t nseg_max) {
    const int64_t ntp = sort_padded_tiles(max_len);
    const int64_t n = (int64_t)(1 << kMaxDigitBits) * (ntp < kSortGroup ? kSortGroup : ntp);
    s->row = scan_row_stride(n);
    s->max_len = max_len;
    s->nseg_max = nseg_max;
    const size_t hb = sizeof(uint32_t) * (size_t)(s->row * nseg_max);
    const size_t bb = sizeof(uint32_t) * (size_t)(scan_num_blocks(n) * nseg_max);
    if (hipMalloc((void**)&s->hist, hb) != hipSuccess || hipMalloc((void**)&s->offs, hb) != hipSuccess ||
        hipMalloc((void**)&s->bsums, bb) != hipSuccess) {
        set_error("hipMalloc of sort scratch failed");
        sort_scratch_free(s);
        return LM_ERR_NOMEM;
    }
    s->bytes = (int64_t)(2 * hb + bb);
    return LM_OK;
}

void sort_scratch_free(SortScratch* s) {
    (void)hipFree(s->hist);
    (void)hipFree(s->offs);
    (void)hipFree(s->bsums);
    s->hist = s->offs = s->bsums = nullptr;
    s->bytes = 0;
}

template <int BITS>
static void launch_scatter(bool stable, dim3 grid, hipStream_t stream, const uint32_t* kin, const uint32_t* vin,
                           uint32_t* kout, uint32_t* vout, int64_t stride, const int64_t* seg_len,
                           int64_t uniform_len, int64_t max_len, int shift, uint32_t dmask, const uint32_t* offs,
                           int64_t row, int64_t ntp) {
    if (stable)
        LM_LAUNCH(K_SORT_SCATTER, (radix_scatter_kernel<BITS, true>), grid, dim3(kSortThreads), 0, stream, kin, vin,
                  kout, vout, stride, seg_len, uniform_len, max_len, shift, dmask, offs, row, ntp);
    else
        LM_LAUNCH(K_SORT_SCATTER, (radix_scatter_kernel<BITS, false>), grid, dim3(kSortThreads), 0, stream, kin, vin,
                  kout, vout, stride, seg_len, uniform_len, max_len, shift, dmask, offs, row, ntp);
}

// One pass over the digit (key >> shift) & dmask (at most kMaxDigitBits
// wide) from keys/vals[cur] into keys/vals[1 - cur]; stable unless asked
// otherwise.
static int radix_pass(uint32_t* keys[2], uint32_t* vals[2], int cur, int64_t stride, int32_t nseg,
                      const int64_t* seg_len, int64_t uniform_len, int64_t max_len, int shift, int bits,
                      SortScratch* sc, hipStream_t stream, bool index_vals = false, bool stable = true) {
    const uint32_t* vin = index_vals ? nullptr : (const uint32_t*)vals[cur];
    const int64_t ntp = sort_padded_tiles(max_len);
    const dim3 grid((unsigned)(ntp / kSortGroup), (unsigned)nseg);
    const uint32_t dmask = (uint32_t)((1ull << bits) - 1ull);
    // the narrowest kernel that holds the digit: 8, 9 or kMaxDigitBits bits
    const int tbits = bits <= 8 ? 8 : bits == 9 ? 9 : kMaxDigitBits;
    const int64_t n = (int64_t)(1 << tbits) * ntp;
    const uint32_t* kin = (const uint32_t*)keys[cur];
    if (tbits == 8)
        LM_LAUNCH(K_SORT_HIST, radix_hist_kernel<8>, grid, dim3(kSortThreads), 0, stream, kin, stride, seg_len,
                  uniform_len, max_len, shift, dmask, sc->hist, sc->row, ntp);
    else if (tbits == 9)
        LM_LAUNCH(K_SORT_HIST, radix_hist_kernel<9>, grid, dim3(kSortThreads), 0, stream, kin, stride, seg_len,
                  uniform_len, max_len, shift, dmask, sc->hist, sc->row, ntp);
    else
        LM_LAUNCH(K_SORT_HIST, radix_hist_kernel<kMaxDigitBits>, grid, dim3(kSortThreads), 0, stream, kin, stride,
                  seg_len, uniform_len, max_len, shift, dmask, sc->hist, sc->row, ntp);
    LM_LAUNCH_CHECK();
    int rc = exclusive_scan_u32_batched(sc->hist, n, sc->row, sc->offs, sc->row, nseg, sc->bsums, nullptr, stream);
    if (rc) return rc;
    const uint32_t* offs = (const uint32_t*)sc->offs;
    if (tbits == 8)
        launch_scatter<8>(stable, grid, stream, kin, vin, keys[1 - cur], vals[1 - cur], stride, seg_len, uniform_len,
                          max_len, shift, dmask, offs, sc->row, ntp);
    else if (tbits == 9)
        launch_scatter<9>(stable, grid, stream, kin, vin, keys[1 - cur], vals[1 - cur], stride, seg_len, uniform_len,
                          max_len, shift, dmask, offs, sc->row, ntp);
    else
        launch_scatter<kMaxDigitBits>(stable, grid, stream, kin, vin, keys[1 - cur], vals[1 - cur], stride, seg_len,
                                      uniform_len, max_len, shift, dmask, offs, sc->row, ntp);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int radix_sort_u32(uint32_t* keys[2], uint32_t* vals[2], int64_t stride, int32_t nseg, const int64_t* seg_len,
                   int64_t uniform_len, int64_t max_len, int key_bits, SortScratch* sc, hipStream_t stream,
                   int* result) {
    LM_REQUIRE(nseg >= 1 && nseg <= sc->nseg_max && max_len <= sc->max_len && stride >= max_len, LM_ERR_ARG,
               "sort request exceeds its scratch");
    LM_REQUIRE(key_bits >= 0 && key_bits <= 32, LM_ERR_ARG, "key_bits out of range");
    int cur = 0;
    if (sort_padded_tiles(max_len) == 0 || key_bits == 0) {
        *result = 0;
        return LM_OK;
    }
    // fewest passes of at most kMaxDigitBits, the narrow kernel when a pass fits 8 bits
    const int passes = (key_bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int per = (key_bits + passes - 1) / passes;
    for (int shift = 0; shift < key_bits; shift += per) {
        const int b = key_bits - shift < per ? key_bits - shift : per;
        int rc = radix_pass(keys, vals, cur, stride, nseg, seg_len, uniform_len, max_len, shift, b, sc, stream);
        if (rc) return rc;
        cur = 1 - cur;
    }
    *result = cur;
    return LM_OK;
}

int radix_partition_u32(uint32_t* keys[2], uint32_t* vals[2], int64_t stride, int32_t nseg, int64_t len,
                        int shift, int bits, SortScratch* sc, hipStream_t stream, const uint32_t** digit_start,
                        int64_t* digit_stride, bool index_vals, bool stable) {
    LM_REQUIRE(nseg >= 1 && nseg <= sc->nseg_max && len <= sc->max_len && stride >= len, LM_ERR_ARG,
               "partition request exceeds its scratch");
    LM_REQUIRE(bits >= 1 && bits <= kMaxDigitBits && shift >= 0 && shift + bits <= 32, LM_ERR_ARG,
               "partition digit out of range");
    LM_REQUIRE(len >= 1, LM_ERR_ARG, "empty partition");
    int rc = radix_pass(keys, vals, 0, stride, nseg, nullptr, len, len, shift, bits, sc, stream, index_vals, stable);
    if (rc) return rc;
    *digit_start = sc->offs;
    *digit_stride = sort_padded_tiles(len);
    return LM_OK;
}

}  // namespace lm

extern "C" {

int lm_sort_pairs_u32(int32_t nseg, const int64_t* seg_len, int64_t uniform_len, int64_t max_len, int64_t stride,
                      int32_t key_bits, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      void* stream) {
    LM_REQUIRE(nseg >= 1 && nseg <= 65535, LM_ERR_ARG, "1..65535 segments");
    LM_REQUIRE(max_len >= 0 && stride >= max_len, LM_ERR_ARG, "stride < max_len");
    LM_REQUIRE(keys && vals && keys_tmp && vals_tmp, LM_ERR_ARG, "NULL buffer");
    lm::SortScratch sc;
    int rc = lm::sort_scratch_alloc(&sc, max_len, nseg);
    if (rc) return rc;
    uint32_t* k[2] = {keys, keys_tmp};
    uint32_t* v[2] = {vals, vals_tmp};
    int res = 0;
    hipStream_t s = (hipStream_t)stream;
    rc = lm::radix_sort_u32(k, v, stride, nseg, seg_len, uniform_len, max_len, key_bits, &sc, s, &res);
    if (rc == LM_OK && res == 1) {  // bring the result back into keys/vals
        const dim3 grid((unsigned)((max_len + lm::kSortThreads - 1) / lm::kSortThreads), (unsigned)nseg);
        if (grid.x > 0) {
            LM_LAUNCH(lm::K_COPY, lm::seg_copy_kernel, grid, dim3(lm::kSortThreads), 0, s, (const uint32_t*)keys_tmp,
                      (const uint32_t*)vals_tmp, keys, vals, stride, seg_len, uniform_len, max_len);
            if (hipGetLastError() != hipSuccess) {
                lm::set_error("result copy failed");
                rc = LM_ERR_HIP;
            }
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess && rc == LM_OK) {
        lm::set_error("sort failed: %s", hipGetErrorString(hipGetLastError()));
        rc = LM_ERR_HIP;
    }
    lm::sort_scratch_free(&sc);
    return rc;
}

}  // extern "C"
