// On-device reductions for long runs: the species count of every recorded
// step (analysis.py:31-35 counts species == ROCK / PAPER / SCISSORS per
// time) and the histogram of all pairwise great-circle distances
// (sandbox/pairwise_distance_histogram_gpu.jl:14-43,
// sandbox/pairwise_distance_histogram_distributed.jl:49-55).
//
// Distance histogram: for every unordered pair i < j, the haversine term
//   a = sin^2(dlat / 2) + cos(lat_i) cos(lat_j) sin^2(dlon / 2)
// (degrees, float32 positions widened to double; sin and cos of degree
// arguments as the fma-Horner polynomials every liblm unit shares, so the
// oracle reproduces a bit for bit) is binned against increasing thresholds:
// bin = number of thresholds <= a.  The caller derives the thresholds from
// distance edges (d = 2 R asin(sqrt(a))), so the bins are distance bins and
// the counts are exact integers.  Work: a triangular grid of 256 x 256 pair
// tiles, the j tile's positions and cosines staged in LDS, one LDS histogram
// per wavefront, one global atomic per non-empty bin per block.
#include "lm_common.h"

namespace lm {

constexpr int kHistTile = 256;
constexpr int kHistMaxBins = 256;

__global__ void __launch_bounds__(kBlock)
species_count_kernel(const int8_t* __restrict__ rec, int64_t n, int64_t stride,
                     unsigned long long* __restrict__ counts) {
    __shared__ uint32_t part[kBlock / 64][4];
    const int64_t s = blockIdx.y;
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int v = rec[s * stride + i];
        c1 += v == 1;
        c2 += v == 2;
        c3 += v == 3;
        c0 += v < 1 || v > 3;
    }
    uint32_t c[4] = {c0, c1, c2, c3};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t x = c[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) part[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += part[w][threadIdx.x];
        if (t) atomicAdd(counts + s * 4 + threadIdx.x, (unsigned long long)t);
    }
}

__global__ void __launch_bounds__(kHistTile)
dist_hist_kernel(int64_t n, const float* __restrict__ lon, const float* __restrict__ lat, int nbins,
                 const double* __restrict__ edges, unsigned long long* __restrict__ hist) {
    const int64_t ti = blockIdx.x, tj = blockIdx.y;
    if (tj < ti) return;  // the lower triangle of the tile grid has no pairs i < j
    __shared__ double sx[kHistTile], sy[kHistTile], sc[kHistTile];
    __shared__ double se[kHistMaxBins];
    __shared__ uint32_t h[kHistTile / 64][kHistMaxBins];
    const int t = threadIdx.x, wave = t >> 6;
    const int64_t j0 = tj * kHistTile, i = ti * kHistTile + t;
    {
        const int64_t j = j0 + t;
        const double x = j < n ? (double)lon[j] : 0.0, y = j < n ? (double)lat[j] : 0.0;
        sx[t] = x;
        sy[t] = y;
        sc[t] = cos_deg(y);
    }
    for (int b = t; b < nbins - 1; b += kHistTile) se[b] = edges[b];
    for (int b = t; b < (kHistTile / 64) * kHistMaxBins; b += kHistTile) (&h[0][0])[b] = 0u;
    __syncthreads();
    if (i < n) {
        const double xi = (double)lon[i], yi = (double)lat[i], ci = cos_deg(yi);
        const int64_t jend = n - j0 < kHistTile ? n - j0 : kHistTile;
        // pairs (i, j) with j > i: the whole tile when tj > ti, past the diagonal when tj == ti
        const int k0 = tj == ti ? t + 1 : 0;
        for (int k = k0; k < (int)jend; ++k) {
            const double s1 = sin_deg(0.5 * (sy[k] - yi));
            const double s2 = sin_deg(0.5 * (sx[k] - xi));
            const double a = s1 * s1 + (ci * sc[k]) * (s2 * s2);
            // bin = number of thresholds <= a (the thresholds increase)
            int lo = 0, hi = nbins - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (se[mid] <= a) lo = mid + 1;
                else hi = mid;
            }
            atomicAdd(&h[wave][lo], 1u);
        }
    }
    __syncthreads();
    for (int b = t; b < nbins; b += kHistTile) {
        uint32_t v = 0;
#pragma unroll
        for (int w = 0; w < kHistTile / 64; ++w) v += h[w][b];
        if (v) atomicAdd(hist + b, (unsigned long long)v);
    }
}

}  // namespace lm

extern "C" {

int lm_species_counts(const int8_t* species, int64_t n, int32_t nsteps, int64_t stride, uint64_t* counts,
                      void* stream) {
    LM_REQUIRE(n >= 0 && nsteps >= 0, LM_ERR_ARG, "negative size");
    LM_REQUIRE(nsteps <= 65535, LM_ERR_ARG, "at most 65535 steps per call");
    LM_REQUIRE(nsteps <= 1 || stride >= n, LM_ERR_ARG, "stride < n");
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(species && counts, LM_ERR_ARG, "NULL pointer");
    unsigned gx = lm::grid_for(n);
    gx = gx > 1024u ? 1024u : gx;
    LM_LAUNCH(lm::K_SPECIES_COUNT, lm::species_count_kernel, dim3(gx, (unsigned)nsteps), dim3(lm::kBlock), 0,
              (hipStream_t)stream, species, n, nsteps <= 1 ? n : stride,
              reinterpret_cast<unsigned long long*>(counts));
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_distance_histogram(int64_t n, const float* lon, const float* lat, int32_t nbins, const double* edges,
                          uint64_t* hist, void* stream) {
    LM_REQUIRE(n >= 0, LM_ERR_ARG, "negative size");
    LM_REQUIRE(nbins >= 1 && nbins <= lm::kHistMaxBins, LM_ERR_ARG, "nbins must be in 1..%d", lm::kHistMaxBins);
    const int64_t ntiles = (n + lm::kHistTile - 1) / lm::kHistTile;
    LM_REQUIRE(ntiles <= 65535, LM_ERR_ARG, "at most %lld points", (long long)65535 * lm::kHistTile);
    if (n < 2) return LM_OK;
    LM_REQUIRE(lon && lat && hist && (nbins == 1 || edges), LM_ERR_ARG, "NULL pointer");
    LM_LAUNCH(lm::K_DIST_HIST, lm::dist_hist_kernel, dim3((unsigned)ntiles, (unsigned)ntiles), dim3(lm::kHistTile),
              0, (hipStream_t)stream, n, lon, lat, (int)nbins, edges, reinterpret_cast<unsigned long long*>(hist));
    LM_LAUNCH_CHECK();
    return LM_OK;
}

}  // extern "C"
