// Fixed-radius pair search in lon/lat degree space, for one or many steps.
//
// Cell list with cells of side h >= radius over a fixed box, stored as a sort
// of the particles by bucket over T = 2^k buckets (bucket = dense cell id mod
// T, so the three cells of a stencil row are three consecutive buckets; when
// the box has fewer than T cells the table is dense and nothing aliases):
//   1. bin_keys     : each particle's bucket
//   2. partition    : one radix pass on the bucket's high bits (super-
//                     buckets; no order inside one is needed)
//   3. bin_table    : bucket starts (every bucket between two neighbouring
//                     sorted keys starts at the later one) and records
//                     {x, y, index} in bucket order
//   4. pair_count   : per particle i, the neighbours j > i within radius over
//                     the 3x3 cell stencil (run in bucket order for locality);
//                     up to kNbr of them are kept in a scratch row
//   5. scan         : row offsets and the pair count P
//   6. pair_emit    : per particle, its neighbours sorted ascending, written at
//                     its row offset; rows with more than kNbr neighbours are
//                     re-searched kNbr at a time.  The result is the (i, j)
//                     lexicographic pair list.
// A batch of S steps (S position snapshots) runs every phase once with
// blockIdx.y = step, each step with its own tables and pair region, so a
// launch carries S times the parallelism of one step.  No global atomics.
// Ensembles: with a run id per particle (independent runs sharing the index
// range), the bucket is (run * cells + cell) mod T and the record carries the
// run, so a stencil visits the cells of its own run and candidates of other
// runs whose cells alias into the same buckets are dropped: pairs never cross
// runs.
// Distances use the float32 positions widened to double, with the same
// operations cKDTree applies: dx*dx + dy*dy <= r*r for p = 2, evaluated as
// written (the library is built without FMA contraction).  For positions of
// similar magnitude the double difference of two floats is exact and so is
// its square.
#include "lm_common.h"
#include "lm_scan.h"
#include "lm_sort.h"

#include <cmath>

namespace lm {

constexpr int kNbr = 8;  // neighbours kept per particle before the selection path

struct CellGrid {
    double x0, y0, inv_h;
    int32_t nx, ny;
    uint32_t mask;   // T - 1
    uint32_t T;
    uint32_t cells;  // nx * ny (mod 2^32): the key distance between runs
};

enum MetricKind { kL2 = 0, kL1 = 1, kLinf = 2, kLp = 3 };

struct Metric {
    int kind;
    double r;    // radius
    double r2;   // radius*radius (p = 2) or radius^p (general p)
    double p;
    float box;   // float32 bound above the radius: |dx| or |dy| beyond it rules a pair out
    // p = 2: a float32 dx*dx + dy*dy at or below sure_in is within the radius,
    // above sure_out beyond it, whatever the float64 test rounds (-inf / +inf:
    // every candidate takes the float64 test)
    float sure_in, sure_out;
};

// Per-step strides of a batch (elements).
struct Strides {
    int64_t pos;    // between position snapshots
    int64_t tab;    // bucket tables (>= T + 1, multiple of 4)
    int64_t rec;    // bucket-ordered records and sort buffers
    int64_t row;    // per-particle counts / offsets (>= n + 1, multiple of 4)
    int64_t nbr;    // neighbour scratch rows
    int64_t pair;   // pair list capacity per step
};

__device__ __forceinline__ void cell_of(const CellGrid& g, float xf, float yf, int* cx, int* cy) {
    const double fx = ((double)xf - g.x0) * g.inv_h;
    const double fy = ((double)yf - g.y0) * g.inv_h;
    int ix = 0, iy = 0;
    if (fx >= 1.0) ix = fx >= (double)(g.nx - 1) ? g.nx - 1 : (int)fx;
    if (fy >= 1.0) iy = fy >= (double)(g.ny - 1) ? g.ny - 1 : (int)fy;
    *cx = ix;
    *cy = iy;
}

__device__ __forceinline__ uint32_t bucket_of(const CellGrid& g, int cx, int cy, uint32_t run) {
    return (run * g.cells + (uint32_t)cy * (uint32_t)g.nx + (uint32_t)cx) & g.mask;
}

// The metric is a template argument of the kernels that test candidates, so
// each instance holds only its own test (the general-p form inlines pow).
template <int KIND>
__device__ __forceinline__ bool within(const Metric& m, float xa, float ya, float xb, float yb) {
    const double dx = (double)xb - (double)xa;
    const double dy = (double)yb - (double)ya;
    if constexpr (KIND == kL2) return dx * dx + dy * dy <= m.r2;
    else if constexpr (KIND == kL1) return fabs(dx) + fabs(dy) <= m.r;
    else if constexpr (KIND == kLinf) return fmax(fabs(dx), fabs(dy)) <= m.r;
    else return pow(fabs(dx), m.p) + pow(fabs(dy), m.p) <= m.r2;
}

// Visit the record slots of the (up to) 3x3 cells around (cx, cy).  Slot
// ranges are clamped to nrec so that even inconsistent bucket starts cannot
// index past the records.
template <typename F>
__device__ __forceinline__ void for_stencil(const CellGrid& g, const uint32_t* __restrict__ start,
                                            uint32_t nrec, int cx, int cy, uint32_t run, F&& visit) {
    const int c0 = cx > 0 ? cx - 1 : 0;
    const int c1 = cx < g.nx - 1 ? cx + 1 : g.nx - 1;
#pragma unroll
    for (int dyc = -1; dyc <= 1; ++dyc) {
        const int row = cy + dyc;
        if (row < 0 || row >= g.ny) continue;
        const uint32_t b0 = bucket_of(g, c0, row, run);
        const uint32_t b1 = bucket_of(g, c1, row, run);
        if (b0 <= b1) {
            const uint32_t e = min(start[b1 + 1], nrec);
            for (uint32_t s = start[b0]; s < e; ++s) visit(s);
        } else {  // the row wraps around the bucket table
            const uint32_t e0 = min(start[g.T], nrec);
            for (uint32_t s = start[b0]; s < e0; ++s) visit(s);
            const uint32_t e1 = min(start[b1 + 1], nrec);
            for (uint32_t s = start[0]; s < e1; ++s) visit(s);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
bin_keys_kernel(int64_t n, const float* __restrict__ x, const float* __restrict__ y,
                const uint16_t* __restrict__ run, CellGrid g, Strides st, uint32_t* __restrict__ keys,
                uint32_t* __restrict__ vals, float2* __restrict__ xy) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t s = blockIdx.y;
    const float xf = x[s * st.pos + i], yf = y[s * st.pos + i];
    int cx, cy;
    cell_of(g, xf, yf, &cx, &cy);
    keys[s * st.rec + i] = bucket_of(g, cx, cy, run ? (uint32_t)run[i] : 0u);
    if (vals) vals[s * st.rec + i] = (uint32_t)i;  // NULL when the radix pass generates the indices
    // interleaved copy: the binning gathers a particle's position with one 8-byte load
    xy[s * st.rec + i] = make_float2(xf, yf);
}

// Second level of the binning: one block per super-bucket (the particles
// whose bucket has the same high bits, made contiguous by one radix pass),
// counting its `fine` buckets in LDS.  It writes their starts (a contiguous
// slice of the table; the last super-bucket also the end, start[T] = n) and
// places the records {x, y, index} of its particles in bucket order (the
// order inside a bucket is free: the pair list is sorted later).  FINE_MAX
// buckets per block: 4096 up to 2^22 buckets, 8192 up to 2^23.
constexpr uint32_t kFineMax = 4096;
constexpr int kPlace = 8;  // particles per thread in flight when placing records
constexpr uint32_t kFineMaxWide = 8192;

template <uint32_t FINE_MAX>
__global__ void __launch_bounds__(kBlock)
bin_table_kernel(int64_t n, const float2* __restrict__ xy, const uint16_t* __restrict__ run, CellGrid g, Strides st,
                 const uint32_t* __restrict__ keys_all,
                 const uint32_t* __restrict__ vals_all, const uint32_t* __restrict__ dstart, int64_t drow,
                 int64_t dstride, uint32_t fine, uint32_t* __restrict__ start_all, float4* __restrict__ recs_all) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    __shared__ uint32_t cnt[FINE_MAX];
    __shared__ uint32_t scan_lds[kBlock / 64];
    const uint32_t d = bx, nsuper = gridDim.x, t = threadIdx.x;
    const int64_t s = by;
    const uint32_t* keys = keys_all + s * st.rec;
    const uint32_t* vals = vals_all + s * st.rec;
    uint32_t* start = start_all + s * st.tab;
    const uint32_t nn = (uint32_t)n;
    uint32_t a = 0, b = nn;
    if (dstart != nullptr) {
        a = min(dstart[s * drow + (int64_t)d * dstride], nn);
        b = d + 1 < nsuper ? min(dstart[s * drow + (int64_t)(d + 1) * dstride], nn) : nn;
        b = max(a, b);
    }
    const uint32_t fmask = fine - 1u;
    for (uint32_t f = t; f < fine; f += kBlock) cnt[f] = 0;
    __syncthreads();
    // kPlace keys per thread in flight at once (indices clamped into the
    // super-bucket, loads pinned ahead of the atomics)
    for (uint32_t e0 = a + t; e0 < b; e0 += kBlock * kPlace) {
        uint32_t kk[kPlace];
#pragma unroll
        for (int u = 0; u < kPlace; ++u) kk[u] = keys[min(e0 + (uint32_t)u * kBlock, b - 1u)];
#pragma unroll
        for (int u = 0; u < kPlace; ++u) lm_pin(kk[u]);
#pragma unroll
        for (int u = 0; u < kPlace; ++u)
            if (e0 + (uint32_t)u * kBlock < b) atomicAdd(&cnt[kk[u] & fmask], 1u);
    }
    __syncthreads();
    // exclusive offsets: thread t owns `per` consecutive buckets
    const uint32_t per = (fine + kBlock - 1) / kBlock;
    const uint32_t f0 = min(t * per, fine), f1 = min(f0 + per, fine);
    uint32_t sum = 0;
    for (uint32_t f = f0; f < f1; ++f) sum += cnt[f];
    uint32_t all;
    uint32_t acc = block_exclusive_scan(sum, scan_lds, &all);
    for (uint32_t f = f0; f < f1; ++f) {
        const uint32_t c = cnt[f];
        cnt[f] = acc;
        acc += c;
    }
    __syncthreads();
    const uint64_t base = (uint64_t)d * fine;
    for (uint32_t f = t; f < fine; f += kBlock)
        if (base + f < (uint64_t)g.T) start[base + f] = a + cnt[f];
    if (d + 1 == nsuper && t == 0) start[g.T] = nn;
    __syncthreads();
    // kPlace particles per thread at once, so their position gathers (random
    // lines) are in flight together
    // two rounds of loads per kPlace particles, each issued at once (indices
    // clamped, loads pinned ahead of their uses): indices and keys, then the
    // positions the indices point at
    for (uint32_t e0 = a + t; e0 < b; e0 += kBlock * kPlace) {
        uint32_t ii[kPlace], kk[kPlace], pp[kPlace];
#pragma unroll
        for (int u = 0; u < kPlace; ++u) {
            const uint32_t e = min(e0 + (uint32_t)u * kBlock, b - 1u);
            ii[u] = vals[e];
            kk[u] = keys[e];
        }
#pragma unroll
        for (int u = 0; u < kPlace; ++u) lm_pin(ii[u], kk[u]);
#pragma unroll
        for (int u = 0; u < kPlace; ++u) {
            const bool live = e0 + (uint32_t)u * kBlock < b && ii[u] < nn;
            pp[u] = live ? a + atomicAdd(&cnt[kk[u] & fmask], 1u) : b;
        }
        float2 pv[kPlace];
#pragma unroll
        for (int u = 0; u < kPlace; ++u) pv[u] = xy[s * st.rec + (ii[u] < nn ? ii[u] : 0u)];
        uint32_t rr[kPlace];
#pragma unroll
        for (int u = 0; u < kPlace; ++u) rr[u] = run ? (uint32_t)run[ii[u] < nn ? ii[u] : 0u] : 0u;
#pragma unroll
        for (int u = 0; u < kPlace; ++u)
            if (pp[u] < b)
                recs_all[s * st.rec + pp[u]] = make_float4(pv[u].x, pv[u].y, __uint_as_float(ii[u]),
                                                           __uint_as_float(rr[u]));
    }
}

// The (up to) six record ranges of the 3x3 stencil around (cx, cy): a row of
// three cells is one bucket range, or two when it wraps around the table.
// All bucket starts are loaded at once.  Ranges are clamped to nrec.
__device__ __forceinline__ void stencil_ranges(const CellGrid& g, const uint32_t* __restrict__ start, uint32_t nrec,
                                               int cx, int cy, uint32_t run, uint32_t (&rb)[6], uint32_t (&re)[6]) {
    const int c0 = cx > 0 ? cx - 1 : 0;
    const int c1 = cx < g.nx - 1 ? cx + 1 : g.nx - 1;
    uint32_t lo[6], hi[6];
    bool on[6];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int row = cy + r - 1;
        const bool valid = row >= 0 && row < g.ny;
        const uint32_t b0 = valid ? bucket_of(g, c0, row, run) : 0u;
        const uint32_t b1 = valid ? bucket_of(g, c1, row, run) : 0u;
        const bool wrap = b0 > b1;
        // first segment [b0, wrap ? T : b1 + 1), second [0, b1 + 1) if wrapped
        lo[2 * r] = b0;
        hi[2 * r] = wrap ? g.T : b1 + 1u;
        on[2 * r] = valid;
        lo[2 * r + 1] = 0u;
        hi[2 * r + 1] = b1 + 1u;
        on[2 * r + 1] = valid && wrap;
    }
    uint32_t sb[6], se[6];
    // the three main segments load unconditionally (index 0 for a row off the
    // grid, then discarded); the wrapped second segments are rare
#pragma unroll
    for (int r = 0; r < 6; r += 2) {
        sb[r] = start[on[r] ? lo[r] : 0u];
        se[r] = start[on[r] ? hi[r] : 0u];
        sb[r] = on[r] ? sb[r] : 0u;
        se[r] = on[r] ? se[r] : 0u;
    }
#pragma unroll
    for (int r = 1; r < 6; r += 2) {
        sb[r] = on[r] ? start[lo[r]] : 0u;
        se[r] = on[r] ? start[hi[r]] : 0u;
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        re[r] = min(se[r], nrec);
        rb[r] = min(sb[r], re[r]);
    }
}

constexpr int kCand = 8;  // candidate records in flight per thread

template <int KIND>
__global__ void __launch_bounds__(kBlock)
pair_count_kernel(int64_t n, CellGrid g, Metric m, Strides st, const uint32_t* __restrict__ start_all,
                  const float4* __restrict__ recs_all, uint32_t* __restrict__ nbr_cnt,
                  uint32_t* __restrict__ nbr) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // a particle's stencil rows lie a few blocks away in bucket order
    const int64_t s0 = (int64_t)bx * kBlock + threadIdx.x;
    if (s0 >= n) return;
    const int64_t s = by;
    const uint32_t* start = start_all + s * st.tab;
    const float4* recs = recs_all + s * st.rec;
    const float4 me = recs[s0];
    uint32_t i = __float_as_uint(me.z), my_run = __float_as_uint(me.w);
    lm_pin(i, my_run);  // the whole record in one load, not split around the branch
    if (i >= (uint64_t)n) return;
    int cx, cy;
    cell_of(g, me.x, me.y, &cx, &cy);
    uint32_t rb[6], re[6];
    stencil_ranges(g, start, (uint32_t)n, cx, cy, my_run, rb, re);
    // the ranges as one sequence: candidate t lies in range r when
    // pre[r] <= t < pre[r + 1], at slot t + dlt[r]
    uint32_t pre[6], dlt[6];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        dlt[r] = rb[r] - pre[r];  // mod 2^32
        if (r < 5) pre[r + 1] = pre[r] + (re[r] - rb[r]);
    }
    const uint32_t total = pre[5] + (re[5] - rb[5]);
    uint32_t cnt = 0;
    uint32_t* row = nbr + s * st.nbr + (int64_t)i * kNbr;
    for (uint32_t t0 = 0; t0 < total; t0 += kCand) {
        float4 o[kCand];
#pragma unroll
        for (int u = 0; u < kCand; ++u) {
            // past the end the lane is masked off (no address-unit work) and
            // the candidate is discarded below
            const uint32_t t = t0 + u;
            uint32_t d = dlt[0];
#pragma unroll
            for (int r = 1; r < 6; ++r) d = t >= pre[r] ? dlt[r] : d;
            o[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (t < total) o[u] = recs[t + d];
        }
        // every load of the batch is issued before any use of the first: the
        // scheduler would otherwise start testing candidate 0 early and wait
        // for it before issuing the rest
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kCand; ++u) {
            const uint32_t j = __float_as_uint(o[u].z);
            const bool cand = (t0 + u < total) & (j > i) & (__float_as_uint(o[u].w) == my_run);
            bool hit;
            if constexpr (KIND == kL2) {
                // A float32 difference of two float32 positions is within one
                // rounding (relative 2^-24) of the exact one, so the float32
                // dx*dx + dy*dy is within 2.4e-7 (relative) of the exact sum of
                // squares, and the float64 test within 2^-52: outside r^2 (1 -/+
                // 1e-6) both decide alike.  Only candidates in that band take
                // the float64 test (a branch almost never taken).
                const float dx = o[u].x - me.x, dy = o[u].y - me.y;
                const float d2 = fmaf(dx, dx, dy * dy);
                hit = cand & (d2 <= m.sure_in);
                if (cand & !(d2 <= m.sure_in) & !(d2 > m.sure_out))
                    hit = within<KIND>(m, me.x, me.y, o[u].x, o[u].y);
            } else {
                // |dx| <= r and |dy| <= r hold for every p >= 1; a float32
                // difference is within one rounding (relative 2^-24) of the exact
                // one and m.box lies above r by more than that, so the prefilter
                // only drops pairs the exact test drops.
                const float ax = fabsf(o[u].x - me.x), ay = fabsf(o[u].y - me.y);
                hit = cand & (ax <= m.box) & (ay <= m.box) & within<KIND>(m, me.x, me.y, o[u].x, o[u].y);
            }
            if (hit & (cnt < kNbr)) row[cnt] = j;
            cnt += hit ? 1u : 0u;
        }
    }
    nbr_cnt[s * st.row + i] = cnt;
}

// Sort the first c (<= K) entries of a scratch row in registers and write
// them as row o.. of the pair list.
template <int K>
__device__ __forceinline__ void emit_sorted(const uint32_t* row, uint32_t c, uint32_t i, uint32_t o,
                                            int32_t* __restrict__ pair_i, int32_t* __restrict__ pair_j) {
    uint32_t v[K];
#pragma unroll
    for (int t = 0; t < K; ++t) v[t] = (uint32_t)t < c ? row[t] : 0xffffffffu;
#pragma unroll
    for (int a = 1; a < K; ++a)
#pragma unroll
        for (int b = a; b > 0; --b) {
            const uint32_t lo = min(v[b - 1], v[b]), hi = max(v[b - 1], v[b]);
            v[b - 1] = lo;
            v[b] = hi;
        }
#pragma unroll
    for (int t = 0; t < K; ++t)
        if ((uint32_t)t < c) {
            pair_i[o + t] = (int32_t)i;
            pair_j[o + t] = (int32_t)v[t];
        }
}

template <int KIND>
__global__ void __launch_bounds__(kBlock)
pair_emit_kernel(int64_t n, const float* __restrict__ x_all, const float* __restrict__ y_all,
                 const uint16_t* __restrict__ run, CellGrid g, Metric m, Strides st, const uint32_t* __restrict__ start_all, const float4* __restrict__ recs_all,
                 const uint32_t* __restrict__ nbr_cnt_all, const uint32_t* __restrict__ nbr_all,
                 const uint32_t* __restrict__ row_off_all, int32_t* __restrict__ pair_i_all,
                 int32_t* __restrict__ pair_j_all) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    const int64_t i64 = (int64_t)bx * kBlock + threadIdx.x;
    if (i64 >= n) return;
    const int64_t s = by;
    const uint32_t* row_off = row_off_all + s * st.row;
    const uint32_t i = (uint32_t)i64;
    // the step's total, this row's count and offset and its first two
    // partners (read whatever the count, then used as it says) load
    // together, before any branch on them
    const uint32_t total = row_off[n];
    const uint32_t* row = nbr_all + s * st.nbr + (int64_t)i * kNbr;
    uint32_t c = nbr_cnt_all[s * st.row + i], o = row_off[i];
    const uint2 ab = *reinterpret_cast<const uint2*>(row);  // rows are 32-byte aligned
    uint32_t r0 = ab.x, r1 = ab.y;
    lm_pin(c, o, r0, r1);
    if ((int64_t)total > st.pair) return;  // caller enlarges and calls again
    if (c == 0) return;
    int32_t* pair_i = pair_i_all + s * st.pair;
    int32_t* pair_j = pair_j_all + s * st.pair;
    if (c <= 2) {
        const uint32_t a = r0, b = c == 2 ? r1 : 0xffffffffu;
        pair_i[o] = (int32_t)i;
        pair_j[o] = (int32_t)min(a, b);
        if (c == 2) {
            pair_i[o + 1] = (int32_t)i;
            pair_j[o + 1] = (int32_t)max(a, b);
        }
        return;
    }
    if (c <= 4) return emit_sorted<4>(row, c, i, o, pair_i, pair_j);
    if (c <= (uint32_t)kNbr) return emit_sorted<kNbr>(row, c, i, o, pair_i, pair_j);
    // Dense neighbourhood: emit in ascending order, 16 at a time, each pass
    // keeping the 16 smallest partners above the last one written.
    constexpr int kSel = 16;
    const uint32_t* start = start_all + s * st.tab;
    const float4* recs = recs_all + s * st.rec;
    const float xf = x_all[s * st.pos + i], yf = y_all[s * st.pos + i];
    const uint32_t my_run = run ? (uint32_t)run[i] : 0u;
    int cx, cy;
    cell_of(g, xf, yf, &cx, &cy);
    uint32_t last = i;
    for (uint32_t t = 0; t < c;) {
        uint32_t v[kSel];
#pragma unroll
        for (int q = 0; q < kSel; ++q) v[q] = 0xffffffffu;
        for_stencil(g, start, (uint32_t)n, cx, cy, my_run, [&](uint32_t q0) {
            const float4 rec = recs[q0];
            uint32_t j = __float_as_uint(rec.z);
            if (j > last && j < v[kSel - 1] && __float_as_uint(rec.w) == my_run && within<KIND>(m, xf, yf, rec.x, rec.y)) {
#pragma unroll
                for (int q = 0; q < kSel; ++q) {  // insert into the sorted register array
                    const uint32_t lo = min(v[q], j), hi = max(v[q], j);
                    v[q] = lo;
                    j = hi;
                }
            }
        });
        const uint32_t take = min((uint32_t)kSel, c - t);
#pragma unroll
        for (int q = 0; q < kSel; ++q)
            if ((uint32_t)q < take) {
                pair_i[o + t + q] = (int32_t)i;
                pair_j[o + t + q] = (int32_t)v[q];
                last = v[q];
            }
        t += take;
    }
}

__global__ void zero_i64_kernel(int64_t* p, int64_t count) {
    for (int64_t k = threadIdx.x; k < count; k += blockDim.x) p[k] = 0;
}

}  // namespace lm

struct lm_pairs_ws {
    int64_t n_max = 0;
    int32_t steps_max = 0;
    double max_radius = 0.0;
    lm::CellGrid grid{};
    int64_t tab_stride = 0, row_stride = 0;
    uint32_t* keys[2] = {nullptr, nullptr};  // [steps][n_max] sort buffers
    uint32_t* vals[2] = {nullptr, nullptr};
    lm::SortScratch sort{};
    int key_bits = 0;              // log2(T)
    uint32_t* start = nullptr;     // [steps][tab]
    uint32_t* scan_tmp = nullptr;  // block sums
    float4* recs = nullptr;        // [steps][n_max]
    float2* xy = nullptr;          // [steps][n_max] positions interleaved, in index order
    uint32_t* nbr_cnt = nullptr;   // [steps][row]
    uint32_t* nbr = nullptr;       // [steps][n_max * kNbr]
    uint32_t* row_off = nullptr;   // [steps][row]
    int64_t bytes = 0;
};

static void pairs_ws_free(lm_pairs_ws* w) {
    if (!w) return;
    for (int q = 0; q < 2; ++q) {
        (void)hipFree(w->keys[q]);
        (void)hipFree(w->vals[q]);
    }
    lm::sort_scratch_free(&w->sort);
    (void)hipFree(w->start);
    (void)hipFree(w->scan_tmp);
    (void)hipFree(w->recs);
    (void)hipFree(w->xy);
    (void)hipFree(w->nbr_cnt);
    (void)hipFree(w->nbr);
    (void)hipFree(w->row_off);
    delete w;
}

template <typename T>
static int ws_alloc(T** p, int64_t count, int64_t* bytes) {
    const size_t b = (size_t)(count < 1 ? 1 : count) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, b);
    if (e != hipSuccess) {
        lm::set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        return LM_ERR_NOMEM;
    }
    *bytes += (int64_t)b;
    return LM_OK;
}

static uint32_t next_pow2_u32(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}

namespace lm {

// Phases 4-6 with the distance test of metric KIND.
template <int KIND>
static int count_and_emit(lm_pairs_ws* ws, int32_t nsteps, int64_t n, const float* lon, const float* lat,
                          const uint16_t* run, const CellGrid& g, const Metric& m, const Strides& st, int64_t* npairs,
                          int32_t* pair_i, int32_t* pair_j, hipStream_t s) {
    const dim3 grid(grid_for(n), (unsigned)nsteps);
    LM_LAUNCH(K_PAIR_COUNT, pair_count_kernel<KIND>, grid, dim3(kBlock), 0, s, n, g, m, st,
              (const uint32_t*)ws->start, (const float4*)ws->recs, ws->nbr_cnt, ws->nbr);
    LM_LAUNCH_CHECK();
    const int rc = exclusive_scan_u32_batched(ws->nbr_cnt, n, st.row, ws->row_off, st.row, nsteps, ws->scan_tmp,
                                              npairs, s);
    if (rc) return rc;
    LM_LAUNCH(K_PAIR_EMIT, pair_emit_kernel<KIND>, grid, dim3(kBlock), 0, s, n, lon, lat, run, g, m, st,
              (const uint32_t*)ws->start, (const float4*)ws->recs, (const uint32_t*)ws->nbr_cnt,
              (const uint32_t*)ws->nbr, (const uint32_t*)ws->row_off, pair_i, pair_j);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

}  // namespace lm

extern "C" {

int lm_pairs_ws_create(int64_t n_max, int32_t steps_max, double lon_min, double lon_max, double lat_min,
                       double lat_max, double max_radius, lm_pairs_ws** out) {
    return lm_pairs_ws_create_ex(n_max, steps_max, lon_min, lon_max, lat_min, lat_max, max_radius, 1.0, out);
}

int lm_pairs_ws_create_ex(int64_t n_max, int32_t steps_max, double lon_min, double lon_max, double lat_min,
                          double lat_max, double max_radius, double cell_scale, lm_pairs_ws** out) {
    LM_REQUIRE(out != nullptr, LM_ERR_ARG, "out is NULL");
    *out = nullptr;
    LM_REQUIRE(n_max >= 0 && n_max < (int64_t)0x7fffffff, LM_ERR_ARG, "n_max out of range");
    LM_REQUIRE(steps_max >= 1 && steps_max <= 65535, LM_ERR_ARG, "steps_max must be in 1..65535");
    LM_REQUIRE(std::isfinite(lon_min) && std::isfinite(lon_max) && std::isfinite(lat_min) &&
                   std::isfinite(lat_max) && lon_max >= lon_min && lat_max >= lat_min,
               LM_ERR_ARG, "invalid bounding box");
    LM_REQUIRE(std::isfinite(max_radius) && max_radius > 0.0, LM_ERR_ARG, "max_radius must be > 0");
    LM_REQUIRE(std::isfinite(cell_scale) && cell_scale >= 1.0 && cell_scale <= 16.0, LM_ERR_ARG,
               "cell_scale must be in [1, 16]");

    lm_pairs_ws* w = new lm_pairs_ws();
    w->n_max = n_max;
    w->steps_max = steps_max;
    w->max_radius = max_radius;
    // cells strictly larger than the radius so rounding in the cell index can
    // never separate two points within the radius by more than one cell
    double h = max_radius * cell_scale * (1.0 + 1e-6);
    const double ext = (lon_max - lon_min) > (lat_max - lat_min) ? (lon_max - lon_min) : (lat_max - lat_min);
    if (ext / h > 30000.0) h = ext / 30000.0;
    lm::CellGrid& g = w->grid;
    g.inv_h = 1.0 / h;
    g.x0 = lon_min - h;
    g.y0 = lat_min - h;
    g.nx = (int32_t)std::floor((lon_max - lon_min) * g.inv_h) + 3;
    g.ny = (int32_t)std::floor((lat_max - lat_min) * g.inv_h) + 3;
    const uint64_t cells = (uint64_t)g.nx * (uint64_t)g.ny;
    // about two buckets per particle: fewer cells alias into a stencil's buckets
    uint64_t want = 2ull * (uint64_t)n_max;
    if (want < 4ull * (uint64_t)g.nx + 4) want = 4ull * (uint64_t)g.nx + 4;
    if (want < 1024) want = 1024;
    uint32_t T = next_pow2_u32(want);
    if ((uint64_t)T >= cells) T = next_pow2_u32(cells);
    if (T < 1024) T = 1024;
    // the binning splits a bucket key into at most 13 bits counted in LDS and one radix pass of at
    // most kMaxDigitBits: at most 2^23 buckets
    const uint32_t t_max = 1u << (13 + lm::kMaxDigitBits);
    if (T > t_max) T = t_max;
    g.T = T;
    g.mask = T - 1;
    g.cells = (uint32_t)((uint64_t)g.nx * (uint64_t)g.ny);
    w->tab_stride = lm::scan_row_stride((int64_t)T);
    w->row_stride = lm::scan_row_stride(n_max);

    const int64_t S = steps_max;
    int64_t nb = lm::scan_num_blocks((int64_t)T);
    const int64_t nb2 = lm::scan_num_blocks(n_max);
    if (nb2 > nb) nb = nb2;
    int rc = LM_OK;
    if ((rc = ws_alloc(&w->keys[0], S * n_max, &w->bytes)) || (rc = ws_alloc(&w->keys[1], S * n_max, &w->bytes)) ||
        (rc = ws_alloc(&w->vals[0], S * n_max, &w->bytes)) || (rc = ws_alloc(&w->vals[1], S * n_max, &w->bytes)) ||
        (rc = lm::sort_scratch_alloc(&w->sort, n_max, steps_max)) ||
        (rc = ws_alloc(&w->start, S * w->tab_stride, &w->bytes)) ||
        (rc = ws_alloc(&w->scan_tmp, S * nb, &w->bytes)) || (rc = ws_alloc(&w->recs, S * n_max, &w->bytes)) ||
        (rc = ws_alloc(&w->xy, S * n_max, &w->bytes)) ||
        (rc = ws_alloc(&w->nbr_cnt, S * w->row_stride, &w->bytes)) ||
        (rc = ws_alloc(&w->nbr, S * n_max * lm::kNbr, &w->bytes)) ||
        (rc = ws_alloc(&w->row_off, S * w->row_stride, &w->bytes))) {
        pairs_ws_free(w);
        return rc;
    }
    w->bytes += w->sort.bytes;
    while (((uint32_t)1 << w->key_bits) < T) ++w->key_bits;
    *out = w;
    return LM_OK;
}

void lm_pairs_ws_destroy(lm_pairs_ws* ws) { pairs_ws_free(ws); }

int64_t lm_pairs_ws_bytes(const lm_pairs_ws* ws) { return ws ? ws->bytes : 0; }

int64_t lm_pairs_ws_buckets(const lm_pairs_ws* ws) { return ws ? (int64_t)ws->grid.T : 0; }

int lm_query_pairs_batch(lm_pairs_ws* ws, int32_t nsteps, int64_t n, const float* lon, const float* lat,
                         int64_t pos_stride, double radius, double p, int32_t* pair_i, int32_t* pair_j,
                         int64_t pair_stride, int64_t* npairs, void* stream) {
    return lm_query_pairs_runs(ws, nsteps, n, lon, lat, pos_stride, nullptr, radius, p, pair_i, pair_j, pair_stride,
                               npairs, stream);
}

int lm_query_pairs_runs(lm_pairs_ws* ws, int32_t nsteps, int64_t n, const float* lon, const float* lat,
                        int64_t pos_stride, const uint16_t* run, double radius, double p, int32_t* pair_i,
                        int32_t* pair_j, int64_t pair_stride, int64_t* npairs, void* stream) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(npairs != nullptr, LM_ERR_ARG, "npairs is NULL");
    LM_REQUIRE(nsteps >= 0 && nsteps <= ws->steps_max, LM_ERR_ARG, "nsteps=%d outside 0..%d of the workspace",
               nsteps, ws->steps_max);
    LM_REQUIRE(n >= 0 && n <= ws->n_max, LM_ERR_ARG, "n=%lld exceeds the workspace (n_max=%lld)",
               (long long)n, (long long)ws->n_max);
    LM_REQUIRE(radius >= 0.0 && radius <= ws->max_radius, LM_ERR_ARG,
               "radius %g outside [0, %g] of the workspace", radius, ws->max_radius);
    LM_REQUIRE(p >= 1.0, LM_ERR_ARG, "Minkowski p must be >= 1 (got %g)", p);
    LM_REQUIRE(pair_stride >= 0, LM_ERR_ARG, "negative pair capacity");
    LM_REQUIRE(nsteps <= 1 || pos_stride >= n, LM_ERR_ARG, "position stride < n");
    hipStream_t s = (hipStream_t)stream;
    if (nsteps == 0) return LM_OK;
    if (n == 0) {
        hipLaunchKernelGGL(lm::zero_i64_kernel, dim3(1), dim3(64), 0, s, npairs, (int64_t)nsteps);
        LM_LAUNCH_CHECK();
        return LM_OK;
    }
    LM_REQUIRE(lon && lat, LM_ERR_ARG, "NULL positions");
    LM_REQUIRE(pair_stride == 0 || (pair_i && pair_j), LM_ERR_ARG, "NULL pair buffers");

    lm::Metric m;
    m.r = radius;
    m.p = p;
    // above r by more than float32 rounding of a difference (2^-24 relative)
    m.box = std::nextafter((float)(radius * (1.0 + 1e-6)), INFINITY);
    m.sure_in = -INFINITY;
    m.sure_out = INFINITY;
    if (p == 2.0) {
        m.kind = lm::kL2;
        m.r2 = radius * radius;
        // far from float32 underflow and overflow, the float32 bounds of the
        // band around r^2 in which candidates take the float64 test
        if (m.r2 >= 1e-30 && m.r2 <= 1e30) {
            m.sure_in = std::nextafter((float)(m.r2 * (1.0 - 1e-6)), -INFINITY);
            m.sure_out = std::nextafter((float)(m.r2 * (1.0 + 1e-6)), INFINITY);
        }
    } else if (p == 1.0) {
        m.kind = lm::kL1;
        m.r2 = radius;
    } else if (std::isinf(p)) {
        m.kind = lm::kLinf;
        m.r2 = radius;
    } else {
        m.kind = lm::kLp;
        m.r2 = std::pow(radius, p);
    }
    const lm::CellGrid& g = ws->grid;
    lm::Strides st;
    st.pos = pos_stride;
    st.tab = ws->tab_stride;
    st.rec = ws->n_max;
    st.row = ws->row_stride;
    st.nbr = ws->n_max * lm::kNbr;
    st.pair = pair_stride;
    // two-level binning: one radix pass on the high bits of the bucket, then
    // a block per super-bucket sorts its fine buckets in LDS
    // 12 fine bits in LDS (13 above 2^22 buckets), the rest in one radix pass
    const int fine_bits = ws->key_bits <= 12 ? ws->key_bits : ws->key_bits <= 22 ? 12 : ws->key_bits - lm::kMaxDigitBits;
    const int super_bits = ws->key_bits - fine_bits;
    const dim3 grid(lm::grid_for(n), (unsigned)nsteps);
    // the particle indices are written only when no radix pass generates them
    LM_LAUNCH(lm::K_BIN_KEYS, lm::bin_keys_kernel, grid, dim3(lm::kBlock), 0, s, n, lon, lat, run, g, st,
              ws->keys[0], super_bits > 0 ? nullptr : ws->vals[0], ws->xy);
    LM_LAUNCH_CHECK();
    const uint32_t* dstart = nullptr;
    int64_t dstride = 0;
    int rc = LM_OK;
    int res = 0;
    if (super_bits > 0) {
        // unstable: the binning places a bucket's records in any order (the
        // pair list is sorted later), so no ranking of equal digits is needed
        rc = lm::radix_partition_u32(ws->keys, ws->vals, st.rec, nsteps, n, fine_bits, super_bits, &ws->sort, s,
                                     &dstart, &dstride, true, false);
        if (rc) return rc;
        res = 1;
    }
    if (fine_bits <= 12) {
        LM_LAUNCH(lm::K_BIN_TABLE, lm::bin_table_kernel<lm::kFineMax>, dim3(1u << super_bits, (unsigned)nsteps),
                  dim3(lm::kBlock), 0, s, n, (const float2*)ws->xy, run, g, st, (const uint32_t*)ws->keys[res],
                  (const uint32_t*)ws->vals[res], dstart, ws->sort.row, dstride, 1u << fine_bits, ws->start, ws->recs);
    } else {
        LM_LAUNCH(lm::K_BIN_TABLE, lm::bin_table_kernel<lm::kFineMaxWide>, dim3(1u << super_bits, (unsigned)nsteps),
                  dim3(lm::kBlock), 0, s, n, (const float2*)ws->xy, run, g, st, (const uint32_t*)ws->keys[res],
                  (const uint32_t*)ws->vals[res], dstart, ws->sort.row, dstride, 1u << fine_bits, ws->start, ws->recs);
    }
    LM_LAUNCH_CHECK();
    switch (m.kind) {
        case lm::kL2: rc = lm::count_and_emit<lm::kL2>(ws, nsteps, n, lon, lat, run, g, m, st, npairs, pair_i, pair_j, s); break;
        case lm::kL1: rc = lm::count_and_emit<lm::kL1>(ws, nsteps, n, lon, lat, run, g, m, st, npairs, pair_i, pair_j, s); break;
        case lm::kLinf:
            rc = lm::count_and_emit<lm::kLinf>(ws, nsteps, n, lon, lat, run, g, m, st, npairs, pair_i, pair_j, s);
            break;
        default: rc = lm::count_and_emit<lm::kLp>(ws, nsteps, n, lon, lat, run, g, m, st, npairs, pair_i, pair_j, s); break;
    }
    return rc;
}

int lm_query_pairs(lm_pairs_ws* ws, int64_t n, const float* lon, const float* lat, double radius, double p,
                   int32_t* pair_i, int32_t* pair_j, int64_t capacity, int64_t* npairs, void* stream) {
    return lm_query_pairs_batch(ws, 1, n, lon, lat, n, radius, p, pair_i, pair_j, capacity, npairs, stream);
}

}  // extern "C"
