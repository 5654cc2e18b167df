// Rock-paper-scissors update over sorted pair lists with the exact result of
// the sequential visit in (i, j) order, for one or many steps.
//
// A pair only reads and writes the species of its two particles, so pairs in
// different connected components of the pair graph commute; inside a component
// the sequential order is kept.  Pairs are sorted by (i, j), so a component's
// pairs in global order are the rows (runs of equal i) of its members taken in
// ascending i.
//
// Everything that depends only on the pairs runs once for a batch of steps
// (blockIdx.y = step), with the parallelism of all steps at once:
//   prepare 1. init     parent[v] = v, local slots cleared
//           2. union    lock-free union-find, hooking the larger root under the
//                       smaller (every root is the smallest index of its
//                       component); the same pass draws each pair's uniform
//                       (Philox keyed by pair rank and step) and keeps its three
//                       comparisons with pRS, pPR, pSP as bits
//           3. claim    each particle of a pair takes a local slot in its
//                       component; each row head counts itself and its row
//           4. scans    head, particle and pair offsets per component
//           5. group    heads and particle ids grouped by component; small
//                       components (<= 8 heads, <= 32 pairs) and the rest listed
//           6. order    heads sorted per component (registers / LDS bitonic) and
//                       each row's position in its component's pair sequence
//           7. fill     pair records (slot_i | bits << 29, slot_j) in component
//                       order
//           8. deps     for the wide components, per 64-pair window, the bitmask
//                       of earlier pairs of the window that share a particle
// Only the replay itself runs per step (two launches):
//   play    narrow      one lane per small component: species into LDS by slot,
//                       pairs replayed in order; particles without pairs copied
//           wide        one wavefront per larger component: species into LDS,
//                       each window replayed in rounds -- a pair fires once every
//                       earlier pair of the window it depends on has fired, which
//                       is exactly the sequential order
// Counters return to zero within a call.
#include "lm_common.h"
#include "lm_scan.h"

namespace lm {

constexpr uint32_t kUnset = 0xffffffffu;
constexpr uint32_t kClaimed = 0xfffffffeu;
constexpr uint32_t kNarrowHeads = 8;
constexpr uint32_t kNarrowPairs = 32;
constexpr uint32_t kNarrowSlots = kNarrowPairs + 1;  // a component of P pairs has <= P + 1 particles
constexpr uint32_t kSortHeads = 2048;                // LDS head sort of the wide tier
constexpr uint32_t kWideSlots = 16384;               // LDS species slots of the wide replay
constexpr uint32_t kSlotMask = (1u << 29) - 1u;
constexpr unsigned kPairGrid = 256;                  // blocks per step for pair-parallel passes
constexpr unsigned kCompGrid = 128;                  // blocks per step for component-parallel passes

// Per-step strides (elements) of a batch.
struct RpsStrides {
    int64_t row;   // per-particle arrays (>= n + 1, multiple of 4)
    int64_t pair;  // per-pair arrays (the pair capacity of a step)
};

// Per-step views of the workspace arrays.
struct RpsArrays {
    uint32_t *parent, *label, *loc, *rowbeg, *rowend, *rowoff;
    uint32_t *hcount, *hstart, *pcount, *pstart, *qcount, *qstart;
    uint32_t *members, *ids, *narrow, *wide, *lists;
    uint8_t* bits;
    uint2* rec;
    unsigned long long* deps;
};

__device__ __forceinline__ uint32_t find_root(const uint32_t* parent_c, uint32_t v) {
    // Plain loads: a stale parent is still an ancestor (pointers only move to
    // smaller indices of the same component), and the CAS in unite() is the
    // only step that must see the current value.
    uint32_t* parent = const_cast<uint32_t*>(parent_c);
    uint32_t p = parent[v];
    while (p != v) {
        const uint32_t gp = parent[p];
        if (gp != p) parent[v] = gp;  // path halving; gp is an ancestor of v
        v = p;
        p = gp;
    }
    return v;
}

__device__ __forceinline__ void unite(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = find_root(parent, a);
        b = find_root(parent, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        // a > b, both looked like roots: hook a under b
        const uint32_t old = atomicCAS(parent + a, a, b);
        if (old == a) return;
        a = old;  // a had been hooked meanwhile; old < a, so this terminates
    }
}

__device__ __forceinline__ int64_t step_pairs(const int64_t* npairs, int64_t s, int64_t cap) {
    const int64_t P = npairs[s];
    return P > cap ? -1 : P;  // an overflowed step is skipped; the caller replays it
}

// One interaction on (si, sj).  bits: bit0 u < pRS, bit1 u < pPR, bit2 u < pSP.
// Cyclic dominance: a beats b iff a == b % 3 + 1 (rock 1 > scissors 3, paper
// 2 > rock 1, scissors 3 > paper 2); the dominant species of the unordered pair
// wins with that pair's forward probability; the loser takes the winner's code.
__device__ __forceinline__ void play_pair(int& si, int& sj, uint32_t bits) {
    if (si == sj || (unsigned)(si - 1) > 2u || (unsigned)(sj - 1) > 2u) return;
    const int sum = si + sj;  // 3: rock-paper, 4: rock-scissors, 5: paper-scissors
    const uint32_t forward = (bits >> (sum == 4 ? 0 : sum == 3 ? 1 : 2)) & 1u;
    const bool i_beats_j = (si == sj % 3 + 1);
    if ((forward != 0) == i_beats_j)
        sj = si;
    else
        si = sj;
}

// Wavefront-aggregated append of v to list (one atomic per wavefront).
__device__ __forceinline__ void wave_append(bool flag, uint32_t v, uint32_t* __restrict__ list,
                                            uint32_t* __restrict__ count) {
    const uint64_t mask = __ballot(flag);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (flag) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = v;
}

#define RPS_STEP_VIEW(A, st, s)                                                                 \
    const int64_t _ro = (s) * (st).row, _po = (s) * (st).pair;                                  \
    uint32_t* parent = (A).parent + _ro;                                                        \
    uint32_t* label = (A).label + _ro;                                                          \
    uint32_t* loc = (A).loc + _ro;                                                              \
    uint32_t* rowbeg = (A).rowbeg + _ro;                                                        \
    uint32_t* rowend = (A).rowend + _ro;                                                        \
    uint32_t* rowoff = (A).rowoff + _ro;                                                        \
    uint32_t* hcount = (A).hcount + _ro;                                                        \
    uint32_t* hstart = (A).hstart + _ro;                                                        \
    uint32_t* pcount = (A).pcount + _ro;                                                        \
    uint32_t* pstart = (A).pstart + _ro;                                                        \
    uint32_t* qcount = (A).qcount + _ro;                                                        \
    uint32_t* qstart = (A).qstart + _ro;                                                        \
    uint32_t* members = (A).members + _ro;                                                      \
    uint32_t* ids = (A).ids + _ro;                                                              \
    uint32_t* narrow = (A).narrow + _ro;                                                        \
    uint32_t* wide = (A).wide + _ro;                                                            \
    uint32_t* lists = (A).lists + 4 * (s);                                                      \
    uint8_t* bits = (A).bits + _po;                                                             \
    uint2* rec = (A).rec + _po;                                                                 \
    unsigned long long* deps = (A).deps + _po;                                                  \
    (void)parent; (void)label; (void)loc; (void)rowbeg; (void)rowend; (void)rowoff;             \
    (void)hcount; (void)hstart; (void)pcount; (void)pstart; (void)qcount; (void)qstart;         \
    (void)members; (void)ids; (void)narrow; (void)wide; (void)lists; (void)bits; (void)rec;     \
    (void)deps

// ---------------------------------------------------------------------------
// prepare
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
rps_init_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    if (v < n) {
        parent[v] = (uint32_t)v;
        loc[v] = kUnset;
    }
    if (v < 4) lists[v] = 0;
}

__global__ void __launch_bounds__(kBlock)
rps_union_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all,
                 const int64_t* __restrict__ npairs, RpsArrays A, RpsStrides st, double pRS, double pPR,
                 double pSP, uint64_t k0, uint64_t k1, uint64_t step0) {
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const int64_t P = step_pairs(npairs, s, st.pair);
    const int32_t* pi = pi_all + _po;
    const int32_t* pj = pj_all + _po;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
        const double u =
            u01_from_u64(philox4x64_10((uint64_t)k, step0 + (uint64_t)s, kStreamPairs, 0, k0, k1).v[0]);
        bits[k] = (uint8_t)((u < pRS ? 1u : 0u) | (u < pPR ? 2u : 0u) | (u < pSP ? 4u : 0u));
        if (a < (uint64_t)n && b < (uint64_t)n && a != b) unite(parent, a, b);
    }
}

__device__ __forceinline__ void claim_slot(uint32_t* loc, uint32_t* label, uint32_t* parent, uint32_t* pcount,
                                           uint32_t x) {
    if (atomicCAS(loc + x, kUnset, kClaimed) == kUnset) {
        const uint32_t L = find_root(parent, x);
        label[x] = L;
        loc[x] = atomicAdd(pcount + L, 1u);
    }
}

__global__ void __launch_bounds__(kBlock)
rps_claim_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all,
                 const int64_t* __restrict__ npairs, RpsArrays A, RpsStrides st) {
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const int64_t P = step_pairs(npairs, s, st.pair);
    const int32_t* pi = pi_all + _po;
    const int32_t* pj = pj_all + _po;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t h = (uint32_t)pi[k], j = (uint32_t)pj[k];
        if (h >= (uint64_t)n || j >= (uint64_t)n) continue;
        if (k == 0 || (uint32_t)pi[k - 1] != h) {  // row head: its row is the run of equal i
            int64_t e = k + 1;
            while (e < P && (uint32_t)pi[e] == h) ++e;
            const uint32_t L = find_root(parent, h);
            rowbeg[h] = (uint32_t)k;
            rowend[h] = (uint32_t)e;
            atomicAdd(hcount + L, 1u);
            atomicAdd(qcount + L, (uint32_t)(e - k));
        }
        claim_slot(loc, label, parent, pcount, h);
        claim_slot(loc, label, parent, pcount, j);
    }
}

// Blocks [0, kPairGrid) group heads by component (pairs in parallel); the rest
// place particle ids at their slots and list each component once (particles in
// parallel).
__global__ void __launch_bounds__(kBlock)
rps_group_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int64_t* __restrict__ npairs,
                 RpsArrays A, RpsStrides st) {
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const int64_t P = step_pairs(npairs, s, st.pair);
    if (P < 0) return;
    if (blockIdx.x < kPairGrid) {
        const int32_t* pi = pi_all + _po;
        for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)kPairGrid * kBlock) {
            const uint32_t h = (uint32_t)pi[k];
            if (h >= (uint64_t)n || !(k == 0 || (uint32_t)pi[k - 1] != h)) continue;
            const uint32_t L = label[h];
            const uint32_t slot = hstart[L] + atomicSub(hcount + L, 1u) - 1u;
            if (slot < (uint64_t)n) members[slot] = h;
        }
        return;
    }
    const int64_t nb = gridDim.x - kPairGrid;
    for (int64_t x = (int64_t)(blockIdx.x - kPairGrid) * kBlock + threadIdx.x; x - threadIdx.x < n;
         x += nb * kBlock) {
        bool to_narrow = false, to_wide = false;
        if (x < n && loc[x] != kUnset) {
            const uint32_t L = label[x];
            const uint32_t slot = pstart[L] + loc[x];
            if (slot < (uint64_t)n) ids[slot] = (uint32_t)x;
            if (L == (uint32_t)x) {
                const uint32_t m = hstart[x + 1] - hstart[x];
                const uint32_t np = qstart[x + 1] - qstart[x];
                to_narrow = m <= kNarrowHeads && np <= kNarrowPairs;
                to_wide = !to_narrow;
                pcount[x] = 0;
                qcount[x] = 0;
            }
        }
        wave_append(to_narrow, (uint32_t)x, narrow, lists + 0);
        wave_append(to_wide, (uint32_t)x, wide, lists + 1);
    }
}

// In-place ascending heapsort of a[0..m) (one lane; O(m log m)).
__device__ void heapsort_u32(uint32_t* a, uint32_t m) {
    auto sift = [&](uint32_t root, uint32_t end) {
        for (;;) {
            uint32_t c = 2 * root + 1;
            if (c >= end) return;
            if (c + 1 < end && a[c + 1] > a[c]) ++c;
            if (a[root] >= a[c]) return;
            const uint32_t t = a[root];
            a[root] = a[c];
            a[c] = t;
            root = c;
        }
    };
    for (uint32_t q = m / 2; q-- > 0;) sift(q, m);
    for (uint32_t e = m; e-- > 1;) {
        const uint32_t t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift(0, e);
    }
}

// 64-lane blocks.  Blocks [0, kCompGrid): one lane per narrow component --
// sort its <= 8 heads in registers and lay its rows out in order.  Blocks from
// kCompGrid on: one wavefront per wide component, heads sorted in LDS (heapsort
// in place beyond kSortHeads), row offsets by a wavefront scan.
__global__ void __launch_bounds__(64)
rps_order_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    __shared__ uint32_t heads[kSortHeads];
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const uint32_t lane = threadIdx.x;
    if (blockIdx.x < kCompGrid) {
        const uint32_t cnt = lists[0];
        for (uint32_t c = blockIdx.x * 64 + lane; c < cnt; c += kCompGrid * 64) {
            const uint32_t L = narrow[c];
            const uint32_t hb = hstart[L], m = min(hstart[L + 1] - hb, kNarrowHeads);
            uint32_t h[kNarrowHeads];
#pragma unroll
            for (uint32_t t = 0; t < kNarrowHeads; ++t) h[t] = t < m ? members[hb + t] : 0xffffffffu;
#pragma unroll
            for (uint32_t a = 1; a < kNarrowHeads; ++a)
#pragma unroll
                for (uint32_t b = a; b > 0; --b) {
                    const uint32_t lo = min(h[b - 1], h[b]), hi = max(h[b - 1], h[b]);
                    h[b - 1] = lo;
                    h[b] = hi;
                }
            uint32_t run = qstart[L];
#pragma unroll
            for (uint32_t t = 0; t < kNarrowHeads; ++t)
                if (t < m) {
                    rowoff[h[t]] = run;
                    run += rowend[h[t]] - rowbeg[h[t]];
                }
        }
        return;
    }
    const uint32_t cnt = lists[1];
    for (uint32_t c = blockIdx.x - kCompGrid; c < cnt; c += gridDim.x - kCompGrid) {
        const uint32_t L = wide[c];
        const uint32_t hb = hstart[L], m = min(hstart[L + 1] - hb, (uint32_t)n);
        uint32_t* hs;
        if (m <= kSortHeads) {
            uint32_t m2 = 1;
            while (m2 < m) m2 <<= 1;
            for (uint32_t t = lane; t < m2; t += 64) heads[t] = t < m ? members[hb + t] : 0xffffffffu;
            __syncthreads();
            for (uint32_t size = 2; size <= m2; size <<= 1) {
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t t = lane; t < m2; t += 64) {
                        const uint32_t u = t ^ stride;
                        if (u > t) {
                            const uint32_t x = heads[t], y = heads[u];
                            if ((x > y) == ((t & size) == 0)) {
                                heads[t] = y;
                                heads[u] = x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            hs = heads;
        } else {
            if (lane == 0) heapsort_u32(members + hb, m);
            __syncthreads();
            hs = members + hb;
        }
        // row offsets: each lane scans a contiguous run of heads
        const uint32_t per = (m + 63) / 64;
        const uint32_t t0 = min(lane * per, m), t1 = min(t0 + per, m);
        uint32_t local = 0;
        for (uint32_t t = t0; t < t1; ++t) local += rowend[hs[t]] - rowbeg[hs[t]];
        uint32_t run = qstart[L] + wave_inclusive_scan(local) - local;
        for (uint32_t t = t0; t < t1; ++t) {
            rowoff[hs[t]] = run;
            run += rowend[hs[t]] - rowbeg[hs[t]];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
rps_fill_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all,
                const int64_t* __restrict__ npairs, RpsArrays A, RpsStrides st) {
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const int64_t P = step_pairs(npairs, s, st.pair);
    const int32_t* pi = pi_all + _po;
    const int32_t* pj = pj_all + _po;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t h = (uint32_t)pi[k], j = (uint32_t)pj[k];
        if (h >= (uint64_t)n || j >= (uint64_t)n) continue;
        const uint32_t pos = rowoff[h] + (uint32_t)k - rowbeg[h];
        if (pos < (uint64_t)P) rec[pos] = make_uint2((loc[h] & kSlotMask) | ((uint32_t)bits[k] << 29), loc[j]);
    }
}

// 64-lane blocks, one wavefront per wide component: for each window of 64
// pairs, the bitmask of earlier lanes that share a particle slot.
__global__ void __launch_bounds__(64)
rps_deps_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    const int64_t s = blockIdx.y;
    RPS_STEP_VIEW(A, st, s);
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = lists[1];
    for (uint32_t c = blockIdx.x; c < cnt; c += gridDim.x) {
        const uint32_t L = wide[c];
        const uint32_t qb = qstart[L], np = qstart[L + 1] - qb;
        for (uint32_t w = 0; w < np; w += 64) {
            const bool live = w + lane < np;
            const uint2 r = live ? rec[qb + w + lane] : make_uint2(kUnset, kUnset - 1u);
            const uint32_t a = live ? (r.x & kSlotMask) : kUnset, b = live ? r.y : kUnset - 1u;
            unsigned long long d = 0ull;
            for (uint32_t e = 1; e < 64; ++e) {
                const uint32_t src = lane >= e ? lane - e : lane;
                const uint32_t a2 = __shfl(a, src, 64), b2 = __shfl(b, src, 64);
                if (live && lane >= e && (a2 == a || a2 == b || b2 == a || b2 == b)) d |= 1ull << (lane - e);
            }
            if (live) deps[qb + w + lane] = d;
        }
    }
}

// ---------------------------------------------------------------------------
// play (one step)
// ---------------------------------------------------------------------------
// Blocks [0, ncopy): particles without pairs keep their species.  The rest:
// one lane per narrow component.
__global__ void __launch_bounds__(kBlock)
rps_play_narrow_kernel(int64_t n, int64_t s, RpsArrays A, RpsStrides st, const int8_t* sp_in, int8_t* sp_out,
                       unsigned ncopy) {
    __shared__ int8_t slots[kBlock * kNarrowSlots];
    RPS_STEP_VIEW(A, st, s);
    if (blockIdx.x < ncopy) {
        // four particles per thread; only particles without pairs are written
        // here (the others belong to a component and are written by its lane)
        if (sp_in == sp_out) return;
        const int64_t x0 = 4 * ((int64_t)blockIdx.x * kBlock + threadIdx.x);
        if (x0 + 4 <= n) {
            const uint4 l4 = *reinterpret_cast<const uint4*>(loc + x0);  // rows are 16-byte aligned
            const int8_t a0 = sp_in[x0], a1 = sp_in[x0 + 1], a2 = sp_in[x0 + 2], a3 = sp_in[x0 + 3];
            if (l4.x == kUnset) sp_out[x0] = a0;
            if (l4.y == kUnset) sp_out[x0 + 1] = a1;
            if (l4.z == kUnset) sp_out[x0 + 2] = a2;
            if (l4.w == kUnset) sp_out[x0 + 3] = a3;
        } else {
            for (int64_t x = x0; x < n; ++x)
                if (loc[x] == kUnset) sp_out[x] = sp_in[x];
        }
        return;
    }
    int8_t* my = slots + threadIdx.x * kNarrowSlots;
    const uint32_t cnt = lists[0];
    const uint32_t nblk = gridDim.x - ncopy;
    for (uint32_t c = (blockIdx.x - ncopy) * kBlock + threadIdx.x; c < cnt; c += nblk * kBlock) {
        const uint32_t L = narrow[c];
        const uint32_t pb = pstart[L], psz = min(pstart[L + 1] - pb, kNarrowSlots);
        const uint32_t qb = qstart[L], np = min(qstart[L + 1] - qb, kNarrowPairs);
        uint32_t id[kNarrowSlots];
#pragma unroll
        for (uint32_t t = 0; t < kNarrowSlots; ++t) id[t] = t < psz ? ids[pb + t] : kUnset;
#pragma unroll
        for (uint32_t t = 0; t < kNarrowSlots; ++t)
            if (t < psz) my[t] = id[t] < (uint64_t)n ? sp_in[id[t]] : (int8_t)0;
        for (uint32_t q0 = 0; q0 < np; q0 += 8) {
            uint2 r[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) r[q] = q0 + q < np ? rec[qb + q0 + q] : make_uint2(0u, 0u);
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                if (q0 + q < np) {
                    const uint32_t a = min(r[q].x & kSlotMask, kNarrowSlots - 1), b = min(r[q].y, kNarrowSlots - 1);
                    int si = my[a], sj = my[b];
                    play_pair(si, sj, r[q].x >> 29);
                    my[a] = (int8_t)si;
                    my[b] = (int8_t)sj;
                }
            }
        }
#pragma unroll
        for (uint32_t t = 0; t < kNarrowSlots; ++t)
            if (t < psz && id[t] < (uint64_t)n) sp_out[id[t]] = my[t];
    }
}

// 64-lane blocks, one wavefront per wide component.
__global__ void __launch_bounds__(64)
rps_play_wide_kernel(int64_t n, int64_t s, RpsArrays A, RpsStrides st, const int8_t* sp_in, int8_t* sp_out) {
    __shared__ int8_t slots[kWideSlots];
    RPS_STEP_VIEW(A, st, s);
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = lists[1];
    for (uint32_t c = blockIdx.x; c < cnt; c += gridDim.x) {
        const uint32_t L = wide[c];
        const uint32_t pb = pstart[L], psz = pstart[L + 1] - pb;
        const uint32_t qb = qstart[L], np = qstart[L + 1] - qb;
        const bool in_lds = psz <= kWideSlots;
        // species by slot: LDS when the component fits, else sp_out itself.
        // Four slots per lane per pass with every load issued before the
        // stores, so a pass costs one memory round trip per dependent level.
        if (in_lds) {
            for (uint32_t t0 = lane; t0 < psz; t0 += 256) {
                uint32_t x[4];
                int8_t v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] = t0 + 64 * q < psz ? ids[pb + t0 + 64 * q] : kUnset;
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = x[q] < (uint64_t)n ? sp_in[x[q]] : (int8_t)0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (t0 + 64 * q < psz) slots[t0 + 64 * q] = v[q];
            }
        } else if (sp_in != sp_out) {
            for (uint32_t t = lane; t < psz; t += 64) {
                const uint32_t x = ids[pb + t];
                if (x < (uint64_t)n) sp_out[x] = sp_in[x];
            }
        }
        __syncthreads();
        uint2 r_next = lane < np ? rec[qb + lane] : make_uint2(0u, 0u);
        unsigned long long d_next = lane < np ? deps[qb + lane] : 0ull;
        for (uint32_t w = 0; w < np; w += 64) {
            const bool live = w + lane < np;
            const uint2 r = r_next;
            const unsigned long long dep = d_next;
            if (w + 64 + lane < np) {  // next window's records, while this one replays
                r_next = rec[qb + w + 64 + lane];
                d_next = deps[qb + w + 64 + lane];
            }
            const uint32_t fb = r.x >> 29;
            const uint32_t a = min(r.x & kSlotMask, psz - 1), b = min(r.y, psz - 1);  // slots are < psz
            const unsigned long long all = __ballot(live);
            unsigned long long done = 0ull;
            bool mine = live;
            while (done != all) {
                const bool ready = mine && (dep & ~done) == 0ull;
                if (ready) {
                    if (in_lds) {
                        int si = slots[a], sj = slots[b];
                        play_pair(si, sj, fb);
                        slots[a] = (int8_t)si;
                        slots[b] = (int8_t)sj;
                    } else {
                        const uint32_t xa = ids[pb + a], xb = ids[pb + b];
                        if (xa < (uint64_t)n && xb < (uint64_t)n) {
                            int si = sp_out[xa], sj = sp_out[xb];
                            play_pair(si, sj, fb);
                            sp_out[xa] = (int8_t)si;
                            sp_out[xb] = (int8_t)sj;
                        }
                    }
                    mine = false;
                }
                done |= __ballot(ready);
                __syncthreads();  // this round's writes before the next round's reads
            }
        }
        if (in_lds) {
            for (uint32_t t0 = lane; t0 < psz; t0 += 256) {
                uint32_t x[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] = t0 + 64 * q < psz ? ids[pb + t0 + 64 * q] : kUnset;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (x[q] < (uint64_t)n) sp_out[x[q]] = slots[t0 + 64 * q];
            }
        }
        __syncthreads();
    }
}

// Component size statistics of one prepared step (diagnostic, on request):
// stats = {components, max pairs in one, max particles in one, total pairs}.
__global__ void __launch_bounds__(kBlock)
rps_stats_kernel(int64_t n, int64_t s, RpsArrays A, RpsStrides st, int64_t* __restrict__ stats) {
    RPS_STEP_VIEW(A, st, s);
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t active = 0, np = 0, psz = 0;
    if (r < n && loc[r] != kUnset && label[r] == (uint32_t)r) {
        active = 1;
        np = qstart[r + 1] - qstart[r];
        psz = pstart[r + 1] - pstart[r];
    }
    uint32_t s_active = active, s_np = np, x_np = np, x_p = psz;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        s_active += __shfl_xor(s_active, d, 64);
        s_np += __shfl_xor(s_np, d, 64);
        x_np = max(x_np, __shfl_xor(x_np, d, 64));
        x_p = max(x_p, __shfl_xor(x_p, d, 64));
    }
    if ((threadIdx.x & 63) == 0 && s_active) {
        atomicAdd((unsigned long long*)&stats[0], (unsigned long long)s_active);
        atomicMax((unsigned long long*)&stats[1], (unsigned long long)x_np);
        atomicMax((unsigned long long*)&stats[2], (unsigned long long)x_p);
        atomicAdd((unsigned long long*)&stats[3], (unsigned long long)s_np);
    }
}

__global__ void zero_i64_small_kernel(int64_t* p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

}  // namespace lm

struct lm_rps_ws {
    int64_t n_max = 0;
    int32_t steps_max = 0;
    int64_t row = 0;          // per-step stride of per-particle arrays
    int64_t pair_cap = 0;     // per-step stride of per-pair arrays
    lm::RpsArrays A{};
    uint32_t* scan_tmp = nullptr;
    int64_t bytes = 0, pair_bytes = 0;
    int64_t last_n = 0;       // particles of the last prepare
    int32_t last_steps = 0;   // steps of the last prepare
    int64_t pair_stride_used = 0;  // per-step pair stride of the last prepare
    bool dirty = false;       // a call stopped midway: counters must be re-zeroed
};

template <typename T>
static int rws_alloc(T** p, int64_t count, int64_t* bytes) {
    const size_t b = (size_t)(count < 1 ? 1 : count) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, b);
    if (e != hipSuccess) {
        lm::set_error("hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        return LM_ERR_NOMEM;
    }
    *bytes += (int64_t)b;
    return LM_OK;
}

static void rps_free_pairs(lm_rps_ws* w) {
    (void)hipFree(w->A.bits);
    (void)hipFree(w->A.rec);
    (void)hipFree(w->A.deps);
    w->A.bits = nullptr;
    w->A.rec = nullptr;
    w->A.deps = nullptr;
    w->bytes -= w->pair_bytes;
    w->pair_bytes = 0;
    w->pair_cap = 0;
}

static void rps_ws_free(lm_rps_ws* w) {
    if (!w) return;
    lm::RpsArrays& A = w->A;
    for (uint32_t* p : {A.parent, A.label, A.loc, A.rowbeg, A.rowend, A.rowoff, A.hcount, A.hstart, A.pcount,
                        A.pstart, A.qcount, A.qstart, A.members, A.ids, A.narrow, A.wide, A.lists})
        (void)hipFree(p);
    rps_free_pairs(w);
    (void)hipFree(w->scan_tmp);
    delete w;
}

// Per-pair arrays sized for `cap` pairs per step (grown, never shrunk).
static int rps_ensure_pairs(lm_rps_ws* w, int64_t cap, hipStream_t s) {
    if (cap <= w->pair_cap) return LM_OK;
    LM_HIP_CHECK(hipStreamSynchronize(s));
    rps_free_pairs(w);
    const int64_t c = cap + cap / 4 + 1024, S = w->steps_max;
    int64_t b = 0;
    int rc;
    if ((rc = rws_alloc(&w->A.bits, S * c, &b)) || (rc = rws_alloc(&w->A.rec, S * c, &b)) ||
        (rc = rws_alloc(&w->A.deps, S * c, &b))) {
        w->bytes += b;
        w->pair_bytes = b;
        rps_free_pairs(w);
        return rc;
    }
    w->bytes += b;
    w->pair_bytes = b;
    w->pair_cap = c;
    return LM_OK;
}

extern "C" {

int lm_rps_ws_create(int64_t n_max, int32_t steps_max, lm_rps_ws** out) {
    LM_REQUIRE(out != nullptr, LM_ERR_ARG, "out is NULL");
    *out = nullptr;
    LM_REQUIRE(n_max >= 0 && n_max < (int64_t)0x1fffffff, LM_ERR_ARG, "n_max out of range");
    LM_REQUIRE(steps_max >= 1 && steps_max <= 65535, LM_ERR_ARG, "steps_max must be in 1..65535");
    lm_rps_ws* w = new lm_rps_ws();
    w->n_max = n_max;
    w->steps_max = steps_max;
    w->row = lm::scan_row_stride(n_max);
    const int64_t S = steps_max, R = S * w->row;
    lm::RpsArrays& A = w->A;
    int rc;
    if ((rc = rws_alloc(&A.parent, R, &w->bytes)) || (rc = rws_alloc(&A.label, R, &w->bytes)) ||
        (rc = rws_alloc(&A.loc, R, &w->bytes)) || (rc = rws_alloc(&A.rowbeg, R, &w->bytes)) ||
        (rc = rws_alloc(&A.rowend, R, &w->bytes)) || (rc = rws_alloc(&A.rowoff, R, &w->bytes)) ||
        (rc = rws_alloc(&A.hcount, R, &w->bytes)) || (rc = rws_alloc(&A.hstart, R, &w->bytes)) ||
        (rc = rws_alloc(&A.pcount, R, &w->bytes)) || (rc = rws_alloc(&A.pstart, R, &w->bytes)) ||
        (rc = rws_alloc(&A.qcount, R, &w->bytes)) || (rc = rws_alloc(&A.qstart, R, &w->bytes)) ||
        (rc = rws_alloc(&A.members, R, &w->bytes)) || (rc = rws_alloc(&A.ids, R, &w->bytes)) ||
        (rc = rws_alloc(&A.narrow, R, &w->bytes)) || (rc = rws_alloc(&A.wide, R, &w->bytes)) ||
        (rc = rws_alloc(&A.lists, 4 * S, &w->bytes)) ||
        (rc = rws_alloc(&w->scan_tmp, S * lm::scan_num_blocks(n_max), &w->bytes))) {
        rps_ws_free(w);
        return rc;
    }
    hipError_t e = hipMemset(A.hcount, 0, sizeof(uint32_t) * (size_t)R);
    if (e == hipSuccess) e = hipMemset(A.pcount, 0, sizeof(uint32_t) * (size_t)R);
    if (e == hipSuccess) e = hipMemset(A.qcount, 0, sizeof(uint32_t) * (size_t)R);
    if (e == hipSuccess) e = hipMemset(A.lists, 0, sizeof(uint32_t) * (size_t)(4 * S));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        lm::set_error("workspace init failed: %s", hipGetErrorString(e));
        rps_ws_free(w);
        return LM_ERR_HIP;
    }
    *out = w;
    return LM_OK;
}

void lm_rps_ws_destroy(lm_rps_ws* ws) { rps_ws_free(ws); }

int64_t lm_rps_ws_bytes(const lm_rps_ws* ws) { return ws ? ws->bytes : 0; }

int lm_rps_prepare(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i, const int32_t* pair_j,
                   int64_t pair_stride, const int64_t* npairs, double pRS, double pPR, double pSP,
                   uint64_t key0, uint64_t key1, uint64_t step0, void* stream) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(nsteps >= 1 && nsteps <= ws->steps_max, LM_ERR_ARG, "nsteps=%d outside 1..%d of the workspace",
               nsteps, ws->steps_max);
    LM_REQUIRE(n >= 0 && n <= ws->n_max, LM_ERR_ARG, "n=%lld exceeds the workspace (n_max=%lld)",
               (long long)n, (long long)ws->n_max);
    LM_REQUIRE(pair_stride >= 0 && pair_stride < (int64_t)0x1fffffff, LM_ERR_ARG, "pair capacity out of range");
    LM_REQUIRE(npairs != nullptr, LM_ERR_ARG, "npairs is NULL");
    LM_REQUIRE(pair_stride == 0 || (pair_i && pair_j), LM_ERR_ARG, "NULL pair buffers");
    hipStream_t s = (hipStream_t)stream;
    int rc = rps_ensure_pairs(ws, pair_stride, s);
    if (rc) return rc;
    ws->last_n = 0;
    ws->last_steps = 0;
    if (n == 0) return LM_OK;
    // The kernels address the caller's pair arrays and the workspace's per-pair
    // arrays with one per-step stride, the caller's pair_stride; the workspace
    // holds steps_max * pair_cap >= steps_max * pair_stride entries.
    LM_REQUIRE(pair_stride <= ws->pair_cap || pair_stride == 0, LM_ERR_ARG, "internal: pair capacity");
    const lm::RpsStrides sp{ws->row, pair_stride};
    const lm::RpsArrays& A = ws->A;
    const int64_t R = (int64_t)ws->steps_max * ws->row;
    if (ws->dirty) {
        LM_HIP_CHECK(hipMemsetAsync(A.hcount, 0, sizeof(uint32_t) * (size_t)R, s));
        LM_HIP_CHECK(hipMemsetAsync(A.pcount, 0, sizeof(uint32_t) * (size_t)R, s));
        LM_HIP_CHECK(hipMemsetAsync(A.qcount, 0, sizeof(uint32_t) * (size_t)R, s));
        ws->dirty = false;
    }
    ws->dirty = true;  // until every launch below went through
    const unsigned S = (unsigned)nsteps;
    LM_LAUNCH(lm::K_RPS_INIT, lm::rps_init_kernel, dim3(lm::grid_for(n), S), dim3(lm::kBlock), 0, s, n, A, sp);
    LM_LAUNCH_CHECK();
    if (pair_stride > 0) {
        LM_LAUNCH(lm::K_RPS_UNION, lm::rps_union_kernel, dim3(lm::kPairGrid, S), dim3(lm::kBlock), 0, s, n,
                  pair_i, pair_j, npairs, A, sp, pRS, pPR, pSP, key0, key1, step0);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_HEADS, lm::rps_claim_kernel, dim3(lm::kPairGrid, S), dim3(lm::kBlock), 0, s, n, pair_i,
                  pair_j, npairs, A, sp);
        LM_LAUNCH_CHECK();
        if ((rc = lm::exclusive_scan_u32_batched(A.hcount, n, ws->row, A.hstart, ws->row, nsteps, ws->scan_tmp,
                                                 nullptr, s)) ||
            (rc = lm::exclusive_scan_u32_batched(A.pcount, n, ws->row, A.pstart, ws->row, nsteps, ws->scan_tmp,
                                                 nullptr, s)) ||
            (rc = lm::exclusive_scan_u32_batched(A.qcount, n, ws->row, A.qstart, ws->row, nsteps, ws->scan_tmp,
                                                 nullptr, s)))
            return rc;
        LM_LAUNCH(lm::K_RPS_SCATTER, lm::rps_group_kernel, dim3(lm::kPairGrid + lm::grid_for(n) / 4 + 1, S),
                  dim3(lm::kBlock), 0, s, n, pair_i, npairs, A, sp);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_ORDER, lm::rps_order_kernel, dim3(lm::kCompGrid + 1024, S), dim3(64), 0, s, n, A, sp);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_FILL, lm::rps_fill_kernel, dim3(lm::kPairGrid, S), dim3(lm::kBlock), 0, s, n, pair_i,
                  pair_j, npairs, A, sp);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_DEPS, lm::rps_deps_kernel, dim3(1024, S), dim3(64), 0, s, n, A, sp);
        LM_LAUNCH_CHECK();
        ws->pair_stride_used = pair_stride;
    } else {
        ws->pair_stride_used = 0;
    }
    ws->last_n = n;
    ws->last_steps = nsteps;
    ws->dirty = false;
    return LM_OK;
}

int lm_rps_play(lm_rps_ws* ws, int32_t step, int64_t n, const int8_t* species_in, int8_t* species_out,
                void* stream) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(n == ws->last_n && step >= 0 && step < ws->last_steps, LM_ERR_ARG,
               "step %d / n %lld do not match the last prepare (%d steps, n %lld)", step, (long long)n,
               ws->last_steps, (long long)ws->last_n);
    if (n == 0) return LM_OK;
    LM_REQUIRE(species_in && species_out, LM_ERR_ARG, "NULL species");
    hipStream_t s = (hipStream_t)stream;
    lm::RpsStrides sp{ws->row, ws->pair_stride_used};
    const unsigned ncopy = lm::grid_for((n + 3) / 4);  // one thread per four particles
    LM_LAUNCH(lm::K_RPS_PLAY, lm::rps_play_narrow_kernel, dim3(ncopy + 512), dim3(lm::kBlock), 0, s, n,
              (int64_t)step, ws->A, sp, species_in, species_out, ncopy);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_PLAY_WIDE, lm::rps_play_wide_kernel, dim3(1024), dim3(64), 0, s, n, (int64_t)step, ws->A,
              sp, species_in, species_out);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_rps_update(lm_rps_ws* ws, int64_t n, int8_t* species, const int32_t* pair_i, const int32_t* pair_j,
                  const int64_t* npairs, int64_t capacity, double pRS, double pPR, double pSP,
                  uint64_t key0, uint64_t key1, uint64_t step, void* stream) {
    int rc = lm_rps_prepare(ws, 1, n, pair_i, pair_j, capacity, npairs, pRS, pPR, pSP, key0, key1, step, stream);
    if (rc) return rc;
    return lm_rps_play(ws, 0, n, species, species, stream);
}

int lm_rps_component_stats(lm_rps_ws* ws, int32_t step, int64_t* out4, void* stream) {
    LM_REQUIRE(ws != nullptr && out4 != nullptr, LM_ERR_ARG, "NULL argument");
    LM_REQUIRE(step >= 0 && (ws->last_steps == 0 || step < ws->last_steps), LM_ERR_ARG, "step out of range");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lm::zero_i64_small_kernel, dim3(1), dim3(64), 0, s, out4, 4);
    LM_LAUNCH_CHECK();
    if (ws->last_n > 0 && ws->last_steps > 0 && ws->pair_stride_used > 0) {
        lm::RpsStrides sp{ws->row, ws->pair_stride_used};
        hipLaunchKernelGGL(lm::rps_stats_kernel, dim3(lm::grid_for(ws->last_n)), dim3(lm::kBlock), 0, s,
                           ws->last_n, (int64_t)step, ws->A, sp, out4);
        LM_LAUNCH_CHECK();
    }
    return LM_OK;
}

}  // extern "C"
