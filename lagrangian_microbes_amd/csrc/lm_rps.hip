// Rock-paper-scissors update over a sorted pair list with the exact result of
// the sequential visit in (i, j) order.
//
// A pair only reads and writes the species of its two particles, so pairs in
// different connected components of the pair graph commute; inside a component
// the sequential order is kept.  Pairs are sorted by (i, j), so a component's
// pairs in global order are the rows (runs of equal i) of its members taken in
// ascending i.  Kernels:
//   1. init        : parent[v] = v, counters cleared
//   2. union       : lock-free union-find over the pairs, always hooking the
//                    larger root under the smaller, so every root is the
//                    smallest index of its component; the same pass draws each
//                    pair's uniform (Philox, keyed by pair rank and step) and
//                    keeps only its three comparisons with pRS, pPR, pSP
//   3. heads       : each row head (first pair of a run of equal i) records its
//                    row [beg, end) and its root, and counts itself per root
//   4. scan        : component offsets over the roots
//   5. scatter     : heads grouped by root; components with more than
//                    kNarrowHeads heads are listed for the wavefront tiers
//   6. play_narrow : one lane per small component: sorts its heads in
//                    registers and replays row after row (a row's partner
//                    species are loaded together: partners within a row are
//                    distinct, so only the row head carries a dependency)
//   7. play_wide   : one wavefront per larger component, staged in LDS: heads
//                    sorted, pairs laid out in order, particles mapped to LDS
//                    slots through a hash table; then 64-pair windows replay in
//                    parallel rounds -- a pair fires once no earlier pending
//                    pair of the window claims one of its particles, which is
//                    exactly the sequential order.  Two instances: up to 128
//                    heads / 512 pairs, then up to 2048 heads / 4096 pairs, with
//                    a one-lane global-memory replay beyond that.
// Scratch counters return to zero within a call.
#include "lm_common.h"
#include "lm_scan.h"

namespace lm {

constexpr uint32_t kNarrowHeads = 8;
constexpr uint32_t kRowPrefetch = 16;  // partner species loaded together per row

__device__ __forceinline__ uint32_t find_root(uint32_t* parent, uint32_t v) {
    // Plain loads: a stale parent is still an ancestor (pointers only move to
    // smaller indices of the same component), and the CAS in unite() is the
    // only step that must see the current value.
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

__device__ __forceinline__ int64_t pair_total(const int64_t* npairs, int64_t capacity) {
    const int64_t P = *npairs;
    return P > capacity ? -1 : P;
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

__global__ void __launch_bounds__(kBlock)
rps_init_kernel(int64_t n, uint32_t* __restrict__ parent, uint32_t* __restrict__ lists) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) parent[v] = (uint32_t)v;
    if (v < 4) lists[v] = 0;
}

__global__ void __launch_bounds__(kBlock)
rps_union_kernel(int64_t n, const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                 const int64_t* __restrict__ npairs, int64_t capacity, uint32_t* parent,
                 uint8_t* __restrict__ bits, double pRS, double pPR, double pSP, uint64_t k0, uint64_t k1,
                 uint64_t step) {
    const int64_t P = pair_total(npairs, capacity);
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
        const double u = u01_from_u64(philox4x64_10((uint64_t)k, step, kStreamPairs, 0, k0, k1).v[0]);
        bits[k] = (uint8_t)((u < pRS ? 1u : 0u) | (u < pPR ? 2u : 0u) | (u < pSP ? 4u : 0u));
        if (a < (uint64_t)n && b < (uint64_t)n) unite(parent, a, b);
    }
}

__global__ void __launch_bounds__(kBlock)
rps_heads_kernel(int64_t n, const int32_t* __restrict__ pi, const int64_t* __restrict__ npairs,
                 int64_t capacity, uint32_t* parent, uint32_t* __restrict__ label,
                 uint32_t* __restrict__ rowbeg, uint32_t* __restrict__ rowend,
                 uint32_t* __restrict__ ccount) {
    const int64_t P = pair_total(npairs, capacity);
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t h = (uint32_t)pi[k];
        if (h >= (uint64_t)n) continue;
        if (k == 0 || (uint32_t)pi[k - 1] != h) {
            const uint32_t L = find_root(parent, h);
            label[h] = L;
            rowbeg[h] = (uint32_t)k;
            atomicAdd(ccount + L, 1u);
        }
        if (k == P - 1 || (uint32_t)pi[k + 1] != h) rowend[h] = (uint32_t)(k + 1);
    }
}

__device__ __forceinline__ void wave_append(bool flag, uint32_t v, uint32_t* __restrict__ list,
                                            uint32_t* __restrict__ count);

// lists[0] / lists[1]: number of roots in mid / large.
__global__ void __launch_bounds__(kBlock)
rps_scatter_kernel(int64_t n, const int32_t* __restrict__ pi, const int64_t* __restrict__ npairs,
                   int64_t capacity, const uint32_t* __restrict__ label, const uint32_t* __restrict__ cstart,
                   uint32_t* __restrict__ ccount, uint32_t* __restrict__ members, uint32_t* __restrict__ lists,
                   uint32_t* __restrict__ mid, uint32_t* __restrict__ large, uint32_t mid_max_heads) {
    const int64_t P = pair_total(npairs, capacity);
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t h = (uint32_t)pi[k];
        bool to_mid = false, to_large = false;
        uint32_t L = 0;
        if (h < (uint64_t)n && (k == 0 || (uint32_t)pi[k - 1] != h)) {
            L = label[h];
            const uint32_t m = cstart[L + 1] - cstart[L];
            const uint32_t old = atomicSub(ccount + L, 1u);
            const uint32_t slot = cstart[L] + old - 1u;
            if (slot < (uint64_t)n) members[slot] = h;
            if (old == m && m > kNarrowHeads) {  // the first head to arrive lists the component once
                to_mid = m <= mid_max_heads;
                to_large = !to_mid;
            }
        }
        wave_append(to_mid, L, mid, lists + 0);
        wave_append(to_large, L, large, lists + 1);
    }
}

// Component size statistics of the last update (diagnostic, run on request):
// stats = {components, max pairs in one, max heads in one, total pairs}.
// Reduced per wavefront first so that few atomics hit the four counters.
__global__ void __launch_bounds__(kBlock)
rps_stats_kernel(int64_t n, const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ members,
                 const uint32_t* __restrict__ rowbeg, const uint32_t* __restrict__ rowend,
                 int64_t* __restrict__ stats) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t active = 0, np = 0, m = 0;
    if (r < n) {
        const uint32_t beg = cstart[r], end = min(cstart[r + 1], (uint32_t)n);
        if (end > beg) {
            active = 1;
            m = end - beg;
            for (uint32_t t = beg; t < end; ++t) np += rowend[members[t]] - rowbeg[members[t]];
        }
    }
    uint32_t s_active = active, s_np = np, x_np = np, x_m = m;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        s_active += __shfl_xor(s_active, d, 64);
        s_np += __shfl_xor(s_np, d, 64);
        x_np = max(x_np, __shfl_xor(x_np, d, 64));
        x_m = max(x_m, __shfl_xor(x_m, d, 64));
    }
    if ((threadIdx.x & 63) == 0 && s_active) {
        atomicAdd((unsigned long long*)&stats[0], (unsigned long long)s_active);
        atomicMax((unsigned long long*)&stats[1], (unsigned long long)x_np);
        atomicMax((unsigned long long*)&stats[2], (unsigned long long)x_m);
        atomicAdd((unsigned long long*)&stats[3], (unsigned long long)s_np);
    }
}

__global__ void zero_i64_kernel(int64_t* p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
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

// Replay row [kb, ke) of head i from global memory, kRowPrefetch partners at
// a time (they are distinct and differ from i, so their species can be loaded
// together; only the head's species is carried from pair to pair).
__device__ __forceinline__ void replay_row(int8_t* __restrict__ species, const int32_t* __restrict__ pj,
                                           const uint8_t* __restrict__ bits, uint32_t i, uint32_t kb,
                                           uint32_t ke, int64_t n) {
    int si = species[i];
    for (uint32_t k0 = kb; k0 < ke; k0 += kRowPrefetch) {
        uint32_t j[kRowPrefetch], b[kRowPrefetch];
        int sj[kRowPrefetch];
#pragma unroll
        for (uint32_t q = 0; q < kRowPrefetch; ++q) {
            const bool live = k0 + q < ke;
            j[q] = live ? (uint32_t)pj[k0 + q] : 0xffffffffu;
            b[q] = live ? bits[k0 + q] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < kRowPrefetch; ++q) sj[q] = j[q] < (uint64_t)n ? species[j[q]] : 0;
#pragma unroll
        for (uint32_t q = 0; q < kRowPrefetch; ++q) {
            if (j[q] < (uint64_t)n) {
                const int before = sj[q];
                play_pair(si, sj[q], b[q]);
                if (sj[q] != before) species[j[q]] = (int8_t)sj[q];
            }
        }
    }
    species[i] = (int8_t)si;
}

__global__ void __launch_bounds__(kBlock)
rps_play_narrow_kernel(int64_t n, int8_t* __restrict__ species, const int32_t* __restrict__ pj,
                       const uint8_t* __restrict__ bits, const int64_t* __restrict__ npairs, int64_t capacity,
                       const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ members,
                       const uint32_t* __restrict__ rowbeg, const uint32_t* __restrict__ rowend) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    if (pair_total(npairs, capacity) < 0) return;
    const uint32_t beg = cstart[r], end = min(cstart[r + 1], (uint32_t)n);
    if (beg >= end || end - beg > kNarrowHeads) return;
    const uint32_t m = end - beg;
    uint32_t h[kNarrowHeads];
#pragma unroll
    for (uint32_t t = 0; t < kNarrowHeads; ++t) h[t] = t < m ? members[beg + t] : 0xffffffffu;
#pragma unroll
    for (uint32_t a = 1; a < kNarrowHeads; ++a)
#pragma unroll
        for (uint32_t b = a; b > 0; --b) {
            const uint32_t lo = min(h[b - 1], h[b]), hi = max(h[b - 1], h[b]);
            h[b - 1] = lo;
            h[b] = hi;
        }
#pragma unroll
    for (uint32_t t = 0; t < kNarrowHeads; ++t)
        if (t < m) replay_row(species, pj, bits, h[t], rowbeg[h[t]], rowend[h[t]], n);
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
    for (uint32_t s = m / 2; s-- > 0;) sift(s, m);
    for (uint32_t e = m; e-- > 1;) {
        const uint32_t t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift(0, e);
    }
}

// ---------------------------------------------------------------------------
// wavefront tier
// ---------------------------------------------------------------------------
constexpr uint32_t kEmpty = 0xffffffffu;

__device__ __forceinline__ uint32_t slot_hash(uint32_t id, uint32_t mask) {
    return ((id * 0x9E3779B1u) >> 7) & mask;
}

// Insert id into the LDS table (or find it); returns its slot.  The table is
// at most half full, so probing ends.
__device__ __forceinline__ uint32_t slot_of(uint32_t* tab, uint32_t mask, uint32_t id) {
    uint32_t h = slot_hash(id, mask);
    for (;;) {
        const uint32_t old = atomicCAS(tab + h, kEmpty, id);
        if (old == kEmpty || old == id) return h;
        h = (h + 1) & mask;
    }
}

template <uint32_t MAXH, uint32_t MAXP, uint32_t HC>
struct WideLds {
    unsigned long long mask[HC];  // lanes of the current window touching the slot
    uint32_t heads[MAXH];
    uint32_t off[MAXH + 1];
    uint32_t pair[MAXP];          // slot_i | slot_j << 16
    uint8_t bits[MAXP];
    uint32_t tab[HC];             // particle id per slot
    int8_t sp[HC];                // species per slot
    uint32_t total;
};

// One 64-lane workgroup per listed component.  Components beyond the caps are
// forwarded to `spill` (if given) or replayed by one lane from global memory.
template <uint32_t MAXH, uint32_t MAXP, uint32_t HC>
__global__ void __launch_bounds__(64)
rps_play_wide_kernel(int64_t n, int8_t* __restrict__ species, const int32_t* __restrict__ pj,
                     const uint8_t* __restrict__ bits, const int64_t* __restrict__ npairs, int64_t capacity,
                     const uint32_t* __restrict__ cstart, uint32_t* __restrict__ members,
                     const uint32_t* __restrict__ rowbeg, const uint32_t* __restrict__ rowend,
                     const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_n,
                     uint32_t* __restrict__ spill, uint32_t* __restrict__ spill_n) {
    static_assert(HC >= 2 * MAXP && (HC & (HC - 1)) == 0 && HC <= 65536, "hash capacity");
    static_assert((MAXH & (MAXH - 1)) == 0, "MAXH must be a power of two");
    __shared__ WideLds<MAXH, MAXP, HC> L;
    const uint32_t lane = threadIdx.x;
    if (pair_total(npairs, capacity) < 0) return;
    const uint32_t count = *list_n;
    for (uint32_t c = blockIdx.x; c < count; c += gridDim.x) {
        const uint32_t root = list[c];
        const uint32_t beg = cstart[root], end = min(cstart[root + 1], (uint32_t)n);
        const uint32_t m = end > beg ? end - beg : 0;
        if (m > MAXH) {
            if (lane == 0) {
                if (spill) {
                    spill[atomicAdd(spill_n, 1u)] = root;
                } else {
                    // beyond every LDS tier: heapsort the heads in place, replay serially
                    uint32_t* a = members + beg;
                    heapsort_u32(a, m);
                    for (uint32_t t = 0; t < m; ++t)
                        replay_row(species, pj, bits, a[t], rowbeg[a[t]], rowend[a[t]], n);
                }
            }
            __syncthreads();
            continue;
        }
        // heads -> LDS, padded to a power of two, bitonic sort
        uint32_t m2 = 1;
        while (m2 < m) m2 <<= 1;
        for (uint32_t t = lane; t < m2; t += 64) L.heads[t] = t < m ? members[beg + t] : kEmpty;
        __syncthreads();
        for (uint32_t size = 2; size <= m2; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = lane; t < m2; t += 64) {
                    const uint32_t u = t ^ stride;
                    if (u > t) {
                        const uint32_t x = L.heads[t], y = L.heads[u];
                        const bool up = (t & size) == 0;
                        if ((x > y) == up) {
                            L.heads[t] = y;
                            L.heads[u] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // row lengths -> offsets: each lane scans a contiguous run of heads,
        // the wavefront scans the run totals
        {
            const uint32_t per = (m + 63) / 64;
            const uint32_t t0 = min(lane * per, m), t1 = min(t0 + per, m);
            uint32_t local = 0;
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t len = rowend[L.heads[t]] - rowbeg[L.heads[t]];
                L.off[t] = len;
                local += len;
            }
            const uint32_t inc = wave_inclusive_scan(local);
            uint32_t run = inc - local;
            for (uint32_t t = t0; t < t1; ++t) {
                const uint32_t len = L.off[t];
                L.off[t] = run;
                run += len;
            }
            if (lane == 63) {
                L.off[m] = inc;
                L.total = inc;
            }
        }
        __syncthreads();
        const uint32_t P = L.total;
        if (P > MAXP) {
            if (lane == 0) {
                if (spill) {
                    spill[atomicAdd(spill_n, 1u)] = root;
                } else {
                    for (uint32_t t = 0; t < m; ++t) {
                        const uint32_t i = L.heads[t];
                        replay_row(species, pj, bits, i, rowbeg[i], rowend[i], n);
                    }
                }
            }
            __syncthreads();
            continue;
        }
        for (uint32_t s = lane; s < HC; s += 64) {
            L.tab[s] = kEmpty;
            L.mask[s] = 0ull;
        }
        __syncthreads();
        // pairs in order, particles mapped to slots
        for (uint32_t t = lane; t < m; t += 64) {
            const uint32_t i = L.heads[t];
            const uint32_t si = slot_of(L.tab, HC - 1, i);
            const uint32_t kb = rowbeg[i], o = L.off[t], len = L.off[t + 1] - o;
            for (uint32_t q = 0; q < len; ++q) {
                const uint32_t sj = slot_of(L.tab, HC - 1, (uint32_t)pj[kb + q]);
                L.pair[o + q] = si | sj << 16;
                L.bits[o + q] = bits[kb + q];
            }
        }
        __syncthreads();
        for (uint32_t s = lane; s < HC; s += 64) {
            const uint32_t id = L.tab[s];
            L.sp[s] = (id != kEmpty && id < (uint64_t)n) ? species[id] : (int8_t)0;
        }
        __syncthreads();
        // Windows of 64 pairs.  Each lane learns, as a bitmask, the earlier
        // lanes of its window that touch one of its particles (every lane ORs
        // its bit into the masks of its two slots, then reads them back); a
        // pair fires in the first round in which all of those are done, which
        // reproduces the sequential order exactly.
        for (uint32_t base = 0; base < P; base += 64) {
            const uint32_t k = base + lane;
            const bool live = k < P;
            const uint32_t rec = live ? L.pair[k] : 0u;
            const uint32_t a = rec & 0xffffu, b = rec >> 16;
            const uint32_t fb = live ? L.bits[k] : 0u;
            const unsigned long long me = 1ull << lane;
            if (live) {
                atomicOr(&L.mask[a], me);
                atomicOr(&L.mask[b], me);
            }
            __syncthreads();
            const unsigned long long deps = live ? ((L.mask[a] | L.mask[b]) & (me - 1ull)) : 0ull;
            __syncthreads();
            if (live) {
                L.mask[a] = 0ull;
                L.mask[b] = 0ull;
            }
            const unsigned long long all = __ballot(live);
            unsigned long long done = 0ull;
            bool mine = live;  // still to play
            while (done != all) {
                const bool ready = mine && (deps & ~done) == 0ull;
                if (ready) {
                    int si = L.sp[a], sj = L.sp[b];
                    play_pair(si, sj, fb);
                    L.sp[a] = (int8_t)si;
                    L.sp[b] = (int8_t)sj;
                    mine = false;
                }
                done |= __ballot(ready);
                __syncthreads();  // this round's species writes before the next round's reads
            }
        }
        for (uint32_t s = lane; s < HC; s += 64) {
            const uint32_t id = L.tab[s];
            if (id != kEmpty && id < (uint64_t)n) species[id] = L.sp[s];
        }
        __syncthreads();
    }
}

}  // namespace lm

struct lm_rps_ws {
    int64_t n_max = 0;
    uint32_t* parent = nullptr;   // [n_max]
    uint32_t* label = nullptr;    // [n_max]
    uint32_t* rowbeg = nullptr;   // [n_max]
    uint32_t* rowend = nullptr;   // [n_max]
    uint32_t* ccount = nullptr;   // [n_max]   zero between calls
    uint32_t* cstart = nullptr;   // [n_max + 1]
    uint32_t* members = nullptr;  // [n_max]
    uint32_t* mid = nullptr;      // [n_max]   roots of mid-size components
    uint32_t* large = nullptr;    // [n_max]   roots of large components
    uint32_t* lists = nullptr;    // [4]       root counts of mid, large
    uint32_t* scan_tmp = nullptr;
    int64_t last_n = 0;           // particles of the last update with pairs grouped
    uint8_t* bits = nullptr;      // [bits_cap] per-pair outcome bits
    int64_t bits_cap = 0;
    int64_t bytes = 0;
    bool dirty = false;           // a call stopped midway: counts must be re-zeroed
};

static void rps_ws_free(lm_rps_ws* w) {
    if (!w) return;
    (void)hipFree(w->parent);
    (void)hipFree(w->label);
    (void)hipFree(w->rowbeg);
    (void)hipFree(w->rowend);
    (void)hipFree(w->ccount);
    (void)hipFree(w->cstart);
    (void)hipFree(w->members);
    (void)hipFree(w->mid);
    (void)hipFree(w->large);
    (void)hipFree(w->lists);
    (void)hipFree(w->scan_tmp);
    (void)hipFree(w->bits);
    delete w;
}

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

namespace {
// LDS tiers of the wavefront replay
constexpr uint32_t kMidHeads = 128, kMidPairs = 512, kMidHash = 2048;
constexpr uint32_t kLargeHeads = 2048, kLargePairs = 4096, kLargeHash = 8192;
constexpr unsigned kWideGrid = 2048;
}  // namespace

extern "C" {

int lm_rps_ws_create(int64_t n_max, lm_rps_ws** out) {
    LM_REQUIRE(out != nullptr, LM_ERR_ARG, "out is NULL");
    *out = nullptr;
    LM_REQUIRE(n_max >= 0 && n_max < (int64_t)0x7fffffff, LM_ERR_ARG, "n_max out of range");
    lm_rps_ws* w = new lm_rps_ws();
    w->n_max = n_max;
    int rc;
    if ((rc = rws_alloc(&w->parent, n_max, &w->bytes)) || (rc = rws_alloc(&w->label, n_max, &w->bytes)) ||
        (rc = rws_alloc(&w->rowbeg, n_max, &w->bytes)) || (rc = rws_alloc(&w->rowend, n_max, &w->bytes)) ||
        (rc = rws_alloc(&w->ccount, n_max + 4, &w->bytes)) ||
        (rc = rws_alloc(&w->cstart, n_max + 4, &w->bytes)) ||
        (rc = rws_alloc(&w->members, n_max, &w->bytes)) || (rc = rws_alloc(&w->mid, n_max, &w->bytes)) ||
        (rc = rws_alloc(&w->large, n_max, &w->bytes)) || (rc = rws_alloc(&w->lists, 4, &w->bytes)) ||
        (rc = rws_alloc(&w->scan_tmp, lm::scan_num_blocks(n_max), &w->bytes))) {
        rps_ws_free(w);
        return rc;
    }
    hipError_t e = hipMemset(w->ccount, 0, sizeof(uint32_t) * (size_t)(n_max + 4));
    if (e == hipSuccess) e = hipMemset(w->lists, 0, sizeof(uint32_t) * 4);
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

int lm_rps_update(lm_rps_ws* ws, int64_t n, int8_t* species, const int32_t* pair_i, const int32_t* pair_j,
                  const int64_t* npairs, int64_t capacity, double pRS, double pPR, double pSP,
                  uint64_t key0, uint64_t key1, uint64_t step, void* stream) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(n >= 0 && n <= ws->n_max, LM_ERR_ARG, "n=%lld exceeds the workspace (n_max=%lld)",
               (long long)n, (long long)ws->n_max);
    LM_REQUIRE(capacity >= 0 && capacity < (int64_t)0xffffffff, LM_ERR_ARG, "capacity out of range");
    LM_REQUIRE(npairs != nullptr, LM_ERR_ARG, "npairs is NULL");
    if (n == 0) return LM_OK;
    LM_REQUIRE(species != nullptr, LM_ERR_ARG, "species is NULL");
    LM_REQUIRE(capacity == 0 || (pair_i && pair_j), LM_ERR_ARG, "NULL pair buffers");
    hipStream_t s = (hipStream_t)stream;
    if (capacity > ws->bits_cap) {  // grows with the caller's pair buffers, not per call
        LM_HIP_CHECK(hipStreamSynchronize(s));
        (void)hipFree(ws->bits);
        ws->bits = nullptr;
        ws->bits_cap = 0;
        const int64_t cap = capacity + capacity / 4 + 1024;
        int rc = rws_alloc(&ws->bits, cap, &ws->bytes);
        if (rc) return rc;
        ws->bits_cap = cap;
    }
    const unsigned gn = lm::grid_for(n);
    if (ws->dirty) {
        LM_HIP_CHECK(hipMemsetAsync(ws->ccount, 0, sizeof(uint32_t) * (size_t)(ws->n_max + 4), s));
        ws->dirty = false;
    }
    ws->dirty = true;  // until every launch below went through
    LM_LAUNCH(lm::K_RPS_INIT, lm::rps_init_kernel, dim3(gn), dim3(lm::kBlock), 0, s, n, ws->parent, ws->lists);
    LM_LAUNCH_CHECK();
    ws->last_n = 0;
    if (capacity == 0) {
        ws->dirty = false;
        return LM_OK;
    }
    LM_LAUNCH(lm::K_RPS_UNION, lm::rps_union_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              pair_j, npairs, capacity, ws->parent, ws->bits, pRS, pPR, pSP, key0, key1, step);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_HEADS, lm::rps_heads_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              npairs, capacity, ws->parent, ws->label, ws->rowbeg, ws->rowend, ws->ccount);
    LM_LAUNCH_CHECK();
    int rc = lm::exclusive_scan_u32(ws->ccount, n, ws->cstart, ws->scan_tmp, nullptr, s);
    if (rc) return rc;
    LM_LAUNCH(lm::K_RPS_SCATTER, lm::rps_scatter_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              npairs, capacity, (const uint32_t*)ws->label, (const uint32_t*)ws->cstart, ws->ccount, ws->members,
              ws->lists, ws->mid, ws->large, kMidHeads);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_PLAY, lm::rps_play_narrow_kernel, dim3(gn), dim3(lm::kBlock), 0, s, n, species, pair_j,
              (const uint8_t*)ws->bits, npairs, capacity, (const uint32_t*)ws->cstart,
              (const uint32_t*)ws->members, (const uint32_t*)ws->rowbeg, (const uint32_t*)ws->rowend);
    LM_LAUNCH_CHECK();
    // the mid tier forwards what does not fit its LDS to the large list, which runs next
    LM_LAUNCH(lm::K_RPS_PLAY_WIDE, (lm::rps_play_wide_kernel<kMidHeads, kMidPairs, kMidHash>), dim3(kWideGrid),
              dim3(64), 0, s, n, species, pair_j, (const uint8_t*)ws->bits, npairs, capacity,
              (const uint32_t*)ws->cstart, ws->members, (const uint32_t*)ws->rowbeg, (const uint32_t*)ws->rowend,
              (const uint32_t*)ws->mid, (const uint32_t*)(ws->lists + 0), ws->large, ws->lists + 1);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_PLAY_LARGE, (lm::rps_play_wide_kernel<kLargeHeads, kLargePairs, kLargeHash>),
              dim3(kWideGrid / 4), dim3(64), 0, s, n, species, pair_j, (const uint8_t*)ws->bits, npairs, capacity,
              (const uint32_t*)ws->cstart, ws->members, (const uint32_t*)ws->rowbeg, (const uint32_t*)ws->rowend,
              (const uint32_t*)ws->large, (const uint32_t*)(ws->lists + 1), (uint32_t*)nullptr, (uint32_t*)nullptr);
    LM_LAUNCH_CHECK();
    ws->last_n = n;
    ws->dirty = false;
    return LM_OK;
}

int lm_rps_component_stats(lm_rps_ws* ws, int64_t* out4, void* stream) {
    LM_REQUIRE(ws != nullptr && out4 != nullptr, LM_ERR_ARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lm::zero_i64_kernel, dim3(1), dim3(64), 0, s, out4, 4);
    LM_LAUNCH_CHECK();
    if (ws->last_n > 0) {
        hipLaunchKernelGGL(lm::rps_stats_kernel, dim3(lm::grid_for(ws->last_n)), dim3(lm::kBlock), 0, s, ws->last_n,
                           (const uint32_t*)ws->cstart, (const uint32_t*)ws->members, (const uint32_t*)ws->rowbeg,
                           (const uint32_t*)ws->rowend, out4);
        LM_LAUNCH_CHECK();
    }
    return LM_OK;
}

}  // extern "C"
