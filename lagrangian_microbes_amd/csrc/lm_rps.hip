// Rock-paper-scissors update over sorted pair lists with the exact result of
// the sequential visit in (i, j) order, for one or many steps.
//
// A pair only reads and writes the species of its two particles, so pairs in
// different connected components of the pair graph commute; inside a component
// the sequential order is kept.
//
// Everything that depends only on the pairs runs once for a batch of steps
// (blockIdx.y = step), with the parallelism of all steps at once:
//   prepare 1. init      touched marks and rows cleared
//           1b. tile     per tile of 2^15 consecutive particles, a union-find
//                        in LDS over the pairs with both ends in the tile;
//                        parent[v] = v's tile-local root
//           2. union     lock-free union-find, hooking the larger root under the
//                        smaller, so every root is the smallest index of its
//                        component, over the pairs the tile pass left (those
//                        that cross tiles); the same pass draws each
//                        pair's uniform
//                        (Philox keyed by pair rank and step), keeps its three
//                        comparisons with pRS, pPR, pSP as bits and records
//                        where each row (the run of pairs with one i) begins
//                        and ends
//           3. keys      which particles have pairs (a row of their own, or a
//                        parent other than themselves: every j ends below itself) and
//                        each such particle's root (any other particle: its
//                        own index, a run of one that is never listed)
//           4. group     particles grouped by root, stably, in two levels (a
//                        radix pass on the root's high bits, then counting in
//                        LDS per super-bucket): each component's particles one
//                        run [pstart, pend) in index order, each particle's
//                        position in that order, and the row lengths in that
//                        order
//           5. scan      row offsets in that order.  Pairs are sorted by
//                        (i, j), so a component's pairs in sequential order are
//                        the rows of its particles in index order: the scan
//                        places every component's pairs in one run, in order
//           7. comps     per particle its slot (position minus its root's) and
//                        the shift from pair rank to run position; each
//                        component listed once, as small (<= 32 pairs) or wide
//           8. fill      pair records (slot_i | bits << 29, slot_j) at their
//                        run positions
//           9. deps      for the wide components, per 64-pair window, the
//                        bitmask of earlier pairs of the window sharing a slot
//          10. levels    for wide components of up to kLevelPairs pairs: each
//                        pair's level (longest chain of earlier pairs sharing a
//                        particle), pairs counting-sorted by level
//   No global atomics but the union's CAS and one add per block for the
//   component lists.
// Only the replay itself runs per step, in one launch of one-wavefront blocks:
//   wide     one wavefront per larger component, species in LDS: level by
//            level (a level's pairs share no particle, so they play at once),
//            or for components past the level limits window by window in
//            dependency rounds
//   small    one lane per component of <= 32 pairs, replayed in order
//   copy     particles without pairs keep their species
#include "lm_common.h"
#include "lm_scan.h"
#include "lm_sort.h"

namespace lm {

constexpr uint32_t kUnset = 0xffffffffu;
constexpr uint32_t kNarrowPairs = 32;
constexpr uint32_t kNarrowSlots = kNarrowPairs + 1;  // a component of P pairs has <= P + 1 particles
constexpr uint32_t kLevelPairs = 4096;               // components of up to this many pairs replay by levels
constexpr uint32_t kLevelSlots = kLevelPairs + 1;    // ... and so have at most this many particles
constexpr uint32_t kSlotMask = (1u << 29) - 1u;
constexpr int kCompItems = 4;                        // sorted positions per thread when listing components

// Per-step strides (elements) of a batch.
struct RpsStrides {
    int64_t row;   // per-particle arrays (>= n + 1, multiple of 4)
    int64_t pair;  // per-pair arrays (the pair capacity of a step)
};

// Workspace arrays; per step s a per-particle array starts at s * row, a
// per-pair array at s * pair.  A component is described by a uint4
// {first pair, pairs, first particle, particles} of its runs.
struct RpsArrays {
    uint32_t* parent;                     // per particle
    uint32_t *pstart, *pend;              // per root: its run of particles [pstart, pend)
    uint32_t *rcnt, *roff;                // per sorted position: row length, run offset
    uint2* info;                          // per particle: {run position - pair rank, slot}
    uint8_t* touched;                     // per particle
    uint2* rows;                          // per particle: its row [x, y) of pairs; y | touched << 31 once keyed
    uint32_t *pkey, *pval, *ptk, *ptv;    // particle sort buffers
    const uint32_t *spk, *ids;            // sorted roots / particle indices (two of the above)
    uint4 *narrow, *wide;                 // component lists (per particle)
    uint32_t* lists;                      // [step][4]: narrow count, wide count
    int64_t* plen;                        // [step]: pairs taken (0 for an overflowed step)
    uint8_t* bits;                        // per pair (by rank)
    uint2* rec;                           // per pair (run order)
    unsigned long long* deps;             // per pair (run order, wide components)
    uint2* lrec;                          // per pair: wide components' records in level order,
                                          // {slot_i, slot_j | win_table << 16}
    uint32_t* lend;                       // per pair: end of each level, from a component's first pair
    uint32_t* depth;                      // per wide component: levels, 0 = windowed replay
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
// Written without branches: with u = s - 1 in {0, 1, 2}, the pair's
// probability bit is k(ui + uj) = {1: pPR (bit 1), 2: pRS (bit 0), 3: pSP
// (bit 2)} (a like pair also lands on 2, and any outcome leaves it unchanged),
// and i beats j iff ui - uj is 1 or -2.  Species outside {1, 2, 3} never
// interact.
__device__ __forceinline__ void play_pair(int& si, int& sj, uint32_t bits) {
    const uint32_t ui = (uint32_t)(si - 1), uj = (uint32_t)(sj - 1);
    const bool valid = (ui < 3u) & (uj < 3u);
    const uint32_t k = (0x2010u >> (4u * ((ui + uj) & 7u))) & 3u;
    const uint32_t forward = (bits >> k) & 1u;
    const uint32_t beats = (0x9u >> ((ui - uj + 2u) & 7u)) & 1u;
    const bool i_wins = forward == beats;
    const int ni = valid && !i_wins ? sj : si;
    const int nj = valid && i_wins ? si : sj;
    si = ni;
    sj = nj;
}

// A pair's outcome as a table: bit 3 (si - 1) + (sj - 1), for si, sj in
// {1, 2, 3}, is set when play_pair leaves both particles with si (i wins, or
// a like pair).  Levelled records carry it (slot_j | table << 16), so that a
// row of the replay, whose chain of dependent operations sets the time of a
// level, decides by one table lookup.
__device__ __forceinline__ uint32_t win_table(uint32_t bits) {
    uint32_t t = 0;
#pragma unroll
    for (int si = 1; si <= 3; ++si)
#pragma unroll
        for (int sj = 1; sj <= 3; ++sj) {
            int a = si, b = sj;
            play_pair(a, b, bits);
            if (a == si && b == si) t |= 1u << (3 * (si - 1) + (sj - 1));
        }
    return t;
}

// Wavefront-aggregated append of v to list (one atomic per wavefront); every
// lane of the wavefront must call it.
__device__ __forceinline__ void wave_append(bool flag, uint4 v, uint4* __restrict__ list,
                                            uint32_t* __restrict__ count, uint32_t cap) {
    const uint64_t mask = __ballot(flag);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    const uint32_t at = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (flag && at < cap) list[at] = v;
}

// ---------------------------------------------------------------------------
// prepare
// ---------------------------------------------------------------------------
// Union-find by tile before the global one: the particles are cut into
// tiles of kTile consecutive indices, and a workgroup joins the pairs whose
// two ends lie in its tile in an LDS forest (a pair search's neighbours are
// spatially close and so, on a lattice start, mostly close in index).  It
// then points every particle of the tile at its tile-local root, the
// smallest index of its part of the component (the init pass leaves the
// parents alone).  The union pass joins only the pairs that cross tiles.  Every link still goes from a larger root to a smaller one, so the
// final roots are the smallest indices of the components, and every pair's j
// ends below itself (its component holds i < j), which is how the keys pass
// tells that j has pairs.
constexpr int kTileBits = 15;
constexpr uint32_t kTile = 1u << kTileBits;  // particles per tile: 128 KB of LDS parents
constexpr int kTileBlock = 1024;

__device__ __forceinline__ uint32_t lds_find(uint32_t* par, uint32_t v) {
    // relaxed workgroup-scope accesses: other lanes move pointers meanwhile,
    // only ever to smaller indices of the same component
    uint32_t p = __hip_atomic_load(par + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != v) {
        const uint32_t gp = __hip_atomic_load(par + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp != p) __hip_atomic_store(par + v, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        v = p;
        p = gp;
    }
    return v;
}

__device__ __forceinline__ void lds_unite(uint32_t* par, uint32_t a, uint32_t b) {
    for (;;) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        uint32_t expect = a;
        if (__hip_atomic_compare_exchange_strong(par + a, &expect, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
        a = expect;  // hooked meanwhile, below a
    }
}

// The first k of a step's list with pair_i[k] >= v (the list is sorted by i).
__device__ __forceinline__ uint32_t first_pair_at_or_above(const int32_t* pi, int64_t P, int64_t v) {
    int64_t lo = 0, hi = P;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)pi[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)lo;
}

// Init, and the tile pass's ranges: off[s][t] (t <= ntiles) for the first
// threads of a step (off == NULL: no pairs, no tiles).
__global__ void __launch_bounds__(kBlock)
rps_init_kernel(int64_t n, RpsArrays A, RpsStrides st, const int32_t* __restrict__ pi_all,
                const int64_t* __restrict__ npairs, uint32_t* __restrict__ off, int ntiles1) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t s = blockIdx.y, ro = s * st.row;
    if (off != nullptr && v < ntiles1) {
        int64_t P = step_pairs(npairs, s, st.pair);
        if (P < 0) P = 0;
        off[s * ntiles1 + v] = first_pair_at_or_above(pi_all + s * st.pair, P, v * (int64_t)kTile);
    }
    if (v < n) {
        A.touched[ro + v] = 0;
        A.rows[ro + v] = make_uint2(0u, 0u);
    }
    if (v < 4) A.lists[4 * s + v] = 0;
}

__global__ void __launch_bounds__(kTileBlock)
rps_tile_union_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all,
                      const int64_t* __restrict__ npairs, RpsArrays A, RpsStrides st,
                      const uint32_t* __restrict__ off, int ntiles1) {
    __shared__ uint32_t par[kTile];
    const int64_t s = blockIdx.y, ro = s * st.row, po = s * st.pair;
    const uint32_t tile = blockIdx.x, base = tile << kTileBits;
    for (uint32_t v = threadIdx.x; v < kTile; v += kTileBlock) par[v] = v;
    __syncthreads();
    const int64_t P = step_pairs(npairs, s, st.pair);
    if (P > 0) {  // an overflowed step keeps every particle apart; its replay is skipped
        const uint32_t k0 = off[s * ntiles1 + tile], k1 = off[s * ntiles1 + tile + 1];
        const int32_t* pi = pi_all + po;
        const int32_t* pj = pj_all + po;
        for (uint32_t k = k0 + threadIdx.x; k < k1; k += kTileBlock) {
            const uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
            if (a < (uint64_t)n && b < (uint64_t)n && a != b && (a >> kTileBits) == tile && (b >> kTileBits) == tile)
                lds_unite(par, a - base, b - base);
        }
    }
    __syncthreads();
    uint32_t* parent = A.parent + ro;
    for (uint32_t v = threadIdx.x; v < kTile; v += kTileBlock)
        if ((uint64_t)base + v < (uint64_t)n) parent[base + v] = base + lds_find(par, v);
}

// The tile pass joined pair k (its ends in one tile, k in that tile's range
// of the list).
__device__ __forceinline__ bool tile_joined(const uint32_t* __restrict__ off_s, uint32_t a, uint32_t b, int64_t k) {
    if ((a >> kTileBits) != (b >> kTileBits)) return false;
    const uint32_t t = a >> kTileBits;
    return (uint32_t)k >= off_s[t] && (uint32_t)k < off_s[t + 1];
}

// Runs: an ensemble of independent runs sharing the index range (run r holds
// particles [run_start[r], run_start[r+1]) and never pairs across runs).  A
// pair of run r draws its uniform as run r would alone: counter (its rank
// among run r's pairs of the step, step, 0, 0), run r's key, run r's p.
struct RunsView {
    const uint16_t* run;       // per particle
    const lm_run_params* rp;   // per run
    const uint32_t* off;       // [step][run]: first pair of the run in the step's list
    int nruns;
};

template <bool RUNS>
__global__ void __launch_bounds__(kBlock)
rps_union_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all,
                 const int64_t* __restrict__ npairs, RpsArrays A, RpsStrides st, double pRS, double pPR,
                 double pSP, uint64_t k0, uint64_t k1, uint64_t step0, RunsView R,
                 const uint32_t* __restrict__ toff, int ntiles1) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    const int64_t s = by, ro = s * st.row, po = s * st.pair;
    const int64_t P = step_pairs(npairs, s, st.pair);
    const uint32_t* off_s = toff + s * ntiles1;
    if (bx == 0 && threadIdx.x == 0) A.plen[s] = P < 0 ? 0 : P;
    const int32_t* pi = pi_all + po;
    const int32_t* pj = pj_all + po;
    uint32_t* parent = A.parent + ro;
    uint8_t* bits = A.bits + po;
    for (int64_t k = (int64_t)bx * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        // the loads this pair needs are issued together (indices clamped, no
        // branch between them) and overlap the draw below; they are pinned
        // after it, so that none sinks into the branches further down
        uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
        uint32_t before = (uint32_t)pi[k > 0 ? k - 1 : 0], after = (uint32_t)pi[k + 1 < P ? k + 1 : k];
        uint64_t rank = (uint64_t)k, q0 = k0, q1 = k1;
        double a_rs = pRS, a_pr = pPR, a_sp = pSP;
        if (RUNS) {
            const uint32_t r = a < (uint64_t)n ? (uint32_t)R.run[a] : 0u;
            const lm_run_params q = R.rp[r];
            rank = (uint64_t)(k - (int64_t)R.off[s * R.nruns + r]);
            q0 = q.key0;
            q1 = q.key1;
            a_rs = q.pRS;
            a_pr = q.pPR;
            a_sp = q.pSP;
        }
        const double u = u01_from_u64(philox4x64_10(rank, step0 + (uint64_t)s, kStreamPairs, 0, q0, q1).v[0]);
        const uint8_t fbits = (uint8_t)((u < a_rs ? 1u : 0u) | (u < a_pr ? 2u : 0u) | (u < a_sp ? 4u : 0u));
        lm_pin(a, b, before, after);
        bits[k] = fbits;
        if (a < (uint64_t)n) {
            uint32_t* row = reinterpret_cast<uint32_t*>(A.rows + ro + a);  // {begin, end}
            if (k == 0 || before != a) row[0] = (uint32_t)k;
            if (k + 1 == P || after != a) row[1] = (uint32_t)(k + 1);
            // a has its row and b ends below itself (joined to a < b), which
            // is how the keys pass tells both have pairs; only a j below its
            // i (never in a list from the pair search) is marked here
            if (b < (uint64_t)n) {
                if (b < a) A.touched[ro + b] = 1;
                if (a != b && !tile_joined(off_s, a, b, k)) unite(parent, a, b);
            }
        }
    }
}

// off[s][r] = first index k of step s's sorted list with pair_i[k] >= run_start[r]
// (binary search; the list is sorted by i and runs are index ranges).
__global__ void __launch_bounds__(kBlock)
rps_run_offsets_kernel(int nsteps, int nruns, const int64_t* __restrict__ run_start, const int32_t* __restrict__ pi_all,
                       const int64_t* __restrict__ npairs, int64_t stride, uint32_t* __restrict__ off) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)nsteps * nruns) return;
    const int64_t s = t / nruns;
    const int r = (int)(t - s * nruns);
    int64_t P = step_pairs(npairs, s, stride);
    if (P < 0) P = 0;
    const int32_t* pi = pi_all + s * stride;
    const int64_t v = run_start[r];
    int64_t lo = 0, hi = P;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)pi[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    off[t] = (uint32_t)lo;
}

// Particles per thread of the keys pass: their root walks run in lockstep,
// so a lane keeps that many dependent loads in flight instead of one.
constexpr int kKeyItems = 4;

__global__ void __launch_bounds__(kBlock)
rps_keys_kernel(int64_t n, RpsArrays A, RpsStrides st, bool write_vals) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    const int64_t s = by, ro = s * st.row;
    const int64_t v0 = (int64_t)bx * kBlock * kKeyItems + threadIdx.x;
    uint32_t x[kKeyItems], q[kKeyItems];
    bool t[kKeyItems], act[kKeyItems];
    // every item's loads are issued together: indices clamped into the step
    // (n >= 1), no branch between them
    uint32_t pv[kKeyItems], ry[kKeyItems], tv[kKeyItems];
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
        const int64_t v = v0 + (int64_t)k * kBlock;
        const int64_t vc = v < n ? v : n - 1;
        pv[k] = A.parent[ro + vc];
        ry[k] = A.rows[ro + vc].y;
        tv[k] = A.touched[ro + vc];
    }
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) lm_pin(pv[k], ry[k], tv[k]);
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
        const int64_t v = v0 + (int64_t)k * kBlock;
        const bool live = v < n;
        // a particle has pairs if it has a row (as some pair's i) or a parent
        // below it (as some pair's j: joined below itself, and parents only move down);
        // one without pairs is keyed by itself: a run of one, never listed
        const uint32_t p = live ? pv[k] : 0u;
        t[k] = live && ((p != (uint32_t)v) | (ry[k] != 0u) | (tv[k] != 0u));
        x[k] = (uint32_t)v;
        q[k] = p;
        act[k] = t[k] && p != (uint32_t)v;
    }
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
        const int64_t v = v0 + (int64_t)k * kBlock;
        if (v < n) A.touched[ro + v] = t[k];  // for the replay's copy role
    }
    // the roots by read-only walks (parents point to smaller indices, so they
    // end), all of a thread's at once: the trees are shallow after the tile pass
    // pass and the union-find, and path halving here would add scattered stores
    for (;;) {
        // q[k] is always a particle index (a parent, or 0 past the end); only
        // the walks still going load, all at once (the others' lanes are
        // masked off and cost the address unit nothing)
        uint32_t nq[kKeyItems];
#pragma unroll
        for (int k = 0; k < kKeyItems; ++k) nq[k] = act[k] ? A.parent[ro + q[k]] : q[k];
#pragma unroll
        for (int k = 0; k < kKeyItems; ++k) lm_pin(nq[k]);
        bool more = false;
#pragma unroll
        for (int k = 0; k < kKeyItems; ++k) {
            if (act[k]) {
                x[k] = q[k];
                q[k] = nq[k];
                act[k] = q[k] != x[k];
            }
            more |= act[k];
        }
        if (!more) break;
    }
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
        const int64_t v = v0 + (int64_t)k * kBlock;
        if (v >= n) continue;
        A.pkey[ro + v] = x[k];
        // the touched flag joins the row record, so the grouping gathers both in one 8-byte load
        // (row ends stay below 2^29; ry[k] is the row end this thread loaded above)
        if (t[k]) reinterpret_cast<uint32_t*>(A.rows + ro + v)[1] = ry[k] | 0x80000000u;
        if (write_vals) A.pval[ro + v] = (uint32_t)v;  // else the radix pass generates the indices
    }
}

// Second level of the grouping by root: one block of four wavefronts per
// super-bucket (the particles whose root has the same high bits, made
// contiguous, in index order, by one stable radix pass) counts its roots'
// particles in LDS and writes each root's run [pstart, pend) and the row
// lengths in that order.  Counting and the offset scan use the whole block.  The placement is
// stable -- a run keeps index order, which the pair order relies on: a pass
// takes 4 x R x 64 consecutive particles, wavefront w the w-th quarter in R
// rounds of 64; a lane's rank among equal roots in its round comes from one
// ballot per root bit, every wavefront ranking at once; then the wavefronts
// read and advance the per-root LDS cursors in turn, w = 0..3, rounds in
// order (the only serialised part), and write their particles at once.
// FB fine bits per block: 10 up to 2^20 particles (one radix pass of at most
// 10 bits covers the rest: 9 at 490k particles, whose per-tile digit
// counters cost half the 10-bit pass's; R = 2, as a super-bucket holds
// ~2^10 particles),
// 12 up to 2^22, 14 up to 2^24 (ensembles; 64 KB of counters).
constexpr uint32_t kGroupFineBits = 10;
constexpr uint32_t kGroupFineBitsWide = 12;
constexpr uint32_t kGroupFineBitsHuge = 14;
constexpr uint32_t kGroupRounds = 8;
constexpr int kGroupWaves = 4;

template <uint32_t FB, uint32_t R = kGroupRounds>
__global__ void __launch_bounds__(64 * kGroupWaves)
rps_group_blk_kernel(int64_t n, RpsArrays A, RpsStrides st, const uint32_t* __restrict__ kin,
                     const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                     const uint32_t* __restrict__ dstart, int64_t drow, int64_t dstride, uint32_t fine) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    static_assert(64 * kGroupWaves == kBlock, "the block scan assumes kBlock threads");
    constexpr uint32_t NT = 64 * kGroupWaves;
    __shared__ uint32_t cnt[1u << FB];
    __shared__ uint32_t scan_lds[kGroupWaves];
    const uint32_t d = bx, nsuper = gridDim.x, t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const int64_t s = by, ro = s * st.row;
    const uint32_t nn = (uint32_t)n;
    kin += ro;
    vin += ro;
    kout += ro;
    vout += ro;
    uint32_t a = 0, b = nn;
    if (dstart != nullptr) {
        a = min(dstart[s * drow + (int64_t)d * dstride], nn);
        b = d + 1 < nsuper ? min(dstart[s * drow + (int64_t)(d + 1) * dstride], nn) : nn;
        b = max(a, b);
    }
    const uint32_t fmask = fine - 1u;
    for (uint32_t f = t; f < fine; f += NT) cnt[f] = 0;
    __syncthreads();
    for (uint32_t c0 = a; c0 < b; c0 += NT * R) {
        uint32_t k[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t e = c0 + NT * r + t;
            k[r] = e < b ? kin[e] : 0u;
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r)
            if (c0 + NT * r + t < b) atomicAdd(&cnt[k[r] & fmask], 1u);
    }
    __syncthreads();
    // exclusive offsets: thread t owns `per` consecutive roots; runs and their ends
    const uint32_t per = (fine + NT - 1) / NT;
    const uint32_t f0 = min(t * per, fine), f1 = min(f0 + per, fine);
    uint32_t sum = 0;
    for (uint32_t f = f0; f < f1; ++f) sum += cnt[f];
    uint32_t all;
    uint32_t acc = block_exclusive_scan(sum, scan_lds, &all);
    const uint64_t base = (uint64_t)d * fine;
    for (uint32_t f = f0; f < f1; ++f) {
        const uint32_t c = cnt[f];
        cnt[f] = acc;
        if (base + f < (uint64_t)nn) {
            A.pstart[ro + base + f] = a + acc;
            A.pend[ro + base + f] = a + acc + c;
        }
        acc += c;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t c0 = a; c0 < b; c0 += NT * R) {  // uniform over the block
        const uint32_t w0 = c0 + 64u * R * wave;     // this wavefront's quarter
        uint32_t k[R], v[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t e = w0 + 64 * r + lane;
            k[r] = e < b ? kin[e] : 0u;
            v[r] = e < b ? vin[e] : kUnset;
        }
        uint8_t tch[R];
        uint32_t rb[R], re[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const bool ok = v[r] < nn;
            const uint2 q = ok ? A.rows[ro + v[r]] : make_uint2(0u, 0u);
            tch[r] = (uint8_t)(q.y >> 31);
            rb[r] = q.x;
            re[r] = q.y & 0x7fffffffu;
        }
        // ranks among equal roots within the wavefront's rounds (no LDS)
        uint32_t rank[R], tot[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const bool valid = w0 + 64 * r + lane < b;
            const uint32_t f = k[r] & fmask;
            unsigned long long m = __ballot(valid);
#pragma unroll
            for (uint32_t bit = 0; bit < FB; ++bit) {
                const bool on = (f >> bit) & 1u;
                const unsigned long long bm = __ballot(valid && on);
                m &= on ? bm : ~bm;
            }
            rank[r] = valid ? (uint32_t)__popcll(m & below) : 0u;
            tot[r] = valid ? (uint32_t)__popcll(m) : 0u;
        }
        // cursors in wavefront order, rounds in order within a wavefront
        uint32_t start[R];
        for (uint32_t w = 0; w < (uint32_t)kGroupWaves; ++w) {
            if (wave == w) {
#pragma unroll
                for (uint32_t r = 0; r < R; ++r) {
                    const bool valid = w0 + 64 * r + lane < b;
                    const uint32_t f = k[r] & fmask;
                    start[r] = valid ? cnt[f] : 0u;
                    // every lane has read the cursor before its leader advances it; the
                    // wavefront's LDS operations complete in program order
                    if (valid && rank[r] == 0) cnt[f] = start[r] + tot[r];
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const bool valid = w0 + 64 * r + lane < b;
            const uint32_t pos = a + start[r] + rank[r];
            if (valid && pos < b) {
                const uint32_t x = v[r];
                kout[pos] = k[r];
                vout[pos] = x;
                const uint32_t len = tch[r] && k[r] < nn && x < nn && re[r] > rb[r] ? re[r] - rb[r] : 0u;
                A.rcnt[ro + pos] = len;
            }
        }
    }
}

// Over the sorted particles: slot and pair shift per particle; every run start
// lists its component.  A block takes kCompItems positions per thread and
// reserves its list entries with one global add per list.
__global__ void __launch_bounds__(kBlock)
rps_comps_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    __shared__ uint32_t cnt[2], base[2];
    const int64_t s = by, ro = s * st.row;
    const uint32_t* spk = A.spk + ro;
    const uint32_t* roff = A.roff + ro;
    // run bounds are clamped to the step's pairs, so that pairs not sorted by i
    // (outside the contract) cannot send the replay past them
    const uint32_t P = (uint32_t)A.plen[s];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint4 d[kCompItems];
    uint32_t at[kCompItems];  // list (bit 31: wide) and local index, or kUnset
    // Every load below is unconditional, from an index clamped into its array
    // (results for items that turn out invalid are discarded), so the
    // dependent gathers of all items are in flight together.
    const uint32_t nn = (uint32_t)n, last = nn - 1u;
    uint32_t L[kCompItems], x[kCompItems];
    bool ok[kCompItems];
#pragma unroll
    for (int r = 0; r < kCompItems; ++r) {
        const int64_t e = ((int64_t)bx * kCompItems + r) * kBlock + threadIdx.x;
        const int64_t ec = e < n ? e : (int64_t)last;
        L[r] = spk[ec];
        x[r] = A.ids[ro + ec];
        ok[r] = e < n && L[r] < nn && x[r] < nn;
        L[r] = L[r] < nn ? L[r] : 0u;
        x[r] = x[r] < nn ? x[r] : 0u;
    }
    uint32_t tch[kCompItems], pb[kCompItems], rbg[kCompItems], ro_e[kCompItems];
#pragma unroll
    for (int r = 0; r < kCompItems; ++r) {
        const int64_t e = ((int64_t)bx * kCompItems + r) * kBlock + threadIdx.x;
        const int64_t ec = e < n ? e : (int64_t)last;
        const uint2 rw = A.rows[ro + x[r]];  // {row begin, row end | touched << 31}
        tch[r] = rw.y >> 31;
        rbg[r] = rw.x;
        pb[r] = A.pstart[ro + L[r]];
        ro_e[r] = roff[ec];
    }
#pragma unroll
    for (int r = 0; r < kCompItems; ++r) {
        const int64_t e = ((int64_t)bx * kCompItems + r) * kBlock + threadIdx.x;
        at[r] = kUnset;
        d[r] = make_uint4(0u, 0u, 0u, 0u);
        if (!ok[r] || !tch[r]) continue;  // past the end or untouched
        A.info[ro + x[r]] = make_uint2(ro_e[r] - rbg[r], (uint32_t)e - pb[r]);
        if ((uint32_t)e == pb[r]) {
            const uint32_t pe = A.pend[ro + L[r]];
            const uint32_t qb = min(ro_e[r], P), qe = min(roff[pe], P);
            d[r] = make_uint4(qb, qe > qb ? qe - qb : 0u, (uint32_t)e, pe - (uint32_t)e);
            const uint32_t wide = d[r].y > kNarrowPairs ? 1u : 0u;
            at[r] = (wide << 31) | atomicAdd(&cnt[wide], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 && cnt[threadIdx.x] > 0) base[threadIdx.x] = atomicAdd(A.lists + 4 * s + threadIdx.x, cnt[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kCompItems; ++r) {
        if (at[r] == kUnset) continue;
        const uint32_t wide = at[r] >> 31, pos = base[wide] + (at[r] & 0x7fffffffu);
        if (pos < (uint64_t)st.row) (wide ? A.wide : A.narrow)[ro + pos] = d[r];
    }
}

// Over the pairs in rank order: records at their run positions.  A pair whose
// j is not a particle index replays as a no-op (both ends on slot_i).
__global__ void __launch_bounds__(kBlock)
rps_fill_kernel(int64_t n, const int32_t* __restrict__ pi_all, const int32_t* __restrict__ pj_all, RpsArrays A,
                RpsStrides st) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    const int64_t s = by, ro = s * st.row, po = s * st.pair;
    const int64_t P = A.plen[s];
    const int32_t* pi = pi_all + po;
    const int32_t* pj = pj_all + po;
    const uint2* info = A.info + ro;
    for (int64_t k = (int64_t)bx * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        // two rounds of loads, each issued together (indices clamped, no
        // branch) and pinned so that none sinks into the branches below
        uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
        uint32_t fb = A.bits[po + k];
        lm_pin(a, b, fb);
        uint2 ia = info[a < (uint64_t)n ? a : 0u], ib = info[b < (uint64_t)n ? b : 0u];
        lm_pin(ia.x, ia.y, ib.y);
        if (a >= (uint64_t)n) continue;
        const uint32_t sb = b < (uint64_t)n ? ib.y : ia.y;
        const uint32_t pos = ia.x + (uint32_t)k;
        if (pos < (uint64_t)P) A.rec[po + pos] = make_uint2((ia.y & kSlotMask) | (fb << 29), sb);
    }
}

// 64-lane blocks, one wavefront per wide component: for each window of 64
// pairs, the bitmask of earlier lanes that share a particle slot.  For a
// component of up to kDepsTable particles, an LDS table of one 64-bit mask
// per slot: every lane ORs its bit into the masks of its two slots, then
// reads them back and keeps the earlier lanes' bits, then clears them (the
// wavefront's LDS operations complete in program order, and every window
// leaves the table zero).  Larger components compare each lane with the 63
// before it by shuffles.  Split like the levels pass below: for a batch of
// many steps, components of up to kDepsSmallPairs pairs with a table of
// kDepsSmallTable slots (2 KB of LDS, so that many blocks stay resident),
// then the larger ones; for a few steps, one launch takes every component.
constexpr int32_t kSplitSteps = 24;  // batches of at least this many steps split the deps and levels passes
enum SizeClass { kAllComponents, kSmallComponents, kLargeComponents };
constexpr uint32_t kDepsTable = 2048;
constexpr uint32_t kDepsSmallPairs = 256;  // the first launch's components: 2 KB of table, more blocks resident
constexpr uint32_t kDepsSmallTable = kDepsSmallPairs + 8;

template <uint32_t TABLE, int CLASS>
__global__ void __launch_bounds__(64)
rps_deps_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    __shared__ unsigned long long mask[TABLE];
    const int64_t s = by, ro = s * st.row, po = s * st.pair;
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = A.lists[4 * s + 1];
    if (bx >= cnt) return;  // uniform: no component for this block
    const uint2* rec = A.rec + po;
    unsigned long long* deps = A.deps + po;
    for (uint32_t t = lane; t < TABLE; t += 64) mask[t] = 0ull;
    __syncthreads();
    const unsigned long long bit = 1ull << lane;
    for (uint32_t c = bx; c < cnt; c += gridDim.x) {
        const uint4 dsc = A.wide[ro + c];
        const uint32_t qb = dsc.x, np = dsc.y, psz = dsc.w;
        if ((CLASS == kSmallComponents && np > kDepsSmallPairs) || (CLASS == kLargeComponents && np <= kDepsSmallPairs))
            continue;  // the other launch's component
        const bool table = psz <= TABLE;  // uniform over the wavefront
        if (np == 0) continue;
        // the next window's records load while this one is processed
        uint2 r_next = rec[qb + min(lane, np - 1u)];
        for (uint32_t w = 0; w < np; w += 64) {
            const bool live = w + lane < np;
            const uint2 r_cur = r_next;
            r_next = rec[qb + min(w + 64u + lane, np - 1u)];
            const uint2 r = live ? r_cur : make_uint2(kUnset, kUnset - 1u);
            const uint32_t a = live ? (r.x & kSlotMask) : kUnset, b = live ? r.y : kUnset - 1u;
            unsigned long long d = 0ull;
            if (table) {
                // slots are < psz <= TABLE; the clamp only keeps a bad record inside the table
                const uint32_t ta = min(a, TABLE - 1u), tb = min(b, TABLE - 1u);
                if (live) {
                    atomicOr(&mask[ta], bit);
                    atomicOr(&mask[tb], bit);
                }
                __builtin_amdgcn_wave_barrier();
                if (live) d = (mask[ta] | mask[tb]) & (bit - 1ull);
                __builtin_amdgcn_wave_barrier();
                if (live) {
                    mask[ta] = 0ull;
                    mask[tb] = 0ull;
                }
                __builtin_amdgcn_wave_barrier();
            } else {
                for (uint32_t e = 1; e < 64; ++e) {
                    const uint32_t src = lane >= e ? lane - e : lane;
                    const uint32_t a2 = __shfl(a, src, 64), b2 = __shfl(b, src, 64);
                    if (live && lane >= e && (a2 == a || a2 == b || b2 == a || b2 == b)) d |= 1ull << (lane - e);
                }
            }
            if (live) deps[qb + w + lane] = d;
        }
    }
}

// 64-lane blocks, one wavefront per wide component of at most kLevelPairs
// pairs: the level of every pair in the dependency order (1 + the largest
// level of an earlier pair sharing a particle; replayed window by window with
// the masks of rps_deps, so it is exactly the longest-path depth), then the
// pairs counting-sorted by level.  Pairs of one level share no particle, so a
// replay may run each level in parallel, levels in order.
//
// For a batch of many steps, two launches of one kernel: components of up
// to kLevelSmallPairs pairs with 1 KB of LDS per block, so that many blocks
// stay resident, then the larger ones with 16 KB; for a few steps, one
// launch takes every component (the larger components' long chains then
// overlap the small ones instead of following them).  The per-slot last level (uint16) is dead once the
// levels are known, so its words then hold the per-level counts and cursors
// as packed uint16 pairs (level l in the (l & 1)-th half of word l >> 1;
// every count and cursor is <= PAIRS, so an add never carries into the
// other half).
__device__ __forceinline__ uint32_t half_of(uint32_t word, uint32_t l) { return (word >> (16u * (l & 1u))) & 0xffffu; }

constexpr uint32_t kLevelSmallPairs = 256;

template <uint32_t PAIRS, int CLASS>
__global__ void __launch_bounds__(64)
rps_levels_kernel(int64_t n, RpsArrays A, RpsStrides st) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD works through a contiguous range of steps: one step's arrays in its L2
    static_assert(PAIRS < 65536u && PAIRS <= kLevelPairs, "level counts must fit 16 bits");
    __shared__ uint32_t words[(PAIRS + 2) / 2];  // last[] during the level phase, then the packed counts
    __shared__ uint16_t lvl[PAIRS];
    uint16_t* last = reinterpret_cast<uint16_t*>(words);
    const int64_t s = by, ro = s * st.row, po = s * st.pair;
    const uint32_t lane = threadIdx.x;
    const uint32_t ncomp = A.lists[4 * s + 1];
    const uint2* rec = A.rec + po;
    const unsigned long long* deps = A.deps + po;
    for (uint32_t c = bx; c < ncomp; c += gridDim.x) {
        const uint4 dsc = A.wide[ro + c];
        const uint32_t qb = dsc.x, np = dsc.y, psz = dsc.w;
        if ((CLASS == kSmallComponents && np > PAIRS) || (CLASS == kLargeComponents && np <= kLevelSmallPairs))
            continue;  // the other launch's component
        if (np > PAIRS || psz > PAIRS + 1 || psz == 0) {
            if (lane == 0) A.depth[ro + c] = 0;  // windowed replay
            continue;
        }
        for (uint32_t t = lane; t < psz; t += 64) last[t] = 0;
        __syncthreads();
        uint32_t dmax = 0;
        // the next window's records and masks load while this one resolves
        // (np >= 1: a listed component has pairs)
        uint2 r_next = rec[qb + min(lane, np - 1u)];
        unsigned long long d_next = deps[qb + min(lane, np - 1u)];
        for (uint32_t w = 0; w < np; w += 64) {
            const bool live = w + lane < np;
            const uint2 r_cur = r_next;
            const unsigned long long d_cur = d_next;
            r_next = rec[qb + min(w + 64u + lane, np - 1u)];
            d_next = deps[qb + min(w + 64u + lane, np - 1u)];
            const uint2 r = live ? r_cur : make_uint2(0u, 0u);
            const unsigned long long dep = live ? d_cur : 0ull;
            const uint32_t a = min(r.x & kSlotMask, psz - 1), b = min(r.y, psz - 1);
            const unsigned long long all = __ballot(live);
            unsigned long long done = 0ull;
            bool mine = live;
            while (done != all) {
                const bool ready = mine && (dep & ~done) == 0ull;
                if (ready) {
                    const uint32_t lv = (uint32_t)max(last[a], last[b]) + 1u;
                    last[a] = (uint16_t)lv;
                    last[b] = (uint16_t)lv;
                    lvl[w + lane] = (uint16_t)lv;
                    dmax = max(dmax, lv);
                    mine = false;
                }
                done |= __ballot(ready);
                __builtin_amdgcn_wave_barrier();
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) dmax = max(dmax, (uint32_t)__shfl_xor(dmax, d, 64));
        __syncthreads();
        // last[] is dead: its words become the per-level counts (dmax <= np <= PAIRS levels)
        for (uint32_t t = lane; t < (dmax + 1u) / 2u; t += 64) words[t] = 0u;
        __syncthreads();
        for (uint32_t q = lane; q < np; q += 64) {
            const uint32_t l = lvl[q] - 1u;
            atomicAdd(&words[l >> 1], 1u << (16u * (l & 1u)));
        }
        __syncthreads();
        // exclusive offsets per level (kept as cursors in the same halves) and level ends
        uint32_t carry = 0;
        for (uint32_t l0 = 0; l0 < dmax; l0 += 64) {
            const uint32_t l = l0 + lane;
            const bool in = l < dmax;
            const uint32_t v = in ? half_of(words[l >> 1], l) : 0u;
            const uint32_t inc = wave_inclusive_scan(v);
            __builtin_amdgcn_wave_barrier();  // every lane has read its half before any half is rewritten
            if (in) {
                last[l] = (uint16_t)(carry + inc - v);
                A.lend[po + qb + l] = carry + inc;
            }
            carry += __shfl(inc, 63, 64);
        }
        __syncthreads();
        // records into level order, four per lane at once (their loads in flight together)
        for (uint32_t q0 = lane; q0 < np; q0 += 256) {
            uint2 rv[4];
            uint32_t pos[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) rv[u] = rec[qb + min(q0 + 64u * u, np - 1u)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t q = q0 + 64u * u;
                pos[u] = kUnset;
                if (q < np) {
                    const uint32_t l = lvl[q] - 1u;
                    pos[u] = half_of(atomicAdd(&words[l >> 1], 1u << (16u * (l & 1u))), l);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (pos[u] < np)
                    A.lrec[po + qb + pos[u]] =
                        make_uint2(rv[u].x & kSlotMask, min(rv[u].y, 0xffffu) | (win_table(rv[u].x >> 29) << 16));
        }
        if (lane == 0) A.depth[ro + c] = dmax;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// play (one step)
// ---------------------------------------------------------------------------
// One launch per step, blocks of one wavefront in three roles, ordered so
// the longest work starts first:
//   [0, nwide)                 one wavefront per wide component (levels, or
//                              windows of dependency rounds past kLevelPairs)
//   [nwide, nwide + nnarrow)   one lane per small component
//   the rest                   particles without pairs keep their species
// A wavefront's LDS operations complete in program order, so no role needs
// a barrier between its LDS writes and later reads by other lanes.
constexpr uint32_t kWaveSlots = (kLevelSlots + 15u) & ~15u;   // LDS species slots of a wide component
constexpr unsigned kPlayWideBlocks = 2048;
constexpr unsigned kPlayNarrowBlocks = 2048;
constexpr int kCopyItems = 8;                                 // particles per thread in the copy role
static_assert(64 * kNarrowSlots <= kWaveSlots, "small components share the slot buffer");

__device__ __forceinline__ void play_slots(int8_t* slots, uint2 r, uint32_t top) {
    const uint32_t a = min(r.x & kSlotMask, top), b = min(r.y, top);
    int si = slots[a], sj = slots[b];
    play_pair(si, sj, r.x >> 29);
    slots[a] = (int8_t)si;
    slots[b] = (int8_t)sj;
}

// play_slots on a levelled record {slot_i, slot_j | table << 16}: the same
// result as play_pair (species outside {1, 2, 3} never interact).
__device__ __forceinline__ void play_slots_table(int8_t* slots, uint2 r, uint32_t top) {
    const uint32_t a = min(r.x, top), b = min(r.y & 0xffffu, top);
    const int si = slots[a], sj = slots[b];
    const uint32_t ui = (uint32_t)(si - 1), uj = (uint32_t)(sj - 1);
    const bool valid = (ui < 3u) & (uj < 3u);
    const uint32_t i_wins = (r.y >> (16u + ((3u * ui + uj) & 15u))) & 1u;  // read only when valid
    const int w = i_wins ? si : sj;
    slots[a] = (int8_t)(valid ? w : si);
    slots[b] = (int8_t)(valid ? w : sj);
}

// Levelled component: species in LDS by slot; the pairs of a level share no
// particle, so a level plays at once, in rows of at most 64 pairs (a wider
// level is several rows, which are independent of each other).  The records
// of kRowBatch rows go straight from memory into registers, the next batch
// loading while the current one plays, so a row costs one LDS round trip
// (its slots) on the critical path; the wavefront's LDS operations complete
// in program order, so a row reads what the rows before it wrote.
constexpr int kRowBatch = 16;

struct RowCursor {
    uint32_t l, lb, lstart, off;  // level, first level held in `ends`, level start, next row within it
};

// Row descriptors of the next kRowBatch rows (count 0 past the end) and their
// records, one per lane.  `ends` holds the ends of levels lb .. lb + 63; it is
// refilled, if needed, before any record load of the batch is issued (a batch
// advances at most kRowBatch levels), and the loads are unconditional
// (clamped), so no wait for them is placed before the batch plays.
__device__ __forceinline__ void next_rows(RowCursor& rc, uint32_t& ends, const uint2* __restrict__ lrec,
                                          const uint32_t* __restrict__ lend, uint32_t np, uint32_t depth,
                                          uint32_t (&cnt)[kRowBatch], uint2 (&rr)[kRowBatch]) {
    const uint32_t lane = threadIdx.x;
    if (rc.l - rc.lb + kRowBatch > 64) {
        rc.lb = rc.l;
        ends = rc.lb + lane < depth ? lend[rc.lb + lane] : np;
    }
    uint32_t pos[kRowBatch];
#pragma unroll
    for (int q = 0; q < kRowBatch; ++q) {
        const uint32_t at = rc.lstart + rc.off;
        const uint32_t le = max(min((uint32_t)__builtin_amdgcn_readlane((int)ends, (int)(rc.l - rc.lb)), np), at + 1u);
        const bool live = at < np;
        pos[q] = live ? at : 0u;
        cnt[q] = live ? min(64u, le - at) : 0u;
        const bool level_done = live && at + cnt[q] >= le;
        rc.off = level_done ? 0u : (live ? rc.off + 64u : rc.off);
        rc.lstart = level_done ? le : rc.lstart;
        rc.l += level_done ? 1u : 0u;
    }
#pragma unroll
    for (int q = 0; q < kRowBatch; ++q) rr[q] = lrec[min(pos[q] + lane, np - 1u)];
}

__device__ __forceinline__ void play_rows(int8_t* slots, const uint32_t (&cnt)[kRowBatch],
                                          const uint2 (&rr)[kRowBatch], uint32_t top) {
    const uint32_t lane = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kRowBatch; ++q) {
        if (lane < cnt[q]) play_slots_table(slots, rr[q], top);
        __builtin_amdgcn_wave_barrier();
    }
}

// A wavefront stages the species of particles ids[0 .. psz) (psz >= 1) into
// LDS slots, or writes the slots back: kStage particles per lane at a time,
// all their id loads in flight together, then all their species gathers (or
// stores, each from its own register), so a component of P particles costs
// ~2 P / (64 kStage) memory latencies rather than two per 64 particles.
// Particles at or past n (none in a valid list) read as 0 and are not written.
constexpr int kStage = 8;

__device__ __forceinline__ void stage_in(int64_t n, const uint32_t* __restrict__ ids, uint32_t psz,
                                         const int8_t* __restrict__ sp_in, int8_t* slots) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t t0 = 0; t0 < psz; t0 += 64u * kStage) {
        uint32_t x[kStage];
        int8_t v[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) x[q] = ids[min(t0 + 64u * q + lane, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q) v[q] = sp_in[x[q] < (uint64_t)n ? x[q] : 0u];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
            const uint32_t t = t0 + 64u * q + lane;
            if (t < psz) slots[t] = x[q] < (uint64_t)n ? v[q] : (int8_t)0;
        }
    }
}

__device__ __forceinline__ void stage_out(int64_t n, const uint32_t* __restrict__ ids, uint32_t psz,
                                          const int8_t* slots, int8_t* __restrict__ sp_out) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t t0 = 0; t0 < psz; t0 += 64u * kStage) {
        uint32_t x[kStage];
        int8_t v[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) x[q] = ids[min(t0 + 64u * q + lane, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q) v[q] = slots[min(t0 + 64u * q + lane, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q)
            if (t0 + 64u * q + lane < psz && x[q] < (uint64_t)n) sp_out[x[q]] = v[q];
    }
}

__device__ void play_levels(int64_t n, const uint32_t* __restrict__ ids, uint32_t psz,
                            const uint2* __restrict__ lrec, const uint32_t* __restrict__ lend, uint32_t np,
                            uint32_t depth, const int8_t* __restrict__ sp_in, int8_t* __restrict__ sp_out,
                            int8_t* slots) {
    const uint32_t lane = threadIdx.x;
    const uint32_t top = psz - 1;  // 1 <= psz <= kLevelSlots
    uint32_t ends = lane < depth ? lend[lane] : np;  // ends of levels lb .. lb + 63
    RowCursor rc{0u, 0u, 0u, 0u};
    uint32_t ca[kRowBatch], cb[kRowBatch];
    uint2 ra[kRowBatch], rb[kRowBatch];
    next_rows(rc, ends, lrec, lend, np, depth, ca, ra);
    stage_in(n, ids, psz, sp_in, slots);
    __builtin_amdgcn_wave_barrier();
    while (ca[0] > 0) {
        next_rows(rc, ends, lrec, lend, np, depth, cb, rb);
        play_rows(slots, ca, ra, top);
        if (cb[0] == 0) break;
        next_rows(rc, ends, lrec, lend, np, depth, ca, ra);
        play_rows(slots, cb, rb, top);
    }
    __builtin_amdgcn_wave_barrier();
    stage_out(n, ids, psz, slots, sp_out);
}

// Component past the level limits: windows of 64 pairs, each replayed in
// dependency rounds (a pair fires once every earlier pair of its window that
// shares a slot has fired); species in LDS when they fit, else in sp_out.
__device__ void play_windows(int64_t n, const uint32_t* __restrict__ ids, uint32_t psz,
                             const uint2* __restrict__ rec, const unsigned long long* __restrict__ deps,
                             uint32_t np, const int8_t* __restrict__ sp_in, int8_t* __restrict__ sp_out,
                             int8_t* slots) {
    const uint32_t lane = threadIdx.x;
    const bool in_lds = psz <= kWaveSlots;
    if (in_lds) {
        stage_in(n, ids, psz, sp_in, slots);
    } else {
        for (uint32_t t = lane; t < psz; t += 64) {
            const uint32_t x = ids[t];
            if (x < (uint64_t)n) sp_out[x] = sp_in[x];
        }
    }
    __syncthreads();  // (one wavefront) the copies above are visible to every lane
    uint2 r_next = lane < np ? rec[lane] : make_uint2(0u, 0u);
    unsigned long long d_next = lane < np ? deps[lane] : 0ull;
    for (uint32_t w = 0; w < np; w += 64) {
        const bool live = w + lane < np;
        const uint2 r = r_next;
        const unsigned long long dep = d_next;
        if (w + 64 + lane < np) {  // next window's records, while this one replays
            r_next = rec[w + 64 + lane];
            d_next = deps[w + 64 + lane];
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
                    const uint32_t xa = ids[a], xb = ids[b];
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
            if (in_lds)
                __builtin_amdgcn_wave_barrier();
            else
                __syncthreads();  // this round's global stores before the next round's loads
        }
    }
    if (in_lds) {
        __builtin_amdgcn_wave_barrier();
        stage_out(n, ids, psz, slots, sp_out);
    }
}

// Small component, one lane: its slots in LDS, its records 8 at a time with
// the next 8 loading while the current ones play.
__device__ __forceinline__ void play_narrow(int64_t n, const uint32_t* __restrict__ ids, const uint2* __restrict__ rec,
                                            uint4 dsc, const int8_t* __restrict__ sp_in,
                                            int8_t* __restrict__ sp_out, int8_t* my) {
    const uint32_t qb = dsc.x, np = min(dsc.y, kNarrowPairs);
    const uint32_t pb = dsc.z, psz = min(dsc.w, kNarrowSlots);
    if (psz == 0) return;
    // every load is clamped inside the component rather than guarded, so
    // none waits behind a branch: the species of kStage slots at a time are
    // gathered together, and the records of the next 8 pairs load while the
    // current 8 play
    uint2 r[8];
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) r[q] = np > 0 ? rec[qb + min(q, np - 1u)] : make_uint2(0u, 0u);
    for (uint32_t t0 = 0; t0 < psz; t0 += kStage) {
        uint32_t x[kStage];
        int8_t v[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) x[q] = ids[pb + min(t0 + q, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q) v[q] = sp_in[x[q] < (uint64_t)n ? x[q] : 0u];
#pragma unroll
        for (int q = 0; q < kStage; ++q)
            if (t0 + q < psz) my[t0 + q] = x[q] < (uint64_t)n ? v[q] : (int8_t)0;
    }
    for (uint32_t q0 = 0; q0 < np; q0 += 8) {
        uint2 rn[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) rn[q] = rec[qb + min(q0 + 8 + q, np - 1u)];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q)
            if (q0 + q < np) play_slots(my, r[q], kNarrowSlots - 1);
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) r[q] = rn[q];
    }
    for (uint32_t t0 = 0; t0 < psz; t0 += kStage) {
        uint32_t x[kStage];
        int8_t v[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) x[q] = ids[pb + min(t0 + q, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q) v[q] = my[min(t0 + q, psz - 1u)];
#pragma unroll
        for (int q = 0; q < kStage; ++q)
            if (t0 + q < psz && x[q] < (uint64_t)n) sp_out[x[q]] = v[q];
    }
}

__global__ void __launch_bounds__(64)
rps_play_kernel(int64_t n, int64_t s, RpsArrays A, RpsStrides st, const int8_t* sp_in, int8_t* sp_out,
                unsigned nwide, unsigned nnarrow) {
    __shared__ int8_t slots[kWaveSlots];
    const int64_t ro = s * st.row, po = s * st.pair;
    const uint32_t lane = threadIdx.x;
    const uint32_t bid = blockIdx.x;
    if (bid < nwide) {
        const uint32_t cnt = A.lists[4 * s + 1];
        for (uint32_t c = bid; c < cnt; c += nwide) {
            const uint4 dsc = A.wide[ro + c];
            const uint32_t qb = dsc.x, np = dsc.y, pb = dsc.z, psz = dsc.w;
            const uint32_t depth = A.depth[ro + c];
            if (depth > 0 && psz >= 1 && psz <= kLevelSlots)
                play_levels(n, A.ids + ro + pb, psz, A.lrec + po + qb, A.lend + po + qb, np, depth, sp_in, sp_out,
                            slots);
            else if (psz >= 1)
                play_windows(n, A.ids + ro + pb, psz, A.rec + po + qb, A.deps + po + qb, np, sp_in, sp_out, slots);
            __syncthreads();  // (one wavefront) LDS free for the next component
        }
        return;
    }
    if (bid < nwide + nnarrow) {
        const uint32_t cnt = A.lists[4 * s + 0];
        int8_t* my = slots + lane * kNarrowSlots;
        for (uint32_t c = (bid - nwide) * 64 + lane; c < cnt; c += nnarrow * 64)
            play_narrow(n, A.ids + ro, A.rec + po, A.narrow[ro + c], sp_in, sp_out, my);
        return;
    }
    // particles without pairs (touched 0) keep their species; the others
    // are written by their component
    if (sp_in == sp_out) return;
    const uint8_t* touched = A.touched + ro;
    const int64_t x0 = (int64_t)kCopyItems * ((int64_t)(bid - nwide - nnarrow) * 64 + lane);
    if (x0 + kCopyItems <= n) {
        static_assert(kCopyItems == 8, "two 4-byte words of flags");
        // step rows are multiples of 4 bytes and x0 of 8, so both words are aligned
        const uint32_t t0 = *reinterpret_cast<const uint32_t*>(touched + x0);
        const uint32_t t1 = *reinterpret_cast<const uint32_t*>(touched + x0 + 4);
        int8_t v[kCopyItems];
#pragma unroll
        for (int q = 0; q < kCopyItems; ++q) v[q] = sp_in[x0 + q];
#pragma unroll
        for (int q = 0; q < kCopyItems; ++q)
            if ((((q < 4 ? t0 : t1) >> (8 * (q & 3))) & 0xffu) == 0u) sp_out[x0 + q] = v[q];
    } else {
        for (int64_t x = x0; x < n; ++x)
            if (touched[x] == 0) sp_out[x] = sp_in[x];
    }
}

// Component size statistics of one prepared step (diagnostic, on request):
// stats = {components, max pairs in one, max particles in one, total pairs,
//          wide components, their pairs, their particles, particles with pairs}.
__global__ void __launch_bounds__(kBlock)
rps_stats_kernel(int64_t s, RpsArrays A, RpsStrides st, int64_t* __restrict__ stats) {
    const int64_t ro = s * st.row;
    const uint32_t c0 = A.lists[4 * s + 0], c1 = A.lists[4 * s + 1];
    uint64_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < c0 + c1; c += gridDim.x * kBlock) {
        const bool wide = c >= c0;
        const uint4 d = wide ? A.wide[ro + (c - c0)] : A.narrow[ro + c];
        v[0] += 1;
        v[1] = v[1] > d.y ? v[1] : d.y;
        v[2] = v[2] > d.w ? v[2] : d.w;
        v[3] += d.y;
        v[4] += wide ? 1 : 0;
        v[5] += wide ? d.y : 0;
        v[6] += wide ? d.w : 0;
        v[7] += d.w;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const uint64_t o = __shfl_xor(v[q], d, 64);
            v[q] = (q == 1 || q == 2) ? (v[q] > o ? v[q] : o) : v[q] + o;
        }
    if ((threadIdx.x & 63) == 0 && v[0]) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q == 1 || q == 2)
                atomicMax((unsigned long long*)&stats[q], (unsigned long long)v[q]);
            else
                atomicAdd((unsigned long long*)&stats[q], (unsigned long long)v[q]);
        }
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
    lm::SortScratch sort{};
    uint32_t* scan_tmp = nullptr;
    int64_t bytes = 0, pair_bytes = 0;
    int64_t last_n = 0;       // particles of the last prepare
    int32_t last_steps = 0;   // steps of the last prepare
    int64_t pair_stride_used = 0;  // per-step pair stride of the last prepare
    uint32_t* run_off = nullptr;   // [steps][runs] first pair of each run (ensembles)
    uint32_t* tile_off = nullptr;  // [steps][tiles1] first pair of each particle tile
    int tiles1 = 0;                // tiles of n_max particles, plus one
    int64_t run_off_cap = 0;
    unsigned play_wide = lm::kPlayWideBlocks;      // replay workgroups on wide components
    unsigned play_narrow = lm::kPlayNarrowBlocks;  // replay workgroups on small components (64 each)
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
    lm::RpsArrays& A = w->A;
    (void)hipFree(A.bits);
    (void)hipFree(A.rec);
    (void)hipFree(A.deps);
    (void)hipFree(A.lrec);
    (void)hipFree(A.lend);
    A.bits = nullptr;
    A.rec = nullptr;
    A.deps = nullptr;
    A.lrec = nullptr;
    A.lend = nullptr;
    w->bytes -= w->pair_bytes;
    w->pair_bytes = 0;
    w->pair_cap = 0;
}

static void rps_ws_free(lm_rps_ws* w) {
    if (!w) return;
    lm::RpsArrays& A = w->A;
    for (uint32_t* p : {A.parent, A.pstart, A.pend, A.rcnt, A.roff, A.pkey, A.pval, A.ptk, A.ptv,
                        A.lists, w->scan_tmp})
        (void)hipFree(p);
    (void)hipFree(A.info);
    (void)hipFree(A.rows);
    (void)hipFree(A.touched);
    (void)hipFree(A.narrow);
    (void)hipFree(A.wide);
    (void)hipFree(A.plen);
    (void)hipFree(A.depth);
    lm::sort_scratch_free(&w->sort);
    (void)hipFree(w->run_off);
    (void)hipFree(w->tile_off);
    rps_free_pairs(w);
    delete w;
}

// Per-pair arrays sized for `cap` pairs per step (grown, never shrunk).
static int rps_ensure_pairs(lm_rps_ws* w, int64_t cap, hipStream_t s) {
    if (cap <= w->pair_cap) return LM_OK;
    LM_HIP_CHECK(hipStreamSynchronize(s));
    rps_free_pairs(w);
    const int64_t c = cap + cap / 4 + 1024, S = w->steps_max;
    lm::RpsArrays& A = w->A;
    int64_t b = 0;
    int rc;
    if ((rc = rws_alloc(&A.bits, S * c, &b)) || (rc = rws_alloc(&A.rec, S * c, &b)) ||
        (rc = rws_alloc(&A.deps, S * c, &b)) || (rc = rws_alloc(&A.lrec, S * c, &b)) ||
        (rc = rws_alloc(&A.lend, S * c, &b))) {
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

static int key_bits_for(int64_t n) {  // bits holding every value in [0, n]
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= n) ++b;
    return b;
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
    w->tiles1 = (int)((n_max + lm::kTile - 1) / lm::kTile) + 1;
    const int64_t S = steps_max, R = S * w->row;
    lm::RpsArrays& A = w->A;
    int rc;
    if ((rc = rws_alloc(&A.parent, R, &w->bytes)) ||
        (rc = rws_alloc(&A.pend, R, &w->bytes)) || (rc = rws_alloc(&A.pstart, R, &w->bytes)) ||
        (rc = rws_alloc(&A.rcnt, R, &w->bytes)) ||
        (rc = rws_alloc(&A.roff, R, &w->bytes)) || (rc = rws_alloc(&A.info, R, &w->bytes)) ||
        (rc = rws_alloc(&A.rows, R, &w->bytes)) ||
        (rc = rws_alloc(&A.touched, R, &w->bytes)) || (rc = rws_alloc(&A.pkey, R, &w->bytes)) ||
        (rc = rws_alloc(&A.pval, R, &w->bytes)) || (rc = rws_alloc(&A.ptk, R, &w->bytes)) ||
        (rc = rws_alloc(&A.ptv, R, &w->bytes)) || (rc = rws_alloc(&A.narrow, R, &w->bytes)) ||
        (rc = rws_alloc(&A.wide, R, &w->bytes)) || (rc = rws_alloc(&A.lists, 4 * S, &w->bytes)) ||
        (rc = rws_alloc(&A.plen, S, &w->bytes)) || (rc = rws_alloc(&A.depth, R, &w->bytes)) ||
        (rc = rws_alloc(&w->scan_tmp, S * lm::scan_num_blocks(n_max), &w->bytes)) ||
        (rc = rws_alloc(&w->tile_off, S * w->tiles1, &w->bytes)) ||
        (rc = lm::sort_scratch_alloc(&w->sort, n_max, steps_max))) {
        rps_ws_free(w);
        return rc;
    }
    w->bytes += w->sort.bytes;
    hipError_t e = hipMemset(A.lists, 0, sizeof(uint32_t) * (size_t)(4 * S));
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

// The tail of a prepare, once the particles are grouped by component: row
// offsets in grouped order, the component lists, the pair records at their
// run positions, the wide components' dependency masks and levels.
static int rps_prepare_tail(lm_rps_ws* ws, int64_t n, unsigned S, const int32_t* pair_i, const int32_t* pair_j,
                            int64_t pair_stride, hipStream_t s) {
    const lm::RpsStrides sp{ws->row, pair_stride};
    lm::RpsArrays& A = ws->A;
    const unsigned nbp = lm::grid_for(n);
    const unsigned npb = lm::grid_for(pair_stride < n ? pair_stride : n);
    int rc = lm::exclusive_scan_u32_batched(A.rcnt, n, ws->row, A.roff, ws->row, (int32_t)S, ws->scan_tmp, nullptr, s);
    if (rc) return rc;
    LM_LAUNCH(lm::K_RPS_COMPS, lm::rps_comps_kernel, dim3((nbp + lm::kCompItems - 1) / lm::kCompItems, S),
              dim3(lm::kBlock), 0, s, n, A, sp);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_FILL, lm::rps_fill_kernel, dim3(npb, S), dim3(lm::kBlock), 0, s, n, pair_i, pair_j, A, sp);
    LM_LAUNCH_CHECK();
    if ((int32_t)S >= lm::kSplitSteps) {
        LM_LAUNCH(lm::K_RPS_DEPS, (lm::rps_deps_kernel<lm::kDepsSmallTable, lm::kSmallComponents>), dim3(1024, S),
                  dim3(64), 0, s, n, A, sp);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_DEPS, (lm::rps_deps_kernel<lm::kDepsTable, lm::kLargeComponents>), dim3(256, S),
                  dim3(64), 0, s, n, A, sp);
    } else {
        LM_LAUNCH(lm::K_RPS_DEPS, (lm::rps_deps_kernel<lm::kDepsTable, lm::kAllComponents>), dim3(1024, S), dim3(64),
                  0, s, n, A, sp);
    }
    LM_LAUNCH_CHECK();
    if ((int32_t)S >= lm::kSplitSteps) {
        LM_LAUNCH(lm::K_RPS_LEVELS, (lm::rps_levels_kernel<lm::kLevelSmallPairs, lm::kSmallComponents>),
                  dim3(1024, S), dim3(64), 0, s, n, A, sp);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_RPS_LEVELS, (lm::rps_levels_kernel<lm::kLevelPairs, lm::kLargeComponents>), dim3(256, S),
                  dim3(64), 0, s, n, A, sp);
    } else {
        LM_LAUNCH(lm::K_RPS_LEVELS, (lm::rps_levels_kernel<lm::kLevelPairs, lm::kAllComponents>), dim3(1024, S),
                  dim3(64), 0, s, n, A, sp);
    }
    LM_LAUNCH_CHECK();
    return LM_OK;
}

static int rps_prepare_impl(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i, const int32_t* pair_j,
                            int64_t pair_stride, const int64_t* npairs, double pRS, double pPR, double pSP,
                            uint64_t key0, uint64_t key1, uint64_t step0, const uint16_t* run, int32_t nruns,
                            const int64_t* run_start, const lm_run_params* params, void* stream) {
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
    lm::RpsArrays& A = ws->A;
    const unsigned S = (unsigned)nsteps;
    const unsigned nbp = lm::grid_for(n);
    const int ntiles = (int)((n + lm::kTile - 1) / lm::kTile);
    LM_REQUIRE(ntiles + 1 <= ws->tiles1, LM_ERR_ARG, "internal: tile offsets");
    LM_REQUIRE((int64_t)nbp * lm::kBlock >= ntiles + 1, LM_ERR_ARG, "internal: tile offsets grid");
    LM_LAUNCH(lm::K_RPS_INIT, lm::rps_init_kernel, dim3(nbp, S), dim3(lm::kBlock), 0, s, n, A, sp, pair_i, npairs,
              pair_stride > 0 ? ws->tile_off : nullptr, ntiles + 1);
    LM_LAUNCH_CHECK();
    if (pair_stride > 0) {
        // grid-stride over a step's pairs: at most one thread per particle (a
        // step has fewer pairs than particles but in dense runs; the capacity,
        // often twice the particles, would launch idle blocks)
        const unsigned npb = lm::grid_for(pair_stride < n ? pair_stride : n);
        LM_LAUNCH(lm::K_RPS_TILE, lm::rps_tile_union_kernel, dim3(ntiles, S), dim3(lm::kTileBlock), 0, s, n, pair_i,
                  pair_j, npairs, A, sp, (const uint32_t*)ws->tile_off, ntiles + 1);
        LM_LAUNCH_CHECK();
        lm::RunsView R{run, params, nullptr, nruns};
        if (run != nullptr) {
            const int64_t need = (int64_t)ws->steps_max * nruns;
            if (ws->run_off_cap < need) {
                (void)hipFree(ws->run_off);
                ws->run_off = nullptr;
                ws->run_off_cap = 0;
                if ((rc = rws_alloc(&ws->run_off, need, &ws->bytes))) return rc;
                ws->run_off_cap = need;
            }
            LM_LAUNCH(lm::K_RPS_INIT, lm::rps_run_offsets_kernel, dim3(lm::grid_for((int64_t)S * nruns)),
                      dim3(lm::kBlock), 0, s, (int)S, (int)nruns, run_start, pair_i, npairs, pair_stride, ws->run_off);
            LM_LAUNCH_CHECK();
            R.off = ws->run_off;
            LM_LAUNCH(lm::K_RPS_UNION, lm::rps_union_kernel<true>, dim3(npb, S), dim3(lm::kBlock), 0, s, n, pair_i,
                      pair_j, npairs, A, sp, pRS, pPR, pSP, key0, key1, step0, R, (const uint32_t*)ws->tile_off,
                      ntiles + 1);
        } else {
            LM_LAUNCH(lm::K_RPS_UNION, lm::rps_union_kernel<false>, dim3(npb, S), dim3(lm::kBlock), 0, s, n, pair_i,
                      pair_j, npairs, A, sp, pRS, pPR, pSP, key0, key1, step0, R, (const uint32_t*)ws->tile_off,
                      ntiles + 1);
        }
        LM_LAUNCH_CHECK();
        // grouping by root in two levels: a radix pass on the root's high
        // bits, then a block per super-bucket counts its roots in LDS
        const int kb = key_bits_for(n);
        const int fb = kb - (int)lm::kGroupFineBits <= lm::kMaxDigitBits ? (int)lm::kGroupFineBits
                       : kb - (int)lm::kGroupFineBitsWide <= lm::kMaxDigitBits ? (int)lm::kGroupFineBitsWide
                                                                               : (int)lm::kGroupFineBitsHuge;
        const int fine_bits = kb < fb ? kb : fb;
        const int super_bits = kb - fine_bits;
        LM_REQUIRE(super_bits <= lm::kMaxDigitBits, LM_ERR_ARG, "at most 2^%d particles",
                   (int)lm::kGroupFineBitsHuge + lm::kMaxDigitBits);
        const unsigned nbk = lm::grid_for((n + lm::kKeyItems - 1) / lm::kKeyItems);
        LM_LAUNCH(lm::K_RPS_KEYS, lm::rps_keys_kernel, dim3(nbk, S), dim3(lm::kBlock), 0, s, n, A, sp, super_bits == 0);
        LM_LAUNCH_CHECK();
        uint32_t* pk[2] = {A.pkey, A.ptk};
        uint32_t* pv[2] = {A.pval, A.ptv};
        const uint32_t* dstart = nullptr;
        int64_t dstride = 0;
        int in = 0;
        if (super_bits > 0) {
            if ((rc = lm::radix_partition_u32(pk, pv, ws->row, nsteps, n, fine_bits, super_bits, &ws->sort, s, &dstart,
                                              &dstride, true)))
                return rc;
            in = 1;
        }
        if (fb == (int)lm::kGroupFineBits) {
            LM_LAUNCH(lm::K_RPS_GROUP, (lm::rps_group_blk_kernel<lm::kGroupFineBits, 4>), dim3(1u << super_bits, S),
                      dim3(64 * lm::kGroupWaves), 0, s, n, A, sp, (const uint32_t*)pk[in], (const uint32_t*)pv[in],
                      pk[1 - in], pv[1 - in], dstart, ws->sort.row, dstride, 1u << fine_bits);
        } else if (fb == (int)lm::kGroupFineBitsWide) {
            LM_LAUNCH(lm::K_RPS_GROUP, lm::rps_group_blk_kernel<lm::kGroupFineBitsWide>, dim3(1u << super_bits, S),
                      dim3(64 * lm::kGroupWaves), 0, s, n, A, sp, (const uint32_t*)pk[in], (const uint32_t*)pv[in],
                      pk[1 - in], pv[1 - in], dstart, ws->sort.row, dstride, 1u << fine_bits);
        } else {
            LM_LAUNCH(lm::K_RPS_GROUP, lm::rps_group_blk_kernel<lm::kGroupFineBitsHuge>, dim3(1u << super_bits, S),
                      dim3(64 * lm::kGroupWaves), 0, s, n, A, sp, (const uint32_t*)pk[in], (const uint32_t*)pv[in],
                      pk[1 - in], pv[1 - in], dstart, ws->sort.row, dstride, 1u << fine_bits);
        }
        LM_LAUNCH_CHECK();
        A.spk = pk[1 - in];
        A.ids = pv[1 - in];
        if ((rc = rps_prepare_tail(ws, n, S, pair_i, pair_j, pair_stride, s))) return rc;
        ws->pair_stride_used = pair_stride;
        ws->last_n = n;
        ws->last_steps = nsteps;
        return LM_OK;
    }
    ws->pair_stride_used = 0;
    ws->last_n = n;
    ws->last_steps = nsteps;
    return LM_OK;
}

int lm_rps_prepare(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i, const int32_t* pair_j,
                   int64_t pair_stride, const int64_t* npairs, double pRS, double pPR, double pSP,
                   uint64_t key0, uint64_t key1, uint64_t step0, void* stream) {
    return rps_prepare_impl(ws, nsteps, n, pair_i, pair_j, pair_stride, npairs, pRS, pPR, pSP, key0, key1, step0,
                            nullptr, 0, nullptr, nullptr, stream);
}

int lm_rps_prepare_runs(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i, const int32_t* pair_j,
                        int64_t pair_stride, const int64_t* npairs, const uint16_t* run, int32_t nruns,
                        const int64_t* run_start, const lm_run_params* params, uint64_t step0, void* stream) {
    LM_REQUIRE(nruns >= 1 && nruns <= 65536, LM_ERR_ARG, "nruns must be in 1..65536");
    LM_REQUIRE(run && run_start && params, LM_ERR_ARG, "NULL run arrays");
    return rps_prepare_impl(ws, nsteps, n, pair_i, pair_j, pair_stride, npairs, 0.0, 0.0, 0.0, 0, 0, step0, run,
                            nruns, run_start, params, stream);
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
    const unsigned nwide = ws->pair_stride_used > 0 ? ws->play_wide : 0u;
    const unsigned nnarrow = ws->pair_stride_used > 0 ? ws->play_narrow : 0u;
    const unsigned ncopy = (unsigned)((n + 64 * lm::kCopyItems - 1) / (64 * lm::kCopyItems));
    LM_LAUNCH(lm::K_RPS_PLAY, lm::rps_play_kernel, dim3(nwide + nnarrow + ncopy), dim3(64), 0, s, n,
              (int64_t)step, ws->A, sp, species_in, species_out, nwide, nnarrow);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_rps_ws_reserve(lm_rps_ws* ws, int64_t pairs_per_step, void* stream) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(pairs_per_step >= 0, LM_ERR_ARG, "negative pair capacity");
    return rps_ensure_pairs(ws, pairs_per_step, (hipStream_t)stream);
}

int lm_rps_ws_set_play_blocks(lm_rps_ws* ws, int32_t wide_blocks, int32_t narrow_blocks) {
    LM_REQUIRE(ws != nullptr, LM_ERR_ARG, "workspace is NULL");
    LM_REQUIRE(wide_blocks >= 1 && wide_blocks <= 65536 && narrow_blocks >= 1 && narrow_blocks <= 65536, LM_ERR_ARG,
               "block counts must be in 1..65536");
    ws->play_wide = (unsigned)wide_blocks;
    ws->play_narrow = (unsigned)narrow_blocks;
    return LM_OK;
}

int lm_rps_update(lm_rps_ws* ws, int64_t n, int8_t* species, const int32_t* pair_i, const int32_t* pair_j,
                  const int64_t* npairs, int64_t capacity, double pRS, double pPR, double pSP,
                  uint64_t key0, uint64_t key1, uint64_t step, void* stream) {
    int rc = lm_rps_prepare(ws, 1, n, pair_i, pair_j, capacity, npairs, pRS, pPR, pSP, key0, key1, step, stream);
    if (rc) return rc;
    return lm_rps_play(ws, 0, n, species, species, stream);
}

int lm_rps_component_stats(lm_rps_ws* ws, int32_t step, int64_t* out8, void* stream) {
    LM_REQUIRE(ws != nullptr && out8 != nullptr, LM_ERR_ARG, "NULL argument");
    LM_REQUIRE(step >= 0 && (ws->last_steps == 0 || step < ws->last_steps), LM_ERR_ARG, "step out of range");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lm::zero_i64_small_kernel, dim3(1), dim3(64), 0, s, out8, 8);
    LM_LAUNCH_CHECK();
    if (ws->last_n > 0 && ws->last_steps > 0 && ws->pair_stride_used > 0) {
        lm::RpsStrides sp{ws->row, ws->pair_stride_used};
        hipLaunchKernelGGL(lm::rps_stats_kernel, dim3(256), dim3(lm::kBlock), 0, s, (int64_t)step, ws->A, sp,
                           out8);
        LM_LAUNCH_CHECK();
    }
    return LM_OK;
}

}  // extern "C"
