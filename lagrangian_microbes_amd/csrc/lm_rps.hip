// Rock-paper-scissors update over a sorted pair list with the exact result of
// the sequential visit in (i, j) order.
//
// A pair only reads and writes the species of its two particles, so pairs in
// different connected components of the pair graph commute; inside a component
// the sequential order is kept.  Pairs are sorted by (i, j), so a component's
// pairs in global order are the rows (runs of equal i) of its members taken in
// ascending i.  Kernels:
//   1. init        : parent[v] = v
//   2. union       : lock-free union-find over the pairs, always hooking the
//                    larger root under the smaller, so every root is the
//                    smallest index of its component
//   3. heads       : each row head h (first pair of a run of equal i) records its
//                    row [beg, end) and its root, and counts itself per root
//   4. scan        : component offsets over the roots
//   5. scatter     : heads grouped by root (order inside a group arbitrary)
//   6. play        : one lane per component sorts its heads ascending and
//                    replays its rows in order with a keyed uniform per pair
//                    (pair rank k, step): independent of visiting order.
// Scratch counters return to zero within a call.
#include "lm_common.h"
#include "lm_scan.h"

namespace lm {

__device__ __forceinline__ uint32_t ld_parent(const uint32_t* p, uint32_t v) {
    return __hip_atomic_load(p + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t find_root(uint32_t* parent, uint32_t v) {
    uint32_t p = ld_parent(parent, v);
    while (p != v) {
        const uint32_t gp = ld_parent(parent, p);
        if (gp != p)  // path halving: gp is an ancestor of v, so this keeps the forest valid
            __hip_atomic_store(parent + v, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        // a > b, both were roots when read: hook a under b
        const uint32_t old = atomicCAS(parent + a, a, b);
        if (old == a) return;
        a = old;  // a was hooked meanwhile; continue from its new parent
    }
}

__device__ __forceinline__ int64_t pair_total(const int64_t* npairs, int64_t capacity) {
    const int64_t P = *npairs;
    return P > capacity ? -1 : P;
}

__global__ void __launch_bounds__(kBlock)
rps_init_kernel(int64_t n, uint32_t* __restrict__ parent, int64_t* __restrict__ stats) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v < n) parent[v] = (uint32_t)v;
    if (v < 4) stats[v] = 0;
}

__global__ void __launch_bounds__(kBlock)
rps_union_kernel(int64_t n, const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
                 const int64_t* __restrict__ npairs, int64_t capacity, uint32_t* parent) {
    const int64_t P = pair_total(npairs, capacity);
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t a = (uint32_t)pi[k], b = (uint32_t)pj[k];
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

__global__ void __launch_bounds__(kBlock)
rps_scatter_kernel(int64_t n, const int32_t* __restrict__ pi, const int64_t* __restrict__ npairs,
                   int64_t capacity, const uint32_t* __restrict__ label, const uint32_t* __restrict__ cstart,
                   uint32_t* __restrict__ ccount, uint32_t* __restrict__ members) {
    const int64_t P = pair_total(npairs, capacity);
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < P; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t h = (uint32_t)pi[k];
        if (h >= (uint64_t)n) continue;
        if (k == 0 || (uint32_t)pi[k - 1] != h) {
            const uint32_t L = label[h];
            const uint32_t slot = cstart[L] + atomicSub(ccount + L, 1u) - 1u;
            if (slot < (uint64_t)n) members[slot] = h;
        }
    }
}

// Sort a[0..m) ascending in place (insertion sort for short groups, heapsort
// for long ones).
__device__ void sort_u32(uint32_t* a, uint32_t m) {
    if (m <= 32) {
        for (uint32_t t = 1; t < m; ++t) {
            const uint32_t v = a[t];
            uint32_t q = t;
            while (q > 0 && a[q - 1] > v) {
                a[q] = a[q - 1];
                --q;
            }
            a[q] = v;
        }
        return;
    }
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
    for (uint32_t e = m - 1; e > 0; --e) {
        const uint32_t t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift(0, e);
    }
}

__global__ void __launch_bounds__(kBlock)
rps_play_kernel(int64_t n, int8_t* __restrict__ species, const int32_t* __restrict__ pj,
                const int64_t* __restrict__ npairs, int64_t capacity, const uint32_t* __restrict__ cstart,
                uint32_t* __restrict__ members, const uint32_t* __restrict__ rowbeg,
                const uint32_t* __restrict__ rowend, double pRS, double pPR, double pSP, uint64_t k0,
                uint64_t k1, uint64_t step, int64_t* __restrict__ stats) {
    const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    if (pair_total(npairs, capacity) < 0) return;
    const uint32_t beg = cstart[r], end = min(cstart[r + 1], (uint32_t)n);
    if (beg >= end) return;
    uint32_t* heads = members + beg;
    const uint32_t m = end - beg;
    sort_u32(heads, m);
    uint32_t npair = 0;
    for (uint32_t t = 0; t < m; ++t) {
        const uint32_t i = heads[t];
        const uint32_t kb = rowbeg[i], ke = rowend[i];
        npair += ke - kb;
        int si = species[i];
        for (uint32_t k = kb; k < ke; ++k) {
            const uint32_t j = (uint32_t)pj[k];
            if (j >= (uint64_t)n) continue;
            const int sj = species[j];
            // Cyclic dominance: a beats b iff a == b % 3 + 1 (rock 1 > scissors 3,
            // paper 2 > rock 1, scissors 3 > paper 2).  The forward winner wins
            // with the probability of the unordered species pair.
            if (si == sj || (unsigned)(si - 1) > 2u || (unsigned)(sj - 1) > 2u) continue;
            const double prob = (si + sj == 3) ? pPR : (si + sj == 4) ? pRS : pSP;
            const philox4x64 rnd = philox4x64_10((uint64_t)k, step, kStreamPairs, 0, k0, k1);
            const bool forward = u01_from_u64(rnd.v[0]) < prob;
            const bool i_beats_j = (si == sj % 3 + 1);
            if (forward == i_beats_j) {
                species[j] = (int8_t)si;
            } else {
                species[i] = (int8_t)sj;
                si = sj;
            }
        }
    }
    atomicAdd((unsigned long long*)&stats[0], 1ull);
    atomicMax((unsigned long long*)&stats[1], (unsigned long long)npair);
    atomicMax((unsigned long long*)&stats[2], (unsigned long long)m);
    atomicAdd((unsigned long long*)&stats[3], (unsigned long long)npair);
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
    uint32_t* scan_tmp = nullptr;
    int64_t* stats = nullptr;     // [4]
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
    (void)hipFree(w->scan_tmp);
    (void)hipFree(w->stats);
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
        (rc = rws_alloc(&w->members, n_max, &w->bytes)) ||
        (rc = rws_alloc(&w->scan_tmp, lm::scan_num_blocks(n_max), &w->bytes)) ||
        (rc = rws_alloc(&w->stats, 4, &w->bytes))) {
        rps_ws_free(w);
        return rc;
    }
    hipError_t e = hipMemset(w->ccount, 0, sizeof(uint32_t) * (size_t)(n_max + 4));
    if (e == hipSuccess) e = hipMemset(w->stats, 0, sizeof(int64_t) * 4);
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
    const unsigned gn = lm::grid_for(n);
    if (ws->dirty) {
        LM_HIP_CHECK(hipMemsetAsync(ws->ccount, 0, sizeof(uint32_t) * (size_t)(ws->n_max + 4), s));
        ws->dirty = false;
    }
    ws->dirty = true;  // until every launch below went through
    LM_LAUNCH(lm::K_RPS_INIT, lm::rps_init_kernel, dim3(gn), dim3(lm::kBlock), 0, s, n, ws->parent, ws->stats);
    LM_LAUNCH_CHECK();
    if (capacity == 0) {
        ws->dirty = false;
        return LM_OK;
    }
    LM_LAUNCH(lm::K_RPS_UNION, lm::rps_union_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              pair_j, npairs, capacity, ws->parent);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_HEADS, lm::rps_heads_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              npairs, capacity, ws->parent, ws->label, ws->rowbeg, ws->rowend, ws->ccount);
    LM_LAUNCH_CHECK();
    int rc = lm::exclusive_scan_u32(ws->ccount, n, ws->cstart, ws->scan_tmp, nullptr, s);
    if (rc) return rc;
    LM_LAUNCH(lm::K_RPS_SCATTER, lm::rps_scatter_kernel, dim3(lm::kStrideGrid), dim3(lm::kBlock), 0, s, n, pair_i,
              npairs, capacity, (const uint32_t*)ws->label, (const uint32_t*)ws->cstart, ws->ccount, ws->members);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(lm::K_RPS_PLAY, lm::rps_play_kernel, dim3(gn), dim3(lm::kBlock), 0, s, n, species, pair_j, npairs,
              capacity, (const uint32_t*)ws->cstart, ws->members, (const uint32_t*)ws->rowbeg,
              (const uint32_t*)ws->rowend, pRS, pPR, pSP, key0, key1, step, ws->stats);
    LM_LAUNCH_CHECK();
    ws->dirty = false;
    return LM_OK;
}

int lm_rps_component_stats(lm_rps_ws* ws, int64_t* out4, void* stream) {
    LM_REQUIRE(ws != nullptr && out4 != nullptr, LM_ERR_ARG, "NULL argument");
    LM_HIP_CHECK(hipMemcpyAsync(out4, ws->stats, 4 * sizeof(int64_t), hipMemcpyDeviceToDevice,
                                (hipStream_t)stream));
    return LM_OK;
}

}  // extern "C"
