// liblm core: error reporting, the device-wide scan, counter-based uniforms
// and a plain copy kernel.
#include "lm_common.h"
#include "lm_scan.h"

#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

namespace lm {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// per-kernel timing
// ---------------------------------------------------------------------------
static const char* const kKernelNames[K_NUM_KERNELS] = {
    "advect_rk4_f64", "field_blend_f64", "bin_keys", "bin_table", "pair_count", "pair_emit",
    "scan_reduce",    "scan_apply",      "rps_init",  "rps_union",   "rps_keys",   "rps_group", "rps_comps",
    "rps_play",       "rps_fill",  "rps_deps",    "rps_levels",    "philox",
    "copy16", "sort_hist", "sort_scatter", "adj_forward", "adj_recompute", "adj_backward", "adj_blend", "adj_slab", "advect_rk4_f32", "field_blend_f32", "species_count", "dist_hist",
    "rps_tile"};

struct TimedLaunch {
    int kid;
    hipEvent_t a, b;
};
static std::atomic<bool> g_timing{false};
static std::mutex g_timing_mu;
static std::vector<TimedLaunch> g_launches;  // pending, in launch order
static std::vector<hipEvent_t> g_event_pool;

bool timing_on() { return g_timing.load(std::memory_order_relaxed); }

static hipEvent_t pool_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void timing_events(int kid, hipEvent_t* start, hipEvent_t* stop) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    TimedLaunch t{kid, pool_event(), pool_event()};
    g_launches.push_back(t);
    *start = t.a;
    *stop = t.b;
}

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
// Segmented: segment g = blockIdx.y scans in[g*in_stride .. + n) into
// out[g*out_stride .. + n], out[g*out_stride + n] = total.
__global__ void __launch_bounds__(kBlock)
scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t n, int64_t in_stride,
                   uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t lds[kBlock / 64];
    in += (int64_t)blockIdx.y * in_stride;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
    if (base + kScanItems <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
        uint4 a = p[0], b = p[1];
        s = a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w;
    } else {
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) s += in[base + k];
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) block_sums[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock)
scan_apply_kernel(const uint32_t* __restrict__ in, int64_t n, int64_t in_stride,
                  const uint32_t* __restrict__ block_sums, uint32_t* __restrict__ out, int64_t out_stride,
                  int64_t* __restrict__ total64) {
    __shared__ uint32_t lds[kBlock / 64];
    in += (int64_t)blockIdx.y * in_stride;
    out += (int64_t)blockIdx.y * out_stride;
    block_sums += (int64_t)blockIdx.y * gridDim.x;
    // offset of this tile = sum of the preceding tiles' sums
    uint32_t pre = 0;
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += kBlock) pre += block_sums[b];
    pre = block_sum(pre, lds);

    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    if (base + kScanItems <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
        uint4 a = p[0], b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = (base + k < n) ? in[base + k] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) tsum += v[k];
    uint32_t tile_total;
    uint32_t run = pre + block_exclusive_scan(tsum, lds, &tile_total);
    if (base + kScanItems <= n) {
        uint32_t o[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) { o[k] = run; run += v[k]; }
        uint4* q = reinterpret_cast<uint4*>(out + base);
        // tiles start at multiples of kScanTile and strides are multiples of
        // 4, so the 16-byte stores are aligned whenever out is
        q[0] = make_uint4(o[0], o[1], o[2], o[3]);
        q[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint32_t total = pre + tile_total;
        out[n] = total;
        if (total64) total64[blockIdx.y] = (int64_t)total;
    }
}

int exclusive_scan_u32_batched(const uint32_t* in, int64_t n, int64_t in_stride, uint32_t* out, int64_t out_stride,
                               int32_t nseg, uint32_t* block_sums, int64_t* total64, hipStream_t stream) {
    const int64_t nb = scan_num_blocks(n);
    LM_REQUIRE(((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0 && (in_stride % 4) == 0 &&
                   (out_stride % 4) == 0,
               LM_ERR_ARG, "scan buffers must be 16-byte aligned");
    LM_REQUIRE(nseg >= 1 && nseg <= 65535, LM_ERR_ARG, "1..65535 scan segments");
    LM_REQUIRE(out_stride >= n + 1, LM_ERR_ARG, "scan output stride too small");
    LM_LAUNCH(K_SCAN_REDUCE, scan_reduce_kernel, dim3((unsigned)nb, (unsigned)nseg), dim3(kBlock), 0, stream, in, n,
              in_stride, block_sums);
    LM_LAUNCH_CHECK();
    LM_LAUNCH(K_SCAN_APPLY, scan_apply_kernel, dim3((unsigned)nb, (unsigned)nseg), dim3(kBlock), 0, stream, in, n,
              in_stride, block_sums, out, out_stride, total64);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int exclusive_scan_u32(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* block_sums,
                       int64_t* total64, hipStream_t stream) {
    const int64_t stride = (n + 4) & ~(int64_t)3;
    return exclusive_scan_u32_batched(in, n, stride, out, stride, 1, block_sums, total64, stream);
}

// ---------------------------------------------------------------------------
// Philox uniforms
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
philox_uniforms_kernel(int64_t count, uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                       uint64_t k0, uint64_t k1, int word, double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    philox4x64 r = philox4x64_10(c0 + (uint64_t)k, c1, c2, c3, k0, k1);
    out[k] = u01_from_u64(r.v[word]);
}

// ---------------------------------------------------------------------------
// copy
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

}  // namespace lm

extern "C" {

int lm_abi_version(void) { return LM_ABI_VERSION; }

int lm_timing_enable(int on) {
    std::lock_guard<std::mutex> lk(lm::g_timing_mu);
    for (auto& t : lm::g_launches) {
        lm::g_event_pool.push_back(t.a);
        lm::g_event_pool.push_back(t.b);
    }
    lm::g_launches.clear();
    lm::g_timing.store(on != 0);
    return LM_OK;
}

int lm_timing_summary(int64_t* counts, double* total_ms, int32_t nkinds) {
    LM_REQUIRE(counts && total_ms && nkinds >= 0, LM_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(lm::g_timing_mu);
    for (int k = 0; k < nkinds; ++k) {
        counts[k] = 0;
        total_ms[k] = 0.0;
    }
    for (auto& t : lm::g_launches) {
        LM_HIP_CHECK(hipEventSynchronize(t.b));
        float ms = 0.f;
        LM_HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
        if (t.kid < nkinds) {
            counts[t.kid] += 1;
            total_ms[t.kid] += (double)ms;
        }
    }
    return lm::K_NUM_KERNELS;
}

const char* lm_timing_kernel_name(int32_t kid) {
    return (kid >= 0 && kid < lm::K_NUM_KERNELS) ? lm::kKernelNames[kid] : "";
}

const char* lm_last_error(void) { return lm::g_err; }

int lm_philox_uniforms(int64_t count, uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                       uint64_t key0, uint64_t key1, int32_t word, double* out, void* stream) {
    LM_REQUIRE(count >= 0, LM_ERR_ARG, "count must be >= 0");
    LM_REQUIRE(word >= 0 && word < 4, LM_ERR_ARG, "word must be in 0..3");
    if (count == 0) return LM_OK;
    LM_REQUIRE(out != nullptr, LM_ERR_ARG, "out is NULL");
    LM_LAUNCH(lm::K_PHILOX, lm::philox_uniforms_kernel, dim3(lm::grid_for(count)), dim3(lm::kBlock), 0,
              (hipStream_t)stream, count, c0, c1, c2, c3, key0, key1, (int)word, out);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_copy_bytes(const void* src, void* dst, int64_t nbytes, void* stream) {
    LM_REQUIRE(nbytes >= 0 && (nbytes % 16) == 0, LM_ERR_ARG, "nbytes must be a multiple of 16");
    LM_REQUIRE(((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 15) == 0, LM_ERR_ARG,
               "copy buffers must be 16-byte aligned");
    const int64_t n16 = nbytes / 16;
    if (n16 == 0) return LM_OK;
    int64_t g = (n16 + lm::kBlock - 1) / lm::kBlock;
    if (g > 8192) g = 8192;
    LM_LAUNCH(lm::K_COPY, lm::copy16_kernel, dim3((unsigned)g), dim3(lm::kBlock), 0, (hipStream_t)stream,
              (const uint4*)src, (uint4*)dst, n16);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

}  // extern "C"
