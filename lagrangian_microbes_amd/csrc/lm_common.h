// Shared device/host helpers for liblm: error reporting, Philox4x64-10,
// the degree cosine used by the velocity unit conversion, and launch sizing.
//
// Every translation unit is compiled with -ffp-contract=off: each fused
// multiply-add below is written out with fma() so that the CPU restatement in
// oracle/lm_oracle.c, which performs the same IEEE operations in the same
// order, produces bit-identical doubles.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>

#include "../../include/lm.h"

namespace lm {

// ----------------------------------------------------------------------------
// error plumbing (thread-local message; no C++ exception crosses the ABI)
// ----------------------------------------------------------------------------
void set_error(const char* fmt, ...);

#define LM_HIP_CHECK(expr)                                                        \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::lm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                  \
            return LM_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define LM_LAUNCH_CHECK()                                                         \
    do {                                                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) {                                                   \
            ::lm::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                  \
            return LM_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define LM_REQUIRE(cond, code, ...)                                               \
    do {                                                                          \
        if (!(cond)) {                                                            \
            ::lm::set_error(__VA_ARGS__);                                         \
            return (code);                                                        \
        }                                                                         \
    } while (0)

// ----------------------------------------------------------------------------
// optional per-kernel timing (lm_timing_enable): launches carry start/stop
// events on their own dispatch packet through hipExtLaunchKernelGGL
// ----------------------------------------------------------------------------
enum KernelId {
    K_ADVECT = 0,
    K_BLEND,
    K_BIN_KEYS,
    K_BIN_TABLE,
    K_PAIR_COUNT,
    K_PAIR_EMIT,
    K_SCAN_REDUCE,
    K_SCAN_APPLY,
    K_RPS_INIT,
    K_RPS_UNION,
    K_RPS_KEYS,
    K_RPS_GROUP,
    K_RPS_COMPS,
    K_RPS_PLAY,
    K_RPS_FILL,
    K_RPS_DEPS,
    K_RPS_LEVELS,
    K_PHILOX,
    K_COPY,
    K_SORT_HIST,
    K_SORT_SCATTER,
    K_ADJ_FORWARD,
    K_ADJ_RECOMPUTE,
    K_ADJ_BACKWARD,
    K_ADJ_BLEND,
    K_ADJ_SLAB,
    K_ADVECT_F32,
    K_BLEND_F32,
    K_SPECIES_COUNT,
    K_DIST_HIST,
    K_RPS_TILE,
    K_NUM_KERNELS
};
bool timing_on();
// Two events to bracket one launch of kernel `kid` (only when timing_on()).
void timing_events(int kid, hipEvent_t* start, hipEvent_t* stop);

#define LM_LAUNCH(kid, kernel, grid, block, shmem, stream, ...)                              \
    do {                                                                                       \
        if (::lm::timing_on()) {                                                               \
            hipEvent_t _lm_a, _lm_b;                                                           \
            ::lm::timing_events((kid), &_lm_a, &_lm_b);                                        \
            hipExtLaunchKernelGGL((kernel), (grid), (block), (shmem), (stream), _lm_a, _lm_b, 0, \
                                  __VA_ARGS__);                                                \
        } else {                                                                               \
            hipLaunchKernelGGL((kernel), (grid), (block), (shmem), (stream), __VA_ARGS__);     \
        }                                                                                      \
    } while (0)

constexpr int kBlock = 256;  // 4 wave64 per workgroup

inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Grid for grid-stride loops over a device-resident count (pairs).  256 CUs
// x 8 workgroups keeps every CU busy without launching empty blocks for the
// capacity that a step does not use.
constexpr unsigned kStrideGrid = 2048;

// ----------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
// as easy as 1, 2, 3"; constants as published with Random123).
// ----------------------------------------------------------------------------
struct philox4x64 {
    uint64_t v[4];
};

__host__ __device__ inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    *lo = a * b;
    *hi = __umul64hi(a, b);
#else
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
#endif
}

__host__ __device__ inline philox4x64 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2,
                                                     uint64_t c3, uint64_t k0, uint64_t k1) {
    const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(M0, c0, &hi0, &lo0);
        mulhilo64(M1, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    philox4x64 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// Uniform double in [0, 1) from the top 53 bits of a 64-bit word: the same
// mapping numpy's Generator.random() applies to a Philox output word.
__host__ __device__ inline double u01_from_u64(uint64_t x) {
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// Counter layout.  Pair interaction uniforms: (pair rank, step, 0, 0).
// Diffusion (random walk) uniforms: (particle, step, 1, 0), word 0 -> lat,
// word 1 -> lon.
constexpr uint64_t kStreamPairs = 0;
constexpr uint64_t kStreamWalk = 1;

#define LM_DEG2RAD 0.017453292519943295
// 1 / (1852 * 60): degrees per metre of arc.
#define LM_INV_M_PER_DEG 0x1.2df73a213fa8bp-17
// Keeps the loads that produced these 32-bit values ahead of this point: the
// compiler may otherwise sink a load into a later branch that alone uses it,
// where it waits on its own, so a group of loads issued together would take
// several memory latencies instead of one.  An empty statement that reads
// and writes each value in a VGPR: it only forces the values to be present
// here (one wait for the group).
__device__ __forceinline__ void lm_pin1(uint32_t& v) { asm volatile("" : "+v"(v)); }

template <typename... T>
__device__ __forceinline__ void lm_pin(T&... v) {
    (lm_pin1(v), ...);
}

// XCD-aware block order.  MI355X has 8 XCDs, each with its own L2, and hands
// workgroups to them round-robin by linear workgroup id.  This maps the
// dispatch order so that XCD x runs the x-th contiguous eighth of the
// logical (blockIdx.x fastest, then blockIdx.y) block range: neighbouring
// logical blocks, which touch neighbouring data, share an L2.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ void xcd_block(uint32_t* bx, uint32_t* by) {
    const uint32_t gx = gridDim.x, total = gridDim.x * gridDim.y;
    const uint32_t id = blockIdx.x + blockIdx.y * gx;
    const uint32_t x = id % kXcds, k = id / kXcds;
    const uint32_t base = total / kXcds, extra = total % kXcds;
    const uint32_t logical = x * base + (x < extra ? x : extra) + k;
    *by = logical / gx;
    *bx = logical % gx;
}

// ----------------------------------------------------------------------------
// cos(lat * pi / 180) for |lat| <= 90: Taylor series of cos about 0 through
// the x^22 term (truncation below 1e-19 for |x| <= pi/2), Horner in z = x^2
// with explicit fma.  Written out so CPU and GPU evaluate the same operations.
// Coefficients c_k = (-1)^k / (2k)!, k = 11 .. 0, as exactly rounded hex
// literals (a macro, so a kernel can also keep them in constant memory).
// ----------------------------------------------------------------------------
#define LM_COS_COEFS                                                                                          \
    -0x1.0ce396db7f853p-70, 0x1.e542ba4020225p-62, -0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45,             \
        -0x1.93974a8c07c9dp-37, 0x1.1eed8eff8d898p-29, -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16,          \
        -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5, -0.5, 1.0
// sin: s_k = (-1)^k / (2k+1)!, k = 8 .. 0.
#define LM_SIN_COEFS                                                                                          \
    0x1.952c77030ad4ap-49, -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33, -0x1.ae64567f544e4p-26,             \
        0x1.71de3a556c734p-19, -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7, -0x1.5555555555555p-3, 1.0

// Horner in z over coefficients c[0 .. N-1] (highest power first).
template <int N>
__host__ __device__ inline double lm_horner(const double* c, double z) {
    double p = c[0];
#pragma unroll
    for (int k = 1; k < N; ++k) p = fma(p, z, c[k]);
    return p;
}

__host__ __device__ inline double cos_deg(double lat_deg) {
    const double x = lat_deg * LM_DEG2RAD;
    constexpr double c[12] = {LM_COS_COEFS};
    return lm_horner<12>(c, x * x);
}

// sin(x * pi / 180) for |x| <= 90: Taylor series through x^17 (truncation
// below 5e-14 at x = pi/2), Horner in z = x^2 with explicit fma; the same
// operations on CPU and GPU.
__host__ __device__ inline double sin_deg(double deg) {
    const double x = deg * LM_DEG2RAD;
    constexpr double c[9] = {LM_SIN_COEFS};
    return lm_horner<9>(c, x * x) * x;
}

}  // namespace lm
