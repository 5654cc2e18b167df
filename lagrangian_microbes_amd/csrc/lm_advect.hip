// RK4 advection of particles through a gridded, linear-in-time velocity
// field (bilinear in lon/lat), FP64.
//
// One lane per particle; particle state is SoA lon[n], lat[n] (f64), read
// once and written once per launch, with nsteps RK4 steps in registers in
// between.  The three stage fields a step samples (t, t+dt/2, t+dt) were
// blended node-wise beforehand (lm_field_blend_f64), so a stage reads four
// interleaved (u,v) nodes = 2 x 32 contiguous bytes, which stay in L2/MALL:
// a 0.25 degree field over 20 x 20 degrees is 0.1 MB per stage time.
//
// Operation order (mirrored operation by operation in oracle/lm_oracle.c):
//   fx  = (x - x0) * inv_dx,          i = trunc(fx) clamped to nx-2, xi  = fx - i
//   fy  = (y - y0) * inv_dy,          j = trunc(fy) clamped to ny-2, eta = fy - j
//   u_b = fma(xi, u10 - u00, u00),   u_t = fma(xi, u11 - u01, u01)
//   u   = fma(eta, u_t - u_b, u_b)   (same for v)
//   u_deg = (u * (1/111120)) / cos_deg(y),   v_deg = v * (1/111120)
//   x1 = fma(u1, dt/2, x) ... x3 = fma(u3, dt, x)
//   x += fma(2, u2 + u3, u1 + u4) * (dt/6) as fma(., dt/6, x)
#include "lm_common.h"

namespace lm {

__device__ __forceinline__ bool sample_uv(const double* __restrict__ f, const lm_grid& g, double x,
                                          double y, double* ud, double* vd) {
    const double fx = (x - g.x0) * g.inv_dx;
    const double fy = (y - g.y0) * g.inv_dy;
    if (!(fx >= 0.0 && fx <= (double)(g.nx - 1) && fy >= 0.0 && fy <= (double)(g.ny - 1)))
        return false;  // outside the grid (or NaN)
    int i = (int)fx;
    int j = (int)fy;
    i = i > g.nx - 2 ? g.nx - 2 : i;
    j = j > g.ny - 2 ? g.ny - 2 : j;
    const double xi = fx - (double)i;
    const double eta = fy - (double)j;
    // node index in 32 bits (the entry points require nx * ny < 2^31): one
    // integer multiply-add instead of a 64-bit product per sample
    const double2* r0 = reinterpret_cast<const double2*>(f) + ((uint32_t)j * (uint32_t)g.nx + (uint32_t)i);
    const double2* r1 = r0 + g.nx;
    const double2 a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    const double ub = fma(xi, b.x - a.x, a.x);
    const double ut = fma(xi, d.x - c.x, c.x);
    const double u = fma(eta, ut - ub, ub);
    const double vb = fma(xi, b.y - a.y, a.y);
    const double vt = fma(xi, d.y - c.y, c.y);
    const double v = fma(eta, vt - vb, vb);
    *ud = (u * LM_INV_M_PER_DEG) / cos_deg(y);
    *vd = v * LM_INV_M_PER_DEG;
    return true;
}

// One RK4 step; returns false (position untouched) if any stage leaves the grid.
__device__ __forceinline__ bool rk4_step(const double* __restrict__ f0, const double* __restrict__ fh,
                                         const double* __restrict__ f1, const lm_grid& g,
                                         double hh, double h, double h6, double* x, double* y) {
    double u1, v1, u2, v2, u3, v3, u4, v4;
    const double x0 = *x, y0 = *y;
    if (!sample_uv(f0, g, x0, y0, &u1, &v1)) return false;
    if (!sample_uv(fh, g, fma(u1, hh, x0), fma(v1, hh, y0), &u2, &v2)) return false;
    if (!sample_uv(fh, g, fma(u2, hh, x0), fma(v2, hh, y0), &u3, &v3)) return false;
    if (!sample_uv(f1, g, fma(u3, h, x0), fma(v3, h, y0), &u4, &v4)) return false;
    const double su = fma(2.0, u2 + u3, u1 + u4);
    const double sv = fma(2.0, v2 + v3, v1 + v4);
    *x = fma(su, h6, x0);
    *y = fma(sv, h6, y0);
    return true;
}

__global__ void __launch_bounds__(kBlock)
advect_rk4_f64_kernel(int64_t n, double* __restrict__ lon, double* __restrict__ lat,
                      const double* __restrict__ fields, int64_t stage_stride, lm_grid g,
                      int nsteps, double hh, double h, double h6,
                      float* __restrict__ rec_lon, float* __restrict__ rec_lat, int64_t rec_stride,
                      double walk_amp, uint64_t k0, uint64_t k1, uint64_t step0,
                      uint32_t* __restrict__ oob_count) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double x = lon[i], y = lat[i];
    for (int s = 0; s < nsteps; ++s) {
        const double* f0 = fields + (int64_t)(2 * s) * stage_stride;
        if (!rk4_step(f0, f0 + stage_stride, f0 + 2 * stage_stride, g, hh, h, h6, &x, &y))
            atomicAdd(oob_count, 1u);
        if (rec_lon) {
            rec_lon[(int64_t)s * rec_stride + i] = (float)x;
            rec_lat[(int64_t)s * rec_stride + i] = (float)y;
        }
        if (walk_amp > 0.0) {
            const philox4x64 r = philox4x64_10((uint64_t)i, step0 + (uint64_t)s, kStreamWalk, 0, k0, k1);
            y = fma(fma(2.0, u01_from_u64(r.v[0]), -1.0), walk_amp, y);
            x = fma(fma(2.0, u01_from_u64(r.v[1]), -1.0), walk_amp, x);
        }
    }
    lon[i] = x;
    lat[i] = y;
}

__global__ void __launch_bounds__(kBlock)
field_blend_f64_kernel(const double* __restrict__ levels, int64_t len, const int32_t* __restrict__ lev0,
                       const double* __restrict__ w, double* __restrict__ out) {
    const int s = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= len) return;
    const double* a = levels + (int64_t)lev0[s] * len;
    const double f0 = a[k], f1 = a[len + k];
    out[(int64_t)s * len + k] = fma(w[s], f1 - f0, f0);
}

// ----------------------------------------------------------------------------
// FP32 path: float32 field, interpolation and velocities; the positions (the
// accumulation x += h/6 (k1 + 2 k2 + 2 k3 + k4), the stage positions and the
// cell lookup) stay float64.  Operation order, mirrored in the oracle's
// oracle_advect_rk4_f32:
//   fx = (x - x0) * inv_dx (f64), i = trunc(fx) clamped, xi = (float)(fx - i)
//   u_b = fmaf(xi, u10 - u00, u00) ... u = fmaf(eta, u_t - u_b, u_b) (f32)
//   u_deg = (u * K) / cos_deg_f32((float)y),  v_deg = v * K  (f32, K = 1/111120)
//   x1 = fma((double)u1, dt/2, x) ... (f64)
//   x += fmaf(2, u2 + u3, u1 + u4) widened, as fma(., dt/6, x) (f64)
// ----------------------------------------------------------------------------
#define LM_DEG2RAD_F 0x1.1df46ap-6f
#define LM_INV_M_PER_DEG_F 0x1.2df73ap-17f

// cos(lat * pi / 180) in float32 for |lat| <= 90: Taylor series through x^12
// (truncation below 7e-9 at x = pi/2), Horner in z = x^2 with fmaf.
__device__ __forceinline__ float cos_deg_f32(float lat_deg) {
    const float x = lat_deg * LM_DEG2RAD_F;
    const float z = x * x;
    float p = 0x1.1eed8ep-29f;
    p = fmaf(p, z, -0x1.27e4fcp-22f);
    p = fmaf(p, z, 0x1.a01a02p-16f);
    p = fmaf(p, z, -0x1.6c16c2p-10f);
    p = fmaf(p, z, 0x1.555556p-5f);
    p = fmaf(p, z, -0.5f);
    return fmaf(p, z, 1.0f);
}

__device__ __forceinline__ bool sample_uv_f32(const float* __restrict__ f, const lm_grid& g, double x, double y,
                                              float* ud, float* vd) {
    const double fx = (x - g.x0) * g.inv_dx;
    const double fy = (y - g.y0) * g.inv_dy;
    if (!(fx >= 0.0 && fx <= (double)(g.nx - 1) && fy >= 0.0 && fy <= (double)(g.ny - 1)))
        return false;  // outside the grid (or NaN)
    int i = (int)fx;
    int j = (int)fy;
    i = i > g.nx - 2 ? g.nx - 2 : i;
    j = j > g.ny - 2 ? g.ny - 2 : j;
    const float xi = (float)(fx - (double)i);
    const float eta = (float)(fy - (double)j);
    const float2* r0 = reinterpret_cast<const float2*>(f) + ((uint32_t)j * (uint32_t)g.nx + (uint32_t)i);
    const float2* r1 = r0 + g.nx;
    const float2 a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    const float ub = fmaf(xi, b.x - a.x, a.x);
    const float ut = fmaf(xi, d.x - c.x, c.x);
    const float u = fmaf(eta, ut - ub, ub);
    const float vb = fmaf(xi, b.y - a.y, a.y);
    const float vt = fmaf(xi, d.y - c.y, c.y);
    const float v = fmaf(eta, vt - vb, vb);
    *ud = (u * LM_INV_M_PER_DEG_F) / cos_deg_f32((float)y);
    *vd = v * LM_INV_M_PER_DEG_F;
    return true;
}

__device__ __forceinline__ bool rk4_step_f32(const float* __restrict__ f0, const float* __restrict__ fh,
                                             const float* __restrict__ f1, const lm_grid& g, double hh, double h,
                                             double h6, double* x, double* y) {
    float u1, v1, u2, v2, u3, v3, u4, v4;
    const double x0 = *x, y0 = *y;
    if (!sample_uv_f32(f0, g, x0, y0, &u1, &v1)) return false;
    if (!sample_uv_f32(fh, g, fma((double)u1, hh, x0), fma((double)v1, hh, y0), &u2, &v2)) return false;
    if (!sample_uv_f32(fh, g, fma((double)u2, hh, x0), fma((double)v2, hh, y0), &u3, &v3)) return false;
    if (!sample_uv_f32(f1, g, fma((double)u3, h, x0), fma((double)v3, h, y0), &u4, &v4)) return false;
    const float su = fmaf(2.0f, u2 + u3, u1 + u4);
    const float sv = fmaf(2.0f, v2 + v3, v1 + v4);
    *x = fma((double)su, h6, x0);
    *y = fma((double)sv, h6, y0);
    return true;
}

__global__ void __launch_bounds__(kBlock)
advect_rk4_f32_kernel(int64_t n, double* __restrict__ lon, double* __restrict__ lat,
                      const float* __restrict__ fields, int64_t stage_stride, lm_grid g, int nsteps, double hh,
                      double h, double h6, float* __restrict__ rec_lon, float* __restrict__ rec_lat,
                      int64_t rec_stride, double walk_amp, uint64_t k0, uint64_t k1, uint64_t step0,
                      uint32_t* __restrict__ oob_count) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double x = lon[i], y = lat[i];
    for (int s = 0; s < nsteps; ++s) {
        const float* f0 = fields + (int64_t)(2 * s) * stage_stride;
        if (!rk4_step_f32(f0, f0 + stage_stride, f0 + 2 * stage_stride, g, hh, h, h6, &x, &y))
            atomicAdd(oob_count, 1u);
        if (rec_lon) {
            rec_lon[(int64_t)s * rec_stride + i] = (float)x;
            rec_lat[(int64_t)s * rec_stride + i] = (float)y;
        }
        if (walk_amp > 0.0) {
            const philox4x64 r = philox4x64_10((uint64_t)i, step0 + (uint64_t)s, kStreamWalk, 0, k0, k1);
            y = fma(fma(2.0, u01_from_u64(r.v[0]), -1.0), walk_amp, y);
            x = fma(fma(2.0, u01_from_u64(r.v[1]), -1.0), walk_amp, x);
        }
    }
    lon[i] = x;
    lat[i] = y;
}

__global__ void __launch_bounds__(kBlock)
field_blend_f32_kernel(const float* __restrict__ levels, int64_t len, const int32_t* __restrict__ lev0,
                       const double* __restrict__ w, float* __restrict__ out) {
    const int s = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= len) return;
    const float* a = levels + (int64_t)lev0[s] * len;
    const float f0 = a[k], f1 = a[len + k];
    out[(int64_t)s * len + k] = fmaf((float)w[s], f1 - f0, f0);
}

}  // namespace lm

extern "C" {

int lm_field_blend_f64(const double* levels, int64_t level_len, const int32_t* lev0, const double* w,
                       int32_t nstage, double* out, void* stream) {
    LM_REQUIRE(level_len >= 0 && nstage >= 0, LM_ERR_ARG, "negative size");
    if (level_len == 0 || nstage == 0) return LM_OK;
    LM_REQUIRE(levels && lev0 && w && out, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(nstage <= 65535, LM_ERR_ARG, "at most 65535 stage times per call");
    dim3 grid(lm::grid_for(level_len), (unsigned)nstage);
    LM_LAUNCH(lm::K_BLEND, lm::field_blend_f64_kernel, grid, dim3(lm::kBlock), 0, (hipStream_t)stream, levels,
              level_len, lev0, w, out);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_advect_rk4_f64(int64_t n, double* lon, double* lat, const double* stage_fields,
                      int64_t stage_stride, const lm_grid* grid, int32_t nsteps, double dt,
                      float* rec_lon, float* rec_lat, int64_t rec_stride, double walk_amp,
                      uint64_t key0, uint64_t key1, uint64_t step0, uint32_t* oob_count,
                      void* stream) {
    LM_REQUIRE(n >= 0 && nsteps >= 0, LM_ERR_ARG, "negative size");
    LM_REQUIRE(grid != nullptr, LM_ERR_ARG, "grid is NULL");
    LM_REQUIRE(grid->nx >= 2 && grid->ny >= 2, LM_ERR_ARG, "grid needs at least 2x2 nodes (got %dx%d)",
               grid->nx, grid->ny);
    LM_REQUIRE((int64_t)grid->nx * grid->ny < ((int64_t)1 << 31), LM_ERR_ARG, "grid has 2^31 nodes or more");
    LM_REQUIRE(stage_stride >= 0, LM_ERR_ARG, "negative stage stride");
    LM_REQUIRE(stage_stride == 0 || stage_stride >= 2 * (int64_t)grid->nx * grid->ny, LM_ERR_ARG,
               "stage stride smaller than one field");
    LM_REQUIRE((rec_lon == nullptr) == (rec_lat == nullptr), LM_ERR_ARG,
               "rec_lon and rec_lat must both be set or both be NULL");
    LM_REQUIRE(rec_lon == nullptr || rec_stride >= n, LM_ERR_ARG, "record stride < n");
    LM_REQUIRE(!(walk_amp < 0.0), LM_ERR_ARG, "walk amplitude must be >= 0");
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(lon && lat && stage_fields && oob_count, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(((uintptr_t)stage_fields & 15) == 0 && (stage_stride % 2) == 0, LM_ERR_ARG,
               "stage fields must be 16-byte aligned");
    const double hh = 0.5 * dt, h6 = dt / 6.0;
    LM_LAUNCH(lm::K_ADVECT, lm::advect_rk4_f64_kernel, dim3(lm::grid_for(n)), dim3(lm::kBlock), 0,
              (hipStream_t)stream, n, lon, lat, stage_fields, stage_stride, *grid, (int)nsteps, hh, dt, h6,
              rec_lon, rec_lat, rec_stride, walk_amp, key0, key1, step0, oob_count);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_field_blend_f32(const float* levels, int64_t level_len, const int32_t* lev0, const double* w,
                       int32_t nstage, float* out, void* stream) {
    LM_REQUIRE(level_len >= 0 && nstage >= 0, LM_ERR_ARG, "negative size");
    if (level_len == 0 || nstage == 0) return LM_OK;
    LM_REQUIRE(levels && lev0 && w && out, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(nstage <= 65535, LM_ERR_ARG, "at most 65535 stage times per call");
    dim3 grid(lm::grid_for(level_len), (unsigned)nstage);
    LM_LAUNCH(lm::K_BLEND_F32, lm::field_blend_f32_kernel, grid, dim3(lm::kBlock), 0, (hipStream_t)stream, levels,
              level_len, lev0, w, out);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

int lm_advect_rk4_f32(int64_t n, double* lon, double* lat, const float* stage_fields, int64_t stage_stride,
                      const lm_grid* grid, int32_t nsteps, double dt, float* rec_lon, float* rec_lat,
                      int64_t rec_stride, double walk_amp, uint64_t key0, uint64_t key1, uint64_t step0,
                      uint32_t* oob_count, void* stream) {
    LM_REQUIRE(n >= 0 && nsteps >= 0, LM_ERR_ARG, "negative size");
    LM_REQUIRE(grid != nullptr, LM_ERR_ARG, "grid is NULL");
    LM_REQUIRE(grid->nx >= 2 && grid->ny >= 2, LM_ERR_ARG, "grid needs at least 2x2 nodes (got %dx%d)",
               grid->nx, grid->ny);
    LM_REQUIRE((int64_t)grid->nx * grid->ny < ((int64_t)1 << 31), LM_ERR_ARG, "grid has 2^31 nodes or more");
    LM_REQUIRE(stage_stride >= 0, LM_ERR_ARG, "negative stage stride");
    LM_REQUIRE(stage_stride == 0 || stage_stride >= 2 * (int64_t)grid->nx * grid->ny, LM_ERR_ARG,
               "stage stride smaller than one field");
    LM_REQUIRE((rec_lon == nullptr) == (rec_lat == nullptr), LM_ERR_ARG,
               "rec_lon and rec_lat must both be set or both be NULL");
    LM_REQUIRE(rec_lon == nullptr || rec_stride >= n, LM_ERR_ARG, "record stride < n");
    LM_REQUIRE(!(walk_amp < 0.0), LM_ERR_ARG, "walk amplitude must be >= 0");
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(lon && lat && stage_fields && oob_count, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(((uintptr_t)stage_fields & 7) == 0 && (stage_stride % 2) == 0, LM_ERR_ARG,
               "stage fields must be 8-byte aligned");
    const double hh = 0.5 * dt, h6 = dt / 6.0;
    LM_LAUNCH(lm::K_ADVECT_F32, lm::advect_rk4_f32_kernel, dim3(lm::grid_for(n)), dim3(lm::kBlock), 0,
              (hipStream_t)stream, n, lon, lat, stage_fields, stage_stride, *grid, (int)nsteps, hh, dt, h6,
              rec_lon, rec_lat, rec_stride, walk_amp, key0, key1, step0, oob_count);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

}  // extern "C"
