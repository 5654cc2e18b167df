// Adjoint of the RK4 advection through a float32 gridded velocity field:
// gradients of a loss on the final positions with respect to the initial
// positions and to every u, v value of every time level.
//
// Field: float32 levels [nlev][ny][nx] of interleaved (u, v).  Stage time s
// (t0 + s dt/2, s = 0 .. 2 nsteps) reads level l = lev0[s] and l + 1 (the one
// level of a steady field, nlev == 1) blended node by node as
// fma(w[s], (double)f1 - (double)f0, (double)f0): the FP64 path's blend of
// the field widened to double.  adj_blend_kernel writes these float64 stage
// fields for a run of stage times into a buffer, so every sample reads four
// ready double2 corners and forward positions are bit-identical to the FP64
// advection of the widened field.
//
// Forward (lm_adjoint_forward): RK4 steps as the FP64 path, in chunks of
// steps whose stage fields fit the caller's buffer, keeping the positions at
// the start of every segment of `ckpt` steps.
//
// Backward (lm_adjoint_backward), segment by segment from the last:
//   blend      the segment's 2 m + 1 stage fields (for the recompute)
//   recompute  from the segment's checkpoint, the start position of every
//              step into scratch, and per block the bounding box of the cells
//              its stage samples touch
//   backward   steps in reverse.  Each step re-samples its four stages from the
//              stored start position (from per-lane bilinear coefficients of
//              the current cell and level pair, gathered from the float32
//              levels, which stay in L2, when the cell or pair changes)
//              keeping each stage's velocity Jacobian and interpolation
//              weights in registers, then sweeps back with
//              lambda = dL/dx at the step's end:
//                a4 = h/6 l, a3 = a2 = h/3 l, a1 = h/6 l, dL/dx0 = l
//                stage 4 (x4 = x0 + h k3):   g = J4^T a4; dx0 += g; a3 += h g
//                stage 3 (x3 = x0 + h/2 k2): g = J3^T a3; dx0 += g; a2 += h/2 g
//                stage 2 (x2 = x0 + h/2 k1): g = J2^T a2; dx0 += g; a1 += h/2 g
//                stage 1 (x0):               g = J1^T a1; dx0 += g
//              and adds a^T dk/df to the field gradient (bilinear corner
//              weights times 1 - w, w for the two levels).  A step with a stage
//              outside the grid left the position unchanged: identity.
// Velocity in degrees per second: ud = u K / cos(lat), vd = v K, K = 1/111120,
// so dud/dlat = K (du/dlat / c + u sin(lat) (pi/180) / c^2).  Recompute and
// backward take 1/cos(lat) from a Newton-refined reciprocal instead of the
// forward's IEEE division: their trajectory only positions the gradient and
// stays within rounding of the forward's.
//
// Field-gradient accumulation: per lane in registers while successive stage
// samples stay in one cell and level pair (a particle crosses a 0.25 degree
// cell in several steps), then per block in an LDS window (the segment's
// level pairs over the block's bounding box; LDS atomics), then one global
// atomic per touched node.  A block whose window does not fit the LDS it was
// launched with adds its moments to a global slab, turned into node values
// at the end.
#include "lm_common.h"

#include <mutex>

namespace lm {

constexpr int kAdjBlock = 256;

// 1 / c by two Newton steps from the hardware reciprocal estimate.
__device__ __forceinline__ double recip(double c) {
    double r = __builtin_amdgcn_rcp(c);
    r = fma(r, fma(-c, r, 1.0), r);
    return fma(r, fma(-c, r, 1.0), r);
}

// Stage fields of stage times [k0, k0 + count) in a buffer, plus the
// schedule the field gradient needs.
struct AdjField {
    const double2* st;   // st[(s - k0) * sstride + node]
    int64_t sstride;     // level_len, or 0 for a steady field (one stage field)
    int k0;
    const int32_t* lev0;
    const double* w;
    int64_t level_len;
    lm_grid g;
};

// Schedule entry s: first level and weight.
struct AdjStage {
    int s, l;
    double w;
};

// Cell and fractional position as the FP64 path computes them; false outside.
__device__ __forceinline__ bool adj_locate(const lm_grid& g, double x, double y, int* i, int* j, double* xi,
                                           double* eta) {
    const double fx = (x - g.x0) * g.inv_dx;
    const double fy = (y - g.y0) * g.inv_dy;
    if (!(fx >= 0.0 && fx <= (double)(g.nx - 1) && fy >= 0.0 && fy <= (double)(g.ny - 1))) return false;
    int ii = (int)fx, jj = (int)fy;
    ii = ii > g.nx - 2 ? g.nx - 2 : ii;
    jj = jj > g.ny - 2 ? g.ny - 2 : jj;
    *i = ii;
    *j = jj;
    *xi = fx - (double)ii;
    *eta = fy - (double)jj;
    return true;
}

// Corners (j,i), (j,i+1), (j+1,i), (j+1,i+1) of stage time s.
__device__ __forceinline__ void adj_corners(const AdjField& F, int s, int i, int j, double2 q[4]) {
    const double2* r0 = F.st + (int64_t)(s - F.k0) * F.sstride + (int64_t)j * F.g.nx + i;
    const double2* r1 = r0 + F.g.nx;
    q[0] = r0[0];
    q[1] = r0[1];
    q[2] = r1[0];
    q[3] = r1[1];
}

// The FP64 path's sample, bit for bit; also returns the cell.
__device__ __forceinline__ bool adj_sample(const AdjField& F, int s, double x, double y, double* ud, double* vd,
                                           int* ci, int* cj) {
    int i, j;
    double xi, eta;
    if (!adj_locate(F.g, x, y, &i, &j, &xi, &eta)) return false;
    double2 q[4];
    adj_corners(F, s, i, j, q);
    const double ub = fma(xi, q[1].x - q[0].x, q[0].x);
    const double ut = fma(xi, q[3].x - q[2].x, q[2].x);
    const double u = fma(eta, ut - ub, ub);
    const double vb = fma(xi, q[1].y - q[0].y, q[0].y);
    const double vt = fma(xi, q[3].y - q[2].y, q[2].y);
    const double v = fma(eta, vt - vb, vb);
    *ud = (u * LM_INV_M_PER_DEG) / cos_deg(y);
    *vd = v * LM_INV_M_PER_DEG;
    *ci = i;
    *cj = j;
    return true;
}

// The same with 1/cos(lat) from recip(): within rounding of adj_sample.
__device__ __forceinline__ bool adj_sample_fast(const AdjField& F, int s, double x, double y, double* ud,
                                                double* vd, int* ci, int* cj) {
    int i, j;
    double xi, eta;
    if (!adj_locate(F.g, x, y, &i, &j, &xi, &eta)) return false;
    double2 q[4];
    adj_corners(F, s, i, j, q);
    const double ub = fma(xi, q[1].x - q[0].x, q[0].x);
    const double ut = fma(xi, q[3].x - q[2].x, q[2].x);
    const double u = fma(eta, ut - ub, ub);
    const double vb = fma(xi, q[1].y - q[0].y, q[0].y);
    const double vt = fma(xi, q[3].y - q[2].y, q[2].y);
    const double v = fma(eta, vt - vb, vb);
    *ud = u * (LM_INV_M_PER_DEG * recip(cos_deg(y)));
    *vd = v * LM_INV_M_PER_DEG;
    *ci = i;
    *cj = j;
    return true;
}

// Cell bounding box of one lane's stage samples.
struct CellBox {
    int jmin, jmax, imin, imax;
    __device__ void reset() {
        jmin = imin = 0x7fffffff;
        jmax = imax = -1;
    }
    __device__ void add(int j, int i) {
        jmin = min(jmin, j);
        jmax = max(jmax, j);
        imin = min(imin, i);
        imax = max(imax, i);
    }
};

// One RK4 step in the FP64 path's operation order (FAST: with
// adj_sample_fast); false (position kept) if a stage leaves the grid.  With
// box != NULL the stage cells of a completed step are added to it.
template <bool FAST>
__device__ __forceinline__ bool adj_rk4(const AdjField& F, int s, double hh, double h, double h6, double* x,
                                       double* y, CellBox* box) {
    double u1, v1, u2, v2, u3, v3, u4, v4;
    int i1, j1, i2, j2, i3, j3, i4, j4;
    const double x0 = *x, y0 = *y;
    auto sample = [&](int st, double px, double py, double* ud, double* vd, int* ci, int* cj) {
        return FAST ? adj_sample_fast(F, st, px, py, ud, vd, ci, cj) : adj_sample(F, st, px, py, ud, vd, ci, cj);
    };
    if (!sample(2 * s, x0, y0, &u1, &v1, &i1, &j1)) return false;
    if (!sample(2 * s + 1, fma(u1, hh, x0), fma(v1, hh, y0), &u2, &v2, &i2, &j2)) return false;
    if (!sample(2 * s + 1, fma(u2, hh, x0), fma(v2, hh, y0), &u3, &v3, &i3, &j3)) return false;
    if (!sample(2 * s + 2, fma(u3, h, x0), fma(v3, h, y0), &u4, &v4, &i4, &j4)) return false;
    const double su = fma(2.0, u2 + u3, u1 + u4);
    const double sv = fma(2.0, v2 + v3, v1 + v4);
    *x = fma(su, h6, x0);
    *y = fma(sv, h6, y0);
    if (box) {
        box->add(j1, i1);
        box->add(j2, i2);
        box->add(j3, i3);
        box->add(j4, i4);
    }
    return true;
}

// A stage linearised for the reverse sweep: the velocity Jacobian (FP64, it
// carries lambda) and, in float32 like the field gradient it feeds, the
// interpolation weights and K / cos(lat).
struct Lin {
    double dux, duy, dvx, dvy;  // d(ud)/dx, d(ud)/dy, d(vd)/dx, d(vd)/dy
    float xi, eta, kc;
    int cell;  // (j << 16) | i
};

// The float32 levels as the backward reads them.  Corners gathered from the
// levels (1 MB for 20 levels of 81 x 81 nodes) rather than from float64
// stage fields (3.5 MB for a 16-step segment) stay in L2.
struct LevelView {
    const float2* glev;  // [nlev][ny][nx]
    int64_t level_len;
    int64_t gstride;  // level_len, or 0 for a steady field
    lm_grid g;
};

// One cell's bilinear coefficients at one level pair, per lane.  At time
// weight w, component c (u, v) of the field at fractional position (xi, eta)
// is C0 + C1 xi + C2 eta + C3 xi eta with Ck = A[4c + k] + w D[4c + k]: A
// from level l's corners (C0 = q00, C1 = q01 - q00, C2 = q10 - q00,
// C3 = q11 - q10 - q01 + q00), D from level l + 1's minus level l's (zero
// for a steady field).  That is the corner blend fma(w, f1 - f0, f0)
// followed by the bilinear sample, reassociated: equal to rounding.  A
// particle stays in one cell and level pair for many stage samples (it
// crosses a 0.25 degree cell in tens of steps), so most samples reuse the
// lane's coefficients: no gather, no blend of corners.
struct CellCoef {
    int l, cell;  // l < 0: empty; cell = (j << 16) | i
    double A[8], D[8];
};

// The coefficients of cell (j, i) at level pair l from the float32 levels.
__device__ __forceinline__ void coef_gather(const LevelView& V, int l, int i, int j, double (&A)[8], double (&D)[8]) {
    const float2* r0 = V.glev + (int64_t)l * V.level_len + (int64_t)j * V.g.nx + i;
    const int64_t ls = V.gstride;
    const int pitch = V.g.nx;
    const float2 a[4] = {r0[0], r0[1], r0[pitch], r0[pitch + 1]};
    const float2 b[4] = {r0[ls], r0[ls + 1], r0[ls + pitch], r0[ls + pitch + 1]};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        double fa[4], fb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            fa[k] = c ? (double)a[k].y : (double)a[k].x;
            fb[k] = c ? (double)b[k].y : (double)b[k].x;
        }
        const double a1 = fa[1] - fa[0], a3 = (fa[3] - fa[2]) - a1;
        const double b1 = fb[1] - fb[0], b3 = (fb[3] - fb[2]) - b1;
        A[4 * c] = fa[0];
        A[4 * c + 1] = a1;
        A[4 * c + 2] = fa[2] - fa[0];
        A[4 * c + 3] = a3;
        D[4 * c] = fb[0] - fa[0];
        D[4 * c + 1] = b1 - a1;
        D[4 * c + 2] = (fb[2] - fb[0]) - (fa[2] - fa[0]);
        D[4 * c + 3] = b3 - a3;
    }
}

// The reverse sweep's gradient window: block-shared LDS (dynamic,
// lds_bytes) over the level pairs [lo, lo + npair) and the cells of the
// block's box (ch x cw cells from (j0, i0)), 16 floats of moment slots per
// (level pair, cell), indexed (p * ch + cj) * cw + ci; used when it fits.
// A flush outside it (every flush of a block whose box does not fit, or one
// that rounding puts just past the recompute's box, which covers every
// sample of the recompute's own trajectory) goes to the global slab.
struct BlockBox {
    int lo, npair, j0, i0, ch, cw;
    bool win;
};

// Index of (level pair l, cell (j, i)) in the block's box, or -1.
__device__ __forceinline__ int box_index(const BlockBox& B, int l, int i, int j) {
    const int p = l - B.lo, cj = j - B.j0, ci = i - B.i0;
    const bool in = (unsigned)p < (unsigned)B.npair && (unsigned)cj < (unsigned)B.ch && (unsigned)ci < (unsigned)B.cw;
    return in ? (p * B.ch + cj) * B.cw + ci : -1;
}

__device__ __forceinline__ void coef_fill(const LevelView& V, int l, int i, int j, CellCoef& cc) {
    coef_gather(V, l, i, j, cc.A, cc.D);
    cc.l = l;
    cc.cell = (j << 16) | i;
}

// cos and sin of a step's start latitude y0 (degrees), from which the
// stage latitudes' are taken by angle addition: for y = y0 + e degrees,
// d = e pi / 180, cos y = cos y0 cos d - sin y0 sin d and sin y = sin y0 cos d
// + cos y0 sin d, with cos d = 1 - d^2/2 + d^4/24 and sin d = d (1 - d^2/6 +
// d^4/120), whose truncation (below d^6/720 and d^7/5040) is under 1e-19 for
// |d| <= 2^-9 rad (0.11 degrees, farther than a particle moves in one step
// at ocean speeds); farther away takes several such rotations.  A step's own
// start latitude is taken the same way from the step after it (the reverse
// sweep's previous one): the chain restarts from the polynomials at each
// segment, so it spans at most one segment's steps and its rounding stays
// within a few units in the last place.  The results agree with cos_deg /
// sin_deg to rounding: the reverse sweep's trajectory, like the
// recompute's, only positions the gradient.
struct Trig0 {
    double y0, c0, s0;
};

__device__ __forceinline__ Trig0 trig_at(double y0) { return Trig0{y0, cos_deg(y0), sin_deg(y0)}; }

__device__ __forceinline__ void trig_near(const Trig0& T, double y, double* c, double* sn) {
    double d = (y - T.y0) * LM_DEG2RAD;
    // Farther than 2^-9 rad (never at ocean speeds; up to pi inside a grid):
    // k equal rotations of at most 2^-9 rad each.  Beyond pi (or NaN) the
    // position is outside every grid and the result unused: one rotation.
    int k = 1;
    if (!(fabs(d) <= 0x1p-9) && fabs(d) <= 3.2) {
        k = (int)ceil(fabs(d) * 512.0);
        d /= (double)k;
    }
    const double d2 = d * d;
    const double cd = fma(fma(d2, 1.0 / 24.0, -0.5), d2, 1.0);
    const double sd = fma(fma(d2, 1.0 / 120.0, -1.0 / 6.0), d2, 1.0) * d;
    double cy = T.c0, sy = T.s0;
    for (int t = 0; t < k; ++t) {
        const double c1 = fma(cy, cd, -(sy * sd));
        sy = fma(sy, cd, cy * sd);
        cy = c1;
    }
    *c = cy;
    *sn = sy;
}

// adj_sample_fast at stage st plus the linearisation: the sample at (x, y)
// from the lane's cell coefficients (refilled when the cell or the level
// pair differs), with cos and sin of the latitude y given (cy, sy).  False
// outside the grid.
__device__ __forceinline__ bool adj_linearise(const LevelView& V, const AdjStage& st, CellCoef& cc, double x,
                                              double y, double cy, double sy, double* ud, double* vd, Lin* L) {
    int i, j;
    double xi, eta;
    if (!adj_locate(V.g, x, y, &i, &j, &xi, &eta)) return false;
    const int cell = (j << 16) | i;
    if (cc.l != st.l || cc.cell != cell) coef_fill(V, st.l, i, j, cc);
    double c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = fma(st.w, cc.D[k], cc.A[k]);
    const double u_xi = fma(eta, c[3], c[1]), v_xi = fma(eta, c[7], c[5]);
    const double u_eta = fma(xi, c[3], c[2]), v_eta = fma(xi, c[7], c[6]);
    const double u = fma(xi, u_xi, fma(eta, c[2], c[0]));
    const double v = fma(xi, v_xi, fma(eta, c[6], c[4]));
    const double rc = recip(cy);
    const double kc = LM_INV_M_PER_DEG * rc;
    *ud = u * kc;
    *vd = v * LM_INV_M_PER_DEG;
    // d(1/cos(lat))/dlat = (pi/180) sin(lat) / cos(lat)^2
    L->dux = kc * (u_xi * V.g.inv_dx);
    L->duy = kc * fma(u * LM_DEG2RAD * rc, sy, u_eta * V.g.inv_dy);
    L->dvx = LM_INV_M_PER_DEG * (v_xi * V.g.inv_dx);
    L->dvy = LM_INV_M_PER_DEG * (v_eta * V.g.inv_dy);
    L->xi = (float)xi;
    L->eta = (float)eta;
    L->kc = (float)kc;
    L->cell = cell;
    return true;
}

// The same at a stage position whose latitude's cos and sin come from T.
__device__ __forceinline__ bool adj_linearise_near(const LevelView& V, const AdjStage& st, CellCoef& cc,
                                                   const Trig0& T, double x, double y, double* ud, double* vd,
                                                   Lin* L) {
    double cy, sy;
    trig_near(T, y, &cy, &sy);
    return adj_linearise(V, st, cc, x, y, cy, sy, ud, vd, L);
}

// Field-gradient accumulation in moment form.  Samples in one cell (j, i)
// at one level pair (l, l + 1) with velocity adjoint g and time weight w
// contribute g (1 - w) b(xi, eta) to level l and g w b(xi, eta) to level
// l + 1, b being the four bilinear corner weights.  Being linear, they are
// summed as the moments S = sum g (1, xi, eta, xi eta) and W = sum g w (...)
// and turned into corner values only once: corner (0,0) of m is
// m0 - m1 - m2 + m3, (0,1) is m1 - m3, (1,0) is m2 - m3, (1,1) is m3, with
// m = S - W for level l and m = W for level l + 1.
//
// Slots: 16 floats, [comp][S, W][4 moments], per level pair and cell: the
// block's LDS window (BlockBox) or, outside it, the global slab of all
// (nlev - 1 or 1) pairs x (ny - 1) x (nx - 1) cells.  Window and slab become
// node gradients at the end: the block's write-out and adj_slab_kernel.
struct GradSink {
    float* slab;  // [pairs][ny - 1][nx - 1][16]
    int cx, cy;   // cells along x, y
};

struct GradCache {
    int l, cell;  // l < 0: empty; cell = (j << 16) | i
    float m[16];  // [comp][S, W][moment]
};

__device__ __forceinline__ float corner_of(const float* m, int dj, int di) {
    const float m3 = m[3];
    if (dj == 0 && di == 0) return (m[0] - m[1]) - (m[2] - m3);
    if (dj == 0) return m[1] - m3;
    if (di == 0) return m[2] - m3;
    return m3;
}

// win: the block's window in LDS (read only when B.win).
__device__ __forceinline__ void cache_flush(GradCache& gc, const BlockBox& B, float* win, const GradSink& sk) {
    if (gc.l < 0) return;
    const int cj = gc.cell >> 16, ci = gc.cell & 0xffff;
    const int t = B.win ? box_index(B, gc.l, ci, cj) : -1;
    if (t >= 0) {
        // workgroup scope, unlike the slab's device-scope atomics: the two
        // loops are not merged into one through a generic pointer, so these
        // stay LDS atomics
        float* slot = win + t * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            __hip_atomic_fetch_add(slot + k, gc.m[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        float* slot = sk.slab + (((int64_t)gc.l * sk.cy + cj) * sk.cx + ci) * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) atomicAdd(slot + k, gc.m[k]);
    }
    gc.l = -1;
}

// Node value of one (u or v) component from the slots around node (nj, ni)
// of level lv: corner (dj, di) of cell (nj - dj, ni - di), the lower part of
// pair lv and the upper part of pair lv - 1 (all of pair 0 when steady).
// slot(p, cj, ci) returns the slot's 8 floats of this component or NULL.
template <typename SlotFn>
__device__ __forceinline__ float node_from_slots(int lv, int nj, int ni, bool steady, SlotFn slot) {
    float acc = 0.0f;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        if (steady && part == 1) continue;
        const int pr = steady ? 0 : lv - part;
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const float* S = slot(pr, nj - dj, ni - di);
                if (!S) continue;
                const float* W = S + 4;
                float mm[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) mm[q] = steady ? S[q] : (part == 0 ? S[q] - W[q] : W[q]);
                acc += corner_of(mm, dj, di);
            }
    }
    return acc;
}

// Adds one stage sample to the cache's moments: velocity adjoints (gu, gv)
// already scaled to m/s, time weight w, fractional position (xi, eta).
__device__ __forceinline__ void cache_accumulate(GradCache& gc, float w, float xi, float eta, float gu, float gv) {
    const float b[4] = {1.0f, xi, eta, xi * eta};
    const float g[2] = {gu, gv};
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        const float gw = g[comp] * w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gc.m[comp * 8 + k] = fmaf(g[comp], b[k], gc.m[comp * 8 + k]);
            gc.m[comp * 8 + 4 + k] = fmaf(gw, b[k], gc.m[comp * 8 + 4 + k]);
        }
    }
}

// Field gradient of one stage sample with adjoint (au, av) of its velocity.
__device__ __forceinline__ void cache_add(GradCache& gc, const BlockBox& B, float* win, const GradSink& sk,
                                          const AdjStage& st, const Lin& L, double au, double av) {
    if (gc.l != st.l || gc.cell != L.cell) {
        cache_flush(gc, B, win, sk);
        gc.l = st.l;
        gc.cell = L.cell;
#pragma unroll
        for (int k = 0; k < 16; ++k) gc.m[k] = 0.0f;
    }
    cache_accumulate(gc, (float)st.w, L.xi, L.eta, (float)au * L.kc, (float)(av * LM_INV_M_PER_DEG));
}

// out[k][node] = blend of stage time k0 + k (k < count, one block row each).
__global__ void __launch_bounds__(kAdjBlock)
adj_blend_kernel(const float2* __restrict__ lev, int64_t level_len, int64_t lstride, const int32_t* __restrict__ lev0,
                 const double* __restrict__ w, int k0, double2* __restrict__ out) {
    const int k = blockIdx.y;
    const int64_t node = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (node >= level_len) return;
    const int s = k0 + k;
    const double wt = w[s];
    const float2* a = lev + (int64_t)lev0[s] * level_len + node;
    const float2 f0 = a[0], f1 = a[lstride];
    out[(int64_t)k * level_len + node] = make_double2(fma(wt, (double)f1.x - (double)f0.x, (double)f0.x),
                                                      fma(wt, (double)f1.y - (double)f0.y, (double)f0.y));
}

// Steps [s_begin, s_end) of the run; the positions at the start of each
// step s with s % ckpt == 0 go to checkpoint s / ckpt.
__global__ void __launch_bounds__(kAdjBlock)
adj_forward_kernel(int64_t n, double* __restrict__ lon, double* __restrict__ lat, AdjField F, int s_begin, int s_end,
                   double hh, double h, double h6, int ckpt, double2* __restrict__ ck, uint32_t* __restrict__ oob) {
    uint32_t bx, by;
    xcd_block(&bx, &by);  // an XCD's blocks cover neighbouring particles: their stage-field cells share its L2
    const int64_t i = (int64_t)bx * kAdjBlock + threadIdx.x;
    if (i >= n) return;
    double x = lon[i], y = lat[i];
    uint32_t out = 0;
    for (int s = s_begin; s < s_end; ++s) {
        if (s % ckpt == 0) ck[(int64_t)(s / ckpt) * n + i] = make_double2(x, y);
        if (!adj_rk4<false>(F, s, hh, h, h6, &x, &y, nullptr)) ++out;
    }
    lon[i] = x;
    lat[i] = y;
    if (out) atomicAdd(oob, out);
}

// Steps s0 .. s0 + m - 1 from the checkpoint ck: the start position of step
// s0 + k into pos[k * n + i]; the block's cell box into box[block].
__global__ void __launch_bounds__(kAdjBlock)
adj_recompute_kernel(int64_t n, const double2* __restrict__ ck, AdjField F, int s0, int m, double hh, double h,
                     double h6, double2* __restrict__ pos, int4* __restrict__ box) {
    __shared__ int sb[4];
    if (threadIdx.x == 0) {
        sb[0] = sb[2] = 0x7fffffff;
        sb[1] = sb[3] = -1;
    }
    __syncthreads();
    uint32_t bx, by;
    xcd_block(&bx, &by);  // the backward maps blocks the same way
    const int64_t i = (int64_t)bx * kAdjBlock + threadIdx.x;
    CellBox cb;
    cb.reset();
    if (i < n) {
        const double2 p = ck[i];
        double x = p.x, y = p.y;
        for (int k = 0; k < m; ++k) {
            pos[(int64_t)k * n + i] = make_double2(x, y);
            adj_rk4<true>(F, s0 + k, hh, h, h6, &x, &y, &cb);
        }
    }
    if (cb.jmax >= 0) {
        atomicMin(&sb[0], cb.jmin);
        atomicMax(&sb[1], cb.jmax);
        atomicMin(&sb[2], cb.imin);
        atomicMax(&sb[3], cb.imax);
    }
    __syncthreads();
    if (threadIdx.x == 0) box[bx] = make_int4(sb[0], sb[1], sb[2], sb[3]);
}

// Steps s0 + m - 1 down to s0, from the recomputed start positions; lambda
// (glon, glat) goes from the end of the segment to its start.  Level pairs
// [lo, lo + npair) are the ones the segment's stages read; lds_bytes is the
// dynamic LDS the launch carries for the block's gradient window.
__global__ void __launch_bounds__(kAdjBlock)
adj_backward_kernel(int64_t n, LevelView V, const int32_t* __restrict__ lev0, const double* __restrict__ wt, int nlev,
                    int s0, int m, double hh, double h, double h6, const double2* __restrict__ pos,
                    const int4* __restrict__ box, double* __restrict__ glon, double* __restrict__ glat,
                    float* __restrict__ gfield, float* __restrict__ slab, int lo, int npair, int lds_bytes) {
    extern __shared__ float lds[];
    uint32_t bx, by;
    xcd_block(&bx, &by);  // as the recompute, whose box of logical block bx this block reads
    const int4 b = box[bx];
    if (b.y < 0) return;  // no lane of the block completed a step inside the grid
    BlockBox B{lo, npair, b.x, b.z, b.y - b.x + 1, b.w - b.z + 1, false};
    // Entries (level pair, cell) of the box, or cap + 1 when more than the
    // window holds: 32-bit saturating products and compares only (a select
    // on a 64-bit compare's result has been compiled with a stale scalar
    // condition, giving a wrong LDS offset).
    const int cap = lds_bytes / 64;
    int ncell = B.ch <= cap ? B.ch : cap + 1;
    ncell = B.cw <= cap / ncell ? ncell * B.cw : cap + 1;
    ncell = npair <= cap / ncell ? ncell * npair : cap + 1;
    B.win = ncell <= cap;
    float* win = lds;
    if (B.win) {
        for (int t = threadIdx.x; t < ncell * 16; t += kAdjBlock) win[t] = 0.0f;
        __syncthreads();
    }
    GradSink sk;
    sk.slab = slab;
    sk.cx = V.g.nx - 1;
    sk.cy = V.g.ny - 1;
    GradCache gc;
    gc.l = -1;
    const int64_t i = (int64_t)bx * kAdjBlock + threadIdx.x;
    if (i < n) {
        double lx = glon[i], ly = glat[i];
        CellCoef cc;
        cc.l = -1;
        cc.cell = 0;
        double2 p = pos[(int64_t)(m - 1) * n + i];
        Trig0 T = trig_at(p.y);
        for (int k = m - 1; k >= 0; --k) {
            const int s = s0 + k;
            const double x0 = p.x, y0 = p.y;
            // the start of the step before, loading while this one is linearised
            if (k > 0) p = pos[(int64_t)(k - 1) * n + i];
            // schedule entries: uniform, read through the restrict pointers with scalar loads
            const AdjStage sa{2 * s, lev0[2 * s], wt[2 * s]}, sb{2 * s + 1, lev0[2 * s + 1], wt[2 * s + 1]},
                sc{2 * s + 2, lev0[2 * s + 2], wt[2 * s + 2]};
            Lin L1, L2, L3, L4;
            double ud, vd, cy, sy;
            trig_near(T, y0, &cy, &sy);
            T = Trig0{y0, cy, sy};
            if (!adj_linearise(V, sa, cc, x0, y0, cy, sy, &ud, &vd, &L1)) continue;
            if (!adj_linearise_near(V, sb, cc, T, fma(ud, hh, x0), fma(vd, hh, y0), &ud, &vd, &L2)) continue;
            if (!adj_linearise_near(V, sb, cc, T, fma(ud, hh, x0), fma(vd, hh, y0), &ud, &vd, &L3)) continue;
            if (!adj_linearise_near(V, sc, cc, T, fma(ud, h, x0), fma(vd, h, y0), &ud, &vd, &L4)) continue;
            const double a4x = h6 * lx, a4y = h6 * ly;
            double a3x = 2.0 * a4x, a3y = 2.0 * a4y;
            double a2x = a3x, a2y = a3y;
            double a1x = a4x, a1y = a4y;
            double gx, gy;
            // stage 4
            gx = fma(a4x, L4.dux, a4y * L4.dvx);
            gy = fma(a4x, L4.duy, a4y * L4.dvy);
            cache_add(gc, B, win, sk, sc, L4, a4x, a4y);
            lx += gx;
            ly += gy;
            a3x = fma(h, gx, a3x);
            a3y = fma(h, gy, a3y);
            // stage 3
            gx = fma(a3x, L3.dux, a3y * L3.dvx);
            gy = fma(a3x, L3.duy, a3y * L3.dvy);
            cache_add(gc, B, win, sk, sb, L3, a3x, a3y);
            lx += gx;
            ly += gy;
            a2x = fma(hh, gx, a2x);
            a2y = fma(hh, gy, a2y);
            // stage 2
            gx = fma(a2x, L2.dux, a2y * L2.dvx);
            gy = fma(a2x, L2.duy, a2y * L2.dvy);
            cache_add(gc, B, win, sk, sb, L2, a2x, a2y);
            lx += gx;
            ly += gy;
            a1x = fma(hh, gx, a1x);
            a1y = fma(hh, gy, a1y);
            // stage 1
            gx = fma(a1x, L1.dux, a1y * L1.dvx);
            gy = fma(a1x, L1.duy, a1y * L1.dvy);
            cache_add(gc, B, win, sk, sa, L1, a1x, a1y);
            lx += gx;
            ly += gy;
        }
        glon[i] = lx;
        glat[i] = ly;
    }
    cache_flush(gc, B, win, sk);
    if (!B.win) return;
    __syncthreads();
    const int nh = B.ch + 1, nw = B.cw + 1, nlw = nlev > 1 ? npair + 1 : 1;
    const int nodes = nlw * nh * nw * 2;
    for (int t = threadIdx.x; t < nodes; t += kAdjBlock) {
        const int comp = t & 1;
        int r = t >> 1;
        const int ni = r % nw;
        r /= nw;
        const int nj = r % nh;
        const int lv = r / nh;
        const float acc = node_from_slots(lv, nj, ni, nlev == 1, [&](int p, int cj, int ci) -> const float* {
            if (p < 0 || p >= npair || cj < 0 || cj >= B.ch || ci < 0 || ci >= B.cw) return nullptr;
            return win + ((p * B.ch + cj) * B.cw + ci) * 16 + comp * 8;
        });
        if (acc != 0.0f)
            atomicAdd(gfield + ((int64_t)(lo + lv) * V.level_len + (int64_t)(B.j0 + nj) * V.g.nx + (B.i0 + ni)) * 2 +
                          comp,
                      acc);
    }
}

// Clears the global slab (a kernel rather than a memset node, so a captured
// backward replays correctly).
__global__ void __launch_bounds__(kAdjBlock) adj_zero_kernel(float4* __restrict__ p, int64_t n4) {
    const int64_t t = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (t < n4) p[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Adds the global slab's node values into gfield: one thread per (level,
// node, component).
__global__ void __launch_bounds__(kAdjBlock)
adj_slab_kernel(const float* __restrict__ slab, int nlev, lm_grid g, float* __restrict__ gfield) {
    const int64_t L = (int64_t)g.nx * g.ny;
    const int64_t t = (int64_t)blockIdx.x * kAdjBlock + threadIdx.x;
    if (t >= (int64_t)nlev * L * 2) return;
    const int comp = (int)(t & 1);
    const int64_t node = (t >> 1) % L;
    const int lv = (int)((t >> 1) / L);
    const int ni = (int)(node % g.nx), nj = (int)(node / g.nx);
    const int cx = g.nx - 1, cy = g.ny - 1, npair = nlev > 1 ? nlev - 1 : 1;
    const float acc = node_from_slots(lv, nj, ni, nlev == 1, [&](int p, int cj, int ci) -> const float* {
        if (p < 0 || p >= npair || cj < 0 || cj >= cy || ci < 0 || ci >= cx) return nullptr;
        return slab + (((int64_t)p * cy + cj) * cx + ci) * 16 + comp * 8;
    });
    if (acc != 0.0f) gfield[t] += acc;
}

}  // namespace lm

namespace {

// Dynamic LDS each backward block may use for its gradient window: 52 KB
// lets three blocks (the waves its registers allow) share a CU's 160 KB; a
// block whose box needs more adds to global memory directly.
constexpr int kAdjWindowBytes = 53248;

int adj_check(int64_t n, int32_t nlev, const lm_grid* grid, const int32_t* lev0_host, int32_t nsteps,
              int32_t ckpt) {
    LM_REQUIRE(n >= 0 && nsteps >= 0, LM_ERR_ARG, "negative size");
    LM_REQUIRE(n <= 0x7fffffffLL * lm::kAdjBlock, LM_ERR_ARG, "too many particles");
    LM_REQUIRE(grid != nullptr, LM_ERR_ARG, "grid is NULL");
    LM_REQUIRE(grid->nx >= 2 && grid->ny >= 2, LM_ERR_ARG, "grid needs at least 2x2 nodes (got %dx%d)", grid->nx,
               grid->ny);
    LM_REQUIRE(grid->nx <= 65536 && grid->ny <= 65536, LM_ERR_ARG, "at most 65536 nodes along each axis");
    LM_REQUIRE(nlev >= 1, LM_ERR_ARG, "at least one time level");
    LM_REQUIRE(ckpt >= 1, LM_ERR_ARG, "checkpoint interval must be >= 1");
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(lev0_host != nullptr, LM_ERR_ARG, "lev0_host is NULL");
    for (int s = 0; s <= 2 * nsteps; ++s) {
        const int l = lev0_host[s];
        LM_REQUIRE(l >= 0 && (nlev == 1 ? l == 0 : l + 1 < nlev), LM_ERR_ARG,
                   "stage time %d: level %d does not fit %d level(s)", s, l, nlev);
    }
    return LM_OK;
}

int64_t level_len_of(const lm_grid* g) { return (int64_t)g->nx * g->ny; }

// Global moment slab: 16 floats per level pair and cell.
int64_t slab_bytes(const lm_grid* g, int32_t nlev) {
    return (int64_t)(nlev > 1 ? nlev - 1 : 1) * (g->nx - 1) * (g->ny - 1) * 64;
}

// Stage fields for stage times [k0, k0 + count) into buf (one field when
// steady); returns the view the kernels read.
int adj_blend(const float* field, int32_t nlev, const lm_grid* grid, const int32_t* lev0, const double* w, int k0,
              int count, double2* buf, hipStream_t s, lm::AdjField* F) {
    const int64_t L = level_len_of(grid);
    const bool steady = nlev == 1;
    const int nblend = steady ? 1 : count;
    for (int a = 0; a < nblend; a += 65535) {
        const int c = nblend - a < 65535 ? nblend - a : 65535;
        LM_LAUNCH(lm::K_ADJ_BLEND, lm::adj_blend_kernel, dim3(lm::grid_for(L, lm::kAdjBlock), (unsigned)c),
                  dim3(lm::kAdjBlock), 0, s, reinterpret_cast<const float2*>(field), L, steady ? (int64_t)0 : L, lev0,
                  w, k0 + a, buf + (int64_t)a * L);
        LM_LAUNCH_CHECK();
    }
    F->st = buf;
    F->sstride = steady ? 0 : L;
    F->k0 = k0;
    F->lev0 = lev0;
    F->w = w;
    F->level_len = L;
    F->g = *grid;
    return LM_OK;
}

// ---------------------------------------------------------------------------
// Static scratch of lm_adjoint_forward_noalloc: the stage fields live in this
// file-scope device array instead of caller memory.  Each device holds its
// own copy at its own address from code-object load to process exit.  On the
// host, the array's name is only a stand-in for the device symbol: host code
// never passes it to a kernel or a copy.  Its device address is looked up
// once per device (hipGetSymbolAddress, with that device current) and kept;
// every call on a device shares that device's array, so calls that overlap
// on one device race.  Nothing in it needs clearing: each chunk's stage
// fields are written by the blend before the advection reads them.
// ---------------------------------------------------------------------------
constexpr int64_t kStaticScratchBytes = 32ll << 20;
constexpr int kStaticScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_static_mu;
void* g_static_ptr[kStaticScratchDevices];

// Device address of `scratch` on the current device.
int static_scratch(void** out) {
    int dev = -1;
    LM_HIP_CHECK(hipGetDevice(&dev));
    LM_REQUIRE(dev >= 0 && dev < kStaticScratchDevices, LM_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (g_static_ptr[dev] == nullptr) {
        void* p = nullptr;
        LM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        LM_REQUIRE(p != nullptr && ((uintptr_t)p & 255) == 0, LM_ERR_HIP, "static scratch has no device address");
        g_static_ptr[dev] = p;
    }
    *out = g_static_ptr[dev];
    return LM_OK;
}

}  // namespace

extern "C" {

int64_t lm_adjoint_scratch_bytes(int64_t n, int32_t ckpt, const lm_grid* grid, int32_t nlev) {
    if (n < 0 || ckpt < 1 || grid == nullptr || grid->nx < 2 || grid->ny < 2 || nlev < 1) return -1;
    const int64_t blocks = (n + lm::kAdjBlock - 1) / lm::kAdjBlock;
    const int64_t stages = nlev == 1 ? 1 : 2 * (int64_t)ckpt + 1;
    return (int64_t)ckpt * n * 16 + blocks * 16 + stages * level_len_of(grid) * 16 + slab_bytes(grid, nlev);
}

int lm_adjoint_forward(int64_t n, double* lon, double* lat, const float* field, int32_t nlev, const lm_grid* grid,
                       const int32_t* lev0, const double* w, const int32_t* lev0_host, int32_t nsteps, double dt,
                       int32_t ckpt, double* checkpoints, void* scratch, int64_t scratch_bytes, uint32_t* oob_count,
                       void* stream) {
    int rc = adj_check(n, nlev, grid, lev0_host, nsteps, ckpt);
    if (rc) return rc;
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(lon && lat && field && lev0 && w && checkpoints && scratch && oob_count, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(((uintptr_t)field & 7) == 0 && ((uintptr_t)checkpoints & 15) == 0 && ((uintptr_t)scratch & 15) == 0,
               LM_ERR_ARG, "field must be 8-byte, checkpoints and scratch 16-byte aligned");
    const int64_t fbytes = level_len_of(grid) * 16;
    // steps per chunk: as many as the scratch holds the stage fields of
    int64_t chunk = nlev == 1 ? nsteps : (scratch_bytes / fbytes - 1) / 2;
    LM_REQUIRE(chunk >= 1 && (nlev > 1 || scratch_bytes >= fbytes), LM_ERR_ARG,
               "scratch holds %lld bytes; one step needs %lld", (long long)scratch_bytes,
               (long long)(nlev == 1 ? fbytes : 3 * fbytes));
    chunk = chunk > nsteps ? nsteps : chunk;
    hipStream_t s = (hipStream_t)stream;
    double2* buf = reinterpret_cast<double2*>(scratch);
    for (int a = 0; a < nsteps; a += (int)chunk) {
        const int b = nsteps - a < chunk ? nsteps : a + (int)chunk;
        lm::AdjField F;
        rc = adj_blend(field, nlev, grid, lev0, w, 2 * a, 2 * (b - a) + 1, buf, s, &F);
        if (rc) return rc;
        LM_LAUNCH(lm::K_ADJ_FORWARD, lm::adj_forward_kernel, dim3(lm::grid_for(n, lm::kAdjBlock)),
                  dim3(lm::kAdjBlock), 0, s, n, lon, lat, F, a, b, 0.5 * dt, dt, dt / 6.0, (int)ckpt,
                  reinterpret_cast<double2*>(checkpoints), oob_count);
        LM_LAUNCH_CHECK();
    }
    return LM_OK;
}

int lm_adjoint_forward_noalloc(int64_t n, double* lon, double* lat, const float* field, int32_t nlev,
                               const lm_grid* grid, const int32_t* lev0, const double* w, const int32_t* lev0_host,
                               int32_t nsteps, double dt, int32_t ckpt, double* checkpoints, uint32_t* oob_count,
                               void* stream) {
    int rc = adj_check(n, nlev, grid, lev0_host, nsteps, ckpt);
    if (rc) return rc;
    if (n == 0 || nsteps == 0) return LM_OK;
    // one step blends 3 stage fields (1 for a steady field) of 16 bytes per node
    const int64_t nodes = level_len_of(grid);
    const int64_t per_node = nlev == 1 ? 16 : 48;
    LM_REQUIRE(nodes <= kStaticScratchBytes / per_node, LM_ERR_ARG,
               "grid of %lld nodes: the static scratch of %lld bytes holds the stage fields of at most %lld nodes",
               (long long)nodes, (long long)kStaticScratchBytes, (long long)(kStaticScratchBytes / per_node));
    LM_REQUIRE(lon && lat && field && lev0 && w && checkpoints && oob_count, LM_ERR_ARG, "NULL pointer");
    void* buf = nullptr;
    rc = static_scratch(&buf);
    if (rc) return rc;
    return lm_adjoint_forward(n, lon, lat, field, nlev, grid, lev0, w, lev0_host, nsteps, dt, ckpt, checkpoints, buf,
                              kStaticScratchBytes, oob_count, stream);
}

int lm_adjoint_backward(int64_t n, const double* checkpoints, const float* field, int32_t nlev, const lm_grid* grid,
                        const int32_t* lev0, const double* w, const int32_t* lev0_host, int32_t nsteps, double dt,
                        int32_t ckpt, void* scratch, int64_t scratch_bytes, double* glon, double* glat, float* gfield,
                        void* stream) {
    int rc = adj_check(n, nlev, grid, lev0_host, nsteps, ckpt);
    if (rc) return rc;
    if (n == 0 || nsteps == 0) return LM_OK;
    LM_REQUIRE(checkpoints && field && lev0 && w && scratch && glon && glat && gfield, LM_ERR_ARG, "NULL pointer");
    LM_REQUIRE(((uintptr_t)field & 7) == 0 && ((uintptr_t)checkpoints & 15) == 0 && ((uintptr_t)scratch & 15) == 0,
               LM_ERR_ARG, "field must be 8-byte, checkpoints and scratch 16-byte aligned");
    const int64_t need = lm_adjoint_scratch_bytes(n, ckpt, grid, nlev);
    LM_REQUIRE(scratch_bytes >= need, LM_ERR_ARG, "scratch holds %lld bytes, %lld needed", (long long)scratch_bytes,
               (long long)need);
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = lm::grid_for(n, lm::kAdjBlock);
    double2* pos = reinterpret_cast<double2*>(scratch);
    int4* box = reinterpret_cast<int4*>(pos + (int64_t)ckpt * n);
    double2* stage = reinterpret_cast<double2*>(box + nb);
    float* slab = reinterpret_cast<float*>(stage + (nlev == 1 ? 1 : 2 * (int64_t)ckpt + 1) * level_len_of(grid));
    const double2* ck = reinterpret_cast<const double2*>(checkpoints);
    const int64_t slab4 = slab_bytes(grid, nlev) / 16;
    LM_LAUNCH(lm::K_ADJ_SLAB, lm::adj_zero_kernel, dim3(lm::grid_for(slab4, lm::kAdjBlock)), dim3(lm::kAdjBlock), 0, s,
              reinterpret_cast<float4*>(slab), slab4);
    LM_LAUNCH_CHECK();
    const double hh = 0.5 * dt, h6 = dt / 6.0;
    lm::LevelView V{};
    V.glev = reinterpret_cast<const float2*>(field);
    V.level_len = level_len_of(grid);
    V.gstride = nlev > 1 ? V.level_len : 0;
    V.g = *grid;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int s0 = seg * ckpt, m = nsteps - s0 < ckpt ? nsteps - s0 : ckpt;
        int lo = 0x7fffffff, hi = -1;  // level pairs the segment's stage times read
        for (int q = 2 * s0; q <= 2 * (s0 + m); ++q) {
            const int l = lev0_host[q];
            lo = l < lo ? l : lo;
            hi = l > hi ? l : hi;
        }
        lm::AdjField F;
        rc = adj_blend(field, nlev, grid, lev0, w, 2 * s0, 2 * m + 1, stage, s, &F);
        if (rc) return rc;
        LM_LAUNCH(lm::K_ADJ_RECOMPUTE, lm::adj_recompute_kernel, dim3(nb), dim3(lm::kAdjBlock), 0, s, n,
                  ck + (int64_t)seg * n, F, s0, m, hh, dt, h6, pos, box);
        LM_LAUNCH_CHECK();
        LM_LAUNCH(lm::K_ADJ_BACKWARD, lm::adj_backward_kernel, dim3(nb), dim3(lm::kAdjBlock),
                  (unsigned)kAdjWindowBytes, s, n, V, lev0, w, (int)nlev, s0, m, hh, dt, h6, pos, box, glon, glat,
                  gfield, slab, lo, hi - lo + 1, kAdjWindowBytes);
        LM_LAUNCH_CHECK();
    }
    LM_LAUNCH(lm::K_ADJ_SLAB, lm::adj_slab_kernel, dim3(lm::grid_for((int64_t)nlev * level_len_of(grid) * 2, lm::kAdjBlock)),
              dim3(lm::kAdjBlock), 0, s, slab, (int)nlev, *grid, gfield);
    LM_LAUNCH_CHECK();
    return LM_OK;
}

}  // extern "C"
