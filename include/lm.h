/*
 * liblm -- C ABI of the MI355X (gfx950) hot path of lagrangian-microbes:
 * RK4 advection through a gridded space-time velocity field, the fixed-radius
 * pair search, and the rock-paper-scissors species update.
 *
 * Conventions
 *  - Every array argument is a caller-owned DEVICE pointer unless stated.
 *  - `stream` is a hipStream_t passed as void* (NULL = the default stream).
 *    Calls are asynchronous on that stream; none of them synchronises with
 *    the host.
 *  - Return value: LM_OK (0) or a negative LM_ERR_* code; lm_last_error()
 *    then holds a thread-local message.  No C++ exception crosses the ABI.
 *  - Particle state is structure-of-arrays: lon[n], lat[n] in degrees
 *    (longitude 0..360 E), species[n] as int8 (1 rock, 2 paper, 3 scissors).
 *  - A velocity field is u,v (m/s) on a uniform lon/lat grid of nx*ny nodes,
 *    stored node-major with u and v interleaved: f[(j*nx + i)*2 + {0,1}],
 *    i along longitude, j along latitude.
 */
#ifndef LM_H
#define LM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LM_ABI_VERSION 6

#define LM_OK 0
#define LM_ERR_ARG (-1)      /* invalid argument                             */
#define LM_ERR_HIP (-2)      /* HIP runtime error                            */
#define LM_ERR_NOMEM (-3)    /* device allocation failed                     */

int lm_abi_version(void);
const char* lm_last_error(void);

/* ------------------------------------------------------------------------ */
/* Velocity field                                                           */
/* ------------------------------------------------------------------------ */

/* Uniform rectilinear lon/lat grid of a velocity field (degrees). */
typedef struct lm_grid {
    int32_t nx, ny;     /* nodes along longitude, latitude (each >= 2)      */
    double x0, y0;      /* longitude, latitude of node (0, 0)               */
    double inv_dx;      /* 1 / longitude spacing                            */
    double inv_dy;      /* 1 / latitude spacing                             */
} lm_grid;

/* Linear-in-time field at a list of stage times, node by node:
 *   out[s][k] = fma(w[s], levels[l+1][k] - levels[l][k], levels[l][k]),
 *   l = lev0[s], for k < level_len (= 2*nx*ny).
 * lev0 (int32[nstage]) and w (double[nstage]) are device arrays.
 * Replaces the time interpolation of parcels.Field(interp_method="linear")
 * that particle_advecter.py:182-184 sets up over the time axis built at
 * particle_advecter.py:171-173. */
int lm_field_blend_f64(const double* levels, int64_t level_len, const int32_t* lev0,
                       const double* w, int32_t nstage, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* RK4 advection                                                            */
/* ------------------------------------------------------------------------ */

/* Advance n particles by nsteps classic RK4 steps of dt seconds.
 * Step s (0-based) samples the velocity at three stage times: field
 * stage_fields + (2s)*stage_stride at t, + (2s+1)*stage_stride at t+dt/2 and
 * + (2s+2)*stage_stride at t+dt (stage_stride = 0 for a steady field).
 * Interpolation is bilinear in lon/lat; m/s become deg/s through
 * 1852*60 m per degree and cos(lat) for the zonal component.
 * After step s the post-step position is stored as float32 into
 * rec_lon/rec_lat + s*rec_stride (both may be NULL), then, if walk_amp > 0,
 * each coordinate takes a uniform random-walk increment
 * (2u-1)*walk_amp with u from Philox4x64-10(counter = (i, step0+s, 1, 0),
 * key = (key0, key1)), word 0 for lat and word 1 for lon.
 * A particle whose stage position leaves the grid keeps its position for that
 * step and increments *oob_count (device uint32).
 * Replaces pset.execute(AdvectionRK4, runtime=dt, dt=dt)
 * (particle_advecter.py:222-223), the per-step record loop
 * (particle_advecter.py:232-235) and the Kh random walk
 * (particle_advecter.py:240-242). */
int lm_advect_rk4_f64(int64_t n, double* lon, double* lat,
                      const double* stage_fields, int64_t stage_stride, const lm_grid* grid,
                      int32_t nsteps, double dt,
                      float* rec_lon, float* rec_lat, int64_t rec_stride,
                      double walk_amp, uint64_t key0, uint64_t key1, uint64_t step0,
                      uint32_t* oob_count, void* stream);

/* FP32 path: the same blend on float32 levels (fmaf((float)w, f1 - f0, f0)),
 * for lm_advect_rk4_f32.  Same call site as lm_field_blend_f64. */
int lm_field_blend_f32(const float* levels, int64_t level_len, const int32_t* lev0,
                       const double* w, int32_t nstage, float* out, void* stream);

/* FP32 path of lm_advect_rk4_f64: same arguments and semantics with float32
 * stage fields.  Interpolation, the m/s -> deg/s conversion (cos(lat) as a
 * float32 polynomial) and the RK4 stage velocities are float32; positions,
 * stage positions, the cell lookup and the accumulation
 * x += h/6 (k1 + 2 k2 + 2 k3 + k4) are float64.  Replaces the same call as
 * lm_advect_rk4_f64 (particle_advecter.py:222-242) for large runs. */
int lm_advect_rk4_f32(int64_t n, double* lon, double* lat,
                      const float* stage_fields, int64_t stage_stride, const lm_grid* grid,
                      int32_t nsteps, double dt,
                      float* rec_lon, float* rec_lat, int64_t rec_stride,
                      double walk_amp, uint64_t key0, uint64_t key1, uint64_t step0,
                      uint32_t* oob_count, void* stream);

/* ------------------------------------------------------------------------ */
/* Fixed-radius pair search                                                 */
/* ------------------------------------------------------------------------ */

typedef struct lm_pairs_ws lm_pairs_ws;

/* Workspace for up to n_max particles and up to steps_max position snapshots
 * per call, expected inside the lon/lat box [lon_min, lon_max] x
 * [lat_min, lat_max] (particles outside are still handled exactly, only more
 * slowly), for radii up to max_radius (deg).  Allocates device memory on the
 * current device (about 80 bytes per particle and snapshot). */
int lm_pairs_ws_create(int64_t n_max, int32_t steps_max, double lon_min, double lon_max,
                       double lat_min, double lat_max, double max_radius, lm_pairs_ws** out);
void lm_pairs_ws_destroy(lm_pairs_ws* ws);
/* Device bytes held by the workspace. */
int64_t lm_pairs_ws_bytes(const lm_pairs_ws* ws);
/* The same with cells of side cell_scale * max_radius (1 <= cell_scale <= 16;
 * 1 is lm_pairs_ws_create): larger cells hold more candidates each but need
 * fewer table entries, a trade the ensemble's autotuner measures.  No
 * reference counterpart (cKDTree has no cell size). */
int lm_pairs_ws_create_ex(int64_t n_max, int32_t steps_max, double lon_min, double lon_max,
                          double lat_min, double lat_max, double max_radius, double cell_scale,
                          lm_pairs_ws** out);
/* Number of cell buckets T of the workspace's cell list (a power of two). */
int64_t lm_pairs_ws_buckets(const lm_pairs_ws* ws);

/* All pairs (i, j), i < j, with ||x_i - x_j||_p <= radius (closed ball) in
 * degree space, where positions are the float32 lon/lat widened to double.
 * p = 2: dx*dx + dy*dy <= radius*radius; p = 1: |dx| + |dy| <= radius;
 * p = +inf: max(|dx|, |dy|) <= radius; other p >= 1:
 * |dx|^p + |dy|^p <= radius^p.
 * Output is sorted lexicographically by (i, j).  *npairs (device int64)
 * receives the pair count P; the pairs are written only if P <= capacity
 * (otherwise the caller enlarges its buffers and calls again).
 * Replaces cKDTree(points).query_pairs(r=interaction_radius,
 * p=interaction_norm) (interaction_simulator.py:93,98) followed by iteration
 * over the returned set (interaction_simulator.py:104). */
int lm_query_pairs(lm_pairs_ws* ws, int64_t n, const float* lon, const float* lat,
                   double radius, double p, int32_t* pair_i, int32_t* pair_j,
                   int64_t capacity, int64_t* npairs, void* stream);

/* The same for nsteps snapshots at once: snapshot s is lon/lat + s*pos_stride,
 * its pairs go to pair_i/pair_j + s*pair_stride (capacity pair_stride) and
 * its count to npairs[s].  One launch per phase covers every snapshot.
 * Replaces the per-step query of interaction_simulator.py:88-98 over a run of
 * steps. */
int lm_query_pairs_batch(lm_pairs_ws* ws, int32_t nsteps, int64_t n, const float* lon,
                         const float* lat, int64_t pos_stride, double radius, double p,
                         int32_t* pair_i, int32_t* pair_j, int64_t pair_stride, int64_t* npairs,
                         void* stream);

/* Ensemble variant: particle k belongs to run run[k] (device uint16[n]); runs
 * are independent, so only pairs of one run are returned.  Each run's pairs
 * are exactly those lm_query_pairs_batch finds on that run alone (indices
 * shifted by the run's first index); the list stays sorted by (i, j).
 * Replaces the query of interaction_simulator.py:88-98 for many runs at
 * once. */
int lm_query_pairs_runs(lm_pairs_ws* ws, int32_t nsteps, int64_t n, const float* lon,
                        const float* lat, int64_t pos_stride, const uint16_t* run, double radius,
                        double p, int32_t* pair_i, int32_t* pair_j, int64_t pair_stride,
                        int64_t* npairs, void* stream);

/* ------------------------------------------------------------------------ */
/* Rock-paper-scissors update                                               */
/* ------------------------------------------------------------------------ */

typedef struct lm_rps_ws lm_rps_ws;

/* Workspace for up to n_max particles and steps_max steps per prepare. */
int lm_rps_ws_create(int64_t n_max, int32_t steps_max, lm_rps_ws** out);
void lm_rps_ws_destroy(lm_rps_ws* ws);
int64_t lm_rps_ws_bytes(const lm_rps_ws* ws);

/* Apply the rock-paper-scissors rule to the sorted pair list, with exactly the
 * result of visiting the pairs one after another in (i, j) order: if the two
 * species differ, draw u (pair of rank k uses word 0 of
 * Philox4x64-10(counter = (k, step, 0, 0), key = (key0, key1)) mapped to
 * [0,1) as (x >> 11) * 2^-53); the forward winner (rock over scissors with
 * pRS, paper over rock with pPR, scissors over paper with pSP) wins iff
 * u < p, and the loser takes the winner's species.  Species outside {1,2,3}
 * never interact.  *npairs is read on the device; if it exceeds capacity the
 * call changes nothing.  The list must be sorted by (i, j), as lm_query_pairs
 * returns it (the reference visits sorted(set(pairs))); pairs with an index
 * outside [0, n) are skipped.
 * Replaces the per-pair loop interaction_simulator.py:104-105 calling
 * rock_paper_scissors_interaction (interactions.py:13-40). */
int lm_rps_update(lm_rps_ws* ws, int64_t n, int8_t* species, const int32_t* pair_i,
                  const int32_t* pair_j, const int64_t* npairs, int64_t capacity,
                  double pRS, double pPR, double pSP, uint64_t key0, uint64_t key1,
                  uint64_t step, void* stream);

/* The same rule over nsteps consecutive steps, split into the part that only
 * depends on the pairs and the part that needs the species:
 *   lm_rps_prepare  for pair lists pair_i/pair_j + s*pair_stride with counts
 *                   npairs[s] (s < nsteps) at global steps step0 + s: groups
 *                   the pairs by connected component in the order they are to
 *                   be visited and draws their uniforms -- all steps at once;
 *   lm_rps_play     step s of the last prepare: species_out = species_in after
 *                   that step's pairs (species_in == species_out is allowed).
 * lm_rps_update(.., step) is lm_rps_prepare(nsteps = 1, step0 = step) followed
 * by lm_rps_play(0, species, species).  A step whose count exceeds
 * pair_stride is left unchanged by play. */
int lm_rps_prepare(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i,
                   const int32_t* pair_j, int64_t pair_stride, const int64_t* npairs,
                   double pRS, double pPR, double pSP, uint64_t key0, uint64_t key1,
                   uint64_t step0, void* stream);
/* Launch shape of lm_rps_play: workgroups (64 lanes) that take wide
 * components (one wavefront each, by level) and that take small components
 * (one lane each); both loop over their components, so any counts >= 1 give
 * the same result (defaults 2048 and 2048).  No reference counterpart (set by
 * the ensemble's autotuner). */
int lm_rps_ws_set_play_blocks(lm_rps_ws* ws, int32_t wide_blocks, int32_t narrow_blocks);

/* Size the workspace's per-pair arrays for pairs_per_step pairs per step now
 * (they are grown, never shrunk), so that a later prepare with that pair
 * capacity does not synchronise `stream` to grow them.  No reference
 * counterpart (buffer management of a pipelined caller). */
int lm_rps_ws_reserve(lm_rps_ws* ws, int64_t pairs_per_step, void* stream);

/* Per-run parameters of an ensemble (a device array, one entry per run): the
 * Philox key and the forward probabilities. */
typedef struct lm_run_params {
    uint64_t key0, key1;
    double pRS, pPR, pSP;
} lm_run_params;

/* Ensemble variant of lm_rps_prepare: run r holds particles
 * [run_start[r], run_start[r+1]) (device int64[nruns+1]; run[k] (device
 * uint16[n]) is the same map per particle) and the pair lists hold pairs of
 * one run each (lm_query_pairs_runs).  A pair of run r draws word 0 of
 * Philox4x64-10(counter = (its rank among run r's pairs of the step,
 * step0 + s, 0, 0), key = params[r].key) and uses params[r]'s p: every run's
 * species evolve exactly as lm_rps_prepare/lm_rps_play give on that run
 * alone.  No reference counterpart: the reference runs ensembles as separate
 * processes (SLURM scripts). */
int lm_rps_prepare_runs(lm_rps_ws* ws, int32_t nsteps, int64_t n, const int32_t* pair_i,
                        const int32_t* pair_j, int64_t pair_stride, const int64_t* npairs,
                        const uint16_t* run, int32_t nruns, const int64_t* run_start,
                        const lm_run_params* params, uint64_t step0, void* stream);

int lm_rps_play(lm_rps_ws* ws, int32_t step, int64_t n, const int8_t* species_in,
                int8_t* species_out, void* stream);

/* Size statistics of the interaction components of step `step` of the last
 * prepare, written to device int64[8]: components with at least one pair,
 * largest component (pairs), largest component (particles), total pairs,
 * components replayed by a wavefront (more than 32 pairs), their pairs, their
 * particles, particles with at least one pair.  Diagnostic only. */
int lm_rps_component_stats(lm_rps_ws* ws, int32_t step, int64_t* out8, void* stream);

/* ------------------------------------------------------------------------ */
/* Counter-based uniforms                                                   */
/* ------------------------------------------------------------------------ */

/* out[k] = word `word` of Philox4x64-10(counter = (c0 + k, c1, c2, c3),
 * key = (key0, key1)) mapped to [0,1) as (x >> 11) * 2^-53, k < count.
 * The same uniforms lm_rps_update and lm_advect_rk4_f64 draw, exposed so
 * they can be fed to a CPU run of the rule.  No reference counterpart: the
 * reference draws from numpy's global generator (interactions.py:20). */
int lm_philox_uniforms(int64_t count, uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                       uint64_t key0, uint64_t key1, int32_t word, double* out, void* stream);

/* ------------------------------------------------------------------------ */
/* Reductions for long runs                                                 */
/* ------------------------------------------------------------------------ */

/* counts[s][c] += number of particles of row s (species + s*stride, int8,
 * n entries) with code c for c = 1 (rock), 2 (paper), 3 (scissors), and
 * with any other code for c = 0 (device uint64 [nsteps][4]).  Replaces the
 * per-time counts of species_count_figure (analysis.py:31-35). */
int lm_species_counts(const int8_t* species, int64_t n, int32_t nsteps, int64_t stride, uint64_t* counts,
                      void* stream);

/* hist[b] += number of unordered pairs i < j whose haversine term
 *   a = sin^2((lat_j - lat_i)/2) + cos(lat_i) cos(lat_j) sin^2((lon_j - lon_i)/2)
 * (degrees; float32 positions widened to double) has exactly b of the
 * nbins - 1 increasing thresholds edges[] <= a (device double).  With
 * thresholds a_b = sin^2(d_b / (2 R)) this is the histogram of the pairwise
 * great-circle distances over the edges d_b.  Replaces the pairwise distance
 * histogram of sandbox/pairwise_distance_histogram_gpu.jl:14-43 (whose
 * binning, round(10 log10(d)), the caller's edges reproduce). */
int lm_distance_histogram(int64_t n, const float* lon, const float* lat, int32_t nbins, const double* edges,
                          uint64_t* hist, void* stream);

/* ------------------------------------------------------------------------ */
/* Adjoint of the RK4 advection with respect to a float32 gridded field and */
/* the initial positions.  No reference counterpart: the reference only     */
/* advects forward (AdvectionRK4 through parcels, particle_advecter.py:     */
/* 177-187,222-223); the forward here is that advection, and the backward   */
/* is its hand-written reverse, for callers that fit u, v to trajectories.  */
/* ------------------------------------------------------------------------ */

/* The field is float32 [nlev][ny][nx][2] (u, v interleaved, m/s).  Stage
 * time k (k = 0 .. 2*nsteps, at t0 + k*dt/2) reads levels lev0[k] and
 * lev0[k] + 1 (level 0 only when nlev == 1) blended with weight w[k]:
 * fma(w, f1 - f0, f0) on the values widened to double, so the positions are
 * bit-identical to lm_advect_rk4_f64 over the widened levels blended by
 * lm_field_blend_f64.  lev0 and w are device arrays of 2*nsteps + 1 entries;
 * lev0_host is a host copy of lev0, checked against nlev before any launch.
 * A step with a stage outside the grid keeps its position and adds one to
 * *oob_count.  Both calls blend the stage times they need into float64
 * stage fields (16 bytes per node) in the caller's scratch. */

/* Advances lon/lat (in place) by nsteps RK4 steps of dt seconds and stores
 * the positions at the start of steps 0, ckpt, 2*ckpt, ... as interleaved
 * (lon, lat) doubles: checkpoints[(s/ckpt)*n + i][2], ceil(nsteps/ckpt)*n*2
 * doubles, 16-byte aligned.  scratch (16-byte aligned) holds the stage
 * fields of as many steps as fit: at least 3 fields (1 when nlev == 1);
 * 2*nsteps + 1 fields run the whole forward in one launch. */
int lm_adjoint_forward(int64_t n, double* lon, double* lat, const float* field, int32_t nlev, const lm_grid* grid,
                       const int32_t* lev0, const double* w, const int32_t* lev0_host, int32_t nsteps, double dt,
                       int32_t ckpt, double* checkpoints, void* scratch, int64_t scratch_bytes, uint32_t* oob_count,
                       void* stream);

/* lm_adjoint_forward without the scratch argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No reference counterpart.  Calls no
 * hipMalloc, hipFree or allocating library routine: the stage fields go to a
 * static device array of 32 MiB per device, so the grid may have at most
 * 699,050 nodes (2,097,152 when nlev == 1); larger grids are rejected with
 * LM_ERR_ARG.  The array is shared by every call on a device and keeps the
 * last call's contents: calls that overlap on one device race.  Results are
 * those of lm_adjoint_forward with the same arguments. */
int lm_adjoint_forward_noalloc(int64_t n, double* lon, double* lat, const float* field, int32_t nlev,
                               const lm_grid* grid, const int32_t* lev0, const double* w, const int32_t* lev0_host,
                               int32_t nsteps, double dt, int32_t ckpt, double* checkpoints, uint32_t* oob_count,
                               void* stream);

/* Bytes of device scratch lm_adjoint_backward needs for n particles,
 * checkpoint interval ckpt and the field's grid and level count (-1 on
 * invalid arguments). */
int64_t lm_adjoint_scratch_bytes(int64_t n, int32_t ckpt, const lm_grid* grid, int32_t nlev);

/* Reverse pass over the run that lm_adjoint_forward made with the same
 * arguments.  On entry glon/glat hold dL/d(final lon, lat); on return
 * dL/d(initial lon, lat).  dL/d(field) is ADDED into gfield (float32, the
 * field's layout; the caller zeroes it).  Each segment of ckpt steps is
 * recomputed from its checkpoint into scratch (16-byte aligned, at least
 * lm_adjoint_scratch_bytes(n, ckpt, grid, nlev) bytes), then swept in
 * reverse. */
int lm_adjoint_backward(int64_t n, const double* checkpoints, const float* field, int32_t nlev, const lm_grid* grid,
                        const int32_t* lev0, const double* w, const int32_t* lev0_host, int32_t nsteps, double dt,
                        int32_t ckpt, void* scratch, int64_t scratch_bytes, double* glon, double* glat, float* gfield,
                        void* stream);

/* ------------------------------------------------------------------------ */
/* Plain streaming copy, used to calibrate hardware byte counters.          */
/* No reference counterpart.                                                */
/* ------------------------------------------------------------------------ */
int lm_copy_bytes(const void* src, void* dst, int64_t nbytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Segmented stable radix sort of (uint32 key, uint32 value) pairs: the     */
/* grouping primitive under the pair search and the RPS grouping.  Exposed  */
/* for testing.  No reference counterpart (the reference sorts with numpy / */
/* set iteration, interaction_simulator.py:74).                             */
/* ------------------------------------------------------------------------ */

/* Sorts nseg segments in place (segment g at keys + g*stride, vals +
 * g*stride) by bits [0, key_bits) of the key, stably.  Segment g holds
 * seg_len[g] elements (device int64 array; uniform_len if seg_len is NULL),
 * at most max_len <= stride.  keys_tmp / vals_tmp are buffers of the same
 * layout.  Allocates its scratch and synchronises the stream. */
int lm_sort_pairs_u32(int32_t nseg, const int64_t* seg_len, int64_t uniform_len, int64_t max_len, int64_t stride,
                      int32_t key_bits, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      void* stream);

/* ------------------------------------------------------------------------ */
/* Per-kernel timing.  No reference counterpart (the reference times phases */
/* with time() around each call, particle_advecter.py:221-244,              */
/* interaction_simulator.py:91-106).                                        */
/* ------------------------------------------------------------------------ */

/* Turn timing on (clearing earlier records) or off.  While on, every kernel
 * launch records start/stop events on its own dispatch. */
int lm_timing_enable(int on);
/* Waits for the recorded launches and writes, per kernel id k < nkinds, the
 * launch count and total milliseconds.  Returns the number of kernel ids. */
int lm_timing_summary(int64_t* counts, double* total_ms, int32_t nkinds);
/* Name of kernel id k ("" if out of range). */
const char* lm_timing_kernel_name(int32_t kid);

#ifdef __cplusplus
}
#endif

#endif /* LM_H */
