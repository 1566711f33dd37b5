/*
 * swe2d_hip.h -- C ABI of the MI355X (gfx950) shallow-water time step.
 *
 * The reference (DiffSWE2d) exposes this path only as a Python object API:
 * SWEConfig (src/config.py:4-28) and SWE2D.spatial_operator / step / forward
 * (src/solver.py:448-779).  It has no FFI.  The entry points below are what a
 * binding of that API needs; each cites the reference function it replaces.
 *
 * Conventions
 *  - All arrays are device pointers, row-major, float32 (_f32) or float64 (_f64).
 *  - State U is [batch, 3, ny, nx] (planes h, hu, hv per member), exactly the
 *    reference layout, so torch tensors pass without copies.
 *  - bed / roughness are [bed_batch, 1, ny, nx] with bed_batch 1 or batch.
 *  - level is the water-level boundary value: one value (level_stride 0) or one
 *    per member (level_stride 1) -- boundary.py:113-147.
 *  - Calls are stream-ordered on `stream` (a hipStream_t) and allocate nothing;
 *    scratch lives in a caller-provided workspace of swe_workspace_bytes().
 *  - Return 0 on success, a negative SWE_ERR_* code otherwise; no exceptions
 *    cross the ABI.  swe_last_error_message() is per calling thread.
 */
#ifndef SWE2D_HIP_H
#define SWE2D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_ABI_VERSION 5

enum {
  SWE_OK = 0,
  SWE_ERR_ARG = -1,
  SWE_ERR_SHAPE = -2,
  SWE_ERR_BC = -3,
  SWE_ERR_HIP = -4,
  SWE_ERR_DT_NONFINITE = -5,
  SWE_ERR_DT_NONPOSITIVE = -6,
  SWE_ERR_CAPACITY = -7
};

/* Boundary type codes; names as in boundary.py:11-18. */
enum {
  SWE_BC_TRANSMISSIVE = 0,
  SWE_BC_PERIODIC = 1,
  SWE_BC_CONSTANT = 2,
  SWE_BC_WATER_LEVEL = 3,
  SWE_BC_WALL = 4,
  SWE_BC_ABSORBING = 5
};

/* Friction models: SWEConfig.frictionmodel (solver.py:648-661). */
enum { SWE_FRICTION_MANNING = 0, SWE_FRICTION_ROUGHNESS_LENGTH = 1 };

typedef struct swe_grid {
  int32_t batch, ny, nx;
  int32_t bed_batch;   /* 1 or batch */
  int32_t rough_batch; /* 1 or batch */
  int32_t pad;
  double dx, dy;       /* SWE2D(dx, dy): solver.py:64-67 */
} swe_grid;

/* SWEConfig field for field (config.py:4-28); strings become codes. */
typedef struct swe_config {
  double gravity, dry_depth, epsilon, cfl, limiter_theta;
  double constant_value;          /* "constant" boundary value */
  int32_t boundary[4];            /* left, right, top, bottom */
  int32_t water_slope_reset;
  int32_t positivity_limiter;
  int32_t apply_friction;
  int32_t friction_model;
  int32_t grad_bed;               /* adjoint calls: also accumulate the bed adjoint (see
                                     swe_adjoint_bed_grad_offset); 0 elsewhere */
  int32_t arithmetic;             /* SWE_ARITH_REFERENCE (default): the reference's rounding.
                                     SWE_ARITH_FAST: FP32 forward steps with division from
                                     the reciprocal estimate and one correction, hardware
                                     square root, h^(1/3) from log2 / exp2, face velocities
                                     without re-division -- each within a few ulps, so
                                     results agree with the reference to a tolerance, not
                                     bit for bit.  Ignored by FP64, by steps that record
                                     k1 / k2 or a tape, by chunked runs and by the adjoint. */
  int32_t fixed_dt;               /* 1: host-driven steps (swe_step_*, swe_step_stage_f32,
                                     swe_step_forced_*, swe_step_record_*, swe_step_noalloc)
                                     take max_dt itself as dt instead of min(CFL dt, max_dt):
                                     the dt-frozen map whose derivative the reference's
                                     detached dt defines (src/solver.py:625-626), for finite-
                                     difference checks of gradients (the reference has no
                                     such switch).  Device-driven runs reject it.  0 (the
                                     reference's behaviour) otherwise. */
} swe_config;

/* swe_config.arithmetic */
enum { SWE_ARITH_REFERENCE = 0, SWE_ARITH_FAST = 1 };

/* Device control block at the head of the workspace (read back by the host). */
/* Source terms of a step (solver.py:520-525, 609-641): a rainfall record
 * (RainfallForcing, forcing.py:5-38: rates [nt, ny, nx] at increasing times
 * [nt], both in the run's dtype, shared by all members), sampled by stage 1 at
 * the step's model time and by stage 2 at time + dt/2; and an infiltration
 * field ([ny, nx] with infil_bstride 0, or [B, ny, nx] with ny*nx), clamped at 0
 * and subtracted from the depth tendency.  `time` is the model time of a
 * host-driven step, or of step 0 of a device-driven run (step k then uses time
 * + the run's elapsed time, rounded to the run's dtype as the reference's
 * U.new_tensor(current_time)). */
typedef struct swe_forcing {
  const void* rain_times;
  const void* rain_rates;
  int32_t rain_nt, pad;
  const void* infiltration;
  int64_t infil_bstride;
  double time;
} swe_forcing;

typedef struct swe_control {
  double elapsed[2];
  double duration, tol, cap, bad_dt;
  int32_t done, steps, status, pad;
  double last_dt;
  int32_t chunked, chunk, n_chunks, pad2;
  double level_f64;
  float level_f32, pad3;
  const void* levels;
} swe_control;

int swe_abi_version(void);
const char* swe_strerror(int code);
const char* swe_last_error_message(void);
size_t swe_control_bytes(void);

/* Workspace size for a grid (dtype_bytes 4 or 8), and its one-time clearing. */
size_t swe_workspace_bytes(const swe_grid* grid, int dtype_bytes);
int swe_workspace_init(void* ws, size_t bytes, void* stream);

/* Tendency dU/dt and wave speeds -- replaces SWE2D.spatial_operator
 * (solver.py:448-531) with ghost cells, both reconstructions and HLL fused.
 * rain / infiltration are optional [ny, nx] rate fields (NULL for none). */
int swe_spatial_operator_f32(const swe_grid* grid, const swe_config* cfg, const float* U, const float* bed,
                             const float* level, int level_stride, const float* rain, const float* infiltration,
                             float* rhs, void* ws, int rows_hint, void* stream);
int swe_spatial_operator_f64(const swe_grid* grid, const swe_config* cfg, const double* U, const double* bed,
                             const double* level, int level_stride, const double* rain,
                             const double* infiltration, double* rhs, void* ws, int rows_hint, void* stream);

/* One adaptive midpoint step -- replaces SWE2D.step (solver.py:572-663).
 * The step's dt lands in swe_control.last_dt; U_out is left untouched when dt
 * is non-finite or non-positive (the host raises FloatingPointError, :627-632).
 * parity alternates 0/1 between consecutive steps on one workspace. */
int swe_step_f32(const swe_grid* grid, const swe_config* cfg, const float* U, const float* bed,
                 const float* roughness, const float* level, int level_stride, double max_dt, int parity,
                 float* U_out, void* ws, int rows_hint, void* stream);
int swe_step_f64(const swe_grid* grid, const swe_config* cfg, const double* U, const double* bed,
                 const double* roughness, const double* level, int level_stride, double max_dt, int parity,
                 double* U_out, void* ws, int rows_hint, void* stream);

/* The two launches of swe_step_f32 separately (stage 1: k1 and wave speeds
 * into the workspace; stage 2: dt, predictor, k2, update, friction), so a
 * caller can bracket each kernel with events on `stream`. */
int swe_step_stage_f32(const swe_grid* grid, const swe_config* cfg, int stage, const float* U, const float* bed,
                       const float* roughness, const float* level, int level_stride, double max_dt, int parity,
                       float* U_out, void* ws, int rows_hint, void* stream);

/* swe_step_f32 / swe_step_f64 (dtype_bytes 4 or 8) without a workspace argument
 * and without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Its scratch (control block,
 * wave-speed banks, first-stage tendency) is a fixed 64 MiB device array that
 * every GPU the library runs on holds until the process ends; the call rejects
 * a grid whose batch * ny * nx exceeds 5,592,128 cells (FP32) or 2,796,064
 * cells (FP64) with SWE_ERR_CAPACITY.  dt_out (a device pointer to one value of
 * the run's dtype) receives the step's dt whether or not it is usable; as in
 * swe_step_*, U_out is left untouched when dt is non-finite or non-positive, and
 * the caller raises as solver.py:627-632 does.  Every call on one GPU shares
 * that GPU's scratch, so calls that overlap on one GPU (several streams or host
 * threads at once) race; calls in stream order do not.  The current device must
 * be the stream's. */
int swe_step_noalloc(const swe_grid* grid, const swe_config* cfg, int dtype_bytes, const void* U, const void* bed,
                     const void* roughness, const void* level, int level_stride, double max_dt, void* U_out,
                     void* dt_out, int rows_hint, void* stream);

/* swe_step_*, additionally recording the two stage tendencies k1 and k2
 * (each [batch, 3, ny, nx]) for the adjoint's checkpoint replay. */
int swe_step_record_f32(const swe_grid* grid, const swe_config* cfg, const float* U, const float* bed,
                        const float* roughness, const float* level, int level_stride, double max_dt, int parity,
                        float* U_out, float* k1_out, float* k2_out, void* ws, int rows_hint, void* stream);
int swe_step_record_f64(const swe_grid* grid, const swe_config* cfg, const double* U, const double* bed,
                        const double* roughness, const double* level, int level_stride, double max_dt, int parity,
                        double* U_out, double* k1_out, double* k2_out, void* ws, int rows_hint, void* stream);

/* Device-driven integration of a duration -- replaces the SWE2D.forward loop
 * (solver.py:719-779) without per-step host synchronisation.  swe_run_begin
 * arms the device clock; each swe_advance call launches n_steps further steps
 * (step k reads U_a when k is even, U_b when odd, and writes the other).  Steps
 * after the duration is reached (or after a bad dt) exit at once.  Optional
 * outputs: running max depth hmax [batch, ny, nx] (:765), per-step dt
 * (:766), member-0 depth and h+z at gauge cells (iy, ix pairs) (:781-801). */
int swe_run_begin(void* ws, double duration, double max_dt_cap, int dtype_bytes, void* stream);
int swe_advance_f32(const swe_grid* grid, const swe_config* cfg, const float* bed, const float* roughness,
                    const float* level, int level_stride, float* U_a, float* U_b, int first_step, int n_steps,
                    float* hmax, float* dt_series, int dt_capacity, const int32_t* gauges, int n_gauges,
                    float* gauge_depth, float* gauge_level, int gauge_capacity, void* ws, int rows_hint,
                    void* stream);
int swe_advance_f64(const swe_grid* grid, const swe_config* cfg, const double* bed, const double* roughness,
                    const double* level, int level_stride, double* U_a, double* U_b, int first_step,
                    int n_steps, double* hmax, double* dt_series, int dt_capacity, const int32_t* gauges,
                    int n_gauges, double* gauge_depth, double* gauge_level, int gauge_capacity, void* ws,
                    int rows_hint, void* stream);

/* swe_step_* / swe_advance_* with source terms (swe_forcing); k1_out and k2_out
 * (both or neither) record the stage tendencies like swe_step_record_*. */
int swe_step_forced_f32(const swe_grid* grid, const swe_config* cfg, const swe_forcing* forcing, const float* U,
                        const float* bed, const float* roughness, const float* level, int level_stride,
                        double max_dt, int parity, float* U_out, float* k1_out, float* k2_out, void* ws,
                        int rows_hint, void* stream);
int swe_step_forced_f64(const swe_grid* grid, const swe_config* cfg, const swe_forcing* forcing, const double* U,
                        const double* bed, const double* roughness, const double* level, int level_stride,
                        double max_dt, int parity, double* U_out, double* k1_out, double* k2_out, void* ws,
                        int rows_hint, void* stream);
int swe_advance_forced_f32(const swe_grid* grid, const swe_config* cfg, const swe_forcing* forcing,
                           const float* bed, const float* roughness, const float* level, int level_stride,
                           float* U_a, float* U_b, int first_step, int n_steps, float* hmax, float* dt_series,
                           int dt_capacity, const int32_t* gauges, int n_gauges, float* gauge_depth,
                           float* gauge_level, int gauge_capacity, void* ws, int rows_hint, void* stream);
int swe_advance_forced_f64(const swe_grid* grid, const swe_config* cfg, const swe_forcing* forcing,
                           const double* bed, const double* roughness, const double* level, int level_stride,
                           double* U_a, double* U_b, int first_step, int n_steps, double* hmax, double* dt_series,
                           int dt_capacity, const int32_t* gauges, int n_gauges, double* gauge_depth,
                           double* gauge_level, int gauge_capacity, void* ws, int rows_hint, void* stream);

/* Dynamic boundary series of a device-driven run -- replaces the per-step
 * interpolation of SWE2D._interpolate_dynamic_boundary (solver.py:391-415) in
 * the forward loop (:751-757) with a one-thread launch before each step.  The
 * level of step k is the series at start_time + the run's elapsed time, in the
 * run's dtype (times and levels arrays of that dtype); it holds for both stages
 * of the step.  step_level / step_index / step_weight (optional, capacity
 * entries) record each step's level, interval index (searchsorted clamped to
 * [1, n - 1]; 0 when n = 1) and interpolation weight, for the adjoint. */
typedef struct swe_dynamic_level {
  const void* times;
  const void* levels;
  int32_t n, pad;
  double start_time;
  void* step_level;
  int32_t* step_index;
  void* step_weight;
} swe_dynamic_level;

/* swe_advance_* / swe_advance_tape_* (tape_k1 and tape_k2 both given: U_a is
 * the state tape as in swe_advance_tape_*) with the boundary level taken from a
 * dynamic series instead of a level argument; forcing is optional (NULL).
 * capacity bounds dt_series, the gauge records and the level records. */
int swe_advance_dynamic_f32(const swe_grid* grid, const swe_config* cfg, const swe_dynamic_level* dyn,
                            const swe_forcing* forcing, const float* bed, const float* roughness, float* U_a,
                            float* U_b, float* tape_k1, float* tape_k2, int tape_first, int first_step, int n_steps,
                            float* hmax, float* dt_series, int capacity, const int32_t* gauges, int n_gauges,
                            float* gauge_depth, float* gauge_level, void* ws, int rows_hint, void* stream);
int swe_advance_dynamic_f64(const swe_grid* grid, const swe_config* cfg, const swe_dynamic_level* dyn,
                            const swe_forcing* forcing, const double* bed, const double* roughness, double* U_a,
                            double* U_b, double* tape_k1, double* tape_k2, int tape_first, int first_step,
                            int n_steps, double* hmax, double* dt_series, int capacity, const int32_t* gauges,
                            int n_gauges, double* gauge_depth, double* gauge_level, void* ws, int rows_hint,
                            void* stream);

/* Chunked run -- replaces the inversion driver's chunk loop
 * (bc_inversion_train.py:332-349, 401-410: forward(U, t_end=chunk,
 * boundary_level=series[c]) per chunk, then h + z at the gauges) with one
 * device-driven run.  swe_run_begin_chunked arms n_chunks chunks of
 * chunk_seconds each with per-chunk levels (device array of the run's dtype);
 * each swe_advance_chunked_* step is stage 1, stage 2 and a one-block clock
 * launch that, at a chunk end, writes obs[chunk][gauge] = h + z (member 0;
 * batch 1) and moves to the next chunk's level.  chunk_series[k] receives the
 * chunk of step k.  With tape_k1/tape_k2 the states and stage tendencies are
 * recorded as in swe_advance_tape_* (U_a is then the state tape, U_b unused);
 * without them U_a/U_b alternate as in swe_advance_*. */
int swe_run_begin_chunked(void* ws, double chunk_seconds, int n_chunks, const void* levels, int dtype_bytes,
                          void* stream);
int swe_advance_chunked_f32(const swe_grid* grid, const swe_config* cfg, const float* bed, const float* roughness,
                            float* U_a, float* U_b, float* tape_k1, float* tape_k2, int tape_first, int first_step,
                            int n_steps, float* dt_series, int32_t* chunk_series, int capacity,
                            const int32_t* gauges, int n_gauges, float* obs, void* ws, int rows_hint, void* stream);
int swe_advance_chunked_f64(const swe_grid* grid, const swe_config* cfg, const double* bed, const double* roughness,
                            double* U_a, double* U_b, double* tape_k1, double* tape_k2, int tape_first,
                            int first_step, int n_steps, double* dt_series, int32_t* chunk_series, int capacity,
                            const int32_t* gauges, int n_gauges, double* obs, void* ws, int rows_hint,
                            void* stream);

/* swe_advance_* that also records the adjoint tape of every step: tape_U holds
 * the states (slot j = step tape_first + j; slot 0 is the input, step k writes
 * slot k - tape_first + 1), tape_k1 / tape_k2 the two stage tendencies (slot
 * k - tape_first).  The layout [tape_U | tape_k1 | tape_k2] of one segment is
 * what swe_segment_reverse_* reads, so a forward that can afford the memory
 * skips the backward's replay. */
int swe_advance_tape_f32(const swe_grid* grid, const swe_config* cfg, const float* bed, const float* roughness,
                         const float* level, int level_stride, float* tape_U, float* tape_k1, float* tape_k2,
                         int tape_first, int first_step, int n_steps, float* hmax, float* dt_series, int dt_capacity,
                         void* ws, int rows_hint, void* stream);
int swe_advance_tape_f64(const swe_grid* grid, const swe_config* cfg, const double* bed, const double* roughness,
                         const double* level, int level_stride, double* tape_U, double* tape_k1, double* tape_k2,
                         int tape_first, int first_step, int n_steps, double* hmax, double* dt_series,
                         int dt_capacity, void* ws, int rows_hint, void* stream);

/* ---- adjoint (reverse mode) --------------------------------------------
 * The reference differentiates step() with torch autograd (dt detached,
 * solver.py:625-626; caller bc_inversion_train.py:332-452).  These entry
 * points are its hand-written replacement.  lam_level receives, per member,
 * the gradient with respect to the water-level boundary value and is
 * ACCUMULATED into (+=), so a reverse sweep over many steps sums it. */
size_t swe_adjoint_workspace_bytes(const swe_grid* grid, int dtype_bytes);
const char* swe_adjoint_last_error_message(void);

/* Byte offset, in the adjoint workspace, of the bed adjoint [batch, ny, nx]
 * (the gradient with respect to SWE2D.bed, through which the reference's
 * autograd reaches dem_correction_parameter, solver.py:100-110).  Every
 * adjoint call made with cfg->grad_bed set ADDS its contribution there; the
 * caller clears it before a backward pass and reads it after. */
size_t swe_adjoint_bed_grad_offset(const swe_grid* grid, int dtype_bytes);

/* L^T: adjoint of swe_spatial_operator with respect to U (and the level). */
int swe_operator_adjoint_f32(const swe_grid* grid, const swe_config* cfg, const float* U, const float* bed,
                             const float* level, int level_stride, const float* lam_rhs, float* lam_U,
                             float* lam_level, void* adj_ws, int rows_hint, void* stream);
int swe_operator_adjoint_f64(const swe_grid* grid, const swe_config* cfg, const double* U, const double* bed,
                             const double* level, int level_stride, const double* lam_rhs, double* lam_U,
                             double* lam_level, void* adj_ws, int rows_hint, void* stream);

/* Adjoint of one step taken with dt: given the step-start state U, the
 * recorded k1, k2 and the adjoint lam_out of the step's output, writes the
 * adjoint lam_U of U (lam_U must not alias the other arrays). */
int swe_step_adjoint_f32(const swe_grid* grid, const swe_config* cfg, const float* U, const float* k1,
                         const float* k2, const float* bed, const float* roughness, const float* level,
                         int level_stride, double dt, const float* lam_out, float* lam_U, float* lam_level,
                         void* adj_ws, int rows_hint, void* stream);
int swe_step_adjoint_f64(const swe_grid* grid, const swe_config* cfg, const double* U, const double* k1,
                         const double* k2, const double* bed, const double* roughness, const double* level,
                         int level_stride, double dt, const double* lam_out, double* lam_U, double* lam_level,
                         void* adj_ws, int rows_hint, void* stream);

/* Reverse sweep over one checkpoint segment in a single call: replays n_steps
 * steps from U_ckpt with the recorded dts (host array; every replayed dt equals
 * the recorded one) onto a caller tape of swe_segment_tape_bytes(), then applies
 * the step adjoints last to first.  lam holds the adjoint of the segment's
 * final state on entry and of U_ckpt on exit; lam_tmp is scratch of the same
 * size; lam_level accumulates.  parity is the forward workspace's step parity. */
size_t swe_segment_tape_bytes(const swe_grid* grid, int n_steps, int dtype_bytes);
int swe_segment_adjoint_f32(const swe_grid* grid, const swe_config* cfg, const float* U_ckpt, const float* bed,
                            const float* roughness, const float* level, int level_stride, const double* dts,
                            int n_steps, int parity, float* lam, float* lam_tmp, float* lam_level, float* tape,
                            void* fwd_ws, void* adj_ws, int rows_hint, void* stream);
int swe_segment_adjoint_f64(const swe_grid* grid, const swe_config* cfg, const double* U_ckpt, const double* bed,
                            const double* roughness, const double* level, int level_stride, const double* dts,
                            int n_steps, int parity, double* lam, double* lam_tmp, double* lam_level, double* tape,
                            void* fwd_ws, void* adj_ws, int rows_hint, void* stream);

/* Reverse sweep of a recorded segment (no replay): tape_U holds the n_steps
 * states the steps start from, tape_k1 / tape_k2 their stage tendencies, as
 * written by swe_advance_tape_*; dts, lam, lam_tmp, lam_level as for
 * swe_segment_adjoint_*. */
int swe_segment_reverse_f32(const swe_grid* grid, const swe_config* cfg, const float* tape_U, const float* tape_k1,
                            const float* tape_k2, const float* bed, const float* roughness, const float* level,
                            int level_stride, const double* dts, int n_steps, float* lam, float* lam_tmp,
                            float* lam_level, void* adj_ws, int rows_hint, void* stream);
int swe_segment_reverse_f64(const swe_grid* grid, const swe_config* cfg, const double* tape_U,
                            const double* tape_k1, const double* tape_k2, const double* bed, const double* roughness,
                            const double* level, int level_stride, const double* dts, int n_steps, double* lam,
                            double* lam_tmp, double* lam_level, void* adj_ws, int rows_hint, void* stream);

/* Reverse of a chunked run's segment (autograd through the inversion driver's
 * chunk loop and gauge misfit, bc_inversion_train.py:332-452): chunk_of_step
 * and ends_chunk are host arrays for the segment's steps (ends_chunk[k] = 1 when
 * step k ended its chunk, whose observation adjoint g_obs[chunk] [n_gauges] is
 * then added to lam's h plane at the gauges before step k is reversed); the
 * level adjoint of step k accumulates into level_grad[chunk_of_step[k]].
 * The _adjoint_ variant replays the segment from U_ckpt onto tape first (as
 * swe_segment_adjoint_*). */
int swe_segment_reverse_chunked_f32(const swe_grid* grid, const swe_config* cfg, const float* tape_U,
                                    const float* tape_k1, const float* tape_k2, const float* bed,
                                    const float* roughness, const float* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps,
                                    const int32_t* gauges, int n_gauges, const float* g_obs, float* lam,
                                    float* lam_tmp, float* level_grad, void* adj_ws, int rows_hint, void* stream);
int swe_segment_reverse_chunked_f64(const swe_grid* grid, const swe_config* cfg, const double* tape_U,
                                    const double* tape_k1, const double* tape_k2, const double* bed,
                                    const double* roughness, const double* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps,
                                    const int32_t* gauges, int n_gauges, const double* g_obs, double* lam,
                                    double* lam_tmp, double* level_grad, void* adj_ws, int rows_hint, void* stream);
int swe_segment_adjoint_chunked_f32(const swe_grid* grid, const swe_config* cfg, const float* U_ckpt,
                                    const float* bed, const float* roughness, const float* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps, int parity,
                                    const int32_t* gauges, int n_gauges, const float* g_obs, float* lam,
                                    float* lam_tmp, float* level_grad, float* tape, void* fwd_ws, void* adj_ws,
                                    int rows_hint, void* stream);
int swe_segment_adjoint_chunked_f64(const swe_grid* grid, const swe_config* cfg, const double* U_ckpt,
                                    const double* bed, const double* roughness, const double* levels,
                                    const double* dts, const int32_t* chunk_of_step, const int32_t* ends_chunk,
                                    int n_steps, int parity, const int32_t* gauges, int n_gauges,
                                    const double* g_obs, double* lam, double* lam_tmp, double* level_grad,
                                    double* tape, void* fwd_ws, void* adj_ws, int rows_hint, void* stream);

/* ---- measurement ---------------------------------------------------------
 * VALU issue-rate probe (no counterpart in the reference): `blocks` blocks of
 * 256 lanes each run `iters` iterations of 16 independent non-packed FP32
 * multiply / add instructions, then write one float per lane to out
 * (blocks * 256 floats).  Timing it gives the FP32 VALU ceiling against which
 * the stage kernels' instruction counts are priced. */
int swe_valu_probe(float* out, int blocks, int iters, void* stream);

/* VALU mix probe (no counterpart in the reference): `blocks` blocks of 256
 * lanes each run `iters` iterations of the stage kernels' x-sweep arithmetic for
 * one cell (primitives, limited increments, face states, HLL face flux, flux
 * differences over dx) on register values, with the kernels' lane shifts, then
 * write one float per lane to out (blocks * 256 floats).  Its issue rate is the
 * VALU ceiling of the kernels' own instruction mix. */
int swe_valu_mix_probe(float* out, int blocks, int iters, void* stream);

/* Division probe: for n operand pairs (device arrays), q_rn = the kernels'
 * FP32 division (reciprocal estimate, one Newton step, two remainder
 * corrections), q_by = the same from a host-rounded reciprocal (the form used
 * for dx and dy), q_ieee = the compiler's IEEE division.  For operands in the
 * normal range the three agree bit for bit.  No counterpart in the reference;
 * a test hook. */
int swe_division_probe(const float* a, const float* b, long long n, float* q_rn, float* q_by, float* q_ieee,
                       void* stream);

/* Square-root probe: s_rn = the kernels' FP32 square root (hardware estimate
 * and a pick among it and its neighbours), s_ieee = the compiler's IEEE square
 * root, for n device arguments; equal bit for bit for arguments >= 2^-96.  No
 * counterpart in the reference; a test hook. */
int swe_sqrt_probe(const float* x, long long n, float* s_rn, float* s_ieee, void* stream);

/* Friction-power probe: y_fast = the kernels' FP32 x^third (a cube root refined
 * in FP64, times x^(third - 1/3)), y_powf = powf(x, third), for n device
 * arguments.  No counterpart in the reference; a test hook. */
int swe_pow_probe(const float* x, long long n, float third, float* y_fast, float* y_powf, void* stream);

/* Lane-shift probe: the kernels' one-lane wavefront shifts on n device values
 * (n a multiple of 256; 64-lane wavefronts over consecutive elements):
 * up[i] = x[i - 1] and dn[i] = x[i + 1] within a wavefront, 0 in the lane with
 * no source; as FP64 (two 32-bit halves) and as FP32 of (float)x.  No counterpart in the reference; a test hook. */
int swe_lane_shift_probe(const double* x, long long n, double* up64, double* dn64, float* up32, float* dn32,
                         void* stream);

/* ---- water volume (FP64) -------------------------------------------------
 * volume[b] = sum over the interior cells of member b of h * dx * dy, summed
 * in FP64 in a fixed order (the same bits on every call).  volume is a device
 * array of grid->batch doubles.  Uses the workspace's stage-tendency region
 * as scratch, so it must not overlap a step on the same workspace.  No
 * counterpart in the reference (which has no volume diagnostic); it is the
 * FP64 mass-conservation sum used by the conservation tests and bench.py. */
int swe_water_volume_f32(const swe_grid* grid, const float* U, double* volume, void* ws, void* stream);
int swe_water_volume_f64(const swe_grid* grid, const double* U, double* volume, void* ws, void* stream);

/* ---- ensemble of forward runs (FP32) -------------------------------------
 * Uneven grids, each with its own adaptive dt, advanced in lockstep by one
 * launch sequence (the reference batches members on one grid with one shared
 * dt, solver.py:528-529; here every scenario keeps its own).  A scenario runs
 * n_chunks calls of SWE2D.forward(U, t_end=chunk_seconds) (solver.py:719-779)
 * with boundary level levels[k] during chunk k, and records h + z at its gauges
 * at the end of every chunk (bc_inversion_train.py:332-410). */
typedef struct swe_scenario {
  int32_t ny, nx;
  double dx, dy;
  float* U_a;                 /* [3, ny, nx]: initial state in, ping-pong buffer */
  float* U_b;                 /* [3, ny, nx] */
  float* k1;                  /* [3, ny, nx] scratch */
  const float* bed;           /* [ny, nx] */
  const float* roughness;     /* [ny, nx] */
  const float* levels;        /* [n_chunks] */
  int32_t n_chunks;
  double chunk_seconds;
  const int32_t* gauges;      /* [n_gauges][2] (iy, ix) */
  int32_t n_gauges;
  float* gauge_out;           /* [n_chunks][n_gauges] */
  int32_t rows_hint;          /* rows per wavefront strip; 0 = heuristic */
} swe_scenario;

size_t swe_ensemble_workspace_bytes(int n_scenarios);
const char* swe_ensemble_last_error_message(void);
/* Copies the scenario table to the device workspace (synchronously) and resets
 * every scenario's clock; returns the number of wavefronts one launch covers. */
int swe_ensemble_init(const swe_config* cfg, int n_scenarios, const swe_scenario* scenarios, void* ws,
                      size_t ws_bytes, long long* total_waves, void* stream);
/* Launches n_steps global steps (3 launches each); finished scenarios exit at
 * once.  Capturable in a HIP graph (parity comes from first_step + i).
 * arithmetic: SWE_ARITH_REFERENCE (the reference's rounding, results equal to
 * batch-1 swe_advance_f32 runs bit for bit) or SWE_ARITH_FAST (the
 * tolerance-bounded arithmetic of swe_config.arithmetic; equal to batch-1 runs
 * with that arithmetic). */
int swe_ensemble_advance(void* ws, int n_scenarios, long long total_waves, int first_step, int n_steps,
                         int arithmetic, void* stream);
size_t swe_ensemble_clock_offset(int n_scenarios);
size_t swe_ensemble_clock_bytes(void);

/* ---- BF16 MFMA kernels of the 1-D convolution surrogate --------------------
 * No reference counterpart (the surrogate warm start is new).  bf16 operands
 * are raw 16-bit words; accumulation is FP32. */
const char* sur_last_error_message(void);
/* C[m][n] = sum_k A[m*lda + k] * B[n*ldb + k]; K % 16 == 0; split-K into FP32
 * partials [splits][M][N] reduced in split order when splits > 1. */
int sur_gemm_nt_bf16(int M, int N, int K, const void* A, long lda, const void* B, long ldb, float* C, long ldc,
                     int splits, float* partials, void* stream);
int sur_im2col(const void* X, int Cin, int Bn, int L, int Kw, int dil, int Kpad, int transpose, void* col,
               void* stream);
int sur_col2im(const float* dcol, int Cin, int Bn, int L, int Kw, int dil, float* dX, void* stream);
int sur_bias_act(const float* C, const float* bias, int M, long N, int relu, float* pre, void* out_bf16,
                 float* out_f32, void* stream);
int sur_act_bwd(const float* g_out, const float* pre, int M, long N, int relu, float* g_pre, void* g_bf16,
                void* g_bf16_t, int ldt, void* stream);
int sur_bias_grad(const float* g, int M, long N, float* db, void* stream);
int sur_cast_bf16(const float* x, long count, void* y, void* stream);
int sur_cast_pad_rows(const float* x, int R, int Cc, int Cpad, void* y, void* stream);
int sur_transpose_bf16(const float* x, int R, int Cc, int Rpad, void* y, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SWE2D_HIP_H */
