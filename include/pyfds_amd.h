/*
 * pyfds_amd C-ABI: MI355X (gfx950) time stepping for pyfds' staggered-grid solvers.
 *
 * The reference (emtpb/pyfds) is pure Python and has no FFI; its drop-in boundary is the
 * Python class protocol.  Each entry point below replaces one piece of that protocol,
 * cited as pyfds/<file>:<line>.  Conventions:
 *   - every call returns 0 on success or a negative FDS_E_* code; fds_last_error() gives a
 *     thread-local message for the last failure on the calling thread;
 *   - the library owns all device memory; host pointers are borrowed for the call only;
 *   - one context per simulation; a context has its own HIP stream and is used from one host
 *     thread at a time;
 *   - all host-side arrays are float64 and flat, index i = ix + nx*iy (pyfds/fields.py:377).
 */
#ifndef PYFDS_AMD_H
#define PYFDS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fds_ctx fds_ctx;

/* models (pyfds/acoustics.py, pyfds/thermal.py, pyfds/nonlinear_acoustics.py) */
enum {
  FDS_ACOUSTIC1D = 1,     /* Acoustic1D         pyfds/acoustics.py:9-52          */
  FDS_ACOUSTIC2D = 2,     /* Acoustic2D         pyfds/acoustics.py:66-128        */
  FDS_ACOUSTIC_AXI = 3,   /* Acoustic3DAxi      pyfds/acoustics.py:142-225       */
  FDS_THERMAL1D = 4,      /* Thermal1D          pyfds/thermal.py:12-51           */
  FDS_THERMAL2D = 5,      /* Thermal2D          pyfds/thermal.py:54-107          */
  FDS_THERMAL_AXI = 6,    /* Thermal3DAxi       pyfds/thermal.py:110-176         */
  FDS_IDEALGAS1D = 7,     /* IdealGas1D         pyfds/nonlinear_acoustics.py:10-89   */
  FDS_ACOUSTIC2ND1D = 8   /* Acoustic2ndOrder1D pyfds/nonlinear_acoustics.py:103-171 */
};

/* numeric forms */
enum {
  FDS_FP64 = 0,        /* reproduces the reference's float64 arithmetic term by term        */
  FDS_FP32 = 1,        /* float32 state and coefficients                                    */
  FDS_FP32_KAHAN = 2   /* float32 state + float32 compensation (Kahan) for accumulated fields */
};

/* error codes */
enum {
  FDS_OK = 0,
  FDS_E_ARG = -1,       /* bad argument (Python: ValueError)                          */
  FDS_E_HIP = -2,       /* HIP runtime error                                          */
  FDS_E_INDEX = -3,     /* index or signal step out of range (Python: IndexError)     */
  FDS_E_STATE = -4,     /* call out of order / unsupported configuration              */
  FDS_E_NONFINITE = -5  /* non-finite value detected                                  */
};

/* coefficient slots (fds_set_coeff); values are computed on the host with the reference's
 * expression order (pyfds/acoustics.py:30-37, 92-107, 181-201; pyfds/thermal.py:32-37,
 * 78-89, 147-158; pyfds/nonlinear_acoustics.py:252-263, 347-357) */
enum {
  /* acoustics 1D/2D/Axi */
  FDS_C_KX = 0, FDS_C_KY = 1,     /* dt/d * c**2 * rho               (a_p_v*)            */
  FDS_C_BX = 2, FDS_C_BY = 3,     /* dt/d / rho                      (a_v*_p)            */
  FDS_C_MX = 4, FDS_C_MY = 5,     /* dt/d**2 * mu / rho              (a_v*_v*)           */
  FDS_C_MC = 6,                   /* Axi: dt/dx * mu / rho  (divided by r in-kernel)     */
  FDS_C_MT = 7,                   /* Axi: dt * mu / rho     (per-step vx/r**2 term)      */
  /* thermal 1D/2D/Axi */
  FDS_C_GX = 0, FDS_C_GY = 1,     /* dt/d / rho / c_p                (a_t_q*)            */
  FDS_C_QX = 2, FDS_C_QY = 3,     /* 1/d * k                         (a_q*_t)            */
  /* nonlinear 1D */
  FDS_C_RHO0 = 0,                 /* static density                                      */
  FDS_C_M = 1,                    /* dt/dx**2 * mu                   (a_v_v)             */
  FDS_C_EOS1 = 2,                 /* IdealGas1D: rho0 * c**2 / gamma; 2nd order: d_rho_p */
  FDS_C_EOS2 = 3,                 /* IdealGas1D: gamma;              2nd order: d_rho2_p / 2 */
  FDS_C_CSQ = 4,                  /* IdealGas1D linear state: c**2                       */
  FDS_NUM_COEFF = 8
};

/* scalars (fds_set_scalars, in this order) */
enum {
  FDS_S_DX = 0,       /* x increment (Axi radius r = ix*dx + dx/2, pyfds/acoustics.py:176) */
  FDS_S_HALF_DX = 1,  /* dx / 2                                                            */
  FDS_S_DTDX = 2,     /* nonlinear 1D: dt/dx          (a_d_v, a_v_p)                       */
  FDS_S_DTDX_4 = 3,   /* nonlinear 1D: dt/dx/2/2      (a_v_v2, central)                    */
  FDS_S_FLAGS = 4,    /* nonlinear 1D: bit0 convective, bit1 nonlinear state               */
  FDS_NUM_SCALARS = 8
};

/* Create a context for one field of nx*ny cells (ny = 1 for 1D).  Replaces the field
 * construction + FieldComponent allocation of pyfds/fields.py:133-150, 247-267, 577-589. */
int fds_create(fds_ctx **out, int model, int form, int64_t nx, int64_t ny, int device);

/* Release every device resource of the context. */
int fds_destroy(fds_ctx *ctx);

/* Per-context scalars, see FDS_S_*.  Part of assemble_matrices (pyfds/acoustics.py:27-38). */
int fds_set_scalars(fds_ctx *ctx, const double *s, int n);

/* Upload one per-cell coefficient array (n = nx*ny) into slot FDS_C_*.  n = 0 clears the slot.
 * Replaces the sparse-matrix assembly of assemble_matrices (pyfds/fields.py:158-365 operators,
 * pyfds/acoustics.py:89-109, pyfds/thermal.py:75-90).  Must be followed by fds_commit. */
int fds_set_coeff(fds_ctx *ctx, int slot, const double *data, int64_t n);

/* Finish a coefficient update: picks the kernel variant (lossless/lossy, x/y aliasing). */
int fds_commit(fds_ctx *ctx);

/* Copy a component's values to / from the device (FieldComponent.values,
 * pyfds/fields.py:585).  comp follows the model's component order
 * (pressure, velocity[_x], velocity_y | temperature, heat_flux[_x], heat_flux_y |
 *  pressure, velocity, density). */
int fds_set_state(fds_ctx *ctx, int comp, const double *data, int64_t n);
int fds_get_state(fds_ctx *ctx, int comp, double *data, int64_t n);

/* Boundary program of one component (FieldComponent.boundaries applied by apply_bounds,
 * pyfds/fields.py:591-600, Boundary.apply pyfds/regions.py:125-145).  The host compiles the
 * boundary list into ncell unique cells; cell u runs entries ptr[u]..ptr[u+1]-1 in list order,
 * each  x = a[e]*x + table[off[e] + step*stride[e]]  (stride 0: scalar value, 1: signal).
 * len[e] is the number of valid steps of entry e (INT64_MAX for scalars).  ncell = 0 clears. */
int fds_set_boundaries(fds_ctx *ctx, int comp, int64_t ncell, const int64_t *cells,
                       const int32_t *ptr, int64_t nent, const double *a, const int64_t *off,
                       const int64_t *stride, const int64_t *len, int64_t ntable,
                       const double *table);

/* Recorded points of one component (FieldComponent.write_outputs, pyfds/fields.py:602-611):
 * m slots, slot k records cells[k] every step after that component's boundaries. */
int fds_set_outputs(fds_ctx *ctx, int comp, int64_t m, const int64_t *cells);

/* Advance nsteps steps starting at pyfds step counter start_step (pyfds/fields.py:87-93 loop
 * with each model's sim_step).  chunk > 0 replays HIP-graph captured chunks of that many
 * steps; chunk = 0 launches step by step.  Output rows 0..nsteps-1 of this run are kept on the
 * device until the next run.  Synchronous. */
int fds_run(fds_ctx *ctx, int64_t start_step, int64_t nsteps, int chunk);

/* GPU time in milliseconds between HIP events recorded on the context's stream before the
 * first and after the last step launch of the last fds_run, or summed over the asynchronous
 * runs taken since the last one enqueued with none in flight (no reference counterpart: the
 * measurement hook for benchmarks and the autotuner). */
int fds_last_run_ms(fds_ctx *ctx, double *ms);

/* Read rows [row0, row0+nrows) of a component's recorded outputs from the last run into
 * dst[nrows][m] (row-major, one row per step). */
int fds_read_outputs(fds_ctx *ctx, int comp, int64_t row0, int64_t nrows, double *dst);

/* Asynchronous runs (no reference counterpart: they let the host turn one run's recorded
 * outputs into pyfds' per-sample lists while the device computes the next run).
 * fds_run_async enqueues fds_run's steps on the context's stream, then a copy of the run's
 * outputs into pinned host memory, and returns at once.  At most two runs may be in flight.
 * fds_run_take waits for the oldest one and copies its outputs (nsteps rows of m values per
 * component, as fds_read_outputs) into dst0..dst2 (null entries and components without
 * outputs are skipped) and adds its time to fds_last_run_ms.  While runs are in flight the
 * setters, fds_run, fds_read_outputs, fds_window_sq_diff and fds_output_device_ptr fail with
 * FDS_E_STATE; fds_get_state waits for the context's own runs and fails while a batch run of
 * the context is in flight. */
int fds_run_async(fds_ctx *ctx, int64_t start_step, int64_t nsteps, int chunk);
int fds_run_take(fds_ctx *ctx, double *dst0, double *dst1, double *dst2);

/* Everything a run of nsteps steps from start_step sets up before its first step, without
 * stepping: event tables, output rows, the chunk graphs it replays (no reference counterpart:
 * lets a caller keep one-time set-up out of a timed region).  A run of at most nsteps steps
 * that follows then starts at once. */
int fds_prepare_run(fds_ctx *ctx, int64_t start_step, int64_t nsteps, int chunk);

/* Tile shape of the 2D step kernels (autotuner hook): index into the compiled shape table,
 * or -1 for the default.  fds_tile_count reports the table size, fds_tile_shape its entries
 * (tx columns x ty rows).  fds_tile_engine returns 0 for an LDS-tile shape and 1 for a
 * row-streaming shape (a workgroup of tx threads marching down a strip of ty rows); streaming
 * shapes serve the 2D models (lossy acoustics one time step per launch on any of them).
 * 1D models (Acoustic1D, Thermal1D, IdealGas1D, Acoustic2ndOrder1D) of up to 2^24 cells
 * default (-1) to the resident 1D engine: workgroups keep the line in registers for many steps
 * per launch (one workgroup up to 512 cells, segments beyond; see fds_set_line_shape); any tile
 * index (and longer lines) puts them on the LDS-tile kernels (one launch per step; nonlinear models on their 256-cell tiles
 * whatever the index). */
int fds_set_tile(fds_ctx *ctx, int tile_index);
int fds_tile_count(void);
int fds_tile_shape(int tile_index, int *tx, int *ty);
int fds_tile_engine(int tile_index);
/* Time steps per launch, cache policy (1: non-temporal state traffic), prefetch depth (input
 * rows a streaming workgroup loads ahead; 0 for LDS-tile shapes) and load path (lds_ring 1:
 * the rows go straight from HBM to LDS rings, no registers in flight; a field whose rows do not
 * start on 16-byte boundaries runs the same tiling with register loads) of a shape.  Shapes
 * with a prefetch depth above 1 serve single lossless fields only. */
int fds_tile_traits(int tile_index, int *steps, int *nontemporal, int *prefetch, int *lds_ring);
/* Strip height (rows per workgroup) of a streaming shape; 0 restores the shape's own height.
 * Ignored by LDS-tile shapes.  For the resident 1D engine the value is the number of cells per
 * thread (1, 2, 4 or 8; 0 = chosen from the line length). */
int fds_set_strip_height(fds_ctx *ctx, int rows);
/* Resident 1D engine layout (no reference counterpart; autotuner and test hook).  threads = 0:
 * one workgroup holds the whole line with cells_per_thread (1, 2, 4, 8) cells per thread.
 * threads > 0: the line is split into segments of threads * cells_per_thread - 2 * ghost owned
 * cells, one workgroup of `threads` threads each, every workgroup also stepping `ghost` cells
 * of each neighbour (at least 2); a launch then runs at most ghost - 1 steps (temporal
 * blocking).  Lines of up to 2^24 cells; longer ones ignore the layout and step on the
 * LDS-tile kernels. */
int fds_set_line_shape(fds_ctx *ctx, int cells_per_thread, int threads, int ghost);
/* Engine that runs a context's steps: 0 LDS tiles, 1 row streaming, 2 resident 1D,
 * 3 nonlinear 1D tiles (no reference counterpart; for tests and benchmarks). */
int fds_engine(fds_ctx *ctx);
/* Time steps one launch of the context's 2D step kernel advances (the streaming shape's fused
 * stage count, else 1; no reference counterpart): runs whose length is a multiple of it use
 * fused launches only. */
int fds_steps_per_launch(fds_ctx *ctx);

/* Batched sweeps (no reference counterpart: pyfds runs one field per Python loop).  A batch
 * groups contexts that are fully set up (coefficients committed, state, boundaries, outputs)
 * and share device and form; fds_batch_run advances all of them nsteps steps from start_step
 * with one launch per step over the tiles of every member (one tile shape for all, index
 * `tile` into the tile table or -1), replayed in HIP-graph chunks.  Afterwards each member's
 * state and outputs read back with fds_get_state / fds_read_outputs as after fds_run. */
typedef struct fds_batch fds_batch;
int fds_batch_create(fds_batch **out, fds_ctx *const *ctxs, int n, int tile);
int fds_batch_run(fds_batch *b, int64_t start_step, int64_t nsteps, int chunk);
int fds_batch_last_run_ms(fds_batch *b, double *ms);
/* Asynchronous batch runs, as fds_run_async / fds_run_take for every member at once: dst holds
 * three destinations per member (member-major: dst[3 * i + comp]); dst = null discards. */
int fds_batch_run_async(fds_batch *b, int64_t start_step, int64_t nsteps, int chunk);
int fds_batch_take(fds_batch *b, double *const *dst);
/* Time steps one launch of the batch kernel advances (as fds_steps_per_launch). */
int fds_batch_steps_per_launch(fds_batch *b);
/* Tile scheduling of a streaming batch (no reference counterpart).  0 (default): one
 * workgroup per tile, the hardware dispatcher hands them out (XCD-banded order).  1: a
 * persistent launch of as many workgroups as the device holds at once, taking tiles from a
 * work queue (a device counter, largest tiles first) until none is left; results are the same
 * bits.  FDS_E_ARG for a batch on the LDS-tile engine. */
int fds_batch_set_schedule(fds_batch *b, int schedule);
int fds_batch_destroy(fds_batch *b);

/* Non-finite detection (no reference counterpart: pyfds propagates NaN and infinity
 * silently; failure detection for long runs).  enable != 0: every run of the context scans the
 * first component's state for NaN and infinity after every graph chunk of steps (the resident
 * 1D engine: after launches at least 50 steps apart) and every component after the run; a
 * batch run scans its members that enable it after the run.  A run that found one completes
 * (state and outputs read back as usual) and returns FDS_E_NONFINITE; fds_last_error names
 * the component and the step count by which it was found. */
int fds_set_nonfinite_check(fds_ctx *ctx, int enable);

/* Flow along x (AcousticFlow2D, pyfds/acoustic_flow.py:49-57 apply_flow): after every step n,
 * each row r with period[r] == 0 or n % |period[r]| == 0 moves one cell to +x (a zero enters at
 * the row start), in p, vx, vy (and their compensation).  period = NULL clears it.  The
 * reference applies the same shift on the host after each sim_step. */
int fds_set_flow(fds_ctx *ctx, const int64_t *period, int64_t nrows);

/* ThermoAcoustic1D on the device (pyfds/coupled_fields.py:10-65: SynchronizedFields.sim_step,
 * pyfds/coupling.py:78-85, with its acoustic-loss BoundaryCoupling, pyfds/coupling.py:107-116).
 * Each of nsteps steps from start_step advances the Acoustic1D context, then the Thermal1D
 * context (same cell count, form and device, both on the resident 1D engine), then adds to the
 * temperature the heat of the acoustic losses,
 *   transfer = (loss * d**2) * dt,  d = a_v_p . (v * rho)  (the acoustic context's backward
 *   factor B applied to v * rho; loss = mu / rho_thermal / c_p, evaluated by the caller),
 * every step with step % stepping == 0; with `accumulate` the transfer is summed every step
 * and the sum added (and restarted) on those steps.  accum (n values, in/out) carries that sum
 * across calls.  Up to 2^24 cells: one workgroup runs both fields up to 512 cells, segments beyond
 * (the call sets both contexts' resident layout).  Outputs and state read back from each
 * context as after fds_run.  The reference applies the coupling on the host after every pair
 * of sim_steps. */
int fds_couple_loss_run(fds_ctx *acoustic, fds_ctx *thermal, const double *rho,
                        const double *loss, double dt, int64_t stepping, int accumulate,
                        double *accum, int64_t start_step, int64_t nsteps);

/* Coupled fields on the device: SynchronizedFields (pyfds/coupling.py:14-90) with the shipped
 * interactions, each step every field's step kernel, then the interactions in list order at
 * that step; nothing returns to the host until the run ends.  The reference applies them on
 * the host after every pair of sim_steps (pyfds/coupling.py:81-87). */
enum {
  FDS_COUPLE_MATERIAL_EXP = 1,     /* MaterialCouplingExponential  pyfds/coupling.py:229-270 */
  FDS_COUPLE_MATERIAL_POWER = 2,   /* MaterialCouplingPowerLaw     pyfds/coupling.py:273-300 */
  FDS_COUPLE_BOUNDARY_LINEAR = 3   /* BoundaryCoupling, transfer = factor * q
                                    *                              pyfds/coupling.py:93-146  */
};
/* material vectors a coupled re-assembly reads (fds_coupling_set_material) */
enum {
  FDS_P_SOUND_VELOCITY = 0, FDS_P_DENSITY = 1, FDS_P_ABSORPTION = 2,  /* acoustics */
  FDS_P_HEAT_CAPACITY = 3, FDS_P_CONDUCTIVITY_X = 4, FDS_P_CONDUCTIVITY_Y = 5  /* thermal */
};
typedef struct {
  int kind;                        /* FDS_COUPLE_*                                            */
  int source_field, source_comp;   /* field index in the coupling, component (model order)    */
  int target_field, target_comp;   /* boundary: target component; material: unused           */
  int target_param;                /* material: FDS_P_* coupled parameter                     */
  double p0, p1;                   /* exp: a, b; power law: power, factor; linear: factor, -  */
  int64_t stepping;                /* apply every stepping-th step (step % stepping == 0)     */
  double threshold;                /* material: relative change to re-assemble (NaN: none)    */
  int additive, accumulate;        /* boundary                                                */
} fds_interaction;
typedef struct fds_coupling fds_coupling;
/* Fields: committed contexts of one device, form (FP64 or FP32) and cell count; at most one
 * material coupling per target field (as in the reference, whose wrapped material_vector
 * supports one).  Material couplings re-assemble acoustic fields (sound velocity, density,
 * absorption) or thermal fields (density, heat capacity, conductivity x / y). */
int fds_coupling_create(fds_coupling **out, fds_ctx *const *fields, int nfields,
                        const fds_interaction *inter, int ninter);
/* Static material vector `param` of field `field` (float64, n = cells): the target field's
 * material_vector without the coupling (pyfds/coupling.py:184-197).  A material coupling's
 * target needs all the vectors its model's assembly reads. */
int fds_coupling_set_material(fds_coupling *cp, int field, int param, const double *values,
                              int64_t n);
/* Expression scalars of a target field, in the reference's order: acoustics dt/dx, dt/dy,
 * dt/dx**2, dt/dy**2, dt; thermal dt/dx, dt/dy, 1/dx, 1/dy. */
int fds_coupling_set_scalars(fds_coupling *cp, int field, const double *s, int n);
/* Interaction state (n = cells float64): a material coupling's last used factors, a boundary
 * coupling's accumulated transfer; set != 0 uploads, 0 downloads.  Both start at zero. */
int fds_coupling_state(fds_coupling *cp, int interaction, int set, double *values, int64_t n);
/* Advance every field nsteps steps from start_step (each field's outputs and state read back
 * as after fds_run).  Re-assemblies of material couplings with a threshold are logged as
 * (step, largest relative change) pairs into log[2 * log_cap]; *nlog = their number. */
int fds_coupling_run(fds_coupling *cp, int64_t start_step, int64_t nsteps, double *log,
                     int64_t log_cap, int64_t *nlog);
int fds_coupling_destroy(fds_coupling *cp);
/* Copy coefficient slot FDS_C_* of a context to the host (no reference counterpart: the
 * host's assembled factors after a device re-assembly). */
int fds_get_coeff(fds_ctx *ctx, int slot, double *data, int64_t n);

/* Learned absorbing boundary (no reference counterpart; pyfds has only the plain method).
 * After every step of an FP32 Acoustic2D context on a square G x G grid, a BF16-MFMA network
 * predicts the outer `layer` cells (16) of p, vx, vy on all four sides from the 8 adjacent
 * interior rows at the current and previous time level.  w1 (bf16 [4][32][5][48]) and w2
 * (bf16 [4][48][5][32]) are device pointers owned by the caller; comp_scale converts p, vx, vy
 * to network units.  w1 = NULL switches the boundary off. */
int fds_set_learned_boundary(fds_ctx *ctx, const uint16_t *w1, const uint16_t *w2, int layer,
                             const float *comp_scale);

/* Training kernels of the learned boundary (device pointers, `stream` = hipStream_t).  1D
 * convolutions along a strip, zero padded: y[b,p,co] = sum_k sum_ci w[side[b],co,k,ci]
 * x[b,p+k-K/2,ci]; x, dy: bf16 [B][P][C]; w: bf16 [S][Co][K][Ci]; wt: bf16 [S][Ci][K][Co];
 * Ci and Co multiples of 16, K odd. */
int fds_lbc_conv_fwd(const uint16_t *x, const uint16_t *w, const int32_t *side, float *y,
                     uint16_t *yb, int B, int P, int Ci, int Co, int K, void *stream);
int fds_lbc_conv_bwd_data(const uint16_t *dy, const uint16_t *wt, const int32_t *side,
                          float *dx, uint16_t *dxb, int B, int P, int Ci, int Co, int K,
                          void *stream);
int fds_lbc_conv_bwd_weight(const uint16_t *dy, const uint16_t *x, float *part, int B, int P,
                            int Ci, int Co, int K, void *stream);
int fds_lbc_reduce_weight(const float *part, const int32_t *side, float *dw, int B, int S, int n,
                          void *stream);
int fds_lbc_mse(const float *y, const float *t, const int32_t *lo, const int32_t *hi,
                const float *wch, uint16_t *dyb, float *loss_part, int nblocks, int B, int P,
                int C, float norm, void *stream);
int fds_lbc_adam(float *w, float *m1, float *m2, const float *g, int n, float lr, float b1,
                 float b2, float eps, int step, void *stream);
int fds_lbc_pack(const float *w, uint16_t *wb, uint16_t *wtb, int S, int Co, int K, int Ci,
                 void *stream);
int fds_lbc_gather(const float *frames, const int32_t *run, const int32_t *step,
                   const int32_t *side, const float *scale, uint16_t *x, float *t, int B, int T,
                   int D, int G, int L, const float *comp_scale, void *stream);
int fds_lbc_band_rms(const float *frames, float *rms, int F, int D, int G, int L, void *stream);

/* Diagnostics in float64 on the device: over a wx x wy window at (ax0, ay0) of context a and
 * (bx0, by0) of context b, num = sum (a - b)^2 and den = sum b^2 of the current states. */
int fds_window_sq_diff(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx, int wy, int ax0,
                       int ay0, int bx0, int by0, double *num, double *den);
/* No counterpart in the reference: fds_window_sq_diff for callers that cannot allocate
 * device memory (code under graph capture, serving loops that forbid hipMalloc).  Same
 * arguments, same numbers bit for bit; it allocates nothing: the 2 x 512 float64 partial sums
 * (one per workgroup, 8 KB, whatever the window size) live in a static device array that each
 * GPU holds until the process ends.  Calls that overlap on one GPU share that array and race:
 * serialise them. */
int fds_window_sq_diff_static(fds_ctx *a, int comp_a, fds_ctx *b, int comp_b, int wx, int wy,
                              int ax0, int ay0, int bx0, int by0, double *num, double *den);

/* Device pointer of a component's recorded outputs ([rows][m] float64) of the last run, for
 * in-process consumers that keep them on the GPU. */
int fds_output_device_ptr(fds_ctx *ctx, int comp, double **ptr, int64_t *rows, int64_t *m);

/* Calibration (no reference counterpart): copies `bytes` bytes device-to-device `reps` times
 * (after one untimed copy) and returns read+write GB/s and milliseconds per copy.  variant 0:
 * grid-stride loop of 16-byte elements (4096 workgroups); 1: one pass, 8 loads in flight per
 * thread; 2: as 1 with non-temporal loads and stores; 3: 16 loads, non-temporal; 4: 4 loads,
 * non-temporal. */
int fds_stream_copy_bench(int64_t bytes, int reps, int variant, double *gbs, double *ms_each);

/* Calibration (no reference counterpart): the memory access pattern of the LDS-ring streaming
 * kernels without their arithmetic, on an nx x ny grid of `elem`-byte elements (8: float64,
 * 4: float32).  One-wave tiles of 64 columns (64 - 2*ghost owned) and strip_h rows, in the
 * step kernels' order, stream rows -ghost-1 .. strip_h+ghost of n_state state arrays and two
 * coefficient arrays through the rings of a `stages`-stage kernel (same LDS footprint, one
 * row ahead) and store the state arrays' owned cells.  n_state 3 (Acoustic2D, isotropic
 * coefficients) or 1 (Thermal2D); stages 2 or 3; nx a multiple of 16 / elem; nx * ny at most
 * 2^28.  Returns milliseconds per launch (after one untimed launch) and the bytes the launch
 * requests (loads and stores).  The step kernel's launch time over this one is how far its
 * arithmetic keeps it from the floor its memory traffic sets. */
int fds_stream_pattern_bench(int nx, int ny, int elem, int stages, int ghost, int strip_h,
                             int n_state, int reps, double *ms_each, double *bytes);

/* Last error message of the calling thread ("" if none). */
const char *fds_last_error(void);

/* Library version string. */
const char *fds_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PYFDS_AMD_H */
