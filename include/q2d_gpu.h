/*
 * q2d_gpu.h — C-ABI of the MI355X explicit finite-volume step.
 *
 * The library replaces one call of the Q2D Fortran host:
 *
 *     call Explicit_Step ( domain(level), External_Function )
 *
 * made from Q2D_solve (src/lib/driver/Wrap_Solve.f90:19-32, body in
 * src/lib/numerics/time/explicit/Mod_Explicit.f90:10-174).  Everything that
 * step does on the hot path — Compute_dt / Set_Global_dt (Mod_dt.f90:10-128),
 * Copy_State (Mod_Explicit.f90:177-197), and per RK stage Fill_Ghost_Cell
 * (Lib_Ghost.f90:16-140), Fluxes (Mod_Fluxes.f90:97-234), BC_Fluxes
 * (Mod_BC_Fluxes.f90:49-271), Ghost_Wall_Extrapolation (Lib_Ghost.f90:366),
 * Compute_Sources (Mod_Sources.f90:79-91) and RK_Newstate
 * (Lib_Newstate.f90:14-161, with Residual_Smoothing, Lib_IRS.f90:17-177, for
 * cases with irs-beta > 0), then Compute_Residual
 * (Mod_Diagnostic.f90:438-469) — runs on the GPU behind these entry points.
 *
 * Memory layouts accepted at the boundary are the reference's own Fortran
 * allocations (Mod_Allocate_Data.f90:100-126), so a Fortran host passes
 * c_loc() of its arrays without repacking:
 *
 *   P      (nprim, 1-gc:ni+gc, 1-gc:nj+gc)        nprim = 4 : rho, u, v, p
 *   dl     (1-gc:ni+gc, 1-gc:nj+gc) % c(2)        -> 2 doubles per cell
 *   M      (1-gc:ni+gc, 1-gc:nj+gc) % c(2,2)      -> 4 doubles per cell, column major
 *   sur    (1-gc:ni+gc, 1-gc:nj+gc)
 *   dir(1)%f(0:ni, 1:nj) {N(2), int_l}            -> 3 doubles per i-face
 *   dir(2)%f(1:ni, 0:nj) {N(2), int_l}            -> 3 doubles per j-face
 *   node   (0:ni, 0:nj) % c(2)                    -> 2 doubles per node (viscous / axisymmetric only)
 *   src_a, src_b (1:ni, 1:nj)                     axis_r_ave / axis_inv_r, or area_geom_term_x / _y
 *
 * All functions return Q2D_OK (0) or a negative Q2D_E_* code; no C++
 * exception crosses the boundary.  q2d_last_error() returns a message for the
 * last failing call on the calling thread.
 *
 * Gas model: one species, either calorically perfect (constant cp, gas constant
 * R: the closed forms of the FLINT calls the reference makes, co_rotot_Rtot,
 * f_ss, f_gamma, H0, E0, prim2cons, cons2prim) or, with q2d_case_desc.thermo,
 * FLINT's native tabulated mode: h, cp, mu and k interpolated linearly from
 * integer-kelvin tables (plain FP64 form).  FLINT itself is not part of the
 * reference tree.
 */
#ifndef Q2D_GPU_H
#define Q2D_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define Q2D_GC    2
#define Q2D_NPRIM 4

/* status codes */
#define Q2D_OK               0
#define Q2D_E_ARG           -1   /* invalid argument / inconsistent problem description */
#define Q2D_E_HIP           -2   /* HIP runtime error */
#define Q2D_E_NOMEM         -3   /* device allocation failed or budget exceeded */
#define Q2D_E_STATE         -4   /* NaN or p<0 after an update (Lib_Newstate.f90:327-331) */
#define Q2D_E_INTERFACE     -5   /* non-physical reconstructed interface state (Lib_Convective.f90:49-54) */
#define Q2D_E_UNSUPPORTED   -6   /* option not implemented on this path */
#define Q2D_E_NODEVICE      -7   /* no usable GPU */

/* limiter ids — Mod_Space.f90:44-78 ('minmod','vanleer',...; default IORD) */
#define Q2D_LIM_IORD      0
#define Q2D_LIM_MINMOD    1
#define Q2D_LIM_VANLEER   2
#define Q2D_LIM_VANALBADA 3
#define Q2D_LIM_MC        4
#define Q2D_LIM_SUPERBEE  5
#define Q2D_LIM_LIMO3     6

/* Riemann solver ids — Mod_Riemann.f90:25-109 */
#define Q2D_RIEMANN_HLLC  0
#define Q2D_RIEMANN_HLLE  1
#define Q2D_RIEMANN_HLLEM 2   /* Lib_Riemann_HLL.f90:355-432 */
#define Q2D_RIEMANN_HLLEPP 3  /* 'HLLE++', Lib_Riemann_HLL.f90:435-530 (uses the shock detector) */
#define Q2D_RIEMANN_HLLCP  4  /* 'HLLC+', Lib_Riemann_HLL.f90:571-603 (uses the shock detector) */

/* source ids — Mod_Sources.f90:16-91 ('axisymmetric' | 'area_variation') */
#define Q2D_SRC_NONE      0
#define Q2D_SRC_AXI       1
#define Q2D_SRC_AREA      2

/* arithmetic forms of the face computation */
#define Q2D_PREC_FP64     0   /* plain: every operation in FP64, reference operation order */
#define Q2D_PREC_MIXED    1   /* reconstruction, limiter and Riemann flux in FP32; state, residual
                                 accumulation, RK update, dt and norms in FP64 */

/* BC codes handled on this path (Mod_BC_Fluxes.f90:93-265) */
#define Q2D_BC_CONNECT    101
#define Q2D_BC_PERIODIC   201
#define Q2D_BC_SYMMETRY   300
#define Q2D_BC_WALL_T     302
#define Q2D_BC_EXTRAP     400
#define Q2D_BC_SUPERSONIC 405
#define Q2D_BC_OUTFLOW    406

/* integration-variables */
#define Q2D_INTVARS_CONS  0
#define Q2D_INTVARS_PRIM  1

/* Numerical scheme of one case: the [Q2D-Numerics]/[Q2D-Physics]/[Q2D-Sources]
 * keys of input.ini after Assign_Setup (Assign_Setup.f90:11-62). */
typedef struct q2d_scheme {
  int32_t n_rk;            /* 1 euler, 2 RK2, 3 RK3  (Lib_RK.f90:8-10) */
  int32_t limiter;         /* Q2D_LIM_* ; 'first-order' reconstruction == IORD */
  int32_t riemann;         /* Q2D_RIEMANN_* ('HLLC', 'HLLE', 'HLLEM', 'HLLE++', 'HLLC+') */
  int32_t time_accurate;   /* 1: global dt (Set_Global_dt), 0: local dt */
  int32_t model;           /* 0 Euler, 1 laminar Navier-Stokes */
  int32_t source;          /* Q2D_SRC_* */
  int32_t precision;       /* Q2D_PREC_* */
  int32_t ramp_iter;       /* cfl-rise-threshold */
  double  cfl, vnn;
  double  gas_R, gas_cp;   /* gas constant and constant cp (table mode: gas_R only) */
  double  mu_ref, T_ref, S_suth, Pr;   /* Sutherland law mu = mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S); k = mu cp / Pr */
  double  irs_beta;        /* irs-beta (Register_Numerics.f90:53): > 0 enables implicit residual smoothing
                              (Assign_Setup.f90:61; Lib_IRS.f90:17-177 from Newstate_Conservative,
                              Lib_Newstate.f90:35-64); 0 leaves it off */
  int32_t space_sd;        /* 1: space-reconstruction = 'MUSCL-SD' (obj_space_scheme%SD, Assign_Setup.f90:
                              49-53): Tramel's pressure detector (Lib_Shock_Detector.f90:14-38) scales the
                              MUSCL slopes of every interior and block-connection face; 0: 'MUSCL'. */
  int32_t integration_vars;  /* integration-variables (Register_Numerics.f90:42, bound by
                              Mod_Newstate.f90:28-36): Q2D_INTVARS_CONS ('cons', Newstate_Conservative);
                              Q2D_INTVARS_PRIM ('prim', Newstate_Primitive, Lib_Newstate.f90:167-310) is
                              rejected with Q2D_E_UNSUPPORTED */
} q2d_scheme;

/* One structured block, in the reference's allocation layout (see top). */
typedef struct q2d_block_desc {
  int32_t ni, nj;
  int32_t case_id;          /* index of the independent solve this block belongs to */
  int32_t reserved;
  const double *P;          /* initial state incl. ghost layers */
  const double *dl;
  const double *M;
  const double *sur;
  const double *face_i;
  const double *face_j;
  const double *node;       /* viscous/axisymmetric only (may be NULL otherwise) */
  const double *src_a;      /* SRC_AXI: axis_r_ave ; SRC_AREA: area_geom_term_x */
  const double *src_b;      /* SRC_AXI: axis_inv_r ; SRC_AREA: area_geom_term_y */
} q2d_block_desc;

/* One boundary-cell record of domain%bc (Advanced_Types_m.f90:78-122, read by
 * IO_BC.f90:163-...).  b/bs are 1-based block numbers inside the record's case;
 * i, j, is, js are 1-based cell indices; f, fs in 1..4.  Records of one case
 * are ordered by block then face, as BC_Fluxes assumes (Mod_BC_Fluxes.f90:79-86).
 * For a connection (101/201) of a laminar case, d11 (+1 or -1) orients the donor
 * along the face (Fill_BC_Ghost_Connection, Lib_Ghost.f90:302-326). */
typedef struct q2d_bc_entry {
  int32_t b, i, j, f, type;
  int32_t bs, is, js, fs, d11, d12, d21, d22;
  int32_t case_id;
  int32_t reserved[2];
  double  mach;             /* 405: inflow Mach number */
  double  T;                /* 405: static temperature ; 302: wall temperature Tw */
  double  p;                /* 405: static pressure ; 406: exit (back) pressure */
  double  rel_fac;          /* 405/406 relaxation factor */
  double  alpha;            /* 405: flow angle [rad]; >= 0.5*DBL_MAX means 'normal' */
} q2d_bc_entry;

/* One independent solve. */
typedef struct q2d_case_desc {
  q2d_scheme scheme;
  int32_t    itermax;       /* iter-threshold */
  int32_t    reserved;
  double     res_threshold; /* res-threshold (absolute, on residual(1)) */
  double     time_threshold;
  double     res_drop;      /* > 0: also stop when residual(1) <= res_drop * residual(1) of the case's
                               first step (a relative drop; serving); 0: off, as in the reference */
  double     time0;         /* domain%time when the run starts (Wrap_Setup.f90:129-152): 0 for a new
                               time-accurate run, the solution's SOLUTIONTIME on a restart of one, -1 for
                               a steady run; time-accurate steps add dtglobal to it, and the
                               time-threshold test reads it */
  const double *thermo;     /* NULL: calorically perfect gas (gas_cp constant, the closed forms).  Else
                               the one species' property tables on integer kelvins, FLINT's native
                               tabulated mode (Load_ThermoTransport.f90:15-71 reads INPUT/thermo.dat and
                               transport.dat): four planes of thermo_n doubles, node k at T = k K, k = 0 ..
                               thermo_n - 1: enthalpy h, cp, viscosity mu, conductivity k (mu and k are
                               read by laminar cases only).  Properties are interpolated linearly, from
                               node int(T) (Lib_Riemann_HLL.f90:257-282); T beyond the last interval
                               extrapolates the end interval.  gas_R stays the species' R = Runiv / W;
                               gas_cp is not read.  Both forms: the mixed form's FP32 faces interpolate
                               the tables in FP32 and its FP64 update inverts e(T) on them.  Copied at
                               create / add_cases. */
  int32_t    thermo_n;      /* nodes per plane (>= 2) when thermo != NULL */
  int32_t    reserved2;
} q2d_case_desc;

typedef struct q2d_problem {
  int32_t nblocks, nbc, ncases, reserved;
  const q2d_block_desc *blocks;    /* blocks of all cases; block k of case c is blocks[case_first_block(c)+k-1] */
  const q2d_bc_entry   *bc;        /* all records, grouped by case, then block, then face */
  const q2d_case_desc  *cases;
} q2d_problem;

/* Per-case result of one explicit step (Mod_Explicit.f90:36-174). */
typedef struct q2d_step_out {
  double  dtglobal;         /* domain%dtglobal */
  double  time;             /* domain%time */
  double  residual[4];      /* residuotot(1:nres): sqrt of the summed squares */
  int32_t iter;             /* domain%iter after the step */
  int32_t todo;             /* 1 continue, 3 end of simulation (iter/res/time thresholds) */
  int32_t status;           /* Q2D_OK or Q2D_E_STATE / Q2D_E_INTERFACE: the stop of the first
                               failure in the reference's order (RK stage, then interface states
                               before the new state, then block, j, i) */
  int32_t fail_b, fail_i, fail_j;   /* that failure's cell or face (block 1-based within case) */
  int32_t reserved[2];
} q2d_step_out;

typedef struct q2d_handle q2d_handle;

/* Library / device information. */
const char *q2d_version(void);
const char *q2d_last_error(void);
int q2d_device_count(int *count);

/* Create the device-resident solver for a problem (all cases). Copies every
 * array to device memory; the caller may free its buffers afterwards.
 * Replaces the device side of the set-up before the time loop: the arrays
 * Allocate_Data fills (Mod_Allocate_Data.f90:86-180), Setup_Metrics
 * (Mod_Metrics.f90:11-148) and the BC list read by IO_BC.f90:163-...; the
 * scheme keys bound by Assign_Setup (Assign_Setup.f90:11-62). */
int q2d_gpu_create(const q2d_problem *prob, q2d_handle **out);
int q2d_gpu_destroy(q2d_handle *h);

/* Continuous batching (no counterpart in the reference, whose domain is fixed
 * at set-up): append the cases of `prob` to a live handle (their case indices
 * start at *first_case; block indices continue after the existing blocks), or
 * remove one case, releasing its device memory.  The other cases' state and
 * results are untouched; a removed case keeps its index and reports todo 0.
 * q2d_gpu_create accepts an empty problem (0 cases) for a handle filled later. */
int q2d_gpu_add_cases(q2d_handle *h, const q2d_problem *prob, int32_t *first_case);
int q2d_gpu_remove_case(q2d_handle *h, int32_t case_id);
/* Case and block counts of the handle (removed ones included) and the bytes of
 * device memory its blocks hold. */
int q2d_gpu_case_count(q2d_handle *h, int32_t *ncases, int32_t *nblocks, int64_t *device_bytes);

/* The keys of one case that the reference's host may change while it runs: every ini-diter steps
 * Q2D_postprocess re-reads input.ini (Read_Inifile_Runtime, Read_Ini.f90:65-88, called from
 * Wrap_Postprocess.f90:288-291), so the next step's Compute_dt and stop test see the new values.  The
 * next q2d_gpu_step / q2d_gpu_run after q2d_gpu_set_runtime evaluates its dt with them (its dt pass runs
 * again; a fused dt from the previous step is discarded).  Keys that would change the handle's memory
 * (irs-beta, the model, the number of RK stages, the arithmetic form) are not runtime keys here. */
typedef struct q2d_runtime {
  double  cfl, vnn;
  int32_t ramp_iter;        /* cfl-rise-threshold */
  int32_t itermax;          /* iter-threshold */
  double  res_threshold, time_threshold, res_drop;
  int32_t limiter;          /* Q2D_LIM_*  (flux-limiter / space-reconstruction first-order) */
  int32_t riemann;          /* Q2D_RIEMANN_* */
  int32_t space_sd;         /* MUSCL-SD on (1) or off (0) */
  int32_t reserved;
} q2d_runtime;
int q2d_gpu_set_runtime(q2d_handle *h, int32_t case_id, const q2d_runtime *rt);

/* Kernel fusion of the mixed form's laminar cases (no counterpart in the reference): visc = 1 (the
 * default) evaluates the diffusive face fluxes and the axisymmetric viscous source inside the laminar
 * stage kernel, in FP32 like the form's convective faces, for cases without the area-variation source and
 * without faces that mix wall and other records; visc = 0 computes them in FP64 in a separate kernel
 * first.  Both are held to the mixed form's tolerance against the FP64 form; the FP64 form is not
 * affected. */
int q2d_gpu_set_fusion(q2d_handle *h, int32_t visc);
int q2d_gpu_get_runtime(q2d_handle *h, int32_t case_id, q2d_runtime *rt);

/* Exactly one Explicit_Step (Mod_Explicit.f90:10-174, called from Q2D_solve,
 * Wrap_Solve.f90:19-32) for every active case; out[] has ncases entries
 * (may be NULL).  Synchronises with the device. */
int q2d_gpu_step(q2d_handle *h, q2d_step_out *out);

/* Allocation-free step (no counterpart in the reference): exactly what q2d_gpu_step does, for callers
 * that cannot allocate device memory while they run (serving loops and harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or allocating library routine and captures no graph: the
 * step's scratch (per-work-item residual-norm and dt partials, per-record boundary fluxes, the mixed
 * form's FP32 per-face array) is taken from one file-scope device array of Q2D_NOALLOC_SCRATCH_BYTES
 * that each GPU holds for the life of the process, instead of the handle's own.  Returns Q2D_E_NOMEM,
 * computing nothing, when the handle's scratch exceeds that size, and Q2D_E_ARG with kernel timing on or
 * another device current.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * race.  Each step writes every part of the scratch it reads, so nothing carries over between calls;
 * q2d_gpu_step, q2d_gpu_run and this entry point may be mixed on one handle. */
#define Q2D_NOALLOC_SCRATCH_BYTES (32u << 20)
int q2d_gpu_step_noalloc(q2d_handle *h, q2d_step_out *out);

/* nsteps explicit steps without host synchronisation in between (replays of
 * captured HIP graphs of eight steps and of one step): the host's time loop around Explicit_Step
 * (main.f90:48 -> Wrap_Solve.f90:24) with its TODO test (Mod_Explicit.f90:
 * 120-168) evaluated on the device: a case whose todo became 3 is frozen.
 * out[] as for step. */
int q2d_gpu_run(q2d_handle *h, int32_t nsteps, q2d_step_out *out);

/* Capture the step graphs that q2d_gpu_run replays (eight steps and one step) now, so that the next
 * q2d_gpu_run spends no host time on capture and instantiation; no step runs and no device memory is
 * allocated.  No counterpart in the reference (for time loops and serving loops that keep graph capture
 * out of their timed or latency-critical region).  Before the handle's first step after create,
 * upload_state, add_cases or set_runtime (which runs eagerly with its dt pass) the call does nothing; any
 * change of the handle's launches (those calls, remove_case, tiling, timing, fusion) drops the graphs
 * again. */
int q2d_gpu_capture(q2d_handle *h);

/* Copy block `block` (0-based global index) state to/from P in the reference
 * layout P(nprim,1-gc:ni+gc,1-gc:nj+gc): what the host needs for a solution
 * write or a restart (Wrap_Postprocess.f90:67-113; domain%blk(b)%P). */
int q2d_gpu_download_state(q2d_handle *h, int32_t block, double *P);
int q2d_gpu_upload_state(q2d_handle *h, int32_t block, const double *P);

/* Local time step of the last step, domain%blk(b)%dtlocal(1:ni,1:nj)
 * (Mod_dt.f90:46). */
int q2d_gpu_download_dt(q2d_handle *h, int32_t block, double *dtlocal);

/* Kernel-time accounting (no counterpart in the reference): while enabled, every
 * stage-kernel launch records its first workgroup start and last wavefront end
 * on the device's constant-rate wall clock (two atomics per workgroup), in a
 * ring of 64 steps that the host reads every 64 steps and at the end of each
 * call; steps still replay the captured graphs.  Enabling (again) clears the
 * statistics.  stats returns the number of timed stage launches and their summed
 * duration in milliseconds. */
int q2d_gpu_set_timing(q2d_handle *h, int32_t enable);
int q2d_gpu_timing_stats(q2d_handle *h, int64_t *launches, double *ms_total);

/* Launch-shape control of the stage kernel (rows per work item, waves per
 * workgroup); no counterpart in the reference.  0 keeps the current value;
 * rows_per_item -1 picks, for every stage launch, the row count whose work
 * items finish soonest on the wavefront slots the device holds at the
 * kernel's occupancy.  A new handle uses 4 waves and its form's default rows
 * (get_tiling reports 0): balanced for FP64, the smaller of balanced and 16 for
 * the mixed form.
 * Used by the tile autotuner. */
int q2d_gpu_set_tiling(q2d_handle *h, int32_t rows_per_item, int32_t waves_per_group);
int q2d_gpu_get_tiling(q2d_handle *h, int32_t *rows_per_item, int32_t *waves_per_group);

/* HBM reference (no counterpart in the reference): time `reps` launches of a plain device copy
 * of nbytes (read nbytes + write nbytes); measures the achievable bandwidth and calibrates HBM
 * byte counters. */
int q2d_hbm_copy(int64_t nbytes, int32_t reps, double *ms_per_copy);

/* Work-item count and cell count of the current launch (no counterpart in the reference). */
int q2d_gpu_info(q2d_handle *h, int64_t *cells, int64_t *work_items);

#ifdef __cplusplus
}
#endif

#endif /* Q2D_GPU_H */
