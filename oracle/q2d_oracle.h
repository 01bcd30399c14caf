/*
 * q2d_oracle.h — CPU restatement of Q2D's explicit finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; nothing in the product links, loads or calls it.  It is used by
 * tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg.
 *
 * The restatement follows the reference Fortran in operation order (every
 * function cites the file:line it restates), in FP64 with floating-point
 * contraction disabled, so that a bit-faithful GPU implementation can be
 * compared to it exactly.
 */
#ifndef Q2D_ORACLE_H
#define Q2D_ORACLE_H

#include "../include/q2d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct q2do q2do;

/* Build the oracle state from the same problem description the C-ABI takes. */
q2do *q2do_create(const q2d_problem *prob);
void  q2do_destroy(q2do *o);

/* One Explicit_Step for every case (Mod_Explicit.f90:10-174). Returns Q2D_OK or
 * the first failure code; out has ncases entries (may be NULL). */
int q2do_step(q2do *o, q2d_step_out *out);

/* State of block `block` (0-based global) in the reference layout. */
int q2do_get_state(q2do *o, int block, double *P);
int q2do_set_state(q2do *o, int block, const double *P);
int q2do_get_dt(q2do *o, int block, double *dtlocal);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int q2do_set_threads(int n);
int q2do_get_threads(void);

/* ---- unit-level entry points (tests pin these against closed forms) ---- */
double q2do_limiter(int id, double x, double y);
/* Reconstruction (Lib_Reconstruction.f90:10-29) of one scalar. */
void q2do_reconstruct(int limiter, double prev, double local, double next, double next2,
                      double dl0, double dl1, double dl2, double dll, double dlr, double beta,
                      double *priml, double *primr);
/* Riemann flux (Lib_Riemann_HLL.f90) for primitive L/R states; F[4] = F_r,F_u,F_v,F_E. */
void q2do_riemann(int id, double gas_R, double gas_cp,
                  const double *primL, const double *primR, double beta, double nx, double ny, double *F);
/* Roe averages (Lib_Riemann_HLL.f90:229-317): out = h0l,h0r,rho,u,v,a,h0. */
void q2do_roe_averages(double gas_R, double gas_cp, const double *primL, const double *primR, double *out);
/* Convective_Flux (Lib_Convective.f90:15-82) for one face: stencil prim[4][4]
 * (cells -1..2), dl[4]; returns the flux added to cell 0 (and subtracted from cell 1). */
int q2do_convective_flux(int limiter, int riemann, double gas_R, double gas_cp,
                         const double *dl, double nx, double ny, double int_l,
                         const double *prim, double *flux);
/* SD_Tramel (Lib_Shock_Detector.f90:14-38) of a cell with pressure p and neighbours i-1, i+1, j-1, j+1. */
double q2do_sd_tramel(double p, double p_im, double p_ip, double p_jm, double p_jp);
/* Residual array R(nprim, 1-gc:ni+gc, 1-gc:nj+gc) of block `block` (0-based global), and
 * Residual_Smoothing (Lib_IRS.f90:17-44) of case `case_id` applied to those arrays as they stand. */
int q2do_get_residual(q2do *o, int block, double *R);
int q2do_set_residual(q2do *o, int block, const double *R);
int q2do_irs_smooth(q2do *o, int case_id, double beta);

/* x^y of the characteristic outflow (exact_pow: the same exactly specified operations as the device's). */
double q2do_pow(double x, double y);
/* Boundary-face fluxes of one record from its cells (no ghost fill), the flux before the face sign:
 * Supersonic_Inflow (Lib_BC_Fluxes_Inflow.f90:318-403) with the inflow direction alpha [rad], and
 * Outflow (Lib_BC_Fluxes_Outflow.f90:89-165).  Return 1 where the reference falls back to BC_Symmetry_Eul. */
int q2do_bc_inflow_flux(int limiter, int riemann, double gas_R, double gas_cp, int face, const double *Bp,
                        const double *Ip, double n1, double n2, double int_l, double mach, double T, double p,
                        double rel_fac, double alpha, double *F);
int q2do_bc_outflow_flux(double gas_R, double gas_cp, int face, const double *Bp, double n1, double n2,
                         double int_l, double p_exit, double rel_fac, double *F);

/* Planted mistakes, for tests that must show they detect them (0 restores the restatement).
 * Returns the previous setting. */
#define Q2DO_PLANT_NONE          0
#define Q2DO_PLANT_AXI_SIGN      1   /* Source_Axi_Eul_Blk with the sign of the source reversed */
#define Q2DO_PLANT_AXI_ENERGY    2   /* Source_Axi_Eul_Blk energy term without + p */
#define Q2DO_PLANT_INFLOW_ANGLE  3   /* Supersonic_Inflow with cos / sin of the flow angle swapped */
#define Q2DO_PLANT_INFLOW_RHO    4   /* Supersonic_Inflow density p / T (gas constant dropped) */
#define Q2DO_PLANT_OUTFLOW_RIEM  5   /* Outflow Riemann invariant with the sign of the Un term reversed */
#define Q2DO_PLANT_OUTFLOW_ISEN  6   /* Outflow isentropic exponent (g-1)/g instead of (g-1)/(2g) */
#define Q2DO_PLANT_RAMP          7   /* CFL ramp with iter - 1 in place of iter */
#define Q2DO_PLANT_TAB_NODE      8   /* table interpolation from node Ti + 1 instead of Ti */
#define Q2DO_PLANT_TAB_TOFE      9   /* table-mode cons2prim: one Newton step, whatever its interval */
int q2do_plant(int which);
/* The runtime keys of one case (as q2d_gpu_set_runtime). */
int q2do_set_runtime(q2do *o, int case_id, const q2d_runtime *rt);
/* Table mode for the unit entry points (riemann, roe_averages, convective_flux, the BC fluxes and the
 * functions below): planes h, cp, mu, k of n nodes (node k at T = k), or NULL for the closed form.  The
 * oracle keeps the pointer, not a copy. */
void q2do_unit_tables(const double *tab, int n);
/* what: 0 h(T=a), 1 cp(T=a), 2 mu(T=a), 3 k(T=a), 4 gamma(T=a), 5 e(T=a), 6 T of e=a from estimate b,
 * 7 speed of sound at p=a, rho=b. */
double q2do_thermo_eval(int what, double gas_R, double gas_cp, double a, double b);
void q2do_prim2cons(double gas_R, double gas_cp, const double *prim, double *cons);
void q2do_cons2prim(double gas_R, double gas_cp, const double *cons, double T0, double *prim);

#ifdef __cplusplus
}
#endif

#endif
