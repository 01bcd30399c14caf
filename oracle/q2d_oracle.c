/*
 * q2d_oracle.c — CPU restatement of Q2D's explicit step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
 * file.  Built with -ffp-contract=off and without fast-math: each expression
 * below keeps the evaluation order of the Fortran it restates (Fortran and C
 * both evaluate a*b*c as (a*b)*c and a+b-c as (a+b)-c), so the FP64 results
 * are those of the reference source read literally.
 *
 * Gas closure (the reference calls the absent FLINT library for these):
 * single species, constant cp, gas constant R, cv = cp - R, gamma = cp/cv.
 *   T(p,rho)   = p / (rho*R)
 *   h(T)       = cp*T            (FLINT h_tab interpolation for constant cp)
 *   a(p,rho)   = sqrt(gamma*p/rho)                                (f_ss)
 *   H0         = cp*T + 0.5*|V|*|V|                               (H0)
 *   E0         = cv*T + 0.5*|V|*|V|                               (E0)
 *   prim2cons  = (rho, rho*u, rho*v, rho*(cv*T + 0.5*(u*u+v*v)))
 *   cons2prim  : u=U2/rho, v=U3/rho, e=U4/rho-0.5*(u*u+v*v), T=e/cv, p=rho*R*T
 * Laminar transport (model 1), the closure of co_k_mi_lam_Wilke(_expr) /
 * f_laminarViscosity for one species:
 *   mu(T)      = mu_ref*(x*sqrt(x))*(T_ref+S)/(T+S),  x = T/T_ref  (Sutherland)
 *   k(mu)      = mu*cp/Pr
 * Parity of these closures with FLINT is unpinned: FLINT is an un-vendored
 * submodule (.gitmodules:4-6) with no pinned revision; its tables interpolate
 * linearly between integer kelvins, which this closed form does not.
 * Everything above the closures is pinned by tests/ (golden Shu-Osher vectors
 * of the reference binary, closed forms, invariants).
 */
#include "q2d_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GC 2
#define NP 4

/* ------------------------------------------------------------------------ */
/* gas closure                                                               */
/* ------------------------------------------------------------------------ */
/* tab: NULL for the calorically perfect closed forms; else the table mode's planes h, cp, mu, k of ntab
 * nodes each (node k at T = k). */
typedef struct { double R, cp, cv, gam, mu_ref, T_ref, S, Pr; const double *tab; int ntab; } gas_t;

static gas_t mkgas(double R, double cp) {
  gas_t g; g.R = R; g.cp = cp; g.cv = cp - R; g.gam = cp / g.cv;
  g.mu_ref = 0.0; g.T_ref = 1.0; g.S = 0.0; g.Pr = 0.72; g.tab = NULL; g.ntab = 0; return g;
}
/* Planted mistakes (tests only, q2do_plant): each makes one restated formula wrong in a way a test of
 * that formula must detect.  0 = none. */
static int g_plant = 0;

/* Table mode — FLINT's native tabulated properties of one species (the library is not in the reference
 * tree; docs/theory/thermo.md:39-41 states the linear interpolation, Lib_Riemann_HLL.f90:257-282 and
 * Lib_Newstate.f90:267-282 show its arithmetic): a property at T is v(Ti) + (v(Ti+1) - v(Ti)) (T - Ti)
 * with Ti = int(T).  Ti is clamped to the table's intervals (T beyond them extrapolates the end interval);
 * T < 0 or NaN takes interval 0. */
enum { TAB_H = 0, TAB_CP = 1, TAB_MU = 2, TAB_K = 3 };
static int tab_node(const gas_t *g, double T) {
  if (!(T >= 1.0)) return 0;
  if (!(T < (double)(g->ntab - 1))) return g->ntab - 2;
  return (int)T;
}
static double tab_interp(const gas_t *g, int q, int Ti, double dT) {
  const double *v = g->tab + (size_t)q * (size_t)g->ntab;
  if (g_plant == Q2DO_PLANT_TAB_NODE) return v[Ti + 1] + (v[Ti + 1] - v[Ti]) * dT;
  return v[Ti] + (v[Ti + 1] - v[Ti]) * dT;
}
static double tabT(const gas_t *g, int q, double T) {
  const int Ti = tab_node(g, T);
  return tab_interp(g, q, Ti, T - (double)Ti);
}
/* f_gamma: cp / (cp - R) at T */
static double f_gamma_T(const gas_t *g, double T) {
  if (!g->tab) return g->gam;
  const double cp = tabT(g, TAB_CP, T);
  return cp / (cp - g->R);
}
/* cons2prim's temperature in table mode (FLINT's cons2prim, absent here; restated as the exact inverse of
 * the interpolated e(T) = h(T) - R T): Newton steps from the caller's estimate T0 (the stage input's
 * p / (rho R), Lib_Newstate.f90:119-123), each with the slope of the current interval, so a step that
 * lands in its own interval is the root of the piecewise-linear e(T); at most 32 steps. */
static double t_of_e(const gas_t *g, double e, double T0) {
  const double *h = g->tab;
  double T = T0;
  for (int it = 0; it < 32; it++) {
    const int Ti = tab_node(g, T);
    const double e0 = h[Ti] - g->R * (double)Ti;
    const double sl = (h[Ti + 1] - h[Ti]) - g->R;
    const double Tn = (double)Ti + (e - e0) / sl;
    if (g_plant == Q2DO_PLANT_TAB_TOFE) return Tn;
    if (Tn >= (double)Ti && Tn <= (double)(Ti + 1)) return Tn;
    if ((Ti == 0 && Tn < 0.0) || (Ti == g->ntab - 2 && Tn > (double)(Ti + 1))) return Tn;
    T = Tn;
  }
  return T;
}

/* x^y for x > 0, from IEEE-754 binary64 +, -, *, /, fma, rint, frexp and ldexp only (each exactly
 * specified), so that any conforming evaluation of this sequence gives the same bits; the device
 * evaluates the same sequence.  Method: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
 * s = (m - 1) / (m + 1), with s and the leading term carried in double-double and the odd series
 * 2 s^3 (1/3 + s^2/5 + ...) in double; ln x = e ln2 + ln m and t = y ln x in double-double (Cody-Waite
 * split of ln 2); exp(t) = 2^k exp(r), k = rint(t / ln 2), r = t - k ln 2, exp(r) by its Taylor series
 * to r^13/13!.  The reference's `**` is the C library's pow; tests hold this within 1 ulp of it. */
static void two_sum(double a, double b, double *s, double *e) {
  double t = a + b, bb = t - a;
  *s = t; *e = (a - (t - bb)) + (b - bb);
}
static double exact_pow(double x, double y) {
  if (!(x > 0.0) || isinf(x) || isinf(y) || isnan(y)) return NAN;
  if (y == 0.0) return 1.0;
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  int e;
  double m = frexp(x, &e);                              /* m in [0.5, 1) */
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  const double num = m - 1.0;                           /* exact (Sterbenz) */
  double den, den_lo;
  two_sum(m, 1.0, &den, &den_lo);
  const double s = num / den;
  const double s_lo = (fma(-s, den, num) - s * den_lo) / den;
  const double s2 = s * s;
  double q = 1.0 / 23.0;
  q = q * s2 + 1.0 / 21.0; q = q * s2 + 1.0 / 19.0; q = q * s2 + 1.0 / 17.0; q = q * s2 + 1.0 / 15.0;
  q = q * s2 + 1.0 / 13.0; q = q * s2 + 1.0 / 11.0; q = q * s2 + 1.0 / 9.0; q = q * s2 + 1.0 / 7.0;
  q = q * s2 + 1.0 / 5.0; q = q * s2 + 1.0 / 3.0;
  const double tail = 2.0 * s * s2 * q;
  double lm, lm_lo;
  two_sum(2.0 * s, 2.0 * s_lo + tail, &lm, &lm_lo);      /* ln m */
  const double de = (double)e;
  double lh, ll;
  two_sum(de * ln2_hi, lm, &lh, &ll);                    /* ln x = lh + ll */
  ll = ll + (lm_lo + de * ln2_lo);
  double th = y * lh;
  double tl = fma(y, lh, -th) + y * ll;                  /* t = y ln x */
  two_sum(th, tl, &th, &tl);
  const double k = rint(th * inv_ln2);
  const double r_hi = th - k * ln2_hi;                   /* exact for |k| < 2^10 (Sterbenz) */
  const double r_lo = tl - k * ln2_lo;
  double r, r_err;
  two_sum(r_hi, r_lo, &r, &r_err);                       /* r + r_err = t - k ln 2 */
  double p = 1.0 / 6227020800.0;                         /* 1/13! */
  p = p * r + 1.0 / 479001600.0; p = p * r + 1.0 / 39916800.0; p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0; p = p * r + 1.0 / 40320.0; p = p * r + 1.0 / 5040.0; p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0; p = p * r + 1.0 / 24.0; p = p * r + 1.0 / 6.0; p = p * r + 0.5;
  const double em1 = r + (r_err + r * r * p);            /* exp(r) - 1 */
  return ldexp(1.0 + em1, (int)k);
}

/* Laminar transport (FLINT f_laminarViscosity / co_k_mi_lam_Wilke, closed form):
 * Sutherland mu = mu_ref * (T/T_ref)^1.5 * (T_ref + S) / (T + S), with (T/T_ref)^1.5 = x*sqrt(x);
 * k = mu * cp / Pr. */
static double mu_lam(const gas_t *g, double T) {
  if (g->tab) return tabT(g, TAB_MU, T);
  double x = T / g->T_ref;
  return g->mu_ref * (x * sqrt(x)) * (g->T_ref + g->S) / (T + g->S);
}
/* conductivity of viscosity mu at temperature T (table mode: the k plane) */
static double k_lam(const gas_t *g, double mu, double T) {
  if (g->tab) return tabT(g, TAB_K, T);
  return mu * g->cp / g->Pr;
}
static double eos_T(const gas_t *g, double p, double rho) { return p / (rho * g->R); }
static double f_ss(const gas_t *g, double p, double rho) {
  if (g->tab) return sqrt(f_gamma_T(g, eos_T(g, p, rho)) * p / rho);
  return sqrt(g->gam * p / rho);
}
/* specific internal energy e(T): cv T, or h(T) - R T in table mode */
static double f_e(const gas_t *g, double T) {
  if (g->tab) return tabT(g, TAB_H, T) - g->R * T;
  return g->cv * T;
}
static double f_H0(const gas_t *g, double p, double rho, double vmag) {
  double T = eos_T(g, p, rho);
  if (g->tab) return tabT(g, TAB_H, T) + 0.5 * vmag * vmag;
  return g->cp * T + 0.5 * vmag * vmag;
}
static double f_E0(const gas_t *g, double p, double rho, double vmag) {
  double T = eos_T(g, p, rho);
  return f_e(g, T) + 0.5 * vmag * vmag;
}
static void prim2cons(const gas_t *g, const double *pr, double *u) {
  double rho = pr[0];
  double T = eos_T(g, pr[3], rho);
  u[0] = rho;
  u[1] = rho * pr[1];
  u[2] = rho * pr[2];
  u[3] = rho * (f_e(g, T) + 0.5 * (pr[1] * pr[1] + pr[2] * pr[2]));
}
/* T0: the temperature estimate table mode starts from (unused by the closed form) */
static void cons2prim(const gas_t *g, const double *u, double *pr, double T0) {
  double rho = u[0];
  double vx = u[1] / rho, vy = u[2] / rho;
  double e = u[3] / rho - 0.5 * (vx * vx + vy * vy);
  double T = g->tab ? t_of_e(g, e, T0) : e / g->cv;
  pr[0] = rho; pr[1] = vx; pr[2] = vy; pr[3] = rho * g->R * T;
}
/* check_gas_state: a reconstructed state is physical when rho>0 and p>0. */
static int check_gas_state(double rho, double p) { return (rho > 0.0 && p > 0.0) ? 0 : 1; }

/* Fortran MIN/MAX of three reals, evaluated left to right. */
static double min3(double a, double b, double c) { double m = a; if (b < m) m = b; if (c < m) m = c; return m; }
static double max3(double a, double b, double c) { double m = a; if (b > m) m = b; if (c > m) m = c; return m; }

/* ------------------------------------------------------------------------ */
/* limiters — Lib_Limiters.f90:26-112                                        */
/* ------------------------------------------------------------------------ */
static double lim_minmod(double x, double y) {          /* :36-46 */
  double r = 0.0;
  if (x * y < 0.0) return r;
  r = x;
  if (fabs(x) > fabs(y)) r = y;
  return r;
}
static double lim_vanleer(double x, double y) {         /* :49-58 */
  double ru = y / (x + 1e-30);
  double phi = (ru + fabs(ru)) / (1.0 + fabs(ru));
  return phi * x;
}
static double lim_vanalbada(double x, double y) {       /* :61-70 */
  double ru = y / (x + 1e-30);
  double phi = (ru * ru + ru) / (1.0 + ru * ru);
  return phi * x;
}
static double lim_mc(double x, double y) {              /* :73-82: maxval([0, minval([2r, (1+r)/2, 2])]) */
  double ru = y / (x + 1e-30);
  double m = min3(2.0 * ru, 0.5 * (1.0 + ru), 2.0);
  double phi = 0.0; if (m > phi) phi = m;
  return phi * x;
}
static double lim_sb(double x, double y) {              /* :85-94: maxval([0, min(2r,1), min(r,2)]) */
  double ru = y / (x + 1e-30);
  double a = 2.0 * ru; if (1.0 < a) a = 1.0;
  double b = ru; if (2.0 < b) b = 2.0;
  return max3(0.0, a, b) * x;
}
static double lim_limo3(double x, double y) {           /* :98-110 */
  double ru = y / (x + 1e-30);
  double t = (2.0 + ru) / 3.0;
  double inner = min3(2.0 * ru, t, 1.6);
  double mid = max3(-0.5 * ru, 0.0, inner);
  double outer = t; if (mid < outer) outer = mid;
  double phi = 0.0; if (outer > phi) phi = outer;
  return phi * x;
}
static double limiter(int id, double x, double y) {
  switch (id) {
    case Q2D_LIM_MINMOD: return lim_minmod(x, y);
    case Q2D_LIM_VANLEER: return lim_vanleer(x, y);
    case Q2D_LIM_VANALBADA: return lim_vanalbada(x, y);
    case Q2D_LIM_MC: return lim_mc(x, y);
    case Q2D_LIM_SUPERBEE: return lim_sb(x, y);
    case Q2D_LIM_LIMO3: return lim_limo3(x, y);
    default: return 0.0;                                /* rlimiter_IORD :141-148 */
  }
}

/* Reconstruction — Lib_Reconstruction.f90:10-29 */
static void reconstruct(int lim, double prev, double local, double next, double next2,
                        double dl0, double dl1, double dl2, double dll, double dlr, double beta,
                        double *pl, double *pr) {
  double s0 = (local - prev) / dl0;
  double s1 = (next - local) / dl1;
  double s2 = (next2 - next) / dl2;
  double sl = limiter(lim, s1, s0);
  double sr = limiter(lim, s2, s1);
  *pl = local + beta * sl * dll;
  *pr = next - beta * sr * dlr;
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers — Lib_Riemann_HLL.f90                                      */
/* ------------------------------------------------------------------------ */
typedef struct { double h0l, h0r, rho, u, v, a, h0; } roe_t;

/* roe_averages :229-317 (one species): the closed form (h = cp T, cp(T) = cp) or the table arithmetic */
static roe_t roe_averages(const gas_t *g, double dl, double dr, double dltot, double drtot,
                          double ul, double vl, double ur, double vr, double pl, double pr) {
  roe_t o;
  double Rl = g->R, Rr = g->R;
  double inv_dltot = 1.0 / dltot;
  double inv_drtot = 1.0 / drtot;
  double Tl = pl * inv_dltot / Rl;
  double Tr = pr * inv_drtot / Rr;
  double srL = sqrt(dltot);
  double srR = sqrt(drtot);
  double inv_sr = 1.0 / (srL + srR);
  o.rho = srL * srR;
  o.u = (srR * ur + srL * ul) * inv_sr;
  o.v = (srR * vr + srL * vl) * inv_sr;
  double R_roe = (srR * Rr + srL * Rl) * inv_sr;
  double T_roe = (srR * Tr + srL * Tl) * inv_sr;
  double h0l = 0.0, h0r = 0.0, cv_roe = 0.0, sum_ei = 0.0;
  if (g->tab) {                                         /* the table arithmetic of :257-292 */
    const int Til = tab_node(g, Tl), Tir = tab_node(g, Tr);
    const double dTl = Tl - (double)Til, dTr = Tr - (double)Tir;
    const double *cp = g->tab + (size_t)TAB_CP * (size_t)g->ntab;
    double invW = g->R;
    double hl = tab_interp(g, TAB_H, Til, dTl);
    double hr = tab_interp(g, TAB_H, Tir, dTr);
    double el = hl - invW * Tl;
    double er = hr - invW * Tr;
    h0l = h0l + hl * dl * inv_dltot;
    h0r = h0r + hr * dr * inv_drtot;
    double d_roe = (srR * dr * inv_drtot + srL * dl * inv_dltot) * inv_sr;
    cv_roe = cv_roe + d_roe * (0.5 * (((cp[Til] + (cp[Til + 1] - cp[Til]) * dTl) + cp[Tir]) +
                                      (cp[Tir + 1] - cp[Tir]) * dTr) - invW);
    sum_ei = sum_ei + d_roe * (srR * er + srL * el) * inv_sr;
  } else {
    double invW = g->R;
    double hl = g->cp * Tl;
    double hr = g->cp * Tr;
    double el = hl - invW * Tl;
    double er = hr - invW * Tr;
    h0l = h0l + hl * dl * inv_dltot;
    h0r = h0r + hr * dr * inv_drtot;
    double d_roe = (srR * dr * inv_drtot + srL * dl * inv_dltot) * inv_sr;
    cv_roe = cv_roe + d_roe * (0.5 * (g->cp + g->cp) - invW);
    sum_ei = sum_ei + d_roe * (srR * er + srL * el) * inv_sr;
  }
  h0l = h0l + 0.5 * (ul * ul + vl * vl);
  h0r = h0r + 0.5 * (ur * ur + vr * vr);
  o.h0 = (srR * h0r + srL * h0l) * inv_sr;
  double gam_roe = 1.0 + R_roe / cv_roe;
  double vel2 = o.u * o.u + o.v * o.v;
  o.a = sqrt((gam_roe - 1.0) * (o.h0 - 0.5 * vel2 + cv_roe * T_roe - sum_ei));
  o.h0l = h0l; o.h0r = h0r;
  return o;
}

/* riemann_HLLC :34-123 */
static void riemann_hllc(const gas_t *g, double dl, double ul, double vl, double pl, double al, double dltot,
                         double dr, double ur, double vr, double pr, double ar, double drtot,
                         double nx, double ny, double *F) {
  double unl = ul * nx + vl * ny;
  double unr = ur * nx + vr * ny;
  roe_t ra = roe_averages(g, dl, dr, dltot, drtot, ul, vl, ur, vr, pl, pr);
  double h0l = ra.h0l, h0r = ra.h0r;
  double un_roe = ra.u * nx + ra.v * ny;
  double e0l = h0l - pl / dltot;
  double e0r = h0r - pr / drtot;
  double S1 = min3(0.0, unl - al, un_roe - ra.a);
  double S4 = max3(0.0, unr + ar, un_roe + ra.a);
  double inv_denom = 1.0 / (dltot * (S1 - unl) - drtot * (S4 - unr));
  double Sstar = (pr - pl + dltot * unl * (S1 - unl) - drtot * unr * (S4 - unr)) * inv_denom;
  double pstar = dltot * (unl - S1) * (unl - Sstar) + pl;
  double F_r, F_u, F_v, F_E;
  if (S1 >= 0.0) {
    F_r = dltot * unl;
    F_u = F_r * ul + pl * nx;
    F_v = F_r * vl + pl * ny;
    F_E = F_r * h0l;
  } else if (Sstar >= 0.0) {
    F_r = dltot * unl;
    F_u = F_r * ul + pl * nx;
    F_v = F_r * vl + pl * ny;
    F_E = F_r * h0l;
    inv_denom = 1.0 / (S1 - Sstar);
    double U1S = dltot * (S1 - unl) * inv_denom;
    double U2S = ((S1 - unl) * dltot * ul + (pstar - pl) * nx) * inv_denom;
    double U3S = ((S1 - unl) * dltot * vl + (pstar - pl) * ny) * inv_denom;
    double U5S = ((S1 - unl) * dltot * e0l - pl * unl + pstar * Sstar) * inv_denom;
    F_r = F_r + S1 * (U1S - dltot);
    F_u = F_u + S1 * (U2S - dltot * ul);
    F_v = F_v + S1 * (U3S - dltot * vl);
    F_E = F_E + S1 * (U5S - dltot * e0l);
  } else if (S4 > 0.0) {
    F_r = drtot * unr;
    F_u = F_r * ur + pr * nx;
    F_v = F_r * vr + pr * ny;
    F_E = F_r * h0r;
    inv_denom = 1.0 / (S4 - Sstar);
    double U1S = drtot * (S4 - unr) * inv_denom;
    double U2S = ((S4 - unr) * drtot * ur + (pstar - pr) * nx) * inv_denom;
    double U3S = ((S4 - unr) * drtot * vr + (pstar - pr) * ny) * inv_denom;
    double U5S = ((S4 - unr) * drtot * e0r - pr * unr + pstar * Sstar) * inv_denom;
    F_r = F_r + S4 * (U1S - drtot);
    F_u = F_u + S4 * (U2S - drtot * ur);
    F_v = F_v + S4 * (U3S - drtot * vr);
    F_E = F_E + S4 * (U5S - drtot * e0r);
  } else {
    F_r = drtot * unr;
    F_u = F_r * ur + pr * nx;
    F_v = F_r * vr + pr * ny;
    F_E = F_r * h0r;
  }
  F[0] = F_r; F[1] = F_u; F[2] = F_v; F[3] = F_E;
}

/* riemann_HLLE :126-169 */
static void riemann_hlle(const gas_t *g, double dl, double ul, double vl, double pl, double al, double dltot,
                         double dr, double ur, double vr, double pr, double ar, double drtot,
                         double nx, double ny, double *F) {
  double unl = ul * nx + vl * ny;
  double unr = ur * nx + vr * ny;
  roe_t ra = roe_averages(g, dl, dr, dltot, drtot, ul, vl, ur, vr, pl, pr);
  double h0l = ra.h0l, h0r = ra.h0r;
  double un_roe = ra.u * nx + ra.v * ny;
  double S1 = min3(0.0, unl - al, un_roe - ra.a);
  double S4 = max3(0.0, unr + ar, un_roe + ra.a);
  double inv_denom = 1.0 / (S4 - S1);
  double FrL = dltot * unl;
  double FrR = drtot * unr;
  F[0] = (S4 * FrL - S1 * FrR + S1 * S4 * (drtot - dltot)) * inv_denom;
  F[1] = (S4 * (FrL * ul + pl * nx) - S1 * (FrR * ur + pr * nx) + S1 * S4 * (drtot * ur - dltot * ul)) * inv_denom;
  F[2] = (S4 * (FrL * vl + pl * ny) - S1 * (FrR * vr + pr * ny) + S1 * S4 * (drtot * vr - dltot * vl)) * inv_denom;
  F[3] = (S4 * (FrL * h0l) - S1 * (FrR * h0r)
          + S1 * S4 * (drtot * (h0r - pr / drtot) - dltot * (h0l - pl / dltot))) * inv_denom;
}

/* fluxes_helper — Lib_Riemann_HLL.f90:334-350: inviscid flux from primitives (one species, H0 closure) */
static void fluxes_helper(const gas_t *g, double p, double r, double u, double v, double nx, double ny, double *F) {
  F[0] = r * (u * nx + v * ny);
  F[1] = F[0] * u + p * nx;
  F[2] = F[0] * v + p * ny;
  F[3] = F[0] * f_H0(g, p, r, sqrt(u * u + v * v));
}

/* riemann_HLLEM — Lib_Riemann_HLL.f90:355-432 (Einfeldt et al. 1991): the HLLE flux (energy through
 * e0 = h0 - p/rho) with the contact and the one 2D shear wave restored in the subsonic range; the
 * detector's beta is not used. */
static void riemann_hllem(const gas_t *g, double dl, double ul, double vl, double pl, double al, double dltot,
                          double dr, double ur, double vr, double pr, double ar, double drtot,
                          double nx, double ny, double *F) {
  roe_t ra = roe_averages(g, dl, dr, dltot, drtot, ul, vl, ur, vr, pl, pr);
  double un_roe = ra.u * nx + ra.v * ny;
  double S1 = min3(0.0, (ul * nx + vl * ny) - al, un_roe - ra.a);
  double S4 = max3(0.0, (ur * nx + vr * ny) + ar, un_roe + ra.a);
  double FL[4], FR[4];
  fluxes_helper(g, pl, dl, ul, vl, nx, ny, FL);
  fluxes_helper(g, pr, dr, ur, vr, nx, ny, FR);
  double e0l = ra.h0l - pl / dltot;
  double e0r = ra.h0r - pr / drtot;
  double inv_denom = 1.0 / (S4 - S1);
  F[0] = (S4 * FL[0] - S1 * FR[0] + S1 * S4 * (drtot - dltot)) * inv_denom;
  F[1] = (S4 * FL[1] - S1 * FR[1] + S1 * S4 * (drtot * ur - dltot * ul)) * inv_denom;
  F[2] = (S4 * FL[2] - S1 * FR[2] + S1 * S4 * (drtot * vr - dltot * vl)) * inv_denom;
  F[3] = (S4 * FL[3] - S1 * FR[3] + S1 * S4 * (drtot * e0r - dltot * e0l)) * inv_denom;
  if (S1 < 0.0 && S4 > 0.0) {
    double drho = drtot - dltot, dp = pr - pl, du = ur - ul, dv = vr - vl;
    double tx = -ny, ty = nx;
    double alpha_c = drho - dp / (ra.a * ra.a);
    double alpha_s = ra.rho * (du * tx + dv * ty);
    double delta = ra.a / (ra.a + fabs(0.5 * (S1 + S4)));
    double C = (S1 * S4) * inv_denom;
    F[0] = F[0] - C * delta * alpha_c;
    F[1] = F[1] - C * delta * (alpha_c * ra.u + alpha_s * tx);
    F[2] = F[2] - C * delta * (alpha_c * ra.v + alpha_s * ty);
    F[3] = F[3] - C * delta * (alpha_c * 0.5 * (ra.u * ra.u + ra.v * ra.v) + alpha_s * (ra.u * tx + ra.v * ty));
  }
}

/* hartenHyman — Lib_Riemann_HLL.f90:320-331 (entropy fix of the acoustic eigenvalues) */
static double harten_hyman(double v, double e) {
  double al = fabs(v);
  double l = al;
  if (al <= e) l = (v * v + e * e) / (2.0 * e);
  return l;
}

/* riemann_HLLEpp — Lib_Riemann_HLL.f90:435-530 (Tramel et al. 2009): the centred flux minus the Roe
 * wave decomposition (contact, one 2D shear wave, two acoustic waves), each wave's speed blended by
 * the detector between the Roe value (beta = 1; Harten-Hyman on the acoustic waves, |un| on the
 * contact and shear) and the HLLE value (beta = 0). */
static void riemann_hllepp(const gas_t *g, double dl, double ul, double vl, double pl, double al, double dltot,
                           double dr, double ur, double vr, double pr, double ar, double drtot, double beta,
                           double nx, double ny, double *F) {
  double unl = ul * nx + vl * ny;
  double unr = ur * nx + vr * ny;
  roe_t ra = roe_averages(g, dl, dr, dltot, drtot, ul, vl, ur, vr, pl, pr);
  double un_roe = ra.u * nx + ra.v * ny;
  double Sm = min3(0.0, unl - al, un_roe - ra.a);
  double Sp = max3(0.0, unr + ar, un_roe + ra.a);
  double inv_denom = 1.0 / (Sp - Sm);
  double l1h = fabs((Sp + Sm) * un_roe * inv_denom - 2.0 * Sp * Sm * inv_denom);
  double l2h = fabs((Sp + Sm) * (un_roe - ra.a) * inv_denom - 2.0 * Sp * Sm * inv_denom);
  double l3h = fabs((Sp + Sm) * (un_roe + ra.a) * inv_denom - 2.0 * Sp * Sm * inv_denom);
  double eps = 0.3 * (fabs(un_roe) + ra.a);
  double l1r = fabs(un_roe);
  double l2r = harten_hyman(un_roe - ra.a, eps);
  double l3r = harten_hyman(un_roe + ra.a, eps);
  double l1 = beta * l1r + (1.0 - beta) * l1h;
  double l2 = beta * l2r + (1.0 - beta) * l2h;
  double l3 = beta * l3r + (1.0 - beta) * l3h;
  double FL[4], FR[4];
  fluxes_helper(g, pl, dl, ul, vl, nx, ny, FL);
  fluxes_helper(g, pr, dr, ur, vr, nx, ny, FR);
  for (int k = 0; k < 4; k++) F[k] = 0.5 * (FL[k] + FR[k]);
  double drho = drtot - dltot, dp = pr - pl, du = ur - ul, dv = vr - vl, dun = unr - unl;
  double tx = -ny, ty = nx;
  double alpha_m = (dp - ra.rho * ra.a * dun) / (2.0 * ra.a * ra.a);
  double alpha_c = drho - dp / (ra.a * ra.a);
  double alpha_s = ra.rho * (du * tx + dv * ty);
  double alpha_p = (dp + ra.rho * ra.a * dun) / (2.0 * ra.a * ra.a);
  F[0] = F[0] - 0.5 * (alpha_c * l1 + alpha_m * l2 + alpha_p * l3);
  F[1] = F[1] - 0.5 * (alpha_c * l1 * ra.u + alpha_s * l1 * tx + alpha_m * l2 * (ra.u - ra.a * nx) +
                       alpha_p * l3 * (ra.u + ra.a * nx));
  F[2] = F[2] - 0.5 * (alpha_c * l1 * ra.v + alpha_s * l1 * ty + alpha_m * l2 * (ra.v - ra.a * ny) +
                       alpha_p * l3 * (ra.v + ra.a * ny));
  F[3] = F[3] - 0.5 * (alpha_c * l1 * 0.5 * (ra.u * ra.u + ra.v * ra.v) + alpha_s * l1 * (ra.u * tx + ra.v * ty) +
                       alpha_m * l2 * (ra.h0 - ra.a * un_roe) + alpha_p * l3 * (ra.h0 + ra.a * un_roe));
}

/* Riemann pointer bound by Assign_Riemann_Solver (Mod_Riemann.f90:25-109).
 * Argument list of riemann_if: (dl,ul,vl,pl,al,dltot, dr,ur,vr,pr,ar,drtot, beta, nx,ny). */
static void riemann(int id, const gas_t *g, double dl, double ul, double vl, double pl, double al, double dltot,
                    double dr, double ur, double vr, double pr, double ar, double drtot, double beta,
                    double nx, double ny, double *F) {
  if (id == Q2D_RIEMANN_HLLEPP) {
    riemann_hllepp(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, beta, nx, ny, F);
    return;
  }
  if (id == Q2D_RIEMANN_HLLCP) {                 /* riemann_HLLCSD — Lib_Riemann_HLL.f90:571-603 */
    if (beta == 0.0) {
      riemann_hlle(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, F);
    } else if (beta >= 1.0) {
      riemann_hllc(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, F);
    } else {
      double FE[4], FC[4];
      riemann_hlle(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, FE);
      riemann_hllc(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, FC);
      for (int k = 0; k < 4; k++) F[k] = beta * FC[k] + (1.0 - beta) * FE[k];
    }
    return;
  }
  if (id == Q2D_RIEMANN_HLLE)
    riemann_hlle(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, F);
  else if (id == Q2D_RIEMANN_HLLEM)
    riemann_hllem(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, F);
  else
    riemann_hllc(g, dl, ul, vl, pl, al, dltot, dr, ur, vr, pr, ar, drtot, nx, ny, F);
}

/* ------------------------------------------------------------------------ */
/* block storage in the reference layout                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int ni, nj, ngi, case_id;
  double *P, *PO, *R;                 /* P(4,1-gc:ni+gc,1-gc:nj+gc) */
  double *dl, *M, *sur;               /* ghost extents */
  double *fi, *fj;                    /* dir(1)%f(0:ni,1:nj), dir(2)%f(1:ni,0:nj): {N1,N2,int_l} */
  double *node;                       /* node(0:ni,0:nj)%c(2) (optional) */
  double *sa, *sb;                    /* (1:ni,1:nj) source geometry (optional) */
  double *dtlocal;                    /* (1:ni,1:nj) */
  double *beta;                       /* (1:ni,1:nj) shock detector of the stage (MUSCL-SD cases) */
} oblk;

static inline int cidx(const oblk *b, int i, int j) { return (j + 1) * b->ngi + (i + 1); }
#define PV(b, v, i, j)  ((b)->P[cidx((b), (i), (j)) * NP + (v)])
#define POV(b, v, i, j) ((b)->PO[cidx((b), (i), (j)) * NP + (v)])
#define RV(b, v, i, j)  ((b)->R[cidx((b), (i), (j)) * NP + (v)])
#define DL(b, d, i, j)  ((b)->dl[cidx((b), (i), (j)) * 2 + (d)])
/* M(i,j)%c(r,c) with r,c in 1..2, column major */
#define MM(b, r, c, i, j) ((b)->M[cidx((b), (i), (j)) * 4 + ((r) - 1) + 2 * ((c) - 1)])
#define SUR(b, i, j)    ((b)->sur[cidx((b), (i), (j))])
#define FI(b, k, i, j)  ((b)->fi[(((j) - 1) * ((b)->ni + 1) + (i)) * 3 + (k)])
#define FJ(b, k, i, j)  ((b)->fj[((j) * (b)->ni + ((i) - 1)) * 3 + (k)])
#define IN(b, i, j)     (((j) - 1) * (b)->ni + ((i) - 1))
#define NODE(b, k, i, j) ((b)->node[((j) * ((b)->ni + 1) + (i)) * 2 + (k)])

/* guide(6,3) — Parameters_m.f90:13-18 */
static const int guide[5][2] = { {0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1} };

typedef struct {
  q2d_case_desc d;
  gas_t gas;
  int first_block, nblocks, first_bc, nbc;
  int iter, iter_from_call;
  double time, dtglobal;
  int ndir;
} ocase;

typedef struct {
  q2d_bc_entry e;
  double cos_a, sin_a;                 /* 405 inflow direction (Compute_Inflow_Direction) */
} obc;

struct q2do {
  int nblocks, nbc, ncases;
  oblk *blk;
  obc *bc;
  ocase *cs;
};

static double *dupd(const double *src, size_t n) {
  double *d = (double *)malloc(n * sizeof(double));
  if (!d) return NULL;
  if (src) memcpy(d, src, n * sizeof(double)); else memset(d, 0, n * sizeof(double));
  return d;
}

q2do *q2do_create(const q2d_problem *prob) {
  if (!prob || prob->nblocks <= 0 || prob->ncases <= 0) return NULL;
  q2do *o = (q2do *)calloc(1, sizeof(q2do));
  o->nblocks = prob->nblocks; o->nbc = prob->nbc; o->ncases = prob->ncases;
  o->blk = (oblk *)calloc((size_t)o->nblocks, sizeof(oblk));
  o->bc = (obc *)calloc((size_t)(o->nbc > 0 ? o->nbc : 1), sizeof(obc));
  o->cs = (ocase *)calloc((size_t)o->ncases, sizeof(ocase));
  for (int b = 0; b < o->nblocks; b++) {
    const q2d_block_desc *d = &prob->blocks[b];
    oblk *k = &o->blk[b];
    k->ni = d->ni; k->nj = d->nj; k->ngi = d->ni + 2 * GC; k->case_id = d->case_id;
    size_t ncg = (size_t)(d->ni + 2 * GC) * (size_t)(d->nj + 2 * GC);
    k->P = dupd(d->P, ncg * NP);
    k->PO = dupd(d->P, ncg * NP);
    k->R = dupd(NULL, ncg * NP);
    k->dl = dupd(d->dl, ncg * 2);
    k->M = dupd(d->M, ncg * 4);
    k->sur = dupd(d->sur, ncg);
    k->fi = dupd(d->face_i, (size_t)(d->ni + 1) * d->nj * 3);
    k->fj = dupd(d->face_j, (size_t)d->ni * (d->nj + 1) * 3);
    k->node = d->node ? dupd(d->node, (size_t)(d->ni + 1) * (d->nj + 1) * 2) : NULL;
    k->sa = d->src_a ? dupd(d->src_a, (size_t)d->ni * d->nj) : NULL;
    k->sb = d->src_b ? dupd(d->src_b, (size_t)d->ni * d->nj) : NULL;
    k->dtlocal = dupd(NULL, (size_t)d->ni * d->nj);
    k->beta = dupd(NULL, (size_t)d->ni * d->nj);
  }
  for (int c = 0; c < o->ncases; c++) {
    ocase *cs = &o->cs[c];
    cs->d = prob->cases[c];
    cs->gas = mkgas(cs->d.scheme.gas_R, cs->d.scheme.gas_cp);
    cs->gas.mu_ref = cs->d.scheme.mu_ref; cs->gas.T_ref = cs->d.scheme.T_ref;
    cs->gas.S = cs->d.scheme.S_suth; cs->gas.Pr = cs->d.scheme.Pr;
    if (cs->d.thermo && cs->d.thermo_n >= 2) {         /* table mode: the case keeps its own copy */
      cs->gas.tab = dupd(cs->d.thermo, (size_t)4 * (size_t)cs->d.thermo_n);
      cs->gas.ntab = cs->d.thermo_n;
    }
    cs->d.thermo = NULL;
    cs->first_block = -1; cs->first_bc = -1;
    cs->dtglobal = 0.0; cs->time = cs->d.time0;      /* Wrap_Setup.f90:129-152 */
  }
  for (int b = 0; b < o->nblocks; b++) {
    ocase *cs = &o->cs[o->blk[b].case_id];
    if (cs->first_block < 0) cs->first_block = b;
    cs->nblocks++;
  }
  for (int i = 0; i < o->nbc; i++) {
    o->bc[i].e = prob->bc[i];
    ocase *cs = &o->cs[prob->bc[i].case_id];
    if (cs->first_bc < 0) cs->first_bc = i;
    cs->nbc++;
  }
  for (int c = 0; c < o->ncases; c++) {
    ocase *cs = &o->cs[c];
    /* Check_Mesh_Type (Lib_Metrics.f90:24-37): ndir from the first block */
    cs->ndir = (o->blk[cs->first_block].nj == 1) ? 1 : 2;
  }
  /* Compute_Inflow_Direction (Lib_BC_Fluxes_Inflow.f90:548-560), evaluated once per record. */
  for (int i = 0; i < o->nbc; i++) {
    q2d_bc_entry *e = &o->bc[i].e;
    if (e->type != Q2D_BC_SUPERSONIC) continue;
    oblk *k = &o->blk[o->cs[e->case_id].first_block + e->b - 1];
    double n1, n2;
    if (e->f <= 2) { int fi = e->i - (e->f % 2); n1 = FI(k, 0, fi, e->j); n2 = FI(k, 1, fi, e->j); }
    else { int fj = e->j - (e->f % 2); n1 = FJ(k, 0, e->i, fj); n2 = FJ(k, 1, e->i, fj); }
    double a = (e->alpha >= 0.5 * DBL_MAX) ? atan(n2 / (n1 + 1e-20)) : e->alpha;
    o->bc[i].cos_a = cos(a); o->bc[i].sin_a = sin(a);
  }
  return o;
}

void q2do_destroy(q2do *o) {
  if (!o) return;
  for (int b = 0; b < o->nblocks; b++) {
    oblk *k = &o->blk[b];
    free(k->P); free(k->PO); free(k->R); free(k->dl); free(k->M); free(k->sur);
    free(k->fi); free(k->fj); free(k->node); free(k->sa); free(k->sb); free(k->dtlocal); free(k->beta);
  }
  for (int c = 0; c < o->ncases; c++) free((void *)o->cs[c].gas.tab);
  free(o->blk); free(o->bc); free(o->cs); free(o);
}

/* ------------------------------------------------------------------------ */
/* Compute_dt — Mod_dt.f90:10-104 ; Set_Global_dt :109-128                   */
/* ------------------------------------------------------------------------ */
static void compute_dt(q2do *o, ocase *cs) {
  const q2d_scheme *s = &cs->d.scheme;
  double dtglobal = cs->dtglobal;
  for (int bb = 0; bb < cs->nblocks; bb++) {
    oblk *b = &o->blk[cs->first_block + bb];
    double dtmin_blk = dtglobal;
#ifdef _OPENMP
#pragma omp parallel for reduction(min : dtmin_blk) schedule(static)
#endif
    for (int j = 1; j <= b->nj; j++) {
      for (int i = 1; i <= b->ni; i++) {
        double rho = PV(b, 0, i, j);
        double sound = f_ss(&cs->gas, PV(b, 3, i, j), rho);
        double dtmin = 1e8;
        for (int d = 1; d <= cs->ndir; d++) {
          double m1 = MM(b, d, 1, i, j), m2 = MM(b, d, 2, i, j);
          double nrm = sqrt(m1 * m1 + m2 * m2);
          double v1 = m1 / nrm, v2 = m2 / nrm;
          double vel_ = fabs(PV(b, 1, i, j) * v1 + PV(b, 2, i, j) * v2);
          double lambda = vel_ + sound;
          double dt = DL(b, d - 1, i, j) / lambda * s->cfl;
          if (dt < dtmin) dtmin = dt;
        }
        if (s->model > 0) {                            /* VNN condition :88-102 (laminar: mie = 0) */
          double mil = mu_lam(&cs->gas, eos_T(&cs->gas, PV(b, 3, i, j), rho));
          double mi = 0.0 + mil;
          for (int d = 1; d <= cs->ndir; d++) {
            double dl = DL(b, d - 1, i, j);
            double dt = rho * (dl * dl) * s->vnn / mi;
            if (dt < dtmin) dtmin = dt;
          }
        }
        double dtcell = dtmin;
        if (cs->iter < s->ramp_iter)
          dtcell = g_plant == Q2DO_PLANT_RAMP ? dtcell * (cs->iter - 1) / s->ramp_iter : dtcell * cs->iter / s->ramp_iter;
        b->dtlocal[IN(b, i, j)] = dtcell;
        if (dtcell < dtmin_blk) dtmin_blk = dtcell;
      }
    }
    if (dtmin_blk < dtglobal) dtglobal = dtmin_blk;
  }
  cs->dtglobal = dtglobal;
}

static void set_global_dt(q2do *o, ocase *cs) {
  for (int bb = 0; bb < cs->nblocks; bb++) {
    oblk *b = &o->blk[cs->first_block + bb];
    for (int k = 0; k < b->ni * b->nj; k++) b->dtlocal[k] = cs->dtglobal;
  }
}

/* ------------------------------------------------------------------------ */
/* Fill_Ghost_Cell — Lib_Ghost.f90:16-140                                    */
/* ------------------------------------------------------------------------ */
static void ghost_connection(oblk *bm, int Im, int Jm, int Fm, const oblk *bs, int Is, int Js, int Fs) {
  /* Ghost_Connection :143-160 */
  for (int g = 1; g <= GC; g++) {
    int Ig = Im - guide[Fm][0] * g, Jg = Jm - guide[Fm][1] * g;
    int Id = Is + guide[Fs][0] * (g - 1), Jd = Js + guide[Fs][1] * (g - 1);
    for (int v = 0; v < NP; v++) PV(bm, v, Ig, Jg) = bs->P[cidx(bs, Id, Jd) * NP + v];
  }
}
static void ghost_symmetry(oblk *b, int Im, int Jm, int Fm) {
  /* Ghost_Symmetry :163-196 */
  int Ig = Im - guide[Fm][0], Jg = Jm - guide[Fm][1];
  double n1, n2;
  if (Fm <= 2) { int Ic = Im - (Fm % 2); n1 = FI(b, 0, Ic, Jm); n2 = FI(b, 1, Ic, Jm); }
  else { int Jc = Jm - (Fm % 2); n1 = FJ(b, 0, Im, Jc); n2 = FJ(b, 1, Im, Jc); }
  double u = PV(b, 1, Im, Jm), v = PV(b, 2, Im, Jm);
  double udotn = u * n1 + v * n2;
  u = u - 2.0 * udotn * n1;
  v = v - 2.0 * udotn * n2;
  PV(b, 1, Ig, Jg) = u; PV(b, 2, Ig, Jg) = v;
  PV(b, 0, Ig, Jg) = PV(b, 0, Im, Jm);
  PV(b, 3, Ig, Jg) = PV(b, 3, Im, Jm);
}
static void ghost_zg(oblk *b, int Im, int Jm, int Fm) {
  /* Ghost_ZG_Extrapolate :199-219 */
  for (int g = 1; g <= GC; g++) {
    int Ig = Im - guide[Fm][0] * g, Jg = Jm - guide[Fm][1] * g;
    for (int v = 0; v < NP; v++) PV(b, v, Ig, Jg) = PV(b, v, Im, Jm);
  }
}
static void ghost_extrap(oblk *b, int Im, int Jm, int Fm) {
  /* Ghost_Extrapolate :270-299 (x_g = 3 x_{g-1} - 3 x_{g-2} + x_{g-3}) */
  int di = -guide[Fm][0], dj = -guide[Fm][1];          /* outward step */
  for (int g = 1; g <= GC; g++) {
    int I0 = Im + di * g, J0 = Jm + dj * g;
    int I1 = Im + di * (g - 1), J1 = Jm + dj * (g - 1);
    int I2 = Im + di * (g - 2), J2 = Jm + dj * (g - 2);
    int I3 = Im + di * (g - 3), J3 = Jm + dj * (g - 3);
    for (int v = 0; v < NP; v++)
      PV(b, v, I0, J0) = 3.0 * PV(b, v, I1, J1) - 3.0 * PV(b, v, I2, J2) + PV(b, v, I3, J3);
  }
}

static void fill_ghost(q2do *o, ocase *cs) {
  for (int k = 0; k < cs->nbc; k++) {
    const q2d_bc_entry *e = &o->bc[cs->first_bc + k].e;
    oblk *bm = &o->blk[cs->first_block + e->b - 1];
    switch (e->type) {
      case 101: case 201: {
        const oblk *bs = &o->blk[cs->first_block + e->bs - 1];
        ghost_connection(bm, e->i, e->j, e->f, bs, e->is, e->js, e->fs);
        break;
      }
      case 300: ghost_symmetry(bm, e->i, e->j, e->f); break;
      case 0: case 401: case 402: case 403: case 404: case 405: case 406: case 407: case 420:
        ghost_zg(bm, e->i, e->j, e->f); break;
      default: ghost_extrap(bm, e->i, e->j, e->f); break;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* SD_Tramel — Lib_Shock_Detector.f90:14-38 (DELTA = 20): p at the cell, p1 / p2 its i-1 / i+1 and      */
/* p3 / p4 its j-1 / j+1 neighbours (pmatrix(0,1), (2,1), (1,0), (1,2)).                               */
/* tanh(x) is evaluated as x - x*x*x/3: here x = 10 phi^3 <= 10 (1/400)^3 = 1.6e-7, where the next     */
/* term of the series, 2x^5/15, is below 1e-28 x, so this is tanh rounded to nearest (the device       */
/* evaluates the same expression; the reference's libm tanh agrees to within one rounding of t).       */
/* ------------------------------------------------------------------------ */
static double sd_tramel(double p, double p1, double p2, double p3, double p4) {
  const double DELTA = 20.0;
  double kp1 = fabs((p2 - 2.0 * p + p1) / (p2 + 2.0 * p + p1));
  double kp2 = fabs((p4 - 2.0 * p + p3) / (p4 + 2.0 * p + p3));
  double sw = kp1 > kp2 ? kp1 : kp2;
  if (sw < 1.0 / DELTA) {
    double phi = sw / DELTA;
    if (phi < 0.0) phi = 0.0;
    double x = 10.0 * phi * phi * phi;
    double t = x - x * x * x / 3.0;
    return 1.0 - t;
  }
  return 0.0;
}

/* ------------------------------------------------------------------------ */
/* Convective_Flux — Lib_Convective.f90:15-82                                */
/* ------------------------------------------------------------------------ */
/* dl[-1..2] -> dl[0..3], prim[c][v] for cells -1..2 -> c = 0..3. Returns 0 or Q2D_E_INTERFACE. */
/* beta_: the reconstruction's (the detector with SD_limiter, else 1); beta_rie: the Riemann solver's (the
 * detector whenever it is computed, else 1) — Lib_Convective.f90:42-46, :70-72 */
static int convective_flux(int lim, int rid, const gas_t *g, const double *dl, double nx, double ny,
                           double int_l, const double (*prim)[NP], double beta_, double beta_rie, double *Flux) {
  double l0 = 0.5 * (dl[0] + dl[1]);
  double l1 = 0.5 * (dl[1] + dl[2]);
  double l2 = 0.5 * (dl[2] + dl[3]);
  double lm = 0.5 * dl[1];
  double lp = 0.5 * dl[2];
  double P1[NP], P4[NP];
  for (int v = 0; v < NP; v++)
    reconstruct(lim, prim[0][v], prim[1][v], prim[2][v], prim[3][v], l0, l1, l2, lm, lp, beta_, &P1[v], &P4[v]);
  if (check_gas_state(P1[0], P1[3]) || check_gas_state(P4[0], P4[3])) return Q2D_E_INTERFACE;
  double rotot1 = P1[0], rotot4 = P4[0];
  double a1 = f_ss(g, P1[3], rotot1);
  double a4 = f_ss(g, P4[3], rotot4);
  double F[4];
  riemann(rid, g, P1[0], P1[1], P1[2], P1[3], a1, rotot1, P4[0], P4[1], P4[2], P4[3], a4, rotot4, beta_rie, nx, ny, F);
  double su = copysign(0.5, F[0]);
  double sp = 0.5 + su;
  double sm = su - 0.5;
  Flux[0] = F[0] * int_l * (sp * (P1[0] / rotot1) - sm * (P4[0] / rotot4));
  Flux[1] = F[1] * int_l;
  Flux[2] = F[2] * int_l;
  Flux[3] = F[3] * int_l;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Laminar viscous terms — Lib_Diffusive.f90, Lib_Fluid.f90 (laminar, mie = 0) */
/* ------------------------------------------------------------------------ */
static const double negativetwothird = -2.0 / 3.0;   /* Lib_Fluid.f90:21 */

/* Stress_Vector_Std (:26-48; VG = V_dot_G, 0 without area variation) / Stress_Vector_Axi (:51-76, VG not
 * used). gv[a][b] = Gradvel(a+1,b+1). */
static void stress_vector(int axi, const double gv[2][2], double n1, double n2, double mil, double r_ave, double vr,
                          double VG, double *S1, double *S2) {
  double mi = mil + 0.0;
  double lam = negativetwothird * mi;
  double Divervel;
  if (axi) {
    if (r_ave <= 1.0e-6) Divervel = gv[0][0] + gv[1][1] + gv[1][1];
    else Divervel = gv[0][0] + gv[1][1] + vr / r_ave;
  } else {
    Divervel = gv[0][0] + gv[1][1] - VG;
  }
  double lamDivervel = lam * Divervel;
  double Tau_xx = lamDivervel + 2.0 * mi * gv[0][0];
  double Tau_yy = lamDivervel + 2.0 * mi * gv[1][1];
  double Tau_xy = mi * (gv[0][1] + gv[1][0]);
  *S1 = Tau_xx * n1 + Tau_xy * n2;
  *S2 = Tau_xy * n1 + Tau_yy * n2;
}

/* V_dot_G_face_avg — Lib_Fluid.f90:550-564: <V>_face . <G>_face */
static double v_dot_g_face(double u1, double v1, double u2, double v2, double gx1, double gy1, double gx2, double gy2) {
  double Va0 = 0.5 * (u1 + u2), Va1 = 0.5 * (v1 + v2);
  double Ga0 = 0.5 * (gx1 + gx2), Ga1 = 0.5 * (gy1 + gy2);
  return Va0 * Ga0 + Va1 * Ga1;
}

/* Area-variation geometric term G of cell (i,j) (area_geom_term_x/y); 0 without area variation. */
static void area_g(const oblk *b, const ocase *cs, int i, int j, double *gx, double *gy) {
  if (cs->d.scheme.source == Q2D_SRC_AREA && b->sa) { *gx = b->sa[IN(b, i, j)]; *gy = b->sb[IN(b, i, j)]; }
  else { *gx = 0.0; *gy = 0.0; }
}

/* Gradient(k,:) <- Gradient(k,:) . M   (matmul(Gradient, M), M(l,c) = MM(l,c)) */
static void grad_times_M(double g[2], const double M[4]) {
  double a = g[0] * M[0] + g[1] * M[1];           /* M(1,1), M(2,1) */
  double c = g[0] * M[2] + g[1] * M[3];           /* M(1,2), M(2,2) */
  g[0] = a; g[1] = c;
}

/* Compute_Diffusive_Flux — Lib_Diffusive.f90:113-199, laminar, one species: Pm the face state, Gr the
 * gradient after the metric transform (rows rho-fraction, u, v, T).  Dm*grad(Y) = 0, so the species flux
 * and the enthalpy-diffusion terms vanish. */
static void compute_diffusive_flux(const ocase *cs, int axi, const double Pm[4], const double Gr[4][2], double n1,
                                   double n2, double int_l, double r_ave, double VG, double *Flux) {
  const gas_t *g = &cs->gas;
  double rho = Pm[0];
  double T = Pm[3] / (rho * g->R);
  double mil = mu_lam(g, T), kl = k_lam(g, mil, T);
  double kappa = kl + 0.0;
  double gv[2][2] = {{Gr[1][0], Gr[1][1]}, {Gr[2][0], Gr[2][1]}};
  double S1, S2;
  stress_vector(axi, gv, n1, n2, mil, r_ave, Pm[2], VG, &S1, &S2);
  Flux[0] = 0.0;
  Flux[1] = int_l * S1;
  Flux[2] = int_l * S2;
  Flux[3] = int_l * ((S1 * Pm[1] + S2 * Pm[2]) + kappa * (Gr[3][0] * n1 + Gr[3][1] * n2) + 0.0);
}

/* Diffusive_Flux + Tangential_Gradient + Compute_Diffusive_Flux — Lib_Diffusive.f90:10-199.
 * P1,P2: the two cells of the face; P3..P6 the tangential stencil; a = direction of the face normal (0/1). */
static void diffusive_flux(const ocase *cs, double n1, double n2, double int_l, const double *P1, const double *P2,
                           const double *P3, const double *P4, const double *P5, const double *P6,
                           const double *M1, const double *M2, double r2, double r1, int a, const double G1[2],
                           const double G2[2], double *Flux) {
  const gas_t *g = &cs->gas;
  int axi = cs->d.scheme.source == Q2D_SRC_AXI;
  int b = 1 - a;
  double r_ave = 0.5 * (r1 + r2);
  double rho1 = P1[0], T1 = P1[3] / (rho1 * g->R);
  double rho2 = P2[0], T2 = P2[3] / (rho2 * g->R);
  double Gr[4][2];                                   /* Gradient(nprim,2) */
  Gr[0][a] = P2[0] / rho2 - P1[0] / rho1;
  Gr[3][a] = T2 - T1;
  Gr[1][a] = P2[1] - P1[1];
  Gr[2][a] = P2[2] - P1[2];
  {                                                  /* Tangential_Gradient :80-110 */
    double r3 = P3[0], r4 = P4[0], r5 = P5[0], r6 = P6[0];
    double T3 = P3[3] / (r3 * g->R), T4 = P4[3] / (r4 * g->R), T5 = P5[3] / (r5 * g->R), T6 = P6[3] / (r6 * g->R);
    Gr[0][b] = (P4[0] / r4 - P3[0] / r3 + P6[0] / r6 - P5[0] / r5) * 0.25;
    Gr[3][b] = (T4 - T3 + T6 - T5) * 0.25;
    Gr[1][b] = (P4[1] - P3[1] + P6[1] - P5[1]) * 0.25;
    Gr[2][b] = (P4[2] - P3[2] + P6[2] - P5[2]) * 0.25;
  }
  double M[4];
  for (int k = 0; k < 4; k++) M[k] = 0.5 * (M1[k] + M2[k]);
  for (int k = 0; k < 4; k++) grad_times_M(Gr[k], M);
  double Pm[4];
  for (int k = 0; k < 4; k++) Pm[k] = 0.5 * (P1[k] + P2[k]);
  /* V.G at the face (Definition A; :61-68) */
  double VG = v_dot_g_face(P1[1], P1[2], P2[1], P2[2], G1[0], G1[1], G2[0], G2[1]);
  compute_diffusive_flux(cs, axi, Pm, (const double(*)[2])Gr, n1, n2, int_l, r_ave, VG, Flux);
}

/* Fluxes_blk_omp_do — Mod_Fluxes.f90:97-234 (convective + laminar diffusive) */
static int fluxes_block(oblk *b, const ocase *cs) {
  const q2d_scheme *s = &cs->d.scheme;
  int err = 0;
  for (int j = 1; j <= b->nj; j++)
    for (int i = 1; i <= b->ni; i++)
      for (int v = 0; v < NP; v++) RV(b, v, i, j) = 0.0;
  /* shock detector on Prim(np, i-1:i+1, j-1:j+1) (Mod_Fluxes.f90:129-138); the faces take the value of
   * their cell i (i-faces, :144-150) or j (j-faces, :191-197), with SD_limiter (Lib_Convective.f90:42-46) */
  const int sd = s->space_sd != 0;
  const int det = sd || s->riemann >= Q2D_RIEMANN_HLLEPP;   /* obj_shock_detector%SD, Assign_Setup.f90:49-53 */
  if (det)
    for (int j = 1; j <= b->nj; j++)
      for (int i = 1; i <= b->ni; i++)
        b->beta[IN(b, i, j)] = sd_tramel(PV(b, 3, i, j), PV(b, 3, i - 1, j), PV(b, 3, i + 1, j), PV(b, 3, i, j - 1),
                                         PV(b, 3, i, j + 1));
  /* i-faces: rows in parallel, each face once (:140-185) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
  for (int j = 1; j <= b->nj; j++) {
    for (int i = 1; i <= b->ni - 1; i++) {
      double dl[4], prim[4][NP], F[NP];
      for (int c = 0; c < 4; c++) {
        dl[c] = DL(b, 0, i - 1 + c, j);
        for (int v = 0; v < NP; v++) prim[c][v] = PV(b, v, i - 1 + c, j);
      }
      err |= -convective_flux(s->limiter, s->riemann, &cs->gas, dl, FI(b, 0, i, j), FI(b, 1, i, j),
                              FI(b, 2, i, j), (const double(*)[NP])prim, sd ? b->beta[IN(b, i, j)] : 1.0,
                              det ? b->beta[IN(b, i, j)] : 1.0, F);
      for (int v = 0; v < NP; v++) { RV(b, v, i, j) = RV(b, v, i, j) + F[v]; RV(b, v, i + 1, j) = RV(b, v, i + 1, j) - F[v]; }
      if (s->model > 0) {
        double D[NP];
        double G1[2], G2[2];
        area_g(b, cs, i, j, &G1[0], &G1[1]);
        area_g(b, cs, i + 1, j, &G2[0], &G2[1]);
        diffusive_flux(cs, FI(b, 0, i, j), FI(b, 1, i, j), FI(b, 2, i, j), &PV(b, 0, i, j), &PV(b, 0, i + 1, j),
                       &PV(b, 0, i, j - 1), &PV(b, 0, i, j + 1), &PV(b, 0, i + 1, j - 1), &PV(b, 0, i + 1, j + 1),
                       &MM(b, 1, 1, i, j), &MM(b, 1, 1, i + 1, j), NODE(b, 1, i, j), NODE(b, 1, i, j - 1), 0, G1, G2, D);
        for (int v = 0; v < NP; v++) { RV(b, v, i, j) = RV(b, v, i, j) - D[v]; RV(b, v, i + 1, j) = RV(b, v, i + 1, j) + D[v]; }
      }
    }
  }
  /* j-faces: columns in parallel (:187-232) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
  for (int i = 1; i <= b->ni; i++) {
    for (int j = 1; j <= b->nj - 1; j++) {
      double dl[4], prim[4][NP], F[NP];
      for (int c = 0; c < 4; c++) {
        dl[c] = DL(b, 1, i, j - 1 + c);
        for (int v = 0; v < NP; v++) prim[c][v] = PV(b, v, i, j - 1 + c);
      }
      err |= -convective_flux(s->limiter, s->riemann, &cs->gas, dl, FJ(b, 0, i, j), FJ(b, 1, i, j),
                              FJ(b, 2, i, j), (const double(*)[NP])prim, sd ? b->beta[IN(b, i, j)] : 1.0,
                              det ? b->beta[IN(b, i, j)] : 1.0, F);
      for (int v = 0; v < NP; v++) { RV(b, v, i, j) = RV(b, v, i, j) + F[v]; RV(b, v, i, j + 1) = RV(b, v, i, j + 1) - F[v]; }
      if (s->model > 0) {
        double D[NP];
        double G1[2], G2[2];
        area_g(b, cs, i, j, &G1[0], &G1[1]);
        area_g(b, cs, i, j + 1, &G2[0], &G2[1]);
        diffusive_flux(cs, FJ(b, 0, i, j), FJ(b, 1, i, j), FJ(b, 2, i, j), &PV(b, 0, i, j), &PV(b, 0, i, j + 1),
                       &PV(b, 0, i - 1, j), &PV(b, 0, i + 1, j), &PV(b, 0, i - 1, j + 1), &PV(b, 0, i + 1, j + 1),
                       &MM(b, 1, 1, i, j), &MM(b, 1, 1, i, j + 1), NODE(b, 1, i, j), NODE(b, 1, i - 1, j), 1, G1, G2, D);
        for (int v = 0; v < NP; v++) { RV(b, v, i, j) = RV(b, v, i, j) - D[v]; RV(b, v, i, j + 1) = RV(b, v, i, j + 1) + D[v]; }
      }
    }
  }
  return err ? Q2D_E_INTERFACE : 0;
}

/* ------------------------------------------------------------------------ */
/* BC fluxes — Mod_BC_Fluxes.f90:49-271                                      */
/* ------------------------------------------------------------------------ */
/* Face_Index — Lib_BC_Fluxes.f90:34-50 */
static void face_geom(const oblk *b, int Fm, int Im, int Jm, int *dir, double *n1, double *n2, double *il) {
  if (Fm <= 2) {
    int fi = Im - (Fm % 2);
    *dir = 1; *n1 = FI(b, 0, fi, Jm); *n2 = FI(b, 1, fi, Jm); *il = FI(b, 2, fi, Jm);
  } else {
    int fj = Jm - (Fm % 2);
    *dir = 2; *n1 = FJ(b, 0, Im, fj); *n2 = FJ(b, 1, Im, fj); *il = FJ(b, 2, Im, fj);
  }
}

/* BC_Connection_Eul — Lib_BC_Fluxes_Connection.f90:24-133 */
static int bc_connection_eul(oblk *b, const ocase *cs, int Im, int Jm, int Fm) {
  const q2d_scheme *s = &cs->d.scheme;
  int modfm = Fm % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  int g1 = guide[Fm][0], g2 = guide[Fm][1];
  int I1 = modfm * (Im - g1 * GC) + modfm1 * (Im + g1);
  int J1 = modfm * (Jm - g2 * GC) + modfm1 * (Jm + g2);
  int I2 = modfm1 * Im + modfm * (Im - g1);
  int J2 = modfm1 * Jm + modfm * (Jm - g2);
  int I3 = modfm1 * (Im - g1) + modfm * Im;
  int J3 = modfm1 * (Jm - g2) + modfm * Jm;
  int I4 = modfm1 * (Im - g1 * GC) + modfm * (Im + g1);
  int J4 = modfm1 * (Jm - g2 * GC) + modfm * (Jm + g2);
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  double Dl1 = DL(b, dir - 1, I1, J1) * 0.5;
  double Dl2 = DL(b, dir - 1, I2, J2) * 0.5;
  double Dl3 = DL(b, dir - 1, I3, J3) * 0.5;
  double Dl4 = DL(b, dir - 1, I4, J4) * 0.5;
  double Dist23 = Dl2 + Dl3;
  double Dist12 = Dl1 + Dl2;
  double Dist34 = Dl3 + Dl4;
  double PL[NP], PR[NP];
  /* SD_limiter: the boundary cell's detector value of this stage (Lib_BC_Fluxes_Connection.f90:64-69) */
  double beta_ = s->space_sd ? b->beta[IN(b, Im, Jm)] : 1.0;
  for (int v = 0; v < NP; v++)
    reconstruct(s->limiter, PV(b, v, I1, J1), PV(b, v, I2, J2), PV(b, v, I3, J3), PV(b, v, I4, J4),
                Dist12, Dist23, Dist34, Dl2, Dl3, beta_, &PL[v], &PR[v]);
  if (check_gas_state(PL[0], PL[3]) || check_gas_state(PR[0], PR[3])) return Q2D_E_INTERFACE;
  double rho_L = PL[0], rho_R = PR[0];
  double a_L = f_ss(&cs->gas, PL[3], rho_L);
  double a_R = f_ss(&cs->gas, PR[3], rho_R);
  double F[4];
  double beta_rie = s->riemann >= Q2D_RIEMANN_HLLEPP ? b->beta[IN(b, Im, Jm)] : 1.0;   /* SD_riemann :103-108 */
  riemann(s->riemann, &cs->gas, PL[0], PL[1], PL[2], PL[3], a_L, rho_L, PR[0], PR[1], PR[2], PR[3], a_R, rho_R, beta_rie, n1, n2, F);
  double su = copysign(0.5, F[0]);
  double sel_L = 0.5 + su, sel_R = su - 0.5;
  double Fmass = F[0] * int_l;
  double Flux[NP];
  Flux[0] = Fmass * (sel_L * PL[0] / rho_L - sel_R * PR[0] / rho_R);
  Flux[1] = F[1] * int_l;
  Flux[2] = F[2] * int_l;
  Flux[3] = F[3] * int_l;
  for (int v = 0; v < NP; v++) RV(b, v, Im, Jm) = RV(b, v, Im, Jm) + Flux[v] * (double)modfm2;
  return 0;
}

/* BC_Symmetry_Eul — Lib_BC_Fluxes_Symmetry.f90:28-124 (Un_wall = 0) */
static void bc_symmetry_eul(oblk *b, const ocase *cs, int Im, int Jm, int Fm) {
  const q2d_scheme *s = &cs->d.scheme;
  int modfm = Fm % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  int Int_i = Im + guide[Fm][0], Int_j = Jm + guide[Fm][1];
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  double Bound_Dl = 0.5 * DL(b, dir - 1, Im, Jm);
  double Int_Dl = 0.5 * DL(b, dir - 1, Int_i, Int_j);
  double Bp[NP], Ip[NP], Gp[NP], P1[NP], P3[NP], Slope[NP], Fp[NP];
  for (int v = 0; v < NP; v++) { Bp[v] = PV(b, v, Im, Jm); Ip[v] = PV(b, v, Int_i, Int_j); }
  double Un = Bp[1] * n1 + Bp[2] * n2;
  for (int v = 0; v < NP; v++) Gp[v] = Bp[v];
  Gp[1] = Gp[1] - 2.0 * Un * n1;
  Gp[2] = Gp[2] - 2.0 * Un * n2;
  double dm = (double)modfm, dm1 = (double)modfm1, dm2 = (double)modfm2;
  for (int v = 0; v < NP; v++) P1[v] = dm * Gp[v] + dm1 * Ip[v];
  double Dist12 = 2.0 * Bound_Dl * dm + (Bound_Dl + Int_Dl) * dm1;
  for (int v = 0; v < NP; v++) P3[v] = dm * Ip[v] + dm1 * Gp[v];
  double Dist23 = (Bound_Dl + Int_Dl) * dm + 2.0 * Bound_Dl * dm1;
  for (int v = 0; v < NP; v++) {
    double s32 = (P3[v] - Bp[v]) / Dist23;
    double s21 = (Bp[v] - P1[v]) / Dist12;
    Slope[v] = limiter(s->limiter, s32, s21);
  }
  double Dist = Bound_Dl * dm2;
  for (int v = 0; v < NP; v++) Fp[v] = Bp[v] + Slope[v] * Dist;
  double rotot = Fp[0];
  double sound = f_ss(&cs->gas, Fp[3], rotot);
  Un = Fp[1] * n1 + Fp[2] * n2;
  double vt1 = Fp[1] - Un * n1, vt2 = Fp[2] - Un * n2;
  double Un1 = Un * dm2;
  double U1x = Un1 * n1 + vt1, U1y = Un1 * n2 + vt2;
  double U4x = -Un1 * n1 + vt1, U4y = -Un1 * n2 + vt2;
  double F[4];
  riemann(s->riemann, &cs->gas, Fp[0], U1x, U1y, Fp[3], sound, rotot, Fp[0], U4x, U4y, Fp[3], sound, rotot, 1.0, n1, n2, F);
  double fu = F[1] * int_l, fv = F[2] * int_l;
  RV(b, 1, Im, Jm) = RV(b, 1, Im, Jm) + dm2 * fu;
  RV(b, 2, Im, Jm) = RV(b, 2, Im, Jm) + dm2 * fv;
}

/* BC_Extrapolation — Lib_BC_Fluxes_Extrapolation.f90:18-49 */
static void bc_extrapolation(oblk *b, const ocase *cs, int Im, int Jm, int Fm) {
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  double rho = PV(b, 0, Im, Jm), u = PV(b, 1, Im, Jm), v = PV(b, 2, Im, Jm), p = PV(b, 3, Im, Jm);
  double Un = u * n1 + v * n2;
  double Flux[NP];
  Flux[0] = rho * Un * int_l;
  double Fmass = Flux[0];
  Flux[1] = Fmass * u + p * int_l * n1;
  Flux[2] = Fmass * v + p * int_l * n2;
  Flux[3] = Fmass * f_H0(&cs->gas, p, rho, sqrt(u * u + v * v));
  double sgn = (double)(1 - 2 * (Fm % 2));
  for (int k = 0; k < NP; k++) RV(b, k, Im, Jm) = RV(b, k, Im, Jm) + sgn * Flux[k];
}

/* Supersonic_Inflow — Lib_BC_Fluxes_Inflow.f90:318-403: the flux of the boundary face from the boundary
 * cell P2, its interior neighbour Ip and the record's inflow values (Mach, T, p, relaxation, direction).
 * Returns 1 when the flow leaves through the face (BC_Sign * Un >= 0, :296-299): the caller then runs
 * BC_Symmetry_Eul (Mod_BC_Fluxes.f90:221). */
static int inflow_flux(const gas_t *g, int lim, int rid, int Fm, const double *P2, const double *Ip, double n1,
                       double n2, double int_l, double mach, double T_in, double p_in, double rel_fac,
                       double cos_a, double sin_a, double *Flux) {
  int modfm = Fm % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  double BC_Sign = (double)modfm2;
  double Un = P2[1] * n1 + P2[2] * n2;
  if (BC_Sign * Un >= 0.0) return 1;
  double flow_Sound = f_ss(g, P2[3], P2[0]);
  double flow_Mach = sqrt(P2[1] * P2[1] + P2[2] * P2[2]) / flow_Sound;
  double Mach_l = rel_fac * mach + (1.0 - rel_fac) * flow_Mach;
  double p0_l = rel_fac * p_in + (1.0 - rel_fac) * P2[3];
  double Sup[NP];
  Sup[3] = p0_l;
  double Sup_T = T_in;
  double Sup_Rgas = 1.0 * g->R;
  double Sup_rho = g_plant == Q2DO_PLANT_INFLOW_RHO ? Sup[3] / Sup_T : Sup[3] / (Sup_Rgas * Sup_T);
  Sup[0] = Sup_rho * 1.0;
  double Sup_Sound = f_ss(g, Sup[3], Sup_rho);
  if (g_plant == Q2DO_PLANT_INFLOW_ANGLE) { double t = cos_a; cos_a = sin_a; sin_a = t; }
  Sup[1] = Mach_l * Sup_Sound * cos_a;
  Sup[2] = Mach_l * Sup_Sound * sin_a;
  double dm = (double)modfm, dm1 = (double)modfm1, dm2 = (double)modfm2;
  double P1[NP], P3[NP], Face[NP], PL[NP], PR[NP];
  for (int v = 0; v < NP; v++) { P1[v] = dm * Sup[v] + dm1 * Ip[v]; P3[v] = dm * Ip[v] + dm1 * Sup[v]; }
  for (int v = 0; v < NP; v++) {
    double d32 = P3[v] - P2[v], d21 = P2[v] - P1[v];
    double sl = limiter(lim, d32, d21);
    Face[v] = P2[v] + 0.5 * sl * dm2;
  }
  for (int v = 0; v < NP; v++) { PL[v] = dm * Sup[v] + dm1 * Face[v]; PR[v] = dm * Face[v] + dm1 * Sup[v]; }
  double rho_L = PL[0], rho_R = PR[0];
  double a_L = f_ss(g, PL[3], rho_L), a_R = f_ss(g, PR[3], rho_R);
  double F[4];
  riemann(rid, g, PL[0], PL[1], PL[2], PL[3], a_L, rho_L, PR[0], PR[1], PR[2], PR[3], a_R, rho_R, 1.0, n1, n2, F);
  double su = copysign(1.0, F[0]);
  double Sel_L = (1.0 + su) / 2.0, Sel_R = (su - 1.0) / 2.0;
  double Fmass = F[0] * int_l;
  Flux[0] = Fmass * (Sel_L * PL[0] / rho_L - Sel_R * PR[0] / rho_R);
  Flux[1] = F[1] * int_l;
  Flux[2] = F[2] * int_l;
  Flux[3] = F[3] * int_l;
  return 0;
}

/* BC_Inlet_Supersonic_Static — Lib_BC_Fluxes_Inflow.f90:272-316 */
static void bc_supersonic_inflow(oblk *b, const ocase *cs, const obc *rec, int Im, int Jm, int Fm) {
  const q2d_scheme *s = &cs->d.scheme;
  const q2d_bc_entry *e = &rec->e;
  int Int_i = Im + guide[Fm][0], Int_j = Jm + guide[Fm][1];
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  double P2[NP], Ip[NP], Flux[NP];
  for (int v = 0; v < NP; v++) { P2[v] = PV(b, v, Im, Jm); Ip[v] = PV(b, v, Int_i, Int_j); }
  if (inflow_flux(&cs->gas, s->limiter, s->riemann, Fm, P2, Ip, n1, n2, int_l, e->mach, e->T, e->p, e->rel_fac,
                  rec->cos_a, rec->sin_a, Flux)) {
    bc_symmetry_eul(b, cs, Im, Jm, Fm);
    return;
  }
  int modfm2 = 1 - 2 * (Fm % 2);
  for (int v = 0; v < NP; v++) RV(b, v, Im, Jm) = RV(b, v, Im, Jm) + (double)modfm2 * Flux[v];
}

/* Outflow — Lib_BC_Fluxes_Outflow.f90:89-165: the flux of the boundary face from the boundary cell Bp.
 * Returns 1 when the flow enters through the face (BC_Sign * Un < 0, :68-72): the caller then runs
 * BC_Symmetry_Eul (Mod_BC_Fluxes.f90:227).  `**` is exact_pow. */
static int outflow_flux(const gas_t *g, int Fm, const double *Bp, double n1, double n2, double int_l,
                        double BC_pexit, double BC_rel_fac, double *Flux) {
  int modfm2 = 1 - 2 * (Fm % 2);
  double BC_Sign = (double)modfm2;
  double Un = Bp[1] * n1 + Bp[2] * n2;
  if (BC_Sign * Un < 0.0) return 1;
  double Bound_rho = Bp[0], Bound_Rgas = g->R;
  double Bound_Sound = f_ss(g, Bp[3], Bound_rho);
  double Bound_Gamma = f_gamma_T(g, eos_T(g, Bp[3], Bound_rho));
  double Riem = g_plant == Q2DO_PLANT_OUTFLOW_RIEM ? Bound_Sound - BC_Sign * 0.5 * (Bound_Gamma - 1.0) * Un
                                                   : Bound_Sound + BC_Sign * 0.5 * (Bound_Gamma - 1.0) * Un;
  double VN1 = Un * n1, VN2 = Un * n2;
  double VT1 = Bp[1] - VN1, VT2 = Bp[2] - VN2;
  double Mach_N = fabs(Un / Bound_Sound);
  double Fp[NP], Face_Un, Face_rho, Face_Rgas;
  if (Mach_N < 1.0 && BC_pexit > 0.0) {
    double p_exit = BC_pexit;
    double delta = 0.5 * (Bound_Gamma - 1.0);
    double ex = Bound_Gamma / (Bound_Gamma - 1.0);
    double den = 1.0 / exact_pow(1.0 + delta, ex);
    double pstar = Bp[3] * den * exact_pow(1.0 + delta * (Mach_N * Mach_N), ex);
    if (pstar >= p_exit) p_exit = pstar;
    double Isentropic = g_plant == Q2DO_PLANT_OUTFLOW_ISEN ? (Bound_Gamma - 1.0) / Bound_Gamma
                                                           : (Bound_Gamma - 1.0) / (2.0 * Bound_Gamma);
    Fp[3] = p_exit * BC_rel_fac + Bp[3] * (1.0 - BC_rel_fac);
    double Face_Sound = Bound_Sound * exact_pow(Fp[3] / Bp[3], Isentropic);
    Face_Un = (Riem - Face_Sound) / (BC_Sign * 0.5 * (Bound_Gamma - 1.0));
    Face_rho = Bound_Gamma * Fp[3] / (Face_Sound * Face_Sound);
    Fp[0] = Bp[0] / Bound_rho * Face_rho;
    Fp[1] = VT1 + Face_Un * n1;
    Fp[2] = VT2 + Face_Un * n2;
    Face_Rgas = Bound_Rgas;
  } else {
    for (int v = 0; v < NP; v++) Fp[v] = Bp[v];
    Face_Un = Un; Face_rho = Bound_rho; Face_Rgas = Bound_Rgas;
  }
  double Face_T = Fp[3] / (Face_Rgas * Face_rho);
  double Face_Enthalpy = 0.0 + Fp[0] / Face_rho * (g->tab ? tabT(g, TAB_H, Face_T) : g->cp * Face_T);
  Flux[0] = Fp[0] * int_l * Face_Un;
  double mdot = Face_rho * int_l * Face_Un;
  Flux[1] = Fp[1] * mdot + Fp[3] * int_l * n1;
  Flux[2] = Fp[2] * mdot + Fp[3] * int_l * n2;
  Flux[3] = mdot * (0.5 * (Fp[1] * Fp[1] + Fp[2] * Fp[2]) + Face_Enthalpy);
  return 0;
}

/* BC_Outflow — Lib_BC_Fluxes_Outflow.f90:36-86 */
static void bc_outflow(oblk *b, const ocase *cs, const obc *rec, int Im, int Jm, int Fm) {
  const q2d_bc_entry *e = &rec->e;
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  double Bp[NP], Flux[NP];
  for (int v = 0; v < NP; v++) Bp[v] = PV(b, v, Im, Jm);
  if (outflow_flux(&cs->gas, Fm, Bp, n1, n2, int_l, e->p, e->rel_fac, Flux)) { bc_symmetry_eul(b, cs, Im, Jm, Fm); return; }
  int modfm2 = 1 - 2 * (Fm % 2);
  for (int v = 0; v < NP; v++) RV(b, v, Im, Jm) = RV(b, v, Im, Jm) + (double)modfm2 * Flux[v];
}

/* r_ave of a boundary face from the nodal radii (Lib_BC_Fluxes_Symmetry.f90:183-192) */
static double face_r_ave(const oblk *b, int Im, int Jm, int Fm) {
  switch (Fm) {
    case 1: return 0.5 * (NODE(b, 1, Im - 1, Jm) + NODE(b, 1, Im - 1, Jm - 1));
    case 2: return 0.5 * (NODE(b, 1, Im, Jm) + NODE(b, 1, Im, Jm - 1));
    case 3: return 0.5 * (NODE(b, 1, Im, Jm - 1) + NODE(b, 1, Im - 1, Jm - 1));
    default: return 0.5 * (NODE(b, 1, Im, Jm) + NODE(b, 1, Im - 1, Jm));
  }
}

/* BC_Connection_Visc — Lib_BC_Fluxes_Connection.f90:135-265 (laminar, one species).  The ghost stencil:
 * Pg(:,1), the donor cell the first ghost layer copies (Ghost_Connection, Lib_Ghost.f90:143-160), and
 * Pg(:,3:4), the donor's neighbours along the face, d11 giving the orientation (Fill_BC_Ghost_Connection,
 * :302-326), read after every ghost cell is filled. */
static void bc_connection_visc(q2do *o, oblk *b, const ocase *cs, const q2d_bc_entry *e) {
  const gas_t *g = &cs->gas;
  int axi = cs->d.scheme.source == Q2D_SRC_AXI;
  int Im = e->i, Jm = e->j, Fm = e->f;
  int modfm2 = 1 - 2 * (Fm % 2);
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  int Ig = Im - guide[Fm][0], Jg = Jm - guide[Fm][1];
  double M[4];
  for (int k = 0; k < 4; k++) M[k] = 0.5 * ((&MM(b, 1, 1, Im, Jm))[k] + (&MM(b, 1, 1, Ig, Jg))[k]);
  const oblk *bs = &o->blk[cs->first_block + e->bs - 1];
  int Is = e->is, Js = e->js, i1, j1, i2, j2;
  if (e->fs <= 2) { i1 = Is; j1 = Js - e->d11; i2 = Is; j2 = Js + e->d11; }
  else { i1 = Is - e->d11; j1 = Js; i2 = Is + e->d11; j2 = Js; }
  const double *Pl = &PV(b, 0, Im, Jm), *Pgh = &PV(bs, 0, Is, Js);
  /* Visc_Variables: (rho_s / rho, u, v, T) */
  double Vl[4], Vg[4], Vip[4], Vim[4], Vjp[4], Vjm[4], Vg3[4], Vg4[4];
  const double *src[8] = {Pl, Pgh, &PV(b, 0, Im + 1, Jm), &PV(b, 0, Im - 1, Jm), &PV(b, 0, Im, Jm + 1),
                          &PV(b, 0, Im, Jm - 1), &PV(bs, 0, i1, j1), &PV(bs, 0, i2, j2)};
  double *dst[8] = {Vl, Vg, Vip, Vim, Vjp, Vjm, Vg3, Vg4};
  for (int k = 0; k < 8; k++) {
    const double *P = src[k];
    dst[k][0] = P[0] / P[0]; dst[k][1] = P[1]; dst[k][2] = P[2]; dst[k][3] = P[3] / (P[0] * g->R);
  }
  double Gl[4][2], Gg[4][2];
  double m2 = (double)modfm2, mm2 = (double)(-modfm2);
  for (int v = 0; v < 4; v++) {
    if (Fm <= 2) {
      Gl[v][0] = (Vg[v] - Vl[v]) * m2;
      Gl[v][1] = (Vjp[v] - Vjm[v]) * 0.5;
      Gg[v][0] = (Vl[v] - Vg[v]) * mm2;
      Gg[v][1] = (Vg4[v] - Vg3[v]) * 0.5;
    } else {
      Gl[v][0] = (Vip[v] - Vim[v]) * 0.5;
      Gl[v][1] = (Vg[v] - Vl[v]) * m2;
      Gg[v][0] = (Vg4[v] - Vg3[v]) * 0.5;
      Gg[v][1] = (Vl[v] - Vg[v]) * mm2;
    }
  }
  for (int v = 0; v < 4; v++) { grad_times_M(Gl[v], M); grad_times_M(Gg[v], M); }
  double Pm[4];
  for (int v = 0; v < 4; v++) Pm[v] = (Pl[v] + Pgh[v]) * 0.5;
  double r_ave = face_r_ave(b, Im, Jm, Fm);
  double Gloc[2];
  area_g(b, cs, Im, Jm, &Gloc[0], &Gloc[1]);
  double VG = v_dot_g_face(Pl[1], Pl[2], Pgh[1], Pgh[2], Gloc[0], Gloc[1], Gloc[0], Gloc[1]);
  double Fl[NP], Fg[NP];
  compute_diffusive_flux(cs, axi, Pm, (const double(*)[2])Gl, n1, n2, int_l, r_ave, VG, Fl);
  compute_diffusive_flux(cs, axi, Pm, (const double(*)[2])Gg, n1, n2, int_l, r_ave, VG, Fg);
  for (int v = 1; v < NP; v++) RV(b, v, Im, Jm) = RV(b, v, Im, Jm) - 0.5 * m2 * (Fl[v] + Fg[v]);
}

/* BC_Symmetry_Visc — Lib_BC_Fluxes_Symmetry.f90:127-236 (laminar) */
static void bc_symmetry_visc(oblk *b, const ocase *cs, int Im, int Jm, int Fm) {
  const gas_t *g = &cs->gas;
  int axi = cs->d.scheme.source == Q2D_SRC_AXI;
  int modfm2 = 1 - 2 * (Fm % 2);
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  const double *M = &MM(b, 1, 1, Im, Jm);
  double rho = PV(b, 0, Im, Jm);
  double mil = mu_lam(g, eos_T(g, PV(b, 3, Im, Jm), rho));
  double V1[2], V3[2], V4[2];
  int a;
  V1[0] = PV(b, 1, Im, Jm); V1[1] = PV(b, 2, Im, Jm);
  if (Fm <= 2) {
    V3[0] = PV(b, 1, Im, Jm - 1); V3[1] = PV(b, 2, Im, Jm - 1);
    V4[0] = PV(b, 1, Im, Jm + 1); V4[1] = PV(b, 2, Im, Jm + 1);
    a = 0;
  } else {
    V3[0] = PV(b, 1, Im - 1, Jm); V3[1] = PV(b, 2, Im - 1, Jm);
    V4[0] = PV(b, 1, Im + 1, Jm); V4[1] = PV(b, 2, Im + 1, Jm);
    a = 1;
  }
  int bb = 1 - a;
  double d1 = V1[0] * n1 + V1[1] * n2, d3 = V3[0] * n1 + V3[1] * n2, d4 = V4[0] * n1 + V4[1] * n2;
  double V2[2] = {V1[0] - 2.0 * d1 * n1, V1[1] - 2.0 * d1 * n2};
  double V5[2] = {V3[0] - 2.0 * d3 * n1, V3[1] - 2.0 * d3 * n2};
  double V6[2] = {V4[0] - 2.0 * d4 * n1, V4[1] - 2.0 * d4 * n2};
  double Gr[2][2];
  for (int k = 0; k < 2; k++) {
    Gr[k][a] = (V2[k] - V1[k]) * (double)modfm2;
    Gr[k][bb] = (V4[k] - V3[k] + V6[k] - V5[k]) * 0.25;
  }
  grad_times_M(Gr[0], M);
  grad_times_M(Gr[1], M);
  double r_ave = face_r_ave(b, Im, Jm, Fm);
  double Vm[2] = {0.5 * (V1[0] + V2[0]), 0.5 * (V1[1] + V2[1])};
  /* V.G at the symmetry face: <V>.G_loc (:196-200) */
  double Gl[2];
  area_g(b, cs, Im, Jm, &Gl[0], &Gl[1]);
  double VG = v_dot_g_face(V1[0], V1[1], V2[0], V2[1], Gl[0], Gl[1], Gl[0], Gl[1]);
  double S1, S2;
  stress_vector(axi, (const double(*)[2])Gr, n1, n2, mil, r_ave, 0.5 * (PV(b, 2, Im, Jm) + V2[1]), VG, &S1, &S2);
  double F1 = int_l * S1, F2 = int_l * S2;
  double F3 = int_l * (Vm[0] * S1 + Vm[1] * S2);
  RV(b, 1, Im, Jm) = RV(b, 1, Im, Jm) - (double)modfm2 * F1;
  RV(b, 2, Im, Jm) = RV(b, 2, Im, Jm) - (double)modfm2 * F2;
  RV(b, 3, Im, Jm) = RV(b, 3, Im, Jm) - (double)modfm2 * F3;
}

/* BC_Wall_Temperature — Lib_BC_Fluxes_Wall_Temperature.f90:20-106 (laminar, no frame rotation) */
static void bc_wall_temperature(oblk *b, const ocase *cs, int Im, int Jm, int Fm, double T_wall) {
  const gas_t *g = &cs->gas;
  int axi = cs->d.scheme.source == Q2D_SRC_AXI;
  int modfm = Fm % 2, modfm2 = 1 - 2 * modfm, modfm3 = 2 * (2 * modfm - 1);
  int dir; double n1, n2, int_l;
  face_geom(b, Fm, Im, Jm, &dir, &n1, &n2, &int_l);
  const double *M = &MM(b, 1, 1, Im, Jm);
  double rho = PV(b, 0, Im, Jm), p = PV(b, 3, Im, Jm);
  double T = p / (rho * g->R);
  double mil = mu_lam(g, T_wall), kl = k_lam(g, mil, T_wall);
  double Gr[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  int d = dir - 1;
  Gr[1][d] = PV(b, 1, Im, Jm) * (double)modfm3;
  Gr[2][d] = PV(b, 2, Im, Jm) * (double)modfm3;
  Gr[3][d] = (T - T_wall) * (double)modfm3;
  for (int k = 0; k < 4; k++) grad_times_M(Gr[k], M);
  double r_ave = face_r_ave(b, Im, Jm, Fm);
  double gv[2][2] = {{Gr[1][0], Gr[1][1]}, {Gr[2][0], Gr[2][1]}};
  double S1, S2;
  stress_vector(axi, gv, n1, n2, mil, r_ave, 0.0, 0.0, &S1, &S2);   /* the wall flux passes no V_dot_G */
  double Flux[4];
  Flux[0] = 0.0;
  Flux[1] = S1 * int_l;
  Flux[2] = S2 * int_l;
  Flux[3] = int_l * kl * (Gr[3][0] * n1 + Gr[3][1] * n2);
  for (int v = 0; v < 4; v++) RV(b, v, Im, Jm) = RV(b, v, Im, Jm) - (double)modfm2 * Flux[v];
}

static int bc_fluxes(q2do *o, ocase *cs) {
  int err = 0;
  for (int k = 0; k < cs->nbc; k++) {
    const obc *rec = &o->bc[cs->first_bc + k];
    const q2d_bc_entry *e = &rec->e;
    oblk *b = &o->blk[cs->first_block + e->b - 1];
    switch (e->type) {
      case 101: case 201: {
        int r = bc_connection_eul(b, cs, e->i, e->j, e->f);
        if (r) err = r;
        if (cs->d.scheme.model > 0) bc_connection_visc(o, b, cs, e);   /* Mod_BC_Fluxes.f90:109-113 */
        break;
      }
      case 300:
        bc_symmetry_eul(b, cs, e->i, e->j, e->f);
        if (cs->d.scheme.model > 0) bc_symmetry_visc(b, cs, e->i, e->j, e->f);
        break;
      case 302:
        bc_symmetry_eul(b, cs, e->i, e->j, e->f);
        if (cs->d.scheme.model > 0) bc_wall_temperature(b, cs, e->i, e->j, e->f, e->T);
        break;
      case 400: bc_extrapolation(b, cs, e->i, e->j, e->f); break;
      case 405: bc_supersonic_inflow(b, cs, rec, e->i, e->j, e->f); break;
      case 406: bc_outflow(b, cs, rec, e->i, e->j, e->f); break;
      default: break;
    }
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* Sources — Lib_Source_Axisymmetry.f90:40-78, Lib_Source_Areavariation.f90:104-144 */
/* ------------------------------------------------------------------------ */
/* Ghost_Wall_Extrapolation — Lib_Ghost.f90:366-399 (wall codes 301..304) */
static void ghost_wall_extrapolation(q2do *o, ocase *cs) {
  for (int k = 0; k < cs->nbc; k++) {
    const q2d_bc_entry *e = &o->bc[cs->first_bc + k].e;
    if (e->type < 301 || e->type > 304) continue;
    oblk *b = &o->blk[cs->first_block + e->b - 1];
    int ig = e->i - guide[e->f][0], jg = e->j - guide[e->f][1];
    PV(b, 1, ig, jg) = -PV(b, 1, e->i, e->j);
    PV(b, 2, ig, jg) = -PV(b, 2, e->i, e->j);
    PV(b, 0, ig, jg) = PV(b, 0, e->i, e->j);
    PV(b, 3, ig, jg) = PV(b, 3, e->i, e->j);
  }
}

/* Source_Axi_Visc_Blk — Lib_Source_Axisymmetry.f90:81-247 (laminar, one species) */
static void source_axi_visc(oblk *b, const ocase *cs) {
  const gas_t *g = &cs->gas;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int j = 1; j <= b->nj; j++) {
    for (int i = 1; i <= b->ni; i++) {
      double rho = PV(b, 0, i, j);
      double rho1 = PV(b, 0, i + 1, j), rho2 = PV(b, 0, i - 1, j), rho3 = PV(b, 0, i, j + 1), rho4 = PV(b, 0, i, j - 1);
      double r_ave = b->sa[IN(b, i, j)], invr = b->sb[IN(b, i, j)];
      double T = PV(b, 3, i, j) / (rho * g->R);
      double T1 = PV(b, 3, i + 1, j) / (rho1 * g->R);
      double T2 = PV(b, 3, i - 1, j) / (rho2 * g->R);
      double T3 = PV(b, 3, i, j + 1) / (rho3 * g->R);
      double T4 = PV(b, 3, i, j - 1) / (rho4 * g->R);
      double mil = mu_lam(g, T), kl = k_lam(g, mil, T);
      double Gr[4][2];
      Gr[0][0] = (PV(b, 0, i + 1, j) / rho1 - PV(b, 0, i - 1, j) / rho2) * 0.5;
      Gr[0][1] = (PV(b, 0, i, j + 1) / rho3 - PV(b, 0, i, j - 1) / rho4) * 0.5;
      Gr[1][0] = (PV(b, 1, i + 1, j) - PV(b, 1, i - 1, j)) * 0.5;
      Gr[2][0] = (PV(b, 2, i + 1, j) - PV(b, 2, i - 1, j)) * 0.5;
      Gr[1][1] = (PV(b, 1, i, j + 1) - PV(b, 1, i, j - 1)) * 0.5;
      Gr[2][1] = (PV(b, 2, i, j + 1) - PV(b, 2, i, j - 1)) * 0.5;
      Gr[3][0] = (T1 - T2) * 0.5;
      Gr[3][1] = (T3 - T4) * 0.5;
      const double *M = &MM(b, 1, 1, i, j);
      for (int k = 0; k < 4; k++) grad_times_M(Gr[k], M);
      double kappa = kl + 0.0;
      /* Stress_Tensor_Axi — Lib_Fluid.f90:163-198 */
      double vr = PV(b, 2, i, j);
      double mi = mil + 0.0;
      double lam = negativetwothird * mi;
      double Divervel = Gr[1][0] + Gr[2][1] + vr / r_ave;
      double lamDivervel = lam * Divervel;
      double Tau_yy = lamDivervel + 2.0 * mi * Gr[2][1];
      double Tau_xy = mi * (Gr[1][1] + Gr[2][0]);
      double Tau_zz = 2.0 * mi * vr / r_ave + lamDivervel;
      double qr = -kappa * Gr[3][1];
      double S[4];
      S[0] = 0.0;                                       /* one species: Dm*grad(Y) = 0 */
      S[1] = Tau_xy * invr;
      S[2] = (Tau_yy - Tau_zz) * invr;
      S[3] = (-qr + PV(b, 1, i, j) * Tau_xy + PV(b, 2, i, j) * Tau_yy + 0.0) * invr;
      double sr = SUR(b, i, j);
      for (int k = 0; k < NP; k++) RV(b, k, i, j) = RV(b, k, i, j) - S[k] * sr;
    }
  }
}

/* Source_Area_Visc_Blk — Lib_Source_Areavariation.f90:147-349 (laminar, one species: the species
 * gradient (rho/rho) is 0 exactly, so the species, enthalpy-diffusion and Hirschfelder terms vanish) */
static void source_area_visc(oblk *b, const ocase *cs) {
  const gas_t *g = &cs->gas;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int j = 1; j <= b->nj; j++) {
    for (int i = 1; i <= b->ni; i++) {
      double Gx = b->sa[IN(b, i, j)], Gy = b->sb[IN(b, i, j)];
      double rho = PV(b, 0, i, j);
      double rho1 = PV(b, 0, i + 1, j), rho2 = PV(b, 0, i - 1, j), rho3 = PV(b, 0, i, j + 1), rho4 = PV(b, 0, i, j - 1);
      double u = PV(b, 1, i, j), v = PV(b, 2, i, j);
      double V_dot_G = u * Gx + v * Gy;
      double T = PV(b, 3, i, j) / (rho * g->R);
      double T1 = PV(b, 3, i + 1, j) / (rho1 * g->R);
      double T2 = PV(b, 3, i - 1, j) / (rho2 * g->R);
      double T3 = PV(b, 3, i, j + 1) / (rho3 * g->R);
      double T4 = PV(b, 3, i, j - 1) / (rho4 * g->R);
      double mil = mu_lam(g, T), kl = k_lam(g, mil, T);
      double Gr[4][2];
      Gr[0][0] = (PV(b, 0, i + 1, j) / rho1 - PV(b, 0, i - 1, j) / rho2) * 0.5;
      Gr[0][1] = (PV(b, 0, i, j + 1) / rho3 - PV(b, 0, i, j - 1) / rho4) * 0.5;
      Gr[1][0] = (PV(b, 1, i + 1, j) - PV(b, 1, i - 1, j)) * 0.5;
      Gr[2][0] = (PV(b, 2, i + 1, j) - PV(b, 2, i - 1, j)) * 0.5;
      Gr[1][1] = (PV(b, 1, i, j + 1) - PV(b, 1, i, j - 1)) * 0.5;
      Gr[2][1] = (PV(b, 2, i, j + 1) - PV(b, 2, i, j - 1)) * 0.5;
      Gr[3][0] = (T1 - T2) * 0.5;
      Gr[3][1] = (T3 - T4) * 0.5;
      const double *M = &MM(b, 1, 1, i, j);
      for (int k = 0; k < 4; k++) grad_times_M(Gr[k], M);
      double kappa = kl + 0.0;
      /* Stress_Tensor_Std with V_dot_G — Lib_Fluid.f90:128-163 */
      double mi = mil + 0.0;
      double lam = negativetwothird * mi;
      double Divervel = Gr[1][0] + Gr[2][1] - V_dot_G;
      double lamDivervel = lam * Divervel;
      double T11 = lamDivervel + 2.0 * mi * Gr[1][0];
      double T22 = lamDivervel + 2.0 * mi * Gr[2][1];
      double T12 = mi * (Gr[1][1] + Gr[2][0]);
      double T33 = lamDivervel + 2.0 * mi * (-V_dot_G);
      double qx = -kappa * Gr[3][0];
      double qy = -kappa * Gr[3][1];
      double S[4];
      S[0] = 0.0;
      S[1] = Gx * (T33 - T11) - Gy * T12;
      S[2] = -Gx * T12 + Gy * (T33 - T22);
      S[3] = -u * (T11 * Gx + T12 * Gy) - v * (T12 * Gx + T22 * Gy) + Gx * qx + Gy * qy;
      double sr = SUR(b, i, j);
      for (int k = 0; k < NP; k++) RV(b, k, i, j) = RV(b, k, i, j) - S[k] * sr;
    }
  }
}

static void sources(q2do *o, ocase *cs) {
  const q2d_scheme *s = &cs->d.scheme;
  if (s->source == Q2D_SRC_NONE) return;
  for (int bb = 0; bb < cs->nblocks; bb++) {
    oblk *b = &o->blk[cs->first_block + bb];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 1; j <= b->nj; j++) {
      for (int i = 1; i <= b->ni; i++) {
        double rho = PV(b, 0, i, j), u = PV(b, 1, i, j), v = PV(b, 2, i, j), p = PV(b, 3, i, j);
        double mod_v = sqrt(u * u + v * v);
        double S[NP];
        if (s->source == Q2D_SRC_AXI) {
          double invr = b->sb[IN(b, i, j)];
          S[0] = -(rho * v * invr);
          S[1] = -(rho * u * v * invr);
          S[2] = -(rho * v * v * invr);
          S[3] = g_plant == Q2DO_PLANT_AXI_ENERGY ? -((rho * f_E0(&cs->gas, p, rho, mod_v)) * v * invr)
                                                  : -((rho * f_E0(&cs->gas, p, rho, mod_v) + p) * v * invr);
          if (g_plant == Q2DO_PLANT_AXI_SIGN)
            for (int k = 0; k < NP; k++) S[k] = -S[k];
        } else {
          if (!b->sa) continue;
          double Gx = b->sa[IN(b, i, j)], Gy = b->sb[IN(b, i, j)];
          double VdG = u * Gx + v * Gy;
          S[0] = rho * VdG;
          S[1] = rho * u * VdG;
          S[2] = rho * v * VdG;
          S[3] = (rho * f_E0(&cs->gas, p, rho, mod_v) + p) * VdG;
        }
        double sr = SUR(b, i, j);
        for (int k = 0; k < NP; k++) RV(b, k, i, j) = RV(b, k, i, j) - S[k] * sr;
      }
    }
    if (s->source == Q2D_SRC_AXI && s->model > 0) source_axi_visc(b, cs);
    if (s->source == Q2D_SRC_AREA && s->model > 0 && b->sa) source_area_visc(b, cs);   /* Lib_Sources.f90:98-115 */
  }
}

/* ------------------------------------------------------------------------ */
/* Implicit residual smoothing — Lib_IRS.f90:17-177                          */
/* ------------------------------------------------------------------------ */
/* Ghost_Residual_Extrapolation (:130-177): one ghost layer per record.  A connection 101 takes the
 * donor cell's residual (:163-169); every other type, periodic 201 included, takes the boundary cell's
 * own (`case default`, :171-172).  Ghosts of a face without records keep what they hold (zero here:
 * the reference leaves that array uninitialised, Mod_Allocate_Data.f90:123). */
static void irs_ghosts(q2do *o, ocase *cs) {
  for (int k = 0; k < cs->nbc; k++) {
    const q2d_bc_entry *e = &o->bc[cs->first_bc + k].e;
    oblk *bm = &o->blk[cs->first_block + e->b - 1];
    int ig = e->i - guide[e->f][0], jg = e->j - guide[e->f][1];
    if (e->type == Q2D_BC_CONNECT) {
      oblk *bs = &o->blk[cs->first_block + e->bs - 1];
      for (int v = 0; v < NP; v++) RV(bm, v, ig, jg) = RV(bs, v, e->is, e->js);
    } else {
      for (int v = 0; v < NP; v++) RV(bm, v, ig, jg) = RV(bm, v, e->i, e->j);
    }
  }
}

/* Residual_Smoothing_Blk (:47-127) along direction d: Res_Star = Res over 0..n+1 (:60-65), then
 * njacobi = 2 sweeps Res_Star_New = (Res + beta*(Res_Star(k-1) + Res_Star(k+1))) / (1 + 2 beta) on the
 * interior (:68-93, stencil :106-125), then Res = Res_Star on the interior (:97-102).  rs and rsn are
 * scratch arrays of the block's full extent. */
static void irs_block(oblk *b, int d, double beta, double *rs, double *rsn) {
  for (int j = 0; j <= b->nj + 1; j++)
    for (int i = 0; i <= b->ni + 1; i++)
      for (int v = 0; v < NP; v++) rs[cidx(b, i, j) * NP + v] = RV(b, v, i, j);
  for (int jacobi = 1; jacobi <= 2; jacobi++) {
    for (int j = 1; j <= b->nj; j++)
      for (int i = 1; i <= b->ni; i++) {
        int i1 = d == 1 ? i - 1 : i, j1 = d == 1 ? j : j - 1;
        int i2 = d == 1 ? i + 1 : i, j2 = d == 1 ? j : j + 1;
        for (int v = 0; v < NP; v++) {
          double r = RV(b, v, i, j), rm = rs[cidx(b, i1, j1) * NP + v], rp = rs[cidx(b, i2, j2) * NP + v];
          rsn[cidx(b, i, j) * NP + v] = (r + beta * (rm + rp)) / (1.0 + 2.0 * beta);
        }
      }
    for (int j = 1; j <= b->nj; j++)
      for (int i = 1; i <= b->ni; i++)
        for (int v = 0; v < NP; v++) rs[cidx(b, i, j) * NP + v] = rsn[cidx(b, i, j) * NP + v];
  }
  for (int j = 1; j <= b->nj; j++)
    for (int i = 1; i <= b->ni; i++)
      for (int v = 0; v < NP; v++) RV(b, v, i, j) = rs[cidx(b, i, j) * NP + v];
}

/* Residual_Smoothing (:17-44): for d = 1..ndir, the ghost extrapolation of every record, then every
 * block's sweeps along d. */
static int irs_smooth(q2do *o, ocase *cs, double beta) {
  size_t mx = 1;
  for (int bb = 0; bb < cs->nblocks; bb++) {
    const oblk *b = &o->blk[cs->first_block + bb];
    size_t n = (size_t)(b->ni + 2 * GC) * (size_t)(b->nj + 2 * GC) * NP;
    if (n > mx) mx = n;
  }
  double *rs = (double *)calloc(mx, sizeof(double)), *rsn = (double *)calloc(mx, sizeof(double));
  if (!rs || !rsn) { free(rs); free(rsn); return Q2D_E_NOMEM; }
  for (int d = 1; d <= cs->ndir; d++) {
    irs_ghosts(o, cs);
    for (int bb = 0; bb < cs->nblocks; bb++) irs_block(&o->blk[cs->first_block + bb], d, beta, rs, rsn);
  }
  free(rs); free(rsn);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Newstate_Conservative — Lib_Newstate.f90:14-161                           */
/* ------------------------------------------------------------------------ */
static const double RKcoeff[3][3] = { {1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0} };

/* Without IRS (beta = 0) each cell scales its residual and updates at once (Compute_Cons_NoIRS,
 * :66-82 and :130-159).  With IRS (beta > 0, Assign_Setup.f90:61) every residual is scaled first
 * (Scale_Residual_Cons, :37-49 and :93-99), then smoothed (Residual_Smoothing, :52), then every cell
 * updates with its smoothed residual (Update_State_Cons_IRS, :55-64 and :101-127). */
static int newstate(q2do *o, ocase *cs, int irk, int *fb, int *fi, int *fj) {
  const q2d_scheme *s = &cs->d.scheme;
  double c = RKcoeff[s->n_rk - 1][irk - 1];
  double strang = 1.0;
  const int irs = s->irs_beta > 0.0;
  int err = 0;
  if (irs) {
    for (int bb = 0; bb < cs->nblocks; bb++) {
      oblk *b = &o->blk[cs->first_block + bb];
      for (int j = 1; j <= b->nj; j++)
        for (int i = 1; i <= b->ni; i++) {
          double dt = b->dtlocal[IN(b, i, j)], sr = SUR(b, i, j);
          for (int v = 0; v < NP; v++) RV(b, v, i, j) = -RV(b, v, i, j) / sr * dt * strang;
        }
    }
    int r = irs_smooth(o, cs, s->irs_beta);
    if (r) return r;
  }
  for (int bb = 0; bb < cs->nblocks; bb++) {
    oblk *b = &o->blk[cs->first_block + bb];
    /* rows in parallel (the reference's !$omp do collapse(2), Lib_Newstate.f90:71-84); each row keeps its
     * first failing column, and the first failing row is taken after the loop: the serial order's first */
    int *row_fail = (int *)calloc((size_t)b->nj + 1, sizeof(int));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 1; j <= b->nj; j++) {
      for (int i = 1; i <= b->ni; i++) {
        double dt = b->dtlocal[IN(b, i, j)], sr = SUR(b, i, j);
        double res[NP], prim[NP], primold[NP], cons[NP], consold[NP];
        for (int v = 0; v < NP; v++) {
          res[v] = irs ? RV(b, v, i, j) : -RV(b, v, i, j) / sr * dt * strang;
          prim[v] = PV(b, v, i, j); primold[v] = POV(b, v, i, j);
        }
        prim2cons(&cs->gas, primold, consold);
        prim2cons(&cs->gas, prim, cons);
        for (int v = 0; v < NP; v++) {                 /* RK_stage — Lib_RK.f90:15-27 */
          double r_ = (cons[v] - consold[v] + res[v]) * c;
          cons[v] = consold[v] + r_;
        }
        cons2prim(&cs->gas, cons, prim, eos_T(&cs->gas, prim[3], prim[0]));   /* :119-123 */
        /* Check_And_Fix_State — Lib_Newstate.f90:316-350 */
        if (isnan(prim[0]) || isnan(prim[1]) || isnan(prim[2]) || isnan(prim[3]) || prim[3] < 0.0) {
          if (!row_fail[j]) row_fail[j] = i;
        }
        if (prim[0] < 1e-20) prim[0] = 1e-20;
        for (int v = 0; v < NP; v++) PV(b, v, i, j) = prim[v];
      }
    }
    for (int j = 1; j <= b->nj && !err; j++)
      if (row_fail[j]) { err = Q2D_E_STATE; *fb = bb + 1; *fi = row_fail[j]; *fj = j; }
    free(row_fail);
  }
  return err;
}

/* Compute_Residual — Mod_Diagnostic.f90:438-469 (row partial sums, rows summed in order) */
static void residual(q2do *o, ocase *cs, double *tot) {
  for (int k = 0; k < 4; k++) tot[k] = 0.0;
  for (int bb = 0; bb < cs->nblocks; bb++) {
    oblk *b = &o->blk[cs->first_block + bb];
    double blk[4] = {0, 0, 0, 0};
    double *rows = (double *)malloc(sizeof(double) * 4 * (size_t)(b->nj + 1));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 1; j <= b->nj; j++) {                 /* row partials in parallel ... */
      double row[4] = {0, 0, 0, 0};
      for (int i = 1; i <= b->ni; i++) {
        for (int v = 0; v < 4; v++) {
          double r = fabs(PV(b, v, i, j) - POV(b, v, i, j));
          row[v] = row[v] + r * r;
        }
      }
      for (int v = 0; v < 4; v++) rows[4 * j + v] = row[v];
    }
    for (int j = 1; j <= b->nj; j++)                   /* ... summed in row order (the same bits) */
      for (int v = 0; v < 4; v++) blk[v] = blk[v] + rows[4 * j + v];
    free(rows);
    for (int v = 0; v < 4; v++) tot[v] = tot[v] + blk[v];
  }
  for (int v = 0; v < 4; v++) tot[v] = sqrt(tot[v]);
}

/* ------------------------------------------------------------------------ */
/* Explicit_Step — Mod_Explicit.f90:10-174                                   */
/* ------------------------------------------------------------------------ */
static int step_case(q2do *o, ocase *cs, q2d_step_out *out) {
  const q2d_scheme *s = &cs->d.scheme;
  int status = 0, fb = 0, fi = 0, fj = 0;
  cs->iter += 1;
  cs->iter_from_call += 1;
  cs->dtglobal = 1e5;
  compute_dt(o, cs);
  if (s->time_accurate) {
    set_global_dt(o, cs);
    cs->time = cs->time + cs->dtglobal;
  }
  for (int bb = 0; bb < cs->nblocks; bb++) {          /* Copy_State :177-197 */
    oblk *b = &o->blk[cs->first_block + bb];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int j = 1; j <= b->nj; j++)
      for (int i = 1; i <= b->ni; i++)
        for (int v = 0; v < NP; v++) POV(b, v, i, j) = PV(b, v, i, j);
  }
  for (int irk = 1; irk <= s->n_rk; irk++) {
    fill_ghost(o, cs);
    for (int bb = 0; bb < cs->nblocks; bb++) {
      int r = fluxes_block(&o->blk[cs->first_block + bb], cs);
      if (r && !status) status = r;
    }
    { int r = bc_fluxes(o, cs); if (r && !status) status = r; }
    if (s->source != Q2D_SRC_NONE) ghost_wall_extrapolation(o, cs);   /* Mod_Explicit.f90:83 */
    sources(o, cs);
    {                                                   /* the first failure stops the reference */
      int tb = 0, ti = 0, tj = 0;
      int r = newstate(o, cs, irk, &tb, &ti, &tj);
      if (r && !status) { status = r; fb = tb; fi = ti; fj = tj; }
    }
  }
  double res[4];
  residual(o, cs, res);
  int endsim = (cs->iter_from_call >= cs->d.itermax) || (res[0] <= cs->d.res_threshold) ||
               (cs->time >= cs->d.time_threshold);
  if (out) {
    out->dtglobal = cs->dtglobal; out->time = cs->time;
    for (int k = 0; k < 4; k++) out->residual[k] = res[k];
    out->iter = cs->iter; out->todo = endsim ? 3 : 1;
    out->status = status; out->fail_b = fb; out->fail_i = fi; out->fail_j = fj;
  }
  return status;
}

int q2do_step(q2do *o, q2d_step_out *out) {
  int st = 0;
  for (int c = 0; c < o->ncases; c++) {
    int r = step_case(o, &o->cs[c], out ? &out[c] : NULL);
    if (r && !st) st = r;
  }
  return st;
}

int q2do_get_state(q2do *o, int block, double *P) {
  if (block < 0 || block >= o->nblocks) return Q2D_E_ARG;
  oblk *b = &o->blk[block];
  memcpy(P, b->P, sizeof(double) * NP * (size_t)(b->ni + 2 * GC) * (size_t)(b->nj + 2 * GC));
  return 0;
}
int q2do_set_state(q2do *o, int block, const double *P) {
  if (block < 0 || block >= o->nblocks) return Q2D_E_ARG;
  oblk *b = &o->blk[block];
  memcpy(b->P, P, sizeof(double) * NP * (size_t)(b->ni + 2 * GC) * (size_t)(b->nj + 2 * GC));
  return 0;
}
int q2do_get_dt(q2do *o, int block, double *dt) {
  if (block < 0 || block >= o->nblocks) return Q2D_E_ARG;
  oblk *b = &o->blk[block];
  memcpy(dt, b->dtlocal, sizeof(double) * (size_t)b->ni * (size_t)b->nj);
  return 0;
}

int q2do_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n; return 1;
#endif
}
int q2do_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- unit-level entry points ---- */
double q2do_limiter(int id, double x, double y) { return limiter(id, x, y); }

void q2do_reconstruct(int lim, double prev, double local, double next, double next2,
                      double dl0, double dl1, double dl2, double dll, double dlr, double beta,
                      double *priml, double *primr) {
  reconstruct(lim, prev, local, next, next2, dl0, dl1, dl2, dll, dlr, beta, priml, primr);
}

/* The unit entry points' gas: closed form, or table mode with the tables of q2do_unit_tables. */
static const double *g_unit_tab = NULL;
static int g_unit_ntab = 0;
static gas_t unit_gas(double R, double cp) {
  gas_t g = mkgas(R, cp);
  if (g_unit_tab) { g.tab = g_unit_tab; g.ntab = g_unit_ntab; }
  return g;
}
void q2do_unit_tables(const double *tab, int n) { g_unit_tab = (tab && n >= 2) ? tab : NULL; g_unit_ntab = n; }

double q2do_thermo_eval(int what, double gas_R, double gas_cp, double a, double b) {
  gas_t g = unit_gas(gas_R, gas_cp);
  switch (what) {
    case 0: return g.tab ? tabT(&g, TAB_H, a) : g.cp * a;
    case 1: return g.tab ? tabT(&g, TAB_CP, a) : g.cp;
    case 2: return mu_lam(&g, a);
    case 3: return k_lam(&g, mu_lam(&g, a), a);
    case 4: return f_gamma_T(&g, a);
    case 5: return f_e(&g, a);
    case 6: return g.tab ? t_of_e(&g, a, b) : a / g.cv;
    case 7: return f_ss(&g, a, b);
    default: return NAN;
  }
}
void q2do_prim2cons(double gas_R, double gas_cp, const double *prim, double *cons) {
  gas_t g = unit_gas(gas_R, gas_cp);
  prim2cons(&g, prim, cons);
}
void q2do_cons2prim(double gas_R, double gas_cp, const double *cons, double T0, double *prim) {
  gas_t g = unit_gas(gas_R, gas_cp);
  cons2prim(&g, cons, prim, T0);
}

void q2do_riemann(int id, double gas_R, double gas_cp, const double *L, const double *Rr,
                  double beta, double nx, double ny, double *F) {
  gas_t g = unit_gas(gas_R, gas_cp);
  double aL = f_ss(&g, L[3], L[0]), aR = f_ss(&g, Rr[3], Rr[0]);
  riemann(id, &g, L[0], L[1], L[2], L[3], aL, L[0], Rr[0], Rr[1], Rr[2], Rr[3], aR, Rr[0], beta, nx, ny, F);
}

void q2do_roe_averages(double gas_R, double gas_cp, const double *L, const double *Rr, double *out) {
  gas_t g = unit_gas(gas_R, gas_cp);
  roe_t r = roe_averages(&g, L[0], Rr[0], L[0], Rr[0], L[1], L[2], Rr[1], Rr[2], L[3], Rr[3]);
  out[0] = r.h0l; out[1] = r.h0r; out[2] = r.rho; out[3] = r.u; out[4] = r.v; out[5] = r.a; out[6] = r.h0;
}

int q2do_convective_flux(int lim, int rid, double gas_R, double gas_cp, const double *dl,
                         double nx, double ny, double int_l, const double *prim, double *flux) {
  gas_t g = unit_gas(gas_R, gas_cp);
  return convective_flux(lim, rid, &g, dl, nx, ny, int_l, (const double(*)[NP])prim, 1.0, 1.0, flux);
}

int q2do_get_residual(q2do *o, int block, double *R) {
  if (block < 0 || block >= o->nblocks) return Q2D_E_ARG;
  oblk *b = &o->blk[block];
  memcpy(R, b->R, sizeof(double) * NP * (size_t)(b->ni + 2 * GC) * (size_t)(b->nj + 2 * GC));
  return 0;
}
int q2do_set_residual(q2do *o, int block, const double *R) {
  if (block < 0 || block >= o->nblocks) return Q2D_E_ARG;
  oblk *b = &o->blk[block];
  memcpy(b->R, R, sizeof(double) * NP * (size_t)(b->ni + 2 * GC) * (size_t)(b->nj + 2 * GC));
  return 0;
}
int q2do_irs_smooth(q2do *o, int case_id, double beta) {
  if (case_id < 0 || case_id >= o->ncases) return Q2D_E_ARG;
  return irs_smooth(o, &o->cs[case_id], beta);
}

double q2do_sd_tramel(double p, double p_im, double p_ip, double p_jm, double p_jp) {
  return sd_tramel(p, p_im, p_ip, p_jm, p_jp);
}

/* Read_Inifile_Runtime's keys of one case (Read_Ini.f90:65-88): the next step's Compute_dt and stop test
 * read them. */
int q2do_set_runtime(q2do *o, int case_id, const q2d_runtime *rt) {
  if (!o || !rt || case_id < 0 || case_id >= o->ncases) return Q2D_E_ARG;
  ocase *cs = &o->cs[case_id];
  q2d_scheme *s = &cs->d.scheme;
  s->cfl = rt->cfl; s->vnn = rt->vnn; s->ramp_iter = rt->ramp_iter;
  s->limiter = rt->limiter; s->riemann = rt->riemann; s->space_sd = rt->space_sd;
  cs->d.itermax = rt->itermax; cs->d.res_threshold = rt->res_threshold;
  cs->d.time_threshold = rt->time_threshold; cs->d.res_drop = rt->res_drop;
  return Q2D_OK;
}

double q2do_pow(double x, double y) { return exact_pow(x, y); }
int q2do_plant(int which) { int old = g_plant; g_plant = which; return old; }

int q2do_bc_inflow_flux(int limiter_id, int riemann_id, double gas_R, double gas_cp, int face, const double *Bp,
                        const double *Ip, double n1, double n2, double int_l, double mach, double T, double p,
                        double rel_fac, double alpha, double *F) {
  gas_t g = unit_gas(gas_R, gas_cp);
  return inflow_flux(&g, limiter_id, riemann_id, face, Bp, Ip, n1, n2, int_l, mach, T, p, rel_fac, cos(alpha),
                     sin(alpha), F);
}
int q2do_bc_outflow_flux(double gas_R, double gas_cp, int face, const double *Bp, double n1, double n2,
                         double int_l, double p_exit, double rel_fac, double *F) {
  gas_t g = unit_gas(gas_R, gas_cp);
  return outflow_flux(&g, face, Bp, n1, n2, int_l, p_exit, rel_fac, F);
}
