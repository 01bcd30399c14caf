// q2d_device.h — per-face and per-cell arithmetic of the explicit step, for CDNA4.
//
// Every function keeps the evaluation order of the Fortran it restates
// (file:line in each comment), so that with T = double and floating-point
// contraction disabled the results are bit-identical to a literal FP64
// reading of the reference.  With T = float the same code is the FP32 face
// computation of the mixed-precision form, where division and square root use
// the hardware reciprocal / square-root instructions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace q2d {

// Gas closure of one case: R, cp, cv, gamma of the calorically perfect closed forms, or, with tab set
// (FP64 only), FLINT's native tabulated mode: planes h, cp, mu, k of ntab nodes each, node k at T = k.
// DV selects the FP64 division of the functions that take the closure (gdiv): 0 the compiler's IEEE
// division, 1 ddiv_fast with the binary exponents of every numerator and quotient folded into ex[0..1].
template <typename T, int DV = 0> struct Gas {
  T R, cp, cv, gam;
  const double* tab = nullptr;
  int ntab = 0;
  int* ex = nullptr;
};

// ---- arithmetic primitives ----------------------------------------------
// FP64 division and square root: the compiler's correctly rounded IEEE operations (the FP64 form's
// results must equal the reference's bit for bit, subnormal round-off included).
__device__ __forceinline__ double ddiv_rn(double a, double b) { return a / b; }
__device__ __forceinline__ double dsqrt_rn(double x) { return __builtin_sqrt(x); }

// a / b by the hardware reciprocal refined by two Newton steps, the quotient and one remainder
// correction (Markstein), then v_div_fixup_f64 for zero, infinite and NaN operands (signed zeros and
// the numerator's NaN included).  This is the compiler's IEEE f64 division without its two
// v_div_scale_f64 steps and with a plain fma for v_div_fmas_f64; those steps rescale operands near
// the ends of the exponent range and are the identity when a = 0 or 2^-969 <= |a|, 2^-1021 <= |b| <=
// 2^1021 (or b = 0) and 2^-1022 <= |a / b| < 2^1023, where both give the IEEE quotient bit for bit.
// Without the scaling steps the sequence does not pass a flag through VCC (one register, which
// serialises the compiler's divisions) and divisions by one denominator share the refined reciprocal.
__device__ __forceinline__ double ddiv_fast(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  double q = a * r;
  const double rem = __builtin_fma(-b, q, a);
  q = __builtin_fma(rem, r, q);
  return __builtin_amdgcn_div_fixup(q, b, a);
}

template <typename T> __device__ __forceinline__ T qdiv(T a, T b);
template <> __device__ __forceinline__ double qdiv<double>(double a, double b) { return ddiv_rn(a, b); }
template <> __device__ __forceinline__ float qdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename T> __device__ __forceinline__ T qsqrt(T a);
template <> __device__ __forceinline__ double qsqrt<double>(double a) { return dsqrt_rn(a); }
template <> __device__ __forceinline__ float qsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }

// Division of the functions that take a gas closure: FP32 as qdiv; FP64 with DV = 0 the IEEE division,
// with DV = 1 ddiv_fast, the exponents (frexp: x = m 2^e, 1/2 <= |m| < 1; 0 for x = 0) of the
// numerator and of the quotient folded into the running minima g.ex[0] and g.ex[1], which the caller
// checks: |a| >= 2^-969 while g.ex[0] >= -968, |q| >= 2^-1020 while g.ex[1] >= -1019.
template <typename T, int DV> __device__ __forceinline__ T gdiv(const Gas<T, DV>& g, T a, T b) {
  if constexpr (sizeof(T) == 4) {
    return qdiv<T>(a, b);
  } else if constexpr (DV == 0) {
    return ddiv_rn(a, b);
  } else {
    const double q = ddiv_fast(a, b);
    const int ea = __builtin_amdgcn_frexp_exp(a), eq = __builtin_amdgcn_frexp_exp(q);
    g.ex[0] = ea < g.ex[0] ? ea : g.ex[0];
    g.ex[1] = eq < g.ex[1] ? eq : g.ex[1];
    return q;
  }
}

template <typename T> __device__ __forceinline__ T qabs(T a) { return __builtin_fabs(a); }
template <> __device__ __forceinline__ float qabs<float>(float a) { return __builtin_fabsf(a); }
template <typename T> __device__ __forceinline__ T qcopysign(T a, T b) { return __builtin_copysign(a, b); }
template <> __device__ __forceinline__ float qcopysign<float>(float a, float b) { return __builtin_copysignf(a, b); }

// Fortran MIN/MAX of three reals, left to right.
template <typename T> __device__ __forceinline__ T min3(T a, T b, T c) { T m = a; if (b < m) m = b; if (c < m) m = c; return m; }
template <typename T> __device__ __forceinline__ T max3(T a, T b, T c) { T m = a; if (b > m) m = b; if (c > m) m = c; return m; }

// ---- property tables (FLINT's native tabulated mode, one species) ---------
// A property at T: v(Ti) + (v(Ti+1) - v(Ti)) (T - Ti), Ti = int(T) (docs/theory/thermo.md:39-41; the
// arithmetic of Lib_Riemann_HLL.f90:257-282), Ti clamped to the table's intervals (T beyond them
// extrapolates the end interval; T < 1 or NaN takes interval 0).
enum { kTabH = 0, kTabCp = 1, kTabMu = 2, kTabK = 3 };
__device__ __forceinline__ int tab_node(int ntab, double T) {
  if (!(T >= 1.0)) return 0;
  if (!(T < (double)(ntab - 1))) return ntab - 2;
  return (int)T;
}
__device__ __forceinline__ double tab_interp(const double* tab, int ntab, int q, int Ti, double dT) {
  const double* v = tab + (long long)q * ntab;
  const double a = v[Ti], b = v[Ti + 1];
  return a + (b - a) * dT;
}
// the same in T's arithmetic (FP32: the mixed form's face arithmetic on the FP64 tables' nodes)
template <typename T> __device__ __forceinline__ T tab_interp_t(const double* tab, int ntab, int q, int Ti, T dT) {
  const double* v = tab + (long long)q * ntab;
  const T a = (T)v[Ti], b = (T)v[Ti + 1];
  return a + (b - a) * dT;
}
template <typename T> __device__ __forceinline__ T tab_at_t(const double* tab, int ntab, int q, T Tk) {
  const int Ti = tab_node(ntab, (double)Tk);
  return tab_interp_t<T>(tab, ntab, q, Ti, Tk - (T)Ti);
}
__device__ __forceinline__ double tab_at(const double* tab, int ntab, int q, double T) {
  const int Ti = tab_node(ntab, T);
  return tab_interp(tab, ntab, q, Ti, T - (double)Ti);
}
// cp / (cp - R) at T (f_gamma)
template <int DV>
__device__ __forceinline__ double tab_gamma(const Gas<double, DV>& g, double T) {
  const double cp = tab_at(g.tab, g.ntab, kTabCp, T);
  return gdiv(g, cp, cp - g.R);
}
// cons2prim's temperature: the exact inverse of the interpolated e(T) = h(T) - R T, by Newton steps from
// the estimate T0 (the stage input's p / (rho R), Lib_Newstate.f90:119-123), each with the slope of its
// interval (a step that lands in its own interval is the root); at most 32 steps.
template <int DV>
__device__ __forceinline__ double tab_t_of_e(const Gas<double, DV>& g, double e, double T0) {
  const double* tab = g.tab;
  const int ntab = g.ntab;
  const double R = g.R;
  double T = T0;
  for (int it = 0; it < 32; ++it) {
    const int Ti = tab_node(ntab, T);
    const double h0 = tab[Ti], h1 = tab[Ti + 1];
    const double e0 = h0 - R * (double)Ti;
    const double sl = (h1 - h0) - R;
    const double Tn = (double)Ti + gdiv(g, e - e0, sl);
    if (Tn >= (double)Ti && Tn <= (double)(Ti + 1)) return Tn;
    if ((Ti == 0 && Tn < 0.0) || (Ti == ntab - 2 && Tn > (double)(Ti + 1))) return Tn;
    T = Tn;
  }
  return T;
}

// ---- gas closure (single species) ------------------------------------------
// a = sqrt(gamma*p/rho)  (f_ss); in table mode gamma at T = p / (rho R)
template <typename T, int DV> __device__ __forceinline__ T sound(const Gas<T, DV>& g, T p, T rho) {
  if constexpr (sizeof(T) == 8) {
    if (g.tab) return dsqrt_rn(gdiv(g, tab_gamma(g, gdiv(g, p, rho * g.R)) * p, rho));
  } else {
    if (g.tab) {                                  // FP32: gamma = cp / (cp - R) at T = p / (rho R)
      const T cp = tab_at_t<T>(g.tab, g.ntab, kTabCp, gdiv(g, p, rho * g.R));
      return qsqrt<T>(gdiv(g, gdiv(g, cp, cp - g.R) * p, rho));
    }
  }
  return qsqrt<T>(gdiv(g, g.gam * p, rho));
}
// specific enthalpy h = cp T (closed form) or h(T) (table), T = p / (rho R)
template <typename T, int DV> __device__ __forceinline__ T enthalpy(const Gas<T, DV>& g, T p, T rho) {
  if constexpr (sizeof(T) == 8) {
    if (g.tab) return tab_at(g.tab, g.ntab, kTabH, gdiv(g, p, rho * g.R));
  } else {
    if (g.tab) return tab_at_t<T>(g.tab, g.ntab, kTabH, gdiv(g, p, rho * g.R));
  }
  return g.cp * gdiv(g, p, rho * g.R);
}
// specific internal energy at temperature T: cv T (closed form) or h(T) - R T (table)
template <int DV> __device__ __forceinline__ double int_energy(const Gas<double, DV>& g, double T) {
  if (g.tab) return tab_at(g.tab, g.ntab, kTabH, T) - g.R * T;
  return g.cv * T;
}

// ---- x^y from exactly specified operations --------------------------------
// The `**` of the characteristic outflow (Lib_BC_Fluxes_Outflow.f90:125-134) for x > 0, built from
// IEEE-754 binary64 +, -, *, /, fma, rint, frexp and ldexp only, so the result does not depend on a
// math library's rounding: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
// with s and the leading term in double-double and the odd series 2 s^3 (1/3 + s^2/5 + ... + s^20/23)
// in double; t = y (e ln2 + ln m) in double-double (ln 2 split hi/lo as in Cody-Waite); exp(t) = 2^k
// exp(r), k = rint(t / ln 2), r = t - k ln 2, exp(r) - 1 by its Taylor series to r^13/13!.  Within one
// ulp of the correctly rounded power (measured against the C library's pow).
__device__ __forceinline__ void dd_two_sum(double a, double b, double& s, double& e) {
  const double t = a + b, bb = t - a;
  s = t; e = (a - (t - bb)) + (b - bb);
}
__device__ __forceinline__ double exact_pow(double x, double y) {
  if (!(x > 0.0) || __builtin_isinf(x) || __builtin_isinf(y) || __builtin_isnan(y)) return __builtin_nan("");
  if (y == 0.0) return 1.0;
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  int e;
  double m = __builtin_frexp(x, &e);
  if (m < 0.70710678118654752440) { m = m * 2.0; e = e - 1; }
  const double num = m - 1.0;
  double den, den_lo;
  dd_two_sum(m, 1.0, den, den_lo);
  const double s = num / den;
  const double s_lo = (__builtin_fma(-s, den, num) - s * den_lo) / den;
  const double s2 = s * s;
  double q = 1.0 / 23.0;
  q = q * s2 + 1.0 / 21.0; q = q * s2 + 1.0 / 19.0; q = q * s2 + 1.0 / 17.0; q = q * s2 + 1.0 / 15.0;
  q = q * s2 + 1.0 / 13.0; q = q * s2 + 1.0 / 11.0; q = q * s2 + 1.0 / 9.0; q = q * s2 + 1.0 / 7.0;
  q = q * s2 + 1.0 / 5.0; q = q * s2 + 1.0 / 3.0;
  const double tail = 2.0 * s * s2 * q;
  double lm, lm_lo;
  dd_two_sum(2.0 * s, 2.0 * s_lo + tail, lm, lm_lo);
  const double de = (double)e;
  double lh, ll;
  dd_two_sum(de * ln2_hi, lm, lh, ll);
  ll = ll + (lm_lo + de * ln2_lo);
  double th = y * lh;
  double tl = __builtin_fma(y, lh, -th) + y * ll;
  dd_two_sum(th, tl, th, tl);
  const double k = __builtin_rint(th * inv_ln2);
  const double r_hi = th - k * ln2_hi;
  const double r_lo = tl - k * ln2_lo;
  double r, r_err;
  dd_two_sum(r_hi, r_lo, r, r_err);
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0; p = p * r + 1.0 / 39916800.0; p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0; p = p * r + 1.0 / 40320.0; p = p * r + 1.0 / 5040.0; p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0; p = p * r + 1.0 / 24.0; p = p * r + 1.0 / 6.0; p = p * r + 0.5;
  const double em1 = r + (r_err + r * r * p);
  return __builtin_ldexp(1.0 + em1, (int)k);
}

// ---- limiters — Lib_Limiters.f90:26-112 ----------------------------------
// g: the closure whose division the limiter uses (gdiv)
template <typename T, int DV> __device__ __forceinline__ T limiter(const Gas<T, DV>& g, int id, T x, T y) {
  const T tiny = T(1e-30);
  switch (id) {
    case 1: {                                   // minmod :36-46
      T r = T(0);
      if (x * y < T(0)) return r;
      r = x;
      if (qabs<T>(x) > qabs<T>(y)) r = y;
      return r;
    }
    case 2: {                                   // vanleer :49-58
      T ru = gdiv(g, y, x + tiny);
      T phi = gdiv(g, ru + qabs<T>(ru), T(1) + qabs<T>(ru));
      return phi * x;
    }
    case 3: {                                   // vanalbada :61-70
      T ru = gdiv(g, y, x + tiny);
      // FP32 (mixed form): |r| held below 1e18 so that r^2 stays finite; where x vanishes r ~ y / 1e-30, and
      // FP64 represents r^2 while FP32 would overflow to inf / inf.  phi(1e18) is 1 to FP32 precision.
      if constexpr (sizeof(T) == 4) ru = __builtin_fminf(__builtin_fmaxf(ru, -1e18f), 1e18f);
      T phi = gdiv(g, ru * ru + ru, T(1) + ru * ru);
      return phi * x;
    }
    case 4: {                                   // MC :73-82
      T ru = gdiv(g, y, x + tiny);
      T m = min3<T>(T(2) * ru, T(0.5) * (T(1) + ru), T(2));
      T phi = T(0); if (m > phi) phi = m;
      return phi * x;
    }
    case 5: {                                   // superbee :85-94
      T ru = gdiv(g, y, x + tiny);
      T a = T(2) * ru; if (T(1) < a) a = T(1);
      T b = ru; if (T(2) < b) b = T(2);
      return max3<T>(T(0), a, b) * x;
    }
    case 6: {                                   // LIMO3 :98-110
      T ru = gdiv(g, y, x + tiny);
      T t = gdiv(g, T(2) + ru, T(3));
      T inner = min3<T>(T(2) * ru, t, T(1.6));
      T mid = max3<T>(T(-0.5) * ru, T(0), inner);
      T outer = t; if (mid < outer) outer = mid;
      T phi = T(0); if (outer > phi) phi = outer;
      return phi * x;
    }
    default: return T(0);                       // IORD :141-148
  }
}

template <typename T> __device__ __forceinline__ T limiter(int id, T x, T y) {
  return limiter(Gas<T>{}, id, x, y);
}

// ---- Riemann solvers — Lib_Riemann_HLL.f90 --------------------------------
// roe_averages :229-317, one species: h = cp*T, cp(T) = cp, or the table arithmetic.
template <typename T, int DV>
__device__ __forceinline__ void roe_averages(const Gas<T, DV>& g, T dltot, T drtot, T ul, T vl, T ur, T vr, T pl, T pr,
                                             T& h0l_o, T& h0r_o, T& u_roe, T& v_roe, T& a_roe) {
  const T dl = dltot, dr = drtot;
  T Rl = g.R, Rr = g.R;
  T inv_dltot = gdiv(g, T(1), dltot);
  T inv_drtot = gdiv(g, T(1), drtot);
  T Tl = gdiv(g, pl * inv_dltot, Rl);
  T Tr = gdiv(g, pr * inv_drtot, Rr);
  T srL = qsqrt<T>(dltot);
  T srR = qsqrt<T>(drtot);
  T inv_sr = gdiv(g, T(1), srL + srR);
  u_roe = (srR * ur + srL * ul) * inv_sr;
  v_roe = (srR * vr + srL * vl) * inv_sr;
  T R_roe = (srR * Rr + srL * Rl) * inv_sr;
  T T_roe = (srR * Tr + srL * Tl) * inv_sr;
  T invW = g.R;
  T hl, hr, cpsum;
  if (g.tab) {                                    // the table arithmetic of :257-292 (FP32: on the FP64 nodes)
    const int Til = tab_node(g.ntab, (double)Tl), Tir = tab_node(g.ntab, (double)Tr);
    const T dTl = Tl - (T)Til, dTr = Tr - (T)Tir;
    hl = tab_interp_t<T>(g.tab, g.ntab, kTabH, Til, dTl);
    hr = tab_interp_t<T>(g.tab, g.ntab, kTabH, Tir, dTr);
    const double* cp = g.tab + (long long)kTabCp * g.ntab;
    const T cl0 = (T)cp[Til], cl1 = (T)cp[Til + 1], cr0 = (T)cp[Tir], cr1 = (T)cp[Tir + 1];
    cpsum = ((cl0 + (cl1 - cl0) * dTl) + cr0) + (cr1 - cr0) * dTr;
  } else {
    hl = g.cp * Tl; hr = g.cp * Tr; cpsum = g.cp + g.cp;
  }
  T el = hl - invW * Tl;
  T er = hr - invW * Tr;
  T h0l = T(0) + hl * dl * inv_dltot;
  T h0r = T(0) + hr * dr * inv_drtot;
  T d_roe = (srR * dr * inv_drtot + srL * dl * inv_dltot) * inv_sr;
  T cv_roe = T(0) + d_roe * (T(0.5) * cpsum - invW);
  T sum_ei = T(0) + d_roe * (srR * er + srL * el) * inv_sr;
  h0l = h0l + T(0.5) * (ul * ul + vl * vl);
  h0r = h0r + T(0.5) * (ur * ur + vr * vr);
  T h0_roe = (srR * h0r + srL * h0l) * inv_sr;
  T gam_roe = T(1) + gdiv(g, R_roe, cv_roe);
  T vel2 = u_roe * u_roe + v_roe * v_roe;
  a_roe = qsqrt<T>((gam_roe - T(1)) * (h0_roe - T(0.5) * vel2 + cv_roe * T_roe - sum_ei));
  h0l_o = h0l; h0r_o = h0r;
}

// riemann_HLLC :34-123, riemann_HLLE :126-169 and riemann_HLLEM :355-432.  F = (F_r, F_u, F_v, F_E).
template <typename T, int RIE, int DV>
__device__ __forceinline__ void riemann(const Gas<T, DV>& g, T ul, T vl, T pl, T al, T dltot,
                                        T ur, T vr, T pr, T ar, T drtot, T nx, T ny, T F[4], T beta = T(1)) {
  T unl = ul * nx + vl * ny;
  T unr = ur * nx + vr * ny;
  T h0l, h0r, u_roe, v_roe, a_roe;
  roe_averages<T>(g, dltot, drtot, ul, vl, ur, vr, pl, pr, h0l, h0r, u_roe, v_roe, a_roe);
  T un_roe = u_roe * nx + v_roe * ny;
  if (RIE == 4) {                               // HLLC+ (riemann_HLLCSD :571-603): HLLC and HLLE blended by beta
    if (beta == T(0)) {
      riemann<T, 1>(g, ul, vl, pl, al, dltot, ur, vr, pr, ar, drtot, nx, ny, F);
    } else if (beta >= T(1)) {
      riemann<T, 0>(g, ul, vl, pl, al, dltot, ur, vr, pr, ar, drtot, nx, ny, F);
    } else {
      T FE[4], FC[4];
      riemann<T, 1>(g, ul, vl, pl, al, dltot, ur, vr, pr, ar, drtot, nx, ny, FE);
      riemann<T, 0>(g, ul, vl, pl, al, dltot, ur, vr, pr, ar, drtot, nx, ny, FC);
#pragma unroll
      for (int k = 0; k < 4; ++k) F[k] = beta * FC[k] + (T(1) - beta) * FE[k];
    }
    return;
  }
  if (RIE == 3) {                               // HLLE++ :435-530: Roe waves, speeds blended HLLE <-> Roe by beta
    const T srL = qsqrt<T>(dltot), srR = qsqrt<T>(drtot);
    const T rho_roe = srL * srR;
    const T h0_roe = (srR * h0r + srL * h0l) * gdiv(g, T(1), srL + srR);
    const T Sm = min3<T>(T(0), unl - al, un_roe - a_roe);
    const T Sp = max3<T>(T(0), unr + ar, un_roe + a_roe);
    const T inv_denom = gdiv(g, T(1), Sp - Sm);
    const T l1h = fabs((Sp + Sm) * un_roe * inv_denom - T(2) * Sp * Sm * inv_denom);
    const T l2h = fabs((Sp + Sm) * (un_roe - a_roe) * inv_denom - T(2) * Sp * Sm * inv_denom);
    const T l3h = fabs((Sp + Sm) * (un_roe + a_roe) * inv_denom - T(2) * Sp * Sm * inv_denom);
    const T eps = T(0.3) * (fabs(un_roe) + a_roe);
    auto hh = [&](T v) {                        // hartenHyman :320-331
      const T av = fabs(v);
      return av <= eps ? gdiv(g, v * v + eps * eps, T(2) * eps) : av;
    };
    const T l1r = fabs(un_roe), l2r = hh(un_roe - a_roe), l3r = hh(un_roe + a_roe);
    const T l1 = beta * l1r + (T(1) - beta) * l1h;
    const T l2 = beta * l2r + (T(1) - beta) * l2h;
    const T l3 = beta * l3r + (T(1) - beta) * l3h;
    auto phys = [&](T p, T r, T u, T v, T Fo[4]) {   // fluxes_helper :334-350
      Fo[0] = r * (u * nx + v * ny);
      Fo[1] = Fo[0] * u + p * nx;
      Fo[2] = Fo[0] * v + p * ny;
      const T vmag = qsqrt<T>(u * u + v * v);
      Fo[3] = Fo[0] * (enthalpy<T>(g, p, r) + T(0.5) * vmag * vmag);
    };
    T FL[4], FR[4];
    phys(pl, dltot, ul, vl, FL);
    phys(pr, drtot, ur, vr, FR);
    const T drho = drtot - dltot, dp = pr - pl, du = ur - ul, dv = vr - vl, dun = unr - unl;
    const T tx = -ny, ty = nx;
    const T alpha_m = gdiv(g, dp - rho_roe * a_roe * dun, T(2) * a_roe * a_roe);
    const T alpha_c = drho - gdiv(g, dp, a_roe * a_roe);
    const T alpha_s = rho_roe * (du * tx + dv * ty);
    const T alpha_p = gdiv(g, dp + rho_roe * a_roe * dun, T(2) * a_roe * a_roe);
    F[0] = T(0.5) * (FL[0] + FR[0]) - T(0.5) * (alpha_c * l1 + alpha_m * l2 + alpha_p * l3);
    F[1] = T(0.5) * (FL[1] + FR[1]) - T(0.5) * (alpha_c * l1 * u_roe + alpha_s * l1 * tx + alpha_m * l2 * (u_roe - a_roe * nx) +
                                                alpha_p * l3 * (u_roe + a_roe * nx));
    F[2] = T(0.5) * (FL[2] + FR[2]) - T(0.5) * (alpha_c * l1 * v_roe + alpha_s * l1 * ty + alpha_m * l2 * (v_roe - a_roe * ny) +
                                                alpha_p * l3 * (v_roe + a_roe * ny));
    F[3] = T(0.5) * (FL[3] + FR[3]) - T(0.5) * (alpha_c * l1 * T(0.5) * (u_roe * u_roe + v_roe * v_roe) +
                                                alpha_s * l1 * (u_roe * tx + v_roe * ty) +
                                                alpha_m * l2 * (h0_roe - a_roe * un_roe) +
                                                alpha_p * l3 * (h0_roe + a_roe * un_roe));
    return;
  }
  if (RIE == 2) {                               // HLLEM :355-432 (the detector's beta is not used)
    T S1 = min3<T>(T(0), unl - al, un_roe - a_roe);
    T S4 = max3<T>(T(0), unr + ar, un_roe + a_roe);
    // fluxes_helper :334-350, H0 = cp T + |V|^2 / 2 with |V| = sqrt(u^2 + v^2)
    auto phys = [&](T p, T r, T u, T v, T Fo[4]) {
      Fo[0] = r * (u * nx + v * ny);
      Fo[1] = Fo[0] * u + p * nx;
      Fo[2] = Fo[0] * v + p * ny;
      const T vmag = qsqrt<T>(u * u + v * v);
      Fo[3] = Fo[0] * (enthalpy<T>(g, p, r) + T(0.5) * vmag * vmag);
    };
    T FL[4], FR[4];
    phys(pl, dltot, ul, vl, FL);
    phys(pr, drtot, ur, vr, FR);
    T e0l = h0l - gdiv(g, pl, dltot);
    T e0r = h0r - gdiv(g, pr, drtot);
    T inv_denom = gdiv(g, T(1), S4 - S1);
    F[0] = (S4 * FL[0] - S1 * FR[0] + S1 * S4 * (drtot - dltot)) * inv_denom;
    F[1] = (S4 * FL[1] - S1 * FR[1] + S1 * S4 * (drtot * ur - dltot * ul)) * inv_denom;
    F[2] = (S4 * FL[2] - S1 * FR[2] + S1 * S4 * (drtot * vr - dltot * vl)) * inv_denom;
    F[3] = (S4 * FL[3] - S1 * FR[3] + S1 * S4 * (drtot * e0r - dltot * e0l)) * inv_denom;
    if (S1 < T(0) && S4 > T(0)) {                 // contact and shear restored in the subsonic range
      const T rho_roe = qsqrt<T>(dltot) * qsqrt<T>(drtot);
      const T drho = drtot - dltot, dp = pr - pl, du = ur - ul, dv = vr - vl;
      const T tx = -ny, ty = nx;
      const T alpha_c = drho - gdiv(g, dp, a_roe * a_roe);
      const T alpha_s = rho_roe * (du * tx + dv * ty);
      const T delta = gdiv(g, a_roe, a_roe + fabs(T(0.5) * (S1 + S4)));
      const T C = (S1 * S4) * inv_denom;
      F[0] = F[0] - C * delta * alpha_c;
      F[1] = F[1] - C * delta * (alpha_c * u_roe + alpha_s * tx);
      F[2] = F[2] - C * delta * (alpha_c * v_roe + alpha_s * ty);
      F[3] = F[3] - C * delta * (alpha_c * T(0.5) * (u_roe * u_roe + v_roe * v_roe) + alpha_s * (u_roe * tx + v_roe * ty));
    }
    return;
  }
  if (RIE == 1) {                               // HLLE
    T S1 = min3<T>(T(0), unl - al, un_roe - a_roe);
    T S4 = max3<T>(T(0), unr + ar, un_roe + a_roe);
    T inv_denom = gdiv(g, T(1), S4 - S1);
    T FrL = dltot * unl;
    T FrR = drtot * unr;
    T S1S4 = S1 * S4;
    F[0] = (S4 * FrL - S1 * FrR + S1S4 * (drtot - dltot)) * inv_denom;
    F[1] = (S4 * (FrL * ul + pl * nx) - S1 * (FrR * ur + pr * nx) + S1S4 * (drtot * ur - dltot * ul)) * inv_denom;
    F[2] = (S4 * (FrL * vl + pl * ny) - S1 * (FrR * vr + pr * ny) + S1S4 * (drtot * vr - dltot * vl)) * inv_denom;
    F[3] = (S4 * (FrL * h0l) - S1 * (FrR * h0r)
            + S1S4 * (drtot * (h0r - gdiv(g, pr, drtot)) - dltot * (h0l - gdiv(g, pl, dltot)))) * inv_denom;
    return;
  }
  T e0l = h0l - gdiv(g, pl, dltot);
  T e0r = h0r - gdiv(g, pr, drtot);
  T S1 = min3<T>(T(0), unl - al, un_roe - a_roe);
  T S4 = max3<T>(T(0), unr + ar, un_roe + a_roe);
  T inv_denom = gdiv(g, T(1), dltot * (S1 - unl) - drtot * (S4 - unr));
  T Sstar = (pr - pl + dltot * unl * (S1 - unl) - drtot * unr * (S4 - unr)) * inv_denom;
  T pstar = dltot * (unl - S1) * (unl - Sstar) + pl;
  // The two star branches share one code path: side s = L when Sstar >= 0, else R.
  bool left = (S1 >= T(0)) || (Sstar >= T(0));
  bool star = (S1 < T(0)) && ((Sstar >= T(0)) || (S4 > T(0)));
  T Sk = left ? S1 : S4;
  T dk = left ? dltot : drtot;
  T uk = left ? ul : ur;
  T vk = left ? vl : vr;
  T pk = left ? pl : pr;
  T unk = left ? unl : unr;
  T h0k = left ? h0l : h0r;
  T e0k = left ? e0l : e0r;
  T F_r = dk * unk;
  T F_u = F_r * uk + pk * nx;
  T F_v = F_r * vk + pk * ny;
  T F_E = F_r * h0k;
  if (star) {
    T inv_d = gdiv(g, T(1), Sk - Sstar);
    T U1S = dk * (Sk - unk) * inv_d;
    T U2S = ((Sk - unk) * dk * uk + (pstar - pk) * nx) * inv_d;
    T U3S = ((Sk - unk) * dk * vk + (pstar - pk) * ny) * inv_d;
    T U5S = ((Sk - unk) * dk * e0k - pk * unk + pstar * Sstar) * inv_d;
    F_r = F_r + Sk * (U1S - dk);
    F_u = F_u + Sk * (U2S - dk * uk);
    F_v = F_v + Sk * (U3S - dk * vk);
    F_E = F_E + Sk * (U5S - dk * e0k);
  }
  F[0] = F_r; F[1] = F_u; F[2] = F_v; F[3] = F_E;
}

// reciprocal used by the mixed form's FP64 update (one division, then products)
__device__ __forceinline__ double rcp64(double a) { return 1.0 / a; }

// ---- conservative <-> primitive (FP64 always) ------------------------------
template <int DV>
__device__ __forceinline__ void prim2cons(const Gas<double, DV>& g, double rho, double u, double v, double p, double U[4]) {
  double T = gdiv(g, p, rho * g.R);
  U[0] = rho;
  U[1] = rho * u;
  U[2] = rho * v;
  U[3] = rho * (int_energy(g, T) + 0.5 * (u * u + v * v));
}
// T0: the temperature estimate of the table mode's inversion (the closed form does not read it)
template <int DV>
__device__ __forceinline__ void cons2prim(const Gas<double, DV>& g, const double U[4], double q[4], double T0) {
  double rho = U[0];
  double vx = gdiv(g, U[1], rho), vy = gdiv(g, U[2], rho);
  double e = gdiv(g, U[3], rho) - 0.5 * (vx * vx + vy * vy);
  double T = g.tab ? tab_t_of_e(g, e, T0) : gdiv(g, e, g.cv);
  q[0] = rho; q[1] = vx; q[2] = vy; q[3] = rho * g.R * T;
}

// ---- mixed form: FP32 face arithmetic --------------------------------------
// The mixed-precision form evaluates reconstruction, limiter and Riemann flux in FP32 from FP32
// geometry.  It is not held to the reference's operation order (it is checked against the FP64
// form within a stated tolerance), so these functions use the closed forms the single-species
// constant-cp closure reduces to, fused multiply-adds, and the hardware reciprocal, reciprocal
// square root and square root.
struct GasF { float gam, gm1, gam_gm1; };          // gamma, gamma - 1, gamma / (gamma - 1) = cp / R

__device__ __forceinline__ float frcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ float frsq(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

// rlimiter(x, y) — Lib_Limiters.f90:26-112.  van Leer in its product form: for r = y/x > 0,
// x (r + |r|) / (1 + |r|) = 2 x y / (x + y); 0 otherwise.  The others as in limiter<float>.
__device__ __forceinline__ float limiter_mx(int id, float x, float y) {
#pragma clang fp contract(on)
  if (id == 1) {                                          // minmod
    if (x * y <= 0.f) return 0.f;
    return __builtin_fabsf(x) > __builtin_fabsf(y) ? y : x;
  }
  if (id == 2) {                                          // van Leer
    const float xy = x * y;
    return xy > 0.f ? 2.f * xy * frcp(x + y) : 0.f;
  }
  return limiter<float>(id, x, y);
}

// The same with the limiter fixed at compile time (no branch on the limiter id).
template <int LIM>
__device__ __forceinline__ float limiter_mxt(float x, float y) {
#pragma clang fp contract(on)
  if constexpr (LIM == 1) {                               // minmod
    if (x * y <= 0.f) return 0.f;
    return __builtin_fabsf(x) > __builtin_fabsf(y) ? y : x;
  } else if constexpr (LIM == 2) {                        // van Leer
    const float xy = x * y;
    return xy > 0.f ? 2.f * xy * frcp(x + y) : 0.f;
  } else {
    return limiter<float>(LIM, x, y);
  }
}

// The same fluxes from the face's length-scaled normal (nx, ny) = int_l * n: the flux through the face
// times its length is homogeneous of degree one in the normal once the sound speeds are scaled by
// int_l = |(nx, ny)| and the HLLC star pressure is formed from unit-normal speeds (the 1/int_l^2
// factor).  Used by the mixed form, whose geometry stores only the scaled normal (8 B per face).
template <int RIE>
__device__ __forceinline__ void flux_mxs(const GasF& g, const float L[4], const float R[4], float nx, float ny,
                                         float F[4]) {
#pragma clang fp contract(on)
  const float l2 = nx * nx + ny * ny;
  const float len = fsqrt(l2);                            // int_l
  const float rsL = frsq(L[0]), rsR = frsq(R[0]);         // 1 / sqrt(rho)
  const float sL = L[0] * rsL, sR = R[0] * rsR;           // sqrt(rho)
  const float irL = rsL * rsL, irR = rsR * rsR;           // 1 / rho
  const float aL = fsqrt(g.gam * L[3] * irL) * len, aR = fsqrt(g.gam * R[3] * irR) * len;
  const float unL = L[1] * nx + L[2] * ny, unR = R[1] * nx + R[2] * ny;
  const float hL = g.gam_gm1 * L[3] * irL + 0.5f * (L[1] * L[1] + L[2] * L[2]);   // h0 = cp T + |u|^2 / 2
  const float hR = g.gam_gm1 * R[3] * irR + 0.5f * (R[1] * R[1] + R[2] * R[2]);
  const float w = frcp(sL + sR);
  const float ut = (sL * L[1] + sR * R[1]) * w, vt = (sL * L[2] + sR * R[2]) * w;
  const float ht = (sL * hL + sR * hR) * w;
  const float at = fsqrt(g.gm1 * (ht - 0.5f * (ut * ut + vt * vt))) * len;
  const float unt = ut * nx + vt * ny;
  const float S1 = __builtin_fminf(__builtin_fminf(0.f, unL - aL), unt - at);
  const float S4 = __builtin_fmaxf(__builtin_fmaxf(0.f, unR + aR), unt + at);
  const float FrL = L[0] * unL, FrR = R[0] * unR;
  if (RIE == 1) {
    const float inv = frcp(S4 - S1), S14 = S1 * S4;
    F[0] = (S4 * FrL - S1 * FrR + S14 * (R[0] - L[0])) * inv;
    F[1] = (S4 * (FrL * L[1] + L[3] * nx) - S1 * (FrR * R[1] + R[3] * nx) + S14 * (R[0] * R[1] - L[0] * L[1])) * inv;
    F[2] = (S4 * (FrL * L[2] + L[3] * ny) - S1 * (FrR * R[2] + R[3] * ny) + S14 * (R[0] * R[2] - L[0] * L[2])) * inv;
    F[3] = (S4 * FrL * hL - S1 * FrR * hR + S14 * ((R[0] * hR - R[3]) - (L[0] * hL - L[3]))) * inv;
  } else {
    const float mL = L[0] * (S1 - unL), mR = R[0] * (S4 - unR);
    const float Sst = (l2 * (R[3] - L[3]) + mL * unL - mR * unR) * frcp(mL - mR);
    const float pst = L[0] * (unL - S1) * (unL - Sst) * frcp(l2) + L[3];
    const bool left = (S1 >= 0.f) || (Sst >= 0.f);
    const bool star = (S1 < 0.f) && ((Sst >= 0.f) || (S4 > 0.f));
    const float Sk = left ? S1 : S4, dk = left ? L[0] : R[0], uk = left ? L[1] : R[1], vk = left ? L[2] : R[2];
    const float pk = left ? L[3] : R[3], unk = left ? unL : unR, hk = left ? hL : hR, irk = left ? irL : irR;
    const float Fr = dk * unk;
    const float Fu = Fr * uk + pk * nx, Fv = Fr * vk + pk * ny, FE = Fr * hk;
    const float id = frcp(Sk - Sst);                      // star-state correction, selected (see flux_mx)
    const float U1 = dk * (Sk - unk) * id;
    const float dp = (pst - pk) * id;
    const float e0k = hk - pk * irk;
    F[0] = star ? Fr + Sk * (U1 - dk) : Fr;
    F[1] = star ? Fu + Sk * (U1 * uk + dp * nx - dk * uk) : Fu;
    F[2] = star ? Fv + Sk * (U1 * vk + dp * ny - dk * vk) : Fv;
    F[3] = star ? FE + Sk * (U1 * e0k + (pst * Sst - pk * unk) * id - dk * e0k) : FE;
  }
}

// HLLC (RIE 0, Batten et al. 1997 as in riemann_HLLC, Lib_Riemann_HLL.f90:34-123) or HLLE (RIE 1,
// Einfeldt, riemann_HLLE :126-169) flux through a face of unit normal (nx, ny), times its length il,
// from primitive states L, R = (rho, u, v, p).  Wave speeds S1 = min(0, unL - aL, un~ - a~),
// S4 = max(0, unR + aR, un~ + a~) with Roe averages (roe_averages :229-317, which for one species
// of constant cp reduces to a~^2 = (gamma - 1) (H~ - |u~|^2 / 2)).
template <int RIE>
__device__ __forceinline__ void flux_mx(const GasF& g, const float L[4], const float R[4], float nx, float ny,
                                        float il, float F[4]) {
#pragma clang fp contract(on)
  const float rsL = frsq(L[0]), rsR = frsq(R[0]);         // 1 / sqrt(rho)
  const float sL = L[0] * rsL, sR = R[0] * rsR;           // sqrt(rho)
  const float irL = rsL * rsL, irR = rsR * rsR;           // 1 / rho
  const float aL = fsqrt(g.gam * L[3] * irL), aR = fsqrt(g.gam * R[3] * irR);
  const float unL = L[1] * nx + L[2] * ny, unR = R[1] * nx + R[2] * ny;
  const float hL = g.gam_gm1 * L[3] * irL + 0.5f * (L[1] * L[1] + L[2] * L[2]);   // h0 = cp T + |u|^2 / 2
  const float hR = g.gam_gm1 * R[3] * irR + 0.5f * (R[1] * R[1] + R[2] * R[2]);
  const float w = frcp(sL + sR);
  const float ut = (sL * L[1] + sR * R[1]) * w, vt = (sL * L[2] + sR * R[2]) * w;
  const float ht = (sL * hL + sR * hR) * w;
  const float at = fsqrt(g.gm1 * (ht - 0.5f * (ut * ut + vt * vt)));
  const float unt = ut * nx + vt * ny;
  const float S1 = __builtin_fminf(__builtin_fminf(0.f, unL - aL), unt - at);
  const float S4 = __builtin_fmaxf(__builtin_fmaxf(0.f, unR + aR), unt + at);
  const float FrL = L[0] * unL, FrR = R[0] * unR;
  if (RIE == 1) {
    const float inv = frcp(S4 - S1), S14 = S1 * S4;
    F[0] = (S4 * FrL - S1 * FrR + S14 * (R[0] - L[0])) * inv;
    F[1] = (S4 * (FrL * L[1] + L[3] * nx) - S1 * (FrR * R[1] + R[3] * nx) + S14 * (R[0] * R[1] - L[0] * L[1])) * inv;
    F[2] = (S4 * (FrL * L[2] + L[3] * ny) - S1 * (FrR * R[2] + R[3] * ny) + S14 * (R[0] * R[2] - L[0] * L[2])) * inv;
    F[3] = (S4 * FrL * hL - S1 * FrR * hR + S14 * ((R[0] * hR - R[3]) - (L[0] * hL - L[3]))) * inv;
  } else {
    const float mL = L[0] * (S1 - unL), mR = R[0] * (S4 - unR);
    const float Sst = (R[3] - L[3] + mL * unL - mR * unR) * frcp(mL - mR);
    const float pst = L[0] * (unL - S1) * (unL - Sst) + L[3];
    const bool left = (S1 >= 0.f) || (Sst >= 0.f);
    const bool star = (S1 < 0.f) && ((Sst >= 0.f) || (S4 > 0.f));
    const float Sk = left ? S1 : S4, dk = left ? L[0] : R[0], uk = left ? L[1] : R[1], vk = left ? L[2] : R[2];
    const float pk = left ? L[3] : R[3], unk = left ? unL : unR, hk = left ? hL : hR, irk = left ? irL : irR;
    const float Fr = dk * unk;
    const float Fu = Fr * uk + pk * nx, Fv = Fr * vk + pk * ny, FE = Fr * hk;
    // star-state correction, evaluated on every lane and selected (no divergent branch; where it is
    // not selected its operands may be non-finite)
    const float id = frcp(Sk - Sst);
    const float U1 = dk * (Sk - unk) * id;                // star-state density
    const float dp = (pst - pk) * id;
    const float e0k = hk - pk * irk;                      // total energy per unit mass
    F[0] = star ? Fr + Sk * (U1 - dk) : Fr;
    F[1] = star ? Fu + Sk * (U1 * uk + dp * nx - dk * uk) : Fu;
    F[2] = star ? Fv + Sk * (U1 * vk + dp * ny - dk * vk) : Fv;
    F[3] = star ? FE + Sk * (U1 * e0k + (pst * Sst - pk * unk) * id - dk * e0k) : FE;
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) F[v] = F[v] * il;
}

// FP64 reciprocal: hardware estimate refined by two Newton steps (within an ulp of 1/a).
__device__ __forceinline__ double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = __builtin_fma(-a, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-a, r, 1.0);
  return __builtin_fma(r, e, r);
}

}  // namespace q2d
