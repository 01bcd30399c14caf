// q2d_kernels.hip — MI355X (gfx950) kernels and C-ABI of the explicit FV step.
//
// Device layout (per block): every per-cell or per-face field is a plane of
// (nj+4) rows x pitch doubles, i fastest; cell (i,j) of the Fortran arrays
// (i,j in 1-gc..n+gc) lives at (j+1)*pitch + (i+1).  The i-face between cells
// i and i+1 (dir(1)%f(i,j)) is stored at the position of cell i, the j-face
// between rows j and j+1 (dir(2)%f(i,j)) at the position of cell (i,j).
// State is SoA (rho,u,v,p planes) in three rotating buffers so that no RK
// stage reads what it writes; Copy_State disappears.
//
// Work decomposition of the stage kernel: one wavefront owns a strip of 60
// interior columns (lanes 2..61; lanes 0,1,62,63 are the i-halo) over a
// segment of rows, and marches up the rows.  The j-direction stencil is kept
// in registers (each row is loaded once), the i-direction stencil is shared
// across lanes with wavefront shuffles.  Each face flux is evaluated once, and
// each cell's limited slope is evaluated once per direction and shared by the
// two faces that use it (the reference evaluates it twice with identical
// operands, so the value is bit-identical).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/q2d_gpu.h"
#include "q2d_device.h"

using namespace q2d;

// Step scratch of q2d_gpu_step_noalloc: a file-scope device array, so that entry point allocates nothing.
// Each GPU the library runs on holds its own copy, at its own address, until the process ends; host code
// uses only the address hipGetSymbolAddress returns for the current device (never the array's name).
static __device__ __align__(256) unsigned char scratch[Q2D_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kWave = 64;
constexpr int kOwn = 60;          // owned columns per strip
constexpr int kMaxCases = 1 << 20;
constexpr int kTimedSteps = 64;                         // timed steps between two reads of the slots
constexpr int kLaunchesPerStep = 3 * 4;                 // stage launches per step (3 stages x 4 variants)
constexpr int kTimedSlots = kLaunchesPerStep * kTimedSteps;
constexpr int kGraphSteps = 8;                          // steps per replay of the multi-step graph
constexpr int kStampLanes = 64;                         // sub-slots per launch slot
// Rows of state the stage kernel's loads run ahead of the row being computed (1 or 2), per form: the
// FP64 form's longer rows leave room for a second row in flight, measured 3 % faster.
#ifndef Q2D_PREFETCH_ROWS_FP64
#define Q2D_PREFETCH_ROWS_FP64 2
#endif
#ifndef Q2D_PREFETCH_ROWS_MIXED
#define Q2D_PREFETCH_ROWS_MIXED 1
#endif
// Fuse the next step's Compute_dt into the last RK stage of Euler steps (saves the separate dt pass
// over the state; the first step after create, upload or add_cases runs the dt pass).  Measured on
// 4 x 1024^2, SSP-RK3: FP64 1.115 -> 1.040 ms per step.
#ifndef Q2D_FUSE_DT_FP64
#define Q2D_FUSE_DT_FP64 1
#endif
constexpr bool kFuseDtFp64 = Q2D_FUSE_DT_FP64 != 0;
// Fused next-step dt in the mixed-form Euler kernel (k_stage_mx): 0.625 -> 0.554 ms per step.
#ifndef Q2D_FUSE_DT_MX
#define Q2D_FUSE_DT_MX 1
#endif
constexpr bool kFuseDtMx = Q2D_FUSE_DT_MX != 0;
#ifndef Q2D_STAGE_MIN_WAVES
#define Q2D_STAGE_MIN_WAVES 1
#endif
// FP64 stage kernel: each work item runs a fast-division pass (ddiv_fast, with checks) first and the
// IEEE division where a check fails.  Measured on 4 x 1024^2, SSP-RK3: 279 -> 240 us per launch when
// every item passes its checks.
#ifndef Q2D_FASTDIV
#define Q2D_FASTDIV 1
#endif

// ---------------------------------------------------------------------------
// device-side descriptors
// ---------------------------------------------------------------------------
struct DevCase {
  int n_rk, limiter, riemann, time_accurate;
  int model, source, precision, ramp_iter;
  int ndir, itermax, item_begin, item_end;
  double cfl, vnn, R, cp, cv, gam;
  double res_threshold, time_threshold;
  double mu_ref, T_ref, S_suth, Pr;     // Sutherland viscosity, Prandtl number (model 1)
  double res_drop;                      // > 0: stop when residual(1) <= res_drop * residual(1) of step 1
  double irs_beta;                      // > 0: implicit residual smoothing (k_irs)
  int irs_begin, irs_end;               // the case's k_irs work items (their partials replace the stage's)
  int space_sd;                         // MUSCL-SD: Tramel's detector scales the slopes (sd_tramel)
  int ntab;                             // table mode: nodes per property plane (0: closed forms)
  int fvisc;                            // mixed laminar case whose stage kernel evaluates the viscous terms
  int mxk;                              // the case's items run k_stage_mx (mixed form: plain MUSCL, HLLC or
                                        // HLLE; laminar ones with fvisc and no residual smoothing)
  int dtmx;                             // its dt is cell_dt_mx's, from FP32 geometry (mixed Euler, mxk laminar)
  int pad_mx;
  const double* tab;                    // table mode: h, cp, mu, k planes (device copy), else null
};

struct DevCtrl {
  unsigned long long dtmin_bits;
  unsigned long long dtmin_next_bits;   // next step's dt, reduced by the last RK stage
  unsigned long long fail_key;
  double dtglobal, time;
  double residual[4];
  int iter, iter_from_call, todo, active;
  int cur, status, pad_ctrl, dtsel;
  double res_first;                     // residual(1) of the case's first step
};

struct DevBlock {
  int ni, nj, pitch, case_id;
  int blk_in_case;
  int fastdiv;                   // FP64: geometry within the fast-division range (k_stage's fast pass)
  int bcf[4];                    // offset of the per-cell BC-record index list of faces 1..4 (-1: none)
  int wallf[4];                  // faces 1..4: 0 no wall record (301-304), 1 every record a wall, 2 mixed
  long long plane;               // doubles per plane
  double* P[3];                  // three state buffers, 4 planes each
  const double *dlx, *dly, *sur;
  const double *fix, *fiy, *fil; // i-face N(1), N(2), int_l
  const double *fjx, *fjy, *fjl; // j-face
  const double *m11, *m21, *m12, *m22;
  const double *sa, *sb;
  double* dtl[2];                // local dt, double-buffered: [dtsel] holds the current step's dt
  const double* nodr;            // node(i,j)%c(2) at cell (i,j), i = 0..ni, j = 0..nj (viscous cases)
  double *vdx, *vey, *vsv;       // 3 planes each: diffusive flux of i-face (x|x+1), of j-face (r|r+1), and
                                 // the axisymmetric viscous source, components rho u, rho v, rho E
  const float4* mgn;             // mixed-form Euler blocks: FP32 geometry (see k_stage_mx)
  const float2* mgs;
  const float4* mgd;
  const float* mhy;
  const float2* msrc;
  const float4* mmm;             // mixed-form laminar blocks: M (c11, c21, c12, c22) in FP32 and the node
  const float* mnod;             // radius node(i,j)%c(2) in FP32, for the viscous terms of k_stage_mx
  double* irs[1];                // IRS cases: the stage's scaled residual (4 planes)
};

struct DevBC {
  int type, blk, i, j;           // blk: global block index
  int f, dblk, is, js;           // donor (101/201)
  int fs, pad0, cid, dpos;       // pad0: d11 of a connection (the donor's orientation along the face);
                                 // cid: global case index of the record's block; dpos: index of the
                                 // record's boundary cell in the per-face arrays (DevBlock::bcf + position)
  double mach, T, p, rel_fac, cos_a, sin_a;
};

struct Item { int blk, x0, y0, y1; };   // rows y0..y1; k_stage_mx items with y0 > y1 march down from row y0
// Per-item results of a step's last stage: residual-norm partials (4) and the fused next-step dt (1).
constexpr int kPartials = 5;

// Laminar transport of one species: Sutherland viscosity and Fourier conductivity k = mu cp / Pr, the
// closed form behind the reference's f_laminarViscosity / co_k_mi_lam_Wilke calls (FP64 in both forms),
// or in table mode the mu and k planes at T (Wilke's rule of one species is the species' value).
__device__ __forceinline__ double mu_lam(const DevCase& C, double T) {
  if (C.tab) return tab_at(C.tab, C.ntab, kTabMu, T);
  const double x = T / C.T_ref;
  return C.mu_ref * (x * sqrt(x)) * (C.T_ref + C.S_suth) / (T + C.S_suth);
}
__device__ __forceinline__ double k_lam(const DevCase& C, double mu, double T) {
  if (C.tab) return tab_at(C.tab, C.ntab, kTabK, T);
  return mu * C.cp / C.Pr;
}
// the case's FP64 gas closure (with its tables in table mode)
__device__ __forceinline__ Gas<double> case_gas(const DevCase& C) {
  Gas<double> g{C.R, C.cp, C.cv, C.gam};
  g.tab = C.tab; g.ntab = C.ntab;
  return g;
}

__device__ __forceinline__ int stage_in(int cur, int irk) { return irk == 1 ? cur : (cur + 1 + ((irk - 2) & 1)) % 3; }
__device__ __forceinline__ int stage_out(int cur, int irk) { return (cur + 1 + ((irk - 1) & 1)) % 3; }

// Global-address-space views of device pointers: loads/stores compile to global_* instead of flat_*
// (flat accesses also count against the LDS counter, which would make every cross-lane shuffle wait
// for outstanding HBM loads).
typedef const __attribute__((address_space(1))) double* GP;
typedef __attribute__((address_space(1))) double* GPW;
__device__ __forceinline__ GP gp(const double* p) { return (GP)p; }
__device__ __forceinline__ GPW gpw(double* p) { return (GPW)p; }

// Neighbour-lane exchange within the wavefront.  With Q2D_DPP the wave-wide DPP shifts of GFX9
// (wave_shl:1 = 0x130: lane l reads lane l+1; wave_shr:1 = 0x138: lane l reads lane l-1) move the
// data inside the VALU with no LDS round trip; the edge lanes read their own value (unused).
#ifndef Q2D_DPP
#define Q2D_DPP 1
#endif
#if Q2D_DPP
__device__ __forceinline__ int dpp_dn(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_up(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ double shdn(double v) {
  long long b = __double_as_longlong(v);
  int lo = dpp_dn((int)(b & 0xffffffffll)), hi = dpp_dn((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double shup(double v) {
  long long b = __double_as_longlong(v);
  int lo = dpp_up((int)(b & 0xffffffffll)), hi = dpp_up((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float shdn(float v) { return __int_as_float(dpp_dn(__float_as_int(v))); }
__device__ __forceinline__ float shup(float v) { return __int_as_float(dpp_up(__float_as_int(v))); }
#else
__device__ __forceinline__ double shdn(double v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ double shup(double v) { return __shfl_up(v, 1, kWave); }
__device__ __forceinline__ float shdn(float v) { return __shfl_down(v, 1, kWave); }
__device__ __forceinline__ float shup(float v) { return __shfl_up(v, 1, kWave); }
#endif

// Per-row data the stage kernel needs for the row it finalises.
struct RowGeo { double dlx, fix, fiy, fil, fjx, fjy, fjl, sur, dtl, po[4], mm[4]; };
__device__ __forceinline__ void load_geo(RowGeo& g, long long o, GP dlx, GP fix, GP fiy, GP fil, GP fjx, GP fjy,
                                         GP fjl, GP sur, GP dtl, GP PO, long long pl, bool want_dtl, GP m11,
                                         bool want_m) {
  g.dlx = dlx[o]; g.fix = fix[o]; g.fiy = fiy[o]; g.fil = fil[o];
  g.fjx = fjx[o]; g.fjy = fjy[o]; g.fjl = fjl[o]; g.sur = sur[o];
  g.dtl = want_dtl ? dtl[o] : 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) g.po[v] = PO[v * pl + o];
  (void)m11; (void)want_m;
}

// Order key of a failure: RK stage (1..3), then kind (0 interface state, 1 new state), block (1..255),
// row, column (1..60000), compared as one integer.  A failing case is frozen at the end of its step
// (k_finish), so only failures within one step are ever compared and the stage within the step orders
// them; create / add_cases reject blocks and cases beyond the field widths.
__device__ __forceinline__ unsigned long long fail_key(int stage, int kind, int b, int j, int i) {
  return ((unsigned long long)(stage & 0x3f) << 42) | ((unsigned long long)kind << 41) |
         ((unsigned long long)(b & 0xff) << 32) | ((unsigned long long)(j & 0xffff) << 16) | (unsigned long long)(i & 0xffff);
}

__device__ __forceinline__ void record_fail(DevCtrl* c, int code, unsigned long long key) {
  atomicMin(&c->fail_key, key);
  atomicMin(&c->status, code);   // codes are negative; min keeps a failure
}

// ---------------------------------------------------------------------------
// step bookkeeping kernels (one thread per case)
// ---------------------------------------------------------------------------
// use_next = 1: the previous step's last stage already reduced this step's dt (from the same
// state Compute_dt reads); use_next = 0: k_dt computes it (first step, or after an upload).
// With use_next the step's dt is already known, so this kernel also finishes it (k_dt_finish's work).
__device__ __forceinline__ void begin_case(DevCtrl* ctrl, const DevCase* cases, int c, int use_next) {
  DevCtrl& k = ctrl[c];
  if (!k.active) return;
  k.iter += 1;                                // Mod_Explicit.f90:38-40
  k.iter_from_call += 1;
  const unsigned long long b1e5 = (unsigned long long)__double_as_longlong(1e5);   // :49
  k.dtmin_bits = use_next ? (k.dtmin_next_bits < b1e5 ? k.dtmin_next_bits : b1e5) : b1e5;
  if (use_next) k.dtsel ^= 1;                 // the plane the last stage wrote becomes current
  k.dtmin_next_bits = b1e5;
  if (use_next) {
    k.dtglobal = __longlong_as_double((long long)k.dtmin_bits);   // Mod_dt.f90:54-55
    if (cases[c].time_accurate) k.time = k.time + k.dtglobal;     // Mod_Explicit.f90:51-54
  }
}

__global__ void k_begin(DevCtrl* ctrl, const DevCase* cases, int ncases, int use_next) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncases) begin_case(ctrl, cases, c, use_next);
}

__global__ void k_dt_finish(DevCtrl* ctrl, const DevCase* cases, int ncases) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncases) return;
  DevCtrl& k = ctrl[c];
  if (!k.active) return;
  k.dtglobal = __longlong_as_double((long long)k.dtmin_bits);     // Mod_dt.f90:54-55
  if (cases[c].time_accurate) k.time = k.time + k.dtglobal;      // Mod_Explicit.f90:51-54
}

// ---------------------------------------------------------------------------
// Compute_dt — Mod_dt.f90:10-104 (one wavefront per work item)
// ---------------------------------------------------------------------------
// The VNN limit of Compute_dt (Mod_dt.f90:88-102; laminar: mie = 0), after the convective limits.  The
// dt pass and the fused dt of the stage kernels' last stage both apply exactly this.
__device__ __forceinline__ double vnn_limit(const DevCase& C, double dtmin, double rho, double p, double dlx,
                                            double dly) {
  const double mil = mu_lam(C, p / (rho * C.R));
  const double mi = 0.0 + mil;
  double dt = rho * (dlx * dlx) * C.vnn / mi;
  if (dt < dtmin) dtmin = dt;
  if (C.ndir == 2) {
    dt = rho * (dly * dly) * C.vnn / mi;
    if (dt < dtmin) dtmin = dt;
  }
  return dtmin;
}

__device__ __forceinline__ double cell_dt(const DevCase& C, const DevCtrl& K, double rho, double u, double v, double p,
                                          double m11, double m21, double m12, double m22, double dlx, double dly) {
  const double snd = sound<double>(case_gas(C), p, rho);
  double dtmin = 1e8;
  {
    const double nrm = sqrt(m11 * m11 + m12 * m12);
    const double v1 = m11 / nrm, v2 = m12 / nrm;
    const double lambda = fabs(u * v1 + v * v2) + snd;
    const double dt = dlx / lambda * C.cfl;
    if (dt < dtmin) dtmin = dt;
  }
  if (C.ndir == 2) {
    const double nrm = sqrt(m21 * m21 + m22 * m22);
    const double v1 = m21 / nrm, v2 = m22 / nrm;
    const double lambda = fabs(u * v1 + v * v2) + snd;
    const double dt = dly / lambda * C.cfl;
    if (dt < dtmin) dtmin = dt;
  }
  if (C.model > 0) dtmin = vnn_limit(C, dtmin, rho, p, dlx, dly);
  if (K.iter < C.ramp_iter) dtmin = dtmin * (double)K.iter / (double)C.ramp_iter;   // Mod_dt.f90:43
  return dtmin;
}

// Compute_dt of a cell of a k_stage_mx case, from its FP32 geometry: half lengths hx, hy and the
// normalised rows of M (gd); FP64 arithmetic, and for laminar cases (VISC) the VNN limit with the lengths
// 2 hx, 2 hy.  The dt pass and the fused dt of k_stage_mx's last stage both evaluate exactly this, so a
// step may take its dt from either.  iter: the iteration being started.
template <bool VISC>
__device__ __forceinline__ double cell_dt_mx(const DevCase& C, int iter, double rho, double u, double v, double p,
                                             float hx, float hy, float g1x, float g1y, float g2x, float g2y) {
  const double snd = sqrt(C.gam * p * rcp_nr(rho));
  double dtmin = 1e8;
  {
    const double lambda = fabs(u * (double)g1x + v * (double)g1y) + snd;
    const double dtd = 2.0 * (double)hx * rcp_nr(lambda) * C.cfl;
    if (dtd < dtmin) dtmin = dtd;
  }
  if (C.ndir == 2) {
    const double lambda = fabs(u * (double)g2x + v * (double)g2y) + snd;
    const double dtd = 2.0 * (double)hy * rcp_nr(lambda) * C.cfl;
    if (dtd < dtmin) dtmin = dtd;
  }
  if (VISC && C.model > 0) dtmin = vnn_limit(C, dtmin, rho, p, 2.0 * (double)hx, 2.0 * (double)hy);
  if (iter < C.ramp_iter) dtmin = dtmin * (double)iter / (double)C.ramp_iter;   // Mod_dt.f90:43
  return dtmin;
}

// Work unit of the dt pass: `count` (<= kDtCells) consecutive cells of one block, in row-major order from
// the block's interior cell `start`.  One wavefront per unit: the block, case and control data are
// wave-uniform (scalar loads), the lanes stream consecutive cells (coalesced row segments), four cells per
// lane in flight, and the unit's min is published with one atomic.
struct DtRange { int blk, start, count, ni; };
constexpr int kDtCells = 1024;

__global__ __launch_bounds__(256) void k_dt(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                            DevCtrl* __restrict__ ctrl, const DtRange* __restrict__ units,
                                            int nunits, int unused) {
  (void)unused;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpg = blockDim.x / kWave;
  const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpg + threadIdx.x / kWave);
  const int nwaves = gridDim.x * wpg;
  for (int u = wave0; u < nunits; u += nwaves) {
    const DtRange U = units[u];
    const DevBlock& B = blocks[U.blk];
    DevCtrl& K = ctrl[B.case_id];
    if (!K.active) continue;
    const DevCase& C = cases[B.case_id];
    const long long pl = B.plane;
    const int pitch = B.pitch;
    GP P = gp(B.P[K.cur]);
    GP m11 = gp(B.m11), m21 = gp(B.m21), m12 = gp(B.m12), m22 = gp(B.m22), dlx = gp(B.dlx), dly = gp(B.dly);
    GPW dtl = gpw(B.dtl[K.dtsel]);
    const bool local = !C.time_accurate;
    double dmin = 1e300;
    for (int c0 = 0; c0 < U.count; c0 += 4 * kWave) {
      long long o[4];
      bool ok[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = c0 + k * kWave + lane;
        ok[k] = c < U.count;
        const int cell = U.start + (ok[k] ? c : 0);
        const int j = cell / U.ni, i = cell - j * U.ni;
        o[k] = (long long)(j + 2) * pitch + (i + 2);
      }
      double d[4];
      if (C.dtmx) {                                  // a k_stage_mx case: that kernel's fused dt formula
        const __attribute__((address_space(1))) float* gd = (const __attribute__((address_space(1))) float*)B.mgd;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          d[k] = cell_dt_mx<true>(C, K.iter, P[o[k]], P[pl + o[k]], P[2 * pl + o[k]], P[3 * pl + o[k]],
                            (float)(0.5 * dlx[o[k]]), (float)(0.5 * dly[o[k]]), gd[4 * o[k]], gd[4 * o[k] + 1],
                            gd[4 * o[k] + 2], gd[4 * o[k] + 3]);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          d[k] = cell_dt(C, K, P[o[k]], P[pl + o[k]], P[2 * pl + o[k]], P[3 * pl + o[k]], m11[o[k]], m21[o[k]],
                         m12[o[k]], m22[o[k]], dlx[o[k]], dly[o[k]]);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!ok[k]) continue;
        if (local) dtl[o[k]] = d[k];
        if (d[k] < dmin) dmin = d[k];
      }
    }
    // wavefront min (order-independent), then one atomic per unit
    for (int sft = 32; sft >= 1; sft >>= 1) {
      const double q = __shfl_xor(dmin, sft, kWave);
      if (q < dmin) dmin = q;
    }
    if (lane == 0 && dmin < 1e300) atomicMin(&K.dtmin_bits, (unsigned long long)__double_as_longlong(dmin));
  }
}

// ---------------------------------------------------------------------------
// Fill_Ghost_Cell — Lib_Ghost.f90:16-140 (one thread per boundary record)
// ---------------------------------------------------------------------------
__constant__ int c_guide[5][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}};   // Parameters_m.f90:13-18

// Ghost layers of one boundary record.  A record reads interior cells (its own block's or the donor's)
// and writes only its own ghost cells, so all records can be filled concurrently.
// For a connection (101/201) the two ghost layers written are also returned in gl[layer-1][var], so that
// the record's own flux uses them from registers.
__device__ __forceinline__ void ghost_fill(const DevBlock* blocks, const DevBC& e, int buf, double gl[2][4]) {
  const DevBlock& B = blocks[e.blk];
  double* P = B.P[buf];
  const long long pl = B.plane;
  const int pitch = B.pitch;
  auto off = [pitch](int i, int j) { return (long long)(j + 1) * pitch + (i + 1); };
  const int g1 = c_guide[e.f][0], g2 = c_guide[e.f][1];
  switch (e.type) {
    case 101:
    case 201: {                                   // Ghost_Connection :143-160
      const DevBlock& S = blocks[e.dblk];
      const double* Ps = S.P[buf];
      const int s1 = c_guide[e.fs][0], s2 = c_guide[e.fs][1];
#pragma unroll
      for (int g = 1; g <= 2; ++g) {
        long long od = (long long)(e.js + s2 * (g - 1) + 1) * S.pitch + (e.is + s1 * (g - 1) + 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) gl[g - 1][v] = Ps[v * S.plane + od];
      }
#pragma unroll
      for (int g = 1; g <= 2; ++g) {
        long long om = off(e.i - g1 * g, e.j - g2 * g);
#pragma unroll
        for (int v = 0; v < 4; ++v) P[v * pl + om] = gl[g - 1][v];
      }
      break;
    }
    case 300: {                                   // Ghost_Symmetry :163-196
      long long om = off(e.i, e.j), og = off(e.i - g1, e.j - g2);
      double n1, n2;
      if (e.f <= 2) { long long of = off(e.i - (e.f % 2), e.j); n1 = B.fix[of]; n2 = B.fiy[of]; }
      else { long long of = off(e.i, e.j - (e.f % 2)); n1 = B.fjx[of]; n2 = B.fjy[of]; }
      double u = P[pl + om], v = P[2 * pl + om];
      double udotn = u * n1 + v * n2;
      u = u - 2.0 * udotn * n1;
      v = v - 2.0 * udotn * n2;
      P[pl + og] = u; P[2 * pl + og] = v;
      P[og] = P[om]; P[3 * pl + og] = P[3 * pl + om];
      break;
    }
    case 0: case 401: case 402: case 403: case 404: case 405: case 406: case 407: case 420: {
      long long om = off(e.i, e.j);              // Ghost_ZG_Extrapolate :199-219
      for (int g = 1; g <= 2; ++g) {
        long long og = off(e.i - g1 * g, e.j - g2 * g);
        for (int v = 0; v < 4; ++v) P[v * pl + og] = P[v * pl + om];
      }
      break;
    }
    default: {                                    // Ghost_Extrapolate :270-299
      const int di = -g1, dj = -g2;
      for (int g = 1; g <= 2; ++g) {
        long long o0 = off(e.i + di * g, e.j + dj * g);
        long long o1 = off(e.i + di * (g - 1), e.j + dj * (g - 1));
        long long o2 = off(e.i + di * (g - 2), e.j + dj * (g - 2));
        long long o3 = off(e.i + di * (g - 3), e.j + dj * (g - 3));
        for (int v = 0; v < 4; ++v)
          P[v * pl + o0] = 3.0 * P[v * pl + o1] - 3.0 * P[v * pl + o2] + P[v * pl + o3];
      }
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// Ghost fill and boundary-face fluxes — Fill_Ghost_Cell + BC_Fluxes (Mod_BC_Fluxes.f90:49-271),
// one thread per boundary record.  The stage kernel adds the stored values to the cells'
// residuals in record order, exactly where the reference adds them.
// ---------------------------------------------------------------------------
// F: convective boundary flux, already carrying the face sign modfm2, added where mask has the bit.
// V: viscous boundary flux as -(modfm2 * Flux), added after F where vmask has the bit (BC 300/302, model 1).
// SD_Tramel — Lib_Shock_Detector.f90:14-38 (DELTA = 20): p at the cell, p1 / p2 at i-1 / i+1, p3 / p4 at
// j-1 / j+1.  tanh(x) for x = 10 phi^3 <= 1.6e-7 is x - x^3/3 (the next term of its series is below
// 1e-28 x): tanh rounded to nearest, the same expression as the oracle's.
// g: the closure whose division it uses (gdiv)
template <int DV>
__device__ __forceinline__ double sd_tramel(const Gas<double, DV>& g, double p, double p1, double p2, double p3,
                                            double p4) {
  const double kp1 = fabs(gdiv(g, p2 - 2.0 * p + p1, p2 + 2.0 * p + p1));
  const double kp2 = fabs(gdiv(g, p4 - 2.0 * p + p3, p4 + 2.0 * p + p3));
  const double sw = kp1 > kp2 ? kp1 : kp2;
  if (!(sw < 1.0 / 20.0)) return 0.0;
  const double phi = gdiv(g, sw, 20.0);
  const double x = 10.0 * phi * phi * phi;
  return 1.0 - (x - gdiv(g, x * x * x, 3.0));
}
__device__ __forceinline__ double sd_tramel(double p, double p1, double p2, double p3, double p4) {
  return sd_tramel(Gas<double>{}, p, p1, p2, p3, p4);
}

struct BCFlux { double F[4]; double V[4]; int mask; int vmask; };

// Cases of a mixed-form handle whose items run k_stage_mx: MUSCL-SD and the HLLEM / HLLE++ / HLLC+
// solvers are not in that kernel, so such cases run the generic FP32-face kernel k_stage<float>; a laminar
// case also needs its viscous terms in the stage kernel (fvisc) and no residual smoothing.
__host__ __device__ __forceinline__ bool mx_kernel_case(int model, int space_sd, int riemann, int fvisc,
                                                        double irs_beta) {
  if (space_sd || riemann >= Q2D_RIEMANN_HLLEM) return false;
  return model == 0 || (fvisc != 0 && !(irs_beta > 0.0));
}

// A record's result, stored at the record's boundary cell (DevBC::dpos): the BCFlux struct (read by the
// FP64, generic FP32-face and laminar stage kernels, which add the faces in record order) and, for the
// cases of k_stage_mx, its convective flux with the mask applied, in FP32 (the precision of that form's
// face fluxes), in the dense per-face array (a float4 per boundary cell, indexed like the per-face record
// index), so that kernel reads a boundary cell's flux with one load and no indirection; k_bcvisc adds a
// laminar record's viscous flux to the same element.
__device__ __forceinline__ void emit_bc(BCFlux* __restrict__ out, float4* __restrict__ dense, int dpos,
                                        const BCFlux& res, bool mx) {
  if (!(dense && mx)) out[dpos] = res;                // k_stage_mx reads only the dense array
  if (dense) {
    float f[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) f[v] = ((res.mask >> v) & 1) ? (float)res.F[v] : 0.f;
    dense[dpos] = make_float4(f[0], f[1], f[2], f[3]);
  }
}

template <typename FT>
__device__ __forceinline__ void riemann_rt(int rid, const Gas<FT>& g, const FT L[4], const FT R[4], FT nx, FT ny, FT F[4],
                                           FT beta = FT(1)) {
  const FT aL = sound<FT>(g, L[3], L[0]);
  const FT aR = sound<FT>(g, R[3], R[0]);
  if (rid == Q2D_RIEMANN_HLLE) riemann<FT, 1>(g, L[1], L[2], L[3], aL, L[0], R[1], R[2], R[3], aR, R[0], nx, ny, F);
  else if (rid == Q2D_RIEMANN_HLLEM) riemann<FT, 2>(g, L[1], L[2], L[3], aL, L[0], R[1], R[2], R[3], aR, R[0], nx, ny, F);
  else if (rid == Q2D_RIEMANN_HLLEPP)
    riemann<FT, 3>(g, L[1], L[2], L[3], aL, L[0], R[1], R[2], R[3], aR, R[0], nx, ny, F, beta);
  else if (rid == Q2D_RIEMANN_HLLCP)
    riemann<FT, 4>(g, L[1], L[2], L[3], aL, L[0], R[1], R[2], R[3], aR, R[0], nx, ny, F, beta);
  else riemann<FT, 0>(g, L[1], L[2], L[3], aL, L[0], R[1], R[2], R[3], aR, R[0], nx, ny, F);
}

template <typename FT>
// phase 0: ghost fill and flux of each record in one pass.  Blocks one cell wide in a direction need
// two passes (phase 1: every record's ghost fill, phase 2: the fluxes), because there a record's
// interior neighbour is a ghost cell of the opposite face's record.
// begin_cases > 0 (the step's first launch, when its dt came from the previous step's last stage): the
// launch also does k_begin's per-case work, which touches no control word that a record reads.
__global__ __launch_bounds__(256) void k_bcflux(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                                DevCtrl* __restrict__ ctrl, const DevBC* __restrict__ bcs, int nbc,
                                                int irk, BCFlux* __restrict__ out, int phase, int begin_cases,
                                                float4* __restrict__ dense) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = t; c < begin_cases; c += gridDim.x * blockDim.x) begin_case(ctrl, cases, c, 1);
  if (t >= nbc) return;
  const DevBC E = bcs[t];
  // the case's descriptors are indexed from the record itself, so their loads do not wait for the block's
  const DevBlock& B = blocks[E.blk];
  const DevCtrl& K = ctrl[E.cid];
  const DevCase& C = cases[E.cid];
  if (!K.active || irk > C.n_rk) return;
  const int buf = stage_in(K.cur, irk);
  const double* P = B.P[buf];
  const long long pl = B.plane;
  const int pitch = B.pitch;
  auto off = [pitch](int i, int j) { return (long long)(j + 1) * pitch + (i + 1); };
  const int Im = E.i, Jm = E.j, Fm = E.f;
  const int modfm = Fm % 2, modfm1 = 1 - modfm, modfm2 = 1 - 2 * modfm;
  const double sgn = (double)modfm2;
  const int g1 = c_guide[Fm][0], g2 = c_guide[Fm][1];
  // the boundary cell and its interior neighbour: interior cells, which no record's ghost fill writes,
  // so they are read before it
  const long long oB = off(Im, Jm), oI = off(Im + g1, Jm + g2);
  double Bp[4], Ip[4];
  if (phase == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) { Bp[v] = P[v * pl + oB]; Ip[v] = P[v * pl + oI]; }
  }
  // Fill_Ghost_Cell for this record: the only ghost cells this record's flux reads (the connection
  // stencil, Lib_BC_Fluxes_Connection.f90:49-62) are its own, written here and kept in registers.
  double gl[2][4];
  if (phase != 2) ghost_fill(blocks, E, buf, gl);
  if (phase == 1) return;
  if (phase == 2) {                                  // after every record's ghost fill
#pragma unroll
    for (int v = 0; v < 4; ++v) { Bp[v] = P[v * pl + oB]; Ip[v] = P[v * pl + oI]; }
#pragma unroll
    for (int g = 1; g <= 2; ++g) {
      const long long og = off(Im - g1 * g, Jm - g2 * g);
#pragma unroll
      for (int v = 0; v < 4; ++v) gl[g - 1][v] = P[v * pl + og];
    }
  }
  const int lim = C.limiter, rid = C.riemann;
  Gas<FT> gf{(FT)C.R, (FT)C.cp, (FT)C.cv, (FT)C.gam};
  gf.tab = C.tab; gf.ntab = C.ntab;
  const Gas<double> gd = case_gas(C);
  // Face_Index — Lib_BC_Fluxes.f90:34-50
  const bool idir = Fm <= 2;
  const long long ofc = idir ? off(Im - modfm, Jm) : off(Im, Jm - modfm);
  // mixed-form Euler blocks and the blocks of laminar k_stage_mx cases read the compact FP32 geometry (face
  // length-scaled normal, half cell sizes): fewer cache lines per record, which matters for i-face records
  // (one line per lane per plane)
  const bool mxg = sizeof(FT) == 4 && B.mgn != nullptr && (C.model == 0 || C.mxk != 0);
  double n1, n2, il;
  if (mxg) {
    const float4 g = B.mgn[ofc];
    const float sx = idir ? g.x : g.z, sy = idir ? g.y : g.w;
    const float l = __builtin_sqrtf(sx * sx + sy * sy);
    n1 = (double)(sx / l); n2 = (double)(sy / l); il = (double)l;
  } else {
    n1 = idir ? B.fix[ofc] : B.fjx[ofc];
    n2 = idir ? B.fiy[ofc] : B.fjy[ofc];
    il = idir ? B.fil[ofc] : B.fjl[ofc];
  }
  const double* dlp = idir ? B.dlx : B.dly;
  auto half_dl = [&](long long o) -> double {        // dl / 2 of the cell at offset o in the face's direction
    if (mxg) return idir ? (double)__uint_as_float((unsigned)((const unsigned long long*)B.mgs)[o]) : (double)B.mhy[o];
    return dlp[o] * 0.5;
  };
  BCFlux res;
  res.mask = 0xF; res.vmask = 0;
  res.V[0] = res.V[1] = res.V[2] = res.V[3] = 0.0;
  int ty = E.type;

  if (ty == 101 || ty == 201) {
    // BC_Connection_Eul — Lib_BC_Fluxes_Connection.f90:24-133
    const int I1 = modfm * (Im - g1 * 2) + modfm1 * (Im + g1), J1 = modfm * (Jm - g2 * 2) + modfm1 * (Jm + g2);
    const int I2 = modfm1 * Im + modfm * (Im - g1), J2 = modfm1 * Jm + modfm * (Jm - g2);
    const int I3 = modfm1 * (Im - g1) + modfm * Im, J3 = modfm1 * (Jm - g2) + modfm * Jm;
    const int I4 = modfm1 * (Im - g1 * 2) + modfm * (Im + g1), J4 = modfm1 * (Jm - g2 * 2) + modfm * (Jm + g2);
    const long long o1 = off(I1, J1), o2 = off(I2, J2), o3 = off(I3, J3), o4 = off(I4, J4);
    const double Dl1 = half_dl(o1), Dl2 = half_dl(o2), Dl3 = half_dl(o3), Dl4 = half_dl(o4);
    const FT Dist23 = (FT)(Dl2 + Dl3), Dist12 = (FT)(Dl1 + Dl2), Dist34 = (FT)(Dl3 + Dl4);
    // MUSCL-SD: the boundary cell's detector value (Lib_BC_Fluxes_Connection.f90:64-69), from the state
    // after every record's ghost fill (such handles run the two-phase launch)
    FT bt = FT(1), brie = FT(1);                       // SD_limiter :64-69, SD_riemann :103-108
    if (C.space_sd || rid >= Q2D_RIEMANN_HLLEPP) {
      const FT d = (FT)sd_tramel(Bp[3], P[3 * pl + off(Im - 1, Jm)], P[3 * pl + off(Im + 1, Jm)],
                                 P[3 * pl + off(Im, Jm - 1)], P[3 * pl + off(Im, Jm + 1)]);
      bt = C.space_sd ? d : FT(1);
      brie = rid >= Q2D_RIEMANN_HLLEPP ? d : FT(1);
    }
    FT L[4], R[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      // ghost layers 2, 1 | cells 1, 2 across the face: the ghost values from ghost_fill's registers,
      // the boundary cell (Bp) and its interior neighbour (Im+g1, Jm+g2) (o4 for odd faces, o1 for even)
      const double p1 = modfm ? gl[1][v] : Ip[v], p2 = modfm ? gl[0][v] : Bp[v];
      const double p3 = modfm ? Bp[v] : gl[0][v], p4 = modfm ? Ip[v] : gl[1][v];
      const FT s0 = qdiv<FT>((FT)(p2 - p1), Dist12);
      const FT s1 = qdiv<FT>((FT)(p3 - p2), Dist23);
      const FT s2 = qdiv<FT>((FT)(p4 - p3), Dist34);
      const FT sl = limiter<FT>(lim, s1, s0);
      const FT sr = limiter<FT>(lim, s2, s1);
      L[v] = (FT)p2 + (bt * sl) * (FT)Dl2;
      R[v] = (FT)p3 - (bt * sr) * (FT)Dl3;
    }
    const bool bad = !(L[0] > FT(0) && L[3] > FT(0) && R[0] > FT(0) && R[3] > FT(0));
    if (bad) record_fail(&ctrl[E.cid], Q2D_E_INTERFACE, fail_key(irk, 0, B.blk_in_case, Jm, Im));
    FT F[4];
    riemann_rt<FT>(rid, gf, L, R, (FT)n1, (FT)n2, F, brie);
    // one species: Sel_L*rho_L/rho_L - Sel_R*rho_R/rho_R == 1, so Flux(1) = F_r*int_l
#pragma unroll
    for (int v = 0; v < 4; ++v) res.F[v] = ((double)F[v] * il) * sgn;
    emit_bc(out, dense, E.dpos, res, C.mxk != 0);
    return;
  }

  if (ty == 400) {
    // BC_Extrapolation — Lib_BC_Fluxes_Extrapolation.f90:18-49
    const double rho = Bp[0], u = Bp[1], v = Bp[2], p = Bp[3];
    const double Un = u * n1 + v * n2;
    double Fb[4];
    Fb[0] = rho * Un * il;
    const double Fmass = Fb[0];
    Fb[1] = Fmass * u + p * il * n1;
    Fb[2] = Fmass * v + p * il * n2;
    const double vmag = sqrt(u * u + v * v);
    const double Tt = p / (rho * gd.R);
    Fb[3] = Fmass * ((gd.tab ? tab_at(gd.tab, gd.ntab, kTabH, Tt) : gd.cp * Tt) + 0.5 * vmag * vmag);
#pragma unroll
    for (int k = 0; k < 4; ++k) res.F[k] = sgn * Fb[k];
    emit_bc(out, dense, E.dpos, res, C.mxk != 0);
    return;
  }
  if (ty == 406) {
    // BC_Outflow + Outflow — Lib_BC_Fluxes_Outflow.f90:36-165
    const double Un = Bp[1] * n1 + Bp[2] * n2;
    if (sgn * Un < 0.0) {
      ty = 300;                                          // error=1 -> BC_Symmetry_Eul (Mod_BC_Fluxes.f90:227)
    } else {
      const double Bound_Sound = sound<double>(gd, Bp[3], Bp[0]);
      const double G = gd.tab ? tab_gamma(gd, Bp[3] / (Bp[0] * gd.R)) : gd.gam;
      const double Riem = Bound_Sound + sgn * 0.5 * (G - 1.0) * Un;
      const double VT1 = Bp[1] - Un * n1, VT2 = Bp[2] - Un * n2;
      const double Mach_N = fabs(Un / Bound_Sound);
      double Fp[4], Face_Un, Face_rho;
      if (Mach_N < 1.0 && E.p > 0.0) {
        double p_exit = E.p;
        const double delta = 0.5 * (G - 1.0);
        const double ex = G / (G - 1.0);
        const double den = 1.0 / exact_pow(1.0 + delta, ex);
        const double pstar = Bp[3] * den * exact_pow(1.0 + delta * (Mach_N * Mach_N), ex);
        if (pstar >= p_exit) p_exit = pstar;
        const double Isen = (G - 1.0) / (2.0 * G);
        Fp[3] = p_exit * E.rel_fac + Bp[3] * (1.0 - E.rel_fac);
        const double Face_Sound = Bound_Sound * exact_pow(Fp[3] / Bp[3], Isen);
        Face_Un = (Riem - Face_Sound) / (sgn * 0.5 * (G - 1.0));
        Face_rho = G * Fp[3] / (Face_Sound * Face_Sound);
        Fp[0] = Bp[0] / Bp[0] * Face_rho;
        Fp[1] = VT1 + Face_Un * n1;
        Fp[2] = VT2 + Face_Un * n2;
      } else {
        Fp[0] = Bp[0]; Fp[1] = Bp[1]; Fp[2] = Bp[2]; Fp[3] = Bp[3];
        Face_Un = Un; Face_rho = Bp[0];
      }
      const double Face_T = Fp[3] / (gd.R * Face_rho);
      const double Hf = 0.0 + Fp[0] / Face_rho * (gd.tab ? tab_at(gd.tab, gd.ntab, kTabH, Face_T) : gd.cp * Face_T);
      double Fb[4];
      Fb[0] = Fp[0] * il * Face_Un;
      const double mdot = Face_rho * il * Face_Un;
      Fb[1] = Fp[1] * mdot + Fp[3] * il * n1;
      Fb[2] = Fp[2] * mdot + Fp[3] * il * n2;
      Fb[3] = mdot * (0.5 * (Fp[1] * Fp[1] + Fp[2] * Fp[2]) + Hf);
#pragma unroll
      for (int k = 0; k < 4; ++k) res.F[k] = sgn * Fb[k];
      emit_bc(out, dense, E.dpos, res, C.mxk != 0);
      return;
    }
  }
  FT L[4], R[4];
  if (ty == 405) {
    // BC_Inlet_Supersonic_Static + Supersonic_Inflow — Lib_BC_Fluxes_Inflow.f90:272-403
    const double Un = Bp[1] * n1 + Bp[2] * n2;
    if (sgn * Un >= 0.0) {
      ty = 300;                                          // error=1 -> BC_Symmetry_Eul (Mod_BC_Fluxes.f90:221)
    } else {
      const double dm = (double)modfm, dm1 = (double)modfm1;
      const double flow_Sound = sound<double>(gd, Bp[3], Bp[0]);
      const double flow_Mach = sqrt(Bp[1] * Bp[1] + Bp[2] * Bp[2]) / flow_Sound;
      const double Mach_l = E.rel_fac * E.mach + (1.0 - E.rel_fac) * flow_Mach;
      double Sup[4];
      Sup[3] = E.rel_fac * E.p + (1.0 - E.rel_fac) * Bp[3];
      const double Sup_rho = Sup[3] / (1.0 * gd.R * E.T);
      Sup[0] = Sup_rho * 1.0;
      const double Sup_Sound = sound<double>(gd, Sup[3], Sup_rho);
      Sup[1] = Mach_l * Sup_Sound * E.cos_a;
      Sup[2] = Mach_l * Sup_Sound * E.sin_a;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double P1 = dm * Sup[v] + dm1 * Ip[v];
        const double P3 = dm * Ip[v] + dm1 * Sup[v];
        const FT d32 = (FT)(P3 - Bp[v]), d21 = (FT)(Bp[v] - P1);
        const FT sl = limiter<FT>(lim, d32, d21);
        const FT Fc = (FT)Bp[v] + FT(0.5) * sl * (FT)sgn;
        L[v] = (FT)dm * (FT)Sup[v] + (FT)dm1 * Fc;
        R[v] = (FT)dm * Fc + (FT)dm1 * (FT)Sup[v];
      }
    }
  }
  if (ty == 300 || ty == 302) {
    // BC_Symmetry_Eul — Lib_BC_Fluxes_Symmetry.f90:28-124 (Un_wall = 0); momentum only
    res.mask = 0x6;
    const double Bound_Dl = half_dl(oB), Int_Dl = half_dl(oI);
    double Gp[4] = {Bp[0], Bp[1], Bp[2], Bp[3]};
    double Un = Bp[1] * n1 + Bp[2] * n2;
    Gp[1] = Gp[1] - 2.0 * Un * n1;
    Gp[2] = Gp[2] - 2.0 * Un * n2;
    const double dm = (double)modfm, dm1 = (double)modfm1;
    const FT Dist12 = (FT)(2.0 * Bound_Dl * dm + (Bound_Dl + Int_Dl) * dm1);
    const FT Dist23 = (FT)((Bound_Dl + Int_Dl) * dm + 2.0 * Bound_Dl * dm1);
    const FT Dist = (FT)(Bound_Dl * sgn);
    FT Fp[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double P1 = dm * Gp[v] + dm1 * Ip[v];
      const double P3 = dm * Ip[v] + dm1 * Gp[v];
      const FT s32 = qdiv<FT>((FT)(P3 - Bp[v]), Dist23);
      const FT s21 = qdiv<FT>((FT)(Bp[v] - P1), Dist12);
      const FT sl = limiter<FT>(lim, s32, s21);
      Fp[v] = (FT)Bp[v] + sl * Dist;
    }
    const FT fn1 = (FT)n1, fn2 = (FT)n2;
    const FT Unf = Fp[1] * fn1 + Fp[2] * fn2;
    const FT vt1 = Fp[1] - Unf * fn1, vt2 = Fp[2] - Unf * fn2;
    const FT Un1 = Unf * (FT)sgn;
    L[0] = Fp[0]; L[1] = Un1 * fn1 + vt1; L[2] = Un1 * fn2 + vt2; L[3] = Fp[3];
    R[0] = Fp[0]; R[1] = -Un1 * fn1 + vt1; R[2] = -Un1 * fn2 + vt2; R[3] = Fp[3];
  }
  FT F[4];
  riemann_rt<FT>(rid, gf, L, R, (FT)n1, (FT)n2, F);
#pragma unroll
  for (int v = 0; v < 4; ++v) res.F[v] = sgn * ((double)F[v] * il);
  if (res.mask == 0x6) { res.F[0] = 0.0; res.F[3] = 0.0; }
  emit_bc(out, dense, E.dpos, res, C.mxk != 0);
}

// ---------------------------------------------------------------------------
// laminar Navier-Stokes terms (model 1), FP64 in both arithmetic forms
// ---------------------------------------------------------------------------
constexpr double kNegTwoThird = -2.0 / 3.0;

// Stress_Vector_Std / Stress_Vector_Axi — Lib_Fluid.f90:26-76, laminar (mi = mi_l + 0).  gUX = Gradient(nu,X),
// gVX = Gradient(nv,X) after the metric transform; VG: V.G of the face (Definition A; 0 without area
// variation), which the Axi form does not use.
__device__ __forceinline__ void stress_vector(bool axi, double gUx, double gUy, double gVx, double gVy, double n1,
                                              double n2, double mil, double r_ave, double vr, double VG, double& S1,
                                              double& S2) {
  const double mi = mil + 0.0;
  const double lam = kNegTwoThird * mi;
  double div;
  if (axi) div = (r_ave <= 1.0e-6) ? (gUx + gVy + gVy) : (gUx + gVy + vr / r_ave);
  else div = gUx + gVy - VG;
  const double ld = lam * div;
  const double txx = ld + 2.0 * mi * gUx;
  const double tyy = ld + 2.0 * mi * gVy;
  const double txy = mi * (gUy + gVx);
  S1 = txx * n1 + txy * n2;
  S2 = txy * n1 + tyy * n2;
}

// V_dot_G_face_avg — Lib_Fluid.f90:550-564: <V>_face . <G>_face
__device__ __forceinline__ double v_dot_g_face(double u1, double v1, double u2, double v2, double gx1, double gy1,
                                               double gx2, double gy2) {
  const double Va0 = 0.5 * (u1 + u2), Va1 = 0.5 * (v1 + v2);
  const double Ga0 = 0.5 * (gx1 + gx2), Ga1 = 0.5 * (gy1 + gy2);
  return Va0 * Ga0 + Va1 * Ga1;
}

// one row of matmul(Gradient, M): (g0, g1) . M, M(l,c) column-major = (m11, m21, m12, m22)
__device__ __forceinline__ void grad_m(double g0, double g1, const double M[4], double& o0, double& o1) {
  o0 = g0 * M[0] + g1 * M[1];
  o1 = g0 * M[2] + g1 * M[3];
}

// One cell's primitive state and temperature T = p / (rho R) (co_rotot_Rtot + the EOS, one species).
struct VC { double r, u, v, p, T; };
__device__ __forceinline__ VC vc_shdn(const VC& c) { return VC{shdn(c.r), shdn(c.u), shdn(c.v), shdn(c.p), shdn(c.T)}; }
__device__ __forceinline__ VC vc_shup(const VC& c) { return VC{shup(c.r), shup(c.u), shup(c.v), shup(c.p), shup(c.T)}; }
__device__ __forceinline__ void vc_temperature(VC& c, double R) { c.T = c.p / (c.r * R); }

// Compute_Diffusive_Flux — Lib_Diffusive.f90:113-199 from the face state (pm0..pm3) and the transformed
// gradients of u, v and T; D = (Flux(nu), Flux(nv), Flux(np)).
__device__ __forceinline__ void diff_flux(const DevCase& C, bool axi, double pm0, double pm1, double pm2, double pm3,
                                          double gUx, double gUy, double gVx, double gVy, double gTx, double gTy,
                                          double n1, double n2, double il, double r_ave, double VG, double D[3]) {
  const double T = pm3 / (pm0 * C.R);
  const double mil = mu_lam(C, T);
  const double kappa = k_lam(C, mil, T) + 0.0;
  double S1, S2;
  stress_vector(axi, gUx, gUy, gVx, gVy, n1, n2, mil, r_ave, pm2, VG, S1, S2);
  D[0] = il * S1;
  D[1] = il * S2;
  D[2] = il * ((S1 * pm1 + S2 * pm2) + kappa * (gTx * n1 + gTy * n2) + 0.0);
}

// Diffusive_Flux + Tangential_Gradient + Compute_Diffusive_Flux — Lib_Diffusive.f90:10-199 for one interior
// face.  c1, c2: the face's two cells; c3..c6: the tangential stencil (c4 - c3 + c6 - c5) / 4; a: 0 for an
// i-face, 1 for a j-face; M: the two cells' metric averaged.  One species: the species flux and the
// enthalpy-diffusion term vanish (grad Y = 0), so D = (Flux(nu), Flux(nv), Flux(np)).
__device__ __forceinline__ void visc_face(const DevCase& C, bool axi, int a, double n1, double n2, double il,
                                          const VC& c1, const VC& c2, const VC& c3, const VC& c4, const VC& c5,
                                          const VC& c6, const double M[4], double r_ave, double VG, double D[3]) {
  const double nU = c2.u - c1.u, nV = c2.v - c1.v, nT = c2.T - c1.T;
  const double tU = (c4.u - c3.u + c6.u - c5.u) * 0.25;
  const double tV = (c4.v - c3.v + c6.v - c5.v) * 0.25;
  const double tT = (c4.T - c3.T + c6.T - c5.T) * 0.25;
  double gUx, gUy, gVx, gVy, gTx, gTy;
  if (a == 0) {
    grad_m(nU, tU, M, gUx, gUy); grad_m(nV, tV, M, gVx, gVy); grad_m(nT, tT, M, gTx, gTy);
  } else {
    grad_m(tU, nU, M, gUx, gUy); grad_m(tV, nV, M, gVx, gVy); grad_m(tT, nT, M, gTx, gTy);
  }
  const double pm0 = 0.5 * (c1.r + c2.r), pm1 = 0.5 * (c1.u + c2.u);
  const double pm2 = 0.5 * (c1.v + c2.v), pm3 = 0.5 * (c1.p + c2.p);
  diff_flux(C, axi, pm0, pm1, pm2, pm3, gUx, gUy, gVx, gVy, gTx, gTy, n1, n2, il, r_ave, VG, D);
}

// Source_Axi_Visc_Blk (Lib_Source_Axisymmetry.f90:125-236) of cell c with its neighbours E, W, N, S (wall
// ghosts already reflected), the cell's metric M, r_ave and 1/r; out = (S(nu), S(nv), S(np)) per unit
// volume.  One species: the species and enthalpy-diffusion terms vanish.
__device__ __forceinline__ void visc_source_axi(const DevCase& C, const VC& cC, const VC& qE, const VC& qW,
                                                const VC& qN, const VC& qS, const double MC[4], double sa,
                                                double sb, double out[3]) {
  const double mil = mu_lam(C, cC.T);
  const double kl = k_lam(C, mil, cC.T);
  double gUx, gUy, gVx, gVy, gTx, gTy;
  grad_m((qE.u - qW.u) * 0.5, (qN.u - qS.u) * 0.5, MC, gUx, gUy);
  grad_m((qE.v - qW.v) * 0.5, (qN.v - qS.v) * 0.5, MC, gVx, gVy);
  grad_m((qE.T - qW.T) * 0.5, (qN.T - qS.T) * 0.5, MC, gTx, gTy);
  (void)gTx;
  const double kappa = kl + 0.0;
  const double mi = mil + 0.0;
  const double lam = kNegTwoThird * mi;
  // Stress_Tensor_Axi — Lib_Fluid.f90:166-198
  const double vr = cC.v;
  const double div = gUx + gVy + vr / sa;
  const double ld2 = lam * div;
  const double tyy = ld2 + 2.0 * mi * gVy;
  const double txy = mi * (gUy + gVx);
  const double tzz = 2.0 * mi * vr / sa + ld2;
  const double qr = -kappa * gTy;
  out[0] = txy * sb;
  out[1] = (tyy - tzz) * sb;
  out[2] = (-qr + cC.u * txy + cC.v * tyy + 0.0) * sb;
}

// ---- the mixed form's FP32 viscous terms (fused into its laminar stage kernel) ------------------------
// The formulas of visc_face / diff_flux / stress_vector / visc_source_axi (Lib_Diffusive.f90:10-199,
// Lib_Fluid.f90:26-76, 166-198, Lib_Source_Axisymmetry.f90:125-236), in FP32 with the hardware reciprocal
// and square root: the mixed form is held to its tolerance against the FP64 oracle, not to the
// reference's operation order.  Table mode is FP64 only, so the closed-form transport applies.
struct VCf { float r, u, v, p, T; };
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ VCf vcf_of(const double q[4], float R) {
  VCf c{(float)q[0], (float)q[1], (float)q[2], (float)q[3], 0.f};
  c.T = c.p * frcp(c.r * R);
  return c;
}
__device__ __forceinline__ VCf vcf_shdn(const VCf& c) { return VCf{shdn(c.r), shdn(c.u), shdn(c.v), shdn(c.p), shdn(c.T)}; }
__device__ __forceinline__ VCf vcf_shup(const VCf& c) { return VCf{shup(c.r), shup(c.u), shup(c.v), shup(c.p), shup(c.T)}; }
struct TransF { float mu_ref, inv_Tref, Tref_S, S, cp_Pr; };   // Sutherland constants, cp / Pr
__device__ __forceinline__ TransF trans_f(const DevCase& C) {
  return TransF{(float)C.mu_ref, (float)(1.0 / C.T_ref), (float)(C.T_ref + C.S_suth), (float)C.S_suth,
                (float)(C.cp / C.Pr)};
}
__device__ __forceinline__ float mu_lam_f(const TransF& t, float T) {
  const float x = T * t.inv_Tref;
  return t.mu_ref * (x * __builtin_amdgcn_sqrtf(x)) * t.Tref_S * frcp(T + t.S);
}
__device__ __forceinline__ void grad_mf(float g0, float g1, const float M[4], float& o0, float& o1) {
  o0 = g0 * M[0] + g1 * M[1];
  o1 = g0 * M[2] + g1 * M[3];
}
// diffusive flux of one interior face (a: 0 i-face, 1 j-face), the stencil as visc_face's
__device__ __forceinline__ void visc_face_f(const TransF& t, float R, bool axi, int a, float n1, float n2, float il,
                                            const VCf& c1, const VCf& c2, const VCf& c3, const VCf& c4,
                                            const VCf& c5, const VCf& c6, const float M[4], float r_ave,
                                            double D[3]) {
  const float nU = c2.u - c1.u, nV = c2.v - c1.v, nT = c2.T - c1.T;
  const float tU = (c4.u - c3.u + c6.u - c5.u) * 0.25f;
  const float tV = (c4.v - c3.v + c6.v - c5.v) * 0.25f;
  const float tT = (c4.T - c3.T + c6.T - c5.T) * 0.25f;
  float gUx, gUy, gVx, gVy, gTx, gTy;
  if (a == 0) {
    grad_mf(nU, tU, M, gUx, gUy); grad_mf(nV, tV, M, gVx, gVy); grad_mf(nT, tT, M, gTx, gTy);
  } else {
    grad_mf(tU, nU, M, gUx, gUy); grad_mf(tV, nV, M, gVx, gVy); grad_mf(tT, nT, M, gTx, gTy);
  }
  const float pm0 = 0.5f * (c1.r + c2.r), pm1 = 0.5f * (c1.u + c2.u);
  const float pm2 = 0.5f * (c1.v + c2.v), pm3 = 0.5f * (c1.p + c2.p);
  const float T = pm3 * frcp(pm0 * R);
  const float mi = mu_lam_f(t, T);
  const float kappa = mi * t.cp_Pr;
  const float lam = -2.0f / 3.0f * mi;
  float div;
  if (axi) div = (r_ave <= 1.0e-6f) ? (gUx + gVy + gVy) : (gUx + gVy + pm2 * frcp(r_ave));
  else div = gUx + gVy;
  const float ld = lam * div;
  const float txx = ld + 2.0f * mi * gUx, tyy = ld + 2.0f * mi * gVy, txy = mi * (gUy + gVx);
  const float S1 = txx * n1 + txy * n2, S2 = txy * n1 + tyy * n2;
  D[0] = (double)(il * S1);
  D[1] = (double)(il * S2);
  D[2] = (double)(il * ((S1 * pm1 + S2 * pm2) + kappa * (gTx * n1 + gTy * n2)));
}
__device__ __forceinline__ void visc_source_axi_f(const TransF& t, const VCf& cC, const VCf& qE, const VCf& qW,
                                                  const VCf& qN, const VCf& qS, const float MC[4], float sa,
                                                  float sb, double out[3]) {
  const float mi = mu_lam_f(t, cC.T);
  const float kappa = mi * t.cp_Pr;
  float gUx, gUy, gVx, gVy, gTx, gTy;
  grad_mf((qE.u - qW.u) * 0.5f, (qN.u - qS.u) * 0.5f, MC, gUx, gUy);
  grad_mf((qE.v - qW.v) * 0.5f, (qN.v - qS.v) * 0.5f, MC, gVx, gVy);
  grad_mf((qE.T - qW.T) * 0.5f, (qN.T - qS.T) * 0.5f, MC, gTx, gTy);
  (void)gTx;
  const float lam = -2.0f / 3.0f * mi;
  const float vr = cC.v, ir = frcp(sa);
  const float ld2 = lam * (gUx + gVy + vr * ir);
  const float tyy = ld2 + 2.0f * mi * gVy;
  const float txy = mi * (gUy + gVx);
  const float tzz = 2.0f * mi * vr * ir + ld2;
  const float qr = -kappa * gTy;
  out[0] = (double)(txy * sb);
  out[1] = (double)((tyy - tzz) * sb);
  out[2] = (double)((-qr + cC.u * txy + cC.v * tyy) * sb);
}

// Interior diffusive face fluxes and the axisymmetric viscous source of every cell of the viscous work
// items.  Cell (x,r) stores the flux of its i-face (x|x+1) (Mod_Fluxes.f90:160-183), of its j-face
// (r|r+1) (:207-230) and its Source_Axi_Visc_Blk term (Lib_Source_Axisymmetry.f90:81-247); the stage
// kernel adds them to the residual in the reference's order.  Like the stage kernel, one wavefront
// marches down a 60-column strip: rows r-1, r, r+1 (state and temperature, metric, node radius) roll
// through registers and the x-1 / x+1 neighbours come from the adjacent lanes (DPP), so every cell's
// state is loaded and its temperature computed once.  The faces read the ghost cells as Fill_Ghost_Cell
// left them; the source reads a wall face's ghost as Ghost_Wall_Extrapolation (Lib_Ghost.f90:366-399)
// rewrites it just before the sources: the boundary cell's own state with the velocity reversed.
__global__ __launch_bounds__(256) void k_visc(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                              const DevCtrl* __restrict__ ctrl, const Item* __restrict__ items,
                                              int item_lo, int item_hi, const DevBC* __restrict__ bcs,
                                              const int* __restrict__ bcface, int irk) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wpg = blockDim.x / kWave;
  const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpg + threadIdx.x / kWave);   // wave-uniform: item and block descriptors load through the scalar cache
  const int nwaves = gridDim.x * wpg;
  for (int it = item_lo + wave0; it < item_hi; it += nwaves) {
    const Item item = items[it];
    const DevBlock& B = blocks[item.blk];
    const DevCtrl& K = ctrl[B.case_id];
    const DevCase& C = cases[B.case_id];
    if (!K.active || irk > C.n_rk || C.fvisc) continue;   // fvisc: the stage kernel evaluates these terms
    const int ni = B.ni, nj = B.nj, pitch = B.pitch;
    const long long pl = B.plane;
    const bool axi = C.source == Q2D_SRC_AXI, area = C.source == Q2D_SRC_AREA;
    const double R = C.R;
    const int x = item.x0 - 2 + lane;                  // Fortran column of this lane
    const int xc = x <= ni + 2 ? x : ni + 2;           // clamped column for loads
    const bool own = lane >= 2 && lane < 2 + kOwn && x <= ni;
    const int rmax = nj + 2;
    GP P = gp(B.P[stage_in(K.cur, irk)]);
    GP m11 = gp(B.m11), m21 = gp(B.m21), m12 = gp(B.m12), m22 = gp(B.m22), nodr = gp(B.nodr);
    GPW vdx = gpw(B.vdx), vey = gpw(B.vey), vsv = gpw(B.vsv);
    auto at = [&](int r) { return (long long)(min(r, rmax) + 1) * pitch + (xc + 1); };
    auto ldc = [&](int r) { const long long o = at(r); return VC{P[o], P[pl + o], P[2 * pl + o], P[3 * pl + o], 0.0}; };
    auto ldm = [&](int r, double M[4]) { const long long o = at(r); M[0] = m11[o]; M[1] = m21[o]; M[2] = m12[o]; M[3] = m22[o]; };
    // wall faces of the boundary cells of this column (BC 301..304)
    auto is_wall = [&](int f, int pos) {               // uniform faces need no record lookup
      if (B.bcf[f] < 0 || B.wallf[f] == 0) return false;
      if (B.wallf[f] == 1) return true;
      const int t = bcs[bcface[B.bcf[f] + pos - 1]].type;
      return t >= 301 && t <= 304;
    };
    const bool src = axi || area;                      // a viscous source term (Lib_Sources.f90)
    const bool wall3 = src && own && is_wall(2, x), wall4 = src && own && is_wall(3, x);
    // area variation: G of rows r (gxC, gyC) and r+1 (gxN, gyN) for the V.G of the j-faces
    GP gsa = gp(B.sa), gsb = gp(B.sb);
    double gxN = area ? gsa[at(item.y0)] : 0.0, gyN = area ? gsb[at(item.y0)] : 0.0;
    // rows y0-1, y0, y0+1 resident; row y0+2 in flight
    VC cS = ldc(item.y0 - 1), cC = ldc(item.y0), cN = ldc(item.y0 + 1), cF = ldc(item.y0 + 2);
    vc_temperature(cS, R); vc_temperature(cC, R); vc_temperature(cN, R);
    double MC[4], MN[4], MF[4];
    ldm(item.y0, MC); ldm(item.y0 + 1, MN); ldm(item.y0 + 2, MF);
    double nodS = nodr[at(item.y0 - 1)], nodC = nodr[at(item.y0)], nodN = nodr[at(item.y0 + 1)];
    for (int r = item.y0; r <= item.y1; ++r) {
      const long long o = at(r);
      // this row's face geometry and source terms (first used after the neighbour shifts)
      const double fix = gp(B.fix)[o], fiy = gp(B.fiy)[o], fil = gp(B.fil)[o];
      const double fjx = gp(B.fjx)[o], fjy = gp(B.fjy)[o], fjl = gp(B.fjl)[o];
      const double sa = src ? gsa[o] : 0.0, sb = src ? gsb[o] : 0.0;   // axi: r_ave, 1/r; area: G_x, G_y
      const double gxC = gxN, gyC = gyN;
      if (area) { gxN = gsa[at(r + 1)]; gyN = gsb[at(r + 1)]; }
      const VC cE = vc_shdn(cC), cW = vc_shup(cC);      // (x+1, r), (x-1, r)
      const VC cSE = vc_shdn(cS);                      // (x+1, r-1)
      const VC cNE = vc_shdn(cN), cNW = vc_shup(cN);    // (x+1, r+1), (x-1, r+1)
      double ME[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) ME[k] = shdn(MC[k]);
      const double nodW = shup(nodC);                  // node(x-1, r)
      const double gxE = shdn(gxC), gyE = shdn(gyC);   // G(x+1, r)
      if (own && x <= ni - 1) {                        // i-face (x | x+1)
        double Ma[4], D[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) Ma[k] = 0.5 * (MC[k] + ME[k]);
        const double VG = v_dot_g_face(cC.u, cC.v, cE.u, cE.v, gxC, gyC, gxE, gyE);   // Lib_Diffusive.f90:61-68
        visc_face(C, axi, 0, fix, fiy, fil, cC, cE, cS, cN, cSE, cNE, Ma, 0.5 * (nodC + nodS), VG, D);
#pragma unroll
        for (int k = 0; k < 3; ++k) vdx[k * pl + o] = D[k];
      }
      if (own && r <= nj - 1) {                        // j-face (r | r+1)
        double Ma[4], D[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) Ma[k] = 0.5 * (MC[k] + MN[k]);
        const double VG = v_dot_g_face(cC.u, cC.v, cN.u, cN.v, gxC, gyC, gxN, gyN);
        visc_face(C, axi, 1, fjx, fjy, fjl, cC, cN, cW, cE, cNW, cNE, Ma, 0.5 * (nodC + nodW), VG, D);
#pragma unroll
        for (int k = 0; k < 3; ++k) vey[k * pl + o] = D[k];
      }
      if (src && own) {
        // Source_Axi_Visc_Blk :125-236 / Source_Area_Visc_Blk, with the wall ghosts reflected
        VC qE = cE, qW = cW, qS = cS, qN = cN;
        const VC refl{cC.r, -cC.u, -cC.v, cC.p, cC.T};
        if (x == 1 && is_wall(0, r)) qW = refl;
        if (x == ni && is_wall(1, r)) qE = refl;
        if (r == 1 && wall3) qS = refl;
        if (r == nj && wall4) qN = refl;
        if (axi) {
          double Sv[3];
          visc_source_axi(C, cC, qE, qW, qN, qS, MC, sa, sb, Sv);
#pragma unroll
          for (int k = 0; k < 3; ++k) vsv[k * pl + o] = Sv[k];
        } else {
        const double mil = mu_lam(C, cC.T);
        const double kl = k_lam(C, mil, cC.T);
        double gUx, gUy, gVx, gVy, gTx, gTy;
        grad_m((qE.u - qW.u) * 0.5, (qN.u - qS.u) * 0.5, MC, gUx, gUy);
        grad_m((qE.v - qW.v) * 0.5, (qN.v - qS.v) * 0.5, MC, gVx, gVy);
        grad_m((qE.T - qW.T) * 0.5, (qN.T - qS.T) * 0.5, MC, gTx, gTy);
        const double kappa = kl + 0.0;
        const double mi = mil + 0.0;
        const double lam = kNegTwoThird * mi;
        {
          // Source_Area_Visc_Blk — Lib_Source_Areavariation.f90:147-349 (one species: the species,
          // enthalpy-diffusion and Hirschfelder terms are 0) with Stress_Tensor_Std and V.G (:128-163)
          const double Gx = sa, Gy = sb;
          const double VdG = cC.u * Gx + cC.v * Gy;
          const double div = gUx + gVy - VdG;
          const double ld2 = lam * div;
          const double t11 = ld2 + 2.0 * mi * gUx;
          const double t22 = ld2 + 2.0 * mi * gVy;
          const double t12 = mi * (gUy + gVx);
          const double t33 = ld2 + 2.0 * mi * (-VdG);
          const double qx = -kappa * gTx, qy = -kappa * gTy;
          vsv[o] = Gx * (t33 - t11) - Gy * t12;
          vsv[pl + o] = -Gx * t12 + Gy * (t33 - t22);
          vsv[2 * pl + o] = -cC.u * (t11 * Gx + t12 * Gy) - cC.v * (t12 * Gx + t22 * Gy) + Gx * qx + Gy * qy;
        }
        }
      }
      // roll the rows; issue row r+3
      cS = cC; cC = cN; cN = cF; vc_temperature(cN, R);
      cF = ldc(r + 3);
#pragma unroll
      for (int k = 0; k < 4; ++k) { MC[k] = MN[k]; MN[k] = MF[k]; }
      ldm(r + 3, MF);
      nodS = nodC; nodC = nodN; nodN = nodr[at(r + 2)];
    }
  }
}

// BC_Symmetry_Visc (Lib_BC_Fluxes_Symmetry.f90:127-236) and BC_Wall_Temperature
// (Lib_BC_Fluxes_Wall_Temperature.f90:20-106, static wall) for the boundary records of viscous cases.
// Runs after every ghost cell of the stage is filled: at a corner the symmetry stencil reads the
// neighbouring face's ghost.  One thread per listed record; the result goes to BCFlux::V.
// A record's viscous flux: into its BCFlux (added after the convective one by the stage kernels that read
// the struct) or, for the cases of k_stage_mx, added in FP32 to the record's element of the dense per-face
// array, which k_bcflux wrote earlier in the stage (one record per element: no other writer).
__device__ __forceinline__ void emit_bcvisc(BCFlux* __restrict__ out, float4* __restrict__ dense, int dpos,
                                            const double V[4], int vmask, bool mx) {
  if (dense && mx) {
    float4 q = dense[dpos];
    if (vmask & 1) q.x += (float)V[0];
    if (vmask & 2) q.y += (float)V[1];
    if (vmask & 4) q.z += (float)V[2];
    if (vmask & 8) q.w += (float)V[3];
    dense[dpos] = q;
    return;
  }
  BCFlux& q = out[dpos];
#pragma unroll
  for (int v = 0; v < 4; ++v) q.V[v] = V[v];
  q.vmask = vmask;
}

__global__ __launch_bounds__(256) void k_bcvisc(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                                const DevCtrl* __restrict__ ctrl, const DevBC* __restrict__ bcs,
                                                const int* __restrict__ vrec, int nvrec, int irk,
                                                BCFlux* __restrict__ out, float4* __restrict__ dense) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nvrec) return;
  const int k = vrec[t];
  const DevBC E = bcs[k];
  const DevBlock& B = blocks[E.blk];
  const DevCtrl& K = ctrl[B.case_id];
  const DevCase& C = cases[B.case_id];
  if (!K.active || irk > C.n_rk) return;
  const double* P = B.P[stage_in(K.cur, irk)];
  const long long pl = B.plane;
  const int pitch = B.pitch;
  auto off = [pitch](int i, int j) { return (long long)(j + 1) * pitch + (i + 1); };
  const int Im = E.i, Jm = E.j, Fm = E.f;
  const int modfm = Fm % 2, modfm2 = 1 - 2 * modfm, modfm3 = 2 * (2 * modfm - 1);
  const double sgn = (double)modfm2;
  const bool axi = C.source == Q2D_SRC_AXI;
  const bool idir = Fm <= 2;
  const long long ofc = idir ? off(Im - modfm, Jm) : off(Im, Jm - modfm);
  const double n1 = idir ? B.fix[ofc] : B.fjx[ofc];
  const double n2 = idir ? B.fiy[ofc] : B.fjy[ofc];
  const double il = idir ? B.fil[ofc] : B.fjl[ofc];
  const long long oB = off(Im, Jm);
  const double M[4] = {B.m11[oB], B.m21[oB], B.m12[oB], B.m22[oB]};
  // face radius from the nodes of the face (:183-192)
  double r_ave;
  switch (Fm) {
    case 1: r_ave = 0.5 * (B.nodr[off(Im - 1, Jm)] + B.nodr[off(Im - 1, Jm - 1)]); break;
    case 2: r_ave = 0.5 * (B.nodr[off(Im, Jm)] + B.nodr[off(Im, Jm - 1)]); break;
    case 3: r_ave = 0.5 * (B.nodr[off(Im, Jm - 1)] + B.nodr[off(Im - 1, Jm - 1)]); break;
    default: r_ave = 0.5 * (B.nodr[off(Im, Jm)] + B.nodr[off(Im - 1, Jm)]); break;
  }
  const double rho = P[oB], u = P[pl + oB], v = P[2 * pl + oB], p = P[3 * pl + oB];
  double F1, F2, F3;
  if (E.type == 300) {
    const double mil = mu_lam(C, p / (rho * C.R));
    const long long o3 = idir ? off(Im, Jm - 1) : off(Im - 1, Jm);
    const long long o4 = idir ? off(Im, Jm + 1) : off(Im + 1, Jm);
    const double V3x = P[pl + o3], V3y = P[2 * pl + o3], V4x = P[pl + o4], V4y = P[2 * pl + o4];
    const double d1 = u * n1 + v * n2, d3 = V3x * n1 + V3y * n2, d4 = V4x * n1 + V4y * n2;
    const double V2x = u - 2.0 * d1 * n1, V2y = v - 2.0 * d1 * n2;
    const double V5x = V3x - 2.0 * d3 * n1, V5y = V3y - 2.0 * d3 * n2;
    const double V6x = V4x - 2.0 * d4 * n1, V6y = V4y - 2.0 * d4 * n2;
    const double aU = (V2x - u) * sgn, aV = (V2y - v) * sgn;                  // Gradient(:,a)
    const double bU = (V4x - V3x + V6x - V5x) * 0.25, bV = (V4y - V3y + V6y - V5y) * 0.25;   // Gradient(:,b)
    double gUx, gUy, gVx, gVy;
    if (idir) { grad_m(aU, bU, M, gUx, gUy); grad_m(aV, bV, M, gVx, gVy); }
    else { grad_m(bU, aU, M, gUx, gUy); grad_m(bV, aV, M, gVx, gVy); }
    const double Vmx = 0.5 * (u + V2x), Vmy = 0.5 * (v + V2y);
    // V.G at the symmetry face: <V> . G_loc (:196-200); 0 without area variation
    const double Gxl = C.source == Q2D_SRC_AREA ? B.sa[oB] : 0.0, Gyl = C.source == Q2D_SRC_AREA ? B.sb[oB] : 0.0;
    const double VG = v_dot_g_face(u, v, V2x, V2y, Gxl, Gyl, Gxl, Gyl);
    double S1, S2;
    stress_vector(axi, gUx, gUy, gVx, gVy, n1, n2, mil, r_ave, 0.5 * (v + V2y), VG, S1, S2);
    F1 = il * S1; F2 = il * S2;
    F3 = il * (Vmx * S1 + Vmy * S2);
  } else if (E.type == 101 || E.type == 201) {
    // BC_Connection_Visc — Lib_BC_Fluxes_Connection.f90:135-265: the mean of the face flux from this side's
    // cell gradient and from the donor side's (Pg(:,1) the donor cell, Pg(:,3:4) its neighbours along the
    // face, oriented by d11: Fill_BC_Ghost_Connection, Lib_Ghost.f90:302-326)
    const DevBlock& S = blocks[E.dblk];
    const double* Ps = S.P[stage_in(K.cur, irk)];
    const long long pls = S.plane;
    auto offs = [&S](int i, int j) { return (long long)(j + 1) * S.pitch + (i + 1); };
    const int d11 = E.pad0;
    int i1, j1, i2, j2;
    if (E.fs <= 2) { i1 = E.is; j1 = E.js - d11; i2 = E.is; j2 = E.js + d11; }
    else { i1 = E.is - d11; j1 = E.js; i2 = E.is + d11; j2 = E.js; }
    const int Ig = Im - c_guide[Fm][0], Jg = Jm - c_guide[Fm][1];
    const long long oG = off(Ig, Jg);
    double Mc[4] = {B.m11[oG], B.m21[oG], B.m12[oG], B.m22[oG]};
#pragma unroll
    for (int k = 0; k < 4; ++k) Mc[k] = 0.5 * (M[k] + Mc[k]);
    // Visc_Variables (u, v, T; the species fraction rho/rho has zero gradient) of the eight stencil cells
    const long long oS = offs(E.is, E.js), oS3 = offs(i1, j1), oS4 = offs(i2, j2);
    const double gr = Ps[oS], gu = Ps[pls + oS], gv = Ps[2 * pls + oS], gpr = Ps[3 * pls + oS];
    auto vvar = [&](const double* Q, long long pq, long long o, double out[3]) {
      const double r = Q[o];
      out[0] = Q[pq + o]; out[1] = Q[2 * pq + o]; out[2] = Q[3 * pq + o] / (r * C.R);
    };
    double Vl[3], Vg[3], Vip[3], Vim[3], Vjp[3], Vjm[3], Vg3[3], Vg4[3];
    vvar(P, pl, oB, Vl); vvar(Ps, pls, oS, Vg);
    vvar(P, pl, off(Im + 1, Jm), Vip); vvar(P, pl, off(Im - 1, Jm), Vim);
    vvar(P, pl, off(Im, Jm + 1), Vjp); vvar(P, pl, off(Im, Jm - 1), Vjm);
    vvar(Ps, pls, oS3, Vg3); vvar(Ps, pls, oS4, Vg4);
    const double m2 = (double)modfm2, mm2 = (double)(-modfm2);
    double gl[3][2], gg[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (idir) {
        gl[k][0] = (Vg[k] - Vl[k]) * m2;   gl[k][1] = (Vjp[k] - Vjm[k]) * 0.5;
        gg[k][0] = (Vl[k] - Vg[k]) * mm2;  gg[k][1] = (Vg4[k] - Vg3[k]) * 0.5;
      } else {
        gl[k][0] = (Vip[k] - Vim[k]) * 0.5; gl[k][1] = (Vg[k] - Vl[k]) * m2;
        gg[k][0] = (Vg4[k] - Vg3[k]) * 0.5; gg[k][1] = (Vl[k] - Vg[k]) * mm2;
      }
    }
    double a[3][2], b[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) { grad_m(gl[k][0], gl[k][1], Mc, a[k][0], a[k][1]); grad_m(gg[k][0], gg[k][1], Mc, b[k][0], b[k][1]); }
    const double pm0 = (rho + gr) * 0.5, pm1 = (u + gu) * 0.5, pm2 = (v + gv) * 0.5, pm3 = (p + gpr) * 0.5;
    const double Gxl = C.source == Q2D_SRC_AREA ? B.sa[oB] : 0.0, Gyl = C.source == Q2D_SRC_AREA ? B.sb[oB] : 0.0;
    const double VG = v_dot_g_face(u, v, gu, gv, Gxl, Gyl, Gxl, Gyl);
    double Dl[3], Dg[3];
    diff_flux(C, axi, pm0, pm1, pm2, pm3, a[0][0], a[0][1], a[1][0], a[1][1], a[2][0], a[2][1], n1, n2, il, r_ave, VG, Dl);
    diff_flux(C, axi, pm0, pm1, pm2, pm3, b[0][0], b[0][1], b[1][0], b[1][1], b[2][0], b[2][1], n1, n2, il, r_ave, VG, Dg);
    double V[4];
    V[0] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) V[c + 1] = -(0.5 * m2 * (Dl[c] + Dg[c]));
    emit_bcvisc(out, dense, E.dpos, V, 0xE, C.mxk != 0);
    return;
  } else {                                           // 302
    const double T = p / (rho * C.R);
    const double mil = mu_lam(C, E.T), kl = k_lam(C, mil, E.T);
    const double m3 = (double)modfm3;
    double gU0 = 0.0, gU1 = 0.0, gV0 = 0.0, gV1 = 0.0, gT0 = 0.0, gT1 = 0.0;
    if (idir) { gU0 = u * m3; gV0 = v * m3; gT0 = (T - E.T) * m3; }
    else { gU1 = u * m3; gV1 = v * m3; gT1 = (T - E.T) * m3; }
    double gUx, gUy, gVx, gVy, gTx, gTy;
    grad_m(gU0, gU1, M, gUx, gUy); grad_m(gV0, gV1, M, gVx, gVy); grad_m(gT0, gT1, M, gTx, gTy);
    double S1, S2;
    stress_vector(axi, gUx, gUy, gVx, gVy, n1, n2, mil, r_ave, 0.0, 0.0, S1, S2);   // the wall passes no V.G
    F1 = S1 * il; F2 = S2 * il;
    F3 = il * kl * (gTx * n1 + gTy * n2);
  }
  double V[4];
  V[0] = -(sgn * 0.0);                               // Flux(1) = 0; BC 302 updates R(:), BC 300 R(nu:np)
  V[1] = -(sgn * F1); V[2] = -(sgn * F2); V[3] = -(sgn * F3);
  emit_bcvisc(out, dense, E.dpos, V, E.type == 302 ? 0xF : 0xE, C.mxk != 0);
}

// ---------------------------------------------------------------------------
// the fused RK-stage kernel
// ---------------------------------------------------------------------------
template <typename FT, int DV>
__device__ __forceinline__ void limiter4(const Gas<FT, DV>& g, int lim, const FT x[4], const FT y[4], FT o[4]) {
  switch (lim) {
    case 1:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 1, x[v], y[v]);
      break;
    case 2:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 2, x[v], y[v]);
      break;
    case 3:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 3, x[v], y[v]);
      break;
    case 4:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 4, x[v], y[v]);
      break;
    case 5:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 5, x[v], y[v]);
      break;
    case 6:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = limiter(g, 6, x[v], y[v]);
      break;
    default:
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = FT(0);
      break;
  }
}

// Convective_Flux — Lib_Convective.f90:46-80 for one face whose left/right
// cell states and limited slopes are known.  Returns true on a non-physical
// interface state (check_gas_state).
template <typename FT, int RIE, int DV>
__device__ __forceinline__ bool face_flux(const Gas<FT, DV>& g, const double qa[4], const FT pa[4], FT lm, const double qb[4],
                                          const FT pb[4], FT lp, FT bt, FT brie, FT nx, FT ny, FT il, FT out[4]) {
  FT qL[4], qR[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    qL[v] = (FT)qa[v] + (bt * pa[v]) * lm;            // priml = local + beta*slopel*dll
    qR[v] = (FT)qb[v] - (bt * pb[v]) * lp;            // primr = next  - beta*sloper*dlr
  }
  const bool bad = !(qL[0] > FT(0) && qL[3] > FT(0) && qR[0] > FT(0) && qR[3] > FT(0));
  const FT aL = sound<FT>(g, qL[3], qL[0]);
  const FT aR = sound<FT>(g, qR[3], qR[0]);
  FT F[4];
  riemann<FT, RIE>(g, qL[1], qL[2], qL[3], aL, qL[0], qR[1], qR[2], qR[3], aR, qR[0], nx, ny, F, brie);
  // one species: sp*(rho1/rotot1) - sm*(rho4/rotot4) == 1 exactly, so Flux(1) = F_r*int_l
#pragma unroll
  for (int v = 0; v < 4; ++v) out[v] = F[v] * il;
  return bad;
}

template <typename FT, int RIE_T, bool FUSE, bool VISC>
__global__ __launch_bounds__(256, Q2D_STAGE_MIN_WAVES) void k_stage(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                               DevCtrl* __restrict__ ctrl, const Item* __restrict__ items, int item_lo,
                                               int item_hi, const int* __restrict__ bcface,
                                               const BCFlux* __restrict__ bcflux, double* __restrict__ partials, int irk,
                                               int xcd_remap, unsigned long long* __restrict__ tstamp, int tslot) {
  // launch timing (only when tstamp is set): earliest workgroup start and latest wavefront end on the
  // device's constant-rate wall clock, in the slot of this launch within the current timed step
  // (tstamp[0] counts steps; slots form a ring of kTimedSteps steps).  Each slot has kStampLanes
  // sub-slots, picked by workgroup, so that the atomics of one launch spread over kStampLanes words.
  unsigned long long* tslotp = nullptr;
  if (tstamp) {
    const int launch = (int)(tstamp[0] % kTimedSteps) * kLaunchesPerStep + tslot;
    tslotp = tstamp + 2 + 2 * ((long long)launch * kStampLanes + (blockIdx.x % kStampLanes));
    if (threadIdx.x == 0) atomicMin(&tslotp[0], (unsigned long long)wall_clock64());
  }
  // RIE_T == -4: every solver of RIE_T == -2 and the cases' property tables (table mode; with FT = float
  // the faces interpolate the tables in FP32 and the update inverts e(T) in FP64, as the FP64 form does)
  constexpr bool TAB = RIE_T == -4;
  const int lane = threadIdx.x & (kWave - 1);
  const int wpg = blockDim.x / kWave;
  __shared__ BCFlux sbq[4][2][16];                     // per wavefront: face 1 and 2 boundary results, 16 rows
  const int wslot = threadIdx.x / kWave;
  int grp = blockIdx.x;
  if (xcd_remap) {                                     // consecutive logical groups on one XCD (8 XCDs, round-robin dispatch)
    const int ng = gridDim.x;
    const int per = ng / 8;
    if (per > 0 && grp < per * 8) grp = (grp % 8) * per + grp / 8;
  }
  const int wave0 = __builtin_amdgcn_readfirstlane(grp * wpg + threadIdx.x / kWave);   // wave-uniform: item and block descriptors load through the scalar cache
  const int nwaves = gridDim.x * wpg;

  for (int it = item_lo + wave0; it < item_hi; it += nwaves) {
    const Item item = items[it];
    const DevBlock& B = blocks[item.blk];
    DevCtrl& K = ctrl[B.case_id];
    const DevCase& C = cases[B.case_id];
    if (!K.active || irk > C.n_rk) continue;
    // RIE_T < 0: the cases' solvers share the launch; the item's case picks one (uniform switch).  HLLE++ is
    // instantiated only with RIE_T == -2, launched for handles that hold an HLLE++ case, so that the other
    // launches keep their register allocation (2 waves per SIMD)
    // dvc = 1: the FP64 fast-division pass (ddiv_fast).  It checks every row it computes and returns false
    // at the first row that fails a check, having committed no failure; the item is then run again with
    // the IEEE division (dvc = 0).  Rows stored before that are the IEEE pass's values already.
    auto item_body = [&](auto riec, auto dvc) -> bool {
      constexpr int RIE = decltype(riec)::value;
      constexpr int DV = decltype(dvc)::value;
      constexpr bool FD = DV != 0;
      static_assert(!FD || sizeof(FT) == 8, "the fast division is an FP64 pass");
      const bool last = (irk == C.n_rk);
      const bool irs = C.irs_beta > 0.0;                 // residual to B.irs[0]; k_irs smooths and updates
      // MUSCL-SD: faces take their cell's detector value; the detector is computed with MUSCL-SD or an
      // HLLE++ / HLLC+ solver (Assign_Setup.f90:49-53).  RIE_T == -3 (HLLC / HLLE launches of handles
      // without the detector) compiles neither.
      const bool sd = RIE_T != -3 && C.space_sd != 0;
      const bool det = RIE_T != -3 && (sd || C.riemann >= Q2D_RIEMANN_HLLEPP);
      const int ni = B.ni, nj = B.nj, pitch = B.pitch;
      const long long pl = B.plane;
      // the case's control words, read once per item (wave-uniform; a read inside the row loop would be a
      // vector load whose wait also drains the row prefetch)
      const int kcur = __builtin_amdgcn_readfirstlane(K.cur), kdtsel = __builtin_amdgcn_readfirstlane(K.dtsel);
      const int kiter = __builtin_amdgcn_readfirstlane(K.iter);
      const int bin = stage_in(kcur, irk), bout = stage_out(kcur, irk);
      GP Pin = gp(B.P[bin]);
      GP PO = gp(B.P[kcur]);
      GPW Pout = gpw(B.P[bout]);
      GP dlxp = gp(B.dlx), dlyp = gp(B.dly), surp = gp(B.sur), dtlp = gp(B.dtl[kdtsel]);
      GP fixp = gp(B.fix), fiyp = gp(B.fiy), filp = gp(B.fil), fjxp = gp(B.fjx), fjyp = gp(B.fjy), fjlp = gp(B.fjl);
      GP m11p = gp(B.m11);
      const int lim = C.limiter;
      const bool tacc = C.time_accurate != 0;
      // Fast pass checks (running minima / maxima of frexp exponents, 0 for 0):
      //  ex_div[0]  every numerator:  a = 0 or |a| >= 2^-969
      //  ex_div[1]  every quotient:   q = 0 or |q| >= 2^-1020 (so the exact quotient is normal)
      //  ex_lo / ex_hi  every state value read (stage input rows, RK base state): rho and p 0 or at
      //                 least 2^-100 in magnitude, every component below 2^100
      // With the block's geometry 0 or within [2^-100, 2^100) (checked at upload, B.fastdiv), every
      // denominator of the path is 0 or within [2^-300, 2^300] (densities, pressures, sound speeds,
      // lengths, their sums, products and cancellations, x + 1e-30 of the limiters) except the
      // limiters' 1 + |r| and 1 + r^2, whose numerators grow with them, and no quotient reaches 2^1023:
      // within these bounds ddiv_fast equals the IEEE division bit for bit.  The next step's dt (the
      // division of values just written) uses the IEEE division in both passes.
      int ex_div[2] = {0, 0}, ex_lo = 0, ex_hi = 0;
      unsigned long long fkey = ~0ull;                  // failures of the fast pass, committed when it holds
      int fcode = 0;
      auto fail = [&](int code, unsigned long long key) {
        if constexpr (FD) {
          fkey = key < fkey ? key : fkey;
          fcode = code < fcode ? code : fcode;
        } else {
          record_fail(&K, code, key);
        }
      };
      auto chk_in = [&](const double* q) {
        if constexpr (FD) {
          const int e0 = __builtin_amdgcn_frexp_exp(q[0]), e1 = __builtin_amdgcn_frexp_exp(q[1]);
          const int e2 = __builtin_amdgcn_frexp_exp(q[2]), e3 = __builtin_amdgcn_frexp_exp(q[3]);
          ex_lo = min(ex_lo, min(e0, e3));
          ex_hi = max(ex_hi, max(max(e0, e1), max(e2, e3)));
        }
      };
      Gas<FT, DV> gf{(FT)C.R, (FT)C.cp, (FT)C.cv, (FT)C.gam};
      if constexpr (TAB) { gf.tab = C.tab; gf.ntab = C.ntab; }
      gf.ex = ex_div;
      Gas<double, DV> gd{C.R, C.cp, C.cv, C.gam};
      if constexpr (TAB) { gd.tab = C.tab; gd.ntab = C.ntab; }
      gd.ex = ex_div;
      const double c_rk = (C.n_rk == 1) ? 1.0 : (C.n_rk == 2 ? (irk == 1 ? 1.0 : 0.5)
                                                              : (irk == 1 ? 1.0 : (irk == 2 ? 0.25 : 2.0 / 3.0)));
      const double dtg = K.dtglobal;
      const int stage_id = irk;

      const int x = item.x0 - 2 + lane;                  // Fortran column of this lane
      const int xc = x <= ni + 2 ? x : ni + 2;           // clamped column for loads
      const bool own = lane >= 2 && lane < 2 + kOwn && x <= ni;
      const bool xdo = lane >= 1 && lane <= 61 && x >= 1 && x <= ni - 1;   // interior face (x|x+1)
      // boundary-face results (BCFlux at each boundary cell): faces 1 and 2 of the strip's edge lanes are
      // staged 16 rows at a time in LDS (one load per lane per 16 rows instead of a load per row on the
      // edge lane's critical path); faces 3 and 4 are read by every owning lane in the first and last rows
      const bool lface = B.bcf[0] >= 0 && item.x0 == 1;
      const bool rface = B.bcf[1] >= 0 && ni < item.x0 + kOwn;
      (void)bcface;
      const int rmax = nj + 2;

      auto at = [&](int r) { return (long long)(r + 1) * pitch + (xc + 1); };

      // ---- pipeline warm-up: rows y0-2, y0-1, y0 ----------------------------
      constexpr int kPrefetchRows = sizeof(FT) == 8 ? Q2D_PREFETCH_ROWS_FP64 : Q2D_PREFETCH_ROWS_MIXED;
      double qA[4], qB[4], qC[4], qN[4], qN2[4];
      double pPrev = 0.0;                                // p of row r-1 (the detector's j-1 neighbour)
      double dlyA, dlyB, dlyC, dlyN, dlyN2 = 0.0;
      FT sAB[4], phiA[4], Gprev[4];
      RowGeo gcur;                                       // row r: i-face and update data, loaded at the top of its iteration
      double fjxN, fjyN, fjlN;                           // j-face geometry of the next row (one row ahead)
      {
        double qm[4];
        const long long o0 = at(item.y0 - 2), o1 = at(item.y0 - 1), o2 = at(item.y0), o3 = at(item.y0 + 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          qm[v] = Pin[v * pl + o0]; qA[v] = Pin[v * pl + o1]; qB[v] = Pin[v * pl + o2]; qN[v] = Pin[v * pl + o3];
        }
        if (kPrefetchRows == 2) {
          const long long o4 = at(min(item.y0 + 2, rmax));
#pragma unroll
          for (int v = 0; v < 4; ++v) qN2[v] = Pin[v * pl + o4];
          dlyN2 = dlyp[o4];
        }
        const double dlm = dlyp[o0];
        chk_in(qm); chk_in(qA); chk_in(qB);
        pPrev = qm[3];
        dlyA = dlyp[o1]; dlyB = dlyp[o2]; dlyN = dlyp[o3];
        { const long long og = at(item.y0 - 1); fjxN = fjxp[og]; fjyN = fjyp[og]; fjlN = fjlp[og]; }
        const FT l0 = (FT)(0.5 * (dlm + dlyA)), l1 = (FT)(0.5 * (dlyA + dlyB));
        FT s0[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          s0[v] = gdiv(gf, (FT)(qA[v] - qm[v]), l0);
          sAB[v] = gdiv(gf, (FT)(qB[v] - qA[v]), l1);
          Gprev[v] = FT(0);
        }
        limiter4(gf, lim, sAB, s0, phiA);
      }
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      double dtnext = 1e300;
      double Eprev[3] = {0.0, 0.0, 0.0};                 // diffusive flux of j-face r-1 (VISC)
      // Mixed laminar cases with fvisc: the diffusive face fluxes and the viscous source are evaluated here
      // in FP32 (the mixed form's face arithmetic) instead of read back from k_visc's FP64 planes: rows
      // r-1, r, r+1 with their temperature, M of rows r .. r+2 and the node radius of rows r-1 .. r+1 roll
      // with the row march.
      constexpr bool kVM = VISC && sizeof(FT) == 4;
      const bool fv = kVM && C.fvisc != 0;
      const bool vaxi = C.source == Q2D_SRC_AXI;
      VCf vP{0.f, 0.f, 0.f, 0.f, 0.f}, vA{0.f, 0.f, 0.f, 0.f, 0.f}, vB{0.f, 0.f, 0.f, 0.f, 0.f};
      float MC[4] = {0.f, 0.f, 0.f, 0.f}, MN[4] = {0.f, 0.f, 0.f, 0.f}, MF[4] = {0.f, 0.f, 0.f, 0.f};
      float nodS = 0.f, nodC = 0.f, nodN = 0.f;
      const float Rf = (float)C.R;
      const TransF trf = trans_f(C);
      auto ldM = [&](int rr, float M[4]) {
        const long long o = at(rr);
        M[0] = (float)gp(B.m11)[o]; M[1] = (float)gp(B.m21)[o]; M[2] = (float)gp(B.m12)[o]; M[3] = (float)gp(B.m22)[o];
      };
      auto is_wall1 = [&](int f) { return B.bcf[f] >= 0 && B.wallf[f] == 1; };   // fvisc: no mixed faces
      if constexpr (kVM) {
        if (fv) {
          vA = vcf_of(qA, Rf);                          // row y0-1
          vB = vcf_of(qB, Rf);                          // row y0
          ldM(item.y0 - 1, MC); ldM(item.y0, MN);
          nodC = (float)gp(B.nodr)[at(item.y0 - 1)];
        }
      }

      for (int r = item.y0 - 1; r <= item.y1; ++r) {
        if ((lface || rface) && r >= item.y0 && ((r - item.y0) & 15) == 0 && lane < 32) {
          // lanes 0-15: face 1, lanes 16-31: face 2, rows r .. r+15 (empty past the item)
          const int kq = lane & 15, fsel = lane >> 4, row = r + kq;
          const bool use = (fsel == 0 ? lface : rface) && row <= item.y1;
          BCFlux& dst = sbq[wslot][fsel][kq];
          if (use) {
            const BCFlux& src = bcflux[B.bcf[fsel] + row - 1];
#pragma unroll
            for (int v = 0; v < 4; ++v) { dst.F[v] = src.F[v]; dst.V[v] = src.V[v]; }
            dst.mask = src.mask; dst.vmask = src.vmask;
          } else {
            dst.mask = 0; dst.vmask = 0;
          }
        }
        // software pipeline: this iteration's row r+2 was loaded one iteration ago; issue row r+3 and
        // the next row's j-face geometry now, to be consumed by the next iteration.  The rest of row r's
        // data (i-face geometry, sur, dt, RK base state) is issued here and first used after the j-face
        // flux, which hides its latency without holding a second row of it in registers.
#pragma unroll
        for (int v = 0; v < 4; ++v) qC[v] = qN[v];
        chk_in(qC);
        dlyC = dlyN;
        gcur.fjx = fjxN; gcur.fjy = fjyN; gcur.fjl = fjlN;
        {
          // row r's own data first: vector-memory waits are in issue order, so data consumed in this
          // iteration is issued before the prefetch consumed in the next one
          if (r >= item.y0) {
            const long long oc = at(r);
            gcur.dlx = dlxp[oc]; gcur.fix = fixp[oc]; gcur.fiy = fiyp[oc]; gcur.fil = filp[oc]; gcur.sur = surp[oc];
            gcur.dtl = (own && !tacc) ? dtlp[oc] : 0.0;
            if (irk != 1 && !irs) {                      // stage 1 reads its RK base state as its input (qA)
#pragma unroll
              for (int v = 0; v < 4; ++v) gcur.po[v] = PO[v * pl + oc];
            }
            if (FUSE && last && !irs) {                  // M of row r for the fused dt (c11, c21, c12, c22)
#pragma unroll
              for (int v = 0; v < 4; ++v) gcur.mm[v] = m11p[v * pl + oc];
            }
          }
          if (kPrefetchRows == 2) {                      // rows r+3 (in flight since the last iteration) and r+4
#pragma unroll
            for (int v = 0; v < 4; ++v) qN[v] = qN2[v];
            dlyN = dlyN2;
            const long long oN = at(min(r + 4, rmax));
#pragma unroll
            for (int v = 0; v < 4; ++v) qN2[v] = Pin[v * pl + oN];
            dlyN2 = dlyp[oN];
          } else {
            const long long oN = at(min(r + 3, rmax));
#pragma unroll
            for (int v = 0; v < 4; ++v) qN[v] = Pin[v * pl + oN];
            dlyN = dlyp[oN];
          }
          const long long og = at(min(r + 1, rmax));
          fjxN = fjxp[og]; fjyN = fjyp[og]; fjlN = fjlp[og];
          if constexpr (kVM) {
            if (fv) { ldM(r + 2, MF); nodN = (float)gp(B.nodr)[at(r + 1)]; }   // consumed by the next iteration
          }
        }
        FT sBC[4], phiB[4], Gcur[4];
        {
          const FT lBC = (FT)(0.5 * (dlyB + dlyC));
#pragma unroll
          for (int v = 0; v < 4; ++v) sBC[v] = gdiv(gf, (FT)(qC[v] - qB[v]), lBC);
          limiter4(gf, lim, sBC, sAB, phiB);            // phi_{r+1} = rlimiter(s(r+1|r+2), s(r|r+1))
        }
        // Tramel's detector of cell (x, r) (Mod_Fluxes.f90:129-138): its i-face x|x+1 and j-face r|r+1
        FT bt = FT(1), brie = FT(1);                     // the reconstruction's and the Riemann solver's
        if (det) {
          const double pc = qA[3], pim = shup(qA[3]), pip = shdn(qA[3]);
          brie = (FT)sd_tramel(gd, pc, pim, pip, pPrev, qB[3]);
          bt = sd ? brie : FT(1);
        }
        double Ecur[3] = {0.0, 0.0, 0.0};
        if (VISC && !fv && own && r >= 1 && r <= nj - 1) {
          const long long oE = at(r);
#pragma unroll
          for (int k = 0; k < 3; ++k) Ecur[k] = gp(B.vey)[k * pl + oE];
        }
        // interior j-face (r | r+1) — Mod_Fluxes.f90:187-232
        if (own && r >= 1 && r <= nj - 1) {
          if (face_flux<FT, RIE>(gf, qA, phiA, (FT)(0.5 * dlyA), qB, phiB, (FT)(0.5 * dlyB), bt, brie, (FT)gcur.fjx,
                                 (FT)gcur.fjy, (FT)gcur.fjl, Gcur))
            fail(Q2D_E_INTERFACE, fail_key(stage_id, 0, B.blk_in_case, r, x));
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) Gcur[v] = FT(0);
        }
        if constexpr (kVM) {
          if (fv) {                                      // diffusive flux of j-face (r | r+1), FP32
            const VCf vAW = vcf_shup(vA), vAE = vcf_shdn(vA), vBW = vcf_shup(vB), vBE = vcf_shdn(vB);
            const float nodW = shup(nodC);               // the lane shifts before the lane-dependent branch
            if (own && r >= 1 && r <= nj - 1) {
              float Ma[4];
#pragma unroll
              for (int k = 0; k < 4; ++k) Ma[k] = 0.5f * (MC[k] + MN[k]);
              visc_face_f(trf, Rf, vaxi, 1, (float)gcur.fjx, (float)gcur.fjy, (float)gcur.fjl, vA, vB, vAW, vAE,
                          vBW, vBE, Ma, 0.5f * (nodC + nodW), Ecur);
            }
          }
        }

        if (r >= item.y0) {
          // ---- i-direction on row r (lanes share the stencil through DPP shifts) ----
          const double dlx = gcur.dlx;
          double qn[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) qn[v] = shdn(qA[v]);
          const double dlxn = shdn(dlx);
          FT sx[4], sxp[4], phx[4], phxn[4];
          {
            const FT l1 = (FT)(0.5 * (dlx + dlxn));
#pragma unroll
            for (int v = 0; v < 4; ++v) sx[v] = gdiv(gf, (FT)(qn[v] - qA[v]), l1);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) sxp[v] = shup(sx[v]);
          limiter4(gf, lim, sx, sxp, phx);
#pragma unroll
          for (int v = 0; v < 4; ++v) phxn[v] = shdn(phx[v]);
          FT Fx[4];
          if (xdo) {                                     // interior i-face (x|x+1) — Mod_Fluxes.f90:140-185
            if (face_flux<FT, RIE>(gf, qA, phx, (FT)(0.5 * dlx), qn, phxn, (FT)(0.5 * dlxn), bt, brie, (FT)gcur.fix,
                                   (FT)gcur.fiy, (FT)gcur.fil, Fx))
              fail(Q2D_E_INTERFACE, fail_key(stage_id, 0, B.blk_in_case, r, x));
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) Fx[v] = FT(0);
          }
          FT Fl[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) Fl[v] = shup(Fx[v]);

          double Dx[3] = {0.0, 0.0, 0.0}, Dl[3] = {0.0, 0.0, 0.0};
          double Sv[3] = {0.0, 0.0, 0.0};                // viscous source of cell (x, r) (fvisc)
          if (VISC) {                                    // diffusive flux of i-face x, and of x-1 from the left lane
            bool done = false;
            if constexpr (kVM) {
              if (fv) {
                // i-face (x | x+1) and the viscous source, as k_visc; lane 1 evaluates lane 2's face x-1.  The
                // lane shifts come before the lane-dependent branches.
                const VCf vE = vcf_shdn(vA), vW = vcf_shup(vA), vSE = vcf_shdn(vP), vNE = vcf_shdn(vB);
                float ME[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) ME[k] = shdn(MC[k]);
                if (lane >= 1 && lane <= 61 && x >= 1 && x <= ni - 1) {
                  float Ma[4];
#pragma unroll
                  for (int k = 0; k < 4; ++k) Ma[k] = 0.5f * (MC[k] + ME[k]);
                  visc_face_f(trf, Rf, vaxi, 0, (float)gcur.fix, (float)gcur.fiy, (float)gcur.fil, vA, vE, vP, vB,
                              vSE, vNE, Ma, 0.5f * (nodC + nodS), Dx);
                }
                if (own && vaxi) {                       // Source_Axi_Visc_Blk :125-236, wall ghosts reflected
                  VCf qE = vE, qW = vW, qS = vP, qN = vB;
                  const VCf refl{vA.r, -vA.u, -vA.v, vA.p, vA.T};
                  if (x == 1 && is_wall1(0)) qW = refl;
                  if (x == ni && is_wall1(1)) qE = refl;
                  if (r == 1 && is_wall1(2)) qS = refl;
                  if (r == nj && is_wall1(3)) qN = refl;
                  const long long oS = at(r);
                  visc_source_axi_f(trf, vA, qE, qW, qN, qS, MC, (float)gp(B.sa)[oS], (float)gp(B.sb)[oS], Sv);
                }
                done = true;
              }
            }
            if (!done) {
              const long long oD = at(r);
#pragma unroll
              for (int k = 0; k < 3; ++k) Dx[k] = gp(B.vdx)[k * pl + oD];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) Dl[k] = shup(Dx[k]);
          }
          if (own) {
            const long long oA = at(r);
            // ---- residual in the reference's accumulation order ------------
            // per face: convective flux (Convective_Flux), then diffusive flux (Diffusive_Flux: Res1 -= F,
            // Res2 += F; its density component is 0)
            double res[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              double rv = 0.0;
              if (x >= 2) {                                   // face x-1 (Mod_Fluxes.f90:140-185)
                rv = 0.0 - (double)Fl[v];
                if (VISC) rv = rv + (v == 0 ? 0.0 : Dl[v - 1]);
              }
              if (x <= ni - 1) {                              // face x
                rv = rv + (double)Fx[v];
                if (VISC) rv = rv - (v == 0 ? 0.0 : Dx[v - 1]);
              }
              if (r >= 2) {                                   // face r-1 (:187-232)
                rv = rv - (double)Gprev[v];
                if (VISC) rv = rv + (v == 0 ? 0.0 : Eprev[v - 1]);
              }
              if (r <= nj - 1) {                              // face r
                rv = rv + (double)Gcur[v];
                if (VISC) rv = rv - (v == 0 ? 0.0 : Ecur[v - 1]);
              }
              res[v] = rv;
            }
            // ---- BC faces 1..4 in record order (Mod_BC_Fluxes.f90:79-86): convective, then viscous ----
            auto add_bc = [&](const BCFlux& q) {
#pragma unroll
              for (int v = 0; v < 4; ++v) res[v] = ((q.mask >> v) & 1) ? res[v] + q.F[v] : res[v];
              if (VISC) {
#pragma unroll
                for (int v = 0; v < 4; ++v) res[v] = ((q.vmask >> v) & 1) ? res[v] + q.V[v] : res[v];
              }
            };
            const int kq = (r - item.y0) & 15;
            if (lface && x == 1) add_bc(sbq[wslot][0][kq]);
            if (rface && x == ni) add_bc(sbq[wslot][1][kq]);
            if (r == 1 && B.bcf[2] >= 0) add_bc(bcflux[B.bcf[2] + x - 1]);
            if (r == nj && B.bcf[3] >= 0) add_bc(bcflux[B.bcf[3] + x - 1]);
            // ---- geometric sources — Lib_Source_Axisymmetry.f90:40-78, Lib_Source_Areavariation.f90:104-144
            const double sr = gcur.sur;
            if (C.source != Q2D_SRC_NONE) {
              const double rho = qA[0], u = qA[1], v = qA[2], p = qA[3];
              const double mod_v = dsqrt_rn(u * u + v * v);
              const double E0 = int_energy(gd, gdiv(gd, p, rho * gd.R)) + 0.5 * mod_v * mod_v;
              double S[4];
              if (C.source == Q2D_SRC_AXI) {
                const double invr = gp(B.sb)[oA];
                S[0] = -(rho * v * invr);
                S[1] = -(rho * u * v * invr);
                S[2] = -(rho * v * v * invr);
                S[3] = -((rho * E0 + p) * v * invr);
              } else {
                const double VdG = u * gp(B.sa)[oA] + v * gp(B.sb)[oA];
                S[0] = rho * VdG;
                S[1] = rho * u * VdG;
                S[2] = rho * v * VdG;
                S[3] = (rho * E0 + p) * VdG;
              }
#pragma unroll
              for (int k = 0; k < 4; ++k) res[k] = res[k] - S[k] * sr;
              if (VISC) {                                 // Source_Axi_Visc_Blk / Source_Area_Visc_Blk, after the Euler source
                if (!fv) {
#pragma unroll
                  for (int k = 0; k < 3; ++k) Sv[k] = gp(B.vsv)[k * pl + oA];
                }
#pragma unroll
                for (int k = 1; k < 4; ++k) res[k] = res[k] - Sv[k - 1] * sr;
              }
            }
            const double dt = tacc ? dtg : gcur.dtl;
            if (irs) {
              // Scale_Residual_Cons (Lib_Newstate.f90:93-99); smoothing and update follow in k_irs
              GPW rs = gpw(B.irs[0]);
              if (sizeof(FT) == 8) {
#pragma unroll
                for (int v = 0; v < 4; ++v) rs[v * pl + oA] = gdiv(gd, -res[v], sr) * dt;
              } else {
                const double dts = dt * rcp64(sr);
#pragma unroll
                for (int v = 0; v < 4; ++v) rs[v * pl + oA] = -res[v] * dts;
              }
            } else {
            // ---- Compute_Cons_NoIRS — Lib_Newstate.f90:131-159 -------------
            if (irk == 1) {                              // RK base state == stage input on stage 1
#pragma unroll
              for (int v = 0; v < 4; ++v) gcur.po[v] = qA[v];
            } else {
              chk_in(gcur.po);
            }
            double Uo[4], U[4], qn2[4];
            prim2cons(gd, gcur.po[0], gcur.po[1], gcur.po[2], gcur.po[3], Uo);
            prim2cons(gd, qA[0], qA[1], qA[2], qA[3], U);
            if (sizeof(FT) == 8 || TAB) {               // (table mode: e(T) inverted on the tables)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const double rr = gdiv(gd, -res[v], sr) * dt;
                U[v] = Uo[v] + (U[v] - Uo[v] + rr) * c_rk;   // RK_stage — Lib_RK.f90:15-27
              }
              cons2prim(gd, U, qn2, TAB ? gdiv(gd, qA[3], qA[0] * gd.R) : 0.0);   // estimate: Lib_Newstate.f90:151-154
            } else {
              // mixed form: FP64 update with one reciprocal of sur and of rho
              const double dts = dt * rcp64(sr);
#pragma unroll
              for (int v = 0; v < 4; ++v) U[v] = Uo[v] + (U[v] - Uo[v] - res[v] * dts) * c_rk;
              const double ir = rcp64(U[0]);
              const double vx = U[1] * ir, vy = U[2] * ir;
              const double e = U[3] * ir - 0.5 * (vx * vx + vy * vy);
              qn2[0] = U[0]; qn2[1] = vx; qn2[2] = vy; qn2[3] = U[0] * (gd.gam - 1.0) * e;
            }
            // Check_And_Fix_State — Lib_Newstate.f90:316-350
            if (__builtin_isnan(qn2[0]) || __builtin_isnan(qn2[1]) || __builtin_isnan(qn2[2]) ||
                __builtin_isnan(qn2[3]) || qn2[3] < 0.0)
              fail(Q2D_E_STATE, fail_key(stage_id, 1, B.blk_in_case, r, x));
            if (qn2[0] < 1e-20) qn2[0] = 1e-20;
#pragma unroll
            for (int v = 0; v < 4; ++v) Pout[v * pl + oA] = qn2[v];
            if (last) {                                        // Compute_Residual — Mod_Diagnostic.f90:452-458
#pragma unroll
              for (int v = 0; v < 4; ++v) { const double d = fabs(qn2[v] - gcur.po[v]); acc[v] = acc[v] + d * d; }
            }
            if (last && FUSE) {
              // Compute_dt of the next step (Mod_dt.f90:59-104) on the state just written; the
              // next step's Compute_dt would read exactly these values.
              Gas<double> gdt{C.R, C.cp, C.cv, C.gam};        // IEEE division: dt of the state just written
              if constexpr (TAB) { gdt.tab = C.tab; gdt.ntab = C.ntab; }
              const double snd = TAB ? sound<double>(gdt, qn2[3], qn2[0]) : dsqrt_rn(ddiv_rn(C.gam * qn2[3], qn2[0]));
              double dtmin = 1e8;
              const double* mm = gcur.mm;                 // c11, c21, c12, c22
              {
                const double m1 = mm[0], m2 = mm[2];
                const double nrm = dsqrt_rn(m1 * m1 + m2 * m2);
                const double v1 = ddiv_rn(m1, nrm), v2 = ddiv_rn(m2, nrm);
                const double lambda = fabs(qn2[1] * v1 + qn2[2] * v2) + snd;
                const double dtd = ddiv_rn(dlx, lambda) * C.cfl;
                if (dtd < dtmin) dtmin = dtd;
              }
              if (C.ndir == 2) {
                const double m1 = mm[1], m2 = mm[3];
                const double nrm = dsqrt_rn(m1 * m1 + m2 * m2);
                const double v1 = ddiv_rn(m1, nrm), v2 = ddiv_rn(m2, nrm);
                const double lambda = fabs(qn2[1] * v1 + qn2[2] * v2) + snd;
                const double dtd = ddiv_rn(dlyA, lambda) * C.cfl;
                if (dtd < dtmin) dtmin = dtd;
              }
              if (VISC && C.model > 0) dtmin = vnn_limit(C, dtmin, qn2[0], qn2[3], dlx, dlyA);
              const int iter_next = kiter + 1;
              if (iter_next < C.ramp_iter) dtmin = dtmin * (double)iter_next / (double)C.ramp_iter;
              if (!tacc) gpw(B.dtl[kdtsel ^ 1])[oA] = dtmin;
              if (dtmin < dtnext) dtnext = dtmin;
            }
            }   // !irs
          }
        }
        // ---- roll the j-stencil -------------------------------------------
        pPrev = qA[3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          qA[v] = qB[v]; qB[v] = qC[v];
          sAB[v] = sBC[v]; phiA[v] = phiB[v]; Gprev[v] = Gcur[v];
        }
        if (VISC) {
#pragma unroll
          for (int k = 0; k < 3; ++k) Eprev[k] = Ecur[k];
        }
        if constexpr (kVM) {
          if (fv) {
            vP = vA; vA = vB;
            vB = vcf_of(qB, Rf);
#pragma unroll
            for (int k = 0; k < 4; ++k) { MC[k] = MN[k]; MN[k] = MF[k]; }
            nodS = nodC; nodC = nodN;
          }
        }
        dlyA = dlyB; dlyB = dlyC;
        if constexpr (FD) {
          const bool bad = ex_div[0] < -968 || ex_div[1] < -1019 || ex_lo < -99 || ex_hi > 100;
          if (__builtin_amdgcn_ballot_w64(bad) != 0ull) return false;
        }
      }
      if constexpr (FD) {
        if (fcode < 0) record_fail(&K, fcode, fkey);
      }
      if (last && !irs) {
        // fixed-pattern wavefront sum -> one partial per work item (deterministic)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double s = acc[v];
          for (int k = 32; k >= 1; k >>= 1) s = s + __shfl_xor(s, k, kWave);
          if (lane == 0) partials[(long long)it * kPartials + v] = s;
        }
      }
      if (last && FUSE && !irs) {
        // next step's dt: the item's exact min, reduced over the case's items by k_finish (no atomics)
        for (int k = 32; k >= 1; k >>= 1) {
          const double o = __shfl_xor(dtnext, k, kWave);
          if (o < dtnext) dtnext = o;
        }
        if (lane == 0) partials[(long long)it * kPartials + 4] = dtnext;
      }
      return true;
    };
    // FP64 items of blocks whose geometry is in the fast-division range run the fast pass first.  Only the
    // Euler instantiations without HLLE++ and table mode hold both passes within 256 VGPRs (2 waves per
    // SIMD); the laminar, HLLE++ and table-mode ones run the IEEE pass alone.
    constexpr bool kFastInst = sizeof(FT) == 8 && Q2D_FASTDIV && !VISC && !TAB && RIE_T != -2;
    auto run_item = [&](auto riec) {
      if constexpr (kFastInst) {
        if (B.fastdiv && item_body(riec, std::integral_constant<int, 1>{})) return;
      }
      item_body(riec, std::integral_constant<int, 0>{});
    };
    if constexpr (RIE_T == -3) {                       // HLLC and HLLE only
      if (C.riemann == Q2D_RIEMANN_HLLE) run_item(std::integral_constant<int, 1>{});
      else run_item(std::integral_constant<int, 0>{});
    } else if constexpr (RIE_T < 0) {
      if (C.riemann == Q2D_RIEMANN_HLLE) run_item(std::integral_constant<int, 1>{});
      else if (C.riemann == Q2D_RIEMANN_HLLEM) run_item(std::integral_constant<int, 2>{});
      else if ((RIE_T == -2 || TAB) && C.riemann == Q2D_RIEMANN_HLLEPP) run_item(std::integral_constant<int, 3>{});
      else if ((RIE_T == -2 || TAB) && C.riemann == Q2D_RIEMANN_HLLCP) run_item(std::integral_constant<int, 4>{});
      else run_item(std::integral_constant<int, 0>{});
    } else {
      run_item(std::integral_constant<int, RIE_T>{});
    }
  }
  if (tslotp && lane == 0) atomicMax(&tslotp[1], (unsigned long long)wall_clock64());
}

// ---------------------------------------------------------------------------
// Mixed-form stage kernel for Euler cases.  The same work items, row march, lane layout, residual
// order and update as k_stage, with the face work in FP32: the j-stencil rolls FP32 copies of the
// FP64 rows, the geometry comes from FP32 planes (one float4 load per face family per row), and the
// face flux is flux_mx.  The update stays FP64 (division-free prim2cons, one refined reciprocal for
// cons2prim, FMA), and so do the residual norm, the dt and their reductions.
//   GI(x, r) = (N(1), N(2), int_l of i-face x|x+1, dl_x / 2 of cell x)
//   GJ(x, r) = (N(1), N(2), int_l of j-face r|r+1, 1 / sur of cell (x, r))
//   HY(x, r) = dl_y / 2,  GD(x, r) = rows 1 and 2 of M, each normalised (Mod_dt.f90:76-84)
// With FUSE the last RK stage also evaluates the next step's Compute_dt (Mod_dt.f90:59-104) on the
// state it writes, so the step needs no separate dt pass.  RIE < 0 (the instantiation launched): the
// Riemann solver is picked per item from its case, so HLLC and HLLE cases share one launch per stage.
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) float* GPF;
typedef const __attribute__((address_space(1))) f4v* GP4;

// 3 waves per SIMD for the mixed stage kernel: the register budget (168 VGPRs) is enforced rather than
// left to the scheduler, which otherwise trades a wave per SIMD for ILP (0.458 -> 0.415 ms per step on
// 4 x 1024^2; a few spilled registers).
#ifndef Q2D_STAGE_MX_MIN_WAVES
#define Q2D_STAGE_MX_MIN_WAVES 3
#endif
// The laminar instantiation (VISC) holds the viscous stencil rows as well: 2 waves per SIMD.
#ifndef Q2D_STAGE_MXV_MIN_WAVES
#define Q2D_STAGE_MXV_MIN_WAVES 2
#endif
// Streaming (non-temporal) access for the data each mixed stage touches once: the new state it writes,
// the RK base state and the i-face, j-face and dt geometry it reads; the L2 keeps the state rows that
// neighbouring items re-read (their halos).  Measured on 4 x 1024^2, SSP-RK3: 0.516 -> 0.508 ms per
// step (state), -> 0.495 (geometry).
#ifndef Q2D_MX_NT_STORE
#define Q2D_MX_NT_STORE 1
#endif
#ifndef Q2D_MX_NT_LOAD
#define Q2D_MX_NT_LOAD 1
#endif
__device__ __forceinline__ f4v mx_stream_load(GP4 p) {
#if Q2D_MX_NT_LOAD
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ f4v mx_load4(const float4* p) { return *(GP4)p; }
typedef const __attribute__((address_space(1))) unsigned long long* GPF2;   // float2 planes, one 8-B load per cell
__device__ __forceinline__ float2 f2_bits(unsigned long long b) {
  return make_float2(__uint_as_float((unsigned)b), __uint_as_float((unsigned)(b >> 32)));
}
__device__ __forceinline__ float2 mx_load2(GPF2 p) { return f2_bits(*p); }
__device__ __forceinline__ float2 mx_stream_load2(GPF2 p) {
#if Q2D_MX_NT_LOAD
  return f2_bits(__builtin_nontemporal_load(p));
#else
  return f2_bits(*p);
#endif
}
// IRS: the launch holds cases with implicit residual smoothing (a separate instantiation, so that the
// register allocation of the plain kernel is not affected by the residual-store branch).
template <int RIE, bool FUSE, bool IRS, bool VISC>
__global__ __launch_bounds__(256, VISC ? Q2D_STAGE_MXV_MIN_WAVES : Q2D_STAGE_MX_MIN_WAVES) void k_stage_mx(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                                  DevCtrl* __restrict__ ctrl, const Item* __restrict__ items, int item_lo,
                                                  int item_hi, const float4* __restrict__ bcdense, double* __restrict__ partials,
                                                  int irk, int xcd_remap, unsigned long long* __restrict__ tstamp,
                                                  int tslot) {
  unsigned long long* tslotp = nullptr;
  if (tstamp) {
    const int launch = (int)(tstamp[0] % kTimedSteps) * kLaunchesPerStep + tslot;
    tslotp = tstamp + 2 + 2 * ((long long)launch * kStampLanes + (blockIdx.x % kStampLanes));
    if (threadIdx.x == 0) atomicMin(&tslotp[0], (unsigned long long)wall_clock64());
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int wpg = blockDim.x / kWave;
  __shared__ f4v sbc[4][2][16];                        // per wavefront: face 1 and 2 boundary fluxes, 16 rows
  const int wslot = threadIdx.x / kWave;
  int grp = blockIdx.x;
  if (xcd_remap) {
    const int ng = gridDim.x;
    const int per = ng / 8;
    if (per > 0 && grp < per * 8) grp = (grp % 8) * per + grp / 8;
  }
  const int wave0 = __builtin_amdgcn_readfirstlane(grp * wpg + threadIdx.x / kWave);
  const int nwaves = gridDim.x * wpg;

  for (int it = item_lo + wave0; it < item_hi; it += nwaves) {
    const Item item = items[it];
    const DevBlock& B = blocks[item.blk];
    DevCtrl& K = ctrl[B.case_id];
    const DevCase& C = cases[B.case_id];
    if (!K.active || irk > C.n_rk) continue;
    // The item body is instantiated per limiter (a uniform switch per item), so the row loop has no
    // limiter branches: straight-line code that the compiler can schedule across the whole row.
    // Direction of the row march (DIR): +1 up from row y0, -1 down from row y1 (Euler items whose Item has
    // y0 > y1, every other row band of a strip).  A band and the next march towards (or away from) their
    // common edge at the same time, so the two halo rows each reads of the other are read by both within
    // a few microseconds and come from the L2.  Every face is evaluated by the same expressions in either
    // direction (left state from its lower row, slopes as (upper - lower) / (h_lower + h_upper), limiter
    // arguments in row order), and each cell adds its faces in the same order, so the state does not
    // depend on the direction.
    auto item_body = [&](auto limc, auto riec, auto dirc) {
      constexpr int LIM = decltype(limc)::value;
      constexpr int RV = decltype(riec)::value;         // Riemann solver of this item's case
      constexpr int DIR = decltype(dirc)::value;
      constexpr bool UP = DIR > 0;
      static_assert(UP || !VISC, "the laminar instantiation marches up only");
      const int y0 = min(item.y0, item.y1), y1 = max(item.y0, item.y1);
      const bool last = (irk == C.n_rk);
      const bool irs = IRS && C.irs_beta > 0.0;          // residual to B.irs[0]; k_irs smooths and updates
      const int ni = B.ni, nj = B.nj, pitch = B.pitch;
      const long long pl = B.plane;
      const int kcur = __builtin_amdgcn_readfirstlane(K.cur), kdtsel = __builtin_amdgcn_readfirstlane(K.dtsel);
      const int kiter = __builtin_amdgcn_readfirstlane(K.iter);   // control words once per item (see k_stage)
      const int bin = stage_in(kcur, irk), bout = stage_out(kcur, irk);
      GP Pin = gp(B.P[bin]);
      GP PO = gp(B.P[kcur]);
      GPW Pout = gpw(B.P[bout]);
      GP dtlp = gp(B.dtl[kdtsel]);
      GPW dtl_next = gpw(B.dtl[kdtsel ^ 1]);
      GP4 gnp = (GP4)B.mgn, gdp = (GP4)B.mgd;
      GPF2 gsp = (GPF2)B.mgs, srcp = (GPF2)B.msrc;
      GPF hyp = (GPF)B.mhy;
      const bool tacc = C.time_accurate != 0;
      const GasF gf{(float)C.gam, (float)(C.gam - 1.0), (float)(C.gam / (C.gam - 1.0))};
      const double gm1 = C.gam - 1.0, rgm1 = 1.0 / (C.gam - 1.0);
      const double c_rk = (C.n_rk == 1) ? 1.0 : (C.n_rk == 2 ? (irk == 1 ? 1.0 : 0.5)
                                                              : (irk == 1 ? 1.0 : (irk == 2 ? 0.25 : 2.0 / 3.0)));
      const double dtg = K.dtglobal;
      const int stage_id = irk;

      const int x = item.x0 - 2 + lane;
      const int xc = x <= ni + 2 ? x : ni + 2;
      const bool own = lane >= 2 && lane < 2 + kOwn && x <= ni;
      const bool xdo = lane >= 1 && lane <= 61 && x >= 1 && x <= ni - 1;
      // boundary-face fluxes (from k_bcflux's dense per-face array): faces 1 and 2 of the strip's edge
      // lanes are staged 16 rows at a time in LDS (one load per lane per 16 rows instead of a load chain
      // per row); faces 3 and 4 are read by every owning lane in the block's first and last rows
      const bool lface = B.bcf[0] >= 0 && item.x0 == 1;
      const bool rface = B.bcf[1] >= 0 && ni < item.x0 + kOwn;
      const int rmax = nj + 2;
      auto at = [&](int r) { return (long long)(r + 1) * pitch + (xc + 1); };
      // Laminar cases (VISC): the diffusive face fluxes and the axisymmetric viscous source in FP32, with the
      // formulas of the generic laminar kernel (visc_face_f, visc_source_axi_f).  The temperatures of rows r
      // and r+1, (u, v, T) of row r-1, M of rows r and r+1 (r+2 in flight) and the node radius of rows r-1
      // and r (r+1 in flight) roll with the stencil; the i-neighbours come through DPP.  A face's diffusive
      // flux D joins its convective flux in the face's slot as F - D (its density component is 0), so the
      // residual, the boundary faces (k_bcvisc adds the records' viscous fluxes to the dense array) and the
      // update are the Euler kernel's; the viscous source joins the Euler source.
      GP4 mmp = (GP4)B.mmm;
      GPF nodp = (GPF)B.mnod;
      const float Rf = (float)C.R;
      TransF trf{0.f, 0.f, 0.f, 0.f, 0.f};
      if constexpr (VISC) trf = trans_f(C);
      const bool vaxi = C.source == Q2D_SRC_AXI;
      float TA = 0.f, TB = 0.f, TC = 0.f, Pu = 0.f, Pv = 0.f, PT = 0.f;
      f4v MA = {0.f, 0.f, 0.f, 0.f}, MB = {0.f, 0.f, 0.f, 0.f}, MF = {0.f, 0.f, 0.f, 0.f};
      float nodS = 0.f, nodA = 0.f, nodN = 0.f;

      // ---- pipeline warm-up: rows ys-2D .. ys+D, ys = y0 - 1 (up) or y1 + 1 (down) ---------------
      // A, B: the march's rows r and r+DIR; N: row r+2 DIR (in flight).  A face is named by its lower row k:
      // (k | k+1); the face of iteration r is (r | r+1) up and (r-1 | r) down.
      const int ys = UP ? y0 - 1 : y1 + 1;               // the warm-up iteration's row
      double qA[4], qB[4], qN[4];                        // FP64 rows r, r+DIR and r+2 DIR (in flight)
      float fA[4], fB[4];                                // their FP32 copies (rows r, r+DIR)
      float hyA, hyB, hyN;
      float sAB[4], phiA[4], Gprev[4];
      f4v gnN = gnp[at(UP ? ys : ys - 1)];               // face normals, one face ahead
      {
        double qm[4];
        const long long o0 = at(ys - DIR), o1 = at(ys), o2 = at(ys + DIR), o3 = at(ys + 2 * DIR);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          qm[v] = Pin[v * pl + o0]; qA[v] = Pin[v * pl + o1]; qB[v] = Pin[v * pl + o2]; qN[v] = Pin[v * pl + o3];
        }
        const float hym = hyp[o0];
        hyA = hyp[o1]; hyB = hyp[o2]; hyN = hyp[o3];
        const float i0 = frcp(hym + hyA), i1 = frcp(hyA + hyB);
        float s0[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          fA[v] = (float)qA[v]; fB[v] = (float)qB[v];
          const float fm = (float)qm[v];
          s0[v] = UP ? (fA[v] - fm) * i0 : (fm - fA[v]) * i0;          // slope between rows ys-DIR and ys
          sAB[v] = UP ? (fB[v] - fA[v]) * i1 : (fA[v] - fB[v]) * i1;   // slope between rows ys and ys+DIR
          Gprev[v] = 0.f;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) phiA[v] = UP ? limiter_mxt<LIM>(sAB[v], s0[v]) : limiter_mxt<LIM>(s0[v], sAB[v]);
      }
      if constexpr (VISC) {                              // rows y0-1 (r) and y0 (r+1) of the first j-face
        TA = fA[3] * frcp(fA[0] * Rf);
        TB = fB[3] * frcp(fB[0] * Rf);
        MA = mmp[at(y0 - 1)];
        MB = mmp[at(y0)];
        nodA = nodp[at(y0 - 1)];
      }
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      double dtnext = 1e300;
      // Failures are kept per lane (the first row of each kind) and recorded once after the rows, so
      // the row loop has no branch around them; the recorded key and status are the same minima.
      constexpr int kNoRow = 0x7fffffff;
      int fail_if_r = kNoRow, fail_st_r = kNoRow;

      // One j-face of the march, between rows r and r+DIR: the FP32 copy and slopes of row r+2 DIR (qC), the
      // limited slopes of row r+DIR and the face flux from rows r, r+DIR (fA, fB) — Mod_Fluxes.f90:187-232.
      // The flux is evaluated on every lane; lanes that own no face of this row discard it.
      float fC[4], sBC[4], phiB[4], Gcur[4];
      auto jface = [&](int r, const double (&qC)[4], float hyC, const f4v& gn) {
        const float iBC = frcp(hyB + hyC);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          fC[v] = (float)qC[v];
          sBC[v] = UP ? (fC[v] - fB[v]) * iBC : (fB[v] - fC[v]) * iBC;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) phiB[v] = UP ? limiter_mxt<LIM>(sBC[v], sAB[v]) : limiter_mxt<LIM>(sAB[v], sBC[v]);
        const int kf = UP ? r : r - 1;                   // the face's lower row
        const bool jdo = own && kf >= 1 && kf <= nj - 1;
        float L[4], R[4], G[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (UP) { L[v] = fA[v] + phiA[v] * hyA; R[v] = fB[v] - phiB[v] * hyB; }
          else { L[v] = fB[v] + phiB[v] * hyB; R[v] = fA[v] - phiA[v] * hyA; }
        }
        const bool bad = jdo && !(L[0] > 0.f && L[3] > 0.f && R[0] > 0.f && R[3] > 0.f);
        fail_if_r = bad ? min(fail_if_r, kf) : fail_if_r;
        flux_mxs<RV>(gf, L, R, gn.z, gn.w, G);
        if constexpr (VISC) {
          // diffusive flux of j-face (r | r+1) (Mod_Fluxes.f90:207-230): cells (x, r), (x, r+1) and their
          // i-neighbours, M averaged over the two cells, the face radius from nodes (x-1, r) and (x, r); the
          // scaled normal carries int_l
          TC = fC[3] * frcp(fC[0] * Rf);
          const VCf a{fA[0], fA[1], fA[2], fA[3], TA}, b{fB[0], fB[1], fB[2], fB[3], TB};
          const VCf aW = vcf_shup(a), aE = vcf_shdn(a), bW = vcf_shup(b), bE = vcf_shdn(b);
          const float nodW = shup(nodA);
          const float Ma[4] = {0.5f * (MA.x + MB.x), 0.5f * (MA.y + MB.y), 0.5f * (MA.z + MB.z), 0.5f * (MA.w + MB.w)};
          double D[3];
          visc_face_f(trf, Rf, vaxi, 1, gn.z, gn.w, 1.0f, a, b, aW, aE, bW, bE, Ma, 0.5f * (nodA + nodW), D);
#pragma unroll
          for (int k = 0; k < 3; ++k) G[k + 1] = G[k + 1] - (float)D[k];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Gcur[v] = jdo ? G[v] : 0.f;
      };
      auto roll = [&](const double (&qC)[4], float hyC) {
        if constexpr (VISC) {
          Pu = fA[1]; Pv = fA[2]; PT = TA; TA = TB; TB = TC;
          MA = MB; MB = MF; nodS = nodA; nodA = nodN;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          qA[v] = qB[v]; qB[v] = qC[v]; fA[v] = fB[v]; fB[v] = fC[v];
          sAB[v] = sBC[v]; phiA[v] = phiB[v]; Gprev[v] = Gcur[v];
        }
        hyA = hyB; hyB = hyC;
      };

      // rows of the march beyond row r+2 DIR (clamped to the ghost rows), and the lower row of the face one
      // iteration ahead
      auto far_row = [&](int r) { return UP ? min(r + 3, rmax) : max(r - 3, -1); };
      auto next_face = [&](int r) { return UP ? min(r + 1, rmax) : max(r - 2, -1); };
      {                                                  // row ys: only its j-face (y0-1 | y0) or (y1 | y1+1)
        const int r = ys;
        double qC[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) qC[v] = qN[v];
        const float hyC = hyN;
        const f4v gn = gnN;
        const long long oN = at(far_row(r));
#pragma unroll
        for (int v = 0; v < 4; ++v) qN[v] = Pin[v * pl + oN];
        hyN = hyp[oN];
        gnN = mx_stream_load(&gnp[at(next_face(r))]);
        if constexpr (VISC) { MF = mmp[at(min(r + 2, rmax))]; nodN = nodp[at(min(r + 1, rmax))]; }
        jface(r, qC, hyC, gn);
        roll(qC, hyC);
      }

      // chunks of 16 rows in march order; the edge lanes' boundary fluxes of a chunk are staged in LDS
      for (int c0 = 0; c0 <= y1 - y0; c0 += 16) {
      const int rlo = UP ? y0 + c0 : max(y0, y1 - c0 - 15);
      const int rhi = UP ? min(y1, y0 + c0 + 15) : y1 - c0;
      if ((lface || rface) && lane < 32) {
        // lanes 0-15: face 1, lanes 16-31: face 2, rows rlo .. rlo+15 (zero past the chunk)
        const int k = lane & 15, fsel = lane >> 4, row = rlo + k;
        const bool use = (fsel == 0 ? lface : rface) && row <= rhi;
        f4v q = {0.f, 0.f, 0.f, 0.f};
        if (use) q = mx_load4(&bcdense[B.bcf[fsel] + row - 1]);
        sbc[wslot][fsel][k] = q;
      }
      for (int n = 0; n <= rhi - rlo; ++n) {
        const int r = UP ? rlo + n : rhi - n;
        // Loads of this iteration, in the order they are consumed (vector-memory waits are in issue
        // order): first row r's own data (i-face geometry, RK base state, local dt, the fused dt's
        // directions; first used after the j-face flux), then row r+3 and row r+1's j-face geometry for the
        // next iteration.  Row r+2 and row r's j-face geometry and 1/sur arrived (issued one iteration ago).
        double qC[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) qC[v] = qN[v];
        const float hyC = hyN;
        const f4v gn = gnN;
        const long long oc = at(r);
        f4v gd = {0.f, 0.f, 0.f, 0.f};
        double po[4] = {0.0, 0.0, 0.0, 0.0};
        double dtl = 0.0;
        const float2 gs = mx_stream_load2(&gsp[oc]);
        if (irk != 1 && !irs) {
#if Q2D_MX_NT_LOAD
#pragma unroll
          for (int v = 0; v < 4; ++v) po[v] = __builtin_nontemporal_load(&PO[v * pl + oc]);
#else
#pragma unroll
          for (int v = 0; v < 4; ++v) po[v] = PO[v * pl + oc];
#endif
        }
        if (!tacc) dtl = dtlp[oc];
        if (FUSE && last && !irs) gd = mx_stream_load(&gdp[oc]);
        {
          const long long oN = at(far_row(r));
#pragma unroll
          for (int v = 0; v < 4; ++v) qN[v] = Pin[v * pl + oN];
          hyN = hyp[oN];
          gnN = mx_stream_load(&gnp[at(next_face(r))]);
          if constexpr (VISC) { MF = mmp[at(min(r + 2, rmax))]; nodN = nodp[at(min(r + 1, rmax))]; }
        }
        jface(r, qC, hyC, gn);

        // ---- i-direction on row r (DPP shifts between lanes) ------------------
        const float hx = gs.x, hxn = shdn(hx);
        float fn[4], sx[4], phx[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) fn[v] = shdn(fA[v]);
        const float ix = frcp(hx + hxn);
#pragma unroll
        for (int v = 0; v < 4; ++v) sx[v] = (fn[v] - fA[v]) * ix;
#pragma unroll
        for (int v = 0; v < 4; ++v) phx[v] = limiter_mxt<LIM>(sx[v], shup(sx[v]));
        float phxn[4];                                   // lane shifts with every lane active
#pragma unroll
        for (int v = 0; v < 4; ++v) phxn[v] = shdn(phx[v]);
        float Fx[4];
        {                                                // interior i-face (x|x+1) — Mod_Fluxes.f90:140-185
          float L[4], R[4], F[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) { L[v] = fA[v] + phx[v] * hx; R[v] = fn[v] - phxn[v] * hxn; }
          const bool bad = xdo && !(L[0] > 0.f && L[3] > 0.f && R[0] > 0.f && R[3] > 0.f);
          fail_if_r = bad ? min(fail_if_r, r) : fail_if_r;
          flux_mxs<RV>(gf, L, R, gn.x, gn.y, F);
          if constexpr (VISC) {
            // diffusive flux of i-face (x | x+1) (Mod_Fluxes.f90:160-183): cells (x, r), (x+1, r) and rows
            // r-1, r+1 with their x+1 neighbours, M averaged over the two cells, the face radius from nodes
            // (x, r-1) and (x, r)
            const VCf a{fA[0], fA[1], fA[2], fA[3], TA}, e = vcf_shdn(a);
            const VCf sS{0.f, Pu, Pv, 0.f, PT}, sN{fB[0], fB[1], fB[2], fB[3], TB};
            const VCf sSE = vcf_shdn(sS), sNE = vcf_shdn(sN);
            const f4v ME = {shdn(MA.x), shdn(MA.y), shdn(MA.z), shdn(MA.w)};
            const float Ma[4] = {0.5f * (MA.x + ME.x), 0.5f * (MA.y + ME.y), 0.5f * (MA.z + ME.z), 0.5f * (MA.w + ME.w)};
            double D[3];
            visc_face_f(trf, Rf, vaxi, 0, gn.x, gn.y, 1.0f, a, e, sS, sN, sSE, sNE, Ma, 0.5f * (nodA + nodS), D);
#pragma unroll
            for (int k = 0; k < 3; ++k) F[k + 1] = F[k + 1] - (float)D[k];
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) Fx[v] = xdo ? F[v] : 0.f;
        }
        float Fl[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) Fl[v] = shup(Fx[v]);

        // ---- boundary faces (BC_Fluxes, Mod_BC_Fluxes.f90:49-271, from k_bcflux): each enters the cell's
        // residual through the face-flux slot of its face (Fl, Fx, Glo, Ghi), carrying the face sign
        const bool bl = lface && x == 1, br = rface && x == ni;
        const bool bb = r == 1 && B.bcf[2] >= 0, bt = r == nj && B.bcf[3] >= 0;
        float Glo[4], Ghi[4];                            // j-faces (r-1 | r) and (r | r+1) of row r
#pragma unroll
        for (int v = 0; v < 4; ++v) { Glo[v] = UP ? Gprev[v] : Gcur[v]; Ghi[v] = UP ? Gcur[v] : Gprev[v]; }
        if (lface || rface) {                            // edge strip: faces 1 / 2 from the staged rows
          const f4v q1 = sbc[wslot][0][r - rlo], q2 = sbc[wslot][1][r - rlo];
          Fl[0] = bl ? -q1.x : Fl[0]; Fl[1] = bl ? -q1.y : Fl[1]; Fl[2] = bl ? -q1.z : Fl[2]; Fl[3] = bl ? -q1.w : Fl[3];
          Fx[0] = br ? q2.x : Fx[0]; Fx[1] = br ? q2.y : Fx[1]; Fx[2] = br ? q2.z : Fx[2]; Fx[3] = br ? q2.w : Fx[3];
        }
        if (bb) {                                        // first row: face 3
          const f4v q = mx_load4(&bcdense[B.bcf[2] + min(max(x, 1), ni) - 1]);
          Glo[0] = -q.x; Glo[1] = -q.y; Glo[2] = -q.z; Glo[3] = -q.w;
        }
        if (bt) {                                        // last row: face 4 (also the first in a one-row block)
          const f4v q = mx_load4(&bcdense[B.bcf[3] + min(max(x, 1), ni) - 1]);
          Ghi[0] = q.x; Ghi[1] = q.y; Ghi[2] = q.z; Ghi[3] = q.w;
        }

        // ---- residual of the cell (every lane; owners store) ----------
        double res[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double rv = 0.0;
          if (x >= 2 || bl) rv = rv - (double)Fl[v];
          if (x <= ni - 1 || br) rv = rv + (double)Fx[v];
          if (r >= 2 || bb) rv = rv - (double)Glo[v];
          if (r <= nj - 1 || bt) rv = rv + (double)Ghi[v];
          res[v] = rv;
        }
        // geometric sources (Lib_Source_Axisymmetry.f90:40-78, Lib_Source_Areavariation.f90:104-144),
        // per unit area: Res/sur - S enters the update
        double S[4] = {0.0, 0.0, 0.0, 0.0};
        if (C.source != Q2D_SRC_NONE) {
          const double rho = qA[0], u = qA[1], v = qA[2], p = qA[3];
          const double rhoE0 = p * rgm1 + 0.5 * rho * (u * u + v * v);
          const float2 sg = mx_load2(&srcp[oc]);                  // FP32 source geometry: (r_ave, 1/r) or (G_x, G_y)
          if (C.source == Q2D_SRC_AXI) {
            const double invr = (double)sg.y;
            S[0] = -(rho * v * invr);
            S[1] = -(rho * u * v * invr);
            S[2] = -(rho * v * v * invr);
            S[3] = -((rhoE0 + p) * v * invr);
            if constexpr (VISC) {
              // Source_Axi_Visc_Blk (Lib_Source_Axisymmetry.f90:125-236), after the Euler source; a wall face's
              // ghost is the boundary cell with its velocity reversed (Ghost_Wall_Extrapolation,
              // Lib_Ghost.f90:366-399)
              const VCf a{fA[0], fA[1], fA[2], fA[3], TA};
              VCf qE = vcf_shdn(a), qW = vcf_shup(a), qS{0.f, Pu, Pv, 0.f, PT}, qN{fB[0], fB[1], fB[2], fB[3], TB};
              const VCf refl{a.r, -a.u, -a.v, a.p, a.T};
              auto wall = [&](int f) { return B.bcf[f] >= 0 && B.wallf[f] == 1; };
              if (x == 1 && wall(0)) qW = refl;
              if (x == ni && wall(1)) qE = refl;
              if (r == 1 && wall(2)) qS = refl;
              if (r == nj && wall(3)) qN = refl;
              const float Mc[4] = {MA.x, MA.y, MA.z, MA.w};
              double Sv[3];
              visc_source_axi_f(trf, a, qE, qW, qN, qS, Mc, sg.x, sg.y, Sv);
#pragma unroll
              for (int k = 0; k < 3; ++k) S[k + 1] = S[k + 1] + Sv[k];
            }
          } else {
            const double VdG = u * (double)sg.x + v * (double)sg.y;
            S[0] = rho * VdG;
            S[1] = rho * u * VdG;
            S[2] = rho * v * VdG;
            S[3] = (rhoE0 + p) * VdG;
          }
        }
        const double dt = tacc ? dtg : dtl;
        const double dts = dt * (double)gs.y;
        if (irs) {
          // Scale_Residual_Cons (Lib_Newstate.f90:93-99); smoothing and update follow in k_irs
          if (own) {
            GPW rs = gpw(B.irs[0]);
#pragma unroll
            for (int v = 0; v < 4; ++v) rs[v * pl + oc] = __builtin_fma(-res[v], dts, S[v] * dt);
          }
        } else {
        // ---- FP64 update: RK_stage (Lib_RK.f90:15-27) on the conservative state ----
        if (irk == 1) {
#pragma unroll
          for (int v = 0; v < 4; ++v) po[v] = qA[v];
        }
        double Uo[4], U[4], qn2[4];
        Uo[0] = po[0]; Uo[1] = po[0] * po[1]; Uo[2] = po[0] * po[2];
        Uo[3] = __builtin_fma(po[3], rgm1, 0.5 * po[0] * (po[1] * po[1] + po[2] * po[2]));
        U[0] = qA[0]; U[1] = qA[0] * qA[1]; U[2] = qA[0] * qA[2];
        U[3] = __builtin_fma(qA[3], rgm1, 0.5 * qA[0] * (qA[1] * qA[1] + qA[2] * qA[2]));
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double inc = __builtin_fma(-res[v], dts, S[v] * dt);
          U[v] = __builtin_fma(U[v] - Uo[v] + inc, c_rk, Uo[v]);
        }
        {
          const double ir = rcp_nr(U[0]);
          const double vx = U[1] * ir, vy = U[2] * ir;
          qn2[0] = U[0]; qn2[1] = vx; qn2[2] = vy;
          qn2[3] = gm1 * __builtin_fma(-0.5, U[1] * vx + U[2] * vy, U[3]);
        }
        // Check_And_Fix_State — Lib_Newstate.f90:316-350
        const bool bad_state = own && (__builtin_isnan(qn2[0]) || __builtin_isnan(qn2[1]) || __builtin_isnan(qn2[2]) ||
                                       __builtin_isnan(qn2[3]) || qn2[3] < 0.0);
        fail_st_r = bad_state ? min(fail_st_r, r) : fail_st_r;
        if (qn2[0] < 1e-20) qn2[0] = 1e-20;
        if (last) {                                    // Compute_Residual — Mod_Diagnostic.f90:452-458
#pragma unroll
          for (int v = 0; v < 4; ++v) { const double d = own ? qn2[v] - po[v] : 0.0; acc[v] = __builtin_fma(d, d, acc[v]); }
        }
        double dtmin = 0.0;
        if (last && FUSE) {
          // next step's Compute_dt on the state just written (Mod_dt.f90:59-104), FP64
          dtmin = cell_dt_mx<VISC>(C, kiter + 1, qn2[0], qn2[1], qn2[2], qn2[3], hx, hyA, gd.x, gd.y, gd.z, gd.w);
          dtnext = (own && dtmin < dtnext) ? dtmin : dtnext;
        }
        if (own) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
#if Q2D_MX_NT_STORE
            __builtin_nontemporal_store(qn2[v], &Pout[v * pl + oc]);
#else
            Pout[v * pl + oc] = qn2[v];
#endif
          }
          if (last && FUSE && !tacc) dtl_next[oc] = dtmin;
        }
        }   // !irs
        roll(qC, hyC);
      }
      }
      if (fail_if_r != kNoRow) record_fail(&K, Q2D_E_INTERFACE, fail_key(stage_id, 0, B.blk_in_case, fail_if_r, x));
      if (fail_st_r != kNoRow) record_fail(&K, Q2D_E_STATE, fail_key(stage_id, 1, B.blk_in_case, fail_st_r, x));
      if (last && !irs) {                                // fixed-pattern wavefront sum (deterministic)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double s = acc[v];
          for (int k = 32; k >= 1; k >>= 1) s = s + __shfl_xor(s, k, kWave);
          if (lane == 0) partials[(long long)it * kPartials + v] = s;
        }
      }
      if (last && FUSE && !irs) {                        // next step's dt: the item's exact min (see k_stage)
        for (int k = 32; k >= 1; k >>= 1) {
          const double o = __shfl_xor(dtnext, k, kWave);
          if (o < dtnext) dtnext = o;
        }
        if (lane == 0) partials[(long long)it * kPartials + 4] = dtnext;
      }
    };
    // RIE < 0: HLLC and HLLE cases share the launch, the solver picked per item (uniform switch)
    auto by_dir = [&](auto limc, auto riec) {
      if constexpr (VISC) item_body(limc, riec, std::integral_constant<int, 1>{});
      else if (item.y0 > item.y1) item_body(limc, riec, std::integral_constant<int, -1>{});
      else item_body(limc, riec, std::integral_constant<int, 1>{});
    };
    auto by_limiter = [&](auto riec) {
      switch (C.limiter) {
        case 1: by_dir(std::integral_constant<int, 1>{}, riec); break;
        case 2: by_dir(std::integral_constant<int, 2>{}, riec); break;
        case 3: by_dir(std::integral_constant<int, 3>{}, riec); break;
        case 4: by_dir(std::integral_constant<int, 4>{}, riec); break;
        case 5: by_dir(std::integral_constant<int, 5>{}, riec); break;
        case 6: by_dir(std::integral_constant<int, 6>{}, riec); break;
        default: by_dir(std::integral_constant<int, 0>{}, riec); break;
      }
    };
    if constexpr (RIE < 0) {
      if (C.riemann == Q2D_RIEMANN_HLLE) by_limiter(std::integral_constant<int, 1>{});
      else by_limiter(std::integral_constant<int, 0>{});
    } else {
      by_limiter(std::integral_constant<int, RIE>{});
    }
  }
  if (tslotp && lane == 0) atomicMax(&tslotp[1], (unsigned long long)wall_clock64());
}

// ---------------------------------------------------------------------------
// Implicit residual smoothing — Lib_IRS.f90:17-177, for cases with irs_beta > 0 (Newstate_Conservative,
// Lib_Newstate.f90:35-64).  The stage kernel stores the scaled residual -R / sur * dt
// (Scale_Residual_Cons) in B.irs[0] instead of updating; k_irs smooths it along i, then along j, and
// applies the RK stage (Update_State_Cons_IRS, :101-127) with the stage kernel's tail: state check,
// residual-norm partials, next step's dt.  A direction is two Jacobi sweeps (Residual_Smoothing_Blk,
// :47-127):
//     s1(k) = (R(k) + beta (R(k-1) + R(k+1))) / (1 + 2 beta)     interior k  (s1 = R on a ghost)
//     s2(k) = (R(k) + beta (s1(k-1) + s1(k+1))) / (1 + 2 beta)
// evaluated in one pass from R(k-2 .. k+2).  Work items are the stage kernel's, one wavefront each: the
// wavefront marches its rows, smooths each row along i as it reads it (neighbours from the adjacent
// lanes, DPP), and keeps the i-smoothed rows r-2 .. r+2 in registers for the j-sweep of row r, so the
// i-smoothed field never goes to memory.  Ghost residuals (Ghost_Residual_Extrapolation, :130-177) are
// formed where they are read: the boundary cell's record gives the donor cell's residual for a
// connection 101 (for the j-direction, the donor's i-smoothed residual, computed there from its own
// row) and the boundary cell's own for every other type (201 included); a face without records gives
// zero (the reference never writes that ghost).
// ---------------------------------------------------------------------------
// Ghost of face f (1..4) beyond boundary cell `pos` (1-based along the face) of block B: scaled residual.
__device__ __forceinline__ void irs_ghost(const DevBlock* __restrict__ blocks, const DevBlock& B,
                                          const DevBC* __restrict__ bc, const int* __restrict__ bcface, int f,
                                          int pos, double out[4]) {
  if (B.bcf[f - 1] < 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = 0.0;
    return;
  }
  const DevBC& e = bc[bcface[B.bcf[f - 1] + pos - 1]];
  const DevBlock& D = e.type == Q2D_BC_CONNECT ? blocks[e.dblk] : B;
  const int gi = e.type == Q2D_BC_CONNECT ? e.is : e.i, gj = e.type == Q2D_BC_CONNECT ? e.js : e.j;
  GP src = gp(D.irs[0]) + (long long)(gj + 1) * D.pitch + (gi + 1);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[v] = src[v * D.plane];
}

// The i-smoothed residual of one cell (i, j) of block B, evaluated by one lane: the same expressions, in
// the same order, as the lane-parallel form in k_irs.
template <bool P2>
__device__ __forceinline__ void irs_i_at(const DevBlock* __restrict__ blocks, const DevBlock& B,
                                         const DevBC* __restrict__ bc, const int* __restrict__ bcface, double beta,
                                         double den, double inv_den, int i, int j, double out[4]) {
  auto idiv = [&](double x) { if constexpr (P2) return x * inv_den; else return x / den; };
  double q[5][4];
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    const int k = i - 2 + d;
    if (k == 0 || k == B.ni + 1) {
      irs_ghost(blocks, B, bc, bcface, k == 0 ? 1 : 2, j, q[d]);
    } else if (k >= 1 && k <= B.ni) {
      GP src = gp(B.irs[0]) + (long long)(j + 1) * B.pitch + (k + 1);
#pragma unroll
      for (int v = 0; v < 4; ++v) q[d][v] = src[v * B.plane];
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) q[d][v] = 0.0;
    }
  }
  const bool inm = i - 1 >= 1, inp = i + 1 <= B.ni;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double sm = inm ? idiv(q[1][v] + beta * (q[0][v] + q[2][v])) : q[1][v];
    const double sp = inp ? idiv(q[3][v] + beta * (q[2][v] + q[4][v])) : q[3][v];
    out[v] = idiv(q[2][v] + beta * (sm + sp));
  }
}

// Compute_dt (Mod_dt.f90:59-104) of a cell of an FP64-geometry block: the form the stage kernel fuses
// into its last stage (with the VNN limit for laminar cases).
__device__ __forceinline__ double cell_dt_fused(const DevCase& C, int iter, const double q[4], const double mm[4],
                                                double dlx, double dly) {
  const double snd = sound<double>(case_gas(C), q[3], q[0]);
  double dtmin = 1e8;
  {
    const double m1 = mm[0], m2 = mm[2];
    const double nrm = sqrt(m1 * m1 + m2 * m2);
    const double v1 = m1 / nrm, v2 = m2 / nrm;
    const double lambda = fabs(q[1] * v1 + q[2] * v2) + snd;
    const double dtd = dlx / lambda * C.cfl;
    if (dtd < dtmin) dtmin = dtd;
  }
  if (C.ndir == 2) {
    const double m1 = mm[1], m2 = mm[3];
    const double nrm = sqrt(m1 * m1 + m2 * m2);
    const double v1 = m1 / nrm, v2 = m2 / nrm;
    const double lambda = fabs(q[1] * v1 + q[2] * v2) + snd;
    const double dtd = dly / lambda * C.cfl;
    if (dtd < dtmin) dtmin = dtd;
  }
  if (C.model > 0) dtmin = vnn_limit(C, dtmin, q[0], q[3], dlx, dly);
  if (iter < C.ramp_iter) dtmin = dtmin * (double)iter / (double)C.ramp_iter;
  return dtmin;
}

__global__ __launch_bounds__(256) void k_irs(const DevBlock* __restrict__ blocks, const DevCase* __restrict__ cases,
                                             DevCtrl* __restrict__ ctrl, const Item* __restrict__ items, int item_lo,
                                             int item_hi, const DevBC* __restrict__ bc, const int* __restrict__ bcface,
                                             double* __restrict__ partials, int irk) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wpg = blockDim.x / kWave;
  const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * wpg + threadIdx.x / kWave);
  const int nwaves = gridDim.x * wpg;
  for (int it = item_lo + wave0; it < item_hi; it += nwaves) {
    const Item item = items[it];
    const DevBlock& B = blocks[item.blk];
    DevCtrl& K = ctrl[B.case_id];
    const DevCase& C = cases[B.case_id];
    if (!K.active || irk > C.n_rk || !(C.irs_beta > 0.0)) continue;
    const bool last = irk == C.n_rk;
    const double beta = C.irs_beta, den = 1.0 + 2.0 * beta;
    // den a power of two (irs-beta 0.5, every steady reference case: den = 2): x / den is exactly
    // x * (1 / den) for every x (both are the correctly rounded x 2^-k), so the sweeps multiply
    const double inv_den = 1.0 / den;
    const bool den_pow2 = (__double_as_longlong(den) & 0x000FFFFFFFFFFFFFll) == 0;
    auto item_body = [&](auto p2c) {
    constexpr bool P2 = decltype(p2c)::value;
    auto idiv = [&](double x) { if constexpr (P2) return x * inv_den; else return x / den; };
    const bool smooth_j = C.ndir == 2;                 // Residual_Smoothing's d = 1..ndir
    const int ni = B.ni, nj = B.nj, pitch = B.pitch;
    const long long pl = B.plane;
    const int kcur = __builtin_amdgcn_readfirstlane(K.cur), kdtsel = __builtin_amdgcn_readfirstlane(K.dtsel);
    const int kiter = __builtin_amdgcn_readfirstlane(K.iter);
    const int bin = stage_in(kcur, irk), bout = stage_out(kcur, irk);
    GP Pin = gp(B.P[bin]), PO = gp(B.P[kcur]);
    GPW Pout = gpw(B.P[bout]);
    GP Rs = gp(B.irs[0]);
    const bool tacc = C.time_accurate != 0;
    const bool mixed = C.precision == Q2D_PREC_MIXED;
    const Gas<double> gd = case_gas(C);
    const double c_rk = (C.n_rk == 1) ? 1.0 : (C.n_rk == 2 ? (irk == 1 ? 1.0 : 0.5)
                                                            : (irk == 1 ? 1.0 : (irk == 2 ? 0.25 : 2.0 / 3.0)));
    const int stage_id = irk;
    const int x = item.x0 - 2 + lane;                  // lanes 0..63: columns x0-2 .. x0+61
    const int xc = min(max(x, 0), ni + 1);
    const bool own = lane >= 2 && lane < 2 + kOwn && x <= ni;
    const bool inner = x >= 1 && x <= ni;
    const bool ghostx = x == 0 || x == ni + 1;
    // scaled residual of interior row k at this lane, the i-ghost at x = 0 / ni + 1
    auto raw_load = [&](int k, double q[4]) {
      if (ghostx) {
        irs_ghost(blocks, B, bc, bcface, x == 0 ? 1 : 2, k, q);
      } else {
        const long long o = (long long)(k + 1) * pitch + (xc + 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) q[v] = Rs[v * pl + o];
      }
    };
    // i-smoothing of one row (meaningful where 1 <= x <= ni); every lane takes part (DPP)
    auto isweep = [&](const double q[4], double out[4]) {
      double s1[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double qm = shup(q[v]), qp = shdn(q[v]);
        s1[v] = inner ? idiv(q[v] + beta * (qm + qp)) : q[v];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double sm = shup(s1[v]), sp = shdn(s1[v]);
        out[v] = idiv(q[v] + beta * (sm + sp));
      }
    };
    auto r1_interior = [&](int k, double out[4]) {
      double q[4];
      raw_load(k, q);
      isweep(q, out);
    };
    // row k of the i-smoothed field, ghost rows 0 and nj+1 included (k is wave-uniform)
    auto r1row = [&](int k, double out[4]) {
      if (k >= 1 && k <= nj) { r1_interior(k, out); return; }
#pragma unroll
      for (int v = 0; v < 4; ++v) out[v] = 0.0;
      if (!smooth_j || (k != 0 && k != nj + 1)) return;
      const int f = k == 0 ? 3 : 4;
      if (B.bcf[f - 1] < 0) return;                   // no records: zero
      r1_interior(k == 0 ? 1 : nj, out);              // zero-gradient: the boundary cell's own
      if (own) {
        const DevBC& e = bc[bcface[B.bcf[f - 1] + x - 1]];
        if (e.type == Q2D_BC_CONNECT) irs_i_at<P2>(blocks, blocks[e.dblk], bc, bcface, beta, den, inv_den, e.is, e.js, out);
      }
    };
    auto sweep1 = [&](int k, const double a[4], const double b[4], const double c[4], double o[4]) {
      const bool in = k >= 1 && k <= nj;
#pragma unroll
      for (int v = 0; v < 4; ++v) o[v] = in ? idiv(b[v] + beta * (a[v] + c[v])) : b[v];
    };
    double Rm[4], Rc[4], Rp[4], sA[4], sB[4];
    {
      double Rm2[4];
      if (smooth_j) { r1row(item.y0 - 2, Rm2); r1row(item.y0 - 1, Rm); }
      r1row(item.y0, Rc);
      if (smooth_j) {
        r1row(item.y0 + 1, Rp);
        sweep1(item.y0 - 1, Rm2, Rm, Rc, sA);
        sweep1(item.y0, Rm, Rc, Rp, sB);
      }
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double dtnext = 1e300;
    // software pipeline: the state of row r + 1 and the residual row r + 3 are loaded one iteration
    // ahead, so each iteration's loads are in flight during the previous one's arithmetic
    double qN[4] = {0.0, 0.0, 0.0, 0.0}, poN[4] = {0.0, 0.0, 0.0, 0.0}, rawN[4] = {0.0, 0.0, 0.0, 0.0};
    auto state_load = [&](int k, double q[4], double po[4]) {
      if (own) {
        const long long o = (long long)(k + 1) * pitch + (xc + 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) { q[v] = Pin[v * pl + o]; po[v] = PO[v * pl + o]; }
      }
    };
    state_load(item.y0, qN, poN);
    if (smooth_j && item.y0 + 2 <= nj) raw_load(item.y0 + 2, rawN);
    for (int r = item.y0; r <= item.y1; ++r) {
      const long long o = (long long)(r + 1) * pitch + (xc + 1);
      double q[4], po[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) { q[v] = qN[v]; po[v] = poN[v]; }
      if (r < item.y1) state_load(r + 1, qN, poN);
      double s2[4];
      if (smooth_j) {
        double Rp2[4], sC[4];
        if (r + 2 <= nj) {                             // interior row r + 2 from the prefetched residual
          double raw[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) raw[v] = rawN[v];
          if (r < item.y1 && r + 3 <= nj) raw_load(r + 3, rawN);
          isweep(raw, Rp2);
        } else {
          r1row(r + 2, Rp2);                           // ghost row nj + 1, or beyond
        }
        sweep1(r + 1, Rc, Rp, Rp2, sC);
#pragma unroll
        for (int v = 0; v < 4; ++v) s2[v] = idiv(Rc[v] + beta * (sA[v] + sC[v]));
#pragma unroll
        for (int v = 0; v < 4; ++v) { sA[v] = sB[v]; sB[v] = sC[v]; Rm[v] = Rc[v]; Rc[v] = Rp[v]; Rp[v] = Rp2[v]; }
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) s2[v] = Rc[v];
        if (r < item.y1) r1row(r + 1, Rc);
      }
      if (own) {
        double Uo[4], U[4], qn2[4];
        prim2cons(gd, po[0], po[1], po[2], po[3], Uo);
        prim2cons(gd, q[0], q[1], q[2], q[3], U);
#pragma unroll
        for (int v = 0; v < 4; ++v) U[v] = Uo[v] + (U[v] - Uo[v] + s2[v]) * c_rk;   // RK_stage — Lib_RK.f90:15-27
        if (!mixed || gd.tab) {                          // (table mode, both forms: e(T) inverted on the tables)
          cons2prim(gd, U, qn2, ddiv_rn(q[3], q[0] * gd.R));   // estimate: Lib_Newstate.f90:119-123
        } else {
          const double ir = rcp64(U[0]);
          const double vx = U[1] * ir, vy = U[2] * ir;
          const double e = U[3] * ir - 0.5 * (vx * vx + vy * vy);
          qn2[0] = U[0]; qn2[1] = vx; qn2[2] = vy; qn2[3] = U[0] * (gd.gam - 1.0) * e;
        }
        // Check_And_Fix_State — Lib_Newstate.f90:316-350
        if (__builtin_isnan(qn2[0]) || __builtin_isnan(qn2[1]) || __builtin_isnan(qn2[2]) ||
            __builtin_isnan(qn2[3]) || qn2[3] < 0.0)
          record_fail(&K, Q2D_E_STATE, fail_key(stage_id, 1, B.blk_in_case, r, x));
        if (qn2[0] < 1e-20) qn2[0] = 1e-20;
#pragma unroll
        for (int v = 0; v < 4; ++v) Pout[v * pl + o] = qn2[v];
        if (last) {                                    // Compute_Residual — Mod_Diagnostic.f90:452-458
#pragma unroll
          for (int v = 0; v < 4; ++v) { const double d = fabs(qn2[v] - po[v]); acc[v] = acc[v] + d * d; }
          // the next step's Compute_dt on the state just written (as the stage kernel's fused form)
          double dtmin;
          if (C.dtmx) {                                // mixed-form Euler block: its FP32 dt geometry
            const float2 gs = mx_load2((GPF2)&B.mgs[o]);
            const f4v g4 = mx_load4(&B.mgd[o]);
            dtmin = cell_dt_mx<true>(C, kiter + 1, qn2[0], qn2[1], qn2[2], qn2[3], gs.x, ((GPF)B.mhy)[o], g4.x, g4.y,
                               g4.z, g4.w);
          } else {
            const double mm[4] = {gp(B.m11)[o], gp(B.m11)[pl + o], gp(B.m11)[2 * pl + o], gp(B.m11)[3 * pl + o]};
            dtmin = cell_dt_fused(C, kiter + 1, qn2, mm, gp(B.dlx)[o], gp(B.dly)[o]);
          }
          if (!tacc) gpw(B.dtl[kdtsel ^ 1])[o] = dtmin;
          if (dtmin < dtnext) dtnext = dtmin;
        }
      }
    }
    if (last) {                                        // the stage kernel's per-item partials (deterministic)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double s = acc[v];
        for (int k = 32; k >= 1; k >>= 1) s = s + __shfl_xor(s, k, kWave);
        if (lane == 0) partials[(long long)it * kPartials + v] = s;
      }
      for (int k = 32; k >= 1; k >>= 1) {
        const double od = __shfl_xor(dtnext, k, kWave);
        if (od < dtnext) dtnext = od;
      }
      if (lane == 0) partials[(long long)it * kPartials + 4] = dtnext;
    }
    };
    if (den_pow2) item_body(std::integral_constant<bool, true>{});
    else item_body(std::integral_constant<bool, false>{});
  }
}

// Timing slots: clear every sub-slot and the step counter.
__global__ void k_tstamp_init(unsigned long long* tstamp, int nsub) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) tstamp[0] = 0ull;
  if (t < nsub) { tstamp[2 + 2 * t] = ~0ull; tstamp[2 + 2 * t + 1] = 0ull; }
}

// Per-case residual reduction and end-of-step control — Mod_Explicit.f90:120-168.
// One workgroup of kFinishThreads per case.  Thread t sums items t, t + kFinishThreads, ... in that order
// (loads of kFinishUnroll items issued together), then a fixed xor-shuffle tree per wavefront and a
// fixed tree over the wavefronts: the same order on every run (bit-stable).
constexpr int kFinishThreads = 1024, kFinishUnroll = 4;
__global__ __launch_bounds__(kFinishThreads) void k_finish(DevCtrl* ctrl, const DevCase* cases, const double* partials,
                                                           int ncases, unsigned long long* tstamp, int fused_dt) {
  const int c = blockIdx.x;
  if (tstamp && c == 0 && threadIdx.x == 0) tstamp[0] = tstamp[0] + 1ull;   // end of a timed step
  if (c >= ncases) return;
  DevCtrl& K = ctrl[c];
  const DevCase& C = cases[c];
  if (!K.active) return;
  __shared__ double sh[5][kFinishThreads / kWave];
  double s[4] = {0, 0, 0, 0};
  double dmin = 1e300;
  // a smoothed case's last stage ends in k_irs, whose items hold the partials
  const int ib = C.irs_beta > 0.0 ? C.irs_begin : C.item_begin, ie = C.irs_beta > 0.0 ? C.irs_end : C.item_end;
  for (int it0 = ib + threadIdx.x; it0 < ie; it0 += kFinishUnroll * kFinishThreads) {
    double a[kFinishUnroll][5];
#pragma unroll
    for (int u = 0; u < kFinishUnroll; ++u) {
      const int it = it0 + u * kFinishThreads;
      const bool in = it < ie;
#pragma unroll
      for (int v = 0; v < 4; ++v) a[u][v] = in ? partials[(long long)it * kPartials + v] : 0.0;
      a[u][4] = (in && fused_dt) ? partials[(long long)it * kPartials + 4] : 1e300;
    }
#pragma unroll
    for (int u = 0; u < kFinishUnroll; ++u) {
#pragma unroll
      for (int v = 0; v < 4; ++v) s[v] = s[v] + a[u][v];
      dmin = fmin(dmin, a[u][4]);
    }
  }
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    for (int k = 32; k >= 1; k >>= 1) s[v] = s[v] + __shfl_xor(s[v], k, kWave);
  for (int k = 32; k >= 1; k >>= 1) dmin = fmin(dmin, __shfl_xor(dmin, k, kWave));
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) sh[v][wv] = s[v];
    sh[4][wv] = dmin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kFinishThreads / kWave; ++w) {
#pragma unroll
      for (int v = 0; v < 4; ++v) sh[v][0] = sh[v][0] + sh[v][w];
      sh[4][0] = fmin(sh[4][0], sh[4][w]);
    }
  }
  if (threadIdx.x == 0) {
    // next step's dt from the fused last stage (Set_Global_dt's min, Mod_dt.f90:109-128)
    if (fused_dt && sh[4][0] < 1e300) K.dtmin_next_bits = (unsigned long long)__double_as_longlong(sh[4][0]);
    for (int v = 0; v < 4; ++v) K.residual[v] = sqrt(sh[v][0]);
    if (K.iter == 1) K.res_first = K.residual[0];
    const bool endsim = (K.iter_from_call >= C.itermax) || (K.residual[0] <= C.res_threshold) ||
                        (K.time >= C.time_threshold) ||
                        (C.res_drop > 0.0 && K.iter > 1 && K.residual[0] <= C.res_drop * K.res_first);
    K.todo = endsim ? 3 : 1;
    K.cur = (K.cur + 1 + ((C.n_rk - 1) & 1)) % 3;          // buffer written by the last stage
    if (endsim || K.status != 0) K.active = 0;
  }
}

// ---------------------------------------------------------------------------
// HBM reference copy (bandwidth and counter calibration): dst = src, 16 B per lane per access.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hbm_copy(const double2* __restrict__ src, double2* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg) { g_err = msg; return code; }

#define HIPCHK(x)                                                                 \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) return fail(Q2D_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

struct HostBlock {
  int ni, nj, pitch, case_id, blk_in_case;
  long long plane;
  double* arena = nullptr;       // one allocation for everything of the block
  size_t arena_bytes = 0;
};

}  // namespace

struct q2d_handle {
  int nblocks = 0, ncases = 0, nbc = 0, max_rk = 1;
  std::vector<HostBlock> hb;
  std::vector<DevBlock> db;
  std::vector<DevCase> dc;
  std::vector<DevBC> dbc;
  std::vector<int> bcface;
  std::vector<Item> items;
  std::vector<Item> irs_items;   // k_irs work items, stored after `items` in d_items (own row count)
  int irs_rows = 0;
  DevBlock* d_blocks = nullptr;
  DevCase* d_cases = nullptr;
  DevCtrl* d_ctrl = nullptr;
  DevBC* d_bc = nullptr;
  int* d_bcface = nullptr;
  Item* d_items = nullptr;
  double* d_partials = nullptr;
  BCFlux* d_bcflux = nullptr;
  float4* d_bcdense = nullptr;   // per-face convective boundary fluxes (mixed form), FP32, one float4 per cell
  int range[5] = {0, 0, 0, 0, 0}; // items sorted into [range[k], range[k+1]) by item_pass()
  std::vector<DtRange> dtr;      // dt launch: units of up to kDtCells cells of one live block
  DtRange* d_dtr = nullptr;
  int dt_total = 0;
  bool any_visc = false;
  bool any_irs = false;          // a live case smooths its residual (k_irs after each stage)
  bool any_hllepp = false;       // a live case uses HLLE++ or HLLC+: the stage kernel's RIE_T = -2 instantiation
  bool any_table = false;        // a live case uses property tables: the FP64 stage kernel's RIE_T = -4 one
  bool fuse_visc = true;         // mixed laminar cases: viscous terms in the stage kernel where possible
  bool any_unfused_visc = false; // a live laminar case needs k_visc
  std::vector<double*> case_tab; // per case: its tables on the device (table mode), else null
  bool any_det_or_hllem = false; // a live case uses HLLEM or the detector: not the RIE_T = -3 instantiation
  bool mixed_generic = false;    // mixed form with a MUSCL-SD / HLLEM / HLLE++ / HLLC+ case: those cases' Euler
                                 // items run the generic FP32-face kernel k_stage<float> (k_stage_mx has no
                                 // room for them), the others keep k_stage_mx; steps run the dt pass
  bool bc_two_phase = false;     // a live block is one cell wide: ghost fill and BC fluxes in two launches
  std::vector<int> vrec;         // boundary records with a viscous flux (BC 300/302 of model-1 cases)
  int* d_vrec = nullptr;
  unsigned long long* d_tstamp = nullptr;   // [0] step counter, then (first start, last end) per stage launch
  int timed_steps_pending = 0;               // timed steps issued since the slots were last read
  double wallclock_khz = 100000.0;
  int tslot = 0;
  std::vector<char> case_live;   // removed cases keep their index but no blocks, records or items
  bool prec_set = false;
  size_t arena_bytes = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;        // one step
  hipGraphExec_t graph_multi = nullptr;  // kGraphSteps steps back to back
  // rows per item: > 0 fixed, -1 balanced per launch, 0 the form's default (FP64: balanced; mixed: 16)
  int rows_per_item = 0, waves_per_group = 4, max_groups = 0, xcd_remap = 1;
  int rows_used[4] = {0, 0, 0, 0};   // rows per item chosen for each stage-kernel variant
  int precision_mixed = 0;
  bool dt_fused_valid = false;   // the last stage of the previous step reduced this step's dt
  bool timing = false;
  long long timed_launches = 0;
  double timed_ms = 0.0;
  int cu_count = 256;
  int device = 0;                // the device current at q2d_gpu_create
  unsigned char* noalloc_scratch = nullptr;   // that device's address of `scratch`
};

namespace {

// The item group of case c.  Mixed form: Euler cases 0 when they run k_stage_mx, 1 when they need the generic
// FP32-face kernel; laminar cases 2 (k_stage_mx's laminar instantiation) or 3 (the generic laminar kernel);
// each group is its own launch.  FP64 form: Euler cases riemann != 0, laminar 2 + (riemann != 0); groups 0
// and 1 share one launch, as do groups 2 and 3.
int item_pass(const q2d_handle* h, int c) {
  const DevCase& C = h->dc[c];
  if (h->precision_mixed) return (C.model > 0 ? 2 : 0) + (C.mxk ? 0 : 1);
  if (C.model > 0) return 2 + (C.riemann != 0 ? 1 : 0);
  return C.riemann != 0 ? 1 : 0;
}

// The stage-kernel instantiation that runs the items of group `pass` (item_pass()).
const void* stage_kernel(const q2d_handle* h, int pass) {
  if (h->precision_mixed) {
    if (pass == 0) return reinterpret_cast<const void*>(&k_stage_mx<-1, kFuseDtMx, false, false>);
    if (pass == 1) return h->any_table ? reinterpret_cast<const void*>(&k_stage<float, -4, false, false>)
                          : h->any_hllepp ? reinterpret_cast<const void*>(&k_stage<float, -2, false, false>)
                                          : reinterpret_cast<const void*>(&k_stage<float, -1, false, false>);
    if (pass == 2) return reinterpret_cast<const void*>(&k_stage_mx<-1, kFuseDtMx, false, true>);
    if (h->any_table) return reinterpret_cast<const void*>(&k_stage<float, -4, kFuseDtMx, true>);
    if (h->any_hllepp) return reinterpret_cast<const void*>(&k_stage<float, -2, kFuseDtMx, true>);
    return h->any_det_or_hllem ? reinterpret_cast<const void*>(&k_stage<float, -1, kFuseDtMx, true>)
                               : reinterpret_cast<const void*>(&k_stage<float, -3, kFuseDtMx, true>);
  }
  if (h->any_table)
    return pass < 2 ? reinterpret_cast<const void*>(&k_stage<double, -4, kFuseDtFp64, false>)
                    : reinterpret_cast<const void*>(&k_stage<double, -4, kFuseDtFp64, true>);
  if (h->any_hllepp)
    return pass < 2 ? reinterpret_cast<const void*>(&k_stage<double, -2, kFuseDtFp64, false>)
                    : reinterpret_cast<const void*>(&k_stage<double, -2, kFuseDtFp64, true>);
  if (h->any_det_or_hllem)
    return pass < 2 ? reinterpret_cast<const void*>(&k_stage<double, -1, kFuseDtFp64, false>)
                    : reinterpret_cast<const void*>(&k_stage<double, -1, kFuseDtFp64, true>);
  return pass < 2 ? reinterpret_cast<const void*>(&k_stage<double, -3, kFuseDtFp64, false>)
                  : reinterpret_cast<const void*>(&k_stage<double, -3, kFuseDtFp64, true>);
}

// The captured step graphs hold the launch arguments; they are dropped whenever those change.
void drop_graphs(q2d_handle* h) {
  if (h->graph) { (void)hipGraphExecDestroy(h->graph); h->graph = nullptr; }
  if (h->graph_multi) { (void)hipGraphExecDestroy(h->graph_multi); h->graph_multi = nullptr; }
}

// Rows per work item for one stage launch, chosen so that the launch finishes soonest on the waves the
// device holds at once.  A launch is one wavefront per item; the hardware starts a new wavefront whenever
// one ends, so the launch time is the makespan of a greedy list schedule of the items, in order, on
// `slots` wavefront slots.  An item of h rows costs h + kItemWarmRows row-times (its pipeline warm-up
// loads rows y0-2 .. y0+1 and computes one extra j-face).  Too few rows waste the warm-up; too many leave
// slots idle at the end of the launch (e.g. 4,608 items of 16 rows on 2,048 slots run in three rounds,
// the last a quarter full).  `strips` lists (60-column strips, rows) per block.  The model ignores that
// the waves of a bandwidth-bound launch share HBM (the waves of a part-full last round run faster), so it
// is one candidate of the tile autotuner, not the default: on 4 blocks of 1024^2 it picks 37 rows (FP64)
// and 25 rows (mixed), which measure 0-4 % slower than 16 rows.
constexpr int kItemWarmRows = 3;
int balanced_rows(const std::vector<std::pair<int, int>>& strips, long long slots) {
  long long strip_rows = 0;
  int max_nj = 1;
  for (auto& s : strips) { strip_rows += (long long)s.first * s.second; max_nj = std::max(max_nj, s.second); }
  if (strip_rows == 0 || slots <= 0) return 16;
  std::vector<int> cand;
  for (int H = 4; H <= 64; ++H) cand.push_back(H);
  for (int H = 72; H <= 256; H += 8) cand.push_back(H);
  for (int H = 288; H <= 1024; H += 32) cand.push_back(H);
  for (int H = 1536; H <= 8192; H += 512) cand.push_back(H);
  int best_h = 16;
  double best = 1e300;
  std::vector<long long> heap;
  for (int H : cand) {
    if (H > max_nj && H != cand.front()) break;
    long long nitems = 0;
    for (auto& s : strips) nitems += (long long)s.first * ((s.second + H - 1) / H);
    long long span = 0;
    const long long total = strip_rows + nitems * kItemWarmRows;
    long long longest = 0;
    for (auto& s : strips) longest = std::max<long long>(longest, std::min(H, s.second) + kItemWarmRows);
    if (nitems <= slots) {                        // one round: the longest item
      span = longest;
    } else if (nitems > 4 * slots) {              // many short items: the schedule is nearly even
      span = (total + slots - 1) / slots + longest;
    } else {                                      // greedy list schedule (min-heap of slot end times)
      heap.assign((size_t)slots, 0);
      std::make_heap(heap.begin(), heap.end(), std::greater<long long>());
      for (auto& s : strips)
        for (int y0 = 1; y0 <= s.second; y0 += H) {
          const long long cost = std::min(H, s.second - y0 + 1) + kItemWarmRows;
          for (int k = 0; k < s.first; ++k) {
            std::pop_heap(heap.begin(), heap.end(), std::greater<long long>());
            heap.back() += cost;
            span = std::max(span, heap.back());
            std::push_heap(heap.begin(), heap.end(), std::greater<long long>());
          }
        }
    }
    // the makespan, plus a small charge for total work (the warm-up rows also cost HBM reads)
    const double score = (double)span + 0.05 * (double)total / (double)slots;
    if (score < best - 1e-9) { best = score; best_h = H; }
  }
  return best_h;
}

int build_items(q2d_handle* h) {
  h->items.clear();
  // items grouped by item_pass() — one kernel instantiation each — then by case
  for (int pass = 0; pass < 4; ++pass) {
    h->range[pass] = (int)h->items.size();
    // Form defaults, measured on 4 x 1024^2: the FP64 stage kernel (2 waves per SIMD, issue-bound)
    // gains 3 % from balanced items; the mixed one (3 waves per SIMD, bandwidth-bound) is 4 % faster
    // with 16 rows than balanced (more, shorter items in flight), unless balanced items are shorter
    // still (a problem too small to fill the device with 16-row items).
    const bool form_default = h->rows_per_item == 0;
    int H = form_default ? -1 : h->rows_per_item;
    if (H <= 0) {                                  // balanced for this launch on this device
      std::vector<std::pair<int, int>> strips;
      for (int c = 0; c < h->ncases; ++c) {
        if (!h->case_live[c] || item_pass(h, c) != pass) continue;
        for (int b = 0; b < h->nblocks; ++b)
          if (h->hb[b].case_id == c) strips.emplace_back((h->hb[b].ni + kOwn - 1) / kOwn, h->hb[b].nj);
      }
      int per_cu = 0;
      const int tpb = kWave * h->waves_per_group;
      if (strips.empty()) H = 16;
      else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stage_kernel(h, pass), tpb, 0) != hipSuccess ||
            per_cu <= 0)
          per_cu = 2 * 4 / h->waves_per_group;    // 2 waves per SIMD
        H = balanced_rows(strips, (long long)per_cu * h->waves_per_group * h->cu_count);
      }
      if (form_default && h->precision_mixed) H = std::min(H, 16);
    }
    h->rows_used[pass] = H;
    for (int c = 0; c < h->ncases; ++c) {
      if (!h->case_live[c] || item_pass(h, c) != pass) continue;
      h->dc[c].item_begin = (int)h->items.size();
      // k_stage_mx's Euler items march every other row band of a strip down (stored with y0 > y1), so that
      // neighbouring bands read their shared halo rows at the same time
      const bool alt = h->precision_mixed && pass == 0;
      for (int b = 0; b < h->nblocks; ++b) {
        if (h->hb[b].case_id != c) continue;
        const int ni = h->hb[b].ni, nj = h->hb[b].nj;
        for (int y0 = 1, band = 0; y0 <= nj; y0 += H, ++band) {
          const int y1 = std::min(nj, y0 + H - 1);
          const bool down = alt && (band & 1);
          for (int x0 = 1; x0 <= ni; x0 += kOwn) h->items.push_back(down ? Item{b, x0, y1, y0} : Item{b, x0, y0, y1});
        }
      }
      h->dc[c].item_end = (int)h->items.size();
    }
  }
  h->range[4] = (int)h->items.size();
  // k_irs items: the same 60-column strips, rows per item balanced for k_irs's own occupancy (the stage
  // kernel's row count leaves k_irs with a part-full last round of wavefronts)
  h->irs_items.clear();
  {
    std::vector<std::pair<int, int>> strips;
    for (int b = 0; b < h->nblocks; ++b)
      if (h->hb[b].arena && h->case_live[h->hb[b].case_id] && h->dc[h->hb[b].case_id].irs_beta > 0.0)
        strips.emplace_back((h->hb[b].ni + kOwn - 1) / kOwn, h->hb[b].nj);
    int H = 16;
    if (!strips.empty()) {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_irs),
                                                       kWave * h->waves_per_group, 0) != hipSuccess || per_cu <= 0)
        per_cu = 3 * 4 / h->waves_per_group;
      H = h->rows_per_item > 0 ? h->rows_per_item
                               : balanced_rows(strips, (long long)per_cu * h->waves_per_group * h->cu_count);
    }
    h->irs_rows = H;
    const int base = (int)h->items.size();
    for (int c = 0; c < h->ncases; ++c) {
      h->dc[c].irs_begin = h->dc[c].irs_end = 0;
      if (!h->case_live[c] || !(h->dc[c].irs_beta > 0.0)) continue;
      h->dc[c].irs_begin = base + (int)h->irs_items.size();
      for (int b = 0; b < h->nblocks; ++b) {
        if (h->hb[b].case_id != c) continue;
        const int ni = h->hb[b].ni, nj = h->hb[b].nj;
        for (int y0 = 1; y0 <= nj; y0 += H)
          for (int x0 = 1; x0 <= ni; x0 += kOwn) h->irs_items.push_back(Item{b, x0, y0, std::min(nj, y0 + H - 1)});
      }
      h->dc[c].irs_end = base + (int)h->irs_items.size();
    }
  }
  h->dtr.clear();
  for (int b = 0; b < h->nblocks; ++b) {
    if (!h->hb[b].arena || !h->case_live[h->hb[b].case_id]) continue;
    const int n = h->hb[b].ni * h->hb[b].nj;
    for (int st = 0; st < n; st += kDtCells) h->dtr.push_back(DtRange{b, st, std::min(kDtCells, n - st), h->hb[b].ni});
  }
  h->dt_total = (int)h->dtr.size();
  (void)hipFree(h->d_dtr); h->d_dtr = nullptr;
  HIPCHK(hipMalloc(&h->d_dtr, sizeof(DtRange) * std::max<size_t>(1, h->dtr.size())));
  if (!h->dtr.empty())
    HIPCHK(hipMemcpy(h->d_dtr, h->dtr.data(), sizeof(DtRange) * h->dtr.size(), hipMemcpyHostToDevice));
  if (h->d_items) { (void)hipFree(h->d_items); h->d_items = nullptr; }
  if (h->d_partials) { (void)hipFree(h->d_partials); h->d_partials = nullptr; }
  const size_t nall = h->items.size() + h->irs_items.size();
  HIPCHK(hipMalloc(&h->d_items, sizeof(Item) * std::max<size_t>(1, nall)));
  HIPCHK(hipMalloc(&h->d_partials, sizeof(double) * kPartials * std::max<size_t>(1, nall)));
  HIPCHK(hipMemcpy(h->d_items, h->items.data(), sizeof(Item) * h->items.size(), hipMemcpyHostToDevice));
  if (!h->irs_items.empty())
    HIPCHK(hipMemcpy(h->d_items + h->items.size(), h->irs_items.data(), sizeof(Item) * h->irs_items.size(),
                     hipMemcpyHostToDevice));
  (void)hipFree(h->d_cases); h->d_cases = nullptr;
  HIPCHK(hipMalloc(&h->d_cases, sizeof(DevCase) * std::max<size_t>(1, h->dc.size())));
  if (!h->dc.empty())
    HIPCHK(hipMemcpy(h->d_cases, h->dc.data(), sizeof(DevCase) * h->dc.size(), hipMemcpyHostToDevice));
  drop_graphs(h);
  return Q2D_OK;
}

// Scratch of one step's kernels: the per-item residual-norm and dt partials (k_stage*, k_irs -> k_finish),
// the per-record boundary fluxes (k_bcflux / k_bcvisc -> k_stage) and, for the mixed form, the dense FP32
// per-face array (k_bcflux -> k_stage_mx; null for FP64 handles).  Every step writes each of them before
// it reads it.
struct StepScratch { double* partials; BCFlux* bcflux; float4* dense; };
StepScratch own_scratch(const q2d_handle* h) { return StepScratch{h->d_partials, h->d_bcflux, h->d_bcdense}; }

// With timing on, each stage launch is bracketed by an event pair on the solver's stream.  The timed steps
// replay a graph captured with these event nodes, so the stage kernels run back to back exactly as in
// the untimed graph; only the event nodes are added.
template <typename FT, bool FUSE>
int launch_stage(q2d_handle* h, int irk, const StepScratch& scr) {
  const int tpb = kWave * h->waves_per_group;
  for (int pass = 0; pass < 4; ++pass) {
    // FP64: groups 0-1 (Euler) and 2-3 (laminar) share one launch; mixed form: every group is its own
    // launch (0, 2: k_stage_mx; 1, 3: the generic FP32-face kernels)
    const bool split = sizeof(FT) == 4;
    if (!split && (pass == 1 || pass == 3)) continue;
    const int lo = h->range[pass], hi = split ? h->range[pass + 1] : h->range[pass + 2];
    if (hi <= lo) continue;
    int groups = (hi - lo + h->waves_per_group - 1) / h->waves_per_group;
    if (h->max_groups > 0) groups = std::min(groups, h->max_groups);
    groups = std::max(groups, 1);
    unsigned long long* ts = nullptr;
    int slot = 0;
    if (h->timing) {
      if (h->tslot >= kLaunchesPerStep) return fail(Q2D_E_HIP, "timing slots exhausted");
      ts = h->d_tstamp; slot = h->tslot++;
    }
    const int* bf = h->d_bcface;
    const BCFlux* bq = scr.bcflux;
#define Q2D_STAGE_LAUNCH(RIE, VISC)                                                                               \
    hipLaunchKernelGGL((k_stage<FT, RIE, FUSE, VISC>), dim3(groups), dim3(tpb), 0, h->stream, h->d_blocks,        \
                       h->d_cases, h->d_ctrl, h->d_items, lo, hi, bf, bq, scr.partials, irk, h->xcd_remap, ts,   \
                       slot)
#define Q2D_STAGE_LAM_MX_LAUNCH(RIE)                                                                              \
    hipLaunchKernelGGL((k_stage<float, RIE, kFuseDtMx, true>), dim3(groups), dim3(tpb), 0, h->stream, h->d_blocks,  \
                       h->d_cases, h->d_ctrl, h->d_items, lo, hi, bf, bq, scr.partials, irk, h->xcd_remap, ts,   \
                       slot)
#define Q2D_STAGE_MX_LAUNCH(RIE, IRS, VISC)                                                                       \
    hipLaunchKernelGGL((k_stage_mx<RIE, kFuseDtMx, IRS, VISC>), dim3(groups), dim3(tpb), 0, h->stream, h->d_blocks,\
                       h->d_cases, h->d_ctrl, h->d_items, lo, hi, scr.dense, scr.partials, irk, h->xcd_remap, ts,     \
                       slot)
    if constexpr (sizeof(FT) == 4) {                   // mixed form: plain Euler items run k_stage_mx
      if (pass == 0 && h->any_irs) Q2D_STAGE_MX_LAUNCH(-1, true, false);
      else if (pass == 0) Q2D_STAGE_MX_LAUNCH(-1, false, false);
      else if (pass == 2) Q2D_STAGE_MX_LAUNCH(-1, false, true);   // laminar k_stage_mx cases (no smoothing)
      else if (h->any_table && pass == 1) Q2D_STAGE_LAUNCH(-4, false);   // property tables
      else if (h->any_table) Q2D_STAGE_LAM_MX_LAUNCH(-4);
      else if (h->any_hllepp && pass == 1) Q2D_STAGE_LAUNCH(-2, false);
      else if (h->any_hllepp) Q2D_STAGE_LAM_MX_LAUNCH(-2);
      else if (pass == 1) Q2D_STAGE_LAUNCH(-1, false);
      else if (h->any_det_or_hllem) Q2D_STAGE_LAM_MX_LAUNCH(-1);
      else Q2D_STAGE_LAM_MX_LAUNCH(-3);                // laminar HLLC / HLLE cases
    } else if (h->any_table) {                         // property tables
      if (pass == 0) Q2D_STAGE_LAUNCH(-4, false);
      else Q2D_STAGE_LAUNCH(-4, true);
    } else if (h->any_hllepp) {                        // HLLE++ / HLLC+ (plain FP64 form only)
      if (pass == 0) Q2D_STAGE_LAUNCH(-2, false);
      else Q2D_STAGE_LAUNCH(-2, true);
    } else if (h->any_det_or_hllem) {                  // HLLEM or the detector
      if (pass == 0) Q2D_STAGE_LAUNCH(-1, false);
      else Q2D_STAGE_LAUNCH(-1, true);
    } else {                                           // HLLC and HLLE only
      if (pass == 0) Q2D_STAGE_LAUNCH(-3, false);
      else Q2D_STAGE_LAUNCH(-3, true);
    }
#undef Q2D_STAGE_LAUNCH
#undef Q2D_STAGE_LAM_MX_LAUNCH
#undef Q2D_STAGE_MX_LAUNCH
  }
  return Q2D_OK;
}

// Clear the launch-timing slots of the first `steps` steps of the ring (all by default) and the step
// counter (a kernel, so it is ordered on the stream).
int reset_timing(q2d_handle* h, int steps = kTimedSteps) {
  const int nsub = std::min(steps, kTimedSteps) * kLaunchesPerStep * kStampLanes;
  hipLaunchKernelGGL(k_tstamp_init, dim3(std::max(1, (nsub + 255) / 256)), dim3(256), 0, h->stream,
                     h->d_tstamp, nsub);
  HIPCHK(hipGetLastError());
  h->timed_steps_pending = 0;
  return Q2D_OK;
}

// Add the durations of the stage launches of the steps timed since the last reset, then reset.
int collect_timing(q2d_handle* h) {
  const int nsteps = h->timed_steps_pending;
  if (nsteps == 0) return Q2D_OK;
  // only the slots of the pending steps were written since the last reset: copy and clear those
  const int nst = std::min(nsteps, kTimedSteps);
  std::vector<unsigned long long> ts(2 + 2 * (size_t)nst * kLaunchesPerStep * kStampLanes);
  HIPCHK(hipMemcpyAsync(ts.data(), h->d_tstamp, sizeof(unsigned long long) * ts.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int k = 0; k < std::min(nsteps, kTimedSteps) * kLaunchesPerStep; ++k) {
    unsigned long long t0 = ~0ull, t1 = 0ull;
    for (int l = 0; l < kStampLanes; ++l) {
      t0 = std::min(t0, ts[2 + 2 * ((size_t)k * kStampLanes + l)]);
      t1 = std::max(t1, ts[2 + 2 * ((size_t)k * kStampLanes + l) + 1]);
    }
    if (t1 < t0) continue;                            // unused slot, or a launch with no work
    h->timed_ms += (double)(t1 - t0) / h->wallclock_khz;
    h->timed_launches += 1;
  }
  return reset_timing(h, nst);
}

// Steps of this handle evaluate the next step's dt in their last stage; otherwise every step runs the dt
// pass.  Every FP64 stage kernel fuses it (the laminar ones with the VNN limit); in the mixed form
// k_stage_mx and the laminar k_stage<float> do, the generic Euler k_stage<float> does not (its cases'
// dt formula is k_stage_mx's, from FP32 geometry).
bool fused_handle(const q2d_handle* h) {
  return h->precision_mixed ? (kFuseDtMx && !h->mixed_generic) : kFuseDtFp64;
}

int launch_step(q2d_handle* h, const StepScratch& scr) {
  if (h->ncases == 0) return Q2D_OK;
  h->tslot = 0;
  const int n_items = (int)h->items.size();
  const int tpb = kWave * h->waves_per_group;
  int groups = (n_items + h->waves_per_group - 1) / h->waves_per_group;
  if (h->max_groups > 0) groups = std::min(groups, h->max_groups);
  groups = std::max(groups, 1);
  const int cthreads = 64, cgroups = (h->ncases + cthreads - 1) / cthreads;
  const int bgroups = (h->nbc + 63) / 64;            // one wavefront per group: spread over every CU
  const int use_next = (h->dt_fused_valid && fused_handle(h)) ? 1 : 0;
  // with the dt already reduced, the first boundary launch also does k_begin's work (one launch fewer)
  const bool fold_begin = use_next && h->nbc > 0;
  if (!fold_begin)
    hipLaunchKernelGGL(k_begin, dim3(cgroups), dim3(cthreads), 0, h->stream, h->d_ctrl, h->d_cases, h->ncases, use_next);
  if (!use_next)
    if (h->dt_total > 0) {                             // one wavefront per unit of kDtCells cells
      const int dg = (h->dt_total + 3) / 4;
      hipLaunchKernelGGL(k_dt, dim3(dg), dim3(256), 0, h->stream, h->d_blocks, h->d_cases, h->d_ctrl, h->d_dtr,
                         h->dt_total, 0);
    }
  if (!use_next)
    hipLaunchKernelGGL(k_dt_finish, dim3(cgroups), dim3(cthreads), 0, h->stream, h->d_ctrl, h->d_cases, h->ncases);
  for (int irk = 1; irk <= h->max_rk; ++irk) {
    if (h->nbc > 0) {
      for (int ph = h->bc_two_phase ? 1 : 0; ph <= (h->bc_two_phase ? 2 : 0); ++ph) {
        const int begin_cases = (fold_begin && irk == 1 && ph <= 1) ? h->ncases : 0;
        if (h->precision_mixed)                        // the dense FP32 face array feeds k_stage_mx
          hipLaunchKernelGGL(k_bcflux<float>, dim3(bgroups), dim3(64), 0, h->stream, h->d_blocks, h->d_cases,
                             h->d_ctrl, h->d_bc, h->nbc, irk, scr.bcflux, ph, begin_cases, scr.dense);
        else
          hipLaunchKernelGGL(k_bcflux<double>, dim3(bgroups), dim3(64), 0, h->stream, h->d_blocks, h->d_cases,
                             h->d_ctrl, h->d_bc, h->nbc, irk, scr.bcflux, ph, begin_cases, scr.dense);
      }
    }
    if (!h->vrec.empty())
      hipLaunchKernelGGL(k_bcvisc, dim3(((int)h->vrec.size() + 255) / 256), dim3(256), 0, h->stream, h->d_blocks,
                         h->d_cases, h->d_ctrl, h->d_bc, h->d_vrec, (int)h->vrec.size(), irk, scr.bcflux, scr.dense);
    if (h->range[4] > h->range[2] && h->any_unfused_visc) {
      int vg = (h->range[4] - h->range[2] + h->waves_per_group - 1) / h->waves_per_group;
      if (h->max_groups > 0) vg = std::min(vg, h->max_groups);
      hipLaunchKernelGGL(k_visc, dim3(std::max(vg, 1)), dim3(tpb), 0, h->stream, h->d_blocks, h->d_cases, h->d_ctrl,
                         h->d_items, h->range[2], h->range[4], h->d_bc, h->d_bcface, irk);
    }
    int rc = h->precision_mixed ? launch_stage<float, false>(h, irk, scr) : launch_stage<double, kFuseDtFp64>(h, irk, scr);
    if (rc != Q2D_OK) return rc;
    if (h->any_irs && !h->irs_items.empty()) {        // residual smoothing along i and j, then the update
      const int n_irs = (int)h->irs_items.size();
      int ig = (n_irs + h->waves_per_group - 1) / h->waves_per_group;
      if (h->max_groups > 0) ig = std::min(ig, h->max_groups);
      hipLaunchKernelGGL(k_irs, dim3(std::max(ig, 1)), dim3(tpb), 0, h->stream, h->d_blocks, h->d_cases, h->d_ctrl,
                         h->d_items, n_items, n_items + n_irs, h->d_bc, h->d_bcface, scr.partials, irk);
    }
  }
  const int fused_dt = fused_handle(h) ? 1 : 0;
  hipLaunchKernelGGL(k_finish, dim3(h->ncases), dim3(kFinishThreads), 0, h->stream, h->d_ctrl, h->d_cases, scr.partials, h->ncases,
                     h->timing ? h->d_tstamp : nullptr, fused_dt);
  HIPCHK(hipGetLastError());
  // the next step may take its dt from this step's last stage only if this step fused it (a handle that
  // loses its last laminar or generic-kernel case switches to fused steps with a dt pass first)
  h->dt_fused_valid = fused_dt != 0;
  return Q2D_OK;
}

int read_out(q2d_handle* h, q2d_step_out* out, int* first_err) {
  *first_err = Q2D_OK;
  if (h->ncases == 0) return Q2D_OK;
  constexpr int kChunk = 64;                         // control words read in chunks of a stack buffer
  DevCtrl ctl[kChunk];
  for (int c = 0; c < h->ncases; ++c) {
    if (c % kChunk == 0) {
      const int n = std::min(kChunk, h->ncases - c);
      HIPCHK(hipMemcpyAsync(ctl, h->d_ctrl + c, sizeof(DevCtrl) * n, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
    }
    DevCtrl& k = ctl[c % kChunk];
    // The reported failure is the first in the reference's order (stage, then fluxes before the new
    // state, then block, row, column): the minimum key; its kind bit names the stop it maps to.
    if (k.status != 0) k.status = ((k.fail_key >> 41) & 1ull) ? Q2D_E_STATE : Q2D_E_INTERFACE;
    if (h->case_live[c] && k.status != 0 && *first_err == Q2D_OK) *first_err = k.status;
    if (!out) continue;
    q2d_step_out& o = out[c];
    memset(&o, 0, sizeof(o));
    if (!h->case_live[c]) continue;               // removed: all zero (todo 0)
    o.dtglobal = k.dtglobal; o.time = k.time;
    for (int v = 0; v < 4; ++v) o.residual[v] = k.residual[v];
    o.iter = k.iter; o.todo = k.todo; o.status = k.status;
    if (k.status != 0) {
      o.fail_b = (int)((k.fail_key >> 32) & 0xff);
      o.fail_j = (int)((k.fail_key >> 16) & 0xffff);
      o.fail_i = (int)(k.fail_key & 0xffff);
    }
  }
  return Q2D_OK;
}

// Fortran layout (nprim fastest) -> SoA planes of the device layout
void pack_state(const HostBlock& b, const double* P, std::vector<double>& out) {
  const int ngi = b.ni + 4, ngj = b.nj + 4;
  out.assign((size_t)b.plane * 4, 0.0);
  for (int j = 0; j < ngj; ++j)
    for (int i = 0; i < ngi; ++i)
      for (int v = 0; v < 4; ++v) out[(size_t)v * b.plane + (size_t)j * b.pitch + i] = P[((size_t)j * ngi + i) * 4 + v];
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* q2d_version(void) { return "q2d-mi355x 0.1 (gfx950)"; }
const char* q2d_last_error(void) { return g_err.c_str(); }

int q2d_device_count(int* count) {
  if (!count) return fail(Q2D_E_ARG, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) { *count = 0; return fail(Q2D_E_NODEVICE, hipGetErrorString(e)); }
  *count = n;
  return Q2D_OK;
}

int q2d_gpu_destroy(q2d_handle* h) {
  if (!h) return Q2D_OK;
  drop_graphs(h);
  for (auto& b : h->hb) if (b.arena) (void)hipFree(b.arena);
  (void)hipFree(h->d_blocks); (void)hipFree(h->d_cases); (void)hipFree(h->d_ctrl); (void)hipFree(h->d_bc);
  (void)hipFree(h->d_bcface); (void)hipFree(h->d_items); (void)hipFree(h->d_partials); (void)hipFree(h->d_bcflux);
  (void)hipFree(h->d_bcdense);
  (void)hipFree(h->d_tstamp);
  (void)hipFree(h->d_vrec); (void)hipFree(h->d_dtr);
  for (double* t : h->case_tab) if (t) (void)hipFree(t);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return Q2D_OK;
}

namespace {

// Rebuild the per-face record index (bcface, DevBlock::bcf), the viscous-record list and the work items
// from the live blocks and records, and upload every descriptor array.  Called after cases are added or
// removed; the captured step graph is invalidated (build_items).
int rebuild_descriptors(q2d_handle* h) {
  h->bcface.clear();
  h->vrec.clear();
  std::vector<std::vector<int>> face_list((size_t)h->nblocks * 4);
  for (int b = 0; b < h->nblocks; ++b) {
    for (int f = 0; f < 4; ++f) h->db[b].bcf[f] = -1;
    if (!h->hb[b].arena) continue;
    face_list[b * 4 + 0].assign(h->hb[b].nj, -1);
    face_list[b * 4 + 1].assign(h->hb[b].nj, -1);
    face_list[b * 4 + 2].assign(h->hb[b].ni, -1);
    face_list[b * 4 + 3].assign(h->hb[b].ni, -1);
  }
  h->nbc = (int)h->dbc.size();
  for (int k = 0; k < h->nbc; ++k) {
    const DevBC& r = h->dbc[k];
    std::vector<int>& fl = face_list[r.blk * 4 + (r.f - 1)];
    const int pos = r.f <= 2 ? r.j - 1 : r.i - 1;
    if (fl[pos] >= 0) return fail(Q2D_E_ARG, "two bc records for one boundary face");
    fl[pos] = k;
    if (h->dc[h->hb[r.blk].case_id].model > 0 && (r.type == 300 || r.type == 302 || r.type == 101 || r.type == 201))
      h->vrec.push_back(k);
  }
  for (int b = 0; b < h->nblocks; ++b)
    for (int f = 0; f < 4; ++f) {
      std::vector<int>& fl = face_list[b * 4 + f];
      bool any = false;
      for (int v : fl) any |= (v >= 0);
      if (!any) continue;
      for (int v : fl) if (v < 0) return fail(Q2D_E_ARG, "boundary face partially covered by bc records");
      h->db[b].bcf[f] = (int)h->bcface.size();
      h->bcface.insert(h->bcface.end(), fl.begin(), fl.end());
      int nwall = 0;
      for (int v : fl) nwall += (h->dbc[v].type >= 301 && h->dbc[v].type <= 304) ? 1 : 0;
      h->db[b].wallf[f] = nwall == 0 ? 0 : (nwall == (int)fl.size() ? 1 : 2);
    }
  for (DevBC& r : h->dbc) r.dpos = h->db[r.blk].bcf[r.f - 1] + (r.f <= 2 ? r.j - 1 : r.i - 1);
  // mixed laminar cases whose stage kernel evaluates the viscous terms itself: no area-variation source
  // (its V.G terms stay in k_visc) and no face with both wall and other records
  h->any_unfused_visc = false;
  for (int c = 0; c < h->ncases; ++c) {
    DevCase& k = h->dc[c];
    bool fv = h->fuse_visc && h->precision_mixed && k.model > 0 && k.source != Q2D_SRC_AREA;
    for (int b = 0; b < h->nblocks && fv; ++b)
      if (h->hb[b].case_id == c)
        for (int f = 0; f < 4; ++f) fv = fv && h->db[b].wallf[f] != 2;
    fv = fv && k.tab == nullptr;                       // table transport: k_visc
    k.fvisc = fv ? 1 : 0;
    if (h->case_live[c] && k.model > 0 && !fv) h->any_unfused_visc = true;
    // table-mode cases run the generic FP32-face kernels, with the dt pass's FP64 table formula
    k.mxk = h->precision_mixed && k.tab == nullptr &&
            mx_kernel_case(k.model, k.space_sd, k.riemann, k.fvisc, k.irs_beta) ? 1 : 0;
    k.dtmx = h->precision_mixed && k.tab == nullptr && (k.model == 0 || k.mxk) ? 1 : 0;
  }
  h->max_rk = 1;
  h->any_visc = false;
  h->any_irs = false;
  h->any_hllepp = false;
  h->any_table = false;
  h->any_det_or_hllem = false;
  h->mixed_generic = false;
  h->bc_two_phase = false;
  for (int b = 0; b < h->nblocks; ++b)
    if (h->hb[b].arena && (h->hb[b].ni < 2 || h->hb[b].nj < 2)) h->bc_two_phase = true;
  for (int c = 0; c < h->ncases; ++c) {
    if (!h->case_live[c]) continue;
    h->max_rk = std::max(h->max_rk, h->dc[c].n_rk);
    if (h->dc[c].model > 0) h->any_visc = true;
    if (h->dc[c].irs_beta > 0.0) h->any_irs = true;
    if (h->dc[c].riemann >= Q2D_RIEMANN_HLLEPP) h->any_hllepp = true;
    if (h->dc[c].tab) h->any_table = true;
    if (h->dc[c].riemann >= Q2D_RIEMANN_HLLEM || h->dc[c].space_sd) h->any_det_or_hllem = true;
    if (h->precision_mixed && (h->dc[c].space_sd || h->dc[c].riemann >= Q2D_RIEMANN_HLLEM || h->dc[c].tab))
      h->mixed_generic = true;
    if (h->dc[c].space_sd || h->dc[c].riemann >= Q2D_RIEMANN_HLLEPP)
      h->bc_two_phase = true;                          // the connection faces' detector reads other records' ghosts
  }
  // device arrays, re-allocated at the current sizes (descriptors are small)
  auto realloc_upload = [&](void** dptr, const void* src, size_t bytes) -> int {
    (void)hipFree(*dptr); *dptr = nullptr;
    HIPCHK(hipMalloc(dptr, std::max<size_t>(bytes, 16)));
    if (bytes) HIPCHK(hipMemcpy(*dptr, src, bytes, hipMemcpyHostToDevice));
    return Q2D_OK;
  };
  int rc;
  if ((rc = realloc_upload((void**)&h->d_blocks, h->db.data(), sizeof(DevBlock) * h->db.size())) != Q2D_OK) return rc;
  if ((rc = realloc_upload((void**)&h->d_bc, h->dbc.data(), sizeof(DevBC) * h->dbc.size())) != Q2D_OK) return rc;
  if ((rc = realloc_upload((void**)&h->d_bcface, h->bcface.data(), sizeof(int) * h->bcface.size())) != Q2D_OK) return rc;
  if ((rc = realloc_upload((void**)&h->d_vrec, h->vrec.data(), sizeof(int) * h->vrec.size())) != Q2D_OK) return rc;
  (void)hipFree(h->d_bcflux); h->d_bcflux = nullptr;
  HIPCHK(hipMalloc(&h->d_bcflux, sizeof(BCFlux) * std::max<size_t>(1, h->dbc.size())));
  (void)hipFree(h->d_bcdense); h->d_bcdense = nullptr;
  if (h->precision_mixed) HIPCHK(hipMalloc(&h->d_bcdense, sizeof(float4) * std::max<size_t>(1, h->bcface.size())));
  return build_items(h);     // also uploads d_cases
}

// Validate `prob` and append its cases, blocks and records to the handle.
int append_cases(q2d_handle* h, const q2d_problem* prob, int32_t* first_case) {
  if (prob->nblocks < 0 || prob->ncases < 0 || prob->nbc < 0) return fail(Q2D_E_ARG, "negative count");
  if ((prob->nblocks > 0 && !prob->blocks) || (prob->ncases > 0 && !prob->cases) || (prob->nbc > 0 && !prob->bc))
    return fail(Q2D_E_ARG, "NULL array");
  if ((prob->nblocks == 0) != (prob->ncases == 0)) return fail(Q2D_E_ARG, "cases without blocks or blocks without cases");
  if ((long long)h->ncases + prob->ncases > kMaxCases) return fail(Q2D_E_ARG, "too many cases");
  if (first_case) *first_case = h->ncases;
  if (prob->ncases == 0) return Q2D_OK;

  // ---- validation (nothing is changed before it passes) ----
  std::vector<int> first_block(prob->ncases, -1), nblk(prob->ncases, 0);
  for (int b = 0; b < prob->nblocks; ++b) {
    const q2d_block_desc& d = prob->blocks[b];
    if (d.case_id < 0 || d.case_id >= prob->ncases) return fail(Q2D_E_ARG, "block case_id out of range");
    if (d.ni < 1 || d.nj < 1 || d.ni > 60000 || d.nj > 60000) return fail(Q2D_E_ARG, "block dimensions out of range");
    if (!d.P || !d.dl || !d.M || !d.sur || !d.face_i || !d.face_j) return fail(Q2D_E_ARG, "missing block array");
    if (first_block[d.case_id] < 0) first_block[d.case_id] = b;
    else if (b != first_block[d.case_id] + nblk[d.case_id]) return fail(Q2D_E_ARG, "blocks of a case must be contiguous");
    nblk[d.case_id]++;
  }
  const int prec = h->prec_set ? (h->precision_mixed ? Q2D_PREC_MIXED : Q2D_PREC_FP64) : prob->cases[0].scheme.precision;
  for (int c = 0; c < prob->ncases; ++c) {
    const q2d_scheme& s = prob->cases[c].scheme;
    if (nblk[c] == 0) return fail(Q2D_E_ARG, "case without blocks");
    if (nblk[c] > 255) return fail(Q2D_E_ARG, "more than 255 blocks in a case");
    if (s.n_rk < 1 || s.n_rk > 3) return fail(Q2D_E_ARG, "n_rk must be 1, 2 or 3");
    if (s.limiter < 0 || s.limiter > 6) return fail(Q2D_E_ARG, "unknown limiter");
    if (s.riemann < 0 || s.riemann > 4) return fail(Q2D_E_ARG, "unknown Riemann solver");

    if (s.model < 0 || s.model > 1) return fail(Q2D_E_UNSUPPORTED, "only the Euler (0) and laminar (1) models are available");
    if (s.model == 1 && !(s.mu_ref > 0)) return fail(Q2D_E_ARG, "laminar model needs mu_ref > 0");
    if (s.model == 1 && (!(s.T_ref > 0) || !(s.Pr > 0) || !(s.S_suth >= 0)))
      return fail(Q2D_E_ARG, "laminar model needs T_ref > 0, S >= 0, Pr > 0");
    if (s.source < 0 || s.source > 2) return fail(Q2D_E_ARG, "unknown source");
    if (s.precision < 0 || s.precision > 1) return fail(Q2D_E_ARG, "unknown precision");
    const q2d_case_desc& cdv = prob->cases[c];
    if (!(s.gas_R > 0) || (!cdv.thermo && !(s.gas_cp > s.gas_R))) return fail(Q2D_E_ARG, "gas constants invalid");
    if (cdv.thermo) {                                  // table mode: e(T) = h(T) - R T must increase
      if (cdv.thermo_n < 2) return fail(Q2D_E_ARG, "thermo tables need thermo_n >= 2");
      for (int k = 0; k < 4 * cdv.thermo_n; ++k)
        if (!std::isfinite(cdv.thermo[k])) return fail(Q2D_E_ARG, "thermo tables hold a non-finite value");
      for (int k = 0; k + 1 < cdv.thermo_n; ++k)
        if (!((cdv.thermo[k + 1] - cdv.thermo[k]) - s.gas_R > 0.0))
          return fail(Q2D_E_ARG, "thermo tables: h(T) - R T must increase on every interval");
    }
    if (s.precision != prec) return fail(Q2D_E_ARG, "all cases of a handle must use one precision");
    if (prob->cases[c].res_drop < 0 || prob->cases[c].res_drop >= 1) return fail(Q2D_E_ARG, "res_drop must be in [0, 1)");
    if (!(s.irs_beta >= 0) || !std::isfinite(s.irs_beta)) return fail(Q2D_E_ARG, "irs_beta must be finite and >= 0");
    if (s.space_sd != 0 && s.space_sd != 1) return fail(Q2D_E_ARG, "space_sd must be 0 or 1");
    if (s.integration_vars == Q2D_INTVARS_PRIM)
      return fail(Q2D_E_UNSUPPORTED, "integration-variables = prim (Newstate_Primitive) is not available on this path");
    if (s.integration_vars != Q2D_INTVARS_CONS) return fail(Q2D_E_ARG, "unknown integration_vars");
    if (!std::isfinite(prob->cases[c].time0)) return fail(Q2D_E_ARG, "time0 must be finite");

    for (int k = 0; k < nblk[c]; ++k) {
      const q2d_block_desc& d = prob->blocks[first_block[c] + k];
      if (s.source == Q2D_SRC_AXI && !d.src_b) return fail(Q2D_E_ARG, "axisymmetric source needs src_b (axis_inv_r)");
      if (s.source == Q2D_SRC_AREA && (!d.src_a || !d.src_b)) return fail(Q2D_E_ARG, "area source needs src_a/src_b");
      if (s.source == Q2D_SRC_AXI && s.model == 1 && (!d.src_a || !d.node))
        return fail(Q2D_E_ARG, "axisymmetric laminar model needs src_a (axis_r_ave) and node");
    }
  }
  const int case0 = h->ncases, block0 = h->nblocks;
  std::vector<DevBC> newbc(prob->nbc);
  std::vector<std::vector<char>> seen(prob->nblocks * 4);
  for (int b = 0; b < prob->nblocks; ++b)
    for (int f = 0; f < 4; ++f) seen[b * 4 + f].assign(f < 2 ? prob->blocks[b].nj : prob->blocks[b].ni, 0);
  int prev_case = -1, prev_blk = 0, prev_face = 0;
  for (int k = 0; k < prob->nbc; ++k) {
    const q2d_bc_entry& e = prob->bc[k];
    if (e.case_id < 0 || e.case_id >= prob->ncases) return fail(Q2D_E_ARG, "bc case_id out of range");
    const int c = e.case_id;
    if (e.b < 1 || e.b > nblk[c] || e.f < 1 || e.f > 4) return fail(Q2D_E_ARG, "bc block/face out of range");
    // records ordered by case, block, face (Mod_BC_Fluxes.f90:79-86)
    if (c < prev_case || (c == prev_case && (e.b < prev_blk || (e.b == prev_blk && e.f < prev_face))))
      return fail(Q2D_E_ARG, "bc records must be ordered by case, block and face");
    prev_case = c; prev_blk = e.b; prev_face = e.f;
    const int lb = first_block[c] + e.b - 1;
    const q2d_block_desc& d = prob->blocks[lb];
    const bool iface = e.f <= 2;
    if (iface) {
      if (e.j < 1 || e.j > d.nj || e.i != (e.f == 1 ? 1 : d.ni)) return fail(Q2D_E_ARG, "bc cell not on its face");
    } else {
      if (e.i < 1 || e.i > d.ni || e.j != (e.f == 3 ? 1 : d.nj)) return fail(Q2D_E_ARG, "bc cell not on its face");
    }
    char& sn = seen[lb * 4 + (e.f - 1)][iface ? e.j - 1 : e.i - 1];
    if (sn) return fail(Q2D_E_ARG, "two bc records for one boundary face");
    sn = 1;
    switch (e.type) {
      case 101: case 201: case 300: case 302: case 400: case 405: case 406:
        break;
      default: return fail(Q2D_E_UNSUPPORTED, "bc type not available on this path");
    }
    if (prob->cases[c].scheme.model > 0 && (e.type == 101 || e.type == 201) && e.d11 != 1 && e.d11 != -1)
      return fail(Q2D_E_ARG, "laminar block connection needs d11 = +1 or -1 (the donor's orientation along the face)");
    DevBC& r = newbc[k];
    memset(&r, 0, sizeof(r));
    r.type = e.type; r.blk = block0 + lb; r.i = e.i; r.j = e.j; r.f = e.f; r.cid = case0 + c;
    if (e.type == 101 || e.type == 201) {
      if (e.bs < 1 || e.bs > nblk[c] || e.fs < 1 || e.fs > 4) return fail(Q2D_E_ARG, "bc donor out of range");
      const q2d_block_desc& sd = prob->blocks[first_block[c] + e.bs - 1];
      // donor cells Is + guide(Fs)*(g-1), g = 1..gc, must be interior cells of the donor
      static const int gd[5][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}};
      for (int g = 1; g <= 2; ++g) {
        int ii = e.is + gd[e.fs][0] * (g - 1), jj = e.js + gd[e.fs][1] * (g - 1);
        if (ii < 1 || ii > sd.ni || jj < 1 || jj > sd.nj) return fail(Q2D_E_ARG, "bc donor cell outside donor block");
      }
      r.dblk = block0 + first_block[c] + e.bs - 1; r.is = e.is; r.js = e.js; r.fs = e.fs;
      r.pad0 = e.d11;                                 // orientation of the donor along the face (d11)
    }
    r.mach = e.mach; r.T = e.T; r.p = e.p; r.rel_fac = e.rel_fac;
    if (e.type == 405) {                                // Compute_Inflow_Direction (Lib_BC_Fluxes_Inflow.f90:548-560)
      double n1, n2;
      if (iface) { int fi = e.i - (e.f % 2); size_t so = ((size_t)(e.j - 1) * (d.ni + 1) + fi) * 3; n1 = d.face_i[so]; n2 = d.face_i[so + 1]; }
      else { int fj = e.j - (e.f % 2); size_t so = ((size_t)fj * d.ni + (e.i - 1)) * 3; n1 = d.face_j[so]; n2 = d.face_j[so + 1]; }
      double a = (e.alpha >= 0.5 * DBL_MAX) ? atan(n2 / (n1 + 1e-20)) : e.alpha;
      r.cos_a = cos(a); r.sin_a = sin(a);
    }
  }
  for (int b = 0; b < prob->nblocks; ++b)
    for (int f = 0; f < 4; ++f) {
      bool any = false, all = true;
      for (char v : seen[b * 4 + f]) { any |= v != 0; all &= v != 0; }
      if (any && !all) return fail(Q2D_E_ARG, "boundary face partially covered by bc records");
    }

  // ---- per-case constants and control ----
  if (!h->prec_set) { h->precision_mixed = prec == Q2D_PREC_MIXED; h->prec_set = true; }
  for (int c = 0; c < prob->ncases; ++c) {
    const q2d_case_desc& cd = prob->cases[c];
    const q2d_scheme& s = cd.scheme;
    DevCase k;
    memset(&k, 0, sizeof(k));
    k.n_rk = s.n_rk; k.limiter = s.limiter; k.riemann = s.riemann; k.time_accurate = s.time_accurate;
    k.model = s.model; k.source = s.source; k.precision = s.precision; k.ramp_iter = s.ramp_iter;
    k.ndir = (prob->blocks[first_block[c]].nj == 1) ? 1 : 2;     // Check_Mesh_Type (Lib_Metrics.f90:24-37)
    k.itermax = cd.itermax; k.cfl = s.cfl; k.vnn = s.vnn;
    k.R = s.gas_R; k.cp = s.gas_cp; k.cv = s.gas_cp - s.gas_R; k.gam = s.gas_cp / k.cv;
    double* dtab = nullptr;
    if (cd.thermo) {                                   // table mode: the case's device copy of its tables
      const size_t tb = sizeof(double) * 4 * (size_t)cd.thermo_n;
      HIPCHK(hipMalloc(&dtab, tb));
      HIPCHK(hipMemcpy(dtab, cd.thermo, tb, hipMemcpyHostToDevice));
      k.tab = dtab; k.ntab = cd.thermo_n;
      if (!(s.gas_cp > s.gas_R)) {                     // closed-form fields unused; kept finite
        k.cp = cd.thermo[cd.thermo_n + 1]; k.cv = k.cp - k.R; k.gam = k.cp / k.cv;
      }
    }
    h->case_tab.push_back(dtab);
    k.res_threshold = cd.res_threshold; k.time_threshold = cd.time_threshold; k.res_drop = cd.res_drop;
    k.mu_ref = s.mu_ref; k.T_ref = s.T_ref; k.S_suth = s.S_suth; k.Pr = s.Pr;
    k.irs_beta = s.irs_beta;
    k.space_sd = s.space_sd;
    h->dc.push_back(k);
    h->case_live.push_back(1);
  }
  {
    DevCtrl* nc = nullptr;
    HIPCHK(hipMalloc(&nc, sizeof(DevCtrl) * (h->ncases + prob->ncases)));
    if (h->ncases > 0) {
      HIPCHK(hipStreamSynchronize(h->stream));
      HIPCHK(hipMemcpy(nc, h->d_ctrl, sizeof(DevCtrl) * h->ncases, hipMemcpyDeviceToDevice));
    }
    std::vector<DevCtrl> ctl(prob->ncases);
    for (int c = 0; c < prob->ncases; ++c) {
      DevCtrl& k = ctl[c];
      memset(&k, 0, sizeof(k)); k.active = 1; k.todo = 1; k.fail_key = ~0ull;
      k.time = prob->cases[c].time0;                   // domain%time at set-up (Wrap_Setup.f90:129-152)
    }
    HIPCHK(hipMemcpy(nc + h->ncases, ctl.data(), sizeof(DevCtrl) * prob->ncases, hipMemcpyHostToDevice));
    (void)hipFree(h->d_ctrl);
    h->d_ctrl = nc;
  }
  h->ncases += prob->ncases;

  // ---- blocks: SoA planes in one arena per block ----
  // planes 0-11: 3 state buffers x 4; 12-24 geometry; 25, 28 local dt; 26, 27 source terms; 29 node radius;
  // viscous blocks add 30-38: diffusive fluxes of i- and j-faces and the viscous source, 3 planes each;
  // mixed-form blocks add 7 planes of FP32 geometry after those (30-36 Euler, 39-45 laminar) and laminar
  // ones 3 more (46-47: M, 48: node radius, FP32); blocks of IRS cases add 4 after all: the stage's scaled
  // residual
  for (int lb = 0; lb < prob->nblocks; ++lb) {
    const q2d_block_desc& d = prob->blocks[lb];
    const bool visc = prob->cases[d.case_id].scheme.model > 0;
    const bool mxgeo = prec == Q2D_PREC_MIXED;            // FP32 geometry planes of k_stage_mx
    const bool mxvis = mxgeo && visc;                     // and its laminar instantiation's M and node radius
    const int g0 = visc ? 39 : 30;                        // first FP32 geometry plane
    const bool irs = prob->cases[d.case_id].scheme.irs_beta > 0;
    const int irs_plane = mxvis ? 49 : (visc ? 39 : (mxgeo ? 37 : 30));
    const int kPlanes = irs_plane + (irs ? 4 : 0);
    h->hb.emplace_back();
    h->db.emplace_back();
    h->nblocks += 1;
    HostBlock& hb = h->hb.back();
    hb.ni = d.ni; hb.nj = d.nj; hb.case_id = case0 + d.case_id; hb.blk_in_case = lb - first_block[d.case_id] + 1;
    hb.pitch = ((d.ni + 4) + 15) / 16 * 16;
    hb.plane = (long long)hb.pitch * (d.nj + 4);
    hb.arena_bytes = sizeof(double) * (size_t)hb.plane * kPlanes;
    if (hipMalloc(&hb.arena, hb.arena_bytes) != hipSuccess) { hb.arena = nullptr; return fail(Q2D_E_NOMEM, "block allocation failed"); }
    h->arena_bytes += hb.arena_bytes;
    const long long pl = hb.plane;
    const int ngi = d.ni + 4, ngj = d.nj + 4, pitch = hb.pitch;
    std::vector<double> host((size_t)pl * kPlanes, 0.0);
    auto plane = [&](int k) { return host.data() + (size_t)k * pl; };
    // state buffer 0 (planes 0..3), buffers 1, 2 initialised to the same values
    for (int j = 0; j < ngj; ++j)
      for (int i = 0; i < ngi; ++i)
        for (int v = 0; v < 4; ++v) {
          double val = d.P[((size_t)j * ngi + i) * 4 + v];
          for (int sb = 0; sb < 3; ++sb) plane(4 * sb + v)[(size_t)j * pitch + i] = val;
        }
    // geometry over the ghost extent
    for (int j = 0; j < ngj; ++j)
      for (int i = 0; i < ngi; ++i) {
        const size_t src = (size_t)j * ngi + i, dst = (size_t)j * pitch + i;
        plane(12)[dst] = d.dl[src * 2 + 0];
        plane(13)[dst] = d.dl[src * 2 + 1];
        plane(14)[dst] = d.sur[src];
        plane(15)[dst] = d.M[src * 4 + 0];   // c(1,1)
        plane(16)[dst] = d.M[src * 4 + 1];   // c(2,1)
        plane(17)[dst] = d.M[src * 4 + 2];   // c(1,2)
        plane(18)[dst] = d.M[src * 4 + 3];   // c(2,2)
      }
    // i-faces f(0:ni,1:nj) at cell (i,j); j-faces f(1:ni,0:nj) at cell (i,j)
    for (int j = 1; j <= d.nj; ++j)
      for (int i = 0; i <= d.ni; ++i) {
        const size_t src = ((size_t)(j - 1) * (d.ni + 1) + i) * 3, dst = (size_t)(j + 1) * pitch + (i + 1);
        plane(19)[dst] = d.face_i[src]; plane(20)[dst] = d.face_i[src + 1]; plane(21)[dst] = d.face_i[src + 2];
      }
    for (int j = 0; j <= d.nj; ++j)
      for (int i = 1; i <= d.ni; ++i) {
        const size_t src = ((size_t)j * d.ni + (i - 1)) * 3, dst = (size_t)(j + 1) * pitch + (i + 1);
        plane(22)[dst] = d.face_j[src]; plane(23)[dst] = d.face_j[src + 1]; plane(24)[dst] = d.face_j[src + 2];
      }
    for (int j = 1; j <= d.nj; ++j)
      for (int i = 1; i <= d.ni; ++i) {
        const size_t src = (size_t)(j - 1) * d.ni + (i - 1), dst = (size_t)(j + 1) * pitch + (i + 1);
        if (d.src_a) plane(26)[dst] = d.src_a[src];
        if (d.src_b) plane(27)[dst] = d.src_b[src];
      }
    if (d.node)                                       // node(0:ni, 0:nj)%c(2) at cell position (i, j)
      for (int j = 0; j <= d.nj; ++j)
        for (int i = 0; i <= d.ni; ++i)
          plane(29)[(size_t)(j + 1) * pitch + (i + 1)] = d.node[((size_t)j * (d.ni + 1) + i) * 2 + 1];
    if (mxgeo) {
      // planes g0, g0+1: GN (float4: int_l * N of the i-face x|x+1 and of the j-face r|r+1), g0+2: GS (float2:
      // dl_x / 2, 1 / sur), g0+3, g0+4: GD (float4), g0+5: HY (float), g0+6: SRC (float2: source geometry),
      // over the ghost extent; laminar blocks: g0+7, g0+8: MM (float4: c11, c21, c12, c22), g0+9: NOD (float)
      float* gn = reinterpret_cast<float*>(plane(g0));
      float* gs = reinterpret_cast<float*>(plane(g0 + 2));
      float* gd = reinterpret_cast<float*>(plane(g0 + 3));
      float* hy = reinterpret_cast<float*>(plane(g0 + 5));
      float* sg = reinterpret_cast<float*>(plane(g0 + 6));
      float* mm = mxvis ? reinterpret_cast<float*>(plane(g0 + 7)) : nullptr;
      float* nd = mxvis ? reinterpret_cast<float*>(plane(g0 + 9)) : nullptr;
      for (int j = 0; j < ngj; ++j)
        for (int i = 0; i < ngi; ++i) {
          const size_t c = (size_t)j * pitch + i;
          gn[4 * c + 0] = (float)(plane(21)[c] * plane(19)[c]); gn[4 * c + 1] = (float)(plane(21)[c] * plane(20)[c]);
          gn[4 * c + 2] = (float)(plane(24)[c] * plane(22)[c]); gn[4 * c + 3] = (float)(plane(24)[c] * plane(23)[c]);
          const double sr = plane(14)[c];
          gs[2 * c + 0] = (float)(0.5 * plane(12)[c]); gs[2 * c + 1] = sr != 0.0 ? (float)(1.0 / sr) : 0.f;
          const double c11 = plane(15)[c], c21 = plane(16)[c], c12 = plane(17)[c], c22 = plane(18)[c];
          const double n1 = sqrt(c11 * c11 + c12 * c12), n2 = sqrt(c21 * c21 + c22 * c22);
          gd[4 * c + 0] = n1 > 0.0 ? (float)(c11 / n1) : 0.f; gd[4 * c + 1] = n1 > 0.0 ? (float)(c12 / n1) : 0.f;
          gd[4 * c + 2] = n2 > 0.0 ? (float)(c21 / n2) : 0.f; gd[4 * c + 3] = n2 > 0.0 ? (float)(c22 / n2) : 0.f;
          hy[c] = (float)(0.5 * plane(13)[c]);
          sg[2 * c + 0] = (float)plane(26)[c]; sg[2 * c + 1] = (float)plane(27)[c];
          if (mxvis) {
            mm[4 * c + 0] = (float)c11; mm[4 * c + 1] = (float)c21; mm[4 * c + 2] = (float)c12; mm[4 * c + 3] = (float)c22;
            nd[c] = (float)plane(29)[c];
          }
        }
    }
    // the FP64 stage kernel's fast-division pass: every geometry value it reads (lengths, area, face
    // normals and lengths, source geometry) is 0 or within [2^-100, 2^100) in magnitude
    bool geo_fast = true;
    for (int k : {12, 13, 14, 19, 20, 21, 22, 23, 24, 26, 27}) {
      const double* g = plane(k);
      for (long long c = 0; c < pl && geo_fast; ++c) {
        const double a = fabs(g[c]);
        if (!(a == 0.0 || (a >= 0x1p-100 && a < 0x1p100))) geo_fast = false;
      }
    }
    HIPCHK(hipMemcpy(hb.arena, host.data(), hb.arena_bytes, hipMemcpyHostToDevice));
    DevBlock& db = h->db.back();
    memset(&db, 0, sizeof(db));
    db.fastdiv = geo_fast ? 1 : 0;
    db.ni = d.ni; db.nj = d.nj; db.pitch = pitch; db.case_id = hb.case_id; db.blk_in_case = hb.blk_in_case;
    db.plane = pl;
    for (int sb = 0; sb < 3; ++sb) db.P[sb] = hb.arena + (size_t)4 * sb * pl;
    double* a = hb.arena;
    db.dlx = a + 12 * pl; db.dly = a + 13 * pl; db.sur = a + 14 * pl;
    db.m11 = a + 15 * pl; db.m21 = a + 16 * pl; db.m12 = a + 17 * pl; db.m22 = a + 18 * pl;
    db.fix = a + 19 * pl; db.fiy = a + 20 * pl; db.fil = a + 21 * pl;
    db.fjx = a + 22 * pl; db.fjy = a + 23 * pl; db.fjl = a + 24 * pl;
    db.dtl[0] = a + 25 * pl; db.dtl[1] = a + 28 * pl; db.sa = a + 26 * pl; db.sb = a + 27 * pl;
    db.nodr = a + 29 * pl;
    if (visc) { db.vdx = a + 30 * pl; db.vey = a + 33 * pl; db.vsv = a + 36 * pl; }
    if (mxgeo) {
      db.mgn = reinterpret_cast<const float4*>(a + g0 * pl); db.mgs = reinterpret_cast<const float2*>(a + (g0 + 2) * pl);
      db.mgd = reinterpret_cast<const float4*>(a + (g0 + 3) * pl); db.mhy = reinterpret_cast<const float*>(a + (g0 + 5) * pl);
      db.msrc = reinterpret_cast<const float2*>(a + (g0 + 6) * pl);
    }
    if (mxvis) {
      db.mmm = reinterpret_cast<const float4*>(a + (g0 + 7) * pl);
      db.mnod = reinterpret_cast<const float*>(a + (g0 + 9) * pl);
    }
    if (irs) db.irs[0] = a + (size_t)irs_plane * pl;
    for (int f = 0; f < 4; ++f) db.bcf[f] = -1;
  }
  h->dbc.insert(h->dbc.end(), newbc.begin(), newbc.end());
  return rebuild_descriptors(h);
}

}  // namespace

namespace {
// Device address of `scratch` for each device, looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

int noalloc_scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(Q2D_E_HIP, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(Q2D_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_addr[dev];
  return Q2D_OK;
}
}  // namespace

int q2d_gpu_create(const q2d_problem* prob, q2d_handle** out) {
  if (!prob || !out) return fail(Q2D_E_ARG, "NULL argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(Q2D_E_NODEVICE, "no HIP device");
  q2d_handle* h = new q2d_handle();
  auto bail = [&](int code) { q2d_gpu_destroy(h); return code; };
  if (hipGetDevice(&h->device) != hipSuccess) return bail(fail(Q2D_E_HIP, "hipGetDevice failed"));
  {
    const int rc = noalloc_scratch_address(h->device, &h->noalloc_scratch);
    if (rc != Q2D_OK) return bail(rc);
  }
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) h->cu_count = prop.multiProcessorCount;
  }
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fail(Q2D_E_HIP, "stream creation failed"));
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      h->wallclock_khz = khz;
  }
  if (hipMalloc(&h->d_tstamp, sizeof(unsigned long long) * (2 + 2 * (size_t)kTimedSlots * kStampLanes)) != hipSuccess)
    return bail(fail(Q2D_E_NOMEM, "timing buffer allocation failed"));
  int rc = append_cases(h, prob, nullptr);
  if (rc == Q2D_OK && h->ncases == 0) rc = rebuild_descriptors(h);
  if (rc != Q2D_OK) {
    std::string msg = g_err;
    q2d_gpu_destroy(h);
    g_err = msg;
    return rc;
  }
  *out = h;
  return Q2D_OK;
}

int q2d_gpu_add_cases(q2d_handle* h, const q2d_problem* prob, int32_t* first_case) {
  if (!h || !prob) return fail(Q2D_E_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  const int nb0 = h->nblocks, nc0 = h->ncases;
  const size_t nbc0 = h->dbc.size();
  int rc = append_cases(h, prob, first_case);
  // the next step runs the dt pass: the added cases have no dt from a previous last stage (the dt pass
  // and the fused dt evaluate the same formula, so the cases already running are not affected)
  h->dt_fused_valid = false;
  if (rc != Q2D_OK && (h->nblocks != nb0 || h->ncases != nc0 || h->dc.size() != (size_t)nc0 ||
                       h->case_tab.size() != (size_t)nc0)) {
    // roll back a partial append: free what was allocated and rebuild from the previous set
    std::string msg = g_err;
    for (int b = nb0; b < h->nblocks; ++b) if (h->hb[b].arena) { (void)hipFree(h->hb[b].arena); h->arena_bytes -= h->hb[b].arena_bytes; }
    for (size_t c = nc0; c < h->case_tab.size(); ++c) if (h->case_tab[c]) (void)hipFree(h->case_tab[c]);
    h->hb.resize(nb0); h->db.resize(nb0); h->nblocks = nb0;
    h->dbc.resize(nbc0);
    h->dc.resize(nc0); h->case_live.resize(nc0); h->case_tab.resize(nc0); h->ncases = nc0;
    (void)rebuild_descriptors(h);
    g_err = msg;
  }
  return rc;
}

int q2d_gpu_remove_case(q2d_handle* h, int32_t case_id) {
  if (!h || case_id < 0 || case_id >= h->ncases) return fail(Q2D_E_ARG, "bad case index");
  if (!h->case_live[case_id]) return fail(Q2D_E_ARG, "case already removed");
  HIPCHK(hipStreamSynchronize(h->stream));
  h->case_live[case_id] = 0;
  if (h->case_tab[case_id]) {                          // its tables (the stream is idle)
    (void)hipFree(h->case_tab[case_id]);
    h->case_tab[case_id] = nullptr;
    h->dc[case_id].tab = nullptr; h->dc[case_id].ntab = 0;
  }
  const int zero = 0;
  HIPCHK(hipMemcpy((char*)(h->d_ctrl + case_id) + offsetof(DevCtrl, active), &zero, sizeof(int), hipMemcpyHostToDevice));
  for (int b = 0; b < h->nblocks; ++b) {
    HostBlock& hb = h->hb[b];
    if (hb.case_id != case_id || !hb.arena) continue;
    (void)hipFree(hb.arena);
    hb.arena = nullptr;
    h->arena_bytes -= hb.arena_bytes;
    DevBlock& db = h->db[b];
    const int ni = db.ni, nj = db.nj;
    memset(&db, 0, sizeof(db));
    db.ni = ni; db.nj = nj; db.case_id = case_id;
  }
  std::vector<DevBC> keep;
  keep.reserve(h->dbc.size());
  for (const DevBC& r : h->dbc) if (h->hb[r.blk].case_id != case_id) keep.push_back(r);
  h->dbc.swap(keep);
  return rebuild_descriptors(h);
}

int q2d_gpu_set_fusion(q2d_handle* h, int32_t visc) {
  if (!h) return fail(Q2D_E_ARG, "NULL argument");
  if (visc != 0 && visc != 1) return fail(Q2D_E_ARG, "visc must be 0 or 1");
  HIPCHK(hipStreamSynchronize(h->stream));
  h->fuse_visc = visc != 0;
  return rebuild_descriptors(h);
}

int q2d_gpu_get_runtime(q2d_handle* h, int32_t case_id, q2d_runtime* rt) {
  if (!h || !rt) return fail(Q2D_E_ARG, "NULL argument");
  if (case_id < 0 || case_id >= h->ncases || !h->case_live[case_id]) return fail(Q2D_E_ARG, "bad case index");
  const DevCase& k = h->dc[case_id];
  memset(rt, 0, sizeof(*rt));
  rt->cfl = k.cfl; rt->vnn = k.vnn; rt->ramp_iter = k.ramp_iter; rt->itermax = k.itermax;
  rt->res_threshold = k.res_threshold; rt->time_threshold = k.time_threshold; rt->res_drop = k.res_drop;
  rt->limiter = k.limiter; rt->riemann = k.riemann; rt->space_sd = k.space_sd;
  return Q2D_OK;
}

int q2d_gpu_set_runtime(q2d_handle* h, int32_t case_id, const q2d_runtime* rt) {
  if (!h || !rt) return fail(Q2D_E_ARG, "NULL argument");
  if (case_id < 0 || case_id >= h->ncases || !h->case_live[case_id]) return fail(Q2D_E_ARG, "bad case index");
  if (!(rt->cfl > 0) || !std::isfinite(rt->cfl) || !(rt->vnn > 0) || !std::isfinite(rt->vnn))
    return fail(Q2D_E_ARG, "cfl and vnn must be finite and > 0");
  if (rt->ramp_iter < 0 || rt->itermax < 0) return fail(Q2D_E_ARG, "ramp_iter and itermax must be >= 0");
  if (rt->res_drop < 0 || rt->res_drop >= 1) return fail(Q2D_E_ARG, "res_drop must be in [0, 1)");
  if (rt->limiter < 0 || rt->limiter > 6) return fail(Q2D_E_ARG, "unknown limiter");
  if (rt->riemann < 0 || rt->riemann > 4) return fail(Q2D_E_ARG, "unknown Riemann solver");
  if (rt->space_sd != 0 && rt->space_sd != 1) return fail(Q2D_E_ARG, "space_sd must be 0 or 1");
  HIPCHK(hipStreamSynchronize(h->stream));
  DevCase& k = h->dc[case_id];
  const bool regroup = k.riemann != rt->riemann || k.space_sd != rt->space_sd;
  k.cfl = rt->cfl; k.vnn = rt->vnn; k.ramp_iter = rt->ramp_iter; k.itermax = rt->itermax;
  k.res_threshold = rt->res_threshold; k.time_threshold = rt->time_threshold; k.res_drop = rt->res_drop;
  k.limiter = rt->limiter; k.riemann = rt->riemann; k.space_sd = rt->space_sd;
  // the next step runs the dt pass with the new keys
  h->dt_fused_valid = false;
  if (regroup) return rebuild_descriptors(h);          // solver set and kernel grouping follow the case
  HIPCHK(hipMemcpy(h->d_cases + case_id, &k, sizeof(DevCase), hipMemcpyHostToDevice));
  return Q2D_OK;
}

int q2d_gpu_case_count(q2d_handle* h, int32_t* ncases, int32_t* nblocks, int64_t* device_bytes) {
  if (!h || !ncases || !nblocks || !device_bytes) return fail(Q2D_E_ARG, "NULL argument");
  *ncases = h->ncases; *nblocks = h->nblocks; *device_bytes = (int64_t)h->arena_bytes;
  return Q2D_OK;
}

// After one more timed step has been issued: read the slots when the ring is full.
int count_timed_step(q2d_handle* h) {
  if (!h->timing) return Q2D_OK;
  if (++h->timed_steps_pending >= kTimedSteps) return collect_timing(h);
  return Q2D_OK;
}

int q2d_gpu_step(q2d_handle* h, q2d_step_out* out) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  int rc = launch_step(h, own_scratch(h));
  if (rc == Q2D_OK) rc = count_timed_step(h);
  if (rc == Q2D_OK && h->timing) rc = collect_timing(h);
  if (rc != Q2D_OK) return rc;
  int err = Q2D_OK;
  rc = read_out(h, out, &err);
  if (rc != Q2D_OK) return rc;
  if (err != Q2D_OK) return fail(err, err == Q2D_E_STATE ? "NaN or p<0 detected" : "Non-physical state detected at the interface");
  return Q2D_OK;
}

namespace {
// Capture `nst` back-to-back steps of the handle's launch sequence into an executable graph.
int capture_steps(q2d_handle* h, int nst, hipGraphExec_t* out) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = Q2D_OK;
  for (int k = 0; k < nst && rc == Q2D_OK; ++k) rc = launch_step(h, own_scratch(h));
  hipError_t ce = hipStreamEndCapture(h->stream, &g);
  if (rc != Q2D_OK) { if (ce == hipSuccess) (void)hipGraphDestroy(g); return rc; }
  if (ce != hipSuccess) return fail(Q2D_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
  hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return fail(Q2D_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  return Q2D_OK;
}
}  // namespace

int q2d_gpu_capture(q2d_handle* h) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  if (h->ncases == 0) return Q2D_OK;
  // the first step after create / upload / add_cases / set_runtime runs eagerly with its dt pass, and the
  // graphs hold the steps after it
  if (fused_handle(h) && !h->dt_fused_valid) return Q2D_OK;
  if (!h->graph_multi) { int rc = capture_steps(h, kGraphSteps, &h->graph_multi); if (rc != Q2D_OK) return rc; }
  if (!h->graph) { int rc = capture_steps(h, 1, &h->graph); if (rc != Q2D_OK) return rc; }
  return Q2D_OK;
}

int q2d_gpu_run(q2d_handle* h, int32_t nsteps, q2d_step_out* out) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  if (nsteps < 0) return fail(Q2D_E_ARG, "nsteps < 0");
  if (h->ncases == 0) return Q2D_OK;
  if (nsteps > 0) {
    if (fused_handle(h) && !h->dt_fused_valid) {      // first step after create/upload: eager, with k_dt
      int rc = launch_step(h, own_scratch(h));
      if (rc == Q2D_OK) rc = count_timed_step(h);
      if (rc != Q2D_OK) return rc;
      --nsteps;
    }
    // Steps replay captured graphs: kGraphSteps steps per launch while that many remain (the replays of
    // one graph run back to back; each graph launch costs ~8 us of idle device time), then single steps.
    // Every kernel checks its case's active flag, so steps after a case stops are no-ops for it.
    auto capture = [h](int nst, hipGraphExec_t* out) { return capture_steps(h, nst, out); };
    while (nsteps >= kGraphSteps) {
      if (!h->graph_multi) { int rc = capture(kGraphSteps, &h->graph_multi); if (rc != Q2D_OK) return rc; }
      if (h->timing && h->timed_steps_pending + kGraphSteps > kTimedSteps) {
        int rc = collect_timing(h);                   // the device's timing ring holds kTimedSteps steps
        if (rc != Q2D_OK) return rc;
      }
      HIPCHK(hipGraphLaunch(h->graph_multi, h->stream));
      if (h->timing) h->timed_steps_pending += kGraphSteps;
      nsteps -= kGraphSteps;
    }
    if (!h->graph && nsteps > 0) { int rc = capture(1, &h->graph); if (rc != Q2D_OK) return rc; }
    for (int s = 0; s < nsteps; ++s) {
      HIPCHK(hipGraphLaunch(h->graph, h->stream));
      int rc = count_timed_step(h);
      if (rc != Q2D_OK) return rc;
    }
    if (h->timing) { int rc = collect_timing(h); if (rc != Q2D_OK) return rc; }
  }
  int err = Q2D_OK;
  int rc = read_out(h, out, &err);
  if (rc != Q2D_OK) return rc;
  if (err != Q2D_OK) return fail(err, err == Q2D_E_STATE ? "NaN or p<0 detected" : "Non-physical state detected at the interface");
  return Q2D_OK;
}

int q2d_gpu_step_noalloc(q2d_handle* h, q2d_step_out* out) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  if (h->timing) return fail(Q2D_E_ARG, "q2d_gpu_step_noalloc records no kernel times: turn timing off");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != h->device) return fail(Q2D_E_ARG, "the handle's device must be current");
  // the handle's step scratch laid out in the file-scope array, each part on a 256-byte boundary
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t n_part = sizeof(double) * kPartials * std::max<size_t>(1, h->items.size() + h->irs_items.size());
  const size_t n_bc = sizeof(BCFlux) * std::max<size_t>(1, h->dbc.size());
  // the dense FP32 array exists for mixed-form handles only: a non-null one tells k_bcflux that the
  // stage kernel reads it instead of the per-record results
  const size_t n_dense = h->precision_mixed ? sizeof(float4) * std::max<size_t>(1, h->bcface.size()) : 0;
  const size_t need = up(n_part) + up(n_bc) + up(n_dense);
  if (need > (size_t)Q2D_NOALLOC_SCRATCH_BYTES)
    return fail(Q2D_E_NOMEM, "step scratch of " + std::to_string(need) + " bytes exceeds the allocation-free maximum of " +
                                 std::to_string((size_t)Q2D_NOALLOC_SCRATCH_BYTES) + " bytes");
  unsigned char* base = h->noalloc_scratch;
  const StepScratch scr{reinterpret_cast<double*>(base), reinterpret_cast<BCFlux*>(base + up(n_part)),
                        h->precision_mixed ? reinterpret_cast<float4*>(base + up(n_part) + up(n_bc)) : nullptr};
  int rc = launch_step(h, scr);
  if (rc != Q2D_OK) return rc;
  int err = Q2D_OK;
  rc = read_out(h, out, &err);
  if (rc != Q2D_OK) return rc;
  if (err != Q2D_OK) return fail(err, err == Q2D_E_STATE ? "NaN or p<0 detected" : "Non-physical state detected at the interface");
  return Q2D_OK;
}

int q2d_gpu_download_state(q2d_handle* h, int32_t block, double* P) {
  if (!h || !P || block < 0 || block >= h->nblocks) return fail(Q2D_E_ARG, "bad argument");
  if (!h->hb[block].arena) return fail(Q2D_E_ARG, "block of a removed case");
  HostBlock& b = h->hb[block];
  DevCtrl k;
  HIPCHK(hipMemcpyAsync(&k, h->d_ctrl + b.case_id, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double> tmp((size_t)b.plane * 4);
  HIPCHK(hipMemcpy(tmp.data(), b.arena + (size_t)4 * k.cur * b.plane, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  const int ngi = b.ni + 4, ngj = b.nj + 4;
  for (int j = 0; j < ngj; ++j)
    for (int i = 0; i < ngi; ++i)
      for (int v = 0; v < 4; ++v) P[((size_t)j * ngi + i) * 4 + v] = tmp[(size_t)v * b.plane + (size_t)j * b.pitch + i];
  return Q2D_OK;
}

int q2d_gpu_upload_state(q2d_handle* h, int32_t block, const double* P) {
  if (!h || !P || block < 0 || block >= h->nblocks) return fail(Q2D_E_ARG, "bad argument");
  if (!h->hb[block].arena) return fail(Q2D_E_ARG, "block of a removed case");
  HostBlock& b = h->hb[block];
  DevCtrl k;
  HIPCHK(hipMemcpyAsync(&k, h->d_ctrl + b.case_id, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double> tmp;
  pack_state(b, P, tmp);
  HIPCHK(hipMemcpy(b.arena + (size_t)4 * k.cur * b.plane, tmp.data(), sizeof(double) * tmp.size(), hipMemcpyHostToDevice));
  h->dt_fused_valid = false;
  return Q2D_OK;
}

int q2d_gpu_download_dt(q2d_handle* h, int32_t block, double* dtlocal) {
  if (!h || !dtlocal || block < 0 || block >= h->nblocks) return fail(Q2D_E_ARG, "bad argument");
  if (!h->hb[block].arena) return fail(Q2D_E_ARG, "block of a removed case");
  HostBlock& b = h->hb[block];
  DevCtrl k;
  HIPCHK(hipMemcpy(&k, h->d_ctrl + b.case_id, sizeof(DevCtrl), hipMemcpyDeviceToHost));
  const DevCase& c = h->dc[b.case_id];
  if (c.time_accurate) {
    for (int i = 0; i < b.ni * b.nj; ++i) dtlocal[i] = k.dtglobal;
    return Q2D_OK;
  }
  std::vector<double> tmp((size_t)b.plane);
  HIPCHK(hipMemcpy(tmp.data(), b.arena + (size_t)(k.dtsel ? 28 : 25) * b.plane, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
  for (int j = 1; j <= b.nj; ++j)
    for (int i = 1; i <= b.ni; ++i) dtlocal[(size_t)(j - 1) * b.ni + (i - 1)] = tmp[(size_t)(j + 1) * b.pitch + (i + 1)];
  return Q2D_OK;
}

int q2d_gpu_set_timing(q2d_handle* h, int32_t enable) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  const bool on = enable != 0;
  if (on != h->timing) drop_graphs(h);             // launch args change
  h->timing = on;
  h->timed_launches = 0; h->timed_ms = 0.0;
  return h->timing ? reset_timing(h) : Q2D_OK;
}

int q2d_gpu_timing_stats(q2d_handle* h, int64_t* launches, double* ms_total) {
  if (!h || !launches || !ms_total) return fail(Q2D_E_ARG, "NULL argument");
  *launches = h->timed_launches; *ms_total = h->timed_ms;
  return Q2D_OK;
}

int q2d_gpu_set_tiling(q2d_handle* h, int32_t rows_per_item, int32_t waves_per_group) {
  if (!h) return fail(Q2D_E_ARG, "NULL handle");
  if (rows_per_item < -1 || rows_per_item > 4096) return fail(Q2D_E_ARG, "rows_per_item out of range");
  if (waves_per_group != 0 && waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4)
    return fail(Q2D_E_ARG, "waves_per_group must be 1, 2 or 4");
  if (rows_per_item != 0) h->rows_per_item = rows_per_item;
  if (waves_per_group > 0) h->waves_per_group = waves_per_group;
  return build_items(h);
}

int q2d_gpu_get_tiling(q2d_handle* h, int32_t* rows_per_item, int32_t* waves_per_group) {
  if (!h || !rows_per_item || !waves_per_group) return fail(Q2D_E_ARG, "NULL argument");
  *rows_per_item = h->rows_per_item; *waves_per_group = h->waves_per_group;
  return Q2D_OK;
}

int q2d_hbm_copy(int64_t nbytes, int32_t reps, double* ms_per_copy) {
  if (nbytes < 16 || reps < 1 || !ms_per_copy) return fail(Q2D_E_ARG, "bad argument");
  const long long n = nbytes / 16;
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  HIPCHK(hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); return fail(Q2D_E_NOMEM, "copy buffer"); }
  HIPCHK(hipMemset(a, 0, n * 16));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // 128 workgroups per CU: a grid of that size measured the highest plain-copy rate (5.0 TB/s against
  // 4.5 TB/s with 32 per CU, on 1 GiB)
  int grid = 256 * 128;
  hipLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, 0, a, b, n);   // warm-up
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / reps;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  (void)hipFree(a); (void)hipFree(b);
  return Q2D_OK;
}

int q2d_gpu_info(q2d_handle* h, int64_t* cells, int64_t* work_items) {
  if (!h || !cells || !work_items) return fail(Q2D_E_ARG, "NULL argument");
  long long n = 0;
  for (auto& b : h->hb) n += (long long)b.ni * b.nj;
  *cells = n; *work_items = (int64_t)h->items.size();
  return Q2D_OK;
}

}  // extern "C"
