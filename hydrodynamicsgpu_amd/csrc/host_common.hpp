// host_common.hpp -- host-side pieces shared by the single-grid and the
// batched C ABI: error reporting, the Runge-Kutta tableaus and their stage
// plans, descriptor validation and per-solver constants.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/hydro_capi.h"
#include "stage_args.hpp"

namespace hydro {
namespace host {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}


#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(HYDRO_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// include/HydroGPU/Integrator/RungeKutta.h:129-294 (alphas(i,j), betas(i,j))
struct Tableau {
  int order = 0;
  double a[4][4] = {};
  double b[4][4] = {};
};

inline bool get_tableau(int id, Tableau& t) {
  t = Tableau();
  auto A = [&](int i, int j, double v) { t.a[i][j] = v; };
  auto B = [&](int i, int j, double v) { t.b[i][j] = v; };
  switch (id) {
  case HYDRO_INTEGRATOR_FORWARDEULER: t.order = 1; A(0, 0, 1); B(0, 0, 1); return true;
  case HYDRO_INTEGRATOR_RK2: t.order = 2; A(0, 0, 1); A(1, 0, 1); B(0, 0, .5); B(1, 1, 1); return true;
  case HYDRO_INTEGRATOR_RK2HEUN:
    t.order = 2; A(0, 0, 1); A(1, 0, 1); B(0, 0, 1); B(1, 0, .5); B(1, 1, .5); return true;
  case HYDRO_INTEGRATOR_RK2RALSTON:
    t.order = 2; A(0, 0, 1); A(1, 0, 1); B(0, 0, 2. / 3.); B(1, 0, 1. / 4.); B(1, 1, 3. / 4.); return true;
  case HYDRO_INTEGRATOR_RK3:
    t.order = 3; A(0, 0, 1); A(1, 0, 1); A(2, 0, 1);
    B(0, 0, .5); B(1, 0, -1); B(1, 1, 2); B(2, 0, 1. / 6.); B(2, 1, 2. / 6.); B(2, 2, 1. / 6.); return true;
  case HYDRO_INTEGRATOR_RK4:
    t.order = 4; A(0, 0, 1); A(1, 0, 1); A(2, 0, 1); A(3, 0, 1);
    B(0, 0, .5); B(1, 1, .5); B(2, 2, 1); B(3, 0, 1. / 6.); B(3, 1, 2. / 6.); B(3, 2, 2. / 6.); B(3, 3, 1. / 6.);
    return true;
  case HYDRO_INTEGRATOR_RK4_3_8THSRULE:
    t.order = 4; A(0, 0, 1); A(1, 0, 1); A(2, 0, 1); A(3, 0, 1);
    B(0, 0, 1. / 3.); B(1, 0, -1. / 3.); B(2, 0, 1); B(2, 1, -1); B(2, 2, 1);
    B(3, 0, 1. / 8.); B(3, 1, 3. / 8.); B(3, 2, 3. / 8.); B(3, 3, 1. / 8.); return true;
  case HYDRO_INTEGRATOR_RK2TVD:
    t.order = 2; A(0, 0, 1); A(1, 0, .5); A(1, 1, .5); B(0, 0, 1); B(1, 1, .5); return true;
  case HYDRO_INTEGRATOR_RK2NONTVD:
    t.order = 2; A(0, 0, 1); A(1, 0, 1); B(0, 0, -20); B(1, 0, 41. / 40.); B(1, 1, -1. / 40.); return true;
  case HYDRO_INTEGRATOR_RK3TVD:
    t.order = 3; A(0, 0, 1); A(1, 0, 3. / 4.); A(1, 1, 1. / 4.); A(2, 0, 1. / 3.); A(2, 1, 2. / 3.);
    B(0, 0, 1); B(1, 1, 1. / 4.); B(2, 2, 2. / 3.); return true;
  case HYDRO_INTEGRATOR_RK4TVD:
    t.order = 4; A(0, 0, 1); A(1, 0, 649. / 1600.); A(1, 1, 951. / 1600.);
    A(2, 0, 53989. / 2500000.); A(2, 1, 4806213. / 20000000.); A(2, 2, 23619. / 32000.);
    A(3, 0, 1. / 5.); A(3, 1, 6127. / 30000.); A(3, 2, 7873. / 30000.); A(3, 3, 1. / 3.);
    B(0, 0, .5); B(1, 0, -10890423. / 25193600.); B(1, 1, 5000. / 7873);
    B(2, 0, -102261. / 5000000.); B(2, 1, -5121. / 20000.); B(2, 2, 7873. / 10000.);
    B(3, 0, 1. / 10.); B(3, 1, 1. / 6.); B(3, 3, 1. / 6.); return true;
  case HYDRO_INTEGRATOR_RK4NONTVD:
    t.order = 4; A(0, 0, 1); A(1, 0, 1); A(2, 0, 1);
    A(3, 0, -1. / 3.); A(3, 1, 1. / 3.); A(3, 2, 2. / 3.); A(3, 3, 1. / 3.);
    B(0, 0, .5); B(1, 1, .5); B(2, 2, 1); B(3, 3, 1. / 6.); return true;
  }
  return false;
}

// Per-stage plan of partial Shu-Osher sums (increment form d_k = u^(k) - u0).
struct StagePlan {
  int pk = -1;
  int nf = 0;
  int fslot[kMaxPartials] = {-1, -1, -1};
  int fload[kMaxPartials] = {0, 0, 0};
  double am[kMaxPartials] = {0, 0, 0};
  double bm[kMaxPartials] = {0, 0, 0};
  double a_self = 0, b_self = 0;
  bool final_stage = false;
  int din = -1, dout = -1;  // increment buffers read (stencil) and written
  int nrep = 0;
  int rep[kMaxPartials] = {0, 0, 0};
  double rc[kMaxPartials] = {0, 0, 0};
  int mode = CM_GENERIC;  // CombineMode the stage kernel runs
};

// Replay plan: valid when every row before the last has the form
// d_{k+1} = alpha[k][k] d_k + beta[k][k] dt L_k (plus its u0 part), i.e. the
// rows mix no earlier stage than their own input (classic RK4, RK2 midpoint,
// Heun, Ralston, TVD-RK2, the reference's TVD-RK3, RK4 non-TVD).  Then
// dt L_k = (d_{k+1} - alpha[k][k] d_k) / beta[k][k] up to rounding, the last
// row is a combination of the stored increments, and no partial-sum buffer
// is read or written: stage k reads d_k (buffer k-1) and writes d_{k+1}
// (buffer k).
inline bool plan_replay(const Tableau& t, std::vector<StagePlan>& out) {
  const int S = t.order;
  if (S < 2 || S - 1 > kMaxPartials) return false;
  for (int k = 0; k < S - 1; ++k) {
    for (int j = 1; j < k; ++j)
      if (t.a[k][j] != 0.0) return false;
    for (int j = 0; j < k; ++j)
      if (t.b[k][j] != 0.0) return false;
    if (t.b[k][k] == 0.0) return false;
  }
  // last row: d_S = sum_j c[j] d_j + beta[S-1][S-1] dt L_{S-1}
  double c[4] = {0, 0, 0, 0};
  for (int j = 1; j < S; ++j) c[j] = t.a[S - 1][j];
  for (int j = 0; j < S - 1; ++j) {
    const double bj = t.b[S - 1][j];
    if (bj == 0.0) continue;
    c[j + 1] += bj / t.b[j][j];
    if (j > 0) c[j] -= bj * t.a[j][j] / t.b[j][j];
  }
  int nrep = 0;
  for (int j = 1; j < S - 1; ++j)
    if (c[j] != 0.0) ++nrep;
  if (nrep > kMaxPartials) return false;
  out.assign(S, StagePlan());
  for (int k = 0; k < S; ++k) {
    StagePlan& p = out[k];
    p.din = k - 1;
    p.dout = k < S - 1 ? k : -1;
    p.b_self = t.b[k][k];
    p.final_stage = k == S - 1;
    if (!p.final_stage) {
      p.a_self = k > 0 ? t.a[k][k] : 0.0;  // coefficient of the stencil input d_k
    } else {
      p.a_self = c[S - 1];
      for (int j = 1; j < S - 1; ++j) {
        if (c[j] == 0.0) continue;
        p.rep[p.nrep] = j - 1;  // buffer holding d_j
        p.rc[p.nrep] = c[j];
        p.nrep++;
      }
    }
  }
  return true;
}

inline bool plan_stages(const Tableau& t, std::vector<StagePlan>& out, int& nslots) {
  const int S = t.order;
  out.assign(S, StagePlan());
  bool live[4] = {false, false, false, false};
  int slot_of[4] = {-1, -1, -1, -1};
  bool used[kMaxPartials] = {false, false, false};
  nslots = 0;
  for (int k = 0; k < S; ++k) {
    StagePlan& p = out[k];
    p.pk = live[k] ? slot_of[k] : -1;
    p.a_self = t.a[k][k];
    p.b_self = t.b[k][k];
    p.final_stage = (k == S - 1);
    p.din = k > 0 ? (k - 1) & 1 : -1;  // ping-pong increment buffers
    p.dout = k < S - 1 ? k & 1 : -1;
    // row k's slot is consumed in this stage; the same thread reads it before
    // any write below, so the slot may be reused at once
    if (live[k]) used[slot_of[k]] = false;
    live[k] = false;
    for (int m = k + 1; m < S; ++m) {
      const bool add = (k > 0 && t.a[m][k] != 0.0) || t.b[m][k] != 0.0;
      if (!add) continue;
      if (p.nf >= kMaxPartials) return false;
      int j = p.nf++;
      if (!live[m]) {
        int s = -1;
        for (int q = 0; q < kMaxPartials; ++q)
          if (!used[q]) { s = q; break; }
        if (s < 0) return false;
        used[s] = true;
        slot_of[m] = s;
        p.fload[j] = 0;
        live[m] = true;
        nslots = std::max(nslots, s + 1);
      } else {
        p.fload[j] = 1;
      }
      p.fslot[j] = slot_of[m];
      p.am[j] = (k > 0) ? t.a[m][k] : 0.0;
      p.bm[j] = t.b[m][k];
    }
  }
  return true;
}

// What the tableau's rows make of a constant derivative L: u0 + c dt L after
// the last row, c = sum over rows of alpha c_k + beta (1 for a consistent
// tableau; include/HydroGPU/Integrator/RungeKutta.h:84-99 with L fixed)
inline double constant_derivative_sum(const Tableau& t) {
  double c[5] = {0, 0, 0, 0, 0};
  for (int i = 1; i <= t.order; ++i) {
    double v = 0;
    for (int k = 0; k < i; ++k) v += t.a[i - 1][k] * c[k] + t.b[i - 1][k];
    c[i] = v;
  }
  return c[t.order];
}

// Fixed-function combination of a stage, or CM_GENERIC when the stage reads
// or writes partial sums.
inline int combine_mode(const StagePlan& p) {
  if (p.pk >= 0 || p.nf > 0) return CM_GENERIC;
  if (p.final_stage) {
    if (p.din < 0) return p.nrep == 0 ? CM_SINGLE : CM_GENERIC;
    return CM_FINAL;
  }
  if (p.nrep != 0) return CM_GENERIC;
  return p.din < 0 ? CM_FIRST : CM_MID;
}

// Rows per wave of the 2D stage kernel.  The kernel is issue-bound, so a
// stage takes about (waves on the busiest SIMD) x (rows + halo cost) wave
// iterations: pick the chunk height that minimises that over the grid's
// strips and rows (a chunk of R output rows costs about R + 3 rows of work:
// its two halo rows on each side load and run the first face pass).  Chunks
// stay small enough for `min_waves` waves per SIMD where the grid allows,
// so that memory latency is hidden behind other waves.
// Rows per y-chunk of the first-order 3D kernel: `columns` = strips x
// z-blocks workgroups per chunk.  At least two workgroups per CU where the
// grid allows (latency hiding), then the fewest rounds x (rows + 1 prologue
// row).
inline int pick_rows_3d_r1(long long columns, int ny, int ncu) {
  int best = ny;
  double best_cost = 1e300;
  const bool enough = columns * ((ny + 7) / 8) >= 2LL * ncu;
  for (int rows = 8; rows <= std::max(8, ny); ++rows) {
    const long long wg = columns * ((ny + rows - 1) / rows);
    if (enough && wg < 2LL * ncu) break;
    const long long rounds = (wg + 2LL * ncu - 1) / (2LL * ncu);
    const double cost = (double)rounds * (std::min(rows, ny) + 1.0);
    if (cost < best_cost * (1.0 - 1e-9)) {
      best_cost = cost;
      best = rows;
    }
  }
  return best;
}

// z-chunk height of the z-march 3D kernel: `columns` (strips x y-blocks)
// workgroups per chunk row, `slots` workgroups resident at once; a chunk of
// P planes marches P + 2 (one halo plane on each side), and the launch takes
// ceil(workgroups / slots) rounds
inline int pick_planes_zm(long long columns, int nz, long long slots) {
  int best = nz;
  double best_cost = 1e300;
  for (int p = 4; p <= nz; ++p) {
    const long long wg = columns * ((nz + p - 1) / p);
    const long long rounds = (wg + slots - 1) / slots;
    const double cost = (double)rounds * (std::min(p, nz) + 2.0);
    if (cost < best_cost * (1.0 - 1e-9)) {
      best_cost = cost;
      best = p;
    }
  }
  return best;
}

inline int pick_rows_2d(long long strips, int ny, int simds, int min_waves) {
  int best = 32;
  double best_cost = 1e300;
  const bool enough = strips * ((ny + 3) / 4) >= (long long)min_waves * simds;
  for (int rows = 4; rows <= 256; ++rows) {
    const long long waves = strips * ((ny + rows - 1) / rows);
    if (enough && waves < (long long)min_waves * simds) break;
    const long long per_simd = (waves + simds - 1) / simds;
    const double cost = (double)per_simd * (std::min(rows, ny) + 3.0);
    if (cost < best_cost * (1.0 - 1e-9)) {
      best_cost = cost;
      best = rows;
    }
  }
  return best;
}

// upper bound on the per-wave minima of the strip-march CFL kernel
constexpr int kDtMaxWaves = 65536;

// include/HydroGPU/toNumericString.h:12-22 (6 significant digits)
inline double kernel_dx(double dx, bool six) {
  if (!six) return dx;
  char buf[64];
  snprintf(buf, sizeof buf, "%.6g", dx);
  return strtod(buf, nullptr);
}

// A host<->device copy ordered on the simulation's (non-blocking) stream and
// complete on return: copies on the legacy null stream are not ordered
// against such streams, and a pageable host-to-device hipMemcpy may return
// before its transfer lands.
inline hipError_t copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

struct DevScalars {
  double dt;
  double t;
  long long steps;
  unsigned health;
  unsigned pad;
};


inline int validate(const hydro_sim_desc* d) {
  if (!d) return fail(HYDRO_E_INVALID, "null descriptor");
  if (d->dim < 1 || d->dim > 3) return fail(HYDRO_E_INVALID, "dim must be 1..3");
  for (int a = 0; a < d->dim; ++a)
    if (d->size[a] < 8) return fail(HYDRO_E_INVALID, "each used axis needs size >= 8 (2+2 ghosts + 4)");
  for (int a = d->dim; a < 3; ++a)
    if (d->size[a] != 1) return fail(HYDRO_E_INVALID, "unused axes must have size 1");
  if (d->equation < HYDRO_EQN_EULER || d->equation > HYDRO_EQN_MAXWELL) return fail(HYDRO_E_INVALID, "bad equation");
  if (d->equation != HYDRO_EQN_EULER && d->solver != HYDRO_SOLVER_ROE)
    return fail(HYDRO_E_UNSUPPORTED, "SRHD and Maxwell use the Roe solver (SRHDRoe, MaxwellRoe)");
  if (d->solver < HYDRO_SOLVER_ROE || d->solver > HYDRO_SOLVER_BURGERS) return fail(HYDRO_E_INVALID, "bad solver");
  if (d->limiter < 0 || d->limiter >= HYDRO_NUM_LIMITERS) return fail(HYDRO_E_INVALID, "bad limiter");
  if (d->integrator < 0 || d->integrator >= HYDRO_NUM_INTEGRATORS) return fail(HYDRO_E_INVALID, "bad integrator");
  if (d->precision != HYDRO_PRECISION_FP64 && d->precision != HYDRO_PRECISION_FP32_MIXED)
    return fail(HYDRO_E_INVALID, "bad precision");
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < 2; ++m)
      if (d->boundary[a][m] < HYDRO_BOUNDARY_NONE || d->boundary[a][m] > HYDRO_BOUNDARY_FREEFLOW)
        return fail(HYDRO_E_INVALID, "bad boundary method");
  if (d->use_fixed_dt && !(d->fixed_dt > 0)) return fail(HYDRO_E_INVALID, "fixed_dt must be > 0");
  // use_gravity acts on the Euler solvers only (the reference's
  // SelfGravitationBehavior solvers); the others ignore it, as there
  if (d->use_gravity && d->equation == HYDRO_EQN_EULER) {
    if (d->dim == 3 && d->solver == HYDRO_SOLVER_HLL)
      return fail(HYDRO_E_UNSUPPORTED, "3D HLL with self-gravity is not supported (2D and 1D are; HLLC and Roe in 3D are)");
    if (d->gauss_seidel_max_iter < 0) return fail(HYDRO_E_INVALID, "gauss_seidel_max_iter must be >= 0");
  }
  return HYDRO_OK;
}

// kernel family: 0 Euler Roe, 1 Euler HLL, 2 Euler HLLC, 3 SRHD Roe, 4 Maxwell Roe,
// 5 EulerBurgers
inline int kernel_kind(const hydro_sim_desc& d) {
  if (d.equation == HYDRO_EQN_SRHD) return 3;
  if (d.equation == HYDRO_EQN_MAXWELL) return 4;
  if (d.solver == HYDRO_SOLVER_BURGERS) return 5;
  return d.solver;
}

// res/MaxwellRoe.cl:11, 30-35, 139: float-typed se/su and inverses, and the
// eigenvalue from the 6-digit renderings of sqrt(eps), sqrt(mu)
// (src/Equation/Maxwell.cpp:58-63)
inline void maxwell_constants(const hydro_sim_desc& d, double mx[5]) {
  const float sqrt_1_2 = 0.7071067811865475727373109293694142252206802368164f;
  const double sp = kernel_dx(std::sqrt(d.maxwell_permittivity), true);
  const double su = kernel_dx(std::sqrt(d.maxwell_permeability), true);
  const float fse = (float)(sp * (double)sqrt_1_2);
  const float fsu = (float)(su * (double)sqrt_1_2);
  mx[0] = fse;
  mx[1] = fsu;
  mx[2] = 1.f / fse;
  mx[3] = 1.f / fsu;
  mx[4] = 1.0 / (sp * su);
}


}  // namespace host
}  // namespace hydro

// a device choice (replaces CLCommon's device/queue pick)
struct hydro_ctx {
  int device;
};
