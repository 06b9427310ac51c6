// stage2d.hip -- fused RK stage for 1D and 2D grids on gfx950.
//
// One kernel evaluates L(u^(k)) (Roe: res/Roe.cl calcDeltaQTilde + calcFlux;
// HLL/HLLC: res/EulerHLL.cl, res/EulerHLLC.cl calcFlux; then
// res/CalcFluxDeriv.cl) and applies the Shu-Osher combination of
// include/HydroGPU/Integrator/RungeKutta.h:84-99 for a block of cells.
//
// Work decomposition: one wavefront owns a strip of 64 consecutive columns
// (60 output columns plus a 2-column halo on each side) and marches up a
// chunk of rows.  x-neighbours are exchanged through full-wave DPP shifts, so
// the x-direction stencil never touches LDS; the y-direction stencil lives in
// registers as a rolling window of the last rows, so every state value is
// read from HBM once per stage (plus the halo overlap).  The frozen
// start-of-step data (Roe eigenbasis, HLL wave speeds) are recomputed from u0
// per interface instead of being stored: O(N) bytes per cell instead of
// O(N^2).
//
// Reference semantics kept:
//  * eigen data frozen at the start of the step (src/Solver/Roe.cpp:82-99,
//    src/Solver/HLL.cpp:225-227);
//  * the Roe limiter ratio uses the neighbour interface's own basis
//    (res/Roe.cl:218-220, 278-290) and the full-step dt (res/Roe.cl:206, 292);
//  * interfaces the reference never writes (index 1 and size-1 along an
//    axis, or any coordinate outside 2..size-2) contribute deltaQTilde = 0.
#include "device_common.hpp"
#include "launch.hpp"
#include "policies.hpp"
#include "phys_srhd.hpp"
#include "phys_maxwell.hpp"
#include "stage_common.hpp"

namespace hydro {

template <class Pol, int DIM, typename T, typename S, bool KAHAN>
__global__ void __launch_bounds__(256) stage2d_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  using Frozen = typename Pol::Frozen;
  using FaceA = typename Pol::FaceA;

  // XCD-aware block order: consecutive logical blocks (neighbouring strips)
  // land on the same XCD and share its L2.
  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int b = blockIdx.x;
  const int lb = (b & 7) * per + (b >> 3);
  const int wave = lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= a.nwaves) return;
  const int lane = threadIdx.x & 63;
  const int strip = wave % a.nstrips;
  const int chunk = wave / a.nstrips;

  const int sx = a.sx, sy = a.sy;
  const int c = strip * kStripOut + lane;      // column (x0 - 2 + lane)
  const int cl = c < sx ? c : sx - 1;           // in-bounds load column
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;
  const bool col_face = c >= 2 && c <= sx - 2;  // the reference writes faces here

  const T gamma = T(a.gamma);
  const PhysParams<T> pp = make_params<T>(a);
  const double dtd = *a.dt;
  const T dt_dx0 = T(dtd * a.inv_dx[0]);
  const T idx0 = T(a.inv_dx[0]);
  StageCombine<T, S, KAHAN, N> comb(a, dtd);

  // x-direction flux divergence of one row: -(F(c+1) - F(c)) / dx
  auto xdiv = [&](const T u[N], const Frozen& fz, T div[N]) {
    T uL[N], pk[Pol::NF], pkL[Pol::NF];
#pragma unroll
    for (int s = 0; s < N; ++s) uL[s] = lane_from_left(u[s]);
    Pol::pack(fz, pk);
#pragma unroll
    for (int i = 0; i < Pol::NF; ++i) pkL[i] = lane_from_left(pk[i]);
    const Frozen fL = Pol::unpack(pkL);
    const FaceA fa = Pol::face_a(fL, fz, uL, u, 0, col_face, pp);
    T dqlo[N], dqhi[N], F[N];
    if constexpr (Pol::kNeighbors) {
#pragma unroll
      for (int s = 0; s < N; ++s) {
        dqlo[s] = lane_from_left(fa.dq[s]);
        dqhi[s] = lane_from_right(fa.dq[s]);
      }
    }
    Pol::face_b(fa, dqlo, dqhi, dt_dx0, 0, a.lim, F);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const T FR = lane_from_right(F[s]);
      div[s] = -(FR - F[s]) * idx0;
    }
  };

  auto load_row = [&](int row, T u[N], Frozen& fz) {
    const long long cell = (long long)row * sx + cl;
    S q0[N];
    comb.load_state(cell, q0, u);
    if constexpr (Pol::kFrozenPrims) {
      double pr[N];
#pragma unroll
      for (int s = 0; s < N; ++s) pr[s] = a.prim[s * a.plane + cell];
      fz = Pol::template frozen<double>(pr, gamma);
    } else {
      fz = Pol::template frozen<S>(q0, gamma);
    }
  };

  if constexpr (DIM == 1) {
    T u[N];
    Frozen fz;
    load_row(0, u, fz);
    T div[N];
    xdiv(u, fz, div);
    if (out_lane) comb.finish((long long)c, div);
  } else {
    const T dt_dy = T(dtd * a.inv_dx[1]);
    const T idy = T(a.inv_dx[1]);
    const int y0 = 2 + chunk * a.rows;
    int y1 = y0 + a.rows;
    if (y1 > sy - 2) y1 = sy - 2;

    // rolling window over rows: state and frozen data of row r-1, the first
    // pass of faces r-1 and r-2, the flux of face r-2 and the x-divergence of
    // rows r-1 and r-2
    T u_prev[N];
    Frozen f_prev;
    FaceA fa_m1, fa_m2;
    T Fy_prev[N], divx_m1[N], divx_m2[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      Fy_prev[s] = divx_m1[s] = divx_m2[s] = T(0);
      fa_m1.dq[s] = fa_m2.dq[s] = T(0);
    }

    load_row(y0 - 2, u_prev, f_prev);

#pragma unroll 1
    for (int r = y0 - 1; r <= y1 + 1; ++r) {
      T u[N];
      Frozen fz;
      load_row(r, u, fz);

      T divx[N];
      if (r >= y0 && r < y1) {
        xdiv(u, fz, divx);
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) divx[s] = T(0);
      }

      // first pass of y-face r (between rows r-1 and r)
      const bool ok = col_face && r >= 2 && r <= sy - 2;
      const FaceA fa = Pol::face_a(f_prev, fz, u_prev, u, 1, ok, pp);

      T Fy[N];
      if (r >= y0 + 1) {
        // flux at face r-1: its own first pass, neighbours r-2 and r
        Pol::face_b(fa_m1, fa_m2.dq, fa.dq, dt_dy, 1, a.lim, Fy);
        if (r >= y0 + 2 && out_lane) {
          T deriv[N];
#pragma unroll
          for (int s = 0; s < N; ++s) deriv[s] = divx_m2[s] - (Fy[s] - Fy_prev[s]) * idy;
          comb.finish((long long)(r - 2) * sx + c, deriv);
        }
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) Fy[s] = T(0);
      }

#pragma unroll
      for (int s = 0; s < N; ++s) {
        Fy_prev[s] = Fy[s];
        divx_m2[s] = divx_m1[s];
        divx_m1[s] = divx[s];
        u_prev[s] = u[s];
      }
      fa_m2 = fa_m1;
      fa_m1 = fa;
      f_prev = fz;
    }
  }
}

template <class Pol, int DIM, typename T, typename S, bool KAHAN>
static void launch_pol(const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  int blocks = (a.nwaves + wpb - 1) / wpb;
  blocks = (blocks + 7) & ~7;  // multiple of 8 for the XCD remap
  hipLaunchKernelGGL((stage2d_kernel<Pol, DIM, T, S, KAHAN>), dim3(blocks), dim3(64 * wpb), 0, st, a);
}

// kind: 0 Euler Roe, 1 Euler HLL, 2 Euler HLLC, 3 SRHD Roe, 4 Maxwell Roe
template <int DIM, typename T, typename S, bool KAHAN>
static void launch_solver(int kind, int lim, const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  if (kind == 1) {
    launch_pol<HLLPolicy<DIM, T, 1>, DIM, T, S, KAHAN>(a, wpb, st);
  } else if (kind == 2) {
    launch_pol<HLLPolicy<DIM, T, 2>, DIM, T, S, KAHAN>(a, wpb, st);
  } else if (kind == 3) {
    switch (lim) {
    case LIM_SUPERBEE: launch_pol<SRHDRoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN>(a, wpb, st); break;
    case LIM_MONOTIZEDCENTRAL:
      launch_pol<SRHDRoePolicy<DIM, T, LIM_MONOTIZEDCENTRAL>, DIM, T, S, KAHAN>(a, wpb, st);
      break;
    default: launch_pol<SRHDRoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN>(a, wpb, st); break;
    }
  } else if (kind == 4) {
    switch (lim) {
    case LIM_SUPERBEE: launch_pol<MaxwellRoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN>(a, wpb, st); break;
    default: launch_pol<MaxwellRoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN>(a, wpb, st); break;
    }
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch_pol<RoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN>(a, wpb, st); break;
    case LIM_MONOTIZEDCENTRAL:
      launch_pol<RoePolicy<DIM, T, LIM_MONOTIZEDCENTRAL>, DIM, T, S, KAHAN>(a, wpb, st);
      break;
    case LIM_MINMOD: launch_pol<RoePolicy<DIM, T, LIM_MINMOD>, DIM, T, S, KAHAN>(a, wpb, st); break;
    default: launch_pol<RoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN>(a, wpb, st); break;
    }
  }
}

void launch_stage2d_f64(int dim, int solver, int lim, const StageArgs<double, double>& a, int wpb,
                        hipStream_t st) {
  if (dim == 1) launch_solver<1, double, double, false>(solver, lim, a, wpb, st);
  else launch_solver<2, double, double, false>(solver, lim, a, wpb, st);
}

void launch_stage2d_mixed(int dim, int solver, int lim, bool kahan, const StageArgs<float, float>& a, int wpb,
                          hipStream_t st) {
  if (dim == 1) {
    if (kahan) launch_solver<1, float, float, true>(solver, lim, a, wpb, st);
    else launch_solver<1, float, float, false>(solver, lim, a, wpb, st);
  } else {
    if (kahan) launch_solver<2, float, float, true>(solver, lim, a, wpb, st);
    else launch_solver<2, float, float, false>(solver, lim, a, wpb, st);
  }
}

}  // namespace hydro
