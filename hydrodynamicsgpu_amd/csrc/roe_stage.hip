// roe_stage.hip -- fused Roe RK stage for 1D and 2D grids on gfx950.
//
// One kernel evaluates L(u^(k)) (res/Roe.cl calcDeltaQTilde + calcFlux,
// res/CalcFluxDeriv.cl) and applies the Shu-Osher combination of
// include/HydroGPU/Integrator/RungeKutta.h:84-99 for a block of cells.
//
// Work decomposition: one wavefront owns a strip of 64 consecutive columns
// (60 output columns plus a 2-column halo on each side) and marches up a
// chunk of rows.  x-neighbours are exchanged through full-wave DPP shifts, so
// the x-direction stencil never touches LDS; the y-direction stencil lives in
// registers as a rolling window of the last rows, so every state value is
// read from HBM once per stage (plus the halo overlap).  The frozen
// start-of-step eigenbasis is recomputed from u0 per face instead of being
// stored (O(N) bytes per cell instead of O(N^2)).
//
// Reference semantics kept:
//  * eigenbasis frozen at the start of the step (src/Solver/Roe.cpp:82-99);
//  * the limiter ratio uses the neighbour interface's own basis
//    (res/Roe.cl:218-220, 278-290);
//  * the full-step dt enters epsilon in every stage (res/Roe.cl:206, 292);
//  * interfaces the reference never writes (index 1 and size-1 along an
//    axis) contribute deltaQTilde = 0 to their neighbours' limiter.
#include "device_common.hpp"
#include "phys_euler.hpp"
#include "stage_args.hpp"
#include "launch.hpp"

namespace hydro {

constexpr int kStripOut = 60;  // output columns per wave (64 lanes - 2*2 halo)

template <int DIM, typename T, int LIM>
struct RoeFlux {
  using P = EulerRoe<DIM, T>;
  static constexpr int N = P::N;

  // res/Roe.cl:253-298 for one interface, in the rotated frame:
  //   Ft_k = lam_k (L qavg)_k - .5 lam_k dq_k (theta + phi(r)(lam_k dt/dx - theta))
  static HYDRO_DEV void limited(const T lam[N], const T lq[N], const T dq[N], const T dqlo[N],
                                const T dqhi[N], T dt_dx, int lim, T ft[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const T l = lam[k];
      const bool up = l >= T(0);
      const T num = up ? dqlo[k] : dqhi[k];
      const T theta = up ? T(1) : T(-1);
      const T r = tdiv(num, dq[k]);
      const T phi = limiter<LIM, T>(lim, r);
      const T eps = l * dt_dx;
      ft[k] = l * lq[k] - T(.5) * (l * dq[k]) * (theta + phi * (eps - theta));
    }
  }
};

template <int DIM, typename T, typename S, bool KAHAN, int LIM>
__global__ void __launch_bounds__(256) roe_stage_kernel(StageArgs<T, S> a) {
  using P = EulerRoe<DIM, T>;
  using Cell = typename P::Cell;
  using Face = typename P::Face;
  constexpr int N = P::N;

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
  const long long pl = a.plane;
  const int c = strip * kStripOut + lane;      // column (x0 - 2 + lane)
  const int cl = c < sx ? c : sx - 1;           // in-bounds load column
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;
  const bool col_face = c >= 2 && c <= sx - 2;  // reference writes faces here

  const T gamma = T(a.gamma);
  const double dtd = *a.dt;
  const T dt_dx0 = T(dtd * a.inv_dx[0]);
  const T idx0 = T(a.inv_dx[0]);

  // x-direction: limited flux at the left face of this lane's column for the
  // row whose values are u (stage state) and cc (frozen cell data); returns
  // -(F(c+1) - F(c)) / dx.
  auto xdiv = [&](const T u[N], const Cell& cc, T div[N]) {
    T uL[N];
#pragma unroll
    for (int s = 0; s < N; ++s) uL[s] = lane_from_left(u[s]);
    Cell cL;
    cL.w = lane_from_left(cc.w);
#pragma unroll
    for (int d = 0; d < DIM; ++d) cL.v[d] = lane_from_left(cc.v[d]);
    cL.H = lane_from_left(cc.H);
    const Face f = P::face(cL, cc, 0, gamma);
    T dlt[N], avg[N], dq[N], lq[N], lam[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      dlt[s] = u[s] - uL[s];
      avg[s] = T(.5) * (u[s] + uL[s]);
    }
    P::left(f, dlt, dq, gamma);
    P::left(f, avg, lq, gamma);
    P::eig(f, lam);
    if (!col_face) {
#pragma unroll
      for (int s = 0; s < N; ++s) dq[s] = T(0);
    }
    T dqlo[N], dqhi[N], ft[N], F[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      dqlo[s] = lane_from_left(dq[s]);
      dqhi[s] = lane_from_right(dq[s]);
    }
    RoeFlux<DIM, T, LIM>::limited(lam, lq, dq, dqlo, dqhi, dt_dx0, a.lim, ft);
    P::right(f, ft, F);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const T FR = lane_from_right(F[s]);
      div[s] = -(FR - F[s]) * idx0;
    }
  };

  // RK combination and store for one output cell (all lanes call it; only
  // out_lane lanes store)
  const T bdt_self = T(a.b_self * dtd);
  const T a_self = T(a.a_self);
  auto finish = [&](int row, const T deriv[N]) {
    if (!out_lane) return;
    const long long cell = (long long)row * sx + c;
    T dk[N];
#pragma unroll
    for (int s = 0; s < N; ++s) dk[s] = a.dk ? a.dk[s * pl + cell] : T(0);
    T acc[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      T v = (a.pk >= 0) ? a.part[a.pk][s * pl + cell] : T(0);
      if (a.dk) v += a_self * dk[s];
      acc[s] = v + bdt_self * deriv[s];
    }
#pragma unroll
    for (int j = 0; j < kMaxPartials; ++j) {
      if (j >= a.nf) break;
      T* pp = a.part[a.fslot[j]];
      const T am = T(a.am[j]);
      const T bm = T(a.bm[j] * dtd);
#pragma unroll
      for (int s = 0; s < N; ++s) {
        T v = a.fload[j] ? pp[s * pl + cell] : T(0);
        if (a.dk) v += am * dk[s];
        pp[s * pl + cell] = v + bm * deriv[s];
      }
    }
    if (a.final_stage) {
#pragma unroll
      for (int s = 0; s < N; ++s) {
        const S hi = a.u0[s * pl + cell];
        if constexpr (KAHAN) {
          // compensated update of (hi, lo) by acc (Fast2Sum)
          const float lo = a.u0lo[s * pl + cell];
          const float y = float(acc[s]) + lo;
          const float t = float(hi) + y;
          const float e = y - (t - float(hi));
          a.unew[s * pl + cell] = S(t);
          a.unewlo[s * pl + cell] = e;
        } else {
          a.unew[s * pl + cell] = S(hi + S(acc[s]));
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < N; ++s) a.dnext[s * pl + cell] = acc[s];
    }
  };

  auto load_row = [&](int row, T u[N], Cell& cc) {
    const long long cell = (long long)row * sx + cl;
    S q0[N];
#pragma unroll
    for (int s = 0; s < N; ++s) q0[s] = a.u0[s * pl + cell];
#pragma unroll
    for (int s = 0; s < N; ++s) u[s] = T(q0[s]) + (a.dk ? a.dk[s * pl + cell] : T(0));
    cc = P::template cell<S>(q0, gamma);
  };

  if constexpr (DIM == 1) {
    T u[N];
    Cell cc;
    load_row(0, u, cc);
    T div[N];
    xdiv(u, cc, div);
    finish(0, div);
  } else {
    const T dt_dy = T(dtd * a.inv_dx[1]);
    const T idy = T(a.inv_dx[1]);
    const int y0 = 2 + chunk * a.rows;
    int y1 = y0 + a.rows;
    if (y1 > sy - 2) y1 = sy - 2;

    // rolling window
    T u_prev[N];
    Cell c_prev;
    T dq_m1[N], dq_m2[N], fb_m1[N], Fy_prev[N], divx_m1[N], divx_m2[N];
    Face fy_m1;
#pragma unroll
    for (int s = 0; s < N; ++s) {
      dq_m1[s] = dq_m2[s] = fb_m1[s] = Fy_prev[s] = divx_m1[s] = divx_m2[s] = T(0);
    }
    fy_m1.H = fy_m1.c = fy_m1.q2 = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) fy_m1.v[d] = T(0);

    load_row(y0 - 2, u_prev, c_prev);

#pragma unroll 1
    for (int r = y0 - 1; r <= y1 + 1; ++r) {
      T u[N];
      Cell cc;
      load_row(r, u, cc);

      T divx[N];
      if (r >= y0 && r < y1) {
        xdiv(u, cc, divx);
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) divx[s] = T(0);
      }

      // y-face r (between rows r-1 and r), rotated frame
      const Face fy = P::face(c_prev, cc, 1, gamma);
      T dq[N], fb[N];
      {
        T dlt[N], avg[N], lq[N], lam[N];
#pragma unroll
        for (int s = 0; s < N; ++s) {
          dlt[s] = u[s] - u_prev[s];
          avg[s] = T(.5) * (u[s] + u_prev[s]);
        }
        P::rotate(dlt, 1);
        P::rotate(avg, 1);
        P::left(fy, dlt, dq, gamma);
        P::left(fy, avg, lq, gamma);
        P::eig(fy, lam);
        const bool ok = col_face && r >= 2 && r <= sy - 2;
#pragma unroll
        for (int s = 0; s < N; ++s) {
          if (!ok) dq[s] = T(0);
          fb[s] = lq[s];
        }
      }

      T Fy[N];
      if (r >= y0 + 1) {
        // finalise the flux at face r-1: its own basis fy_m1, neighbours r-2, r
        T lam[N], ft[N];
        P::eig(fy_m1, lam);
        RoeFlux<DIM, T, LIM>::limited(lam, fb_m1, dq_m1, dq_m2, dq, dt_dy, a.lim, ft);
        P::right(fy_m1, ft, Fy);
        P::rotate(Fy, 1);
        if (r >= y0 + 2) {
          T deriv[N];
#pragma unroll
          for (int s = 0; s < N; ++s) deriv[s] = divx_m2[s] - (Fy[s] - Fy_prev[s]) * idy;
          finish(r - 2, deriv);
        }
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) Fy[s] = T(0);
      }

      // shift the window
#pragma unroll
      for (int s = 0; s < N; ++s) {
        dq_m2[s] = dq_m1[s];
        dq_m1[s] = dq[s];
        fb_m1[s] = fb[s];
        Fy_prev[s] = Fy[s];
        divx_m2[s] = divx_m1[s];
        divx_m1[s] = divx[s];
        u_prev[s] = u[s];
      }
      fy_m1 = fy;
      c_prev = cc;
    }
  }
}

template <int DIM, typename T, typename S, bool KAHAN>
static void launch_dim(int lim, const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  const int waves = a.nwaves;
  int blocks = (waves + wpb - 1) / wpb;
  blocks = (blocks + 7) & ~7;  // multiple of 8 for the XCD remap
  dim3 grid(blocks), block(64 * wpb);
  switch (lim) {
  case LIM_SUPERBEE:
    hipLaunchKernelGGL((roe_stage_kernel<DIM, T, S, KAHAN, LIM_SUPERBEE>), grid, block, 0, st, a);
    break;
  case LIM_MONOTIZEDCENTRAL:
    hipLaunchKernelGGL((roe_stage_kernel<DIM, T, S, KAHAN, LIM_MONOTIZEDCENTRAL>), grid, block, 0, st, a);
    break;
  case LIM_MINMOD:
    hipLaunchKernelGGL((roe_stage_kernel<DIM, T, S, KAHAN, LIM_MINMOD>), grid, block, 0, st, a);
    break;
  default:
    hipLaunchKernelGGL((roe_stage_kernel<DIM, T, S, KAHAN, LIM_RUNTIME>), grid, block, 0, st, a);
    break;
  }
}

void launch_roe_stage_f64(int dim, int lim, const StageArgs<double, double>& a, int wpb, hipStream_t st) {
  if (dim == 1) launch_dim<1, double, double, false>(lim, a, wpb, st);
  else launch_dim<2, double, double, false>(lim, a, wpb, st);
}

void launch_roe_stage_mixed(int dim, int lim, bool kahan, const StageArgs<float, float>& a, int wpb,
                            hipStream_t st) {
  if (dim == 1) {
    if (kahan) launch_dim<1, float, float, true>(lim, a, wpb, st);
    else launch_dim<1, float, float, false>(lim, a, wpb, st);
  } else {
    if (kahan) launch_dim<2, float, float, true>(lim, a, wpb, st);
    else launch_dim<2, float, float, false>(lim, a, wpb, st);
  }
}

}  // namespace hydro
