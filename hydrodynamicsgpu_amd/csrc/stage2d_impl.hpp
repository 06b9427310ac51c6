// stage2d_impl.hpp -- fused RK stage for 1D and 2D grids on gfx950.
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
// registers as a window over the last rows, so every state value is read
// from HBM once per stage (plus the halo overlap).  The march is unrolled by
// three so the window rotates by register renaming instead of copies.  The
// frozen start-of-step data (Roe eigenbasis, HLL wave speeds) are recomputed
// from u0 per interface instead of being stored: O(N) bytes per cell instead
// of O(N^2).
//
// Reference semantics kept:
//  * eigen data frozen at the start of the step (src/Solver/Roe.cpp:82-99,
//    src/Solver/HLL.cpp:225-227);
//  * the Roe limiter ratio uses the neighbour interface's own basis
//    (res/Roe.cl:218-220, 278-290) and the full-step dt (res/Roe.cl:206, 292);
//  * interfaces the reference never writes (index 1 and size-1 along an
//    axis, or any coordinate outside 2..size-2) contribute deltaQTilde = 0;
//  * solid cells (policies with kSolid) take a zero flux divergence
//    (res/CalcFluxDeriv.cl:28-30).
#pragma once
#include "device_common.hpp"
#include "launch.hpp"
#include "policies.hpp"
#include "phys_srhd.hpp"
#include "phys_maxwell.hpp"
#include "stage_common.hpp"

namespace hydro {

// One wavefront's share of a stage: columns strip*60-2 .. strip*60+61 of
// rows 2 + chunk*rows .. (1D: the single row).  strip and chunk are
// wave-uniform; dtd is the step's dt (*a.dt unless the caller holds it).
template <class Pol, int DIM, typename T, typename S, bool KAHAN, int MODE>
HYDRO_DEV void stage2d_wave(const StageArgs<T, S>& a, int strip, int chunk, double dtd) {
  constexpr int N = Pol::N;
  using Frozen = typename Pol::Frozen;
  using Stage = typename Pol::Stage;
  using FaceA = typename Pol::FaceA;
  static_assert(sizeof(T) == sizeof(S), "stage state and increments share one layout");
  const int lane = threadIdx.x & 63;
  const int sx = a.sx, sy = a.sy;
  const long long pl = a.plane;
  const int c = strip * kStripOut + lane;       // column (x0 - 2 + lane)
  const uint32_t cl = c < sx ? c : sx - 1;      // in-bounds load column
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;
  const bool col_face = c >= 2 && c <= sx - 2;  // the reference writes faces here

  const PhysParams<T> pp = make_params<T>(a);
  const T gamma = pp.gamma;
  const T dt_dx0 = T(dtd * a.inv_dx[0]);
  const T idx0 = T(a.inv_dx[0]);
  const StageCombine<T, S, KAHAN, N, MODE> comb(a, dtd);

  using Comb = StageCombine<T, S, KAHAN, N, MODE>;
  struct Row {
    Stage st;   // stage state (and what the policy derives from it per cell)
    Frozen fz;  // frozen start-of-step data
  };
  // global loads of one row, issued one march step ahead of their use
  struct Raw {
    S q0[N];
    T dk[N];
    S pr[Pol::kFrozenPrims ? N : 1];
    S pot;  // the cell's potential (policies that read one), else 0
    S sol;  // the cell's solid mask (policies that read one), else 0
  };
  constexpr bool kPot = PolicyPotential<Pol>::value;
  constexpr bool kSol = PolicySolid<Pol>::value;

  auto fetch = [&](int row, Raw& w) {
    const long long e = (long long)__builtin_amdgcn_readfirstlane(row) * sx;  // wave-uniform
    comb.load_u0(e, cl, w.q0);
    comb.load_dk(e, cl, w.dk);
    if constexpr (Pol::kFrozenPrims) {
#pragma unroll
      for (int s = 0; s < N; ++s) w.pr[s] = ldp(a.prim, pl, s, e, cl);
    }
    if constexpr (kPot) w.pot = ldp(a.pot, pl, 0, e, cl);
    else w.pot = S(0);
    if constexpr (kSol) w.sol = ldp(a.solid, pl, 0, e, cl);
    else w.sol = S(0);
  };
  auto build = [&](const Raw& w, Row& R) {
    T u[N];
    comb.stage_state(w.q0, w.dk, u);
    R.st = stage_of<Pol>(u, w.pot, gamma);
    if constexpr (Pol::kFrozenPrims) R.fz = Pol::template frozen<S>(w.pr, gamma);
    else R.fz = frozen_of<Pol>(w.q0, w.pot, w.sol, gamma);
  };
  auto load_row = [&](int row, Row& R) {
    Raw w;
    fetch(row, w);
    build(w, R);
  };

  // x-direction flux divergence of one row: -(F(c+1) - F(c)) / dx
  auto xdiv = [&](const Row& R, T div[N]) {
    T ps[Pol::NS], pk[Pol::NF], pkL[Pol::NF];
    Pol::pack_stage(R.st, ps);
#pragma unroll
    for (int i = 0; i < Pol::NS; ++i) ps[i] = lane_from_left(ps[i]);
    Pol::pack(R.fz, pk);
#pragma unroll
    for (int i = 0; i < Pol::NF; ++i) pkL[i] = lane_from_left(pk[i]);
    const Frozen fL = Pol::unpack(pkL);
    const FaceA fa = Pol::face_a(fL, R.fz, Pol::unpack_stage(ps), R.st, 0, col_face, pp);
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
    for (int s = 0; s < N; ++s) div[s] = (F[s] - lane_from_right(F[s])) * idx0;
  };

  if constexpr (DIM == 1) {
    // the combination's operands are loaded with the row, ahead of the flux
    Raw w;
    fetch(0, w);
    typename Comb::Fin fin;
    comb.fetch_fin_rest(0, cl, w.q0, w.dk, fin);
    Row R;
    build(w, R);
    T div[N];
    xdiv(R, div);
    if (solid_cell_of<Pol>(w.sol)) {
#pragma unroll
      for (int s = 0; s < N; ++s) div[s] = T(0);
    }
    if (out_lane) comb.finish(0, c, fin, div);
  } else {
    const T dt_dy = T(dtd * a.inv_dx[1]);
    const T idy = T(a.inv_dx[1]);
    const int y0 = 2 + chunk * a.rows;
    int y1 = y0 + a.rows;
    if (y1 > sy - 2) y1 = sy - 2;
    // face r lies between rows r-1 and r; the reference writes faces 2..sy-2
    auto yok = [&](int r) { return col_face && r >= 2 && r <= sy - 2; };

    // Window at the top of step(r): P = row r-1, FA1 = first pass of face
    // r-1, dq2 = deltaQTilde of face r-2, acc = partial derivative of row r-2
    // (its x-divergence plus the flux through its lower face), nx = the
    // loads of row r (in flight), w1 / w2 = the loads of rows r-1 / r-2 (the
    // combination of row r-2 reuses its u0 and d_k instead of reloading).
    Row P;
    FaceA FA1;
    T dq2[N], acc[N];
    Raw nx, w1, w2;
    {
      Row C;
      Raw w0;
      fetch(y0 - 2, w0);
      fetch(y0 - 1, w1);
      build(w0, P);
      fetch(y0, w0);
      build(w1, C);
      const FaceA Fm = Pol::face_a(P.fz, C.fz, P.st, C.st, 1, yok(y0 - 1), pp);
      fetch(y0 + 1, w1);
      build(w0, P);
      w2 = w0;  // row y0
      FA1 = Pol::face_a(C.fz, P.fz, C.st, P.st, 1, yok(y0), pp);
#pragma unroll
      for (int s = 0; s < N; ++s) dq2[s] = Fm.dq[s];
      fetch(y0 + 2, nx);
      build(w1, C);  // row y0 + 1
      const FaceA F0 = Pol::face_a(P.fz, C.fz, P.st, C.st, 1, yok(y0 + 1), pp);
      T G[N], dx[N];
      Pol::face_b(FA1, dq2, F0.dq, dt_dy, 1, a.lim, G);  // flux through face y0
      xdiv(P, dx);
#pragma unroll
      for (int s = 0; s < N; ++s) {
        acc[s] = dx[s] + G[s] * idy;
        dq2[s] = FA1.dq[s];
      }
      FA1 = F0;
      P = C;
    }

    auto step = [&](int r, bool more) {
      const Raw cur = nx;
      // prefetch the next row (clamped: the last step reloads its own row)
      fetch(r + 1 <= y1 + 1 ? r + 1 : y1 + 1, nx);
      // operands of this step's output row r-2 (u0 and d_k from the window)
      const long long eo = (long long)__builtin_amdgcn_readfirstlane(r - 2) * sx;
      typename Comb::Fin fin;
      comb.fetch_fin_rest(eo, cl, w2.q0, w2.dk, fin);
      Row C;
      build(cur, C);
      const FaceA F0 = Pol::face_a(P.fz, C.fz, P.st, C.st, 1, yok(r), pp);
      T G[N];
      Pol::face_b(FA1, dq2, F0.dq, dt_dy, 1, a.lim, G);  // flux through face r-1
      if (out_lane) {
        T deriv[N];
#pragma unroll
        for (int s = 0; s < N; ++s) deriv[s] = acc[s] - G[s] * idy;
        if (solid_cell_of<Pol>(w2.sol)) {
#pragma unroll
          for (int s = 0; s < N; ++s) deriv[s] = T(0);
        }
        comb.finish(eo, c, fin, deriv);
      }
      if (more) {
        T dx[N];
        xdiv(P, dx);
#pragma unroll
        for (int s = 0; s < N; ++s) acc[s] = dx[s] + G[s] * idy;
      }
#pragma unroll
      for (int s = 0; s < N; ++s) dq2[s] = FA1.dq[s];
      FA1 = F0;
      P = C;
      w2 = w1;
      w1 = cur;
    };

    int r = y0 + 2;
#pragma unroll 1
    for (; r + 2 <= y1; r += 3) {
      step(r, true);
      step(r + 1, true);
      step(r + 2, true);
    }
#pragma unroll 1
    for (; r <= y1; ++r) step(r, true);
    step(y1 + 1, false);
  }
}

template <class Pol, int DIM, typename T, typename S, bool KAHAN, int MODE>
__global__ void __launch_bounds__(256) stage2d_kernel(StageArgs<T, S> a) {
  // XCD-aware block order: consecutive logical blocks (neighbouring strips)
  // land on the same XCD and share its L2.  The wave index is wave-uniform;
  // readfirstlane lets the compiler keep the row arithmetic in scalars.
  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int b = blockIdx.x;
  const int lb = (b & 7) * per + (b >> 3);
  const int wave = __builtin_amdgcn_readfirstlane(lb * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
  if (wave >= a.nwaves) return;
  stage2d_wave<Pol, DIM, T, S, KAHAN, MODE>(a, wave % a.nstrips, wave / a.nstrips, *a.dt);
}

template <class Pol, int DIM, typename T, typename S, bool KAHAN, int MODE>
static void launch_pol_mode(const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  int blocks = (a.nwaves + wpb - 1) / wpb;
  blocks = (blocks + 7) & ~7;  // multiple of 8 for the XCD remap
  hipLaunchKernelGGL((stage2d_kernel<Pol, DIM, T, S, KAHAN, MODE>), dim3(blocks), dim3(64 * wpb), 0, st, a);
}

// Fixed-function combination modes are instantiated for the 2D policies
// with compile-time limiters; everything else runs CM_GENERIC.
template <class Pol, int DIM, typename T, typename S, bool KAHAN, bool MODES>
static void launch_pol(int mode, const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  if constexpr (!KAHAN) {  // generic stage evaluation (plain storage only)
    if (mode == CM_DERIV) {
      launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_DERIV>(a, wpb, st);
      return;
    }
  }
  if constexpr (MODES) {
    switch (mode) {
    case CM_FIRST: launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_FIRST>(a, wpb, st); return;
    case CM_MID: launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_MID>(a, wpb, st); return;
    case CM_FINAL: launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_FINAL>(a, wpb, st); return;
    case CM_SINGLE: launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_SINGLE>(a, wpb, st); return;
    default: break;
    }
  }
  launch_pol_mode<Pol, DIM, T, S, KAHAN, CM_GENERIC>(a, wpb, st);
}

// kind: 0 Euler Roe, 1 Euler HLL, 2 Euler HLLC, 3 SRHD Roe, 4 Maxwell Roe,
// 5 EulerBurgers (advection part)
template <int DIM, typename T, typename S, bool KAHAN>
static void launch_solver(int kind, int lim, int mode, const StageArgs<T, S>& a, int wpb, hipStream_t st) {
  constexpr bool M = DIM == 2;
  if (a.solid && kind == 0) {
    launch_pol<SolidRoePolicy<DIM, T>, DIM, T, S, KAHAN, false>(mode, a, wpb, st);
  } else if (a.solid && kind == 5) {
    launch_pol<SolidBurgersPolicy<DIM, T>, DIM, T, S, KAHAN, false>(mode, a, wpb, st);
  } else if (a.pot && (kind == 0 || kind == 1)) {
    // a non-zero gravitational potential: the runtime-limiter Roe form, or HLL
    if (kind == 0) launch_pol<RoePolicy<DIM, T, LIM_RUNTIME, true>, DIM, T, S, KAHAN, false>(mode, a, wpb, st);
    else launch_pol<HLLPolicy<DIM, T, 1, true>, DIM, T, S, KAHAN, false>(mode, a, wpb, st);
  } else if (kind == 5) {
    if (lim == LIM_SUPERBEE) launch_pol<BurgersPolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
    else launch_pol<BurgersPolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN, false>(mode, a, wpb, st);
  } else if (kind == 1) {
    launch_pol<HLLPolicy<DIM, T, 1>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
  } else if (kind == 2) {
    launch_pol<HLLPolicy<DIM, T, 2>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
  } else if (kind == 3) {
    switch (lim) {
    case LIM_SUPERBEE: launch_pol<SRHDRoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN, M>(mode, a, wpb, st); break;
    case LIM_MONOTIZEDCENTRAL:
      launch_pol<SRHDRoePolicy<DIM, T, LIM_MONOTIZEDCENTRAL>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
      break;
    default: launch_pol<SRHDRoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN, false>(mode, a, wpb, st); break;
    }
  } else if (kind == 4) {
    switch (lim) {
    case LIM_SUPERBEE:
      launch_pol<MaxwellRoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
      break;
    default: launch_pol<MaxwellRoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN, false>(mode, a, wpb, st); break;
    }
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch_pol<RoePolicy<DIM, T, LIM_SUPERBEE>, DIM, T, S, KAHAN, M>(mode, a, wpb, st); break;
    case LIM_MONOTIZEDCENTRAL:
      launch_pol<RoePolicy<DIM, T, LIM_MONOTIZEDCENTRAL>, DIM, T, S, KAHAN, M>(mode, a, wpb, st);
      break;
    case LIM_MINMOD: launch_pol<RoePolicy<DIM, T, LIM_MINMOD>, DIM, T, S, KAHAN, M>(mode, a, wpb, st); break;
    default: launch_pol<RoePolicy<DIM, T, LIM_RUNTIME>, DIM, T, S, KAHAN, false>(mode, a, wpb, st); break;
    }
  }
}

}  // namespace hydro
