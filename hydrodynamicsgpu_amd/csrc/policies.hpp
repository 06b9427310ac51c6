// policies.hpp -- face solvers plugged into the fused stage kernels.
//
// A policy turns (frozen start-of-step data, stage state) of the two cells of
// an interface into the interface flux in grid order.  It runs in two passes
// so the Roe limiter can see its neighbouring interfaces:
//   face_a: per-interface work that needs only the interface's two cells
//           (Roe: frozen basis, deltaQTilde = L (qR - qL), L (qL + qR));
//   face_b: the flux, given the neighbours' deltaQTilde (Roe) or nothing
//           (HLL/HLLC, first order: res/EulerHLL.cl:317-335 keeps the
//           limiter disabled).
#pragma once
#include <type_traits>

#include "device_common.hpp"
#include "phys_euler.hpp"

namespace hydro {

// Policies whose faces read a gravitational potential declare
// kPotential = true and take the cell's potential in frozen_pot and
// stage_pot; the kernels load one potential plane for them.
template <class P, class = void>
struct PolicyPotential : std::false_type {};
template <class P>
struct PolicyPotential<P, std::void_t<decltype(P::kPotential)>> : std::bool_constant<P::kPotential> {};

// Policies that read the solid mask (kSolid = true) take it in frozen_solid:
// one value per cell whose bit 0 is the cell's own solid flag and bits
// 1+2a / 2+2a those of its lower / upper neighbour along axis a.
template <class P, class = void>
struct PolicySolid : std::false_type {};
template <class P>
struct PolicySolid<P, std::void_t<decltype(P::kSolid)>> : std::bool_constant<P::kSolid> {};

template <class Pol, typename S, typename T>
HYDRO_DEV typename Pol::Frozen frozen_of(const S* q0, S pot, T gamma) {
  if constexpr (PolicyPotential<Pol>::value) {
    return Pol::template frozen_pot<S>(q0, T(pot), gamma);
  } else {
    (void)pot;
    return Pol::template frozen<S>(q0, gamma);
  }
}
template <class Pol, typename S, typename T>
HYDRO_DEV typename Pol::Frozen frozen_of(const S* q0, S pot, S sol, T gamma) {
  if constexpr (PolicySolid<Pol>::value) return Pol::template frozen_solid<S>(q0, T(pot), T(sol), gamma);
  else return frozen_of<Pol>(q0, pot, gamma);
}
template <class Pol, typename S>
HYDRO_DEV bool solid_cell_of(S sol) {
  if constexpr (PolicySolid<Pol>::value) return (int(sol) & 1) != 0;
  else return false;
}
template <class Pol, typename T, typename S>
HYDRO_DEV typename Pol::Stage stage_of(const T* u, S pot, T gamma) {
  if constexpr (PolicyPotential<Pol>::value) {
    return Pol::stage_pot(u, T(pot), gamma);
  } else {
    (void)pot;
    return Pol::stage(u, gamma);
  }
}

// Per-cell stage data a policy's face_a reads, built once per cell by
// stage() and moved between lanes as NS values (pack_stage/unpack_stage).
// The Roe family reads the stage state itself.
template <int N, typename T>
struct StageIsState {
  struct Stage {
    T u[N];
  };
  static constexpr int NS = N;
  static HYDRO_DEV Stage stage(const T u[N], T) {
    Stage c;
#pragma unroll
    for (int s = 0; s < N; ++s) c.u[s] = u[s];
    return c;
  }
  static HYDRO_DEV void pack_stage(const Stage& c, T o[N]) {
#pragma unroll
    for (int s = 0; s < N; ++s) o[s] = c.u[s];
  }
  static HYDRO_DEV Stage unpack_stage(const T o[N]) { return stage(o, T(0)); }
};

// minmod(a, b): the argument of smaller magnitude when a and b have the same
// sign, else 0 -- the median of (a, b, 0), one v_med3_f32 in FP32.
template <typename T>
HYDRO_DEV T minmod(T a, T b) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_fmed3f(a, b, T(0));
  else return tmax(tmin(a, b), tmin(T(0), tmax(a, b)));
}

// Half the limited slope, dq phi(r) / 2 with r = num / dq (res/Roe.cl:272-298).
// The limiters with compile-time piecewise-linear forms skip the division:
// dq phi(num / dq) is homogeneous in (dq, num), so
//   MC       (phi = max(0, min(2r, (1+r)/2, 2))):  minmod(dq, minmod((dq+num)/4, num))
//   Superbee (phi = max(0, min(2r, 1), min(r, 2))): the larger in magnitude of
//            minmod(num, dq/2) and minmod(num/2, dq), which carry dq's sign or are 0
//   MinMod   (phi = max(0, min(1, r))):             minmod(dq, num) / 2
// dq = 0 gives 0, which is what the product gives for the reference's r =
// +-inf or 0/0 (phi stays finite under fmin/fmax).  The halving is exact.
template <int LIM, typename T>
HYDRO_DEV T half_limited_slope(int lim, T num, T d) {
  if constexpr (LIM == LIM_MONOTIZEDCENTRAL) {
    (void)lim;
    return minmod(d, minmod((d + num) * T(.25), num));
  } else if constexpr (LIM == LIM_SUPERBEE) {
    (void)lim;
    const T a = minmod(num, T(.5) * d), b = minmod(T(.5) * num, d);
    return copysign(tmax(tabs(a), tabs(b)), d);
  } else if constexpr (LIM == LIM_MINMOD) {
    (void)lim;
    return T(.5) * minmod(d, num);
  } else {
    return (T(.5) * d) * limiter<LIM, T>(lim, tdiv(num, d));
  }
}

// The characteristic flux of one wave k of an interface (res/Roe.cl:272-298):
//   l qavg - 1/2 l dq (theta + phi(r) (eps - theta)),
// theta = +1 for l >= 0 else -1, eps = l dt/dx, r from the upwind
// neighbour's dq (dqlo for l >= 0, dqhi otherwise).  With l theta = |l| and
// lq2 = L (qL + qR) = 2 L qavg it is evaluated as
//   (l/2) lq2 - (|l/2| dq + (dq phi(r) / 2) (l eps - |l|)),
// every halving exact, so the product is the same expression term by term.
template <int LIM, typename T>
HYDRO_DEV T wave_flux(int lim, T l, T lq2, T d, T dqlo, T dqhi, T dt_dx) {
  const bool up = l >= T(0);
  const T num = up ? dqlo : dqhi;
  const T hl = T(.5) * l;
  const T hd = half_limited_slope<LIM, T>(lim, num, d);
  const T w = l * (l * dt_dx) - tabs(l);
  return hl * lq2 - (tabs(hl) * d + hd * w);
}

// ------------------------------------------------------------------ Roe
template <int DIM, typename T, int LIM, bool POT = false>
struct RoePolicy : StageIsState<DIM + 2, T> {
  using P = EulerRoe<DIM, T, POT>;
  static constexpr int N = P::N;
  using Stage = typename StageIsState<N, T>::Stage;
  static constexpr bool kNeighbors = true;
  static constexpr bool kFrozenPrims = false;
  static constexpr bool kPotential = POT;
  using Frozen = typename P::Cell;
  static constexpr int NF = DIM + 2 + (POT ? 1 : 0);  // w, v[DIM], H (, phi)

  struct FaceA {
    typename P::Face f;
    T dq[N];  // deltaQTilde of this interface (0 where the reference never writes it)
    T lq2[N]; // L * (qL + qR), twice L * qavg
  };

  template <typename S>
  static HYDRO_DEV Frozen frozen(const S q0[N], T gamma) {
    return P::template cell<S>(q0, gamma);
  }
  template <typename S>
  static HYDRO_DEV Frozen frozen_pot(const S q0[N], T phi, T gamma) {
    return P::template cell<S>(q0, gamma, phi);
  }
  static HYDRO_DEV Stage stage_pot(const T u[N], T, T gamma) { return StageIsState<N, T>::stage(u, gamma); }
  static HYDRO_DEV void pack(const Frozen& c, T o[NF]) {
    o[0] = c.w;
#pragma unroll
    for (int d = 0; d < DIM; ++d) o[1 + d] = c.v[d];
    o[DIM + 1] = c.H;
    if constexpr (POT) o[DIM + 2] = c.phi;
  }
  static HYDRO_DEV Frozen unpack(const T o[NF]) {
    Frozen c;
    c.w = o[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) c.v[d] = o[1 + d];
    c.H = o[DIM + 1];
    if constexpr (POT) c.phi = o[DIM + 2];
    else c.phi = T(0);
    return c;
  }

  // res/Roe.cl:77-150 (deltaQTilde) and the L qavg half of res/Roe.cl:264-270 (as L (qL + qR))
  static HYDRO_DEV FaceA face_a(const Frozen& cL, const Frozen& cR, const Stage& sL, const Stage& sR, int side,
                                bool valid, const PhysParams<T>& pp) {
    const T gamma = pp.gamma;
    FaceA a;
    a.f = P::face(cL, cR, side, gamma);
    T dlt[N], sum[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      dlt[s] = sR.u[s] - sL.u[s];
      sum[s] = sR.u[s] + sL.u[s];
    }
    P::rotate(dlt, side);
    P::rotate(sum, side);
    P::left(a.f, dlt, a.dq, gamma);
    P::left(a.f, sum, a.lq2, gamma);
    if (!valid) {
#pragma unroll
      for (int s = 0; s < N; ++s) a.dq[s] = T(0);
    }
    return a;
  }

  // res/Roe.cl:272-298: limited characteristic flux, then R back to grid order
  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int side, int lim,
                               T F[N]) {
    T lam[N], ft[N];
    P::eig(a.f, lam);
#pragma unroll
    for (int k = 0; k < N; ++k) ft[k] = wave_flux<LIM, T>(lim, lam[k], a.lq2[k], a.dq[k], dqlo[k], dqhi[k], dt_dx);
    P::right(a.f, ft, F);
    P::rotate(F, side);
  }
};

// ------------------------------------------------------------------ Roe, solid cells
// The SOLID build of EulerRoe (src/Solver/EulerRoe.cpp:11-26).  Across a face
// between a solid and a fluid cell the solid side is replaced by the fluid
// cell:
//  * eigenbasis (res/EulerRoe.cl:48-65): the fluid cell's density and
//    velocity, with component side+1 of the velocity 4-vector (vx, vy, vz, 0)
//    negated (y on x-faces, z on y-faces, the padding on z-faces), and the
//    solid cell's own total energy and potential;
//  * deltaQTilde and the flux average (res/Roe.cl:110-124, 235-248): the
//    fluid cell's stage state with the normal momentum negated;
//  * the limiter ratio is 1 when the cell two over on the upwind side is
//    solid (res/Roe.cl:249-250, 278-290);
// and solid cells take no flux divergence (res/CalcFluxDeriv.cl:28-30; the
// kernels zero their derivative).  The frozen data are the raw (rho, v, E,
// phi) and the mask, the Roe averages formed per face after the mirror.
// The potential is always read (a zero plane gives the plain Roe values:
// x - 0 is x); the limiter is the runtime one.
template <int DIM, typename T>
struct SolidRoePolicy : StageIsState<DIM + 2, T> {
  using P = EulerRoe<DIM, T, true>;
  static constexpr int N = P::N;
  static constexpr int E = DIM + 1;
  using Stage = typename StageIsState<N, T>::Stage;
  static constexpr bool kNeighbors = true;
  static constexpr bool kFrozenPrims = false;
  static constexpr bool kPotential = true;
  static constexpr bool kSolid = true;
  struct Frozen {
    T rho;
    T v[DIM];
    T en, phi, mask;
  };
  static constexpr int NF = DIM + 4;
  struct FaceA {
    typename P::Face f;
    T dq[N];
    T lq2[N];
    bool one_lo, one_hi;  // the ratio is 1 for waves moving right / left
  };

  template <typename S>
  static HYDRO_DEV Frozen frozen_solid(const S q0[N], T phi, T mask, T) {
    Frozen c;
    c.rho = T(q0[0]);
#pragma unroll
    for (int a = 0; a < DIM; ++a) c.v[a] = tdiv(T(q0[1 + a]), c.rho);
    c.en = T(q0[E]);
    c.phi = phi;
    c.mask = mask;
    return c;
  }
  static HYDRO_DEV Stage stage_pot(const T u[N], T, T gamma) { return StageIsState<N, T>::stage(u, gamma); }
  static HYDRO_DEV void pack(const Frozen& c, T o[NF]) {
    o[0] = c.rho;
#pragma unroll
    for (int a = 0; a < DIM; ++a) o[1 + a] = c.v[a];
    o[DIM + 1] = c.en;
    o[DIM + 2] = c.phi;
    o[DIM + 3] = c.mask;
  }
  static HYDRO_DEV Frozen unpack(const T o[NF]) {
    Frozen c;
    c.rho = o[0];
#pragma unroll
    for (int a = 0; a < DIM; ++a) c.v[a] = o[1 + a];
    c.en = o[DIM + 1];
    c.phi = o[DIM + 2];
    c.mask = o[DIM + 3];
    return c;
  }

  static HYDRO_DEV FaceA face_a(const Frozen& cL, const Frozen& cR, const Stage& sL, const Stage& sR, int side,
                                bool valid, const PhysParams<T>& pp) {
    const T gamma = pp.gamma;
    const int mL = int(cL.mask), mR = int(cR.mask);
    const bool sl = (mL & 1) != 0, sr = (mR & 1) != 0;
    const bool mirL = sl && !sr, mirR = sr && !sl;
    T rL = cL.rho, rR = cR.rho, vL[DIM], vR[DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const bool flip = a == side + 1;
      vL[a] = mirL ? (flip ? -cR.v[a] : cR.v[a]) : cL.v[a];
      vR[a] = mirR ? (flip ? -cL.v[a] : cL.v[a]) : cR.v[a];
    }
    if (mirL) rL = rR;
    if (mirR) rR = rL;
    FaceA a;
    a.f = P::face(P::cell_rv(rL, vL, cL.en, cL.phi, gamma), P::cell_rv(rR, vR, cR.en, cR.phi, gamma), side,
                  gamma);
    T dlt[N], sum[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const bool normal = s == side + 1;
      T uL = sL.u[s], uR = sR.u[s];
      if (mirL) uL = normal ? -sR.u[s] : sR.u[s];
      if (mirR) uR = normal ? -sL.u[s] : sL.u[s];
      dlt[s] = uR - uL;
      sum[s] = uR + uL;
    }
    P::rotate(dlt, side);
    P::rotate(sum, side);
    P::left(a.f, dlt, a.dq, gamma);
    P::left(a.f, sum, a.lq2, gamma);
    if (!valid) {
#pragma unroll
      for (int s = 0; s < N; ++s) a.dq[s] = T(0);
    }
    a.one_lo = ((mL >> (1 + 2 * side)) & 1) != 0;
    a.one_hi = ((mR >> (2 + 2 * side)) & 1) != 0;
    return a;
  }

  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int side, int lim,
                               T F[N]) {
    T lam[N], ft[N];
    P::eig(a.f, lam);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const T l = lam[k], d = a.dq[k];
      const bool up = l >= T(0);
      const T r = (up ? a.one_lo : a.one_hi) ? T(1) : tdiv(up ? dqlo[k] : dqhi[k], d);
      const T hl = T(.5) * l;
      const T hd = (T(.5) * d) * limiter<LIM_RUNTIME, T>(lim, r);
      const T w = l * (l * dt_dx) - tabs(l);
      ft[k] = hl * a.lq2[k] - (tabs(hl) * d + hd * w);
    }
    P::right(a.f, ft, F);
    P::rotate(F, side);
  }
};

// ------------------------------------------------------------------ HLL / HLLC
// POT (HLL only): the cells' potentials enter the pressures and the
// Roe-averaged sound speed (res/EulerHLL.cl:56-93, 226-240); HLLC reads none
// (res/EulerHLLC.cl).
template <int DIM, typename T, int SOLVER, bool POT = false>  // SOLVER: 1 HLL, 2 HLLC
struct HLLPolicy {
  static_assert(!POT || SOLVER == 1, "only HLL reads the potential");
  static constexpr int N = DIM + 2;
  static constexpr int E = DIM + 1;
  static constexpr bool kNeighbors = false;
  static constexpr bool kFrozenPrims = false;
  static constexpr bool kPotential = POT;
  struct Frozen {
    T rho, p, cs, w, H;
    T v[DIM];
    T phi;  // POT only
  };
  static constexpr int NF = DIM + 5 + (POT ? 1 : 0);
  struct FaceA {
    T F[N];
    T dq[N];  // unused (no limiter); kept so the kernels stay generic
  };
  // stage state with its velocity, pressure and total enthalpy, computed once
  // per cell (each cell takes part in 2*DIM faces)
  struct Stage {
    T q[N];
    T v[DIM];
    T p, h;
  };
  static constexpr int NS = N + DIM + 2;

  static HYDRO_DEV Stage stage(const T u[N], T gamma) { return stage_pot(u, T(0), gamma); }
  static HYDRO_DEV Stage stage_pot(const T u[N], T phi, T gamma) {
    Stage c;
    const T d = u[0];
    const T inv = trcp(d);
    T vsq = T(0);
#pragma unroll
    for (int s = 0; s < N; ++s) c.q[s] = u[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      c.v[k] = tdiv(u[1 + k], d);
      vsq += c.v[k] * c.v[k];
    }
    const T eT = u[E] * inv;
    if constexpr (POT) c.p = (gamma - T(1)) * d * (eT - T(.5) * vsq - phi);
    else c.p = (gamma - T(1)) * d * (eT - T(.5) * vsq);
    c.h = eT + c.p * inv;
    return c;
  }
  static HYDRO_DEV void pack_stage(const Stage& c, T o[NS]) {
#pragma unroll
    for (int s = 0; s < N; ++s) o[s] = c.q[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) o[N + k] = c.v[k];
    o[N + DIM] = c.p;
    o[N + DIM + 1] = c.h;
  }
  static HYDRO_DEV Stage unpack_stage(const T o[NS]) {
    Stage c;
#pragma unroll
    for (int s = 0; s < N; ++s) c.q[s] = o[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) c.v[k] = o[N + k];
    c.p = o[N + DIM];
    c.h = o[N + DIM + 1];
    return c;
  }

  template <typename S>
  static HYDRO_DEV Frozen frozen(const S q0[N], T gamma) {
    return frozen_pot<S>(q0, T(0), gamma);
  }
  template <typename S>
  static HYDRO_DEV Frozen frozen_pot(const S q0[N], T phi, T gamma) {
    Frozen c;
    c.phi = phi;
    c.rho = T(q0[0]);
    const T inv = trcp(c.rho);
    T vsq = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      c.v[d] = tdiv(T(q0[1 + d]), c.rho);
      vsq += c.v[d] * c.v[d];
    }
    const T eT = T(q0[E]) * inv;
    if constexpr (POT) c.p = (gamma - T(1)) * c.rho * (eT - T(.5) * vsq - phi);
    else c.p = (gamma - T(1)) * c.rho * (eT - T(.5) * vsq);
    c.H = eT + c.p * inv;
    c.cs = tsqrt((gamma - T(1)) * (c.H - T(.5) * vsq));
    c.w = SOLVER == 1 ? tsqrt(c.rho) : T(0);  // only HLL's Roe averages weigh by sqrt(rho)
    return c;
  }
  // What a face reads of its neighbour's frozen data: HLLC's wave-speed
  // estimate needs rho, p, c and the normal velocity only; HLL's Roe
  // averages need all of it.  Kernels that move a neighbour's frozen data
  // (lane shifts, LDS) move these NFS values.
  static constexpr int NFS = SOLVER == 2 ? 4 : NF;
  static HYDRO_DEV void pack_side(const Frozen& c, int side, T o[NFS]) {
    if constexpr (SOLVER == 2) {
      o[0] = c.rho; o[1] = c.p; o[2] = c.cs; o[3] = c.v[side];
    } else {
      pack(c, o);
    }
  }
  static HYDRO_DEV Frozen unpack_side(const T o[NFS], int side) {
    if constexpr (SOLVER == 2) {
      Frozen c{};
      c.rho = o[0]; c.p = o[1]; c.cs = o[2]; c.v[side] = o[3];
      return c;
    } else {
      return unpack(o);
    }
  }
  static HYDRO_DEV void pack(const Frozen& c, T o[NF]) {
    o[0] = c.rho; o[1] = c.p; o[2] = c.cs; o[3] = c.w; o[4] = c.H;
#pragma unroll
    for (int d = 0; d < DIM; ++d) o[5 + d] = c.v[d];
    if constexpr (POT) o[5 + DIM] = c.phi;
  }
  static HYDRO_DEV Frozen unpack(const T o[NF]) {
    Frozen c;
    c.rho = o[0]; c.p = o[1]; c.cs = o[2]; c.w = o[3]; c.H = o[4];
#pragma unroll
    for (int d = 0; d < DIM; ++d) c.v[d] = o[5 + d];
    if constexpr (POT) c.phi = o[5 + DIM];
    else c.phi = T(0);
    return c;
  }

  // frozen wave speeds of one interface (res/EulerHLL.cl:56-103,
  // res/EulerHLLC.cl:54-97), normal = `side`
  static HYDRO_DEV void speeds(const Frozen& L, const Frozen& R, int side, T gamma, T& sl, T& sr) {
    const T vLn = L.v[side], vRn = R.v[side];
    if constexpr (SOLVER == 1) {
      const T norm = trcp(L.w + R.w);
      T vsq = T(0), vn = T(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const T v = (L.w * L.v[d] + R.w * R.v[d]) * norm;
        vsq += v * v;
        if (d == side) vn = v;
      }
      const T H = (L.w * L.H + R.w * R.H) * norm;
      T c;
      if constexpr (POT) c = tsqrt((gamma - T(1)) * (H - T(.5) * vsq - (L.w * L.phi + R.w * R.phi) * norm));
      else c = tsqrt((gamma - T(1)) * (H - T(.5) * vsq));
      sl = tmin(vLn - L.cs, vn - c);
      sr = tmax(vRn + R.cs, vn + c);
    } else {
      const T ps = tmax(T(0), T(.5) * (L.p + R.p) - T(.5) * (vRn - vLn) * T(.5) * (L.rho + R.rho) * T(.5) * (L.cs + R.cs));
      const T k = (gamma + T(1)) / (T(2) * gamma);
      const T qL = ps <= L.p ? T(1) : tsqrt(T(1) + k * (tdiv(ps, L.p) - T(1)));
      const T qR = ps <= R.p ? T(1) : tsqrt(T(1) + k * (tdiv(ps, R.p) - T(1)));
      sl = vLn - L.cs * qL;
      sr = vRn + R.cs * qR;
    }
  }

  // res/EulerHLL.cl:185-343, res/EulerHLLC.cl:178-419 (HLLC_METHOD 2)
  static HYDRO_DEV FaceA face_a(const Frozen& cL, const Frozen& cR, const Stage& sL, const Stage& sR, int side,
                                bool valid, const PhysParams<T>& pp) {
    const T gamma = pp.gamma;
    (void)valid;
    FaceA a;
    T sl, sr;
    speeds(cL, cR, side, gamma, sl, sr);
    // both cells in the frame whose component 0 is the face normal
    T qL[N], qR[N], vL[DIM], vR[DIM];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      qL[s] = sL.q[s];
      qR[s] = sR.q[s];
      a.dq[s] = T(0);
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      vL[k] = sL.v[k];
      vR[k] = sR.v[k];
    }
    EulerRoe<DIM, T>::rotate(qL, side);
    EulerRoe<DIM, T>::rotate(qR, side);
    if (side != 0) {
      T t = vL[0]; vL[0] = vL[side]; vL[side] = t;
      t = vR[0]; vR[0] = vR[side]; vR[side] = t;
    }
    const T dL = qL[0], dR = qR[0];
    const T pL = sL.p, pR = sR.p;
    const T hL = sL.h, hR = sR.h;
    T fL[N], fR[N];
    fL[0] = dL * vL[0];
    fR[0] = dR * vR[0];
    fL[1] = dL * vL[0] * vL[0] + pL;
    fR[1] = dR * vR[0] * vR[0] + pR;
#pragma unroll
    for (int d = 1; d < DIM; ++d) {
      fL[1 + d] = dL * vL[d] * vL[0];
      fR[1 + d] = dR * vR[d] * vR[0];
    }
    fL[E] = dL * hL * vL[0];
    fR[E] = dR * hR * vR[0];
    T F[N];
    if constexpr (SOLVER == 1) {
      if (T(0) <= sl) {
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] = fL[s];
      } else if (sr <= T(0)) {
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] = fR[s];
      } else {
        const T inv = trcp(sr - sl);
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] = (sr * fL[s] - sl * fR[s] + sl * sr * (qR[s] - qL[s])) * inv;
      }
    } else {
      const T sStar = tdiv(dR * vR[0] * (sr - vR[0]) - dL * vL[0] * (sl - vL[0]) + pL - pR,
                           dR * (sr - vR[0]) - dL * (sl - vL[0]));
      // branch order of res/EulerHLLC.cl:276-364
      if (T(0) <= sl) {
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] = fL[s];
      } else if (!(sStar <= T(0) && T(0) <= sr) && !(T(0) <= sStar) && sr <= T(0)) {
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] = fR[s];
      } else {
        const T pLR = T(.5) * (pL + pR + dL * (sl - vL[0]) * (sStar - vL[0]) + dR * (sr - vR[0]) * (sStar - vR[0]));
        const bool left = T(0) <= sStar;
        const T sK = left ? sl : sr;
        const T inv = trcp(sK - sStar);
#pragma unroll
        for (int s = 0; s < N; ++s) {
          const T q = left ? qL[s] : qR[s];
          const T f = left ? fL[s] : fR[s];
          F[s] = sStar * (sK * q - f);
        }
        F[1] += sK * pLR;
        F[E] += sK * pLR * sStar;
#pragma unroll
        for (int s = 0; s < N; ++s) F[s] *= inv;
      }
    }
    EulerRoe<DIM, T>::rotate(F, side);
#pragma unroll
    for (int s = 0; s < N; ++s) a.F[s] = F[s];
    return a;
  }

  static HYDRO_DEV void face_b(const FaceA& a, const T*, const T*, T, int, int, T F[N]) {
#pragma unroll
    for (int s = 0; s < N; ++s) F[s] = a.F[s];
  }
};

// ------------------------------------------------------------------ EulerBurgers
// res/EulerBurgers.cl:81-281 (advection part of the split Euler step): every
// state is carried by the interface velocity u = (v_L + v_R) / 2 (the mean of
// the two cells' normal velocities) in upwind form, plus a limited
// correction built from the state's own raw jumps -- no eigen-decomposition
// and no frozen start-of-step data:
//   F = u q_up + (dq phi(r) / 2) |u| (1 - |u dt/dx|),
// q_up = q_L for u >= 0 else q_R, r = the upwind neighbour face's jump over
// this face's jump.  The reference's .5 u ((1+theta) q_L + (1-theta) q_R) is
// u q_up exactly (theta = +-1: the factors are 2 and 0, every product and the
// halving exact) for finite states.  Jumps are raw state differences on every
// face, so no face reads as zero (unlike the Roe family's never-written
// deltaQTilde slots).
template <int DIM, typename T, int LIM>
struct BurgersPolicy {
  static constexpr int N = DIM + 2;
  static constexpr bool kNeighbors = true;
  static constexpr bool kFrozenPrims = false;
  struct Frozen {};
  static constexpr int NF = 1;  // one placeholder value (device code has no zero-length arrays)
  // stage state and its velocities, divided once per cell (the reference
  // divides per face: the same quotient)
  struct Stage {
    T u[N];
    T v[DIM];
  };
  static constexpr int NS = N + DIM;
  struct FaceA {
    T dq[N];  // q_R - q_L
    T q[N];   // the upwind state
    T ui;     // interface velocity
  };

  static HYDRO_DEV Stage stage(const T u[N], T) {
    Stage c;
#pragma unroll
    for (int s = 0; s < N; ++s) c.u[s] = u[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) c.v[k] = tdiv(u[1 + k], u[0]);
    return c;
  }
  static HYDRO_DEV void pack_stage(const Stage& c, T o[NS]) {
#pragma unroll
    for (int s = 0; s < N; ++s) o[s] = c.u[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) o[N + k] = c.v[k];
  }
  static HYDRO_DEV Stage unpack_stage(const T o[NS]) {
    Stage c;
#pragma unroll
    for (int s = 0; s < N; ++s) c.u[s] = o[s];
#pragma unroll
    for (int k = 0; k < DIM; ++k) c.v[k] = o[N + k];
    return c;
  }
  template <typename S>
  static HYDRO_DEV Frozen frozen(const S*, T) { return Frozen{}; }
  static HYDRO_DEV void pack(const Frozen&, T o[NF]) { o[0] = T(0); }
  static HYDRO_DEV Frozen unpack(const T*) { return Frozen{}; }

  // res/EulerBurgers.cl:98-133 (interface velocity) and the jumps of
  // res/EulerBurgers.cl:216-235
  static HYDRO_DEV FaceA face_a(const Frozen&, const Frozen&, const Stage& sL, const Stage& sR, int side, bool,
                                const PhysParams<T>&) {
    FaceA a;
    a.ui = T(.5) * (sL.v[side] + sR.v[side]);
    const bool up = a.ui >= T(0);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      a.dq[s] = sR.u[s] - sL.u[s];
      a.q[s] = up ? sL.u[s] : sR.u[s];
    }
    return a;
  }

  // res/EulerBurgers.cl:236-263: the ratio from the upwind neighbour face's
  // jump (dqlo for u >= 0, dqhi otherwise), the correction term in the
  // reference's order ((dq phi / 2) |u|) (1 - |u dt/dx|)
  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int, int lim, T F[N]) {
    const bool up = a.ui >= T(0);
    const T au = tabs(a.ui);
    const T w = T(1) - tabs(a.ui * dt_dx);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const T num = up ? dqlo[s] : dqhi[s];
      const T hd = half_limited_slope<LIM, T>(lim, num, a.dq[s]);
      F[s] = a.ui * a.q[s] + (hd * au) * w;
    }
  }
};

// The SOLID build of EulerBurgers (src/Solver/EulerBurgers.cpp:29-58),
// advection part:
//  * interface velocity (res/EulerBurgers.cl:119-127): a solid cell facing a
//    fluid one takes the negated normal velocity of the fluid cell, so the
//    face velocity is exactly 0 and the face carries no flux (unless the
//    limited term is not finite, as in the reference);
//  * flux states (res/EulerBurgers.cl:209-236): the solid side takes the
//    fluid state, normal momentum kept and every other state negated; a
//    solid cell two over repeats the near state with the normal momentum
//    negated, so that jump is 2 q (normal momentum) or 0.
// The upwind neighbour's jump is the raw jump of the neighbour face plus
// this face's own mirror correction (zero unless this face is a wall, where
// the face velocity is 0).  The frozen data are the solid mask alone.
template <int DIM, typename T>
struct SolidBurgersPolicy : BurgersPolicy<DIM, T, LIM_RUNTIME> {
  using Base = BurgersPolicy<DIM, T, LIM_RUNTIME>;
  using Stage = typename Base::Stage;
  static constexpr int N = Base::N;
  static constexpr bool kSolid = true;
  struct Frozen {
    T mask;
  };
  static constexpr int NF = 1;
  struct FaceA {
    T dq[N];  // raw q_R - q_L (the neighbours' ratios)
    T dm[N];  // this face's jump after the mirror
    T ql[N], qr[N];  // this face's states after the mirror
    T cl[N], cr[N];  // q_L' - q_L and q_R - q_R'
    T ui;
    bool one_lo, one_hi;
  };

  template <typename S>
  static HYDRO_DEV Frozen frozen_solid(const S*, T, T mask, T) {
    return Frozen{mask};
  }
  static HYDRO_DEV void pack(const Frozen& c, T o[NF]) { o[0] = c.mask; }
  static HYDRO_DEV Frozen unpack(const T o[NF]) { return Frozen{o[0]}; }

  static HYDRO_DEV FaceA face_a(const Frozen& cL, const Frozen& cR, const Stage& sL, const Stage& sR, int side,
                                bool, const PhysParams<T>&) {
    const int mL = int(cL.mask), mR = int(cR.mask);
    const bool sl = (mL & 1) != 0, sr = (mR & 1) != 0;
    const bool mirL = sl && !sr, mirR = sr && !sl;
    FaceA a;
    T vL = sL.v[side], vR = sR.v[side];
    if (mirL) vL = -vR;
    if (mirR) vR = -vL;
    a.ui = T(.5) * (vL + vR);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const bool normal = s == side + 1;
      T qL = sL.u[s], qR = sR.u[s];
      if (mirL) qL = normal ? sR.u[s] : -sR.u[s];
      if (mirR) qR = normal ? sL.u[s] : -sL.u[s];
      a.dq[s] = sR.u[s] - sL.u[s];
      a.dm[s] = qR - qL;
      a.ql[s] = qL;
      a.qr[s] = qR;
      a.cl[s] = qL - sL.u[s];
      a.cr[s] = sR.u[s] - qR;
    }
    a.one_lo = ((mL >> (1 + 2 * side)) & 1) != 0;
    a.one_hi = ((mR >> (2 + 2 * side)) & 1) != 0;
    return a;
  }

  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int side, int lim, T F[N]) {
    const bool up = a.ui >= T(0);
    const T au = tabs(a.ui);
    const T w = T(1) - tabs(a.ui * dt_dx);
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const bool normal = s == side + 1;
      T num;
      if (up) num = a.one_lo ? a.ql[s] - (normal ? -a.ql[s] : a.ql[s]) : a.cl[s] + dqlo[s];
      else num = a.one_hi ? (normal ? -a.qr[s] : a.qr[s]) - a.qr[s] : dqhi[s] + a.cr[s];
      const T hd = half_limited_slope<LIM_RUNTIME, T>(lim, num, a.dm[s]);
      F[s] = a.ui * (up ? a.ql[s] : a.qr[s]) + (hd * au) * w;
    }
  }
};

}  // namespace hydro
