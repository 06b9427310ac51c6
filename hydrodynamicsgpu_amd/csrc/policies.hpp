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

template <class Pol, typename S, typename T>
HYDRO_DEV typename Pol::Frozen frozen_of(const S* q0, S pot, T gamma) {
  if constexpr (PolicyPotential<Pol>::value) {
    return Pol::template frozen_pot<S>(q0, T(pot), gamma);
  } else {
    (void)pot;
    return Pol::template frozen<S>(q0, gamma);
  }
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

}  // namespace hydro
