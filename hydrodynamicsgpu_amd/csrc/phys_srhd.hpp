// phys_srhd.hpp -- special-relativistic hydrodynamics for the Roe solver.
//
// State order (src/Equation/SRHD.cpp:36-40): D, S[dim], tau.  Primitives
// (res/SRHDCommon.cl): rho, v[dim], eInt, kept in FP64.
// FP32 arithmetic divides by reciprocal multiplies (tdiv / trcp, ~1 ulp);
// FP64 keeps the correctly rounded division.
// The interface basis follows res/SRHDRoe.cl:21-301: arithmetic mean of the
// primitives (not a Roe average), Marti-Mueller eigenvalues and analytic left
// and right eigenvectors built for the normal rotated into x, then rotated
// back by -90 degrees (side 1: (x,y) -> (y,-x); side 2: (x,z) -> (z,-x)).
#pragma once
#include "device_common.hpp"
#include "policies.hpp"

namespace hydro {

template <int DIM, typename T, int LIM>
struct SRHDRoePolicy : StageIsState<DIM + 2, T> {
  static constexpr int N = DIM + 2;
  using Stage = typename StageIsState<N, T>::Stage;
  static constexpr int E = DIM + 1;
  static constexpr bool kNeighbors = true;
  static constexpr bool kFrozenPrims = true;

  struct Frozen {
    T rho, eInt;
    T v[DIM];
  };
  static constexpr int NF = DIM + 2;

  struct FaceA {
    T lam[N];
    T R[N][N];   // right eigenvectors, rotated frame: R[row][col]
    T dq[N];
    T lq2[N];
  };

  template <typename P>
  static HYDRO_DEV Frozen frozen(const P prim[N], T) {
    Frozen f;
    f.rho = T(prim[0]);
#pragma unroll
    for (int d = 0; d < DIM; ++d) f.v[d] = T(prim[1 + d]);
    f.eInt = T(prim[E]);
    return f;
  }
  static HYDRO_DEV void pack(const Frozen& f, T o[NF]) {
    o[0] = f.rho;
    o[1] = f.eInt;
#pragma unroll
    for (int d = 0; d < DIM; ++d) o[2 + d] = f.v[d];
  }
  static HYDRO_DEV Frozen unpack(const T o[NF]) {
    Frozen f;
    f.rho = o[0];
    f.eInt = o[1];
#pragma unroll
    for (int d = 0; d < DIM; ++d) f.v[d] = o[2 + d];
    return f;
  }

  // conserved vector, grid order -> rotated frame (x <- side, side <- -x)
  template <typename V>
  static HYDRO_DEV void rot_in(V x[N], int side) {
    if (side != 0) {
      const V t = x[1];
      x[1] = x[1 + side];
      x[1 + side] = -t;
    }
  }
  // rotated frame -> grid order (inverse of rot_in)
  template <typename V>
  static HYDRO_DEV void rot_out(V y[N], int side) {
    if (side != 0) {
      const V t = y[1];
      y[1] = -y[1 + side];
      y[1 + side] = t;
    }
  }

  struct Mean {
    T v[3];  // rotated velocity (x normal), unused components 0
    T vSq, W, W2, oneOverW, h, cs, csSq, lmin, lmax;
  };

  static HYDRO_DEV Mean mean(const Frozen& L, const Frozen& R, int side, T gamma) {
    Mean m;
    const T rho = T(.5) * (L.rho + R.rho);
    const T eInt = T(.5) * (L.eInt + R.eInt);
    T v[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[d] = T(.5) * (L.v[d] + R.v[d]);
    m.vSq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const T oneOverW2 = T(1) - m.vSq;
    m.oneOverW = tsqrt(oneOverW2);
    m.W = trcp(m.oneOverW);
    m.W2 = trcp(oneOverW2);
    const T P = (gamma - T(1)) * rho * eInt;
    m.h = T(1) + eInt + tdiv(P, rho);
    const T P_over_rho_h = tdiv(P, rho * m.h);
    if (side == 1) { m.v[0] = v[1]; m.v[1] = -v[0]; m.v[2] = v[2]; }
    else if (side == 2) { m.v[0] = v[2]; m.v[1] = v[1]; m.v[2] = -v[0]; }
    else { m.v[0] = v[0]; m.v[1] = v[1]; m.v[2] = v[2]; }
    const T vx = m.v[0];
    const T vxSq = vx * vx;
    m.csSq = gamma * P_over_rho_h;
    m.cs = tsqrt(m.csSq);
    const T discr = tsqrt((T(1) - m.vSq) * ((T(1) - m.vSq * m.csSq) - vxSq * (T(1) - m.csSq)));
    const T den = trcp(T(1) - m.vSq * m.csSq);
    m.lmin = (vx * (T(1) - m.csSq) - m.cs * discr) * den;
    m.lmax = (vx * (T(1) - m.csSq) + m.cs * discr) * den;
    return m;
  }

  // res/SRHDRoe.cl:124-240 in the rotated frame
  static HYDRO_DEV void basis(const Mean& m, T lam[N], T Lm[N][N], T Rm[N][N]) {
    const T vx = m.v[0], vy = m.v[1], vz = m.v[2];
    const T h = m.h, W = m.W, W2 = m.W2, vSq = m.vSq;
    const T hW = h * W, hSq = h * h, vxSq = vx * vx;
    const T lmin = m.lmin, lmax = m.lmax;
    lam[0] = lmin;
#pragma unroll
    for (int a = 1; a <= DIM; ++a) lam[a] = vx;
    lam[E] = lmax;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) Lm[i][j] = Rm[i][j] = T(0);
    const T Kappa = h;
    const T AMinus = tdiv(T(1) - vxSq, T(1) - vx * lmin);
    const T APlus = tdiv(T(1) - vxSq, T(1) - vx * lmax);
    // right eigenvectors: column j
    Rm[0][0] = T(1);
    Rm[1][0] = hW * AMinus * lmin;
    if (DIM > 1) Rm[2][0] = hW * vy;
    if (DIM > 2) Rm[3][0] = hW * vz;
    Rm[E][0] = hW * AMinus - T(1);
    Rm[0][1] = m.oneOverW;
    Rm[1][1] = vx;
    if (DIM > 1) Rm[2][1] = vy;
    if (DIM > 2) Rm[3][1] = vz;
    Rm[E][1] = T(1) - m.oneOverW;
    if (DIM > 1) {
      Rm[0][2] = W * vy;
      Rm[1][2] = T(2) * h * W2 * vx * vy;
      Rm[2][2] = h * (T(1) + T(2) * W2 * vy * vy);
      if (DIM > 2) Rm[3][2] = T(2) * h * W2 * vy * vz;
      Rm[E][2] = (T(2) * hW - T(1)) * W * vy;
    }
    if (DIM > 2) {
      Rm[0][3] = W * vz;
      Rm[1][3] = T(2) * h * W2 * vx * vz;
      Rm[2][3] = T(2) * h * W2 * vy * vz;
      Rm[3][3] = h * (T(1) + T(2) * W2 * vz * vz);
      Rm[E][3] = (T(2) * hW - T(1)) * W * vz;
    }
    Rm[0][E] = T(1);
    Rm[1][E] = hW * APlus * lmax;
    if (DIM > 1) Rm[2][E] = hW * vy;
    if (DIM > 2) Rm[3][E] = hW * vz;
    Rm[E][E] = hW * APlus - T(1);

    // left eigenvectors: row i
    const T Delta = hSq * hW * (Kappa - T(1)) * (T(1) - vxSq) * (APlus * lmax - AMinus * lmin);
    const T k2 = T(2) * Kappa - T(1);
    const T tq = W2 * (vSq - vxSq) * k2;
    const T hSq_Delta = tdiv(hSq, Delta);
    T scale = hSq_Delta;
    Lm[0][0] = scale * (hW * APlus * (vx - lmax) - vx - tq * (vx - APlus * lmax) + Kappa * APlus * lmax);
    Lm[0][1] = scale * (T(1) + tq * (T(1) - APlus) - Kappa * APlus);
    if (DIM > 1) Lm[0][2] = scale * (W2 * vy * k2 * APlus * (vx - lmax));
    if (DIM > 2) Lm[0][3] = scale * (W2 * vz * k2 * APlus * (vx - lmax));
    Lm[0][E] = scale * (-vx - tq * (vx - APlus * lmax) + Kappa * APlus * lmax);
    scale = tdiv(W, Kappa - T(1));
    Lm[1][0] = scale * (h - W);
    Lm[1][1] = scale * (W * vx);
    if (DIM > 1) Lm[1][2] = scale * (W * vy);
    if (DIM > 2) Lm[1][3] = scale * (W * vz);
    Lm[1][E] = scale * (-W);
    if (DIM > 1) {
      scale = trcp(h * (T(1) - vxSq));
      Lm[2][0] = scale * (-vy);
      Lm[2][1] = scale * (vx * vy);
      Lm[2][2] = scale * (T(1) - vxSq);
      Lm[2][E] = scale * (-vy);
      if (DIM > 2) {
        Lm[3][0] = scale * (-vz);
        Lm[3][1] = scale * (vx * vz);
        Lm[3][3] = scale * (T(1) - vxSq);
        Lm[3][E] = scale * (-vz);
      }
    }
    scale = -hSq_Delta;
    Lm[E][0] = scale * (hW * AMinus * (vx - lmin) - vx - tq * (vx - AMinus * lmin) + Kappa * AMinus * lmin);
    Lm[E][1] = scale * (T(1) + tq * (T(1) - AMinus) - Kappa * AMinus);
    if (DIM > 1) Lm[E][2] = scale * (W2 * vy * k2 * AMinus * (vx - lmin));
    if (DIM > 2) Lm[E][3] = scale * (W2 * vz * k2 * AMinus * (vx - lmin));
    Lm[E][E] = scale * (-vx - tq * (vx - AMinus * lmin) + Kappa * AMinus * lmin);
  }

  static HYDRO_DEV FaceA face_a(const Frozen& cL, const Frozen& cR, const Stage& sL, const Stage& sR, int side,
                                bool valid, const PhysParams<T>& pp) {
    const T* uL = sL.u;
    const T* uR = sR.u;
    const T gamma = pp.gamma;
    FaceA a;
    const Mean m = mean(cL, cR, side, gamma);
    T Lm[N][N];
    basis(m, a.lam, Lm, a.R);
    T dlt[N], sum[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      dlt[s] = uR[s] - uL[s];
      sum[s] = uR[s] + uL[s];
    }
    rot_in(dlt, side);
    rot_in(sum, side);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T sd = T(0), sa = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        sd += Lm[i][j] * dlt[j];
        sa += Lm[i][j] * sum[j];
      }
      a.dq[i] = valid ? sd : T(0);
      a.lq2[i] = sa;
    }
    return a;
  }

  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int side, int lim,
                               T F[N]) {
    T ft[N];
#pragma unroll
    for (int k = 0; k < N; ++k) ft[k] = wave_flux<LIM, T>(lim, a.lam[k], a.lq2[k], a.dq[k], dqlo[k], dqhi[k], dt_dx);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = T(0);
#pragma unroll
      for (int j = 0; j < N; ++j) s += a.R[i][j] * ft[j];
      F[i] = s;
    }
    rot_out(F, side);
  }
};

// con2prim of one cell (res/SRHDCommon.cl:113-196) in FP64.  Returns false
// when the Newton iteration does not converge or leaves the physical range
// (v^2 >= 1, which would overflow the Lorentz factor); the old primitives
// are kept then, as the reference keeps them on non-convergence.
struct SRHDDefs {
  double gamma;
  int max_iter;
  double stop_eps, vel_eps, pmin_eps;
  double rho_min, rho_max, eint_max;
  double D_min, D_max, tau_min, tau_max;
};

//
// WARM: start the Newton iteration from the cell's previous pressure
// (gamma-1) rho eInt, clamped to [PMin, PMax], instead of the reference's
// bracket midpoint.  Both starts converge to the same root; the stopping
// test (relative change < stop_eps, quadratic convergence) leaves them
// within about stop_eps^2 of it, far below FP32 resolution, and the warm
// start needs a few iterations instead of tens.  The FP32-mixed path uses
// it; the FP64 path keeps the reference's start.
template <int DIM, bool WARM = false>
HYDRO_DEV bool srhd_con2prim(const SRHDDefs& d, const double cons[DIM + 2], double prim[DIM + 2]) {
  constexpr int E = DIM + 1;
  const double gamma = d.gamma;
  const double D = cons[0];
  double S[3] = {0, 0, 0};
  for (int a = 0; a < DIM; ++a) S[a] = cons[1 + a];
  const double tau = cons[E];
  const double SLen = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  const double PMin = fmax(SLen - tau - D + SLen * d.vel_eps, d.pmin_eps);
  double PMax = (gamma - 1.) * tau;
  PMax = fmax(PMax, PMin);
  double P = .5 * (PMin + PMax);
  if constexpr (WARM) {
    const double Pold = (gamma - 1.) * prim[0] * prim[E];
    if (Pold == Pold) P = fmin(fmax(Pold, PMin), PMax);
  }
  for (int iter = 0; iter < d.max_iter; ++iter) {
    const double vLen = SLen / (tau + D + P);
    const double vSq = vLen * vLen;
    if (!(vSq < 1.0)) return false;  // Lorentz-factor overflow guard
    double W = 1. / sqrt(1. - vSq);
    double eInt = (tau + D * (1. - W) + P * (1. - W * W)) / (D * W);
    double rho = D / W;
    const double f = (gamma - 1.) * rho * eInt - P;
    const double csSq = (gamma - 1.) * (tau + D * (1. - W) + P) / (tau + D + P);
    const double df_dP = vSq * csSq - 1.;
    double newP = P - f / df_dP;
    newP = fmax(newP, PMin);
    const double PError = fabs(1. - newP / P);
    P = newP;
    if (PError < d.stop_eps) {
      double v[3];
      for (int a = 0; a < 3; ++a) v[a] = S[a] / (tau + D + P);
      const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (!(vv < 1.0)) return false;
      W = 1. / sqrt(1. - vv);
      rho = D / W;
      rho = fmax(rho, d.rho_min);
      rho = fmin(rho, d.rho_max);
      eInt = P / (rho * (gamma - 1.));
      eInt = fmin(eInt, d.eint_max);
      prim[0] = rho;
      for (int a = 0; a < DIM; ++a) prim[1 + a] = v[a];
      prim[E] = eInt;
      return true;
    }
  }
  return false;
}

}  // namespace hydro
