// phys_maxwell.hpp -- Maxwell's equations for the Roe solver.
//
// State order (src/Equation/Maxwell.cpp:39-46): E[3], B[3].  The basis is
// constant: eigenvalues -1/sqrt(eps mu) (x2), 0 (x2), +1/sqrt(eps mu) (x2)
// (res/MaxwellRoe.cl:117-157); the eigenvector transforms are the closed forms
// of res/MaxwellRoe.cl:13-109, with their FP32-typed constants (se, su and
// their inverses are `const float` there, so L*R = I only to ~1e-7), and the
// characteristic differences are formed separately, L qR - L qL
// (ROE_EIGENFIELD_TRANSFORM_SEPARATE, src/Equation/Maxwell.cpp:66-70).
#pragma once
#include "device_common.hpp"
#include "policies.hpp"

namespace hydro {

template <int DIM, typename T, int LIM>
struct MaxwellRoePolicy : StageIsState<6, T> {
  static constexpr int N = 6;
  using Stage = typename StageIsState<N, T>::Stage;
  static constexpr bool kNeighbors = true;
  static constexpr bool kFrozenPrims = false;
  struct Frozen {
    T unused;
  };
  static constexpr int NF = 1;
  struct FaceA {
    T dq[N];
    T lq2[N];
    T se, su, lam;
  };

  template <typename S>
  static HYDRO_DEV Frozen frozen(const S*, T) {
    return Frozen{T(0)};
  }
  static HYDRO_DEV void pack(const Frozen& f, T o[NF]) { o[0] = f.unused; }
  static HYDRO_DEV Frozen unpack(const T o[NF]) { return Frozen{o[0]}; }

  // res/MaxwellRoe.cl:13-36: swap the normal into x, then project
  static HYDRO_DEV void left(const T x[N], int side, T ise, T isu, T r[N]) {
    T e0 = x[0], e1 = x[1], e2 = x[2], b0 = x[3], b1 = x[4], b2 = x[5];
    if (side == 1) {
      T t = e0; e0 = e1; e1 = t;
      t = b0; b0 = b1; b1 = t;
    } else if (side == 2) {
      T t = e0; e0 = e2; e2 = t;
      t = b0; b0 = b2; b2 = t;
    }
    r[0] = e2 * ise + b1 * isu;
    r[1] = e1 * -ise + b2 * isu;
    r[2] = e0 * -ise + b0 * isu;
    r[3] = e0 * ise + b0 * isu;
    r[4] = e1 * ise + b2 * isu;
    r[5] = e2 * -ise + b1 * isu;
  }

  // res/MaxwellRoe.cl:61-109: reconstruct, then swap the output components
  // exactly as the reference does (side 2 exchanges slots 0<->4 and 3<->5)
  static HYDRO_DEV void right(const T y[N], int side, T se, T su, T r[N]) {
    r[0] = y[2] * -se + y[3] * se;
    r[1] = y[1] * -se + y[4] * se;
    r[2] = y[0] * se + y[5] * -se;
    r[3] = y[2] * su + y[3] * su;
    r[4] = y[0] * su + y[5] * su;
    r[5] = y[1] * su + y[4] * su;
    if (side == 1) {
      T t = r[0]; r[0] = r[1]; r[1] = t;
      t = r[3]; r[3] = r[4]; r[4] = t;
    } else if (side == 2) {
      T t = r[0]; r[0] = r[4]; r[4] = t;
      t = r[3]; r[3] = r[5]; r[5] = t;
    }
  }

  // res/Roe.cl:126-138 and 254-263 (separate transforms)
  static HYDRO_DEV FaceA face_a(const Frozen&, const Frozen&, const Stage& sL, const Stage& sR, int side, bool valid,
                                const PhysParams<T>& pp) {
    FaceA a;
    T tl[N], tr[N];
    left(sL.u, side, pp.mx[2], pp.mx[3], tl);
    left(sR.u, side, pp.mx[2], pp.mx[3], tr);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      a.dq[k] = valid ? tr[k] - tl[k] : T(0);
      a.lq2[k] = tr[k] + tl[k];
    }
    a.se = pp.mx[0];
    a.su = pp.mx[1];
    a.lam = pp.mx[4];
    return a;
  }

  static HYDRO_DEV void face_b(const FaceA& a, const T dqlo[N], const T dqhi[N], T dt_dx, int side, int lim,
                               T F[N]) {
    const T lam[N] = {-a.lam, -a.lam, T(0), T(0), a.lam, a.lam};
    T ft[N];
#pragma unroll
    for (int k = 0; k < N; ++k) ft[k] = wave_flux<LIM, T>(lim, lam[k], a.lq2[k], a.dq[k], dqlo[k], dqhi[k], dt_dx);
    right(ft, side, a.se, a.su, F);
  }
};

}  // namespace hydro
