// phys_euler.hpp -- Euler equations for the Roe and HLL/HLLC face solvers.
//
// State order (src/Equation/Euler.cpp:36-40): rho, rho*v[dim], E.
// Roe face data follow res/EulerRoe.cl:18-259: sqrt(rho)-weighted averages of
// velocity and total enthalpy, c^2 = (gamma-1)(H - v^2/2 - phi), the normal
// rotated into x.  The eigenvector matrices are not materialised: L*x and R*y
// are applied in closed form from (v, H, c), which is algebraically the dense
// product res/RoeEigenfieldLinear.cl:16-74 performs with the stored matrices.
//
// FP32: one v_rsq_f32 gives both sqrt(rho) and 1/rho per cell, and both c and
// 1/c^2 per face (rsq is accurate to ~1 ulp).  FP64 keeps correctly rounded
// sqrt and division, as the reference's FP64 kernels have them.
#pragma once
#include "device_common.hpp"

namespace hydro {

// POT: the cells carry a gravitational potential phi (res/EulerRoe.cl:56-94):
// the pressure subtracts it from the specific energy and the face's sound
// speed subtracts its Roe average; without POT phi is zero and never read.
template <int DIM, typename T, bool POT = false>
struct EulerRoe {
  static constexpr int N = DIM + 2;
  static constexpr int E = DIM + 1;
  static constexpr bool kFast = sizeof(T) == 4;

  // frozen per-cell data from the start-of-step state
  struct Cell {
    T w;        // sqrt(rho)
    T v[DIM];   // velocity
    T H;        // total specific enthalpy
    T phi;      // potential (POT only)
  };

  // frozen per-face data, normal rotated into component 0 of v
  struct Face {
    T v[DIM];   // v[0] normal, v[1..] tangential (rotated order)
    T H, c, q2; // enthalpy, sound speed, |v|^2
    T invD;     // 1 / (2 c^2)
  };

  template <typename S>
  static HYDRO_DEV Cell cell(const S q[N], T gamma, T phi = T(0)) {
    Cell c;
    c.phi = phi;
    const T rho = T(q[0]);
    T inv, vsq = T(0);
    if constexpr (kFast) {
      const T rs = __builtin_amdgcn_rsqf(rho);
      c.w = rho * rs;
      inv = rs * rs;
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        c.v[a] = T(q[1 + a]) * inv;
        vsq += c.v[a] * c.v[a];
      }
    } else {
      c.w = tsqrt(rho);
      inv = T(1) / rho;
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        c.v[a] = T(q[1 + a]) / rho;
        vsq += c.v[a] * c.v[a];
      }
    }
    // H = E/rho + p/rho with p = (gamma-1) rho (E/rho - v^2/2 - phi)
    const T eT = T(q[E]) * inv;
    if constexpr (POT) c.H = eT + (gamma - T(1)) * (eT - T(.5) * vsq - phi);
    else c.H = eT + (gamma - T(1)) * (eT - T(.5) * vsq);
    return c;
  }

  // the same frozen data from density, velocity, total energy and potential
  // given separately (a solid cell's mirrored density and velocity with its
  // own energy and potential)
  static HYDRO_DEV Cell cell_rv(T rho, const T v[DIM], T E_, T phi, T gamma) {
    Cell c;
    c.phi = phi;
    T inv, vsq = T(0);
    if constexpr (kFast) {
      const T rs = __builtin_amdgcn_rsqf(rho);
      c.w = rho * rs;
      inv = rs * rs;
    } else {
      c.w = tsqrt(rho);
      inv = T(1) / rho;
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      c.v[a] = v[a];
      vsq += v[a] * v[a];
    }
    const T eT = E_ * inv;
    c.H = eT + (gamma - T(1)) * (eT - T(.5) * vsq - phi);
    return c;
  }

  static HYDRO_DEV Face face(const Cell& L, const Cell& R, int side, T gamma) {
    Face f;
    const T norm = trcp(L.w + R.w);
    T v[DIM];
    T q2 = T(0);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      v[a] = (L.w * L.v[a] + R.w * R.v[a]) * norm;
      q2 += v[a] * v[a];
    }
    f.H = (L.w * L.H + R.w * R.H) * norm;
    f.q2 = q2;
    T c2;
    if constexpr (POT) c2 = (f.H - T(.5) * q2 - (L.w * L.phi + R.w * R.phi) * norm) * (gamma - T(1));
    else c2 = (f.H - T(.5) * q2) * (gamma - T(1));
    if constexpr (kFast) {
      const T rs = __builtin_amdgcn_rsqf(c2);
      f.c = c2 * rs;
      f.invD = T(.5) * rs * rs;
    } else {
      f.c = tsqrt(c2);
      f.invD = T(.5) / (f.c * f.c);
    }
    // rotate: normal first, the normal slot swapped with x (EulerRoe.cl:101-110)
#pragma unroll
    for (int a = 0; a < DIM; ++a) f.v[a] = v[a];
    if (side != 0) {
      f.v[0] = v[side];
      f.v[side] = v[0];
    }
    return f;
  }

  static HYDRO_DEV void eig(const Face& f, T lam[N]) {
    lam[0] = f.v[0] - f.c;
#pragma unroll
    for (int a = 1; a <= DIM; ++a) lam[a] = f.v[0];
    lam[E] = f.v[0] + f.c;
  }

  // x in the rotated frame -> characteristic amplitudes (L * x)
  static HYDRO_DEV void left(const Face& f, const T x[N], T out[N], T gamma) {
    const T g1 = gamma - T(1);
    T s = T(.5) * f.q2 * x[0] - f.v[0] * x[1] + x[E];
#pragma unroll
    for (int a = 1; a < DIM; ++a) s -= f.v[a] * x[1 + a];
    const T A = g1 * f.invD * s;
    const T B = f.c * f.invD * (f.v[0] * x[0] - x[1]);
    out[0] = A + B;
    out[1] = x[0] - T(2) * A;
#pragma unroll
    for (int a = 1; a < DIM; ++a) out[1 + a] = x[1 + a] - f.v[a] * x[0];
    out[E] = A - B;
  }

  // characteristic -> conserved, rotated frame (R * y)
  static HYDRO_DEV void right(const Face& f, const T y[N], T out[N]) {
    const T s = y[0] + y[1] + y[E];
    const T d = y[E] - y[0];
    out[0] = s;
    out[1] = f.v[0] * s + f.c * d;
    T e = f.H * (y[0] + y[E]) + f.c * f.v[0] * d + T(.5) * f.q2 * y[1];
#pragma unroll
    for (int a = 1; a < DIM; ++a) {
      out[1 + a] = f.v[a] * s + y[1 + a];
      e += f.v[a] * y[1 + a];
    }
    out[E] = e;
  }

  // conserved vector in grid order -> rotated order for `side` (and back:
  // the permutation swaps momentum components 1 and 1+side, an involution)
  template <typename V>
  static HYDRO_DEV void rotate(V x[N], int side) {
    if (side != 0) {
      V t = x[1];
      x[1] = x[1 + side];
      x[1 + side] = t;
    }
  }
};

}  // namespace hydro
