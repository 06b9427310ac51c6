// phys_euler.hpp -- Euler equations for the Roe and HLL/HLLC face solvers.
//
// State order (src/Equation/Euler.cpp:36-40): rho, rho*v[dim], E.
// Roe face data follow res/EulerRoe.cl:18-259: sqrt(rho)-weighted averages of
// velocity and total enthalpy, c^2 = (gamma-1)(H - v^2/2 - phi), the normal
// rotated into x.  The eigenvector matrices are not materialised: L*x and R*y
// are applied in closed form from (v, H, c), which is algebraically the dense
// product res/RoeEigenfieldLinear.cl:16-74 performs with the stored matrices.
#pragma once
#include "device_common.hpp"

namespace hydro {

template <int DIM, typename T>
struct EulerRoe {
  static constexpr int N = DIM + 2;
  static constexpr int E = DIM + 1;

  // frozen per-cell data from the start-of-step state
  struct Cell {
    T w;        // sqrt(rho)
    T v[DIM];   // velocity
    T H;        // total specific enthalpy
  };

  // frozen per-face data, normal rotated into component 0 of v
  struct Face {
    T v[DIM];   // v[0] normal, v[1..] tangential (rotated order)
    T H, c, q2; // enthalpy, sound speed, |v|^2
  };

  template <typename S>
  static HYDRO_DEV Cell cell(const S q[N], T gamma) {
    Cell c;
    const T rho = T(q[0]);
    const T inv = trcp(rho);
    T vsq = T(0);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      c.v[a] = tdiv(T(q[1 + a]), rho);
      vsq += c.v[a] * c.v[a];
    }
    const T eT = T(q[E]) * inv;
    const T eInt = eT - T(.5) * vsq;
    const T p = (gamma - T(1)) * rho * eInt;
    c.H = eT + p * inv;
    c.w = tsqrt(rho);
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
    f.c = tsqrt((f.H - T(.5) * q2) * (gamma - T(1)));
    // rotate: normal first, then the others in their original order with the
    // normal slot swapped (EulerRoe.cl:101-110 swaps x<->side)
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
    const T invD = T(.5) * trcp(f.c * f.c);
    T s = T(.5) * f.q2 * x[0] - f.v[0] * x[1] + x[E];
#pragma unroll
    for (int a = 1; a < DIM; ++a) s -= f.v[a] * x[1 + a];
    const T A = g1 * invD * s;
    const T B = f.c * invD * (f.v[0] * x[0] - x[1]);
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
