// stage_common.hpp -- the Shu-Osher combination shared by the fused stage
// kernels (include/HydroGPU/Integrator/RungeKutta.h:84-99 in increment form).
#pragma once
#include "device_common.hpp"
#include "stage_args.hpp"

namespace hydro {

constexpr int kStripOut = 60;  // output columns per wave (64 lanes - 2*2 halo)

template <typename T, typename S>
HYDRO_DEV PhysParams<T> make_params(const StageArgs<T, S>& a) {
  PhysParams<T> p;
  p.gamma = T(a.gamma);
#pragma unroll
  for (int i = 0; i < 5; ++i) p.mx[i] = T(a.mx[i]);
  return p;
}

template <typename T, typename S, bool KAHAN, int N>
struct StageCombine {
  const StageArgs<T, S>& a;
  const T bdt_self, a_self;
  const double dtd;

  HYDRO_DEV StageCombine(const StageArgs<T, S>& args, double dt)
      : a(args), bdt_self(T(args.b_self * dt)), a_self(T(args.a_self)), dtd(dt) {}

  // u0 (storage precision) and the stage state u = u0 + d_k (arithmetic precision)
  HYDRO_DEV void load_state(long long cell, S q0[N], T u[N]) const {
    const long long pl = a.plane;
#pragma unroll
    for (int s = 0; s < N; ++s) q0[s] = a.u0[s * pl + cell];
#pragma unroll
    for (int s = 0; s < N; ++s) u[s] = T(q0[s]) + (a.dk ? a.dk[s * pl + cell] : T(0));
  }

  // d_{k+1} = P_k + alpha[k][k] d_k + dt beta[k][k] L_k, the partial sums of
  // later rows, and on the last stage u0 + d_S (compensated when KAHAN).
  HYDRO_DEV void finish(long long cell, const T deriv[N]) const {
    const long long pl = a.plane;
    T dk[N];
#pragma unroll
    for (int s = 0; s < N; ++s) dk[s] = a.dk ? a.dk[s * pl + cell] : T(0);
    T acc[N];
#pragma unroll
    for (int s = 0; s < N; ++s) {
      T v = (a.pk >= 0) ? a.part[a.pk][s * pl + cell] : T(0);
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nrep) v += T(a.rc[j]) * a.rep[j][s * pl + cell];
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
          // compensated update of (hi, lo) by acc (Fast2Sum: |hi| >= |lo + acc|)
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
  }
};

}  // namespace hydro
