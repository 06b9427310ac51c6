// run1d.hip -- whole updates of a small 1D Euler grid in one launch.
//
// A 1D grid of about a thousand cells gives each kernel of a step (the
// boundary, the CFL march, the dt tail and S fused stages) only a handful of
// wavefronts, so a step costs S + 3 kernel launches of a few microseconds
// each and almost no work.  Here one workgroup of 16 wavefronts runs every
// phase of every step itself -- the same device functions the
// launch-per-kernel path runs (boundary_cell, dt_march_wave, finalize_dt,
// stage2d_wave) on the same buffers -- with a workgroup barrier between the
// phases, so results are bit-identical and a step costs a few barriers.
// Waves take the strips of a phase round-robin.  Writes of one phase are
// visible to the next after the barrier: every wave of a workgroup runs on
// one CU and shares its write-through L1.  The clock (t, step count, health)
// stays in thread 0's registers and dt reaches the stages through LDS; both
// are written back to device memory when the launch ends.
#include "aux_device.hpp"
#include "stage2d_impl.hpp"

namespace hydro {

template <class Pol, int SOLVER, typename T, typename S, bool KAHAN>
__global__ void __launch_bounds__(kRun1dThreads) run1d_kernel(Run1dArgs<T, S> r) {
  constexpr int NW = kRun1dThreads / 64;
  __shared__ double wmin[NW];
  __shared__ int wbad[NW];
  __shared__ int advanced;
  __shared__ double sdt;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  auto boundary = [&](int p) {
    if (threadIdx.x < 4) boundary_cell<S>(r.b, r.U[p], KAHAN ? r.lo[p] : nullptr, (long long)threadIdx.x);
  };
  // the clock, live in thread 0
  double t = 0.0, dt = 0.0;
  long long steps = 0;
  unsigned health = 0;
  if (threadIdx.x == 0) {
    t = *r.f.t;
    steps = *r.f.steps;
    health = *r.f.health;
    dt = *r.f.dt;
  }
  int par = 0;
#pragma unroll 1
  for (int n = 0; n < r.nsteps; ++n) {
    // Solver::update: boundary, then the CFL minimum and the dt tail
    boundary(par);
    __syncthreads();
    double m = INFINITY;
    bool bad = false;
    if (r.march) {
#pragma unroll 1
      for (int w = wave; w < r.dt.nstrips; w += NW) dt_merge(dt_march_wave<S, 1, SOLVER>(r.dt, r.U[par], w, 0), m, bad);
    }
    if (lane == 0) {
      wmin[wave] = m;
      wbad[wave] = bad ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double mm = INFINITY;
      bool nan = false;
      for (int i = 0; i < NW; ++i) {
        nan |= wbad[i] != 0;
        mm = fmin(mm, wmin[i]);
      }
      advanced = finalize_step(r.f, mm, nan, t, steps, health, dt) ? 1 : 0;
      sdt = dt;
    }
    __syncthreads();
    if (!advanced) {
      // t_end reached (clipped runs): the remaining updates are no-ops.  The
      // launch-per-kernel path would leave both buffers holding this state
      // with its ghost cells just written; copy it so either parity reads it.
      const long long cells = (long long)r.b.plane * r.b.nstates;
      S* dst = r.U[par ^ 1];
      const S* src = r.U[par];
      for (long long i = threadIdx.x; i < cells; i += kRun1dThreads) dst[i] = src[i];
      if (KAHAN)
        for (long long i = threadIdx.x; i < cells; i += kRun1dThreads) r.lo[par ^ 1][i] = r.lo[par][i];
      break;
    }
    // the RK stages
#pragma unroll 1
    for (int k = 0; k < r.nstages; ++k) {
      StageArgs<T, S> a = r.st[k];
      if (par) {
        S* u0 = const_cast<S*>(a.u0);
        a.u0 = a.unew;
        a.unew = u0;
        float* l0 = const_cast<float*>(a.u0lo);
        a.u0lo = a.unewlo;
        a.unewlo = l0;
      }
#pragma unroll 1
      for (int w = wave; w < a.nstrips; w += NW) stage2d_wave<Pol, 1, T, S, KAHAN, CM_GENERIC>(a, w, 0, sdt);
      __syncthreads();
    }
    par ^= 1;
    if (r.post_boundary) {
      boundary(par);
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    *r.f.t = t;
    *r.f.steps = steps;
    *r.f.health = health;
    *r.f.dt = dt;
  }
}

template <class Pol, int SOLVER, typename T, typename S, bool KAHAN>
static void launch_run1d_pol(const Run1dArgs<T, S>& a, hipStream_t st) {
  hipLaunchKernelGGL((run1d_kernel<Pol, SOLVER, T, S, KAHAN>), dim3(1), dim3(kRun1dThreads), 0, st, a);
}

template <typename T, typename S, bool KAHAN>
static void launch_run1d_kind(int kind, int lim, const Run1dArgs<T, S>& a, hipStream_t st) {
  if (kind == 1) {
    launch_run1d_pol<HLLPolicy<1, T, 1>, 1, T, S, KAHAN>(a, st);
  } else if (kind == 2) {
    launch_run1d_pol<HLLPolicy<1, T, 2>, 2, T, S, KAHAN>(a, st);
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch_run1d_pol<RoePolicy<1, T, LIM_SUPERBEE>, 0, T, S, KAHAN>(a, st); break;
    case LIM_MONOTIZEDCENTRAL: launch_run1d_pol<RoePolicy<1, T, LIM_MONOTIZEDCENTRAL>, 0, T, S, KAHAN>(a, st); break;
    case LIM_MINMOD: launch_run1d_pol<RoePolicy<1, T, LIM_MINMOD>, 0, T, S, KAHAN>(a, st); break;
    default: launch_run1d_pol<RoePolicy<1, T, LIM_RUNTIME>, 0, T, S, KAHAN>(a, st); break;
    }
  }
}

void launch_run1d_f64(int kind, int lim, const Run1dArgs<double, double>& a, hipStream_t st) {
  launch_run1d_kind<double, double, false>(kind, lim, a, st);
}
void launch_run1d_f32(int kind, int lim, bool kahan, const Run1dArgs<float, float>& a, hipStream_t st) {
  if (kahan) launch_run1d_kind<float, float, true>(kind, lim, a, st);
  else launch_run1d_kind<float, float, false>(kind, lim, a, st);
}

}  // namespace hydro
