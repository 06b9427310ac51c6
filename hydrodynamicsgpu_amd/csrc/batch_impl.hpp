// batch_impl.hpp -- the fused RK stage over a batch of independent grids.
//
// One persistent launch per stage covers every member: a fixed grid of
// wavefronts pulls tasks (a 2D member's strip x row chunk, or a run of a 1D
// member's strips) from a device counter until the list is empty, and runs
// the single-grid stage body (stage2d_wave) on the member's slice of the
// arenas with the member's own dt.  Tasks are listed largest first, so the
// dynamic pull balances uneven members.
#pragma once
#include "batch_args.hpp"
#include "stage2d_impl.hpp"

namespace hydro {

template <typename T, typename S>
HYDRO_DEV void shift_args(StageArgs<T, S>& a, long long off) {
  auto sh = [off](auto* p) { return p ? p + off : p; };
  a.u0 = sh(a.u0);
  a.prim = sh(a.prim);
  a.u0lo = sh(a.u0lo);
  a.dk = sh(a.dk);
  a.dnext = sh(a.dnext);
  a.unew = sh(a.unew);
  a.unewlo = sh(a.unewlo);
  a.pkp = sh(a.pkp);
#pragma unroll
  for (int j = 0; j < kMaxPartials; ++j) {
    a.fp[j] = sh(a.fp[j]);
    a.rep[j] = sh(a.rep[j]);
  }
}

template <class Pol1, class Pol2, typename T, typename S, bool KAHAN, int MODE>
__global__ void __launch_bounds__(256) batch_stage_kernel(BatchStageArgs<T, S> b) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned t = 0;
    if (lane == 0) t = atomicAdd(b.counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);  // lane 0's ticket, wave-uniform
    if (t >= (unsigned)b.ntasks) break;
    const BatchTask task = b.tasks[t];
    const BatchMember& m = b.members[task.member];
    StageArgs<T, S> a = b.tmpl;
    shift_args(a, m.off);
    a.sx = m.sx;
    a.sy = m.sy;
    a.sz = 1;
    a.plane = m.plane;
    a.inv_dx[0] = m.inv_dx[0];
    a.inv_dx[1] = m.inv_dx[1];
    a.dt = b.dts + (long long)task.member * kBatchDtStride;
    if (task.dim == 2) {
      stage2d_wave<Pol2, 2, T, S, KAHAN, MODE>(a, task.a, task.b, *a.dt);
    } else {
#pragma unroll 1
      for (int st = task.a; st < task.a + task.b; ++st) stage2d_wave<Pol1, 1, T, S, KAHAN, MODE>(a, st, 0, *a.dt);
    }
  }
}

template <template <int, typename, int> class PolT, int X, typename T, typename S, bool KAHAN, int MODE>
static void launch_batch_mode(const BatchStageArgs<T, S>& a, int blocks, hipStream_t st) {
  hipLaunchKernelGGL((batch_stage_kernel<PolT<1, T, X>, PolT<2, T, X>, T, S, KAHAN, MODE>), dim3(blocks), dim3(256),
                     0, st, a);
}

// fixed-function combination modes for the Roe limiters with compile-time
// forms and for HLL/HLLC; the runtime-limiter Roe uses the generic form
template <template <int, typename, int> class PolT, int X, typename T, typename S, bool KAHAN>
static void launch_batch_pol(int mode, const BatchStageArgs<T, S>& a, int blocks, hipStream_t st) {
  switch (mode) {
  case CM_FIRST: launch_batch_mode<PolT, X, T, S, KAHAN, CM_FIRST>(a, blocks, st); return;
  case CM_MID: launch_batch_mode<PolT, X, T, S, KAHAN, CM_MID>(a, blocks, st); return;
  case CM_FINAL: launch_batch_mode<PolT, X, T, S, KAHAN, CM_FINAL>(a, blocks, st); return;
  case CM_SINGLE: launch_batch_mode<PolT, X, T, S, KAHAN, CM_SINGLE>(a, blocks, st); return;
  default: launch_batch_mode<PolT, X, T, S, KAHAN, CM_GENERIC>(a, blocks, st); return;
  }
}

template <typename T, typename S, bool KAHAN>
static void launch_batch_solver(int solver, int lim, int mode, const BatchStageArgs<T, S>& a, int blocks,
                                hipStream_t st) {
  if (solver == 1) {
    launch_batch_pol<HLLPolicy, 1, T, S, KAHAN>(mode, a, blocks, st);
  } else if (solver == 2) {
    launch_batch_pol<HLLPolicy, 2, T, S, KAHAN>(mode, a, blocks, st);
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch_batch_pol<RoePolicy, LIM_SUPERBEE, T, S, KAHAN>(mode, a, blocks, st); break;
    case LIM_MONOTIZEDCENTRAL: launch_batch_pol<RoePolicy, LIM_MONOTIZEDCENTRAL, T, S, KAHAN>(mode, a, blocks, st); break;
    case LIM_MINMOD: launch_batch_pol<RoePolicy, LIM_MINMOD, T, S, KAHAN>(mode, a, blocks, st); break;
    default: launch_batch_mode<RoePolicy, LIM_RUNTIME, T, S, KAHAN, CM_GENERIC>(a, blocks, st); break;
    }
  }
}

}  // namespace hydro
