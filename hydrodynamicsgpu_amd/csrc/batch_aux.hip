// batch_aux.hip -- per-step kernels of the batched launch: the composed
// boundary of every member, the CFL strip march over the members' tasks and
// the per-member findMinTimestep tail.
#include "aux_device.hpp"
#include "batch_args.hpp"

namespace hydro {

// member = blockIdx.y; one thread per ghost cell of that member
template <typename S>
__global__ void __launch_bounds__(256) batch_boundary_kernel(BatchBoundaryArgs b, S* u, float* lo) {
  const BatchMember& m = b.members[blockIdx.y];
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m.ghosts) return;
  BoundaryArgs a{};
  a.dim = m.dim;
  a.sx = m.sx;
  a.sy = m.sy;
  a.sz = 1;
  a.plane = m.plane;
  a.nstates = m.nstates;
  for (int ax = 0; ax < 3; ++ax)
    for (int mm = 0; mm < 2; ++mm) {
      a.method[ax][mm] = ax < 2 ? m.method[ax][mm] : (signed char)-1;
      a.negmask[ax][mm] = ax < 2 ? m.negmask[ax][mm] : 0u;
    }
  boundary_cell<S>(a, u + m.off, lo ? lo + m.off : nullptr, g);
}

void launch_batch_boundary_f64(const BatchBoundaryArgs& b, int count, double* u, hipStream_t st) {
  hipLaunchKernelGGL((batch_boundary_kernel<double>), dim3((b.max_ghosts + 255) / 256, count), dim3(256), 0, st, b,
                     u, (float*)nullptr);
}
void launch_batch_boundary_f32(const BatchBoundaryArgs& b, int count, float* u, float* lo, hipStream_t st) {
  hipLaunchKernelGGL((batch_boundary_kernel<float>), dim3((b.max_ghosts + 255) / 256, count), dim3(256), 0, st, b, u,
                     lo);
}

// one wave per CFL task; the task's minimum (NaN if a speed was not finite)
template <typename S, int SOLVER>
__global__ void __launch_bounds__(256) batch_dt_kernel(BatchDtArgs b, const S* u) {
  const int t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (t >= b.ntasks) return;
  const BatchTask task = b.tasks[t];
  const BatchMember& m = b.members[task.member];
  DtMarchArgs a{};
  a.dim = m.dim;
  a.sx = m.sx;
  a.sy = m.sy;
  a.plane = m.plane;
  a.kdx[0] = m.kdx[0];
  a.kdx[1] = m.kdx[1];
  a.gamma = b.gamma;
  a.solver = SOLVER;
  a.last_x = m.last_x;
  a.last_y = m.last_y;
  a.rows = kDtRows;
  const S* u0 = u + m.off;
  double mn;
  if (task.dim == 2) {
    mn = dt_march_wave<S, 2, SOLVER>(a, u0, task.a, task.b);
  } else {
    mn = INFINITY;
    bool bad = false;
#pragma unroll 1
    for (int st = task.a; st < task.a + task.b; ++st) dt_merge(dt_march_wave<S, 1, SOLVER>(a, u0, st, 0), mn, bad);
    mn = dt_flag(mn, bad);
  }
  if ((threadIdx.x & 63) == 0) b.taskmin[t] = mn;
}

template <typename S>
static void batch_dt(const BatchDtArgs& a, const S* u, hipStream_t st) {
  dim3 g((a.ntasks + 3) / 4), bl(256);
  if (a.solver == 0) hipLaunchKernelGGL((batch_dt_kernel<S, 0>), g, bl, 0, st, a, u);
  else if (a.solver == 1) hipLaunchKernelGGL((batch_dt_kernel<S, 1>), g, bl, 0, st, a, u);
  else hipLaunchKernelGGL((batch_dt_kernel<S, 2>), g, bl, 0, st, a, u);
}
void launch_batch_dt_f64(const BatchDtArgs& a, const double* u, hipStream_t st) { batch_dt<double>(a, u, st); }
void launch_batch_dt_f32(const BatchDtArgs& a, const float* u, hipStream_t st) { batch_dt<float>(a, u, st); }

// block per member: the minimum over the member's task minima, then
// findMinTimestep's tail on that member's scalars; block 0 also zeroes the
// stage task counters for this step
__global__ void __launch_bounds__(256) batch_finalize_kernel(BatchFinalizeArgs f) {
  __shared__ double sh[4];
  __shared__ int anynan;
  const int mi = blockIdx.x;
  const BatchMember& m = f.members[mi];
  if (threadIdx.x == 0) anynan = 0;
  __syncthreads();
  double v = INFINITY;
  bool nan = false;
  const int ntasks = f.use_fixed ? 0 : m.dt_ntasks;  // no CFL launch with a fixed dt
  for (int i = threadIdx.x; i < ntasks; i += blockDim.x) {
    dt_merge(f.taskmin[m.dt_task0 + i], v, nan);
  }
  if (nan) atomicOr(&anynan, 1);
  v = block_min(v, sh);
  if (threadIdx.x == 0) {
    double* sc = f.scalars + (long long)mi * kBatchDtStride;
    FinalizeArgs a{};
    a.cfl = m.cfl;
    a.use_fixed = f.use_fixed;
    a.fixed_dt = f.fixed_dt;
    a.clip = f.clip;
    a.t_end = m.t_end;
    a.dt = sc;
    a.t = sc + 1;
    a.steps = reinterpret_cast<long long*>(sc + 2);
    a.health = reinterpret_cast<unsigned*>(sc + 3);
    finalize_dt(a, v, anynan != 0);
    if (mi == 0)
      for (int k = 0; k < f.ncounters; ++k) f.counters[k] = 0u;
  }
}

void launch_batch_finalize(const BatchFinalizeArgs& f, int count, hipStream_t st) {
  hipLaunchKernelGGL(batch_finalize_kernel, dim3(count), dim3(256), 0, st, f);
}

}  // namespace hydro
