// batch_args.hpp -- descriptors of the batched launch (many independent
// grids advanced together), shared by the host driver and the kernels.
#pragma once
#include <stdint.h>

#include "launch.hpp"
#include "stage_args.hpp"

namespace hydro {

// One member grid.  Every arena (states, increments, partial sums) holds the
// members back to back with the same layout, so a member's planes start at
// element `off` of any arena.
struct BatchMember {
  int dim, sx, sy, nstates;
  long long off;        // first element of the member in every arena
  long long plane;      // elements between the member's state planes
  double inv_dx[2];     // 1 / kernel dx, stage kernels
  double kdx[2];        // kernel dx, CFL
  double cfl, t_end;
  int last_x, last_y;   // last cell with a CFL dt per axis
  int dt_task0, dt_ntasks;  // the member's CFL tasks
  int ghosts;           // ghost cells (boundary threads)
  int pad;
  signed char method[2][2];  // boundary method per axis side (launch.hpp BoundaryArgs)
  unsigned negmask[2][2];
};

// A unit of work for one wavefront.  2D: strip a, row chunk b.  1D: strips
// a .. a+b-1 of the single row.
struct BatchTask {
  int member, dim, a, b;
};

constexpr int kBatchDtStride = 4;  // doubles per member in the scalar array

// fused RK stage over every member (persistent: waves pull tasks)
template <typename T, typename S>
struct BatchStageArgs {
  StageArgs<T, S> tmpl;        // arena base pointers and the stage plan
  const BatchMember* members;
  const BatchTask* tasks;
  int ntasks;
  unsigned* counter;           // task counter of this stage (reset every step)
  const double* dts;           // member m's dt at dts[m * kBatchDtStride]
};

// CFL tasks and the per-member tail
struct BatchDtArgs {
  const BatchMember* members;
  const BatchTask* tasks;
  int ntasks;
  double gamma;
  int solver;
  double* taskmin;
};

struct BatchFinalizeArgs {
  const BatchMember* members;
  const double* taskmin;
  double* scalars;     // per member: dt, t, steps (as double bits), health
  int use_fixed, clip;
  double fixed_dt;
  unsigned* counters;  // stage task counters, zeroed here
  int ncounters;
};

struct BatchBoundaryArgs {
  const BatchMember* members;
  int max_ghosts;
};

void launch_batch_boundary_f64(const BatchBoundaryArgs& b, int count, double* u, hipStream_t st);
void launch_batch_boundary_f32(const BatchBoundaryArgs& b, int count, float* u, float* lo, hipStream_t st);
void launch_batch_dt_f64(const BatchDtArgs& a, const double* u, hipStream_t st);
void launch_batch_dt_f32(const BatchDtArgs& a, const float* u, hipStream_t st);
void launch_batch_finalize(const BatchFinalizeArgs& f, int count, hipStream_t st);
// solver: 0 Roe, 1 HLL, 2 HLLC; mode: CombineMode; blocks of 4 waves
void launch_batch_stage_f64(int solver, int lim, int mode, const BatchStageArgs<double, double>& a, int blocks,
                            hipStream_t st);
void launch_batch_stage_f32(int solver, int lim, int mode, const BatchStageArgs<float, float>& a, int blocks,
                            hipStream_t st);
void launch_batch_stage_f32k(int solver, int lim, int mode, const BatchStageArgs<float, float>& a, int blocks,
                             hipStream_t st);

}  // namespace hydro
