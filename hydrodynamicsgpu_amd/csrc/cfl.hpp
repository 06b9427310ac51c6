// cfl.hpp -- the CFL timestep pass of one grid (calcCellTimestep +
// findMinTimestep, res/Roe.cl:4-65, res/EulerHLL.cl:142-175,
// res/Common.cl:9-40), shared by the simulations of the C ABI and the
// allocation-free entry point: which kernel runs, its decomposition and how
// many per-wave (or per-block) minima it writes.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/hydro_capi.h"
#include "launch.hpp"

namespace hydro {
namespace host {

struct CflGrid {
  const hydro_sim_desc* d = nullptr;
  int dim = 1;
  int sz[3] = {1, 1, 1};
  long long plane = 0;   // elements between state planes
  double kdx[3] = {1, 1, 1};
  double cfl = 0.5;
  bool mixed = false;
  int dt_rows = 0;       // 2D strip march: rows per wave (0: chosen from the grid size)
  int dt_blocks = 1024;  // blocks of the generic CFL kernel
  const void* pot = nullptr;  // potential plane the speeds read (Euler Roe, HLL, EulerBurgers), or null
  const void* solid = nullptr;  // solid mask plane (Euler Roe, EulerBurgers), or null; Roe then reads pot too
};

// sizes, kernel dx (src/HydroGPUApp.cpp:224-226 and the 6-digit rendering of
// src/Solver/Solver.cpp:272-280) and cfl as a float (HydroGPUApp.h:84)
CflGrid cfl_grid(const hydro_sim_desc& d, long long plane);

// the 1D/2D strip march's decomposition, minima into `minima`
DtMarchArgs cfl_march_args(const CflGrid& g, double* minima);

// the number of minima the pass writes (at most max(dt_blocks, kDtMaxWaves))
long long cfl_minima(const CflGrid& g);

// CFL kernel + dt tail on `st`: U are the state planes (float for the
// FP32-mixed form), prim the SRHD primitive planes (FP64 when prim_f64,
// else FP32: the FP32-mixed simulation's storage); f.blockmin and
// f.nblocks are filled here
int enqueue_cfl(const CflGrid& g, const void* U, const void* prim, bool prim_f64, double* minima, FinalizeArgs f,
                hipStream_t st);

}  // namespace host
}  // namespace hydro
