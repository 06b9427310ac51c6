// HydroGPU/Solver/FiniteVolumeSolver.h -- flux-divergence solvers
// (include/HydroGPU/Solver/FiniteVolumeSolver.h, src/Solver/FiniteVolumeSolver.cpp:17-26).
#pragma once
#include "HydroGPU/Solver/Solver.h"

namespace HydroGPU {
namespace Solver {

struct FiniteVolumeSolver : public Solver {
  typedef Solver Super;
  using Super::Super;
};

}  // namespace Solver
}  // namespace HydroGPU
