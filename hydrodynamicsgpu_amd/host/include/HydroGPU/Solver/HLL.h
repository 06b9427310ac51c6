// HydroGPU/Solver/HLL.h -- HLL-family solvers: frozen per-step wave speeds,
// first-order flux (include/HydroGPU/Solver/HLL.h:9-28, src/Solver/HLL.cpp).
#pragma once
#include "HydroGPU/Solver/FiniteVolumeSolver.h"

namespace HydroGPU {
namespace Solver {

struct HLL : public FiniteVolumeSolver {
  typedef FiniteVolumeSolver Super;
  using Super::Super;
};

}  // namespace Solver
}  // namespace HydroGPU
