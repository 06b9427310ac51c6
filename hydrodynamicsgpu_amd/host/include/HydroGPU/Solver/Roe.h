// HydroGPU/Solver/Roe.h -- Roe solvers: frozen per-step eigenbasis, limited
// characteristic flux (include/HydroGPU/Solver/Roe.h:13-40, src/Solver/Roe.cpp).
#pragma once
#include "HydroGPU/Solver/FiniteVolumeSolver.h"

namespace HydroGPU {
namespace Solver {

struct Roe : public FiniteVolumeSolver {
  typedef FiniteVolumeSolver Super;
  using Super::Super;

 protected:
  int getSolverId() const override { return HYDRO_SOLVER_ROE; }
};

}  // namespace Solver
}  // namespace HydroGPU
