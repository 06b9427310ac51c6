// HydroGPU/Solver/EulerBurgers.h -- (src/Solver/EulerBurgers.cpp:8-100,
// res/EulerBurgers.cl): the split Euler step -- advection of every state
// with the interface velocity, then momentum and work diffusion from the
// pressure, each pass integrated by the configured integrator with a
// boundary after it.  The three passes run fused on the device behind
// step(dt) (hydro_sim_step); the callback form of integrate() is not offered
// for this solver.
#pragma once
#include "HydroGPU/Solver/FiniteVolumeSolver.h"

namespace HydroGPU {
namespace Solver {

struct EulerBurgers : public FiniteVolumeSolver {
  typedef FiniteVolumeSolver Super;
  using Super::Super;
  std::string name() const override { return "EulerBurgers"; }

 protected:
  void createEquation() override;
  int getSolverId() const override { return HYDRO_SOLVER_BURGERS; }
};

}  // namespace Solver
}  // namespace HydroGPU
