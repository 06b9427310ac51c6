// HydroGPU/Solver/EulerHLL.h -- (src/Solver/EulerHLL.cpp:8-28)
#pragma once
#include "HydroGPU/Solver/HLL.h"

namespace HydroGPU {
namespace Solver {

struct EulerHLL : public HLL {
  typedef HLL Super;
  using Super::Super;
  std::string name() const override { return "EulerHLL"; }

 protected:
  void createEquation() override;
  int getSolverId() const override { return HYDRO_SOLVER_HLL; }
};

}  // namespace Solver
}  // namespace HydroGPU
