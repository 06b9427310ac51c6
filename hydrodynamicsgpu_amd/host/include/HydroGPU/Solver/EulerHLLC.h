// HydroGPU/Solver/EulerHLLC.h -- (src/Solver/EulerHLLC.cpp:7-9, res/EulerHLLC.cl)
#pragma once
#include "HydroGPU/Solver/EulerHLL.h"

namespace HydroGPU {
namespace Solver {

struct EulerHLLC : public EulerHLL {
  typedef EulerHLL Super;
  using Super::Super;
  std::string name() const override { return "EulerHLLC"; }

 protected:
  int getSolverId() const override { return HYDRO_SOLVER_HLLC; }
};

}  // namespace Solver
}  // namespace HydroGPU
