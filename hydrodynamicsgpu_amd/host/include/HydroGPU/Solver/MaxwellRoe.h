// HydroGPU/Solver/MaxwellRoe.h -- (src/Solver/MaxwellRoe.cpp:8-46)
#pragma once
#include "HydroGPU/Solver/Roe.h"

namespace HydroGPU {
namespace Solver {

struct MaxwellRoe : public Roe {
  typedef Roe Super;
  using Super::Super;
  std::string name() const override { return "MaxwellRoe"; }

 protected:
  void createEquation() override;
};

}  // namespace Solver
}  // namespace HydroGPU
