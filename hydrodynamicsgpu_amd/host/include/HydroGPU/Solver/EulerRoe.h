// HydroGPU/Solver/EulerRoe.h -- (src/Solver/EulerRoe.cpp:8-34)
#pragma once
#include "HydroGPU/Solver/Roe.h"

namespace HydroGPU {
namespace Solver {

struct EulerRoe : public Roe {
  typedef Roe Super;
  using Super::Super;
  std::string name() const override { return "EulerRoe"; }

 protected:
  void createEquation() override;
};

}  // namespace Solver
}  // namespace HydroGPU
