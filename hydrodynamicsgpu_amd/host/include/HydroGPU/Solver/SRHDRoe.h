// HydroGPU/Solver/SRHDRoe.h -- (src/Solver/SRHDRoe.cpp:9-100)
#pragma once
#include "HydroGPU/Solver/Roe.h"

namespace HydroGPU {
namespace Solver {

struct SRHDRoe : public Roe {
  typedef Roe Super;
  using Super::Super;
  std::string name() const override { return "SRHDRoe"; }
  // primitives (rho, v, eInt) in the state layout (src/Solver/SRHDRoe.cpp:9-13)
  std::vector<real> getPrimitives();

 protected:
  void createEquation() override;
};

}  // namespace Solver
}  // namespace HydroGPU
