// HydroGPU/Equation/Maxwell.h -- Maxwell equation plug-in
// (include/HydroGPU/Equation/Maxwell.h, src/Equation/Maxwell.cpp:18-85).
#pragma once
#include "HydroGPU/Equation/Equation.h"
#include "hydro_capi.h"

namespace HydroGPU {
namespace Equation {

struct Maxwell : public Equation {
  typedef Equation Super;

 public:
  explicit Maxwell(HydroGPUApp* app_);
  int stateGetBoundaryKernelForBoundaryMethod(int dim, int state, int minmax) override;
  std::string name() const override { return "Maxwell"; }
  int equationId() const override { return HYDRO_EQN_MAXWELL; }
};

}  // namespace Equation
}  // namespace HydroGPU
