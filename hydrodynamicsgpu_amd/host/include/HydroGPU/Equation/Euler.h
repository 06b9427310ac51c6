// HydroGPU/Equation/Euler.h -- Euler equation plug-in
// (include/HydroGPU/Equation/Euler.h, src/Equation/Euler.cpp:12-93).
#pragma once
#include "HydroGPU/Equation/Equation.h"
#include "hydro_capi.h"

namespace HydroGPU {
namespace Equation {

struct Euler : public Equation {
  typedef Equation Super;

 protected:
  int dim;

 public:
  Euler(HydroGPUApp* app_, int dim_ = -1);
  int stateGetBoundaryKernelForBoundaryMethod(int dim, int state, int minmax) override;
  void readStateCell(real* state, const real* source) override;
  int numReadStateChannels() override;
  std::string name() const override { return "Euler"; }
  int equationId() const override { return HYDRO_EQN_EULER; }
};

}  // namespace Equation
}  // namespace HydroGPU
