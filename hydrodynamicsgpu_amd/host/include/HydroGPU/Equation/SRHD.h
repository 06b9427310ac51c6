// HydroGPU/Equation/SRHD.h -- SRHD equation plug-in
// (include/HydroGPU/Equation/SRHD.h, src/Equation/SRHD.cpp:11-98).
#pragma once
#include "HydroGPU/Equation/Equation.h"
#include "hydro_capi.h"

namespace HydroGPU {
namespace Equation {

struct SRHD : public Equation {
  typedef Equation Super;

 public:
  explicit SRHD(HydroGPUApp* app_);
  int stateGetBoundaryKernelForBoundaryMethod(int dim, int state, int minmax) override;
  void readStateCell(real* state, const real* source) override;
  int numReadStateChannels() override;
  std::string name() const override { return "SRHD"; }
  int equationId() const override { return HYDRO_EQN_SRHD; }
};

}  // namespace Equation
}  // namespace HydroGPU
