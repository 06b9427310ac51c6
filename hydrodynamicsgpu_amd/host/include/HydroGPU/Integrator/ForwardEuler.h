// HydroGPU/Integrator/ForwardEuler.h -- u += dt L(u)
// (include/HydroGPU/Integrator/ForwardEuler.h, src/Integrator/ForwardEuler.cpp:8-32).
#pragma once
#include "HydroGPU/Integrator/Integrator.h"
#include "hydro_capi.h"

namespace HydroGPU {
namespace Integrator {

struct ForwardEuler : public Integrator {
  typedef Integrator Super;
  using Super::Super;
  int integratorId() const override { return HYDRO_INTEGRATOR_FORWARDEULER; }
  std::string name() const override { return "ForwardEuler"; }
  int numStages() const override { return 1; }
};

}  // namespace Integrator
}  // namespace HydroGPU
