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
  using Super::integrate;
  // src/Integrator/ForwardEuler.cpp:20-32: zero the derivative, callback,
  // state += dt * deriv
  void integrate(real dt, std::function<void(void* derivBuffer)> callback) override;
};

}  // namespace Integrator
}  // namespace HydroGPU
