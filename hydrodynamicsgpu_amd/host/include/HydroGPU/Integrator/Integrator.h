// HydroGPU/Integrator/Integrator.h -- time integrator interface
// (include/HydroGPU/Integrator/Integrator.h:13-17).
//
// The reference's integrate(dt, callback) runs the stages on the host and
// calls back into the solver for each derivative.  Here the stages are fused
// into the device stage kernels: an integrator names its tableau, and
// integrate(dt) advances the solver's state one step with it.  There is no
// per-stage callback (derivative evaluation happens inside the kernels).
#pragma once
#include <string>

#include "HydroGPU/Shared/Common.h"

namespace HydroGPU {
namespace Solver {
struct Solver;
}
namespace Integrator {

struct Integrator {
  HydroGPU::Solver::Solver* solver;
  explicit Integrator(HydroGPU::Solver::Solver* solver_) : solver(solver_) {}
  virtual ~Integrator() {}
  // one step of size dt: every stage, then the solver's per-step tail
  virtual void integrate(real dt);
  // the integrator id of the C ABI (HYDRO_INTEGRATOR_*)
  virtual int integratorId() const = 0;
  virtual std::string name() const = 0;
  // number of derivative evaluations per step
  virtual int numStages() const = 0;
};

}  // namespace Integrator
}  // namespace HydroGPU
