// HydroGPU/Integrator/Integrator.h -- time integrator interface
// (include/HydroGPU/Integrator/Integrator.h:13-17).
//
// Two forms.  integrate(dt) advances the solver's state one step with the
// stages fused into the device stage kernels (the integrator names its
// tableau).  integrate(dt, callback) is the reference's form: the
// Runge-Kutta combination runs on the host side over device buffers
// (hydro_sim_buffer_*, hydro_sim_mult_add) and the callback fills a zeroed
// derivative buffer for the state buffer's current contents -- typically
// Solver::calcDeriv plus any source terms of the caller's own.
#pragma once
#include <functional>
#include <string>
#include <vector>

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
  // include/HydroGPU/Integrator/RungeKutta.h:59-125: u(0) = state; per stage
  // the callback fills a zeroed derivative buffer (a device buffer of the
  // simulation, see hydro_sim_buffer_create); the state becomes
  // sum_k alphas(i-1,k) u(k) + dt betas(i-1,k) L(u(k)), zero coefficients
  // skipped.  Frames the step with hydro_sim_init_step / finish_step, so it
  // also advances the simulated time.  Stage buffers are created on first use
  // and belong to the simulation.
  virtual void integrate(real dt, std::function<void(void* derivBuffer)> callback);
  // the integrator id of the C ABI (HYDRO_INTEGRATOR_*)
  virtual int integratorId() const = 0;
  virtual std::string name() const = 0;
  // number of derivative evaluations per step
  virtual int numStages() const = 0;

 protected:
  // stage states u(k) and derivatives L(u(k)) of the host-side form (null
  // where no later row reads them)
  std::vector<void*> stateBuffers, derivBuffers;
};

}  // namespace Integrator
}  // namespace HydroGPU
