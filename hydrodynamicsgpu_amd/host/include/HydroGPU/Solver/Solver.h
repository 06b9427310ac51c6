// HydroGPU/Solver/Solver.h -- base of the finite-volume solvers
// (include/HydroGPU/Solver/Solver.h:18-158, src/Solver/Solver.cpp).
//
// Same virtual interface as the reference: init, resetState, boundary,
// initStep, calcTimestep, step, update.  The bodies call the C ABI of
// libhydro_hip.so (include/hydro_capi.h) instead of binding OpenCL kernels;
// device buffers live behind the opaque hydro_sim handle.  Errors throw
// HydroGPU::Exception, as the reference throws Common::Exception.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "HydroGPU/Exception.h"
#include "HydroGPU/Integrator/Integrator.h"
#include "HydroGPU/Shared/Common.h"
#include "HydroGPU/Solver/ISolver.h"
#include "hydro_capi.h"

namespace HydroGPU {
struct HydroGPUApp;
namespace Solver {

struct Solver : public ISolver {
  typedef ISolver Super;

  HydroGPUApp* app;
  std::shared_ptr<HydroGPU::Integrator::Integrator> integrator;
  std::shared_ptr<Equation::Equation> equation;
  int frame = 0;

  explicit Solver(HydroGPUApp* app_);
  ~Solver() override;
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // src/Solver/Solver.cpp:48-205: equation, device buffers, then the
  // integrator picked by app->integratorName (throws for an unknown name)
  void init() override;

 protected:
  virtual void createEquation() {}
  // HYDRO_SOLVER_* of the flux solver
  virtual int getSolverId() const = 0;

 public:
  std::shared_ptr<Equation::Equation> getEquation() const override { return equation; }
  int numStates();
  int getVolume();
  // ghost cells of the state (src/Solver/Solver.cpp:446-464)
  virtual void boundary();

 protected:
  // the frozen start-of-step data are taken inside calcTimestep / step on the
  // device, so initStep has nothing left to do (src/Solver/Solver.cpp:518-519)
  virtual void initStep() {}
  // cfl * min cell dt (FP64), freezing the start-of-step eigen data
  virtual real calcTimestep();
  // all integrator stages plus the solver's per-step tail; advances the
  // simulated time by dt
  virtual void step(real dt);
  friend struct HydroGPU::Integrator::Integrator;

 public:
  // Roe::calcDeriv / HLL::calcDeriv (src/Solver/Roe.cpp:101-112): adds the
  // flux derivative of the state buffer's current contents to derivBuffer,
  // with the frozen start-of-step data and the full-step dt; the callback of
  // the host-side integrate(dt, callback) form
  virtual void calcDeriv(void* derivBuffer, real dt);

 public:
  // src/Solver/Solver.cpp:521-545: boundary -> initStep -> calcTimestep (or
  // fixedDT) -> step
  virtual void update();
  // src/Solver/Solver.cpp:345-392: initState per cell centre (ghosts
  // included), readStateCell, upload
  void resetState() override;

  // Converter::fromGPU / toGPU: the state in the reference's layout
  // state[s + numStates() * cell]
  std::vector<real> getState();
  void setState(const std::vector<real>& state);

  // src/Solver/Solver.cpp:547-589: one FITS image per state channel named
  // `<state name><index>.fits` at the first free index (getSaveIndex), the
  // full grid with ghosts, as 32-bit floats
  virtual std::vector<std::string> getSaveChannelNames();
  int getSaveIndex();
  virtual void save();

  // no counterpart in the reference: a checkpoint of the state (with the
  // SRHD primitives) and the clock, and the exact resume from one (the
  // descriptor must describe the same grid and equation)
  void saveCheckpoint(const std::string& path);
  void loadCheckpoint(const std::string& path);

  // no counterpart in the reference: `nsteps` updates with dt kept on the
  // device (no per-step host synchronisation); simulated time and step count;
  // FP64 interior totals; health flags; the C handle
  void run(int nsteps);
  real getTime();
  long long getSteps();
  std::vector<real> totals();
  unsigned health();
  hydro_sim* handle() const { return sim; }
  // the descriptor init() built from the app (for callers that drive the C ABI)
  const hydro_sim_desc& desc() const { return desc_; }

 private:
  hydro_ctx* ctx = nullptr;
  hydro_sim* sim = nullptr;
  hydro_sim_desc desc_{};
};

}  // namespace Solver
}  // namespace HydroGPU
