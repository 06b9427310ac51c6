// Solver base and the concrete solvers: the reference's virtual interface
// (src/Solver/Solver.cpp, Roe.cpp, HLL.cpp, EulerRoe.cpp, EulerHLL.cpp,
// EulerHLLC.cpp, EulerBurgers.cpp, SRHDRoe.cpp, MaxwellRoe.cpp) with bodies that call the C ABI
// of libhydro_hip.so.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>

#include "HydroGPU/Equation/Euler.h"
#include "HydroGPU/Equation/Maxwell.h"
#include "HydroGPU/Equation/SRHD.h"
#include "HydroGPU/HydroGPUApp.h"
#include "HydroGPU/Integrator/ForwardEuler.h"
#include "HydroGPU/Integrator/RungeKutta.h"
#include "HydroGPU/Shared/Files.h"
#include "HydroGPU/Solver/EulerHLL.h"
#include "HydroGPU/Solver/EulerHLLC.h"
#include "HydroGPU/Solver/EulerBurgers.h"
#include "HydroGPU/Solver/EulerRoe.h"
#include "HydroGPU/Solver/MaxwellRoe.h"
#include "HydroGPU/Solver/SRHDRoe.h"

namespace HydroGPU {

void checkStatus(int status, const char* where) {
  if (status != HYDRO_OK) throw Exception() << where << ": " << hydro_last_error();
}

namespace {

// res/SlopeLimiter.cl:5-68, in the order of the HYDRO_LIMITER_* ids
const char* const kLimiterNames[HYDRO_NUM_LIMITERS] = {
    "DonorCell", "LaxWendroff", "BeamWarming", "Fromm", "CHARM", "HCUS", "HQUICK",
    "Koren", "MinMod", "Oshker", "Ospre", "Smart", "Sweby", "UMIST", "VanAlbada1",
    "VanAlbada2", "VanLeer", "MonotizedCentral", "Superbee", "BarthJespersen"};

int limiterId(const std::string& name) {
  for (int i = 0; i < HYDRO_NUM_LIMITERS; ++i)
    if (name == kLimiterNames[i]) return i;
  throw Exception() << "unknown slope limiter " << name;
}

// a `defs` entry as the kernels see it: Lua's tostring ("%.14g"), which is
// how the reference renders defs into the program source
// (src/Solver/Solver.cpp:289-307)
real def(const HydroGPUApp* app, const char* key, real fallback) {
  auto it = app->defs.find(key);
  if (it == app->defs.end()) return fallback;
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.14g", it->second);
  return std::strtod(buf, nullptr);
}

}  // namespace

namespace Integrator {

void Integrator::integrate(real dt) { solver->step(dt); }

void Integrator::integrate(real dt, std::function<void(void*)> callback) {
  hydro_sim* sim = solver->handle();
  int order = 0;
  double A[16], B[16];
  checkStatus(hydro_integrator_tableau(integratorId(), &order, A, B), "hydro_integrator_tableau");
  auto alpha = [&](int i, int k) { return A[4 * i + k]; };
  auto beta = [&](int i, int k) { return B[4 * i + k]; };
  // RungeKutta::RungeKutta (RungeKutta.h:28-51): a stage buffer exists only
  // if some later row reads it
  if (stateBuffers.empty()) {
    stateBuffers.assign(order, nullptr);
    derivBuffers.assign(order, nullptr);
    for (int i = 0; i < order; ++i) {
      bool su = false, sd = false;
      for (int m = i; m < order; ++m) {
        su |= alpha(m, i) != 0;
        sd |= beta(m, i) != 0;
      }
      if (su) checkStatus(hydro_sim_buffer_create(sim, &stateBuffers[i]), "hydro_sim_buffer_create");
      if (sd) checkStatus(hydro_sim_buffer_create(sim, &derivBuffers[i]), "hydro_sim_buffer_create");
    }
  }
  void* state = nullptr;
  checkStatus(hydro_sim_state_buffer(sim, &state), "hydro_sim_state_buffer");
  checkStatus(hydro_sim_init_step(sim), "hydro_sim_init_step");
  // RungeKutta::integrate (RungeKutta.h:59-125)
  if (stateBuffers[0]) checkStatus(hydro_sim_buffer_copy(sim, stateBuffers[0], state), "hydro_sim_buffer_copy");
  if (derivBuffers[0]) {
    checkStatus(hydro_sim_buffer_zero(sim, derivBuffers[0]), "hydro_sim_buffer_zero");
    callback(derivBuffers[0]);
  }
  for (int i = 1; i <= order; ++i) {
    checkStatus(hydro_sim_buffer_zero(sim, state), "hydro_sim_buffer_zero");
    for (int k = 0; k < i; ++k) {
      if (alpha(i - 1, k) != 0)
        checkStatus(hydro_sim_mult_add(sim, state, state, stateBuffers[k], alpha(i - 1, k)), "hydro_sim_mult_add");
      if (beta(i - 1, k) != 0)
        checkStatus(hydro_sim_mult_add(sim, state, state, derivBuffers[k], beta(i - 1, k) * dt), "hydro_sim_mult_add");
    }
    if (i < order) {
      if (stateBuffers[i]) checkStatus(hydro_sim_buffer_copy(sim, stateBuffers[i], state), "hydro_sim_buffer_copy");
      if (derivBuffers[i]) {
        checkStatus(hydro_sim_buffer_zero(sim, derivBuffers[i]), "hydro_sim_buffer_zero");
        callback(derivBuffers[i]);
      }
    }
  }
  checkStatus(hydro_sim_finish_step(sim, dt), "hydro_sim_finish_step");
}

void ForwardEuler::integrate(real dt, std::function<void(void*)> callback) {
  hydro_sim* sim = solver->handle();
  if (derivBuffers.empty()) {
    derivBuffers.assign(1, nullptr);
    checkStatus(hydro_sim_buffer_create(sim, &derivBuffers[0]), "hydro_sim_buffer_create");
  }
  void* state = nullptr;
  checkStatus(hydro_sim_state_buffer(sim, &state), "hydro_sim_state_buffer");
  checkStatus(hydro_sim_init_step(sim), "hydro_sim_init_step");
  checkStatus(hydro_sim_buffer_zero(sim, derivBuffers[0]), "hydro_sim_buffer_zero");
  callback(derivBuffers[0]);
  checkStatus(hydro_sim_mult_add(sim, state, state, derivBuffers[0], dt), "hydro_sim_mult_add");
  checkStatus(hydro_sim_finish_step(sim, dt), "hydro_sim_finish_step");
}

}  // namespace Integrator

namespace Solver {

Solver::Solver(HydroGPUApp* app_) : app(app_) {}

Solver::~Solver() {
  if (sim) hydro_sim_destroy(sim);
  if (ctx) hydro_ctx_destroy(ctx);
}

void Solver::init() {
  createEquation();
  if (!equation) throw Exception() << name() << " has no equation";

  // src/Solver/Solver.cpp:178-204
  typedef std::map<std::string, std::function<std::shared_ptr<Integrator::Integrator>()>> gensMap_t;
  gensMap_t gens;
#define MAKE_INTEGRATOR(I) gens[#I] = [this]() { return std::make_shared<Integrator::I>(this); }
  MAKE_INTEGRATOR(ForwardEuler);
  MAKE_INTEGRATOR(RungeKutta2);
  MAKE_INTEGRATOR(RungeKutta2Heun);
  MAKE_INTEGRATOR(RungeKutta2Ralston);
  MAKE_INTEGRATOR(RungeKutta3);
  MAKE_INTEGRATOR(RungeKutta4);
  MAKE_INTEGRATOR(RungeKutta4_3_8thsRule);
  MAKE_INTEGRATOR(RungeKutta2TVD);
  MAKE_INTEGRATOR(RungeKutta2NonTVD);
  MAKE_INTEGRATOR(RungeKutta3TVD);
  MAKE_INTEGRATOR(RungeKutta4TVD);
  MAKE_INTEGRATOR(RungeKutta4NonTVD);
#undef MAKE_INTEGRATOR
  auto i = gens.find(app->integratorName);
  if (i == gens.end()) throw Exception() << "failed to find integrator named " << app->integratorName;
  integrator = i->second();

  hydro_sim_desc d;
  hydro_sim_desc_default(&d);
  d.dim = app->dim;
  for (int a = 0; a < 3; ++a) {
    d.size[a] = a < app->dim ? app->size[a] : 1;
    d.xmin[a] = app->xmin[a];
    d.xmax[a] = app->xmax[a];
    for (int m = 0; m < 2; ++m) d.boundary[a][m] = app->boundaryMethods[a][m];
  }
  d.equation = equation->equationId();
  d.solver = getSolverId();
  d.limiter = limiterId(app->slopeLimiter);
  d.integrator = integrator->integratorId();
  d.precision = app->precision == PRECISION_FP32_MIXED ? HYDRO_PRECISION_FP32_MIXED : HYDRO_PRECISION_FP64;
  d.cfl = app->cfl;
  d.use_fixed_dt = app->useFixedDT ? 1 : 0;
  d.fixed_dt = app->fixedDT;
  d.gamma = def(app, "idealGas_heatCapacityRatio", d.gamma);
  d.srhd_solve_prim_max_iter = (int)def(app, "srhd_solvePrimMaxIter", d.srhd_solve_prim_max_iter);
  d.srhd_solve_prim_stop_epsilon = def(app, "srhd_solvePrimStopEpsilon", d.srhd_solve_prim_stop_epsilon);
  d.srhd_solve_prim_vel_epsilon = def(app, "srhd_solvePrimVelEpsilon", d.srhd_solve_prim_vel_epsilon);
  d.srhd_solve_prim_pmin_epsilon = def(app, "srhd_solvePrimPMinEpsilon", d.srhd_solve_prim_pmin_epsilon);
  d.srhd_rho_min = def(app, "srhd_rhoMin", d.srhd_rho_min);
  d.srhd_rho_max = def(app, "srhd_rhoMax", d.srhd_rho_max);
  d.srhd_eint_max = def(app, "srhd_eIntMax", d.srhd_eint_max);
  d.srhd_D_min = def(app, "srhd_DMin", d.srhd_D_min);
  d.srhd_D_max = def(app, "srhd_DMax", d.srhd_D_max);
  d.srhd_tau_min = def(app, "srhd_tauMin", d.srhd_tau_min);
  d.srhd_tau_max = def(app, "srhd_tauMax", d.srhd_tau_max);
  d.maxwell_permittivity = def(app, "maxwell_permittivity", d.maxwell_permittivity);
  d.maxwell_permeability = def(app, "maxwell_permeability", d.maxwell_permeability);
  d.maxwell_conductivity = def(app, "maxwell_conductivity", d.maxwell_conductivity);
  d.flags = app->flags;
  d.use_gravity = app->useGravity ? 1 : 0;
  d.gauss_seidel_max_iter = app->gaussSeidelMaxIter;
  d.self_grav_gravitational_constant =
      def(app, "selfGrav_gravitationalConstant", d.self_grav_gravitational_constant);

  if (!ctx) checkStatus(hydro_ctx_create(app->device, &ctx), "hydro_ctx_create");
  if (sim) {
    hydro_sim_destroy(sim);
    sim = nullptr;
  }
  checkStatus(hydro_sim_create(ctx, &d, &sim), "hydro_sim_create");
  desc_ = d;
  frame = 0;
}

int Solver::numStates() { return (int)equation->states.size(); }

int Solver::getVolume() { return app->size[0] * app->size[1] * app->size[2]; }

void Solver::boundary() { checkStatus(hydro_sim_boundary(sim), "boundary"); }

real Solver::calcTimestep() {
  double dt = 0;
  checkStatus(hydro_sim_calc_timestep(sim, &dt), "calcTimestep");
  return dt;
}

void Solver::step(real dt) { checkStatus(hydro_sim_step(sim, dt), "step"); }

void Solver::calcDeriv(void* derivBuffer, real dt) {
  checkStatus(hydro_sim_calc_deriv(sim, dt, derivBuffer), "calcDeriv");
}

void Solver::update() {
  boundary();
  initStep();
  const real dt = app->useFixedDT ? app->fixedDT : calcTimestep();
  integrator->integrate(dt);
  ++frame;
}

void Solver::resetState() {
  if (!app->initState) throw Exception() << "expected initState to be defined in config file";
  if (!sim) throw Exception() << "resetState before init";
  const int n = numStates();
  std::vector<real> stateVec((size_t)getVolume() * n);
  std::vector<real> potentialVec((size_t)getVolume(), 0.0);
  std::vector<unsigned char> solidVec((size_t)getVolume(), 0);
  bool anyPotential = false, anySolid = false;
  std::vector<real> cellResults;
  const int nch = equation->numReadStateChannels();
  int flattenedIndex = 0;
  int index[3];
  for (index[2] = 0; index[2] < app->size[2]; ++index[2]) {
    for (index[1] = 0; index[1] < app->size[1]; ++index[1]) {
      for (index[0] = 0; index[0] < app->size[0]; ++index[0], ++flattenedIndex) {
        real pos[3];
        for (int i = 0; i < 3; ++i)
          pos[i] = real(app->xmax[i] - app->xmin[i]) * (real(index[i]) + .5) / real(app->size[i]) + real(app->xmin[i]);
        cellResults = app->initState(pos[0], pos[1], pos[2]);
        if ((int)cellResults.size() < nch)
          throw Exception() << "initState returned " << cellResults.size() << " values, expected " << nch;
        equation->readStateCell(stateVec.data() + (size_t)flattenedIndex * n, cellResults.data());
        // SelfGravitationBehavior::Converter::setValues: the potential is the
        // next-to-last channel (include/HydroGPU/Solver/SelfGravitationBehavior.h:64-68)
        if ((int)cellResults.size() >= nch + 2) {
          potentialVec[flattenedIndex] = cellResults[cellResults.size() - 2];
          anyPotential = anyPotential || potentialVec[flattenedIndex] != 0.0;
          // ... and the solid flag the last, converted to char (:69-73)
          solidVec[flattenedIndex] = static_cast<long long>(cellResults.back()) != 0;
          anySolid = anySolid || solidVec[flattenedIndex];
        }
      }
    }
  }
  setState(stateVec);
  // the Euler solvers' SelfGravitation::resetState: the potential (or the
  // relaxed guess with useGravity) and its energy
  if (desc_.equation == HYDRO_EQN_EULER && (anyPotential || desc_.use_gravity))
    checkStatus(hydro_sim_set_potential(sim, anyPotential ? potentialVec.data() : nullptr, potentialVec.size()),
                "resetState");
  // ... and the solid flags (src/Solver/SelfGravitation.cpp:75)
  if (desc_.equation == HYDRO_EQN_EULER && anySolid)
    checkStatus(hydro_sim_set_solid(sim, solidVec.data(), solidVec.size()), "resetState");
}

std::vector<real> Solver::getState() {
  std::vector<real> v((size_t)getVolume() * numStates());
  checkStatus(hydro_sim_get_state(sim, v.data(), v.size()), "getState");
  return v;
}

void Solver::setState(const std::vector<real>& state) {
  checkStatus(hydro_sim_set_state(sim, state.data(), state.size()), "setState");
}

void Solver::run(int nsteps) {
  checkStatus(hydro_sim_run(sim, nsteps), "run");
  frame += nsteps;
}

real Solver::getTime() {
  double t = 0;
  checkStatus(hydro_sim_time(sim, &t, nullptr, nullptr), "getTime");
  return t;
}

long long Solver::getSteps() {
  long long n = 0;
  checkStatus(hydro_sim_time(sim, nullptr, &n, nullptr), "getSteps");
  return n;
}

std::vector<real> Solver::totals() {
  std::vector<real> v(numStates());
  checkStatus(hydro_sim_totals(sim, v.data()), "totals");
  return v;
}

unsigned Solver::health() {
  unsigned h = 0;
  checkStatus(hydro_sim_health(sim, &h), "health");
  return h;
}

std::vector<std::string> Solver::getSaveChannelNames() { return equation->states; }

int Solver::getSaveIndex() {
  const std::string first = getSaveChannelNames()[0];
  for (int i = 0; i < 1000000; ++i)
    if (!fileExists(first + std::to_string(i) + ".fits")) return i;
  throw Exception() << "failed to find available save filename";
}

void Solver::save() {
  const int saveIndex = getSaveIndex();
  const std::vector<real> state = getState();
  const std::vector<std::string> names = getSaveChannelNames();
  const int n = numStates();
  const int sx = app->size[0], sy = app->size[1], sz = app->size[2];
  std::vector<float> image((size_t)sx * sy * sz);
  for (int channel = 0; channel < (int)names.size(); ++channel) {
    for (size_t cell = 0; cell < image.size(); ++cell) image[cell] = (float)state[channel + (size_t)n * cell];
    const std::string filename = names[channel] + std::to_string(saveIndex) + ".fits";
    std::cout << "saving file " << filename << std::endl;
    writeFits(filename, image, sx, sy, sz);
  }
}

void Solver::saveCheckpoint(const std::string& path) {
  Checkpoint c;
  c.desc = desc_;
  c.numStates = numStates();
  c.volume = getVolume();
  checkStatus(hydro_sim_time(sim, &c.t, &c.steps, nullptr), "saveCheckpoint");
  c.state.resize((size_t)c.numStates * (size_t)c.volume);
  if (desc_.precision == HYDRO_PRECISION_FP32_MIXED && (desc_.flags & HYDRO_FLAG_KAHAN)) {
    c.low.resize(c.state.size());
    checkStatus(hydro_sim_get_storage(sim, c.state.data(), c.low.data(), c.state.size()), "saveCheckpoint");
  } else {
    c.state = getState();
  }
  if (desc_.equation == HYDRO_EQN_SRHD) {
    c.prims.resize(c.state.size());
    checkStatus(hydro_sim_get_primitives(sim, c.prims.data(), c.prims.size()), "saveCheckpoint");
  }
  writeCheckpoint(path, c);
}

void Solver::loadCheckpoint(const std::string& path) {
  const Checkpoint c = readCheckpoint(path);
  if (c.numStates != numStates() || c.volume != getVolume() || c.desc.equation != desc_.equation ||
      c.desc.dim != desc_.dim)
    throw Exception() << path << " holds a different grid or equation";
  std::vector<double> state = c.state;
  const float* low = nullptr;
  if (!c.low.empty()) {
    if (desc_.precision == HYDRO_PRECISION_FP32_MIXED && (desc_.flags & HYDRO_FLAG_KAHAN))
      low = c.low.data();
    else
      for (size_t i = 0; i < state.size(); ++i) state[i] += (double)c.low[i];
  }
  checkStatus(hydro_sim_restore(sim, state.data(), low, c.prims.empty() ? nullptr : c.prims.data(), state.size(),
                                c.t, c.steps),
              "loadCheckpoint");
  frame = (int)c.steps;
}

void EulerRoe::createEquation() { equation = std::make_shared<Equation::Euler>(app); }
void EulerHLL::createEquation() { equation = std::make_shared<Equation::Euler>(app); }
void EulerBurgers::createEquation() { equation = std::make_shared<Equation::Euler>(app); }
void SRHDRoe::createEquation() { equation = std::make_shared<Equation::SRHD>(app); }
void MaxwellRoe::createEquation() { equation = std::make_shared<Equation::Maxwell>(app); }

std::vector<real> SRHDRoe::getPrimitives() {
  std::vector<real> v((size_t)getVolume() * numStates());
  checkStatus(hydro_sim_get_primitives(handle(), v.data(), v.size()), "getPrimitives");
  return v;
}

}  // namespace Solver

HydroGPUApp::HydroGPUApp() {
#define MAKE_SOLVER(S) SolverGenPair(#S, [this]() -> SolverPtr { return std::make_shared<Solver::S>(this); })
  solverGensForEqns = {
      SolverEqnsPair("Euler", {MAKE_SOLVER(EulerBurgers), MAKE_SOLVER(EulerHLL), MAKE_SOLVER(EulerHLLC),
                               MAKE_SOLVER(EulerRoe)}),
      SolverEqnsPair("SRHD", {MAKE_SOLVER(SRHDRoe)}),
      SolverEqnsPair("Maxwell", {MAKE_SOLVER(MaxwellRoe)}),
  };
#undef MAKE_SOLVER
}

HydroGPUApp::SolverPtr HydroGPUApp::createSolver() {
  for (const SolverEqnsPair& p : solverGensForEqns)
    for (const SolverGenPair& q : p.generators)
      if (q.name == solverName) {
        solver = q.func();
        solver->init();
        solver->resetState();
        return solver;
      }
  throw Exception() << "unknown solver " << solverName;
}

}  // namespace HydroGPU
