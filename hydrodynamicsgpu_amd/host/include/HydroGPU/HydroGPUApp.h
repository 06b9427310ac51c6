// HydroGPU/HydroGPUApp.h -- the run configuration the solvers read.
//
// The reference's HydroGPUApp (include/HydroGPU/HydroGPUApp.h:35-140) is a GL
// application that reads these fields from res/config.lua and owns the solver
// registry (src/HydroGPUApp.cpp:90-119, 330-349).  Here it keeps only what
// the update path reads: the same field names and meanings, the `defs`
// table, `initState` as a callable instead of a Lua function, and the solver
// registry.  No window, GL or Lua.
#pragma once
#include <array>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "HydroGPU/Shared/Common.h"

namespace HydroGPU {

namespace Solver {
struct Solver;
}

// include/HydroGPU/Boundary/Boundary.h:8-14
enum {
  BOUNDARY_METHOD_NONE = -1,
  BOUNDARY_METHOD_PERIODIC,
  BOUNDARY_METHOD_MIRROR,
  BOUNDARY_METHOD_FREEFLOW,
  NUM_BOUNDARY_METHODS
};

// device precision of the update path (no counterpart in the reference,
// which is FP64 throughout)
enum Precision { PRECISION_FP64 = 0, PRECISION_FP32_MIXED = 1 };

struct HydroGPUApp {
  // config (src/HydroGPUApp.cpp:164-230; res/config.lua)
  std::string solverName = "EulerRoe";
  std::string integratorName = "ForwardEuler";
  std::string slopeLimiter = "Superbee";
  int dim = 1;
  std::array<int, 4> size{{1, 1, 1, 1}};  // cells per axis, ghosts included
  std::array<real, 4> xmin{{-.5, -.5, -.5, 0}};
  std::array<real, 4> xmax{{.5, .5, .5, 0}};
  bool useFixedDT = false;
  real fixedDT = .001;
  float cfl = .5f;  // a float in the reference (include/HydroGPU/HydroGPUApp.h:84)
  int boundaryMethods[3][2] = {{BOUNDARY_METHOD_FREEFLOW, BOUNDARY_METHOD_FREEFLOW},
                               {BOUNDARY_METHOD_FREEFLOW, BOUNDARY_METHOD_FREEFLOW},
                               {BOUNDARY_METHOD_FREEFLOW, BOUNDARY_METHOD_FREEFLOW}};
  bool useGravity = false;      // relax the potential from the density (src/HydroGPUApp.cpp:77, 207)
  int gaussSeidelMaxIter = 20;  // relaxation sweeps per potential solve (src/HydroGPUApp.cpp:78, 208)
  // the `defs` table (res/config.lua:70-93): idealGas_heatCapacityRatio,
  // srhd_*, maxwell_*; missing keys take the reference defaults
  std::map<std::string, real> defs;
  // initState(x, y, z) -> channel values of one cell (res/initConds.lua;
  // buildStateEuler order: rho, rho v[3], B[3], E, potential, solid)
  std::function<std::vector<real>(real, real, real)> initState;

  // no counterpart in the reference
  Precision precision = PRECISION_FP64;
  unsigned flags = 0x3u;  // HYDRO_FLAG_REFERENCE: 6-digit dx, gravity identity pass
  int device = 0;

  typedef std::shared_ptr<Solver::Solver> SolverPtr;
  typedef std::function<SolverPtr()> SolverGenFunc;
  struct SolverGenPair {
    SolverGenPair(const std::string& name_, SolverGenFunc func_) : name(name_), func(func_) {}
    std::string name;
    SolverGenFunc func;
  };
  struct SolverEqnsPair {
    SolverEqnsPair(const std::string& name_, const std::vector<SolverGenPair>& generators_)
        : name(name_), generators(generators_) {}
    std::string name;
    std::vector<SolverGenPair> generators;
  };
  // equation -> solvers, as src/HydroGPUApp.cpp:90-119 (the solvers of this
  // build: Euler Roe/HLL/HLLC, SRHD Roe, Maxwell Roe)
  std::vector<SolverEqnsPair> solverGensForEqns;
  SolverPtr solver;

  HydroGPUApp();
  virtual ~HydroGPUApp() {}

  // boundaryMethods(dim, minmax) accessor of the reference's tensor
  int getBoundaryMethod(int d, int minmax) const { return boundaryMethods[d][minmax]; }

  // src/HydroGPUApp.cpp:330-349: pick `solverName` from the registry, then
  // init() and resetState().  Throws Exception for an unknown name.
  SolverPtr createSolver();
};

}  // namespace HydroGPU
