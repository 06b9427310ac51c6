// HydroGPU/Equation/Equation.h -- equation plug-in interface
// (include/HydroGPU/Equation/Equation.h:15-34).
//
// An equation names its states, maps boundary methods to per-state ghost
// operations and converts the initState channels into a state vector.  The
// OpenCL program-assembly hooks of the reference (getProgramSources,
// buildEnumCode, the GL texture/vector-field setup) have no role here: the
// kernels are compiled ahead of time and select the equation by id.
#pragma once
#include <string>
#include <vector>

#include "HydroGPU/Shared/Common.h"

namespace HydroGPU {
struct HydroGPUApp;
namespace Equation {

// include/HydroGPU/Boundary/Boundary.h:17-24
enum {
  BOUNDARY_KERNEL_NONE = -1,
  BOUNDARY_KERNEL_PERIODIC,
  BOUNDARY_KERNEL_MIRROR,
  BOUNDARY_KERNEL_REFLECT,
  BOUNDARY_KERNEL_FREEFLOW,
  NUM_BOUNDARY_KERNELS
};

struct Equation {
 protected:
  HydroGPUApp* app;

 public:
  std::vector<std::string> displayVariables;
  std::vector<std::string> vectorFieldVars;
  std::vector<std::string> boundaryMethods;
  std::vector<std::string> states;

  explicit Equation(HydroGPUApp* app_);
  virtual ~Equation() {}
  // ghost operation for (axis, state, min/max) under the app's boundary methods
  virtual int stateGetBoundaryKernelForBoundaryMethod(int dim, int state, int minmax) = 0;
  // channels of one cell -> state vector (src/Equation/Equation.cpp:30-34)
  virtual void readStateCell(real* state, const real* source);
  virtual int numReadStateChannels();
  virtual std::string name() const = 0;
  // the equation id of the C ABI (HYDRO_EQN_*)
  virtual int equationId() const = 0;
};

}  // namespace Equation
}  // namespace HydroGPU
