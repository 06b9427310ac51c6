// Equation plug-ins: state names, boundary-method mapping and initState
// channel conversion (src/Equation/Equation.cpp, Euler.cpp, SRHD.cpp,
// Maxwell.cpp).
#include "HydroGPU/Equation/Euler.h"
#include "HydroGPU/Equation/Maxwell.h"
#include "HydroGPU/Equation/SRHD.h"
#include "HydroGPU/Exception.h"
#include "HydroGPU/HydroGPUApp.h"

namespace HydroGPU {
namespace Equation {

Equation::Equation(HydroGPUApp* app_) : app(app_) {}

void Equation::readStateCell(real* state, const real* source) {
  for (int i = 0; i < (int)states.size(); ++i) state[i] = source[i];
}

int Equation::numReadStateChannels() { return (int)states.size(); }

namespace {
// the mapping shared by Euler and SRHD (src/Equation/Euler.cpp:63-75,
// src/Equation/SRHD.cpp:63-75): MIRROR reflects the momentum normal to the axis
int fluidBoundaryKernel(HydroGPUApp* app, int dim, int state, int minmax) {
  switch (app->getBoundaryMethod(dim, minmax)) {
  case BOUNDARY_METHOD_NONE: return BOUNDARY_KERNEL_NONE;
  case BOUNDARY_METHOD_PERIODIC: return BOUNDARY_KERNEL_PERIODIC;
  case BOUNDARY_METHOD_MIRROR: return dim + 1 == state ? BOUNDARY_KERNEL_REFLECT : BOUNDARY_KERNEL_MIRROR;
  case BOUNDARY_METHOD_FREEFLOW: return BOUNDARY_KERNEL_FREEFLOW;
  }
  throw Exception() << "got an unknown boundary method " << app->getBoundaryMethod(dim, minmax) << " for dim " << dim;
}
}  // namespace

Euler::Euler(HydroGPUApp* app_, int dim_) : Super(app_), dim(dim_) {
  if (dim == -1) dim = app->dim;
  if (dim < 1 || dim > 3) throw Exception() << "Equation::Euler must have a dim between 1 and 3.  Got " << dim;
  displayVariables = {"DENSITY", "VELOCITY", "VELOCITY_X"};
  if (dim > 1) displayVariables.push_back("VELOCITY_Y");
  if (dim > 2) displayVariables.push_back("VELOCITY_Z");
  for (const char* v : {"PRESSURE", "ENERGY_INTERNAL", "ENERGY_KINETIC", "ENERGY_TOTAL", "POTENTIAL"})
    displayVariables.push_back(v);
  boundaryMethods = {"PERIODIC", "MIRROR", "FREEFLOW"};
  states = {"DENSITY", "MOMENTUM_X"};
  if (dim > 1) states.push_back("MOMENTUM_Y");
  if (dim > 2) states.push_back("MOMENTUM_Z");
  states.push_back("ENERGY_TOTAL");
  vectorFieldVars = {"VELOCITY", "MOMENTUM", "GRAVITY"};
  if (dim == 3) vectorFieldVars.push_back("VORTICITY");
}

int Euler::stateGetBoundaryKernelForBoundaryMethod(int d, int state, int minmax) {
  return fluidBoundaryKernel(app, d, state, minmax);
}

// channels: rho, rho v[3], B[3], E, ... (res/util.lua:36-80); the energy is
// channel 7 (src/Equation/Euler.cpp:77-90)
void Euler::readStateCell(real* state, const real* source) {
  state[0] = source[0];
  state[1] = source[1];
  if (dim > 1) state[2] = source[2];
  if (dim > 2) state[3] = source[3];
  state[states.size() - 1] = source[7];
}

int Euler::numReadStateChannels() { return 8; }

SRHD::SRHD(HydroGPUApp* app_) : Super(app_) {
  displayVariables = {"DENSITY", "VELOCITY_X"};
  if (app->dim > 1) displayVariables.push_back("VELOCITY_Y");
  if (app->dim > 2) displayVariables.push_back("VELOCITY_Z");
  for (const char* v : {"VELOCITY_MAGN", "E_INTERNAL", "P", "H", "D", "S_X"}) displayVariables.push_back(v);
  if (app->dim > 1) displayVariables.push_back("S_Y");
  if (app->dim > 2) displayVariables.push_back("S_Z");
  displayVariables.push_back("S_MAGN");
  displayVariables.push_back("TAU");
  boundaryMethods = {"PERIODIC", "MIRROR", "FREEFLOW"};
  states = {"REST_MASS_DENSITY", "MOMENTUM_DENSITY_X"};
  if (app->dim > 1) states.push_back("MOMENTUM_DENSITY_Y");
  if (app->dim > 2) states.push_back("MOMENTUM_DENSITY_Z");
  states.push_back("TOTAL_ENERGY_DENSITY");
  vectorFieldVars = {"VELOCITY", "MOMENTUM", "VORTICITY"};
}

int SRHD::stateGetBoundaryKernelForBoundaryMethod(int d, int state, int minmax) {
  return fluidBoundaryKernel(app, d, state, minmax);
}

// Newtonian channels as for Euler (src/Equation/SRHD.cpp:79-95); the
// conversion to (D, S, tau) happens at upload (res/SRHDCommon.cl:24-92)
void SRHD::readStateCell(real* state, const real* source) {
  state[0] = source[0];
  state[1] = source[1];
  if (app->dim > 1) state[2] = source[2];
  if (app->dim > 2) state[3] = source[3];
  state[states.size() - 1] = source[7];
}

int SRHD::numReadStateChannels() { return 8; }

Maxwell::Maxwell(HydroGPUApp* app_) : Super(app_) {
  displayVariables = {"ELECTRIC", "ELECTRIC_X", "ELECTRIC_Y", "ELECTRIC_Z",
                      "MAGNETIC", "MAGNETIC_X", "MAGNETIC_Y", "MAGNETIC_Z"};
  boundaryMethods = {"PERIODIC", "MIRROR", "FREEFLOW"};
  states = {"ELECTRIC_X", "ELECTRIC_Y", "ELECTRIC_Z", "MAGNETIC_X", "MAGNETIC_Y", "MAGNETIC_Z"};
  vectorFieldVars = {"ELECTRIC", "MAGNETIC", "POYNTING"};
}

// MIRROR reflects the normal E and B components (src/Equation/Maxwell.cpp:73-85)
int Maxwell::stateGetBoundaryKernelForBoundaryMethod(int d, int state, int minmax) {
  switch (app->getBoundaryMethod(d, minmax)) {
  case BOUNDARY_METHOD_NONE: return BOUNDARY_KERNEL_NONE;
  case BOUNDARY_METHOD_PERIODIC: return BOUNDARY_KERNEL_PERIODIC;
  case BOUNDARY_METHOD_MIRROR:
    return (state == d || state == 3 + d) ? BOUNDARY_KERNEL_REFLECT : BOUNDARY_KERNEL_MIRROR;
  case BOUNDARY_METHOD_FREEFLOW: return BOUNDARY_KERNEL_FREEFLOW;
  }
  throw Exception() << "got an unknown boundary method " << app->getBoundaryMethod(d, minmax) << " for dim " << d;
}

}  // namespace Equation
}  // namespace HydroGPU
