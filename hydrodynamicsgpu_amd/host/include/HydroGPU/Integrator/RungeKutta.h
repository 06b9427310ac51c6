// HydroGPU/Integrator/RungeKutta.h -- the Runge-Kutta family
// (include/HydroGPU/Integrator/RungeKutta.h:16-294).
//
// RungeKutta<Tableau> with the reference's eleven tableau names.  The
// coefficients themselves live with the device stage planner (as written in
// the reference, including the three tableaus whose stability polynomials
// differ from their names'); a tableau tag carries the C-ABI id and the
// stage count.
#pragma once
#include "HydroGPU/Integrator/Integrator.h"
#include "hydro_capi.h"

namespace HydroGPU {
namespace Integrator {

template <typename Tableau>
struct RungeKutta : public Integrator {
  typedef Integrator Super;
  using Super::Super;
  int integratorId() const override { return Tableau::id; }
  std::string name() const override { return Tableau::name(); }
  int numStages() const override { return Tableau::order; }
};

#define HYDROGPU_RK_TABLEAU(NAME, ID, ORDER)                \
  struct NAME##Tableau {                                    \
    enum { id = ID, order = ORDER };                        \
    static std::string name() { return #NAME; }             \
  };                                                        \
  typedef RungeKutta<NAME##Tableau> NAME;

HYDROGPU_RK_TABLEAU(RungeKutta2, HYDRO_INTEGRATOR_RK2, 2)  // RungeKutta.h:129-140
HYDROGPU_RK_TABLEAU(RungeKutta2Heun, HYDRO_INTEGRATOR_RK2HEUN, 2)  // :142-153
HYDROGPU_RK_TABLEAU(RungeKutta2Ralston, HYDRO_INTEGRATOR_RK2RALSTON, 2)  // :155-166
HYDROGPU_RK_TABLEAU(RungeKutta3, HYDRO_INTEGRATOR_RK3, 3)  // :168-181
HYDROGPU_RK_TABLEAU(RungeKutta4, HYDRO_INTEGRATOR_RK4, 4)  // :183-198
HYDROGPU_RK_TABLEAU(RungeKutta4_3_8thsRule, HYDRO_INTEGRATOR_RK4_3_8THSRULE, 4)  // :200-215
HYDROGPU_RK_TABLEAU(RungeKutta2TVD, HYDRO_INTEGRATOR_RK2TVD, 2)  // :219-230
HYDROGPU_RK_TABLEAU(RungeKutta2NonTVD, HYDRO_INTEGRATOR_RK2NONTVD, 2)  // :232-243
HYDROGPU_RK_TABLEAU(RungeKutta3TVD, HYDRO_INTEGRATOR_RK3TVD, 3)  // :245-258
HYDROGPU_RK_TABLEAU(RungeKutta4TVD, HYDRO_INTEGRATOR_RK4TVD, 4)  // :260-275
HYDROGPU_RK_TABLEAU(RungeKutta4NonTVD, HYDRO_INTEGRATOR_RK4NONTVD, 4)  // :279-294
#undef HYDROGPU_RK_TABLEAU

}  // namespace Integrator
}  // namespace HydroGPU
