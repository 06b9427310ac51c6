// stage2d_f64.hip -- instantiations of the 1D/2D fused stage kernel
// (double arithmetic).
#include "stage2d_impl.hpp"

namespace hydro {

void launch_stage2d_f64(int dim, int solver, int lim, int mode, const StageArgs<double, double>& a, int wpb,
                        hipStream_t st) {
  if (dim == 1) launch_solver<1, double, double, false>(solver, lim, mode, a, wpb, st);
  else launch_solver<2, double, double, false>(solver, lim, mode, a, wpb, st);
}

}  // namespace hydro
