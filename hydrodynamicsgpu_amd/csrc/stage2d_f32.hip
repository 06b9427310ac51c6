// stage2d_f32.hip -- instantiations of the 1D/2D fused stage kernel
// (float arithmetic).
#include "stage2d_impl.hpp"

namespace hydro {

void launch_stage2d_f32(int dim, int solver, int lim, int mode, const StageArgs<float, float>& a, int wpb,
                        hipStream_t st) {
  if (dim == 1) launch_solver<1, float, float, false>(solver, lim, mode, a, wpb, st);
  else launch_solver<2, float, float, false>(solver, lim, mode, a, wpb, st);
}

}  // namespace hydro
