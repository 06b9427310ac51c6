// stage2d_f32k.hip -- instantiations of the 1D/2D fused stage kernel
// (float arithmetic, compensated FP32 storage).
#include "stage2d_impl.hpp"

namespace hydro {

void launch_stage2d_f32k(int dim, int solver, int lim, int mode, const StageArgs<float, float>& a, int wpb,
                         hipStream_t st) {
  if (dim == 1) launch_solver<1, float, float, true>(solver, lim, mode, a, wpb, st);
  else launch_solver<2, float, float, true>(solver, lim, mode, a, wpb, st);
}

}  // namespace hydro
