// batch_f32k.hip -- instantiations of the batched stage kernel
// (float arithmetic, compensated FP32 storage).
#include "batch_impl.hpp"

namespace hydro {

void launch_batch_stage_f32k(int solver, int lim, int mode, const BatchStageArgs<float, float>& a, int blocks,
                             hipStream_t st) {
  launch_batch_solver<float, float, true>(solver, lim, mode, a, blocks, st);
}

}  // namespace hydro
