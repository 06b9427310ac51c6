// batch_f32.hip -- instantiations of the batched stage kernel
// (float arithmetic).
#include "batch_impl.hpp"

namespace hydro {

void launch_batch_stage_f32(int solver, int lim, int mode, const BatchStageArgs<float, float>& a, int blocks,
                            hipStream_t st) {
  launch_batch_solver<float, float, false>(solver, lim, mode, a, blocks, st);
}

}  // namespace hydro
