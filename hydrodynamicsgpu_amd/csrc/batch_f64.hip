// batch_f64.hip -- instantiations of the batched stage kernel
// (double arithmetic).
#include "batch_impl.hpp"

namespace hydro {

void launch_batch_stage_f64(int solver, int lim, int mode, const BatchStageArgs<double, double>& a, int blocks,
                            hipStream_t st) {
  launch_batch_solver<double, double, false>(solver, lim, mode, a, blocks, st);
}

}  // namespace hydro
