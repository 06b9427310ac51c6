// stage_args.hpp -- kernel argument blocks shared by the host driver and the
// device code.
#pragma once
#include <stdint.h>

namespace hydro {

// Maximum number of Shu-Osher rows whose partial sums can be in flight.
constexpr int kMaxPartials = 3;

// How a stage combines its derivative with the stored increments.  The host
// picks the mode from the stage plan; all but GENERIC are fixed-function
// forms without runtime branches in the kernel.
enum CombineMode {
  CM_GENERIC = 0,  // partial sums of later rows (pkp / fp[] below)
  CM_FIRST = 1,    // k = 0, not last:   d_1 = b dt L_0
  CM_MID = 2,      // 0 < k < S-1:       d_{k+1} = a d_k + b dt L_k
  CM_FINAL = 3,    // k = S-1 > 0:       u_new = u0 + a d_k + sum_j rc_j d_{j+1} + b dt L_k
  CM_SINGLE = 4,   // single stage:      u_new = u0 + b dt L_0
  CM_DERIV = 5,    // derivative only:   dnext += L(dk), dk a full state, frozen data of u0
};

// One fused RK stage: L(u^(k)) with the frozen start-of-step data, then the
// Shu-Osher combination.  States are carried as increments over the
// start-of-step state u0:  u^(k) = u0 + d_k,  d_0 = 0, and row m of the
// tableau becomes  d_{m+1} = sum_k (alpha[m][k] d_k + dt beta[m][k] L_k).
template <typename T, typename S>
struct StageArgs {
  const S* u0;          // start-of-step state planes (FP64, or FP32 high words)
  const S* prim;        // SRHD: start-of-step primitive planes (storage precision), else null
  const S* pot;         // gravitational potential plane (policies with kPotential), else null
  const S* solid;       // solid mask plane (policies with kSolid), else null
  const float* u0lo;    // FP32 low words (compensated storage), or null
  const T* dk;          // d_k planes, null for k == 0
  T* dnext;             // d_{k+1} planes (non-final stage)
  S* unew;              // final stage: u0 + d_S
  float* unewlo;        // final stage low words (compensated), or null
  // CM_GENERIC: partial sum of this row (read), and of later rows (updated)
  const T* pkp;                    // null when row k has no partial sum
  int nf;                          // later rows carried
  T* fp[kMaxPartials];             // partial-sum buffer per later row
  int fload[kMaxPartials];         // 1 if that buffer already holds a partial sum
  double am[kMaxPartials];         // alpha[m][k] per later row
  double bm[kMaxPartials];         // beta[m][k] per later row
  double a_self, b_self;           // coefficient of d_k, and beta[k][k]
  // CM_FINAL: replayed increments, acc += rc[j] * rep[j] (tableaus whose
  // earlier rows each depend on one L_k only, so dt L_k = d_{k+1} / beta[k][k])
  int nrep;
  const T* rep[kMaxPartials];
  double rc[kMaxPartials];
  int final_stage;
  const double* dt;          // device-resident step size
  int sx, sy, sz;
  long long plane;           // elements between state planes
  double inv_dx[3];          // 1 / kernel dx per axis
  double gamma;
  double mx[5];              // Maxwell constants: se, su, ise, isu, eigenvalue
  int lim;                   // runtime limiter id (generic instantiation)
  int nstrips, rows, nchunks;  // wave -> (strip, chunk) decomposition
  int nwaves;
  int zwaves;                // first-order 3D kernel: waves (z-planes) per workgroup
  int zshare;                // first-order 3D kernel: bit MODE set -> rows shared through LDS (4 waves)
  int zmarch;                // first-order 3D kernel: 0 y-march over z-planes; else z-march, cells per lane
};

}  // namespace hydro
