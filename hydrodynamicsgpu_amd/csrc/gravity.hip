// gravity.hip -- self-gravity of the Euler solvers (src/Solver/SelfGravitation.cpp,
// res/SelfGravitation.cl):
//   * one relaxation sweep of the potential's Poisson equation, fused with
//     the potential's boundary (SelfGravitation::potentialBoundary, mirror
//     sides copy as free-flow) that the reference applies before each sweep;
//   * one RK stage of the gravity source pass (calcGravityDeriv) on the
//     momentum and energy planes;
//   * the initial potential added into the total energy (resetState).
// The potential is one plane in the state's precision.  The reference's
// sweep updates the potential in place, so a work-item reads neighbours old
// or new depending on the order work-items run; the sweep here reads the
// previous sweep everywhere (ping-pong planes, Jacobi order), the
// order-independent form of the same update.
#include "device_common.hpp"
#include "launch.hpp"
#include "stage_common.hpp"
#include "aux_device.hpp"

namespace hydro {

// cell -> the cell whose value its ghost takes under the potential's
// boundary (the composed x, y, z maps of res/Common.cl with the gravity
// boundary kinds); returns false for interior cells and for ghosts no side
// writes
HYDRO_DEV bool potential_ghost_source(const GravityArgs& g, int x, int y, int z, long long& src) {
  int sc[3] = {x, y, z};
  const int n[3] = {g.sx, g.sy, g.sz};
  bool ghost = false, write = false;
  for (int a = 0; a < g.dim; ++a) {
    const int k = sc[a];
    if (k >= 2 && k < n[a] - 2) continue;
    ghost = true;
    const int m = g.method[a][k < 2 ? 0 : 1];
    if (m < 0) continue;
    write = true;
    sc[a] = ghost_source(m, k, n[a]);
  }
  src = sc[0] + (long long)g.sx * (sc[1] + (long long)g.sy * sc[2]);
  return ghost && write;
}

// res/SelfGravitation.cl:20-63, one sweep phi_old -> phi_new on every cell:
// interior cells relax  phi = (4 pi G rho - sum_j phi_j / dx_j^2) / diag
// over the interior neighbours j (the others count as zero), ghost cells take
// the boundary's value from phi_old.
template <typename S, int DIM>
__global__ void __launch_bounds__(256) potential_relax_kernel(GravityArgs g, const S* rho, const S* phi, S* out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long vol = (long long)g.sx * g.sy * g.sz;
  if (c >= vol) return;
  const int x = (int)(c % g.sx);
  const int y = DIM > 1 ? (int)((c / g.sx) % g.sy) : 0;
  const int z = DIM > 2 ? (int)(c / ((long long)g.sx * g.sy)) : 0;
  const int i[3] = {x, y, z};
  const int n[3] = {g.sx, g.sy, g.sz};
  bool interior = true;
#pragma unroll
  for (int a = 0; a < DIM; ++a) interior = interior && i[a] >= 2 && i[a] < n[a] - 2;
  if (!interior) {
    long long src;
    out[c] = potential_ghost_source(g, x, y, z, src) ? phi[src] : phi[c];
    return;
  }
  const long long step[3] = {1, (long long)g.sx, (long long)g.sx * g.sy};
  double skew = 0.;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const double h2 = g.kdx[a] * g.kdx[a];
    if (i[a] < n[a] - 3) skew += double(phi[c + step[a]]) / h2;
    if (i[a] > 2) skew += double(phi[c - step[a]]) / h2;
  }
  out[c] = S((g.four_pi_g * double(rho[c]) - skew) / g.diag);
}

// One RK stage of the gravity source pass on the interior cells
// (res/SelfGravitation.cl:65-102): with gradient = (phi(c+e_a) - phi(c-e_a))
// / (2 dx_a) and gravity = -gradient,
//   deriv(m_a) = -rho gravity,  deriv(E) = -sum_a rho gravity m_a,
// rho and m of the stage state.  a is the stage argument block advanced by
// one plane (momentum and energy: DIM + 1 planes); density is unchanged by
// the pass and the last stage copies it.
template <typename T, typename S, int DIM, bool KAHAN, int MODE>
__global__ void __launch_bounds__(256) gravity_stage_kernel(StageArgs<T, S> a, GravityArgs g, const S* u0, float* unewlo,
                                                            const float* u0lo, S* unew, const S* phi) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const int nx = g.sx - 4, ny = DIM > 1 ? g.sy - 4 : 1, nz = DIM > 2 ? g.sz - 4 : 1;
  const int xchunks = (nx + 63) / 64;
  const int row = wave / xchunks;
  if (row >= ny * nz) return;
  const int x = 2 + (wave % xchunks) * 64 + (int)(threadIdx.x & 63);
  if (x > g.sx - 3) return;
  const int y = DIM > 1 ? 2 + row % ny : 0;
  const int z = DIM > 2 ? 2 + row / ny : 0;
  const long long e = ((long long)z * g.sy + y) * g.sx;
  const long long c = e + x;
  const long long step[3] = {1, (long long)g.sx, (long long)g.sx * g.sy};
  const double dtd = *a.dt;
  constexpr int NC = DIM + 1;
  const StageCombine<T, S, KAHAN, NC, MODE> comb(a, dtd);
  const T density = T(u0[c]);
  // the stage state's momentum: u0 + d_k
  T dk[NC];
  comb.load_dk(e, (uint32_t)x, dk);
  T deriv[NC];
  T derivE = T(0);
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    T m = T(u0[(1 + k) * g.plane + c]);
    if constexpr (StageCombine<T, S, KAHAN, NC, MODE>::kReadsDk) m = m + dk[k];
    const T gradient = (T(phi[c + step[k]]) - T(phi[c - step[k]])) * T(g.inv_2dx[k]);
    const T gravity = -gradient;
    deriv[k] = -(density * gravity);
    derivE -= density * gravity * m;
  }
  deriv[DIM] = derivE;
  comb.finish(e, (uint32_t)x, deriv);
  constexpr bool kFinal = MODE == CM_FINAL || MODE == CM_SINGLE;
  if (kFinal || (MODE == CM_GENERIC && a.final_stage)) {
    unew[c] = u0[c];
    if constexpr (KAHAN) unewlo[c] = u0lo[c];
  }
}

// the initial potential added into the total energy of every cell
// (src/Solver/SelfGravitation.cpp:89-94); compensated storage adds it to the
// (high, low) pair
template <typename S>
__global__ void __launch_bounds__(256) add_potential_energy_kernel(long long vol, S* E, float* lo, const S* phi) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= vol) return;
  if (lo) {
    const float hi = float(E[c]);
    const float yv = float(phi[c]) + lo[c];
    const float t = hi + yv;
    E[c] = S(t);
    lo[c] = yv - (t - hi);
  } else {
    E[c] = E[c] + phi[c];
  }
}

// dst = sign * src (sign +1 or -1: exact)
template <typename S>
__global__ void __launch_bounds__(256) copy_plane_kernel(long long n, const S* src, S* dst, S sign) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) dst[c] = sign * src[c];
}

static dim3 cell_blocks(long long n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename S>
static void relax_launch(const GravityArgs& g, const S* rho, const S* phi, S* out, hipStream_t st) {
  const long long vol = (long long)g.sx * g.sy * g.sz;
  if (g.dim == 1) hipLaunchKernelGGL((potential_relax_kernel<S, 1>), cell_blocks(vol), dim3(256), 0, st, g, rho, phi, out);
  else if (g.dim == 2) hipLaunchKernelGGL((potential_relax_kernel<S, 2>), cell_blocks(vol), dim3(256), 0, st, g, rho, phi, out);
  else hipLaunchKernelGGL((potential_relax_kernel<S, 3>), cell_blocks(vol), dim3(256), 0, st, g, rho, phi, out);
}
void launch_potential_relax_f64(const GravityArgs& g, const double* rho, const double* phi, double* out, hipStream_t st) {
  relax_launch<double>(g, rho, phi, out, st);
}
void launch_potential_relax_f32(const GravityArgs& g, const float* rho, const float* phi, float* out, hipStream_t st) {
  relax_launch<float>(g, rho, phi, out, st);
}

template <typename T, bool KAHAN, int DIM>
static void gstage_dim(int mode, const StageArgs<T, T>& a, const GravityArgs& g, const T* u0, const float* u0lo,
                       T* unew, float* unewlo, const T* phi, hipStream_t st) {
  const int nx = g.sx - 4, ny = DIM > 1 ? g.sy - 4 : 1, nz = DIM > 2 ? g.sz - 4 : 1;
  const long long waves = (long long)((nx + 63) / 64) * ny * nz;
  const dim3 grid((unsigned)((waves + 3) / 4)), blk(256);
#define HYDRO_GST(MODE)                                                                                      \
  hipLaunchKernelGGL((gravity_stage_kernel<T, T, DIM, KAHAN, MODE>), grid, blk, 0, st, a, g, u0, unewlo, u0lo, \
                     unew, phi)
  switch (mode) {
  case CM_FIRST: HYDRO_GST(CM_FIRST); break;
  case CM_MID: HYDRO_GST(CM_MID); break;
  case CM_FINAL: HYDRO_GST(CM_FINAL); break;
  case CM_SINGLE: HYDRO_GST(CM_SINGLE); break;
  default: HYDRO_GST(CM_GENERIC); break;
  }
#undef HYDRO_GST
}
template <typename T, bool KAHAN>
static void gstage_launch(int mode, const StageArgs<T, T>& a, const GravityArgs& g, const T* u0, const float* u0lo,
                          T* unew, float* unewlo, const T* phi, hipStream_t st) {
  if (g.dim == 1) gstage_dim<T, KAHAN, 1>(mode, a, g, u0, u0lo, unew, unewlo, phi, st);
  else if (g.dim == 2) gstage_dim<T, KAHAN, 2>(mode, a, g, u0, u0lo, unew, unewlo, phi, st);
  else gstage_dim<T, KAHAN, 3>(mode, a, g, u0, u0lo, unew, unewlo, phi, st);
}
void launch_gravity_stage_f64(int mode, const StageArgs<double, double>& a, const GravityArgs& g, const double* u0,
                              double* unew, const double* phi, hipStream_t st) {
  gstage_launch<double, false>(mode, a, g, u0, nullptr, unew, nullptr, phi, st);
}
void launch_gravity_stage_f32(int mode, const StageArgs<float, float>& a, const GravityArgs& g, const float* u0,
                              const float* u0lo, float* unew, float* unewlo, const float* phi, hipStream_t st) {
  if (u0lo) gstage_launch<float, true>(mode, a, g, u0, u0lo, unew, unewlo, phi, st);
  else gstage_launch<float, false>(mode, a, g, u0, nullptr, unew, nullptr, phi, st);
}

void launch_add_potential_energy_f64(long long vol, double* E, const double* phi, hipStream_t st) {
  hipLaunchKernelGGL((add_potential_energy_kernel<double>), cell_blocks(vol), dim3(256), 0, st, vol, E,
                     (float*)nullptr, phi);
}
void launch_add_potential_energy_f32(long long vol, float* E, float* lo, const float* phi, hipStream_t st) {
  hipLaunchKernelGGL((add_potential_energy_kernel<float>), cell_blocks(vol), dim3(256), 0, st, vol, E, lo, phi);
}
void launch_copy_plane_f64(long long n, const double* src, double* dst, bool negate, hipStream_t st) {
  hipLaunchKernelGGL((copy_plane_kernel<double>), cell_blocks(n), dim3(256), 0, st, n, src, dst, negate ? -1.0 : 1.0);
}
void launch_copy_plane_f32(long long n, const float* src, float* dst, bool negate, hipStream_t st) {
  hipLaunchKernelGGL((copy_plane_kernel<float>), cell_blocks(n), dim3(256), 0, st, n, src, dst, negate ? -1.f : 1.f);
}

}  // namespace hydro
