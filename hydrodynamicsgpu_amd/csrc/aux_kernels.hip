// aux_kernels.hip -- ghost fill, CFL reduction, HLL wave speeds, totals.
#include <type_traits>

#include "device_common.hpp"
#include "phys_euler.hpp"
#include "phys_srhd.hpp"
#include "policies.hpp"
#include "launch.hpp"
#include "stage_common.hpp"
#include "aux_device.hpp"

namespace hydro {

template <typename S>
__global__ void __launch_bounds__(256) boundary_kernel(BoundaryArgs b, S* planes, float* lo, S* prim) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ghost_count(b.dim, b.sx, b.sy, b.sz)) return;
  boundary_cell<S>(b, planes, lo, g);
  // SRHDRoe::boundary fills the primitives with the same maps
  // (src/Solver/SRHDRoe.cpp:79-96); same launch, independent planes
  if (prim) boundary_cell<S>(b, prim, nullptr, g);
}

template <typename S>
__global__ void copy_ghosts_kernel(int dim, int sx, int sy, int sz, long long plane, int n, const S* src,
                                   S* dst) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ghost_count(dim, sx, sy, sz)) return;
  int x, y, z;
  if (!ghost_cell((unsigned)g, dim, sx, sy, sz, x, y, z)) return;
  const long long cell = x + (long long)sx * (y + (long long)sy * z);
  for (int s = 0; s < n; ++s) dst[s * plane + cell] = src[s * plane + cell];
}

static long long host_ghost_count(int dim, int sx, int sy, int sz) {
  if (dim == 1) return 4;
  if (dim == 2) return 4LL * sx + 4LL * (sy - 4);
  return 4LL * sx * sy + 4LL * sx * (sz - 4) + 4LL * (sy - 4) * (sz - 4);
}

void launch_boundary_f64(const BoundaryArgs& b, double* planes, double* prim, hipStream_t st) {
  const long long g = host_ghost_count(b.dim, b.sx, b.sy, b.sz);
  hipLaunchKernelGGL((boundary_kernel<double>), dim3((unsigned)((g + 255) / 256)), dim3(256), 0, st, b, planes,
                     (float*)nullptr, prim);
}
void launch_boundary_f32(const BoundaryArgs& b, float* planes, float* lo, float* prim, hipStream_t st) {
  const long long g = host_ghost_count(b.dim, b.sx, b.sy, b.sz);
  hipLaunchKernelGGL((boundary_kernel<float>), dim3((unsigned)((g + 255) / 256)), dim3(256), 0, st, b, planes, lo,
                     prim);
}

void launch_copy_ghosts_f64(int dim, int sx, int sy, int sz, long long plane, int n, const double* src,
                            double* dst, hipStream_t st) {
  const long long g = host_ghost_count(dim, sx, sy, sz);
  hipLaunchKernelGGL((copy_ghosts_kernel<double>), dim3((unsigned)((g + 255) / 256)), dim3(256), 0, st, dim,
                     sx, sy, sz, plane, n, src, dst);
}
void launch_copy_ghosts_f32(int dim, int sx, int sy, int sz, long long plane, int n, const float* src,
                            float* dst, hipStream_t st) {
  const long long g = host_ghost_count(dim, sx, sy, sz);
  hipLaunchKernelGGL((copy_ghosts_kernel<float>), dim3((unsigned)((g + 255) / 256)), dim3(256), 0, st, dim,
                     sx, sy, sz, plane, n, src, dst);
}

// ---------------------------------------------------------------- CFL
template <typename S, int DIM>
HYDRO_DEV void load_q(const S* u, long long plane, long long cell, S q[DIM + 2]) {
#pragma unroll
  for (int s = 0; s < DIM + 2; ++s) q[s] = u[s * plane + cell];
}

// Roe: res/Roe.cl:4-65 on the frozen basis (lambda_max of the left face,
// lambda_min of the right face, |.| + 1e-9), interior cells only.
// HLL/HLLC: res/EulerHLL.cl:142-175 on the frozen wave speeds of
// res/EulerHLL.cl:17-115 / res/EulerHLLC.cl:15-108, cells 2..S-2; the right
// face of cell S-2 is never written by the reference and reads as 0.
// an SRHD primitive of the block-minima CFL kernel, FP64 or FP32 planes
HYDRO_DEV double prim_at(const DtArgs& a, long long i) {
  return a.prim_f32 ? double(((const float*)a.prim)[i]) : ((const double*)a.prim)[i];
}

template <typename S, int DIM>
__global__ void __launch_bounds__(256) dt_kernel(DtArgs a, const S* u0) {
  __shared__ double sh[4];
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int hi = (a.solver == 0) ? 3 : 2;  // cells 2 .. S-hi
  const int nx = sx - 2 - hi + 1;
  const int ny = DIM > 1 ? sy - 2 - hi + 1 : 1;
  const int nz = DIM > 2 ? sz - 2 - hi + 1 : 1;
  const long long total = (long long)nx * ny * nz;
  const S gamma = S(a.gamma);
  const long long step[3] = {1, (long long)sx, (long long)sx * sy};
  const int cs[3] = {sx, sy, sz};
  double m = INFINITY;
  bool bad = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = 2 + (int)(i % nx);
    const int y = DIM > 1 ? 2 + (int)((i / nx) % ny) : 0;
    const int z = DIM > 2 ? 2 + (int)(i / ((long long)nx * ny)) : 0;
    const int ci[3] = {x, y, z};
    const long long cell = x + (long long)sx * (y + (long long)sy * z);
    S q[DIM + 2], qm[DIM + 2], qp[DIM + 2];
    load_q<S, DIM>(u0, a.plane, cell, q);
    double r = INFINITY;
#pragma unroll
    for (int side = 0; side < DIM; ++side) {
      const bool right_ok = ci[side] + 1 <= cs[side] - 2;
      load_q<S, DIM>(u0, a.plane, cell - step[side], qm);
      load_q<S, DIM>(u0, a.plane, right_ok ? cell + step[side] : cell, qp);
      double d;
      if (a.eqn == 1) {
        // SRHD Roe: Marti-Mueller speeds of the mean primitives (res/SRHDRoe.cl:76-118)
        using P = SRHDRoePolicy<DIM, double, 0>;
        double pc[DIM + 2], pm[DIM + 2], pp[DIM + 2];
#pragma unroll
        for (int s = 0; s < DIM + 2; ++s) {
          pc[s] = prim_at(a, s * a.plane + cell);
          pm[s] = prim_at(a, s * a.plane + cell - step[side]);
          pp[s] = prim_at(a, s * a.plane + cell + step[side]);
        }
        const auto fc = P::template frozen<double>(pc, a.gamma);
        const auto mL = P::mean(P::template frozen<double>(pm, a.gamma), fc, side, a.gamma);
        const auto mR = P::mean(fc, P::template frozen<double>(pp, a.gamma), side, a.gamma);
        bad |= !(mL.lmax == mL.lmax) || !(mR.lmin == mR.lmin);
        const double lmax = fmax(0.0, mL.lmax);
        const double lmin = fmin(0.0, mR.lmin);
        d = a.kdx[side] / (fabs(lmax - lmin) + 1e-9);
      } else if (a.eqn == 2) {
        // Maxwell: constant speeds +-1/sqrt(eps mu) (res/MaxwellRoe.cl:139-146)
        d = a.kdx[side] / (fabs(fmax(0.0, a.maxwell_lambda) - fmin(0.0, -a.maxwell_lambda)) + 1e-9);
      } else if (a.solver == 0 && a.solid) {
        // solid cells (res/Roe.cl:38-44): no limit across a side touching a
        // solid cell; a solid lower neighbour enters the lower face mirrored
        // (res/EulerRoe.cl:48-65): this cell's density and velocity, velocity
        // component side+1 negated, its own energy and potential
        using P = EulerRoe<DIM, S, true>;
        const S* pt = static_cast<const S*>(a.pot);
        const int ms = int(static_cast<const S*>(a.solid)[cell]);
        if ((ms & 1) || ((ms >> (2 + 2 * side)) & 1)) {
          d = INFINITY;
        } else {
          const long long cm = cell - step[side], cp = cell + step[side];
          const auto cc = P::template cell<S>(q, gamma, pt[cell]);
          typename P::Cell cl;
          if ((ms >> (1 + 2 * side)) & 1) {
            S v[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
              v[k] = q[1 + k] / q[0];
              if (k == side + 1) v[k] = -v[k];
            }
            cl = P::cell_rv(q[0], v, qm[DIM + 1], pt[cm], gamma);
          } else {
            cl = P::template cell<S>(qm, gamma, pt[cm]);
          }
          const auto fL = P::face(cl, cc, side, gamma);
          const auto fR = P::face(cc, P::template cell<S>(qp, gamma, pt[cp]), side, gamma);
          const double smax = double(fL.v[0] + fL.c), smin = double(fR.v[0] - fR.c);
          bad |= !(smax == smax) || !(smin == smin);
          const double lmax = fmax(0.0, smax);
          const double lmin = fmin(0.0, smin);
          d = a.kdx[side] / (fabs(lmax - lmin) + 1e-9);
        }
      } else if (a.solver == 0 && a.pot) {
        // with the cells' potentials (res/EulerRoe.cl:56-94)
        using P = EulerRoe<DIM, S, true>;
        const S* pt = static_cast<const S*>(a.pot);
        const long long cm = cell - step[side], cp = right_ok ? cell + step[side] : cell;
        const auto cc = P::template cell<S>(q, gamma, pt[cell]);
        const auto fL = P::face(P::template cell<S>(qm, gamma, pt[cm]), cc, side, gamma);
        const auto fR = P::face(cc, P::template cell<S>(qp, gamma, pt[cp]), side, gamma);
        const double smax = double(fL.v[0] + fL.c), smin = double(fR.v[0] - fR.c);
        bad |= !(smax == smax) || !(smin == smin);
        const double lmax = fmax(0.0, smax);
        const double lmin = fmin(0.0, smin);
        d = a.kdx[side] / (fabs(lmax - lmin) + 1e-9);
      } else if (a.solver == 0) {
        using P = EulerRoe<DIM, S>;
        const auto cc = P::template cell<S>(q, gamma);
        const auto fL = P::face(P::template cell<S>(qm, gamma), cc, side, gamma);
        const auto fR = P::face(cc, P::template cell<S>(qp, gamma), side, gamma);
        const double smax = double(fL.v[0] + fL.c), smin = double(fR.v[0] - fR.c);
        bad |= !(smax == smax) || !(smin == smin);
        const double lmax = fmax(0.0, smax);
        const double lmin = fmin(0.0, smin);
        d = a.kdx[side] / (fabs(lmax - lmin) + 1e-9);
      } else {
        S slL, srL, slR, srR;
        if (a.solver == 1 && a.pot) {
          // res/EulerHLL.cl:56-93 with the cells' potentials
          using P = HLLPolicy<DIM, S, 1, true>;
          const S* pt = static_cast<const S*>(a.pot);
          const long long cm = cell - step[side], cp = right_ok ? cell + step[side] : cell;
          const auto cc = P::template frozen_pot<S>(q, pt[cell], gamma);
          P::speeds(P::template frozen_pot<S>(qm, pt[cm], gamma), cc, side, gamma, slL, srL);
          P::speeds(cc, P::template frozen_pot<S>(qp, pt[cp], gamma), side, gamma, slR, srR);
        } else if (a.solver == 1) {
          using P = HLLPolicy<DIM, S, 1>;
          const auto cc = P::template frozen<S>(q, gamma);
          P::speeds(P::template frozen<S>(qm, gamma), cc, side, gamma, slL, srL);
          P::speeds(cc, P::template frozen<S>(qp, gamma), side, gamma, slR, srR);
        } else {
          using P = HLLPolicy<DIM, S, 2>;
          const auto cc = P::template frozen<S>(q, gamma);
          P::speeds(P::template frozen<S>(qm, gamma), cc, side, gamma, slL, srL);
          P::speeds(cc, P::template frozen<S>(qp, gamma), side, gamma, slR, srR);
        }
        bad |= !(srL == srL) || !(slR == slR);
        const double minL = fmin(0.0, right_ok ? double(slR) : 0.0);
        const double maxL = fmax(0.0, double(srL));
        d = a.kdx[side] / (maxL - minL);
      }
      bad |= !(d == d);
      r = fmin(r, d);
    }
    m = fmin(m, r);
  }
  m = block_min(m, sh);
  // a NaN in the block is flagged for the finaliser's health check
  __shared__ int anybad;
  if (threadIdx.x == 0) anybad = 0;
  __syncthreads();
  if (bad) atomicOr(&anybad, 1);
  __syncthreads();
  if (threadIdx.x == 0) a.blockmin[blockIdx.x] = dt_flag(m, anybad != 0);
}

void launch_dt_euler_f64(const DtArgs& a, const double* u0, hipStream_t st) {
  dim3 g(a.nblocks), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((dt_kernel<double, 1>), g, b, 0, st, a, u0);
  else if (a.dim == 2) hipLaunchKernelGGL((dt_kernel<double, 2>), g, b, 0, st, a, u0);
  else hipLaunchKernelGGL((dt_kernel<double, 3>), g, b, 0, st, a, u0);
}
void launch_dt_euler_f32(const DtArgs& a, const float* u0, hipStream_t st) {
  dim3 g(a.nblocks), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((dt_kernel<float, 1>), g, b, 0, st, a, u0);
  else if (a.dim == 2) hipLaunchKernelGGL((dt_kernel<float, 2>), g, b, 0, st, a, u0);
  else hipLaunchKernelGGL((dt_kernel<float, 3>), g, b, 0, st, a, u0);
}

// Solver::findMinTimestep tail (src/Solver/Solver.cpp:510-515): dt = min * cfl;
// plus fixed dt, clipping to t_end, time/step bookkeeping and health flags.
__global__ void __launch_bounds__(1024) finalize_dt_kernel(FinalizeArgs f) {
  __shared__ double sh[16];
  double m = INFINITY;
  bool nan = false;
  // eight independent loads in flight per thread per round
  for (int i0 = threadIdx.x; i0 < f.nblocks; i0 += 8 * blockDim.x) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * blockDim.x;
      v[k] = i < f.nblocks ? f.blockmin[i] : INFINITY;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) dt_merge(v[k], m, nan);
  }
  __shared__ int anynan;
  if (threadIdx.x == 0) anynan = 0;
  __syncthreads();
  if (nan) atomicOr(&anynan, 1);
  m = block_min(m, sh);
  if (threadIdx.x == 0) finalize_dt(f, m, anynan != 0);
}

void launch_finalize_dt(const FinalizeArgs& f, hipStream_t st) {
  hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, st, f);
}

template <typename S, int DIM, int SOLVER, typename L = S>
__global__ void __launch_bounds__(256) dt_march_kernel(DtMarchArgs a, const L* u0) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (wave >= a.nwaves) return;
  const double m = dt_march_wave<S, DIM, SOLVER, L>(a, u0, wave % a.nstrips, wave / a.nstrips);
  if ((threadIdx.x & 63) == 0) a.wavemin[wave] = m;
}

template <typename S>
static void dt_march(const DtMarchArgs& a, const S* u0, hipStream_t st) {
  const int blocks = (a.nwaves + 3) / 4;
  dim3 g(blocks), b(256);
#define LAUNCH(D, SV) hipLaunchKernelGGL((dt_march_kernel<S, D, SV>), g, b, 0, st, a, u0)
  if (a.dim == 1) {
    if (a.solver == 0) LAUNCH(1, 0); else if (a.solver == 1) LAUNCH(1, 1); else LAUNCH(1, 2);
  } else {
    if (a.solver == 0) LAUNCH(2, 0); else if (a.solver == 1) LAUNCH(2, 1); else LAUNCH(2, 2);
  }
#undef LAUNCH
}
void launch_dt_march_f64(const DtMarchArgs& a, const double* u0, hipStream_t st) { dt_march<double>(a, u0, st); }
void launch_dt_march_f32(const DtMarchArgs& a, const float* u0, hipStream_t st) { dt_march<float>(a, u0, st); }

// SRHD: the primitive planes (FP64, or FP32 with FP32 speeds), speeds in
// FP64 or (FP32-mixed) FP32
void launch_dt_march_srhd(const DtMarchArgs& a, const float* prim, hipStream_t st) {
  dim3 g((a.nwaves + 3) / 4), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((dt_march_kernel<float, 1, 3, float>), g, b, 0, st, a, prim);
  else hipLaunchKernelGGL((dt_march_kernel<float, 2, 3, float>), g, b, 0, st, a, prim);
}
void launch_dt_march_srhd(const DtMarchArgs& a, const double* prim, bool fp32, hipStream_t st) {
  dim3 g((a.nwaves + 3) / 4), b(256);
  if (fp32) {
    if (a.dim == 1) hipLaunchKernelGGL((dt_march_kernel<float, 1, 3, double>), g, b, 0, st, a, prim);
    else hipLaunchKernelGGL((dt_march_kernel<float, 2, 3, double>), g, b, 0, st, a, prim);
  } else {
    if (a.dim == 1) hipLaunchKernelGGL((dt_march_kernel<double, 1, 3, double>), g, b, 0, st, a, prim);
    else hipLaunchKernelGGL((dt_march_kernel<double, 2, 3, double>), g, b, 0, st, a, prim);
  }
}

// wave -> (strip, row chunk, z-chunk), z-chunks slowest
template <typename S, int SOLVER>
__global__ void __launch_bounds__(256) dt_march3_kernel(DtMarchArgs a, const S* u0) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  if (wave >= a.nwaves) return;
  const int rest = wave / a.nstrips;
  const double m = dt_march3_wave<S, SOLVER>(a, u0, wave % a.nstrips, rest % a.nchunks, rest / a.nchunks);
  if ((threadIdx.x & 63) == 0) a.wavemin[wave] = m;
}

template <typename S>
static void dt_march3(const DtMarchArgs& a, const S* u0, hipStream_t st) {
  dim3 g((a.nwaves + 3) / 4), b(256);
  if (a.solver == 0) hipLaunchKernelGGL((dt_march3_kernel<S, 0>), g, b, 0, st, a, u0);
  else if (a.solver == 1) hipLaunchKernelGGL((dt_march3_kernel<S, 1>), g, b, 0, st, a, u0);
  else hipLaunchKernelGGL((dt_march3_kernel<S, 2>), g, b, 0, st, a, u0);
}
void launch_dt_march3_f64(const DtMarchArgs& a, const double* u0, hipStream_t st) { dt_march3<double>(a, u0, st); }
void launch_dt_march3_f32(const DtMarchArgs& a, const float* u0, hipStream_t st) { dt_march3<float>(a, u0, st); }

// ---------------------------------------------------------------- step head
// boundary + CFL march in one launch (HeadArgs).  The marching workgroups
// read every ghost cell through the boundary's map, so the ghost-writing
// workgroups run beside them; each marching workgroup leaves its minimum
// for the dt tail.
template <typename S, int DIM, int SOLVER>
__global__ void __launch_bounds__(256) step_head_kernel(HeadArgs h, S* planes, float* lo, S* prim) {
  // the marching workgroups take the low block indices (dispatched first;
  // the short ghost-writing workgroups fill the slots they leave)
  const int bb = (int)blockIdx.x - h.mblocks;
  if (bb >= 0) {
    const long long g = (long long)bb * blockDim.x + threadIdx.x;
    if (g >= ghost_count(h.b.dim, h.b.sx, h.b.sy, h.b.sz)) return;
    boundary_cell<S>(h.b, planes, lo, g);
    if (prim) boundary_cell<S>(h.b, prim, nullptr, g);
    return;
  }
  // XCD-aware order (mblocks is a multiple of 8): consecutive logical blocks
  // -- neighbouring strips, whose halo columns and rows overlap -- run on
  // the same XCD and share its L2
  const int per = h.mblocks >> 3;
  const int mb = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
  const int wave = __builtin_amdgcn_readfirstlane(mb * 4 + (int)(threadIdx.x >> 6));
  double m = INFINITY;
  if (wave < h.a.nwaves) {
    if constexpr (DIM == 3) {
      const int rest = wave / h.a.nstrips;
      m = dt_march3_wave<S, SOLVER, true>(h.a, planes, wave % h.a.nstrips, rest % h.a.nchunks, rest / h.a.nchunks,
                                          &h.b);
    } else if constexpr (SOLVER == 3) {
      m = dt_march_wave<S, DIM, 3, S, true>(h.a, prim, wave % h.a.nstrips, wave / h.a.nstrips, &h.b);
    } else {
      m = dt_march_wave<S, DIM, SOLVER, S, true>(h.a, planes, wave % h.a.nstrips, wave / h.a.nstrips, &h.b);
    }
  }
  __shared__ double wmin[4];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = INFINITY;
    bool nan = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) dt_merge(wmin[i], bm, nan);
    h.a.wavemin[mb] = dt_flag(bm, nan);
  }
}

template <typename S>
static void step_head(const HeadArgs& h, S* planes, float* lo, S* prim, hipStream_t st) {
  dim3 g(h.bblocks + h.mblocks), b(256);
  const int dim = h.a.dim, sv = h.a.solver;
#define LAUNCH(D, SV) hipLaunchKernelGGL((step_head_kernel<S, D, SV>), g, b, 0, st, h, planes, lo, prim)
  if (dim == 3) {
    if (sv == 0) LAUNCH(3, 0); else if (sv == 1) LAUNCH(3, 1); else LAUNCH(3, 2);
  } else if (dim == 2) {
    if (sv == 0) LAUNCH(2, 0); else if (sv == 1) LAUNCH(2, 1); else if (sv == 2) LAUNCH(2, 2); else LAUNCH(2, 3);
  } else {
    if (sv == 0) LAUNCH(1, 0); else if (sv == 1) LAUNCH(1, 1); else if (sv == 2) LAUNCH(1, 2); else LAUNCH(1, 3);
  }
#undef LAUNCH
}
void launch_step_head(const HeadArgs& h, void* planes, float* lo, void* prim, bool fp32, hipStream_t st) {
  if (fp32) step_head<float>(h, (float*)planes, lo, (float*)prim, st);
  else step_head<double>(h, (double*)planes, nullptr, (double*)prim, st);
}

// ---------------------------------------------------------------- totals
template <typename S>
__global__ void __launch_bounds__(256) totals_kernel(int dim, int sx, int sy, int sz, long long plane, int n,
                                                     const S* u, const float* lo, double* partial) {
  __shared__ double sh[256];
  const int nx = sx - 4, ny = dim > 1 ? sy - 4 : 1, nz = dim > 2 ? sz - 4 : 1;
  const long long total = (long long)nx * ny * nz;
  for (int s = 0; s < n; ++s) {
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
      const int x = 2 + (int)(i % nx);
      const int y = dim > 1 ? 2 + (int)((i / nx) % ny) : 0;
      const int z = dim > 2 ? 2 + (int)(i / ((long long)nx * ny)) : 0;
      const long long cell = x + (long long)sx * (y + (long long)sy * z);
      double v = double(u[s * plane + cell]);
      if (lo) v += double(lo[s * plane + cell]);
      acc += v;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long long)blockIdx.x * n + s] = sh[0];
    __syncthreads();
  }
}

void launch_totals_f64(int dim, int sx, int sy, int sz, long long plane, int n, const double* u, double* partial,
                       int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((totals_kernel<double>), dim3(nblocks), dim3(256), 0, st, dim, sx, sy, sz, plane, n, u,
                     (const float*)nullptr, partial);
}
void launch_totals_f32(int dim, int sx, int sy, int sz, long long plane, int n, const float* u, const float* lo,
                       double* partial, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((totals_kernel<float>), dim3(nblocks), dim3(256), 0, st, dim, sx, sy, sz, plane, n, u, lo,
                     partial);
}

}  // namespace hydro

namespace hydro {

// ---------------------------------------------------------------- SRHD
// SRHDRoe::step tail (src/Solver/SRHDRoe.cpp:62-73): constrainState on every
// cell (res/SRHDCommon.cl:94-109), then updatePrimitives on cells 2..S-2 of
// every axis (res/SRHDCommon.cl:113-196), in FP64.
// One thread per cell: x from the block's 256-column slice, y and z from the
// grid (no index division).
template <typename S, int DIM>
__global__ void __launch_bounds__(256) con2prim_kernel(Con2PrimArgs a, S* u, float* lo, S* prim) {
  constexpr int N = DIM + 2, E = DIM + 1;
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (x >= a.sx) return;
  const int y = (int)blockIdx.y, z = (int)blockIdx.z;
  const long long cell = x + (long long)a.sx * (y + (long long)a.sy * z);
  const SRHDDefs defs{a.gamma, a.max_iter, a.stop_eps, a.vel_eps, a.pmin_eps, a.rho_min, a.rho_max,
                      a.eint_max, a.D_min, a.D_max, a.tau_min, a.tau_max};
  // ghost cells (not written by the stages) first take the start-of-step
  // values when the caller defers that copy to here
  bool ghost = x < 2 || x > a.sx - 3;
  if (DIM > 1) ghost = ghost || y < 2 || y > a.sy - 3;
  if (DIM > 2) ghost = ghost || z < 2 || z > a.sz - 3;
  const bool copy = ghost && a.ghost_src;
  const S* src = copy ? (const S*)a.ghost_src : u;
  const float* srclo = copy ? a.ghost_src_lo : lo;
  double q[N];
#pragma unroll
  for (int s = 0; s < N; ++s) {
    const S h = src[s * a.plane + cell];
    q[s] = double(h);
    float l = 0.f;
    if (lo) {
      l = srclo[s * a.plane + cell];
      q[s] += double(l);
    }
    if (copy) {
      u[s * a.plane + cell] = h;
      if (lo) lo[s * a.plane + cell] = l;
    }
  }
  const double D = fmin(fmax(q[0], a.D_min), a.D_max);
  const double tau = fmin(fmax(q[E], a.tau_min), a.tau_max);
  if (D != q[0]) {
    q[0] = D;
    u[cell] = S(D);
    if (lo) lo[cell] = float(D - double(S(D)));
  }
  if (tau != q[E]) {
    q[E] = tau;
    u[E * a.plane + cell] = S(tau);
    if (lo) lo[E * a.plane + cell] = float(tau - double(S(tau)));
  }
  bool act = x >= 2 && x <= a.sx - 2;
  if (DIM > 1) act = act && y >= 2 && y <= a.sy - 2;
  if (DIM > 2) act = act && z >= 2 && z <= a.sz - 2;
  if (!act) return;
  double p[N];
#pragma unroll
  for (int s = 0; s < N; ++s) p[s] = double(prim[s * a.plane + cell]);
  // FP32 storage: warm-started Newton (same root, see srhd_con2prim); the
  // primitives are stored in the state's precision (FP32: what the FP32
  // fluxes and speeds read)
  if (srhd_con2prim<DIM, sizeof(S) == 4>(defs, q, p)) {
#pragma unroll
    for (int s = 0; s < N; ++s) prim[s * a.plane + cell] = S(p[s]);
  } else {
    atomicOr(a.health, 4u);
  }
}

// ---------------------------------------------------------------- buffers
// The host-side integrator's buffer operations (include/HydroGPU/Integrator/
// RungeKutta.h:59-125): multAdd a = b + c * k (res/Common.cl:201-210) and a
// zero fill (Solver::CL::zero, src/Solver/Solver.cpp:22-32), as kernels so a
// captured graph holds no memset node.  Grid-stride loops over n elements.
template <typename V>
__global__ void __launch_bounds__(256) mult_add_kernel(V* a, const V* b, const V* c, V k, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) a[i] = b[i] + c[i] * k;
}
__global__ void __launch_bounds__(256) zero_words_kernel(uint32_t* p, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0u;
}
static unsigned stride_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}
void launch_mult_add_f64(double* a, const double* b, const double* c, double k, long long n, hipStream_t st) {
  hipLaunchKernelGGL((mult_add_kernel<double>), dim3(stride_blocks(n)), dim3(256), 0, st, a, b, c, k, n);
}
void launch_mult_add_f32(float* a, const float* b, const float* c, float k, long long n, hipStream_t st) {
  hipLaunchKernelGGL((mult_add_kernel<float>), dim3(stride_blocks(n)), dim3(256), 0, st, a, b, c, k, n);
}
void launch_zero_words(void* p, long long nwords, hipStream_t st) {
  hipLaunchKernelGGL(zero_words_kernel, dim3(stride_blocks(nwords)), dim3(256), 0, st, (uint32_t*)p, nwords);
}

// ---------------------------------------------------------------- layouts
// The reference's array of structures state[s + N*cell] (FP64, what
// Converter::toGPU / fromGPU move, src/Solver/Solver.cpp:333-343) to and from
// the state planes of the storage type: FP32 storage takes hi = float(v) and,
// when compensated, lo = float(v - hi); reading back gives hi + lo in FP64.
template <typename S>
__global__ void __launch_bounds__(256) aos_to_planes_kernel(const double* aos, int n, long long vol, long long plane,
                                                            S* hi, float* lo) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < vol; c += (long long)gridDim.x * 256) {
    for (int k = 0; k < n; ++k) {
      const double v = aos[k + (long long)n * c];
      const S h = S(v);
      hi[k * plane + c] = h;
      if (lo) lo[k * plane + c] = float(v - double(h));
    }
  }
}
template <typename S>
__global__ void __launch_bounds__(256) planes_to_aos_kernel(const S* hi, const float* lo, int n, long long vol,
                                                            long long plane, double* aos) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < vol; c += (long long)gridDim.x * 256) {
    for (int k = 0; k < n; ++k) {
      double v = double(hi[k * plane + c]);
      if (lo) v += double(lo[k * plane + c]);
      aos[k + (long long)n * c] = v;
    }
  }
}
void launch_aos_to_planes_f64(const double* aos, int n, long long vol, long long plane, double* u, hipStream_t st) {
  hipLaunchKernelGGL((aos_to_planes_kernel<double>), dim3(stride_blocks(vol)), dim3(256), 0, st, aos, n, vol, plane,
                     u, (float*)nullptr);
}
void launch_aos_to_planes_f32(const double* aos, int n, long long vol, long long plane, float* hi, float* lo,
                              hipStream_t st) {
  hipLaunchKernelGGL((aos_to_planes_kernel<float>), dim3(stride_blocks(vol)), dim3(256), 0, st, aos, n, vol, plane,
                     hi, lo);
}
void launch_planes_to_aos_f64(const double* u, int n, long long vol, long long plane, double* aos, hipStream_t st) {
  hipLaunchKernelGGL((planes_to_aos_kernel<double>), dim3(stride_blocks(vol)), dim3(256), 0, st, u,
                     (const float*)nullptr, n, vol, plane, aos);
}
void launch_planes_to_aos_f32(const float* hi, const float* lo, int n, long long vol, long long plane, double* aos,
                              hipStream_t st) {
  hipLaunchKernelGGL((planes_to_aos_kernel<float>), dim3(stride_blocks(vol)), dim3(256), 0, st, hi, lo, n, vol,
                     plane, aos);
}

template <typename S>
static void con2prim(const Con2PrimArgs& a, S* u, float* lo, S* prim, hipStream_t st) {
  dim3 g((a.sx + 255) / 256, a.dim > 1 ? a.sy : 1, a.dim > 2 ? a.sz : 1), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((con2prim_kernel<S, 1>), g, b, 0, st, a, u, lo, prim);
  else if (a.dim == 2) hipLaunchKernelGGL((con2prim_kernel<S, 2>), g, b, 0, st, a, u, lo, prim);
  else hipLaunchKernelGGL((con2prim_kernel<S, 3>), g, b, 0, st, a, u, lo, prim);
}
void launch_con2prim_f64(const Con2PrimArgs& a, double* u, double* prim, hipStream_t st) {
  con2prim<double>(a, u, nullptr, prim, st);
}
void launch_con2prim_f32(const Con2PrimArgs& a, float* u, float* lo, float* prim, hipStream_t st) {
  con2prim<float>(a, u, lo, prim, st);
}

}  // namespace hydro
