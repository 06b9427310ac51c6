// burgers.hip -- the passes of the EulerBurgers step that are not flux
// divergences (src/Solver/EulerBurgers.cpp:61-100, res/EulerBurgers.cl):
//   * the CFL pass (calcCellTimestep: dx / (c + |v|) per axis, the smallest,
//     per-block minima for the shared dt tail);
//   * the pressure of a state (computePressure) on cells 1 .. size-2;
//   * one RK stage of the momentum pass (diffuseMomentum on the pressure of
//     the stage state), which also leaves the pressure of the next stage
//     state, so every stage reads one pressure plane;
//   * the work pass (diffuseWork), one in-place update of the energy.
// The advection pass runs the fused stage kernels with BurgersPolicy.
//
// Layout and combination are those of the fused stage kernels: state planes,
// stage increments d_k over the pass's start state u0, Shu-Osher rows in
// increment form.  Only the momentum planes change in the momentum pass, so
// its stages combine DIM planes (the argument block is shifted by one plane:
// plane 0 of the combination is the x-momentum).  One thread per cell; a
// wavefront covers 64 consecutive cells of one row, so a row's element
// offset is wave-uniform (buffer-resource addressing of StageCombine).
#include "device_common.hpp"
#include "launch.hpp"
#include "stage_common.hpp"
#include "aux_device.hpp"

namespace hydro {

// computePressure of one cell (res/EulerBurgers.cl:311-329):
// (gamma - 1) rho (E/rho - |v|^2 / 2 - phi)
template <int DIM, typename T>
HYDRO_DEV T burgers_pressure(T rho, const T m[DIM], T E, T gamma, T phi) {
  T vsq = T(0);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const T v = tdiv(m[a], rho);
    vsq += v * v;
  }
  const T eint = tdiv(E, rho) - T(.5) * vsq - phi;
  return (gamma - T(1)) * rho * eint;
}

// the potential of a cell (zero without one)
template <typename S>
HYDRO_DEV S pot_at(const void* pot, long long c) {
  return pot ? static_cast<const S*>(pot)[c] : S(0);
}

// the solid mask of a cell (0 without one): bit 0 the cell itself, bits
// 1+2a / 2+2a its lower / upper neighbour along axis a
template <typename S>
HYDRO_DEV int mask_at(const void* solid, long long c) {
  return solid ? int(static_cast<const S*>(solid)[c]) : 0;
}

// ---------------------------------------------------------------- CFL
// res/EulerBurgers.cl:19-79: cells 2 .. size-3, dx / (c + |v|) with
// c = sqrt(gamma (gamma - 1) e_int), the smallest over the axes (fmin).
template <typename S, int DIM>
__global__ void __launch_bounds__(256) burgers_dt_kernel(DtArgs a, const S* u) {
  __shared__ double sh[4];
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int nx = sx - 4;
  const int ny = DIM > 1 ? sy - 4 : 1;
  const int nz = DIM > 2 ? sz - 4 : 1;
  const long long total = (long long)nx * ny * nz;
  const S gamma = S(a.gamma);
  double m = INFINITY;
  bool bad = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = 2 + (int)(i % nx);
    const int y = DIM > 1 ? 2 + (int)((i / nx) % ny) : 0;
    const int z = DIM > 2 ? 2 + (int)(i / ((long long)nx * ny)) : 0;
    const long long cell = x + (long long)sx * (y + (long long)sy * z);
    if (mask_at<S>(a.solid, cell) & 1) continue;  // res/EulerBurgers.cl:42-47: INFINITY
    const S rho = u[cell];
    const S E = u[(DIM + 1) * a.plane + cell];
    S v[DIM];
    S vsq = S(0);
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      v[k] = tdiv(u[(1 + k) * a.plane + cell], rho);
      vsq += v[k] * v[k];
    }
    const S eint = tdiv(E, rho) - S(.5) * vsq - pot_at<S>(a.pot, cell);
    const S cs = tsqrt(gamma * (gamma - S(1)) * eint);
    double r = a.kdx[0] / double(cs + tabs(v[0]));
    bad |= !(r == r);
#pragma unroll
    for (int k = 1; k < DIM; ++k) {
      const double d = a.kdx[k] / double(cs + tabs(v[k]));
      bad |= !(d == d);
      r = fmin(r, d);
    }
    m = fmin(m, r);
  }
  m = block_min(m, sh);
  __shared__ int anybad;
  if (threadIdx.x == 0) anybad = 0;
  __syncthreads();
  if (bad) atomicOr(&anybad, 1);
  __syncthreads();
  if (threadIdx.x == 0) a.blockmin[blockIdx.x] = dt_flag(m, anybad != 0);
}

void launch_dt_burgers_f64(const DtArgs& a, const double* u, hipStream_t st) {
  dim3 g(a.nblocks), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((burgers_dt_kernel<double, 1>), g, b, 0, st, a, u);
  else if (a.dim == 2) hipLaunchKernelGGL((burgers_dt_kernel<double, 2>), g, b, 0, st, a, u);
  else hipLaunchKernelGGL((burgers_dt_kernel<double, 3>), g, b, 0, st, a, u);
}
void launch_dt_burgers_f32(const DtArgs& a, const float* u, hipStream_t st) {
  dim3 g(a.nblocks), b(256);
  if (a.dim == 1) hipLaunchKernelGGL((burgers_dt_kernel<float, 1>), g, b, 0, st, a, u);
  else if (a.dim == 2) hipLaunchKernelGGL((burgers_dt_kernel<float, 2>), g, b, 0, st, a, u);
  else hipLaunchKernelGGL((burgers_dt_kernel<float, 3>), g, b, 0, st, a, u);
}

// Thread -> cell of the row-wise kernels below: cells lo .. size-1-lo on
// every used axis, 64 consecutive x per wavefront.
struct CellMap {
  int lo;         // first coordinate on each axis (1: pressure, 2: interior)
  int nx, ny, nz; // cells per axis
  int xchunks;    // wavefronts per row
};
template <int DIM>
HYDRO_DEV bool map_cell(const BurgersArgs& b, const CellMap& m, int& x, int& y, int& z, long long& e) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const int xc = wave % m.xchunks;
  const int row = wave / m.xchunks;
  if (row >= m.ny * m.nz) return false;
  y = DIM > 1 ? m.lo + row % m.ny : 0;
  z = DIM > 2 ? m.lo + row / m.ny : 0;
  e = ((long long)z * b.sy + y) * b.sx;
  x = m.lo + xc * 64 + (int)(threadIdx.x & 63);
  return true;
}
inline CellMap host_cell_map(const BurgersArgs& b, int lo) {
  CellMap m;
  m.lo = lo;
  m.nx = b.sx - 2 * lo;
  m.ny = b.dim > 1 ? b.sy - 2 * lo : 1;
  m.nz = b.dim > 2 ? b.sz - 2 * lo : 1;
  m.xchunks = (m.nx + 63) / 64;
  return m;
}
inline unsigned map_blocks(const CellMap& m) {
  const long long waves = (long long)m.xchunks * m.ny * m.nz;
  return (unsigned)((waves + 3) / 4);
}

// ---------------------------------------------------------------- pressure
// computePressure on cells 1 .. size-2 of every used axis into p0; cells with
// a coordinate 1 or size-2 (whose state no pass stage changes) into p1 as
// well, so the two pressure planes the stages alternate between agree there.
template <typename S, int DIM>
__global__ void __launch_bounds__(256) burgers_pressure_kernel(BurgersArgs b, CellMap m, const S* u, S* p0, S* p1) {
  int x, y, z;
  long long e;
  if (!map_cell<DIM>(b, m, x, y, z, e)) return;
  if (x > b.sx - 2) return;
  const long long c = e + x;
  S mm[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) mm[a] = u[(1 + a) * b.plane + c];
  const S p = burgers_pressure<DIM, S>(u[c], mm, u[(DIM + 1) * b.plane + c], S(b.gamma), pot_at<S>(b.pot, c));
  p0[c] = p;
  bool edge = x == 1 || x == b.sx - 2;
  if (DIM > 1) edge = edge || y == 1 || y == b.sy - 2;
  if (DIM > 2) edge = edge || z == 1 || z == b.sz - 2;
  if (edge) p1[c] = p;
}

// ---------------------------------------------------------------- momentum
// One RK stage of the momentum pass on the interior cells:
//   deriv(m_a) = -(.5 (p(c + e_a) - p(c - e_a))) / dx_a    (res/EulerBurgers.cl:350-406)
// with p the stage state's pressure (pin), combined as the fused stage
// kernels combine.  a is the stage argument block shifted by one plane.  A
// stage that is not the last leaves the pressure of the next stage state
// u0 + d_{k+1} in pout; the last copies density and energy (which the pass
// leaves unchanged) into the new state.
template <typename T, typename S, int DIM, bool KAHAN, int MODE>
__global__ void __launch_bounds__(256) burgers_momentum_kernel(StageArgs<T, S> a, BurgersArgs b, CellMap m,
                                                               const T* pin, T* pout) {
  int x, y, z;
  long long e;
  if (!map_cell<DIM>(b, m, x, y, z, e)) return;
  if (x > b.sx - 3) return;
  const long long c = e + x;
  const long long step[3] = {1, (long long)b.sx, (long long)b.sx * b.sy};
  const double dtd = *a.dt;
  const StageCombine<T, S, KAHAN, DIM, MODE> comb(a, dtd);
  T deriv[DIM];
  // solid cells (res/EulerBurgers.cl:372-400): none for a solid cell; a solid
  // neighbour reads as this cell's pressure
  const int ms = mask_at<S>(b.solid, c);
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const T pL = (ms >> (1 + 2 * k)) & 1 ? pin[c] : pin[c - step[k]];
    const T pR = (ms >> (2 + 2 * k)) & 1 ? pin[c] : pin[c + step[k]];
    deriv[k] = (ms & 1) ? T(0) : -(T(.5) * (pR - pL)) * T(b.inv_dx[k]);
  }
  T dn[DIM];  // d_{k+1} of this cell, as stored
  comb.finish(e, (uint32_t)x, deriv, dn);
  constexpr bool kFinal = MODE == CM_FINAL || MODE == CM_SINGLE;
  const bool final_stage = kFinal || (MODE == CM_GENERIC && a.final_stage);
  const S* u0 = static_cast<const S*>(b.u0);
  if (!final_stage) {
    // the next stage state's momentum: u0 + d_{k+1}
    T mm[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) mm[k] = T(u0[(1 + k) * b.plane + c]) + dn[k];
    pout[c] = burgers_pressure<DIM, T>(T(u0[c]), mm, T(u0[(DIM + 1) * b.plane + c]), T(b.gamma),
                                       T(pot_at<S>(b.pot, c)));
  } else {
    S* un = static_cast<S*>(b.unew);
    un[c] = u0[c];
    un[(DIM + 1) * b.plane + c] = u0[(DIM + 1) * b.plane + c];
    if constexpr (KAHAN) {
      b.unewlo[c] = b.u0lo[c];
      b.unewlo[(DIM + 1) * b.plane + c] = b.u0lo[(DIM + 1) * b.plane + c];
    }
  }
}

// ---------------------------------------------------------------- work
// The work pass (res/EulerBurgers.cl:408-489) on the interior cells, in place:
//   W = -sum_a .5 (p(c+e_a) v_a(c+e_a) - p(c-e_a) v_a(c-e_a)) / dx_a
// with velocities of the current state and pressures of the last momentum
// stage state.  Density and momentum do not change during the pass, so W is
// the same in every stage and the pass is E += (c_tab dt) W, c_tab the sum
// the tableau's rows give a constant derivative (1 for every consistent
// tableau).  Each cell writes only its own energy and reads only its
// neighbours' density and momentum: in place is safe.
template <typename T, typename S, int DIM, bool KAHAN>
__global__ void __launch_bounds__(256) burgers_work_kernel(BurgersArgs b, CellMap m, S* u, float* lo, const T* p) {
  int x, y, z;
  long long e;
  if (!map_cell<DIM>(b, m, x, y, z, e)) return;
  if (x > b.sx - 3) return;
  const long long c = e + x;
  const long long step[3] = {1, (long long)b.sx, (long long)b.sx * b.sy};
  // solid cells (res/EulerBurgers.cl:425-447): none for a solid cell; a solid
  // neighbour's own velocity negated, and this cell's pressure
  const int ms = mask_at<S>(b.solid, c);
  if (ms & 1) return;
  T W = T(0);
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const long long l = c - step[k], r = c + step[k];
    const bool sl = (ms >> (1 + 2 * k)) & 1, sr = (ms >> (2 + 2 * k)) & 1;
    T vL = tdiv(T(u[(1 + k) * b.plane + l]), T(u[l]));
    T vR = tdiv(T(u[(1 + k) * b.plane + r]), T(u[r]));
    if (sl) vL = -vL;
    if (sr) vR = -vR;
    const T pL = sl ? p[c] : p[l];
    const T pR = sr ? p[c] : p[r];
    W -= (T(.5) * (pR * vR - pL * vL)) * T(b.inv_dx[k]);
  }
  const T acc = T(b.work_coef * *b.dt) * W;
  const long long ie = (DIM + 1) * b.plane + c;
  if constexpr (KAHAN) {
    const float hi = float(u[ie]);
    const float yv = float(acc) + lo[ie];
    const float t = hi + yv;
    u[ie] = S(t);
    lo[ie] = yv - (t - hi);
  } else {
    u[ie] = u[ie] + S(acc);
  }
}

// ---------------------------------------------------------------- launchers
template <typename S>
static void pressure_launch(const BurgersArgs& b, const S* u, S* p0, S* p1, hipStream_t st) {
  const CellMap m = host_cell_map(b, 1);
  const dim3 g(map_blocks(m)), blk(256);
  if (b.dim == 1) hipLaunchKernelGGL((burgers_pressure_kernel<S, 1>), g, blk, 0, st, b, m, u, p0, p1);
  else if (b.dim == 2) hipLaunchKernelGGL((burgers_pressure_kernel<S, 2>), g, blk, 0, st, b, m, u, p0, p1);
  else hipLaunchKernelGGL((burgers_pressure_kernel<S, 3>), g, blk, 0, st, b, m, u, p0, p1);
}
void launch_burgers_pressure_f64(const BurgersArgs& b, const double* u, double* p0, double* p1, hipStream_t st) {
  pressure_launch<double>(b, u, p0, p1, st);
}
void launch_burgers_pressure_f32(const BurgersArgs& b, const float* u, float* p0, float* p1, hipStream_t st) {
  pressure_launch<float>(b, u, p0, p1, st);
}

template <typename T, bool KAHAN, int DIM>
static void momentum_dim(int mode, const StageArgs<T, T>& a, const BurgersArgs& b, const T* pin, T* pout,
                         hipStream_t st) {
  const CellMap m = host_cell_map(b, 2);
  const dim3 g(map_blocks(m)), blk(256);
#define HYDRO_MOM(MODE) \
  hipLaunchKernelGGL((burgers_momentum_kernel<T, T, DIM, KAHAN, MODE>), g, blk, 0, st, a, b, m, pin, pout)
  switch (mode) {
  case CM_FIRST: HYDRO_MOM(CM_FIRST); break;
  case CM_MID: HYDRO_MOM(CM_MID); break;
  case CM_FINAL: HYDRO_MOM(CM_FINAL); break;
  case CM_SINGLE: HYDRO_MOM(CM_SINGLE); break;
  default: HYDRO_MOM(CM_GENERIC); break;
  }
#undef HYDRO_MOM
}
template <typename T, bool KAHAN>
static void momentum_launch(int mode, const StageArgs<T, T>& a, const BurgersArgs& b, const T* pin, T* pout,
                            hipStream_t st) {
  if (b.dim == 1) momentum_dim<T, KAHAN, 1>(mode, a, b, pin, pout, st);
  else if (b.dim == 2) momentum_dim<T, KAHAN, 2>(mode, a, b, pin, pout, st);
  else momentum_dim<T, KAHAN, 3>(mode, a, b, pin, pout, st);
}
void launch_burgers_momentum_f64(int mode, const StageArgs<double, double>& a, const BurgersArgs& b,
                                 const double* pin, double* pout, hipStream_t st) {
  momentum_launch<double, false>(mode, a, b, pin, pout, st);
}
void launch_burgers_momentum_f32(int mode, bool kahan, const StageArgs<float, float>& a, const BurgersArgs& b,
                                 const float* pin, float* pout, hipStream_t st) {
  if (kahan) momentum_launch<float, true>(mode, a, b, pin, pout, st);
  else momentum_launch<float, false>(mode, a, b, pin, pout, st);
}

template <typename T, bool KAHAN>
static void work_launch(const BurgersArgs& b, T* u, float* lo, const T* p, hipStream_t st) {
  const CellMap m = host_cell_map(b, 2);
  const dim3 g(map_blocks(m)), blk(256);
  if (b.dim == 1) hipLaunchKernelGGL((burgers_work_kernel<T, T, 1, KAHAN>), g, blk, 0, st, b, m, u, lo, p);
  else if (b.dim == 2) hipLaunchKernelGGL((burgers_work_kernel<T, T, 2, KAHAN>), g, blk, 0, st, b, m, u, lo, p);
  else hipLaunchKernelGGL((burgers_work_kernel<T, T, 3, KAHAN>), g, blk, 0, st, b, m, u, lo, p);
}
void launch_burgers_work_f64(const BurgersArgs& b, double* u, const double* p, hipStream_t st) {
  work_launch<double, false>(b, u, nullptr, p, st);
}
void launch_burgers_work_f32(const BurgersArgs& b, float* u, float* lo, const float* p, hipStream_t st) {
  if (lo) work_launch<float, true>(b, u, lo, p, st);
  else work_launch<float, false>(b, u, nullptr, p, st);
}

}  // namespace hydro
