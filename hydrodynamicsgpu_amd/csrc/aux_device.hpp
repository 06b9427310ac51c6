// aux_device.hpp -- device functions of the per-step kernels that both the
// single-grid and the batched launches use: ghost enumeration and the
// composed boundary, the CFL strip march and the findMinTimestep tail.
#pragma once
#include <type_traits>

#include "device_common.hpp"
#include "launch.hpp"
#include "phys_euler.hpp"
#include "phys_srhd.hpp"
#include "policies.hpp"
#include "stage_common.hpp"

namespace hydro {

// ---------------------------------------------------------------- boundary
// res/Common.cl:49-194 applies, per axis in order x -> y -> z, per state, the
// min side then the max side, each pass covering every line of the grid
// (ghost lines of the other axes included).  Every source cell of those maps
// is interior along the pass's axis, so the value a ghost cell ends with is
// the composition of the per-axis index maps: for each axis on which the
// cell is a ghost, replace that coordinate by its source and multiply by -1
// for a reflect.  One thread per ghost cell evaluates that composition
// directly, so the whole boundary is one launch with no ordering between
// threads (sources are never written by the launch).
// method: 0 periodic, 1 mirror (a reflect is a mirror whose state is
// negated), 2 freeflow (chained copy: both ghosts take the first interior
// cell), -1 none.
HYDRO_DEV int ghost_source(int method, int k, int n) {
  if (k < 2) {
    if (method == 0) return n - 4 + k;   // 0 <- n-4, 1 <- n-3
    if (method == 2) return 2;           // 1 <- 2, 0 <- 1
    return 3 - k;                        // 0 <- 3, 1 <- 2
  }
  if (method == 0) return k - (n - 4);   // n-2 <- 2, n-1 <- 3
  if (method == 2) return n - 3;         // n-2 <- n-3, n-1 <- n-2
  return 2 * n - 5 - k;                  // n-1 <- n-4, n-2 <- n-3
}

// The same composition seen from a reader: coordinate k along `axis` of a
// cell whose post-boundary value is wanted becomes the coordinate that holds
// that value now, and the axis' reflect mask is folded into `neg`.  Interior
// coordinates and sides without a method (NONE: the ghost keeps its value)
// map to themselves.  A kernel that reads ghosts this way does not need the
// boundary to have run, so both can share one launch.
HYDRO_DEV int mapped_coord(const BoundaryArgs& b, int axis, int k, unsigned& neg) {
  const int n = axis == 0 ? b.sx : (axis == 1 ? b.sy : b.sz);
  if (k >= 2 && k < n - 2) return k;
  const int side = k < 2 ? 0 : 1;
  const int m = b.method[axis][side];
  if (m < 0) return k;
  neg ^= b.negmask[axis][side];
  return ghost_source(m, k, n);
}

// ---------------------------------------------------------------- ghosts
// Enumerates exactly the ghost cells (some coordinate in {0,1,S-2,S-1}) as
// coordinates; 32-bit index arithmetic (a ghost count fits in 31 bits for
// any grid that fits in memory), so no 64-bit division is emitted.
HYDRO_DEV bool ghost_cell(unsigned g, int dim, int sx, int sy, int sz, int& x, int& y, int& z) {
  auto pick = [](unsigned k, int n) { return k < 2 ? (int)k : n - 4 + (int)k; };  // k in 0..3
  const unsigned ux = (unsigned)sx, uy = (unsigned)sy, uz = (unsigned)sz;
  y = 0;
  z = 0;
  if (dim == 1) {
    if (g >= 4u) return false;
    x = pick(g, sx);
    return true;
  }
  if (dim == 2) {
    const unsigned rows = 4u * ux;
    if (g < rows) {
      y = pick(g / ux, sy);
      x = (int)(g % ux);
      return true;
    }
    g -= rows;
    if (g >= 4u * (uy - 4u)) return false;
    y = 2 + (int)(g >> 2);
    x = pick(g & 3u, sx);
    return true;
  }
  const unsigned plane = ux * uy;
  if (g < 4u * plane) {
    const unsigned k = g / plane, r = g - k * plane;
    z = pick(k, sz);
    y = (int)(r / ux);
    x = (int)(r - (unsigned)y * ux);
    return true;
  }
  unsigned h = g - 4u * plane;
  const unsigned rows = 4u * ux;
  if (h < rows * (uz - 4u)) {
    const unsigned zz = h / rows, r = h - zz * rows;
    z = 2 + (int)zz;
    const unsigned yy = r / ux;
    y = pick(yy, sy);
    x = (int)(r - yy * ux);
    return true;
  }
  h -= rows * (uz - 4u);
  const unsigned cols = 4u * (uy - 4u);
  if (h >= cols * (uz - 4u)) return false;
  const unsigned zz = h / cols, r = h - zz * cols;
  z = 2 + (int)zz;
  y = 2 + (int)(r >> 2);
  x = pick(r & 3u, sx);
  return true;
}

HYDRO_DEV long long ghost_count(int dim, int sx, int sy, int sz) {
  if (dim == 1) return 4;
  if (dim == 2) return 4LL * sx + 4LL * (sy - 4);
  return 4LL * sx * sy + 4LL * sx * (sz - 4) + 4LL * (sy - 4) * (sz - 4);
}

// ghost cell number g of a grid (g < ghost_count)
template <typename S>
HYDRO_DEV void boundary_cell(const BoundaryArgs& b, S* planes, float* lo, long long g) {
  int x, y, z;
  if (!ghost_cell((unsigned)g, b.dim, b.sx, b.sy, b.sz, x, y, z)) return;
  const long long cell = x + (long long)b.sx * (y + (long long)b.sy * z);
  int sc[3] = {x, y, z};
  const int n[3] = {b.sx, b.sy, b.sz};
  // one source cell for every state (mirror and reflect share the index
  // map); per-state signs from the reflect masks of the ghost axes
  unsigned neg = 0;
  bool write = false;
  for (int a = 0; a < b.dim; ++a) {
    const int k = sc[a];
    if (k >= 2 && k < n[a] - 2) continue;
    const int side = k < 2 ? 0 : 1;
    const int m = b.method[a][side];
    if (m < 0) continue;
    write = true;
    neg ^= b.negmask[a][side];
    sc[a] = ghost_source(m, k, n[a]);
  }
  if (!write) return;
  const long long src = sc[0] + (long long)b.sx * (sc[1] + (long long)b.sy * sc[2]);
  S v[8];
  float w[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (s < b.nstates) {
      v[s] = planes[s * b.plane + src];
      if (lo) w[s] = lo[s * b.plane + src];
    }
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (s < b.nstates) {
      const bool ng = (neg >> s) & 1u;
      planes[s * b.plane + cell] = ng ? -v[s] : v[s];
      if (lo) lo[s * b.plane + cell] = ng ? -w[s] : w[s];
    }
  }
}

// ---------------------------------------------------------------- reductions
HYDRO_DEV double block_min(double v, double* sh) {
  // wave min via shuffles, then across the block's waves through LDS
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int i = 1; i < nw; ++i) v = fmin(v, sh[i]);
  }
  return v;
}

// Solver::findMinTimestep tail (src/Solver/Solver.cpp:510-515): dt = min * cfl,
// or the fixed dt; clipping to t_end, time/step bookkeeping, health bit 0.
// With clipping, an update that starts at t >= t_end is a no-op: dt = 0 and
// the clock, the step count and the health bits stay as they are.  Returns
// false for such a no-op.  The clock is passed in registers here and kept in
// device memory by finalize_dt.
HYDRO_DEV bool finalize_step(const FinalizeArgs& f, double m, bool anynan, double& t, long long& steps,
                             unsigned& health, double& dt) {
  if (f.clip && !(t < f.t_end)) {
    dt = 0.0;
    return false;
  }
  dt = f.use_fixed ? f.fixed_dt : m * f.cfl;
  if (f.clip && t + dt > f.t_end) dt = f.t_end - t;
  if (anynan || !(dt > 0.0) || dt == INFINITY) health |= 1u;
  t = t + dt;
  steps = steps + 1;
  return true;
}

HYDRO_DEV bool finalize_dt(const FinalizeArgs& f, double m, bool anynan) {
  double t = *f.t, dt;
  long long steps = *f.steps;
  unsigned health = *f.health;
  const bool adv = finalize_step(f, m, anynan, t, steps, health, dt);
  *f.dt = dt;
  if (adv) {
    *f.health = health;
    *f.t = t;
    *f.steps = steps;
  }
  return adv;
}

// The per-cell CFL pieces of one solver: frozen data of a cell, its packed
// form for lane shifts, the (min, max) wave speed of a face, and one axis'
// rate of a cell -- the reciprocal of its dt -- from the left face's maximal
// and the right face's minimal speed.  Roe (res/Roe.cl:4-65) and SRHD Roe
// (SOLVER 3, res/SRHDRoe.cl:76-118): (|max(0, lmax) - min(0, lmin)| + 1e-9)
// / dx; HLL/HLLC: (max(0, sR) - min(0, sL)) / dx with the never-written
// right face of the last cell read as 0 (res/EulerHLL.cl:142-175).  The
// reference's min over cells of dx / (...) is 1 / max of these rates; the
// march takes one reciprocal per wave instead of one division per cell.
template <typename S, int DIM, int SOLVER>
struct DtSolver {
  using Roe = EulerRoe<DIM, S>;
  using HLL = HLLPolicy<DIM, S, SOLVER == 2 ? 2 : 1>;
  using SR = SRHDRoePolicy<DIM, S, 0>;
  using Frozen = typename std::conditional<
      SOLVER == 0, typename Roe::Cell,
      typename std::conditional<SOLVER == 3, typename SR::Frozen, typename HLL::Frozen>::type>::type;
  static constexpr bool kRoeForm = SOLVER == 0 || SOLVER == 3;
  static constexpr int N = DIM + 2;
  static constexpr int NF = SOLVER == 0 ? RoePolicy<DIM, S, 0>::NF : (SOLVER == 3 ? SR::NF : HLL::NF);

  static HYDRO_DEV Frozen frozen(const S q[N], S gamma) {
    if constexpr (SOLVER == 0) return Roe::template cell<S>(q, gamma);
    else if constexpr (SOLVER == 3) return SR::template frozen<S>(q, gamma);
    else return HLL::template frozen<S>(q, gamma);
  }
  static HYDRO_DEV void pack(const Frozen& f, S o[NF]) {
    if constexpr (SOLVER == 0) RoePolicy<DIM, S, 0>::pack(f, o);
    else if constexpr (SOLVER == 3) SR::pack(f, o);
    else HLL::pack(f, o);
  }
  static HYDRO_DEV Frozen unpack(const S o[NF]) {
    if constexpr (SOLVER == 0) return RoePolicy<DIM, S, 0>::unpack(o);
    else if constexpr (SOLVER == 3) return SR::unpack(o);
    else return HLL::unpack(o);
  }
  static HYDRO_DEV void speeds(const Frozen& L, const Frozen& R, int side, S gamma, S& lmin, S& lmax) {
    if constexpr (SOLVER == 0) {
      const auto f = Roe::face(L, R, side, gamma);
      lmin = f.v[0] - f.c;
      lmax = f.v[0] + f.c;
    } else if constexpr (SOLVER == 3) {
      const auto m = SR::mean(L, R, side, gamma);
      lmin = m.lmin;
      lmax = m.lmax;
    } else {
      HLL::speeds(L, R, side, gamma, lmin, lmax);
    }
  }
  static HYDRO_DEV double side_rate(S lmaxL, S lminR, bool right_ok, double inv_dx, bool& bad) {
    bad |= !(lmaxL == lmaxL) || (right_ok && !(lminR == lminR));
    if constexpr (kRoeForm) {
      const double lmax = fmax(0.0, double(lmaxL));
      const double lmin = fmin(0.0, double(lminR));
      return (fabs(lmax - lmin) + 1e-9) * inv_dx;
    } else {
      const double maxL = fmax(0.0, double(lmaxL));
      const double minL = fmin(0.0, right_ok ? double(lminR) : 0.0);
      return (maxL - minL) * inv_dx;
    }
  }
};

// Partial CFL minima carry "some entry was NaN" in the sign bit: the value
// is the minimum over the finite entries (>= 0, or +inf).  This port skips
// NaN entries on purpose (fmin) and raises a health bit; the reference's
// reduction (res/Common.cl:9-40) is a ternary (acc < el ? acc : el), whose
// result with a NaN entry depends on where the NaN falls in the order of
// the reduction.  The flag only feeds the health bit.  dt_flag encodes, dt_merge folds one into a running
// (minimum, flag) pair.
HYDRO_DEV double dt_flag(double dt, bool nan) { return nan ? copysign(dt, -1.0) : dt; }
HYDRO_DEV void dt_merge(double v, double& m, bool& nan) {
  nan |= signbit(v) || !(v == v);
  m = fmin(m, fabs(v));
}

// A wave's largest rate -> its smallest dt, flagged when a speed was not
// finite (the finite entries still count)
HYDRO_DEV double wave_dt(double rate, bool bad) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) rate = fmax(rate, __shfl_xor(rate, o, 64));
  return dt_flag(1.0 / rate, __any(bad));
}

// ---------------------------------------------------------------- CFL, 1D/2D
// The CFL of the Euler Roe / HLL / HLLC solvers for 1D and 2D grids as a
// strip march (the decomposition of the fused stage kernels): a wave owns 64
// columns and a chunk of rows, computes each cell's frozen data once, the
// x-face between lane-1 and lane through a DPP shift and the y-face against
// the previous row held in registers; the next row's loads are issued before
// the current row's arithmetic.  Per cell: the left face's maximal and the
// right face's minimal wave speed give dx / (|max(0, lmax) - min(0, lmin)|
// + 1e-9) for Roe (res/Roe.cl:4-65) and dx / (max(0, sR) - min(0, sL)) with
// the never-written right face of cell S-2 read as 0 for HLL/HLLC
// (res/EulerHLL.cl:142-175).  Each wave writes its minimum (NaN when a
// speed was not finite) to its own slot; finalize_dt_kernel reduces them.
// One wave's minimum over columns strip*60+2 .. strip*60+61 of rows
// 2 + chunk*rows .. (1D: the single row); NaN when a wave speed of one of
// its cells was not finite.  strip and chunk are wave-uniform.
// SOLVER 3 is SRHD Roe: `u0` is then the FP64 primitive planes and the
// speeds are the Marti-Mueller eigenvalues of the mean primitives
// (res/SRHDRoe.cl:76-118), with the Roe form of the cell dt.
// L: the element type in memory (the FP32-mixed SRHD march reads the FP64
// primitives and computes its speeds in FP32).  MAP: ghost cells are read
// through the boundary's composed map `gm` (mapped_coord), so the values are
// the post-boundary ones while the boundary itself runs in the same launch.
template <typename S, int DIM, int SOLVER, typename L = S, bool MAP = false>
HYDRO_DEV double dt_march_wave(const DtMarchArgs& a, const L* u0, int strip, int chunk,
                               const BoundaryArgs* gm = nullptr) {
  using D = DtSolver<S, DIM, SOLVER>;
  using Frozen = typename D::Frozen;
  constexpr int N = DIM + 2, NF = D::NF;
  const int lane = threadIdx.x & 63;
  const int sx = a.sx, sy = a.sy;
  const int c = strip * 60 + lane;
  const uint32_t cl = c < sx ? c : sx - 1;
  const bool out_col = lane >= 2 && lane < 62 && c <= a.last_x;
  const bool right_ok_x = c + 1 <= sx - 2;
  const S gamma = S(a.gamma);
  const double idx0 = 1.0 / a.kdx[0];
  bool bad = false;
  unsigned negc = 0;
  uint32_t cs = cl;
  if constexpr (MAP) cs = (uint32_t)mapped_coord(*gm, 0, (int)cl, negc);

  auto fetch = [&](int row, S q[N]) {
    unsigned neg = negc;
    if constexpr (MAP && DIM > 1) row = mapped_coord(*gm, 1, row, neg);
    const long long e = (long long)__builtin_amdgcn_readfirstlane(row) * sx;
#pragma unroll
    for (int s = 0; s < N; ++s) q[s] = S(ldp(u0, a.plane, s, e, cs));
    if constexpr (MAP) {
#pragma unroll
      for (int s = 0; s < N; ++s) q[s] = ((neg >> s) & 1u) ? -q[s] : q[s];
    }
  };
  auto x_rate = [&](const Frozen& f) {
    S pk[NF], pkL[NF];
    D::pack(f, pk);
#pragma unroll
    for (int i = 0; i < NF; ++i) pkL[i] = lane_from_left(pk[i]);
    S lmin, lmax;
    D::speeds(D::unpack(pkL), f, 0, gamma, lmin, lmax);  // face c: cells c-1, c
    const S lminR = lane_from_right(lmin);                // face c+1
    return D::side_rate(lmax, lminR, right_ok_x, idx0, bad);
  };

  double m = 0.0;  // largest rate of the wave's cells
  if constexpr (DIM == 1) {
    S q[N];
    fetch(0, q);
    const double d = x_rate(D::frozen(q, gamma));
    if (out_col) m = d;
  } else {
    const double idx1 = 1.0 / a.kdx[1];
    // a chunk is at most kDtRows rows: all its rows (plus one below and one
    // above) are loaded up front, so the wave waits for memory once
    const int y0 = 2 + chunk * a.rows;
    int y1 = y0 + a.rows;
    if (y1 > a.last_y + 1) y1 = a.last_y + 1;
    S q[kDtRows + 2][N];
#pragma unroll
    for (int i = 0; i < kDtRows + 2; ++i) fetch(y0 - 1 + i <= y1 ? y0 - 1 + i : y1, q[i]);
    Frozen cur = D::frozen(q[1], gamma);
    S lminY, lmaxY_prev;
    D::speeds(D::frozen(q[0], gamma), cur, 1, gamma, lminY, lmaxY_prev);  // face y0: bottom face of row y0
    double xd = x_rate(cur);
#pragma unroll
    for (int i = 2; i < kDtRows + 2; ++i) {
      const int r = y0 - 1 + i;  // face r: top face of row r-1
      if (r > y1) break;
      const Frozen prev = cur;
      cur = D::frozen(q[i], gamma);
      S lmaxY;
      D::speeds(prev, cur, 1, gamma, lminY, lmaxY);
      const double yd = D::side_rate(lmaxY_prev, lminY, r <= sy - 2, idx1, bad);
      if (out_col) m = fmax(m, fmax(xd, yd));
      if (r < y1) xd = x_rate(cur);
      lmaxY_prev = lmaxY;
    }
  }
  return wave_dt(m, bad && out_col);
}

// ---------------------------------------------------------------- CFL, 3D
// The CFL of the Euler solvers on 3D grids as a z-march: a wave owns 64
// columns (60 output columns) of R consecutive rows and marches a
// chunk of z-planes.  Per plane it loads the rows (and one halo row below
// and above), builds each cell's frozen data once, takes the x-faces through
// DPP shifts and the y-faces between consecutive rows, and the z-face
// against the previous plane's frozen data held in registers; a plane's
// rates are complete one plane later, with the minimal speed of its upper
// z-face.  Every face speed is computed once per wave (the y halo rows and
// the planes below and above the chunk once more).  Cells 2 .. last_x /
// last_y / last_z of every axis; per cell the same face speeds and side
// rates as the reference's calcCellTimestep, so the minimum is the same.
// MAP: ghost reads through the boundary's composed map, as in dt_march_wave.
template <typename S, int SOLVER, bool MAP = false>
HYDRO_DEV double dt_march3_wave(const DtMarchArgs& a, const S* u0, int strip, int ychunk, int zchunk,
                                const BoundaryArgs* gm = nullptr) {
  using D = DtSolver<S, 3, SOLVER>;
  using Frozen = typename D::Frozen;
  constexpr int N = 5, NF = D::NF, R = sizeof(S) == 4 ? kDt3Rows : kDt3RowsF64;
  const int lane = threadIdx.x & 63;
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int c = strip * 60 + lane;
  const uint32_t cl = c < sx ? c : sx - 1;
  const bool out_col = lane >= 2 && lane < 62 && c <= a.last_x;
  const bool right_ok_x = c + 1 <= sx - 2;
  const S gamma = S(a.gamma);
  const long long psz = (long long)sx * sy;
  bool bad = false;
  unsigned negc = 0;
  uint32_t cs = cl;
  if constexpr (MAP) cs = (uint32_t)mapped_coord(*gm, 0, (int)cl, negc);

  auto fetch = [&](int plane, int row, S q[N]) {
    unsigned neg = negc;
    if constexpr (MAP) {
      plane = mapped_coord(*gm, 2, plane, neg);
      row = mapped_coord(*gm, 1, row, neg);
    }
    const long long e = (long long)__builtin_amdgcn_readfirstlane(plane) * psz +
                        (long long)__builtin_amdgcn_readfirstlane(row) * sx;
#pragma unroll
    for (int s = 0; s < N; ++s) q[s] = ldp(u0, a.plane, s, e, cs);
    if constexpr (MAP) {
#pragma unroll
      for (int s = 0; s < N; ++s) q[s] = ((neg >> s) & 1u) ? -q[s] : q[s];
    }
  };
  const double idx[3] = {1.0 / a.kdx[0], 1.0 / a.kdx[1], 1.0 / a.kdx[2]};

  // rows y0 .. y0+nr-1 are this wave's output rows; row slot i holds row
  // y0-1+i (0: the halo row below, nr+1: the halo row above; slots past it
  // repeat row y0+nr and are not used)
  const int y0 = 2 + ychunk * R;
  const int nr = a.last_y - y0 + 1 < R ? a.last_y - y0 + 1 : R;
  const int z0 = 2 + zchunk * a.zrows;
  const int z1 = a.last_z < z0 + a.zrows - 1 ? a.last_z : z0 + a.zrows - 1;  // output planes z0 .. z1
  auto slot_row = [&](int i) { return i <= nr + 1 ? y0 - 1 + i : y0 + nr; };

  Frozen fp[R];         // frozen data of the previous plane, output rows
  double pend[R];       // rate of the previous plane's cells without its upper z-face
  S lmaxz[R];           // largest speed of the previous plane's lower z-face
  {
    S q[N];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      fetch(z0 - 1, slot_row(i + 1), q);
      fp[i] = D::frozen(q, gamma);
      pend[i] = 0.0;
      lmaxz[i] = S(0);
    }
  }
  double m = 0.0;  // largest rate of the wave's cells
#pragma unroll 1
  for (int z = z0; z <= z1 + 1; ++z) {
    const bool out_plane = z <= z1;
    S q[R + 2][N];
#pragma unroll
    for (int i = 0; i < R + 2; ++i)
      if (out_plane || (i >= 1 && i <= R)) fetch(z, slot_row(i), q[i]);
    // z-faces between planes z-1 and z: the upper face of plane z-1 (whose
    // rates it completes) and the lower face of plane z
    Frozen fz[R];
    S lminz[R], lmaxz_new[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      fz[i] = D::frozen(q[i + 1], gamma);
      D::speeds(fp[i], fz[i], 2, gamma, lminz[i], lmaxz_new[i]);
    }
    if (z > z0) {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        if (i >= nr) break;
        const double zr = D::side_rate(lmaxz[i], lminz[i], z <= sz - 2, idx[2], bad);
        if (out_col) m = fmax(m, fmax(pend[i], zr));
      }
    }
    if (!out_plane) break;
    // x- and y-faces of plane z
    Frozen lo = D::frozen(q[0], gamma);
    S lminy, lmaxy;
    D::speeds(lo, fz[0], 1, gamma, lminy, lmaxy);  // lower y-face of row y0
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const Frozen up = i + 1 < R ? fz[i + 1] : D::frozen(q[R + 1], gamma);
      S lminy_up, lmaxy_up;
      D::speeds(fz[i], up, 1, gamma, lminy_up, lmaxy_up);  // upper y-face of row y0+i
      if (i < nr) {
        S pk[NF], pkL[NF];
        D::pack(fz[i], pk);
#pragma unroll
        for (int k = 0; k < NF; ++k) pkL[k] = lane_from_left(pk[k]);
        S lminx, lmaxx;
        D::speeds(D::unpack(pkL), fz[i], 0, gamma, lminx, lmaxx);  // face c: cells c-1, c
        const S lminxR = lane_from_right(lminx);                  // face c+1
        const double xr = D::side_rate(lmaxx, lminxR, right_ok_x, idx[0], bad);
        const double yr = D::side_rate(lmaxy, lminy_up, y0 + i + 1 <= sy - 2, idx[1], bad);
        pend[i] = fmax(xr, yr);
      }
      lmaxy = lmaxy_up;
      fp[i] = fz[i];
      lmaxz[i] = lmaxz_new[i];
    }
  }
  return wave_dt(m, bad && out_col);
}

}  // namespace hydro
