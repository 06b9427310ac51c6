// stage3d.hip -- fused RK stage for 3D grids on gfx950.
//
// Roe family (limited, stencil radius 2): a workgroup of W wavefronts covers
// a 64-column strip (x, lanes; DPP shifts) by W consecutive rows (y, one row
// per wave; LDS exchange) and marches a chunk of z-planes with the z stencil
// held in registers.  The first and last 2 rows of the workgroup are halo
// rows that compute only what their neighbours need.
// HLL/HLLC (first order, radius 1): stage3d_r1_kernel below, and its
// variant stage3d_r1s_kernel that shares the rows of a workgroup's planes
// through LDS (chosen per combination mode).
// Kernel arithmetic and reference semantics are those of the 2D kernel
// (stage2d_impl.hpp).
#include "device_common.hpp"
#include "launch.hpp"
#include "policies.hpp"
#include "phys_srhd.hpp"
#include "phys_maxwell.hpp"
#include "stage_common.hpp"

namespace hydro {

template <class Pol, typename T, typename S, bool KAHAN, int W, int MODE>
__global__ void __launch_bounds__(64 * W) stage3d_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  constexpr int NF = Pol::NF;
  constexpr int NS = Pol::NS;
  constexpr int NA = NS + NF;
  constexpr int H = Pol::kNeighbors ? 2 : 1;
  using Frozen = typename Pol::Frozen;
  using Stage = typename Pol::Stage;
  using FaceA = typename Pol::FaceA;

  __shared__ T shA[2][W][NA][64];  // stage state + frozen data per row, double-buffered by plane
  __shared__ T shB[Pol::kNeighbors ? W : 1][N][64];  // deltaQTilde of the y-faces (Roe)
  __shared__ T shC[W][N][64];      // y-face fluxes

  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (lb >= a.nwaves) return;  // nwaves counts workgroups here (uniform per block)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform row
  const int strip = lb % a.nstrips;
  const int rest = lb / a.nstrips;
  const int ych = rest % a.nchunks;
  const int zch = rest / a.nchunks;

  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const long long pl = a.plane;
  const int c = strip * kStripOut + lane;
  const uint32_t cl = c < sx ? c : sx - 1;
  const int yb = 2 + ych * (W - 2 * H);
  const int y = yb - H + w;
  const int yl = y < sy ? y : sy - 1;
  const bool out_row = w >= H && w < W - H && y <= sy - 3;
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;
  const bool col_face = c >= 2 && c <= sx - 2;
  const bool row_face = y >= 2 && y <= sy - 2;

  const T gamma = T(a.gamma);
  const PhysParams<T> pp = make_params<T>(a);
  const double dtd = *a.dt;
  const T dt_dx = T(dtd * a.inv_dx[0]), dt_dy = T(dtd * a.inv_dx[1]), dt_dz = T(dtd * a.inv_dx[2]);
  const T idx = T(a.inv_dx[0]), idy = T(a.inv_dx[1]), idz = T(a.inv_dx[2]);
  const StageCombine<T, S, KAHAN, N, MODE> comb(a, dtd);
  const long long psz = (long long)sx * sy;

  auto load = [&](int z, Stage& st, Frozen& fz, bool& solid) {
    const long long e = (long long)z * psz + (long long)yl * sx;  // wave-uniform
    S q0[N];
    T dk[N], u[N];
    comb.load_u0(e, cl, q0);
    comb.load_dk(e, cl, dk);
    comb.stage_state(q0, dk, u);
    S pot = S(0), sol = S(0);
    if constexpr (PolicyPotential<Pol>::value) pot = ldp(a.pot, pl, 0, e, cl);
    if constexpr (PolicySolid<Pol>::value) sol = ldp(a.solid, pl, 0, e, cl);
    solid = solid_cell_of<Pol>(sol);
    st = stage_of<Pol>(u, pot, gamma);
    if constexpr (Pol::kFrozenPrims) {
      S pr[N];
#pragma unroll
      for (int s = 0; s < N; ++s) pr[s] = ldp(a.prim, pl, s, e, cl);
      fz = Pol::template frozen<S>(pr, gamma);
    } else {
      fz = frozen_of<Pol>(q0, pot, sol, gamma);
    }
  };

  const int z0 = 2 + zch * a.rows;
  int z1 = z0 + a.rows;
  if (z1 > sz - 2) z1 = sz - 2;

  Stage u_prev;
  Frozen f_prev;
  FaceA fa_m1, fa_m2;
  T Fz_prev[N], div_m1[N], div_m2[N];
#pragma unroll
  for (int s = 0; s < N; ++s) {
    Fz_prev[s] = div_m1[s] = div_m2[s] = T(0);
    fa_m1.dq[s] = fa_m2.dq[s] = T(0);
  }
  bool solid_m1, solid_m2 = false;  // solid flags of this row's cells in planes zn-1, zn-2
  load(z0 - 2, u_prev, f_prev, solid_m1);

#pragma unroll 1
  for (int zn = z0 - 1; zn <= z1 + 1; ++zn) {
    Stage u;
    Frozen fz;
    bool solid;
    load(zn, u, fz, solid);
    const bool out_plane = zn >= z0 && zn < z1;
    const int buf = zn & 1;

    // x-direction (DPP), output rows of output planes only
    T div[N];
#pragma unroll
    for (int s = 0; s < N; ++s) div[s] = T(0);
    if (out_plane && out_row) {
      T uL[NS], pk[NF], pkL[NF];
      Pol::pack_stage(u, uL);
#pragma unroll
      for (int s = 0; s < NS; ++s) uL[s] = lane_from_left(uL[s]);
      Pol::pack(fz, pk);
#pragma unroll
      for (int i = 0; i < NF; ++i) pkL[i] = lane_from_left(pk[i]);
      const FaceA fa = Pol::face_a(Pol::unpack(pkL), fz, Pol::unpack_stage(uL), u, 0, col_face, pp);
      T dqlo[N], dqhi[N], F[N];
      if constexpr (Pol::kNeighbors) {
#pragma unroll
        for (int s = 0; s < N; ++s) {
          dqlo[s] = lane_from_left(fa.dq[s]);
          dqhi[s] = lane_from_right(fa.dq[s]);
        }
      }
      Pol::face_b(fa, dqlo, dqhi, dt_dx, 0, a.lim, F);
#pragma unroll
      for (int s = 0; s < N; ++s) div[s] = -(lane_from_right(F[s]) - F[s]) * idx;
    }

    // y-direction through LDS (whole workgroup, uniform barriers)
    {
      T pk[NF], ps[NS];
      Pol::pack(fz, pk);
      Pol::pack_stage(u, ps);
#pragma unroll
      for (int s = 0; s < NS; ++s) shA[buf][w][s][lane] = ps[s];
#pragma unroll
      for (int i = 0; i < NF; ++i) shA[buf][w][NS + i][lane] = pk[i];
    }
    __syncthreads();
    FaceA fy;
    {
      const int ws = w > 0 ? w - 1 : 0;
      T uS[NS], pkS[NF];
#pragma unroll
      for (int s = 0; s < NS; ++s) uS[s] = shA[buf][ws][s][lane];
#pragma unroll
      for (int i = 0; i < NF; ++i) pkS[i] = shA[buf][ws][NS + i][lane];
      const bool ok = w > 0 && col_face && row_face && zn >= 2 && zn <= sz - 2;
      fy = Pol::face_a(Pol::unpack(pkS), fz, Pol::unpack_stage(uS), u, 1, ok, pp);
    }
    T Fy[N];
    if constexpr (Pol::kNeighbors) {
#pragma unroll
      for (int s = 0; s < N; ++s) shB[w][s][lane] = fy.dq[s];
      __syncthreads();
      const int ws = w > 0 ? w - 1 : 0;
      const int wn = w < W - 1 ? w + 1 : W - 1;
      T dqlo[N], dqhi[N];
#pragma unroll
      for (int s = 0; s < N; ++s) {
        dqlo[s] = shB[ws][s][lane];
        dqhi[s] = shB[wn][s][lane];
      }
      Pol::face_b(fy, dqlo, dqhi, dt_dy, 1, a.lim, Fy);
    } else {
      Pol::face_b(fy, nullptr, nullptr, dt_dy, 1, a.lim, Fy);
    }
#pragma unroll
    for (int s = 0; s < N; ++s) shC[w][s][lane] = Fy[s];
    __syncthreads();
    if (out_plane && out_row) {
      const int wn = w < W - 1 ? w + 1 : W - 1;
#pragma unroll
      for (int s = 0; s < N; ++s) div[s] -= (shC[wn][s][lane] - Fy[s]) * idy;
    }

    // z-direction in registers: first pass of face zn, flux of face zn-1,
    // then cells of plane zn-2
    const bool okz = col_face && row_face && zn >= 2 && zn <= sz - 2;
    const FaceA fa = Pol::face_a(f_prev, fz, u_prev, u, 2, okz, pp);
    T Fz[N];
    if (zn >= z0 + 1) {
      Pol::face_b(fa_m1, fa_m2.dq, fa.dq, dt_dz, 2, a.lim, Fz);
      if (zn >= z0 + 2 && out_row && out_lane) {
        T deriv[N];
#pragma unroll
        for (int s = 0; s < N; ++s) deriv[s] = solid_m2 ? T(0) : div_m2[s] - (Fz[s] - Fz_prev[s]) * idz;
        comb.finish((long long)(zn - 2) * psz + (long long)y * sx, c, deriv);
      }
    } else {
#pragma unroll
      for (int s = 0; s < N; ++s) Fz[s] = T(0);
    }
#pragma unroll
    for (int s = 0; s < N; ++s) {
      Fz_prev[s] = Fz[s];
      div_m2[s] = div_m1[s];
      div_m1[s] = div[s];
    }
    u_prev = u;
    fa_m2 = fa_m1;
    fa_m1 = fa;
    f_prev = fz;
    solid_m2 = solid_m1;
    solid_m1 = solid;
  }
}

// First-order face solvers (HLL, HLLC: stencil radius 1).  A workgroup of
// W waves covers one 64-column strip (x: lanes, DPP), W consecutive z-planes
// (one per wave) and a chunk of rows that every wave marches in y with the
// y stencil in registers, as the 2D kernel does.  Each wave computes the
// flux through the lower z-face of its own plane (loading the plane below,
// which the wave beneath loads too: an L2 hit) and hands it to the wave
// below through LDS, so every z-face is computed once; the top wave of the
// workgroup only supplies that face (W-1 output planes per workgroup).  One
// barrier per row.
template <class Pol, typename T, typename S, bool KAHAN, int MODE, int W>
__global__ void __launch_bounds__(64 * W) stage3d_r1_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  using Frozen = typename Pol::Frozen;
  using Comb = StageCombine<T, S, KAHAN, N, MODE>;
  __shared__ T shz[2][W][N][64];  // lower z-face fluxes per wave, double-buffered by row

  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (lb >= a.nwaves) return;  // nwaves counts workgroups here (uniform per block)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int nzb = (sz - 4 + W - 2) / (W - 1);
  const int zblk = lb % nzb;  // z fastest: neighbouring blocks share their boundary plane in L2
  const int rest = lb / nzb;
  const int strip = rest % a.nstrips;
  const int ych = rest / a.nstrips;
  const int p = 2 + zblk * (W - 1) + w;    // this wave's plane
  const bool ow = w < W - 1 && p <= sz - 3;  // output plane (wave-uniform)
  const bool zw = p <= sz - 2;               // its lower face is read by an output wave
  const long long psz = (long long)sx * sy;
  const int c = strip * kStripOut + lane;
  const uint32_t cl = c < sx ? c : sx - 1;
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;

  const PhysParams<T> pp = make_params<T>(a);
  const T gamma = pp.gamma;
  const double dtd = *a.dt;
  const T idx = T(a.inv_dx[0]), idy = T(a.inv_dx[1]), idz = T(a.inv_dx[2]);
  const Comb comb(a, dtd);

  struct Raw {
    S q0[N];
    T dk[N];
  };
  using Stage = typename Pol::Stage;
  struct Row {
    Stage st;
    Frozen fz;
  };
  auto fetch = [&](int plane, int row, Raw& r) {
    const long long e = (long long)__builtin_amdgcn_readfirstlane(plane) * psz +
                        (long long)__builtin_amdgcn_readfirstlane(row) * sx;
    comb.load_u0(e, cl, r.q0);
    comb.load_dk(e, cl, r.dk);
  };
  auto build = [&](const Raw& r, Row& R) {
    T u[N];
    comb.stage_state(r.q0, r.dk, u);
    R.st = Pol::stage(u, gamma);
    R.fz = Pol::template frozen<S>(r.q0, gamma);
  };
  auto flux = [&](const Row& L, const Row& R, int side, T F[N]) {
    const auto fa = Pol::face_a(L.fz, R.fz, L.st, R.st, side, true, pp);
    Pol::face_b(fa, nullptr, nullptr, T(0), side, 0, F);
  };
  auto xdiv = [&](const Row& R, T div[N]) {
    T ps[Pol::NS], pk[Pol::NF], pkL[Pol::NF];
    Pol::pack_stage(R.st, ps);
#pragma unroll
    for (int i = 0; i < Pol::NS; ++i) ps[i] = lane_from_left(ps[i]);
    Pol::pack(R.fz, pk);
#pragma unroll
    for (int i = 0; i < Pol::NF; ++i) pkL[i] = lane_from_left(pk[i]);
    Row L;
    L.st = Pol::unpack_stage(ps);
    L.fz = Pol::unpack(pkL);
    T F[N];
    flux(L, R, 0, F);
#pragma unroll
    for (int s = 0; s < N; ++s) div[s] = (F[s] - lane_from_right(F[s])) * idx;
  };

  const int pown = p < sz ? p : sz - 1;  // clamped load planes (idle waves past the grid)
  const int plow = p - 1 < sz ? p - 1 : sz - 1;
  const int y0 = 2 + ych * a.rows;
  int y1 = y0 + a.rows;
  if (y1 > sy - 2) y1 = sy - 2;

  // window: P = own row r-1 and its loads w1, Gp = flux through face r-1
  // (lower face of row r-1), acc = x and z divergence of row r-1
  Row P;
  Raw w1, nx, nz;
  T Gp[N], acc[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Gp[s] = acc[s] = T(0);
  if (ow) {
    fetch(pown, y0 - 1, w1);
    build(w1, P);
  }
  if (zw) {
    fetch(pown, y0, nx);
    fetch(plow, y0, nz);
  }
  // one row step; zstep: row r is an output row (its z-face and x-face run,
  // with the barrier), else only face r closes row r-1.  buf alternates.
  auto step = [&](int r, bool zstep, int buf) {
    const Raw cur = nx, curz = nz;
    if (zw && zstep) {  // prefetch row r+1 (<= y1 <= sy-2)
      fetch(pown, r + 1, nx);
      fetch(plow, r + 1, nz);
    }
    const long long eo = (long long)__builtin_amdgcn_readfirstlane(pown) * psz +
                         (long long)__builtin_amdgcn_readfirstlane(r - 1) * sx;
    // the final and generic combinations load more operands (replayed
    // increments, partial sums): they are fetched just before use, which
    // keeps them out of the registers live across the flux work
    constexpr bool kLateFin = MODE == CM_FINAL || MODE == CM_GENERIC;
    typename Comb::Fin fin;
    if (!kLateFin && ow && r > y0) comb.fetch_fin_rest(eo, cl, w1.q0, w1.dk, fin);
    Row C;
    if (zw) build(cur, C);
    T G[N], Fz[N], xd[N];
    if (ow) flux(P, C, 1, G);  // face r: rows r-1 | r
    if (zstep) {
      if (zw) {
        Row Z;
        build(curz, Z);
        flux(Z, C, 2, Fz);  // lower z-face of row r
#pragma unroll
        for (int s = 0; s < N; ++s) shz[buf][w][s][lane] = Fz[s];
      }
      if (ow) xdiv(C, xd);
      __syncthreads();
      if (ow) {
#pragma unroll
        for (int s = 0; s < N; ++s) xd[s] += (Fz[s] - shz[buf][w + 1][s][lane]) * idz;
      }
    }
    if (ow) {
      if (r > y0 && out_lane) {
        T deriv[N];
#pragma unroll
        for (int s = 0; s < N; ++s) deriv[s] = acc[s] + (Gp[s] - G[s]) * idy;
        if (kLateFin) comb.fetch_fin_rest(eo, cl, w1.q0, w1.dk, fin);
        comb.finish(eo, c, fin, deriv);
      }
#pragma unroll
      for (int s = 0; s < N; ++s) {
        Gp[s] = G[s];
        acc[s] = xd[s];
      }
      P = C;
      w1 = cur;
    }
  };
  // unrolled by two so the window rotates by renaming and the LDS buffer
  // index is static
  int r = y0;
#pragma unroll 1
  for (; r + 1 < y1; r += 2) {
    step(r, true, 0);
    step(r + 1, true, 1);
  }
  if (r < y1) {
    step(r, true, 0);
    ++r;
  }
  step(y1, false, 0);
}

// The same decomposition with the rows shared through LDS: each wave builds
// its own row (stage state and frozen data) once and publishes it; the wave
// above takes it as the lower neighbour of its z-face instead of loading and
// building the plane below itself (only wave 0 loads the plane beneath the
// workgroup).  The z-face flux then exists only after the row barrier, so a
// wave hands it to the wave below one row later: row r-1 is completed at
// step r, as in stage3d_r1_kernel, with the same arithmetic in the same
// order (bit-identical results).  Loads and frozen-data builds per workgroup
// and row: W + 1 instead of 2W.
template <class Pol, typename T, typename S, bool KAHAN, int MODE, int W>
__global__ void __launch_bounds__(64 * W) stage3d_r1s_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  constexpr int NS = Pol::NS, NF = Pol::NF;
  using Frozen = typename Pol::Frozen;
  using Comb = StageCombine<T, S, KAHAN, N, MODE>;
  __shared__ T shr[2][W - 1][NS + NF][64];  // published rows of waves 0..W-2, by row parity
  __shared__ T shz[2][W][N][64];            // lower z-face fluxes, by row parity

  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (lb >= a.nwaves) return;  // nwaves counts workgroups here (uniform per block)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int nzb = (sz - 4 + W - 2) / (W - 1);
  const int zblk = lb % nzb;
  const int rest = lb / nzb;
  const int strip = rest % a.nstrips;
  const int ych = rest / a.nstrips;
  const int p = 2 + zblk * (W - 1) + w;      // this wave's plane
  const bool ow = w < W - 1 && p <= sz - 3;  // output plane (wave-uniform)
  const bool zw = p <= sz - 2;               // its lower face is read by an output wave
  const long long psz = (long long)sx * sy;
  const int c = strip * kStripOut + lane;
  const uint32_t cl = c < sx ? c : sx - 1;
  const bool out_lane = lane >= 2 && lane < 62 && c <= sx - 3;

  const PhysParams<T> pp = make_params<T>(a);
  const T gamma = pp.gamma;
  const double dtd = *a.dt;
  const T idx = T(a.inv_dx[0]), idy = T(a.inv_dx[1]), idz = T(a.inv_dx[2]);
  const Comb comb(a, dtd);

  struct Raw {
    S q0[N];
    T dk[N];
  };
  using Stage = typename Pol::Stage;
  struct Row {
    Stage st;
    Frozen fz;
  };
  auto fetch = [&](int plane, int row, Raw& r) {
    const long long e = (long long)__builtin_amdgcn_readfirstlane(plane) * psz +
                        (long long)__builtin_amdgcn_readfirstlane(row) * sx;
    comb.load_u0(e, cl, r.q0);
    comb.load_dk(e, cl, r.dk);
  };
  auto build = [&](const Raw& r, Row& R) {
    T u[N];
    comb.stage_state(r.q0, r.dk, u);
    R.st = Pol::stage(u, gamma);
    R.fz = Pol::template frozen<S>(r.q0, gamma);
  };
  auto flux = [&](const Row& L, const Row& R, int side, T F[N]) {
    const auto fa = Pol::face_a(L.fz, R.fz, L.st, R.st, side, true, pp);
    Pol::face_b(fa, nullptr, nullptr, T(0), side, 0, F);
  };
  auto xdiv = [&](const Row& R, T div[N]) {
    T ps[NS], pk[NF], pkL[NF];
    Pol::pack_stage(R.st, ps);
#pragma unroll
    for (int i = 0; i < NS; ++i) ps[i] = lane_from_left(ps[i]);
    Pol::pack(R.fz, pk);
#pragma unroll
    for (int i = 0; i < NF; ++i) pkL[i] = lane_from_left(pk[i]);
    Row L;
    L.st = Pol::unpack_stage(ps);
    L.fz = Pol::unpack(pkL);
    T F[N];
    flux(L, R, 0, F);
#pragma unroll
    for (int s = 0; s < N; ++s) div[s] = (F[s] - lane_from_right(F[s])) * idx;
  };

  const int pown = p < sz ? p : sz - 1;  // clamped load planes (idle waves past the grid)
  const int plow = p - 1 < sz ? p - 1 : sz - 1;
  const int y0 = 2 + ych * a.rows;
  int y1 = y0 + a.rows;
  if (y1 > sy - 2) y1 = sy - 2;

  // window: P = own row r-1 and its loads w1; Gp = flux through face r-1;
  // xp = x-divergence of row r-1; Fzp = flux through row r-1's lower z-face
  Row P;
  Raw w1, nx, nz;
  T Gp[N], xp[N], Fzp[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Gp[s] = xp[s] = Fzp[s] = T(0);
  if (ow) {
    fetch(pown, y0 - 1, w1);
    build(w1, P);
  }
  if (zw) {
    fetch(pown, y0, nx);
    if (w == 0) fetch(plow, y0, nz);
  }
  // one row step, one barrier.  zstep: row r is an output row (its z-face
  // and x-face run); otherwise (r = y1) face r only closes row r-1.
  auto step = [&](int r, bool zstep, int buf) {
    const Raw cur = nx, curz = nz;
    if (zw && zstep) {  // prefetch row r+1 (<= y1 <= sy-2)
      fetch(pown, r + 1, nx);
      if (w == 0 && r + 1 < y1) fetch(plow, r + 1, nz);
    }
    const long long eo = (long long)__builtin_amdgcn_readfirstlane(pown) * psz +
                         (long long)__builtin_amdgcn_readfirstlane(r - 1) * sx;
    constexpr bool kLateFin = MODE == CM_FINAL || MODE == CM_GENERIC;
    typename Comb::Fin fin;
    if (!kLateFin && ow && r > y0) comb.fetch_fin_rest(eo, cl, w1.q0, w1.dk, fin);
    Row C;
    if (zw) build(cur, C);
    if (zstep && zw && w < W - 1) {  // publish the row for the wave above
      T pk[NS + NF];
      Pol::pack_stage(C.st, pk);
      Pol::pack(C.fz, pk + NS);
#pragma unroll
      for (int i = 0; i < NS + NF; ++i) shr[buf][w][i][lane] = pk[i];
    }
    T G[N], xd[N];
    if (ow) flux(P, C, 1, G);  // face r: rows r-1 | r
    if (zstep && ow) xdiv(C, xd);
    __syncthreads();
    if (ow) {
      if (r > y0 && out_lane) {
        // row r-1: its upper z-face is the lower face the wave above computed
        // one step ago
        T acc[N], deriv[N];
#pragma unroll
        for (int s = 0; s < N; ++s) {
          acc[s] = xp[s];
          acc[s] += (Fzp[s] - shz[buf ^ 1][w + 1][s][lane]) * idz;
          deriv[s] = acc[s] + (Gp[s] - G[s]) * idy;
        }
        if (kLateFin) comb.fetch_fin_rest(eo, cl, w1.q0, w1.dk, fin);
        comb.finish(eo, c, fin, deriv);
      }
#pragma unroll
      for (int s = 0; s < N; ++s) {
        Gp[s] = G[s];
        xp[s] = xd[s];
      }
      P = C;
      w1 = cur;
    }
    if (zstep && zw) {
      // lower z-face of row r: the plane below is the wave beneath's row
      // (wave 0: loaded)
      Row Z;
      if (w == 0) {
        build(curz, Z);
      } else {
        T pk[NS + NF];
#pragma unroll
        for (int i = 0; i < NS + NF; ++i) pk[i] = shr[buf][w - 1][i][lane];
        Z.st = Pol::unpack_stage(pk);
        Z.fz = Pol::unpack(pk + NS);
      }
      T Fz[N];
      flux(Z, C, 2, Fz);
#pragma unroll
      for (int s = 0; s < N; ++s) {
        shz[buf][w][s][lane] = Fz[s];
        Fzp[s] = Fz[s];
      }
    }
  };
  // unrolled by two so the LDS buffer index is static
  int r = y0;
#pragma unroll 1
  for (; r + 1 < y1; r += 2) {
    step(r, true, 0);
    step(r + 1, true, 1);
  }
  if (r < y1) {
    step(r, true, 0);
    ++r;
  }
  step(y1, false, (y1 - y0) & 1);
}

// Neighbour exchange along x of per-lane cell arrays.  CPL = 1: lane l
// holds strip cell l; CPL = 2: lane l holds cells l and 64 + l (two
// coalesced 64-wide halves), so the left neighbour of cell 64 is lane 63's
// first cell and the right neighbour of cell 63 is lane 0's second cell:
// full-wave rotations with the halves swapped at the wrapping lane.
template <int CPL>
struct XLanes;
template <>
struct XLanes<1> {
  template <typename T, int M>
  static HYDRO_DEV void from_left(const T v[1][M], T o[1][M]) {
#pragma unroll
    for (int i = 0; i < M; ++i) o[0][i] = lane_from_left(v[0][i]);
  }
  template <typename T, int M>
  static HYDRO_DEV void from_right(const T v[1][M], T o[1][M]) {
#pragma unroll
    for (int i = 0; i < M; ++i) o[0][i] = lane_from_right(v[0][i]);
  }
};
template <>
struct XLanes<2> {
  template <typename T, int M>
  static HYDRO_DEV void from_left(const T v[2][M], T o[2][M]) {
    const bool last = (threadIdx.x & 63) == 63;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      o[0][i] = lane_from_left(v[0][i]);
      o[1][i] = lane_rot_from_left(last ? v[0][i] : v[1][i]);
    }
  }
  template <typename T, int M>
  static HYDRO_DEV void from_right(const T v[2][M], T o[2][M]) {
    const bool first = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      o[0][i] = lane_rot_from_right(first ? v[1][i] : v[0][i]);
      o[1][i] = lane_from_right(v[1][i]);
    }
  }
};

// z-march form of the first-order stage (HLL, HLLC).  A workgroup of W
// waves covers one strip of 64 * CPL columns (x: lanes, CPL cells per lane,
// DPP exchange) and W consecutive rows, one per wave: rows yb-1 .. yb+W-2,
// of which yb .. yb+W-3 are output rows (the first and last are halo rows).
// It marches a chunk of z-planes with the z stencil in registers: per plane
// every wave loads and builds its own row once, publishes it in LDS,
// computes the flux through its lower y-face from the row of the wave
// below and publishes that; an output row takes its upper y-face from the
// wave above.  A plane's cells are complete one plane later, when the flux
// through their upper z-face exists.  Two barriers per plane.  Each cell is
// loaded and built once per workgroup (W / (W-2) per output cell), every
// face is computed once, and the divergence is summed in the order of
// stage3d_r1_kernel, so both give the same bits.
template <class Pol, typename T, typename S, bool KAHAN, int MODE, int W, int CPL>
__global__ void __launch_bounds__(64 * W) stage3d_zm_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  constexpr int NS = Pol::NS, NP = NS + Pol::NFS;  // a neighbour's stage data + the frozen data a face reads
  constexpr int XW = 64 * CPL;  // strip width in cells
  constexpr int XO = XW - 2;    // output columns per strip
  using Frozen = typename Pol::Frozen;
  using Stage = typename Pol::Stage;
  using Comb = StageCombine<T, S, KAHAN, N, MODE>;
  __shared__ T shr[W][NP][XW];  // each wave's row of the current plane, packed
  __shared__ T shf[W][N][XW];   // flux through each wave's lower y-face

  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (lb >= a.nwaves) return;  // nwaves counts workgroups here (uniform per block)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int nzc = (sz - 4 + a.rows - 1) / a.rows;
  const int zch = lb % nzc;  // z fastest: neighbouring chunks share their halo planes in L2
  const int rest = lb / nzc;
  const int yblk = rest % a.nchunks;
  const int strip = rest / a.nchunks;
  const int y = 1 + yblk * (W - 2) + w;  // this wave's row
  const int yl = y < sy ? y : sy - 1;
  const bool orow = w >= 1 && w <= W - 2 && y <= sy - 3;  // output row (wave-uniform)
  const long long psz = (long long)sx * sy;
  const int base = 1 + strip * XO;
  int c[CPL];
  uint32_t cl[CPL];
  bool oc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int p = k * 64 + lane;
    c[k] = base + p;
    cl[k] = c[k] < sx ? c[k] : sx - 1;
    oc[k] = p >= 1 && p <= XW - 2 && c[k] <= sx - 3;
  }

  const PhysParams<T> pp = make_params<T>(a);
  const T gamma = pp.gamma;
  const double dtd = *a.dt;
  const T idx = T(a.inv_dx[0]), idy = T(a.inv_dx[1]), idz = T(a.inv_dx[2]);
  const Comb comb(a, dtd);

  struct Raw {
    S q0[N];
    T dk[N];
  };
  struct Cell {
    Stage st;
    Frozen fz;
  };
  auto erow = [&](int plane) {
    return (long long)__builtin_amdgcn_readfirstlane(plane) * psz + (long long)__builtin_amdgcn_readfirstlane(yl) * sx;
  };
  auto fetch = [&](int plane, Raw r[CPL]) {
    const long long e = erow(plane);
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      comb.load_u0(e, cl[k], r[k].q0);
      comb.load_dk(e, cl[k], r[k].dk);
    }
  };
  auto build = [&](const Raw r[CPL], Cell C[CPL]) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      T u[N];
      comb.stage_state(r[k].q0, r[k].dk, u);
      C[k].st = Pol::stage(u, gamma);
      C[k].fz = Pol::template frozen<S>(r[k].q0, gamma);
    }
  };
  auto pack = [&](const Cell& C, int side, T o[NP]) {
    Pol::pack_stage(C.st, o);
    Pol::pack_side(C.fz, side, o + NS);
  };
  auto unpack = [&](const T o[NP], int side, Cell& C) {
    C.st = Pol::unpack_stage(o);
    C.fz = Pol::unpack_side(o + NS, side);
  };
  auto flux = [&](const Cell& L, const Cell& R, int side, T F[N]) {
    const auto fa = Pol::face_a(L.fz, R.fz, L.st, R.st, side, true, pp);
    Pol::face_b(fa, nullptr, nullptr, T(0), side, 0, F);
  };
  auto xdiv = [&](const Cell C[CPL], T div[CPL][N]) {
    T pk[CPL][NP], pl[CPL][NP];
#pragma unroll
    for (int k = 0; k < CPL; ++k) pack(C[k], 0, pk[k]);
    XLanes<CPL>::template from_left<T, NP>(pk, pl);
    T F[CPL][N], Fr[CPL][N];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      Cell L;
      unpack(pl[k], 0, L);
      flux(L, C[k], 0, F[k]);
    }
    XLanes<CPL>::template from_right<T, N>(F, Fr);
#pragma unroll
    for (int k = 0; k < CPL; ++k)
#pragma unroll
      for (int s = 0; s < N; ++s) div[k][s] = (F[k][s] - Fr[k][s]) * idx;
  };

  const int z0 = 2 + zch * a.rows;
  int z1 = z0 + a.rows;
  if (z1 > sz - 2) z1 = sz - 2;  // output planes z0 .. z1-1; z0-1 and z1 are halo planes

  // registers: P = the row at plane zn-1 and its loads rp; per cell of
  // plane zn-1: x-divergence xp, y-face difference yd, lower z-face flux fzl
  Cell P[CPL];
  Raw rp[CPL], nx[CPL];
  T xp[CPL][N], yd[CPL][N], fzl[CPL][N];
#pragma unroll
  for (int k = 0; k < CPL; ++k)
#pragma unroll
    for (int s = 0; s < N; ++s) xp[k][s] = yd[k][s] = fzl[k][s] = T(0);
  fetch(z0 - 1, rp);
  build(rp, P);
  fetch(z0, nx);
  constexpr bool kLateFin = MODE == CM_FINAL || MODE == CM_GENERIC;

#pragma unroll 1
  for (int zn = z0; zn <= z1; ++zn) {
    Raw cur[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) cur[k] = nx[k];
    if (zn < z1) fetch(zn + 1, nx);  // prefetch the next plane (<= z1 <= sz-2)
    typename Comb::Fin fin[CPL];
    const long long ep = erow(zn - 1);
    if (!kLateFin && orow && zn > z0) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) comb.fetch_fin_rest(ep, cl[k], rp[k].q0, rp[k].dk, fin[k]);
    }
    Cell C[CPL];
    build(cur, C);
    // face zn: lower z-face of plane zn, upper z-face of plane zn-1
    T Fz[CPL][N];
    if (orow) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) flux(P[k], C[k], 2, Fz[k]);
      if (zn > z0) {
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          if (!oc[k]) continue;
          T deriv[N];
#pragma unroll
          for (int s = 0; s < N; ++s) {
            T acc = xp[k][s];
            acc += (fzl[k][s] - Fz[k][s]) * idz;
            deriv[s] = acc + yd[k][s] * idy;
          }
          if (kLateFin) comb.fetch_fin_rest(ep, cl[k], rp[k].q0, rp[k].dk, fin[k]);
          comb.finish(ep, c[k], fin[k], deriv);
        }
      }
    }
    if (zn < z1) {  // plane zn is an output plane: its x- and y-faces
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        T pk[NP];
        pack(C[k], 1, pk);
#pragma unroll
        for (int i = 0; i < NP; ++i) shr[w][i][k * 64 + lane] = pk[i];
      }
      __syncthreads();
      T G[CPL][N];
      if (w >= 1) {
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          T pk[NP];
#pragma unroll
          for (int i = 0; i < NP; ++i) pk[i] = shr[w - 1][i][k * 64 + lane];
          Cell L;
          unpack(pk, 1, L);
          flux(L, C[k], 1, G[k]);
#pragma unroll
          for (int s = 0; s < N; ++s) shf[w][s][k * 64 + lane] = G[k][s];
        }
      }
      if (orow) xdiv(C, xp);
      __syncthreads();
      if (orow) {
#pragma unroll
        for (int k = 0; k < CPL; ++k)
#pragma unroll
          for (int s = 0; s < N; ++s) {
            yd[k][s] = G[k][s] - shf[w + 1][s][k * 64 + lane];
            fzl[k][s] = Fz[k][s];
          }
      }
    }
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      P[k] = C[k];
      rp[k] = cur[k];
    }
  }
}

template <class Pol, typename T, typename S, bool KAHAN, int MODE>
static void launch3_r1_mode(const StageArgs<T, S>& a, hipStream_t st) {
  const int blocks = (a.nwaves + 7) & ~7;  // multiple of 8 for the XCD remap
  if (a.zmarch) {
    // z-march form: (waves per workgroup, cells per lane) as chosen on the
    // host; the FP32 forms with 1024 threads or 2 cells per lane fit the
    // 160 KB of LDS, the FP64 form takes 8 waves and 1 cell per lane
    if constexpr (sizeof(T) == 4) {
      if (a.zwaves == 16 && a.zmarch == 1) {
        hipLaunchKernelGGL((stage3d_zm_kernel<Pol, T, S, KAHAN, MODE, 16, 1>), dim3(blocks), dim3(64 * 16), 0, st, a);
        return;
      }
      if (a.zwaves == 8 && a.zmarch == 2) {
        hipLaunchKernelGGL((stage3d_zm_kernel<Pol, T, S, KAHAN, MODE, 8, 2>), dim3(blocks), dim3(64 * 8), 0, st, a);
        return;
      }
    }
    hipLaunchKernelGGL((stage3d_zm_kernel<Pol, T, S, KAHAN, MODE, 8, 1>), dim3(blocks), dim3(64 * 8), 0, st, a);
    return;
  }
  if (a.zwaves == 4 && ((a.zshare >> MODE) & 1))
    hipLaunchKernelGGL((stage3d_r1s_kernel<Pol, T, S, KAHAN, MODE, 4>), dim3(blocks), dim3(64 * 4), 0, st, a);
  else if (a.zwaves == 4)
    hipLaunchKernelGGL((stage3d_r1_kernel<Pol, T, S, KAHAN, MODE, 4>), dim3(blocks), dim3(64 * 4), 0, st, a);
  else
    hipLaunchKernelGGL((stage3d_r1_kernel<Pol, T, S, KAHAN, MODE, 8>), dim3(blocks), dim3(64 * 8), 0, st, a);
}

template <class Pol, typename T, typename S, bool KAHAN>
static void launch3_r1(int mode, const StageArgs<T, S>& a, hipStream_t st) {
  switch (mode) {
  case CM_FIRST: launch3_r1_mode<Pol, T, S, KAHAN, CM_FIRST>(a, st); return;
  case CM_MID: launch3_r1_mode<Pol, T, S, KAHAN, CM_MID>(a, st); return;
  case CM_FINAL: launch3_r1_mode<Pol, T, S, KAHAN, CM_FINAL>(a, st); return;
  case CM_SINGLE: launch3_r1_mode<Pol, T, S, KAHAN, CM_SINGLE>(a, st); return;
  case CM_DERIV:
    if constexpr (!KAHAN) {
      launch3_r1_mode<Pol, T, S, KAHAN, CM_DERIV>(a, st);
      return;
    }
    [[fallthrough]];
  default: launch3_r1_mode<Pol, T, S, KAHAN, CM_GENERIC>(a, st); return;
  }
}

template <class Pol, typename T, typename S, bool KAHAN, int W>
static void launch3(int mode, const StageArgs<T, S>& a, hipStream_t st) {
  int blocks = (a.nwaves + 7) & ~7;  // nwaves = workgroups; multiple of 8 for the XCD remap
  if constexpr (!KAHAN) {
    if (mode == CM_DERIV) {
      hipLaunchKernelGGL((stage3d_kernel<Pol, T, S, KAHAN, W, CM_DERIV>), dim3(blocks), dim3(64 * W), 0, st, a);
      return;
    }
  }
  hipLaunchKernelGGL((stage3d_kernel<Pol, T, S, KAHAN, W, CM_GENERIC>), dim3(blocks), dim3(64 * W), 0, st, a);
}

template <typename T, typename S, bool KAHAN, int W>
static void launch3_solver(int kind, int lim, int mode, const StageArgs<T, S>& a, hipStream_t st) {
  if (kind == 1) {
    launch3_r1<HLLPolicy<3, T, 1>, T, S, KAHAN>(mode, a, st);
  } else if (kind == 2) {
    launch3_r1<HLLPolicy<3, T, 2>, T, S, KAHAN>(mode, a, st);
  } else if (kind == 3) {
    launch3<SRHDRoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(mode, a, st);
  } else if (kind == 4) {
    launch3<MaxwellRoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(mode, a, st);
  } else if (kind == 5 && a.solid) {
    launch3<SolidBurgersPolicy<3, T>, T, S, KAHAN, W>(mode, a, st);
  } else if (kind == 5) {
    launch3<BurgersPolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(mode, a, st);
  } else if (kind == 0 && a.solid) {
    launch3<SolidRoePolicy<3, T>, T, S, KAHAN, W>(mode, a, st);
  } else if (kind == 0 && a.pot) {
    launch3<RoePolicy<3, T, LIM_RUNTIME, true>, T, S, KAHAN, W>(mode, a, st);
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch3<RoePolicy<3, T, LIM_SUPERBEE>, T, S, KAHAN, W>(mode, a, st); break;
    case LIM_MONOTIZEDCENTRAL: launch3<RoePolicy<3, T, LIM_MONOTIZEDCENTRAL>, T, S, KAHAN, W>(mode, a, st); break;
    case LIM_MINMOD: launch3<RoePolicy<3, T, LIM_MINMOD>, T, S, KAHAN, W>(mode, a, st); break;
    default: launch3<RoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(mode, a, st); break;
    }
  }
}

// output rows per workgroup of the Roe-family kernel: 2 halo rows per side
// (limiter neighbours)
int stage3d_rows_per_block(int waves) { return waves - 2 * 2; }

// the Roe-family kernel runs the generic combination (or the derivative
// alone); the first-order kernel takes every mode
void launch_stage3d_f64(int solver, int lim, int mode, const StageArgs<double, double>& a, hipStream_t st) {
  launch3_solver<double, double, false, kStage3dWaves>(solver, lim, mode, a, st);
}

void launch_stage3d_mixed(int solver, int lim, int mode, bool kahan, const StageArgs<float, float>& a,
                          hipStream_t st) {
  // the Roe family in 16-row workgroups when a.zwaves says so (the
  // first-order kernels read zwaves themselves)
  const bool w16 = solver != 1 && solver != 2 && a.zwaves == 16;
  if (kahan) {
    if (w16) launch3_solver<float, float, true, 16>(solver, lim, mode, a, st);
    else launch3_solver<float, float, true, kStage3dWaves>(solver, lim, mode, a, st);
  } else {
    if (w16) launch3_solver<float, float, false, 16>(solver, lim, mode, a, st);
    else launch3_solver<float, float, false, kStage3dWaves>(solver, lim, mode, a, st);
  }
}

}  // namespace hydro
