// stage3d.hip -- fused RK stage for 3D grids on gfx950.
//
// A workgroup of W wavefronts covers a 64-column strip (x, lanes; DPP
// shifts) by W consecutive rows (y, one row per wave; LDS exchange) and
// marches a chunk of z-planes with the z stencil held in registers.  The
// first and last H rows of the workgroup are halo rows (H = 2 for Roe, whose
// limiter needs the neighbouring interfaces, 1 for HLL/HLLC); they compute
// only what their neighbours need.  Kernel arithmetic and reference
// semantics are those of stage2d.hip.
#include "device_common.hpp"
#include "launch.hpp"
#include "policies.hpp"
#include "phys_srhd.hpp"
#include "phys_maxwell.hpp"
#include "stage_common.hpp"

namespace hydro {

template <class Pol, typename T, typename S, bool KAHAN, int W>
__global__ void __launch_bounds__(64 * W) stage3d_kernel(StageArgs<T, S> a) {
  constexpr int N = Pol::N;
  constexpr int NF = Pol::NF;
  constexpr int NA = N + NF;
  constexpr int H = Pol::kNeighbors ? 2 : 1;
  using Frozen = typename Pol::Frozen;
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
  const StageCombine<T, S, KAHAN, N, CM_GENERIC> comb(a, dtd);
  const long long psz = (long long)sx * sy;

  auto load = [&](int z, T u[N], Frozen& fz) {
    const long long e = (long long)z * psz + (long long)yl * sx;  // wave-uniform
    S q0[N];
    T dk[N];
    comb.load_u0(e, cl, q0);
    comb.load_dk(e, cl, dk);
    comb.stage_state(q0, dk, u);
    if constexpr (Pol::kFrozenPrims) {
      double pr[N];
#pragma unroll
      for (int s = 0; s < N; ++s) pr[s] = ldp(a.prim, pl, s, e, cl);
      fz = Pol::template frozen<double>(pr, gamma);
    } else {
      fz = Pol::template frozen<S>(q0, gamma);
    }
  };

  const int z0 = 2 + zch * a.rows;
  int z1 = z0 + a.rows;
  if (z1 > sz - 2) z1 = sz - 2;

  T u_prev[N];
  Frozen f_prev;
  FaceA fa_m1, fa_m2;
  T Fz_prev[N], div_m1[N], div_m2[N];
#pragma unroll
  for (int s = 0; s < N; ++s) {
    Fz_prev[s] = div_m1[s] = div_m2[s] = T(0);
    fa_m1.dq[s] = fa_m2.dq[s] = T(0);
  }
  load(z0 - 2, u_prev, f_prev);

#pragma unroll 1
  for (int zn = z0 - 1; zn <= z1 + 1; ++zn) {
    T u[N];
    Frozen fz;
    load(zn, u, fz);
    const bool out_plane = zn >= z0 && zn < z1;
    const int buf = zn & 1;

    // x-direction (DPP), output rows of output planes only
    T div[N];
#pragma unroll
    for (int s = 0; s < N; ++s) div[s] = T(0);
    if (out_plane && out_row) {
      T uL[N], pk[NF], pkL[NF];
#pragma unroll
      for (int s = 0; s < N; ++s) uL[s] = lane_from_left(u[s]);
      Pol::pack(fz, pk);
#pragma unroll
      for (int i = 0; i < NF; ++i) pkL[i] = lane_from_left(pk[i]);
      const FaceA fa = Pol::face_a(Pol::unpack(pkL), fz, uL, u, 0, col_face, pp);
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
      T pk[NF];
      Pol::pack(fz, pk);
#pragma unroll
      for (int s = 0; s < N; ++s) shA[buf][w][s][lane] = u[s];
#pragma unroll
      for (int i = 0; i < NF; ++i) shA[buf][w][N + i][lane] = pk[i];
    }
    __syncthreads();
    FaceA fy;
    {
      const int ws = w > 0 ? w - 1 : 0;
      T uS[N], pkS[NF];
#pragma unroll
      for (int s = 0; s < N; ++s) uS[s] = shA[buf][ws][s][lane];
#pragma unroll
      for (int i = 0; i < NF; ++i) pkS[i] = shA[buf][ws][N + i][lane];
      const bool ok = w > 0 && col_face && row_face && zn >= 2 && zn <= sz - 2;
      fy = Pol::face_a(Pol::unpack(pkS), fz, uS, u, 1, ok, pp);
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
        for (int s = 0; s < N; ++s) deriv[s] = div_m2[s] - (Fz[s] - Fz_prev[s]) * idz;
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
      u_prev[s] = u[s];
    }
    fa_m2 = fa_m1;
    fa_m1 = fa;
    f_prev = fz;
  }
}

template <class Pol, typename T, typename S, bool KAHAN, int W>
static void launch3(const StageArgs<T, S>& a, hipStream_t st) {
  int blocks = (a.nwaves + 7) & ~7;  // nwaves = workgroups; multiple of 8 for the XCD remap
  hipLaunchKernelGGL((stage3d_kernel<Pol, T, S, KAHAN, W>), dim3(blocks), dim3(64 * W), 0, st, a);
}

template <typename T, typename S, bool KAHAN, int W>
static void launch3_solver(int kind, int lim, const StageArgs<T, S>& a, hipStream_t st) {
  if (kind == 1) {
    launch3<HLLPolicy<3, T, 1>, T, S, KAHAN, W>(a, st);
  } else if (kind == 2) {
    launch3<HLLPolicy<3, T, 2>, T, S, KAHAN, W>(a, st);
  } else if (kind == 3) {
    launch3<SRHDRoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(a, st);
  } else if (kind == 4) {
    launch3<MaxwellRoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(a, st);
  } else {
    switch (lim) {
    case LIM_SUPERBEE: launch3<RoePolicy<3, T, LIM_SUPERBEE>, T, S, KAHAN, W>(a, st); break;
    case LIM_MONOTIZEDCENTRAL: launch3<RoePolicy<3, T, LIM_MONOTIZEDCENTRAL>, T, S, KAHAN, W>(a, st); break;
    case LIM_MINMOD: launch3<RoePolicy<3, T, LIM_MINMOD>, T, S, KAHAN, W>(a, st); break;
    default: launch3<RoePolicy<3, T, LIM_RUNTIME>, T, S, KAHAN, W>(a, st); break;
    }
  }
}

// halo rows: 2 for the Roe family (limiter neighbours), 1 for HLL/HLLC
int stage3d_rows_per_block(int kind) { return kStage3dWaves - 2 * ((kind == 1 || kind == 2) ? 1 : 2); }

void launch_stage3d_f64(int solver, int lim, int mode, const StageArgs<double, double>& a, hipStream_t st) {
  (void)mode;  // 3D stages use the generic combination
  launch3_solver<double, double, false, kStage3dWaves>(solver, lim, a, st);
}

void launch_stage3d_mixed(int solver, int lim, int mode, bool kahan, const StageArgs<float, float>& a,
                          hipStream_t st) {
  (void)mode;
  if (kahan) launch3_solver<float, float, true, kStage3dWaves>(solver, lim, a, st);
  else launch3_solver<float, float, false, kStage3dWaves>(solver, lim, a, st);
}

}  // namespace hydro
