// launch.hpp -- host-side launchers of the device kernels (one translation
// unit per kernel family, called by the C ABI in capi.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include "stage_args.hpp"

namespace hydro {

// fused RK stage, 1D/2D (stage2d_*.hip).  solver: 0 Euler Roe, 1 HLL,
// 2 HLLC, 3 SRHD Roe, 4 Maxwell Roe; mode: CombineMode of the stage.
// f32: FP32 storage and arithmetic; f32k: the same with compensated storage.
void launch_stage2d_f64(int dim, int solver, int lim, int mode, const StageArgs<double, double>& a, int wpb,
                        hipStream_t st);
void launch_stage2d_f32(int dim, int solver, int lim, int mode, const StageArgs<float, float>& a, int wpb,
                        hipStream_t st);
void launch_stage2d_f32k(int dim, int solver, int lim, int mode, const StageArgs<float, float>& a, int wpb,
                         hipStream_t st);

// fused RK stage, 3D (stage3d.hip).  Roe family: workgroups of a.zwaves
// rows (8, or 16 for FP32, whose LDS rows are half as wide), 2 halo rows on
// each side
constexpr int kStage3dWaves = 8;
// the default workgroup height of the Roe-family kernel
inline int stage3d_roe_default_waves(bool fp32) { return fp32 ? 16 : 8; }
// first-order 3D kernel: waves (z-planes) per workgroup (4 or 8), one of
// them supplies only the top z-face.  Four measured faster in both
// precisions on the 256^3 blast (more workgroups per CU hide the barrier).
inline int stage3d_r1_waves(bool) { return 4; }
// output rows per workgroup of the Roe-family kernel with `waves` rows
int stage3d_rows_per_block(int waves);
void launch_stage3d_f64(int solver, int lim, int mode, const StageArgs<double, double>& a, hipStream_t st);
void launch_stage3d_mixed(int solver, int lim, int mode, bool kahan, const StageArgs<float, float>& a,
                          hipStream_t st);

// ghost fill of every axis in one launch (res/Common.cl:49-194 composed),
// SoA planes.  method[axis][minmax]: 0 periodic, 1 mirror, 2 freeflow,
// -1 none; negmask[axis][minmax]: bit s set when state s is reflected
// (negated) rather than mirrored
struct BoundaryArgs {
  int dim, sx, sy, sz;
  long long plane;
  int nstates;
  signed char method[3][2];
  unsigned negmask[3][2];
};
// prim: SRHD primitive planes filled by the same launch, or null
void launch_boundary_f64(const BoundaryArgs& b, double* planes, double* prim, hipStream_t st);
void launch_boundary_f32(const BoundaryArgs& b, float* planes, float* lo, float* prim, hipStream_t st);

// copy ghost cells of every plane from src to dst (the cells the stage
// kernels do not write)
void launch_copy_ghosts_f64(int dim, int sx, int sy, int sz, long long plane, int n, const double* src,
                            double* dst, hipStream_t st);
void launch_copy_ghosts_f32(int dim, int sx, int sy, int sz, long long plane, int n, const float* src,
                            float* dst, hipStream_t st);

// CFL: per-block minima of dt_cell, then a one-block finalise
struct DtArgs {
  int dim, sx, sy, sz;
  long long plane;
  double kdx[3];
  double gamma;
  int solver;            // 0 Roe, 1 HLL, 2 HLLC
  int eqn;               // 0 Euler, 1 SRHD, 2 Maxwell
  const void* prim;      // SRHD primitive planes
  int prim_f32;          // ... in FP32 (else FP64)
  double maxwell_lambda; // Maxwell: 1/sqrt(eps mu)
  const void* pot;       // Euler Roe / HLL / EulerBurgers: potential plane (state precision), or null
  const void* solid;     // Euler Roe / EulerBurgers: solid mask plane (state precision), or null;
                         // Roe with a mask reads pot too
  int nblocks;
  double* blockmin;      // [nblocks]
};

// SRHD: constrainState on every cell, then con2prim (FP64) on cells 2..S-2
// (res/SRHDCommon.cl:94-196); health bit 2 on non-convergence
struct Con2PrimArgs {
  int dim, sx, sy, sz;
  long long plane;
  double gamma;
  int max_iter;
  double stop_eps, vel_eps, pmin_eps;
  double rho_min, rho_max, eint_max;
  double D_min, D_max, tau_min, tau_max;
  unsigned* health;
  // when set: the state planes (and low words) whose ghost cells the new
  // state takes before the clamp (the deferred ghost copy, in this launch)
  const void* ghost_src;
  const float* ghost_src_lo;
};
void launch_con2prim_f64(const Con2PrimArgs& a, double* u, double* prim, hipStream_t st);
void launch_con2prim_f32(const Con2PrimArgs& a, float* u, float* lo, float* prim, hipStream_t st);
void launch_dt_euler_f64(const DtArgs& a, const double* u0, hipStream_t st);
void launch_dt_euler_f32(const DtArgs& a, const float* u0, hipStream_t st);

struct FinalizeArgs {
  const double* blockmin;
  int nblocks;
  double cfl;
  int use_fixed;
  double fixed_dt;
  int clip;
  double t_end;
  double* dt;       // out: this step's dt
  double* t;        // in/out: simulated time
  long long* steps; // in/out
  unsigned* health; // in/out
};
void launch_finalize_dt(const FinalizeArgs& f, hipStream_t st);

// CFL of Euler Roe/HLL/HLLC on 1D/2D grids (strip march): one minimum per
// wave into wavemin[nwaves] (NaN where a speed was not finite), reduced by
// launch_finalize_dt over nblocks = nwaves
constexpr int kDtRows = 10;  // rows per wave of the strip-march CFL kernel (at most)
constexpr int kDt3Rows = 8;     // rows per wave of the 3D CFL z-march (FP32)
constexpr int kDt3RowsF64 = 4;  // ... (FP64: twice the registers per value)
struct DtMarchArgs {
  int dim, sx, sy, sz;
  long long plane;
  double kdx[3];
  double gamma;
  int solver;                   // 0 Roe, 1 HLL, 2 HLLC, 3 SRHD Roe (reads the primitives)
  int last_x, last_y, last_z;   // last cell with a dt: S-3 (Roe), S-2 (HLL)
  int nstrips, rows, nchunks, nwaves;
  int zrows, nzchunks;          // 3D: planes per z-chunk and their count
  double* wavemin;
};
void launch_dt_march_f64(const DtMarchArgs& a, const double* u0, hipStream_t st);
void launch_dt_march_f32(const DtMarchArgs& a, const float* u0, hipStream_t st);
// SRHD (solver 3) over FP64 primitive planes; fp32: speeds in FP32
void launch_dt_march_srhd(const DtMarchArgs& a, const double* prim, bool fp32, hipStream_t st);
// SRHD over FP32 primitive planes (the FP32-mixed simulation's storage), FP32 speeds
void launch_dt_march_srhd(const DtMarchArgs& a, const float* prim, hipStream_t st);
// 3D (Euler) z-march: waves = strips x row chunks x z-chunks, strips fastest
void launch_dt_march3_f64(const DtMarchArgs& a, const double* u0, hipStream_t st);
void launch_dt_march3_f32(const DtMarchArgs& a, const float* u0, hipStream_t st);

// The head of an update in one launch: boundary and CFL strip march (1D/2D,
// or 3D Euler).  Workgroups 0 .. mblocks-1 march, reading ghost cells
// through the boundary's composed map so that they need not wait for the
// bblocks others, which write the ghost cells (and the SRHD primitives'
// ghosts when prim is set); each marching workgroup leaves one minimum in
// a.wavemin for the dt tail (launch_finalize_dt over mblocks minima).
struct HeadArgs {
  BoundaryArgs b;
  DtMarchArgs a;
  int bblocks, mblocks;
};
void launch_step_head(const HeadArgs& h, void* planes, float* lo, void* prim, bool fp32, hipStream_t st);

// Whole updates of a small 1D Euler grid in one workgroup (run1d.hip): per
// step the boundary, the CFL march, the dt tail and every RK stage, with
// workgroup barriers between them.  st[k] are the stage arguments with
// u0 = U[0], unew = U[1]; odd steps swap the two.
constexpr int kRun1dMaxStages = 4;
// One CU's arithmetic bounds the single-workgroup run: beyond these strip
// counts (measured crossover) the launch-per-kernel path, which spreads a
// stage over many CUs, is faster.
constexpr int kRun1dMaxStripsF64 = 24;
constexpr int kRun1dMaxStripsF32 = 48;
constexpr int kRun1dThreads = 1024;
template <typename T, typename S>
struct Run1dArgs {
  StageArgs<T, S> st[kRun1dMaxStages];
  int nstages;
  int nsteps;
  S* U[2];
  float* lo[2];         // compensated storage low words, or null
  BoundaryArgs b;
  int march;            // 0: fixed dt (no CFL pass)
  DtMarchArgs dt;       // the 1D strip march (wavemin unused)
  FinalizeArgs f;       // the dt tail (blockmin unused)
  int post_boundary;    // boundary on the new state after each update
};
// kind: 0 Roe, 1 HLL, 2 HLLC (Euler)
void launch_run1d_f64(int kind, int lim, const Run1dArgs<double, double>& a, hipStream_t st);
void launch_run1d_f32(int kind, int lim, bool kahan, const Run1dArgs<float, float>& a, hipStream_t st);

// EulerBurgers (burgers.hip): the CFL pass (per-block minima, as the
// generic CFL kernel leaves them), and the pressure passes of its split step
struct BurgersArgs {
  int dim, sx, sy, sz;
  long long plane;
  double gamma;
  double inv_dx[3];   // 1 / kernel dx per axis
  const double* dt;   // device-resident step size
  double work_coef;   // what the tableau's rows sum a constant derivative to
  const void* pot;    // the potential plane (state precision), or null
  const void* solid;  // the solid mask plane (state precision), or null
  // momentum pass: its start state and the new state (unshifted planes;
  // low words of compensated storage, or null)
  const void* u0;
  const float* u0lo;
  void* unew;
  float* unewlo;
};
void launch_dt_burgers_f64(const DtArgs& a, const double* u, hipStream_t st);
void launch_dt_burgers_f32(const DtArgs& a, const float* u, hipStream_t st);
// computePressure of u on cells 1 .. size-2 into p0, and the cells with a
// coordinate 1 or size-2 into p1 too
void launch_burgers_pressure_f64(const BurgersArgs& b, const double* u, double* p0, double* p1, hipStream_t st);
void launch_burgers_pressure_f32(const BurgersArgs& b, const float* u, float* p0, float* p1, hipStream_t st);
// one RK stage of the momentum pass: a is the stage argument block with every
// plane pointer advanced by one plane (the x-momentum first); reads the stage
// state's pressure pin, writes the next stage state's pressure pout
void launch_burgers_momentum_f64(int mode, const StageArgs<double, double>& a, const BurgersArgs& b,
                                 const double* pin, double* pout, hipStream_t st);
void launch_burgers_momentum_f32(int mode, bool kahan, const StageArgs<float, float>& a, const BurgersArgs& b,
                                 const float* pin, float* pout, hipStream_t st);
// the work pass on u in place (lo: compensated storage low words, or null),
// pressures p of the last momentum stage state
void launch_burgers_work_f64(const BurgersArgs& b, double* u, const double* p, hipStream_t st);
void launch_burgers_work_f32(const BurgersArgs& b, float* u, float* lo, const float* p, hipStream_t st);

// self-gravity (gravity.hip)
struct GravityArgs {
  int dim, sx, sy, sz;
  long long plane;
  double kdx[3];            // kernel dx per axis
  double diag;              // -2 sum_a 1 / kdx_a^2
  double four_pi_g;         // 4 pi G
  double inv_2dx[3];        // 1 / (2 kdx_a)
  signed char method[3][2]; // the potential's boundary per side: 0 periodic, 2 free-flow, -1 none
  const double* dt;         // device-resident step size
};
// one relaxation sweep phi -> out (every cell: interior relaxed, ghosts by
// the potential's boundary from phi); rho is the density plane
void launch_potential_relax_f64(const GravityArgs& g, const double* rho, const double* phi, double* out,
                                hipStream_t st);
void launch_potential_relax_f32(const GravityArgs& g, const float* rho, const float* phi, float* out,
                                hipStream_t st);
// one RK stage of the gravity source pass: a advanced by one plane (momentum
// and energy), u0 / unew the pass's start and new state (unshifted; low
// words of compensated storage or null), phi the potential plane
void launch_gravity_stage_f64(int mode, const StageArgs<double, double>& a, const GravityArgs& g, const double* u0,
                              double* unew, const double* phi, hipStream_t st);
void launch_gravity_stage_f32(int mode, const StageArgs<float, float>& a, const GravityArgs& g, const float* u0,
                              const float* u0lo, float* unew, float* unewlo, const float* phi, hipStream_t st);
// E += phi on every cell (lo: compensated low words, or null)
void launch_add_potential_energy_f64(long long vol, double* E, const double* phi, hipStream_t st);
void launch_add_potential_energy_f32(long long vol, float* E, float* lo, const float* phi, hipStream_t st);
// dst = src (or -src) over n elements
void launch_copy_plane_f64(long long n, const double* src, double* dst, bool negate, hipStream_t st);
void launch_copy_plane_f32(long long n, const float* src, float* dst, bool negate, hipStream_t st);

// buffer operations of host-side integrators: a = b + c * k over n
// elements (multAdd), and a zero fill of n 32-bit words
void launch_mult_add_f64(double* a, const double* b, const double* c, double k, long long n, hipStream_t st);
void launch_mult_add_f32(float* a, const float* b, const float* c, float k, long long n, hipStream_t st);
void launch_zero_words(void* p, long long nwords, hipStream_t st);

// the reference's FP64 array of structures aos[s + n*cell] (a device
// buffer) to / from the state planes (FP32: hi, and the low words when lo
// is not null)
void launch_aos_to_planes_f64(const double* aos, int n, long long vol, long long plane, double* u, hipStream_t st);
void launch_aos_to_planes_f32(const double* aos, int n, long long vol, long long plane, float* hi, float* lo,
                              hipStream_t st);
void launch_planes_to_aos_f64(const double* u, int n, long long vol, long long plane, double* aos, hipStream_t st);
void launch_planes_to_aos_f32(const float* hi, const float* lo, int n, long long vol, long long plane, double* aos,
                              hipStream_t st);

// FP64 interior sums, per block: partial[b * n + s]
void launch_totals_f64(int dim, int sx, int sy, int sz, long long plane, int n, const double* u,
                       double* partial, int nblocks, hipStream_t st);
void launch_totals_f32(int dim, int sx, int sy, int sz, long long plane, int n, const float* u,
                       const float* lo, double* partial, int nblocks, hipStream_t st);

}  // namespace hydro
