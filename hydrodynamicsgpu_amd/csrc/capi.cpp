// capi.cpp -- C ABI of the MI355X update path (include/hydro_capi.h).
//
// Owns the device buffers of one grid and sequences the kernels of a step:
//   boundary (every ghost cell, one launch) -> CFL minima per wave ->
//   finalise dt on the device -> S fused RK stages (frozen start-of-step
//   data recomputed per stage) -> [SRHD: constrain + con2prim].
// The new state's ghost cells are copied only when something reads them.
// dt never leaves the device on the run() path (captured graphs, or one
// single-workgroup launch for small 1D grids); the only host sync is the
// one a caller asks for (update() returns dt, get_state, totals).  The
// generic entry points at the end serve host-side integrators.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hydro_capi.h"
#include "cfl.hpp"
#include "host_common.hpp"
#include "launch.hpp"
#include "stage_args.hpp"

using namespace hydro;

thread_local std::string hydro::host::g_err;

namespace {

using namespace hydro::host;

}  // namespace



namespace hydro {
namespace host {

CflGrid cfl_grid(const hydro_sim_desc& d, long long plane) {
  CflGrid g;
  g.d = &d;
  g.dim = d.dim;
  for (int a = 0; a < 3; ++a) g.sz[a] = a < d.dim ? d.size[a] : 1;
  g.plane = plane;
  for (int a = 0; a < 3; ++a) {
    // src/HydroGPUApp.cpp:224-226: (xmax - xmin) / (float)size
    const double szf = a < d.dim ? (double)(float)d.size[a] : 1.0;
    const double dx = (d.xmax[a] - d.xmin[a]) / szf;
    g.kdx[a] = kernel_dx(dx, (d.flags & HYDRO_FLAG_REF_DX_6DIGITS) != 0);
  }
  g.cfl = (double)(float)d.cfl;  // HydroGPUApp.h:84 stores cfl as float
  g.mixed = d.precision == HYDRO_PRECISION_FP32_MIXED;
  return g;
}

DtMarchArgs cfl_march_args(const CflGrid& g, double* minima) {
  const hydro_sim_desc& d = *g.d;
  DtMarchArgs a{};
  a.dim = g.dim;
  a.sx = g.sz[0];
  a.sy = g.sz[1];
  a.plane = g.plane;
  a.kdx[0] = g.kdx[0];
  a.kdx[1] = g.kdx[1];
  a.gamma = d.gamma;
  a.solver = d.equation == HYDRO_EQN_SRHD ? 3 : d.solver;
  const int hi = d.solver == HYDRO_SOLVER_ROE ? 3 : 2;  // cells 2 .. S-hi
  a.last_x = g.sz[0] - hi;
  a.last_y = g.dim > 1 ? g.sz[1] - hi : 0;
  a.nstrips = (a.last_x - 2) / 60 + 1;
  if (g.dim > 1) {
    const long long ny = a.last_y - 1;
    // up to kDtRows rows per wave, fewer when that leaves SIMDs idle
    a.rows = kDtRows;
    while (a.rows > 2 && (long long)a.nstrips * ((ny + a.rows - 1) / a.rows) < 4096) --a.rows;
    if (g.dt_rows > 0) {
      const int r = std::min(kDtRows, std::max(2, g.dt_rows));
      if ((long long)a.nstrips * ((ny + r - 1) / r) <= kDtMaxWaves) a.rows = r;  // slots in the minima
    }
    a.nchunks = (int)((ny + a.rows - 1) / a.rows);
  } else {
    a.rows = 1;
    a.nchunks = 1;
  }
  a.nwaves = a.nstrips * a.nchunks;
  a.wavemin = minima;
  return a;
}

namespace {

// which CFL kernel a grid uses: 1 = 1D/2D strip march, 2 = 3D strip march,
// 3 = block minima, 0 = none (fixed dt)
int cfl_path(const CflGrid& g) {
  const hydro_sim_desc& d = *g.d;
  if (d.use_fixed_dt) return 0;
  if (d.solver == HYDRO_SOLVER_BURGERS) return 3;  // cell-local speeds: block minima
  if (g.pot || g.solid) return 3;  // speeds with the cells' potentials or solid cells: the generic kernel
  const int hi = d.solver == HYDRO_SOLVER_ROE ? 3 : 2;
  const long long strips = (g.sz[0] - hi - 2) / 60 + 1;
  const bool srhd = d.equation == HYDRO_EQN_SRHD;
  if (g.dim <= 2 && (d.equation == HYDRO_EQN_EULER || srhd)) {
    const long long waves = strips * (g.dim > 1 ? (g.sz[1] - hi - 1 + kDtRows - 1) / kDtRows : 1);
    if (waves <= kDtMaxWaves) return 1;
  }
  if (g.dim == 3 && d.equation == HYDRO_EQN_EULER &&
      strips * ((g.sz[1] - hi - 1 + kDt3RowsF64 - 1) / kDt3RowsF64) <= kDtMaxWaves)
    return 2;
  return 3;
}

DtMarchArgs march3_args(const CflGrid& g, double* minima) {
  const hydro_sim_desc& d = *g.d;
  DtMarchArgs a{};
  a.dim = 3;
  a.sx = g.sz[0];
  a.sy = g.sz[1];
  a.sz = g.sz[2];
  a.plane = g.plane;
  for (int i = 0; i < 3; ++i) a.kdx[i] = g.kdx[i];
  a.gamma = d.gamma;
  a.solver = d.solver;
  const int hi = d.solver == HYDRO_SOLVER_ROE ? 3 : 2;  // cells 2 .. S-hi
  a.last_x = g.sz[0] - hi;
  a.last_y = g.sz[1] - hi;
  a.last_z = g.sz[2] - hi;
  a.nstrips = (a.last_x - 2) / 60 + 1;
  const long long ny = a.last_y - 1, nz = a.last_z - 1;
  a.rows = g.mixed ? kDt3Rows : kDt3RowsF64;  // the device function's R
  a.nchunks = (int)((ny + a.rows - 1) / a.rows);
  // z-chunks of 8 planes or more: about four waves per SIMD, and no more
  // waves than minima slots
  a.zrows = 8;
  auto waves = [&](int zr) { return (long long)a.nstrips * a.nchunks * ((nz + zr - 1) / zr); };
  while (a.zrows < nz && (waves(a.zrows) > 4096 || waves(a.zrows) > kDtMaxWaves)) a.zrows *= 2;
  a.nzchunks = (int)((nz + a.zrows - 1) / a.zrows);
  a.nwaves = (int)waves(a.zrows);
  a.wavemin = minima;
  return a;
}

}  // namespace

long long cfl_minima(const CflGrid& g) {
  switch (cfl_path(g)) {
  case 1: return cfl_march_args(g, nullptr).nwaves;
  case 2: return march3_args(g, nullptr).nwaves;
  case 3: return g.dt_blocks;
  default: return 0;
  }
}

int enqueue_cfl(const CflGrid& g, const void* U, const void* prim, bool prim_f64, double* minima, FinalizeArgs f,
                hipStream_t st) {
  const hydro_sim_desc& d = *g.d;
  f.blockmin = minima;
  f.nblocks = 0;
  switch (cfl_path(g)) {
  case 1: {
    // strip march, one minimum per wave, then the one-block tail
    const DtMarchArgs a = cfl_march_args(g, minima);
    if (d.equation == HYDRO_EQN_SRHD && prim_f64) launch_dt_march_srhd(a, (const double*)prim, g.mixed, st);
    else if (d.equation == HYDRO_EQN_SRHD) launch_dt_march_srhd(a, (const float*)prim, st);
    else if (g.mixed) launch_dt_march_f32(a, (const float*)U, st);
    else launch_dt_march_f64(a, (const double*)U, st);
    f.nblocks = a.nwaves;
    break;
  }
  case 2: {
    // 3D strip march: waves = strips x row chunks x planes
    const DtMarchArgs a = march3_args(g, minima);
    if (g.mixed) launch_dt_march3_f32(a, (const float*)U, st);
    else launch_dt_march3_f64(a, (const double*)U, st);
    f.nblocks = a.nwaves;
    break;
  }
  case 3: {
    DtArgs a{};
    a.dim = g.dim;
    a.sx = g.sz[0];
    a.sy = g.sz[1];
    a.sz = g.sz[2];
    a.plane = g.plane;
    for (int i = 0; i < 3; ++i) a.kdx[i] = g.kdx[i];
    a.gamma = d.gamma;
    a.solver = d.solver;
    a.eqn = d.equation;
    a.prim = prim;
    a.prim_f32 = prim && !prim_f64;
    double mx[5];
    maxwell_constants(d, mx);
    a.maxwell_lambda = mx[4];
    a.nblocks = g.dt_blocks;
    a.blockmin = minima;
    a.pot = g.pot;
    a.solid = g.solid;
    if (d.solver == HYDRO_SOLVER_BURGERS) {
      // EulerBurgers::calcTimestep (src/Solver/EulerBurgers.cpp:61-65)
      if (g.mixed) launch_dt_burgers_f32(a, (const float*)U, st);
      else launch_dt_burgers_f64(a, (const double*)U, st);
    } else if (g.mixed) {
      launch_dt_euler_f32(a, (const float*)U, st);
    } else {
      launch_dt_euler_f64(a, (const double*)U, st);
    }
    f.nblocks = g.dt_blocks;
    break;
  }
  default: break;
  }
  launch_finalize_dt(f, st);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

}  // namespace host
}  // namespace hydro

struct hydro_sim {
  hydro_ctx* ctx = nullptr;
  hydro_sim_desc d{};
  int dim = 1, N = 3;
  int sz[3] = {1, 1, 1};
  long long vol = 0, plane = 0;
  bool mixed = false, kahan = false;
  double kdx[3] = {1, 1, 1};
  double cfl = 0.5;
  hipStream_t stream = nullptr;
  Tableau tab;
  std::vector<StagePlan> plan;
  int nslots = 0;
  // device buffers (element type double for FP64, float for FP32-mixed)
  void* U[2] = {nullptr, nullptr};
  float* Ulo[2] = {nullptr, nullptr};
  void* D[3] = {nullptr, nullptr, nullptr};  // stage increments
  int ndbuf = 0;
  bool replay = false;
  void* P[kMaxPartials] = {nullptr, nullptr, nullptr};
  void* prim = nullptr;  // SRHD primitive planes (the state's precision: double, or float when mixed)
  void* pres[2] = {nullptr, nullptr};  // EulerBurgers: pressure planes of the momentum pass's stage states
  // Euler: the gravitational potential (pot[0]; pot[1] is the relaxation's
  // second plane), in the state's precision; pot_on when it is non-zero or
  // relaxed from the density (use_gravity): then the kernels read it and
  // every update adds the gravity pass
  void* pot[2] = {nullptr, nullptr};
  bool pot_on = false;
  // Euler: the solid mask plane (state precision; per cell bit 0 the cell's
  // own solid flag, bits 1+2a / 2+2a its lower / upper neighbour's along
  // axis a), read by the Roe and EulerBurgers kernels when solid_on
  void* solid = nullptr;
  bool solid_on = false;
  double work_coef = 1.0;              // EulerBurgers: the tableau's sum for a constant derivative
  double* blockmin = nullptr;
  double* partial = nullptr;
  DevScalars* sc = nullptr;
  DevScalars* host_sc = nullptr;  // pinned
  int cur = 0;
  bool ghosts_pending = false;  // U[cur]'s ghost cells still to be copied from U[cur ^ 1]
  int dt_blocks = 1024;
  int tot_blocks = 256;
  // stage decomposition
  int rows = 32, wpb = 4;
  int zwaves = 8;  // first-order 3D kernel: waves per workgroup
  int rwaves = 8;  // 3D Roe-family kernel: rows (waves) per workgroup, 8 or 16 (FP32)
  int zshare = 0;  // first-order 3D kernel: modes whose rows go through LDS (bit per CombineMode)
  int zmarch = 0;  // first-order 3D kernel: z-march form with this many cells per lane (0: y-march form)
  int form3d = 0;  // first-order 3D kernel: one of the kForm3d* forms
  int ncu = 256;   // compute units of the device
  int dt_rows = 0;  // 2D CFL strip march: rows per wave (0: chosen from the grid size)
  CflGrid cg;       // the CFL pass's view of this grid
  // measurement: event pairs around fused stage launches
  bool profiling = false;
  // two captured updates per starting parity (ping-pong state), replayed by run()
  hipGraphExec_t graph2[2] = {nullptr, nullptr};
  // longer captures: graph_updates updates (even) per submission
  int graph_updates = 2;
  hipGraphExec_t graphN[2] = {nullptr, nullptr};
  bool graph_pending[2] = {false, false};  // ghosts_pending after the captured pair
  bool use_graphs = true;
  bool use_run1d = true;  // small 1D Euler grids: whole runs in one workgroup
  bool use_head = true;   // boundary + CFL march of an update in one launch
  // generic stage evaluation: buffers handed out to a host-side integrator,
  // and whether U[cur ^ 1] holds the frozen start-of-step state
  std::vector<void*> user_bufs;
  bool frozen = false;
  // device staging of the reference's FP64 array-of-structures state, for
  // the layout change on the GPU (allocated on first use)
  double* aos_dev = nullptr;
  // measurement: event pairs bracketing the phases of updates
  struct PhaseEvent {
    int phase;
    hipEvent_t e0, e1;
  };
  std::vector<PhaseEvent> ev;
};

namespace {

size_t esize(const hydro_sim* s) { return s->mixed ? sizeof(float) : sizeof(double); }

// ---- phase timing (hydro_sim_set_profiling)
// roctx ranges through the ROCm tracer library, loaded on first use
struct Roctx {
  int (*push)(const char*) = nullptr;
  int (*pop)() = nullptr;
};
const Roctx& roctx() {
  static const Roctx r = [] {
    Roctx x;
    void* h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("librocprofiler-sdk-roctx.so", RTLD_NOW | RTLD_LOCAL);
    if (h) {
      x.push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
      x.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
      if (!x.push || !x.pop) x.push = nullptr, x.pop = nullptr;
    }
    return x;
  }();
  return r;
}
const char* const kPhaseRange[HYDRO_NUM_PHASES] = {
    "hydro:boundary", "hydro:cfl", "hydro:head", "hydro:dt_tail", "hydro:stages",
    "hydro:con2prim", "hydro:pressure", "hydro:ghosts", "hydro:totals", "hydro:gravity"};

// Brackets the work a scope enqueues on the sim's stream with two events
// (and a roctx range) while profiling is enabled; nothing otherwise.
class PhaseScope {
 public:
  PhaseScope(hydro_sim* s, int phase) : s_(s), phase_(phase) {
    if (!s_->profiling) return;
    if (hipEventCreate(&e0_) != hipSuccess || hipEventRecord(e0_, s_->stream) != hipSuccess) {
      if (e0_) hipEventDestroy(e0_);
      e0_ = nullptr;
      return;
    }
    if (roctx().push) roctx().push(kPhaseRange[phase_]);
  }
  ~PhaseScope() {
    if (!e0_) return;
    if (roctx().pop) roctx().pop();
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e1) == hipSuccess && hipEventRecord(e1, s_->stream) == hipSuccess) {
      s_->ev.push_back({phase_, e0_, e1});
    } else {
      hipEventDestroy(e0_);
      if (e1) hipEventDestroy(e1);
    }
  }
  PhaseScope(const PhaseScope&) = delete;
  PhaseScope& operator=(const PhaseScope&) = delete;

 private:
  hydro_sim* s_;
  int phase_;
  hipEvent_t e0_ = nullptr;
};

// Euler::stateGetBoundaryKernelForBoundaryMethod (src/Equation/Euler.cpp:63-75),
// SRHD (src/Equation/SRHD.cpp:63-75), Maxwell (src/Equation/Maxwell.cpp:73-85)
BoundaryArgs boundary_args(const hydro_sim* s) {
  BoundaryArgs b{};
  b.dim = s->dim;
  b.sx = s->sz[0];
  b.sy = s->sz[1];
  b.sz = s->sz[2];
  b.plane = s->plane;
  b.nstates = s->N;
  for (int axis = 0; axis < 3; ++axis)
    for (int mm = 0; mm < 2; ++mm) {
      signed char m = -1;
      unsigned neg = 0;
      if (axis < s->dim) {
        switch (s->d.boundary[axis][mm]) {
        case HYDRO_BOUNDARY_PERIODIC: m = 0; break;
        case HYDRO_BOUNDARY_MIRROR:
          m = 1;
          if (s->d.equation == HYDRO_EQN_MAXWELL)  // src/Equation/Maxwell.cpp:80: normal E and B
            neg = (1u << axis) | (1u << (3 + axis));
          else  // src/Equation/Euler.cpp:70, src/Equation/SRHD.cpp:70: normal momentum
            neg = 1u << (axis + 1);
          break;
        case HYDRO_BOUNDARY_FREEFLOW: m = 2; break;
        default: m = -1;
        }
      }
      b.method[axis][mm] = m;
      b.negmask[axis][mm] = neg;
    }
  return b;
}

// first-order 3D stage forms: y-march over 4 z-planes per workgroup;
// z-march with one cell per lane (16-wave workgroups in FP32, 8 in FP64),
// two cells per lane (FP32, 8 waves), one cell per lane in 8 waves
enum { kForm3dYMarch = 0, kForm3dZm1 = 1, kForm3dZm2 = 2, kForm3dZm1w8 = 3 };

// applies a form with its default decomposition (zwaves: the y-march form's
// workgroup height, 0 for the default)
void set_form3d(hydro_sim* s, int form, int zwaves = 0) {
  s->form3d = form;
  if (form == kForm3dYMarch) {
    const long long strips = (s->sz[0] - 4 + 59) / 60;
    s->zmarch = 0;
    s->zwaves = zwaves ? zwaves : stage3d_r1_waves(!s->mixed);
    const long long nzb = (s->sz[2] - 4 + s->zwaves - 2) / (s->zwaves - 1);
    s->rows = pick_rows_3d_r1(strips * nzb, s->sz[1] - 4, s->ncu * (8 / s->zwaves));
    return;
  }
  s->zmarch = (form == kForm3dZm2 && s->mixed) ? 2 : 1;
  s->zwaves = (form == kForm3dZm1 && s->mixed) ? 16 : 8;
  const int cpl = s->zmarch;
  const long long zstrips = (s->sz[0] - 4 + 64 * cpl - 3) / (64 * cpl - 2);
  const long long yblocks = (s->sz[1] - 4 + s->zwaves - 3) / (s->zwaves - 2);
  const long long lds = (long long)s->zwaves * (3 * s->N + 4) * 64 * cpl * (long long)esize(s);
  const long long per_cu = std::max(1LL, std::min(160LL * 1024 / lds, 32LL / s->zwaves));
  s->rows = pick_planes_zm(zstrips * yblocks, s->sz[2] - 4, s->ncu * per_cu);
}

// z-planes per chunk of the 3D Roe-family kernel with workgroups of `waves`
// rows: the longest chunks (64 planes at most) that still give about four
// workgroups per CU
int pick_planes_3d_roe(const hydro_sim* s, int waves) {
  const long long strips = (s->sz[0] - 4 + 59) / 60;
  const int ry = stage3d_rows_per_block(waves);
  const long long ych = (s->sz[1] - 4 + ry - 1) / ry;
  int rows = 64;
  while (rows > 8 && strips * ych * ((s->sz[2] - 4 + rows - 1) / rows) < 1024) rows >>= 1;
  return rows;
}

// true when the boundary rewrites every ghost cell (no side is NONE)
bool boundary_covers_ghosts(const hydro_sim* s) {
  for (int a = 0; a < s->dim; ++a)
    for (int m = 0; m < 2; ++m)
      if (s->d.boundary[a][m] == HYDRO_BOUNDARY_NONE) return false;
  return true;
}

// The stage kernels write interior cells only; the new state's ghost cells
// are those of u0 (their derivative is zero).  Copying them is deferred until
// something reads them: the next boundary() rewrites every ghost cell from
// interior values, so in a run of updates the copy is never needed.
int materialize_ghosts(hydro_sim* s) {
  if (!s->ghosts_pending) return HYDRO_OK;
  PhaseScope ph(s, HYDRO_PHASE_GHOSTS);
  const int prev = s->cur ^ 1;
  if (s->mixed) {
    launch_copy_ghosts_f32(s->dim, s->sz[0], s->sz[1], s->sz[2], s->plane, s->N, (const float*)s->U[prev],
                           (float*)s->U[s->cur], s->stream);
    if (s->kahan)
      launch_copy_ghosts_f32(s->dim, s->sz[0], s->sz[1], s->sz[2], s->plane, s->N, s->Ulo[prev], s->Ulo[s->cur],
                             s->stream);
  } else {
    launch_copy_ghosts_f64(s->dim, s->sz[0], s->sz[1], s->sz[2], s->plane, s->N, (const double*)s->U[prev],
                           (double*)s->U[s->cur], s->stream);
  }
  HIP_TRY(hipGetLastError());
  s->ghosts_pending = false;
  return HYDRO_OK;
}

int enqueue_boundary(hydro_sim* s) {
  PhaseScope ph(s, HYDRO_PHASE_BOUNDARY);
  if (boundary_covers_ghosts(s)) {
    s->ghosts_pending = false;
  } else {
    const int rc = materialize_ghosts(s);
    if (rc) return rc;
  }
  const BoundaryArgs b = boundary_args(s);
  // SRHDRoe::boundary fills the primitives with the same kernels afterwards
  // (src/Solver/SRHDRoe.cpp:79-96): the same launch here
  if (s->mixed)
    launch_boundary_f32(b, (float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr, (float*)s->prim, s->stream);
  else
    launch_boundary_f64(b, (double*)s->U[s->cur], (double*)s->prim, s->stream);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

// The CFL strip march of a 1D/2D Euler or SRHD grid: waves = strips x row chunks
DtMarchArgs march_args(const hydro_sim* s) { return cfl_march_args(s->cg, s->blockmin); }

// The dt tail; advance_time false leaves the clock and the step count in a
// scratch copy (calc_timestep)
FinalizeArgs finalize_args(hydro_sim* s, bool advance_time) {
  FinalizeArgs f{};
  f.blockmin = s->blockmin;
  f.nblocks = s->d.use_fixed_dt ? 0 : s->dt_blocks;
  f.cfl = s->cfl;
  f.use_fixed = s->d.use_fixed_dt;
  f.fixed_dt = s->d.fixed_dt;
  f.clip = (s->d.flags & HYDRO_FLAG_CLIP_TO_TEND) ? 1 : 0;
  f.t_end = s->d.t_end;
  f.dt = &s->sc->dt;
  f.t = advance_time ? &s->sc->t : &s->sc[1].t;
  f.steps = advance_time ? &s->sc->steps : &s->sc[1].steps;
  f.health = &s->sc->health;
  return f;
}

// SRHD primitive planes: stored in the state's precision (the FP32-mixed
// form's fluxes and speeds read them as FP32; con2prim computes in FP64)
size_t prim_bytes(const hydro_sim* s) {
  return (s->mixed ? sizeof(float) : sizeof(double)) * (size_t)s->plane * (size_t)s->N;
}

int upload_prims(hydro_sim* s, const std::vector<double>& planes) {
  if (!s->mixed)
    return copy_sync(s->prim, planes.data(), prim_bytes(s), hipMemcpyHostToDevice, s->stream) == hipSuccess
               ? HYDRO_OK
               : fail(HYDRO_E_HIP, "copy (primitives)");
  std::vector<float> f(planes.begin(), planes.end());
  return copy_sync(s->prim, f.data(), prim_bytes(s), hipMemcpyHostToDevice, s->stream) == hipSuccess
             ? HYDRO_OK
             : fail(HYDRO_E_HIP, "copy (primitives)");
}

int download_prims(hydro_sim* s, std::vector<double>& planes) {
  planes.assign((size_t)s->plane * s->N, 0.0);
  if (!s->mixed)
    return copy_sync(planes.data(), s->prim, prim_bytes(s), hipMemcpyDeviceToHost, s->stream) == hipSuccess
               ? HYDRO_OK
               : fail(HYDRO_E_HIP, "copy (primitives)");
  std::vector<float> f(planes.size());
  if (copy_sync(f.data(), s->prim, prim_bytes(s), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    return fail(HYDRO_E_HIP, "copy (primitives)");
  std::copy(f.begin(), f.end(), planes.begin());
  return HYDRO_OK;
}

int enqueue_dt(hydro_sim* s, bool advance_time) {
  PhaseScope ph(s, HYDRO_PHASE_CFL);
  return enqueue_cfl(s->cg, s->U[s->cur], s->prim, !s->mixed, s->blockmin, finalize_args(s, advance_time), s->stream);
}

template <typename T>
void fill_stage_args(hydro_sim* s, int k, StageArgs<T, T>& a) {
  const StagePlan& p = s->plan[k];
  a.u0 = (const T*)s->U[s->cur];
  a.u0lo = s->kahan ? s->Ulo[s->cur] : nullptr;
  a.dk = p.din < 0 ? nullptr : (const T*)s->D[p.din];
  a.dnext = p.dout < 0 ? nullptr : (T*)s->D[p.dout];
  a.nrep = p.nrep;
  for (int j = 0; j < kMaxPartials; ++j) {
    a.rep[j] = j < p.nrep ? (const T*)s->D[p.rep[j]] : nullptr;
    a.rc[j] = p.rc[j];
  }
  a.unew = (T*)s->U[s->cur ^ 1];
  a.unewlo = s->kahan ? s->Ulo[s->cur ^ 1] : nullptr;
  a.pkp = p.pk < 0 ? nullptr : (const T*)s->P[p.pk];
  for (int j = 0; j < kMaxPartials; ++j) {
    a.fp[j] = j < p.nf ? (T*)s->P[p.fslot[j]] : nullptr;
    a.fload[j] = p.fload[j];
    a.am[j] = p.am[j];
    a.bm[j] = p.bm[j];
  }
  a.nf = p.nf;
  a.a_self = p.a_self;
  a.b_self = p.b_self;
  a.final_stage = p.final_stage ? 1 : 0;
  a.dt = &s->sc->dt;
  a.sx = s->sz[0];
  a.sy = s->sz[1];
  a.sz = s->sz[2];
  a.plane = s->plane;
  for (int i = 0; i < 3; ++i) a.inv_dx[i] = 1.0 / s->kdx[i];
  a.gamma = s->d.gamma;
  a.prim = (const T*)s->prim;
  const int kind = kernel_kind(s->d);
  // the solid Roe form always reads the potential (a zero plane without one)
  a.pot = ((s->pot_on && (kind == 0 || kind == 1)) || (s->solid_on && kind == 0)) ? (const T*)s->pot[0] : nullptr;
  a.solid = (s->solid_on && (kind == 0 || kind == 5)) ? (const T*)s->solid : nullptr;
  maxwell_constants(s->d, a.mx);
  a.lim = s->d.limiter;
  const int interior_x = s->sz[0] - 4;
  a.nstrips = (interior_x + 59) / 60;
  if (s->dim == 1) {
    a.rows = 1;
    a.nchunks = 1;
    a.nwaves = a.nstrips;
  } else if (s->dim == 2) {
    a.rows = s->rows;
    a.nchunks = (s->sz[1] - 4 + a.rows - 1) / a.rows;
    a.nwaves = a.nstrips * a.nchunks;
  } else if (kernel_kind(s->d) == 1 || kernel_kind(s->d) == 2) {
    // 3D first order: rows = y rows per chunk, workgroups of W-1 z-planes
    a.rows = s->rows;
    a.nchunks = (s->sz[1] - 4 + a.rows - 1) / a.rows;
    a.zwaves = s->zwaves;
    a.zshare = s->zshare;
    a.zmarch = s->zmarch;
    if (s->zmarch) {
      // z-march: strips of 64 * cells-per-lane - 2 output columns, y-blocks
      // of W - 2 output rows, z-chunks of `rows` planes
      const int cpl = s->zmarch;
      a.nstrips = (interior_x + 64 * cpl - 3) / (64 * cpl - 2);
      a.nchunks = (s->sz[1] - 4 + s->zwaves - 3) / (s->zwaves - 2);
      const int nzc = (s->sz[2] - 4 + a.rows - 1) / a.rows;
      a.nwaves = a.nstrips * a.nchunks * nzc;
    } else {
      const int nzb = (s->sz[2] - 4 + s->zwaves - 2) / (s->zwaves - 1);
      a.nwaves = a.nstrips * a.nchunks * nzb;
    }
  } else {
    // 3D Roe family: rows = z-planes per workgroup, nchunks = y-chunks, nwaves = workgroups
    const int ry = stage3d_rows_per_block(s->rwaves);
    a.zwaves = s->rwaves;
    a.rows = s->rows;
    a.nchunks = (s->sz[1] - 4 + ry - 1) / ry;
    const int nz = (s->sz[2] - 4 + a.rows - 1) / a.rows;
    a.nwaves = a.nstrips * a.nchunks * nz;
  }
}

// one fused stage launch in the given combination mode
void launch_stage(hydro_sim* s, int mode, const StageArgs<float, float>& a) {
  const int kind = kernel_kind(s->d);
  if (s->dim == 3) launch_stage3d_mixed(kind, s->d.limiter, mode, s->kahan, a, s->stream);
  else if (s->kahan) launch_stage2d_f32k(s->dim, kind, s->d.limiter, mode, a, s->wpb, s->stream);
  else launch_stage2d_f32(s->dim, kind, s->d.limiter, mode, a, s->wpb, s->stream);
}
void launch_stage(hydro_sim* s, int mode, const StageArgs<double, double>& a) {
  const int kind = kernel_kind(s->d);
  if (s->dim == 3) launch_stage3d_f64(kind, s->d.limiter, mode, a, s->stream);
  else launch_stage2d_f64(s->dim, kind, s->d.limiter, mode, a, s->wpb, s->stream);
}

// SRHDRoe::step tail: constrainState + updatePrimitives on the current state
// (src/Solver/SRHDRoe.cpp:62-73)
int enqueue_srhd_tail(hydro_sim* s) {
  Con2PrimArgs c{};
  c.dim = s->dim;
  c.sx = s->sz[0];
  c.sy = s->sz[1];
  c.sz = s->sz[2];
  c.plane = s->plane;
  c.gamma = s->d.gamma;
  c.max_iter = s->d.srhd_solve_prim_max_iter;
  c.stop_eps = s->d.srhd_solve_prim_stop_epsilon;
  c.vel_eps = s->d.srhd_solve_prim_vel_epsilon;
  c.pmin_eps = s->d.srhd_solve_prim_pmin_epsilon;
  c.rho_min = s->d.srhd_rho_min;
  c.rho_max = s->d.srhd_rho_max;
  c.eint_max = s->d.srhd_eint_max;
  c.D_min = s->d.srhd_D_min;
  c.D_max = s->d.srhd_D_max;
  c.tau_min = s->d.srhd_tau_min;
  c.tau_max = s->d.srhd_tau_max;
  c.health = &s->sc->health;
  PhaseScope ph(s, HYDRO_PHASE_CON2PRIM);
  // constrainState touches every cell, ghosts included: the deferred ghost
  // copy from the start-of-step state happens in the same launch
  if (s->ghosts_pending) {
    c.ghost_src = s->U[s->cur ^ 1];
    c.ghost_src_lo = s->kahan ? s->Ulo[s->cur ^ 1] : nullptr;
    s->ghosts_pending = false;
  }
  if (s->mixed)
    launch_con2prim_f32(c, (float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr, (float*)s->prim, s->stream);
  else
    launch_con2prim_f64(c, (double*)s->U[s->cur], (double*)s->prim, s->stream);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

// EulerBurgers::step after the advection pass (src/Solver/EulerBurgers.cpp:80-99):
// boundary, the momentum pass (computePressure + diffuseMomentum per stage),
// boundary, the work pass (diffuseWork with the pressure the last momentum
// stage computed), boundary.  The zero-potential pass between the advection
// boundary and the momentum pass is skipped as for EulerRoe (its derivative
// is +0; its ForwardEuler boundary repeats the one just applied).
// ---- self-gravity (src/Solver/SelfGravitation.cpp)
GravityArgs gravity_args(const hydro_sim* s) {
  GravityArgs g{};
  g.dim = s->dim;
  g.sx = s->sz[0];
  g.sy = s->sz[1];
  g.sz = s->sz[2];
  g.plane = s->plane;
  double diagsum = 0.;
  for (int a = 0; a < 3; ++a) {
    g.kdx[a] = s->kdx[a];
    g.inv_2dx[a] = 1.0 / (2.0 * s->kdx[a]);
    if (a < s->dim) diagsum += 1. / (s->kdx[a] * s->kdx[a]);
  }
  g.diag = -2. * diagsum;
  // res/SelfGravitation.cl:60 (OpenCL's M_PI for double)
  g.four_pi_g = 4. * 3.141592653589793115997963468544185161590576171875 * s->d.self_grav_gravitational_constant;
  // SelfGravitationBehavior::gravityGetBoundaryKernelForBoundaryMethod
  // (include/HydroGPU/Equation/SelfGravitationBehavior.h:22-34): mirror
  // sides copy as free-flow
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < 2; ++m) {
      signed char k = -1;
      if (a < s->dim) {
        switch (s->d.boundary[a][m]) {
        case HYDRO_BOUNDARY_PERIODIC: k = 0; break;
        case HYDRO_BOUNDARY_MIRROR:
        case HYDRO_BOUNDARY_FREEFLOW: k = 2; break;
        default: k = -1;
        }
      }
      g.method[a][m] = k;
    }
  g.dt = &s->sc->dt;
  return g;
}

// n relaxation sweeps (each with the potential's boundary before it) from
// the density of the current state; the potential ends in pot[0]
int enqueue_relax(hydro_sim* s, int n) {
  if (n <= 0) return HYDRO_OK;
  const GravityArgs g = gravity_args(s);
  for (int i = 0; i < n; ++i) {
    if (s->mixed)
      launch_potential_relax_f32(g, (const float*)s->U[s->cur], (const float*)s->pot[i & 1], (float*)s->pot[(i + 1) & 1],
                                 s->stream);
    else
      launch_potential_relax_f64(g, (const double*)s->U[s->cur], (const double*)s->pot[i & 1],
                                 (double*)s->pot[(i + 1) & 1], s->stream);
  }
  if (n & 1) {
    if (s->mixed) launch_copy_plane_f32(s->vol, (const float*)s->pot[1], (float*)s->pot[0], false, s->stream);
    else launch_copy_plane_f64(s->vol, (const double*)s->pot[1], (double*)s->pot[0], false, s->stream);
  }
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

// SelfGravitation::applyPotential (src/Solver/SelfGravitation.cpp:99-124):
// the integrator over the gravity source, with gaussSeidelMaxIter
// relaxation sweeps per stage when use_gravity.  Density is unchanged by the
// pass, so every stage relaxes against the current state's density.
// ForwardEuler keeps the boundary the callback applies (the Runge-Kutta
// rebuild of the stage state discards it).
int enqueue_gravity_pass(hydro_sim* s) {
  PhaseScope ph(s, HYDRO_PHASE_GRAVITY);
  const bool fe = s->d.integrator == HYDRO_INTEGRATOR_FORWARDEULER;
  if (fe) {
    const int rc = enqueue_boundary(s);
    if (rc) return rc;
  }
  // the input's ghost cells are valid (after a boundary) or still pending
  // from the buffer the output is written into, which then holds them
  const bool in_pending = s->ghosts_pending;
  const GravityArgs g = gravity_args(s);
  const long long pl = s->plane;
  auto shift = [pl](auto p) { return p ? p + pl : p; };
  for (int k = 0; k < s->tab.order; ++k) {
    if (s->d.use_gravity) {
      const int rc = enqueue_relax(s, s->d.gauss_seidel_max_iter);
      if (rc) return rc;
    }
    const int mode = s->plan[k].mode;
    if (s->mixed) {
      StageArgs<float, float> a{};
      fill_stage_args<float>(s, k, a);
      a.u0 = shift(a.u0); a.u0lo = shift(a.u0lo); a.dk = shift(a.dk); a.dnext = shift(a.dnext);
      a.unew = shift(a.unew); a.unewlo = shift(a.unewlo); a.pkp = shift(a.pkp);
      for (int j = 0; j < kMaxPartials; ++j) { a.fp[j] = shift(a.fp[j]); a.rep[j] = shift(a.rep[j]); }
      launch_gravity_stage_f32(mode, a, g, (const float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr,
                               (float*)s->U[s->cur ^ 1], s->kahan ? s->Ulo[s->cur ^ 1] : nullptr,
                               (const float*)s->pot[0], s->stream);
    } else {
      StageArgs<double, double> a{};
      fill_stage_args<double>(s, k, a);
      a.u0 = shift(a.u0); a.dk = shift(a.dk); a.dnext = shift(a.dnext);
      a.unew = shift(a.unew); a.pkp = shift(a.pkp);
      for (int j = 0; j < kMaxPartials; ++j) { a.fp[j] = shift(a.fp[j]); a.rep[j] = shift(a.rep[j]); }
      launch_gravity_stage_f64(mode, a, g, (const double*)s->U[s->cur], (double*)s->U[s->cur ^ 1],
                               (const double*)s->pot[0], s->stream);
    }
    HIP_TRY(hipGetLastError());
  }
  s->cur ^= 1;
  s->ghosts_pending = !in_pending;
  return HYDRO_OK;
}

// the momentum pass: the pressure of its start state, then one launch per
// stage; stage k reads pres[k & 1] and leaves pres[(k + 1) & 1]
int enqueue_burgers_momentum(hydro_sim* s, BurgersArgs b) {
  PhaseScope ph(s, HYDRO_PHASE_PRESSURE);
  if (s->mixed)
    launch_burgers_pressure_f32(b, (const float*)s->U[s->cur], (float*)s->pres[0], (float*)s->pres[1], s->stream);
  else
    launch_burgers_pressure_f64(b, (const double*)s->U[s->cur], (double*)s->pres[0], (double*)s->pres[1],
                                s->stream);
  HIP_TRY(hipGetLastError());
  b.u0 = s->U[s->cur];
  b.u0lo = s->kahan ? s->Ulo[s->cur] : nullptr;
  b.unew = s->U[s->cur ^ 1];
  b.unewlo = s->kahan ? s->Ulo[s->cur ^ 1] : nullptr;
  const long long pl = s->plane;
  auto shift = [pl](auto p) { return p ? p + pl : p; };
  for (int k = 0; k < s->tab.order; ++k) {
    const int mode = s->plan[k].mode;
    if (s->mixed) {
      StageArgs<float, float> a{};
      fill_stage_args<float>(s, k, a);
      a.u0 = shift(a.u0); a.u0lo = shift(a.u0lo); a.dk = shift(a.dk); a.dnext = shift(a.dnext);
      a.unew = shift(a.unew); a.unewlo = shift(a.unewlo); a.pkp = shift(a.pkp);
      for (int j = 0; j < kMaxPartials; ++j) { a.fp[j] = shift(a.fp[j]); a.rep[j] = shift(a.rep[j]); }
      launch_burgers_momentum_f32(mode, s->kahan, a, b, (const float*)s->pres[k & 1], (float*)s->pres[(k + 1) & 1],
                                  s->stream);
    } else {
      StageArgs<double, double> a{};
      fill_stage_args<double>(s, k, a);
      a.u0 = shift(a.u0); a.dk = shift(a.dk); a.dnext = shift(a.dnext);
      a.unew = shift(a.unew); a.pkp = shift(a.pkp);
      for (int j = 0; j < kMaxPartials; ++j) { a.fp[j] = shift(a.fp[j]); a.rep[j] = shift(a.rep[j]); }
      launch_burgers_momentum_f64(mode, a, b, (const double*)s->pres[k & 1], (double*)s->pres[(k + 1) & 1],
                                  s->stream);
    }
    HIP_TRY(hipGetLastError());
  }
  s->cur ^= 1;
  s->ghosts_pending = true;  // interior written; the ghost cells follow from the boundary
  return HYDRO_OK;
}

// the work pass in place on the current state, with the pressure of the
// momentum pass's last stage state
int enqueue_burgers_work(hydro_sim* s, const BurgersArgs& b) {
  PhaseScope ph(s, HYDRO_PHASE_PRESSURE);
  const int plast = (s->tab.order - 1) & 1;
  if (s->mixed)
    launch_burgers_work_f32(b, (float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr,
                            (const float*)s->pres[plast], s->stream);
  else
    launch_burgers_work_f64(b, (double*)s->U[s->cur], (const double*)s->pres[plast], s->stream);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

int enqueue_burgers_pressure_passes(hydro_sim* s) {
  BurgersArgs b{};
  b.dim = s->dim;
  b.sx = s->sz[0];
  b.sy = s->sz[1];
  b.sz = s->sz[2];
  b.plane = s->plane;
  b.gamma = s->d.gamma;
  for (int i = 0; i < 3; ++i) b.inv_dx[i] = 1.0 / s->kdx[i];
  b.dt = &s->sc->dt;
  b.work_coef = s->work_coef;
  b.pot = s->pot_on ? s->pot[0] : nullptr;
  b.solid = s->solid_on ? s->solid : nullptr;
  int rc = enqueue_boundary(s);
  // the gravity pass between the advection boundary and the momentum pass
  // (src/Solver/EulerBurgers.cpp:80-82)
  if (!rc && s->pot_on) rc = enqueue_gravity_pass(s);
  if (!rc) rc = materialize_ghosts(s);  // computePressure reads the ghost layer
  if (!rc) rc = enqueue_burgers_momentum(s, b);
  if (!rc) rc = enqueue_boundary(s);
  if (!rc) rc = enqueue_burgers_work(s, b);
  if (!rc) rc = enqueue_boundary(s);
  return rc;
}

int enqueue_stages(hydro_sim* s) {
  {
    const int rc = materialize_ghosts(s);  // the stencil reads u0's ghost cells
    if (rc) return rc;
  }
  s->frozen = false;  // U[cur ^ 1] becomes the new state
  const int S = s->tab.order;
  for (int k = 0; k < S; ++k) {
    PhaseScope ph(s, HYDRO_PHASE_STAGES);
    const int mode = s->plan[k].mode;
    if (s->mixed) {
      StageArgs<float, float> a{};
      fill_stage_args<float>(s, k, a);
      launch_stage(s, mode, a);
    } else {
      StageArgs<double, double> a{};
      fill_stage_args<double>(s, k, a);
      launch_stage(s, mode, a);
    }
    HIP_TRY(hipGetLastError());
  }
  // ghost cells of the new state are those of u0 (the RK combination leaves
  // them unchanged: their derivative is zero); copied when first read
  s->cur ^= 1;
  s->ghosts_pending = true;  // SRHD: copied inside the con2prim launch below
  // ForwardEuler's zero-gravity pass refreshes the ghosts after the update
  // (src/Solver/SelfGravitation.cpp:122 + src/Integrator/ForwardEuler.cpp:29-31)
  // (EulerBurgers applies the boundary next in any case; with a potential
  // the gravity pass below runs for real and keeps it itself)
  if ((s->d.flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS) && s->d.integrator == HYDRO_INTEGRATOR_FORWARDEULER &&
      s->d.equation == HYDRO_EQN_EULER && s->d.solver != HYDRO_SOLVER_BURGERS && !s->pot_on) {
    int rc = enqueue_boundary(s);
    if (rc) return rc;
  }
  // EulerRoe / EulerHLL(C)::step -> SelfGravitation::applyPotential with a
  // potential (src/Solver/EulerRoe.cpp:31-34, EulerHLL.cpp:25-28)
  if (s->pot_on && s->d.solver != HYDRO_SOLVER_BURGERS) {
    int rc = enqueue_gravity_pass(s);
    if (rc) return rc;
  }
  // SRHDRoe::step tail: constrainState + updatePrimitives (src/Solver/SRHDRoe.cpp:62-73)
  if (s->d.equation == HYDRO_EQN_SRHD) return enqueue_srhd_tail(s);
  if (s->d.solver == HYDRO_SOLVER_BURGERS) return enqueue_burgers_pressure_passes(s);
  return HYDRO_OK;
}

// A run of a small 1D Euler grid: every update of up to kRun1dChunk steps in
// one single-workgroup launch (run1d.hip) instead of S + 3 launches per
// update; the same device code on the same buffers, so the same results.
constexpr int kRun1dChunk = 256;

bool run1d_eligible(const hydro_sim* s) {
  if (!s->use_run1d || s->profiling || s->dim != 1 || s->d.equation != HYDRO_EQN_EULER) return false;
  if (s->pot_on || (s->solid_on && s->d.solver == HYDRO_SOLVER_ROE)) return false;  // HLL/HLLC ignore solids
  if (s->d.solver == HYDRO_SOLVER_BURGERS) return false;
  if (!boundary_covers_ghosts(s) || s->tab.order > kRun1dMaxStages) return false;
  const int strips = (s->sz[0] - 4 + 59) / 60;  // 60 output columns per strip
  return strips <= (s->mixed ? kRun1dMaxStripsF32 : kRun1dMaxStripsF64);
}

template <typename T>
int run1d(hydro_sim* s, int nsteps) {
  s->frozen = false;
  Run1dArgs<T, T> r{};
  static_assert(sizeof(r) <= 4096, "kernel arguments fit the kernarg segment");
  r.nstages = s->tab.order;
  r.b = boundary_args(s);
  r.march = s->d.use_fixed_dt ? 0 : 1;
  r.dt = march_args(s);
  r.f = finalize_args(s, true);
  // ForwardEuler's zero-gravity pass refreshes the ghosts after the update
  // (src/Solver/SelfGravitation.cpp:122 + src/Integrator/ForwardEuler.cpp:29-31)
  r.post_boundary =
      (s->d.flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS) && s->d.integrator == HYDRO_INTEGRATOR_FORWARDEULER ? 1 : 0;
  for (int i = 0; i < nsteps;) {
    const int c = std::min(nsteps - i, kRun1dChunk);
    for (int k = 0; k < r.nstages; ++k) fill_stage_args<T>(s, k, r.st[k]);
    r.U[0] = (T*)s->U[s->cur];
    r.U[1] = (T*)s->U[s->cur ^ 1];
    r.lo[0] = s->kahan ? s->Ulo[s->cur] : nullptr;
    r.lo[1] = s->kahan ? s->Ulo[s->cur ^ 1] : nullptr;
    r.nsteps = c;
    if constexpr (sizeof(T) == 8) launch_run1d_f64(kernel_kind(s->d), s->d.limiter, r, s->stream);
    else launch_run1d_f32(kernel_kind(s->d), s->d.limiter, s->kahan, r, s->stream);
    HIP_TRY(hipGetLastError());
    s->cur ^= c & 1;
    i += c;
  }
  // as after enqueue_update: the new state's ghost cells are u0's until
  // copied, unless the post-update boundary rewrote them
  s->ghosts_pending = !r.post_boundary;
  return HYDRO_OK;
}

// ---- generic stage evaluation helpers
bool known_buffer(const hydro_sim* s, const void* p) {
  if (p && p == s->U[s->cur]) return true;
  return p && std::find(s->user_bufs.begin(), s->user_bufs.end(), p) != s->user_bufs.end();
}

int generic_supported(const hydro_sim* s) {
  if (s->kahan)
    return fail(HYDRO_E_UNSUPPORTED, "generic stage evaluation needs plain storage (no HYDRO_FLAG_KAHAN)");
  if (s->d.solver == HYDRO_SOLVER_BURGERS)
    return fail(HYDRO_E_UNSUPPORTED,
                "EulerBurgers integrates three passes per step; its step runs fused (hydro_sim_step / update)");
  return HYDRO_OK;
}

// deriv += L(U[cur]) with the frozen start-of-step data of U[cur ^ 1] and dt
template <typename T>
int enqueue_deriv(hydro_sim* s, double dt, void* deriv) {
  // dt into the scratch scalars (the limiter's full-step dt), clock untouched
  FinalizeArgs f{};
  f.use_fixed = 1;
  f.fixed_dt = dt;
  f.dt = &s->sc[1].dt;
  f.t = &s->sc[1].t;
  f.steps = &s->sc[1].steps;
  f.health = &s->sc[1].health;
  launch_finalize_dt(f, s->stream);
  StageArgs<T, T> a{};
  fill_stage_args<T>(s, 0, a);
  a.u0 = (const T*)s->U[s->cur ^ 1];
  a.u0lo = nullptr;
  a.dk = (const T*)s->U[s->cur];
  a.dnext = (T*)deriv;
  a.unew = nullptr;
  a.unewlo = nullptr;
  a.pkp = nullptr;
  a.nf = 0;
  a.nrep = 0;
  for (int j = 0; j < kMaxPartials; ++j) {
    a.fp[j] = nullptr;
    a.rep[j] = nullptr;
  }
  a.final_stage = 0;
  a.a_self = 0.0;
  a.b_self = 1.0;
  a.dt = &s->sc[1].dt;
  launch_stage(s, CM_DERIV, a);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

// The head of an update: boundary and CFL march as one launch, then the dt
// tail; for the strip-march CFL paths (1D/2D Euler and SRHD, 3D Euler)
bool head_eligible(const hydro_sim* s) {
  const int path = cfl_path(s->cg);
  return s->use_head && (path == 1 || path == 2);
}

int enqueue_head(hydro_sim* s) {
  if (boundary_covers_ghosts(s)) {
    s->ghosts_pending = false;
  } else {
    const int rc = materialize_ghosts(s);
    if (rc) return rc;
  }
  HeadArgs h{};
  h.b = boundary_args(s);
  h.a = cfl_path(s->cg) == 1 ? cfl_march_args(s->cg, s->blockmin) : march3_args(s->cg, s->blockmin);
  h.mblocks = (((h.a.nwaves + 3) / 4) + 7) & ~7;  // a multiple of 8 for the XCD-aware order
  long long ghosts = 4;
  if (s->dim == 2) ghosts = 4LL * s->sz[0] + 4LL * (s->sz[1] - 4);
  if (s->dim == 3)
    ghosts = 4LL * s->sz[0] * s->sz[1] + 4LL * s->sz[0] * (s->sz[2] - 4) + 4LL * (s->sz[1] - 4) * (s->sz[2] - 4);
  h.bblocks = (int)((ghosts + 255) / 256);
  {
    PhaseScope ph(s, HYDRO_PHASE_HEAD);
    launch_step_head(h, s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr, s->prim, s->mixed, s->stream);
  }
  FinalizeArgs f = finalize_args(s, true);
  f.nblocks = h.mblocks;
  {
    PhaseScope ph(s, HYDRO_PHASE_DT_TAIL);
    launch_finalize_dt(f, s->stream);
  }
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

int enqueue_update(hydro_sim* s) {
  int rc;
  if (head_eligible(s)) {
    rc = enqueue_head(s);
  } else {
    rc = enqueue_boundary(s);
    if (!rc) rc = enqueue_dt(s, true);
  }
  if (rc) return rc;
  return enqueue_stages(s);
}

}  // namespace

extern "C" {

void hydro_sim_desc_default(hydro_sim_desc* d) {
  std::memset(d, 0, sizeof *d);
  d->dim = 1;
  d->size[0] = 256;
  d->size[1] = d->size[2] = 1;
  for (int a = 0; a < 3; ++a) {
    d->xmin[a] = -.5;
    d->xmax[a] = .5;
    d->boundary[a][0] = d->boundary[a][1] = HYDRO_BOUNDARY_FREEFLOW;
  }
  d->equation = HYDRO_EQN_EULER;
  d->solver = HYDRO_SOLVER_ROE;
  d->limiter = HYDRO_LIMITER_SUPERBEE;
  d->integrator = HYDRO_INTEGRATOR_FORWARDEULER;
  d->precision = HYDRO_PRECISION_FP64;
  d->cfl = 0.5;
  d->fixed_dt = .001;
  d->gamma = 1.4;
  d->srhd_solve_prim_max_iter = 1000;
  d->srhd_solve_prim_stop_epsilon = 1e-7;
  d->srhd_solve_prim_vel_epsilon = 1e-15;
  d->srhd_solve_prim_pmin_epsilon = 1e-16;
  d->srhd_rho_min = 1e-15;
  d->srhd_rho_max = 1e20;
  d->srhd_eint_max = 1e20;
  d->srhd_D_min = 1e-15;
  d->srhd_D_max = 1e20;
  d->srhd_tau_min = 1e-15;
  d->srhd_tau_max = 1e20;
  d->maxwell_permittivity = d->maxwell_permeability = d->maxwell_conductivity = 1.0;
  d->flags = HYDRO_FLAG_REFERENCE;
  d->use_gravity = 0;
  d->gauss_seidel_max_iter = 20;
  d->self_grav_gravitational_constant = 1.0;
}

const char* hydro_last_error(void) { return g_err.c_str(); }
const char* hydro_version(void) { return "hydro-mi355x 0.1 (gfx950)"; }

int hydro_ctx_create(int device, hydro_ctx** out) {
  if (!out) return fail(HYDRO_E_INVALID, "null out");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(HYDRO_E_INVALID, "no such device");
  HIP_TRY(hipSetDevice(device));
  hydro_ctx* c = new hydro_ctx;
  c->device = device;
  *out = c;
  return HYDRO_OK;
}

void hydro_ctx_destroy(hydro_ctx* ctx) { delete ctx; }

void hydro_sim_destroy(hydro_sim* s) {
  if (!s) return;
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto& g : s->graph2)
    if (g) hipGraphExecDestroy(g);
  for (auto& g : s->graphN)
    if (g) hipGraphExecDestroy(g);
  for (auto& p : s->ev) {
    hipEventDestroy(p.e0);
    hipEventDestroy(p.e1);
  }
  for (int i = 0; i < 2; ++i) {
    if (s->U[i]) hipFree(s->U[i]);
    if (s->Ulo[i]) hipFree(s->Ulo[i]);
  }
  for (int i = 0; i < 3; ++i)
    if (s->D[i]) hipFree(s->D[i]);
  for (int j = 0; j < kMaxPartials; ++j)
    if (s->P[j]) hipFree(s->P[j]);
  for (void* p : s->user_bufs) hipFree(p);
  if (s->aos_dev) hipFree(s->aos_dev);
  if (s->prim) hipFree(s->prim);
  for (void* p : s->pres)
    if (p) hipFree(p);
  for (void* p : s->pot)
    if (p) hipFree(p);
  if (s->solid) hipFree(s->solid);
  if (s->blockmin) hipFree(s->blockmin);
  if (s->partial) hipFree(s->partial);
  if (s->sc) hipFree(s->sc);
  if (s->host_sc) hipHostFree(s->host_sc);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

// ------------------------------------------------------------ autotuning
// Work decompositions of the fused stage kernels: rows per wave (2D), rows
// or planes per chunk (3D), waves per workgroup (2D), the first-order 3D
// form.  The winners of hydro_sim_autotune are kept per grid shape for the
// process and, with HYDRO_TUNE_FILE set, in that file (one line per shape:
// key, then rows wpb form zwaves ms), which later processes read at create.
namespace {

struct TuneParams {
  int rows, wpb, form, zwaves;
  double ms;
};
std::mutex g_tune_mu;
std::map<std::string, TuneParams> g_tune;
bool g_tune_loaded = false;

std::string tune_key(const hydro_sim* s) {
  std::ostringstream k;
  k << "d" << s->dim << "_" << s->sz[0] << "x" << s->sz[1] << "x" << s->sz[2] << "_e" << s->d.equation << "_s"
    << s->d.solver << "_l" << s->d.limiter << "_i" << s->d.integrator << "_p" << s->d.precision << "_k"
    << (s->kahan ? 1 : 0);
  return k.str();
}

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = getenv("HYDRO_TUNE_FILE");
  if (!path) return;
  std::ifstream f(path);
  std::string key;
  TuneParams p{};
  while (f >> key >> p.rows >> p.wpb >> p.form >> p.zwaves >> p.ms) g_tune[key] = p;
}

void tune_store(const std::string& key, const TuneParams& p) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_load_locked();
  g_tune[key] = p;
  if (const char* path = getenv("HYDRO_TUNE_FILE")) {
    std::ofstream f(path, std::ios::app);
    f << key << " " << p.rows << " " << p.wpb << " " << p.form << " " << p.zwaves << " " << p.ms << "\n";
  }
}

bool tune_lookup(const std::string& key, TuneParams& p) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  tune_load_locked();
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  p = it->second;
  return true;
}

bool first_order_3d(const hydro_sim* s) {
  return s->dim == 3 && (kernel_kind(s->d) == 1 || kernel_kind(s->d) == 2);
}
bool roe_family_3d(const hydro_sim* s) { return s->dim == 3 && !first_order_3d(s); }

// the 4th decomposition parameter: waves per workgroup of the 3D kernel
TuneParams current_params(const hydro_sim* s) {
  return TuneParams{s->rows, s->wpb, s->form3d, roe_family_3d(s) ? s->rwaves : s->zwaves, 0.0};
}

// the planes the CFL pass reads: the potential (Roe, HLL, EulerBurgers with
// one; the solid Roe form always) and the solid mask (Roe, EulerBurgers)
void set_cfl_planes(hydro_sim* s) {
  const int sv = s->d.solver;
  const bool roe = sv == HYDRO_SOLVER_ROE && s->d.equation == HYDRO_EQN_EULER;
  const bool solid = s->solid_on && (roe || sv == HYDRO_SOLVER_BURGERS);
  s->cg.solid = solid ? s->solid : nullptr;
  const bool pot = (s->pot_on && (sv == HYDRO_SOLVER_ROE || sv == HYDRO_SOLVER_HLL || sv == HYDRO_SOLVER_BURGERS)) ||
                   (solid && roe);
  s->cg.pot = pot ? s->pot[0] : nullptr;
}

void drop_graphs(hydro_sim* s) {
  for (auto& g : s->graph2)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  for (auto& g : s->graphN)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void apply_params(hydro_sim* s, const TuneParams& p) {
  if (first_order_3d(s)) {
    const int form = (p.form == kForm3dZm2 && !s->mixed) ? kForm3dZm1w8 : p.form;
    set_form3d(s, form, form == kForm3dYMarch ? p.zwaves : 0);
  }
  if (roe_family_3d(s)) s->rwaves = (s->mixed && p.zwaves == 16) ? 16 : 8;
  s->rows = std::max(1, p.rows);
  s->wpb = std::min(4, std::max(1, p.wpb));
  drop_graphs(s);
}

}  // namespace

int hydro_sim_create(hydro_ctx* ctx, const hydro_sim_desc* d, hydro_sim** out) {
  if (!ctx || !out) return fail(HYDRO_E_INVALID, "null argument");
  int rc = validate(d);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  hydro_sim* s = new hydro_sim;
  s->ctx = ctx;
  s->d = *d;
  s->dim = d->dim;
  s->N = d->equation == HYDRO_EQN_MAXWELL ? 6 : d->dim + 2;
  for (int a = 0; a < 3; ++a) s->sz[a] = a < d->dim ? d->size[a] : 1;
  s->vol = (long long)s->sz[0] * s->sz[1] * s->sz[2];
  s->plane = (s->vol + 63) & ~63LL;
  s->mixed = d->precision == HYDRO_PRECISION_FP32_MIXED;
  s->kahan = s->mixed && (d->flags & HYDRO_FLAG_KAHAN);
  {
    const CflGrid g = cfl_grid(*d, s->plane);
    for (int a = 0; a < 3; ++a) s->kdx[a] = g.kdx[a];
    s->cfl = g.cfl;
  }
  const char* plan_env = getenv("HYDRO_RK_PLAN");
  const bool allow_replay = !(plan_env && std::string(plan_env) == "partial");
  if (get_tableau(d->integrator, s->tab) && allow_replay && plan_replay(s->tab, s->plan)) {
    s->replay = true;
    s->nslots = 0;
    s->ndbuf = s->tab.order - 1;
  } else if (get_tableau(d->integrator, s->tab) && plan_stages(s->tab, s->plan, s->nslots)) {
    s->ndbuf = s->tab.order > 1 ? 2 : 0;
  } else {
    delete s;
    return fail(HYDRO_E_INVALID, "integrator not supported");
  }
  // HYDRO_COMBINE=generic runs every stage through the generic combination
  const char* comb_env = getenv("HYDRO_COMBINE");
  const bool force_generic = comb_env && std::string(comb_env) == "generic";
  for (auto& p : s->plan) p.mode = force_generic ? CM_GENERIC : combine_mode(p);
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu <= 0)
    ncu = 256;
  s->ncu = ncu;
  if (s->dim == 2) {
    int min_waves = 3;
    if (const char* e = getenv("HYDRO_MIN_WAVES_PER_SIMD")) min_waves = std::max(1, atoi(e));
    s->rows = pick_rows_2d((s->sz[0] - 4 + 59) / 60, s->sz[1] - 4, 4 * ncu, min_waves);
  } else if (s->dim == 3 && (kernel_kind(*d) == 1 || kernel_kind(*d) == 2)) {
    // rows shared through LDS: measured faster for the FP64 first stage only
    // (256^3 HLLC blast: 794 vs 891 us; FP32 and the later FP64 stages hold
    // more registers than they save loads)
    s->zshare = s->mixed ? 0 : (1 << CM_FIRST) | (1 << CM_SINGLE);
    if (const char* e = getenv("HYDRO_3D_ZSHARE")) s->zshare = (int)strtol(e, nullptr, 0);
    // the z-march form: FP32 two cells per lane in 8-wave workgroups (zm2)
    // or one cell per lane in 16-wave workgroups (zm1); FP64 one cell per
    // lane in 8-wave workgroups (LDS holds 94 KB per workgroup in each).
    // Measured on the 256^3 HLLC blast (stage average, FP32 compensated /
    // FP64): y-march 523 / 963 us, zm1 520 / 953, zm2 400 / -, zm1w8 540 /
    // 955.  FP32 takes zm2 (two strips of 126 columns cover the 252
    // interior columns, against five strips of 60); FP64 keeps the y-march.
    int form = s->mixed ? kForm3dZm2 : kForm3dYMarch;
    if (const char* e = getenv("HYDRO_3D_FORM")) {
      const std::string f = e;
      form = f == "zm1" ? kForm3dZm1 : f == "zm2" ? kForm3dZm2 : f == "zm1w8" ? kForm3dZm1w8 : kForm3dYMarch;
    }
    set_form3d(s, form);
    if (form == kForm3dYMarch) {
      if (const char* e = getenv("HYDRO_3D_ZWAVES")) {
        s->zwaves = atoi(e) == 4 ? 4 : 8;
        set_form3d(s, form, s->zwaves);
      }
    } else if (const char* e = getenv("HYDRO_3D_PLANES")) {
      s->rows = std::max(1, atoi(e));
    }
  } else if (s->dim == 3) {
    s->rwaves = stage3d_roe_default_waves(s->mixed);
    if (const char* e = getenv("HYDRO_3D_ROE_WAVES")) s->rwaves = (s->mixed && atoi(e) == 16) ? 16 : 8;
    s->rows = pick_planes_3d_roe(s, s->rwaves);
  }
  if (const char* e = getenv("HYDRO_ROWS_PER_WAVE")) s->rows = std::max(1, atoi(e));
  if (const char* e = getenv("HYDRO_GRAPHS")) s->use_graphs = atoi(e) != 0;
  if (const char* e = getenv("HYDRO_RUN1D")) s->use_run1d = atoi(e) != 0;
  if (const char* e = getenv("HYDRO_STEP_HEAD")) s->use_head = atoi(e) != 0;
  if (const char* e = getenv("HYDRO_WAVES_PER_BLOCK")) s->wpb = std::min(4, std::max(1, atoi(e)));
  if (const char* e = getenv("HYDRO_DT_ROWS")) s->dt_rows = atoi(e);
  if (const char* e = getenv("HYDRO_GRAPH_UPDATES")) s->graph_updates = std::max(2, atoi(e) & ~1);
  s->cg = cfl_grid(s->d, s->plane);
  s->cg.dt_rows = s->dt_rows;
  s->cg.dt_blocks = s->dt_blocks;
  {
    // a decomposition tuned earlier for this grid shape (this process or
    // HYDRO_TUNE_FILE)
    // (decomposition switches in the environment take precedence)
    bool env = false;
    for (const char* k : {"HYDRO_ROWS_PER_WAVE", "HYDRO_WAVES_PER_BLOCK", "HYDRO_3D_FORM", "HYDRO_3D_PLANES",
                          "HYDRO_3D_ZWAVES", "HYDRO_MIN_WAVES_PER_SIMD"})
      env = env || getenv(k) != nullptr;
    TuneParams p;
    if (!env && s->dim >= 2 && tune_lookup(tune_key(s), p)) apply_params(s, p);
  }

  auto cleanup = [&](int code, const std::string& m) {
    hydro_sim_destroy(s);
    return fail(code, m);
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "stream create failed");
  const size_t es = esize(s);
  const size_t bytes = es * s->plane * s->N;
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&s->U[i], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (state)");
    if (hipMemset(s->U[i], 0, bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
    if (s->kahan) {
      if (hipMalloc((void**)&s->Ulo[i], sizeof(float) * s->plane * s->N) != hipSuccess)
        return cleanup(HYDRO_E_HIP, "out of device memory (state low words)");
      hipMemset(s->Ulo[i], 0, sizeof(float) * s->plane * s->N);
    }
  }
  // increment buffers: their ghost cells stay 0 (stages write interior cells only)
  for (int i = 0; i < s->ndbuf; ++i) {
    if (hipMalloc(&s->D[i], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (stages)");
    if (hipMemset(s->D[i], 0, bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
  }
  for (int j = 0; j < s->nslots; ++j) {
    if (hipMalloc(&s->P[j], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (partials)");
    hipMemset(s->P[j], 0, bytes);
  }
  if (d->equation == HYDRO_EQN_SRHD) {
    if (hipMalloc((void**)&s->prim, prim_bytes(s)) != hipSuccess)
      return cleanup(HYDRO_E_HIP, "out of device memory (primitives)");
    hipMemset(s->prim, 0, prim_bytes(s));
  }
  if (d->equation == HYDRO_EQN_EULER) {
    // src/Solver/SelfGravitation.cpp:14-18: the potential buffer (two planes
    // here: the relaxation sweeps alternate between them)
    for (auto& p : s->pot) {
      if (hipMalloc(&p, es * s->plane) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (potential)");
      hipMemset(p, 0, es * s->plane);
    }
  }
  if (d->solver == HYDRO_SOLVER_BURGERS) {
    // src/Solver/EulerBurgers.cpp:14-22: the pressure buffer (two planes
    // here: the momentum stages alternate between them)
    for (auto& p : s->pres) {
      if (hipMalloc(&p, es * s->plane) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (pressure)");
      hipMemset(p, 0, es * s->plane);
    }
    s->work_coef = constant_derivative_sum(s->tab);
  }
  if (hipMalloc((void**)&s->blockmin, sizeof(double) * std::max(s->dt_blocks, kDtMaxWaves)) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "alloc");

  if (hipMalloc((void**)&s->partial, sizeof(double) * s->tot_blocks * 8) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "alloc");
  if (hipMalloc((void**)&s->sc, sizeof(DevScalars) * 2) != hipSuccess) return cleanup(HYDRO_E_HIP, "alloc");
  hipMemset(s->sc, 0, sizeof(DevScalars) * 2);
  if (hipHostMalloc((void**)&s->host_sc, sizeof(DevScalars), hipHostMallocDefault) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "pinned alloc");
  if (hipDeviceSynchronize() != hipSuccess) return cleanup(HYDRO_E_HIP, "sync");
  *out = s;
  return HYDRO_OK;
}

int hydro_sim_num_states(const hydro_sim* s) { return s ? s->N : 0; }
size_t hydro_sim_volume(const hydro_sim* s) { return s ? (size_t)s->vol : 0; }
void* hydro_sim_stream(hydro_sim* s) { return s ? (void*)s->stream : nullptr; }

// AoS (reference layout) -> the device planes of the current state buffer;
// FP32 storage keeps the residual as the low word when compensated, or
// takes the low words given (`low`, compensated storage only)
static int ensure_aos_staging(hydro_sim* s) {
  if (!s->aos_dev) HIP_TRY(hipMalloc((void**)&s->aos_dev, sizeof(double) * (size_t)s->N * (size_t)s->vol));
  return HYDRO_OK;
}

static int upload_state(hydro_sim* s, const double* aos, const float* low = nullptr) {
  s->frozen = false;
  const int N = s->N;
  const long long V = s->vol;
  if (!low) {
    // the array of structures goes over as it is; the planes are built on the GPU
    int rc = ensure_aos_staging(s);
    if (rc) return rc;
    // in stream order: the layout kernel reads the staging after the copy
    HIP_TRY(hipMemcpyAsync(s->aos_dev, aos, sizeof(double) * (size_t)N * (size_t)V, hipMemcpyHostToDevice,
                           s->stream));
    if (s->mixed)
      launch_aos_to_planes_f32(s->aos_dev, N, V, s->plane, (float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr,
                               s->stream);
    else
      launch_aos_to_planes_f64(s->aos_dev, N, V, s->plane, (double*)s->U[s->cur], s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->ghosts_pending = false;
    return HYDRO_OK;
  }
  if (s->mixed) {
    std::vector<float> hi((size_t)s->plane * N, 0.f), lo((size_t)s->plane * N, 0.f);
    for (long long c = 0; c < V; ++c)
      for (int k = 0; k < N; ++k) {
        const double v = aos[k + (size_t)N * c];
        const float h = (float)v;
        hi[(size_t)k * s->plane + c] = h;
        lo[(size_t)k * s->plane + c] = low ? low[k + (size_t)N * c] : (float)(v - (double)h);
      }
    HIP_TRY(copy_sync(s->U[s->cur], hi.data(), hi.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    if (s->kahan) HIP_TRY(copy_sync(s->Ulo[s->cur], lo.data(), lo.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
  } else {
    std::vector<double> buf((size_t)s->plane * N, 0.0);
    for (long long c = 0; c < V; ++c)
      for (int k = 0; k < N; ++k) buf[(size_t)k * s->plane + c] = aos[k + (size_t)N * c];
    HIP_TRY(copy_sync(s->U[s->cur], buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  }
  s->ghosts_pending = false;
  return HYDRO_OK;
}

int hydro_sim_set_state(hydro_sim* s, const double* aos, size_t n) {
  if (!s || !aos) return fail(HYDRO_E_INVALID, "null argument");
  if (n != (size_t)s->N * (size_t)s->vol) return fail(HYDRO_E_INVALID, "state size must be N * volume");
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int N = s->N;
  const long long V = s->vol;
  std::vector<double> conv;
  if (s->d.equation == HYDRO_EQN_SRHD) {
    // initVariables (res/SRHDCommon.cl:24-92): Newtonian (rho, rho v, E) ->
    // conserved (D, S, tau) and primitives (rho, v, eInt), every cell, FP64
    conv.resize((size_t)N * V);
    std::vector<double> pr((size_t)s->plane * N, 0.0);
    const double gamma = s->d.gamma;
    const int D = s->dim, E = D + 1;
    for (long long c = 0; c < V; ++c) {
      const double* st = aos + (size_t)N * c;
      double* out = conv.data() + (size_t)N * c;
      const double rho = st[0];
      double v[3] = {0, 0, 0};
      for (int a = 0; a < D; ++a) v[a] = st[1 + a] / st[0];
      const double vSq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      const double eInt = st[E] / rho - .5 * vSq;
      const double P = (gamma - 1.) * rho * eInt;
      const double h = 1. + eInt + P / rho;
      const double WSq = 1. / (1. - vSq);
      const double W = std::sqrt(WSq);
      out[0] = rho * W;
      for (int a = 0; a < D; ++a) out[1 + a] = rho * h * WSq * v[a];
      out[E] = rho * h * WSq - P - rho * W;
      pr[c] = rho;
      for (int a = 0; a < D; ++a) pr[(size_t)(1 + a) * s->plane + c] = v[a];
      pr[(size_t)E * s->plane + c] = eInt;
    }
    {
      const int rc = upload_prims(s, pr);
      if (rc) return rc;
    }
    aos = conv.data();
  }
  const int rc = upload_state(s, aos);
  if (rc) return rc;
  // a new initial state restarts the clock
  DevScalars z{};
  HIP_TRY(copy_sync(s->sc, &z, sizeof z, hipMemcpyHostToDevice, s->stream));
  return HYDRO_OK;
}

int hydro_sim_restore(hydro_sim* s, const double* aos, const float* lo, const double* prims, size_t n, double t,
                      long long steps) {
  if (!s || !aos) return fail(HYDRO_E_INVALID, "null argument");
  if (n != (size_t)s->N * (size_t)s->vol) return fail(HYDRO_E_INVALID, "state size must be N * volume");
  if (lo && !s->kahan) return fail(HYDRO_E_UNSUPPORTED, "low words exist for compensated FP32 storage only");
  if (s->prim && !prims) return fail(HYDRO_E_INVALID, "an SRHD state needs its primitives");
  if (!s->prim && prims) return fail(HYDRO_E_UNSUPPORTED, "primitives exist for SRHD only");
  if (!(t >= 0.0) || steps < 0) return fail(HYDRO_E_INVALID, "time and step count must be >= 0");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (prims) {
    std::vector<double> pr((size_t)s->plane * s->N, 0.0);
    for (long long c = 0; c < s->vol; ++c)
      for (int k = 0; k < s->N; ++k) pr[(size_t)k * s->plane + c] = prims[k + (size_t)s->N * c];
    const int rc = upload_prims(s, pr);
    if (rc) return rc;
  }
  const int rc = upload_state(s, aos, lo);
  if (rc) return rc;
  DevScalars z{};
  z.t = t;
  z.steps = steps;
  HIP_TRY(copy_sync(s->sc, &z, sizeof z, hipMemcpyHostToDevice, s->stream));
  return HYDRO_OK;
}

int hydro_sim_get_state(hydro_sim* s, double* aos, size_t n) {
  if (!s || !aos) return fail(HYDRO_E_INVALID, "null argument");
  if (n != (size_t)s->N * (size_t)s->vol) return fail(HYDRO_E_INVALID, "state size must be N * volume");
  {
    const int rc = materialize_ghosts(s);
    if (rc) return rc;
  }
  const int N = s->N;
  const long long V = s->vol;
  // the planes become the array of structures on the GPU (FP64: hi + lo for
  // compensated storage), then one copy back
  int rc = ensure_aos_staging(s);
  if (rc) return rc;
  if (s->mixed)
    launch_planes_to_aos_f32((const float*)s->U[s->cur], s->kahan ? s->Ulo[s->cur] : nullptr, N, V, s->plane,
                             s->aos_dev, s->stream);
  else
    launch_planes_to_aos_f64((const double*)s->U[s->cur], N, V, s->plane, s->aos_dev, s->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(aos, s->aos_dev, sizeof(double) * (size_t)N * (size_t)V, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return HYDRO_OK;
}

int hydro_sim_get_storage(hydro_sim* s, double* hi, float* lo, size_t n) {
  if (!s || !hi) return fail(HYDRO_E_INVALID, "null argument");
  if (n != (size_t)s->N * (size_t)s->vol) return fail(HYDRO_E_INVALID, "state size must be N * volume");
  if (!s->kahan) {
    const int rc = hydro_sim_get_state(s, hi, n);
    if (rc) return rc;
    if (lo) std::fill(lo, lo + n, 0.f);
    return HYDRO_OK;
  }
  {
    const int rc = materialize_ghosts(s);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int N = s->N;
  std::vector<float> h((size_t)s->plane * N), l((size_t)s->plane * N);
  HIP_TRY(copy_sync(h.data(), s->U[s->cur], h.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(copy_sync(l.data(), s->Ulo[s->cur], l.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  for (long long c = 0; c < s->vol; ++c)
    for (int k = 0; k < N; ++k) {
      hi[k + (size_t)N * c] = h[(size_t)k * s->plane + c];
      if (lo) lo[k + (size_t)N * c] = l[(size_t)k * s->plane + c];
    }
  return HYDRO_OK;
}

// SelfGravitation::resetState's solidBuffer upload
// (src/Solver/SelfGravitation.cpp:56-75): one flag per cell, kept as a mask
// plane with the neighbours' flags folded in.
int hydro_sim_set_solid(hydro_sim* s, const unsigned char* solid, size_t n) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (s->d.equation != HYDRO_EQN_EULER) return fail(HYDRO_E_UNSUPPORTED, "solid cells belong to the Euler solvers");
  if (solid && n != (size_t)s->vol) return fail(HYDRO_E_INVALID, "solid mask size must be the volume");
  bool any = false;
  if (solid)
    for (size_t i = 0; i < n && !any; ++i) any = solid[i] != 0;
  if (any && !s->solid) {
    if (hipMalloc(&s->solid, esize(s) * (size_t)s->plane) != hipSuccess)
      return fail(HYDRO_E_HIP, "out of device memory (solid mask)");
  }
  if (any) {
    const long long step[3] = {1, s->sz[0], (long long)s->sz[0] * s->sz[1]};
    std::vector<double> md((size_t)s->plane, 0.0);
    for (int z = 0; z < s->sz[2]; ++z)
      for (int y = 0; y < s->sz[1]; ++y)
        for (int x = 0; x < s->sz[0]; ++x) {
          const int ci[3] = {x, y, z};
          const long long c = x + step[1] * y + step[2] * z;
          int m = solid[c] ? 1 : 0;
          for (int a = 0; a < s->dim; ++a) {
            if (ci[a] > 0 && solid[c - step[a]]) m |= 1 << (1 + 2 * a);
            if (ci[a] < s->sz[a] - 1 && solid[c + step[a]]) m |= 1 << (2 + 2 * a);
          }
          md[c] = m;
        }
    const size_t bytes = esize(s) * (size_t)s->plane;
    hipError_t e;
    if (s->mixed) {
      std::vector<float> mf(md.begin(), md.end());
      e = copy_sync(s->solid, mf.data(), bytes, hipMemcpyHostToDevice, s->stream);
    } else {
      e = copy_sync(s->solid, md.data(), bytes, hipMemcpyHostToDevice, s->stream);
    }
    if (e != hipSuccess) return fail(HYDRO_E_HIP, "copy (solid mask)");
  }
  s->solid_on = any;
  set_cfl_planes(s);
  drop_graphs(s);  // captured updates with the other kernel forms
  return HYDRO_OK;
}

int hydro_sim_set_potential(hydro_sim* s, const double* pot, size_t n) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (s->d.equation != HYDRO_EQN_EULER) return fail(HYDRO_E_UNSUPPORTED, "a potential belongs to the Euler solvers");
  if (pot && n != (size_t)s->vol) return fail(HYDRO_E_INVALID, "potential size must be the volume");
  bool nonzero = s->d.use_gravity != 0;
  if (pot && !s->d.use_gravity)
    for (size_t i = 0; i < n && !nonzero; ++i) nonzero = pot[i] != 0.0;
  if (nonzero && s->dim == 3 && s->d.solver == HYDRO_SOLVER_HLL)
    return fail(HYDRO_E_UNSUPPORTED, "3D HLL with a gravitational potential is not supported (2D and 1D are; HLLC "
                                     "and Roe in 3D are)");
  int rc = materialize_ghosts(s);  // every cell's energy gains the potential
  if (rc) return rc;
  PhaseScope ph(s, HYDRO_PHASE_GRAVITY);
  // phi0: the guess -rho (src/Solver/SelfGravitation.cpp:48-52) or the given potential
  const size_t bytes = esize(s) * (size_t)s->plane;
  if (s->d.use_gravity) {
    if (s->mixed) launch_copy_plane_f32(s->vol, (const float*)s->U[s->cur], (float*)s->pot[0], true, s->stream);
    else launch_copy_plane_f64(s->vol, (const double*)s->U[s->cur], (double*)s->pot[0], true, s->stream);
    HIP_TRY(hipGetLastError());
  } else {
    std::vector<double> pd((size_t)s->plane, 0.0);
    if (pot) std::copy(pot, pot + n, pd.begin());
    if (s->mixed) {
      std::vector<float> pf(pd.begin(), pd.end());
      if (copy_sync(s->pot[0], pf.data(), bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess)
        return fail(HYDRO_E_HIP, "copy (potential)");
    } else if (copy_sync(s->pot[0], pd.data(), bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
      return fail(HYDRO_E_HIP, "copy (potential)");
    }
  }
  // :89-94: the initial potential into every cell's total energy
  const long long E = (long long)(s->dim + 1) * s->plane;
  if (s->mixed)
    launch_add_potential_energy_f32(s->vol, (float*)s->U[s->cur] + E, s->kahan ? s->Ulo[s->cur] + E : nullptr,
                                    (const float*)s->pot[0], s->stream);
  else
    launch_add_potential_energy_f64(s->vol, (double*)s->U[s->cur] + E, (const double*)s->pot[0], s->stream);
  HIP_TRY(hipGetLastError());
  // :81-88: 100 x gaussSeidelMaxIter sweeps towards the steady potential
  if (s->d.use_gravity) {
    for (int tries = 0; tries < 100; ++tries) {
      rc = enqueue_relax(s, s->d.gauss_seidel_max_iter);
      if (rc) return rc;
    }
  }
  s->pot_on = nonzero;
  set_cfl_planes(s);
  drop_graphs(s);  // captured updates without (or with) the gravity pass
  HIP_TRY(hipStreamSynchronize(s->stream));
  return HYDRO_OK;
}

int hydro_sim_get_potential(hydro_sim* s, double* pot, size_t n) {
  if (!s || !pot) return fail(HYDRO_E_INVALID, "null argument");
  if (s->d.equation != HYDRO_EQN_EULER) return fail(HYDRO_E_UNSUPPORTED, "a potential belongs to the Euler solvers");
  if (n != (size_t)s->vol) return fail(HYDRO_E_INVALID, "potential size must be the volume");
  if (s->mixed) {
    std::vector<float> f((size_t)s->vol);
    if (copy_sync(f.data(), s->pot[0], sizeof(float) * f.size(), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
      return fail(HYDRO_E_HIP, "copy (potential)");
    std::copy(f.begin(), f.end(), pot);
  } else if (copy_sync(pot, s->pot[0], sizeof(double) * n, hipMemcpyDeviceToHost, s->stream) != hipSuccess) {
    return fail(HYDRO_E_HIP, "copy (potential)");
  }
  return HYDRO_OK;
}

// Converter::fromGPU's solidBuffer read-back (include/HydroGPU/Solver/
// SelfGravitationBehavior.h:83-88): bit 0 of the mask plane per cell
int hydro_sim_get_solid(hydro_sim* s, unsigned char* solid, size_t n) {
  if (!s || !solid) return fail(HYDRO_E_INVALID, "null argument");
  if (s->d.equation != HYDRO_EQN_EULER) return fail(HYDRO_E_UNSUPPORTED, "solid cells belong to the Euler solvers");
  if (n != (size_t)s->vol) return fail(HYDRO_E_INVALID, "solid mask size must be the volume");
  if (!s->solid_on) {
    std::fill(solid, solid + n, (unsigned char)0);
    return HYDRO_OK;
  }
  std::vector<double> m(n);
  if (s->mixed) {
    std::vector<float> f(n);
    if (copy_sync(f.data(), s->solid, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream) != hipSuccess)
      return fail(HYDRO_E_HIP, "copy (solid mask)");
    std::copy(f.begin(), f.end(), m.begin());
  } else if (copy_sync(m.data(), s->solid, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream) != hipSuccess) {
    return fail(HYDRO_E_HIP, "copy (solid mask)");
  }
  for (size_t i = 0; i < n; ++i) solid[i] = (unsigned char)(((int)m[i]) & 1);
  return HYDRO_OK;
}

int hydro_sim_get_primitives(hydro_sim* s, double* aos, size_t n) {
  if (!s || !aos) return fail(HYDRO_E_INVALID, "null argument");
  if (!s->prim) return fail(HYDRO_E_UNSUPPORTED, "primitives exist for SRHD only");
  if (n != (size_t)s->N * (size_t)s->vol) return fail(HYDRO_E_INVALID, "size must be N * volume");
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::vector<double> buf;
  {
    const int rc = download_prims(s, buf);
    if (rc) return rc;
  }
  for (long long c = 0; c < s->vol; ++c)
    for (int k = 0; k < s->N; ++k) aos[k + (size_t)s->N * c] = buf[(size_t)k * s->plane + c];
  return HYDRO_OK;
}

int hydro_sim_boundary(hydro_sim* s) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  return enqueue_boundary(s);
}

int hydro_sim_calc_timestep(hydro_sim* s, double* dt_out) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  int rc = materialize_ghosts(s);  // the CFL reads the ghost neighbours of edge cells
  if (rc) return rc;
  // the scratch clock starts at the real one (clipping to t_end reads it)
  HIP_TRY(hipMemcpyAsync(&s->sc[1], &s->sc[0], sizeof(DevScalars), hipMemcpyDeviceToDevice, s->stream));
  rc = enqueue_dt(s, false);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(s->host_sc, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (dt_out) *dt_out = s->host_sc->dt;
  return HYDRO_OK;
}

int hydro_sim_step(hydro_sim* s, double dt) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  // dt onto the device and the clock advanced by it (the tail of
  // findMinTimestep with a given dt), then the stages
  FinalizeArgs f{};
  f.nblocks = 0;
  f.use_fixed = 1;
  f.fixed_dt = dt;
  f.dt = &s->sc->dt;
  f.t = &s->sc->t;
  f.steps = &s->sc->steps;
  f.health = &s->sc->health;
  {
    PhaseScope ph(s, HYDRO_PHASE_DT_TAIL);
    launch_finalize_dt(f, s->stream);
  }
  HIP_TRY(hipGetLastError());
  return enqueue_stages(s);
}

int hydro_sim_update(hydro_sim* s, double* dt_out) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  int rc = enqueue_update(s);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(s->host_sc, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (dt_out) *dt_out = s->host_sc->dt;
  return HYDRO_OK;
}

// Records n updates (n even) into an instantiated graph.  The capture
// records the launches without running them and leaves the host-side
// bookkeeping where n updates leave it.
static int capture_updates(hydro_sim* s, int n, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = HYDRO_OK;
  for (int k = 0; k < n && !rc; ++k) rc = enqueue_update(s);
  const hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(HYDRO_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(HYDRO_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return HYDRO_OK;
}

// Captures the n-update graph for both buffer parities when the current
// one is missing, so a later run that starts on the other parity (after an
// odd number of updates) replays instead of capturing.  The captures only
// record launches; the host-side parity and ghost bookkeeping are put back.
static int ensure_graphs(hydro_sim* s, int n, hipGraphExec_t g[2]) {
  const int par = s->cur;
  if (g[par]) return HYDRO_OK;
  const bool pending = s->ghosts_pending;
  for (int k = 0; k < 2; ++k) {
    const int p = par ^ k;
    if (g[p]) continue;
    s->cur = p;
    const int rc = capture_updates(s, n, &g[p]);
    s->graph_pending[p] = s->ghosts_pending;
    s->cur = par;
    s->ghosts_pending = pending;
    if (rc) return rc;
  }
  return HYDRO_OK;
}

int hydro_sim_run(hydro_sim* s, int nsteps) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (nsteps > 0 && run1d_eligible(s)) return s->mixed ? run1d<float>(s, nsteps) : run1d<double>(s, nsteps);
  // Runs of graph_updates updates, then pairs, replay captured hipGraphs (one
  // submission instead of a launch per kernel); an odd step, or profiling,
  // runs the launches directly.  An even number of updates returns the state
  // to the same buffer parity and the same deferred-ghost bookkeeping, so a
  // capture is valid for every replay.
  // (with a NONE side the ghost copy depends on the bookkeeping at replay
  // time, so those grids launch directly)
  const bool graphs = s->use_graphs && !s->profiling && nsteps >= 2 && boundary_covers_ghosts(s);
  int i = 0;
  if (graphs) {
    const int par = s->cur;
    const int big = s->graph_updates;
    if (big > 2 && nsteps >= big) {
      const int rc = ensure_graphs(s, big, s->graphN);
      if (rc) return rc;
      for (; i + big <= nsteps; i += big) HIP_TRY(hipGraphLaunch(s->graphN[par], s->stream));
    }
    if (i + 2 <= nsteps) {
      const int rc = ensure_graphs(s, 2, s->graph2);
      if (rc) return rc;
      for (; i + 2 <= nsteps; i += 2) HIP_TRY(hipGraphLaunch(s->graph2[par], s->stream));
    }
    s->ghosts_pending = s->graph_pending[par];
  }
  for (; i < nsteps; ++i) {
    int rc = enqueue_update(s);
    if (rc) return rc;
  }
  return HYDRO_OK;
}

int hydro_sim_sync(hydro_sim* s) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return HYDRO_OK;
}

int hydro_sim_time(hydro_sim* s, double* t, long long* steps, double* last_dt) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(s->host_sc, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (t) *t = s->host_sc->t;
  if (steps) *steps = s->host_sc->steps;
  if (last_dt) *last_dt = s->host_sc->dt;
  return HYDRO_OK;
}

int hydro_sim_health(hydro_sim* s, unsigned* flags_out) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(s->host_sc, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (flags_out) *flags_out = s->host_sc->health;
  return HYDRO_OK;
}

int hydro_sim_set_profiling(hydro_sim* s, int enable) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (auto& p : s->ev) {
    hipEventDestroy(p.e0);
    hipEventDestroy(p.e1);
  }
  s->ev.clear();
  s->profiling = enable != 0;
  return HYDRO_OK;
}

int hydro_sim_stage_timing(hydro_sim* s, double* avg_ms, long long* launches) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  double tot = 0.0;
  long long n = 0;
  for (auto& p : s->ev) {
    if (p.phase != HYDRO_PHASE_STAGES) continue;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.e0, p.e1));
    tot += ms;
    ++n;
  }
  if (launches) *launches = n;
  if (avg_ms) *avg_ms = n ? tot / (double)n : 0.0;
  return HYDRO_OK;
}

int hydro_sim_phase_timing(hydro_sim* s, double* total_ms, long long* launches) {
  if (!s || !total_ms) return fail(HYDRO_E_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int i = 0; i < HYDRO_NUM_PHASES; ++i) {
    total_ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  for (auto& p : s->ev) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.e0, p.e1));
    total_ms[p.phase] += ms;
    if (launches) ++launches[p.phase];
  }
  return HYDRO_OK;
}

int hydro_sim_decomposition(hydro_sim* s, int* params_out) {
  if (!s || !params_out) return fail(HYDRO_E_INVALID, "null argument");
  params_out[0] = s->rows;
  params_out[1] = s->wpb;
  params_out[2] = first_order_3d(s) ? s->form3d : 0;
  params_out[3] = first_order_3d(s) ? s->zwaves : roe_family_3d(s) ? s->rwaves : 0;
  return HYDRO_OK;
}

int hydro_sim_autotune(hydro_sim* s, int updates, int* params_out, double* ms_per_update_out) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (updates <= 0) updates = 4;
  if (s->dim == 1) {  // one row per wave; small grids run whole updates in one workgroup
    if (params_out) hydro_sim_decomposition(s, params_out);
    if (ms_per_update_out) *ms_per_update_out = 0.0;
    return HYDRO_OK;
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  int rc = materialize_ghosts(s);
  if (rc) return rc;
  // the trial updates run on a copy: state, low words, primitives, clocks
  const size_t bytes = esize(s) * (size_t)s->plane * s->N;
  void* saveU = nullptr;
  float* saveLo = nullptr;
  double* savePrim = nullptr;
  HIP_TRY(hipMalloc(&saveU, bytes));
  HIP_TRY(hipMemcpyAsync(saveU, s->U[s->cur], bytes, hipMemcpyDeviceToDevice, s->stream));
  if (s->kahan) {
    HIP_TRY(hipMalloc((void**)&saveLo, sizeof(float) * s->plane * s->N));
    HIP_TRY(hipMemcpyAsync(saveLo, s->Ulo[s->cur], sizeof(float) * s->plane * s->N, hipMemcpyDeviceToDevice,
                           s->stream));
  }
  if (s->prim) {
    HIP_TRY(hipMalloc((void**)&savePrim, prim_bytes(s)));
    HIP_TRY(hipMemcpyAsync(savePrim, s->prim, prim_bytes(s), hipMemcpyDeviceToDevice, s->stream));
  }
  void* savePot = nullptr;  // the relaxation sweeps of the trials move the potential
  if (s->pot_on && s->d.use_gravity) {
    HIP_TRY(hipMalloc(&savePot, esize(s) * (size_t)s->plane));
    HIP_TRY(hipMemcpyAsync(savePot, s->pot[0], esize(s) * (size_t)s->plane, hipMemcpyDeviceToDevice, s->stream));
  }
  DevScalars sc0[2];
  HIP_TRY(hipMemcpyAsync(sc0, s->sc, sizeof sc0, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int cur0 = s->cur;
  auto restore = [&]() -> int {
    s->cur = cur0;
    s->ghosts_pending = false;
    HIP_TRY(hipMemcpyAsync(s->U[s->cur], saveU, bytes, hipMemcpyDeviceToDevice, s->stream));
    if (saveLo)
      HIP_TRY(hipMemcpyAsync(s->Ulo[s->cur], saveLo, sizeof(float) * s->plane * s->N, hipMemcpyDeviceToDevice,
                             s->stream));
    if (savePrim)
      HIP_TRY(hipMemcpyAsync(s->prim, savePrim, prim_bytes(s), hipMemcpyDeviceToDevice, s->stream));
    if (savePot)
      HIP_TRY(hipMemcpyAsync(s->pot[0], savePot, esize(s) * (size_t)s->plane, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->sc, sc0, sizeof sc0, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return HYDRO_OK;
  };

  // candidates: the current choice first, then the grid around it
  const TuneParams base = current_params(s);
  std::vector<TuneParams> cand{base};
  auto add = [&](TuneParams p) {
    for (const auto& c : cand)
      if (c.rows == p.rows && c.wpb == p.wpb && c.form == p.form && c.zwaves == p.zwaves) return;
    cand.push_back(p);
  };
  if (s->dim == 2) {
    for (int r : {base.rows, 8, 12, 16, 20, 24, 32})
      for (int w : {1, 2, 4}) add(TuneParams{r, w, 0, base.zwaves, 0.0});
  } else if (first_order_3d(s)) {
    std::vector<int> forms{kForm3dYMarch, kForm3dZm1, kForm3dZm1w8};
    if (s->mixed) forms.push_back(kForm3dZm2);
    for (int f : forms) {
      for (int zw : (f == kForm3dYMarch ? std::vector<int>{4, 8} : std::vector<int>{0})) {
        set_form3d(s, f, zw);
        const int r0 = s->rows;
        add(TuneParams{r0, base.wpb, f, s->zwaves, 0.0});
        if (f != kForm3dYMarch) {
          add(TuneParams{std::max(4, r0 / 2), base.wpb, f, s->zwaves, 0.0});
          add(TuneParams{std::min(s->sz[2] - 4, 2 * r0), base.wpb, f, s->zwaves, 0.0});
        }
      }
    }
  } else {
    // 3D Roe family: z-planes per chunk x rows per workgroup (16 for FP32 only: LDS)
    for (int w : s->mixed ? std::vector<int>{8, 16} : std::vector<int>{8})
      for (int r : {base.rows, pick_planes_3d_roe(s, w), 8, 16, 32, 64}) add(TuneParams{r, base.wpb, 0, w, 0.0});
  }

  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  TuneParams best = base;
  best.ms = 1e30;
  for (const TuneParams& c : cand) {
    apply_params(s, c);
    if ((rc = restore())) break;
    if ((rc = enqueue_update(s))) break;  // warm-up
    if (hipEventRecord(e0, s->stream) != hipSuccess) { rc = fail(HYDRO_E_HIP, "hipEventRecord"); break; }
    for (int k = 0; k < updates && !rc; ++k) rc = enqueue_update(s);
    if (rc) break;
    if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(HYDRO_E_HIP, "hipEventRecord");
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / updates;
    if (per < best.ms) {
      best = c;
      best.ms = per;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  apply_params(s, rc ? base : best);
  const int rc2 = restore();
  hipFree(saveU);
  if (saveLo) hipFree(saveLo);
  if (savePrim) hipFree(savePrim);
  if (savePot) hipFree(savePot);
  if (rc) return rc;
  if (rc2) return rc2;
  tune_store(tune_key(s), best);
  if (params_out) hydro_sim_decomposition(s, params_out);
  if (ms_per_update_out) *ms_per_update_out = best.ms;
  return HYDRO_OK;
}

int hydro_sim_totals(hydro_sim* s, double* totals_out) {
  if (!s || !totals_out) return fail(HYDRO_E_INVALID, "null argument");
  {
    PhaseScope ph(s, HYDRO_PHASE_TOTALS);
    if (s->mixed)
      launch_totals_f32(s->dim, s->sz[0], s->sz[1], s->sz[2], s->plane, s->N, (const float*)s->U[s->cur],
                        s->kahan ? s->Ulo[s->cur] : nullptr, s->partial, s->tot_blocks, s->stream);
    else
      launch_totals_f64(s->dim, s->sz[0], s->sz[1], s->sz[2], s->plane, s->N, (const double*)s->U[s->cur],
                        s->partial, s->tot_blocks, s->stream);
  }
  HIP_TRY(hipGetLastError());
  std::vector<double> part((size_t)s->tot_blocks * s->N);
  HIP_TRY(hipMemcpyAsync(part.data(), s->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  double cellvol = 1.0;
  for (int a = 0; a < s->dim; ++a)
    cellvol *= (s->d.xmax[a] - s->d.xmin[a]) / (double)(float)s->d.size[a];
  for (int k = 0; k < s->N; ++k) {
    double acc = 0.0;
    for (int b = 0; b < s->tot_blocks; ++b) acc += part[(size_t)b * s->N + k];
    totals_out[k] = acc * cellvol;
  }
  return HYDRO_OK;
}

// ---- generic stage evaluation (host-side integrators)

size_t hydro_sim_element_size(const hydro_sim* s) { return s ? esize(s) : 0; }

int hydro_integrator_tableau(int integrator, int* order, double* alpha, double* beta) {
  Tableau t;
  if (!order || !alpha || !beta) return fail(HYDRO_E_INVALID, "null argument");
  if (!get_tableau(integrator, t)) return fail(HYDRO_E_INVALID, "unknown integrator");
  *order = t.order;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) {
      alpha[4 * i + k] = t.a[i][k];
      beta[4 * i + k] = t.b[i][k];
    }
  return HYDRO_OK;
}

int hydro_sim_buffer_create(hydro_sim* s, void** buf) {
  if (!s || !buf) return fail(HYDRO_E_INVALID, "null argument");
  const size_t bytes = esize(s) * (size_t)s->plane * s->N;
  void* p = nullptr;
  HIP_TRY(hipMalloc(&p, bytes));
  // zeroed in the simulation's stream order (a null-stream memset is not
  // ordered against the non-blocking stream that uses the buffer next)
  hipError_t e = hipMemsetAsync(p, 0, bytes, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    hipFree(p);
    return fail(HYDRO_E_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  }
  s->user_bufs.push_back(p);
  *buf = p;
  return HYDRO_OK;
}

int hydro_sim_buffer_destroy(hydro_sim* s, void* buf) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  auto it = std::find(s->user_bufs.begin(), s->user_bufs.end(), buf);
  if (it == s->user_bufs.end()) return fail(HYDRO_E_INVALID, "not a buffer of this simulation");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipFree(buf));
  s->user_bufs.erase(it);
  return HYDRO_OK;
}

int hydro_sim_state_buffer(hydro_sim* s, void** buf) {
  if (!s || !buf) return fail(HYDRO_E_INVALID, "null argument");
  const int rc = materialize_ghosts(s);
  if (rc) return rc;
  *buf = s->U[s->cur];
  return HYDRO_OK;
}

int hydro_sim_init_step(hydro_sim* s) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  int rc = generic_supported(s);
  if (rc) return rc;
  rc = materialize_ghosts(s);
  if (rc) return rc;
  // the frozen start-of-step state: the fused path's scratch buffer holds it
  HIP_TRY(hipMemcpyAsync(s->U[s->cur ^ 1], s->U[s->cur], esize(s) * (size_t)s->plane * s->N,
                         hipMemcpyDeviceToDevice, s->stream));
  s->frozen = true;
  return HYDRO_OK;
}

int hydro_sim_calc_deriv(hydro_sim* s, double dt, void* deriv) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  int rc = generic_supported(s);
  if (rc) return rc;
  if (!s->frozen) return fail(HYDRO_E_INVALID, "hydro_sim_init_step must precede hydro_sim_calc_deriv");
  if (!known_buffer(s, deriv) || deriv == s->U[s->cur])
    return fail(HYDRO_E_INVALID, "the derivative must be a buffer from hydro_sim_buffer_create");
  return s->mixed ? enqueue_deriv<float>(s, dt, deriv) : enqueue_deriv<double>(s, dt, deriv);
}

int hydro_sim_buffer_copy(hydro_sim* s, void* dst, const void* src) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (!known_buffer(s, dst) || !known_buffer(s, src)) return fail(HYDRO_E_INVALID, "unknown buffer");
  if (dst == src) return HYDRO_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, esize(s) * (size_t)s->plane * s->N, hipMemcpyDeviceToDevice, s->stream));
  return HYDRO_OK;
}

int hydro_sim_buffer_zero(hydro_sim* s, void* buf) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (!known_buffer(s, buf)) return fail(HYDRO_E_INVALID, "unknown buffer");
  launch_zero_words(buf, (long long)(esize(s) / 4) * s->plane * s->N, s->stream);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

int hydro_sim_mult_add(hydro_sim* s, void* a, const void* b, const void* c, double k) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  if (!known_buffer(s, a) || !known_buffer(s, b) || !known_buffer(s, c)) return fail(HYDRO_E_INVALID, "unknown buffer");
  const long long n = s->plane * s->N;
  if (s->mixed) launch_mult_add_f32((float*)a, (const float*)b, (const float*)c, (float)k, n, s->stream);
  else launch_mult_add_f64((double*)a, (const double*)b, (const double*)c, k, n, s->stream);
  HIP_TRY(hipGetLastError());
  return HYDRO_OK;
}

int hydro_sim_finish_step(hydro_sim* s, double dt) {
  if (!s) return fail(HYDRO_E_INVALID, "null argument");
  int rc = generic_supported(s);
  if (rc) return rc;
  // the clock advanced by dt (the tail of findMinTimestep with a given dt)
  FinalizeArgs f{};
  f.use_fixed = 1;
  f.fixed_dt = dt;
  f.dt = &s->sc->dt;
  f.t = &s->sc->t;
  f.steps = &s->sc->steps;
  f.health = &s->sc->health;
  launch_finalize_dt(f, s->stream);
  HIP_TRY(hipGetLastError());
  s->frozen = false;
  s->ghosts_pending = false;  // the integrator combined every cell, ghosts included
  if (s->d.equation == HYDRO_EQN_SRHD) return enqueue_srhd_tail(s);
  return HYDRO_OK;
}

}  // extern "C"
