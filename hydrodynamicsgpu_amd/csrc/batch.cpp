// batch.cpp -- C ABI of the batched launch: many independent Euler grids of
// uneven sizes (1D and 2D mixed) advanced together.
//
// The reference runs one grid per program (src/Solver/Solver.cpp:133-153,
// 261-307: one OpenCL program compiled per configuration).  Here every member
// lives in shared arenas and one step is five launches for the whole batch:
// the composed boundary of every member, the CFL strip march over all
// members' tasks, the per-member dt tail, then one persistent launch per RK
// stage whose waves pull (member, strip, row chunk) tasks.  Each member keeps
// its own dt, time and step count; the arithmetic per cell is the
// single-grid path's, so a member's results equal a lone grid's bit for bit.
// Pairs of steps are captured in a hipGraph and replayed.  An autotuner picks
// the row-chunk height of 2D tasks and the strip count of 1D tasks by timing
// candidate decompositions on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch_args.hpp"
#include "host_common.hpp"

using namespace hydro;
using namespace hydro::host;

struct hydro_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int count = 0;
  std::vector<hydro_sim_desc> d;
  int solver = 0, limiter = 0, integrator = 0;
  bool mixed = false, kahan = false;
  unsigned flags = 0;
  int use_fixed = 0;
  double fixed_dt = 0;
  double gamma = 1.4;
  Tableau tab;
  std::vector<StagePlan> plan;
  int nslots = 0, ndbuf = 0;
  std::vector<BatchMember> mem;
  long long total = 0;  // elements per arena
  void* U[2] = {nullptr, nullptr};
  float* Ulo[2] = {nullptr, nullptr};
  void* D[3] = {nullptr, nullptr, nullptr};
  void* P[kMaxPartials] = {nullptr, nullptr, nullptr};
  BatchMember* d_mem = nullptr;
  BatchTask* d_tasks = nullptr;
  int ntasks = 0, cap_tasks = 0;
  BatchTask* d_dtasks = nullptr;
  int ndtasks = 0, cap_dtasks = 0;
  double* d_taskmin = nullptr;
  unsigned* d_counters = nullptr;
  double* sc = nullptr;       // per member: dt, t, steps, health (kBatchDtStride doubles)
  double* host_sc = nullptr;  // pinned mirror
  int cur = 0;
  bool ghosts_pending = false;
  int rows = 16, kstrips = 4, blocks = 1024;
  int max_ghosts = 0;
  hipGraphExec_t graph2[2] = {nullptr, nullptr};
  bool graph_pending[2] = {false, false};  // ghosts_pending after the captured pair
  bool use_graphs = true;
};

namespace {

size_t esize(const hydro_batch* b) { return b->mixed ? sizeof(float) : sizeof(double); }

long long ghost_cells(int dim, int sx, int sy) { return dim == 1 ? 4 : 4LL * sx + 4LL * (sy - 4); }

void drop_graphs(hydro_batch* b) {
  for (auto& g : b->graph2)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// (re)build the stage and CFL task lists for row-chunk height `rows` (2D)
// and `k` strips per task (1D); stage tasks largest first
int build_tasks(hydro_batch* b, int rows, int k) {
  std::vector<BatchTask> st, dt;
  std::vector<double> cost;
  for (int m = 0; m < b->count; ++m) {
    BatchMember& M = b->mem[m];
    const int nstrips = (M.sx - 4 + 59) / 60;
    if (M.dim == 2) {
      const int ny = M.sy - 4;
      for (int c = 0; c * rows < ny; ++c)
        for (int s = 0; s < nstrips; ++s) {
          st.push_back({m, 2, s, c});
          cost.push_back(60.0 * (std::min(rows, ny - c * rows) + 3));
        }
    } else {
      for (int s = 0; s < nstrips; s += k) {
        const int n = std::min(k, nstrips - s);
        st.push_back({m, 1, s, n});
        cost.push_back(60.0 * 0.6 * n);
      }
    }
    // CFL tasks: the strip march's decomposition (at most kDtRows rows)
    M.dt_task0 = (int)dt.size();
    const int dstrips = (M.last_x - 2) / 60 + 1;
    if (M.dim == 2) {
      const int ny = M.last_y - 1;
      for (int c = 0; c * kDtRows < ny; ++c)
        for (int s = 0; s < dstrips; ++s) dt.push_back({m, 2, s, c});
    } else {
      for (int s = 0; s < dstrips; s += k) dt.push_back({m, 1, s, std::min(k, dstrips - s)});
    }
    M.dt_ntasks = (int)dt.size() - M.dt_task0;
  }
  std::vector<int> order(st.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<BatchTask> sorted(st.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = st[order[i]];
  if ((int)sorted.size() > b->cap_tasks) {
    if (b->d_tasks) hipFree(b->d_tasks);
    b->d_tasks = nullptr;
    HIP_TRY(hipMalloc((void**)&b->d_tasks, sizeof(BatchTask) * sorted.size()));
    b->cap_tasks = (int)sorted.size();
  }
  if ((int)dt.size() > b->cap_dtasks) {
    if (b->d_dtasks) hipFree(b->d_dtasks);
    if (b->d_taskmin) hipFree(b->d_taskmin);
    b->d_dtasks = nullptr;
    b->d_taskmin = nullptr;
    HIP_TRY(hipMalloc((void**)&b->d_dtasks, sizeof(BatchTask) * dt.size()));
    HIP_TRY(hipMalloc((void**)&b->d_taskmin, sizeof(double) * dt.size()));
    b->cap_dtasks = (int)dt.size();
  }
  HIP_TRY(hipMemcpyAsync(b->d_tasks, sorted.data(), sizeof(BatchTask) * sorted.size(), hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_dtasks, dt.data(), sizeof(BatchTask) * dt.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_mem, b->mem.data(), sizeof(BatchMember) * b->count, hipMemcpyHostToDevice,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->ntasks = (int)sorted.size();
  b->ndtasks = (int)dt.size();
  b->rows = rows;
  b->kstrips = k;
  drop_graphs(b);
  return HYDRO_OK;
}

template <typename T>
void fill_template(hydro_batch* b, int k, StageArgs<T, T>& a) {
  const StagePlan& p = b->plan[k];
  a = StageArgs<T, T>{};
  a.u0 = (const T*)b->U[b->cur];
  a.u0lo = b->kahan ? b->Ulo[b->cur] : nullptr;
  a.dk = p.din < 0 ? nullptr : (const T*)b->D[p.din];
  a.dnext = p.dout < 0 ? nullptr : (T*)b->D[p.dout];
  a.nrep = p.nrep;
  for (int j = 0; j < kMaxPartials; ++j) {
    a.rep[j] = j < p.nrep ? (const T*)b->D[p.rep[j]] : nullptr;
    a.rc[j] = p.rc[j];
  }
  a.unew = (T*)b->U[b->cur ^ 1];
  a.unewlo = b->kahan ? b->Ulo[b->cur ^ 1] : nullptr;
  a.pkp = p.pk < 0 ? nullptr : (const T*)b->P[p.pk];
  for (int j = 0; j < kMaxPartials; ++j) {
    a.fp[j] = j < p.nf ? (T*)b->P[p.fslot[j]] : nullptr;
    a.fload[j] = p.fload[j];
    a.am[j] = p.am[j];
    a.bm[j] = p.bm[j];
  }
  a.nf = p.nf;
  a.a_self = p.a_self;
  a.b_self = p.b_self;
  a.final_stage = p.final_stage ? 1 : 0;
  a.gamma = b->gamma;
  a.lim = b->limiter;
  a.rows = b->rows;
  a.sz = 1;
}

int enqueue_step(hydro_batch* b) {
  BatchBoundaryArgs bb{b->d_mem, b->max_ghosts};
  if (b->mixed) launch_batch_boundary_f32(bb, b->count, (float*)b->U[b->cur], b->kahan ? b->Ulo[b->cur] : nullptr,
                                          b->stream);
  else launch_batch_boundary_f64(bb, b->count, (double*)b->U[b->cur], b->stream);
  b->ghosts_pending = false;  // every side has a method: the boundary rewrites all ghosts
  if (!b->use_fixed) {
    BatchDtArgs da{b->d_mem, b->d_dtasks, b->ndtasks, b->gamma, b->solver, b->d_taskmin};
    if (b->mixed) launch_batch_dt_f32(da, (const float*)b->U[b->cur], b->stream);
    else launch_batch_dt_f64(da, (const double*)b->U[b->cur], b->stream);
  }
  const int S = b->tab.order;
  BatchFinalizeArgs fa{b->d_mem, b->d_taskmin, b->sc, b->use_fixed,
                       (b->flags & HYDRO_FLAG_CLIP_TO_TEND) ? 1 : 0, b->fixed_dt, b->d_counters, S};
  launch_batch_finalize(fa, b->count, b->stream);
  for (int k = 0; k < S; ++k) {
    const int mode = b->plan[k].mode;
    if (b->mixed) {
      BatchStageArgs<float, float> a{};
      fill_template<float>(b, k, a.tmpl);
      a.members = b->d_mem;
      a.tasks = b->d_tasks;
      a.ntasks = b->ntasks;
      a.counter = b->d_counters + k;
      a.dts = b->sc;
      if (b->kahan) launch_batch_stage_f32k(b->solver, b->limiter, mode, a, b->blocks, b->stream);
      else launch_batch_stage_f32(b->solver, b->limiter, mode, a, b->blocks, b->stream);
    } else {
      BatchStageArgs<double, double> a{};
      fill_template<double>(b, k, a.tmpl);
      a.members = b->d_mem;
      a.tasks = b->d_tasks;
      a.ntasks = b->ntasks;
      a.counter = b->d_counters + k;
      a.dts = b->sc;
      launch_batch_stage_f64(b->solver, b->limiter, mode, a, b->blocks, b->stream);
    }
  }
  HIP_TRY(hipGetLastError());
  b->cur ^= 1;
  b->ghosts_pending = true;
  // ForwardEuler's zero-gravity pass refreshes the ghosts after the update
  // (src/Solver/SelfGravitation.cpp:122 + src/Integrator/ForwardEuler.cpp:29-31)
  if ((b->flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS) && b->integrator == HYDRO_INTEGRATOR_FORWARDEULER) {
    if (b->mixed)
      launch_batch_boundary_f32(bb, b->count, (float*)b->U[b->cur], b->kahan ? b->Ulo[b->cur] : nullptr, b->stream);
    else
      launch_batch_boundary_f64(bb, b->count, (double*)b->U[b->cur], b->stream);
    HIP_TRY(hipGetLastError());
    b->ghosts_pending = false;
  }
  return HYDRO_OK;
}

// the new state's ghost cells are those of the state the step started from
int materialize_ghosts(hydro_batch* b) {
  if (!b->ghosts_pending) return HYDRO_OK;
  const int prev = b->cur ^ 1;
  for (const BatchMember& m : b->mem) {
    if (b->mixed) {
      launch_copy_ghosts_f32(m.dim, m.sx, m.sy, 1, m.plane, m.nstates, (const float*)b->U[prev] + m.off,
                             (float*)b->U[b->cur] + m.off, b->stream);
      if (b->kahan)
        launch_copy_ghosts_f32(m.dim, m.sx, m.sy, 1, m.plane, m.nstates, b->Ulo[prev] + m.off, b->Ulo[b->cur] + m.off,
                               b->stream);
    } else {
      launch_copy_ghosts_f64(m.dim, m.sx, m.sy, 1, m.plane, m.nstates, (const double*)b->U[prev] + m.off,
                             (double*)b->U[b->cur] + m.off, b->stream);
    }
  }
  HIP_TRY(hipGetLastError());
  b->ghosts_pending = false;
  return HYDRO_OK;
}

// Records two steps into an instantiated graph (launches only; the host-side
// parity and ghost bookkeeping end where two steps leave them).
int capture_pair(hydro_batch* b, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(b);
  if (!rc) rc = enqueue_step(b);
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
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

int run_steps(hydro_batch* b, int nsteps, bool graphs) {
  int i = 0;
  if (graphs && nsteps >= 2) {
    const int par = b->cur;
    if (!b->graph2[par]) {
      // both buffer parities at once, so a run that starts after an odd
      // number of steps replays instead of capturing (every side of every
      // member has a boundary method: the captured steps do not depend on
      // the ghost bookkeeping)
      const bool pending = b->ghosts_pending;
      for (int k = 0; k < 2; ++k) {
        const int p = par ^ k;
        if (b->graph2[p]) continue;
        b->cur = p;
        const int rc = capture_pair(b, &b->graph2[p]);
        b->graph_pending[p] = b->ghosts_pending;
        b->cur = par;
        b->ghosts_pending = pending;
        if (rc) return rc;
      }
    }
    for (; i + 2 <= nsteps; i += 2) HIP_TRY(hipGraphLaunch(b->graph2[par], b->stream));
    b->ghosts_pending = b->graph_pending[par];
  }
  for (; i < nsteps; ++i) {
    const int rc = enqueue_step(b);
    if (rc) return rc;
  }
  return HYDRO_OK;
}

int check_member(const hydro_batch* b, int m) {
  if (!b) return fail(HYDRO_E_INVALID, "null batch");
  if (m < 0 || m >= b->count) return fail(HYDRO_E_INVALID, "member index out of range");
  return HYDRO_OK;
}

}  // namespace

extern "C" {

int hydro_batch_create(hydro_ctx* ctx, const hydro_sim_desc* descs, int count, hydro_batch** out) {
  if (!ctx || !descs || !out) return fail(HYDRO_E_INVALID, "null argument");
  if (count < 1) return fail(HYDRO_E_INVALID, "a batch needs at least one member");
  const hydro_sim_desc& d0 = descs[0];
  for (int m = 0; m < count; ++m) {
    const hydro_sim_desc& d = descs[m];
    if (int rc = validate(&d)) return rc;
    if (d.equation != HYDRO_EQN_EULER || d.solver == HYDRO_SOLVER_BURGERS)
      return fail(HYDRO_E_UNSUPPORTED, "batched members are Euler grids (Roe, HLL or HLLC)");
    if (d.dim > 2) return fail(HYDRO_E_UNSUPPORTED, "batched members are 1D or 2D");
    if (d.use_gravity) return fail(HYDRO_E_UNSUPPORTED, "batched members have no self-gravity (use_gravity)");
    if (d.solver != d0.solver || d.limiter != d0.limiter || d.integrator != d0.integrator ||
        d.precision != d0.precision || d.flags != d0.flags || d.use_fixed_dt != d0.use_fixed_dt ||
        (d.use_fixed_dt && d.fixed_dt != d0.fixed_dt) || d.gamma != d0.gamma)
      return fail(HYDRO_E_INVALID,
                  "batch members share solver, limiter, integrator, precision, flags, gamma and fixed dt");
    for (int a = 0; a < d.dim; ++a)
      for (int s = 0; s < 2; ++s)
        if (d.boundary[a][s] == HYDRO_BOUNDARY_NONE)
          return fail(HYDRO_E_UNSUPPORTED, "batched members need a boundary method on every side");
  }
  hydro_batch* b = new hydro_batch;
  auto cleanup = [&](int code, const std::string& msg) {
    hydro_batch_destroy(b);
    return fail(code, msg);
  };
  b->device = ctx->device;
  b->count = count;
  b->d.assign(descs, descs + count);
  b->solver = d0.solver;
  b->limiter = d0.limiter;
  b->integrator = d0.integrator;
  b->mixed = d0.precision == HYDRO_PRECISION_FP32_MIXED;
  b->kahan = b->mixed && (d0.flags & HYDRO_FLAG_KAHAN);
  b->flags = d0.flags;
  b->use_fixed = d0.use_fixed_dt;
  b->fixed_dt = d0.fixed_dt;
  b->gamma = d0.gamma;
  if (!get_tableau(b->integrator, b->tab)) return cleanup(HYDRO_E_INVALID, "integrator not supported");
  if (plan_replay(b->tab, b->plan)) {
    b->ndbuf = b->tab.order - 1;
  } else if (plan_stages(b->tab, b->plan, b->nslots)) {
    b->ndbuf = b->tab.order > 1 ? 2 : 0;
  } else {
    return cleanup(HYDRO_E_INVALID, "integrator not supported");
  }
  for (auto& p : b->plan) p.mode = combine_mode(p);

  b->mem.resize(count);
  long long off = 0;
  for (int m = 0; m < count; ++m) {
    const hydro_sim_desc& d = descs[m];
    BatchMember& M = b->mem[m];
    std::memset(&M, 0, sizeof M);
    M.dim = d.dim;
    M.sx = d.size[0];
    M.sy = d.dim > 1 ? d.size[1] : 1;
    M.nstates = d.dim + 2;
    M.plane = ((long long)M.sx * M.sy + 63) & ~63LL;
    M.off = off;
    off += M.plane * M.nstates;
    for (int a = 0; a < 2; ++a) {
      const double szf = a < d.dim ? (double)(float)d.size[a] : 1.0;
      const double kdx = kernel_dx((d.xmax[a] - d.xmin[a]) / szf, (d.flags & HYDRO_FLAG_REF_DX_6DIGITS) != 0);
      M.kdx[a] = kdx;
      M.inv_dx[a] = 1.0 / kdx;
    }
    M.cfl = (double)(float)d.cfl;
    M.t_end = d.t_end;
    const int hi = d.solver == HYDRO_SOLVER_ROE ? 3 : 2;
    M.last_x = M.sx - hi;
    M.last_y = d.dim > 1 ? M.sy - hi : 0;
    M.ghosts = (int)ghost_cells(d.dim, M.sx, M.sy);
    b->max_ghosts = std::max(b->max_ghosts, M.ghosts);
    for (int a = 0; a < 2; ++a)
      for (int s = 0; s < 2; ++s) {
        signed char meth = -1;
        unsigned neg = 0;
        if (a < d.dim) {
          switch (d.boundary[a][s]) {
          case HYDRO_BOUNDARY_PERIODIC: meth = 0; break;
          case HYDRO_BOUNDARY_MIRROR: meth = 1; neg = 1u << (a + 1); break;  // src/Equation/Euler.cpp:70
          case HYDRO_BOUNDARY_FREEFLOW: meth = 2; break;
          default: meth = -1;
          }
        }
        M.method[a][s] = meth;
        M.negmask[a][s] = neg;
      }
  }
  b->total = off;
  if (hipSetDevice(b->device) != hipSuccess) return cleanup(HYDRO_E_HIP, "hipSetDevice");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "stream create failed");
  const size_t bytes = esize(b) * (size_t)b->total;
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&b->U[i], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (states)");
    if (hipMemset(b->U[i], 0, bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
    if (b->kahan) {
      if (hipMalloc((void**)&b->Ulo[i], sizeof(float) * b->total) != hipSuccess)
        return cleanup(HYDRO_E_HIP, "out of device memory (low words)");
      if (hipMemset(b->Ulo[i], 0, sizeof(float) * b->total) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
    }
  }
  // increments and partial sums: their ghost cells stay 0 (stages write interior cells only)
  for (int i = 0; i < b->ndbuf; ++i) {
    if (hipMalloc(&b->D[i], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (stages)");
    if (hipMemset(b->D[i], 0, bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
  }
  for (int j = 0; j < b->nslots; ++j) {
    if (hipMalloc(&b->P[j], bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "out of device memory (partials)");
    if (hipMemset(b->P[j], 0, bytes) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
  }
  if (hipMalloc((void**)&b->d_mem, sizeof(BatchMember) * count) != hipSuccess) return cleanup(HYDRO_E_HIP, "alloc");
  if (hipMalloc((void**)&b->d_counters, sizeof(unsigned) * 4) != hipSuccess) return cleanup(HYDRO_E_HIP, "alloc");
  if (hipMemset(b->d_counters, 0, sizeof(unsigned) * 4) != hipSuccess) return cleanup(HYDRO_E_HIP, "memset");
  if (hipMalloc((void**)&b->sc, sizeof(double) * kBatchDtStride * count) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "alloc");
  if (hipMemset(b->sc, 0, sizeof(double) * kBatchDtStride * count) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "memset");
  if (hipHostMalloc((void**)&b->host_sc, sizeof(double) * kBatchDtStride * count) != hipSuccess)
    return cleanup(HYDRO_E_HIP, "pinned alloc");
  // the zero fills above ran on the null stream: complete them before the
  // batch's own stream touches the arenas
  if (hipDeviceSynchronize() != hipSuccess) return cleanup(HYDRO_E_HIP, "sync");
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess || ncu <= 0) ncu = 256;
  b->blocks = 4 * ncu;  // 4 waves per SIMD
  if (const char* e = getenv("HYDRO_GRAPHS")) b->use_graphs = atoi(e) != 0;
  if (int rc = build_tasks(b, 16, 4)) {
    hydro_batch_destroy(b);
    return rc;
  }
  *out = b;
  return HYDRO_OK;
}

void hydro_batch_destroy(hydro_batch* b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (int i = 0; i < 2; ++i) {
    if (b->U[i]) hipFree(b->U[i]);
    if (b->Ulo[i]) hipFree(b->Ulo[i]);
  }
  for (auto* p : b->D)
    if (p) hipFree(p);
  for (auto* p : b->P)
    if (p) hipFree(p);
  if (b->d_mem) hipFree(b->d_mem);
  if (b->d_tasks) hipFree(b->d_tasks);
  if (b->d_dtasks) hipFree(b->d_dtasks);
  if (b->d_taskmin) hipFree(b->d_taskmin);
  if (b->d_counters) hipFree(b->d_counters);
  if (b->sc) hipFree(b->sc);
  if (b->host_sc) hipHostFree(b->host_sc);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

int hydro_batch_set_state(hydro_batch* b, int member, const double* aos, size_t n) {
  if (int rc = check_member(b, member)) return rc;
  if (!aos) return fail(HYDRO_E_INVALID, "null state");
  const BatchMember& M = b->mem[member];
  const long long vol = (long long)M.sx * M.sy;
  const int N = M.nstates;
  if (n != (size_t)N * (size_t)vol) return fail(HYDRO_E_INVALID, "state size must be N * volume of the member");
  // settle the other members' carried-over ghost cells before this one is replaced
  if (int rc = materialize_ghosts(b)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->mixed) {
    std::vector<float> hi((size_t)M.plane * N, 0.f), lo((size_t)M.plane * N, 0.f);
    for (long long c = 0; c < vol; ++c)
      for (int k = 0; k < N; ++k) {
        const double v = aos[k + (size_t)N * c];
        const float h = (float)v;
        hi[(size_t)k * M.plane + c] = h;
        lo[(size_t)k * M.plane + c] = (float)(v - (double)h);
      }
    HIP_TRY(copy_sync((float*)b->U[b->cur] + M.off, hi.data(), hi.size() * sizeof(float), hipMemcpyHostToDevice, b->stream));
    if (b->kahan)
      HIP_TRY(copy_sync(b->Ulo[b->cur] + M.off, lo.data(), lo.size() * sizeof(float), hipMemcpyHostToDevice, b->stream));
  } else {
    std::vector<double> buf((size_t)M.plane * N, 0.0);
    for (long long c = 0; c < vol; ++c)
      for (int k = 0; k < N; ++k) buf[(size_t)k * M.plane + c] = aos[k + (size_t)N * c];
    HIP_TRY(copy_sync((double*)b->U[b->cur] + M.off, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  }
  // a new initial state restarts the member's clock
  double z[kBatchDtStride] = {0, 0, 0, 0};
  HIP_TRY(copy_sync(b->sc + (size_t)member * kBatchDtStride, z, sizeof z, hipMemcpyHostToDevice, b->stream));
  return HYDRO_OK;
}

int hydro_batch_get_state(hydro_batch* b, int member, double* aos, size_t n) {
  if (int rc = check_member(b, member)) return rc;
  if (!aos) return fail(HYDRO_E_INVALID, "null state");
  const BatchMember& M = b->mem[member];
  const long long vol = (long long)M.sx * M.sy;
  const int N = M.nstates;
  if (n != (size_t)N * (size_t)vol) return fail(HYDRO_E_INVALID, "state size must be N * volume of the member");
  if (int rc = materialize_ghosts(b)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->mixed) {
    std::vector<float> hi((size_t)M.plane * N), lo;
    HIP_TRY(copy_sync(hi.data(), (float*)b->U[b->cur] + M.off, hi.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (b->kahan) {
      lo.resize(hi.size());
      HIP_TRY(copy_sync(lo.data(), b->Ulo[b->cur] + M.off, lo.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    }
    for (long long c = 0; c < vol; ++c)
      for (int k = 0; k < N; ++k) {
        double v = hi[(size_t)k * M.plane + c];
        if (b->kahan) v += (double)lo[(size_t)k * M.plane + c];
        aos[k + (size_t)N * c] = v;
      }
  } else {
    std::vector<double> buf((size_t)M.plane * N);
    HIP_TRY(copy_sync(buf.data(), (double*)b->U[b->cur] + M.off, buf.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    for (long long c = 0; c < vol; ++c)
      for (int k = 0; k < N; ++k) aos[k + (size_t)N * c] = buf[(size_t)k * M.plane + c];
  }
  return HYDRO_OK;
}

int hydro_batch_run(hydro_batch* b, int nsteps, int use_graph) {
  if (!b) return fail(HYDRO_E_INVALID, "null batch");
  if (nsteps < 0) return fail(HYDRO_E_INVALID, "nsteps must be >= 0");
  return run_steps(b, nsteps, use_graph != 0 && b->use_graphs);
}

int hydro_batch_sync(hydro_batch* b) {
  if (!b) return fail(HYDRO_E_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return HYDRO_OK;
}

int hydro_batch_member_dt(hydro_batch* b, int member, double* dt_out) {
  double t;
  long long steps;
  return hydro_batch_member_time(b, member, &t, &steps, dt_out);
}

int hydro_batch_member_time(hydro_batch* b, int member, double* t, long long* steps, double* last_dt) {
  if (int rc = check_member(b, member)) return rc;
  HIP_TRY(hipMemcpyAsync(b->host_sc, b->sc, sizeof(double) * kBatchDtStride * b->count, hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  const double* s = b->host_sc + (size_t)member * kBatchDtStride;
  if (last_dt) *last_dt = s[0];
  if (t) *t = s[1];
  if (steps) std::memcpy(steps, s + 2, sizeof(long long));
  return HYDRO_OK;
}

int hydro_batch_member_health(hydro_batch* b, int member, unsigned* flags_out) {
  if (int rc = check_member(b, member)) return rc;
  HIP_TRY(hipMemcpyAsync(b->host_sc, b->sc, sizeof(double) * kBatchDtStride * b->count, hipMemcpyDeviceToHost,
                         b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (flags_out) std::memcpy(flags_out, b->host_sc + (size_t)member * kBatchDtStride + 3, sizeof(unsigned));
  return HYDRO_OK;
}

long long hydro_batch_interior_cells(const hydro_batch* b) {
  if (!b) return 0;
  long long n = 0;
  for (const BatchMember& m : b->mem) n += (long long)(m.sx - 4) * (m.dim > 1 ? m.sy - 4 : 1);
  return n;
}

int hydro_batch_set_decomposition(hydro_batch* b, int rows, int strips) {
  if (!b) return fail(HYDRO_E_INVALID, "null batch");
  if (rows < 1 || strips < 1) return fail(HYDRO_E_INVALID, "rows and strips must be >= 1");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return build_tasks(b, rows, strips);
}

int hydro_batch_autotune(hydro_batch* b, int* rows_out, int* strips_out, double* ms_per_step_out) {
  if (!b) return fail(HYDRO_E_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (int rc = materialize_ghosts(b)) return rc;
  // the trial steps run on a copy: save the state, clocks and parity
  const size_t bytes = esize(b) * (size_t)b->total;
  void* saveU = nullptr;
  float* saveLo = nullptr;
  std::vector<double> save_sc((size_t)kBatchDtStride * b->count);
  HIP_TRY(hipMalloc(&saveU, bytes));
  HIP_TRY(hipMemcpyAsync(saveU, b->U[b->cur], bytes, hipMemcpyDeviceToDevice, b->stream));
  if (b->kahan) {
    HIP_TRY(hipMalloc((void**)&saveLo, sizeof(float) * b->total));
    HIP_TRY(hipMemcpyAsync(saveLo, b->Ulo[b->cur], sizeof(float) * b->total, hipMemcpyDeviceToDevice, b->stream));
  }
  HIP_TRY(hipMemcpyAsync(save_sc.data(), b->sc, sizeof(double) * save_sc.size(), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  const int cur0 = b->cur;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int best_r = b->rows, best_k = b->kstrips;
  float best = 1e30f;
  int rc = HYDRO_OK;
  const int rows_c[] = {8, 16, 32, 64};
  const int k_c[] = {1, 2, 4, 8, 16};
  for (int r : rows_c) {
    for (int k : k_c) {
      if ((rc = build_tasks(b, r, k))) break;
      if ((rc = run_steps(b, 1, false))) break;  // warm-up
      HIP_TRY(hipEventRecord(e0, b->stream));
      if ((rc = run_steps(b, 4, false))) break;
      HIP_TRY(hipEventRecord(e1, b->stream));
      HIP_TRY(hipEventSynchronize(e1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_r = r;
        best_k = k;
      }
    }
    if (rc) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  // restore
  b->cur = cur0;
  b->ghosts_pending = false;
  HIP_TRY(hipMemcpyAsync(b->U[b->cur], saveU, bytes, hipMemcpyDeviceToDevice, b->stream));
  if (b->kahan)
    HIP_TRY(hipMemcpyAsync(b->Ulo[b->cur], saveLo, sizeof(float) * b->total, hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->sc, save_sc.data(), sizeof(double) * save_sc.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  hipFree(saveU);
  if (saveLo) hipFree(saveLo);
  if (rc) return rc;
  if ((rc = build_tasks(b, best_r, best_k))) return rc;
  if (rows_out) *rows_out = best_r;
  if (strips_out) *strips_out = best_k;
  if (ms_per_step_out) *ms_per_step_out = best / 4.0;
  return HYDRO_OK;
}

}  // extern "C"
