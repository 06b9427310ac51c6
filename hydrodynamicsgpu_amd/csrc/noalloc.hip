// noalloc.hip -- the CFL timestep of hydro_sim_calc_timestep for a state the
// caller holds in device memory, without allocating anything.
//
// The pass needs scratch memory beyond its input and output: one minimum
// per wave (or per block) of the CFL kernel and the clock the dt tail reads
// and advances.  Here that memory is the file-scope device array `scratch`
// below (kScratchBytes per GPU, held until the process ends).  Its device
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and only that address reaches kernels.  Every call on one
// GPU shares the array: calls that overlap on one GPU (two streams, two
// threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/hydro_capi.h"
#include "cfl.hpp"
#include "host_common.hpp"
#include "launch.hpp"

using namespace hydro;
using namespace hydro::host;

namespace {

// layout: the clock (DevScalars, padded to 256 bytes), then the minima
constexpr size_t kClockBytes = 256;
constexpr long long kScratchMinima = kDtMaxWaves;
constexpr size_t kScratchBytes = kClockBytes + sizeof(double) * (size_t)kScratchMinima;
static_assert(sizeof(DevScalars) <= kClockBytes, "clock must fit its slot");
static_assert(kScratchBytes == HYDRO_NOALLOC_SCRATCH_BYTES, "header and library disagree on the scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// the clock of one call: t as given, no steps, no health bits (a kernel,
// not hipMemsetAsync, so the call can be captured into a graph)
__global__ void noalloc_clock_kernel(DevScalars* c, double t) {
  if (threadIdx.x == 0) {
    c->dt = 0.0;
    c->t = t;
    c->steps = 0;
    c->health = 0u;
    c->pad = 0u;
  }
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current device
int scratch_address(unsigned char** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(HYDRO_E_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(HYDRO_E_HIP, "hipGetSymbolAddress returned a null address for the scratch array");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return HYDRO_OK;
}

}  // namespace

extern "C" {

int hydro_calc_timestep_noalloc(const hydro_sim_desc* d, const void* state, size_t plane, const double* prims,
                                double t, double* dt_out, unsigned* flags_out, void* stream) {
  if (int rc = validate(d)) return rc;
  if (!state || !dt_out) return fail(HYDRO_E_INVALID, "null argument");
  if (d->use_gravity && d->equation == HYDRO_EQN_EULER)
    return fail(HYDRO_E_UNSUPPORTED, "the allocation-free timestep does not read a gravitational potential");
  const long long vol = (long long)d->size[0] * d->size[1] * d->size[2];
  if ((long long)plane < vol) return fail(HYDRO_E_INVALID, "plane stride must be at least the cell count");
  if (d->equation == HYDRO_EQN_SRHD && !prims && !d->use_fixed_dt)
    return fail(HYDRO_E_INVALID, "the SRHD timestep reads the primitives");
  if (d->equation != HYDRO_EQN_SRHD && prims) return fail(HYDRO_E_UNSUPPORTED, "primitives exist for SRHD only");
  if (!(t >= 0.0)) return fail(HYDRO_E_INVALID, "time must be >= 0");
  CflGrid g = cfl_grid(*d, (long long)plane);
  const long long need = cfl_minima(g);
  if (need > kScratchMinima)
    return fail(HYDRO_E_INVALID, "the CFL pass of this grid needs " + std::to_string(need) +
                                     " minima, more than the scratch array holds (" +
                                     std::to_string(kScratchMinima) + ")");
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  DevScalars* clock = reinterpret_cast<DevScalars*>(base);
  double* minima = reinterpret_cast<double*>(base + kClockBytes);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(noalloc_clock_kernel, dim3(1), dim3(64), 0, st, clock, t);
  HIP_TRY(hipGetLastError());
  FinalizeArgs f{};
  f.cfl = g.cfl;
  f.use_fixed = d->use_fixed_dt;
  f.fixed_dt = d->fixed_dt;
  f.clip = (d->flags & HYDRO_FLAG_CLIP_TO_TEND) ? 1 : 0;
  f.t_end = d->t_end;
  f.dt = dt_out;
  f.t = &clock->t;
  f.steps = &clock->steps;
  f.health = flags_out ? flags_out : &clock->health;
  return enqueue_cfl(g, state, prims, true, minima, f, st);
}

}  // extern "C"
