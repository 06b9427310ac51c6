// device_common.hpp -- shared device helpers for the gfx950 update path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HYDRO_DEV __device__ __forceinline__

namespace hydro {

// ---------------------------------------------------------------- lane shifts
// Full-wave DPP row shifts (wave_shr:1 / wave_shl:1 on CDNA): lane l receives
// lane l-1 (shr) or lane l+1 (shl).  Edge lanes receive 0; callers only use
// lanes whose neighbour exists.
HYDRO_DEV float lane_from_left(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
HYDRO_DEV float lane_from_right(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}
HYDRO_DEV double lane_from_left(double v) {
  int2 p = __builtin_bit_cast(int2, v);
  p.x = __builtin_amdgcn_update_dpp(0, p.x, 0x138, 0xF, 0xF, true);
  p.y = __builtin_amdgcn_update_dpp(0, p.y, 0x138, 0xF, 0xF, true);
  return __builtin_bit_cast(double, p);
}
HYDRO_DEV double lane_from_right(double v) {
  int2 p = __builtin_bit_cast(int2, v);
  p.x = __builtin_amdgcn_update_dpp(0, p.x, 0x130, 0xF, 0xF, true);
  p.y = __builtin_amdgcn_update_dpp(0, p.y, 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(double, p);
}

// Full-wave DPP rotations (wave_ror:1 / wave_rol:1): lane l receives lane
// l-1 (ror) or lane l+1 (rol), modulo 64, so lane 0 receives lane 63 and
// lane 63 lane 0.
HYDRO_DEV float lane_rot_from_left(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x13C, 0xF, 0xF, true));
}
HYDRO_DEV float lane_rot_from_right(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF, true));
}
HYDRO_DEV double lane_rot_from_left(double v) {
  int2 p = __builtin_bit_cast(int2, v);
  p.x = __builtin_amdgcn_update_dpp(0, p.x, 0x13C, 0xF, 0xF, true);
  p.y = __builtin_amdgcn_update_dpp(0, p.y, 0x13C, 0xF, 0xF, true);
  return __builtin_bit_cast(double, p);
}
HYDRO_DEV double lane_rot_from_right(double v) {
  int2 p = __builtin_bit_cast(int2, v);
  p.x = __builtin_amdgcn_update_dpp(0, p.x, 0x134, 0xF, 0xF, true);
  p.y = __builtin_amdgcn_update_dpp(0, p.y, 0x134, 0xF, 0xF, true);
  return __builtin_bit_cast(double, p);
}

// ---------------------------------------------------------------- math
template <typename T> HYDRO_DEV T tmin(T a, T b);
template <typename T> HYDRO_DEV T tmax(T a, T b);
template <> HYDRO_DEV float tmin(float a, float b) { return fminf(a, b); }
template <> HYDRO_DEV float tmax(float a, float b) { return fmaxf(a, b); }
template <> HYDRO_DEV double tmin(double a, double b) { return fmin(a, b); }
template <> HYDRO_DEV double tmax(double a, double b) { return fmax(a, b); }
// FP32: the hardware square root (v_sqrt_f32, ~1 ulp) without the
// correctly-rounded refinement sequence; the FP32 path's tolerance covers it.
HYDRO_DEV float tsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
HYDRO_DEV double tsqrt(double a) { return __builtin_sqrt(a); }
HYDRO_DEV float tabs(float a) { return fabsf(a); }
HYDRO_DEV double tabs(double a) { return fabs(a); }

// Reciprocal used where the FP32-mixed path tolerates 1 ulp (limiter ratio,
// per-cell 1/rho): v_rcp_f32.  FP64 keeps the correctly rounded division.
HYDRO_DEV float trcp(float a) { return __builtin_amdgcn_rcpf(a); }
HYDRO_DEV double trcp(double a) { return 1.0 / a; }
HYDRO_DEV float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
HYDRO_DEV double tdiv(double a, double b) { return a / b; }

// ---------------------------------------------------------------- limiters
// res/SlopeLimiter.cl:5-68; OpenCL min/max on floating point are fmin/fmax
// (NaN-ignoring), which keeps phi finite for r = 0/0.
enum {
  LIM_DONORCELL = 0, LIM_LAXWENDROFF, LIM_BEAMWARMING, LIM_FROMM, LIM_CHARM, LIM_HCUS,
  LIM_HQUICK, LIM_KOREN, LIM_MINMOD, LIM_OSHKER, LIM_OSPRE, LIM_SMART, LIM_SWEBY, LIM_UMIST,
  LIM_VANALBADA1, LIM_VANALBADA2, LIM_VANLEER, LIM_MONOTIZEDCENTRAL, LIM_SUPERBEE,
  LIM_BARTHJESPERSEN, LIM_RUNTIME = -1
};

template <typename T>
HYDRO_DEV T limiter_eval(int lim, T r) {
  const T z = T(0), one = T(1), two = T(2);
  switch (lim) {
  case LIM_DONORCELL: return z;
  case LIM_LAXWENDROFF: return one;
  case LIM_BEAMWARMING: return r;
  case LIM_FROMM: return T(.5) * (one + r);
  case LIM_CHARM: return tmax(z, r) * (T(3) * r + one) / ((r + one) * (r + one));
  case LIM_HCUS: return tmax(z, T(1.5) * (r + tabs(r)) / (r + two));
  case LIM_HQUICK: return tmax(z, two * (r + tabs(r)) / (r + T(3)));
  case LIM_KOREN: return tmax(z, tmin(two * r, tmin((one + two * r) / T(3), two)));
  case LIM_MINMOD: return tmax(z, tmin(r, one));
  case LIM_OSHKER: return tmax(z, tmin(r, T(1.5)));
  case LIM_OSPRE: return T(.5) * (r * r + r) / (r * r + r + one);
  case LIM_SMART: return tmax(z, tmin(two * r, tmin(T(.25) + T(.75) * r, T(4))));
  case LIM_SWEBY: return tmax(z, tmax(tmin(T(1.5) * r, one), tmin(r, T(1.5))));
  case LIM_UMIST: return tmax(z, tmin(tmin(two * r, T(.75) + T(.25) * r), tmin(T(.25) + T(.75) * r, two)));
  case LIM_VANALBADA1: return (r * r + r) / (r * r + one);
  case LIM_VANALBADA2: return two * r / (r * r + one);
  case LIM_VANLEER: return tmax(z, r) * two / (one + r);
  case LIM_MONOTIZEDCENTRAL: return tmax(z, tmin(two, tmin(T(.5) * (one + r), two * r)));
  case LIM_SUPERBEE: return tmax(z, tmax(tmin(one, two * r), tmin(two, r)));
  case LIM_BARTHJESPERSEN: return T(.5) * (r + one) * tmin(one, tmin(T(4) * r / (r + one), T(4) / (r + one)));
  }
  return z;
}

// compile-time limiter when LIM >= 0, runtime switch otherwise
template <int LIM, typename T>
HYDRO_DEV T limiter(int lim_rt, T r) {
  if constexpr (LIM >= 0) return limiter_eval<T>(LIM, r);
  else return limiter_eval<T>(lim_rt, r);
}

// ---------------------------------------------------------------- parameters
// Runtime physical constants of a face solver (the reference bakes them into
// the program source as `defs`, src/Solver/Solver.cpp:289-307).
template <typename T>
struct PhysParams {
  T gamma;
  T mx[5];  // Maxwell: se, su, ise, isu, eigenvalue (res/MaxwellRoe.cl:30-35, 139)
};

// ---------------------------------------------------------------- indexing
struct Grid {
  int sx, sy, sz;       // sizes incl. ghosts
  long long plane;      // elements between state planes (SoA)
};

}  // namespace hydro
