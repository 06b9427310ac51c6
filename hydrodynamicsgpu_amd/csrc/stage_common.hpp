// stage_common.hpp -- the Shu-Osher combination shared by the fused stage
// kernels (include/HydroGPU/Integrator/RungeKutta.h:84-99 in increment form),
// and the addressing helpers they use.
#pragma once
#include "device_common.hpp"
#include "stage_args.hpp"

namespace hydro {

constexpr int kStripOut = 60;  // output columns per wave (64 lanes - 2*2 halo)

// Addressing: element (plane s, uniform offset e, column col) of a plane
// set.  base + s*plane + e is wave-uniform: it becomes a raw buffer resource
// in scalar registers (readfirstlane pins it there), and the column enters
// as a 32-bit byte offset, so each access is one buffer_load/store with no
// 64-bit vector address arithmetic.
constexpr int kBufferWord3 = 0x00020000;  // 32-bit data format, raw access (gfx9 family)

template <typename V>
HYDRO_DEV __amdgpu_buffer_rsrc_t row_rsrc(const V* p) {
  const uint64_t x = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  void* q = reinterpret_cast<void*>((uint64_t(hi) << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(q, 0, -1, kBufferWord3);
}
template <typename V>
HYDRO_DEV V ldp(const V* base, long long plane, int s, long long e, uint32_t col) {
  const __amdgpu_buffer_rsrc_t r = row_rsrc(base + (s * plane + e));
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, col * 4u, 0, 0));
  } else {
    static_assert(sizeof(V) == 8, "4- or 8-byte elements");
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, col * 8u, 0, 0));
  }
}
template <typename V>
HYDRO_DEV void stp(V* base, long long plane, int s, long long e, uint32_t col, V v) {
  const __amdgpu_buffer_rsrc_t r = row_rsrc(base + (s * plane + e));
  if constexpr (sizeof(V) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, col * 4u, 0, 0);
  } else {
    static_assert(sizeof(V) == 8, "4- or 8-byte elements");
    using U2 = __attribute__((ext_vector_type(2))) uint32_t;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, col * 8u, 0, 0);
  }
}

template <typename T, typename S>
HYDRO_DEV PhysParams<T> make_params(const StageArgs<T, S>& a) {
  PhysParams<T> p;
  p.gamma = T(a.gamma);
#pragma unroll
  for (int i = 0; i < 5; ++i) p.mx[i] = T(a.mx[i]);
  return p;
}

// MODE is a CombineMode.  Coefficients are converted to the arithmetic type
// once per wave; a cell is (wave-uniform element offset e, column col).
template <typename T, typename S, bool KAHAN, int N, int MODE>
struct StageCombine {
  static constexpr bool kReadsDk = MODE == CM_MID || MODE == CM_FINAL || MODE == CM_GENERIC || MODE == CM_DERIV;
  const StageArgs<T, S>& a;
  const long long pl;
  T bdt, aself;
  T rc[kMaxPartials], am[kMaxPartials], bm[kMaxPartials];

  HYDRO_DEV StageCombine(const StageArgs<T, S>& args, double dt) : a(args), pl(args.plane) {
    bdt = T(args.b_self * dt);
    aself = T(args.a_self);
#pragma unroll
    for (int j = 0; j < kMaxPartials; ++j) {
      rc[j] = T(args.rc[j]);
      am[j] = T(args.am[j]);
      bm[j] = T(args.bm[j] * dt);
    }
  }

  // Loads of one cell that the stage state needs: u0 (storage precision)
  // and d_k.  Issued ahead of use so their latency overlaps compute.
  HYDRO_DEV void load_u0(long long e, uint32_t col, S q0[N]) const {
#pragma unroll
    for (int s = 0; s < N; ++s) q0[s] = ldp(a.u0, pl, s, e, col);
  }
  HYDRO_DEV void load_dk(long long e, uint32_t col, T dk[N]) const {
    if constexpr (MODE == CM_MID || MODE == CM_FINAL || MODE == CM_DERIV) {
#pragma unroll
      for (int s = 0; s < N; ++s) dk[s] = ldp(a.dk, pl, s, e, col);
    } else if constexpr (MODE == CM_GENERIC) {
      if (a.dk) {
#pragma unroll
        for (int s = 0; s < N; ++s) dk[s] = ldp(a.dk, pl, s, e, col);
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) dk[s] = T(0);
      }
    }
  }
  // stage state u = u0 + d_k in arithmetic precision (CM_DERIV: d_k is the
  // stage state itself)
  HYDRO_DEV void stage_state(const S q0[N], const T dk[N], T u[N]) const {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      if constexpr (MODE == CM_DERIV) u[s] = dk[s];
      else if constexpr (kReadsDk) u[s] = T(q0[s]) + dk[s];
      else u[s] = T(q0[s]);
    }
  }

  // Operands of the combination of one cell, loaded ahead of the flux work.
  struct Fin {
    T dk[N];
    T rep[kMaxPartials][N];
    S hi[N];
    float lo[N];
    T pk[N];
    T fv[kMaxPartials][N];
  };

  // Operands of the combination of a cell whose u0 and d_k the caller
  // already holds (the row march keeps them): only the rest is loaded.
  HYDRO_DEV void fetch_fin_rest(long long e, uint32_t col, const S q0[N], const T dk[N], Fin& f) const {
    constexpr bool kFinal = MODE == CM_FINAL || MODE == CM_SINGLE;
#pragma unroll
    for (int s = 0; s < N; ++s) {
      if constexpr (kReadsDk) f.dk[s] = dk[s];
      f.hi[s] = q0[s];
    }
    if constexpr (MODE == CM_FINAL || MODE == CM_GENERIC) {
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nrep) {
#pragma unroll
          for (int s = 0; s < N; ++s) f.rep[j][s] = ldp(a.rep[j], pl, s, e, col);
        }
    }
    if constexpr (MODE == CM_GENERIC) {
#pragma unroll
      for (int s = 0; s < N; ++s) f.pk[s] = a.pkp ? ldp(a.pkp, pl, s, e, col) : T(0);
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nf) {
#pragma unroll
          for (int s = 0; s < N; ++s) f.fv[j][s] = a.fload[j] ? ldp((const T*)a.fp[j], pl, s, e, col) : T(0);
        }
    }
    if constexpr (KAHAN) {
      if (kFinal || (MODE == CM_GENERIC && a.final_stage)) {
#pragma unroll
        for (int s = 0; s < N; ++s) f.lo[s] = ldp(a.u0lo, pl, s, e, col);
      }
    }
  }

  HYDRO_DEV void fetch_fin(long long e, uint32_t col, Fin& f) const {
    constexpr bool kFinal = MODE == CM_FINAL || MODE == CM_SINGLE;
    if constexpr (kReadsDk) load_dk(e, col, f.dk);
    if constexpr (MODE == CM_FINAL || MODE == CM_GENERIC) {
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nrep) {
#pragma unroll
          for (int s = 0; s < N; ++s) f.rep[j][s] = ldp(a.rep[j], pl, s, e, col);
        }
    }
    if constexpr (MODE == CM_GENERIC) {
#pragma unroll
      for (int s = 0; s < N; ++s) f.pk[s] = a.pkp ? ldp(a.pkp, pl, s, e, col) : T(0);
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nf) {
#pragma unroll
          for (int s = 0; s < N; ++s) f.fv[j][s] = a.fload[j] ? ldp((const T*)a.fp[j], pl, s, e, col) : T(0);
        }
    }
    if (kFinal || (MODE == CM_GENERIC && a.final_stage)) {
#pragma unroll
      for (int s = 0; s < N; ++s) {
        f.hi[s] = ldp(a.u0, pl, s, e, col);
        if constexpr (KAHAN) f.lo[s] = ldp(a.u0lo, pl, s, e, col);
      }
    }
  }

  HYDRO_DEV void store_final(long long e, uint32_t col, const Fin& f, const T acc[N]) const {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      const S hi = f.hi[s];
      if constexpr (KAHAN) {
        // compensated update of (hi, lo) by acc (Fast2Sum: |hi| >= |lo + acc|)
        const float y = float(acc[s]) + f.lo[s];
        const float t = float(hi) + y;
        const float err = y - (t - float(hi));
        stp(a.unew, pl, s, e, col, S(t));
        stp(a.unewlo, pl, s, e, col, err);
      } else {
        stp(a.unew, pl, s, e, col, S(hi + S(acc[s])));
      }
    }
  }

  // combine the derivative of a cell into the increment / state; a stage
  // that is not the last also hands the increment it stores to dnext_out
  // when given
  HYDRO_DEV void finish(long long e, uint32_t col, const Fin& f, const T deriv[N], T* dnext_out = nullptr) const {
    if constexpr (MODE == CM_DERIV) {
      // calcFluxDeriv accumulates into the caller's derivative buffer
      // (res/CalcFluxDeriv.cl:7-44: deriv -= flux differences)
#pragma unroll
      for (int s = 0; s < N; ++s) stp(a.dnext, pl, s, e, col, ldp((const T*)a.dnext, pl, s, e, col) + deriv[s]);
    } else if constexpr (MODE == CM_FIRST) {
#pragma unroll
      for (int s = 0; s < N; ++s) {
        const T v = bdt * deriv[s];
        stp(a.dnext, pl, s, e, col, v);
        if (dnext_out) dnext_out[s] = v;
      }
    } else if constexpr (MODE == CM_MID) {
#pragma unroll
      for (int s = 0; s < N; ++s) {
        const T v = aself * f.dk[s] + bdt * deriv[s];
        stp(a.dnext, pl, s, e, col, v);
        if (dnext_out) dnext_out[s] = v;
      }
    } else if constexpr (MODE == CM_SINGLE) {
      T acc[N];
#pragma unroll
      for (int s = 0; s < N; ++s) acc[s] = bdt * deriv[s];
      store_final(e, col, f, acc);
    } else if constexpr (MODE == CM_FINAL) {
      T acc[N];
#pragma unroll
      for (int s = 0; s < N; ++s) acc[s] = aself * f.dk[s] + bdt * deriv[s];
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nrep) {
#pragma unroll
          for (int s = 0; s < N; ++s) acc[s] += rc[j] * f.rep[j][s];
        }
      store_final(e, col, f, acc);
    } else {
      // the fixed-function forms' order (d_k and L first), then the partial
      // sum: a stage gives the same bits whichever form runs it
      T acc[N];
#pragma unroll
      for (int s = 0; s < N; ++s) {
        acc[s] = aself * f.dk[s] + bdt * deriv[s];
        if (a.pkp) acc[s] += f.pk[s];
      }
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nrep) {
#pragma unroll
          for (int s = 0; s < N; ++s) acc[s] += rc[j] * f.rep[j][s];
        }
#pragma unroll
      for (int j = 0; j < kMaxPartials; ++j)
        if (j < a.nf) {
#pragma unroll
          for (int s = 0; s < N; ++s) stp(a.fp[j], pl, s, e, col, f.fv[j][s] + am[j] * f.dk[s] + bm[j] * deriv[s]);
        }
      if (a.final_stage) {
        store_final(e, col, f, acc);
      } else {
#pragma unroll
        for (int s = 0; s < N; ++s) {
          stp(a.dnext, pl, s, e, col, acc[s]);
          if (dnext_out) dnext_out[s] = acc[s];
        }
      }
    }
  }

  // load the operands, then combine (for callers that do not prefetch)
  HYDRO_DEV void finish(long long e, uint32_t col, const T deriv[N], T* dnext_out = nullptr) const {
    Fin f;
    fetch_fin(e, col, f);
    finish(e, col, f, deriv, dnext_out);
  }
};

}  // namespace hydro
