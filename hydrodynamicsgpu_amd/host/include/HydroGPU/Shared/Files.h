// HydroGPU/Shared/Files.h -- the on-disk forms of a grid: FITS images as
// Solver::save writes them (src/Solver/Solver.cpp:547-589) and this build's
// checkpoint record (resume).  Both layouts are shared byte for byte with
// hydrodynamicsgpu_amd/io.py.
#pragma once
#include <string>
#include <vector>

#include "hydro_capi.h"

namespace HydroGPU {

// A single primary-HDU FITS image: BITPIX -32 (big-endian IEEE floats),
// NAXIS1 = sx (fastest), NAXIS2 = sy, and NAXIS3 = sz when sz > 1.
void writeFits(const std::string& path, const std::vector<float>& data, int sx, int sy, int sz);

// Checkpoint: magic "HGPUCKP1", u32 version 1, u32 N, u64 volume, f64 t,
// i64 steps, u32 flags (bit 0 primitives, bit 1 low words),
// u32 sizeof(hydro_sim_desc), the descriptor record, f64 state[N*V]
// (reference AoS layout; the high words of compensated FP32 storage),
// f32 low[N*V] (flag bit 1), f64 prims[N*V] (flag bit 0); little-endian.
struct Checkpoint {
  hydro_sim_desc desc{};
  int numStates = 0;
  long long volume = 0;
  double t = 0;
  long long steps = 0;
  std::vector<double> state;
  std::vector<float> low;     // empty unless compensated FP32 storage
  std::vector<double> prims;  // empty unless SRHD
};
void writeCheckpoint(const std::string& path, const Checkpoint& c);
Checkpoint readCheckpoint(const std::string& path);

bool fileExists(const std::string& path);

}  // namespace HydroGPU
