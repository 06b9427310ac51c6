// Files.cpp -- FITS images and checkpoint records (HydroGPU/Shared/Files.h).
#include "HydroGPU/Shared/Files.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>

#include "HydroGPU/Exception.h"

namespace HydroGPU {

bool fileExists(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  return (bool)f;
}

namespace {

std::string card(const std::string& key, const std::string& value, const std::string& comment = "") {
  char buf[96];
  std::snprintf(buf, sizeof buf, "%-8s= %20s", key.c_str(), value.c_str());
  std::string s(buf);
  if (!comment.empty()) s += " / " + comment;
  s.resize(80, ' ');
  return s;
}

template <typename T>
void put(std::string& out, T v) {
  char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));  // the formats are little-endian, as the host
  out.append(b, sizeof(T));
}

template <typename T>
T get(const std::string& in, size_t& pos) {
  if (pos + sizeof(T) > in.size()) throw Exception() << "checkpoint truncated";
  T v;
  std::memcpy(&v, in.data() + pos, sizeof(T));
  pos += sizeof(T);
  return v;
}

}  // namespace

void writeFits(const std::string& path, const std::vector<float>& data, int sx, int sy, int sz) {
  if ((long long)data.size() != (long long)sx * sy * sz) throw Exception() << "image size mismatch";
  std::string head;
  head += card("SIMPLE", "T", "conforms to FITS standard");
  head += card("BITPIX", "-32", "IEEE single precision");
  head += card("NAXIS", std::to_string(sz > 1 ? 3 : 2));
  head += card("NAXIS1", std::to_string(sx));
  head += card("NAXIS2", std::to_string(sy));
  if (sz > 1) head += card("NAXIS3", std::to_string(sz));
  head += card("EXTEND", "T");
  std::string end = "END";
  end.resize(80, ' ');
  head += end;
  head.resize((head.size() + 2879) / 2880 * 2880, ' ');
  std::string body(data.size() * 4, '\0');
  for (size_t i = 0; i < data.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &data[i], 4);
    body[4 * i + 0] = (char)(u >> 24);
    body[4 * i + 1] = (char)(u >> 16);
    body[4 * i + 2] = (char)(u >> 8);
    body[4 * i + 3] = (char)u;
  }
  body.resize((body.size() + 2879) / 2880 * 2880, '\0');
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Exception() << "cannot write " << path;
  f << head << body;
  if (!f) throw Exception() << "write failed: " << path;
}

void writeCheckpoint(const std::string& path, const Checkpoint& c) {
  const size_t n = (size_t)c.numStates * (size_t)c.volume;
  if (c.state.size() != n) throw Exception() << "checkpoint state size mismatch";
  if (!c.prims.empty() && c.prims.size() != n) throw Exception() << "checkpoint primitives size mismatch";
  if (!c.low.empty() && c.low.size() != n) throw Exception() << "checkpoint low words size mismatch";
  std::string out("HGPUCKP1", 8);
  put<uint32_t>(out, 1);
  put<uint32_t>(out, (uint32_t)c.numStates);
  put<uint64_t>(out, (uint64_t)c.volume);
  put<double>(out, c.t);
  put<int64_t>(out, (int64_t)c.steps);
  put<uint32_t>(out, (c.prims.empty() ? 0u : 1u) | (c.low.empty() ? 0u : 2u));
  put<uint32_t>(out, (uint32_t)sizeof(hydro_sim_desc));
  out.append(reinterpret_cast<const char*>(&c.desc), sizeof(hydro_sim_desc));
  out.append(reinterpret_cast<const char*>(c.state.data()), n * sizeof(double));
  if (!c.low.empty()) out.append(reinterpret_cast<const char*>(c.low.data()), n * sizeof(float));
  if (!c.prims.empty()) out.append(reinterpret_cast<const char*>(c.prims.data()), n * sizeof(double));
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Exception() << "cannot write " << path;
  f << out;
  if (!f) throw Exception() << "write failed: " << path;
}

Checkpoint readCheckpoint(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Exception() << "cannot read " << path;
  const std::string in((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (in.size() < 8 || in.compare(0, 8, "HGPUCKP1") != 0) throw Exception() << path << " is not a checkpoint";
  size_t pos = 8;
  Checkpoint c;
  if (get<uint32_t>(in, pos) != 1u) throw Exception() << "unknown checkpoint version";
  c.numStates = (int)get<uint32_t>(in, pos);
  c.volume = (long long)get<uint64_t>(in, pos);
  c.t = get<double>(in, pos);
  c.steps = (long long)get<int64_t>(in, pos);
  const uint32_t flags = get<uint32_t>(in, pos);
  const bool prims = (flags & 1u) != 0, low = (flags & 2u) != 0;
  if (get<uint32_t>(in, pos) != sizeof(hydro_sim_desc)) throw Exception() << "descriptor record size mismatch";
  if (pos + sizeof(hydro_sim_desc) > in.size()) throw Exception() << "checkpoint truncated";
  std::memcpy(&c.desc, in.data() + pos, sizeof(hydro_sim_desc));
  pos += sizeof(hydro_sim_desc);
  const size_t n = (size_t)c.numStates * (size_t)c.volume;
  if (in.size() != pos + n * (sizeof(double) * (prims ? 2 : 1) + (low ? sizeof(float) : 0)))
    throw Exception() << "checkpoint size mismatch";
  c.state.resize(n);
  std::memcpy(c.state.data(), in.data() + pos, n * sizeof(double));
  pos += n * sizeof(double);
  if (low) {
    c.low.resize(n);
    std::memcpy(c.low.data(), in.data() + pos, n * sizeof(float));
    pos += n * sizeof(float);
  }
  if (prims) {
    c.prims.resize(n);
    std::memcpy(c.prims.data(), in.data() + pos, n * sizeof(double));
  }
  return c;
}

}  // namespace HydroGPU
