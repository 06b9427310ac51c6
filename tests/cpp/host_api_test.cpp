// host_api_test.cpp -- the C++ host layer (HydroGPU::Solver & co.) driven
// the way the reference's HydroGPUApp drives its solvers, checked against the
// oracle (the FP64 CPU restatement, linked as test infrastructure).
//
//   host_api_test cpu   configuration errors that throw before any device work
//   host_api_test gpu   Sod (Euler Roe, TVD-RK2, Superbee, free-flow, CFL 0.5,
//                       to t = 0.2) and short runs of every solver class,
//                       each compared with the oracle on the same descriptor;
//                       checkpoint resume and Solver::save (in the cwd)
//   host_api_test fits <path> <sx> <sy> <sz>
//                       writes the image data[i] = .5 i - 3 (for the
//                       byte-for-byte comparison with io.py)
//   host_api_test ckpt-write <path> / ckpt-read <path>
//                       writes a synthetic checkpoint / prints one's summary
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "HydroGPU/Equation/Equation.h"
#include "HydroGPU/HydroGPUApp.h"
#include "HydroGPU/Shared/Files.h"
#include "HydroGPU/Solver/Solver.h"
#include "../../oracle/hydro_oracle.h"

using HydroGPU::real;

static int failures = 0;

#define CHECK(cond, ...)                    \
  do {                                      \
    if (!(cond)) {                          \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);             \
      std::printf("\n");                    \
      ++failures;                           \
    }                                       \
  } while (0)

template <typename F>
static bool throws(F f, const char* expect) {
  try {
    f();
  } catch (const HydroGPU::Exception& e) {
    if (std::strstr(e.what(), expect)) return true;
    std::printf("  threw '%s', expected '%s'\n", e.what(), expect);
    return false;
  }
  return false;
}

// buildStateEuler (res/util.lua:36-80): rho, rho v, B, E
static std::vector<real> euler_cell(real rho, real vx, real vy, real p, real gamma) {
  std::vector<real> c(10, 0.0);
  c[0] = rho;
  c[1] = rho * vx;
  c[2] = rho * vy;
  c[7] = rho * .5 * (vx * vx + vy * vy) + p / (gamma - 1.0);
  return c;
}

static real relerr(const std::vector<real>& a, const std::vector<real>& b) {
  real d = 0, m = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    d = std::fmax(d, std::fabs(a[i] - b[i]));
    m = std::fmax(m, std::fabs(b[i]));
  }
  return d / m;
}

static int run_cpu() {
  HydroGPU::HydroGPUApp app;
  app.solverName = "MHDBurgers";  // in the reference's registry, not in this build
  CHECK(throws([&] { app.createSolver(); }, "unknown solver MHDBurgers"), "unknown solver");
  app.solverName = "EulerRoe";
  app.integratorName = "BackwardEulerConjugateGradient";
  app.size = {{64, 1, 1, 1}};
  CHECK(throws([&] { app.createSolver(); }, "failed to find integrator named BackwardEulerConjugateGradient"),
        "unknown integrator");
  app.integratorName = "RungeKutta4";
  app.slopeLimiter = "NoSuchLimiter";
  CHECK(throws([&] { app.createSolver(); }, "unknown slope limiter NoSuchLimiter"), "unknown limiter");
  app.slopeLimiter = "Superbee";
  std::printf("%s cpu checks\n", failures ? "FAILED" : "passed");
  return failures ? 1 : 0;
}

// the state resetState() uploads: initState at every cell centre (ghosts
// included), through the equation's readStateCell (before any device-side
// conversion, which the oracle applies on its own for SRHD)
static std::vector<real> initial_state(HydroGPU::HydroGPUApp& app, HydroGPU::Solver::Solver& s) {
  const int n = s.numStates();
  std::vector<real> st((size_t)s.getVolume() * n);
  int flat = 0;
  for (int k = 0; k < app.size[2]; ++k)
    for (int j = 0; j < app.size[1]; ++j)
      for (int i = 0; i < app.size[0]; ++i, ++flat) {
        const int idx[3] = {i, j, k};
        real pos[3];
        for (int a = 0; a < 3; ++a)
          pos[a] = real(app.xmax[a] - app.xmin[a]) * (real(idx[a]) + .5) / real(app.size[a]) + real(app.xmin[a]);
        const std::vector<real> c = app.initState(pos[0], pos[1], pos[2]);
        s.getEquation()->readStateCell(st.data() + (size_t)flat * n, c.data());
      }
  return st;
}

// one solver run through the host classes, the same descriptor through the oracle
static void compare(HydroGPU::HydroGPUApp& app, int updates, real tol, const char* what) {
  auto solver = app.createSolver();
  oracle_sim* o = oracle_create(&solver->desc(), 0);
  std::vector<real> q0 = initial_state(app, *solver);
  oracle_set_state(o, q0.data(), q0.size());
  real t = 0;
  for (int i = 0; i < updates; ++i) {
    solver->update();
    t += oracle_update(o);
  }
  std::vector<real> ref(q0.size());
  oracle_get_state(o, ref.data(), ref.size());
  const real e = relerr(solver->getState(), ref);
  CHECK(e < tol, "%s: rel err %.3e >= %.1e", what, e, tol);
  CHECK(std::fabs(solver->getTime() - t) <= 1e-12 * t, "%s: time %.17g vs oracle %.17g", what, solver->getTime(), t);
  CHECK(solver->health() == 0, "%s: health %u", what, solver->health());
  std::printf("%-34s %4d updates  rel err %.2e  t %.6f\n", what, updates, e, solver->getTime());
  oracle_destroy(o);
}

static int run_gpu() {
  const real gamma = 1.4;
  {
    // Sod: the 1D CPU-runnable configuration
    HydroGPU::HydroGPUApp app;
    app.solverName = "EulerRoe";
    app.integratorName = "RungeKutta2TVD";
    app.slopeLimiter = "Superbee";
    app.dim = 1;
    app.size = {{1024, 1, 1, 1}};
    app.xmin[0] = -1;
    app.xmax[0] = 1;
    app.cfl = .5f;
    app.defs["idealGas_heatCapacityRatio"] = gamma;
    app.initState = [&](real x, real, real) {
      return x < 0 ? euler_cell(1.0, 0, 0, 1.0, gamma) : euler_cell(.125, 0, 0, .1, gamma);
    };
    auto solver = app.createSolver();
    oracle_sim* o = oracle_create(&solver->desc(), 0);
    std::vector<real> q0 = solver->getState();
    oracle_set_state(o, q0.data(), q0.size());
    int steps = 0;
    while (solver->getTime() < 0.2) {
      solver->update();
      oracle_update(o);
      ++steps;
    }
    std::vector<real> ref(q0.size());
    oracle_get_state(o, ref.data(), ref.size());
    const real e = relerr(solver->getState(), ref);
    CHECK(e < 1e-10, "Sod: rel err %.3e", e);
    CHECK(std::fabs(oracle_time(o) - solver->getTime()) < 1e-12, "Sod: time");
    std::printf("%-34s %4d updates  rel err %.2e  t %.6f\n", "EulerRoe Sod 1024 RK2TVD", steps, e, solver->getTime());
    // run(n) (no per-step synchronisation) continues where update() left off
    solver->run(5);
    for (int i = 0; i < 5; ++i) oracle_update(o);
    oracle_get_state(o, ref.data(), ref.size());
    CHECK(relerr(solver->getState(), ref) < 1e-10, "Sod run(5)");
    CHECK(solver->getSteps() == steps + 5, "step count %lld", solver->getSteps());
    oracle_destroy(o);
    // checkpoint: save, run on, reload, run the same steps again -> the same
    // bits and clock
    solver->saveCheckpoint("sod.ckpt");
    solver->run(7);
    const std::vector<real> a = solver->getState();
    const real ta = solver->getTime();
    solver->loadCheckpoint("sod.ckpt");
    CHECK(solver->getSteps() == steps + 5, "restored step count %lld", solver->getSteps());
    solver->run(7);
    CHECK(solver->getState() == a, "checkpoint resume differs");
    CHECK(solver->getTime() == ta, "checkpoint resume time");
    // Solver::save: one FITS image per state channel at the first free index
    solver->save();
    solver->save();
    for (const char* f : {"DENSITY0.fits", "MOMENTUM_X0.fits", "ENERGY_TOTAL0.fits", "DENSITY1.fits"})
      CHECK(HydroGPU::fileExists(f), "missing %s", f);
    CHECK(!HydroGPU::fileExists("MOMENTUM_Y0.fits"), "1D grid saved a y channel");
  }
  auto kh = [&](real x, real y, real) {
    const bool in = std::fabs(y) < .25;
    return euler_cell(in ? 2.0 : 1.0, in ? .5 : -.5, .01 * std::sin(4 * M_PI * x), 2.5, gamma);
  };
  for (const char* name : {"EulerRoe", "EulerHLL", "EulerHLLC", "EulerBurgers"}) {
    HydroGPU::HydroGPUApp app;
    app.solverName = name;
    app.integratorName = "RungeKutta3TVD";
    app.slopeLimiter = "MonotizedCentral";
    app.dim = 2;
    app.size = {{72, 40, 1, 1}};
    for (int a = 0; a < 2; ++a)
      for (int m = 0; m < 2; ++m) app.boundaryMethods[a][m] = HydroGPU::BOUNDARY_METHOD_PERIODIC;
    app.initState = kh;
    compare(app, 6, 1e-10, (std::string(name) + " KH 72x40 RK3TVD").c_str());
  }
  {
    HydroGPU::HydroGPUApp app;
    app.solverName = "SRHDRoe";
    app.integratorName = "RungeKutta2";
    app.dim = 2;
    app.size = {{40, 36, 1, 1}};
    app.xmin = {{0, 0, 0, 0}};
    app.xmax = {{1, 1, 1, 0}};
    app.defs["idealGas_heatCapacityRatio"] = 5. / 3.;
    app.initState = [&](real x, real y, real) {
      const bool r = x > .5, u = y > .5;
      return euler_cell(r == u ? 1.0 : .5, r ? .3 : -.2, u ? .1 : -.3, r ? 1.0 : .4, 5. / 3.);
    };
    compare(app, 4, 1e-9, "SRHDRoe quadrants 40x36 RK2");
  }
  {
    HydroGPU::HydroGPUApp app;
    app.solverName = "MaxwellRoe";
    app.integratorName = "RungeKutta4";
    app.dim = 1;
    app.size = {{128, 1, 1, 1}};
    for (int m = 0; m < 2; ++m) app.boundaryMethods[0][m] = HydroGPU::BOUNDARY_METHOD_PERIODIC;
    app.initState = [&](real x, real, real) {
      std::vector<real> c(6, 0.0);
      c[1] = std::sin(2 * M_PI * x);  // E_y
      c[5] = std::sin(2 * M_PI * x);  // B_z
      return c;
    };
    compare(app, 5, 1e-10, "MaxwellRoe wave 128 RK4");
  }
  // the potential and solid channels of initState (the last two of ten,
  // include/HydroGPU/Solver/SelfGravitationBehavior.h:64-74) reach the
  // simulation through resetState: a solid disk and a linear potential
  {
    HydroGPU::HydroGPUApp app;
    app.solverName = "EulerRoe";
    app.integratorName = "RungeKutta2";
    app.dim = 2;
    app.size = {{48, 40, 1, 1}};
    const real gamma = 1.4;
    auto cyl = [&](real x, real y, real) {
      std::vector<real> c = euler_cell(1.0, .2, 0.0, 1.0, gamma);
      c[8] = .1 * y;
      c[9] = x * x + y * y < .04 ? 1.0 : 0.0;
      return c;
    };
    app.initState = cyl;
    auto solver = app.createSolver();  // createSolver runs resetState
    oracle_sim* o = oracle_create(&solver->desc(), 0);
    std::vector<real> q0 = initial_state(app, *solver);
    std::vector<real> pot((size_t)solver->getVolume());
    std::vector<unsigned char> solid(pot.size());
    int flat = 0;
    for (int j = 0; j < app.size[1]; ++j)
      for (int i = 0; i < app.size[0]; ++i, ++flat) {
        const real x = real(app.xmax[0] - app.xmin[0]) * (i + .5) / app.size[0] + app.xmin[0];
        const real y = real(app.xmax[1] - app.xmin[1]) * (j + .5) / app.size[1] + app.xmin[1];
        const std::vector<real> c = cyl(x, y, 0);
        pot[flat] = c[8];
        solid[flat] = c[9] != 0;
      }
    oracle_set_state(o, q0.data(), q0.size());
    oracle_set_potential(o, pot.data(), pot.size());
    oracle_set_solid(o, solid.data(), solid.size());
    for (int i = 0; i < 4; ++i) {
      solver->update();
      oracle_update(o);
    }
    std::vector<real> ref(q0.size());
    oracle_get_state(o, ref.data(), ref.size());
    const real e = relerr(solver->getState(), ref);
    CHECK(e < 1e-10, "potential + solid channels: rel err %.3e", e);
    std::printf("%-34s %4d updates  rel err %.2e\n", "EulerRoe potential + solid 48x40", 4, e);
    oracle_destroy(o);
  }
  // the reference's host-side form, Integrator::integrate(dt, callback) with
  // Solver::calcDeriv as the callback (Roe::step, src/Solver/Roe.cpp:94-99),
  // for a Runge-Kutta tableau and for ForwardEuler
  for (const char* integ : {"RungeKutta4", "ForwardEuler"}) {
    HydroGPU::HydroGPUApp app;
    app.solverName = "EulerRoe";
    app.integratorName = integ;
    app.slopeLimiter = "MonotizedCentral";
    app.dim = 2;
    app.size = {{48, 40, 1, 1}};
    for (int a = 0; a < 2; ++a)
      for (int m = 0; m < 2; ++m) app.boundaryMethods[a][m] = HydroGPU::BOUNDARY_METHOD_PERIODIC;
    app.initState = kh;
    auto solver = app.createSolver();
    oracle_sim* o = oracle_create(&solver->desc(), 0);
    std::vector<real> q0 = initial_state(app, *solver);
    oracle_set_state(o, q0.data(), q0.size());
    for (int i = 0; i < 3; ++i) {
      solver->boundary();
      double dt = 0;
      HydroGPU::checkStatus(hydro_sim_calc_timestep(solver->handle(), &dt), "calc_timestep");
      solver->integrator->integrate(dt, [&](void* deriv) { solver->calcDeriv(deriv, dt); });
      const double dto = oracle_update(o);
      CHECK(std::fabs(dt - dto) <= 1e-12 * dto, "%s callback form: dt %.17g vs %.17g", integ, dt, dto);
    }
    // the oracle's ForwardEuler ends with the zero-gravity pass's boundary
    solver->boundary();
    oracle_boundary(o);
    std::vector<real> ref(q0.size());
    oracle_get_state(o, ref.data(), ref.size());
    const real e = relerr(solver->getState(), ref);
    CHECK(e < 1e-10, "%s callback form: rel err %.3e", integ, e);
    CHECK(solver->getSteps() == 3, "%s callback form: steps %lld", integ, solver->getSteps());
    std::printf("%-34s %4d updates  rel err %.2e\n", (std::string(integ) + " integrate(dt, callback)").c_str(), 3, e);
    oracle_destroy(o);
  }
  std::printf("%s gpu checks\n", failures ? "FAILED" : "passed");
  return failures ? 1 : 0;
}

static int run_files(int argc, char** argv) {
  const std::string mode = argv[1];
  if (mode == "fits" && argc == 6) {
    const int sx = std::atoi(argv[3]), sy = std::atoi(argv[4]), sz = std::atoi(argv[5]);
    std::vector<float> img((size_t)sx * sy * sz);
    for (size_t i = 0; i < img.size(); ++i) img[i] = .5f * (float)i - 3.f;
    HydroGPU::writeFits(argv[2], img, sx, sy, sz);
    return 0;
  }
  if (mode == "ckpt-write" && argc == 3) {
    HydroGPU::Checkpoint c;
    hydro_sim_desc_default(&c.desc);
    c.desc.dim = 2;
    c.desc.size[0] = 12;
    c.desc.size[1] = 10;
    c.desc.equation = HYDRO_EQN_SRHD;
    c.numStates = 4;
    c.volume = 120;
    c.t = 0.125;
    c.steps = 7;
    c.state.resize(480);
    c.prims.resize(480);
    c.low.resize(480);
    for (size_t i = 0; i < 480; ++i) {
      c.state[i] = 1.0 + 1e-3 * (double)i;
      c.prims[i] = -2.0 + 1e-2 * (double)i;
      c.low[i] = 1e-9f * (float)i;
    }
    HydroGPU::writeCheckpoint(argv[2], c);
    return 0;
  }
  if (mode == "ckpt-read" && argc == 3) {
    const HydroGPU::Checkpoint c = HydroGPU::readCheckpoint(argv[2]);
    double s = 0, p = 0, l = 0;
    for (double v : c.state) s += v;
    for (double v : c.prims) p += v;
    for (float v : c.low) l += v;
    std::printf("dim %d size %d %d %d equation %d states %d volume %lld t %.17g steps %lld sum %.17g prims %zu %.17g "
                "low %zu %.17g\n",
                c.desc.dim, c.desc.size[0], c.desc.size[1], c.desc.size[2], c.desc.equation, c.numStates, c.volume,
                c.t, c.steps, s, c.prims.size(), p, c.low.size(), l);
    return 0;
  }
  std::printf("usage: host_api_test cpu|gpu|fits <path> <sx> <sy> <sz>|ckpt-write <path>|ckpt-read <path>\n");
  return 2;
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  try {
    if (mode == "gpu") return run_gpu();
    if (mode != "cpu") return run_files(argc, argv);
    return run_cpu();
  } catch (const std::exception& e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 1;
  }
}
