/*
 * hydro_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C FP64 restatement of the reference's finite-volume update path:
 * the OpenCL kernels of res/ *.cl and the host dispatch order of
 * src/Solver/ *.cpp + include/HydroGPU/Integrator/RungeKutta.h.  Buffers keep the
 * reference layouts (state[s + N*cell], per-interface data at [k + K*(side +
 * DIM*cell)]) and every kernel loops over the full NDRange with the kernel's own
 * bounds test, so the never-written interface slots, the frozen eigenbasis and
 * the once-per-step boundary behave as in the reference.  Never-written slots
 * start at 0.0 (the reference never clears them, src/Solver/Roe.cpp:11-16).
 *
 * Built with -ffp-contract=off: every a*b+c is two roundings, in the order the
 * OpenCL source writes it.  OpenCL min/max on doubles are taken as fmin/fmax
 * (NaN-ignoring), which is what the limiter relies on for r = 0/0.
 *
 * Solid cells (solidBuffer, src/Solver/SelfGravitation.cpp:17, filled from
 * the initState's last channel, include/HydroGPU/Solver/SelfGravitationBehavior.h:
 * 70-74) enter only the two solvers compiled with SOLID: EulerRoe
 * (src/Solver/EulerRoe.cpp:11-26) and EulerBurgers (src/Solver/EulerBurgers.cpp:
 * 29-58).  The HLL family and the other equations never read the buffer.
 */
#include "hydro_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct oracle_sim {
  hydro_sim_desc d;
  unsigned opts;
  int dim, N, sz[3], step[3], vol;
  int ev_size;          /* EIGEN_TRANSFORM_STRUCT_SIZE */
  double dx[3];         /* host dx (src/HydroGPUApp.cpp:224-226) */
  double kdx[3];        /* kernel DX/DY/DZ macro values */
  double cfl;           /* float-rounded (HydroGPUApp.h:84) */
  double gamma;
  double* state;
  double* eigval;       /* Roe: N per interface; HLL: N per interface */
  double* eigvec;       /* Roe: ev_size per interface */
  double* dqt;          /* deltaQTilde */
  double* flux;
  double* dtbuf;
  double* potential;
  double* pot_next;     /* the relaxation's second buffer */
  int pot_on;           /* a non-zero potential or use_gravity: the gravity pass runs */
  double* prim;         /* SRHD primitives */
  double* ifvel;        /* EulerBurgers: interface velocities, one per side and cell */
  double* pressure;     /* EulerBurgers: computePressure's buffer */
  unsigned char* solid; /* solidBuffer: 1 for a solid cell (NULL: none) */
  double* rk_state[4];
  double* rk_deriv[4];
  double* fe_deriv;
  int order;
  double alpha[4][4], beta[4][4];
  double t;
  long long steps;
};

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------- limiter */

/* res/SlopeLimiter.cl:5-68.  min/max are fmin/fmax. */
double oracle_slope_limiter(int lim, double r) {
  switch (lim) {
  case HYDRO_LIMITER_DONORCELL: return 0.;
  case HYDRO_LIMITER_LAXWENDROFF: return 1.;
  case HYDRO_LIMITER_BEAMWARMING: return r;
  case HYDRO_LIMITER_FROMM: return .5 * (1. + r);
  case HYDRO_LIMITER_CHARM: return fmax(0., r) * (3. * r + 1.) / ((r + 1.) * (r + 1.));
  case HYDRO_LIMITER_HCUS: return fmax(0., 1.5 * (r + fabs(r)) / (r + 2.));
  case HYDRO_LIMITER_HQUICK: return fmax(0., 2. * (r + fabs(r)) / (r + 3.));
  case HYDRO_LIMITER_KOREN: return fmax(0., fmin(2. * r, fmin((1. + 2. * r) / 3., 2.)));
  case HYDRO_LIMITER_MINMOD: return fmax(0., fmin(r, 1.));
  case HYDRO_LIMITER_OSHKER: return fmax(0., fmin(r, 1.5));
  case HYDRO_LIMITER_OSPRE: return .5 * (r * r + r) / (r * r + r + 1.);
  case HYDRO_LIMITER_SMART: return fmax(0., fmin(2. * r, fmin(.25 + .75 * r, 4.)));
  case HYDRO_LIMITER_SWEBY: return fmax(0., fmax(fmin(1.5 * r, 1.), fmin(r, 1.5)));
  case HYDRO_LIMITER_UMIST: return fmax(0., fmin(fmin(2. * r, .75 + .25 * r), fmin(.25 + .75 * r, 2.)));
  case HYDRO_LIMITER_VANALBADA1: return (r * r + r) / (r * r + 1.);
  case HYDRO_LIMITER_VANALBADA2: return 2. * r / (r * r + 1.);
  case HYDRO_LIMITER_VANLEER: return fmax(0., r) * 2. / (1. + r);
  case HYDRO_LIMITER_MONOTIZEDCENTRAL: return fmax(0., fmin(2., fmin(.5 * (1. + r), 2. * r)));
  case HYDRO_LIMITER_SUPERBEE: return fmax(0., fmax(fmin(1., 2. * r), fmin(2., r)));
  case HYDRO_LIMITER_BARTHJESPERSEN: return .5 * (r + 1.) * fmin(1., fmin(4. * r / (r + 1.), 4. / (r + 1.)));
  }
  return 0.;
}

/* include/HydroGPU/toNumericString.h:12-22: ostream default precision (6
 * significant digits), read back by the OpenCL compiler as a double literal. */
double oracle_kernel_dx(double dx, int six_digits) {
  if (!six_digits) return dx;
  char buf[64];
  snprintf(buf, sizeof buf, "%.6g", dx);
  return strtod(buf, NULL);
}

/* ---------------------------------------------------------------- tableaus */

/* include/HydroGPU/Integrator/RungeKutta.h:129-294 (alphas(i,j), betas(i,j)). */
int oracle_rk_tableau(int integ, int* order, double* A, double* B) {
  memset(A, 0, 16 * sizeof(double));
  memset(B, 0, 16 * sizeof(double));
#define SETA(i, j, v) A[(i) * 4 + (j)] = (v)
#define SETB(i, j, v) B[(i) * 4 + (j)] = (v)
  switch (integ) {
  case HYDRO_INTEGRATOR_RK2: /* :129-139 */
    *order = 2; SETA(0, 0, 1); SETA(1, 0, 1);
    SETB(0, 0, .5); SETB(1, 1, 1); return 0;
  case HYDRO_INTEGRATOR_RK2HEUN: /* :142-152 */
    *order = 2; SETA(0, 0, 1); SETA(1, 0, 1);
    SETB(0, 0, 1); SETB(1, 0, .5); SETB(1, 1, .5); return 0;
  case HYDRO_INTEGRATOR_RK2RALSTON: /* :155-165 */
    *order = 2; SETA(0, 0, 1); SETA(1, 0, 1);
    SETB(0, 0, 2. / 3.); SETB(1, 0, 1. / 4.); SETB(1, 1, 3. / 4.); return 0;
  case HYDRO_INTEGRATOR_RK3: /* :168-180 */
    *order = 3; SETA(0, 0, 1); SETA(1, 0, 1); SETA(2, 0, 1);
    SETB(0, 0, .5); SETB(1, 0, -1); SETB(1, 1, 2);
    SETB(2, 0, 1. / 6.); SETB(2, 1, 2. / 6.); SETB(2, 2, 1. / 6.); return 0;
  case HYDRO_INTEGRATOR_RK4: /* :183-197 */
    *order = 4; SETA(0, 0, 1); SETA(1, 0, 1); SETA(2, 0, 1); SETA(3, 0, 1);
    SETB(0, 0, .5); SETB(1, 1, .5); SETB(2, 2, 1);
    SETB(3, 0, 1. / 6.); SETB(3, 1, 2. / 6.); SETB(3, 2, 2. / 6.); SETB(3, 3, 1. / 6.); return 0;
  case HYDRO_INTEGRATOR_RK4_3_8THSRULE: /* :200-214 */
    *order = 4; SETA(0, 0, 1); SETA(1, 0, 1); SETA(2, 0, 1); SETA(3, 0, 1);
    SETB(0, 0, 1. / 3.); SETB(1, 0, -1. / 3.); SETB(2, 0, 1); SETB(2, 1, -1); SETB(2, 2, 1);
    SETB(3, 0, 1. / 8.); SETB(3, 1, 3. / 8.); SETB(3, 2, 3. / 8.); SETB(3, 3, 1. / 8.); return 0;
  case HYDRO_INTEGRATOR_RK2TVD: /* :219-229 */
    *order = 2; SETA(0, 0, 1); SETA(1, 0, .5); SETA(1, 1, .5);
    SETB(0, 0, 1); SETB(1, 1, .5); return 0;
  case HYDRO_INTEGRATOR_RK2NONTVD: /* :232-242 */
    *order = 2; SETA(0, 0, 1); SETA(1, 0, 1);
    SETB(0, 0, -20); SETB(1, 0, 41. / 40.); SETB(1, 1, -1. / 40.); return 0;
  case HYDRO_INTEGRATOR_RK3TVD: /* :245-257 (row 2 combines u(1), as written) */
    *order = 3; SETA(0, 0, 1); SETA(1, 0, 3. / 4.); SETA(1, 1, 1. / 4.);
    SETA(2, 0, 1. / 3.); SETA(2, 1, 2. / 3.);
    SETB(0, 0, 1); SETB(1, 1, 1. / 4.); SETB(2, 2, 2. / 3.); return 0;
  case HYDRO_INTEGRATOR_RK4TVD: /* :260-274 */
    *order = 4; SETA(0, 0, 1);
    SETA(1, 0, 649. / 1600.); SETA(1, 1, 951. / 1600.);
    SETA(2, 0, 53989. / 2500000.); SETA(2, 1, 4806213. / 20000000.); SETA(2, 2, 23619. / 32000.);
    SETA(3, 0, 1. / 5.); SETA(3, 1, 6127. / 30000.); SETA(3, 2, 7873. / 30000.); SETA(3, 3, 1. / 3.);
    SETB(0, 0, .5);
    SETB(1, 0, -10890423. / 25193600.); SETB(1, 1, 5000. / 7873);
    SETB(2, 0, -102261. / 5000000.); SETB(2, 1, -5121. / 20000.); SETB(2, 2, 7873. / 10000.);
    SETB(3, 0, 1. / 10.); SETB(3, 1, 1. / 6.); SETB(3, 3, 1. / 6.); return 0;
  case HYDRO_INTEGRATOR_RK4NONTVD: /* :279-293 */
    *order = 4; SETA(0, 0, 1); SETA(1, 0, 1); SETA(2, 0, 1);
    SETA(3, 0, -1. / 3.); SETA(3, 1, 1. / 3.); SETA(3, 2, 2. / 3.); SETA(3, 3, 1. / 3.);
    SETB(0, 0, .5); SETB(1, 1, .5); SETB(2, 2, 1); SETB(3, 3, 1. / 6.); return 0;
  case HYDRO_INTEGRATOR_FORWARDEULER:
    *order = 1; SETA(0, 0, 1); SETB(0, 0, 1); return 0;
  }
#undef SETA
#undef SETB
  return -1;
}

/* ---------------------------------------------------------------- helpers */

static inline int idx3(const oracle_sim* s, int x, int y, int z) {
  return x + s->sz[0] * (y + s->sz[1] * z);
}

/* the "2 <= i < SIZE-e" active test of the kernels, on every axis of the grid */
static inline int active(const oracle_sim* s, const int i[3], int e) {
  for (int a = 0; a < s->dim; ++a)
    if (i[a] < 2 || i[a] >= s->sz[a] - e) return 0;
  return 1;
}

static inline void decode(const oracle_sim* s, int cell, int i[3]) {
  i[0] = cell % s->sz[0];
  i[1] = (cell / s->sz[0]) % s->sz[1];
  i[2] = cell / (s->sz[0] * s->sz[1]);
}

static double dot4(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

/* ------------------------------------------------ Euler Roe eigen-basis */

/* res/EulerRoe.cl:18-259 for one interface; VELOCITY() from
 * res/include/HydroGPU/Euler.h:6-12 (momentum / density per component). */
void oracle_euler_eigenbasis(int dim, int side, double gamma, const double* qL, const double* qR,
                             double potL, double potR, double* eigval, double* evl, double* evr) {
  oracle_euler_eigenbasis_solid(dim, side, gamma, qL, qR, potL, potR, 0, 0, eigval, evl, evr);
}

/* The same with the cells' solid flags (res/EulerRoe.cl:48-65): a solid cell
 * facing a fluid one takes the fluid cell's density and velocity with
 * velocity component side+1 of the 4-vector (vx, vy, vz, 0) negated -- the
 * tangential y for x-faces, z for y-faces, the padding for z-faces -- while
 * its total energy and potential stay its own. */
void oracle_euler_eigenbasis_solid(int dim, int side, double gamma, const double* qL, const double* qR,
                                   double potL, double potR, int solidL, int solidR, double* eigval,
                                   double* evl, double* evr) {
  const int N = dim + 2;
  const int E = dim + 1;
  double densityL = qL[0], densityR = qR[0];
  double velocityL[4] = {0, 0, 0, 0}, velocityR[4] = {0, 0, 0, 0};
  for (int a = 0; a < dim; ++a) {
    velocityL[a] = qL[1 + a] / qL[0];
    velocityR[a] = qR[1 + a] / qR[0];
  }
  if (solidL && !solidR) {
    densityL = densityR;
    for (int a = 0; a < 4; ++a) velocityL[a] = velocityR[a];
    velocityL[side + 1] = -velocityR[side + 1];
  }
  if (solidR && !solidL) {
    densityR = densityL;
    for (int a = 0; a < 4; ++a) velocityR[a] = velocityL[a];
    velocityR[side + 1] = -velocityL[side + 1];
  }
  double invDensityL = 1. / densityL;
  double energyTotalL = qL[E] * invDensityL;
  double energyKineticL = .5 * dot4(velocityL, velocityL);
  double energyInternalL = energyTotalL - energyKineticL - potL;
  double pressureL = (gamma - 1.) * densityL * energyInternalL;
  double enthalpyTotalL = energyTotalL + pressureL * invDensityL;
  double roeWeightL = sqrt(densityL);

  double invDensityR = 1. / densityR;
  double energyTotalR = qR[E] * invDensityR;
  double energyKineticR = .5 * dot4(velocityR, velocityR);
  double energyInternalR = energyTotalR - energyKineticR - potR;
  double pressureR = (gamma - 1.) * densityR * energyInternalR;
  double enthalpyTotalR = energyTotalR + pressureR * invDensityR;
  double roeWeightR = sqrt(densityR);

  double roeWeightNormalization = 1. / (roeWeightL + roeWeightR);
  double velocity[4];
  for (int a = 0; a < 4; ++a)
    velocity[a] = (roeWeightL * velocityL[a] + roeWeightR * velocityR[a]) * roeWeightNormalization;
  double enthalpyTotal = (roeWeightL * enthalpyTotalL + roeWeightR * enthalpyTotalR) * roeWeightNormalization;
  double energyPotential = (roeWeightL * potL + roeWeightR * potR) * roeWeightNormalization;
  double velocitySq = dot4(velocity, velocity);
  double speedOfSound = sqrt((enthalpyTotal - .5 * velocitySq - energyPotential) * (gamma - 1.));

  /* rotate the normal into x (EulerRoe.cl:101-110) */
  if (side == 1) { double t = velocity[0]; velocity[0] = velocity[1]; velocity[1] = t; }
  else if (side == 2) { double t = velocity[0]; velocity[0] = velocity[2]; velocity[2] = t; }
  const double vx = velocity[0], vy = velocity[1], vz = velocity[2];

  eigval[0] = vx - speedOfSound;
  for (int a = 1; a <= dim; ++a) eigval[a] = vx;
  eigval[E] = vx + speedOfSound;

#define R(i, j) evr[(i) + N * (j)]
#define L(i, j) evl[(i) + N * (j)]
  memset(evr, 0, sizeof(double) * N * N);
  memset(evl, 0, sizeof(double) * N * N);
  R(0, 0) = 1.; R(1, 0) = vx - speedOfSound;
  if (dim > 1) R(2, 0) = vy;
  if (dim > 2) R(3, 0) = vz;
  R(E, 0) = enthalpyTotal - speedOfSound * vx;
  R(0, 1) = 1.; R(1, 1) = vx;
  if (dim > 1) R(2, 1) = vy;
  if (dim > 2) R(3, 1) = vz;
  R(E, 1) = .5 * velocitySq;
  if (dim > 1) {
    R(0, 2) = 0.; R(1, 2) = 0.; R(2, 2) = 1.;
    if (dim > 2) R(3, 2) = 0.;
    R(E, 2) = vy;
  }
  if (dim > 2) {
    R(0, 3) = 0.; R(1, 3) = 0.; R(2, 3) = 0.; R(3, 3) = 1.; R(E, 3) = vz;
  }
  R(0, E) = 1.; R(1, E) = vx + speedOfSound;
  if (dim > 1) R(2, E) = vy;
  if (dim > 2) R(3, E) = vz;
  R(E, E) = enthalpyTotal + speedOfSound * vx;

  double invDenom = .5 / (speedOfSound * speedOfSound);
  L(0, 0) = (.5 * (gamma - 1.) * velocitySq + speedOfSound * vx) * invDenom;
  L(0, 1) = -(speedOfSound + (gamma - 1.) * vx) * invDenom;
  if (dim > 1) L(0, 2) = -(gamma - 1.) * vy * invDenom;
  if (dim > 2) L(0, 3) = -(gamma - 1.) * vz * invDenom;
  L(0, E) = (gamma - 1.) * invDenom;
  L(1, 0) = 1. - (gamma - 1.) * velocitySq * invDenom;
  L(1, 1) = (gamma - 1.) * vx * 2. * invDenom;
  if (dim > 1) L(1, 2) = (gamma - 1.) * vy * 2. * invDenom;
  if (dim > 2) L(1, 3) = (gamma - 1.) * vz * 2. * invDenom;
  L(1, E) = -(gamma - 1.) * 2. * invDenom;
  if (dim > 1) {
    L(2, 0) = -vy; L(2, 1) = 0.; L(2, 2) = 1.;
    if (dim > 2) L(2, 3) = 0.;
    L(2, E) = 0.;
  }
  if (dim > 2) {
    L(3, 0) = -vz; L(3, 1) = 0.; L(3, 2) = 0.; L(3, 3) = 1.; L(3, E) = 0.;
  }
  L(E, 0) = (.5 * (gamma - 1.) * velocitySq - speedOfSound * vx) * invDenom;
  L(E, 1) = (speedOfSound - (gamma - 1.) * vx) * invDenom;
  if (dim > 1) L(E, 2) = -(gamma - 1.) * vy * invDenom;
  if (dim > 2) L(E, 3) = -(gamma - 1.) * vz * invDenom;
  L(E, E) = (gamma - 1.) * invDenom;

  /* un-rotate: swap the momentum columns of L and rows of R (EulerRoe.cl:228-255) */
  if (side >= 1) {
    int a = 1, b = 1 + side;
    for (int i = 0; i < N; ++i) {
      double t = L(i, a); L(i, a) = L(i, b); L(i, b) = t;
      t = R(a, i); R(a, i) = R(b, i); R(b, i) = t;
    }
  }
#undef R
#undef L
}

/* ------------------------------------------------ SRHD eigen-basis */

/* res/SRHDRoe.cl:21-301 for one interface (arithmetic-mean primitives). */
void oracle_srhd_eigenbasis(int dim, int side, double gamma, const double* primL, const double* primR,
                            double* eigval, double* evl, double* evr) {
  const int N = dim + 2;
  const int E = dim + 1;
  double rhoL = primL[0], rhoR = primR[0];
  double vL[4] = {0, 0, 0, 0}, vR[4] = {0, 0, 0, 0};
  for (int a = 0; a < dim; ++a) { vL[a] = primL[1 + a]; vR[a] = primR[1 + a]; }
  double eIntL = primL[E], eIntR = primR[E];

  double rho = .5 * (rhoL + rhoR);
  double eInt = .5 * (eIntL + eIntR);
  double v[4];
  for (int a = 0; a < 4; ++a) v[a] = .5 * (vL[a] + vR[a]);
  double vSq = dot4(v, v);
  double oneOverW2 = 1. - vSq;
  double oneOverW = sqrt(oneOverW2);
  double W = 1. / oneOverW;
  double W2 = 1. / oneOverW2;
  double P = (gamma - 1.) * rho * eInt;
  double h = 1. + eInt + P / rho;
  double P_over_rho_h = P / (rho * h);

  if (side == 1) { double t[4] = {v[1], -v[0], v[2], 0.}; memcpy(v, t, sizeof t); }
  else if (side == 2) { double t[4] = {v[2], v[1], -v[0], 0.}; memcpy(v, t, sizeof t); }

  double hW = h * W;
  double hSq = h * h;
  double vxSq = v[0] * v[0];
  double csSq = gamma * P_over_rho_h;
  double cs = sqrt(csSq);
  double discr = sqrt((1. - vSq) * ((1. - vSq * csSq) - vxSq * (1. - csSq)));
  double lambdaMin = (v[0] * (1. - csSq) - cs * discr) / (1. - vSq * csSq);
  double lambdaMax = (v[0] * (1. - csSq) + cs * discr) / (1. - vSq * csSq);

  eigval[0] = lambdaMin;
  for (int a = 1; a <= dim; ++a) eigval[a] = v[0];
  eigval[E] = lambdaMax;

#define R(i, j) evr[(i) + N * (j)]
#define L(i, j) evl[(i) + N * (j)]
  memset(evr, 0, sizeof(double) * N * N);
  memset(evl, 0, sizeof(double) * N * N);
  double Kappa = h;
  double AMinus = (1. - vxSq) / (1. - v[0] * lambdaMin);
  double APlus = (1. - vxSq) / (1. - v[0] * lambdaMax);
  R(0, 0) = 1.;
  R(1, 0) = hW * AMinus * lambdaMin;
  if (dim > 1) R(2, 0) = hW * v[1];
  if (dim > 2) R(3, 0) = hW * v[2];
  R(E, 0) = hW * AMinus - 1.;
  R(0, 1) = oneOverW;
  R(1, 1) = v[0];
  if (dim > 1) R(2, 1) = v[1];
  if (dim > 2) R(3, 1) = v[2];
  R(E, 1) = 1. - oneOverW;
  if (dim > 1) {
    R(0, 2) = W * v[1];
    R(1, 2) = 2. * h * W2 * v[0] * v[1];
    R(2, 2) = h * (1. + 2. * W2 * v[1] * v[1]);
    if (dim > 2) R(3, 2) = 2. * h * W2 * v[1] * v[2];
    R(E, 2) = (2. * hW - 1.) * W * v[1];
  }
  if (dim > 2) {
    R(0, 3) = W * v[2];
    R(1, 3) = 2. * h * W2 * v[0] * v[2];
    R(2, 3) = 2. * h * W2 * v[1] * v[2];
    R(3, 3) = h * (1. + 2. * W2 * v[2] * v[2]);
    R(E, 3) = (2. * hW - 1.) * W * v[2];
  }
  R(0, E) = 1.;
  R(1, E) = hW * APlus * lambdaMax;
  if (dim > 1) R(2, E) = hW * v[1];
  if (dim > 2) R(3, E) = hW * v[2];
  R(E, E) = hW * APlus - 1.;

  double Delta = hSq * hW * (Kappa - 1.) * (1. - vxSq) * (APlus * lambdaMax - AMinus * lambdaMin);
  double scale = hSq / Delta;
  L(0, 0) = scale * (hW * APlus * (v[0] - lambdaMax) - v[0] - W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (v[0] - APlus * lambdaMax) + Kappa * APlus * lambdaMax);
  L(0, 1) = scale * (1. + W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (1. - APlus) - Kappa * APlus);
  if (dim > 1) L(0, 2) = scale * (W2 * v[1] * (2. * Kappa - 1.) * APlus * (v[0] - lambdaMax));
  if (dim > 2) L(0, 3) = scale * (W2 * v[2] * (2. * Kappa - 1.) * APlus * (v[0] - lambdaMax));
  L(0, E) = scale * (-v[0] - W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (v[0] - APlus * lambdaMax) + Kappa * APlus * lambdaMax);
  scale = W / (Kappa - 1.);
  L(1, 0) = scale * (h - W);
  L(1, 1) = scale * (W * v[0]);
  if (dim > 1) L(1, 2) = scale * (W * v[1]);
  if (dim > 2) L(1, 3) = scale * (W * v[2]);
  L(1, E) = scale * (-W);
  if (dim > 1) {
    scale = 1. / (h * (1. - vxSq));
    L(2, 0) = scale * (-v[1]);
    L(2, 1) = scale * (v[0] * v[1]);
    L(2, 2) = scale * (1. - vxSq);
    if (dim > 2) L(2, 3) = 0.;
    L(2, E) = scale * (-v[1]);
  }
  if (dim > 2) {
    L(3, 0) = scale * (-v[2]);
    L(3, 1) = scale * (v[0] * v[2]);
    L(3, 2) = 0.;
    L(3, 3) = scale * (1. - vxSq);
    L(3, E) = scale * (-v[2]);
  }
  scale = -hSq / Delta;
  L(E, 0) = scale * (hW * AMinus * (v[0] - lambdaMin) - v[0] - W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (v[0] - AMinus * lambdaMin) + Kappa * AMinus * lambdaMin);
  L(E, 1) = scale * (1. + W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (1. - AMinus) - Kappa * AMinus);
  if (dim > 1) L(E, 2) = scale * (W2 * v[1] * (2. * Kappa - 1.) * AMinus * (v[0] - lambdaMin));
  if (dim > 2) L(E, 3) = scale * (W2 * v[2] * (2. * Kappa - 1.) * AMinus * (v[0] - lambdaMin));
  L(E, E) = scale * (-v[0] - W2 * (vSq - vxSq) * (2. * Kappa - 1.) * (v[0] - AMinus * lambdaMin) + Kappa * AMinus * lambdaMin);

  /* rotate back (SRHDRoe.cl:266-300): rows of L: x <- -other, other <- x;
   * columns of R: x <- -other, other <- x */
  if (side >= 1) {
    int a = 1, b = 1 + side;
    for (int i = 0; i < N; ++i) {
      double t = L(i, a); L(i, a) = -L(i, b); L(i, b) = t;
      t = R(a, i); R(a, i) = -R(b, i); R(b, i) = t;
    }
  }
#undef R
#undef L
}

/* ------------------------------------------------ Maxwell transforms */

/* res/MaxwellRoe.cl:11-109: float-typed constants se, su, ise, isu */
static void maxwell_consts(const oracle_sim* s, double* se, double* su, double* ise, double* isu) {
  const float sqrt_1_2 = 0.7071067811865475727373109293694142252206802368164f;
  double sqrt_perm = oracle_kernel_dx(sqrt(s->d.maxwell_permittivity), 1);
  double sqrt_perb = oracle_kernel_dx(sqrt(s->d.maxwell_permeability), 1);
  float fse = (float)(sqrt_perm * (double)sqrt_1_2);
  float fsu = (float)(sqrt_perb * (double)sqrt_1_2);
  float fise = 1.f / fse;
  float fisu = 1.f / fsu;
  *se = fse; *su = fsu; *ise = fise; *isu = fisu;
}

static void maxwell_left(const oracle_sim* s, double* res, const double* in, int side) {
  double e[3] = {in[0], in[1], in[2]}, b[3] = {in[3], in[4], in[5]};
  if (side == 1) { double t = e[0]; e[0] = e[1]; e[1] = t; t = b[0]; b[0] = b[1]; b[1] = t; }
  else if (side == 2) { double t = e[0]; e[0] = e[2]; e[2] = t; t = b[0]; b[0] = b[2]; b[2] = t; }
  double se, su, ise, isu;
  maxwell_consts(s, &se, &su, &ise, &isu);
  res[0] = e[2] * ise + b[1] * isu;
  res[1] = e[1] * -ise + b[2] * isu;
  res[2] = e[0] * -ise + b[0] * isu;
  res[3] = e[0] * ise + b[0] * isu;
  res[4] = e[1] * ise + b[2] * isu;
  res[5] = e[2] * -ise + b[1] * isu;
}

static void maxwell_right(const oracle_sim* s, double* res, const double* in, int side) {
  double se, su, ise, isu;
  maxwell_consts(s, &se, &su, &ise, &isu);
  const double* e = in;
  const double* b = in + 3;
  res[0] = e[2] * -se + b[0] * se;
  res[1] = e[1] * -se + b[1] * se;
  res[2] = e[0] * se + b[2] * -se;
  res[3] = e[2] * su + b[0] * su;
  res[4] = e[0] * su + b[2] * su;
  res[5] = e[1] * su + b[1] * su;
  double t;
  if (side == 1) { t = res[0]; res[0] = res[1]; res[1] = t; t = res[3]; res[3] = res[4]; res[4] = t; }
  else if (side == 2) { t = res[0]; res[0] = res[4]; res[4] = t; t = res[3]; res[3] = res[5]; res[5] = t; }
}

/* ------------------------------------------------ HLL / HLLC */

static void hll_prims(int dim, double gamma, const double* q, double pot, double* vel, double* velSq,
                      double* energyTotal, double* pressure, double* enthalpyTotal) {
  const int E = dim + 1;
  double invDensity = 1. / q[0];
  vel[0] = vel[1] = vel[2] = vel[3] = 0.;
  for (int a = 0; a < dim; ++a) vel[a] = q[1 + a] / q[0];
  *velSq = dot4(vel, vel);
  *energyTotal = q[E] * invDensity;
  double energyKinetic = .5 * *velSq;
  double energyInternal = *energyTotal - energyKinetic - pot;
  *pressure = (gamma - 1.) * q[0] * energyInternal;
  *enthalpyTotal = *energyTotal + *pressure * invDensity;
}

static void rotate_in(int dim, int side, const double* src, double* dst) {
  const int N = dim + 2;
  for (int i = 0; i < N; ++i) dst[i] = src[i];
  double t = dst[1]; dst[1] = dst[1 + side]; dst[1 + side] = t;
}

/* res/EulerHLL.cl:17-115 (the cells' potentials potL, potR) and
 * res/EulerHLLC.cl:15-108 (no potential: pass zeros) */
void oracle_euler_hll_speeds_pot(int solver, int dim, int side, double gamma, const double* srcL,
                                 const double* srcR, double potL, double potR, double* eigval) {
  const int E = dim + 1;
  if (solver != HYDRO_SOLVER_HLL) potL = potR = 0.; /* res/EulerHLLC.cl reads no potential */
  double qL[5] = {0}, qR[5] = {0};
  rotate_in(dim, side, srcL, qL);
  rotate_in(dim, side, srcR, qR);
  double vL[4], vR[4], vSqL, vSqR, eTL, eTR, pL, pR, hL, hR;
  hll_prims(dim, gamma, qL, potL, vL, &vSqL, &eTL, &pL, &hL);
  hll_prims(dim, gamma, qR, potR, vR, &vSqR, &eTR, &pR, &hR);
  double speedOfSoundL = sqrt((gamma - 1.) * (hL - .5 * vSqL));
  double speedOfSoundR = sqrt((gamma - 1.) * (hR - .5 * vSqR));
  double roeWeightL = sqrt(qL[0]);
  double roeWeightR = sqrt(qR[0]);
  double roeWeightNormalization = 1. / (roeWeightL + roeWeightR);
  double velocity[4];
  for (int a = 0; a < 4; ++a)
    velocity[a] = (roeWeightL * vL[a] + roeWeightR * vR[a]) * roeWeightNormalization;
  double sl, sr;
  if (solver == HYDRO_SOLVER_HLL) {
    double enthalpyTotal = (roeWeightL * hL + roeWeightR * hR) * roeWeightNormalization;
    double energyPotential = (roeWeightL * potL + roeWeightR * potR) * roeWeightNormalization;
    double velocitySq = dot4(velocity, velocity);
    double speedOfSound = sqrt((gamma - 1.) * (enthalpyTotal - .5 * velocitySq - energyPotential));
    double eigenvaluesMinL = vL[0] - speedOfSoundL;
    double eigenvaluesMaxR = vR[0] + speedOfSoundR;
    double eigenvaluesMin = velocity[0] - speedOfSound;
    double eigenvaluesMax = velocity[0] + speedOfSound;
    sl = fmin(eigenvaluesMinL, eigenvaluesMin);
    sr = fmax(eigenvaluesMaxR, eigenvaluesMax);
  } else {
    double pressureStar = fmax(0., .5 * (pL + pR) - .5 * (vR[0] - vL[0]) * .5 * (qL[0] + qR[0]) * .5 * (speedOfSoundL + speedOfSoundR));
    double qLf = pressureStar <= pL ? 1. : sqrt(1. + (gamma + 1.) / (2. * gamma) * (pressureStar / pL - 1.));
    double qRf = pressureStar <= pR ? 1. : sqrt(1. + (gamma + 1.) / (2. * gamma) * (pressureStar / pR - 1.));
    sl = vL[0] - speedOfSoundL * qLf;
    sr = vR[0] + speedOfSoundR * qRf;
  }
  eigval[0] = sl;
  for (int a = 1; a <= dim; ++a) eigval[a] = velocity[0];
  eigval[E] = sr;
}

void oracle_euler_hll_speeds(int solver, int dim, int side, double gamma, const double* srcL,
                             const double* srcR, double* eigval) {
  oracle_euler_hll_speeds_pot(solver, dim, side, gamma, srcL, srcR, 0., 0., eigval);
}

/* res/EulerHLL.cl:185-343 (the cells' potentials in the pressures,
 * :226-240) and res/EulerHLLC.cl:178-419 (HLLC_METHOD 2, no potential) */
int oracle_euler_hll_flux_pot(int solver, int dim, int side, double gamma, const double* srcL,
                              const double* srcR, double potL, double potR, const double* eigval, double* flux) {
  const int N = dim + 2, E = dim + 1;
  if (solver != HYDRO_SOLVER_HLL) potL = potR = 0.; /* res/EulerHLLC.cl reads no potential */
  double qL[5] = {0}, qR[5] = {0};
  rotate_in(dim, side, srcL, qL);
  rotate_in(dim, side, srcR, qR);
  double vL[4], vR[4], vSqL, vSqR, eTL, eTR, pL, pR, hL, hR;
  hll_prims(dim, gamma, qL, potL, vL, &vSqL, &eTL, &pL, &hL);
  hll_prims(dim, gamma, qR, potR, vR, &vSqR, &eTR, &pR, &hR);
  double dL = qL[0], dR = qR[0];
  double sl = eigval[0];
  double sr = (solver == HYDRO_SOLVER_HLL) ? eigval[N - 1] : eigval[E];
  double fL[5], fR[5];
  fL[0] = dL * vL[0];
  fL[1] = dL * vL[0] * vL[0] + pL;
  if (dim > 1) fL[2] = dL * vL[1] * vL[0];
  if (dim > 2) fL[3] = dL * vL[2] * vL[0];
  fL[E] = dL * hL * vL[0];
  fR[0] = dR * vR[0];
  fR[1] = dR * vR[0] * vR[0] + pR;
  if (dim > 1) fR[2] = dR * vR[1] * vR[0];
  if (dim > 2) fR[3] = dR * vR[2] * vR[0];
  fR[E] = dR * hR * vR[0];
  double f[5] = {0, 0, 0, 0, 0};
  int written = 1;
  if (solver == HYDRO_SOLVER_HLL) {
    if (0. <= sl) {
      for (int i = 0; i < N; ++i) f[i] = fL[i];
    } else if (sl <= 0. && 0. <= sr) {
      double invDenom = 1. / (sr - sl);
      for (int i = 0; i < N; ++i) f[i] = (sr * fL[i] - sl * fR[i] + sl * sr * (qR[i] - qL[i])) * invDenom;
    } else if (sr <= 0.) {
      for (int i = 0; i < N; ++i) f[i] = fR[i];
    } else written = 0;
  } else {
    double sStar = (dR * vR[0] * (sr - vR[0]) - dL * vL[0] * (sl - vL[0]) + pL - pR)
                   / (dR * (sr - vR[0]) - dL * (sl - vL[0]));
    if (0. <= sl) {
      for (int i = 0; i < N; ++i) f[i] = fL[i];
    } else if (sl <= 0. && 0. <= sStar) {
      double pLR = .5 * (pL + pR + dL * (sl - vL[0]) * (sStar - vL[0]) + dR * (sr - vR[0]) * (sStar - vR[0]));
      f[0] = sStar * (sl * qL[0] - fL[0]) / (sl - sStar);
      f[1] = (sStar * (sl * qL[1] - fL[1]) + sl * pLR) / (sl - sStar);
      if (dim > 1) f[2] = sStar * (sl * qL[2] - fL[2]) / (sl - sStar);
      if (dim > 2) f[3] = sStar * (sl * qL[3] - fL[3]) / (sl - sStar);
      f[E] = (sStar * (sl * qL[E] - fL[E]) + sl * pLR * sStar) / (sl - sStar);
    } else if (sStar <= 0. && 0. <= sr) {
      double pLR = .5 * (pL + pR + dL * (sl - vL[0]) * (sStar - vL[0]) + dR * (sr - vR[0]) * (sStar - vR[0]));
      f[0] = sStar * (sr * qR[0] - fR[0]) / (sr - sStar);
      f[1] = (sStar * (sr * qR[1] - fR[1]) + sr * pLR) / (sr - sStar);
      if (dim > 1) f[2] = sStar * (sr * qR[2] - fR[2]) / (sr - sStar);
      if (dim > 2) f[3] = sStar * (sr * qR[3] - fR[3]) / (sr - sStar);
      f[E] = (sStar * (sr * qR[E] - fR[E]) + sr * pLR * sStar) / (sr - sStar);
    } else if (sr <= 0.) {
      for (int i = 0; i < N; ++i) f[i] = fR[i];
    } else written = 0;
  }
  if (!written) return 0; /* no branch taken (NaN speeds): the slot keeps its old value */
  double t = f[1]; f[1] = f[1 + side]; f[1 + side] = t;
  for (int i = 0; i < N; ++i) flux[i] = f[i];
  return 1;
}

int oracle_euler_hll_flux(int solver, int dim, int side, double gamma, const double* srcL,
                          const double* srcR, const double* eigval, double* flux) {
  return oracle_euler_hll_flux_pot(solver, dim, side, gamma, srcL, srcR, 0., 0., eigval, flux);
}

/* ------------------------------------------------ SRHD cell kernels */

/* res/SRHDCommon.cl:24-92 for one cell */
void oracle_srhd_prim2cons(int dim, double gamma, const double* state, double* cons, double* prim) {
  const int E = dim + 1;
  double rho = state[0];
  double v[4] = {0, 0, 0, 0};
  for (int a = 0; a < dim; ++a) v[a] = state[1 + a] / state[0];
  double ETotalClassic = state[E];
  double vSq = dot4(v, v);
  double eKinClassic = .5 * vSq;
  double eTotalClassic = ETotalClassic / rho;
  double eInt = eTotalClassic - eKinClassic;
  double P = (gamma - 1.) * rho * eInt;
  double h = 1. + eInt + P / rho;
  double WSq = 1. / (1. - vSq);
  double W = sqrt(WSq);
  double D = rho * W;
  double tau = rho * h * WSq - P - D;
  prim[0] = rho;
  for (int a = 0; a < dim; ++a) prim[1 + a] = v[a];
  prim[E] = eInt;
  cons[0] = D;
  for (int a = 0; a < dim; ++a) cons[1 + a] = rho * h * WSq * v[a];
  cons[E] = tau;
}

/* res/SRHDCommon.cl:113-196 for one cell: Newton on P, old prims kept when
 * the iteration does not converge */
int oracle_srhd_con2prim(const hydro_sim_desc* d, int dim, const double* state, double* prim) {
  const int E = dim + 1;
  const double gamma = d->gamma;
  double D = state[0];
  double S[4] = {0, 0, 0, 0};
  for (int a = 0; a < dim; ++a) S[a] = state[1 + a];
  double tau = state[E];
  double SLen = sqrt(dot4(S, S));
  double PMin = fmax(SLen - tau - D + SLen * d->srhd_solve_prim_vel_epsilon, d->srhd_solve_prim_pmin_epsilon);
  double PMax = (gamma - 1.) * tau;
  PMax = fmax(PMax, PMin);
  double P = .5 * (PMin + PMax);
  for (int iter = 0; iter < d->srhd_solve_prim_max_iter; ++iter) {
    double vLen = SLen / (tau + D + P);
    double vSq = vLen * vLen;
    double W = 1. / sqrt(1. - vSq);
    double eInt = (tau + D * (1. - W) + P * (1. - W * W)) / (D * W);
    double rho = D / W;
    double f = (gamma - 1.) * rho * eInt - P;
    double csSq = (gamma - 1.) * (tau + D * (1. - W) + P) / (tau + D + P);
    double df_dP = vSq * csSq - 1.;
    double newP = P - f / df_dP;
    newP = fmax(newP, PMin);
    double PError = fabs(1. - newP / P);
    P = newP;
    if (PError < d->srhd_solve_prim_stop_epsilon) {
      double v[4];
      for (int a = 0; a < 4; ++a) v[a] = S[a] / (tau + D + P);
      W = 1. / sqrt(1. - dot4(v, v));
      rho = D / W;
      rho = fmax(rho, d->srhd_rho_min);
      rho = fmin(rho, d->srhd_rho_max);
      eInt = P / (rho * (gamma - 1.));
      eInt = fmin(eInt, d->srhd_eint_max);
      prim[0] = rho;
      for (int a = 0; a < dim; ++a) prim[1 + a] = v[a];
      prim[E] = eInt;
      return iter + 1;
    }
  }
  return -1;
}

/* ------------------------------------------------ create / destroy */

static int eqn_num_states(int eqn, int dim) {
  if (eqn == HYDRO_EQN_MAXWELL) return 6;
  return dim + 2;
}

oracle_sim* oracle_create(const hydro_sim_desc* d, unsigned opts) {
  if (!d || d->dim < 1 || d->dim > 3) return NULL;
  oracle_sim* s = (oracle_sim*)calloc(1, sizeof(oracle_sim));
  s->d = *d;
  s->opts = opts;
  s->dim = d->dim;
  s->N = eqn_num_states(d->equation, d->dim);
  for (int a = 0; a < 3; ++a) s->sz[a] = (a < d->dim) ? d->size[a] : 1;
  s->step[0] = 1;
  s->step[1] = s->sz[0];
  s->step[2] = s->sz[0] * s->sz[1];
  s->vol = s->sz[0] * s->sz[1] * s->sz[2];
  for (int a = 0; a < 3; ++a) {
    /* src/HydroGPUApp.cpp:224-226: (xmax - xmin) / (float)size */
    double sz = (a < d->dim) ? (double)(float)d->size[a] : 1.;
    s->dx[a] = (d->xmax[a] - d->xmin[a]) / sz;
    s->kdx[a] = oracle_kernel_dx(s->dx[a], (d->flags & HYDRO_FLAG_REF_DX_6DIGITS) != 0);
  }
  s->cfl = (double)(float)d->cfl;
  s->gamma = d->gamma;
  const int N = s->N, V = s->vol, D = s->dim;
  s->ev_size = (d->equation == HYDRO_EQN_MAXWELL) ? 1 : 2 * N * N;
  s->state = (double*)calloc((size_t)N * V, sizeof(double));
  s->eigval = (double*)calloc((size_t)N * V * D, sizeof(double));
  if (d->solver == HYDRO_SOLVER_ROE) {
    s->eigvec = (double*)calloc((size_t)s->ev_size * V * D, sizeof(double));
    s->dqt = (double*)calloc((size_t)N * V * D, sizeof(double));
  }
  s->flux = (double*)calloc((size_t)N * V * D, sizeof(double));
  s->dtbuf = (double*)malloc(sizeof(double) * V * D);
  for (int i = 0; i < V * D; ++i) s->dtbuf[i] = DBL_MAX; /* Solver.cpp:217-222 */
  s->potential = (double*)calloc((size_t)V, sizeof(double));
  s->pot_next = (double*)calloc((size_t)V, sizeof(double));
  if (d->equation == HYDRO_EQN_SRHD) s->prim = (double*)calloc((size_t)N * V, sizeof(double));
  if (d->solver == HYDRO_SOLVER_BURGERS) {
    /* src/Solver/EulerBurgers.cpp:14-22: both zeroed at creation */
    s->ifvel = (double*)calloc((size_t)V * D, sizeof(double));
    s->pressure = (double*)calloc((size_t)V, sizeof(double));
  }
  if (oracle_rk_tableau(d->integrator, &s->order, &s->alpha[0][0], &s->beta[0][0]) != 0) {
    oracle_destroy(s);
    return NULL;
  }
  if (d->integrator == HYDRO_INTEGRATOR_FORWARDEULER) {
    s->fe_deriv = (double*)calloc((size_t)N * V, sizeof(double));
  } else {
    for (int i = 0; i < s->order; ++i) {
      s->rk_state[i] = (double*)calloc((size_t)N * V, sizeof(double));
      s->rk_deriv[i] = (double*)calloc((size_t)N * V, sizeof(double));
    }
  }
  return s;
}

void oracle_destroy(oracle_sim* s) {
  if (!s) return;
  free(s->state); free(s->eigval); free(s->eigvec); free(s->dqt); free(s->flux);
  free(s->dtbuf); free(s->potential); free(s->pot_next); free(s->prim); free(s->fe_deriv); free(s->ifvel); free(s->pressure);
  free(s->solid);
  for (int i = 0; i < 4; ++i) { free(s->rk_state[i]); free(s->rk_deriv[i]); }
  free(s);
}

int oracle_num_states(const oracle_sim* s) { return s->N; }
size_t oracle_volume(const oracle_sim* s) { return (size_t)s->vol; }
double oracle_time(const oracle_sim* s) { return s->t; }

/* Solver::resetState upload; SRHD runs initVariables (SRHDRoe.cpp:40-45);
 * Euler adds the potential into the total energy
 * (src/Solver/SelfGravitation.cpp:89-94). */
int oracle_set_state(oracle_sim* s, const double* aos, size_t n) {
  if (n != (size_t)s->N * s->vol) return -1;
  memcpy(s->state, aos, n * sizeof(double));
  if (s->d.equation == HYDRO_EQN_SRHD) {
    for (int c = 0; c < s->vol; ++c) {
      double in[5], cons[5], prim[5];
      memcpy(in, s->state + (size_t)s->N * c, sizeof(double) * s->N);
      oracle_srhd_prim2cons(s->dim, s->gamma, in, cons, prim);
      memcpy(s->state + (size_t)s->N * c, cons, sizeof(double) * s->N);
      memcpy(s->prim + (size_t)s->N * c, prim, sizeof(double) * s->N);
    }
  }
  s->t = 0.;
  s->steps = 0;
  return 0;
}

static void gravity_relax(oracle_sim* s, int sweeps);

/* SelfGravitation::resetState (src/Solver/SelfGravitation.cpp:35-97) after
 * the state upload: the initial potential phi0 (`pot`, NULL: zero; the
 * guess -rho when useGravity), 100 x gaussSeidelMaxIter relaxation sweeps
 * when useGravity, then phi0 added into every cell's total energy. */
int oracle_set_potential(oracle_sim* s, const double* pot, size_t n) {
  if ((pot && n != (size_t)s->vol) || s->d.equation != HYDRO_EQN_EULER) return -1;
  const int V = s->vol;
  double* phi0 = (double*)malloc((size_t)V * sizeof(double));
  for (int c = 0; c < V; ++c)
    phi0[c] = s->d.use_gravity ? -s->state[(size_t)s->N * c] : (pot ? pot[c] : 0.);
  memcpy(s->potential, phi0, (size_t)V * sizeof(double));
  int nonzero = s->d.use_gravity;
  for (int c = 0; c < V; ++c) nonzero |= phi0[c] != 0.;
  s->pot_on = nonzero;
  if (s->d.use_gravity)
    for (int tries = 0; tries < 100; ++tries) gravity_relax(s, s->d.gauss_seidel_max_iter);
  for (int c = 0; c < V; ++c) s->state[(size_t)s->N * c + s->dim + 1] += phi0[c];
  free(phi0);
  return 0;
}

/* SelfGravitation::resetState writes solidBuffer (src/Solver/SelfGravitation.cpp:75) */
int oracle_set_solid(oracle_sim* s, const unsigned char* solid, size_t n) {
  if (solid && n != (size_t)s->vol) return -1;
  free(s->solid);
  s->solid = NULL;
  if (!solid) return 0;
  s->solid = (unsigned char*)malloc((size_t)s->vol);
  for (int c = 0; c < s->vol; ++c) s->solid[c] = solid[c] != 0;
  return 0;
}

/* the solid flag of a cell as the running solver's kernels see it: only the
 * SOLID builds (EulerRoe, EulerBurgers) read solidBuffer */
static inline int solid_at(const oracle_sim* s, int c) {
  return s->solid && s->d.equation == HYDRO_EQN_EULER &&
         (s->d.solver == HYDRO_SOLVER_ROE || s->d.solver == HYDRO_SOLVER_BURGERS) && s->solid[c];
}

int oracle_get_potential(const oracle_sim* s, double* pot, size_t n) {
  if (n != (size_t)s->vol) return -1;
  memcpy(pot, s->potential, n * sizeof(double));
  return 0;
}

int oracle_get_state(const oracle_sim* s, double* aos, size_t n) {
  if (n != (size_t)s->N * s->vol) return -1;
  memcpy(aos, s->state, n * sizeof(double));
  return 0;
}

int oracle_get_primitives(const oracle_sim* s, double* aos, size_t n) {
  if (!s->prim || n != (size_t)s->N * s->vol) return -1;
  memcpy(aos, s->prim, n * sizeof(double));
  return 0;
}

/* ------------------------------------------------ boundary */

enum { BK_PERIODIC = 0, BK_MIRROR, BK_REFLECT, BK_FREEFLOW };

/* Euler/SRHD (src/Equation/Euler.cpp:63-75, src/Equation/SRHD.cpp:63-75) and
 * Maxwell (src/Equation/Maxwell.cpp:73-85) boundary kernel per state */
static int boundary_kernel(const oracle_sim* s, int axis, int state, int minmax) {
  switch (s->d.boundary[axis][minmax]) {
  case HYDRO_BOUNDARY_PERIODIC: return BK_PERIODIC;
  case HYDRO_BOUNDARY_MIRROR:
    if (s->d.equation == HYDRO_EQN_MAXWELL)
      return (state == axis || state == 3 + axis) ? BK_REFLECT : BK_MIRROR;
    return (axis + 1 == state) ? BK_REFLECT : BK_MIRROR;
  case HYDRO_BOUNDARY_FREEFLOW: return BK_FREEFLOW;
  }
  return -1;
}

/* SelfGravitationBehavior::gravityGetBoundaryKernelForBoundaryMethod
 * (include/HydroGPU/Equation/SelfGravitationBehavior.h:22-34): the
 * potential's ghosts copy (mirror becomes free-flow, nothing is negated) */
static int potential_boundary_kernel(const oracle_sim* s, int axis, int state, int minmax) {
  (void)state;
  switch (s->d.boundary[axis][minmax]) {
  case HYDRO_BOUNDARY_PERIODIC: return BK_PERIODIC;
  case HYDRO_BOUNDARY_MIRROR: return BK_FREEFLOW;
  case HYDRO_BOUNDARY_FREEFLOW: return BK_FREEFLOW;
  }
  return -1;
}

/* res/Common.cl:49-194 applied over the boundary NDRange of
 * src/Solver/Solver.cpp:412-444; in 1D only work-item 0 is in bounds.
 * N interleaved states per cell, each state's kernel from `kernel`. */
static void boundary_buffer_n(oracle_sim* s, double* buf, int N,
                              int (*kernel)(const oracle_sim*, int, int, int)) {
  for (int axis = 0; axis < s->dim; ++axis) {
    int o1 = (axis == 0) ? 1 : 0;
    int o2 = (axis == 2) ? 1 : 2;
    int n1 = s->sz[o1], n2 = s->sz[o2];
    int S = s->sz[axis];
    int st = s->step[axis];
    for (int j = 0; j < N; ++j) {
      for (int minmax = 0; minmax < 2; ++minmax) {
        int k = kernel(s, axis, j, minmax);
        if (k < 0) continue;
        for (int b = 0; b < n2; ++b) {
          for (int a = 0; a < n1; ++a) {
            int i3[3] = {0, 0, 0};
            i3[o1] = a;
            i3[o2] = b;
            i3[axis] = 0;
            int base = idx3(s, i3[0], i3[1], i3[2]);
#define B(i) buf[j + N * (base + st * (i))]
            if (k == BK_PERIODIC) {
              if (!minmax) { B(0) = B(S - 4); B(1) = B(S - 3); }
              else { B(S - 2) = B(2); B(S - 1) = B(3); }
            } else if (k == BK_MIRROR) {
              if (!minmax) { B(0) = B(3); B(1) = B(2); }
              else { B(S - 1) = B(S - 4); B(S - 2) = B(S - 3); }
            } else if (k == BK_REFLECT) {
              if (!minmax) { B(0) = -B(3); B(1) = -B(2); }
              else { B(S - 1) = -B(S - 4); B(S - 2) = -B(S - 3); }
            } else {
              if (!minmax) { B(1) = B(2); B(0) = B(1); }
              else { B(S - 2) = B(S - 3); B(S - 1) = B(S - 2); }
            }
#undef B
          }
        }
      }
    }
  }
}

static void boundary_buffer(oracle_sim* s, double* buf) { boundary_buffer_n(s, buf, s->N, boundary_kernel); }

void oracle_boundary(oracle_sim* s) {
  boundary_buffer(s, s->state);
  if (s->prim) boundary_buffer(s, s->prim); /* SRHDRoe.cpp:79-96 */
}

/* SelfGravitation::potentialBoundary (src/Solver/SelfGravitation.cpp:126-141) */
static void potential_boundary(oracle_sim* s) { boundary_buffer_n(s, s->potential, 1, potential_boundary_kernel); }

/* OpenCL's M_PI for double */
#define ORACLE_PI 3.141592653589793115997963468544185161590576171875

/* res/SelfGravitation.cl:20-63: one Jacobi-style relaxation sweep of
 * nabla^2 phi = 4 pi G rho on the interior cells (neighbours outside the
 * interior count as zero).  The reference updates the potential in place,
 * so its sweep reads neighbours old or new depending on the order its
 * work-items run; this restatement reads the previous sweep everywhere. */
static void gravity_relax_sweep(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
  const double* phi = s->potential;
  double* next = s->pot_next;
  memcpy(next, phi, (size_t)V * sizeof(double));
  double diagsum = 0.;
  for (int a = 0; a < D; ++a) diagsum += 1. / (s->kdx[a] * s->kdx[a]);
  const double diag = -2. * diagsum;
  const double G = s->d.self_grav_gravitational_constant;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2)) continue;
    double skewSum = 0.;
    for (int a = 0; a < D; ++a) {
      const double h2 = s->kdx[a] * s->kdx[a];
      if (i[a] < s->sz[a] - 3) skewSum += phi[c + s->step[a]] / h2;
      if (i[a] > 2) skewSum += phi[c - s->step[a]] / h2;
    }
    const double density = s->state[(size_t)N * c];
    next[c] = (4. * ORACLE_PI * G * density - skewSum) / diag;
  }
  s->potential = next;
  s->pot_next = (double*)phi;
}

/* gaussSeidelMaxIter x (potential boundary, relaxation sweep)
 * (src/Solver/SelfGravitation.cpp:112-117, 83-88) */
static void gravity_relax(oracle_sim* s, int sweeps) {
  for (int k = 0; k < sweeps; ++k) {
    potential_boundary(s);
    gravity_relax_sweep(s);
  }
}

/* ------------------------------------------------ Roe kernels */

/* calcEigenBasis: EulerRoe.cl:261-271, SRHDRoe.cl:303-312, MaxwellRoe.cl:117-157 */
static void calc_eigen_basis(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      int prev = c - s->step[side];
      size_t iface = (size_t)side + D * c;
      double* ev = s->eigval + N * iface;
      if (s->d.equation == HYDRO_EQN_EULER) {
        double* evl = s->eigvec + s->ev_size * iface;
        oracle_euler_eigenbasis_solid(D, side, s->gamma, s->state + N * prev, s->state + N * c,
                                      s->potential[prev], s->potential[c], solid_at(s, prev), solid_at(s, c), ev,
                                      evl, evl + N * N);
      } else if (s->d.equation == HYDRO_EQN_SRHD) {
        double* evl = s->eigvec + s->ev_size * iface;
        oracle_srhd_eigenbasis(D, side, s->gamma, s->prim + N * prev, s->prim + N * c, ev, evl, evl + N * N);
      } else {
        double sqrt_perm = oracle_kernel_dx(sqrt(s->d.maxwell_permittivity), 1);
        double sqrt_perb = oracle_kernel_dx(sqrt(s->d.maxwell_permeability), 1);
        double e = 1. / (sqrt_perm * sqrt_perb);
        ev[0] = -e; ev[1] = -e; ev[2] = 0.; ev[3] = 0.; ev[4] = e; ev[5] = e;
      }
    }
  }
}

/* res/Roe.cl:4-65 */
static void roe_calc_cell_timestep(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2)) {
      for (int side = 0; side < D; ++side) s->dtbuf[side + D * c] = INFINITY;
      continue;
    }
    for (int side = 0; side < D; ++side) {
      /* res/Roe.cl:38-44: no limit across a side that touches a solid cell */
      if (solid_at(s, c) || solid_at(s, c + s->step[side])) {
        s->dtbuf[side + D * c] = INFINITY;
        continue;
      }
      const double* evL = s->eigval + (size_t)N * (side + D * c);
      const double* evR = s->eigval + (size_t)N * (side + D * (c + s->step[side]));
      double maxLambda = fmax(0., evL[N - 1]);
      double minLambda = fmin(0., evR[0]);
      s->dtbuf[side + D * c] = s->kdx[side] / (fabs(maxLambda - minLambda) + 1e-9);
    }
  }
}

/* res/EulerHLL.cl:142-175, res/EulerHLLC.cl:135-168 */
static void hll_calc_cell_timestep(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
  const int top = D + 1; /* DIM+1 (EulerHLL.cl:168) == EULER_DIM+1 (EulerHLLC.cl:161) */
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) { s->dtbuf[c] = INFINITY; continue; }
    double result = INFINITY;
    for (int side = 0; side < D; ++side) {
      /* res/Roe.cl:38-44: no limit across a side that touches a solid cell */
      if (solid_at(s, c) || solid_at(s, c + s->step[side])) {
        s->dtbuf[side + D * c] = INFINITY;
        continue;
      }
      const double* evL = s->eigval + (size_t)N * (side + D * c);
      const double* evR = s->eigval + (size_t)N * (side + D * (c + s->step[side]));
      double minLambda = fmin(0., evR[0]);
      double maxLambda = fmax(0., evL[top]);
      double dum = s->kdx[side] / (maxLambda - minLambda);
      result = fmin(result, dum);
    }
    s->dtbuf[c] = result;
  }
}

/* Solver::findMinTimestep (src/Solver/Solver.cpp:466-516): the tree of
 * findMinTimestep passes (res/Common.cl:9-40) is an exact min. */
static double find_min_timestep(oracle_sim* s) {
  double m = INFINITY;
  const int n = s->vol * s->dim;
  for (int i = 0; i < n; ++i) {
    double e = s->dtbuf[i];
    if (e < m) m = e;
  }
  return m * s->cfl;
}

static void left_transform(oracle_sim* s, double* res, const double* ev, const double* in, int side) {
  if (s->d.equation == HYDRO_EQN_MAXWELL) { maxwell_left(s, res, in, side); return; }
  const int N = s->N;
  /* res/RoeEigenfieldLinear.cl:16-27 */
  for (int i = 0; i < N; ++i) {
    double sum = 0.;
    for (int j = 0; j < N; ++j) sum += ev[i + N * j] * in[j];
    res[i] = sum;
  }
}

static void right_transform(oracle_sim* s, double* res, const double* ev, const double* in, int side) {
  if (s->d.equation == HYDRO_EQN_MAXWELL) { maxwell_right(s, res, in, side); return; }
  const int N = s->N;
  const double* m = ev + N * N;
  for (int i = 0; i < N; ++i) {
    double sum = 0.;
    for (int j = 0; j < N; ++j) sum += m[i + N * j] * in[j];
    res[i] = sum;
  }
}

static const double* eigvec_at(oracle_sim* s, size_t iface) {
  if (s->d.equation == HYDRO_EQN_MAXWELL) return NULL;
  return s->eigvec + s->ev_size * iface;
}

/* res/Roe.cl:110-124 and 235-248: across a face between a solid and a fluid
 * cell the solid side takes the fluid cell's state with the normal momentum
 * negated.  qL/qR are redirected to the copies mL/mR when mirrored. */
static void mirror_wall(const oracle_sim* s, int l, int r, int side, const double** qL, const double** qR,
                        double* mL, double* mR) {
  const int N = s->N;
  const int sl = solid_at(s, l), sr = solid_at(s, r);
  if (sl && !sr) {
    for (int k = 0; k < N; ++k) mL[k] = (*qR)[k];
    mL[side + 1] = -mL[side + 1];
    *qL = mL;
  } else if (sr && !sl) {
    for (int k = 0; k < N; ++k) mR[k] = (*qL)[k];
    mR[side + 1] = -mR[side + 1];
    *qR = mR;
  }
}

/* res/Roe.cl:77-168 */
static void calc_delta_q_tilde(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
  const int separate = (s->d.equation == HYDRO_EQN_MAXWELL);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      int prev = c - s->step[side];
      size_t iface = (size_t)side + D * c;
      const double* ev = eigvec_at(s, iface);
      double* dq = s->dqt + N * iface;
      double mL[6], mR[6];
      const double* qL = s->state + (size_t)N * prev;
      const double* qR = s->state + (size_t)N * c;
      mirror_wall(s, prev, c, side, &qL, &qR, mL, mR);
      if (separate) {
        double tl[6], tr[6];
        left_transform(s, tl, ev, qL, side);
        left_transform(s, tr, ev, qR, side);
        for (int k = 0; k < N; ++k) dq[k] = tr[k] - tl[k];
      } else {
        double d[6], o[6];
        for (int k = 0; k < N; ++k) d[k] = qR[k] - qL[k];
        left_transform(s, o, ev, d, side);
        for (int k = 0; k < N; ++k) dq[k] = o[k];
      }
    }
  }
}

/* res/Roe.cl:183-320 */
static void roe_calc_flux(oracle_sim* s, double dt) {
  const int N = s->N, D = s->dim, V = s->vol;
  const int separate = (s->d.equation == HYDRO_EQN_MAXWELL);
  const int lim = s->d.limiter;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      double dt_dx = dt / s->kdx[side];
      int indexL = c - s->step[side];
      int indexR2 = c + s->step[side];
      size_t iface = (size_t)side + D * c;
      const double* dqL = s->dqt + (size_t)N * (side + D * indexL);
      const double* dq = s->dqt + (size_t)N * iface;
      const double* dqR = s->dqt + (size_t)N * (side + D * indexR2);
      const double* eigenvalues = s->eigval + N * iface;
      const double* ev = eigvec_at(s, iface);
      const double* qL = s->state + (size_t)N * indexL;
      const double* qR = s->state + (size_t)N * c;
      double mL[6], mR[6];
      mirror_wall(s, indexL, c, side, &qL, &qR, mL, mR);
      /* res/Roe.cl:249-250, 278-290: a solid cell two over sets the ratio to 1 */
      const int solidL2 = solid_at(s, indexL - s->step[side]);
      const int solidR2 = solid_at(s, indexR2);
      double fluxTilde[6];
      if (separate) {
        double tl[6], tr[6];
        left_transform(s, tl, ev, qL, side);
        left_transform(s, tr, ev, qR, side);
        for (int k = 0; k < N; ++k) fluxTilde[k] = .5 * (tr[k] + tl[k]);
      } else {
        double avg[6];
        for (int k = 0; k < N; ++k) avg[k] = .5 * (qR[k] + qL[k]);
        left_transform(s, fluxTilde, ev, avg, side);
      }
      for (int k = 0; k < N; ++k) {
        double eigenvalue = eigenvalues[k];
        fluxTilde[k] *= eigenvalue;
        double rTilde, theta;
        if (eigenvalue >= 0.) {
          rTilde = dqL[k] / dq[k];
          theta = 1.;
          if (solidL2) rTilde = 1.;
        } else {
          rTilde = dqR[k] / dq[k];
          theta = -1.;
          if (solidR2) rTilde = 1.;
        }
        double phi = oracle_slope_limiter(lim, rTilde);
        double epsilon = eigenvalue * dt_dx;
        double deltaFluxTilde = eigenvalue * dq[k];
        fluxTilde[k] -= .5 * deltaFluxTilde * (theta + phi * (epsilon - theta));
      }
      right_transform(s, s->flux + (size_t)N * iface, ev, fluxTilde, side);
    }
  }
}

/* res/CalcFluxDeriv.cl:7-44 */
static void calc_flux_deriv(oracle_sim* s, double* deriv) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2)) continue;
    if (solid_at(s, c)) continue; /* res/CalcFluxDeriv.cl:28-30 */
    double* dv = deriv + (size_t)N * c;
    for (int side = 0; side < D; ++side) {
      size_t iface = (size_t)side + D * c;
      size_t ifaceNext = iface + (size_t)D * s->step[side];
      const double* fL = s->flux + N * iface;
      const double* fR = s->flux + N * ifaceNext;
      for (int j = 0; j < N; ++j) {
        double deltaFlux = fR[j] - fL[j];
        dv[j] -= deltaFlux / s->kdx[side];
      }
    }
  }
}

/* HLL::initStep -> calcEigenvalues (res/EulerHLL.cl:117-138) */
static void hll_calc_eigenvalues(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      int prev = c - s->step[side];
      const int pot = s->d.solver == HYDRO_SOLVER_HLL;
      oracle_euler_hll_speeds_pot(s->d.solver, D, side, s->gamma, s->state + (size_t)N * prev,
                                  s->state + (size_t)N * c, pot ? s->potential[prev] : 0.,
                                  pot ? s->potential[c] : 0., s->eigval + (size_t)N * (side + D * c));
    }
  }
}

/* res/EulerHLL.cl:345-369 */
static void hll_calc_flux(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      int prev = c - s->step[side];
      size_t iface = (size_t)side + D * c;
      /* EulerHLL reads the potential, EulerHLLC does not */
      const int pot = s->d.solver == HYDRO_SOLVER_HLL;
      oracle_euler_hll_flux_pot(s->d.solver, D, side, s->gamma, s->state + (size_t)N * prev,
                                s->state + (size_t)N * c, pot ? s->potential[prev] : 0., pot ? s->potential[c] : 0.,
                                s->eigval + N * iface, s->flux + N * iface);
    }
  }
}

/* ------------------------------------------------ EulerBurgers */

/* velocity, squared speed and internal energy of one Euler cell
 * (res/EulerBurgers.cl:48-66, 311-327), with the cell's potential */
static double burgers_eint(const oracle_sim* s, const double* st, double pot, double* vel) {
  const int D = s->dim;
  double density = st[0];
  double energyTotal = st[D + 1];
  double velocitySq = 0.;
  for (int a = 0; a < D; ++a) {
    vel[a] = st[1 + a] / density;
    velocitySq += vel[a] * vel[a];
  }
  double specificEnergyTotal = energyTotal / density;
  double specificEnergyKinetic = .5 * velocitySq;
  return specificEnergyTotal - specificEnergyKinetic - pot;
}

/* res/EulerBurgers.cl:19-79: dx / (c + |v|) per axis, the smallest */
static void burgers_calc_cell_timestep(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
  const double gamma = s->gamma;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2) || solid_at(s, c)) { s->dtbuf[c] = INFINITY; continue; } /* EulerBurgers.cl:42-47 */
    double vel[3];
    double specificEnergyInternal = burgers_eint(s, s->state + (size_t)N * c, s->potential[c], vel);
    double speedOfSound = sqrt(gamma * (gamma - 1.) * specificEnergyInternal);
    double result = s->kdx[0] / (speedOfSound + fabs(vel[0]));
    for (int a = 1; a < D; ++a) result = fmin(result, s->kdx[a] / (speedOfSound + fabs(vel[a])));
    s->dtbuf[c] = result;
  }
}

/* res/EulerBurgers.cl:81-144: the mean of the two cells' normal velocities
 * on interfaces 2 .. size-2 of every axis */
static void burgers_calc_interface_velocity(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      int l = c - s->step[side];
      double densityL = s->state[(size_t)N * l];
      double velocityL = s->state[side + 1 + (size_t)N * l] / densityL;
      double densityR = s->state[(size_t)N * c];
      double velocityR = s->state[side + 1 + (size_t)N * c] / densityR;
      /* res/EulerBurgers.cl:119-127 */
      const int solidL = solid_at(s, l), solidR = solid_at(s, c);
      if (solidL && !solidR) velocityL = -velocityR;
      else if (solidR && !solidL) velocityR = -velocityL;
      s->ifvel[side + D * c] = .5 * (velocityL + velocityR);
    }
  }
}

/* res/EulerBurgers.cl:146-281: upwind advective flux of every state with the
 * limited correction, the ratio from the upwind neighbour's jump */
static void burgers_calc_flux(oracle_sim* s, double dt) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    for (int side = 0; side < D; ++side) {
      double dt_dx = dt / s->kdx[side];
      int iR = c, iL = c - s->step[side], iL2 = iL - s->step[side], iR2 = c + s->step[side];
      size_t iface = (size_t)side + D * c;
      double interfaceVelocity = s->ifvel[iface];
      double* flux = s->flux + (size_t)N * iface;
      const int solidL2 = solid_at(s, iL2), solidL = solid_at(s, iL);
      const int solidR = solid_at(s, iR), solidR2 = solid_at(s, iR2);
      for (int j = 0; j < N; ++j) {
        double stateR2 = s->state[j + (size_t)N * iR2];
        double stateR = s->state[j + (size_t)N * iR];
        double stateL = s->state[j + (size_t)N * iL];
        double stateL2 = s->state[j + (size_t)N * iL2];
        /* res/EulerBurgers.cl:209-236: a solid neighbour takes the fluid
         * state, the normal momentum kept and every other state negated; a
         * solid cell two over repeats the near state, normal momentum negated */
        const int normal = j == 1 + side;
        if (solidL && solidR) {
        } else if (solidL) {
          stateL = normal ? stateR : -stateR;
        } else if (solidR) {
          stateR = normal ? stateL : -stateL;
        }
        if (solidL2) {
          stateL2 = stateL;
          if (normal) stateL2 = -stateL2;
        }
        if (solidR2) {
          stateR2 = stateR;
          if (normal) stateR2 = -stateR2;
        }
        double deltaStateL = stateL - stateL2;
        double deltaState = stateR - stateL;
        double deltaStateR = stateR2 - stateR;
        double theta, stateSlopeRatio;
        if (interfaceVelocity >= 0.) {
          theta = 1.;
          stateSlopeRatio = deltaStateL / deltaState;
        } else {
          theta = -1.;
          stateSlopeRatio = deltaStateR / deltaState;
        }
        double phi = oracle_slope_limiter(s->d.limiter, stateSlopeRatio);
        flux[j] = .5 * interfaceVelocity * ((1. + theta) * stateL + (1. - theta) * stateR)
                  + .5 * deltaState * phi * fabs(interfaceVelocity) * (1. - fabs(interfaceVelocity * dt_dx));
      }
    }
  }
}

/* res/EulerBurgers.cl:283-348 (no artificial viscosity): cells 1 .. size-2 */
static void burgers_compute_pressure(oracle_sim* s) {
  const int N = s->N, D = s->dim, V = s->vol;
  const double gamma = s->gamma;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    int in = 1;
    for (int a = 0; a < D; ++a) in &= i[a] >= 1 && i[a] < s->sz[a] - 1;
    if (!in || solid_at(s, c)) continue; /* EulerBurgers.cl:307-309: the buffer keeps its old value */
    double vel[3];
    const double* st = s->state + (size_t)N * c;
    double specificEnergyInternal = burgers_eint(s, st, s->potential[c], vel);
    s->pressure[c] = (gamma - 1.) * st[0] * specificEnergyInternal;
  }
}

/* res/EulerBurgers.cl:350-406: deriv[momentum] -= (pR - pL) / (2 dx) */
static void burgers_diffuse_momentum(oracle_sim* s, double* deriv) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2) || solid_at(s, c)) continue;
    double* dv = deriv + (size_t)N * c;
    for (int side = 0; side < D; ++side) {
      double pressureL = s->pressure[c - s->step[side]];
      double pressureR = s->pressure[c + s->step[side]];
      /* EulerBurgers.cl:372-375: a solid neighbour reads as this cell's pressure */
      if (solid_at(s, c - s->step[side])) pressureL = s->pressure[c];
      if (solid_at(s, c + s->step[side])) pressureR = s->pressure[c];
      dv[1 + side] -= .5 * (pressureR - pressureL) / s->kdx[side];
    }
  }
}

/* res/EulerBurgers.cl:408-489: deriv[energy] -= (pR vR - pL vL) / (2 dx),
 * velocities from the current state, pressures from the buffer as the last
 * computePressure left it */
static void burgers_diffuse_work(oracle_sim* s, double* deriv) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2) || solid_at(s, c)) continue;
    double* dv = deriv + (size_t)N * c;
    double deltaEnergyTotal = 0.;
    for (int side = 0; side < D; ++side) {
      int l = c - s->step[side], r = c + s->step[side];
      double velocityL = s->state[1 + side + (size_t)N * l] / s->state[(size_t)N * l];
      double velocityR = s->state[1 + side + (size_t)N * r] / s->state[(size_t)N * r];
      double pressureL = s->pressure[l];
      double pressureR = s->pressure[r];
      /* EulerBurgers.cl:442-447: a solid neighbour's own velocity negated,
       * and this cell's pressure */
      if (solid_at(s, l)) velocityL = -velocityL;
      if (solid_at(s, r)) velocityR = -velocityR;
      if (solid_at(s, l)) pressureL = s->pressure[c];
      if (solid_at(s, r)) pressureR = s->pressure[c];
      deltaEnergyTotal -= .5 * (pressureR * velocityR - pressureL * velocityL) / s->kdx[side];
    }
    dv[D + 1] += deltaEnergyTotal;
  }
}

/* res/SelfGravitation.cl:65-102 */
static void calc_gravity_deriv(oracle_sim* s, double* deriv) {
  const int N = s->N, D = s->dim, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 2)) continue;
    double* dv = deriv + (size_t)N * c;
    const double* st = s->state + (size_t)N * c;
    double density = st[0];
    double derivEnergyTotal = 0.;
    for (int side = 0; side < D; ++side) {
      int prev = c - s->step[side];
      int next = c + s->step[side];
      double gradient = (s->potential[next] - s->potential[prev]) / (2. * s->kdx[side]);
      double gravity = -gradient;
      dv[side + 1] -= density * gravity;
      derivEnergyTotal -= density * gravity * st[side + 1];
    }
    dv[D + 1] += derivEnergyTotal;
  }
}

/* res/SRHDCommon.cl:94-109 (every cell) */
static void srhd_constrain_state(oracle_sim* s) {
  const int N = s->N, E = s->dim + 1;
  for (int c = 0; c < s->vol; ++c) {
    double* st = s->state + (size_t)N * c;
    st[0] = fmax(st[0], s->d.srhd_D_min);
    st[E] = fmax(st[E], s->d.srhd_tau_min);
    st[0] = fmin(st[0], s->d.srhd_D_max);
    st[E] = fmin(st[E], s->d.srhd_tau_max);
  }
}

/* res/SRHDCommon.cl:113-196 over its NDRange */
static void srhd_update_primitives(oracle_sim* s) {
  const int N = s->N, V = s->vol;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < V; ++c) {
    int i[3];
    decode(s, c, i);
    if (!active(s, i, 1)) continue;
    oracle_srhd_con2prim(&s->d, s->dim, s->state + (size_t)N * c, s->prim + (size_t)N * c);
  }
}

/* ------------------------------------------------ integrators */

typedef enum { CB_HYDRO = 0, CB_GRAVITY, CB_MAXWELL_SOURCE, CB_BURGERS_MOMENTUM, CB_BURGERS_WORK } callback_t;

static void roe_calc_deriv(oracle_sim* s, double* deriv, double dt) {
  /* src/Solver/Roe.cpp:101-112 */
  calc_delta_q_tilde(s);
  roe_calc_flux(s, dt);
  calc_flux_deriv(s, deriv);
}

static void hll_deriv(oracle_sim* s, double* deriv) {
  /* src/Solver/HLL.cpp:234-241 */
  hll_calc_flux(s);
  calc_flux_deriv(s, deriv);
}

static void run_callback(oracle_sim* s, callback_t cb, double* deriv, double dt) {
  if (cb == CB_HYDRO) {
    if (s->d.solver == HYDRO_SOLVER_ROE) {
      roe_calc_deriv(s, deriv, dt);
    } else if (s->d.solver == HYDRO_SOLVER_BURGERS) {
      /* src/Solver/EulerBurgers.cpp:70-78 */
      burgers_calc_interface_velocity(s);
      burgers_calc_flux(s, dt);
      calc_flux_deriv(s, deriv);
    } else {
      hll_deriv(s, deriv);
    }
  } else if (cb == CB_BURGERS_MOMENTUM) {
    /* src/Solver/EulerBurgers.cpp:85-91 */
    burgers_compute_pressure(s);
    burgers_diffuse_momentum(s, deriv);
  } else if (cb == CB_BURGERS_WORK) {
    /* src/Solver/EulerBurgers.cpp:94-98: computePressure not called again */
    burgers_diffuse_work(s, deriv);
  } else if (cb == CB_GRAVITY) {
    /* src/Solver/SelfGravitation.cpp:111-123 */
    if (s->d.use_gravity) gravity_relax(s, s->d.gauss_seidel_max_iter);
    calc_gravity_deriv(s, deriv);
    oracle_boundary(s);
  } else {
    /* res/MaxwellRoe.cl:159-188: returns before touching deriv */
  }
}

/* multAdd (res/Common.cl:201-210): result = a + b * c */
static void mult_add(double* result, const double* a, const double* b, double c, size_t n) {
  for (size_t i = 0; i < n; ++i) result[i] = a[i] + b[i] * c;
}

/* RungeKutta<Tableau>::integrate (include/HydroGPU/Integrator/RungeKutta.h:59-125)
 * and ForwardEuler::integrate (src/Integrator/ForwardEuler.cpp:19-32) */
static void integrate(oracle_sim* s, double dt, callback_t cb) {
  const size_t len = (size_t)s->N * s->vol;
  if (s->d.integrator == HYDRO_INTEGRATOR_FORWARDEULER) {
    memset(s->fe_deriv, 0, len * sizeof(double));
    run_callback(s, cb, s->fe_deriv, dt);
    mult_add(s->state, s->state, s->fe_deriv, dt, len);
    return;
  }
  const int order = s->order;
  int needed = 0;
  for (int m = 0; m < order; ++m) needed |= s->alpha[m][0] != 0;
  if (needed) memcpy(s->rk_state[0], s->state, len * sizeof(double));
  needed = 0;
  for (int m = 0; m < order; ++m) needed |= s->beta[m][0] != 0;
  if (needed) {
    memset(s->rk_deriv[0], 0, len * sizeof(double));
    run_callback(s, cb, s->rk_deriv[0], dt);
  }
  for (int i = 1; i <= order; ++i) {
    memset(s->state, 0, len * sizeof(double));
    for (int k = 0; k < i; ++k) {
      if (s->alpha[i - 1][k] != 0) mult_add(s->state, s->state, s->rk_state[k], s->alpha[i - 1][k], len);
      if (s->beta[i - 1][k] != 0) mult_add(s->state, s->state, s->rk_deriv[k], s->beta[i - 1][k] * dt, len);
    }
    if (i < order) {
      needed = 0;
      for (int m = i; m < order; ++m) needed |= s->alpha[m][i] != 0;
      if (needed) memcpy(s->rk_state[i], s->state, len * sizeof(double));
      needed = 0;
      for (int m = i; m < order; ++m) needed |= s->beta[m][i] != 0;
      if (needed) {
        memset(s->rk_deriv[i], 0, len * sizeof(double));
        run_callback(s, cb, s->rk_deriv[i], dt);
      }
    }
  }
}

/* ------------------------------------------------ solver driver */

/* Solver::initStep: HLL freezes its wave speeds (src/Solver/HLL.cpp:225-227);
 * Roe has nothing to do before calcTimestep. */
void oracle_init_step(oracle_sim* s) {
  if (s->d.solver == HYDRO_SOLVER_HLL || s->d.solver == HYDRO_SOLVER_HLLC) hll_calc_eigenvalues(s);
}

double oracle_calc_timestep(oracle_sim* s) {
  if (s->d.solver == HYDRO_SOLVER_ROE) {
    /* Roe::calcTimestep (src/Solver/Roe.cpp:88-92) */
    calc_eigen_basis(s);
    roe_calc_cell_timestep(s);
  } else if (s->d.solver == HYDRO_SOLVER_BURGERS) {
    /* EulerBurgers::calcTimestep (src/Solver/EulerBurgers.cpp:61-65) */
    burgers_calc_cell_timestep(s);
  } else {
    /* HLL::calcTimestep (src/Solver/HLL.cpp:229-232); eigenvalues from initStep */
    hll_calc_cell_timestep(s);
  }
  return find_min_timestep(s);
}

void oracle_step(oracle_sim* s, double dt) {
  if (s->d.solver == HYDRO_SOLVER_BURGERS) {
    /* EulerBurgers::step (src/Solver/EulerBurgers.cpp:67-100): advection,
       boundary, the zero-potential pass, momentum diffusion, boundary, work
       diffusion, boundary */
    integrate(s, dt, CB_HYDRO);
    oracle_boundary(s);
    if ((s->d.flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS) || s->pot_on) integrate(s, dt, CB_GRAVITY);
    integrate(s, dt, CB_BURGERS_MOMENTUM);
    oracle_boundary(s);
    integrate(s, dt, CB_BURGERS_WORK);
    oracle_boundary(s);
    return;
  }
  integrate(s, dt, CB_HYDRO);
  if (s->d.equation == HYDRO_EQN_EULER && ((s->d.flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS) || s->pot_on)) {
    /* EulerRoe::step / EulerHLL::step -> SelfGravitation::applyPotential */
    integrate(s, dt, CB_GRAVITY);
  } else if (s->d.equation == HYDRO_EQN_MAXWELL && (s->d.flags & HYDRO_FLAG_GRAVITY_IDENTITY_PASS)) {
    integrate(s, dt, CB_MAXWELL_SOURCE); /* MaxwellRoe::step */
  } else if (s->d.equation == HYDRO_EQN_SRHD) {
    srhd_constrain_state(s); /* SRHDRoe::step (src/Solver/SRHDRoe.cpp:62-73) */
    srhd_update_primitives(s);
  }
}

/* Solver::update (src/Solver/Solver.cpp:521-545) */
double oracle_update(oracle_sim* s) {
  oracle_boundary(s);
  /* HYDRO_FLAG_CLIP_TO_TEND (this build's own flag): once t_end is reached an
     update only refreshes the ghost cells; the clock and the step count stay */
  if ((s->d.flags & HYDRO_FLAG_CLIP_TO_TEND) && !(s->t < s->d.t_end)) return 0.0;
  oracle_init_step(s); /* Solver::initStep */
  double dt;
  if (s->d.use_fixed_dt) {
    dt = s->d.fixed_dt;
    if (s->d.solver == HYDRO_SOLVER_ROE && (s->opts & ORACLE_OPT_FRESH_BASIS_FIXED_DT)) calc_eigen_basis(s);
  } else {
    dt = oracle_calc_timestep(s);
  }
  if ((s->d.flags & HYDRO_FLAG_CLIP_TO_TEND) && s->t + dt > s->d.t_end) dt = s->d.t_end - s->t;
  oracle_step(s, dt);
  s->t += dt;
  s->steps++;
  return dt;
}
