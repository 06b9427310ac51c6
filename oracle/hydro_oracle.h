/*
 * hydro_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference's update path (KennethYeah/HydrodynamicsGPU
 * res/ *.cl kernels + the Solver/Integrator dispatch order).  It is the checker
 * the HIP path is compared against; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links it.
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#include <stddef.h>
#include "../include/hydro_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* options beyond the descriptor */
#define ORACLE_OPT_FRESH_BASIS_FIXED_DT 0x1u /* with use_fixed_dt, compute the Roe
   eigenbasis from the start-of-step state (the reference leaves it stale,
   src/Solver/Roe.cpp:88-92 is skipped by src/Solver/Solver.cpp:528) */

oracle_sim* oracle_create(const hydro_sim_desc* d, unsigned opts);
void oracle_destroy(oracle_sim* s);
int oracle_num_states(const oracle_sim* s);
size_t oracle_volume(const oracle_sim* s);
int oracle_set_state(oracle_sim* s, const double* aos, size_t n);
int oracle_get_state(const oracle_sim* s, double* aos, size_t n);
int oracle_get_primitives(const oracle_sim* s, double* aos, size_t n);
int oracle_set_potential(oracle_sim* s, const double* pot, size_t n);
int oracle_get_potential(const oracle_sim* s, double* pot, size_t n);
/* solidBuffer (n = volume, nonzero = solid); NULL clears it.  Read by the
 * EulerRoe and EulerBurgers updates only. */
int oracle_set_solid(oracle_sim* s, const unsigned char* solid, size_t n);
void oracle_boundary(oracle_sim* s);
void oracle_init_step(oracle_sim* s);
double oracle_calc_timestep(oracle_sim* s);
void oracle_step(oracle_sim* s, double dt);
double oracle_update(oracle_sim* s);
double oracle_time(const oracle_sim* s);
void oracle_set_threads(int n);

/* kernel-level entry points used by the unit tests */
double oracle_slope_limiter(int limiter, double r);
double oracle_kernel_dx(double dx, int six_digits);
int oracle_rk_tableau(int integrator, int* order, double* alphas /*4x4*/, double* betas /*4x4*/);
/* Euler Roe eigenbasis of one interface; qL, qR have N=dim+2 entries;
 * outputs eigval[N], evl[N*N], evr[N*N] (column-major m[i + N*j]) */
void oracle_euler_eigenbasis(int dim, int side, double gamma, const double* qL, const double* qR,
                             double potL, double potR, double* eigval, double* evl, double* evr);
/* the same with the two cells' solid flags (the wall mirror) */
void oracle_euler_eigenbasis_solid(int dim, int side, double gamma, const double* qL, const double* qR,
                                   double potL, double potR, int solidL, int solidR, double* eigval,
                                   double* evl, double* evr);
/* SRHD eigenbasis of one interface from primitives (rho, v[dim], eInt) */
void oracle_srhd_eigenbasis(int dim, int side, double gamma, const double* primL, const double* primR,
                            double* eigval, double* evl, double* evr);
/* HLL / HLLC one-interface flux (no limiter), states rotated as in the kernels;
 * sl, sr are the frozen wave speeds written by calcEigenvalues */
void oracle_euler_hll_speeds(int solver, int dim, int side, double gamma, const double* qL,
                             const double* qR, double* eigval);
/* returns 0 when no branch of the kernel is taken (NaN speeds): flux untouched */
int oracle_euler_hll_flux(int solver, int dim, int side, double gamma, const double* qL,
                          const double* qR, const double* eigval, double* flux);
/* the same with the cells' potentials (EulerHLL; HLLC takes zeros) */
void oracle_euler_hll_speeds_pot(int solver, int dim, int side, double gamma, const double* qL,
                                 const double* qR, double potL, double potR, double* eigval);
int oracle_euler_hll_flux_pot(int solver, int dim, int side, double gamma, const double* qL,
                              const double* qR, double potL, double potR, const double* eigval, double* flux);
/* SRHD con2prim of one cell; returns iterations used or -1 if not converged
 * (prim keeps the initial guess then) */
int oracle_srhd_con2prim(const hydro_sim_desc* d, int dim, const double* cons, double* prim);
void oracle_srhd_prim2cons(int dim, double gamma, const double* newtonian, double* cons, double* prim);

#ifdef __cplusplus
}
#endif

#endif
