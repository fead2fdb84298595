/*
 * mph_oracle.h -- CPU restatement of the reference MPH-I time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and the
 * CPU baseline of bench.py.  The product (libmph_hip.so, mphimplicit_amd) never links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * It restates source/main.cpp of Masahiro-Kondo-AIST/MphImplicit in plain C++
 * (FP64, same loop and summation order, no FMA contraction), with one switch the
 * reference does not have: the summation order of the conjugate-residual dot
 * products (MPHO_DOT_*).  With MPHO_DOT_SERIAL it is bit-identical to the serial
 * reference build; with MPHO_DOT_TREE it uses the fixed tree order of the HIP
 * solver, so the GPU path can be compared with it bit for bit.
 */
#ifndef MPH_ORACLE_H
#define MPH_ORACLE_H

#include "../include/mph_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPHO_DOT_SERIAL 0 /* sum += x[i]*y[i], i ascending (myDdot, main.cpp:4064-4075)  */
#define MPHO_DOT_TREE 1   /* the HIP solver's two-level tree                              */
#define MPHO_DOT_OMP 2    /* OpenMP reduction(+) like the Mph_gcc builds (not repeatable) */

typedef struct mpho_scene mpho_scene;

/* initialise + the init sequence of source/main.cpp:353-375; NULL on failure */
mpho_scene* mpho_create(const mph_params* p, const mph_grid* g, int dot_mode);
int mpho_step(mpho_scene* s, int nsteps, mph_step_stats* st);
int mpho_run_phase(mpho_scene* s, int phase);
int mpho_field_size(mpho_scene* s, const char* name, size_t* bytes, int* elem_bytes);
int mpho_read(mpho_scene* s, const char* name, void* out, size_t bytes);
int mpho_write(mpho_scene* s, const char* name, const void* in, size_t bytes);
int mpho_scalar(mpho_scene* s, const char* name, double* value);
/* "ConvergenceCheck" (0/1): the reference built with CONVERGENCE_CHECK (main.cpp:46,
 * 4348-4351, 4383-4397).  Each solve then keeps (rr, r.s) of every iteration that
 * did not stop (field ConvergenceHistory) and the true residual's r.r and r.s after
 * the loop (scalars CheckRR, CheckRS). */
int mpho_set_option(mpho_scene* s, const char* name, double value);
/* s = M^-1 q with the current multigrid hierarchy (preconditionWithMultiGrid) */
int mpho_precondition(mpho_scene* s, const double* q, double* out);
void mpho_destroy(mpho_scene* s);
const char* mpho_last_error(void);

/* stand-alone kernels for unit tests */
double mpho_dot(int mode, long n, const double* x, const double* y);
void mpho_bitonic_sort(int power, int* key, int* val); /* in place, 2^power entries */
void mpho_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
