/*
 * mhit_oracle.h -- CPU FP64 restatement of the MHIT36 timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the GPU library is compared
 * with; it is never linked into, loaded by, or called from the product
 * library (libmhit36.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * It follows multi/main.f90:209-250, 322-397, 440-1007 and
 * multi/particles.f90:78-122, 364-380 of aroccon/MHIT36_mixedp at one rank,
 * with halos replaced by periodic index wrap.  Parameters use mhit_params
 * from include/mhit36.h.
 */
#ifndef MHIT_ORACLE_H
#define MHIT_ORACLE_H

#include <stdint.h>
#include "../include/mhit36.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

oracle_ctx* oracle_create(const mhit_params* p);
void   oracle_destroy(oracle_ctx* c);
int    oracle_set_field(oracle_ctx* c, int fieldn, const double* in);
int    oracle_get_field(oracle_ctx* c, int fieldn, double* out);
int    oracle_set_tracers(oracle_ctx* c, int64_t n, const double* x,
                          const double* y, const double* z);
int    oracle_get_tracers(oracle_ctx* c, double* x, double* y, double* z,
                          double* up, double* vp, double* wp);
int    oracle_set_ab2_state(oracle_ctx* c, double alpha, double beta,
                            double gumax);
int    oracle_step(oracle_ctx* c, int nsteps);
int    oracle_diag(oracle_ctx* c, mhit_diag* d);
int    oracle_set_threads(int nthreads);

/* sub-operators on the current state (for pinning each one separately)   */
#define ORACLE_NS_CONV  1
#define ORACLE_NS_VISC  2
#define ORACLE_NS_FORCE 4
int    oracle_ns_rhs(oracle_ctx* c, int parts, double* ru, double* rv,
                     double* rw);
#define ORACLE_PHI_CONV    1
#define ORACLE_PHI_DIFF    2
#define ORACLE_PHI_SHARPEN 4
int    oracle_acdi_rhs(oracle_ctx* c, int parts, double gamma, double* rphi);
int    oracle_csf(oracle_ctx* c, double* fx, double* fy, double* fz);
int    oracle_divergence(oracle_ctx* c, double* rhsp);
int    oracle_interp(oracle_ctx* c, double* up, double* vp, double* wp);
int    oracle_poisson(oracle_ctx* c, const double* rhs, double* p);

/* Drop statistics of post/unionfind_periodic/main.f90 (mask phi >= thr,
 * 6- or 26-point connectivity with periodic wrap) plus d_eq of
 * post/DropCountSize/calculate_deq.f90.  Drops are numbered in the order of
 * their first cell (z, y, x scan); each writes 10 values to out: cells,
 * first cell, centre x y z, mean u v w, volume, d_eq.  u, v, w may be NULL
 * (velocities then 0).  Returns the number of drops in *nd.               */
int    oracle_drop_stats(int n, const double* phi, const double* u,
                         const double* v, const double* w, double thr,
                         int conn, int64_t maxd, double* out, int64_t* nd);

/* FFT building blocks: complex interleaved (re,im), unnormalised,
 * sign = -1 forward (e^{-i}), +1 inverse.                                  */
void   oracle_fft1d(int n, const double* in, double* out, int sign);
void   oracle_fft3d(int n, const double* in, double* out, int sign);
void   oracle_fft1d_f32(int n, const float* in, float* out, int sign);

#ifdef __cplusplus
}
#endif
#endif
