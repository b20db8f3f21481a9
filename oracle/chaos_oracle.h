/*
 * chaos_oracle.h -- CPU restatement of the ChaosLibrary per-point hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in chaoslibrary_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links or calls
 * it.  Every function cites the reference file:line it restates (paths are
 * relative to the reference checkout).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, explicit fma() where the
 * kernels use fma, no fast-math).  The reference itself cannot be built here
 * (CUDA-only sources, <conio.h>, curand), so there is no oracle/_ref.
 *
 * Parity pins (which test pins which function) are listed per function below
 * as "pinned by: ...".  A function marked "parity unpinned" has no check
 * beyond its own restatement.
 */
#ifndef CHAOS_ORACLE_H
#define CHAOS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* System ids (same numbering as include/chaoslibrary.h). */
enum {
    ORC_SYS_CHAMELEON   = 0, /* systems/systems.cuh:11-39   */
    ORC_SYS_ROSSLER_CD  = 1, /* systems/systems.cuh:41-63   */
    ORC_SYS_BASINS      = 2, /* systems/systems.cuh:67-83   */
    ORC_SYS_BASINS_2    = 3, /* systems/systems.cuh:85-101  */
    ORC_SYS_DAMIR       = 4, /* systems/systems.cuh:103-178 */
    ORC_SYS_ROSSLER_RK4 = 5, /* new: classic RK4 of x'=-y-z, y'=x+a y, z'=b+z(x-c) */
    ORC_SYS_LORENZ_RK4  = 6, /* new: classic RK4 of x'=s(y-x), y'=x(r-z)-y, z'=xy-b z */
    ORC_SYS_COUNT       = 7
};

/* Semantics: the reference's quirks, or the benchmark-config semantics. */
enum { ORC_SEM_REFERENCE = 0, ORC_SEM_CONFIG = 1 };

/* Divergence / fixed-point check profiles (see orc_loop). */
enum { ORC_CHK_BIF_REF = 0, ORC_CHK_BASINS_REF = 1, ORC_CHK_CONFIG = 2 };

/* Portable sin / cos / exp / atan the reference systems' steps use (the
 * kernels evaluate the same operations; see chaos_oracle.c). */
double orc_sin(double x);
double orc_cos(double x);
double orc_exp(double x);
double orc_atan(double x);

int  orc_dims(int sys, int* nx, int* na);
void orc_step(int sys, double* X, const double* a, double h);
void orc_steps(int sys, double* X, const double* a, double h, long long n);
/* Batched: states[n][nx] advanced nsteps each with params[n][na] (OpenMP). */
void orc_integrate(int sys, int n, const double* params, double* states,
                   long long nsteps, double h);

double orc_grid_value(int k, int n, double start, double end);
double orc_grid_value_pow(int idx, int n, double start, double end, int dim);

/* The divergence test of a record loop (see orc_loop): 1 if v counts as
 * diverged under the check profile and maxValue. */
int orc_diverged(double v, double maxValue, int check);

int orc_loop(int sys, double* X, const double* a, double h, int amount,
             int preScaller, int writableVar, double maxValue, int check,
             double* data);

int orc_peak_finder(const double* data, int amountOfPoints, double* outPeaks,
                    double* timeOfPeaks, double h);

/* 1-D bifurcation: per point p the output lines are counts[p] (or -1 for a
 * fixed point, written as one line "p, first[p], 0").  peaks/intervals are
 * concatenated in point order; returns total kept peaks or -1 if capacity
 * is too small.  */
long long orc_bif1d(int sys, int sem, double tMax, int nPts, double h,
                    const double* ic, int nx, const double* ranges,
                    const int* idx, int writableVar, double maxValue,
                    double transientTime, const double* values, int na,
                    int preScaller, int* counts, double* first,
                    double* peaks, double* intervals, long long capacity);

/* 2-D bifurcation map out[nPts*nPts] (row = 2nd swept parameter). */
int orc_bif2d(int sys, int sem, double tMax, int nPts, double h,
              const double* ic, int nx, const double* ranges, const int* idx,
              int writableVar, double maxValue, double transientTime,
              const double* values, int na, int preScaller, double eps,
              int* out, int row_begin, int row_end);

int orc_dbscan_ref(double* data, const double* intervals, double* helpful,
                   int amountOfPeaks, int sizeOfHelpful, double eps);
int orc_dbscan1d_count(const double* values, int n, double eps);

/* LLE map: out[j*nPts + i] = LLE(A_i, B_j) for rows j in [row_begin,row_end) */
uint64_t orc_splitmix64(uint64_t* state);
void orc_unit_vector(uint64_t seed, long long point, int nx, double* v);
double orc_lle_point(int sys, int sem, double tMax, double NT, double h,
                     double eps, const double* ic, int nx, const double* a,
                     double transientTime, uint64_t seed, long long point);
int orc_lle2d(int sys, int sem, double tMax, double NT, int nPts, double h,
              double eps, const double* ic, int nx, const double* ranges,
              const int* idx, double transientTime, const double* values,
              int na, uint64_t seed, double* out, int row_begin, int row_end);

/* Basins: per-IC features and flags. */
int orc_basins_features(int sys, int sem, double tMax, int nPts, double h,
                        const double* ic, int nx, const double* ranges,
                        const int* idx, int writableVar, double maxValue,
                        double transientTime, const double* values, int na,
                        int preScaller, double* f0, double* f1, double* f2,
                        int* flags, int point_begin, int point_end);
int orc_basins_cluster_ref(const double* f0, const double* f1,
                           const int* flags, int n, double eps, int* labels);
double orc_det_sum(const double* x, const int* mask, long long n);
int orc_zscore(double* f0, double* f1, double* f2, const int* flags, long long n,
               double* moments /* [6] mean0,std0,mean1,std1,mean2,std2 */);
int orc_dbscan_bruteforce(const double* pts, int n, int dim, double eps,
                          int minPts, const int* include, int* labels);
/* Same result as orc_dbscan_bruteforce with grid-binned neighbour searches
 * (for millions of points). */
int orc_dbscan_grid(const double* pts, int n, int dim, double eps,
                    int minPts, const int* include, int* labels);

/* Legacy entry points (debri/, namespace old_library). */
double orc_legacy_lle_point(int sys, int sem, int form, double tMax, double NT, double h,
                            double eps, const double* X0, int nx, const double* a,
                            double maxValue, double transientTime, uint64_t seed,
                            long long point);
int orc_legacy_lle(int sys, int sem, int form, int dims, double tMax, double NT, int nPts,
                   double h, double eps, const double* ic, int nx, const double* ranges,
                   const int* idx, double maxValue, double transientTime,
                   const double* values, int na, uint64_t seed, double* out,
                   int point_begin, int point_end);
void orc_gram_schmidt(const double* a, double* b, int n, double* den);
int orc_legacy_ls_point(int sys, int sem, double tMax, double NT, double h, double eps,
                        const double* X0, int nx, const double* a, double maxValue,
                        double transientTime, uint64_t seed, long long point, double* out);
int orc_legacy_ls(int sys, int sem, int dims, double tMax, double NT, int nPts, double h,
                  double eps, const double* ic, int nx, const double* ranges, const int* idx,
                  double maxValue, double transientTime, const double* values, int na,
                  uint64_t seed, double* out, int point_begin, int point_end);
long long orc_time_domain(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                          const double* ranges, const int* idx, int writableVar, double maxValue,
                          double transientTime, const double* values, int na, int preScaller,
                          double* out, int* flags);

long long orc_legacy_bif1d_ic(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                              const double* ranges, const int* idx, int writableVar, double maxValue,
                              double transientTime, const double* values, int na, int preScaller,
                              int* counts, double* peaks, long long capacity);
int orc_legacy_bif2d_ic(int sys, int sem, double tMax, int nPts, double h, const double* ic, int nx,
                        const double* ranges, const int* idx, int writableVar, double maxValue,
                        double transientTime, const double* values, int na, int preScaller, double eps,
                        int* out);

/* CPU baseline: LLE rows over OpenMP threads; returns threads used. */
int orc_lle2d_omp(int sys, int sem, double tMax, double NT, int nPts, double h,
                  double eps, const double* ic, int nx, const double* ranges,
                  const int* idx, double transientTime, const double* values,
                  int na, uint64_t seed, double* out, int row_begin,
                  int row_end, int row_stride, int threads);

#ifdef __cplusplus
}
#endif
#endif
