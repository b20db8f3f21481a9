// basins.hpp -- Basins stages shared by the one-shot entry points and the
// batched engine.
#pragma once

#include <string>

#include "runtime.hpp"
#include "sweep.hpp"

namespace chaos {

// Validated sweep arguments for an IC grid (outputs left unset).
SweepArgs make_basins_args(const chaos_options& o, double tMax, int nPts, double h, int nx, const double* ic,
                           const double* ranges, const int* idx, int writableVar, double maxValue,
                           double transientTime, const double* values, int na, int preScaller, double eps);

// Labels from the per-IC features: REFERENCE single linkage on (f0, f1) with
// the fixed-point sign rule; CONFIG 3-D DBSCAN of the z-scored features
// (formed on each read; the raw features stay as they are).
// zbuf (CONFIG only): NULL, or device memory of basins_zbuf_bytes(n) into
// which the sweep wrote its blocks' first moments (basins_sweep_partials).
void basins_cluster(CallCtx& ctx, bool ref, long long n, const double* f0, const double* f1, const double* f2,
                    const int* flags, double eps, int minPts, int* labels, void* zbuf = nullptr);
size_t basins_zbuf_bytes(long long n);
// Points the CONFIG sweep's moment outputs (zsum, zcnt, zmm, zb) into zbuf.
void basins_sweep_partials(SweepArgs& A, void* zbuf);

// The files basinsOfAttraction_2 writes: OUT (labels), OUT_1.csv / OUT_2.csv
// (features 0, 1), OUT_3.csv (flags) and, for CONFIG, OUT_4.csv (feature 2).
// The headers go first (the reference writes them before computing); the
// bodies append.  feat holds nfeat planes of nPts^2 doubles.
void write_basins_headers(const std::string& base, const double* ranges, int nfeat, int output);
void write_basins_bodies(const std::string& base, const double* ranges, int nPts, int nfeat, const int* labels,
                         const double* feat, const int* flags, int output);

}  // namespace chaos
