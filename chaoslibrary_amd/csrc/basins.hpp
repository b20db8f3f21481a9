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
// the fixed-point sign rule; CONFIG z-score (on a copy) + 3-D DBSCAN.
void basins_cluster(CallCtx& ctx, bool ref, long long n, const double* f0, const double* f1, const double* f2,
                    const int* flags, double eps, int minPts, int* labels);

// The files basinsOfAttraction_2 writes: OUT (labels), OUT_1.csv / OUT_2.csv
// (features 0, 1), OUT_3.csv (flags) and, for CONFIG, OUT_4.csv (feature 2).
// The headers go first (the reference writes them before computing); the
// bodies append.  feat holds nfeat planes of nPts^2 doubles.
void write_basins_headers(const std::string& base, const double* ranges, int nfeat, int output);
void write_basins_bodies(const std::string& base, const double* ranges, int nPts, int nfeat, const int* labels,
                         const double* feat, const int* flags, int output);

}  // namespace chaos
