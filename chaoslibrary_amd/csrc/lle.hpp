// lle.hpp -- launch arguments of the LLE map kernels.
#pragma once

#include "runtime.hpp"
#include "systems.hpp"

namespace chaos {

struct LleArgs {
    double values[8];
    double ic[4];
    double rA0, rA1, rB0, rB1;
    int idxA, idxB;
    int nPts;
    int S;      // transient steps (one trajectory)
    int NTp;    // steps per renormalisation window
    int W;      // windows
    long long row0;       // first B row of this launch
    long long n_points;   // lanes in this launch
    double eps, tMax;
    StepConst sc;
    unsigned long long seed;
};

LleArgs make_lle_args(const chaos_options& o, double tMax, double NT, int nPts, double h, double eps,
                      const double* ic, int nx, const double* ranges, const int* idx,
                      double transientTime, const double* values, int na);
double lle_steps_per_point(const LleArgs& p);
void launch_lle_rows(CallCtx& ctx, int system, LleArgs p, long long row0, long long rows, double* d_out);

class TextFile;
// The file LLE2D writes for an nPts x nPts map v: CSV into f (already open,
// or NULL) and/or the binary sidecar <path>.bin, as output (CHAOS_OUT_*) asks.
void write_lle_file(TextFile* f, const char* path, const double* ranges, int nPts, const double* v, int output);

}  // namespace chaos
