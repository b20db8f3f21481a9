// sweep.hpp -- arguments of the transient+record sweep kernels that feed the
// Bifurcation and Basins analyses.
#pragma once

#include "peaks.hpp"
#include "runtime.hpp"
#include "systems.hpp"

namespace chaos {

enum SweepKind : int {
    KIND_BIF = 0,          // store every peak (value, time) -> Bifurcation 1D/2D
    KIND_BASINS_REF = 1,   // mean peak / mean interval (+ sentinels)
    KIND_BASINS_CFG = 2,   // mean X[0], mean X[NX-1], peak std
    KIND_BIF_VALS = 3      // store every peak's value (times left 0) -> the CONFIG peak-count map
};

// Kinds that store peaks in the pool.
constexpr bool stores_peaks(int kind) { return kind == KIND_BIF || kind == KIND_BIF_VALS; }

enum CheckMode : int {
    CHK_BIF_REF = 0,     // v!=v || v==v+1 ; (double)fabsf((float)v) > max ; sqrtf(fabsf((float)t)) < 1e-9
    CHK_BASINS_REF = 1,  // isnan||isinf     ; (double)fabsf((float)v) > max ; sqrt(fabs(t)) < 1e-9
    CHK_CONFIG = 2       // !isfinite        ; fabs(v) > max                 ; sqrt(t) < 1e-9
};

struct SweepArgs {
    double values[8];
    double ic[4];
    double r[4];          // ranges (dim 0: r[0..1], dim 1: r[2..3])
    int idx0, idx1;       // swept index into values (Bifurcation) or X (Basins); idx1 < 0 for 1-D
    int nPts;             // grid side
    int dims;             // 1 or 2
    long long n_points;   // points of this launch: nPts or nPts^2, or a range of them
    long long p0;         // grid index of the launch's point 0 (a multiple of 64; outputs are indexed from it)
    int S;                // transient steps
    int B;                // recorded samples
    int pre;              // steps per sample (Basins; Bifurcation records every step)
    int check;            // CheckMode
    double maxValue;      // divergence bound (0 = off)
    double L;             // |v| < L guarantees "not diverged" (fast path)
    double hp;            // interval unit: h * preScaller
    int ic_sweep;         // Bifurcation: swept entries are initial conditions (legacy *IC sweeps)
    int legacy_transient; // Bifurcation: transient checked on X[0] for NaN/inf only, a fixed point still records
    StepConst sc;
    // outputs (per point)
    int* flags;
    int* counts;          // peaks found (including the first, later dropped)
    double* first;        // first recorded sample
    double* f0;           // features (Basins)
    double* f1;
    double* f2;
    PeakPool pool;        // Bifurcation peak store
    // Basins CONFIG, one-shot sweep: first moments of the masked features
    // (flag != 0) per 256-point block for the z-score (zsum [3][zb], zcnt
    // [zb], zmm [6][zb], zb = the launch's blocks), or zsum NULL
    double* zsum;
    int* zcnt;
    unsigned long long* zmm;
    long long zb;
};

// Fast-path bound for the divergence test of each check mode.
double divergence_fast_bound(int check, double maxValue);

void launch_sweep(CallCtx& ctx, int system, int writableVar, int kind, const SweepArgs& A);

}  // namespace chaos
