// bifurcation.hpp -- Bifurcation stages shared by the one-shot entry points
// and the batched engine.
#pragma once

#include <vector>

#include "peaks.hpp"
#include "runtime.hpp"
#include "sweep.hpp"

namespace chaos {

struct BifInputs {
    double tMax, h, maxValue, transientTime, eps;
    int nPts, nx, na, writableVar, preScaller, dims;
    int legacy_ic;  // legacy bifurcation1DIC / 2DIC: initial conditions swept, legacy loop semantics
    const double *ic, *ranges, *values;
    const int* idx;
};

// Device results of one Bifurcation sweep (segments of flag-1 points).
struct BifDevice {
    long long n = 0;         // points
    long long total = 0;     // stored peaks of flag-1 points (first peak included)
    double hp = 0;           // interval unit
    int samples = 0;         // recorded samples per point (bounds every peak-time gap); 0: unknown
    bool need_times = true;  // scatter the peak times too (intervals)
    bool need_vals = true;   // scatter the peak values (the REFERENCE 2-D dbscan reads times only)
    bool need_host = true;   // host copies of flags / counts / first peaks (1-D results)
    DevBuf<int> flags, counts;
    DevBuf<double> first;
    DevBuf<long long> off;   // segment offsets
    DevBuf<double> vals;
    DevBuf<unsigned> times;
    std::vector<int> h_flags, h_counts;
    std::vector<double> h_first;
};

// Chunk pool backing a PeakPool view.
struct PeakPoolBuf {
    DevBuf<PeakEntry> entries;
    DevBuf<int> meta;  // PeakPool::meta
    long long nchunks = 0;
    long long cap = 0;  // chunks the buffers hold
    void alloc(hipStream_t st, long long chunks);
    // A store of `chunks` for the next sweep: the buffers of an earlier one
    // when they are large enough (a grid swept in ranges allocates its store
    // once), else new ones.
    void prepare(hipStream_t st, long long chunks);
    PeakPool view() const;
};

// Validated sweep arguments (outputs and pool left unset).
SweepArgs make_bif_args(const chaos_options& o, const BifInputs& in);
// Initial pool size for n points (overflow is detected and retried).
long long bif_initial_chunks(const SweepArgs& A);
// Exact pool need from per-point counts (each chunk switch wastes < 64 slots).
long long bif_exact_chunks(const std::vector<int>& counts, long long n);
// Reads flags/counts/first back, builds segment offsets and scatters the
// pooled peaks into per-point segments.  Returns false on pool overflow.
bool bif_post(CallCtx& ctx, BifDevice& R, PeakPoolBuf& pool);
// Per-point 2-D analysis (REFERENCE dbscan / CONFIG 1-D DBSCAN count).
void bif2d_analysis(CallCtx& ctx, int semantics, BifDevice& R, double eps, int* d_out);
// Host assembly of the 1-D outputs (first peak dropped, intervals).
void bif1d_collect(CallCtx& ctx, BifDevice& R, int* counts, double* first, double* peaks, double* intervals,
                   long long capacity, long long* total);

class TextFile;
// The bifurcation2D map file (two range lines, then the int map): CSV into f
// (already open, or NULL) and/or the binary sidecar <path>.bin.
void write_int_map_file(TextFile* f, const char* path, const double* ranges, int nPts, const int* v, int output);
// The bifurcation1D body after its range line: per point "p, x_first, 0" for
// flag -1, else "p, peak, interval" per kept peak, at precision 12.
void write_bif1d_body(TextFile& f, const double* ranges, int nPts, const int* counts, const double* first,
                      const double* peaks, const double* intervals);

}  // namespace chaos
