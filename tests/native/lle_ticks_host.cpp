// Host build of the LLE lane program (chaoslibrary_amd/csrc/lanes.hpp
// LleProg), for the CPU tests of tick-split invariance and of the progress
// guard.  Compiled by tests/test_lle_ticks_cpu.py with hipcc --cuda-host-only;
// the program's methods are __host__ __device__, so this runs the same source
// the kernels run.
#include <vector>

#include "lanes.hpp"

using namespace chaos;

namespace {

LleArgs make_args(const double* values, const double* ic, const double* ranges, const int* idx, int nPts, int S,
                  int NTp, int W, double eps, double tMax, double h, unsigned long long seed)
{
    LleArgs p{};
    for (int k = 0; k < 3; ++k) {
        p.values[k] = values[k];
        p.ic[k] = ic[k];
    }
    p.rA0 = ranges[0];
    p.rA1 = ranges[1];
    p.rB0 = ranges[2];
    p.rB1 = ranges[3];
    p.idxA = idx[0];
    p.idxB = idx[1];
    p.nPts = nPts;
    p.S = S;
    p.NTp = NTp;
    p.W = W;
    p.eps = eps;
    p.tMax = tMax;
    p.sc = make_step_const(h);
    p.seed = seed;
    return p;
}

// Point (i, j) with its ticks split at cuts[] (ascending, inside (0, total)):
// the lane state is parked and reloaded between segments as the engine does.
template <int SYS>
double run_split(const LleArgs& p, int i, int j, const long long* cuts, int ncuts, bool dir_stored, int* status)
{
    using P = LleProg<SYS>;
    std::vector<double> mem((size_t)(P::ND + P::NDIR), 0.0);
    const LaneStore st{mem.data(), nullptr, 1};
    const long long point = (long long)j * p.nPts + i;
    if (dir_stored) {
        double d[P::NX];
        P::offset(p, point, d);
        for (int k = 0; k < P::NX; ++k) mem[P::ND + k] = d[k];
    }
    const long long total = P::total(p);
    P L;
    L.params(p, i, j);
    L.init(p);
    long long g = 0;
    *status = 0;
    for (int c = 0; c <= ncuts; ++c) {
        const long long e = c < ncuts ? cuts[c] : total;
        if (g > 0) L.load(st, 0);
        const bool ok = dir_stored ? L.template ticks<true>(p, point, g, e, P::dir_of(st, 0), 1)
                                   : L.template ticks<false>(p, point, g, e);
        if (!ok) *status = -1;
        L.store(st, 0);
        g = e;
    }
    return L.result(p);
}

}  // namespace

extern "C" double host_lle_split(int sys, const double* values, const double* ic, const double* ranges,
                                 const int* idx, int nPts, int S, int NTp, int W, double eps, double tMax, double h,
                                 unsigned long long seed, int i, int j, const long long* cuts, int ncuts,
                                 int dir_stored, int* status)
{
    const LleArgs p = make_args(values, ic, ranges, idx, nPts, S, NTp, W, eps, tMax, h, seed);
    if (sys == SYS_LORENZ_RK4) return run_split<SYS_LORENZ_RK4>(p, i, j, cuts, ncuts, dir_stored != 0, status);
    return run_split<SYS_ROSSLER_RK4>(p, i, j, cuts, ncuts, dir_stored != 0, status);
}

// The tick call a kernel makes with an LLE block that was never filled in
// (all zero, NTp = 0) or with a reversed range: returns the guard's verdict
// (1 = advanced, 0 = refused) and the ticks the window loop would take per
// trip with these arguments (<= 0 means an unguarded loop never ends).
extern "C" int host_lle_guard(int NTp, int S, long long g0, long long g1, long long* trip)
{
    LleArgs p{};
    p.NTp = NTp;
    p.S = S;
    p.eps = 1e-8;
    p.sc = make_step_const(0.01);
    LleProg<SYS_ROSSLER_RK4> L;
    L.params(p, 0, 0);
    L.init(p);
    // what the unguarded window loop would advance in its first trip
    const long long left = g1 - g0;
    const int local = (NTp > 0 && g0 >= S) ? (int)((g0 - S) % NTp) : 0;
    *trip = (NTp - local) < left ? (NTp - local) : left;
    const double d[3] = {1e-8, 0.0, 0.0};
    return L.template ticks<true>(p, 0, g0, g1, d, 1) ? 1 : 0;
}
