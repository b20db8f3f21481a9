// lanes.hpp -- resumable per-lane programs shared by the one-shot kernels and
// the batched engine.
//
// Every program advances one grid point over a range of "ticks" [g0, g1) of
// its job and can stop anywhere: the one-shot kernels run [0, total) in one
// launch, the engine runs a few thousand ticks per launch and parks the lane
// state in HBM in between.  Both paths execute the same arithmetic in the
// same order, so a point's result does not depend on how its work was split.
#pragma once

#include <hip/hip_runtime.h>

#include "lle.hpp"
#include "peaks.hpp"
#include "sweep.hpp"
#include "systems.hpp"

namespace chaos {

// ---------------------------------------------------------------------------
// shared checks (Bifurcation/src/bifurcationCUDA.cu:445-466,
// Basins/src/basinsCUDA.cu:350-376)
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool diverged_exact(double v, double M, int check)
{
    if (check == CHK_BIF_REF) {
        if (v != v || v == v + 1.0) return true;
        if (M != 0 && (double)fabsf((float)v) > M) return true;
    } else if (check == CHK_BASINS_REF) {
        if (isnan(v) || isinf(v)) return true;
        if (M != 0 && (double)fabsf((float)v) > M) return true;
    } else {
        if (!isfinite(v)) return true;
        if (M != 0 && fabs(v) > M) return true;
    }
    return false;
}

__device__ __forceinline__ bool fixed_point(double t, int check)
{
    if (check == CHK_BIF_REF) return (double)sqrtf(fabsf((float)t)) < 1e-9;
    if (check == CHK_BASINS_REF) return sqrt(fabs(t)) < 1e-9;
    return sqrt(t) < 1e-9;
}

template <int NX>
__device__ __forceinline__ double step_delta2(const double* X, const double* xp)
{
    double t = 0;
#pragma unroll
    for (int j = 0; j < NX; ++j) t += ((X[j] - xp[j]) * (X[j] - xp[j]));
    return t;
}

// X[wv] for a compile-time wv (WVT >= 0) or a run-time one (WVT < 0).
template <int NX, int WVT>
__device__ __forceinline__ double pick(const double* X, int wv)
{
    if constexpr (WVT >= 0) {
        return X[WVT];
    } else {
        double v = X[0];
#pragma unroll
        for (int k = 1; k < NX; ++k) v = (wv == k) ? X[k] : v;
        return v;
    }
}

// Lane state parked in HBM between engine launches (SoA, k-major).
struct LaneStore {
    double* d;     // [nd][np]
    int* i;        // [ni][np]
    long long np;  // lanes (multiple of 64)
};

// ---------------------------------------------------------------------------
// LLE: Benettin with two trajectories (LLE/src/LLECUDA.cu:56-166)
// ticks [0, S): transient on X; tick S starts with the perturbation; ticks
// [S, S + W*NTp) step X and Xp once each, renormalising after every NTp.
// Host-callable as well, so the tick-split invariance and the progress guard
// are checked on the CPU.
// ---------------------------------------------------------------------------
template <int SYS>
struct LleProg {
    using Sy = System<SYS>;
    static constexpr int NX = Sy::NX, NA = Sy::NA;
    static constexpr int ND = 2 * NX + 1;   // parked state: X, Xp, sum
    static constexpr int NDIR = NX;         // engine: eps * direction, after the parked state
    double X[NX], Xp[NX], a[NA], sum;

    __host__ __device__ __forceinline__ static long long total(const LleArgs& p)
    {
        return (long long)p.S + (long long)p.W * p.NTp;
    }

    __host__ __device__ __forceinline__ void params(const LleArgs& p, int i, int j)
    {
        build_params<NA>(a, p.values, p.idxA, grid_value(i, p.nPts, p.rA0, p.rA1), p.idxB,
                         grid_value(j, p.nPts, p.rB0, p.rB1));
    }
    __host__ __device__ __forceinline__ void init(const LleArgs& p)
    {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            X[k] = p.ic[k];
            Xp[k] = 0.0;
        }
        sum = 0.0;
    }
    // The perturbation added at tick S: eps * v with v the point's seeded
    // unit vector (LLE/src/LLECUDA.cu:88-102 draws it from curand instead).
    __host__ __device__ __forceinline__ static void offset(const LleArgs& p, long long point, double* d)
    {
        double v[NX];
        unit_vector<NX>(p.seed, point, v);
#pragma unroll
        for (int k = 0; k < NX; ++k) d[k] = p.eps * v[k];
    }

    // Advance over ticks [g0, g1).  The perturbation at tick S comes from
    // offset() (DIR_STORED false) or is read from dir[k * dstride] (the
    // engine stores it when the job starts, which keeps the unit-vector code
    // out of its tile loop).  Returns false without stepping when the range
    // cannot make progress -- NTp < 1 would leave every window empty and the
    // window loop would never advance -- so a caller handed a malformed job
    // reports it instead of spinning.
    template <bool DIR_STORED = false>
    __host__ __device__ __forceinline__ bool ticks(const LleArgs& p, long long point, long long g0, long long g1,
                                                   const double* dir = nullptr, long long dstride = 0)
    {
        if (p.NTp < 1 || p.S < 0 || g0 < 0 || g1 < g0) return false;
        long long g = g0;
        if (g < p.S) {
            const long long e = g1 < p.S ? g1 : p.S;
            for (long long s = g; s < e; ++s) Sy::step(X, a, p.sc);
            g = e;
        }
        if (g == p.S && g < g1) {
            double d[NX];
            if (DIR_STORED) {
#pragma unroll
                for (int k = 0; k < NX; ++k) d[k] = dir[k * dstride];
            } else {
                offset(p, point, d);
            }
#pragma unroll
            for (int k = 0; k < NX; ++k) Xp[k] = X[k] + d[k];
        }
        int local = g < g1 ? (int)((g - p.S) % p.NTp) : 0;  // position in the first window
        while (g < g1) {
            const long long left = g1 - g;
            const int n = (p.NTp - local) < left ? (p.NTp - local) : (int)left;  // >= 1: local < NTp
            for (int s = 0; s < n; ++s) {
                Sy::step(X, a, p.sc);
                Sy::step(Xp, a, p.sc);
            }
            g += n;
            const bool window_end = local + n == p.NTp;
            local = 0;
            if (window_end) {
                double d2 = 0.0;
#pragma unroll
                for (int l = 0; l < NX; ++l) {
                    const double diff = X[l] - Xp[l];
                    d2 = d2 + diff * diff;
                }
                const double dist = sqrt(d2) / p.eps;
                sum = sum + log(dist);
#pragma unroll
                for (int l = 0; l < NX; ++l) Xp[l] = X[l] - ((X[l] - Xp[l]) / dist);
            }
        }
        return true;
    }
    __host__ __device__ __forceinline__ double result(const LleArgs& p) const { return sum / p.tMax; }

    __host__ __device__ __forceinline__ void load(const LaneStore& L, long long q)
    {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            X[k] = L.d[k * L.np + q];
            Xp[k] = L.d[(NX + k) * L.np + q];
        }
        sum = L.d[2 * NX * L.np + q];
    }
    __host__ __device__ __forceinline__ void store(const LaneStore& L, long long q) const
    {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            L.d[k * L.np + q] = X[k];
            L.d[(NX + k) * L.np + q] = Xp[k];
        }
        L.d[2 * NX * L.np + q] = sum;
    }
    // Stored perturbation of lane q (engine): dir[k * np] for k < NX.
    __host__ __device__ __forceinline__ static const double* dir_of(const LaneStore& L, long long q)
    {
        return L.d + ND * L.np + q;
    }
};

// ---------------------------------------------------------------------------
// Sweep (Bifurcation / Basins): ticks [0, S) transient, [S, S+B) samples
// (each sample = pre steps).  Kernels and engine must keep all 64 lanes of
// a wave in the loop (append_peaks is a wave-level operation).
// ---------------------------------------------------------------------------
template <int SYS, int KIND>
struct SweepProg {
    using Sy = System<SYS>;
    static constexpr int NX = Sy::NX, NA = Sy::NA;
    static constexpr int ND = NX + 8;  // X, p, dn, first, last, s0, s1, sx, sz
    static constexpr int NI = 4;       // pstart, cnt, prev_t, bits
    double X[NX], a[NA];
    static constexpr bool BIF = stores_peaks(KIND);
    // peak times: the intervals of the bifurcation lists and of the
    // reference basins features; the other kinds use peak values only
    WavePeakDetector<KIND == KIND_BIF || KIND == KIND_BASINS_REF> det;
    int cnt, prev_t;
    unsigned long long alive;  // wave mask of the lanes still integrating
    bool fixed;
    double s0, s1, sx, sz, first, last;

    __device__ __forceinline__ static long long total(const SweepArgs& A) { return (long long)A.S + A.B; }

    __device__ __forceinline__ static void grid(const SweepArgs& A, long long pq, double& g0, double& g1)
    {
        const int i = (int)(pq % A.nPts);
        const int j = A.dims == 2 ? (int)(pq / A.nPts) : 0;
        g0 = grid_value(i, A.nPts, A.r[0], A.r[1]);
        g1 = A.dims == 2 ? grid_value(j, A.nPts, A.r[2], A.r[3]) : 0.0;
    }
    // Parameter vector of point pq.
    __device__ __forceinline__ void params(const SweepArgs& A, long long pq)
    {
        if (BIF && !A.ic_sweep) {
            // swept entries of the parameter vector (bifurcationCUDA.cu:282-284)
            double g0, g1;
            grid(A, pq, g0, g1);
            build_params<NA>(a, A.values, A.idx0, g0, A.idx1, g1);
        } else {
#pragma unroll
            for (int k = 0; k < NA; ++k) a[k] = A.values[k];
        }
    }
    // Parameters and the initial lane state of point pq.
    __device__ __forceinline__ void init(const SweepArgs& A, long long pq, bool in_range)
    {
        params(A, pq);
        if (BIF && !A.ic_sweep) {
#pragma unroll
            for (int k = 0; k < NX; ++k) X[k] = A.ic[k];
        } else {
            // swept entries of the initial condition (basinsCUDA.cu:253-262,
            // debri/cudaLibrary.cu:1022-1062)
            double g0, g1;
            grid(A, pq, g0, g1);
#pragma unroll
            for (int k = 0; k < NX; ++k) X[k] = (k == A.idx0) ? g0 : ((k == A.idx1) ? g1 : A.ic[k]);
        }
        det.reset();
        cnt = 0;
        prev_t = 0;
        alive = wave_mask(in_range);
        fixed = false;
        s0 = s1 = sx = sz = first = last = 0.0;
    }

    // Divergence test of the recorded variable.  The band test |v| < L is
    // exact-safe (no lane inside it can fail the exact test); the exact test
    // runs only when some lane of the wave is outside it, behind a uniform
    // branch.  Callers keep the wave's control flow uniform.
    template <int WVT>
    __device__ __forceinline__ void check(const SweepArgs& A, int wv)
    {
        const double v = pick<NX, WVT>(X, wv);
        const unsigned long long out = wave_mask(!(fabs(v) < A.L)) & alive;
        if (out) alive &= ~wave_mask(lane_bit(out) && diverged_exact(v, A.maxValue, A.check));
    }

    // One recorded sample t: peak detection on the sample, the kind's
    // per-peak work, then preScaller steps and the divergence check.  BODY:
    // 2 <= t < B - 1 (detector without the warm-up test, no first/last
    // sample); PRE1: preScaller == 1.
    template <int WVT, bool BODY, bool PRE1>
    __device__ __forceinline__ void sample(const SweepArgs& A, int wv, WaveCursor& cur, int wave, int t)
    {
        const double v = pick<NX, WVT>(X, wv);
        if (!BODY && t == 0) first = v;
        if (KIND == KIND_BASINS_CFG) {
            sx = sx + X[0];
            sz = sz + X[NX - 1];
        }
        double pv = 0.0;
        int pt = 0;
        const unsigned long long em = (BODY ? det.feed_body(t - 1, v, pv, pt) : det.feed(t, v, pv, pt)) & alive;
        const bool e = lane_bit(em);
        // em is wave-uniform: a wave with no peak at this sample skips the
        // per-peak work altogether
        if (em) {
            if (BIF) {
                append_peaks(A.pool, cur, wave, em, pv, pt, cnt);
                if (e) ++cnt;
            } else if (KIND == KIND_BASINS_REF) {
                // running sums of the kept peaks and intervals
                // (avgPeakFinderCUDA, basinsCUDA.cu:841-861)
                if (e) {
                    if (cnt >= 1) {
                        s0 = s0 + pv;
                        s1 = s1 + (double)(pt - prev_t) * A.hp;
                    }
                    prev_t = pt;
                    ++cnt;
                }
            } else {
                if (e) {
                    if (cnt >= 1) {
                        s0 = s0 + pv;
                        s1 = fma(pv, pv, s1);
                    }
                    ++cnt;
                }
            }
        }
        if (!BODY && t == A.B - 1) {
            last = v;
            double xp[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) xp[k] = X[k];
            for (int r = 0; r < A.pre; ++r) Sy::step(X, a, A.sc);
            check<WVT>(A, wv);
            fixed = lane_bit(alive) && fixed_point(step_delta2<NX>(X, xp), A.check);
        } else {
            if (PRE1) Sy::step(X, a, A.sc);
            else
                for (int r = 0; r < A.pre; ++r) Sy::step(X, a, A.sc);
            check<WVT>(A, wv);
        }
    }

    template <int WVT>
    __device__ __forceinline__ void ticks(const SweepArgs& A, int wv, WaveCursor& cur, int wave, long long g0,
                                          long long g1)
    {
        long long g = g0;
        // ---- transient ----
        if (BIF && A.legacy_transient) {
            // legacy IC sweeps (debri/cudaLibrary.cu:1063-1070): the transient
            // tests X[0] for NaN/inf only (maxValue 0) and a fixed point at its
            // end is still recorded
            if (g < A.S) {
                const long long e = g1 < A.S ? g1 : A.S;
                for (long long s = g; s < e; ++s) {
                    Sy::step(X, a, A.sc);
                    alive &= ~wave_mask(!isfinite(X[0]));
                }
                g = e;
            }
        } else if (BIF) {
            // checked every step; a transient that ends diverged or at a
            // fixed point is not recorded and reports 0 (bifurcationCUDA.cu:334-342)
            if (g < A.S - 1) {
                const long long e = g1 < A.S - 1 ? g1 : A.S - 1;
                for (long long s = g; s < e; ++s) {
                    Sy::step(X, a, A.sc);
                    check<WVT>(A, wv);
                }
                g = e;
            }
            if (g == A.S - 1 && g < g1) {
                double xp[NX];
#pragma unroll
                for (int k = 0; k < NX; ++k) xp[k] = X[k];
                Sy::step(X, a, A.sc);
                check<WVT>(A, wv);
                alive &= ~wave_mask(fixed_point(step_delta2<NX>(X, xp), A.check));
                ++g;
            }
        } else if (g < A.S) {
            // no checks during the Basins transient (basinsCUDA.cu:264-271)
            const long long e = g1 < A.S ? g1 : A.S;
            for (long long s = g; s < e; ++s) Sy::step(X, a, A.sc);
            g = e;
        }
        // ---- record ----
        // head (t < 2: detector warm-up, first sample), body, tail (t = B-1)
        const int B = A.B;
        int t = (int)(g - A.S);
        const int te = (int)(g1 - A.S);
        const int body_end = te < B - 1 ? te : B - 1;
        for (; t < te && t < 2; ++t) sample<WVT, false, false>(A, wv, cur, wave, t);
        if (A.pre == 1) {
            // two samples per trip: the detector history and the state
            // rotate through alternating registers instead of being copied
            for (; t + 1 < body_end; t += 2) {
                sample<WVT, true, true>(A, wv, cur, wave, t);
                sample<WVT, true, true>(A, wv, cur, wave, t + 1);
            }
            for (; t < body_end; ++t) sample<WVT, true, true>(A, wv, cur, wave, t);
        } else {
            for (; t < body_end; ++t) sample<WVT, true, false>(A, wv, cur, wave, t);
        }
        for (; t < te; ++t) sample<WVT, false, false>(A, wv, cur, wave, t);
    }

    __device__ __forceinline__ void finish(const SweepArgs& A, long long q) const
    {
        const int flag = this->flag();
        if (A.flags) A.flags[q] = flag;
        if (BIF) {
            A.counts[q] = cnt;
            A.first[q] = first;
        } else if (KIND == KIND_BASINS_REF) {
            double f0, f1;
            if (flag == 0) {
                f0 = 999;
                f1 = 999;
            } else if (flag == -1) {
                f0 = last;
                f1 = -1.0;
            } else {
                const int P = cnt > 1 ? cnt - 1 : 0;
                if (P <= 0) {
                    f0 = 1000;
                    f1 = 1000;
                } else {
                    f0 = s0 / P;
                    f1 = s1 / P;
                }
            }
            A.f0[q] = f0;
            A.f1[q] = f1;
            if (A.f2) A.f2[q] = 0.0;
            if (A.counts) A.counts[q] = cnt;
        } else {
            double f[3];
            config_features(A, flag, f);
            A.f0[q] = f[0];
            A.f1[q] = f[1];
            A.f2[q] = f[2];
            if (A.counts) A.counts[q] = cnt;
        }
    }

    // Lane flag (1 alive, -1 fixed point, 0 diverged).
    __device__ __forceinline__ int flag() const { return lane_bit(alive) ? (fixed ? -1 : 1) : 0; }

    // CONFIG features (mean X[0], mean X[NX-1], population std of the kept
    // peaks), all 0 for a diverged lane.
    __device__ __forceinline__ void config_features(const SweepArgs& A, int flag, double* f) const
    {
        const int B = A.B;
        f[0] = f[1] = f[2] = 0.0;
        if (flag != 0) {
            f[0] = sx / B;
            f[1] = sz / B;
            const int n = cnt > 1 ? cnt - 1 : 0;
            if (n >= 2) {
                const double m = s0 / n;
                const double var = s1 / n - m * m;
                f[2] = var > 0 ? sqrt(var) : 0.0;
            }
        }
    }

    __device__ __forceinline__ void load(const LaneStore& L, long long q)
    {
#pragma unroll
        for (int k = 0; k < NX; ++k) X[k] = L.d[k * L.np + q];
        const double* d = L.d + NX * L.np + q;
        det.p = d[0];
        det.dn = d[L.np];
        first = d[2 * L.np];
        last = d[3 * L.np];
        s0 = d[4 * L.np];
        s1 = d[5 * L.np];
        sx = d[6 * L.np];
        sz = d[7 * L.np];
        det.pstart = L.i[q];
        cnt = L.i[L.np + q];
        prev_t = L.i[2 * L.np + q];
        const int bits = L.i[3 * L.np + q];
        alive = wave_mask(bits & 1);
        fixed = bits & 2;
        det.plat = wave_mask(bits & 4);
        det.skip = wave_mask(bits & 8);
    }
    __device__ __forceinline__ void store(const LaneStore& L, long long q) const
    {
#pragma unroll
        for (int k = 0; k < NX; ++k) L.d[k * L.np + q] = X[k];
        double* d = L.d + NX * L.np + q;
        d[0] = det.p;
        d[L.np] = det.dn;
        d[2 * L.np] = first;
        d[3 * L.np] = last;
        d[4 * L.np] = s0;
        d[5 * L.np] = s1;
        d[6 * L.np] = sx;
        d[7 * L.np] = sz;
        L.i[q] = det.pstart;
        L.i[L.np + q] = cnt;
        L.i[2 * L.np + q] = prev_t;
        L.i[3 * L.np + q] =
            (lane_bit(alive) ? 1 : 0) | (fixed ? 2 : 0) | (lane_bit(det.plat) ? 4 : 0) | (lane_bit(det.skip) ? 8 : 0);
    }
};

}  // namespace chaos
