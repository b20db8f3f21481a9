// sweep.hip -- fused transient + record + online analysis, one lane per grid
// point, for the Bifurcation and Basins paths.
//
// Replaces the reference's three-kernel pipeline per chunk
// (calculateTransTimeCUDA -> calculateDiscreteModelCUDA -> peakFinderCUDA /
// avgPeakFinderCUDA, Bifurcation/src/bifurcationCUDA.cu:258-371,
// Basins/src/basinsCUDA.cu:233-314,818-864), which writes every recorded
// sample to HBM (points x steps doubles) and reads it back.  Here the sample
// stream never leaves VGPRs: the peak detector runs online and only the
// peaks (Bifurcation) or a few running sums (Basins) are produced.
#include <cmath>

#include "dispatch.hpp"
#include "sweep.hpp"

namespace chaos {

// Divergence test of one step (Bifurcation/src/bifurcationCUDA.cu:445-453,
// Basins/src/basinsCUDA.cu:350-361), exact per check mode.
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

// Last-step fixed-point test (Bifurcation/src/bifurcationCUDA.cu:456-466,
// Basins/src/basinsCUDA.cu:364-376).
__device__ __forceinline__ bool fixed_point(double t, int check)
{
    if (check == CHK_BIF_REF) return (double)sqrtf(fabsf((float)t)) < 1e-9;
    if (check == CHK_BASINS_REF) return sqrt(fabs(t)) < 1e-9;
    return sqrt(t) < 1e-9;
}

// One VALU compare per step on the common path; the exact test only runs
// when some live lane leaves the proven-safe band |v| < L.
template <int WV>
__device__ __forceinline__ void check_step(const double* X, const SweepArgs& A, bool& alive)
{
    const double v = X[WV];
    if (alive && !(fabs(v) < A.L)) {
        if (diverged_exact(v, A.maxValue, A.check)) alive = false;
    }
}

template <int NX>
__device__ __forceinline__ double step_delta2(const double* X, const double* xp)
{
    double t = 0;
#pragma unroll
    for (int j = 0; j < NX; ++j) t += ((X[j] - xp[j]) * (X[j] - xp[j]));
    return t;
}

template <int SYS, int WV, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Occupancy<SYS>::waves)))
void sweep_kernel(const SweepArgs A)
{
    using Sy = System<SYS>;
    constexpr int NX = Sy::NX, NA = Sy::NA;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = (int)(q >> 6);
    const bool in_range = q < A.n_points;
    bool alive = in_range;
    const long long pq = in_range ? q : 0;
    const int i = (int)(pq % A.nPts);
    const int j = A.dims == 2 ? (int)(pq / A.nPts) : 0;
    const double g0 = grid_value(i, A.nPts, A.r[0], A.r[1]);
    const double g1 = A.dims == 2 ? grid_value(j, A.nPts, A.r[2], A.r[3]) : 0.0;

    double a[NA], X[NX];
    if (KIND == KIND_BIF) {
        // swept entries of the parameter vector (bifurcationCUDA.cu:282-284)
        build_params<NA>(a, A.values, A.idx0, g0, A.idx1, g1);
#pragma unroll
        for (int k = 0; k < NX; ++k) X[k] = A.ic[k];
    } else {
        // swept entries of the initial condition (basinsCUDA.cu:253-262)
#pragma unroll
        for (int k = 0; k < NA; ++k) a[k] = A.values[k];
#pragma unroll
        for (int k = 0; k < NX; ++k) X[k] = (k == A.idx0) ? g0 : ((k == A.idx1) ? g1 : A.ic[k]);
    }

    // ---- transient ----
    if (KIND == KIND_BIF) {
        // checked every step; a transient that ends diverged or at a fixed
        // point is not recorded and reports 0 (bifurcationCUDA.cu:334-342)
        if (A.S > 0) {
            for (int s = 0; s < A.S - 1; ++s) {
                Sy::step(X, a, A.sc);
                check_step<WV>(X, A, alive);
            }
            double xp[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) xp[k] = X[k];
            Sy::step(X, a, A.sc);
            check_step<WV>(X, A, alive);
            if (alive && fixed_point(step_delta2<NX>(X, xp), A.check)) alive = false;
        }
    } else {
        // no checks during the Basins transient (basinsCUDA.cu:264-271)
        for (int s = 0; s < A.S; ++s) Sy::step(X, a, A.sc);
    }

    // ---- record ----
    PeakDetector det;
    det.reset();
    WaveCursor cur{A.pool.nchunks, kChunk};
    int cnt = 0, prev_t = 0;
    bool fixed = false;
    double s0 = 0.0, s1 = 0.0, sx = 0.0, sz = 0.0, first = 0.0, last = 0.0;
    const int B = A.B;
    for (int t = 0; t < B; ++t) {
        const double v = X[WV];
        if (t == 0) first = v;
        if (KIND == KIND_BASINS_CFG) {
            sx = sx + X[0];
            sz = sz + X[NX - 1];
        }
        double pv = 0.0;
        int pt = 0;
        const bool e = det.feed(t, v, pv, pt) && alive;
        if (KIND == KIND_BIF) {
            append_peaks(A.pool, cur, wave, e, pv, pt, cnt);
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
        if (t == B - 1) {
            last = v;
            double xp[NX];
#pragma unroll
            for (int k = 0; k < NX; ++k) xp[k] = X[k];
            for (int r = 0; r < A.pre; ++r) Sy::step(X, a, A.sc);
            check_step<WV>(X, A, alive);
            fixed = alive && fixed_point(step_delta2<NX>(X, xp), A.check);
        } else {
            for (int r = 0; r < A.pre; ++r) Sy::step(X, a, A.sc);
            check_step<WV>(X, A, alive);
        }
    }

    if (KIND == KIND_BIF) close_chunk(A.pool, cur);
    if (!in_range) return;
    const int flag = alive ? (fixed ? -1 : 1) : 0;
    if (A.flags) A.flags[q] = flag;
    if (KIND == KIND_BIF) {
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
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        if (flag != 0) {
            f0 = sx / B;
            f1 = sz / B;
            const int n = cnt > 1 ? cnt - 1 : 0;
            if (n >= 2) {
                const double m = s0 / n;
                const double var = s1 / n - m * m;
                f2 = var > 0 ? sqrt(var) : 0.0;
            }
        }
        A.f0[q] = f0;
        A.f1[q] = f1;
        A.f2[q] = f2;
        if (A.counts) A.counts[q] = cnt;
    }
}

double divergence_fast_bound(int check, double maxValue)
{
    // |v| < 2^53 keeps v == v + 1.0 false; |v| < M (1 - 2^-22) keeps the
    // float-rounded |v| <= M.  Non-finite v always fails the bound.
    double L = INFINITY;
    if (check == CHK_BIF_REF) L = 9007199254740992.0;
    if (maxValue != 0) {
        const double m = check == CHK_CONFIG ? maxValue : maxValue * (1.0 - 0x1p-22);
        L = m > 0 ? (m < L ? m : L) : 0.0;
    }
    return L;
}

void launch_sweep(CallCtx& ctx, int system, int writableVar, int kind, const SweepArgs& A)
{
    const int block = 256;
    // a whole number of waves so every wave's lanes share one control flow
    const unsigned grid = ceil_div((unsigned long long)A.n_points, block);
    int nx = 0, na = 0;
    chaos_system_dims(system, &nx, &na);
    CHAOS_REQUIRE(writableVar >= 0 && writableVar < nx, "writableVar %d out of range [0,%d)", writableVar, nx);
    dispatch_system(system, [&](auto tag) {
        constexpr int S = decltype(tag)::value;
        constexpr int NXs = System<S>::NX;
        auto go = [&](auto wvtag) {
            constexpr int WV = decltype(wvtag)::value;
            if constexpr (WV < NXs) {
                if (kind == KIND_BIF)
                    hipLaunchKernelGGL((sweep_kernel<S, WV, KIND_BIF>), dim3(grid), dim3(block), 0, ctx.stream(), A);
                else if (kind == KIND_BASINS_REF)
                    hipLaunchKernelGGL((sweep_kernel<S, WV, KIND_BASINS_REF>), dim3(grid), dim3(block), 0,
                                       ctx.stream(), A);
                else
                    hipLaunchKernelGGL((sweep_kernel<S, WV, KIND_BASINS_CFG>), dim3(grid), dim3(block), 0,
                                       ctx.stream(), A);
            }
        };
        if (writableVar == 0) go(std::integral_constant<int, 0>{});
        else if (writableVar == 1) go(std::integral_constant<int, 1>{});
        else go(std::integral_constant<int, 2>{});
    });
    CHAOS_HIP(hipGetLastError());
    ctx.add_launch();
}

}  // namespace chaos
