// sweep.hip -- fused transient + record + online analysis, one lane per grid
// point, for the Bifurcation and Basins paths.
//
// Replaces the reference's three-kernel pipeline per chunk
// (calculateTransTimeCUDA -> calculateDiscreteModelCUDA -> peakFinderCUDA /
// avgPeakFinderCUDA, Bifurcation/src/bifurcationCUDA.cu:258-371,
// Basins/src/basinsCUDA.cu:233-314,818-864), which writes every recorded
// sample to HBM (points x steps doubles) and reads it back.  Here the sample
// stream never leaves VGPRs: the peak detector runs online and only the
// peaks (Bifurcation) or a few running sums (Basins) are produced.  The
// per-lane program (SweepProg) is shared with the batched engine.
#include <cfloat>
#include <cmath>

#include "dispatch.hpp"
#include "lanes.hpp"
#include "moments.hpp"
#include "sweep.hpp"

namespace chaos {

template <int SYS, int WV, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Occupancy<SYS>::sweep_min, 8)))
void sweep_kernel(const SweepArgs A)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int wave = (int)(q >> 6);
    const bool in_range = q < A.n_points;
    SweepProg<SYS, KIND> L;
    // parameters from the grid index, outputs at the launch's own index q
    L.init(A, A.p0 + (in_range ? q : 0), in_range);
    WaveCursor cur{A.pool.nchunks, kChunk};
    // every lane of the wave stays in the loop (wave-level peak appends)
    L.template ticks<WV>(A, WV, cur, wave, 0, SweepProg<SYS, KIND>::total(A));
    if (stores_peaks(KIND)) close_chunk(A.pool, cur);
    if (in_range) L.finish(A, q);
    if constexpr (KIND == KIND_BASINS_CFG) {
        // the z-score's first pass, from the features still in registers
        if (A.zsum) {
            const int flag = in_range ? L.flag() : 0;
            double f[3];
            L.config_features(A, flag, f);
            block256_moments(f, 3, flag != 0, A.zb, A.zsum, A.zcnt, A.zmm);
        }
    }
}

double divergence_fast_bound(int check, double maxValue)
{
    // |v| < 2^53 keeps v == v + 1.0 false; |v| < M (1 - 2^-22) keeps the
    // float-rounded |v| <= M.  Non-finite v always fails the bound.  The
    // float-cast tests also need |v| < FLT_MAX: a larger double rounds to
    // inf in (float)v, which exceeds any finite maxValue.
    double L = INFINITY;
    if (check == CHK_BIF_REF) L = 9007199254740992.0;
    if (maxValue != 0) {
        double m = check == CHK_CONFIG ? maxValue : maxValue * (1.0 - 0x1p-22);
        if (check != CHK_CONFIG && m > (double)FLT_MAX) m = (double)FLT_MAX;
        L = m > 0 ? (m < L ? m : L) : 0.0;
    }
    return L;
}

void launch_sweep(CallCtx& ctx, int system, int writableVar, int kind, const SweepArgs& A)
{
    // A grid of at most one wavefront per CU (a 1-D sweep of 1000 points: 16
    // wavefronts) runs them one per CU in 64-thread blocks, so that a
    // latency-bound wavefront does not share its CU's scalar unit and
    // instruction issue with three others: 31.9 -> 28.3 ms for that sweep.
    // (Grids of a few wavefronts per CU gained nothing and sometimes lost.)
    // The basins epilogue's moments need the 256-thread blocks.
    const bool spread = A.zsum == nullptr &&
                        ceil_div((unsigned long long)A.n_points, 64) <=
                            (unsigned long long)device_props(ctx.device()).multiProcessorCount;
    const int block = spread ? 64 : 256;
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
                else if (kind == KIND_BIF_VALS)
                    hipLaunchKernelGGL((sweep_kernel<S, WV, KIND_BIF_VALS>), dim3(grid), dim3(block), 0,
                                       ctx.stream(), A);
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
