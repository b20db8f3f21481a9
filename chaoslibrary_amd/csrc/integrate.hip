// integrate.hip -- plain trajectory advance: one lane per trajectory,
// X <- step^nsteps(X) for the selected system (the loop of
// LLE/src/LLECUDA.cu:38-49 / Bifurcation/src/bifurcationCUDA.cu:435-443
// without the analyses).  Used for the bit-exact trajectory checks.
#include "dispatch.hpp"

namespace chaos {

template <int SYS>
__global__ __launch_bounds__(256) void integrate_kernel(int n, const double* __restrict__ params,
                                                        double* __restrict__ states, long long nsteps,
                                                        StepConst sc)
{
    using S = System<SYS>;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double a[S::NA], X[S::NX];
#pragma unroll
    for (int k = 0; k < S::NA; ++k) a[k] = params[(size_t)p * S::NA + k];
#pragma unroll
    for (int k = 0; k < S::NX; ++k) X[k] = states[(size_t)p * S::NX + k];
    for (long long s = 0; s < nsteps; ++s) S::step(X, a, sc);
#pragma unroll
    for (int k = 0; k < S::NX; ++k) states[(size_t)p * S::NX + k] = X[k];
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_integrate(const chaos_options* opt, int n, const double* params, double* states,
                               long long nsteps, double h)
{
    return guarded("chaos_integrate", [&] {
        chaos_options o = resolve_options(opt);
        int nx = 0, na = 0;
        if (chaos_system_dims(o.system, &nx, &na) != CHAOS_OK) fail(CHAOS_E_INVALID, "unknown system %d", o.system);
        CHAOS_REQUIRE(n >= 0, "n must be >= 0");
        CHAOS_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
        CHAOS_REQUIRE(n == 0 || (params && states), "NULL params/states");
        if (n == 0) return;
        CallCtx ctx(o);
        DevBuf<double> dp, ds;
        const double* d_params = params;
        double* d_states = states;
        if (!o.device_buffers) {
            dp.alloc((size_t)n * na);
            ds.alloc((size_t)n * nx);
            CHAOS_HIP(hipMemcpyAsync(dp.get(), params, dp.bytes(), hipMemcpyHostToDevice, ctx.stream()));
            CHAOS_HIP(hipMemcpyAsync(ds.get(), states, ds.bytes(), hipMemcpyHostToDevice, ctx.stream()));
            d_params = dp.get();
            d_states = ds.get();
        }
        const int block = 256;
        const unsigned grid = ceil_div((unsigned long long)n, block);
        const StepConst sc = make_step_const(h);
        ctx.begin();
        dispatch_system(o.system, [&](auto tag) {
            constexpr int S = decltype(tag)::value;
            hipLaunchKernelGGL(integrate_kernel<S>, dim3(grid), dim3(block), 0, ctx.stream(), n, d_params,
                               d_states, nsteps, sc);
        });
        CHAOS_HIP(hipGetLastError());
        ctx.add_launch();
        ctx.add_kernel_ms(ctx.end());
        ctx.add_steps((double)n * (double)nsteps);
        if (!o.device_buffers)
            CHAOS_HIP(hipMemcpyAsync(states, d_states, ds.bytes(), hipMemcpyDeviceToHost, ctx.stream()));
        ctx.sync();
        ctx.finish();
    });
}
