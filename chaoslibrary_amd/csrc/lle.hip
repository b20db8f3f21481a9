// lle.hip -- largest Lyapunov exponent maps (Benettin, two trajectories per
// lane), the path of LLE/src/LLECUDA.cu:56-166 (kernels) and :205-428 (host).
//
// One lane per (A_i, B_j) grid point.  The base and perturbed trajectories
// are stepped in the same loop iteration, which gives the scheduler two
// independent FP64 dependency chains per lane; all state stays in VGPRs and
// nothing but the final FP64 exponent touches HBM (the reference keeps the
// state in shared memory and hands it between two kernels via semi_result).
#include <cmath>
#include <memory>
#include <vector>

#include "csv.hpp"
#include "dispatch.hpp"
#include "lanes.hpp"
#include "lle.hpp"

namespace chaos {

// Benettin loop for one point.  Shared by the one-shot kernel below and the
// batched engine, which stores (X, Xp, sum) between launches.
template <int SYS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Occupancy<SYS>::waves)))
void lle_map_kernel(const LleArgs p, double* __restrict__ out)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n_points) return;
    const int i = (int)(q % p.nPts);
    const int j = (int)(p.row0 + q / p.nPts);
    // transient (LLE/src/LLECUDA.cu:86), perturbation (:88-102) and the
    // Benettin windows (:147-163) in one pass
    LleProg<SYS> L;
    L.params(p, i, j);
    L.init(p);
    L.ticks(p, (long long)j * p.nPts + i, 0, LleProg<SYS>::total(p));
    out[q] = L.result(p);
}

LleArgs make_lle_args(const chaos_options& o, double tMax, double NT, int nPts, double h, double eps,
                      const double* ic, int nx, const double* ranges, const int* idx,
                      double transientTime, const double* values, int na)
{
    check_system(o.system, nx, na);
    CHAOS_REQUIRE(nPts >= 1, "nPts must be >= 1 (got %d)", nPts);
    CHAOS_REQUIRE(h > 0 && std::isfinite(h), "h must be positive");
    CHAOS_REQUIRE(ic && ranges && idx && values, "NULL input array");
    CHAOS_REQUIRE(idx[0] >= 0 && idx[0] < na && idx[1] >= 0 && idx[1] < na,
                  "indicesOfMutVars out of range [0,%d)", na);
    CHAOS_REQUIRE(eps > 0, "eps must be positive");
    CHAOS_REQUIRE(tMax > 0, "tMax must be positive");
    LleArgs p{};
    for (int k = 0; k < na; ++k) p.values[k] = values[k];
    for (int k = 0; k < nx; ++k) p.ic[k] = ic[k];
    p.rA0 = ranges[0];
    p.rA1 = ranges[1];
    p.rB0 = ranges[2];
    p.rB1 = ranges[3];
    p.idxA = idx[0];
    p.idxB = idx[1];
    p.nPts = nPts;
    p.S = horizon_steps(transientTime, h, o.semantics);
    p.NTp = horizon_steps(NT, h, o.semantics);
    const int all = horizon_steps(tMax, h, o.semantics);
    CHAOS_REQUIRE(p.NTp >= 1, "NT/h must be >= 1");
    // The reference runs amountOfCalcBlocks + 1 = all/NTp + 2 windows
    // (LLE/src/LLECUDA.cu:227,147); CONFIG runs exactly all/NTp.
    p.W = (o.semantics == CHAOS_SEM_REFERENCE) ? all / p.NTp + 2 : all / p.NTp;
    p.eps = eps;
    p.tMax = tMax;
    p.sc = make_step_const(h);
    p.seed = o.seed;
    return p;
}

double lle_steps_per_point(const LleArgs& p) { return (double)p.S + 2.0 * p.W * (double)p.NTp; }

void launch_lle_rows(CallCtx& ctx, int system, LleArgs p, long long row0, long long rows, double* d_out)
{
    p.row0 = row0;
    p.n_points = rows * p.nPts;
    const int block = ctx.opt().block > 0 ? ctx.opt().block : 256;
    CHAOS_REQUIRE(block % 64 == 0 && block <= 256, "block must be a multiple of 64 and <= 256");
    const unsigned grid = ceil_div(p.n_points, block);
    dispatch_system(system, [&](auto tag) {
        constexpr int S = decltype(tag)::value;
        hipLaunchKernelGGL(lle_map_kernel<S>, dim3(grid), dim3(block), 0, ctx.stream(), p, d_out);
    });
    CHAOS_HIP(hipGetLastError());
    ctx.add_launch();
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_lle2d(const chaos_options* opt, double tMax, double NT, int nPts, double h,
                           double eps, const double* initialConditions, int amountOfInitialConditions,
                           const double* ranges, const int* indicesOfMutVars, int writableVar,
                           double maxValue, double transientTime, const double* values,
                           int amountOfValues, double* out)
{
    (void)writableVar;
    (void)maxValue;  // never read by the reference either (LLE/src/LLECUDA.cu:205-428)
    return guarded("chaos_lle2d", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(out, "out is NULL");
        LleArgs p = make_lle_args(o, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions,
                                  ranges, indicesOfMutVars, transientTime, values, amountOfValues);
        CallCtx ctx(o);
        const long long n = (long long)nPts * nPts;
        DevBuf<double> tmp;
        double* d_out = out;
        if (!o.device_buffers) {
            tmp.alloc((size_t)n);
            d_out = tmp.get();
        }
        ctx.begin();
        launch_lle_rows(ctx, o.system, p, 0, nPts, d_out);
        const double ms = ctx.end();
        ctx.add_kernel_ms(ms);
        ctx.add_steps(lle_steps_per_point(p) * (double)n);
        if (!o.device_buffers)
            CHAOS_HIP(hipMemcpyAsync(out, d_out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream()));
        ctx.sync();
        ctx.finish();
    });
}

namespace chaos {

void write_lle_file(TextFile* f, const char* path, const double* ranges, int nPts, const double* v, int output)
{
    // two range lines, then the map at the default precision with NaN
    // written as 0 (LLE/src/LLECUDA.cu:252-253, 383-398)
    MapSpec m;
    m.ranges = ranges;
    m.header_lines = 2;
    m.nPts = nPts;
    m.count = (long long)nPts * nPts;
    m.nan_to_zero = true;
    m.dv = v;
    if (f) {
        write_map_text(*f, m);
        f->close();
    }
    if (out_bin(output)) write_map_bin(std::string(path) + ".bin", m);
}

}  // namespace chaos

extern "C" int chaos_LLE2D(const chaos_options* opt, double tMax, double NT, int nPts, double h,
                           double eps, const double* initialConditions, int amountOfInitialConditions,
                           const double* ranges, const int* indicesOfMutVars, int writableVar,
                           double maxValue, double transientTime, const double* values,
                           int amountOfValues, const char* OUT_FILE_PATH)
{
    return guarded("chaos_LLE2D", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        const long long n = (long long)nPts * (nPts > 0 ? nPts : 0);
        // Open first: the reference exits before computing when the file
        // cannot be opened (LLE/src/LLECUDA.cu:245-251).
        std::unique_ptr<TextFile> f;
        if (out_csv(o.output)) f.reset(new TextFile(OUT_FILE_PATH));
        std::vector<double> res((size_t)n);
        int rc = chaos_lle2d(&o, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
                             indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                             res.data());
        if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
        write_lle_file(f.get(), OUT_FILE_PATH, ranges, nPts, res.data(), o.output);
    });
}
