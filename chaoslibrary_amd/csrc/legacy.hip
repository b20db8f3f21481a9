// legacy.hip -- the reference's legacy entry points (namespace old_library,
// debri/hostLibrary.cu) that sweep the same integrator: LLE over a 1-D
// parameter line or a 1-D / 2-D initial-condition grid, the Lyapunov
// spectrum (LS1D / LS2D) and the time-domain recorder.
//
// One lane per grid point with the state in VGPRs, as everywhere else.  The
// reference's kernels (debri/cudaLibrary.cu:1897-2137 LLE, :2189-2331 LS,
// :879-953 time domain) keep the state in shared memory, re-seed curand or
// sinf for the directions and cut the grid into memory-sized chunks; here
// the directions come from the seeded SplitMix64 unit vector and a map is one
// launch.
#include <cmath>
#include <string>
#include <vector>

#include "csv.hpp"
#include "dispatch.hpp"
#include "lanes.hpp"

namespace chaos {

struct LegacyArgs {
    double values[kMaxNA];
    double ic[kMaxNX];
    double r[4];           // ranges of the swept entries
    int idx0, idx1;        // swept indices (parameters or state), idx1 = -1 in 1-D
    int dims;              // 1 or 2
    int nPts;
    long long n_points;    // nPts or nPts^2
    long long S, NTp, W;   // transient steps, steps per window, windows
    double eps, tMax, maxValue;
    StepConst sc;
    unsigned long long seed;
};

// The legacy loop's early stop (debri/cudaLibrary.cu:472-503 with
// writableVar 0): a step that leaves X[0] NaN/inf or fabsf(X[0]) > maxValue
// ends the current call's steps.
__device__ __forceinline__ bool legacy_bad(double v, double M)
{
    if (isnan(v) || isinf(v)) return true;
    return M != 0 && (double)fabsf((float)v) > M;
}

template <int SYS>
__device__ __forceinline__ void legacy_run(double* X, const double* a, const LegacyArgs& A, long long n)
{
    for (long long s = 0; s < n; ++s) {
        System<SYS>::step(X, a, A.sc);
        if (legacy_bad(X[0], A.maxValue)) return;
    }
}

// Parameters and state of legacy point p (getValueByIdx,
// debri/cudaLibrary.cu:1182-1186: dim 0 = p % nPts, dim 1 = p / nPts).
template <int SYS, bool IC>
__device__ __forceinline__ void legacy_point(const LegacyArgs& A, long long p, double* X, double* a)
{
    constexpr int NX = System<SYS>::NX, NA = System<SYS>::NA;
    const double g0 = grid_value((int)(p % A.nPts), A.nPts, A.r[0], A.r[1]);
    const double g1 = A.dims == 2 ? grid_value((int)((p / A.nPts) % A.nPts), A.nPts, A.r[2], A.r[3]) : 0.0;
    const int i1 = A.dims == 2 ? A.idx1 : -1;
    if (IC) {
        build_params<NA>(a, A.values, -1, 0.0, -1, 0.0);
#pragma unroll
        for (int k = 0; k < NX; ++k) X[k] = (k == A.idx0) ? g0 : ((k == i1) ? g1 : A.ic[k]);
    } else {
        build_params<NA>(a, A.values, A.idx0, g0, i1, g1);
#pragma unroll
        for (int k = 0; k < NX; ++k) X[k] = A.ic[k];
    }
}

// Legacy LLE (FORM 0: LLEKernelCUDA, debri/cudaLibrary.cu:1897-2031,
// parameters swept; FORM 1: LLEKernelICCUDA, :2034-2137, initial conditions
// swept).  W windows of NTp steps on X then on Y, each with its own early
// stop; FORM 0 d = sqrt(sum((1/eps)(x-y))^2), Y = X - (X-Y)/d; FORM 1
// d = sqrt(sum(x-y)^2)/eps, Y = X - (X-Y)*(1/d); result sum(log d)/tMax.
template <int SYS, int FORM>
__global__ __launch_bounds__(256) void legacy_lle_kernel(const LegacyArgs A, double* __restrict__ out)
{
    constexpr int NX = System<SYS>::NX, NA = System<SYS>::NA;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n_points) return;
    double X[NX], Y[NX], a[NA], v[NX];
    legacy_point<SYS, FORM == 1>(A, p, X, a);
    legacy_run<SYS>(X, a, A, A.S);
    unit_vector<NX>(A.seed, p, v);
#pragma unroll
    for (int i = 0; i < NX; ++i) Y[i] = X[i] + v[i] * A.eps;
    double result = 0;
    for (long long w = 0; w < A.W; ++w) {
        // X and Y advance independently; interleaving their steps is the
        // same arithmetic as running them one after the other
        bool gx = true, gy = true;
        for (long long s = 0; s < A.NTp; ++s) {
            if (gx) {
                System<SYS>::step(X, a, A.sc);
                gx = !legacy_bad(X[0], A.maxValue);
            }
            if (gy) {
                System<SYS>::step(Y, a, A.sc);
                gy = !legacy_bad(Y[0], A.maxValue);
            }
        }
        double t = 0;
        if (FORM == 0) {
#pragma unroll
            for (int l = 0; l < NX; ++l) {
                double t2 = 1 / A.eps;
                t2 = t2 * (X[l] - Y[l]);
                t += t2 * t2;
            }
            t = sqrt(t);
            result += log(t);
#pragma unroll
            for (int j = 0; j < NX; ++j) Y[j] = X[j] - ((X[j] - Y[j]) / t);
        } else {
#pragma unroll
            for (int l = 0; l < NX; ++l) t += (X[l] - Y[l]) * (X[l] - Y[l]);
            t = sqrt(t) / A.eps;
            result += log(t);
            if (t != 0) t = 1 / t;
#pragma unroll
            for (int j = 0; j < NX; ++j) Y[j] = X[j] - ((X[j] - Y[j]) * t);
        }
    }
    out[p] = result / A.tMax;
}

// Classical Gram-Schmidt of the N rows of a into b, then normalisation
// (debri/cudaLibrary.cu:2141-2187): numerators from the original row, the
// earlier rows unnormalised; den[i] = norm of row i (0 rows stay 0).
template <int N>
__device__ __forceinline__ void gram_schmidt(const double (&a)[N][N], double (&b)[N][N], double (&den)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
        for (int j = 0; j < N; ++j) b[i][j] = a[i][j];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double num = 0, dn = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                num += a[i][k] * b[j][k];
                dn += b[j][k] * b[j][k];
            }
            const double frac = dn == 0 ? 0 : num / dn;
#pragma unroll
            for (int k = 0; k < N; ++k) b[i][k] -= frac * b[j][k];
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double d = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) d += b[i][j] * b[i][j];
        d = sqrt(d);
#pragma unroll
        for (int j = 0; j < N; ++j) b[i][j] = d == 0 ? 0 : b[i][j] / d;
        den[i] = d;
    }
}

// Lyapunov spectrum (LSKernelCUDA, debri/cudaLibrary.cu:2189-2331): NX
// perturbed trajectories renormalised by Gram-Schmidt every window;
// out[p*NX + k] = sum_k log(den_k/eps) / tMax.  Initial directions: seeded
// unit vectors at point*4 + j (the reference's sinf rows cannot be matched).
template <int SYS>
__global__ __launch_bounds__(256) void legacy_ls_kernel(const LegacyArgs A, double* __restrict__ out)
{
    constexpr int NX = System<SYS>::NX, NA = System<SYS>::NA;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n_points) return;
    double X[NX], a[NA], Y[NX][NX], Z[NX][NX], den[NX], res[NX];
    legacy_point<SYS, false>(A, p, X, a);
#pragma unroll
    for (int j = 0; j < NX; ++j) {
        unit_vector<NX>(A.seed, p * 4 + j, Z[j]);
        res[j] = 0;
    }
    legacy_run<SYS>(X, a, A, A.S);
    gram_schmidt<NX>(Z, Y, den);
#pragma unroll
    for (int j = 0; j < NX; ++j)
#pragma unroll
        for (int i = 0; i < NX; ++i) Y[j][i] = Y[j][i] * A.eps + X[i];
    for (long long w = 0; w < A.W; ++w) {
        bool gx = true, gy[NX];
#pragma unroll
        for (int j = 0; j < NX; ++j) gy[j] = true;
        for (long long s = 0; s < A.NTp; ++s) {
            if (gx) {
                System<SYS>::step(X, a, A.sc);
                gx = !legacy_bad(X[0], A.maxValue);
            }
#pragma unroll
            for (int j = 0; j < NX; ++j)
                if (gy[j]) {
                    System<SYS>::step(Y[j], a, A.sc);
                    gy[j] = !legacy_bad(Y[j][0], A.maxValue);
                }
        }
#pragma unroll
        for (int k = 0; k < NX; ++k)
#pragma unroll
            for (int l = 0; l < NX; ++l) Y[k][l] = Y[k][l] - X[l];
        gram_schmidt<NX>(Y, Z, den);
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            res[k] += log(den[k] / A.eps);
#pragma unroll
            for (int j = 0; j < NX; ++j) Y[k][j] = X[j] + Z[k][j] * A.eps;
        }
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) out[p * NX + k] = res[k] / A.tMax;
}

// Time-domain recorder (calculateDiscreteModelCUDA with
// loopCalculateDiscreteModel_int, debri/cudaLibrary.cu:879-953, :604-680):
// transient of S iterations of `pre` steps, checked on X[wv]; only flag 1
// records B samples X[wv] (before each iteration) into out[p*B + t].  The
// caller zero-fills out: samples after a divergence stay 0.
template <int SYS>
__global__ __launch_bounds__(256) void time_domain_kernel(const LegacyArgs A, int wv, int pre, long long B,
                                                          double* __restrict__ out, int* __restrict__ flags)
{
    constexpr int NX = System<SYS>::NX, NA = System<SYS>::NA;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n_points) return;
    double X[NX], a[NA], xp[NX];
    legacy_point<SYS, false>(A, p, X, a);
    auto loop = [&](long long n, double* data) -> int {
        if (n <= 0) return 1;
        for (long long i = 0; i < n; ++i) {
#pragma unroll
            for (int k = 0; k < NX; ++k) xp[k] = X[k];
            if (data) data[i] = pick<NX, -1>(X, wv);
            for (int r = 0; r < pre; ++r) System<SYS>::step(X, a, A.sc);
            if (diverged_exact(pick<NX, -1>(X, wv), A.maxValue, CHK_BASINS_REF)) return 0;
        }
        return fixed_point(step_delta2<NX>(X, xp), CHK_BASINS_REF) ? -1 : 1;
    };
    int flag = loop(A.S, nullptr);
    if (flag == 1) flag = loop(B, out + p * B);
    flags[p] = flag;
}

__global__ void zero_kernel(double* a, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

// Legacy horizons: `int x = T / h` (debri/hostLibrary.cu:3129-3136) or
// llround in CONFIG semantics.
long long legacy_count(double q, int sem) { return sem == CHAOS_SEM_REFERENCE ? (long long)q : std::llround(q); }

LegacyArgs make_legacy_args(const chaos_options& o, double tMax, double NT, int nPts, double h, double eps,
                            const double* ic, int nx, const double* ranges, const int* idx, double maxValue,
                            double transientTime, const double* values, int na, int dims, bool sweep_ic)
{
    check_system(o.system, nx, na);
    CHAOS_REQUIRE(dims == 1 || dims == 2, "dims must be 1 or 2");
    CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2 (got %d)", nPts);
    CHAOS_REQUIRE(h > 0 && std::isfinite(h), "h must be positive");
    CHAOS_REQUIRE(tMax > 0 && NT > 0, "tMax and NT must be positive");
    CHAOS_REQUIRE(eps > 0, "eps must be positive");
    CHAOS_REQUIRE(ic && ranges && idx && values, "NULL input array");
    const int lim = sweep_ic ? nx : na;
    for (int d = 0; d < dims; ++d)
        CHAOS_REQUIRE(idx[d] >= 0 && idx[d] < lim, "indicesOfMutVars[%d] out of range [0,%d)", d, lim);
    LegacyArgs A{};
    for (int k = 0; k < na; ++k) A.values[k] = values[k];
    for (int k = 0; k < nx; ++k) A.ic[k] = ic[k];
    for (int k = 0; k < 2 * dims; ++k) A.r[k] = ranges[k];
    A.idx0 = idx[0];
    A.idx1 = dims == 2 ? idx[1] : -1;
    A.dims = dims;
    A.nPts = nPts;
    A.n_points = dims == 2 ? (long long)nPts * nPts : nPts;
    A.S = legacy_count(transientTime / h, o.semantics);
    A.NTp = legacy_count(NT / h, o.semantics);
    A.W = legacy_count(tMax / NT, o.semantics);
    CHAOS_REQUIRE(A.NTp >= 1, "NT/h must be >= 1");
    CHAOS_REQUIRE(A.S >= 0 && A.W >= 0, "negative horizon");
    A.eps = eps;
    A.tMax = tMax;
    A.maxValue = maxValue;
    A.sc = make_step_const(h);
    A.seed = o.seed;
    return A;
}

// Runs the kernel picked by launch(tag) on the call stream and returns the
// result to host or device memory (opt.device_buffers).
template <class Launch>
void run_legacy(CallCtx& ctx, const chaos_options& o, long long count, double* out, double steps, Launch&& launch)
{
    DevBuf<double> tmp;
    double* d_out = out;
    if (!o.device_buffers) {
        tmp.alloc((size_t)count);
        d_out = tmp.get();
    }
    ctx.begin();
    dispatch_system(o.system, [&](auto tag) { launch(tag, d_out); });
    CHAOS_HIP(hipGetLastError());
    ctx.add_launch();
    ctx.add_kernel_ms(ctx.end());
    ctx.add_steps(steps);
    if (!o.device_buffers)
        CHAOS_HIP(hipMemcpyAsync(out, d_out, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    ctx.finish();
}

void legacy_lle(const chaos_options& o, double tMax, double NT, int nPts, double h, double eps,
                const double* ic, int nx, const double* ranges, const int* idx, double maxValue,
                double transientTime, const double* values, int na, int dims, bool sweep_ic, double* out)
{
    CHAOS_REQUIRE(out, "out is NULL");
    const LegacyArgs A = make_legacy_args(o, tMax, NT, nPts, h, eps, ic, nx, ranges, idx, maxValue, transientTime,
                                          values, na, dims, sweep_ic);
    CallCtx ctx(o);
    const unsigned grid = ceil_div(A.n_points, 256);
    const double steps = (double)A.n_points * ((double)A.S + 2.0 * (double)A.W * (double)A.NTp);
    run_legacy(ctx, o, A.n_points, out, steps, [&](auto tag, double* d_out) {
        constexpr int S = decltype(tag)::value;
        if (sweep_ic)
            hipLaunchKernelGGL((legacy_lle_kernel<S, 1>), dim3(grid), dim3(256), 0, ctx.stream(), A, d_out);
        else
            hipLaunchKernelGGL((legacy_lle_kernel<S, 0>), dim3(grid), dim3(256), 0, ctx.stream(), A, d_out);
    });
}

void legacy_ls(const chaos_options& o, double tMax, double NT, int nPts, double h, double eps, const double* ic,
               int nx, const double* ranges, const int* idx, double maxValue, double transientTime,
               const double* values, int na, int dims, double* out)
{
    CHAOS_REQUIRE(out, "out is NULL");
    const LegacyArgs A = make_legacy_args(o, tMax, NT, nPts, h, eps, ic, nx, ranges, idx, maxValue, transientTime,
                                          values, na, dims, false);
    CallCtx ctx(o);
    const unsigned grid = ceil_div(A.n_points, 256);
    const double steps = (double)A.n_points * ((double)A.S + (nx + 1.0) * (double)A.W * (double)A.NTp);
    run_legacy(ctx, o, A.n_points * nx, out, steps, [&](auto tag, double* d_out) {
        constexpr int S = decltype(tag)::value;
        hipLaunchKernelGGL((legacy_ls_kernel<S>), dim3(grid), dim3(256), 0, ctx.stream(), A, d_out);
    });
}

// The 1-D legacy writers: "<value>, <v1>[, <v2>...]\n" per point.
void write_lines(const std::string& path, const LegacyArgs& A, const double* v, int per, int precision)
{
    TextFile f(path);
    for (long long k = 0; k < A.n_points; ++k) {
        f.put_g(grid_value((int)k, A.nPts, A.r[0], A.r[1]), precision);
        for (int j = 0; j < per; ++j) {
            f.put(", ", 2);
            f.put_g(v[k * per + j], precision);
        }
        f.put("\n", 1);
    }
    f.close();
}

}  // namespace

}  // namespace chaos

using namespace chaos;

extern "C" {

int chaos_legacy_lle(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
                     const double* initialConditions, int amountOfInitialConditions, const double* ranges,
                     const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
                     const double* values, int amountOfValues, int dims, int sweep_ic, double* out)
{
    (void)writableVar;  // the legacy kernels test X[0] whatever writableVar says
    return guarded("chaos_legacy_lle", [&] {
        legacy_lle(resolve_options(opt), tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
                   indicesOfMutVars, maxValue, transientTime, values, amountOfValues, dims, sweep_ic != 0, out);
    });
}

static int legacy_lle_file(const char* fn, const chaos_options* opt, double tMax, double NT, int nPts, double h,
                           double eps, const double* ic, int nIC, const double* ranges, const int* idx,
                           double maxValue, double transientTime, const double* values, int nV, int dims,
                           bool sweep_ic, const char* path)
{
    return guarded(fn, [&] {
        CHAOS_REQUIRE(path, "OUT_FILE_PATH is NULL");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        const LegacyArgs A = make_legacy_args(o, tMax, NT, nPts, h, eps, ic, nIC, ranges, idx, maxValue,
                                              transientTime, values, nV, dims, sweep_ic);
        std::vector<double> v((size_t)A.n_points);
        legacy_lle(o, tMax, NT, nPts, h, eps, ic, nIC, ranges, idx, maxValue, transientTime, values, nV, dims,
                   sweep_ic, v.data());
        if (dims == 1) {
            // debri/hostLibrary.cu:3291-3299 (LLE1D), :3508-3516 (LLE1DIC):
            // "<value>, <lle>\n" at precision 12
            write_lines(path, A, v.data(), 1, 12);
        } else {
            // debri/hostLibrary.cu:3795-3803 (LLE2DIC): the 2-D layout, no
            // header, default precision
            MapSpec m;
            m.nPts = nPts;
            m.count = A.n_points;
            m.dv = v.data();
            if (out_csv(o.output)) {
                TextFile f(path);
                write_map_text(f, m);
                f.close();
            }
            if (out_bin(o.output)) write_map_bin(std::string(path) + ".bin", m);
        }
    });
}

int chaos_LLE1D(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
                const double* initialConditions, int amountOfInitialConditions, const double* ranges,
                const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
                const double* values, int amountOfValues, const char* OUT_FILE_PATH)
{
    (void)writableVar;
    return legacy_lle_file("chaos_LLE1D", opt, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions,
                           ranges, indicesOfMutVars, maxValue, transientTime, values, amountOfValues, 1, false,
                           OUT_FILE_PATH);
}

int chaos_LLE1DIC(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
                  const double* initialConditions, int amountOfInitialConditions, const double* ranges,
                  const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
                  const double* values, int amountOfValues, const char* OUT_FILE_PATH)
{
    (void)writableVar;
    return legacy_lle_file("chaos_LLE1DIC", opt, tMax, NT, nPts, h, eps, initialConditions,
                           amountOfInitialConditions, ranges, indicesOfMutVars, maxValue, transientTime, values,
                           amountOfValues, 1, true, OUT_FILE_PATH);
}

int chaos_LLE2DIC(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
                  const double* initialConditions, int amountOfInitialConditions, const double* ranges,
                  const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
                  const double* values, int amountOfValues, const char* OUT_FILE_PATH)
{
    (void)writableVar;
    return legacy_lle_file("chaos_LLE2DIC", opt, tMax, NT, nPts, h, eps, initialConditions,
                           amountOfInitialConditions, ranges, indicesOfMutVars, maxValue, transientTime, values,
                           amountOfValues, 2, true, OUT_FILE_PATH);
}

int chaos_legacy_ls(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
                    const double* initialConditions, int amountOfInitialConditions, const double* ranges,
                    const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
                    const double* values, int amountOfValues, int dims, double* out)
{
    (void)writableVar;
    return guarded("chaos_legacy_ls", [&] {
        legacy_ls(resolve_options(opt), tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
                  indicesOfMutVars, maxValue, transientTime, values, amountOfValues, dims, out);
    });
}

static int legacy_ls_file(const char* fn, const chaos_options* opt, double tMax, double NT, int nPts, double h,
                          double eps, const double* ic, int nIC, const double* ranges, const int* idx,
                          double maxValue, double transientTime, const double* values, int nV, int dims,
                          const char* path)
{
    return guarded(fn, [&] {
        CHAOS_REQUIRE(path, "OUT_FILE_PATH is NULL");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        const LegacyArgs A = make_legacy_args(o, tMax, NT, nPts, h, eps, ic, nIC, ranges, idx, maxValue,
                                              transientTime, values, nV, dims, false);
        std::vector<double> v((size_t)A.n_points * nIC);
        legacy_ls(o, tMax, NT, nPts, h, eps, ic, nIC, ranges, idx, maxValue, transientTime, values, nV, dims,
                  v.data());
        if (dims == 1) {
            // debri/hostLibrary.cu:3931-3935 (LS1D): "<value>, <l1>, ... <lNX>\n",
            // default precision
            write_lines(path, A, v.data(), nIC, 6);
            return;
        }
        // debri/hostLibrary.cu:4034-4038, 4078-4090 (LS2D): one file per exponent,
        // OUT_FILE_PATH + "<k>.csv", two range lines then the 2-D layout
        std::vector<double> col((size_t)A.n_points);
        for (int k = 0; k < nIC; ++k) {
            for (long long p = 0; p < A.n_points; ++p) col[p] = v[p * nIC + k];
            const std::string file = std::string(path) + std::to_string(k + 1) + ".csv";
            MapSpec m;
            m.ranges = ranges;
            m.header_lines = 2;
            m.nPts = nPts;
            m.count = A.n_points;
            m.dv = col.data();
            if (out_csv(o.output)) {
                TextFile f(file);
                write_map_text(f, m);
                f.close();
            }
            if (out_bin(o.output)) write_map_bin(file + ".bin", m);
        }
    });
}

int chaos_LS1D(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
               const double* initialConditions, int amountOfInitialConditions, const double* ranges,
               const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
               const double* values, int amountOfValues, const char* OUT_FILE_PATH)
{
    (void)writableVar;
    return legacy_ls_file("chaos_LS1D", opt, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions,
                          ranges, indicesOfMutVars, maxValue, transientTime, values, amountOfValues, 1,
                          OUT_FILE_PATH);
}

int chaos_LS2D(const chaos_options* opt, double tMax, double NT, int nPts, double h, double eps,
               const double* initialConditions, int amountOfInitialConditions, const double* ranges,
               const int* indicesOfMutVars, int writableVar, double maxValue, double transientTime,
               const double* values, int amountOfValues, const char* OUT_FILE_PATH)
{
    (void)writableVar;
    return legacy_ls_file("chaos_LS2D", opt, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions,
                          ranges, indicesOfMutVars, maxValue, transientTime, values, amountOfValues, 2,
                          OUT_FILE_PATH);
}

int chaos_time_domain(const chaos_options* opt, double tMax, int nPts, double h, int amountOfInitialConditions,
                      const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                      int writableVar, double maxValue, double transientTime, const double* values,
                      int amountOfValues, int preScaller, double* out, int* flags, long long capacity,
                      long long* samples)
{
    return guarded("chaos_time_domain", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(samples, "samples is NULL");
        CHAOS_REQUIRE(preScaller >= 1, "preScaller must be >= 1");
        CHAOS_REQUIRE(writableVar >= 0 && writableVar < amountOfInitialConditions, "writableVar out of range");
        // NT = h: one step per "window" (unused here), eps only validated
        LegacyArgs A = make_legacy_args(o, tMax, h, nPts, h, 1.0, initialConditions, amountOfInitialConditions,
                                        ranges, indicesOfMutVars, maxValue, transientTime, values, amountOfValues,
                                        1, false);
        // debri/hostLibrary.cu:5111-5115: B = tMax/h/preScaller and the
        // transient count transientTime/h, both truncated
        const long long B = legacy_count(tMax / h / preScaller, o.semantics);
        *samples = B;
        const long long total = (long long)nPts * B;
        if (!out) return;
        if (total > capacity) fail(CHAOS_E_CAPACITY, "need capacity %lld (%d points x %lld samples)", total, nPts, B);
        CallCtx ctx(o);
        DevBuf<double> d_tmp;
        DevBuf<int> d_flags(nPts);
        double* d_out = out;
        if (!o.device_buffers) {
            d_tmp.alloc((size_t)std::max<long long>(total, 1));
            d_out = d_tmp.get();
        }
        hipStream_t st = ctx.stream();
        ctx.begin();
        if (total > 0) hipLaunchKernelGGL(zero_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, st, d_out, total);
        dispatch_system(o.system, [&](auto tag) {
            constexpr int S = decltype(tag)::value;
            hipLaunchKernelGGL((time_domain_kernel<S>), dim3(ceil_div(nPts, 256)), dim3(256), 0, st, A,
                               writableVar, preScaller, B, d_out, d_flags.get());
        });
        CHAOS_HIP(hipGetLastError());
        ctx.add_launch(2);
        ctx.add_kernel_ms(ctx.end());
        ctx.add_steps((double)nPts * (double)(A.S + B) * preScaller);
        if (!o.device_buffers && total > 0)
            CHAOS_HIP(hipMemcpyAsync(out, d_out, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, st));
        if (flags)
            CHAOS_HIP(hipMemcpyAsync(flags, d_flags.get(), nPts * sizeof(int),
                                     o.device_buffers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        ctx.sync();
        ctx.finish();
    });
}

int chaos_TimeDomainCalculation(const chaos_options* opt, double tMax, int nPts, double h,
                                int amountOfInitialConditions, const double* initialConditions,
                                const double* ranges, const int* indicesOfMutVars, int writableVar,
                                double maxValue, double transientTime, const double* values, int amountOfValues,
                                int preScaller, const char* OUT_FILE_PATH)
{
    return guarded("chaos_TimeDomainCalculation", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        long long B = 0;
        int rc = chaos_time_domain(&o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                   indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                   preScaller, nullptr, nullptr, 0, &B);
        if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
        std::vector<double> v((size_t)std::max<long long>(1, (long long)nPts * B));
        rc = chaos_time_domain(&o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                               indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                               preScaller, v.data(), nullptr, (long long)v.size(), &B);
        if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
        // debri/hostLibrary.cu:5283-5309: every sample at precision 16
        // followed by ", ", each point's row closed by " \n"
        TextFile f(OUT_FILE_PATH);
        for (int k = 0; k < nPts; ++k) {
            for (long long i = 0; i < B; ++i) {
                f.put_g(v[(size_t)k * B + i], 16);
                f.put(", ", 2);
            }
            f.put(" \n", 2);
        }
        f.close();
    });
}

}  // extern "C"
