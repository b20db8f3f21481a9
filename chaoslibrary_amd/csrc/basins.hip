// basins.hip -- basins of attraction over a 2-D grid of initial conditions
// (Basins/src/basinsCUDA.cu:442-813).
//
// sweep_kernel<KIND_BASINS_*> produces per-IC features and flags in one
// launch; the labels come from dbscan_device:
//   REFERENCE: (mean peak, mean interval) with the reference's sentinels,
//              single linkage at <= eps over non-divergent points, negative
//              labels for components that hold a fixed point;
//   CONFIG:    (mean X[0], mean X[NX-1], peak std) z-scored over
//              non-divergent points, DBSCAN(eps, minPts) in 3-D.
#include <cmath>
#include <string>
#include <vector>

#include "csv.hpp"
#include "dbscan.hpp"
#include "dispatch.hpp"
#include "basins.hpp"
#include "sweep.hpp"

using namespace chaos;

namespace chaos {

SweepArgs make_basins_args(const chaos_options& o, double tMax, int nPts, double h, int nx, const double* ic,
                           const double* ranges, const int* idx, int writableVar, double maxValue,
                           double transientTime, const double* values, int na, int preScaller, double eps)
{
    check_system(o.system, nx, na);
    CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2 (got %d)", nPts);
    CHAOS_REQUIRE(h > 0 && std::isfinite(h), "h must be positive");
    CHAOS_REQUIRE(preScaller >= 1, "preScaller must be >= 1");
    CHAOS_REQUIRE(ic && ranges && idx && values, "NULL input array");
    CHAOS_REQUIRE(idx[0] >= 0 && idx[0] < nx && idx[1] >= 0 && idx[1] < nx,
                  "indicesOfMutVars index the state vector here (basinsCUDA.cu:259) and must be in [0,%d)", nx);
    CHAOS_REQUIRE(writableVar >= 0 && writableVar < nx, "writableVar out of range");
    CHAOS_REQUIRE(eps > 0, "eps must be positive");
    const long long n = (long long)nPts * nPts;
    CHAOS_REQUIRE(n <= (1LL << 31) - 64, "grid too large");
    SweepArgs A{};
    for (int k = 0; k < na; ++k) A.values[k] = values[k];
    for (int k = 0; k < nx; ++k) A.ic[k] = ic[k];
    for (int k = 0; k < 4; ++k) A.r[k] = ranges[k];
    A.idx0 = idx[0];
    A.idx1 = idx[1];
    A.nPts = nPts;
    A.dims = 2;
    A.n_points = n;
    A.S = horizon_steps(transientTime, h, o.semantics);
    A.B = record_samples(tMax, h, preScaller, o.semantics);
    CHAOS_REQUIRE(A.B >= 1, "tMax/h/preScaller must give at least one recorded sample");
    A.pre = preScaller;
    A.check = o.semantics == CHAOS_SEM_REFERENCE ? CHK_BASINS_REF : CHK_CONFIG;
    A.maxValue = maxValue;
    A.L = divergence_fast_bound(A.check, maxValue);
    A.hp = h * preScaller;  // basinsCUDA.cu:684
    A.sc = make_step_const(h);
    A.pool.nchunks = 0;
    return A;
}

// CONFIG clustering buffer: the encoded box (256 bytes) and the z-score
// scratch.
size_t basins_zbuf_bytes(long long n) { return 256 + zscore_scratch_bytes(n, 3); }

void basins_sweep_partials(SweepArgs& A, void* zbuf)
{
    const ZPartials z = zscore_sweep_partials(static_cast<char*>(zbuf) + 256, A.n_points, 3);
    A.zsum = z.sum;
    A.zcnt = z.cnt;
    A.zmm = z.mm;
    A.zb = z.nb;
}

void basins_cluster(CallCtx& ctx, bool ref, long long n, const double* f0, const double* f1, const double* f2,
                    const int* flags, double eps, int minPts, int* labels, void* zbuf)
{
    if (ref) {
        const double* pts[2] = {f0, f1};
        DbscanParams prm{2, eps, 1, true};
        dbscan_device(ctx, pts, flags, n, prm, labels);
    } else {
        // clustered by their z-scores, formed on each read from the raw
        // features and the device moments (no normalised copy); the moment
        // passes leave the clustering's bounding box, and the clustering's
        // one synchronisation comes before the buffer is released
        DevBuf<char> own;
        char* buf = static_cast<char*>(zbuf);
        if (!buf) {
            own.alloc(basins_zbuf_bytes(n));
            buf = own.get();
        }
        unsigned long long* box = (unsigned long long*)buf;
        const double* src[3] = {f0, f1, f2};
        const double* mom = zscore_moments_device(ctx, src, 3, flags, n, box, buf + 256, zbuf != nullptr);
        DbscanParams prm{3, eps, minPts, false};
        dbscan_device(ctx, src, flags, n, prm, labels, nullptr, 0, box, mom);
    }
}

}  // namespace chaos

extern "C" int chaos_basins2d(const chaos_options* opt, double tMax, int nPts, double h, int amountOfInitialConditions,
                              const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                              int writableVar, double maxValue, double transientTime, const double* values,
                              int amountOfValues, int preScaller, double eps, int* labels, double* features,
                              int* flags)
{
    return guarded("chaos_basins2d", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(labels, "labels is NULL");
        SweepArgs A = make_basins_args(o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                       indicesOfMutVars, writableVar, maxValue, transientTime, values,
                                       amountOfValues, preScaller, eps);
        const long long n = A.n_points;
        const bool ref = o.semantics == CHAOS_SEM_REFERENCE;
        CallCtx ctx(o);
        DevBuf<int> d_flags(n), d_labels;
        DevBuf<double> d_f(3 * (size_t)n);
        DevBuf<char> zbuf;
        if (!ref) {
            // the sweep forms the z-score's first moments in its epilogue
            zbuf.alloc(basins_zbuf_bytes(n));
            basins_sweep_partials(A, zbuf.get());
        }
        A.flags = d_flags.get();
        A.f0 = d_f.get();
        A.f1 = d_f.get() + n;
        A.f2 = d_f.get() + 2 * n;
        int* out_labels = labels;
        if (!o.device_buffers) {
            d_labels.alloc(n);
            out_labels = d_labels.get();
        }
        ctx.begin();
        launch_sweep(ctx, o.system, writableVar, ref ? KIND_BASINS_REF : KIND_BASINS_CFG, A);
        ctx.add_kernel_ms(ctx.end());
        ctx.add_steps((double)n * ((double)A.S + (double)A.B * A.pre));
        ctx.begin();
        basins_cluster(ctx, ref, n, A.f0, A.f1, A.f2, A.flags, eps, o.min_pts, out_labels, ref ? nullptr : zbuf.get());
        ctx.add_analysis_ms(ctx.end());
        const hipMemcpyKind k = o.device_buffers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (!o.device_buffers)
            CHAOS_HIP(hipMemcpyAsync(labels, out_labels, n * sizeof(int), k, ctx.stream()));
        if (features)
            CHAOS_HIP(hipMemcpyAsync(features, d_f.get(), (ref ? 2 : 3) * n * sizeof(double), k, ctx.stream()));
        if (flags) CHAOS_HIP(hipMemcpyAsync(flags, d_flags.get(), n * sizeof(int), k, ctx.stream()));
        ctx.sync();
        ctx.finish();
    });
}

namespace chaos {

void write_basins_headers(const std::string& base, const double* ranges, int nfeat, int output)
{
    // all four (five) files get the two range lines first, at the precision
    // 15 the reference sets before writing them (basinsCUDA.cu:557-590)
    if (!out_csv(output)) return;
    TextFile f(base);
    write_range_header(f, ranges, 2, 15);
    f.close();
    for (int k = 1; k <= nfeat + 1; ++k) {
        TextFile g(base + "_" + std::to_string(k) + ".csv");
        write_range_header(g, ranges, 2, 15);
        g.close();
    }
}

void write_basins_bodies(const std::string& base, const double* ranges, int nPts, int nfeat, const int* labels,
                         const double* feat, const int* flags, int output)
{
    const long long n = (long long)nPts * nPts;
    auto emit = [&](const std::string& path, const double* dv, const int* iv) {
        MapSpec m;
        m.ranges = ranges;
        m.header_precision = 15;
        m.nPts = nPts;
        m.count = n;
        m.precision = 15;
        m.dv = dv;
        m.iv = iv;
        if (out_csv(output)) {  // the headers were written before
            TextFile g(path, true);
            write_map_text(g, m);
            g.close();
        }
        m.header_lines = 2;
        if (out_bin(output)) write_map_bin(path + ".bin", m);
    };
    // OUT: labels; _1: feature 0, _2: feature 1, _3: flags (:733-792);
    // CONFIG adds _4: feature 2
    emit(base, nullptr, labels);
    for (int k = 0; k < 2; ++k) emit(base + "_" + std::to_string(k + 1) + ".csv", feat + k * n, nullptr);
    emit(base + "_3.csv", nullptr, flags);
    if (nfeat > 2) emit(base + "_4.csv", feat + 2 * n, nullptr);
}

}  // namespace chaos

extern "C" int chaos_basinsOfAttraction_2(const chaos_options* opt, double tMax, int nPts, double h,
                                          int amountOfInitialConditions, const double* initialConditions,
                                          const double* ranges, const int* indicesOfMutVars, int writableVar,
                                          double maxValue, double transientTime, const double* values,
                                          int amountOfValues, int preScaller, double eps,
                                          const char* OUT_FILE_PATH)
{
    return guarded("chaos_basinsOfAttraction_2", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        const bool ref = o.semantics == CHAOS_SEM_REFERENCE;
        const long long n = (long long)nPts * nPts;
        const std::string base(OUT_FILE_PATH);
        const int nfeat = ref ? 2 : 3;
        write_basins_headers(base, ranges, nfeat, o.output);
        std::vector<int> labels(n), flags(n);
        std::vector<double> feat(nfeat * n);
        int rc = chaos_basins2d(&o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                preScaller, eps, labels.data(), feat.data(), flags.data());
        if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
        write_basins_bodies(base, ranges, nPts, nfeat, labels.data(), feat.data(), flags.data(), o.output);
    });
}
