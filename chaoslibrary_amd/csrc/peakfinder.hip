// peakfinder.hip -- the online peak detector applied to caller-supplied
// sample rows: Bifurcation/src/bifurcationCUDA.cu:469-514 (peakFinder) with
// one lane per row.  Row r's kept peaks (first dropped) and intervals go to
// peaks[r*len + k], intervals[r*len + k]; counts[r] = number kept.
#include "peaks.hpp"
#include "runtime.hpp"

namespace chaos {

__global__ void peak_rows_kernel(const double* __restrict__ data, int rows, int len, double h, int* __restrict__ counts,
                                 double* __restrict__ peaks, double* __restrict__ intervals)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const double* d = data + (size_t)r * len;
    PeakDetector det;
    det.reset();
    int found = 0, prev_t = 0;
    for (int t = 0; t < len; ++t) {
        double v = 0.0;
        int tt = 0;
        if (det.feed(t, d[t], v, tt)) {
            if (found >= 1) {
                const size_t o = (size_t)r * len + (found - 1);
                peaks[o] = v;
                intervals[o] = (double)(tt - prev_t) * h;
            }
            prev_t = tt;
            ++found;
        }
    }
    counts[r] = found > 1 ? found - 1 : 0;
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_peak_finder(const chaos_options* opt, const double* data, int rows, int len, double h,
                                 int* counts, double* peaks, double* intervals)
{
    return guarded("chaos_peak_finder", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(rows >= 0 && len >= 0, "rows and len must be >= 0");
        CHAOS_REQUIRE(rows == 0 || (data && counts && peaks && intervals), "NULL buffer");
        if (rows == 0) return;
        CallCtx ctx(o);
        const size_t nd = (size_t)rows * (size_t)(len > 0 ? len : 1);
        DevBuf<double> dd, dp, di;
        DevBuf<int> dc;
        const double* d_data = data;
        double *d_pk = peaks, *d_iv = intervals;
        int* d_cnt = counts;
        if (!o.device_buffers) {
            dd.alloc(nd);
            dp.alloc(nd);
            di.alloc(nd);
            dc.alloc(rows);
            if (len > 0) CHAOS_HIP(hipMemcpyAsync(dd.get(), data, (size_t)rows * len * 8, hipMemcpyHostToDevice, ctx.stream()));
            d_data = dd.get();
            d_pk = dp.get();
            d_iv = di.get();
            d_cnt = dc.get();
        }
        ctx.begin();
        hipLaunchKernelGGL(peak_rows_kernel, dim3(ceil_div(rows, 64)), dim3(64), 0, ctx.stream(), d_data, rows, len, h,
                           d_cnt, d_pk, d_iv);
        CHAOS_HIP(hipGetLastError());
        ctx.add_launch();
        ctx.add_analysis_ms(ctx.end());
        if (!o.device_buffers) {
            CHAOS_HIP(hipMemcpyAsync(counts, d_cnt, rows * sizeof(int), hipMemcpyDeviceToHost, ctx.stream()));
            if (len > 0) {
                CHAOS_HIP(hipMemcpyAsync(peaks, d_pk, (size_t)rows * len * 8, hipMemcpyDeviceToHost, ctx.stream()));
                CHAOS_HIP(hipMemcpyAsync(intervals, d_iv, (size_t)rows * len * 8, hipMemcpyDeviceToHost, ctx.stream()));
            }
        }
        ctx.sync();
        ctx.finish();
    });
}
