// bifurcation.hip -- Bifurcation 1-D and 2-D (Bifurcation/src/bifurcationCUDA.cu).
//
// Pipeline per call:
//   sweep_kernel<KIND_BIF>   transient + record + online peaks; every peak is
//                            appended to its wave's chunk (ballot + mbcnt)
//   kept_kernel + scan       per-point segment offsets (flag == 1 points)
//   scatter_kernel           chunk entries -> per-point contiguous (value, time)
//   1-D: host assembles the kept peaks / intervals (first peak dropped)
//   2-D: one wave per point, points listed on the device by power-of-two
//        size class -- CONFIG: sort (bitonic over 64 lanes x R registers up
//        to 2048 peaks, in LDS up to 4096, in global memory beyond) + gap
//        count = 1-D DBSCAN(eps, minPts=1) cluster count; REFERENCE: the
//        reference's per-point dbscan on the kept intervals.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <vector>

#include "csv.hpp"
#include "dispatch.hpp"
#include "peak_ranges.hpp"
#include "scan.hpp"
#include "bifurcation.hpp"
#include "sweep.hpp"

namespace chaos {

__global__ void kept_kernel(const int* __restrict__ flags, const int* __restrict__ counts,
                            long long* __restrict__ kept, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) kept[p] = flags[p] == 1 ? counts[p] : 0;
}

__global__ void scatter_kernel(const PeakPool pool, unsigned used, const int* __restrict__ flags,
                               const long long* __restrict__ off, double* __restrict__ vals,
                               unsigned* __restrict__ times, long long n)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned c = (unsigned)(g / kChunk);
    const int s = (int)(g % kChunk);
    if (c >= used || s >= pool.fill()[c]) return;
    const PeakEntry e = pool.entries[(size_t)c * kChunk + s];
    const long long p = (long long)pool.owner()[c] * 64 + (e.kl & 63u);
    if (p >= n || flags[p] != 1) return;
    const long long pos = off[p] + (e.kl >> 6);
    if (vals) vals[pos] = e.v;
    if (times) times[pos] = e.t;
}

void PeakPoolBuf::alloc(hipStream_t st, long long chunks)
{
    entries.alloc((size_t)chunks * kChunk);
    meta.alloc((size_t)(2 * chunks + 2));
    cap = nchunks = chunks;
    // head and overflow start at zero
    CHAOS_HIP(hipMemsetAsync(meta.get() + 2 * chunks, 0, 2 * sizeof(int), st));
}

void PeakPoolBuf::prepare(hipStream_t st, long long chunks)
{
    if (!meta.get() || chunks > cap) {
        alloc(st, chunks);
        return;
    }
    nchunks = chunks;
    // head and overflow (after this store's owner and fill words) start at zero
    CHAOS_HIP(hipMemsetAsync(meta.get() + 2 * chunks, 0, 2 * sizeof(int), st));
}

PeakPool PeakPoolBuf::view() const
{
    PeakPool p;
    p.entries = entries.get();
    p.meta = meta.get();
    p.nchunks = (unsigned)nchunks;
    return p;
}

SweepArgs make_bif_args(const chaos_options& o, const BifInputs& in)
{
    check_system(o.system, in.nx, in.na);
    CHAOS_REQUIRE(in.nPts >= 1, "nPts must be >= 1 (got %d)", in.nPts);
    CHAOS_REQUIRE(in.h > 0 && std::isfinite(in.h), "h must be positive");
    CHAOS_REQUIRE(in.preScaller >= 1, "preScaller must be >= 1");
    CHAOS_REQUIRE(in.ic && in.ranges && in.values && in.idx, "NULL input array");
    const int lim = in.legacy_ic ? in.nx : in.na;
    for (int d = 0; d < in.dims; ++d)
        CHAOS_REQUIRE(in.idx[d] >= 0 && in.idx[d] < lim, "indicesOfMutVars[%d]=%d out of range [0,%d)", d,
                      in.idx[d], lim);
    CHAOS_REQUIRE(in.writableVar >= 0 && in.writableVar < in.nx, "writableVar out of range");
    const long long n = in.dims == 2 ? (long long)in.nPts * in.nPts : in.nPts;
    CHAOS_REQUIRE(n <= (1LL << 31) - 64, "grid too large");
    SweepArgs A{};
    for (int k = 0; k < in.na; ++k) A.values[k] = in.values[k];
    for (int k = 0; k < in.nx; ++k) A.ic[k] = in.ic[k];
    for (int k = 0; k < 2 * in.dims; ++k) A.r[k] = in.ranges[k];
    A.idx0 = in.idx[0];
    A.idx1 = in.dims == 2 ? in.idx[1] : -1;
    A.nPts = in.nPts;
    A.dims = in.dims;
    A.n_points = n;
    A.S = horizon_steps(in.transientTime, in.h, o.semantics);
    A.B = record_samples(in.tMax, in.h, in.preScaller, o.semantics);
    CHAOS_REQUIRE(A.B >= 1, "tMax/h/preScaller must give at least one recorded sample");
    A.pre = 1;  // Bifurcation steps once per sample (bifurcationCUDA.cu:435-443)
    A.check = o.semantics == CHAOS_SEM_REFERENCE ? CHK_BIF_REF : CHK_CONFIG;
    if (in.legacy_ic) {
        // the legacy loop (debri/cudaLibrary.cu:604-680) steps preScaller
        // times per sample and tests isnan/isinf and fabsf(x) > maxValue
        A.pre = in.preScaller;
        A.check = o.semantics == CHAOS_SEM_REFERENCE ? CHK_BASINS_REF : CHK_CONFIG;
        A.ic_sweep = 1;
        A.legacy_transient = 1;
    }
    A.maxValue = in.maxValue;
    A.L = divergence_fast_bound(A.check, in.maxValue);
    A.hp = in.h * (double)in.preScaller;  // bifurcationCUDA.cu:196,726
    A.sc = make_step_const(in.h);
    return A;
}

// Device bytes of one chunk of the peak store (entries + its two meta words).
constexpr size_t kChunkBytes = kChunk * sizeof(PeakEntry) + 2 * sizeof(int);

// Cap on the device memory of one sweep's peak store: CHAOS_PEAK_STORE_BYTES
// when the process sets it, else none.
static size_t peak_store_cap()
{
    const char* s = std::getenv("CHAOS_PEAK_STORE_BYTES");
    if (!s || !*s) return SIZE_MAX;
    const long long v = std::atoll(s);
    return v > 0 ? (size_t)v : SIZE_MAX;
}

// First guess of the peaks per point: one per 32 samples (at most one per
// two is possible); an overflow is detected on the device and the sweep
// rerun with the exact per-wave need.
static double first_guess_per_point(const SweepArgs& A)
{
    return std::min((double)A.B / 2.0 + 1.0, std::max(256.0, (double)A.B / 32.0));
}

// Chunks the first-guess store may take: a sixth of the device memory left
// (47 GB on a 288 GB device: a 1024^2 peak-count map in one pass, 2048^2 in
// three), or CHAOS_PEAK_STORE_BYTES.  A store that size stays in the device
// allocator's cache (a quarter of the device) beside the values scattered
// out of it, so repeated calls reuse it: mapping a 70 GB store anew took
// 1.2-1.8 s per call.
static long long first_guess_budget_chunks()
{
    const size_t budget = std::min((size_t)((double)dev_mem_available() / 6.0), peak_store_cap());
    return (long long)(budget / kChunkBytes);
}

long long bif_initial_chunks(const SweepArgs& A)
{
    const long long n = A.n_points, waves = (n + 63) / 64;
    const long long want = (long long)std::ceil((double)n * first_guess_per_point(A) / (kChunk - 64)) + waves;
    return std::max<long long>(waves, std::min(want, first_guess_budget_chunks()));
}

long long bif_exact_chunks(const std::vector<int>& counts, long long n)
{
    long long need = 0;
    for (long long w = 0; w < n; w += 64) {
        long long s = 0;
        for (long long p = w; p < std::min(n, w + 64); ++p) s += counts[p];
        need += wave_chunks(s, kChunk);
    }
    return need;
}

// Exact passes over ranges of whole waves that each fit `budget`
// (exact_ranges, peak_ranges.hpp); fails when one wave alone does not.
static std::vector<long long> exact_pass_ranges(const std::vector<int>& counts, long long n, size_t budget)
{
    size_t wave_bytes = 0;
    std::vector<long long> starts = exact_ranges(counts.data(), n, budget, kChunk, kChunkBytes, &wave_bytes);
    if (starts.empty())
        fail(CHAOS_E_NOMEM, "the peaks of 64 points need %zu bytes of device memory, more than the %zu available",
             wave_bytes, budget);
    return starts;
}

bool bif_post(CallCtx& ctx, BifDevice& R, PeakPoolBuf& pool)
{
    const long long n = R.n;
    unsigned used = 0;
    int overflow = 0;
    const PeakPool v = pool.view();
    CHAOS_HIP(hipMemcpyAsync(&used, v.head(), sizeof used, hipMemcpyDeviceToHost, ctx.stream()));
    CHAOS_HIP(hipMemcpyAsync(&overflow, v.overflow(), sizeof overflow, hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    if (overflow || R.need_host) {
        // the 1-D results are assembled on the host; an overflow sizes the
        // rerun's pool from the exact per-point counts
        R.h_counts.resize(n);
        CHAOS_HIP(hipMemcpyAsync(R.h_counts.data(), R.counts.get(), n * sizeof(int), hipMemcpyDeviceToHost,
                                 ctx.stream()));
        if (R.need_host) {
            R.h_flags.resize(n);
            R.h_first.resize(n);
            CHAOS_HIP(hipMemcpyAsync(R.h_flags.data(), R.flags.get(), n * sizeof(int), hipMemcpyDeviceToHost,
                                     ctx.stream()));
            CHAOS_HIP(hipMemcpyAsync(R.h_first.data(), R.first.get(), n * sizeof(double), hipMemcpyDeviceToHost,
                                     ctx.stream()));
        }
        ctx.sync();
    }
    if (overflow) return false;
    ctx.begin();
    DevBuf<long long> kept((size_t)n);
    R.off.alloc(n);
    hipLaunchKernelGGL(kept_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), R.flags.get(),
                       R.counts.get(), kept.get(), n);
    CHAOS_HIP(hipGetLastError());
    R.total = exclusive_scan(ctx.stream(), kept.get(), R.off.get(), n);
    if (R.need_vals) R.vals.alloc((size_t)std::max(1LL, R.total));
    else R.vals.release();
    if (R.need_times) R.times.alloc((size_t)std::max(1LL, R.total));
    else R.times.release();
    if (used > (unsigned)pool.nchunks) used = (unsigned)pool.nchunks;
    if (used > 0) {
        hipLaunchKernelGGL(scatter_kernel, dim3(ceil_div((unsigned long long)used * kChunk, 256)), dim3(256), 0,
                           ctx.stream(), pool.view(), used, R.flags.get(), R.off.get(), R.vals.get(), R.times.get(),
                           n);
        CHAOS_HIP(hipGetLastError());
    }
    ctx.add_analysis_ms(ctx.end());
    ctx.add_launch(2);
    ctx.add_peaks(R.total);
    return true;
}

// One sweep of the points [A.p0, A.p0 + A.n_points) into R with a peak store
// of nchunks chunks (in `pool`, which the call's passes share); false when
// the store overflowed (R.h_counts then hold every point's exact count).
static bool sweep_pass(CallCtx& ctx, const chaos_options& o, const BifInputs& in, SweepArgs A, BifDevice& R,
                       long long nchunks, PeakPoolBuf& pool)
{
    const long long n = A.n_points;
    R.n = n;
    R.hp = A.hp;
    R.samples = A.B;
    R.flags.alloc(n);
    R.counts.alloc(n);
    R.first.alloc(n);
    A.flags = R.flags.get();
    A.counts = R.counts.get();
    A.first = R.first.get();
    pool.prepare(ctx.stream(), nchunks);
    A.pool = pool.view();
    ctx.begin();
    // a map that clusters peak values only leaves the detector's time logic out
    launch_sweep(ctx, o.system, in.writableVar, R.need_times ? KIND_BIF : KIND_BIF_VALS, A);
    ctx.add_kernel_ms(ctx.end());
    ctx.add_steps((double)n * ((double)A.S + (double)A.B * A.pre));
    return bif_post(ctx, R, pool);
}

// Sweeps the grid and hands each pass's results to use(R, p0), p0 the grid
// index of the pass's first point; R holds the last pass.
//   * First passes: the whole grid; or, with `ranges` and when the
//     first-guess store of the whole grid exceeds its budget (a quarter of
//     the device memory left, or CHAOS_PEAK_STORE_BYTES), ranges of whole
//     waves whose first guesses fit.
//   * A first pass whose store overflowed runs again with the exact store:
//     over its whole range, or, with `ranges` and when that store and its
//     post stages need more than 60 % of the device memory left (or than
//     CHAOS_PEAK_STORE_BYTES), over ranges of whole waves that each fit.
// The grid is cut by the memory as the reference sizes its chunks of the
// grid by the free memory (Bifurcation/src/bifurcationCUDA.cu:64-70).
template <class Use>
static void run_bif(CallCtx& ctx, const chaos_options& o, const BifInputs& in, BifDevice& R, bool ranges, Use&& use)
{
    const SweepArgs A = make_bif_args(o, in);
    const long long n = A.n_points;
    const double wave_guess = 64.0 * first_guess_per_point(A) / (kChunk - 64) + 1.0;
    const std::vector<long long> first = ranges ? first_pass_cuts(n, wave_guess, first_guess_budget_chunks())
                                                : std::vector<long long>{0, n};
    const bool single = first.size() == 2;
    PeakPoolBuf pool;  // one store for every pass, grown when a pass needs more
    // R for a pass over the whole grid, else a record of the pass's own
    auto record = [&](BifDevice& own, bool whole) -> BifDevice& {
        if (whole) return R;
        own.need_times = R.need_times;
        own.need_vals = R.need_vals;
        own.need_host = R.need_host;
        return own;
    };
    auto args = [&](long long a, long long b) {
        SweepArgs Ak = A;
        Ak.p0 = a;
        Ak.n_points = b - a;
        return Ak;
    };
    for (size_t k = 0; k + 1 < first.size(); ++k) {
        const long long a = first[k], b = first[k + 1];
        BifDevice own;
        BifDevice& dst = record(own, single);
        const SweepArgs Ak = args(a, b);
        if (sweep_pass(ctx, o, in, Ak, dst, bif_initial_chunks(Ak), pool)) {
            use(dst, a);
            continue;
        }
        // exact passes over this range from its per-point counts
        const std::vector<int> counts = dst.h_counts;
        dst.flags.release();
        dst.counts.release();
        dst.first.release();
        pool = PeakPoolBuf();  // the exact passes size the store anew
        const long long m = b - a;
        const size_t budget = std::min((size_t)(0.6 * (double)dev_mem_available()), peak_store_cap());
        const std::vector<long long> cuts = ranges ? exact_pass_ranges(counts, m, budget) : std::vector<long long>{0, m};
        std::vector<long long> need(cuts.size() - 1);
        for (size_t j = 0; j + 1 < cuts.size(); ++j)
            need[j] = bif_exact_chunks(std::vector<int>(counts.begin() + cuts[j], counts.begin() + cuts[j + 1]),
                                       cuts[j + 1] - cuts[j]);
        pool.prepare(ctx.stream(), *std::max_element(need.begin(), need.end()));
        for (size_t j = 0; j + 1 < cuts.size(); ++j) {
            const long long c = cuts[j], d = cuts[j + 1];
            BifDevice own2;
            BifDevice& dst2 = record(own2, single && cuts.size() == 2);
            CHAOS_REQUIRE(sweep_pass(ctx, o, in, args(a + c, a + d), dst2, need[j], pool),
                          "peak store overflow after resize");
            use(dst2, a + c);
        }
    }
}

// ---------------------------------------------------------------------------
// 2-D analyses, one wave (64-thread block) per point
// ---------------------------------------------------------------------------

// In-place ascending bitonic sort of M (power of two) doubles shared by the
// block's threads; sync() orders the stages.
template <class T>
__device__ __forceinline__ void bitonic_sort(T* v, int M)
{
    for (int size = 2; size <= M; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int idx = threadIdx.x; idx < (M >> 1); idx += blockDim.x) {
                const int lo = 2 * idx - (idx & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const T a = v[lo], b = v[hi];
                if ((a > b) == up) {
                    v[lo] = b;
                    v[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

// 1 + #{sorted gaps > eps} over v[0..P) (sorted).
__device__ __forceinline__ int gap_count(const double* v, int P, double eps)
{
    int gaps = 0;
    for (int k = 1 + threadIdx.x; k < P; k += blockDim.x) gaps += (v[k] - v[k - 1] > eps) ? 1 : 0;
    // block of one wave: reduce with ballot-free shuffle sum
    for (int o = 32; o > 0; o >>= 1) gaps += __shfl_down(gaps, o);
    return 1 + __shfl(gaps, 0);
}

// CONFIG 2-D, every point: -1 for a fixed point, else 0; the points with
// kept peaks are overwritten by the 1-D DBSCAN kernels.
__global__ void flag_out_kernel(const int* __restrict__ flags, int* __restrict__ out, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = flags[p] == -1 ? -1 : 0;
}

// CONFIG 2-D: number of 1-D DBSCAN(eps, minPts=1) clusters of the kept peak
// values (first peak dropped) of the listed points, which all have
// 1 <= P <= M kept peaks: one wave per point, LDS sized for this size class
// (the launches go by powers of two, so a map's few long segments do not
// set the LDS footprint -- and the occupancy -- of all the others).
__global__ __launch_bounds__(64) void dbscan1d_list_kernel(const int* __restrict__ list, int nlist,
                                                           const int* __restrict__ counts,
                                                           const long long* __restrict__ off,
                                                           const double* __restrict__ vals, double eps, int M,
                                                           int* __restrict__ out)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    if (b >= nlist) return;
    const int p = list[b];
    const int P = counts[p] - 1;
    const double* v = vals + off[p] + 1;
    for (int k = threadIdx.x; k < M; k += blockDim.x) lds[k] = k < P ? v[k] : INFINITY;
    __syncthreads();
    bitonic_sort(lds, M);
    const int c = gap_count(lds, P, eps);
    if (threadIdx.x == 0) out[p] = c;
}

// Bitonic sort of a wave's 64 * R values held in registers: element
// e = lane * R + r.  Pairs closer than R sit in one lane (compare-exchange
// of registers); farther pairs are one lane-xor apart and swap through a
// shuffle, each lane keeping the minimum or the maximum.
template <int R, typename V>
__device__ __forceinline__ void wave_bitonic(V (&v)[R])
{
    const int lane = threadIdx.x & 63;
    constexpr int M = 64 * R;
#pragma unroll
    for (int size = 2; size <= M; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= R) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = lane * R + r;
                    const V o = __shfl_xor(v[r], stride / R);
                    const bool take_min = ((e & stride) == 0) == ((e & size) == 0);
                    const bool less = v[r] < o;
                    const V mn = less ? v[r] : o;
                    const V mx = less ? o : v[r];
                    v[r] = take_min ? mn : mx;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if ((r & stride) == 0) {
                        const bool up = ((lane * R + r) & size) == 0;
                        const V a = v[r], b = v[r + stride];
                        const bool gt = a > b;
                        const V lo = gt ? b : a, hi = gt ? a : b;
                        v[r] = up ? lo : hi;
                        v[r + stride] = up ? hi : lo;
                    }
                }
            }
        }
    }
}

// CONFIG 2-D count for the listed points with P <= 64 * R kept peaks, the
// sort held in registers (no LDS): one wave per point, then
// 1 + #{k : v[k] - v[k-1] > eps} over the sorted values.
template <int R>
__global__ __launch_bounds__(64) void dbscan1d_reg_kernel(const int* __restrict__ list, int nlist,
                                                          const int* __restrict__ counts,
                                                          const long long* __restrict__ off,
                                                          const double* __restrict__ vals, double eps,
                                                          int* __restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= nlist) return;
    const int p = list[b];
    const int P = counts[p] - 1;
    const double* src = vals + off[p] + 1;
    const int lane = threadIdx.x & 63;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = lane * R + r;
        v[r] = e < P ? src[e] : INFINITY;
    }
    wave_bitonic<R>(v);
    const double prev = __shfl_up(v[R - 1], 1);  // last element of the lane below
    int gaps = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = lane * R + r;
        const double before = r == 0 ? prev : v[r - 1];
        if (e >= 1 && e < P && v[r] - before > eps) ++gaps;
    }
    for (int o = 32; o > 0; o >>= 1) gaps += __shfl_down(gaps, o);
    if (lane == 0) out[p] = 1 + gaps;
}

__global__ __launch_bounds__(64) void dbscan1d_big_kernel(const int* __restrict__ counts,
                                                          const long long* __restrict__ off,
                                                          const double* __restrict__ vals,
                                                          const int* __restrict__ big, int nbig,
                                                          unsigned char* __restrict__ scratch,
                                                          const long long* __restrict__ scratch_off, double eps,
                                                          int* __restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= nbig) return;
    const long long p = big[b];
    const int P = counts[p] - 1;
    int M = 1;
    while (M < P) M <<= 1;
    double* w = reinterpret_cast<double*>(scratch + scratch_off[b]);
    const double* v = vals + off[p] + 1;
    for (int k = threadIdx.x; k < M; k += blockDim.x) w[k] = k < P ? v[k] : INFINITY;
    __syncthreads();
    bitonic_sort(w, M);
    const int c = gap_count(w, P, eps);
    if (threadIdx.x == 0) out[p] = c;
}

// REFERENCE 2-D: Bifurcation/src/bifurcationCUDA.cu:516-572 over the kept
// intervals of one point, one wave per point.
//
// The reference walks every k per pass (O(P^2)).  Here the decisions are the
// same, made in bulk:
//  * distance() sees zeroed peak values, so it is hypotf(0, dy) =
//    |float(dy)| (:575-583) and depends only on the two intervals; an
//    interval is (t[k+1]-t[k]) * hp, a function of the integer sample gap
//    dt.  Indices are grouped into buckets of equal dt, sorted by dt.
//  * a pass centred on y_i assigns every unassigned candidate (k < P-1, the
//    reference's `k < amountOfPeaks - 1`) whose interval is within eps: whole
//    buckets, and since |float(y_b - y_i)| grows monotonically away from
//    y_i, a contiguous run of them around y_i.  Buckets left without
//    unassigned candidates are unlinked (union-find "nearest live bucket"
//    links), so a pass costs the buckets it empties plus two.
//  * helpful[k < P] only goes from 0 to a cluster number, so a second pass
//    centred on the same interval value finds nothing; such passes (most
//    stack pops) are skipped, and runs of them are taken 64 at a time: a pop
//    reads slot P+NumNeibor-1, clears it and moves the loop index, exactly as
//    the reference's quirky stack does.
// T is the index type of the scratch arrays (16-bit in LDS, 32-bit in global
// memory); keys holds C >= P slots (C a power of two) of type K: 64-bit
// (dt << 32) | k, or 32-bit (dt << 16) | k when every gap dt and index k fits
// 16 bits -- the same order in half the LDS, so more waves share a CU.
template <typename T, typename K>
struct RefScratch {
    K* keys;                   // (dt << shift) | k, sorted
    T* h;                      // [2P] the reference's helpfulArray
    T* bucket_of;              // [P] bucket of index k
    T* bstart;                 // [nb + 1] first sorted position of bucket b
    T* lp;                     // [nb] nearest live bucket at or left of b (or b-1 chain)
    T* rp;                     // [nb] same to the right
    unsigned char* expd;       // [nb] a pass centred on bucket b has run
};

template <typename T>
__device__ __forceinline__ int ref_find_left(T* lp, int b)
{
    while (b >= 0) {
        const int p = __builtin_amdgcn_readfirstlane((int)lp[b]);
        if (p == b) return b;
        if (p >= 0) {
            const int g = __builtin_amdgcn_readfirstlane((int)lp[p]);
            if (g != p && threadIdx.x == 0) lp[b] = (T)g;  // path halving
        }
        b = p;
    }
    return -1;
}

template <typename T>
__device__ __forceinline__ int ref_find_right(T* rp, int b, int nb)
{
    while (b < nb) {
        const int p = __builtin_amdgcn_readfirstlane((int)rp[b]);
        if (p == b) return b;
        if (p < nb) {
            const int g = __builtin_amdgcn_readfirstlane((int)rp[p]);
            if (g != p && threadIdx.x == 0) rp[b] = (T)g;
        }
        b = p;
    }
    return nb;
}

// RS > 0: the keys are sorted in registers (wave_bitonic over 64 * RS >= C
// slots, a power of two), then the first C stored in order; RS = 0: sorted
// in place in the scratch array (C a power of two).
template <typename T, typename K, int RS>
__device__ int dbscan_ref_wave(const unsigned* __restrict__ t, int P, double hp, double eps, RefScratch<T, K> S,
                               int C)
{
    constexpr int SH = (int)sizeof(K) * 4;
    constexpr K LO = (K(1) << SH) - 1;
    const int lane = threadIdx.x;
    K* keys = S.keys;
    if constexpr (RS > 0) {
        K kv[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int e = lane * RS + r;
            kv[r] = e < P ? ((K)(t[e + 1] - t[e]) << SH) | (K)e : ~K(0);
        }
        wave_bitonic<RS>(kv);
#pragma unroll
        for (int r = 0; r < RS; ++r)
            if (lane * RS + r < C) keys[lane * RS + r] = kv[r];
    } else {
        for (int s = lane; s < C; s += 64)
            keys[s] = s < P ? ((K)(t[s + 1] - t[s]) << SH) | (K)s : ~K(0);
    }
    for (int k = lane; k < 2 * P; k += 64) S.h[k] = 0;
    __syncthreads();
    if constexpr (RS == 0) bitonic_sort(keys, C);
    // buckets of equal dt over the sorted positions [0, P)
    int nb = 0;
    for (int base = 0; base < P; base += 64) {
        const int s = base + lane;
        bool head = false;
        if (s < P) head = s == 0 || (keys[s] >> SH) != (keys[s - 1] >> SH);
        const unsigned long long m = __ballot(head);
        const int bid = nb + (int)__popcll(m & ((2ull << lane) - 1ull)) - 1;
        if (s < P) {
            if (head) S.bstart[bid] = (T)s;
            S.bucket_of[(int)(keys[s] & LO)] = (T)bid;
        }
        nb += (int)__popcll(m);
    }
    __syncthreads();
    if (lane == 0) S.bstart[nb] = (T)P;
    for (int b = lane; b < nb; b += 64) S.lp[b] = 0;
    __syncthreads();
    // live: the bucket holds a candidate (k < P-1)
    for (int s = lane; s < P; s += 64)
        if ((int)(keys[s] & LO) < P - 1) S.lp[(int)S.bucket_of[(int)(keys[s] & LO)]] = 1;
    __syncthreads();
    for (int b = lane; b < nb; b += 64) {
        const bool live = S.lp[b] != 0;
        S.lp[b] = (T)(live ? b : b - 1);
        S.rp[b] = (T)(live ? b : b + 1);
        S.expd[b] = 0;
    }
    __syncthreads();

    int cluster = 0, num = 0;
    // every unassigned candidate of bucket b joins the cluster; b is unlinked
    auto assign = [&](int b, int i) -> int {
        const int s0 = S.bstart[b], s1 = S.bstart[b + 1];
        int found = 0;
        for (int s = s0 + lane; s - lane < s1; s += 64) {
            bool hit = false;
            if (s < s1) {
                const int k = (int)(keys[s] & LO);
                if (k < P - 1 && k != i && S.h[k] == 0) {
                    S.h[k] = (T)cluster;
                    S.h[P + k] = (T)k;
                    hit = true;
                }
            }
            found += (int)__popcll(__ballot(hit));
        }
        if (lane == 0 && (int)S.lp[b] == b) {
            S.lp[b] = (T)(b - 1);
            S.rp[b] = (T)(b + 1);
        }
        __syncthreads();
        return found;
    };
    auto pass = [&](int i) {
        const int b0 = __builtin_amdgcn_readfirstlane((int)S.bucket_of[i]);
        if (__builtin_amdgcn_readfirstlane((int)S.expd[b0])) return;  // found nothing
        if (lane == 0) S.expd[b0] = 1;
        const double yi = (double)((long long)t[i + 1] - (long long)t[i]) * hp;
        auto qualifies = [&](int b) {
            const double yb = (double)(long long)(keys[(int)S.bstart[b]] >> SH) * hp;
            const double d = (yb == yi) ? 0.0 : (double)fabsf((float)(yb - yi));
            return __builtin_amdgcn_readfirstlane((int)(d < eps)) != 0;
        };
        int found = 0;
        if (__builtin_amdgcn_readfirstlane((int)S.lp[b0]) == b0 && qualifies(b0)) found += assign(b0, i);
        for (int b = ref_find_left(S.lp, b0 - 1); b >= 0 && qualifies(b); b = ref_find_left(S.lp, b - 1))
            found += assign(b, i);
        for (int b = ref_find_right(S.rp, b0 + 1, nb); b < nb && qualifies(b); b = ref_find_right(S.rp, b + 1, nb))
            found += assign(b, i);
        num += found;
    };

    int i = 0;
    while (i < P) {
        if (num >= 1) {
            // the next up-to-64 pops: slot P+num-1-q holds index num-1-q or 0
            const int cnt = num < 64 ? num : 64;
            int v = 0;
            bool eff = false;
            if (lane < cnt) {
                v = (int)S.h[P + num - 1 - lane];
                eff = S.expd[(int)S.bucket_of[v]] == 0;
            }
            const unsigned long long em = __ballot(eff);
            const int q0 = em ? (int)__builtin_ctzll(em) : cnt;
            if (lane < q0) S.h[P + num - 1 - lane] = 0;
            if (q0 > 0) {
                // q0 pops whose passes find nothing: the loop index ends at
                // the last popped value, then steps on
                i = __shfl(v, q0 - 1) + 1;
                num -= q0;
                __syncthreads();
                continue;
            }
            i = __builtin_amdgcn_readfirstlane(v);
            if (lane == 0) S.h[P + num - 1] = 0;
            num -= 1;
            __syncthreads();
            pass(i);
            ++i;
        } else {
            // next index still unassigned: a new cluster starts there
            int j = P;
            for (int base = i; base < P; base += 64) {
                const int k = base + lane;
                const unsigned long long zm = __ballot(k < P && S.h[k] == 0);
                if (zm) {
                    j = base + (int)__builtin_ctzll(zm);
                    break;
                }
            }
            if (j >= P) break;
            i = j;
            ++cluster;
            if (lane == 0) S.h[i] = (T)cluster;
            __syncthreads();
            pass(i);
            ++i;
        }
    }
    return cluster - 1;
}

// Outputs that need no clustering: flag -1 -> -1, 0 -> 0 (dbscanCUDA,
// :387-395), and P <= 1 kept intervals -> P (:521-525).
__global__ void dbscan_ref_trivial_kernel(const int* __restrict__ flags, const int* __restrict__ counts,
                                          int* __restrict__ out, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int fl = flags[p];
    const int P = counts[p] > 1 ? counts[p] - 1 : 0;
    if (fl != 1) out[p] = fl == -1 ? -1 : 0;
    else if (P <= 1) out[p] = P;
}

// LDS bytes of dbscan_ref_lds_kernel for capacity C: keys of key_bytes, 16-bit
// helpful[2C], bucket_of, bstart (C + 1), lp, rp and the expanded flags.
constexpr size_t ref_lds_bytes(int C, int key_bytes)
{
    return (size_t)C * (key_bytes + 2 * 2 + 2 + 2 + 2 + 2 + 1) + 2 + 8;
}

template <typename T, typename K>
__device__ __forceinline__ RefScratch<T, K> ref_scratch(void* base, int C)
{
    RefScratch<T, K> S;
    S.keys = static_cast<K*>(base);
    S.h = reinterpret_cast<T*>(S.keys + C);
    S.bucket_of = S.h + 2 * C;
    S.bstart = S.bucket_of + C;
    S.lp = S.bstart + C + 1;
    S.rp = S.lp + C;
    S.expd = reinterpret_cast<unsigned char*>(S.rp + C);
    return S;
}

// The listed points (2 <= P <= C kept intervals each), scratch in LDS; RS > 0
// sorts the keys in registers (64 * RS >= C).
template <typename K, int RS>
__global__ __launch_bounds__(64) void dbscan_ref_lds_kernel(const int* __restrict__ list, int nlist,
                                                            const int* __restrict__ counts,
                                                            const long long* __restrict__ off,
                                                            const unsigned* __restrict__ times, double hp,
                                                            double eps, int C, int* __restrict__ out)
{
    extern __shared__ unsigned long long lds_keys[];
    const int b = blockIdx.x;
    if (b >= nlist) return;
    const int p = list[b];
    const int P = counts[p] - 1;
    const int c =
        dbscan_ref_wave<short, K, RS>(times + off[p], P, hp, eps, ref_scratch<short, K>(lds_keys, C), C);
    if (threadIdx.x == 0) out[p] = c;
}

// Points with more kept intervals than the largest LDS class: the same with
// 32-bit scratch in global memory, ref_global_bytes(C) at the given offsets.
constexpr long long ref_global_bytes(long long C) { return C * (8 + 2 * 4 + 4 + 4 + 4 + 4 + 1) + 4 + 15; }

__global__ __launch_bounds__(64) void dbscan_ref_big_kernel(const int* __restrict__ big, int nbig,
                                                            const int* __restrict__ counts,
                                                            const long long* __restrict__ off,
                                                            const unsigned* __restrict__ times, double hp,
                                                            double eps, unsigned char* __restrict__ scratch,
                                                            const long long* __restrict__ soff,
                                                            int* __restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= nbig) return;
    const long long p = big[b];
    const int P = counts[p] - 1;
    int C = 1;
    while (C < P) C <<= 1;
    const int c = dbscan_ref_wave<int, unsigned long long, 0>(times + off[p], P, hp, eps,
                                                           ref_scratch<int, unsigned long long>(scratch + soff[b], C), C);
    if (threadIdx.x == 0) out[p] = c;
}

// Size classes of the per-point 2-D analysis: class k < ncls holds the
// flag-1 points with min_p <= P <= cap[k] kept peaks (and above cap[k-1]);
// class ncls the longer ones, each with a scratch region of big_bytes(P).
struct ClassCaps {
    int cap[8];
    int ncls;
    int min_p;
    int big_mode;  // 0: next_pow2(P) doubles (CONFIG), 1: ref_global_bytes (REFERENCE)
};

__device__ __forceinline__ long long class_big_bytes(const ClassCaps& K, int P)
{
    long long C = 1;
    while (C < P) C <<= 1;
    const long long b = K.big_mode == 0 ? C * (long long)sizeof(double) : ref_global_bytes(C);
    return (b + 15) / 16 * 16;
}

// Lists every point in its class (order within a class is arbitrary: each
// point's result depends on its own peaks only).  Wave-aggregated appends;
// every lane of the grid takes part in the ballots.
__global__ __launch_bounds__(256) void class_lists_kernel(const int* __restrict__ flags,
                                                          const int* __restrict__ counts, long long n,
                                                          const ClassCaps K, int* __restrict__ lists,
                                                          long long* __restrict__ soff, int* __restrict__ lcount,
                                                          unsigned long long* __restrict__ big_bytes)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1, P = 0;
    if (p < n && flags[p] == 1) {
        P = counts[p] - 1;
        if (P >= K.min_p) {
            c = 0;
            while (c < K.ncls && K.cap[c] < P) ++c;
        }
    }
    const int lane = threadIdx.x & 63;
    for (int k = 0; k <= K.ncls; ++k) {
        const unsigned long long m = __ballot(c == k);
        if (m == 0ull) continue;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&lcount[k], __popcll(m));
        base = __shfl(base, leader);
        if (c == k) {
            const int pos = base + (int)__popcll(m & ((1ull << lane) - 1ull));
            lists[(long long)k * n + pos] = (int)p;
            if (k == K.ncls) soff[pos] = (long long)atomicAdd(big_bytes, (unsigned long long)class_big_bytes(K, P));
        }
    }
}

__global__ void zero_class_counts_kernel(int* __restrict__ lcount, int m, unsigned long long* __restrict__ big)
{
    if ((int)threadIdx.x < m) lcount[threadIdx.x] = 0;
    if (threadIdx.x == 0) *big = 0ull;
}

namespace {
struct ClassLists {
    DevBuf<int> lists;         // [(ncls + 1) * n]
    DevBuf<long long> soff;    // [n] scratch offsets of the long ones
    DevBuf<int> lcount;        // [ncls + 1]
    DevBuf<unsigned long long> big;
    int count[9] = {0};
    unsigned long long big_bytes = 0;
};
}  // namespace

// Builds the class lists on the device and reads back their sizes.
static void class_lists(CallCtx& ctx, const BifDevice& R, const ClassCaps& K, ClassLists& L)
{
    const long long n = R.n;
    L.lists.alloc((size_t)(K.ncls + 1) * n);
    L.soff.alloc((size_t)n);
    L.lcount.alloc((size_t)K.ncls + 1);
    L.big.alloc(1);
    hipLaunchKernelGGL(zero_class_counts_kernel, dim3(1), dim3(64), 0, ctx.stream(), L.lcount.get(), K.ncls + 1,
                       L.big.get());
    hipLaunchKernelGGL(class_lists_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), R.flags.get(),
                       R.counts.get(), n, K, L.lists.get(), L.soff.get(), L.lcount.get(), L.big.get());
    CHAOS_HIP(hipGetLastError());
    CHAOS_HIP(hipMemcpyAsync(L.count, L.lcount.get(), sizeof(int) * (K.ncls + 1), hipMemcpyDeviceToHost,
                             ctx.stream()));
    CHAOS_HIP(hipMemcpyAsync(&L.big_bytes, L.big.get(), sizeof L.big_bytes, hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    ctx.add_launch(3);
}

void bif2d_analysis(CallCtx& ctx, int semantics, BifDevice& R, double eps, int* d_out)
{
    const long long n = R.n;
    ctx.begin();
    ClassLists L;
    if (semantics == CHAOS_SEM_REFERENCE) {
        hipLaunchKernelGGL(dbscan_ref_trivial_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(),
                           R.flags.get(), R.counts.get(), d_out, n);
        CHAOS_HIP(hipGetLastError());
        // Sample gaps are below the recorded sample count and indices below
        // C, so 16 bits each suffice when the record holds fewer than 65536
        // samples.  32-bit keys sort in registers, so the LDS classes need not
        // be powers of two: C = 256 ... 2048 in six steps (17 bytes per slot,
        // 4.4 to 34.8 KiB), each point in the smallest that holds it.  64-bit
        // keys sort in LDS: C = 256, 1024, 2048 (21 bytes per slot).
        const bool keys32 = R.samples > 0 && R.samples < 65536;
        const ClassCaps K = keys32 ? ClassCaps{{256, 512, 768, 1024, 1536, 2048}, 6, 2, 1}
                                   : ClassCaps{{256, 1024, 2048}, 3, 2, 1};
        class_lists(ctx, R, K, L);
        for (int c = 0; c < K.ncls; ++c) {
            if (L.count[c] == 0) continue;
            const int C = K.cap[c];
            const int* lst = L.lists.get() + (long long)c * n;
            auto go = [&](auto kern, int key_bytes) {
                hipLaunchKernelGGL(kern, dim3((unsigned)L.count[c]), dim3(64), ref_lds_bytes(C, key_bytes),
                                   ctx.stream(), lst, L.count[c], R.counts.get(), R.off.get(), R.times.get(), R.hp,
                                   eps, C, d_out);
            };
            // 32-bit keys are sorted in registers (64 lanes x the power of two >= C)
            if (!keys32) go(dbscan_ref_lds_kernel<unsigned long long, 0>, 8);
            else if (C <= 256) go(dbscan_ref_lds_kernel<unsigned, 4>, 4);
            else if (C <= 512) go(dbscan_ref_lds_kernel<unsigned, 8>, 4);
            else if (C <= 1024) go(dbscan_ref_lds_kernel<unsigned, 16>, 4);
            else go(dbscan_ref_lds_kernel<unsigned, 32>, 4);
            CHAOS_HIP(hipGetLastError());
            ctx.add_launch();
        }
        if (L.count[K.ncls] > 0) {
            DevBuf<unsigned char> scratch((size_t)L.big_bytes);
            hipLaunchKernelGGL(dbscan_ref_big_kernel, dim3((unsigned)L.count[K.ncls]), dim3(64), 0, ctx.stream(),
                               L.lists.get() + (long long)K.ncls * n, L.count[K.ncls], R.counts.get(), R.off.get(),
                               R.times.get(), R.hp, eps, scratch.get(), L.soff.get(), d_out);
            CHAOS_HIP(hipGetLastError());
            ctx.sync();
        }
    } else {
        hipLaunchKernelGGL(flag_out_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), R.flags.get(), d_out,
                           n);
        // classes M = 64 << c: up to 2048 values the sort runs in registers
        // (64 * R per wave), 4096 in LDS (32 KiB), longer in global memory
        const ClassCaps K{{64, 128, 256, 512, 1024, 2048, 4096}, 7, 1, 0};
        class_lists(ctx, R, K, L);
        for (int c = 0; c < K.ncls; ++c) {
            const int cnt = L.count[c];
            if (cnt == 0) continue;
            const int* lst = L.lists.get() + (long long)c * n;
            auto reg = [&](auto rtag) {
                constexpr int RV = decltype(rtag)::value;
                hipLaunchKernelGGL((dbscan1d_reg_kernel<RV>), dim3((unsigned)cnt), dim3(64), 0, ctx.stream(), lst,
                                   cnt, R.counts.get(), R.off.get(), R.vals.get(), eps, d_out);
            };
            switch (c) {
            case 0: reg(std::integral_constant<int, 1>{}); break;
            case 1: reg(std::integral_constant<int, 2>{}); break;
            case 2: reg(std::integral_constant<int, 4>{}); break;
            case 3: reg(std::integral_constant<int, 8>{}); break;
            case 4: reg(std::integral_constant<int, 16>{}); break;
            case 5: reg(std::integral_constant<int, 32>{}); break;
            default:
                hipLaunchKernelGGL(dbscan1d_list_kernel, dim3((unsigned)cnt), dim3(64),
                                   (size_t)K.cap[c] * sizeof(double), ctx.stream(), lst, cnt, R.counts.get(),
                                   R.off.get(), R.vals.get(), eps, K.cap[c], d_out);
            }
            CHAOS_HIP(hipGetLastError());
            ctx.add_launch();
        }
        if (L.count[K.ncls] > 0) {
            DevBuf<unsigned char> scratch((size_t)L.big_bytes);
            hipLaunchKernelGGL(dbscan1d_big_kernel, dim3((unsigned)L.count[K.ncls]), dim3(64), 0, ctx.stream(),
                               R.counts.get(), R.off.get(), R.vals.get(), L.lists.get() + (long long)K.ncls * n,
                               L.count[K.ncls], scratch.get(), L.soff.get(), eps, d_out);
            CHAOS_HIP(hipGetLastError());
            ctx.sync();
        }
    }
    ctx.sync();  // the list buffers go back to the allocator at the end of this scope
    ctx.add_analysis_ms(ctx.end());
    ctx.add_launch();
}

void bif1d_collect(CallCtx& ctx, BifDevice& R, int* counts, double* first, double* peaks, double* intervals,
                   long long capacity, long long* total)
{
    std::vector<double> v((size_t)R.total);
    std::vector<unsigned> t((size_t)R.total);
    std::vector<long long> off((size_t)R.n);
    if (R.total > 0) {
        CHAOS_HIP(hipMemcpyAsync(v.data(), R.vals.get(), R.total * 8, hipMemcpyDeviceToHost, ctx.stream()));
        CHAOS_HIP(hipMemcpyAsync(t.data(), R.times.get(), R.total * 4, hipMemcpyDeviceToHost, ctx.stream()));
    }
    CHAOS_HIP(hipMemcpyAsync(off.data(), R.off.get(), R.n * 8, hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    // Reference output rules (bifurcationCUDA.cu:359-367,498-510,748-763):
    // flag -1 -> one line with the first sample; flag 0 -> nothing;
    // otherwise the peaks after the first, intervals (t[k+1]-t[k])*h*pre.
    long long k_out = 0;
    for (long long p = 0; p < R.n; ++p) {
        first[p] = R.h_first[p];
        const int fl = R.h_flags[p];
        if (fl == -1) {
            counts[p] = -1;
            continue;
        }
        if (fl != 1 || R.h_counts[p] < 2) {
            counts[p] = 0;
            continue;
        }
        const int P = R.h_counts[p] - 1;
        counts[p] = P;
        for (int k = 0; k < P; ++k, ++k_out) {
            if (k_out < capacity) {
                const long long e = off[p] + k;
                peaks[k_out] = v[e + 1];
                intervals[k_out] = (double)((long long)t[e + 1] - (long long)t[e]) * R.hp;
            }
        }
    }
    *total = k_out;
}

static BifInputs bif_inputs(double tMax, int nPts, double h, int nx, const double* ic, const double* ranges,
                            const int* idx, int writableVar, double maxValue, double transientTime,
                            const double* values, int na, int preScaller, double eps, int dims,
                            int legacy_ic = 0)
{
    BifInputs in{};
    in.legacy_ic = legacy_ic;
    in.tMax = tMax;
    in.nPts = nPts;
    in.h = h;
    in.nx = nx;
    in.ic = ic;
    in.ranges = ranges;
    in.idx = idx;
    in.writableVar = writableVar;
    in.maxValue = maxValue;
    in.transientTime = transientTime;
    in.values = values;
    in.na = na;
    in.preScaller = preScaller;
    in.eps = eps;
    in.dims = dims;
    return in;
}

}  // namespace chaos

using namespace chaos;

namespace chaos {

void write_int_map_file(TextFile* f, const char* path, const double* ranges, int nPts, const int* v, int output)
{
    MapSpec m;
    m.ranges = ranges;
    m.header_lines = 2;
    m.nPts = nPts;
    m.count = (long long)nPts * nPts;
    m.iv = v;
    if (f) {
        write_map_text(*f, m);
        f->close();
    }
    if (out_bin(output)) write_map_bin(std::string(path) + ".bin", m);
}

void write_bif1d_body(TextFile& f, const double* ranges, int nPts, const int* counts, const double* first,
                      const double* peaks, const double* intervals)
{
    long long k = 0;
    for (int p = 0; p < nPts; ++p) {
        const double x = grid_value(p, nPts, ranges[0], ranges[1]);
        if (counts[p] == -1) {
            f.put_g(x, 12);
            f.put(", ", 2);
            f.put_g(first[p], 12);
            f.put(", 0\n");
            continue;
        }
        for (int j = 0; j < counts[p]; ++j, ++k) {
            f.put_g(x, 12);
            f.put(", ", 2);
            f.put_g(peaks[k], 12);
            f.put(", ", 2);
            f.put_g(intervals[k], 12);
            f.put("\n", 1);
        }
    }
}

static void bif1d_impl(const chaos_options* opt, double tMax, int nPts, double h, int nIC, const double* ic,
                       const double* ranges, const int* idx, int writableVar, double maxValue, double transientTime,
                       const double* values, int nV, int preScaller, int* counts, double* first, double* peaks,
                       double* intervals, long long capacity, long long* total, int legacy_ic)
{
    chaos_options o = resolve_options(opt);
    CHAOS_REQUIRE(counts && first && total, "NULL output");
    CHAOS_REQUIRE(capacity == 0 || (peaks && intervals), "NULL peaks/intervals with capacity > 0");
    BifInputs in = bif_inputs(tMax, nPts, h, nIC, ic, ranges, idx, writableVar, maxValue, transientTime, values, nV,
                              preScaller, 0.0, 1, legacy_ic);
    CallCtx ctx(o);
    BifDevice R;
    // the 1-D results are assembled from one pass over the whole grid
    run_bif(ctx, o, in, R, false, [](BifDevice&, long long) {});
    bif1d_collect(ctx, R, counts, first, peaks, intervals, capacity, total);
    const long long k_out = *total;
    ctx.finish();
    if (k_out > capacity) fail(CHAOS_E_CAPACITY, "need capacity %lld, got %lld", k_out, capacity);
}

static void bif2d_impl(const chaos_options* opt, double tMax, int nPts, double h, int nIC, const double* ic,
                       const double* ranges, const int* idx, int writableVar, double maxValue, double transientTime,
                       const double* values, int nV, int preScaller, double eps, int* out, int legacy_ic)
{
    chaos_options o = resolve_options(opt);
    CHAOS_REQUIRE(out, "out is NULL");
    BifInputs in = bif_inputs(tMax, nPts, h, nIC, ic, ranges, idx, writableVar, maxValue, transientTime, values, nV,
                              preScaller, eps, 2, legacy_ic);
    CallCtx ctx(o);
    BifDevice R;
    R.need_times = o.semantics == CHAOS_SEM_REFERENCE;  // the CONFIG count reads peak values only
    R.need_vals = o.semantics != CHAOS_SEM_REFERENCE;   // the reference dbscan reads peak times only
    R.need_host = false;                                // the 2-D analysis lists its points on the device
    const long long n = (long long)nPts * nPts;
    DevBuf<int> tmp;
    int* d_out = out;
    // each pass (the whole grid, or a range of it when its peaks do not fit
    // the device at once) writes its part of the map
    run_bif(ctx, o, in, R, true, [&](BifDevice& Rk, long long p0) {
        if (!o.device_buffers && !tmp.get()) {
            tmp.alloc((size_t)n);
            d_out = tmp.get();
        }
        bif2d_analysis(ctx, o.semantics, Rk, eps, d_out + p0);
    });
    if (!o.device_buffers)
        CHAOS_HIP(hipMemcpyAsync(out, d_out, n * sizeof(int), hipMemcpyDeviceToHost, ctx.stream()));
    ctx.sync();
    ctx.finish();
}

// Host results of a 1-D sweep in one pass when the peak buffers are large
// enough (one peak per 32 samples), else a second, exactly sized pass.
static long long bif1d_host(const chaos_options* opt, double tMax, int nPts, double h, int nIC, const double* ic,
                            const double* ranges, const int* idx, int writableVar, double maxValue,
                            double transientTime, const double* values, int nV, int preScaller,
                            std::vector<int>& counts, std::vector<double>& first, std::vector<double>& pk,
                            std::vector<double>& iv, int legacy_ic)
{
    const chaos_options o = resolve_options(opt);
    counts.assign(nPts, 0);
    first.assign(nPts, 0.0);
    const long long B = std::max(1, record_samples(tMax, h, std::max(1, preScaller), o.semantics));
    long long cap = (long long)nPts * std::min(B / 2 + 1, std::max(256LL, B / 32));
    for (int attempt = 0; attempt < 2; ++attempt) {
        pk.assign((size_t)std::max(1LL, cap), 0.0);
        iv.assign((size_t)std::max(1LL, cap), 0.0);
        long long total = 0;
        try {
            bif1d_impl(opt, tMax, nPts, h, nIC, ic, ranges, idx, writableVar, maxValue, transientTime, values, nV,
                       preScaller, counts.data(), first.data(), pk.data(), iv.data(), cap, &total, legacy_ic);
            pk.resize(total);
            iv.resize(total);
            return total;
        } catch (const Error& e) {
            if (e.code != CHAOS_E_CAPACITY || attempt == 1) throw;
            cap = total;
        }
    }
    return 0;
}

}  // namespace chaos

extern "C" int chaos_bif1d(const chaos_options* opt, double tMax, int nPts, double h, int amountOfInitialConditions,
                           const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                           int writableVar, double maxValue, double transientTime, const double* values,
                           int amountOfValues, int preScaller, int* counts, double* first, double* peaks,
                           double* intervals, long long capacity, long long* total)
{
    return guarded("chaos_bif1d", [&] {
        bif1d_impl(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, counts, first, peaks,
                   intervals, capacity, total, 0);
    });
}

extern "C" int chaos_legacy_bif1d_ic(const chaos_options* opt, double tMax, int nPts, double h,
                                     int amountOfInitialConditions, const double* initialConditions,
                                     const double* ranges, const int* indicesOfMutVars, int writableVar,
                                     double maxValue, double transientTime, const double* values,
                                     int amountOfValues, int preScaller, int* counts, double* peaks,
                                     long long capacity, long long* total)
{
    return guarded("chaos_legacy_bif1d_ic", [&] {
        CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2");
        std::vector<double> first(nPts), iv((size_t)std::max(1LL, capacity));
        bif1d_impl(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, counts, first.data(),
                   peaks, capacity > 0 ? iv.data() : nullptr, capacity, total, 1);
    });
}

extern "C" int chaos_bifurcation1DIC(const chaos_options* opt, double tMax, int nPts, double h,
                                     int amountOfInitialConditions, const double* initialConditions,
                                     const double* ranges, const int* indicesOfMutVars, int writableVar,
                                     double maxValue, double transientTime, const double* values,
                                     int amountOfValues, int preScaller, const char* OUT_FILE_PATH)
{
    return guarded("chaos_bifurcation1DIC", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2");
        TextFile f(OUT_FILE_PATH);
        std::vector<int> counts;
        std::vector<double> first, pk, iv;
        bif1d_host(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, counts, first, pk, iv,
                   1);
        // debri/hostLibrary.cu:1447-1456: "<ic>, <peak>\n" per kept peak at
        // precision 12, nothing for fixed points or diverged points, no header
        long long k = 0;
        for (int p = 0; p < nPts; ++p) {
            const double x = grid_value(p, nPts, ranges[0], ranges[1]);
            for (int j = 0; j < counts[p]; ++j, ++k) {
                f.put_g(x, 12);
                f.put(", ", 2);
                f.put_g(pk[k], 12);
                f.put("\n", 1);
            }
        }
        f.close();
    });
}

extern "C" int chaos_bifurcation1D(const chaos_options* opt, double tMax, int nPts, double h,
                                   int amountOfInitialConditions, const double* initialConditions,
                                   const double* ranges, const int* indicesOfMutVars, int writableVar,
                                   double maxValue, double transientTime, const double* values,
                                   int amountOfValues, int preScaller, const char* OUT_FILE_PATH)
{
    return guarded("chaos_bifurcation1D", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(nPts >= 1, "nPts must be >= 1");
        TextFile f(OUT_FILE_PATH);
        std::vector<int> counts;
        std::vector<double> first, pk, iv;
        bif1d_host(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, counts, first, pk, iv,
                   0);
        // header at default precision, body at 12 digits (:645-647,746-763)
        write_range_header(f, ranges, 1, 6);
        write_bif1d_body(f, ranges, nPts, counts.data(), first.data(), pk.data(), iv.data());
        f.close();
    });
}

extern "C" int chaos_bif2d(const chaos_options* opt, double tMax, int nPts, double h, int amountOfInitialConditions,
                           const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                           int writableVar, double maxValue, double transientTime, const double* values,
                           int amountOfValues, int preScaller, double eps, int* out)
{
    return guarded("chaos_bif2d", [&] {
        bif2d_impl(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps, out, 0);
    });
}

extern "C" int chaos_legacy_bif2d_ic(const chaos_options* opt, double tMax, int nPts, double h,
                                     int amountOfInitialConditions, const double* initialConditions,
                                     const double* ranges, const int* indicesOfMutVars, int writableVar,
                                     double maxValue, double transientTime, const double* values,
                                     int amountOfValues, int preScaller, double eps, int* out)
{
    return guarded("chaos_legacy_bif2d_ic", [&] {
        CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2");
        bif2d_impl(opt, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps, out, 1);
    });
}

extern "C" int chaos_bifurcation2DIC(const chaos_options* opt, double tMax, int nPts, double h,
                                     int amountOfInitialConditions, const double* initialConditions,
                                     const double* ranges, const int* indicesOfMutVars, int writableVar,
                                     double maxValue, double transientTime, const double* values,
                                     int amountOfValues, int preScaller, double eps, const char* OUT_FILE_PATH)
{
    return guarded("chaos_bifurcation2DIC", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(nPts >= 2, "nPts must be >= 2");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        std::unique_ptr<TextFile> f;
        if (out_csv(o.output)) f.reset(new TextFile(OUT_FILE_PATH));
        std::vector<int> res((size_t)nPts * nPts);
        bif2d_impl(&o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, transientTime, values, amountOfValues, preScaller, eps, res.data(), 1);
        // debri/hostLibrary.cu:2934-2935, 3060-3073: two range lines at
        // default precision, then the map
        write_int_map_file(f.get(), OUT_FILE_PATH, ranges, nPts, res.data(), o.output);
    });
}

extern "C" int chaos_bifurcation2D(const chaos_options* opt, double tMax, int nPts, double h,
                                   int amountOfInitialConditions, const double* initialConditions,
                                   const double* ranges, const int* indicesOfMutVars, int writableVar,
                                   double maxValue, double transientTime, const double* values,
                                   int amountOfValues, int preScaller, double eps, const char* OUT_FILE_PATH)
{
    return guarded("chaos_bifurcation2D", [&] {
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(nPts >= 1, "nPts must be >= 1");
        chaos_options o = resolve_options(opt);
        o.device_buffers = 0;
        std::unique_ptr<TextFile> f;
        if (out_csv(o.output)) f.reset(new TextFile(OUT_FILE_PATH));
        std::vector<int> res((size_t)nPts * nPts);
        int rc = chaos_bif2d(&o, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                             indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                             preScaller, eps, res.data());
        if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
        write_int_map_file(f.get(), OUT_FILE_PATH, ranges, nPts, res.data(), o.output);  // header :109-110
    });
}
