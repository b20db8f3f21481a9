// scan.hip -- exclusive prefix sum of int64 counts (two-level, recursive),
// used to turn per-point peak counts into segment offsets.
#include "scan.hpp"

namespace chaos {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void scan_tile_kernel(const long long* __restrict__ in,
                                                                 long long* __restrict__ out,
                                                                 long long* __restrict__ tile_sums, long long n)
{
    __shared__ long long s[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
    long long v[kScanPer];
    long long sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums
    for (int off = 1; off < kScanThreads; off <<= 1) {
        long long t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(long long* __restrict__ out,
                                                                const long long* __restrict__ tile_off, long long n)
{
    const long long i = (long long)blockIdx.x * kScanThreads + threadIdx.x;
    if (i < n) out[i] += tile_off[i / kScanTile];
}

static void scan_rec(hipStream_t st, const long long* in, long long* out, long long n)
{
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    DevBuf<long long> sums((size_t)tiles), offs((size_t)tiles);
    hipLaunchKernelGGL(scan_tile_kernel, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, out, sums.get(), n);
    CHAOS_HIP(hipGetLastError());
    if (tiles > 1) {
        scan_rec(st, sums.get(), offs.get(), tiles);
        hipLaunchKernelGGL(scan_add_kernel, dim3(ceil_div(n, kScanThreads)), dim3(kScanThreads), 0, st, out,
                           offs.get(), n);
        CHAOS_HIP(hipGetLastError());
    }
    CHAOS_HIP(hipStreamSynchronize(st));  // sums/offs are freed on return
}

long long exclusive_scan(hipStream_t st, const long long* d_in, long long* d_out, long long n)
{
    if (n <= 0) return 0;
    scan_rec(st, d_in, d_out, n);
    long long last_in = 0, last_out = 0;
    CHAOS_HIP(hipMemcpyAsync(&last_in, d_in + n - 1, sizeof(long long), hipMemcpyDeviceToHost, st));
    CHAOS_HIP(hipMemcpyAsync(&last_out, d_out + n - 1, sizeof(long long), hipMemcpyDeviceToHost, st));
    CHAOS_HIP(hipStreamSynchronize(st));
    return last_in + last_out;
}

}  // namespace chaos
