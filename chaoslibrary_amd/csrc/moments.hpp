// moments.hpp -- pieces of the deterministic z-score moments shared by the
// basins sweep kernel (which forms its blocks' first moments in its epilogue)
// and the clustering's moment kernels.
//
// The sum of n values is the complete pairwise tree over adjacent pairs of
// the values padded with +0 to P = max(1024, the power of two >= n); every
// aligned run of 2^k values is a subtree, so a block may sum its own run and
// the runs are combined by the same tree later.  IEEE addition is
// commutative, so a wavefront butterfly (xor 1, 2, ..., 32) forms exactly the
// tree's sums of aligned runs of 2, 4, ..., 64 lanes.
#pragma once
#include <cstring>

#include "runtime.hpp"

namespace chaos {

// Order-preserving map double -> u64 (for atomicMin/atomicMax on doubles).
__host__ __device__ __forceinline__ unsigned long long ord_enc(double v)
{
    unsigned long long u;
    memcpy(&u, &v, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__host__ __device__ __forceinline__ double ord_dec(unsigned long long e)
{
    const unsigned long long u = (e >> 63) ? (e & 0x7fffffffffffffffull) : ~e;
    double v;
    memcpy(&v, &u, 8);
    return v;
}

// Sum of the 64 lanes' values by the pairwise tree (every lane gets it).
__device__ __forceinline__ double wave_tree_sum(double v)
{
    for (int o = 1; o < 64; o <<= 1) v = v + __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned long long wave_min_enc(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_enc(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// First moments of a 256-thread block's points (thread t holds point
// 256 b + t): the tree sums of the k values f[j] over lanes with m set (the
// others add +0), the count of such lanes and the encoded minimum / maximum
// of each value, written to sum[j * nb + b], cnt[b], mm[(2 j) * nb + b] and
// mm[(2 j + 1) * nb + b].  Every thread of the block must call it.
__device__ __forceinline__ void block256_moments(const double* f, int k, bool m, long long nb, double* sum, int* cnt,
                                                 unsigned long long* mm)
{
    __shared__ double s_sum[3][4];
    __shared__ unsigned long long s_mm[6][4];
    __shared__ int s_cnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j < k; ++j) {
        const double s = wave_tree_sum(m ? f[j] : 0.0);
        const unsigned long long e = ord_enc(f[j]);
        const unsigned long long lo = wave_min_enc(m ? e : ~0ull), hi = wave_max_enc(m ? e : 0ull);
        if (lane == 0) {
            s_sum[j][wave] = s;
            s_mm[2 * j][wave] = lo;
            s_mm[2 * j + 1][wave] = hi;
        }
    }
    const int c = __popcll(__ballot(m));
    if (lane == 0) s_cnt[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long b = blockIdx.x;
        for (int j = 0; j < k; ++j) {
            sum[j * nb + b] = (s_sum[j][0] + s_sum[j][1]) + (s_sum[j][2] + s_sum[j][3]);
            unsigned long long lo = s_mm[2 * j][0], hi = s_mm[2 * j + 1][0];
            for (int w = 1; w < 4; ++w) {
                lo = s_mm[2 * j][w] < lo ? s_mm[2 * j][w] : lo;
                hi = s_mm[2 * j + 1][w] > hi ? s_mm[2 * j + 1][w] : hi;
            }
            mm[(2 * j) * nb + b] = lo;
            mm[(2 * j + 1) * nb + b] = hi;
        }
        cnt[b] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    }
}

}  // namespace chaos
