// dbscan.hpp -- grid DBSCAN over 2-D / 3-D feature points on the device, and
// the deterministic z-score that feeds it.
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace chaos {

struct DbscanParams {
    int dim;          // 2 or 3
    double eps;       // neighbourhood radius, test sqrt(sum d^2) <= eps
    int minPts;       // core threshold (self counted)
    bool sign_rule;   // Basins REFERENCE: components holding a flag -1 point are negative
};

// Scratch of one clustering call: everything the kernels use besides the
// caller's points, include flags and labels, carved from one device buffer
// in 256-byte aligned pieces: a control block; per cell (H slots, a power of
// two) key, count, smallest member, segment start and cursor, smallest core,
// member box and the core-cell list; per point slot, cell list, parent,
// label root, candidate list (and smallest fixed point under the sign rule),
// pieces of split core tests;
// per 64-point word root bits and their prefix sums; scan temporaries.
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr size_t scan_tmp_ints(long long n)
{
    return n <= 1024 ? 1 : 2 * (size_t)((n + 1023) / 1024) + scan_tmp_ints((n + 1023) / 1024);
}

// Capacity of the list of split core-test pieces (16 bytes each).
constexpr long long core_piece_cap(long long n) { return n / 16 + 64; }

constexpr size_t dbscan_ws_bytes_c(long long n, long long H, bool sign)
{
    const long long nw = (n + 63) / 64;
    size_t t = 256;  // control block
    t += align256(8 * (size_t)H) + 5 * align256(4 * (size_t)H) + align256(48 * (size_t)H) + align256(4 * (size_t)H);
    t += 5 * align256(4 * (size_t)n) + (sign ? align256(4 * (size_t)n) : 0);
    t += align256(16 * (size_t)core_piece_cap(n));
    t += align256(8 * (size_t)nw) + align256(4 * (size_t)nw);
    if (sign) t += align256(8 * (size_t)nw) + align256(4 * (size_t)nw);
    t += align256(4 * scan_tmp_ints(H > nw ? H : nw));
    return t;
}

size_t dbscan_ws_bytes(long long n, long long H, bool sign_rule);
// Cell-table sizes for n points: the first attempt is sized for the few
// distinct cells clustered features occupy (a table more than half full is
// redone at the second size, which holds any n points).
long long dbscan_first_cells(long long n);
long long dbscan_max_cells(long long n);

// pts: dim device arrays of n doubles; include: device flags or NULL (all
// points used; with sign_rule, include[i] == -1 marks a fixed point).
// Writes labels[n] (device).  Deterministic: clusters are numbered 1..K by
// smallest core index (sign_rule: negatives -1..-K by smallest fixed-point
// index); a border point joins the cluster of its smallest-index core
// neighbour; noise and excluded points get 0.  Returns K (positives).
// ws / ws_bytes: caller scratch (a call that needs more than ws_bytes fails
// with CHAOS_E_CAPACITY); NULL takes the scratch from the device allocator.
// box: device [6] order-preserving encodings of the included points'
// bounding box (min x, y, z, max x, y, z), or NULL to compute it.
// zmom: device [2 * dim] (mean, sd) per coordinate: the points are raw
// features and are clustered by their z-scores, formed on each read with the
// z-score's own expression (box then bounds the z-scores); NULL clusters the
// values as given.
// Synchronises the stream once (twice when the first cell table fills).
int dbscan_device(CallCtx& ctx, const double* const* pts, const int* include, long long n,
                  const DbscanParams& prm, int* labels, void* ws = nullptr, size_t ws_bytes = 0,
                  const unsigned long long* box = nullptr, const double* zmom = nullptr);

// Moments of the z-score of the k feature arrays over the points with
// mask != 0, with no host round trip: sums by the pairwise tree of the
// points padded to max(1024, the power of two >= n) -- the host
// restatement's order -- formed in blocks and finished on the device.
// Returns the device array [mean_0, sd_0, ..., mean_{k-1}, sd_{k-1}, N]
// inside scratch; when box is not NULL, also writes the encoded bounding box
// of the z-scored masked points (device [6], the form dbscan_device takes).
// scratch: device memory of zscore_scratch_bytes(n, k) that stays valid
// until the stream has run the kernels that read the moments.
// sweep_partials: the first-moment partials of 256-point blocks are in
// scratch already (a basins sweep wrote them through zscore_sweep_partials),
// so the first pass over the features is skipped.
size_t zscore_scratch_bytes(long long n, int k);
const double* zscore_moments_device(CallCtx& ctx, const double* const* src, int k, const int* mask, long long n,
                                    unsigned long long* box, void* scratch, bool sweep_partials = false);

// Where a 256-thread kernel over the n points writes its blocks' first
// moments inside z-score scratch (block256_moments, moments.hpp): sums
// [k][nb], counts [nb], encoded minimum / maximum [2k][nb].
struct ZPartials {
    double* sum;
    int* cnt;
    unsigned long long* mm;
    long long nb;
};
ZPartials zscore_sweep_partials(void* scratch, long long n, int k);
}  // namespace chaos
