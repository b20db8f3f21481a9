// dbscan.hpp -- grid DBSCAN over 2-D / 3-D feature points on the device.
#pragma once

#include "runtime.hpp"

namespace chaos {

struct DbscanParams {
    int dim;          // 2 or 3
    double eps;       // neighbourhood radius, test sqrt(sum d^2) <= eps
    int minPts;       // core threshold (self counted)
    bool sign_rule;   // Basins REFERENCE: components holding a flag -1 point are negative
};

// pts: dim device arrays of n doubles; include: device flags (point used iff
// include[i] != 0; with sign_rule, include[i] == -1 marks a fixed point).
// Writes labels[n] (device).  Deterministic: clusters are numbered 1..K by
// smallest core index (sign_rule: negatives -1..-K by smallest fixed-point
// index); a border point joins the cluster of its smallest-index core
// neighbour; noise and excluded points get 0.  Returns K (positives).
int dbscan_device(CallCtx& ctx, const double* const* pts, const int* include, long long n,
                  const DbscanParams& prm, int* labels);

// Deterministic fixed-order sum (blocks of 1024 reduced by the pairwise tree
// stride 512..1, block sums added in order) of mask ? f(x) : 0 where f is x
// (sq == false) or (x - mean)^2 (sq == true).
double det_sum_device(CallCtx& ctx, const double* x, const int* mask, long long n, bool sq, double mean);

// In-place z-score of the k feature arrays over points with mask != 0.
void zscore_device(CallCtx& ctx, const double* const* src, double* const* dst, int k, const int* mask, long long n,
                   double* moments);

}  // namespace chaos
