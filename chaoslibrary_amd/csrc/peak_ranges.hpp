// peak_ranges.hpp -- how a peak-count sweep is cut into ranges of whole
// 64-point waves by the device memory, as plain host C++ (no HIP types), so
// that the rule runs and is tested on machines without a GPU.
//
// The peak store hands each wave chunks of `chunk` entries; a wave's exact
// need for s peaks is ceil(s / (chunk - 64)) + 1 chunks (one chunk switch
// can strand up to 64 entries, and the last chunk is open).
#pragma once

#include <cstddef>
#include <vector>

namespace chaos {

inline long long wave_chunks(long long s, int chunk) { return (s + (chunk - 64) - 1) / (chunk - 64) + 1; }

// Device bytes of an exact pass over one wave of `points` points holding s
// peaks: its chunks, 32 bytes per stored peak (the scattered values and
// times, the 2-D analysis's scratch) and 128 per point (flags, counts,
// offsets, class lists).
inline size_t wave_pass_bytes(long long s, long long points, int chunk, size_t chunk_bytes)
{
    return (size_t)wave_chunks(s, chunk) * chunk_bytes + 32 * (size_t)s + 128 * (size_t)points;
}

// Exact passes over the points [0, n) with per-point peak counts: ranges of
// whole waves, each as long as its bytes fit `budget`.  Returns the range
// starts followed by n ([s[k], s[k + 1]) is range k), or an empty vector
// when one wave alone needs more than the budget (*wave_bytes: its need).
inline std::vector<long long> exact_ranges(const int* counts, long long n, size_t budget, int chunk,
                                           size_t chunk_bytes, size_t* wave_bytes)
{
    std::vector<long long> starts{0};
    size_t bytes = 0;
    for (long long w = 0; w < n; w += 64) {
        const long long e = w + 64 < n ? w + 64 : n;
        long long s = 0;
        for (long long p = w; p < e; ++p) s += counts[p];
        const size_t wb = wave_pass_bytes(s, e - w, chunk, chunk_bytes);
        if (wb > budget) {
            if (wave_bytes) *wave_bytes = wb;
            return {};
        }
        if (bytes + wb > budget) {
            starts.push_back(w);
            bytes = 0;
        }
        bytes += wb;
    }
    starts.push_back(n);
    return starts;
}

// First passes over [0, n): the whole grid when its first-guess store
// (wave_guess chunks per wave) fits budget_chunks, else ranges of equally
// many whole waves that do (at least one wave each).  Same form as
// exact_ranges.
inline std::vector<long long> first_pass_cuts(long long n, double wave_guess, long long budget_chunks)
{
    const long long waves = (n + 63) / 64;
    long long per = (long long)((double)budget_chunks / wave_guess);
    if (per < 1) per = 1;
    std::vector<long long> cuts{0};
    for (long long w = per; w < waves; w += per) cuts.push_back(64 * w);
    cuts.push_back(n);
    return cuts;
}

}  // namespace chaos
