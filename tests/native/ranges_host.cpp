// Host build of the peak-count sweep's range planning (chaoslibrary_amd/
// csrc/peak_ranges.hpp), for the CPU tests in tests/test_peak_ranges_cpu.py.
#include <cstring>

#include "peak_ranges.hpp"

// Writes the range starts and n into out (capacity cap); returns how many,
// or -1 when one wave needs more than the budget (*wave_bytes: its need).
extern "C" long long exact_ranges_host(const int* counts, long long n, unsigned long long budget, int chunk,
                                       unsigned long long chunk_bytes, long long* out, long long cap,
                                       unsigned long long* wave_bytes)
{
    size_t wb = 0;
    const std::vector<long long> s = chaos::exact_ranges(counts, n, budget, chunk, chunk_bytes, &wb);
    if (s.empty()) {
        *wave_bytes = wb;
        return -1;
    }
    for (size_t k = 0; k < s.size() && (long long)k < cap; ++k) out[k] = s[k];
    return (long long)s.size();
}

extern "C" long long first_pass_cuts_host(long long n, double wave_guess, long long budget_chunks, long long* out,
                                          long long cap)
{
    const std::vector<long long> s = chaos::first_pass_cuts(n, wave_guess, budget_chunks);
    for (size_t k = 0; k < s.size() && (long long)k < cap; ++k) out[k] = s[k];
    return (long long)s.size();
}

extern "C" unsigned long long wave_pass_bytes_host(long long s, long long points, int chunk,
                                                   unsigned long long chunk_bytes)
{
    return chaos::wave_pass_bytes(s, points, chunk, chunk_bytes);
}
