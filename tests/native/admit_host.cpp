// Host build of the job server's admission rule (chaoslibrary_amd/csrc/
// admission.hpp), for the CPU tests in tests/test_admission_cpu.py.
#include "admission.hpp"

extern "C" int admit_host(const long long* footprint, const long long* tiles, int n, int free_slots,
                          long long* hbm_used, long long hbm_budget, long long* active_tiles, long long max_tiles)
{
    return chaos::admit_fifo(footprint, tiles, n, free_slots, hbm_used, hbm_budget, active_tiles, max_tiles);
}

extern "C" int family_waves_host(long long tiles, int take, long long grid_waves)
{
    return chaos::family_waves(tiles, take, grid_waves);
}
