// admission.hpp -- the job server's admission rule, as plain host C++ (no HIP
// types), so that the rule runs and is tested on machines without a GPU.
#pragma once

namespace chaos {

// Pending jobs enter in submission order while
//   * a job-table slot is free,
//   * the job's HBM footprint fits the budget beside the footprints still
//     charged (jobs integrating, and finished jobs whose post stages and
//     result copies have not yet released their memory), and
//   * its tiles fit the device queue beside the active jobs' tiles.
// The first pending job that does not fit stops admission, so a large job
// is not overtaken by later small ones (FIFO, no starvation).
// footprint[i], tiles[i]: the i-th pending job in submission order, i < n.
// Returns how many of them enter now and charges them to *hbm_used and
// *active_tiles.
inline int admit_fifo(const long long* footprint, const long long* tiles, int n, int free_slots, long long* hbm_used,
                      long long hbm_budget, long long* active_tiles, long long max_tiles)
{
    int k = 0;
    while (k < n && k < free_slots) {
        if (*hbm_used + footprint[k] > hbm_budget) break;
        if (*active_tiles + tiles[k] > max_tiles) break;
        *hbm_used += footprint[k];
        *active_tiles += tiles[k];
        ++k;
    }
    return k;
}

}  // namespace chaos
