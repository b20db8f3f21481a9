// admission.hpp -- the job server's admission rule and round sizing, as plain
// host C++ (no HIP types), so that they run and are tested on machines
// without a GPU.
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

// Wavefronts of a family kernel that take part in a round: one per claim of
// `take` tiles, at most the grid's `grid_waves` (0 for no tiles).  The
// lowest wave indices are the first blocks the dispatcher spreads over the
// CUs, so a family with fewer claims than the grid has wavefronts runs them
// spread as a one-shot grid of that size is.  When every wavefront of the
// grid raced for fewer claims, the ones that started first took them and
// some SIMDs held more of the busy ones than others: a lone 300 x 300 LLE
// job ran 1.5x longer than its one-shot kernel.  (Fewer wavefronts than the
// grid when there are more claims -- the same number of claims for each --
// was slower: the grid's blocks then no longer divide evenly over the CUs.)
inline int family_waves(long long tiles, int take, long long grid_waves)
{
    if (tiles <= 0) return 0;
    const long long claims = (tiles + take - 1) / take;
    return (int)(claims < grid_waves ? claims : grid_waves);
}

}  // namespace chaos
