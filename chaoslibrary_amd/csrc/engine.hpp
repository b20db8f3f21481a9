// engine.hpp -- batched engine: device job table, tile queue, host scheduler.
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bifurcation.hpp"
#include "lanes.hpp"
#include "lle.hpp"
#include "runtime.hpp"
#include "sweep.hpp"

namespace chaos {

enum EngineProg : int { PROG_LLE = 0, PROG_BIF = 1, PROG_BASINS_REF = 2, PROG_BASINS_CFG = 3 };

// Device-resident description of one admitted job (one slot of the table).
struct JobDev {
    int prog;       // EngineProg
    int lorenz;     // 0: Rossler RK4, 1: Lorenz RK4
    int wv;         // writableVar
    int pad;
    long long np;        // lanes (multiple of 64)
    long long n_points;  // grid points
    long long total;     // ticks of the whole job
    LaneStore st;        // parked lane state
    double* lle_out;     // LLE output
    WaveCursor* cursors; // Bifurcation: per-group peak-store cursor
    LleArgs lle;
    SweepArgs sw;
};

// One unit of queued work: ticks [g0, g1) of a 64-point group of a job.
struct Tile {
    int slot;
    int group;
    long long g0, g1;
};

struct JobRec {
    long long id = 0;
    chaos_job spec{};
    std::atomic<int> status{0};  // 0 pending, 1 active, 3 finishing, 2 done, -1 failed
    std::string error;           // failure message (status -1)
    int slot = -1;
    int prog = 0, lorenz = 0, reran = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double rk4_steps = 0;
    long long footprint = 0;
    long long n = 0, np = 0, groups = 0, total = 0, done = 0, ticks_per_launch = 0, pool_chunks = 0;
    LleArgs lle{};
    SweepArgs sw{};
    // device working set (released when the job finishes)
    DevBuf<double> dstate, out, feats;
    DevBuf<int> istate, flags;
    DevBuf<WaveCursor> cursors;
    BifDevice bif;
    PeakPoolBuf pool;
    // results
    std::vector<double> r_lle, r_first, r_peaks, r_intervals, r_feats;
    std::vector<int> r_map, r_counts, r_labels, r_flags;
};

class Engine {
public:
    Engine(int device, const chaos_engine_params& prm);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    long long submit(const chaos_job& j);
    int step();
    void drain();
    JobRec& job(long long id);
    void release(long long id);
    chaos_server_stats stats() const;
    double now_ms() const;
    bool idle() const { return pending_.empty() && active_.empty() && finishing_.load() == 0; }

private:
    void plan(JobRec& r);
    void admit();
    void allocate(JobRec& r);
    void finalize(JobRec& r, CallCtx& ctx);
    void capture();
    void enqueue_launch();
    void finisher();                 // worker thread body
    int take_completed();            // jobs finished since the last call
    void wait_finisher_progress();   // block until a finishing job completes


    chaos_engine_params prm_;
    int device_ = 0;
    std::unique_ptr<CallCtx> ctx_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    DevBuf<JobDev> d_jobs_;
    DevBuf<Tile> d_tiles_;
    DevBuf<int> d_ntiles_;
    DevBuf<unsigned> d_head_;
    JobDev* h_jobs_ = nullptr;   // pinned
    Tile* h_tiles_ = nullptr;    // pinned
    int* h_ntiles_ = nullptr;    // pinned
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    int grid_ = 0;
    std::deque<long long> pending_;
    std::vector<long long> active_;
    std::map<long long, std::unique_ptr<JobRec>> jobs_;
    std::vector<int> free_slots_;
    long long next_id_ = 1, hbm_budget_ = 0, hbm_used_ = 0, active_groups_ = 0;
    long long launches_ = 0, graph_replays_ = 0, tiles_ = 0;
    double busy_ms_ = 0, finalize_ms_ = 0, steps_done_ = 0, t0_ = 0;
    // Finishing stages (peak post-processing, clustering, result copies) run
    // on a worker thread with its own stream, overlapping the next rounds.
    // mu_ guards finq_, hbm_used_, finalize_ms_, steps_done_ and stop_.
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::deque<JobRec*> finq_;
    std::atomic<int> finishing_{0}, completed_{0};
    bool stop_ = false;
    hipStream_t fstream_ = nullptr;
    std::thread worker_;
};

}  // namespace chaos
