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

#include "admission.hpp"
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
// A job's groups sit consecutively in its family's queue; a wave claims
// `take` of them at once (RoundCtl.take).
struct Tile {
    int slot;
    int group;
    long long g0, g1;
};

// Kernel families: one persistent kernel per (program, system, recorded
// variable), each compiled on its own: 2 LLE families (Rossler, Lorenz) and
// 2 systems x 3 sweep kinds x 3 recorded variables.
constexpr int kFamilies = 2 + 2 * 3 * 3;

__host__ __device__ inline int engine_family(int prog, int lorenz, int wv)
{
    return prog == PROG_LLE ? lorenz : 2 + ((lorenz * 3 + (prog - 1)) * 3 + wv);
}

// Per-round queue of one family: its tiles are tiles[off, off + n), taken
// in order through the atomic head by the family kernel's wavefronts of
// global index < waves (the others return at once).
struct FamQueue {
    int off;
    int n;
    unsigned head;
    int waves;  // family_waves (admission.hpp)
};

// Everything the host hands the device per round in one copy: the family
// queues (heads at 0), the error word (0) and the tiles a wave claims per
// atomic.  Kernels set error bits instead of looping on a malformed tile.
enum EngineError : int {
    ENG_ERR_FAMILY = 1,    // a tile's job is not of the kernel's family
    ENG_ERR_RANGE = 2,     // tick or group range outside the job
    ENG_ERR_PROGRESS = 4   // a program could not advance its tick range
};
struct RoundCtl {
    FamQueue fam[kFamilies];
    int err;
    int take;
    int pad[2];
};

struct JobRec {
    long long id = 0;
    chaos_job spec{};
    std::atomic<int> status{0};  // 0 pending, 1 active, 3 finishing, 2 done, -1 failed
    std::string error;           // failure message (status -1)
    int slot = -1;
    int prog = 0, lorenz = 0, reran = 0, family = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double rk4_steps = 0;
    long long footprint = 0;
    long long n = 0, np = 0, groups = 0, tiles = 0, total = 0, done = 0, ticks_per_launch = 0, pool_chunks = 0;
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
    int take_completed();            // jobs finished (or failed) since the last call or step()

private:
    void plan(JobRec& r);
    void admit();
    void allocate(JobRec& r);
    void finalize(JobRec& r, CallCtx& ctx);
    hipGraphExec_t capture(unsigned mask);
    void enqueue_round(unsigned mask);
    void launch_family(int f, hipStream_t s);
    void fail_active(const std::string& msg);
    void finisher(int k);            // body of worker thread k
    void wait_finisher_progress();   // block until a finishing job completes


    chaos_engine_params prm_;
    int device_ = 0;
    std::unique_ptr<CallCtx> ctx_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    DevBuf<JobDev> d_jobs_;
    DevBuf<Tile> d_tiles_;
    DevBuf<RoundCtl> d_ctl_;
    JobDev* h_jobs_ = nullptr;   // pinned
    Tile* h_tiles_ = nullptr;    // pinned
    RoundCtl* h_ctl_ = nullptr;  // pinned
    int* h_err_ = nullptr;       // pinned: the round's error word, copied back
    // family kernels of one round run on up to kLanes streams (the tail of
    // one family's queue overlaps the next family's kernel)
    static constexpr int kLanes = 4;
    hipStream_t aux_[kLanes - 1] = {};
    hipEvent_t fork_ = nullptr, join_[kLanes - 1] = {};
    std::map<unsigned, hipGraphExec_t> graphs_;  // family set -> captured round
    int grid_ = 0, gpt_ = 1;     // persistent grid; groups per tile
    long long active_tiles_ = 0;
    std::deque<long long> pending_;
    std::vector<long long> active_;
    std::map<long long, std::unique_ptr<JobRec>> jobs_;
    std::vector<int> free_slots_;
    long long next_id_ = 1, hbm_budget_ = 0, hbm_used_ = 0;
    long long launches_ = 0, graph_replays_ = 0, tiles_ = 0;
    double busy_ms_ = 0, finalize_ms_ = 0, steps_done_ = 0, t0_ = 0;
    // Finishing stages (peak post-processing, clustering, result copies) run
    // on kFinishers worker threads, each with its own stream, overlapping the
    // next rounds; the jobs waiting for them keep their HBM charged, so
    // finishing that falls behind stalls admission.  A finishing stream must
    // not share a hardware queue with a busy integration stream, or its
    // kernels wait behind whole persistent rounds: the process needs
    // GPU_MAX_HW_QUEUES >= 8 (HIP's default is 4; the Python package raises
    // it to 16 before HIP starts).
    // mu_ guards finq_, hbm_used_, finalize_ms_, steps_done_ and stop_.
    static constexpr int kFinishers = 2;
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::deque<JobRec*> finq_;
    std::atomic<int> finishing_{0}, completed_{0};
    bool stop_ = false;
    hipStream_t fstreams_[kFinishers] = {};
    std::vector<std::thread> workers_;
};

}  // namespace chaos
