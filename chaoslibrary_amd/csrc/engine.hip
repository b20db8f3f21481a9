// engine.hip -- batched execution of many sweeps in persistent kernels.
//
// Work unit: a tile = (job slot, run of 64-point groups, tick range).  Each
// round carries one tile per run of groups of every active job.  Jobs fall
// into families (program x system x recorded variable); every family has its
// own persistent kernel, compiled alone so that it keeps the one-shot
// kernels' register budget, and its own queue in the round's tile list.
// The family's wavefronts (as many as it has claims, at most its persistent
// grid) pull tiles from their family's queue (one atomic per claim of
// `take` tiles), resume the groups' lane state from HBM, advance it, and
// park it again (or write the job's outputs when its last tick is done).  The per-lane programs are
// the ones the one-shot kernels run (lanes.hpp), so a job's results do not
// depend on how it was batched.  The host side admits jobs FIFO while job
// slots and the HBM budget allow, replays each round from a hipGraph
// captured per set of families, checks the round's device error word before
// it advances any job, and finishes completed jobs with the same
// post-processing stages as the one-shot entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "basins.hpp"
#include "csv.hpp"
#include "bifurcation.hpp"
#include "dispatch.hpp"
#include "engine.hpp"
#include "lanes.hpp"
#include "lle.hpp"

namespace chaos {

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------

__device__ __forceinline__ void report(int* err, int bits)
{
    if (__lane_id() == 0) atomicOr(err, bits);
}

// One 64-point group of an LLE job over ticks [T.g0, T.g1).
template <int SYS>
__device__ __forceinline__ void lle_group(const JobDev& J, const Tile& T, long long group, int lane, int* err)
{
    const LleArgs& p = J.lle;
    const long long q = group * 64 + lane;
    const bool valid = q < J.n_points;
    const long long qq = valid ? q : 0;
    const int i = (int)(qq % p.nPts), j = (int)(qq / p.nPts);
    LleProg<SYS> L;
    L.params(p, i, j);
    if (T.g0 == 0) L.init(p);
    else L.load(J.st, q);
    // the perturbation was stored with the job (lle_offset_kernel)
    if (!L.template ticks<true>(p, 0, T.g0, T.g1, LleProg<SYS>::dir_of(J.st, q), J.st.np)) {
        report(err, ENG_ERR_PROGRESS);
        return;
    }
    if (T.g1 == J.total) {
        if (valid) J.lle_out[q] = L.result(p);
    } else {
        L.store(J.st, q);
    }
}

// One 64-point group of a sweep job; WV is the recorded variable.
template <int SYS, int KIND, int WV>
__device__ __forceinline__ void sweep_group(const JobDev& J, const Tile& T, long long group, int lane)
{
    const SweepArgs& A = J.sw;
    const long long q = group * 64 + lane;
    const bool in_range = q < A.n_points;
    const long long qq = in_range ? q : 0;
    SweepProg<SYS, KIND> L;
    WaveCursor cur{A.pool.nchunks, kChunk};
    if (T.g0 == 0) {
        L.init(A, qq, in_range);
    } else {
        L.params(A, qq);
        L.load(J.st, q);
        if (KIND == KIND_BIF) {
            const WaveCursor c = J.cursors[group];
            cur.chunk = __builtin_amdgcn_readfirstlane(c.chunk);
            cur.fill = __builtin_amdgcn_readfirstlane(c.fill);
        }
    }
    L.template ticks<WV>(A, WV, cur, (int)group, T.g0, T.g1);
    if (T.g1 == J.total) {
        if (KIND == KIND_BIF) close_chunk(A.pool, cur);
        if (in_range) L.finish(A, q);
    } else {
        L.store(J.st, q);
        if (KIND == KIND_BIF && lane == 0) J.cursors[group] = cur;
    }
}

// Tile checks shared by the families: the tile's job must be of the
// kernel's family, its tick range inside [0, total] and its groups inside
// the job.  A failed check sets an error bit and skips the tile; the host
// reports the round as failed.
__device__ __forceinline__ bool tile_ok(const JobDev& J, const Tile& T, int fam, int* err)
{
    if (engine_family(J.prog, J.lorenz, J.prog == PROG_LLE ? 0 : J.wv) != fam) {
        report(err, ENG_ERR_FAMILY);
        return false;
    }
    if (T.g0 < 0 || T.g1 <= T.g0 || T.g1 > J.total || T.group < 0 || (long long)(T.group + 1) * 64 > J.np) {
        report(err, ENG_ERR_RANGE);
        return false;
    }
    return true;
}

// Persistent loop of one family: every taking part wave (global index <
// Q->waves, family_waves) claims `take` consecutive tiles of its family's
// queue per atomic (points per tile = 64 x take) until the queue is empty;
// each wave reaches the exit test after every claim, so the grid drains once
// the last tile is taken.
template <class Body>
__device__ __forceinline__ void drain_family(const Tile* __restrict__ tiles, RoundCtl* __restrict__ ctl, int fam,
                                             Body&& body)
{
    const int lane = threadIdx.x & 63;
    FamQueue* Q = &ctl->fam[fam];
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= Q->waves) return;
    const int off = Q->off, n = Q->n, take = ctl->take;
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = atomicAdd(&Q->head, (unsigned)take);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if ((int)t >= n) break;
        const int e = min((int)t + take, n);
        for (int u = (int)t; u < e; ++u) body(tiles[off + u], lane);
    }
}

template <int SYS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void engine_lle_kernel(const JobDev* __restrict__ jobs, const Tile* __restrict__ tiles, RoundCtl* __restrict__ ctl)
{
    constexpr int FAM = SYS == SYS_LORENZ_RK4 ? 1 : 0;
    drain_family(tiles, ctl, FAM, [&](const Tile& T, int lane) {
        const JobDev& J = jobs[T.slot];
        if (!tile_ok(J, T, FAM, &ctl->err)) return;
        lle_group<SYS>(J, T, T.group, lane, &ctl->err);
    });
}

// The sweep programs need up to 80 VGPRs here (lane state load/store on
// top of the one-shot kernels' loop): 6 waves per SIMD, no spills.
template <int SYS, int KIND, int WV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
void engine_sweep_kernel(const JobDev* __restrict__ jobs, const Tile* __restrict__ tiles, RoundCtl* __restrict__ ctl)
{
    constexpr int FAM = 2 + (((SYS == SYS_LORENZ_RK4 ? 1 : 0) * 3 + KIND) * 3 + WV);
    drain_family(tiles, ctl, FAM, [&](const Tile& T, int lane) {
        const JobDev& J = jobs[T.slot];
        if (!tile_ok(J, T, FAM, &ctl->err)) return;
        sweep_group<SYS, KIND, WV>(J, T, T.group, lane);
    });
}

// Stores eps * (seeded unit vector) of every lane of an LLE job beside its
// parked state, once, when the job is admitted: the tile loop then adds it
// at tick S instead of generating it (the generator's registers would
// otherwise sit in the tile loop).  Same value as the one-shot kernel's
// perturbation (LleProg::offset).
template <int SYS>
__global__ void lle_offset_kernel(const JobDev* __restrict__ jobs, int slot)
{
    const JobDev& J = jobs[slot];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= J.np) return;
    const long long qq = q < J.n_points ? q : 0;
    const int i = (int)(qq % J.lle.nPts), j = (int)(qq / J.lle.nPts);
    double d[LleProg<SYS>::NX];
    LleProg<SYS>::offset(J.lle, (long long)j * J.lle.nPts + i, d);
    double* dir = J.st.d + LleProg<SYS>::ND * J.st.np + q;
    for (int k = 0; k < LleProg<SYS>::NX; ++k) dir[k * J.st.np] = d[k];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static double steady_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static chaos_options job_options(const chaos_job& j, int device, hipStream_t st)
{
    chaos_options o = resolve_options(nullptr);
    o.system = j.system;
    o.semantics = j.semantics;
    o.seed = j.seed;
    o.min_pts = j.min_pts > 0 ? j.min_pts : 20;
    o.device = device;
    o.stream = st;
    o.device_buffers = 0;
    o.block = 0;
    o.stats = nullptr;
    return o;
}

Engine::Engine(int device, const chaos_engine_params& prm) : prm_(prm), device_(device)
{
    CHAOS_REQUIRE(prm_.block == 64 || prm_.block == 128 || prm_.block == 256, "block must be 64, 128 or 256");
    CHAOS_REQUIRE(prm_.waves_per_simd >= 1 && prm_.waves_per_simd <= 8, "waves_per_simd must be in [1,8]");
    CHAOS_REQUIRE(prm_.steps_per_launch >= 2, "steps_per_launch must be >= 2");
    CHAOS_REQUIRE(prm_.max_jobs >= 1 && prm_.max_tiles >= 1, "max_jobs and max_tiles must be >= 1");
    if (prm_.points_per_tile == 0) prm_.points_per_tile = 64;
    CHAOS_REQUIRE(prm_.points_per_tile == 64 || prm_.points_per_tile == 128 || prm_.points_per_tile == 256,
                  "points_per_tile must be 64, 128 or 256");
    gpt_ = prm_.points_per_tile / 64;
    chaos_options o = resolve_options(nullptr);
    o.device = device;
    ctx_.reset(new CallCtx(o));
    CHAOS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));  // default priority
    o.stream = stream_;
    ctx_.reset(new CallCtx(o));
    CHAOS_HIP(hipEventCreate(&ev0_));
    CHAOS_HIP(hipEventCreate(&ev1_));
    for (int k = 0; k < kLanes - 1; ++k) {
        CHAOS_HIP(hipStreamCreateWithFlags(&aux_[k], hipStreamNonBlocking));  // default priority
        CHAOS_HIP(hipEventCreateWithFlags(&join_[k], hipEventDisableTiming));
    }
    CHAOS_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    d_jobs_.alloc(prm_.max_jobs);
    d_tiles_.alloc(prm_.max_tiles);
    d_ctl_.alloc(1);
    CHAOS_HIP(hipHostMalloc((void**)&h_jobs_, sizeof(JobDev) * prm_.max_jobs, hipHostMallocDefault));
    CHAOS_HIP(hipHostMalloc((void**)&h_tiles_, sizeof(Tile) * prm_.max_tiles, hipHostMallocDefault));
    CHAOS_HIP(hipHostMalloc((void**)&h_ctl_, sizeof(RoundCtl), hipHostMallocDefault));
    CHAOS_HIP(hipHostMalloc((void**)&h_err_, sizeof(int), hipHostMallocDefault));
    *h_ctl_ = RoundCtl{};
    *h_err_ = 0;
    for (int s = prm_.max_jobs - 1; s >= 0; --s) free_slots_.push_back(s);
    const hipDeviceProp_t& props = device_props(device);
    const int waves = props.multiProcessorCount * 4 * prm_.waves_per_simd;
    grid_ = std::max(1, waves / (prm_.block / 64));
    const size_t free_b = dev_mem_available();
    hbm_budget_ = (long long)(prm_.hbm_fraction * (double)free_b);
    for (hipStream_t& fs : fstreams_)
        CHAOS_HIP(hipStreamCreateWithFlags(&fs, hipStreamNonBlocking));  // default priority
    t0_ = steady_ms();
    for (int k = 0; k < kFinishers; ++k) workers_.emplace_back([this, k] { finisher(k); });
}

// Worker k: takes finished jobs in completion order and finishes them on its
// stream.  A failure marks the job failed with its message; the engine keeps
// serving the others.
void Engine::finisher(int k)
{
    (void)hipSetDevice(device_);
    hipStream_t fstream = fstreams_[k];
    std::unique_ptr<CallCtx> fctx;
    for (;;) {
        JobRec* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !finq_.empty(); });
            if (finq_.empty()) return;
            r = finq_.front();
            finq_.pop_front();
        }
        const double t0 = steady_ms();
        int st = 2;
        std::string err;
        try {
            if (!fctx) {
                chaos_options o = resolve_options(nullptr);
                o.device = device_;
                o.stream = fstream;
                fctx.reset(new CallCtx(o));
            }
            finalize(*r, *fctx);
        } catch (const std::exception& e) {
            st = -1;
            err = e.what();
            (void)hipStreamSynchronize(fstream);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            finalize_ms_ += steady_ms() - t0;
            hbm_used_ -= r->footprint;
            if (st == 2) steps_done_ += r->rk4_steps;
            r->error = err;
            r->t_done = now_ms();
            r->status = st;
            --finishing_;
            ++completed_;
        }
        cv_.notify_all();
    }
}

int Engine::take_completed() { return completed_.exchange(0); }

void Engine::wait_finisher_progress()
{
    std::unique_lock<std::mutex> lk(mu_);
    const int before = finishing_.load();
    if (before == 0) return;
    cv_.wait(lk, [&] { return finishing_.load() < before; });
}

Engine::~Engine()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& w : workers_)
        if (w.joinable()) w.join();
    for (hipStream_t fs : fstreams_)
        if (fs) (void)hipStreamDestroy(fs);
    jobs_.clear();
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
    if (h_jobs_) (void)hipHostFree(h_jobs_);
    if (h_tiles_) (void)hipHostFree(h_tiles_);
    if (h_ctl_) (void)hipHostFree(h_ctl_);
    if (h_err_) (void)hipHostFree(h_err_);
    for (int k = 0; k < kLanes - 1; ++k) {
        if (join_[k]) (void)hipEventDestroy(join_[k]);
        if (aux_[k]) (void)hipStreamDestroy(aux_[k]);
    }
    if (fork_) (void)hipEventDestroy(fork_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    d_jobs_.release();
    d_tiles_.release();
    d_ctl_.release();
    ctx_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
}

double Engine::now_ms() const { return steady_ms() - t0_; }

// Validate and size a job (no device work).
void Engine::plan(JobRec& r)
{
    const chaos_job& j = r.spec;
    CHAOS_REQUIRE(j.system == CHAOS_SYS_ROSSLER_RK4 || j.system == CHAOS_SYS_LORENZ_RK4,
                  "the engine runs the RK4 systems only (got system %d); use the one-shot entry points", j.system);
    const chaos_options o = job_options(j, device_, stream_);
    int nx = 0, na = 0;
    chaos_system_dims(j.system, &nx, &na);
    r.lorenz = j.system == CHAOS_SYS_LORENZ_RK4 ? 1 : 0;
    const long long K = prm_.steps_per_launch;
    if (j.kind == CHAOS_JOB_LLE2D) {
        r.lle = make_lle_args(o, j.tMax, j.NT, j.nPts, j.h, j.eps, j.initialConditions, j.amountOfInitialConditions,
                              j.ranges, j.indicesOfMutVars, j.transientTime, j.values, j.amountOfValues);
        r.lle.row0 = 0;
        r.n = (long long)j.nPts * j.nPts;
        r.lle.n_points = r.n;
        r.total = (long long)r.lle.S + (long long)r.lle.W * r.lle.NTp;
        r.ticks_per_launch = std::max<long long>(1, K / 2);
        r.rk4_steps = (double)r.n * ((double)r.lle.S + 2.0 * r.lle.W * r.lle.NTp);
        r.prog = PROG_LLE;
    } else if (j.kind == CHAOS_JOB_BIF1D || j.kind == CHAOS_JOB_BIF2D) {
        BifInputs in{};
        in.tMax = j.tMax;
        in.nPts = j.nPts;
        in.h = j.h;
        in.nx = j.amountOfInitialConditions;
        in.ic = j.initialConditions;
        in.ranges = j.ranges;
        in.idx = j.indicesOfMutVars;
        in.writableVar = j.writableVar;
        in.maxValue = j.maxValue;
        in.transientTime = j.transientTime;
        in.values = j.values;
        in.na = j.amountOfValues;
        in.preScaller = j.preScaller;
        in.eps = j.eps;
        in.dims = j.kind == CHAOS_JOB_BIF2D ? 2 : 1;
        r.sw = make_bif_args(o, in);
        r.n = r.sw.n_points;
        r.total = (long long)r.sw.S + r.sw.B;
        r.ticks_per_launch = K;
        r.rk4_steps = (double)r.n * ((double)r.sw.S + (double)r.sw.B);
        r.prog = PROG_BIF;
        r.pool_chunks = bif_initial_chunks(r.sw);
    } else if (j.kind == CHAOS_JOB_BASINS2D) {
        r.sw = make_basins_args(o, j.tMax, j.nPts, j.h, j.amountOfInitialConditions, j.initialConditions, j.ranges,
                                j.indicesOfMutVars, j.writableVar, j.maxValue, j.transientTime, j.values,
                                j.amountOfValues, j.preScaller, j.eps);
        r.n = r.sw.n_points;
        r.total = (long long)r.sw.S + r.sw.B;
        r.ticks_per_launch = std::max<long long>(1, K / std::max(1, r.sw.pre));
        r.rk4_steps = (double)r.n * ((double)r.sw.S + (double)r.sw.B * r.sw.pre);
        r.prog = j.semantics == CHAOS_SEM_REFERENCE ? PROG_BASINS_REF : PROG_BASINS_CFG;
    } else {
        fail(CHAOS_E_INVALID, "unknown job kind %d", j.kind);
    }
    CHAOS_REQUIRE(j.writableVar >= 0 && j.writableVar < nx, "writableVar %d out of range [0,%d)", j.writableVar, nx);
    r.family = engine_family(r.prog, r.lorenz, r.prog == PROG_LLE ? 0 : j.writableVar);
    r.groups = (r.n + 63) / 64;
    r.np = r.groups * 64;
    r.tiles = r.groups;
    CHAOS_REQUIRE(r.tiles <= prm_.max_tiles, "job needs %lld tiles, queue holds %d", r.tiles, prm_.max_tiles);
    // HBM footprint: lane state + outputs + per-kind working sets
    long long fp = 0;
    if (r.prog == PROG_LLE) {
        fp = r.np * (3 * nx + 1) * 8 + r.n * 8;  // parked state + stored perturbation, map
    } else {
        fp = r.np * ((nx + 8) * 8 + 4 * 4);
        if (r.prog == PROG_BIF)
            fp += r.groups * (long long)sizeof(WaveCursor) + r.n * 16 + r.pool_chunks * (kChunk * 16 + 8) +
                  r.n * 8 + r.pool_chunks * kChunk * 12;
        else
            fp += r.n * (3 * 8 + 4 + 4) + r.n * 200;  // features, flags, labels, clustering scratch
    }
    r.footprint = fp;
}

long long Engine::submit(const chaos_job& j)
{
    std::unique_ptr<JobRec> r(new JobRec());
    r->spec = j;
    plan(*r);
    CHAOS_REQUIRE(r->footprint <= hbm_budget_, "job footprint %lld bytes exceeds the HBM budget %lld", r->footprint,
                  hbm_budget_);
    r->id = next_id_++;
    r->t_submit = now_ms();
    const long long id = r->id;
    pending_.push_back(id);
    jobs_[id] = std::move(r);
    return id;
}

void Engine::allocate(JobRec& r)
{
    const int nx = r.lorenz ? System<SYS_LORENZ_RK4>::NX : System<SYS_ROSSLER_RK4>::NX;
    JobDev& D = h_jobs_[r.slot];
    D = JobDev{};
    D.prog = r.prog;
    D.lorenz = r.lorenz;
    D.wv = r.spec.writableVar;
    D.np = r.np;
    D.n_points = r.n;
    D.total = r.total;
    if (r.prog == PROG_LLE) {
        r.dstate.alloc((size_t)r.np * (3 * nx + 1));
        r.out.alloc((size_t)r.n);
        D.st = LaneStore{r.dstate.get(), nullptr, r.np};
        D.lle = r.lle;
        D.lle_out = r.out.get();
    } else {
        r.dstate.alloc((size_t)r.np * (nx + 8));
        r.istate.alloc((size_t)r.np * 4);
        D.st = LaneStore{r.dstate.get(), r.istate.get(), r.np};
        D.sw = r.sw;
        if (r.prog == PROG_BIF) {
            r.bif.n = r.n;
            r.bif.hp = r.sw.hp;
            r.bif.samples = r.sw.B;
            r.bif.need_times = !(r.spec.kind == CHAOS_JOB_BIF2D && r.spec.semantics == CHAOS_SEM_CONFIG);
            r.bif.need_vals = !(r.spec.kind == CHAOS_JOB_BIF2D && r.spec.semantics == CHAOS_SEM_REFERENCE);
            r.bif.need_host = r.spec.kind == CHAOS_JOB_BIF1D;
            r.bif.flags.alloc(r.n);
            r.bif.counts.alloc(r.n);
            r.bif.first.alloc(r.n);
            r.pool.alloc(stream_, r.pool_chunks);
            r.cursors.alloc((size_t)r.groups);
            D.sw.flags = r.bif.flags.get();
            D.sw.counts = r.bif.counts.get();
            D.sw.first = r.bif.first.get();
            D.sw.pool = r.pool.view();
            D.cursors = r.cursors.get();
        } else {
            r.flags.alloc(r.n);
            r.feats.alloc(3 * (size_t)r.n);
            D.sw.flags = r.flags.get();
            D.sw.f0 = r.feats.get();
            D.sw.f1 = r.feats.get() + r.n;
            D.sw.f2 = r.feats.get() + 2 * r.n;
            D.sw.counts = nullptr;
        }
    }
    CHAOS_HIP(hipMemcpyAsync(d_jobs_.get() + r.slot, &D, sizeof(JobDev), hipMemcpyHostToDevice, stream_));
    if (r.prog == PROG_LLE) {
        const unsigned g = ceil_div((unsigned long long)r.np, 256);
        if (r.lorenz)
            hipLaunchKernelGGL(lle_offset_kernel<SYS_LORENZ_RK4>, dim3(g), dim3(256), 0, stream_, d_jobs_.get(), r.slot);
        else
            hipLaunchKernelGGL(lle_offset_kernel<SYS_ROSSLER_RK4>, dim3(g), dim3(256), 0, stream_, d_jobs_.get(),
                               r.slot);
        CHAOS_HIP(hipGetLastError());
    }
}

void Engine::admit()
{
    const int m = (int)std::min(pending_.size(), free_slots_.size());
    if (m == 0) return;
    std::vector<long long> fp(m), tl(m);
    for (int i = 0; i < m; ++i) {
        const JobRec& r = *jobs_[pending_[i]];
        fp[i] = r.footprint;
        tl[i] = r.tiles;
    }
    int k;
    {
        // the worker thread only lowers hbm_used_, so the rule sees at most
        // the footprints still charged
        std::lock_guard<std::mutex> lk(mu_);
        k = admit_fifo(fp.data(), tl.data(), m, (int)free_slots_.size(), &hbm_used_, hbm_budget_, &active_tiles_,
                       prm_.max_tiles);
    }
    for (int i = 0; i < k; ++i) {
        JobRec& r = *jobs_[pending_.front()];
        r.slot = free_slots_.back();
        free_slots_.pop_back();
        r.status = 1;
        r.t_admit = now_ms();
        try {
            allocate(r);
        } catch (...) {
            // this job and the ones after it stay pending, uncharged
            free_slots_.push_back(r.slot);
            r.slot = -1;
            r.status = 0;
            std::lock_guard<std::mutex> lk(mu_);
            for (int j = i; j < k; ++j) {
                hbm_used_ -= fp[j];
                active_tiles_ -= tl[j];
            }
            throw;
        }
        active_.push_back(r.id);
        pending_.pop_front();
    }
}

namespace {

template <int F>
struct FamilyKernel {
    static constexpr int sweep = F - 2;
    static constexpr int sys = (F < 2 ? F : sweep / 9) == 1 ? SYS_LORENZ_RK4 : SYS_ROSSLER_RK4;
    static constexpr int kind = (sweep / 3) % 3;
    static constexpr int wv = sweep % 3;
    static void launch(int grid, int block, hipStream_t s, const JobDev* jobs, const Tile* tiles, RoundCtl* ctl)
    {
        if constexpr (F < 2)
            hipLaunchKernelGGL(engine_lle_kernel<sys>, dim3(grid), dim3(block), 0, s, jobs, tiles, ctl);
        else
            hipLaunchKernelGGL((engine_sweep_kernel<sys, kind, wv>), dim3(grid), dim3(block), 0, s, jobs, tiles, ctl);
    }
};

template <int... F>
void launch_family_of(int f, std::integer_sequence<int, F...>, int grid, int block, hipStream_t s,
                      const JobDev* jobs, const Tile* tiles, RoundCtl* ctl)
{
    ((f == F ? FamilyKernel<F>::launch(grid, block, s, jobs, tiles, ctl) : void()), ...);
}

}  // namespace

void Engine::launch_family(int f, hipStream_t s)
{
    launch_family_of(f, std::make_integer_sequence<int, kFamilies>{}, grid_, prm_.block, s, d_jobs_.get(),
                     d_tiles_.get(), d_ctl_.get());
    CHAOS_HIP(hipGetLastError());
}

// The per-round launch sequence: the tile list and the round control (queue
// heads and error word at 0: a copy, no memset node) to the device, the
// kernels of the round's families spread over the lanes' streams (fork /
// join by events), the error word back to pinned host memory.
void Engine::enqueue_round(unsigned mask)
{
    CHAOS_HIP(hipMemcpyAsync(d_tiles_.get(), h_tiles_, sizeof(Tile) * (size_t)prm_.max_tiles, hipMemcpyHostToDevice,
                             stream_));
    CHAOS_HIP(hipMemcpyAsync(d_ctl_.get(), h_ctl_, sizeof(RoundCtl), hipMemcpyHostToDevice, stream_));
    int used = 0;
    for (int f = 0; f < kFamilies; ++f)
        if (mask >> f & 1u) ++used;
    const int lanes = std::min(used, (int)kLanes);
    if (lanes > 1) {
        CHAOS_HIP(hipEventRecord(fork_, stream_));
        for (int k = 0; k < lanes - 1; ++k) CHAOS_HIP(hipStreamWaitEvent(aux_[k], fork_, 0));
    }
    int i = 0;
    for (int f = 0; f < kFamilies; ++f) {
        if (!(mask >> f & 1u)) continue;
        const int lane = i++ % lanes;
        launch_family(f, lane == 0 ? stream_ : aux_[lane - 1]);
    }
    if (lanes > 1) {
        for (int k = 0; k < lanes - 1; ++k) {
            CHAOS_HIP(hipEventRecord(join_[k], aux_[k]));
            CHAOS_HIP(hipStreamWaitEvent(stream_, join_[k], 0));
        }
    }
    CHAOS_HIP(hipMemcpyAsync(h_err_, &d_ctl_.get()->err, sizeof(int), hipMemcpyDeviceToHost, stream_));
}

hipGraphExec_t Engine::capture(unsigned mask)
{
    auto it = graphs_.find(mask);
    if (it != graphs_.end()) return it->second;
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    CHAOS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    enqueue_round(mask);
    CHAOS_HIP(hipStreamEndCapture(stream_, &g));
    const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CHAOS_HIP(ie);
    graphs_[mask] = e;
    return e;
}

// A device error fails every job still integrating (their lane state can no
// longer be trusted) and the call that ran the round.
void Engine::fail_active(const std::string& msg)
{
    for (long long id : active_) {
        JobRec& r = *jobs_[id];
        r.error = msg;
        r.t_done = now_ms();
        r.status = -1;
        active_tiles_ -= r.tiles;
        free_slots_.push_back(r.slot);
        r.slot = -1;
        std::lock_guard<std::mutex> lk(mu_);
        hbm_used_ -= r.footprint;
    }
    active_.clear();
    fail(CHAOS_E_HIP, "%s", msg.c_str());
}

void Engine::finalize(JobRec& r, CallCtx& ctx)
{
    hipStream_t st = ctx.stream();
    const chaos_job& j = r.spec;
    if (r.prog == PROG_LLE) {
        r.r_lle.resize(r.n);
        CHAOS_HIP(hipMemcpyAsync(r.r_lle.data(), r.out.get(), r.n * 8, hipMemcpyDeviceToHost, st));
        CHAOS_HIP(hipStreamSynchronize(st));
    } else if (r.prog == PROG_BIF) {
        bool ok = bif_post(ctx, r.bif, r.pool);
        if (!ok) {
            // peak store overflow: rerun this job alone with the exact pool
            r.reran = 1;
            chaos_options o = job_options(j, device_, st);
            if (j.kind == CHAOS_JOB_BIF1D) {
                r.r_counts.resize(r.n);
                r.r_first.resize(r.n);
                long long total = 0;
                int rc = chaos_bif1d(&o, j.tMax, j.nPts, j.h, j.amountOfInitialConditions, j.initialConditions,
                                     j.ranges, j.indicesOfMutVars, j.writableVar, j.maxValue, j.transientTime,
                                     j.values, j.amountOfValues, j.preScaller, r.r_counts.data(), r.r_first.data(),
                                     nullptr, nullptr, 0, &total);
                if (rc != CHAOS_OK && rc != CHAOS_E_CAPACITY) throw Error(rc, chaos_last_error());
                r.r_peaks.resize(std::max(1LL, total));
                r.r_intervals.resize(std::max(1LL, total));
                rc = chaos_bif1d(&o, j.tMax, j.nPts, j.h, j.amountOfInitialConditions, j.initialConditions, j.ranges,
                                 j.indicesOfMutVars, j.writableVar, j.maxValue, j.transientTime, j.values,
                                 j.amountOfValues, j.preScaller, r.r_counts.data(), r.r_first.data(), r.r_peaks.data(),
                                 r.r_intervals.data(), total, &total);
                if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
                r.r_peaks.resize(total);
                r.r_intervals.resize(total);
            } else {
                r.r_map.resize(r.n);
                int rc = chaos_bif2d(&o, j.tMax, j.nPts, j.h, j.amountOfInitialConditions, j.initialConditions,
                                     j.ranges, j.indicesOfMutVars, j.writableVar, j.maxValue, j.transientTime,
                                     j.values, j.amountOfValues, j.preScaller, j.eps, r.r_map.data());
                if (rc != CHAOS_OK) throw Error(rc, chaos_last_error());
            }
        } else if (j.kind == CHAOS_JOB_BIF1D) {
            r.r_counts.resize(r.n);
            r.r_first.resize(r.n);
            r.r_peaks.resize(std::max(1LL, r.bif.total));
            r.r_intervals.resize(std::max(1LL, r.bif.total));
            long long total = 0;
            bif1d_collect(ctx, r.bif, r.r_counts.data(), r.r_first.data(), r.r_peaks.data(), r.r_intervals.data(),
                          r.bif.total, &total);
            r.r_peaks.resize(total);
            r.r_intervals.resize(total);
        } else {
            DevBuf<int> map((size_t)r.n);
            bif2d_analysis(ctx, j.semantics, r.bif, j.eps, map.get());
            r.r_map.resize(r.n);
            CHAOS_HIP(hipMemcpyAsync(r.r_map.data(), map.get(), r.n * 4, hipMemcpyDeviceToHost, st));
            CHAOS_HIP(hipStreamSynchronize(st));
        }
    } else {
        const bool ref = r.prog == PROG_BASINS_REF;
        DevBuf<int> labels((size_t)r.n);
        basins_cluster(ctx, ref, r.n, r.feats.get(), r.feats.get() + r.n, r.feats.get() + 2 * r.n, r.flags.get(),
                       j.eps, j.min_pts > 0 ? j.min_pts : 20, labels.get());
        const int nf = ref ? 2 : 3;
        r.r_labels.resize(r.n);
        r.r_flags.resize(r.n);
        r.r_feats.resize(nf * r.n);
        CHAOS_HIP(hipMemcpyAsync(r.r_labels.data(), labels.get(), r.n * 4, hipMemcpyDeviceToHost, st));
        CHAOS_HIP(hipMemcpyAsync(r.r_flags.data(), r.flags.get(), r.n * 4, hipMemcpyDeviceToHost, st));
        CHAOS_HIP(hipMemcpyAsync(r.r_feats.data(), r.feats.get(), nf * r.n * 8, hipMemcpyDeviceToHost, st));
        CHAOS_HIP(hipStreamSynchronize(st));
    }
    // release the device working set
    r.dstate.release();
    r.istate.release();
    r.out.release();
    r.cursors.release();
    r.flags.release();
    r.feats.release();
    r.bif = BifDevice();
    r.pool = PeakPoolBuf();
}

int Engine::step()
{
    admit();
    // tiles grouped by family: family f owns tiles [off_f, off_f + n_f)
    int cnt[kFamilies] = {};
    for (long long id : active_) {
        const JobRec& r = *jobs_[id];
        cnt[r.family] += (int)r.tiles;
    }
    int fill[kFamilies];
    unsigned mask = 0;
    int nt = 0;
    const long long grid_waves = (long long)grid_ * (prm_.block / 64);
    for (int f = 0; f < kFamilies; ++f) {
        h_ctl_->fam[f] = FamQueue{nt, cnt[f], 0u, family_waves(cnt[f], gpt_, grid_waves)};
        fill[f] = nt;
        nt += cnt[f];
        if (cnt[f]) mask |= 1u << f;
    }
    h_ctl_->err = 0;
    h_ctl_->take = gpt_;
    for (long long id : active_) {
        const JobRec& r = *jobs_[id];
        const long long g0 = r.done, g1 = std::min(r.total, r.done + r.ticks_per_launch);
        for (long long g = 0; g < r.groups; ++g) h_tiles_[fill[r.family]++] = Tile{r.slot, (int)g, g0, g1};
    }
    if (nt == 0) {
        // nothing to integrate: wait for a job being finished (it may free
        // the HBM a pending job needs)
        wait_finisher_progress();
        return take_completed();
    }
    *h_err_ = 0;
    CHAOS_HIP(hipEventRecord(ev0_, stream_));
    if (prm_.use_graph) {
        CHAOS_HIP(hipGraphLaunch(capture(mask), stream_));
        ++graph_replays_;
    } else {
        enqueue_round(mask);
    }
    CHAOS_HIP(hipEventRecord(ev1_, stream_));
    CHAOS_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CHAOS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    busy_ms_ += ms;
    ++launches_;
    tiles_ += nt;
    // the jobs advance only when the device confirms the round
    if (*h_err_ != 0)
        fail_active("engine round failed on the device (error bits " + std::to_string(*h_err_) +
                    ": 1 tile of another family, 2 tile range outside its job, 4 program made no progress)");
    std::vector<long long> still;
    still.reserve(active_.size());
    std::vector<JobRec*> done;
    for (long long id : active_) {
        JobRec& r = *jobs_[id];
        r.done = std::min(r.total, r.done + r.ticks_per_launch);
        if (r.done < r.total) {
            still.push_back(id);
            continue;
        }
        // integration finished: the slot and tile budget are free now, the
        // HBM footprint stays charged until the worker has finished the job
        r.status = 3;
        active_tiles_ -= r.tiles;
        free_slots_.push_back(r.slot);
        r.slot = -1;
        done.push_back(&r);
    }
    active_.swap(still);
    if (!done.empty()) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (JobRec* r : done) finq_.push_back(r);
            finishing_ += (int)done.size();
        }
        cv_.notify_all();
    }
    return take_completed();
}

void Engine::drain()
{
    while (!pending_.empty() || !active_.empty()) {
        const size_t before = pending_.size() + active_.size();
        step();
        if (active_.empty() && !pending_.empty() && pending_.size() + active_.size() == before &&
            finishing_.load() == 0)
            fail(CHAOS_E_BUSY, "pending job cannot be admitted");
    }
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return finishing_.load() == 0; });
}

JobRec& Engine::job(long long id)
{
    auto it = jobs_.find(id);
    if (it == jobs_.end()) fail(CHAOS_E_NOTFOUND, "unknown job id %lld", id);
    return *it->second;
}

void Engine::release(long long id)
{
    JobRec& r = job(id);
    CHAOS_REQUIRE(r.status == 2 || r.status == -1, "job %lld is still running", id);
    jobs_.erase(id);
}

chaos_server_stats Engine::stats() const
{
    std::lock_guard<std::mutex> lk(mu_);
    chaos_server_stats s{};
    s.launches = launches_;
    s.graph_replays = graph_replays_;
    s.tiles = tiles_;
    s.busy_ms = busy_ms_;
    s.finalize_ms = finalize_ms_;
    s.rk4_steps = steps_done_;
    s.active_jobs = (int)active_.size();
    s.pending_jobs = (int)pending_.size();
    s.hbm_in_use = hbm_used_;
    s.hbm_budget = hbm_budget_;
    s.finishing_jobs = finishing_.load();
    return s;
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_engine_params_default(chaos_engine_params* p)
{
    if (!p) return CHAOS_E_INVALID;
    // the autotuner's choice on an MI355X (power of a short saturated
    // served stream at full horizons; 10000-40000 steps per launch tie
    // within 0.5 %, 40000 completing the most work per second)
    p->block = 256;
    p->waves_per_simd = 4;
    p->steps_per_launch = 40000;
    p->max_jobs = 256;
    p->max_tiles = 65536;
    p->use_graph = 1;
    p->hbm_fraction = 0.6;
    p->points_per_tile = 64;
    p->reserved = 0;
    return CHAOS_OK;
}

extern "C" int chaos_server_create(int device, const chaos_engine_params* p, void** server)
{
    return guarded("chaos_server_create", [&] {
        CHAOS_REQUIRE(server, "server is NULL");
        chaos_engine_params prm;
        chaos_engine_params_default(&prm);
        if (p) prm = *p;
        *server = new Engine(device, prm);
    });
}

extern "C" int chaos_server_destroy(void* server)
{
    return guarded("chaos_server_destroy", [&] { delete static_cast<Engine*>(server); });
}

extern "C" int chaos_server_submit(void* server, const chaos_job* job, long long* job_id)
{
    return guarded("chaos_server_submit", [&] {
        CHAOS_REQUIRE(server && job && job_id, "NULL argument");
        *job_id = static_cast<Engine*>(server)->submit(*job);
    });
}

extern "C" int chaos_server_step(void* server, int* completed)
{
    return guarded("chaos_server_step", [&] {
        CHAOS_REQUIRE(server, "server is NULL");
        const int c = static_cast<Engine*>(server)->step();
        if (completed) *completed = c;
    });
}

extern "C" int chaos_server_drain(void* server)
{
    return guarded("chaos_server_drain", [&] {
        CHAOS_REQUIRE(server, "server is NULL");
        static_cast<Engine*>(server)->drain();
    });
}

extern "C" int chaos_server_job_info(void* server, long long job_id, chaos_job_info* info)
{
    return guarded("chaos_server_job_info", [&] {
        CHAOS_REQUIRE(server && info, "NULL argument");
        const JobRec& r = static_cast<Engine*>(server)->job(job_id);
        *info = chaos_job_info{};
        info->status = r.status;
        info->kind = r.spec.kind;
        info->t_submit_ms = r.t_submit;
        info->t_admit_ms = r.t_admit;
        info->t_done_ms = r.t_done;
        info->rk4_steps = r.rk4_steps;
        info->footprint_bytes = r.footprint;
        info->n_points = r.n;
        info->result_len = r.spec.kind == CHAOS_JOB_BIF1D ? (long long)r.r_peaks.size() : r.n;
        info->reran = r.reran;
    });
}

extern "C" int chaos_server_stats_get(void* server, chaos_server_stats* stats)
{
    return guarded("chaos_server_stats_get", [&] {
        CHAOS_REQUIRE(server && stats, "NULL argument");
        *stats = static_cast<Engine*>(server)->stats();
    });
}

static const JobRec& done_job(void* server, long long id, int kind)
{
    CHAOS_REQUIRE(server, "server is NULL");
    const JobRec& r = static_cast<Engine*>(server)->job(id);
    CHAOS_REQUIRE(r.status != -1, "job %lld failed: %s", id, r.error.c_str());
    CHAOS_REQUIRE(r.status == 2, "job %lld is not finished", id);
    CHAOS_REQUIRE(r.spec.kind == kind, "job %lld is of kind %d", id, r.spec.kind);
    return r;
}

extern "C" int chaos_server_result_lle(void* server, long long job_id, double* out)
{
    return guarded("chaos_server_result_lle", [&] {
        const JobRec& r = done_job(server, job_id, CHAOS_JOB_LLE2D);
        CHAOS_REQUIRE(out, "out is NULL");
        std::copy(r.r_lle.begin(), r.r_lle.end(), out);
    });
}

extern "C" int chaos_server_result_bif2d(void* server, long long job_id, int* out)
{
    return guarded("chaos_server_result_bif2d", [&] {
        const JobRec& r = done_job(server, job_id, CHAOS_JOB_BIF2D);
        CHAOS_REQUIRE(out, "out is NULL");
        std::copy(r.r_map.begin(), r.r_map.end(), out);
    });
}

extern "C" int chaos_server_result_bif1d(void* server, long long job_id, int* counts, double* first, double* peaks,
                                         double* intervals, long long capacity, long long* total)
{
    return guarded("chaos_server_result_bif1d", [&] {
        const JobRec& r = done_job(server, job_id, CHAOS_JOB_BIF1D);
        CHAOS_REQUIRE(counts && first && total, "NULL output");
        std::copy(r.r_counts.begin(), r.r_counts.end(), counts);
        std::copy(r.r_first.begin(), r.r_first.end(), first);
        *total = (long long)r.r_peaks.size();
        if (*total > capacity) fail(CHAOS_E_CAPACITY, "need capacity %lld", *total);
        std::copy(r.r_peaks.begin(), r.r_peaks.end(), peaks);
        std::copy(r.r_intervals.begin(), r.r_intervals.end(), intervals);
    });
}

extern "C" int chaos_server_result_basins(void* server, long long job_id, int* labels, double* features, int* flags)
{
    return guarded("chaos_server_result_basins", [&] {
        const JobRec& r = done_job(server, job_id, CHAOS_JOB_BASINS2D);
        CHAOS_REQUIRE(labels, "labels is NULL");
        std::copy(r.r_labels.begin(), r.r_labels.end(), labels);
        if (features) std::copy(r.r_feats.begin(), r.r_feats.end(), features);
        if (flags) std::copy(r.r_flags.begin(), r.r_flags.end(), flags);
    });
}

extern "C" int chaos_server_write_result(void* server, long long job_id, const char* OUT_FILE_PATH, int output)
{
    return guarded("chaos_server_write_result", [&] {
        CHAOS_REQUIRE(server, "server is NULL");
        CHAOS_REQUIRE(OUT_FILE_PATH, "OUT_FILE_PATH is NULL");
        CHAOS_REQUIRE(output >= 0 && output <= (CHAOS_OUT_CSV | CHAOS_OUT_BINARY), "invalid output %d", output);
        const JobRec& r = static_cast<Engine*>(server)->job(job_id);
        CHAOS_REQUIRE(r.status == 2, "job %lld is not finished", job_id);
        const chaos_job& j = r.spec;
        switch (j.kind) {
        case CHAOS_JOB_LLE2D: {
            std::unique_ptr<TextFile> f;
            if (out_csv(output)) f.reset(new TextFile(OUT_FILE_PATH));
            write_lle_file(f.get(), OUT_FILE_PATH, j.ranges, j.nPts, r.r_lle.data(), output);
            break;
        }
        case CHAOS_JOB_BIF2D: {
            std::unique_ptr<TextFile> f;
            if (out_csv(output)) f.reset(new TextFile(OUT_FILE_PATH));
            write_int_map_file(f.get(), OUT_FILE_PATH, j.ranges, j.nPts, r.r_map.data(), output);
            break;
        }
        case CHAOS_JOB_BIF1D: {
            TextFile f(OUT_FILE_PATH);
            write_range_header(f, j.ranges, 1, 6);
            write_bif1d_body(f, j.ranges, j.nPts, r.r_counts.data(), r.r_first.data(), r.r_peaks.data(),
                             r.r_intervals.data());
            f.close();
            break;
        }
        case CHAOS_JOB_BASINS2D: {
            const int nfeat = j.semantics == CHAOS_SEM_REFERENCE ? 2 : 3;
            write_basins_headers(OUT_FILE_PATH, j.ranges, nfeat, output);
            write_basins_bodies(OUT_FILE_PATH, j.ranges, j.nPts, nfeat, r.r_labels.data(), r.r_feats.data(),
                                r.r_flags.data(), output);
            break;
        }
        default:
            fail(CHAOS_E_INVALID, "job %lld has unknown kind %d", job_id, j.kind);
        }
    });
}

extern "C" int chaos_server_release(void* server, long long job_id)
{
    return guarded("chaos_server_release", [&] {
        CHAOS_REQUIRE(server, "server is NULL");
        static_cast<Engine*>(server)->release(job_id);
    });
}
