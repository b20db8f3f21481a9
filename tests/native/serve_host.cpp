// Host build of the served-stream loop (chaoslibrary_amd/csrc/serve_loop.hpp)
// over a stand-in engine, for tests/test_serve_loop_cpu.py.
//
// The stand-in integrates every admitted job for a given number of rounds,
// then hands it to a "finishing thread" that completes it between two calls
// of the loop: the jobs finishing after a round are completed when the loop
// next asks whether the engine is idle -- the interleaving in which a real
// finishing thread beats the loop to its idle check.
#include <deque>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "serve_loop.hpp"

namespace {

struct Rec {
    int status = 0;  // 0 pending, 1 integrating, 3 finishing, 2 done, -1 failed
    std::string error;
    int rounds_left = 0;
    bool fails = false;
};

struct StandIn {
    std::map<long long, Rec> recs;
    std::deque<long long> pending;
    std::vector<long long> active, finishing;
    int completed = 0;
    long long next_id = 1, calls = 0, released = 0;
    long long call_limit = 0;

    void finish_all()
    {
        for (long long id : finishing) {
            Rec& r = recs[id];
            r.status = r.fails ? -1 : 2;
            if (r.fails) r.error = "stand-in failure";
            ++completed;
        }
        finishing.clear();
    }
    void count_call()
    {
        if (++calls > call_limit) throw std::runtime_error("loop did not end");
    }
    long long submit(const std::pair<int, bool>& job)
    {
        const long long id = next_id++;
        Rec r;
        r.rounds_left = job.first;
        r.fails = job.second;
        recs[id] = r;
        pending.push_back(id);
        return id;
    }
    bool idle()
    {
        count_call();
        finish_all();  // the finishing thread got there first
        return pending.empty() && active.empty() && finishing.empty();
    }
    int take_completed()
    {
        const int k = completed;
        completed = 0;
        return k;
    }
    int step()
    {
        count_call();
        for (long long id : pending) {
            recs[id].status = 1;
            active.push_back(id);
        }
        pending.clear();
        if (active.empty()) finish_all();  // nothing to integrate: wait for the finishing thread
        std::vector<long long> still;
        for (long long id : active) {
            Rec& r = recs[id];
            if (--r.rounds_left > 0) {
                still.push_back(id);
            } else {
                r.status = 3;
                finishing.push_back(id);
            }
        }
        active.swap(still);
        return take_completed();
    }
    const Rec& job(long long id) { return recs.at(id); }
    void release(long long) { ++released; }
};

}  // namespace

// Serves n jobs (arrival_ms ascending, rounds[k] >= 1 rounds of integration,
// job fail_job fails when >= 0) on a clock that advances 1 ms per loop step.
// Returns 0 when every job was served (done_ms filled), 1 when the failure
// reached on_fail, 2 when the loop did not end within 100000 engine calls.
extern "C" int serve_stand_in(int n, const double* arrival_ms, const int* rounds, int fail_job, double* done_ms,
                              long long* released)
{
    StandIn eng;
    eng.call_limit = 100000;
    std::vector<std::pair<int, bool>> jobs;
    std::vector<double> arr(arrival_ms, arrival_ms + n), done;
    for (int k = 0; k < n; ++k) jobs.emplace_back(rounds[k], k == fail_job);
    double clock = 0.0;
    try {
        chaos::serve_stream(
            eng, jobs, arr, done, [&] { return clock += 1.0; },
            [&](double t) {
                if (clock < t) clock = t;
            },
            0.0, [](size_t i, const std::string& err) { throw std::logic_error(std::to_string(i) + ": " + err); });
    } catch (const std::logic_error&) {
        return 1;
    } catch (const std::runtime_error&) {
        return 2;
    }
    for (int k = 0; k < n; ++k) done_ms[k] = done[k];
    *released = eng.released;
    return 0;
}

// The loop as it was before the fix (an idle engine: wait for the next
// arrival, count nothing), over the same stand-in: returns 2 when it does not
// end -- the stand-in reproduces the interleaving that left a stream's last
// job uncounted.
extern "C" int serve_stand_in_without_idle_count(int n, const double* arrival_ms, const int* rounds)
{
    StandIn eng;
    eng.call_limit = 100000;
    std::vector<long long> ids(n, -1);
    std::vector<char> done(n, 0);
    size_t next = 0, finished = 0;
    double clock = 0.0;
    try {
        while (finished < (size_t)n) {
            clock += 1.0;
            while (next < (size_t)n && arrival_ms[next] <= clock)
                ids[next] = eng.submit({rounds[next], false}), ++next;
            if (eng.idle()) {
                if (next < (size_t)n && clock < arrival_ms[next]) clock = arrival_ms[next];
                continue;
            }
            if (eng.step() > 0)
                for (size_t i = 0; i < next; ++i)
                    if (!done[i] && eng.job(ids[i]).status == 2) done[i] = 1, ++finished;
        }
    } catch (const std::runtime_error&) {
        return 2;
    }
    return 0;
}
