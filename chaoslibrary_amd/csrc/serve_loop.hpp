// serve_loop.hpp -- the host loop that serves a stream of jobs through the
// job server, as a template over the engine, so that its bookkeeping runs
// and is tested on machines without a GPU (with a stand-in engine).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <string>
#include <vector>

namespace chaos {

// Serves jobs[k], arriving at arrival_ms[k] (ascending, ms after the start),
// through eng: each job is submitted once its arrival time has passed, the
// engine is stepped while it has work, and done_ms[k] receives the time
// (ms after the start) at which job k was seen done.  Returns once every job
// is done; a failed job goes to on_fail(k, message), which must not return.
//
// Eng provides: long long submit(const Job&); bool idle(); int step();
// int take_completed(); job(id) with .status (2 done, -1 failed) and .error;
// void release(long long id).  now_ms() is the clock relative to the start,
// sleep_until(t) waits until now_ms() >= t; progress lines go to stderr
// every progress_s seconds (0: none).
template <class Eng, class Job, class Now, class Sleep, class OnFail>
void serve_stream(Eng& eng, const std::vector<Job>& jobs, const std::vector<double>& arrival_ms,
                  std::vector<double>& done_ms, Now&& now_ms, Sleep&& sleep_until, double progress_s,
                  OnFail&& on_fail)
{
    const size_t N = jobs.size();
    done_ms.assign(N, 0.0);
    std::vector<long long> ids(N, -1);
    std::vector<char> done(N, 0);
    size_t next = 0, finished = 0;
    double next_report = progress_s > 0 ? progress_s * 1000.0 : INFINITY;
    while (finished < N) {
        const double now = now_ms();
        if (now >= next_report) {
            std::fprintf(stderr, "stream engine: %.0f s, %zu of %zu jobs arrived, %zu done\n", now / 1000.0, next, N,
                         finished);
            next_report += progress_s * 1000.0;
        }
        while (next < N && arrival_ms[next] <= now) ids[next] = eng.submit(jobs[next]), ++next;
        int k = 0;
        if (eng.idle()) {
            // Nothing to integrate.  A finishing thread may have completed
            // jobs since the last step: count them first -- skipping them
            // here left the last job of a stream uncounted and this loop
            // waiting for it for ever.
            k = eng.take_completed();
            if (k == 0) {
                if (next < N) sleep_until(arrival_ms[next]);
                continue;
            }
        } else {
            k = eng.step();
        }
        if (k > 0) {
            const double t = now_ms();
            for (size_t i = 0; i < next; ++i) {
                if (done[i]) continue;
                const auto& r = eng.job(ids[i]);
                if (r.status == -1) on_fail(i, std::string(r.error));
                if (r.status == 2) {
                    done_ms[i] = t;
                    done[i] = 1;
                    eng.release(ids[i]);
                    ++finished;
                }
            }
        }
    }
}

}  // namespace chaos
