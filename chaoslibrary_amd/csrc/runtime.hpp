// runtime.hpp -- host-side plumbing shared by every entry point: error
// reporting (status codes + thread-local message), device buffers, the
// per-call context (device, stream, timing) and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/chaoslibrary.h"

namespace chaos {

// Internal failure carrying a chaos_status; converted to a return code at
// the C-ABI boundary by guarded().
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...);
void hip_check(hipError_t e, const char* what, const char* file, int line);
void set_last_error(const std::string& msg);

#define CHAOS_HIP(expr) ::chaos::hip_check((expr), #expr, __FILE__, __LINE__)
#define CHAOS_REQUIRE(cond, ...)                                  \
    do {                                                          \
        if (!(cond)) ::chaos::fail(CHAOS_E_INVALID, __VA_ARGS__); \
    } while (0)

// Run body(); map exceptions to status codes and the thread-local message.
template <class F>
int guarded(const char* fn, F&& body) noexcept
{
    try {
        body();
        return CHAOS_OK;
    } catch (const Error& e) {
        set_last_error(std::string(fn) + ": " + e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error(std::string(fn) + ": host allocation failed");
        return CHAOS_E_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(std::string(fn) + ": " + e.what());
        return CHAOS_E_INVALID;
    } catch (...) {
        set_last_error(std::string(fn) + ": unknown failure");
        return CHAOS_E_INVALID;
    }
}

// Options with defaults filled in (opt == NULL -> process defaults).
chaos_options resolve_options(const chaos_options* opt);

// Caching device allocator on the current device: freed blocks are kept in
// size classes and handed out again instead of going back to hipFree (which
// synchronises the device) -- per-call and per-job scratch is allocated and
// freed thousands of times per second by the job server.  A block must only
// be freed once the work using it has completed (every caller synchronises
// its stream first).  On allocation failure the cache is emptied and the
// allocation retried.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
void dev_cache_trim();
// Free HBM on the current device plus the bytes parked in the cache.
size_t dev_mem_available();

// Device buffer owned by one call or job (cached hipMalloc on the current device).
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void alloc(size_t n)
    {
        release();
        if (n == 0) return;
        p_ = static_cast<T*>(dev_alloc(n * sizeof(T)));
        n_ = n;
    }
    void release()
    {
        if (p_) dev_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Per-call execution context: selects the device, picks the stream and
// records HIP events for the timing report.
class CallCtx {
public:
    // library_stream_ok false: with no caller stream the call runs on the
    // null stream instead of creating the library stream.  Timing events are
    // created at the first begin().
    explicit CallCtx(const chaos_options& o, bool library_stream_ok = true);
    ~CallCtx();
    CallCtx(const CallCtx&) = delete;
    CallCtx& operator=(const CallCtx&) = delete;

    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }
    const chaos_options& opt() const { return opt_; }
    const hipDeviceProp_t& props() const;

    // Timed region on the call stream: begin() ... end() adds device ms to
    // the named accumulator of chaos_stats.
    void begin();
    double end();            // returns ms of this region (synchronises)
    void sync();
    void add_kernel_ms(double ms) { kernel_ms_ += ms; }
    void add_analysis_ms(double ms) { analysis_ms_ += ms; }
    void add_steps(double s) { steps_ += s; }
    void add_launch(int n = 1) { launches_ += n; }
    void add_peaks(long long n) { peaks_ += n; }
    void finish();          // fill opt.stats

private:
    chaos_options opt_;
    int device_ = 0;
    int prev_device_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double kernel_ms_ = 0, analysis_ms_ = 0, steps_ = 0;
    long long peaks_ = 0;
    int launches_ = 0;
    double t0_ = 0;
};

double wall_ms();
hipStream_t library_stream(int device);
const hipDeviceProp_t& device_props(int device);

inline unsigned ceil_div(unsigned long long a, unsigned long long b) { return (unsigned)((a + b - 1) / b); }

// Horizon conversions: the reference truncates int(T/h)
// (Bifurcation/src/bifurcationCUDA.cu:60-62, LLE/src/LLECUDA.cu:223-226);
// CONFIG semantics rounds to nearest.
int horizon_steps(double T, double h, int semantics);
int record_samples(double tMax, double h, int preScaller, int semantics);

void check_system(int system, int nx, int na);

}  // namespace chaos
