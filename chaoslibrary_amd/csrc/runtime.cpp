// runtime.cpp -- error reporting, options, streams, timing, device helpers.
#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "systems.hpp"

namespace chaos {

namespace {
thread_local std::string g_last_error;
std::mutex g_mu;
chaos_options g_defaults;
bool g_defaults_init = false;
std::vector<hipStream_t> g_streams;
std::vector<hipDeviceProp_t> g_props;
std::vector<char> g_props_ok;

void init_defaults_locked()
{
    if (g_defaults_init) return;
    g_defaults.system = CHAOS_SYS_CHAMELEON;  // systems/systems.cuh:9
    g_defaults.semantics = CHAOS_SEM_REFERENCE;
    g_defaults.seed = 12345ULL;
    g_defaults.device = 0;
    g_defaults.device_buffers = 0;
    g_defaults.stream = nullptr;
    g_defaults.min_pts = 20;
    g_defaults.block = 0;
    g_defaults.stats = nullptr;
    g_defaults.output = CHAOS_OUT_CSV;
    g_defaults.reserved = 0;
    g_defaults_init = true;
}
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// caching device allocator
// ---------------------------------------------------------------------------

namespace {
struct DevCache {
    std::map<size_t, std::vector<void*>> free_blocks;  // size class -> blocks
    size_t cached_bytes = 0;
    size_t cap = 0;  // bytes kept: a quarter of the device's memory, at least kCacheCap
};
std::mutex g_alloc_mu;
std::map<int, DevCache> g_cache;                         // device -> cache
std::map<void*, std::pair<int, size_t>> g_live;          // block -> (device, class)
constexpr size_t kCacheCap = 16ull << 30;                // least bytes kept per device

size_t size_class(size_t bytes)
{
    if (bytes <= 512) return 512;
    if (bytes <= (64ull << 20)) {
        size_t c = 512;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t g = 64ull << 20;
    return (bytes + g - 1) / g * g;
}

void trim_locked(int device)
{
    DevCache& c = g_cache[device];
    for (auto& kv : c.free_blocks)
        for (void* p : kv.second) (void)hipFree(p);
    c.free_blocks.clear();
    c.cached_bytes = 0;
}
}  // namespace

void* dev_alloc(size_t bytes)
{
    int dev = 0;
    CHAOS_HIP(hipGetDevice(&dev));
    const size_t cls = size_class(bytes);
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    DevCache& c = g_cache[dev];
    // the block's own class, or above 64 MiB the smallest cached class at
    // most an eighth larger (buffers sized by data, e.g. the peak values of
    // one grid range after another, differ by more than 64 MiB)
    const size_t most = cls > (64ull << 20) ? cls + cls / 8 : cls;
    for (auto it = c.free_blocks.lower_bound(cls); it != c.free_blocks.end() && it->first <= most; ++it) {
        if (it->second.empty()) continue;
        void* p = it->second.back();
        it->second.pop_back();
        c.cached_bytes -= it->first;
        g_live[p] = {dev, it->first};
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim_locked(dev);
        e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(CHAOS_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", cls, hipGetErrorString(e));
    }
    g_live[p] = {dev, cls};
    return p;
}

void dev_free(void* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        (void)hipFree(p);
        return;
    }
    const int dev = it->second.first;
    const size_t cls = it->second.second;
    g_live.erase(it);
    DevCache& c = g_cache[dev];
    if (c.cap == 0) {
        // the quarter lets a repeated full-size map reuse its peak store
        // (tens of GB); a failed hipMalloc empties the cache and retries
        size_t free_b = 0, total_b = 0;
        c.cap = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? std::max(kCacheCap, total_b / 4) : kCacheCap;
    }
    if (c.cached_bytes + cls > c.cap) {
        // room by dropping cached blocks no larger than this one, largest
        // first (mapping a large block again is what costs: 1.2-1.8 s for
        // 70 GB); when that cannot make room, this block goes instead
        size_t droppable = 0;
        for (const auto& kv : c.free_blocks)
            if (kv.first <= cls) droppable += kv.first * kv.second.size();
        if (c.cached_bytes - droppable + cls > c.cap) {
            (void)hipFree(p);
            return;
        }
        for (auto it = c.free_blocks.upper_bound(cls); c.cached_bytes + cls > c.cap && it != c.free_blocks.begin();) {
            --it;
            while (!it->second.empty() && c.cached_bytes + cls > c.cap) {
                (void)hipFree(it->second.back());
                it->second.pop_back();
                c.cached_bytes -= it->first;
            }
        }
    }
    c.free_blocks[cls].push_back(p);
    c.cached_bytes += cls;
}

size_t dev_mem_available()
{
    int dev = 0;
    CHAOS_HIP(hipGetDevice(&dev));
    size_t free_b = 0, total_b = 0;
    CHAOS_HIP(hipMemGetInfo(&free_b, &total_b));
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    auto it = g_cache.find(dev);
    return free_b + (it == g_cache.end() ? 0 : it->second.cached_bytes);
}

void dev_cache_trim()
{
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    for (auto& kv : g_cache) trim_locked(kv.first);
}

[[noreturn]] void fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

void hip_check(hipError_t e, const char* what, const char* file, int line)
{
    if (e == hipSuccess) return;
    int code = (e == hipErrorOutOfMemory) ? CHAOS_E_NOMEM
               : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? CHAOS_E_NODEVICE
                                                                        : CHAOS_E_HIP;
    fail(code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

// Stats sink of this thread's calls made without options (the C++ API).
static thread_local chaos_stats* t_thread_stats = nullptr;

chaos_options resolve_options(const chaos_options* opt)
{
    std::lock_guard<std::mutex> lk(g_mu);
    init_defaults_locked();
    chaos_options o = opt ? *opt : g_defaults;
    if (!opt) o.stats = t_thread_stats;
    if (o.min_pts <= 0) o.min_pts = 20;
    return o;
}

double wall_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

hipStream_t library_stream(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_streams.size() <= device) g_streams.resize(device + 1, nullptr);
    if (!g_streams[device]) {
        hipStream_t s;
        CHAOS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
        g_streams[device] = s;
    }
    return g_streams[device];
}

const hipDeviceProp_t& device_props(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_props.size() <= device) {
        g_props.resize(device + 1);
        g_props_ok.resize(device + 1, 0);
    }
    if (!g_props_ok[device]) {
        CHAOS_HIP(hipGetDeviceProperties(&g_props[device], device));
        g_props_ok[device] = 1;
    }
    return g_props[device];
}

CallCtx::CallCtx(const chaos_options& o, bool library_stream_ok) : opt_(o), device_(o.device)
{
    t0_ = wall_ms();
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) fail(CHAOS_E_NODEVICE, "no HIP device available");
    if (device_ < 0 || device_ >= count) fail(CHAOS_E_INVALID, "device %d out of range (%d devices)", device_, count);
    CHAOS_HIP(hipGetDevice(&prev_device_));
    CHAOS_HIP(hipSetDevice(device_));
    stream_ = o.stream ? static_cast<hipStream_t>(o.stream) : (library_stream_ok ? library_stream(device_) : nullptr);
}

CallCtx::~CallCtx()
{
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    (void)hipSetDevice(prev_device_);
}

const hipDeviceProp_t& CallCtx::props() const { return device_props(device_); }

void CallCtx::begin()
{
    if (!ev0_) CHAOS_HIP(hipEventCreate(&ev0_));
    if (!ev1_) CHAOS_HIP(hipEventCreate(&ev1_));
    CHAOS_HIP(hipEventRecord(ev0_, stream_));
}

double CallCtx::end()
{
    CHAOS_HIP(hipEventRecord(ev1_, stream_));
    CHAOS_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CHAOS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    return ms;
}

void CallCtx::sync() { CHAOS_HIP(hipStreamSynchronize(stream_)); }

void CallCtx::finish()
{
    if (!opt_.stats) return;
    chaos_stats* s = opt_.stats;
    s->kernel_ms = kernel_ms_;
    s->analysis_ms = analysis_ms_;
    s->total_ms = wall_ms() - t0_;
    s->rk4_steps = steps_;
    s->peaks = peaks_;
    s->launches = launches_;
}

int horizon_steps(double T, double h, int semantics)
{
    const double q = T / h;
    if (semantics == CHAOS_SEM_REFERENCE) return (int)q;
    return (int)std::llround(q);
}

int record_samples(double tMax, double h, int preScaller, int semantics)
{
    const double q = tMax / h / preScaller;
    if (semantics == CHAOS_SEM_REFERENCE) return (int)q;
    return (int)std::llround(q);
}

void check_system(int system, int nx, int na)
{
    int snx = 0, sna = 0;
    if (chaos_system_dims(system, &snx, &sna) != CHAOS_OK) fail(CHAOS_E_INVALID, "unknown system id %d", system);
    if (nx != snx) fail(CHAOS_E_INVALID, "system %d has %d state variables, got amountOfInitialConditions=%d", system, snx, nx);
    if (na != sna) fail(CHAOS_E_INVALID, "system %d has %d parameters, got amountOfValues=%d", system, sna, na);
}

}  // namespace chaos

using namespace chaos;

extern "C" {

void chaos_options_init(chaos_options* opt)
{
    if (!opt) return;
    *opt = resolve_options(nullptr);
}

int chaos_api_version(void) { return CHAOS_API_VERSION; }

const char* chaos_last_error(void) { return g_last_error.c_str(); }

int chaos_system_dims(int system, int* nx, int* na)
{
    static const int NX[SYS_COUNT] = {System<0>::NX, System<1>::NX, System<2>::NX, System<3>::NX,
                                      System<4>::NX, System<5>::NX, System<6>::NX};
    static const int NA[SYS_COUNT] = {System<0>::NA, System<1>::NA, System<2>::NA, System<3>::NA,
                                      System<4>::NA, System<5>::NA, System<6>::NA};
    if (system < 0 || system >= SYS_COUNT) {
        set_last_error("chaos_system_dims: unknown system id");
        return CHAOS_E_INVALID;
    }
    if (nx) *nx = NX[system];
    if (na) *na = NA[system];
    return CHAOS_OK;
}

int chaos_set_default_options(const chaos_options* opt)
{
    if (!opt) return CHAOS_E_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    init_defaults_locked();
    g_defaults = *opt;
    g_defaults.stats = nullptr;
    return CHAOS_OK;
}

int chaos_set_thread_stats(chaos_stats* stats)
{
    t_thread_stats = stats;
    return CHAOS_OK;
}

int chaos_device_count(int* count)
{
    return guarded("chaos_device_count", [&] {
        CHAOS_REQUIRE(count, "count is NULL");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

int chaos_device_malloc(void** ptr, size_t bytes, int device)
{
    return guarded("chaos_device_malloc", [&] {
        CHAOS_REQUIRE(ptr, "ptr is NULL");
        int prev = 0;
        CHAOS_HIP(hipGetDevice(&prev));
        CHAOS_HIP(hipSetDevice(device));
        hipError_t e = hipMalloc(ptr, bytes);
        (void)hipSetDevice(prev);
        CHAOS_HIP(e);
    });
}

int chaos_device_free(void* ptr)
{
    return guarded("chaos_device_free", [&] { CHAOS_HIP(hipFree(ptr)); });
}

int chaos_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    return guarded("chaos_memcpy_h2d", [&] { CHAOS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

int chaos_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    return guarded("chaos_memcpy_d2h", [&] { CHAOS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

int chaos_device_synchronize(int device)
{
    return guarded("chaos_device_synchronize", [&] {
        int prev = 0;
        CHAOS_HIP(hipGetDevice(&prev));
        CHAOS_HIP(hipSetDevice(device));
        hipError_t e = hipDeviceSynchronize();
        (void)hipSetDevice(prev);
        CHAOS_HIP(e);
    });
}

}  // extern "C"
