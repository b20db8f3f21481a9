// dbscan_noalloc.hip -- chaos_dbscan without any allocation: the clustering
// scratch is a static device array of the library, for callers that cannot
// allocate (code under hipGraph capture, tight serving loops).
//
// The array holds the workspace of CHAOS_DBSCAN_NOALLOC_MAX_POINTS points
// with a cell table of 2^18 slots (dbscan_ws_bytes_c; about 48 MB).  Every
// GPU the library runs on holds its own copy until the process ends, and
// every call on one GPU uses that GPU's copy: calls that overlap on one GPU
// race.  The kernels initialise every part of it they read, so a call does
// not depend on what an earlier call left behind.
#include <climits>
#include <mutex>

#include "../../include/chaoslibrary.h"
#include "dbscan.hpp"

namespace chaos {

constexpr long long kNoAllocPoints = CHAOS_DBSCAN_NOALLOC_MAX_POINTS;
constexpr long long kNoAllocCells = 1LL << 18;
constexpr size_t kNoAllocBytes = dbscan_ws_bytes_c(kNoAllocPoints, kNoAllocCells, true);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};  // device address of `scratch`, per device ordinal
}  // namespace

// Device address of the current device's copy of `scratch`, looked up once
// per device (with that device current).  In host code `scratch` itself is
// only a host-side stand-in: kernels and copies get this address.
static void* scratch_address(int device)
{
    CHAOS_REQUIRE(device >= 0 && device < 64, "device ordinal %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        CHAOS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CHAOS_REQUIRE(p != nullptr, "the scratch array has no device address");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_dbscan_noalloc(const chaos_options* opt, const double* x, const double* y, const double* z,
                                    long long n, int dim, double eps, int min_pts, int sign_rule, const int* include,
                                    int* labels, int* clusters)
{
    return guarded("chaos_dbscan_noalloc", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(n >= 0, "n out of range");
        if (n > kNoAllocPoints)
            fail(CHAOS_E_CAPACITY, "%lld points exceed the allocation-free limit of %lld", n, kNoAllocPoints);
        CHAOS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        CHAOS_REQUIRE(sign_rule == 0 || dim == 2, "sign_rule needs dim 2");
        CHAOS_REQUIRE(x && y && (dim == 2 || z), "coordinate array is NULL");
        CHAOS_REQUIRE(labels, "labels is NULL");
        CHAOS_REQUIRE(sign_rule || min_pts >= 1, "min_pts must be >= 1");
        if (clusters) *clusters = 0;
        if (n == 0) return;
        // the caller's stream, or the null stream: no library stream is created
        CallCtx ctx(o, /*library_stream_ok=*/false);
        void* ws = scratch_address(ctx.device());
        const double* pts[3] = {x, y, z};
        DbscanParams prm{dim, eps, sign_rule ? 1 : min_pts, sign_rule != 0};
        const int K = dbscan_device(ctx, pts, include, n, prm, labels, ws, kNoAllocBytes);
        if (clusters) *clusters = K;
    });
}
