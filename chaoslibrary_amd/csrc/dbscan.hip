// dbscan.hip -- grid DBSCAN for the Basins labels and the deterministic
// z-score reduction.
//
// Replaces the reference's host-driven single linkage
// (Basins/src/basinsCUDA.cu:35-179, 866-923: one kernel launch and host sync
// per visited point, O(N) launches and O(N^2) work) with a fixed sequence of
// launches:
//   bbox -> hash cells of side eps/sqrt(dim) -> per-cell point lists
//   core test (a cell holding >= minPts points is all-core: its diagonal is
//   below eps; otherwise count neighbours in the 5^dim cell block, early exit)
//   union-find: cell members to the cell's smallest core index, then one
//   witness pair per neighbouring cell pair; links always go larger -> smaller
//   index, so every root is its component's smallest core index
//   border points take their smallest-index core neighbour's root
//   numbering by mark + exclusive scan over point indices.
#include <climits>
#include <cmath>
#include <vector>

#include "dbscan.hpp"
#include "scan.hpp"

namespace chaos {

constexpr unsigned long long kEmpty = ~0ull;

struct GridDev {
    const double* x[3];
    int dim;
    double lo[3];
    double s;                  // cell side
    double eps;
    double prune2;             // (eps * (1 + 1e-7))^2 for conservative box pruning
    unsigned long long hmask;  // hash capacity - 1
    unsigned long long* hkey;
    unsigned long long* hcount;
    long long* hstart;
    int* hcur;
    int* hrep;      // smallest core index in the cell (INT_MAX if none)
    int* pslot;     // slot of each point's cell
    int* cellpts;   // points grouped by cell
    const int* include;
    int* core;
    int* parent;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long pack_cell(int dim, const long long* c)
{
    if (dim == 3) return (unsigned long long)c[0] | ((unsigned long long)c[1] << 21) | ((unsigned long long)c[2] << 42);
    return (unsigned long long)c[0] | ((unsigned long long)c[1] << 31);
}

__device__ __forceinline__ void cell_coords(const GridDev& G, long long p, long long* c)
{
    for (int d = 0; d < G.dim; ++d) c[d] = (long long)floor((G.x[d][p] - G.lo[d]) / G.s);
}

__device__ __forceinline__ int lookup(const GridDev& G, unsigned long long key)
{
    unsigned long long h = mix64(key) & G.hmask;
    while (true) {
        const unsigned long long k = G.hkey[h];
        if (k == key) return (int)h;
        if (k == kEmpty) return -1;
        h = (h + 1) & G.hmask;
    }
}

__device__ __forceinline__ bool within(const GridDev& G, long long p, long long q)
{
    double d2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double df = G.x[d][p] - G.x[d][q];
        d2 = d2 + df * df;
    }
    return sqrt(d2) <= G.eps;
}

// Lower bound (squared) of the distance from point p to the box of cell c.
__device__ __forceinline__ double box_gap2(const GridDev& G, long long p, const long long* c)
{
    double g2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double lo = G.lo[d] + (double)c[d] * G.s;
        const double hi = lo + G.s;
        const double v = G.x[d][p];
        const double g = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        g2 += g * g;
    }
    return g2;
}

// Visit every cell of the 5^dim block around p's cell that may hold a point
// within eps: f(slot, coords).  Stops early when f returns false.
template <class F>
__device__ __forceinline__ void for_neighbour_cells(const GridDev& G, long long p, F&& f)
{
    long long c0[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    cell_coords(G, p, c0);
    const int r2 = G.dim == 3 ? 2 : 0;
    for (int dz = -r2; dz <= r2; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                c[0] = c0[0] + dx;
                c[1] = c0[1] + dy;
                c[2] = c0[2] + dz;
                bool ok = c[0] >= 0 && c[1] >= 0 && (G.dim == 2 || c[2] >= 0);
                if (!ok) continue;
                if (box_gap2(G, p, c) > G.prune2) continue;
                const int slot = lookup(G, pack_cell(G.dim, c));
                if (slot < 0) continue;
                if (!f(slot, c)) return;
            }
}

__device__ __forceinline__ int uf_load(const int* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ int uf_find(const GridDev& G, int x)
{
    while (true) {
        const int px = uf_load(&G.parent[x]);
        if (px == x) return x;
        x = px;
    }
}

__device__ __forceinline__ void uf_union(const GridDev& G, int a, int b)
{
    while (true) {
        a = uf_find(G, a);
        b = uf_find(G, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&G.parent[a], a, b) == a) return;
    }
}

__global__ void bbox_kernel(const GridDev G, long long n, double* out)
{
    __shared__ double smin[3][256], smax[3][256];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
        if (!G.include[p]) continue;
        for (int d = 0; d < G.dim; ++d) {
            mn[d] = fmin(mn[d], G.x[d][p]);
            mx[d] = fmax(mx[d], G.x[d][p]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        smin[d][threadIdx.x] = mn[d];
        smax[d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int d = 0; d < 3; ++d) {
                smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + o]);
                smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            out[blockIdx.x * 6 + d] = smin[d][0];
            out[blockIdx.x * 6 + 3 + d] = smax[d][0];
        }
}

__global__ void insert_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (!G.include[p]) {
        G.pslot[p] = -1;
        return;
    }
    long long c[3] = {0, 0, 0};
    cell_coords(G, p, c);
    const unsigned long long key = pack_cell(G.dim, c);
    unsigned long long h = mix64(key) & G.hmask;
    while (true) {
        const unsigned long long prev = atomicCAS(&G.hkey[h], kEmpty, key);
        if (prev == kEmpty || prev == key) break;
        h = (h + 1) & G.hmask;
    }
    G.pslot[p] = (int)h;
    atomicAdd(&G.hcount[h], 1ull);
}

__global__ void fill_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = G.pslot[p];
    if (s < 0) return;
    const long long pos = G.hstart[s] + atomicAdd(&G.hcur[s], 1);
    G.cellpts[pos] = (int)p;
}

__global__ void core_kernel(const GridDev G, long long n, int minPts)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = G.pslot[p];
    if (s < 0) {
        G.core[p] = 0;
        G.parent[p] = -1;
        return;
    }
    bool core = (long long)G.hcount[s] >= minPts;
    if (!core) {
        int cnt = 0;
        for_neighbour_cells(G, p, [&](int slot, const long long*) {
            const long long b = G.hstart[slot], e = b + (long long)G.hcount[slot];
            for (long long k = b; k < e; ++k)
                if (within(G, p, G.cellpts[k]) && ++cnt >= minPts) return false;
            return true;
        });
        core = cnt >= minPts;
    }
    G.core[p] = core ? 1 : 0;
    G.parent[p] = core ? (int)p : -1;
    if (core) atomicMin(&G.hrep[s], (int)p);
}

__global__ void link_cell_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !G.core[p]) return;
    G.parent[p] = G.hrep[G.pslot[p]];  // the cell's smallest core index (all-pairs within eps)
}

__global__ void link_cross_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !G.core[p]) return;
    const unsigned long long own = G.hkey[G.pslot[p]];
    for_neighbour_cells(G, p, [&](int slot, const long long*) {
        if (G.hkey[slot] <= own) return true;          // each cell pair from one side
        const int r2 = G.hrep[slot];
        if (r2 == INT_MAX) return true;                // no core point there
        if (uf_find(G, (int)p) == uf_find(G, r2)) return true;
        const long long b = G.hstart[slot], e = b + (long long)G.hcount[slot];
        for (long long k = b; k < e; ++k) {
            const int q = G.cellpts[k];
            if (G.core[q] && within(G, p, q)) {
                uf_union(G, (int)p, q);
                break;
            }
        }
        return true;
    });
}

__global__ void flatten_kernel(const GridDev G, long long n, int* root)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    root[p] = G.core[p] ? uf_find(G, (int)p) : -1;
}

__global__ void border_kernel(const GridDev G, long long n, const int* root, int* labroot)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (G.pslot[p] < 0) {
        labroot[p] = -1;
        return;
    }
    if (G.core[p]) {
        labroot[p] = root[p];
        return;
    }
    int best = INT_MAX;
    for_neighbour_cells(G, p, [&](int slot, const long long*) {
        const long long b = G.hstart[slot], e = b + (long long)G.hcount[slot];
        for (long long k = b; k < e; ++k) {
            const int q = G.cellpts[k];
            if (q < best && G.core[q] && within(G, p, q)) best = q;
        }
        return true;
    });
    labroot[p] = best == INT_MAX ? -1 : root[best];
}

__global__ void mark_roots_kernel(const int* root, long long n, const int* include, bool sign_rule, int* minfix,
                                  long long* markP, long long* markN)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    markP[p] = 0;
    if (markN) markN[p] = 0;
    if (sign_rule && root[p] >= 0 && include[p] == -1) atomicMin(&minfix[root[p]], (int)p);
}

__global__ void mark_roots2_kernel(const int* root, long long n, bool sign_rule, const int* minfix, long long* markP,
                                   long long* markN)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (root[p] != (int)p) return;
    if (sign_rule && minfix[p] != INT_MAX) markN[minfix[p]] = 1;
    else markP[p] = 1;
}

__global__ void label_kernel(const int* labroot, long long n, bool sign_rule, const int* minfix,
                             const long long* scanP, const long long* scanN, int* labels)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = labroot[p];
    if (r < 0) {
        labels[p] = 0;
        return;
    }
    if (sign_rule && minfix[r] != INT_MAX) labels[p] = -(int)(scanN[minfix[r]] + 1);
    else labels[p] = (int)(scanP[r] + 1);
}

__global__ void fill_int_kernel(int* a, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* a, long long n, unsigned long long v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

int dbscan_device(CallCtx& ctx, const double* const* pts, const int* include, long long n, const DbscanParams& prm,
                  int* labels)
{
    CHAOS_REQUIRE(prm.dim == 2 || prm.dim == 3, "dbscan dim must be 2 or 3");
    CHAOS_REQUIRE(prm.eps > 0 && std::isfinite(prm.eps), "eps must be positive");
    CHAOS_REQUIRE(prm.minPts >= 1, "minPts must be >= 1");
    hipStream_t st = ctx.stream();
    const unsigned T = 256;
    const unsigned gn = ceil_div(n, T);
    GridDev G{};
    for (int d = 0; d < 3; ++d) G.x[d] = d < prm.dim ? pts[d] : pts[0];
    G.dim = prm.dim;
    G.eps = prm.eps;
    G.include = include;

    // bounding box of the included points
    const unsigned nb = std::min<unsigned>(gn, 2048);
    DevBuf<double> bb((size_t)nb * 6);
    hipLaunchKernelGGL(bbox_kernel, dim3(nb), dim3(256), 0, st, G, n, bb.get());
    CHAOS_HIP(hipGetLastError());
    std::vector<double> hb((size_t)nb * 6);
    CHAOS_HIP(hipMemcpyAsync(hb.data(), bb.get(), hb.size() * 8, hipMemcpyDeviceToHost, st));
    CHAOS_HIP(hipStreamSynchronize(st));
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (unsigned b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], hb[b * 6 + d]);
            hi[d] = std::fmax(hi[d], hb[b * 6 + 3 + d]);
        }
    if (!(lo[0] <= hi[0])) {  // nothing included
        hipLaunchKernelGGL(fill_int_kernel, dim3(gn), dim3(T), 0, st, labels, n, 0);
        CHAOS_HIP(hipGetLastError());
        return 0;
    }
    // cell side eps/sqrt(dim), shrunk so that rounding can never put two
    // points of one cell farther apart than eps
    G.s = prm.eps / std::sqrt((double)prm.dim) * (1.0 - 1e-7);
    G.prune2 = (prm.eps * (1 + 1e-7)) * (prm.eps * (1 + 1e-7));
    const double cell_lim = prm.dim == 3 ? (double)(1 << 21) - 4 : 2147483000.0;
    for (int d = 0; d < prm.dim; ++d) {
        CHAOS_REQUIRE(std::isfinite(lo[d]) && std::isfinite(hi[d]), "non-finite feature values");
        G.lo[d] = lo[d];
        const double cells = std::floor((hi[d] - lo[d]) / G.s) + 1;
        if (cells >= cell_lim)
            fail(CHAOS_E_UNSUPPORTED, "feature range %.3g too wide for eps %.3g in dim %d", hi[d] - lo[d], prm.eps, d);
    }
    unsigned long long H = 1024;
    while (H < 2ull * (unsigned long long)n) H <<= 1;
    G.hmask = H - 1;
    DevBuf<unsigned long long> hkey(H), hcount(H);
    DevBuf<long long> hstart(H);
    DevBuf<int> hcur(H), hrep(H), pslot(n), cellpts(n), core(n), parent(n), root(n), labroot(n);
    G.hkey = hkey.get();
    G.hcount = hcount.get();
    G.hstart = hstart.get();
    G.hcur = hcur.get();
    G.hrep = hrep.get();
    G.pslot = pslot.get();
    G.cellpts = cellpts.get();
    G.core = core.get();
    G.parent = parent.get();
    const unsigned gh = ceil_div(H, T);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(gh), dim3(T), 0, st, G.hkey, (long long)H, kEmpty);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(gh), dim3(T), 0, st, G.hcount, (long long)H, 0ull);
    hipLaunchKernelGGL(fill_int_kernel, dim3(gh), dim3(T), 0, st, G.hcur, (long long)H, 0);
    hipLaunchKernelGGL(fill_int_kernel, dim3(gh), dim3(T), 0, st, G.hrep, (long long)H, INT_MAX);
    hipLaunchKernelGGL(insert_kernel, dim3(gn), dim3(T), 0, st, G, n);
    CHAOS_HIP(hipGetLastError());
    exclusive_scan(st, reinterpret_cast<const long long*>(G.hcount), G.hstart, (long long)H);
    hipLaunchKernelGGL(fill_kernel, dim3(gn), dim3(T), 0, st, G, n);
    hipLaunchKernelGGL(core_kernel, dim3(gn), dim3(T), 0, st, G, n, prm.minPts);
    hipLaunchKernelGGL(link_cell_kernel, dim3(gn), dim3(T), 0, st, G, n);
    hipLaunchKernelGGL(link_cross_kernel, dim3(gn), dim3(T), 0, st, G, n);
    hipLaunchKernelGGL(flatten_kernel, dim3(gn), dim3(T), 0, st, G, n, root.get());
    hipLaunchKernelGGL(border_kernel, dim3(gn), dim3(T), 0, st, G, n, root.get(), labroot.get());
    CHAOS_HIP(hipGetLastError());

    DevBuf<int> minfix(n);
    DevBuf<long long> markP(n), scanP(n), markN, scanN;
    if (prm.sign_rule) {
        markN.alloc(n);
        scanN.alloc(n);
    }
    hipLaunchKernelGGL(fill_int_kernel, dim3(gn), dim3(T), 0, st, minfix.get(), n, INT_MAX);
    hipLaunchKernelGGL(mark_roots_kernel, dim3(gn), dim3(T), 0, st, root.get(), n, include, prm.sign_rule,
                       minfix.get(), markP.get(), markN.get());
    hipLaunchKernelGGL(mark_roots2_kernel, dim3(gn), dim3(T), 0, st, root.get(), n, prm.sign_rule, minfix.get(),
                       markP.get(), markN.get());
    CHAOS_HIP(hipGetLastError());
    const long long K = exclusive_scan(st, markP.get(), scanP.get(), n);
    if (prm.sign_rule) exclusive_scan(st, markN.get(), scanN.get(), n);
    hipLaunchKernelGGL(label_kernel, dim3(gn), dim3(T), 0, st, labroot.get(), n, prm.sign_rule, minfix.get(),
                       scanP.get(), scanN.get(), labels);
    CHAOS_HIP(hipGetLastError());
    CHAOS_HIP(hipStreamSynchronize(st));
    ctx.add_launch(16);
    return (int)K;
}

// ---------------------------------------------------------------------------
// deterministic reductions and z-score
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(512) void det_block_sum_kernel(const double* __restrict__ x, const int* __restrict__ mask,
                                                            long long n, bool sq, double mean,
                                                            double* __restrict__ out)
{
    __shared__ double s[512];
    const long long base = (long long)blockIdx.x * 1024;
    auto val = [&](long long i) -> double {
        if (i >= n || (mask && mask[i] == 0)) return 0.0;
        if (!sq) return x[i];
        const double d = x[i] - mean;
        return d * d;
    };
    const int t = threadIdx.x;
    s[t] = val(base + t) + val(base + t + 512);  // stride-512 level
    __syncthreads();
    for (int stride = 256; stride > 0; stride >>= 1) {
        if (t < stride) s[t] = s[t] + s[t + stride];
        __syncthreads();
    }
    if (t == 0) out[blockIdx.x] = s[0];
}

double det_sum_device(CallCtx& ctx, const double* x, const int* mask, long long n, bool sq, double mean)
{
    if (n <= 0) return 0.0;
    const long long nb = (n + 1023) / 1024;
    DevBuf<double> bs((size_t)nb);
    hipLaunchKernelGGL(det_block_sum_kernel, dim3((unsigned)nb), dim3(512), 0, ctx.stream(), x, mask, n, sq, mean,
                       bs.get());
    CHAOS_HIP(hipGetLastError());
    std::vector<double> h((size_t)nb);
    CHAOS_HIP(hipMemcpyAsync(h.data(), bs.get(), nb * 8, hipMemcpyDeviceToHost, ctx.stream()));
    CHAOS_HIP(hipStreamSynchronize(ctx.stream()));
    double tot = 0.0;
    for (long long b = 0; b < nb; ++b) tot = tot + h[b];  // block sums in order
    return tot;
}

__global__ void count_mask_kernel(const int* mask, long long n, unsigned long long* cnt)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool m = i < n && mask[i] != 0;
    const unsigned long long b = __ballot(m);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(cnt, (unsigned long long)__popcll(b));
}

__global__ void normalize_kernel(double* x, const int* mask, long long n, double mean, double sd)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || mask[i] == 0) return;
    x[i] = sd > 0 ? (x[i] - mean) / sd : 0.0;
}

void zscore_device(CallCtx& ctx, double* const* f, int k, const int* mask, long long n, double* moments)
{
    DevBuf<unsigned long long> c(1);
    CHAOS_HIP(hipMemsetAsync(c.get(), 0, 8, ctx.stream()));
    hipLaunchKernelGGL(count_mask_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), mask, n, c.get());
    unsigned long long N = 0;
    CHAOS_HIP(hipMemcpyAsync(&N, c.get(), 8, hipMemcpyDeviceToHost, ctx.stream()));
    CHAOS_HIP(hipStreamSynchronize(ctx.stream()));
    if (N == 0) return;
    for (int j = 0; j < k; ++j) {
        const double mean = det_sum_device(ctx, f[j], mask, n, false, 0.0) / (double)N;
        const double sd = std::sqrt(det_sum_device(ctx, f[j], mask, n, true, mean) / (double)N);
        if (moments) {
            moments[2 * j] = mean;
            moments[2 * j + 1] = sd;
        }
        hipLaunchKernelGGL(normalize_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), f[j], mask, n, mean,
                           sd);
        CHAOS_HIP(hipGetLastError());
    }
}

}  // namespace chaos
