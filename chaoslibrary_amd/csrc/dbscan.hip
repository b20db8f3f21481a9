// dbscan.hip -- grid DBSCAN for the Basins labels and the deterministic
// z-score reduction.
//
// Replaces the reference's host-driven single linkage
// (Basins/src/basinsCUDA.cu:35-179, 866-923: one kernel launch and host sync
// per visited point, O(N) launches and O(N^2) work) with a fixed sequence of
// launches:
//   bbox -> hash cells of side eps/sqrt(dim) -> per-cell point lists
//   per-cell exact bounding box of the member points: a neighbour cell whose
//   box lies wholly beyond eps of a point is skipped, one wholly within eps is
//   taken as a whole, only the rest are scanned point by point
//   core test: a cell holding >= minPts points is all-core (its diagonal is
//   below eps); every other point is tested by one wavefront that sums whole
//   neighbour cells and scans the undecided ones 64 points at a time
//   union-find: cell members to the cell's smallest core index, then one
//   witness pair per neighbouring cell pair, searched by one wavefront per
//   cell in 64 x 64 tiles, nearest neighbour rings first so that most far
//   pairs are already connected; links always go larger -> smaller index, so every
//   root is its component's smallest core index
//   border points take their smallest-index core neighbour's root
//   numbering by mark + exclusive scan over point indices.
// Crowded cells (a basin that converges to a fixed point puts millions of
// points into one cell) would serialise the per-point atomics on a few
// addresses: lanes of a wavefront that share a cell combine their update and
// one lane issues it.
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
    unsigned long long* bbox;  // [6][H]: per-cell min x,y,z then max x,y,z (order-preserving encoding)
    long long H;
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

// Order-preserving map double -> u64 (for atomicMin/atomicMax on doubles).
__device__ __forceinline__ unsigned long long ord_enc(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double ord_dec(unsigned long long e)
{
    const unsigned long long u = (e >> 63) ? (e & 0x7fffffffffffffffull) : ~e;
    return __longlong_as_double((long long)u);
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

// Relation of point p to the member points of cell `slot`, from the cell's
// exact bounding box: 0 = none within eps, 2 = all within eps, 1 = undecided.
// The gap and far distances are formed with the same operations, in the same
// order, as within(); rounding is monotone, so every member's computed
// distance lies between them and both verdicts agree with within() exactly.
__device__ __forceinline__ int box_relation(const GridDev& G, long long p, int slot)
{
    double g2 = 0.0, f2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double lo = ord_dec(G.bbox[d * G.H + slot]);
        const double hi = ord_dec(G.bbox[(3 + d) * G.H + slot]);
        const double v = G.x[d][p];
        const double g = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        const double f = fmax(v - lo, hi - v);
        g2 = g2 + g * g;
        f2 = f2 + f * f;
    }
    if (sqrt(g2) > G.eps) return 0;
    if (sqrt(f2) <= G.eps) return 2;
    return 1;
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

// Slot of `key` in the open-addressing table, inserting it if absent.  A
// non-empty slot never changes, so a plain load that sees a key is final; an
// empty one is claimed by compare-and-swap.
__device__ __forceinline__ int hash_insert(const GridDev& G, unsigned long long key)
{
    unsigned long long h = mix64(key) & G.hmask;
    while (true) {
        const unsigned long long k = G.hkey[h];
        if (k == key) return (int)h;
        if (k == kEmpty) {
            const unsigned long long prev = atomicCAS(&G.hkey[h], kEmpty, key);
            if (prev == kEmpty || prev == key) return (int)h;
        }
        h = (h + 1) & G.hmask;
    }
}

// Wavefront groups: up to kAggGroups rounds, each taking the lowest pending
// lane as leader and every pending lane with the leader's key as its group.
// Lanes left pending afterwards (many distinct keys) update on their own.
constexpr int kAggGroups = 4;

// Grid of the wavefront-per-item kernels (blocks of 256): 8192 waves, a full
// complement for 256 CUs; the waves stride over the device-side item count.
constexpr unsigned kWaveBlocks = 2048;

__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << __lane_id()) - 1ull; }

__global__ void insert_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < n && G.include[p];
    unsigned long long key = kEmpty;
    if (valid) {
        long long c[3] = {0, 0, 0};
        cell_coords(G, p, c);
        key = pack_cell(G.dim, c);
    }
    const int lane = __lane_id();
    bool pending = valid;
    int slot = -1;
    for (int it = 0; it < kAggGroups; ++it) {
        const unsigned long long act = __ballot(pending);
        if (!act) break;
        const int leader = __builtin_ctzll(act);
        const unsigned long long lk = __shfl(key, leader);
        const bool mine = pending && key == lk;
        const unsigned long long grp = __ballot(mine);
        int h = 0;
        if (lane == leader) {
            h = hash_insert(G, lk);
            atomicAdd(&G.hcount[h], (unsigned long long)__popcll(grp));
        }
        h = __shfl(h, leader);
        if (mine) {
            slot = h;
            pending = false;
        }
    }
    if (pending) {
        slot = hash_insert(G, key);
        atomicAdd(&G.hcount[slot], 1ull);
    }
    if (p < n) G.pslot[p] = valid ? slot : -1;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// Places every included point in its cell's segment of cellpts and grows the
// cell's bounding box.
__global__ void fill_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = p < n ? G.pslot[p] : -1;
    const bool valid = s >= 0;
    unsigned long long e[3] = {0, 0, 0};
    if (valid)
        for (int d = 0; d < G.dim; ++d) e[d] = ord_enc(G.x[d][p]);
    const int lane = __lane_id();
    bool pending = valid;
    for (int it = 0; it < kAggGroups; ++it) {
        const unsigned long long act = __ballot(pending);
        if (!act) break;
        const int leader = __builtin_ctzll(act);
        const int ls = __shfl(s, leader);
        const bool mine = pending && s == ls;
        const unsigned long long grp = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&G.hcur[ls], __popcll(grp));
        base = __shfl(base, leader);
        if (mine) G.cellpts[G.hstart[ls] + base + __popcll(grp & lanes_below())] = (int)p;
        for (int d = 0; d < G.dim; ++d) {
            const unsigned long long mn = wave_min_u64(mine ? e[d] : ~0ull);
            const unsigned long long mx = wave_max_u64(mine ? e[d] : 0ull);
            if (lane == leader) {
                atomicMin(&G.bbox[d * G.H + ls], mn);
                atomicMax(&G.bbox[(3 + d) * G.H + ls], mx);
            }
        }
        if (mine) pending = false;
    }
    if (pending) {
        G.cellpts[G.hstart[s] + atomicAdd(&G.hcur[s], 1)] = (int)p;
        for (int d = 0; d < G.dim; ++d) {
            atomicMin(&G.bbox[d * G.H + s], e[d]);
            atomicMax(&G.bbox[(3 + d) * G.H + s], e[d]);
        }
    }
}

// Core test, dense part: a cell holding >= minPts points is all-core (its
// diagonal is below eps).  Points of sparser cells go to a candidate list for
// the wavefront-per-point test.
__global__ void core_dense_kernel(const GridDev G, long long n, int minPts, int* cand, int* ncand)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = p < n ? G.pslot[p] : -1;
    const bool dense = s >= 0 && (long long)G.hcount[s] >= minPts;
    const bool sparse = s >= 0 && !dense;
    if (p < n) {
        G.core[p] = dense ? 1 : 0;
        G.parent[p] = dense ? (int)p : -1;
    }
    const int lane = __lane_id();
    const unsigned long long sb = __ballot(sparse);
    int base = 0;
    if (lane == 0 && sb) base = atomicAdd(ncand, __popcll(sb));
    base = __shfl(base, 0);
    if (sparse) cand[base + __popcll(sb & lanes_below())] = (int)p;
    // hrep = smallest core index of the cell.  Indices rise with the lane, so
    // a group's leader holds the group's smallest index.
    bool pending = dense;
    for (int it = 0; it < kAggGroups; ++it) {
        const unsigned long long act = __ballot(pending);
        if (!act) break;
        const int leader = __builtin_ctzll(act);
        const int ls = __shfl(s, leader);
        const bool mine = pending && s == ls;
        if (lane == leader) atomicMin(&G.hrep[ls], (int)p);
        if (mine) pending = false;
    }
    if (pending) atomicMin(&G.hrep[s], (int)p);
}

__device__ __forceinline__ long long wave_sum_ll(long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Neighbour cell o (0 <= o < 5^dim) of cell c0: coordinates, false if any is
// negative (no such cell).
__device__ __forceinline__ bool neighbour_cell(const GridDev& G, const long long* c0, int o, long long* c)
{
    c[0] = c0[0] + o % 5 - 2;
    c[1] = c0[1] + (o / 5) % 5 - 2;
    c[2] = G.dim == 3 ? c0[2] + o / 25 - 2 : 0;
    return c[0] >= 0 && c[1] >= 0 && c[2] >= 0;
}

__device__ __forceinline__ void unpack_cell(int dim, unsigned long long key, long long* c)
{
    if (dim == 3) {
        c[0] = (long long)(key & 0x1FFFFFull);
        c[1] = (long long)((key >> 21) & 0x1FFFFFull);
        c[2] = (long long)(key >> 42);
    } else {
        c[0] = (long long)(key & 0x7FFFFFFFull);
        c[1] = (long long)(key >> 31);
        c[2] = 0;
    }
}

// Core test, sparse part: one wavefront per candidate point.  The lanes look
// up the 5^dim neighbour cells, add whole cells whose box lies within eps and
// scan the undecided ones 64 points at a time, stopping at minPts.
__global__ void core_sparse_kernel(const GridDev G, const int* cand, const int* ncand, int minPts)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int nc = *ncand;
    const int noff = G.dim == 3 ? 125 : 25;
    for (long long i = wave; i < nc; i += nw) {
        const int p = cand[i];
        long long c0[3] = {0, 0, 0};
        cell_coords(G, p, c0);
        int sl[2] = {-1, -1}, rl[2] = {0, 0};
        for (int j = 0; j < 2; ++j) {
            const int o = lane + 64 * j;
            long long c[3];
            if (o >= noff || !neighbour_cell(G, c0, o, c)) continue;
            if (box_gap2(G, p, c) > G.prune2) continue;
            const int slot = lookup(G, pack_cell(G.dim, c));
            if (slot < 0) continue;
            sl[j] = slot;
            rl[j] = box_relation(G, p, slot);
        }
        long long whole = 0;
        for (int j = 0; j < 2; ++j)
            if (rl[j] == 2) whole += (long long)G.hcount[sl[j]];
        long long cnt = wave_sum_ll(whole);
        for (int j = 0; j < 2 && cnt < minPts; ++j) {
            unsigned long long m = __ballot(rl[j] == 1);
            while (m && cnt < minPts) {
                const int ld = __builtin_ctzll(m);
                m &= m - 1;
                const int slot = __shfl(sl[j], ld);
                const long long b = G.hstart[slot], e = b + (long long)G.hcount[slot];
                for (long long k = b; k < e && cnt < minPts; k += 64) {
                    const bool hit = k + lane < e && within(G, p, G.cellpts[k + lane]);
                    cnt += __popcll(__ballot(hit));
                }
            }
        }
        if (cnt >= minPts && lane == 0) {
            G.core[p] = 1;
            G.parent[p] = p;
            atomicMin(&G.hrep[G.pslot[p]], p);
        }
    }
}

__global__ void link_cell_kernel(const GridDev G, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !G.core[p]) return;
    G.parent[p] = G.hrep[G.pslot[p]];  // the cell's smallest core index (all-pairs within eps)
}

// Cells holding at least one core point, compacted.
__global__ void cell_list_kernel(const GridDev G, int* cells, int* ncells)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool has = s < G.H && G.hkey[s] != kEmpty && G.hrep[s] != INT_MAX;
    const int lane = __lane_id();
    const unsigned long long m = __ballot(has);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(ncells, __popcll(m));
    base = __shfl(base, 0);
    if (has) cells[base + __popcll(m & lanes_below())] = (int)s;
}

// May some point of cell a lie within eps of some point of cell b?  (Exact
// boxes; same monotone-rounding argument as box_relation.)
__device__ __forceinline__ bool boxes_may_touch(const GridDev& G, int a, int b)
{
    double g2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double loA = ord_dec(G.bbox[d * G.H + a]), hiA = ord_dec(G.bbox[(3 + d) * G.H + a]);
        const double loB = ord_dec(G.bbox[d * G.H + b]), hiB = ord_dec(G.bbox[(3 + d) * G.H + b]);
        const double g = loB > hiA ? loB - hiA : (loA > hiB ? loA - hiB : 0.0);
        g2 = g2 + g * g;
    }
    return sqrt(g2) <= G.eps;
}

// One witness pair (core p in cell [a0, a1), core q in cell b, within eps)
// found by the whole wavefront -- 64 points of a against 64 of b per tile --
// and united.  Returns at the first witness.
__device__ void link_cell_pair(const GridDev& G, int a, long long a0, long long a1, int b)
{
    const int lane = __lane_id();
    const long long b0 = G.hstart[b], b1 = b0 + (long long)G.hcount[b];
    const int rb = G.hrep[b];
    for (long long ka = a0; ka < a1; ka += 64) {
        int p = -1, rel = 0;
        double px[3] = {0.0, 0.0, 0.0};
        if (ka + lane < a1) {
            p = G.cellpts[ka + lane];
            if (G.core[p]) rel = box_relation(G, p, b);
        }
        const unsigned long long m2 = __ballot(rel == 2);
        if (m2) {                                  // all of b within eps of p, rb among them
            if (lane == __builtin_ctzll(m2)) uf_union(G, p, rb);
            return;
        }
        const bool cand = rel == 1;
        if (!__ballot(cand)) continue;
        if (cand)
            for (int d = 0; d < G.dim; ++d) px[d] = G.x[d][p];
        for (long long kb = b0; kb < b1; kb += 64) {
            int q = -1;
            bool qok = false;
            double qx[3] = {0.0, 0.0, 0.0};
            if (kb + lane < b1) {
                q = G.cellpts[kb + lane];
                qok = G.core[q] != 0 && box_relation(G, q, a) != 0;
                if (qok)
                    for (int d = 0; d < G.dim; ++d) qx[d] = G.x[d][q];
            }
            unsigned long long mq = __ballot(qok);
            while (mq) {
                const int l = __builtin_ctzll(mq);
                mq &= mq - 1;
                double d2 = 0.0;
                for (int d = 0; d < G.dim; ++d) {
                    const double df = px[d] - __shfl(qx[d], l);
                    d2 = d2 + df * df;
                }
                const unsigned long long mh = __ballot(cand && sqrt(d2) <= G.eps);
                if (mh) {
                    const int qq = __shfl(q, l);
                    if (lane == __builtin_ctzll(mh)) uf_union(G, p, qq);
                    return;
                }
            }
        }
    }
}

// Ring of neighbour offset o: 0 = shares a face, 1 = rest of the 3^dim
// block, 2 = outer shell of the 5^dim block.
__device__ __forceinline__ int offset_ring(const GridDev& G, int o)
{
    const int dx = abs(o % 5 - 2), dy = abs((o / 5) % 5 - 2), dz = G.dim == 3 ? abs(o / 25 - 2) : 0;
    const int cheb = max(dx, max(dy, dz));
    if (cheb == 2) return 2;
    return dx + dy + dz == 1 ? 0 : 1;
}

// Cross-cell links: one wavefront per core-holding cell a, over its
// neighbour cells b of one ring with a larger key (each pair from one side),
// skipping pairs already connected or whose boxes are farther apart than eps.
// The rings run as three launches, nearest first: in a dense region the face
// links connect the cells, and the farther pairs then pass the root check.
__global__ void link_pairs_kernel(const GridDev G, const int* cells, const int* ncells, int ring)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int nc = *ncells;
    const int noff = G.dim == 3 ? 125 : 25;
    for (long long i = wave; i < nc; i += nw) {
        const int a = cells[i];
        const unsigned long long ka = G.hkey[a];
        const int ra = G.hrep[a];
        long long c0[3];
        unpack_cell(G.dim, ka, c0);
        int sl[2] = {-1, -1};
        for (int j = 0; j < 2; ++j) {
            const int o = lane + 64 * j;
            long long c[3];
            if (o >= noff || offset_ring(G, o) != ring || !neighbour_cell(G, c0, o, c)) continue;
            const unsigned long long kb = pack_cell(G.dim, c);
            if (kb <= ka) continue;
            const int slot = lookup(G, kb);
            if (slot < 0 || G.hrep[slot] == INT_MAX || !boxes_may_touch(G, a, slot)) continue;
            sl[j] = slot;
        }
        const long long a0 = G.hstart[a], a1 = a0 + (long long)G.hcount[a];
        for (int j = 0; j < 2; ++j) {
            unsigned long long m = __ballot(sl[j] >= 0);
            while (m) {
                const int ld = __builtin_ctzll(m);
                m &= m - 1;
                const int b = __shfl(sl[j], ld);
                int same = 0;
                if (lane == 0) same = uf_find(G, ra) == uf_find(G, G.hrep[b]);
                if (__shfl(same, 0)) continue;
                link_cell_pair(G, a, a0, a1, b);
            }
        }
    }
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
        const int rel = box_relation(G, p, slot);
        if (rel == 0) return true;
        if (rel == 2) {                                // smallest core index of the cell
            best = min(best, G.hrep[slot]);
            return true;
        }
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
    DevBuf<unsigned long long> bbox(6 * H);
    G.bbox = bbox.get();
    G.H = (long long)H;
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
    hipLaunchKernelGGL(fill_u64_kernel, dim3(3 * gh), dim3(T), 0, st, G.bbox, 3 * (long long)H, ~0ull);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(3 * gh), dim3(T), 0, st, G.bbox + 3 * H, 3 * (long long)H, 0ull);
    hipLaunchKernelGGL(insert_kernel, dim3(gn), dim3(T), 0, st, G, n);
    CHAOS_HIP(hipGetLastError());
    exclusive_scan(st, reinterpret_cast<const long long*>(G.hcount), G.hstart, (long long)H);
    hipLaunchKernelGGL(fill_kernel, dim3(gn), dim3(T), 0, st, G, n);
    DevBuf<int> cand(n), cells(n), counters(2);  // counters: candidates, cells
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, counters.get(), 2LL, 0);
    hipLaunchKernelGGL(core_dense_kernel, dim3(gn), dim3(T), 0, st, G, n, prm.minPts, cand.get(), counters.get());
    hipLaunchKernelGGL(core_sparse_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G, cand.get(), counters.get(),
                       prm.minPts);
    hipLaunchKernelGGL(link_cell_kernel, dim3(gn), dim3(T), 0, st, G, n);
    hipLaunchKernelGGL(cell_list_kernel, dim3(gh), dim3(T), 0, st, G, cells.get(), counters.get() + 1);
    for (int ring = 0; ring < 3; ++ring)
        hipLaunchKernelGGL(link_pairs_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G, cells.get(), counters.get() + 1,
                           ring);
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
    ctx.add_launch(21);
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

// y = z-scored x on the masked points, x elsewhere (y may be x).
__global__ void normalize_kernel(const double* x, double* y, const int* mask, long long n, double mean, double sd)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    y[i] = mask[i] == 0 ? v : (sd > 0 ? (v - mean) / sd : 0.0);
}

void zscore_device(CallCtx& ctx, const double* const* src, double* const* dst, int k, const int* mask, long long n,
                   double* moments)
{
    // the block sums of all k features come back in one copy per moment
    // (two round trips in all), summed on the host in block order
    const long long nb = (n + 1023) / 1024;
    DevBuf<unsigned long long> c(1);
    DevBuf<double> bs((size_t)(k * nb));
    std::vector<double> h((size_t)(k * nb));
    unsigned long long N = 0;
    CHAOS_HIP(hipMemsetAsync(c.get(), 0, 8, ctx.stream()));
    hipLaunchKernelGGL(count_mask_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), mask, n, c.get());
    std::vector<double> mean(k), sd(k);
    for (int pass = 0; pass < 2; ++pass) {
        for (int j = 0; j < k; ++j)
            hipLaunchKernelGGL(det_block_sum_kernel, dim3((unsigned)nb), dim3(512), 0, ctx.stream(), src[j], mask, n,
                               pass == 1, pass == 1 ? mean[j] : 0.0, bs.get() + j * nb);
        CHAOS_HIP(hipGetLastError());
        if (pass == 0) CHAOS_HIP(hipMemcpyAsync(&N, c.get(), 8, hipMemcpyDeviceToHost, ctx.stream()));
        CHAOS_HIP(hipMemcpyAsync(h.data(), bs.get(), k * nb * 8, hipMemcpyDeviceToHost, ctx.stream()));
        CHAOS_HIP(hipStreamSynchronize(ctx.stream()));
        if (N == 0) break;
        for (int j = 0; j < k; ++j) {
            double tot = 0.0;
            for (long long b = 0; b < nb; ++b) tot = tot + h[j * nb + b];
            if (pass == 0) mean[j] = tot / (double)N;
            else sd[j] = std::sqrt(tot / (double)N);
        }
    }
    for (int j = 0; j < k; ++j) {
        if (N > 0 && moments) {
            moments[2 * j] = mean[j];
            moments[2 * j + 1] = sd[j];
        }
        if (N > 0)
            hipLaunchKernelGGL(normalize_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx.stream(), src[j], dst[j],
                               mask, n, mean[j], sd[j]);
        else if (dst[j] != src[j])
            CHAOS_HIP(hipMemcpyAsync(dst[j], src[j], n * 8, hipMemcpyDeviceToDevice, ctx.stream()));
        CHAOS_HIP(hipGetLastError());
    }
}

}  // namespace chaos

using namespace chaos;

extern "C" int chaos_dbscan(const chaos_options* opt, const double* x, const double* y, const double* z, long long n,
                            int dim, double eps, int min_pts, int sign_rule, const int* include, int* labels,
                            int* clusters)
{
    return guarded("chaos_dbscan", [&] {
        chaos_options o = resolve_options(opt);
        CHAOS_REQUIRE(n >= 0 && n < INT_MAX, "n out of range");
        CHAOS_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
        CHAOS_REQUIRE(sign_rule == 0 || dim == 2, "sign_rule needs dim 2");
        CHAOS_REQUIRE(x && y && (dim == 2 || z), "coordinate array is NULL");
        CHAOS_REQUIRE(labels, "labels is NULL");
        CHAOS_REQUIRE(sign_rule || min_pts >= 1, "min_pts must be >= 1");
        if (clusters) *clusters = 0;
        if (n == 0) return;
        CallCtx ctx(o);
        hipStream_t st = ctx.stream();
        const bool dev = o.device_buffers != 0;
        DevBuf<double> d_pts;
        DevBuf<int> d_inc, d_lab;
        const double* pts[3] = {x, y, z};
        const int* inc = include;
        int* lab = labels;
        if (!dev) {
            d_pts.alloc((size_t)dim * n);
            for (int d = 0; d < dim; ++d) {
                CHAOS_HIP(hipMemcpyAsync(d_pts.get() + d * n, pts[d], n * 8, hipMemcpyHostToDevice, st));
                pts[d] = d_pts.get() + d * n;
            }
            d_lab.alloc(n);
            lab = d_lab.get();
        }
        if (!inc || !dev) {
            d_inc.alloc(n);
            if (inc) CHAOS_HIP(hipMemcpyAsync(d_inc.get(), inc, n * sizeof(int), hipMemcpyHostToDevice, st));
            else hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, d_inc.get(), n, 1);
            inc = d_inc.get();
        }
        ctx.begin();
        DbscanParams prm{dim, eps, sign_rule ? 1 : min_pts, sign_rule != 0};
        const int K = dbscan_device(ctx, pts, inc, n, prm, lab);
        ctx.add_analysis_ms(ctx.end());
        if (!dev) CHAOS_HIP(hipMemcpyAsync(labels, lab, n * sizeof(int), hipMemcpyDeviceToHost, st));
        ctx.sync();
        if (clusters) *clusters = K;
        ctx.finish();
    });
}
