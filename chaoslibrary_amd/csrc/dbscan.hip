// dbscan.hip -- grid DBSCAN for the Basins labels and the deterministic
// z-score that feeds it.
//
// Replaces the reference's host-driven single linkage
// (Basins/src/basinsCUDA.cu:35-179, 866-923: one kernel launch and host sync
// per visited point, O(N) launches and O(N^2) work) with a fixed sequence of
// launches and one synchronisation at the end:
//   bounding box -> hash cells of side eps/sqrt(dim): per-cell count, smallest
//   member index and exact bounding box of the members, aggregated per block
//   in LDS first (a basin that converges to a fixed point puts millions of
//   points into one cell; per-point atomics on it would serialise)
//   per-cell point lists (block-reserved ranges of each cell's segment)
//   core test: a cell holding >= minPts points is all-core (its diagonal is
//   below eps); every other point is tested by one wavefront that sums whole
//   neighbour cells (exact boxes within eps) and scans the undecided ones 64
//   points at a time
//   union-find: cell members to the cell's smallest core index, then one
//   witness pair per neighbouring cell pair, searched by one wavefront per
//   cell in 64 x 64 tiles, nearest neighbour rings first; links always go
//   larger -> smaller index, so every root is its component's smallest core
//   index
//   border points take their smallest-index core neighbour's root
//   numbering: one bit per root in 64-point words, popcount prefix sums.
// All scratch lives in one workspace (dbscan_ws_bytes), so the same kernels
// serve the allocating entry point and chaos_dbscan_noalloc.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "dbscan.hpp"
#include "moments.hpp"

namespace chaos {

constexpr unsigned long long kEmpty = ~0ull;

enum DbFlags : int {
    DB_FULL = 1,       // cell table more than half full (redo with the larger table)
    DB_RANGE = 2,      // feature range too wide for eps (cell coordinates overflow the key)
    DB_NONFINITE = 4   // non-finite coordinate among the included points
};

// Device control block at the start of the workspace.
struct DbCtl {
    unsigned long long box[6];  // encoded min x, y, z, max x, y, z of the included points
    double lo[3];               // decoded minimum: the grid origin
    int ncells;                 // distinct cells claimed
    int ncand;                  // points of sparse cells (core candidates)
    int ncore_cells;            // cells holding a core point
    int flags;                  // DbFlags
    int K;                      // positive clusters
    int any;                    // some point is included
    int npieces;                // pieces of split core tests claimed
};

struct GridDev {
    const double* x[3];
    const double* zmom;  // device [mean_0, sd_0, mean_1, sd_1, ...] to z-score x on access, or NULL
    const int* include;  // NULL: every point included
    int dim;
    int sign;            // fixed-point sign rule
    double s;            // cell side
    double eps;
    double prune2;       // (eps * (1 + 1e-7))^2 for conservative box pruning
    long long n, nw;     // points, 64-point words
    unsigned long long hmask;
    long long H;
    // cell tables [H]
    unsigned long long* hkey;
    unsigned* hcount;
    int* hmin;     // smallest member index
    int* hstart;   // segment of the cell in cellpts
    int* hcur;     // fill cursor of the segment
    int* hrep;     // smallest core index (INT_MAX if none)
    unsigned long long* bbox;  // [6][H]: member min x, y, z, max x, y, z (order-preserving encoding)
    int* cells;    // compacted cells holding a core point
    // point tables [n]
    int* pslot;    // cell slot, -1 when excluded
    int* cellpts;  // points grouped by cell
    int* parent;   // union-find parent of core points, -1 for non-core points
    int* labroot;  // component root of the point's label, -1 for none
    int* cand;     // points of sparse cells
    int* minfix;   // sign rule: smallest fixed-point core index per root
    int4* pieces;  // split core tests: {point, first member, end member, 0} ranges of cellpts
    long long piece_cap;
    // numbering [nw]
    unsigned long long* bitsP;  // roots of positive clusters
    unsigned long long* bitsN;  // smallest fixed point of each negative cluster
    int* rankP;
    int* rankN;
    int* scan_tmp;
    DbCtl* ctl;
};

// Coordinate d of point p: the stored value, or its z-score when the points
// are raw features with their moments in zmom -- the same operations as the
// z-score of the host restatement, so the same bits as a normalised copy.
__device__ __forceinline__ double coord(const GridDev& G, int d, long long p)
{
    const double v = G.x[d][p];
    if (!G.zmom) return v;
    const double sd = G.zmom[2 * d + 1];
    return sd > 0 ? (v - G.zmom[2 * d]) / sd : 0.0;
}

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------

static_assert(sizeof(DbCtl) <= 256, "control block fits its 256-byte slot");

static void carve(GridDev& G, char* base, long long n, long long H, bool sign)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += align256(bytes);
        return p;
    };
    const long long nw = (n + 63) / 64;
    G.ctl = (DbCtl*)take(256);
    G.hkey = (unsigned long long*)take(8 * (size_t)H);
    G.hcount = (unsigned*)take(4 * (size_t)H);
    G.hmin = (int*)take(4 * (size_t)H);
    G.hstart = (int*)take(4 * (size_t)H);
    G.hcur = (int*)take(4 * (size_t)H);
    G.hrep = (int*)take(4 * (size_t)H);
    G.bbox = (unsigned long long*)take(48 * (size_t)H);
    G.cells = (int*)take(4 * (size_t)H);
    G.pslot = (int*)take(4 * (size_t)n);
    G.cellpts = (int*)take(4 * (size_t)n);
    G.parent = (int*)take(4 * (size_t)n);
    G.labroot = (int*)take(4 * (size_t)n);
    G.cand = (int*)take(4 * (size_t)n);
    G.minfix = sign ? (int*)take(4 * (size_t)n) : nullptr;
    G.piece_cap = core_piece_cap(n);
    G.pieces = (int4*)take(16 * (size_t)G.piece_cap);
    G.bitsP = (unsigned long long*)take(8 * (size_t)nw);
    G.rankP = (int*)take(4 * (size_t)nw);
    G.bitsN = sign ? (unsigned long long*)take(8 * (size_t)nw) : nullptr;
    G.rankN = sign ? (int*)take(4 * (size_t)nw) : nullptr;
    G.scan_tmp = (int*)take(4 * scan_tmp_ints(H > nw ? H : nw));
    G.H = H;
    G.hmask = (unsigned long long)H - 1;
    G.n = n;
    G.nw = nw;
}

size_t dbscan_ws_bytes(long long n, long long H, bool sign_rule) { return dbscan_ws_bytes_c(n, H, sign_rule); }

long long dbscan_max_cells(long long n)
{
    long long H = 1024;
    while (H < 2 * n) H <<= 1;
    return H;
}

long long dbscan_first_cells(long long n)
{
    const long long H = dbscan_max_cells(n);
    return H < (1LL << 18) ? H : (1LL << 18);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ bool included(const GridDev& G, long long p)
{
    return G.include == nullptr || G.include[p] != 0;
}

__device__ __forceinline__ unsigned long long pack_cell(int dim, const long long* c)
{
    if (dim == 3) return (unsigned long long)c[0] | ((unsigned long long)c[1] << 21) | ((unsigned long long)c[2] << 42);
    return (unsigned long long)c[0] | ((unsigned long long)c[1] << 31);
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

__device__ __forceinline__ void cell_coords(const GridDev& G, long long p, long long* c)
{
    for (int d = 0; d < G.dim; ++d) c[d] = (long long)floor((coord(G, d, p) - G.ctl->lo[d]) / G.s);
}


__device__ __forceinline__ int lookup(const GridDev& G, unsigned long long key)
{
    unsigned long long h = mix64(key) & G.hmask;
    for (long long probe = 0; probe < G.H; ++probe) {
        const unsigned long long k = G.hkey[h];
        if (k == key) return (int)h;
        if (k == kEmpty) return -1;
        h = (h + 1) & G.hmask;
    }
    return -1;
}

// Slot of `key`, inserting it if absent.  A non-empty slot never changes, so
// a plain load that sees a key is final; an empty one is claimed by
// compare-and-swap.  A table more than half full is flagged (the host redoes
// the clustering with the larger table); a full one returns -1.
__device__ __forceinline__ int hash_insert(const GridDev& G, unsigned long long key)
{
    unsigned long long h = mix64(key) & G.hmask;
    for (long long probe = 0; probe < G.H; ++probe) {
        const unsigned long long k = G.hkey[h];
        if (k == key) return (int)h;
        if (k == kEmpty) {
            const unsigned long long prev = atomicCAS(&G.hkey[h], kEmpty, key);
            if (prev == kEmpty) {
                if ((long long)atomicAdd(&G.ctl->ncells, 1) + 1 > (G.H >> 1)) atomicOr(&G.ctl->flags, DB_FULL);
                return (int)h;
            }
            if (prev == key) return (int)h;
        }
        h = (h + 1) & G.hmask;
    }
    atomicOr(&G.ctl->flags, DB_FULL);
    return -1;
}

__device__ __forceinline__ bool within(const GridDev& G, long long p, long long q)
{
    double d2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double df = coord(G, d, p) - coord(G, d, q);
        d2 = d2 + df * df;
    }
    return sqrt(d2) <= G.eps;
}

// Lower bound (squared) of the distance from point p to the box of cell c.
__device__ __forceinline__ double box_gap2(const GridDev& G, long long p, const long long* c)
{
    double g2 = 0.0;
    for (int d = 0; d < G.dim; ++d) {
        const double lo = G.ctl->lo[d] + (double)c[d] * G.s;
        const double hi = lo + G.s;
        const double v = coord(G, d, p);
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
        const double v = coord(G, d, p);
        const double g = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        const double f = fmax(v - lo, hi - v);
        g2 = g2 + g * g;
        f2 = f2 + f * f;
    }
    if (sqrt(g2) > G.eps) return 0;
    if (sqrt(f2) <= G.eps) return 2;
    return 1;
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

__device__ __forceinline__ bool is_core(const GridDev& G, int p) { return uf_load(&G.parent[p]) >= 0; }

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

__device__ __forceinline__ int wave_min_int(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << __lane_id()) - 1ull; }

// Wavefront groups: up to kAggGroups rounds, each taking the lowest pending
// lane as leader and every pending lane with the leader's key as its group.
// Lanes left pending afterwards (many distinct keys) update on their own.
constexpr int kAggGroups = 4;

// Per-block chunks of the aggregated kernels: 256 threads x kPer points,
// consecutive point indices per pass, and an LDS table of kLds cells.
constexpr int kPer = 8;
constexpr int kChunkPts = 256 * kPer;
constexpr int kLds = 512;

// Grid of the wavefront-per-item kernels (blocks of 256): 8192 waves, a full
// complement for 256 CUs; the waves stride over the device-side item count.
constexpr unsigned kWaveBlocks = 2048;

// Block-wide append of the lanes with `flag` to list[*counter ...]: one
// global atomic per block.  All threads of the block must call it.
__device__ __forceinline__ void block_append(bool flag, int value, int* list, int* counter)
{
    __shared__ int s_wave[4];
    __shared__ int s_base;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(flag);
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int off = s_base;
    for (int w = 0; w < wave; ++w) off += s_wave[w];
    if (flag) list[off + __popcll(m & lanes_below())] = value;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Cell tables to empty; with `all`, also the control block (box from `box`
// or empty), the fixed-point minima and the negative-root bits.
__global__ void init_kernel(const GridDev G, const unsigned long long* box, int all)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = t; i < G.H; i += stride) {
        G.hkey[i] = kEmpty;
        G.hcount[i] = 0u;
        G.hmin[i] = INT_MAX;
        G.hcur[i] = 0;
        for (int d = 0; d < 3; ++d) {
            G.bbox[d * G.H + i] = ~0ull;
            G.bbox[(3 + d) * G.H + i] = 0ull;
        }
    }
    if (!all) {
        if (t == 0) {
            G.ctl->ncells = 0;
            G.ctl->ncand = 0;
            G.ctl->ncore_cells = 0;
            G.ctl->flags = 0;
            G.ctl->K = 0;
            G.ctl->npieces = 0;
        }
        return;
    }
    if (G.sign) {
        for (long long i = t; i < G.n; i += stride) G.minfix[i] = INT_MAX;
        for (long long i = t; i < G.nw; i += stride) G.bitsN[i] = 0ull;
    }
    if (t == 0) {
        DbCtl c{};
        for (int d = 0; d < 3; ++d) {
            c.box[d] = box ? box[d] : ~0ull;
            c.box[3 + d] = box ? box[3 + d] : 0ull;
        }
        *G.ctl = c;
    }
}

// Bounding box of the included points (encoded, atomic min / max per block).
__global__ __launch_bounds__(256) void bbox_kernel(const GridDev G)
{
    __shared__ unsigned long long smin[3][4], smax[3][4];
    unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < G.n;
         p += (long long)gridDim.x * blockDim.x) {
        if (!included(G, p)) continue;
        for (int d = 0; d < G.dim; ++d) {
            const unsigned long long e = ord_enc(coord(G, d, p));
            mn[d] = e < mn[d] ? e : mn[d];
            mx[d] = e > mx[d] ? e : mx[d];
        }
    }
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (int d = 0; d < 3; ++d) {
        mn[d] = wave_min_u64(mn[d]);
        mx[d] = wave_max_u64(mx[d]);
        if (lane == 0) {
            smin[d][wave] = mn[d];
            smax[d][wave] = mx[d];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        unsigned long long a = smin[d][0], b = smax[d][0];
        for (int w = 1; w < 4; ++w) {
            a = smin[d][w] < a ? smin[d][w] : a;
            b = smax[d][w] > b ? smax[d][w] : b;
        }
        if (d < G.dim && a != ~0ull) {
            atomicMin(&G.ctl->box[d], a);
            atomicMax(&G.ctl->box[3 + d], b);
        }
    }
}

// Grid origin from the box; range and finiteness checks.
__global__ void setup_kernel(const GridDev G, double cell_lim)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DbCtl* c = G.ctl;
    c->any = c->box[0] != ~0ull;
    for (int d = 0; d < 3; ++d) c->lo[d] = 0.0;
    if (!c->any) return;
    for (int d = 0; d < G.dim; ++d) {
        const double lo = ord_dec(c->box[d]), hi = ord_dec(c->box[3 + d]);
        if (!isfinite(lo) || !isfinite(hi)) {
            c->flags |= DB_NONFINITE;
            continue;
        }
        c->lo[d] = lo;
        if (floor((hi - lo) / G.s) + 1 >= cell_lim) c->flags |= DB_RANGE;
    }
}

// Slot of `key` in the block's LDS table, claiming it if absent; -1 if full.
__device__ __forceinline__ int lds_claim(unsigned long long* s_key, unsigned long long key)
{
    int h = (int)(mix64(key) & (kLds - 1));
    for (int probe = 0; probe < kLds; ++probe) {
        const unsigned long long k = s_key[h];
        if (k == key) return h;
        if (k == kEmpty) {
            const unsigned long long prev = atomicCAS(&s_key[h], kEmpty, key);
            if (prev == kEmpty || prev == key) return h;
        }
        h = (h + 1) & (kLds - 1);
    }
    return -1;
}

// Cells of the included points: count, smallest member index and member box
// per cell, aggregated over the wavefront, then over the block's chunk in
// LDS, then added to the global table once per cell and block.
__global__ __launch_bounds__(256) void insert_kernel(const GridDev G)
{
    __shared__ unsigned long long s_key[kLds];
    __shared__ unsigned long long s_box[6][kLds];
    __shared__ unsigned s_cnt[kLds];
    __shared__ int s_min[kLds];
    __shared__ int s_slot[kLds];
    const int tid = threadIdx.x, lane = __lane_id();
    for (int i = tid; i < kLds; i += 256) {
        s_key[i] = kEmpty;
        s_cnt[i] = 0u;
        s_min[i] = INT_MAX;
        for (int d = 0; d < 3; ++d) {
            s_box[d][i] = ~0ull;
            s_box[3 + d][i] = 0ull;
        }
    }
    __syncthreads();
    // per point: LDS entry (>= 0), a global slot s encoded as -2 - s, or -1
    int loc[kPer];
    const long long base = (long long)blockIdx.x * kChunkPts;
    for (int k = 0; k < kPer; ++k) {
        const long long p = base + (long long)k * 256 + tid;
        const bool valid = p < G.n && included(G, p);
        unsigned long long key = kEmpty, e[3] = {0ull, 0ull, 0ull};
        if (valid) {
            long long c[3] = {0, 0, 0};
            cell_coords(G, p, c);
            key = pack_cell(G.dim, c);
            for (int d = 0; d < G.dim; ++d) e[d] = ord_enc(coord(G, d, p));
        }
        int mine_loc = -1;
        bool pending = valid;
        for (int it = 0; it <= kAggGroups; ++it) {
            const unsigned long long act = __ballot(pending);
            if (!act) break;
            // the last round: every lane left updates on its own
            const bool solo = it == kAggGroups;
            const int leader = __builtin_ctzll(act);
            const unsigned long long lk = __shfl(key, leader);
            const bool mine = pending && (solo || key == lk);
            const unsigned long long grp = __ballot(mine);
            const bool act_lane = solo ? mine : lane == leader;
            const unsigned cnt = solo ? 1u : (unsigned)__popcll(grp);
            unsigned long long mn[3], mx[3];
            for (int d = 0; d < 3; ++d) {
                mn[d] = solo ? e[d] : wave_min_u64(mine ? e[d] : ~0ull);
                mx[d] = solo ? e[d] : wave_max_u64(mine ? e[d] : 0ull);
            }
            int l = -1;
            if (act_lane) {
                const unsigned long long kk = solo ? key : lk;
                // the leader is the group's lowest lane: its index is the group's smallest
                const int pmin = (int)p;
                l = lds_claim(s_key, kk);
                if (l >= 0) {
                    atomicAdd(&s_cnt[l], cnt);
                    atomicMin(&s_min[l], pmin);
                    for (int d = 0; d < G.dim; ++d) {
                        atomicMin(&s_box[d][l], mn[d]);
                        atomicMax(&s_box[3 + d][l], mx[d]);
                    }
                } else {
                    const int slot = hash_insert(G, kk);
                    if (slot >= 0) {
                        atomicAdd(&G.hcount[slot], cnt);
                        atomicMin(&G.hmin[slot], pmin);
                        for (int d = 0; d < G.dim; ++d) {
                            atomicMin(&G.bbox[d * G.H + slot], mn[d]);
                            atomicMax(&G.bbox[(3 + d) * G.H + slot], mx[d]);
                        }
                    }
                    l = -2 - slot;  // -1 when the table is full
                }
            }
            if (!solo) l = __shfl(l, leader);
            if (mine) {
                mine_loc = l;
                pending = false;
            }
        }
        loc[k] = valid ? mine_loc : -1;
    }
    __syncthreads();
    for (int i = tid; i < kLds; i += 256) {
        const unsigned long long key = s_key[i];
        int slot = -1;
        if (key != kEmpty) {
            slot = hash_insert(G, key);
            if (slot >= 0) {
                atomicAdd(&G.hcount[slot], s_cnt[i]);
                atomicMin(&G.hmin[slot], s_min[i]);
                for (int d = 0; d < G.dim; ++d) {
                    atomicMin(&G.bbox[d * G.H + slot], s_box[d][i]);
                    atomicMax(&G.bbox[(3 + d) * G.H + slot], s_box[3 + d][i]);
                }
            }
        }
        s_slot[i] = slot;
    }
    __syncthreads();
    for (int k = 0; k < kPer; ++k) {
        const long long p = base + (long long)k * 256 + tid;
        if (p >= G.n) break;
        const int l = loc[k];
        G.pslot[p] = l >= 0 ? s_slot[l] : (l <= -2 ? -2 - l : -1);
    }
}

// Slot of an int key in an LDS table of kLds entries (-1 = empty).
__device__ __forceinline__ int lds_claim_int(int* s_key, int key)
{
    int h = (int)(mix64((unsigned long long)key) & (kLds - 1));
    for (int probe = 0; probe < kLds; ++probe) {
        const int k = s_key[h];
        if (k == key) return h;
        if (k == -1) {
            const int prev = atomicCAS(&s_key[h], -1, key);
            if (prev == -1 || prev == key) return h;
        }
        h = (h + 1) & (kLds - 1);
    }
    return -1;
}

// Points into their cells' segments of cellpts: ranks within the block's
// chunk per cell in LDS, one global reservation per cell and block.  Also the
// dense part of the core test: a cell holding >= minPts points is all-core
// (its diagonal is below eps) and its members link straight to its smallest
// member; points of sparser cells go to the candidate list.
__global__ __launch_bounds__(256) void place_kernel(const GridDev G, int minPts)
{
    __shared__ int s_slot[kLds];
    __shared__ int s_cnt[kLds];
    __shared__ int s_base[kLds];
    const int tid = threadIdx.x, lane = __lane_id();
    for (int i = tid; i < kLds; i += 256) {
        s_slot[i] = -1;
        s_cnt[i] = 0;
    }
    __syncthreads();
    int loc[kPer], rank[kPer], sl[kPer];  // loc: LDS entry, -2 = placed directly, -1 = none
    const long long base = (long long)blockIdx.x * kChunkPts;
    for (int k = 0; k < kPer; ++k) {
        const long long p = base + (long long)k * 256 + tid;
        const int s = p < G.n ? G.pslot[p] : -1;
        sl[k] = s;
        int my_loc = -1, my_rank = 0;
        bool pending = s >= 0;
        for (int it = 0; it <= kAggGroups; ++it) {
            const unsigned long long act = __ballot(pending);
            if (!act) break;
            const bool solo = it == kAggGroups;
            const int leader = __builtin_ctzll(act);
            const int ls = __shfl(s, leader);
            const bool mine = pending && (solo || s == ls);
            const unsigned long long grp = __ballot(mine);
            int l = -1, r0 = 0;
            if (solo ? mine : lane == leader) {
                const int cnt = solo ? 1 : __popcll(grp);
                const int key = solo ? s : ls;
                l = lds_claim_int(s_slot, key);
                if (l >= 0) {
                    r0 = atomicAdd(&s_cnt[l], cnt);
                } else {
                    r0 = G.hstart[key] + atomicAdd(&G.hcur[key], cnt);
                    l = -2;
                }
            }
            if (!solo) {
                l = __shfl(l, leader);
                r0 = __shfl(r0, leader);
            }
            if (mine) {
                my_loc = l;
                my_rank = r0 + (solo ? 0 : __popcll(grp & lanes_below()));
                pending = false;
            }
        }
        loc[k] = my_loc;
        rank[k] = my_rank;
    }
    __syncthreads();
    for (int i = tid; i < kLds; i += 256) {
        const int s = s_slot[i];
        s_base[i] = s >= 0 ? G.hstart[s] + atomicAdd(&G.hcur[s], s_cnt[i]) : 0;
    }
    __syncthreads();
    for (int k = 0; k < kPer; ++k) {
        const long long p = base + (long long)k * 256 + tid;
        const int l = loc[k];
        if (l >= 0) G.cellpts[s_base[l] + rank[k]] = (int)p;
        else if (l == -2) G.cellpts[rank[k]] = (int)p;
        const int s = sl[k];
        const bool dense = s >= 0 && (long long)G.hcount[s] >= minPts;
        if (p < G.n) G.parent[p] = dense ? G.hmin[s] : -1;
        block_append(s >= 0 && !dense, (int)p, G.cand, &G.ctl->ncand);  // every thread of the block
    }
}

// Smallest core index of the dense cells (all members core: their smallest
// member); INT_MAX elsewhere until the sparse core test.
__global__ void cell_init_kernel(const GridDev G, int minPts)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G.H) return;
    G.hrep[s] = (G.hkey[s] != kEmpty && (long long)G.hcount[s] >= minPts) ? G.hmin[s] : INT_MAX;
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

// Members per lane and pass of the wavefront cell scans: kScanU member
// indices are loaded, then their coordinates gathered, before any test, so a
// wavefront keeps 64 x kScanU gathers in flight (a cell of 10^5 members is
// about 200 passes).
constexpr int kScanU = 8;

// The up-to-2 x 64 neighbour cells of candidate p that may hold a point within
// eps, one per lane and j: slot and box relation (0 = none / no cell).
__device__ __forceinline__ void candidate_cells(const GridDev& G, int p, int* sl, int* rl)
{
    const int lane = __lane_id();
    const int noff = G.dim == 3 ? 125 : 25;
    long long c0[3] = {0, 0, 0};
    cell_coords(G, p, c0);
    for (int j = 0; j < 2; ++j) {
        sl[j] = -1;
        rl[j] = 0;
        const int o = lane + 64 * j;
        long long c[3];
        if (o >= noff || !neighbour_cell(G, c0, o, c)) continue;
        if (box_gap2(G, p, c) > G.prune2) continue;
        const int slot = lookup(G, pack_cell(G.dim, c));
        if (slot < 0) continue;
        sl[j] = slot;
        rl[j] = box_relation(G, p, slot);
    }
}

// Exclusive prefix sum over the lanes of a wavefront.
__device__ __forceinline__ int wave_excl_sum(int v)
{
    const int lane = __lane_id();
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x - v;
}

// Members of one piece of a split core test (8 passes of a wavefront); a
// candidate whose undecided cells hold more members than this is split.
constexpr int kPieceMembers = 64 * kScanU * 8;
// Members of each undecided cell a split candidate scans itself first (one
// pass): a core candidate usually reaches minPts there.
constexpr int kHeadMembers = 64 * kScanU;

// Hits of point p among members [b + first, min(e, b + last)) of every
// undecided cell (rl == 1) of the lane's slots, added to cnt, stopping once
// cnt reaches minPts; wavefront-uniform.
__device__ __forceinline__ long long scan_undecided(const GridDev& G, int p, const int* sl, const int* rl,
                                                    long long first, long long last, long long cnt, int minPts)
{
    const int lane = __lane_id();
    for (int j = 0; j < 2 && cnt < minPts; ++j) {
        unsigned long long m = __ballot(rl[j] == 1);
        while (m && cnt < minPts) {
            const int ld = __builtin_ctzll(m);
            m &= m - 1;
            const int slot = __shfl(sl[j], ld);
            const long long c0 = G.hstart[slot], c1 = c0 + (long long)G.hcount[slot];
            const long long b = c0 + first, e = c1 - c0 > last ? c0 + last : c1;
            for (long long k = b; k < e && cnt < minPts; k += 64 * kScanU) {
                int q[kScanU];
#pragma unroll
                for (int u = 0; u < kScanU; ++u) {
                    const long long kk = k + 64 * u + lane;
                    q[u] = kk < e ? G.cellpts[kk] : -1;
                }
#pragma unroll
                for (int u = 0; u < kScanU; ++u) cnt += __popcll(__ballot(q[u] >= 0 && within(G, p, q[u])));
            }
        }
    }
    return cnt;
}

// Core test, sparse part: one wavefront per candidate point.  The lanes look
// up the 5^dim neighbour cells, add whole cells whose box lies within eps and
// scan the undecided ones 64 x kScanU points a pass, stopping at minPts.  A
// candidate whose undecided cells hold more than kPieceMembers members (a
// point beside a cell of a converged basin: 10^5..10^6 members) scans the
// first kHeadMembers of each itself; if that does not decide it, the rest
// is split: its count so far goes to labroot[p] (free until the border
// pass) and the remaining members go to the piece list for
// core_piece_kernel.  labroot[p] = -1 marks a candidate decided here.  A
// full list falls back to the scan in this wavefront.
__global__ __launch_bounds__(256) void core_sparse_kernel(const GridDev G, int minPts)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int nc = G.ctl->ncand;
    for (long long i = wave; i < nc; i += nw) {
        const int p = G.cand[i];
        int sl[2], rl[2];
        candidate_cells(G, p, sl, rl);
        long long whole = 0, und = 0;
        for (int j = 0; j < 2; ++j) {
            if (rl[j] == 2) whole += (long long)G.hcount[sl[j]];
            if (rl[j] == 1) und += (long long)G.hcount[sl[j]];
        }
        long long cnt = wave_sum_ll(whole);
        const bool heavy = cnt < minPts && wave_sum_ll(und) > kPieceMembers;
        cnt = scan_undecided(G, p, sl, rl, 0, heavy ? kHeadMembers : LLONG_MAX, cnt, minPts);
        if (heavy && cnt < minPts) {
            int npc = 0;
            for (int j = 0; j < 2; ++j)
                if (rl[j] == 1) {
                    const long long rest = (long long)G.hcount[sl[j]] - kHeadMembers;
                    if (rest > 0) npc += (int)((rest + kPieceMembers - 1) / kPieceMembers);
                }
            const int off = wave_excl_sum(npc);
            const int tot = __shfl(off + npc, 63);
            // claim tot slots only while they fit, so that every slot below
            // the counter is written
            int base = -1;
            if (lane == 0) {
                int cur = __hip_atomic_load(&G.ctl->npieces, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while ((long long)cur + tot <= G.piece_cap) {
                    const int prev = atomicCAS(&G.ctl->npieces, cur, cur + tot);
                    if (prev == cur) {
                        base = cur;
                        break;
                    }
                    cur = prev;
                }
            }
            base = __shfl(base, 0);
            if (base >= 0) {
                int w = base + off;
                for (int j = 0; j < 2; ++j) {
                    if (rl[j] != 1) continue;
                    const long long c0 = G.hstart[sl[j]], e = c0 + (long long)G.hcount[sl[j]];
                    for (long long k = c0 + kHeadMembers; k < e; k += kPieceMembers)
                        G.pieces[w++] = make_int4(p, (int)k, (int)(k + kPieceMembers < e ? k + kPieceMembers : e), 0);
                }
                if (lane == 0) G.labroot[p] = (int)cnt;
                continue;
            }
            cnt = scan_undecided(G, p, sl, rl, kHeadMembers, LLONG_MAX, cnt, minPts);
        }
        if (lane == 0) {
            G.labroot[p] = -1;
            if (cnt >= minPts) {
                G.parent[p] = p;
                atomicMin(&G.hrep[G.pslot[p]], p);
            }
        }
    }
}

// Pieces of the split core tests, one wavefront per piece: the hits among
// the piece's members, added to the candidate's count after every pass; the
// piece stops once the candidate has reached minPts.
__global__ __launch_bounds__(256) void core_piece_kernel(const GridDev G, int minPts)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const long long np = G.ctl->npieces;  // <= piece_cap, every slot below it written
    for (long long i = wave; i < np; i += nw) {
        const int4 pc = G.pieces[i];
        const int p = pc.x;
        for (long long k = pc.y; k < pc.z; k += 64 * kScanU) {
            if (__builtin_amdgcn_readfirstlane(uf_load(&G.labroot[p])) >= minPts) break;
            int q[kScanU];
#pragma unroll
            for (int u = 0; u < kScanU; ++u) {
                const long long kk = k + 64 * u + lane;
                q[u] = kk < pc.z ? G.cellpts[kk] : -1;
            }
            int hits = 0;
#pragma unroll
            for (int u = 0; u < kScanU; ++u) hits += __popcll(__ballot(q[u] >= 0 && within(G, p, q[u])));
            if (lane == 0 && hits) atomicAdd(&G.labroot[p], hits);
        }
    }
}

// Verdicts of the split core tests (labroot[p] >= 0 marks a split candidate).
__global__ void core_split_kernel(const GridDev G, int minPts)
{
    const int nc = G.ctl->ncand;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int p = G.cand[i];
        if (G.labroot[p] >= minPts) {
            G.parent[p] = p;
            atomicMin(&G.hrep[G.pslot[p]], p);
        }
    }
}

// Border search of the non-core candidates (every non-core point of a cell
// lies in a sparse cell, so the candidate list holds them all): the smallest
// core index within eps, one wavefront per point, and its component root.
// Runs after border_kernel, which leaves -1 for these points.
__global__ __launch_bounds__(256) void border_sparse_kernel(const GridDev G)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int nc = G.ctl->ncand;
    for (long long i = wave; i < nc; i += nw) {
        const int p = G.cand[i];
        if (G.parent[p] >= 0) continue;  // core (uniform over the wavefront)
        int sl[2], rl[2];
        candidate_cells(G, p, sl, rl);
        int best = INT_MAX;
        for (int j = 0; j < 2; ++j)
            if (rl[j] == 2) best = min(best, G.hrep[sl[j]]);
        for (int j = 0; j < 2; ++j) {
            unsigned long long m = __ballot(rl[j] == 1);
            while (m) {
                const int ld = __builtin_ctzll(m);
                m &= m - 1;
                const int slot = __shfl(sl[j], ld);
                const long long b = G.hstart[slot], e = b + (long long)G.hcount[slot];
                for (long long k = b; k < e; k += 64 * kScanU) {
                    int q[kScanU];
#pragma unroll
                    for (int u = 0; u < kScanU; ++u) {
                        const long long kk = k + 64 * u + lane;
                        q[u] = kk < e ? G.cellpts[kk] : INT_MAX;
                    }
#pragma unroll
                    for (int u = 0; u < kScanU; ++u)
                        if (q[u] < best && G.parent[q[u]] >= 0 && within(G, p, q[u])) best = q[u];
                }
            }
        }
        best = wave_min_int(best);
        if (lane == 0 && best != INT_MAX) G.labroot[p] = uf_find(G, best);
    }
}

// Core points of sparse cells link to their cell's smallest core index (all
// members of a cell lie within eps of each other).
__global__ void link_sparse_kernel(const GridDev G)
{
    const int nc = G.ctl->ncand;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
         i += (long long)gridDim.x * blockDim.x) {
        const int p = G.cand[i];
        if (G.parent[p] >= 0) G.parent[p] = G.hrep[G.pslot[p]];
    }
}

// Cells holding at least one core point, compacted.
__global__ __launch_bounds__(256) void cell_list_kernel(const GridDev G)
{
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool has = s < G.H && G.hkey[s] != kEmpty && G.hrep[s] != INT_MAX;
    block_append(has, (int)s, G.cells, &G.ctl->ncore_cells);
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
            if (is_core(G, p)) rel = box_relation(G, p, b);
        }
        const unsigned long long m2 = __ballot(rel == 2);
        if (m2) {                                  // all of b within eps of p, rb among them
            if (lane == __builtin_ctzll(m2)) uf_union(G, p, rb);
            return;
        }
        const bool cand = rel == 1;
        if (!__ballot(cand)) continue;
        if (cand)
            for (int d = 0; d < G.dim; ++d) px[d] = coord(G, d, p);
        for (long long kb = b0; kb < b1; kb += 64) {
            int q = -1;
            bool qok = false;
            double qx[3] = {0.0, 0.0, 0.0};
            if (kb + lane < b1) {
                q = G.cellpts[kb + lane];
                qok = is_core(G, q) && box_relation(G, q, a) != 0;
                if (qok)
                    for (int d = 0; d < G.dim; ++d) qx[d] = coord(G, d, q);
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
__global__ __launch_bounds__(256) void link_pairs_kernel(const GridDev G, int ring)
{
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nw = ((long long)gridDim.x * blockDim.x) >> 6;
    const int nc = G.ctl->ncore_cells;
    const int noff = G.dim == 3 ? 125 : 25;
    for (long long i = wave; i < nc; i += nw) {
        const int a = G.cells[i];
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

// Label root of every core point (its component root; -1 for the others,
// whose border search border_sparse_kernel does next), the positive-root bit words
// (one wavefront = one 64-point word) and, under the sign rule, each
// component's smallest fixed-point core index.
__global__ __launch_bounds__(256) void border_kernel(const GridDev G)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int root = -1;
    bool isroot = false;
    if (p < G.n) {
        const int s = G.pslot[p];
        if (s >= 0) {
            const int par = G.parent[p];
            if (par >= 0) {
                root = uf_find(G, (int)p);
                isroot = par == (int)p;
                if (G.sign && G.include && G.include[p] == -1) atomicMin(&G.minfix[root], (int)p);
            }  // non-core: border_sparse_kernel
        }
        G.labroot[p] = root;
    }
    const unsigned long long bits = __ballot(isroot);
    if (__lane_id() == 0 && (p >> 6) < G.nw) G.bitsP[p >> 6] = bits;
}

// Sign rule: a root whose component holds a fixed point moves from the
// positive numbering to the negative one, keyed by that fixed point.
__global__ void sign_move_kernel(const GridDev G)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.n) return;
    const unsigned long long bit = 1ull << (p & 63);
    if (!(G.bitsP[p >> 6] & bit)) return;
    const int f = G.minfix[p];
    if (f == INT_MAX) return;
    atomicAnd(&G.bitsP[p >> 6], ~bit);
    atomicOr(&G.bitsN[f >> 6], 1ull << (f & 63));
}

__global__ void word_count_kernel(const GridDev G)
{
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= G.nw) return;
    G.rankP[w] = __popcll(G.bitsP[w]);
    if (G.sign) G.rankN[w] = __popcll(G.bitsN[w]);
}

__device__ __forceinline__ int bit_rank(const unsigned long long* bits, const int* rank, int x)
{
    return rank[x >> 6] + __popcll(bits[x >> 6] & ((1ull << (x & 63)) - 1ull));
}

__global__ void label_kernel(const GridDev G, int* labels)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= G.n) return;
    const int r = G.labroot[p];
    int lab = 0;
    if (r >= 0) {
        const int f = G.sign ? G.minfix[r] : INT_MAX;
        lab = f != INT_MAX ? -(bit_rank(G.bitsN, G.rankN, f) + 1) : bit_rank(G.bitsP, G.rankP, r) + 1;
    }
    labels[p] = lab;
}

__global__ void finish_kernel(const GridDev G)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long w = G.nw - 1;
    G.ctl->K = w >= 0 ? G.rankP[w] + __popcll(G.bitsP[w]) : 0;
}

// Exclusive prefix sums of int32 (in place allowed): tiles of 1024 scanned in
// LDS, tile sums scanned recursively in tmp.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 4;
constexpr int kScanTile = kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void scan_tile_kernel(const int* in, int* out, int* tile_sums,
                                                                 long long n)
{
    __shared__ int s[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(int* out, const int* tile_off, long long n)
{
    const long long i = (long long)blockIdx.x * kScanThreads + threadIdx.x;
    if (i < n) out[i] += tile_off[i / kScanTile];
}

static void scan_i32(hipStream_t st, const int* in, int* out, long long n, int* tmp)
{
    if (n <= 0) return;
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    int* sums = tmp;
    hipLaunchKernelGGL(scan_tile_kernel, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, out, sums, n);
    if (tiles > 1) {
        int* offs = tmp + tiles;
        scan_i32(st, sums, offs, tiles, tmp + 2 * tiles);
        hipLaunchKernelGGL(scan_add_kernel, dim3(ceil_div(n, kScanThreads)), dim3(kScanThreads), 0, st, out, offs, n);
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

// One pass of the pipeline on a carved workspace; returns the control block.
static DbCtl run_pipeline(CallCtx& ctx, GridDev& G, const DbscanParams& prm, int* labels,
                          const unsigned long long* box)
{
    hipStream_t st = ctx.stream();
    const unsigned T = 256;
    const long long n = G.n;
    const unsigned gn = ceil_div(n, T);
    const unsigned gh = ceil_div(G.H, T);
    const unsigned gi = std::min<unsigned>(std::max(gh, gn), 4096u);
    const double cell_lim = prm.dim == 3 ? (double)(1 << 21) - 4 : 2147483000.0;
    hipLaunchKernelGGL(init_kernel, dim3(gi), dim3(T), 0, st, G, box, 1);
    if (!box) hipLaunchKernelGGL(bbox_kernel, dim3(std::min<unsigned>(gn, 1024u)), dim3(T), 0, st, G);
    hipLaunchKernelGGL(setup_kernel, dim3(1), dim3(64), 0, st, G, cell_lim);
    const unsigned gc = ceil_div(n, kChunkPts);
    hipLaunchKernelGGL(insert_kernel, dim3(gc), dim3(T), 0, st, G);
    scan_i32(st, reinterpret_cast<const int*>(G.hcount), G.hstart, G.H, G.scan_tmp);
    hipLaunchKernelGGL(place_kernel, dim3(gc), dim3(T), 0, st, G, prm.minPts);
    hipLaunchKernelGGL(cell_init_kernel, dim3(gh), dim3(T), 0, st, G, prm.minPts);
    hipLaunchKernelGGL(core_sparse_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G, prm.minPts);
    hipLaunchKernelGGL(core_piece_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G, prm.minPts);
    hipLaunchKernelGGL(core_split_kernel, dim3(1024), dim3(T), 0, st, G, prm.minPts);
    hipLaunchKernelGGL(link_sparse_kernel, dim3(1024), dim3(T), 0, st, G);
    hipLaunchKernelGGL(cell_list_kernel, dim3(gh), dim3(T), 0, st, G);
    for (int ring = 0; ring < 3; ++ring)
        hipLaunchKernelGGL(link_pairs_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G, ring);
    hipLaunchKernelGGL(border_kernel, dim3(gn), dim3(T), 0, st, G);
    hipLaunchKernelGGL(border_sparse_kernel, dim3(kWaveBlocks), dim3(T), 0, st, G);
    if (G.sign) hipLaunchKernelGGL(sign_move_kernel, dim3(gn), dim3(T), 0, st, G);
    hipLaunchKernelGGL(word_count_kernel, dim3(ceil_div(G.nw, T)), dim3(T), 0, st, G);
    scan_i32(st, G.rankP, G.rankP, G.nw, G.scan_tmp);
    if (G.sign) scan_i32(st, G.rankN, G.rankN, G.nw, G.scan_tmp);
    hipLaunchKernelGGL(label_kernel, dim3(gn), dim3(T), 0, st, G, labels);
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, st, G);
    CHAOS_HIP(hipGetLastError());
    ctx.add_launch(24);
    DbCtl c{};
    CHAOS_HIP(hipMemcpyAsync(&c, G.ctl, sizeof(DbCtl), hipMemcpyDeviceToHost, st));
    CHAOS_HIP(hipStreamSynchronize(st));
    return c;
}

int dbscan_device(CallCtx& ctx, const double* const* pts, const int* include, long long n, const DbscanParams& prm,
                  int* labels, void* ws, size_t ws_bytes, const unsigned long long* box, const double* zmom)
{
    CHAOS_REQUIRE(prm.dim == 2 || prm.dim == 3, "dbscan dim must be 2 or 3");
    CHAOS_REQUIRE(prm.eps > 0 && std::isfinite(prm.eps), "eps must be positive");
    CHAOS_REQUIRE(prm.minPts >= 1, "minPts must be >= 1");
    CHAOS_REQUIRE(n >= 1 && n < INT_MAX, "dbscan point count out of range");
    GridDev G{};
    for (int d = 0; d < 3; ++d) G.x[d] = d < prm.dim ? pts[d] : pts[0];
    G.zmom = zmom;
    G.include = include;
    G.dim = prm.dim;
    G.sign = prm.sign_rule ? 1 : 0;
    G.eps = prm.eps;
    // cell side eps/sqrt(dim), shrunk so that rounding can never put two
    // points of one cell farther apart than eps
    G.s = prm.eps / std::sqrt((double)prm.dim) * (1.0 - 1e-7);
    G.prune2 = (prm.eps * (1 + 1e-7)) * (prm.eps * (1 + 1e-7));
    long long H = dbscan_first_cells(n);
    for (int attempt = 0; attempt < 2; ++attempt) {
        const size_t need = dbscan_ws_bytes(n, H, prm.sign_rule);
        DevBuf<char> own;
        char* base = static_cast<char*>(ws);
        if (!ws) {
            own.alloc(need);
            base = own.get();
        } else if (need > ws_bytes) {
            fail(CHAOS_E_CAPACITY, "clustering needs %zu bytes of scratch (%lld cells), %zu available", need, H,
                 ws_bytes);
        }
        carve(G, base, n, H, prm.sign_rule);
        const DbCtl c = run_pipeline(ctx, G, prm, labels, box);
        if (c.flags & DB_NONFINITE) fail(CHAOS_E_INVALID, "non-finite feature values");
        if (c.flags & DB_RANGE) fail(CHAOS_E_UNSUPPORTED, "feature range too wide for eps %.3g", prm.eps);
        if (std::getenv("CHAOS_DBSCAN_TRACE"))
            std::fprintf(stderr, "dbscan: %lld points, %d cells of %lld slots, %d sparse-cell points, "
                         "%d core cells, %d clusters, %d split pieces of %lld, flags %d\n", n, c.ncells, H,
                         c.ncand, c.ncore_cells, c.K, c.npieces, core_piece_cap(n), c.flags);
        if (!(c.flags & DB_FULL)) return c.K;
        if (H >= dbscan_max_cells(n)) fail(CHAOS_E_HIP, "cell table full at %lld cells", H);
        H = dbscan_max_cells(n);
    }
    fail(CHAOS_E_HIP, "clustering did not settle");
}

// ---------------------------------------------------------------------------
// deterministic z-score
// ---------------------------------------------------------------------------

// Block b sums its 1024 points by the pairwise tree (thread t adds points
// 2t and 2t+1, the wavefronts' butterflies give the 128-point runs, then the
// 8 runs pairwise); pass 0 sums the values, counts the masked points and
// takes their encoded minimum and maximum, pass 1 sums the squared
// deviations from the pass-0 mean.  Masked-out points add +0.
__global__ __launch_bounds__(512) void moment_block_kernel(const double* x0, const double* x1, const double* x2,
                                                           const int* __restrict__ mask, long long n, int k, int pass,
                                                           const double* __restrict__ mom, double* __restrict__ part,
                                                           int* __restrict__ cnt, unsigned long long* __restrict__ pmm)
{
    __shared__ double s_w[3][8];
    __shared__ unsigned long long s_mm[6][8];
    __shared__ int s_c[8];
    const double* xs[3] = {x0, x1, x2};
    const int t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
    const long long nb = gridDim.x;
    const long long i0 = (long long)blockIdx.x * 1024 + 2 * t, i1 = i0 + 1;
    const bool m0 = i0 < n && (!mask || mask[i0] != 0), m1 = i1 < n && (!mask || mask[i1] != 0);
    for (int j = 0; j < k; ++j) {
        const double* x = xs[j];
        const double v0 = m0 ? x[i0] : 0.0, v1 = m1 ? x[i1] : 0.0;
        double a0 = v0, a1 = v1;
        if (pass == 1) {
            const double mean = mom[2 * j];
            const double d0 = v0 - mean, d1 = v1 - mean;
            a0 = m0 ? d0 * d0 : 0.0;
            a1 = m1 ? d1 * d1 : 0.0;
        }
        const double sw = wave_tree_sum(a0 + a1);
        if (lane == 0) s_w[j][wave] = sw;
        if (pass == 0) {
            unsigned long long lo = ~0ull, hi = 0ull;
            if (m0) lo = hi = ord_enc(v0);
            if (m1) {
                const unsigned long long e = ord_enc(v1);
                lo = e < lo ? e : lo;
                hi = e > hi ? e : hi;
            }
            lo = wave_min_enc(lo);
            hi = wave_max_enc(hi);
            if (lane == 0) {
                s_mm[2 * j][wave] = lo;
                s_mm[2 * j + 1][wave] = hi;
            }
        }
    }
    if (pass == 0) {
        const int c = __popcll(__ballot(m0)) + __popcll(__ballot(m1));
        if (lane == 0) s_c[wave] = c;
    }
    __syncthreads();
    if (t == 0) {
        const long long b = blockIdx.x;
        for (int j = 0; j < k; ++j) {
            const double* w = s_w[j];
            part[j * nb + b] = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
            if (pass == 0) {
                unsigned long long lo = s_mm[2 * j][0], hi = s_mm[2 * j + 1][0];
                for (int u = 1; u < 8; ++u) {
                    lo = s_mm[2 * j][u] < lo ? s_mm[2 * j][u] : lo;
                    hi = s_mm[2 * j + 1][u] > hi ? s_mm[2 * j + 1][u] : hi;
                }
                pmm[(2 * j) * nb + b] = lo;
                pmm[(2 * j + 1) * nb + b] = hi;
            }
        }
        if (pass == 0) {
            int tot = 0;
            for (int u = 0; u < 8; ++u) tot += s_c[u];
            cnt[b] = tot;
        }
    }
}

// The z-score of value v with moments (mean, sd): the host restatement's
// expression (sd 0: every value maps to 0).
__device__ __forceinline__ double zval(double v, double mean, double sd) { return sd > 0 ? (v - mean) / sd : 0.0; }

// Values per LDS tree (16 KB of doubles).
constexpr int kTreeChunk = 2048;

// Pairwise tree of w values s[0..w) (w a power of two <= kTreeChunk) by a
// 256-thread block: level by level, pairs (2t, 2t+1) -> t.  Every thread of
// the block calls it and gets the sum; s is overwritten.
__device__ double lds_tree(double* s, int w)
{
    for (int h = w >> 1; h >= 1; h >>= 1) {
        double v[kTreeChunk / 512];
#pragma unroll
        for (int u = 0; u < kTreeChunk / 512; ++u) {
            const int t = threadIdx.x + 256 * u;
            v[u] = t < h ? s[2 * t] + s[2 * t + 1] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kTreeChunk / 512; ++u) {
            const int t = threadIdx.x + 256 * u;
            if (t < h) s[t] = v[u];
        }
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// 1024-point runs of the padded tree for n points: P / 1024 with P = max(1024,
// the power of two >= n).
__host__ __device__ inline long long moment_runs(long long n)
{
    long long P = 1024;
    while (P < n) P <<= 1;
    return P / 1024;
}

// mom = [mean_0, sd_0, ..., mean_{k-1}, sd_{k-1}, N]; rmm = encoded raw
// minimum and maximum per feature.  The block partials in part / cnt / pmm
// cover runs of `span` points (1024: moment_block_kernel; 256: the basins
// sweep's epilogue), nb of them; the tree goes on from there over the
// 1024-point runs (four 256-point runs pairwise) up to the padded whole.
// Pass 0: N, the means and rmm; pass 1: the standard deviations and, when
// box is not NULL, the encoded bounding box of the z-scored masked points (z
// is monotone in v, so the box is the z-scores of the raw extremes; empty:
// min ~0, max 0).  One block of 256 threads.
__global__ __launch_bounds__(256) void moment_final_kernel(const double* __restrict__ part, const int* __restrict__ cnt,
                                                           const unsigned long long* __restrict__ pmm, long long nb,
                                                           int span, long long n, int k, int pass,
                                                           double* __restrict__ mom, unsigned long long* __restrict__ rmm,
                                                           unsigned long long* __restrict__ box)
{
    __shared__ double s[kTreeChunk];
    __shared__ double s_chunk[1024];
    __shared__ long long s_n[256];
    __shared__ unsigned long long s_mm[6][256];
    const int t = threadIdx.x;
    if (pass == 0) {
        // the count (an integer sum) and the raw extremes, in any order
        long long c = 0;
        unsigned long long mm[6] = {~0ull, 0ull, ~0ull, 0ull, ~0ull, 0ull};
        for (long long b = t; b < nb; b += 256) {
            c += cnt[b];
            for (int f = 0; f < k; ++f) {
                const unsigned long long lo = pmm[(2 * f) * nb + b], hi = pmm[(2 * f + 1) * nb + b];
                mm[2 * f] = lo < mm[2 * f] ? lo : mm[2 * f];
                mm[2 * f + 1] = hi > mm[2 * f + 1] ? hi : mm[2 * f + 1];
            }
        }
        s_n[t] = c;
        for (int f = 0; f < 6; ++f) s_mm[f][t] = mm[f];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) {
                s_n[t] += s_n[t + o];
                for (int f = 0; f < 6; f += 2) {
                    s_mm[f][t] = s_mm[f][t + o] < s_mm[f][t] ? s_mm[f][t + o] : s_mm[f][t];
                    s_mm[f + 1][t] = s_mm[f + 1][t + o] > s_mm[f + 1][t] ? s_mm[f + 1][t + o] : s_mm[f + 1][t];
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            mom[2 * k] = (double)s_n[0];
            for (int f = 0; f < 2 * k; ++f) rmm[f] = s_mm[f][0];
        }
        __syncthreads();
    }
    const long long runs = moment_runs(n);
    const double N = pass == 0 ? (double)s_n[0] : mom[2 * k];
    for (int j = 0; j < k; ++j) {
        const double* pj = part + j * nb;
        auto leaf = [&](long long r) -> double {  // the 1024-point run r
            if (span == 1024) return r < nb ? pj[r] : 0.0;
            const long long q = 4 * r;
            const double a = q < nb ? pj[q] : 0.0, b = q + 1 < nb ? pj[q + 1] : 0.0;
            const double c = q + 2 < nb ? pj[q + 2] : 0.0, d = q + 3 < nb ? pj[q + 3] : 0.0;
            return (a + b) + (c + d);
        };
        double tot;
        if (runs <= kTreeChunk) {
            for (int u = t; u < runs; u += 256) s[u] = leaf(u);
            __syncthreads();
            tot = lds_tree(s, (int)runs);
        } else {
            const long long nch = runs / kTreeChunk;  // <= 1024 for n < 2^31
            for (long long ch = 0; ch < nch; ++ch) {
                for (int u = t; u < kTreeChunk; u += 256) s[u] = leaf(ch * kTreeChunk + u);
                __syncthreads();
                const double cs = lds_tree(s, kTreeChunk);
                if (t == 0) s_chunk[ch] = cs;
            }
            __syncthreads();
            for (int u = t; u < nch; u += 256) s[u] = s_chunk[u];
            __syncthreads();
            tot = lds_tree(s, (int)nch);
        }
        if (t == 0) {
            if (pass == 0) {
                mom[2 * j] = N > 0 ? tot / N : 0.0;
            } else {
                const double sd = N > 0 ? sqrt(tot / N) : 0.0;
                mom[2 * j + 1] = sd;
                if (box) {
                    const unsigned long long lo = rmm[2 * j], hi = rmm[2 * j + 1];
                    const bool any = lo != ~0ull;
                    box[j] = any ? ord_enc(zval(ord_dec(lo), mom[2 * j], sd)) : ~0ull;
                    box[3 + j] = any ? ord_enc(zval(ord_dec(hi), mom[2 * j], sd)) : 0ull;
                }
            }
        }
    }
    if (pass == 1 && box && t < 3 && t >= k) {
        box[t] = ~0ull;
        box[3 + t] = 0ull;
    }
}

namespace {

struct ZLayout {
    double* part;
    int* cnt;
    unsigned long long* pmm;
    double* mom;
    unsigned long long* rmm;
    size_t bytes;
};

// Partials sized for 256-point blocks (the larger form), then the moments
// and the raw extremes.
ZLayout zlayout(void* scratch, long long n, int k)
{
    const long long nb = (n + 255) / 256;
    char* p = static_cast<char*>(scratch);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* q = p ? p + off : nullptr;
        off += align256(bytes);
        return q;
    };
    ZLayout z{};
    z.part = (double*)take(8 * (size_t)k * nb);
    z.cnt = (int*)take(4 * (size_t)nb);
    z.pmm = (unsigned long long*)take(8 * (size_t)(2 * k) * nb);
    z.mom = (double*)take(8 * (size_t)(2 * k + 1));
    z.rmm = (unsigned long long*)take(8 * (size_t)(2 * k));
    z.bytes = off;
    return z;
}

}  // namespace

size_t zscore_scratch_bytes(long long n, int k) { return zlayout(nullptr, n, k).bytes; }

ZPartials zscore_sweep_partials(void* scratch, long long n, int k)
{
    const ZLayout z = zlayout(scratch, n, k);
    return ZPartials{z.part, z.cnt, z.pmm, (n + 255) / 256};
}

const double* zscore_moments_device(CallCtx& ctx, const double* const* src, int k, const int* mask, long long n,
                                    unsigned long long* box, void* scratch, bool sweep_partials)
{
    CHAOS_REQUIRE(k >= 1 && k <= 3, "z-score of 1 to 3 features");
    CHAOS_REQUIRE(n >= 1 && n < (1LL << 31), "z-score point count out of range");
    CHAOS_REQUIRE(scratch, "z-score scratch is NULL");
    hipStream_t st = ctx.stream();
    const ZLayout z = zlayout(scratch, n, k);
    const long long nb1024 = (n + 1023) / 1024, nb256 = (n + 255) / 256;
    const double* s[3] = {src[0], k > 1 ? src[1] : src[0], k > 2 ? src[2] : src[0]};
    int launches = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool from_sweep = pass == 0 && sweep_partials;
        if (!from_sweep) {
            hipLaunchKernelGGL(moment_block_kernel, dim3((unsigned)nb1024), dim3(512), 0, st, s[0], s[1], s[2], mask,
                               n, k, pass, z.mom, z.part, z.cnt, z.pmm);
            ++launches;
        }
        hipLaunchKernelGGL(moment_final_kernel, dim3(1), dim3(256), 0, st, z.part, z.cnt, z.pmm,
                           from_sweep ? nb256 : nb1024, from_sweep ? 256 : 1024, n, k, pass, z.mom, z.rmm,
                           pass == 1 ? box : nullptr);
        ++launches;
    }
    CHAOS_HIP(hipGetLastError());
    ctx.add_launch(launches);
    return z.mom;
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
            if (inc) {
                d_inc.alloc(n);
                CHAOS_HIP(hipMemcpyAsync(d_inc.get(), inc, n * sizeof(int), hipMemcpyHostToDevice, st));
                inc = d_inc.get();
            }
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
