// peaks.hpp -- online peak detection over a lane's sample stream, plus the
// wavefront-compacted peak store.
//
// The detector replays the reference peakFinder
// (Bifurcation/src/bifurcationCUDA.cu:469-514, same text in
// Basins/src/basinsCUDA.cu:382-440) one sample at a time, in registers,
// instead of re-reading a stored trajectory:
//   candidate at i:   d[i]-d[i-1] > 1e-13  &&  d[i] >= d[i+1]
//   plateau scan j:   d[j] < d[j+1]          -> reject, next candidate i = j+2
//                     d[j]-d[j+1] > 1e-13    -> peak (d[j], trunc((i+j)/2)),
//                                               next candidate i = j+2
// (the reference's i = j+1 followed by the loop's ++i skips index j+1).
// All comparisons are on exact differences of adjacent samples:
// d[c]-d[c+1] == -(d[c+1]-d[c]) and d[c] >= d[c+1] <=> d[c+1]-d[c] <= 0 for
// finite samples, so one subtraction per sample feeds all three tests.
#pragma once

#include <hip/hip_runtime.h>

namespace chaos {

struct PeakDetector {
    double p;      // last sample d[t-1]
    double dn;     // d[t-1] - d[t-2]
    int pstart;    // plateau start index
    bool plat;     // scanning a plateau
    bool skip;     // next index is not a candidate

    __device__ __forceinline__ void reset()
    {
        p = 0.0;
        dn = 0.0;
        pstart = 0;
        plat = false;
        skip = false;
    }

    // Feed sample d[t] = cur (t >= 0).  For t >= 2 the index c = t-1 is
    // examined; on a peak, returns true with its value (d[c]) and time.
    __device__ __forceinline__ bool feed(int t, double cur, double& val, int& time)
    {
        const double dnx = cur - p;  // d[c+1] - d[c]
        bool emit = false;
        if (t >= 2) {
            const int c = t - 1;
            const bool c1 = dn > 1e-13;     // d[c] - d[c-1] > 1e-13
            const bool c2 = dnx <= 0.0;     // d[c] >= d[c+1]
            const bool c3 = dnx < -1e-13;   // d[c] - d[c+1] > 1e-13
            if (plat) {
                if (!c2) {
                    plat = false;
                    skip = true;
                } else if (c3) {
                    emit = true;
                    val = p;
                    time = (pstart + c) >> 1;
                    plat = false;
                    skip = true;
                }
            } else if (skip) {
                skip = false;
            } else if (c1 && c2) {
                if (c3) {
                    emit = true;
                    val = p;
                    time = c;
                    skip = true;
                } else {
                    plat = true;
                    pstart = c;
                }
            }
        }
        dn = dnx;
        p = cur;
        return emit;
    }

    // feed() for t >= 2 (c = t - 1) without branches: the state machine as
    // lane predicates, so the hot record loop carries no divergent control
    // flow.  Same transitions as feed():
    //   plateau scan or fresh candidate ("act") emits iff d[c]-d[c+1] > 1e-13;
    //   it stays / becomes a plateau iff the step is flat within 1e-13;
    //   skip follows an emit or a rejected plateau.
    __device__ __forceinline__ bool feed_body(int c, double cur, double& val, int& time)
    {
        const double dnx = cur - p;
        const bool c1 = dn > 1e-13;
        const bool c2 = dnx <= 0.0;
        const bool c3 = dnx < -1e-13;
        const bool cand = !plat && !skip && c1 && c2;
        const bool act = plat || cand;
        const bool emit = act && c3;
        const bool flat = c2 && !c3;
        val = p;
        time = plat ? (pstart + c) >> 1 : c;
        pstart = cand && flat ? c : pstart;
        skip = (plat && !flat) || (cand && c3);
        plat = act && flat;
        dn = dnx;
        p = cur;
        return emit;
    }
};

// Lane mask <-> lane predicate: a compare written straight into a wave mask,
// and a wave mask used as a lane predicate, both without VALU work.  Every
// lane of the wave must be active where a mask is formed.
__device__ __forceinline__ unsigned long long wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool lane_bit(unsigned long long m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// PeakDetector for a whole wave, as the sweeps use it: the same transitions
// with the two state flags held as wave masks (SGPRs).  The three tests are
// compares written into masks, the state machine runs on the scalar unit and
// the emit mask feeds the wave's append directly -- a per-lane bool state
// costs two VALU instructions per sample to form each ballot.  All 64 lanes
// must take every call (wave-uniform control flow).  TIMES: track the
// plateau start and report peak times (not needed when only peak values are
// used).
template <bool TIMES>
struct WavePeakDetector {
    double p;                 // last sample d[t-1]
    double dn;                // d[t-1] - d[t-2]
    int pstart;               // plateau start index
    unsigned long long plat;  // lanes scanning a plateau
    unsigned long long skip;  // lanes whose next index is not a candidate

    __device__ __forceinline__ void reset()
    {
        p = 0.0;
        dn = 0.0;
        pstart = 0;
        plat = 0ull;
        skip = 0ull;
    }

    // t >= 2, c = t - 1: PeakDetector::feed_body over the wave; returns the
    // mask of lanes with a peak at c (value val, time time).
    __device__ __forceinline__ unsigned long long feed_body(int c, double cur, double& val, int& time)
    {
        const double dnx = cur - p;
        const unsigned long long c1 = wave_mask(dn > 1e-13);
        const unsigned long long c2 = wave_mask(dnx <= 0.0);
        const unsigned long long c3 = wave_mask(dnx < -1e-13);
        const unsigned long long cand = ~plat & ~skip & c1 & c2;
        const unsigned long long act = plat | cand;
        const unsigned long long emit = act & c3;
        const unsigned long long flat = c2 & ~c3;
        val = p;
        if (TIMES) {
            // a peak's time is its sample index, or the middle of the
            // plateau for a lane that ends one; plateaus are rare, so both
            // lane updates sit behind wave-uniform branches
            time = c;
            if (plat & c3) time = lane_bit(plat) ? (pstart + c) >> 1 : c;
            const unsigned long long ps = cand & flat;
            if (ps) pstart = lane_bit(ps) ? c : pstart;
        } else {
            time = 0;
        }
        skip = (plat & ~flat) | (cand & c3);
        plat = act & flat;
        dn = dnx;
        p = cur;
        return emit;
    }

    // Any t: samples 0 and 1 only fill the history.
    __device__ __forceinline__ unsigned long long feed(int t, double cur, double& val, int& time)
    {
        if (t >= 2) return feed_body(t - 1, cur, val, time);
        dn = cur - p;
        p = cur;
        val = 0.0;
        time = 0;
        return 0ull;
    }
};

// One stored peak: value, sample index, and (sequence number << 6 | lane).
struct PeakEntry {
    double v;
    unsigned t;
    unsigned kl;
};

constexpr int kChunk = 256;  // entries per chunk (4 KiB)

// Pool of fixed-size chunks handed out to waves by an atomic head.  Each
// wave appends its lanes' peaks in time order with one ballot + mbcnt per
// sample that produced any peak, so a chunk holds one wave's entries densely.
// The bookkeeping sits in one array meta[2 * nchunks + 2] (owner, fill,
// head, overflow), so a sweep loop holds one pointer for it, not four.
struct PeakPool {
    PeakEntry* entries;     // [nchunks * kChunk]
    int* meta;              // owner[nchunks], fill[nchunks], head, overflow
    unsigned nchunks;
    __host__ __device__ int* owner() const { return meta; }                    // global wave id per chunk
    __host__ __device__ int* fill() const { return meta + nchunks; }           // entries used per chunk
    __host__ __device__ unsigned* head() const { return (unsigned*)(meta + 2 * (size_t)nchunks); }  // next free chunk
    __host__ __device__ int* overflow() const { return meta + 2 * (size_t)nchunks + 1; }  // pool ran out
};

struct WaveCursor {
    unsigned chunk;  // current chunk (nchunks when none / exhausted)
    int fill;        // entries used in it
};

__device__ __forceinline__ unsigned lane_prefix(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ int first_lane(unsigned long long active) { return __ffsll((long long)active) - 1; }

// Close the cursor's chunk (record its fill) -- call from all active lanes.
__device__ __forceinline__ void close_chunk(const PeakPool& pool, WaveCursor& cur)
{
    const unsigned long long act = __ballot(1);
    if (cur.chunk < pool.nchunks && (int)__lane_id() == first_lane(act)) pool.fill()[cur.chunk] = cur.fill;
}

// Append the entries of the lanes in mask m (wave-uniform control flow
// required).
__device__ __forceinline__ void append_peaks(const PeakPool& pool, WaveCursor& cur, int wave,
                                             unsigned long long m, double v, int t, int k)
{
    if (m == 0ull) return;
    const bool emit = lane_bit(m);
    const int n = __popcll(m);
    if (cur.fill + n > kChunk) {
        const unsigned long long act = __ballot(1);
        const bool leader = (int)__lane_id() == first_lane(act);
        if (cur.chunk < pool.nchunks && leader) pool.fill()[cur.chunk] = cur.fill;
        unsigned nc = 0;
        if (leader) nc = atomicAdd(pool.head(), 1u);
        // the leader is the first active lane: the cursor stays wave-uniform (SGPRs)
        nc = __builtin_amdgcn_readfirstlane(nc);
        if (nc < pool.nchunks) {
            if (leader) pool.owner()[nc] = wave;
        } else {
            if (leader) atomicExch(pool.overflow(), 1);
            nc = pool.nchunks;
        }
        cur.chunk = nc;
        cur.fill = 0;
    }
    if (emit && cur.chunk < pool.nchunks) {
        // wave-uniform chunk base + a 32-bit byte offset inside the chunk
        const unsigned off = ((unsigned)cur.fill + lane_prefix(m)) * (unsigned)sizeof(PeakEntry);
        char* base = (char*)(pool.entries + (size_t)cur.chunk * kChunk);
        PeakEntry e;
        e.v = v;
        e.t = (unsigned)t;
        e.kl = ((unsigned)k << 6) | (unsigned)__lane_id();
        *(PeakEntry*)(base + off) = e;
    }
    cur.fill += n;
}

}  // namespace chaos
