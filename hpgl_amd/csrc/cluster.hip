// cluster.hip - the cell-binned neighbour search with the reference's exact
// neighbour order (sugarbox_indexed_neighbour_lookup.h:66-123).
//
// Per target the reference gathers the informed cells of its 27 bins
// (bins k, j, i ascending; cells ascending inside a bin), keeps those with
// covariance > C0/100, sorts them with libstdc++ std::sort by descending
// covariance (unstable introsort: ties come out in an order fixed by the
// algorithm's steps) and keeps the first max_n.  This kernel reproduces that
// order exactly:
//
//  * keys are dense ranks of the host-computed double covariances
//    (RankTable), so comparisons are those of the reference.  An element is
//    packed as (rank << lb) | j, where j is the candidate's position in the
//    raw 27-bin concatenation (lb bits), so a candidate is 4 bytes in LDS;
//    the cell is decoded from j only for the max_n elements that are kept;
//  * the introsort is replayed step by step, one wavefront per target, with
//    each __unguarded_partition done in parallel.  In that partition the
//    k-th left stop (first elements with key <= pivot, scanning up) is
//    swapped with the k-th right stop (key >= pivot, scanning down) while
//    the left stop is below the right stop, and the cut is
//    min(l_{s+1}, r_s) for s swaps.  Two cursors ballot the stop masks of
//    64-position chunks from both ends, 64 stop pairs are laid out at a
//    time and their (disjoint) swaps performed together;
//  * ranges that start at or beyond max_n are never partitioned: they
//    cannot change the first max_n positions, which after the final
//    insertion sort are the stable sort of the partition blocks they fall
//    in.  The first min(n, max_n + 16) positions are therefore finished by a
//    stable rank sort;
//  * a range whose depth budget (2 floor(log2 n)) runs out is heap-sorted
//    by one lane with libstdc++'s make_heap/sort_heap steps.
#include "counters.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

namespace hpgl_amd {

namespace {

// Candidates a wavefront holds in LDS (dynamic LDS, 4 bytes each, so the
// host sizes it to the job's expected lists and occupancy follows): a
// multiple of 64, at most 4096 (the register-held partition covers 64
// chunks); longer lists go to global scratch.
constexpr int kMinCap = 256, kMaxCap = 4096;

// The 27 bins of a target in gather order (k, j, i ascending) as 21
// segments of the bin-sorted arrays: the 9 bins of the plane below, one
// segment each, keeping the cells with z >= z_target - span; the 3 rows of
// the target's plane, each three adjacent bins in one segment; the 9 bins
// of the plane above, keeping z <= z_target + span.  A bin's cells ascend
// by cell index, hence by z, so the kept cells are a suffix (below) or a
// prefix (above) and the order of the candidates is unchanged; the cells
// dropped have offsets outside the rank table (rank 0).
constexpr int kSegs = 21, kMaxSegs = 32;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr int kWavesPerGroup = 2;

// A wavefront orders its own accesses (LDS and the global scratch lists)
// with a workgroup-scope fence, which waits for them to complete; the
// workgroup is the wavefront itself, so no barrier is needed.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Mask of the lanes whose predicate holds, straight from the compare.
__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Fixed LDS of a wavefront; the elements of a list that fits the capacity
// live in the dynamic LDS.
struct WaveLds {
    // positions of the 64 stop pairs being swapped: left stops in [0, 64),
    // right stops in [64, 128); [128, 192) takes stores that are not needed
    int pairs[192];
    // candidate segments (kSegs, then kMaxSegs - kSegs sentinels): first
    // raw index of each (segment sizes padded to multiples of 64) and the
    // offset from raw index to position in the bin-sorted arrays
    int seg_raw[kMaxSegs];
    int seg_off[kMaxSegs];
};

constexpr int kGatherUnrollMax = 4;  // 64-candidate groups whose loads are in flight together
// Targets gathered together when consecutive list entries share their bins.
constexpr int kTargetsPerRun = kClusterRunSlack + 1;

struct Keyed {
    int lb;
    __device__ uint32_t operator()(uint32_t e) const { return e >> lb; }
};

// libstdc++ heap steps (stl_heap.h) on e[first, last) with
// comp(a, b) = key(a) > key(b).  Run by a single lane.
template <typename E>
__device__ void push_heap_(E *e, long hole, long top, uint32_t value, Keyed key) {
    long parent = (hole - 1) / 2;
    while (hole > top && key(e[parent]) > key(value)) {
        e[hole] = e[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    e[hole] = value;
}
template <typename E>
__device__ void adjust_heap_(E *e, long hole, long len, uint32_t value, Keyed key) {
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (key(e[second]) > key(e[second - 1])) second--;
        e[hole] = e[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        e[hole] = e[second - 1];
        hole = second - 1;
    }
    push_heap_(e, hole, top, value, key);
}
template <typename E>
__device__ void heap_sort_(E *e, long len, Keyed key) {
    if (len >= 2) {  // __make_heap
        long parent = (len - 2) / 2;
        while (true) {
            uint32_t v = e[parent];
            adjust_heap_(e, parent, len, v, key);
            if (parent == 0) break;
            parent--;
        }
    }
    long last = len;  // __sort_heap
    while (last > 1) {
        --last;
        uint32_t v = e[last];
        e[last] = e[0];
        adjust_heap_(e, 0, last, v, key);
    }
}

// Per-lane select on a wave-uniform lane mask: t where bit lane of m is
// set, else f (a single v_cndmask with the mask as its condition).
__device__ inline int lane_select(uint64_t m, int t, int f) {
    int r;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(f), "v"(t), "s"(m));
    return r;
}

// base[idx] for idx < 2^30: a 32-bit byte offset, so the load takes the
// scalar-base + vector-offset form (no 64-bit address arithmetic).
__device__ inline uint32_t load_u32(const uint32_t *base, uint32_t idx) {
    return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + (idx << 2));
}

// Number of set bits of the wave-uniform mask m below this lane (plus base).
__device__ inline int bits_below(uint64_t m, int base = 0) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)base));
}

// std::__unguarded_partition_pivot on e[lo, hi) (e in LDS or in global
// scratch).  With pivot key kp the left scan stops at elements with
// key <= kp, the right scan at key >= kp; the k-th left stop L_k is swapped
// with the k-th right stop R_k (counted from the end) while L_k < R_k, and
// for s swaps the cut is min(L_s, R_{s-1}).
//
// Two cursors produce the stops 64 pairs at a time: the left one walks
// chunks of 64 positions upwards from the range start, the right one
// downwards from its end; entering a chunk, a cursor ballots its stop mask
// and keeps the stops not yet used in a scalar mask.  Swaps only touch
// positions both scans have passed, so a mask taken before earlier swaps
// is exact for every position below R_{s-1}, and the stops it reports at
// or beyond R_{s-1} change neither the swaps nor min(L_s, R_{s-1}).
template <typename S>
__device__ __forceinline__ int partition_(uint32_t *e, S &sh, int lo, int hi, Keyed key) {
    const int lane = threadIdx.x & 63;
    uint32_t kp;
    {  // __move_median_to_first(lo, lo+1, mid, hi-1)
        const int a = lo + 1, bm = lo + (hi - lo) / 2, c = hi - 1;
        const uint32_t ka = __builtin_amdgcn_readfirstlane(key(e[a]));
        const uint32_t kb = __builtin_amdgcn_readfirstlane(key(e[bm]));
        const uint32_t kc = __builtin_amdgcn_readfirstlane(key(e[c]));
        int sw;
        if (ka > kb) {
            if (kb > kc)
                sw = bm;
            else if (ka > kc)
                sw = c;
            else
                sw = a;
        } else if (ka > kc) {
            sw = a;
        } else if (kb > kc) {
            sw = c;
        } else {
            sw = bm;
        }
        kp = sw == a ? ka : sw == bm ? kb : kc;
        wave_sync();
        if (lane == 0) {
            const uint32_t t0 = e[lo];
            e[lo] = e[sw];
            e[sw] = t0;
        }
        wave_sync();
    }
    // whole elements against thresholds: key <= kp iff e <= t_le, key >= kp
    // iff e >= t_ge
    const uint32_t t_ge = kp << key.lb, t_le = t_ge | ((1u << key.lb) - 1u);
    const int first = lo + 1;
    const int nch = (hi - first + 63) >> 6;
    const int p0 = first + lane;  // this lane's position in chunk 0
    int cl = -1, cr = nch;        // chunks the cursors are in
    int pcl = p0 - 64, pcr = p0 + (nch << 6);  // this lane's position in them
    uint64_t lrem = 0, rrem = 0;  // their stops not yet paired
    int r_prev = 0x7fffffff;      // R_{s-1} of the rounds before
    for (;;) {
        // left stops L_{s} .. L_{s+63} into pairs[0, 64), in position order:
        // use the stops left in the current chunk, then enter the next one
        int fl = 0;
        for (;;) {
            const int idx = bits_below(lrem, fl);
            const uint64_t took = lrem & ballot((unsigned)idx < 64u);
            sh.pairs[lane_select(took, idx, 128 + lane)] = pcl;
            lrem ^= took;
            fl += __popcll(took);
            if (fl == 64 || cl + 1 >= nch) break;
            ++cl;
            pcl += 64;
            lrem = ballot(e[min(pcl, hi - 1)] <= t_le) & ballot(pcl < hi);
        }
        // right stops R_{s} .. R_{s+63} into pairs[64, 128), in descending
        // position order
        int fr = 0;
        for (;;) {
            const int idx = fr + __popcll(rrem) - 1 - bits_below(rrem);
            const uint64_t took = rrem & ballot((unsigned)idx < 64u);
            sh.pairs[lane_select(took, 64 + idx, 128 + lane)] = pcr;
            rrem ^= took;
            fr += __popcll(took);
            if (fr == 64 || cr <= 0) break;
            --cr;
            pcr -= 64;
            rrem = ballot(e[min(pcr, hi - 1)] >= t_ge) & ballot(pcr < hi);
        }
        wave_sync();
        const int m = min(fl, fr);
        const int l_at = sh.pairs[lane], r_at = sh.pairs[64 + lane];
        const bool in = lane < m;
        const int pl = in ? l_at : first, pr = in ? r_at : first;
        const bool c = in && pl < pr;
        const uint32_t tl = e[pl], tr = e[pr];
        if (c) {
            e[pl] = tr;
            e[pr] = tl;
        }
        const int cnt = __popcll(ballot(c));
        if (cnt < 64) {
            const int a = cnt < fl ? __builtin_amdgcn_readlane(l_at, cnt) : 0x7fffffff;
            const int bq = cnt >= 1 ? __builtin_amdgcn_readlane(r_at, cnt - 1) : r_prev;
            wave_sync();
            return min(a, bq);
        }
        r_prev = __builtin_amdgcn_readlane(r_at, 63);
        wave_sync();
    }
}

// Lane i receives v of lane src_i.
__device__ inline int lane_fetch(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
// Lane dst_i receives v of lane i (dst a permutation of the lanes).
__device__ inline int lane_send(int v, int dst) { return __builtin_amdgcn_ds_permute(dst << 2, v); }

// The introsort steps on e[lo, hi), hi - lo <= 64, and on every range they
// split off that starts below K, with lane i holding e[lo + i] in a
// register.  Each __unguarded_partition is two ballots for the stops, one
// compaction of the left stops and one of the right stops into lane order
// (lane j then holds L_j and R_j), and one fetch of every swapped lane's
// partner.  d is the depth budget of [lo, hi); a range that exhausts it is
// heap-sorted in memory.  Returns false on a state libstdc++'s algorithm
// cannot reach.
template <typename S>
__device__ __forceinline__ bool small_sort_(uint32_t *e, S &sh, int lo, int hi, int d, int K, Keyed key) {
    const int lane = threadIdx.x & 63;
    const int n = hi - lo;
    wave_sync();
    uint32_t v = lane < n ? e[lo + lane] : 0u;
    // pending ranges relative to lo: lane i holds stack entry i
    int s_l = 0, s_h = n, s_d = d;
    int top = 1;
    while (top > 0) {
        --top;
        int l = __builtin_amdgcn_readlane(s_l, top), h = __builtin_amdgcn_readlane(s_h, top),
            dd = __builtin_amdgcn_readlane(s_d, top);
        while (h - l > 16) {
            if (dd == 0) {
                if (lane < n) e[lo + lane] = v;
                wave_sync();
                if (lane == 0) heap_sort_(e + lo + l, h - l, key);
                wave_sync();
                v = lane < n ? e[lo + lane] : 0u;
                break;
            }
            --dd;
            // __move_median_to_first(l, l+1, mid, h-1)
            const int a = l + 1, bm = l + (h - l) / 2, c = h - 1;
            const uint32_t ka = key(__builtin_amdgcn_readlane(v, a)), kb = key(__builtin_amdgcn_readlane(v, bm)),
                           kc = key(__builtin_amdgcn_readlane(v, c));
            int sw;
            if (ka > kb) {
                if (kb > kc)
                    sw = bm;
                else if (ka > kc)
                    sw = c;
                else
                    sw = a;
            } else if (ka > kc) {
                sw = a;
            } else if (kb > kc) {
                sw = c;
            } else {
                sw = bm;
            }
            const uint32_t kp = sw == a ? ka : sw == bm ? kb : kc;
            const uint32_t vl = __builtin_amdgcn_readlane(v, l), vs = __builtin_amdgcn_readlane(v, sw);
            v = lane == l ? vs : lane == sw ? vl : v;
            // left stops key <= kp, right stops key >= kp, in [l + 1, h),
            // by whole elements against thresholds
            const uint32_t t_ge = kp << key.lb, t_le = t_ge | ((1u << key.lb) - 1u);
            const uint64_t range = (h >= 64 ? ~0ull : (1ull << h) - 1) & ~((1ull << (l + 1)) - 1);
            const uint64_t lm = ballot(v <= t_le) & range, rm = ballot(v >= t_ge) & range;
            const int nl = __popcll(lm), nr = __popcll(rm);
            const int bl = bits_below(lm), br = bits_below(rm);
            const bool in_l = (lm >> lane) & 1u, in_r = (rm >> lane) & 1u;
            const int il = bl, jr = nr - 1 - br;  // L index; R index from the top
            const int l_at = lane_send(lane, in_l ? il : nl + lane - bl);
            const int r_at = lane_send(lane, in_r ? jr : nr + lane - br);
            const int sc = __popcll(ballot(lane < min(nl, nr) && l_at < r_at));
            const int to_r = lane_fetch(r_at, il), to_l = lane_fetch(l_at, jr & 63);
            const int src = (in_l && il < sc) ? to_r : (in_r && jr < sc) ? to_l : lane;
            v = (uint32_t)lane_fetch((int)v, src);
            const int cl = sc < nl ? __builtin_amdgcn_readlane(l_at, sc) : 0x7fffffff;
            const int cr = sc >= 1 ? __builtin_amdgcn_readlane(r_at, sc - 1) : 0x7fffffff;
            const int cut = min(cl, cr);
            if (cut <= l || cut >= h || top >= 64) return false;
            if (lo + cut < K) {
                if (lane == top) {
                    s_l = cut;
                    s_h = h;
                    s_d = dd;
                }
                ++top;
            }
            h = cut;
        }
    }
    if (lane < n) e[lo + lane] = v;
    wave_sync();
    return true;
}

// Returns false if the replay reached a state libstdc++'s algorithm cannot
// reach (a cut outside the range or a deeper pending stack than 64
// ranges); the caller then reports an internal error instead of looping.
template <typename S>
__device__ __forceinline__ bool topk_sort_(uint32_t *e, S &sh, int n, int K, Keyed key) {
    const int lane = threadIdx.x & 63;
    if (n <= 1) return true;
    const int lg = 31 - __clz(n);
    // pending ranges: lane i holds stack entry i
    int s_lo = 0, s_hi = n, s_d = 2 * lg;
    int top = 1;
    while (top > 0) {
        --top;
        int lo = __builtin_amdgcn_readlane(s_lo, top), hi = __builtin_amdgcn_readlane(s_hi, top),
            d = __builtin_amdgcn_readlane(s_d, top);
        while (hi - lo > 16) {
            if (hi - lo <= 64) {  // this range and all it splits off, in registers
                if (!small_sort_(e, sh, lo, hi, d, K, key)) return false;
                break;
            }
            if (d == 0) {
                wave_sync();
                if (lane == 0) heap_sort_(e + lo, hi - lo, key);
                wave_sync();
                break;
            }
            --d;
            const int cut = partition_(e, sh, lo, hi, key);
            if (cut <= lo || cut >= hi || top >= 64) return false;
            if (cut < K) {
                if (lane == top) {
                    s_lo = cut;
                    s_hi = hi;
                    s_d = d;
                }
                ++top;
            }
            hi = cut;
        }
    }
    wave_sync();
    return true;
}

// Stable rank sort of e[0, P) by descending key; writes the first K cells,
// decoding each kept element's raw candidate index back to its cell.
// P = min(n, K + 16) <= 80: lane i holds positions i and 64 + i.  When the
// keys leave 7 bits free (local bits >= 7), e[0, P) is overwritten with
// (key << 7) | (127 - position), so that one unsigned compare orders by
// descending key with ties in position order, as the stable sort does.
template <typename E, typename S>
__device__ void finish_(E *e, int n, int K, int *out, const S &sh, const int *bin_cells, Keyed key) {
    const int lane = threadIdx.x & 63;
    const int P = min(n, K + 16);
    const bool two = P > 64;
    const uint32_t el0 = lane < P ? e[lane] : 0u;
    const uint32_t el1 = two && 64 + lane < P ? e[64 + lane] : 0u;
    int r0 = 0, r1 = 0;
    if (key.lb >= 7) {
        const uint32_t c0 = (key(el0) << 7) | (127u - (uint32_t)lane);
        const uint32_t c1 = (key(el1) << 7) | (63u - (uint32_t)lane);
        // each lane rewrites only its own positions, then all read them
        if (lane < P) e[lane] = c0;
        if (two && 64 + lane < P) e[64 + lane] = c1;
        wave_sync();
        if (two) {
#pragma unroll 8
            for (int q = 0; q < P; ++q) {
                const uint32_t cq = e[q];
                r0 += cq > c0;
                r1 += cq > c1;
            }
        } else {
#pragma unroll 8
            for (int q = 0; q < P; ++q) r0 += e[q] > c0;
        }
    } else {
        const uint32_t k0 = key(el0), k1 = key(el1);
        for (int q = 0; q < P; ++q) {
            const uint32_t kq = key(e[q]);
            r0 += (kq > k0) || (kq == k0 && q < lane);
            r1 += (kq > k1) || (kq == k1 && q < 64 + lane);
        }
    }
    const uint32_t lmask = (1u << key.lb) - 1u;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = 64 * h + lane;
        const int r = h ? r1 : r0;
        if (p < P && r < K) {
            // the segment of raw index j: the last one starting at or
            // before j (empty segments start where the next one does)
            const int j = (int)((h ? el1 : el0) & lmask);
            int g = 0;
#pragma unroll
            for (int step = kMaxSegs / 2; step >= 1; step >>= 1)
                if (sh.seg_raw[g + step] <= j) g += step;
            out[r] = bin_cells[sh.seg_off[g] + j];
        }
    }
}

// G consecutive rank-table entries from byte offset off, one vector load
// (offsets are dword aligned; G = 2 and 4 need only dword alignment for
// global loads).
template <int G>
struct RankVec {
    uint32_t v[G];
};
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
template <int G>
__device__ inline RankVec<G> load_ranks(const uint32_t *base, uint32_t off) {
    RankVec<G> r;
    const char *p = reinterpret_cast<const char *>(base) + off;
    if constexpr (G == 1) {
        r.v[0] = *reinterpret_cast<const uint32_t *>(p);
    } else if constexpr (G == 2) {
        const u32x2_a4 w = *reinterpret_cast<const u32x2_a4 *>(p);
        r.v[0] = w.x;
        r.v[1] = w.y;
    } else {
        static_assert(G % 4 == 0, "groups of 1, 2 or a multiple of 4 targets");
#pragma unroll
        for (int k = 0; k < G; k += 4) {
            const u32x4_a4 w = *reinterpret_cast<const u32x4_a4 *>(p + 4 * k);
            r.v[k] = w.x;
            r.v[k + 1] = w.y;
            r.v[k + 2] = w.z;
            r.v[k + 3] = w.w;
        }
    }
    return r;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Gather of one run of nt <= G targets that share their 27 bins: the same
// y, z and x bin, at x = x0 .. x0 + nt - 1.  Their candidates are the same
// cells, and a candidate's rank-table entries for the nt targets are nt
// consecutive entries (the x offset drops by one per target), so one
// coordinate load and one G-entry rank read per lane serve all of them.
//
// The candidates are numbered densely in gather order (raw index j = the
// position in the concatenation of the 21 segments), 64 to a group; a
// group that straddles segment boundaries takes each lane's offset into the
// bin-sorted arrays from the segment its j falls in.  A candidate's entry
// is its packed table coordinate plus the target's base; candidates outside
// the covariance box in x or y (the z trim keeps only the box's layers) are
// dropped by one packed 16-bit test, and read the zero entries before the
// table instead.  Run members whose own offset leaves the box read the
// table's zero columns in x.
//
// Target 0 keeps the candidates with rank > 0 in raw order at e0 (LDS up
// to cap, then the spill scratch ge, if any), target g > 0 at
// gs + (g - 1) cap (global scratch, up to cap); returns the list lengths in
// pos[0 .. nt).
template <int G, bool SPILLED>
__device__ __forceinline__ void gather_run_(const ClusterArgs &a, WaveLds &sh, int x, int y, int z, int nt,
                                            uint32_t *e0, int cap, uint32_t *ge, uint32_t *gs, int *pos) {
    const int lane = threadIdx.x & 63;
    const BinArgs &b = a.bins;
    const int bx = x / b.r[0], by = y / b.r[1], bz = z / b.r[2];
    const int s2 = a.span[2];
    // lane g < kSegs: segment g's first position in the bin-sorted arrays
    // and its length (bins outside the bin box hold no cells; bx + 1 <=
    // m0 - 1 always holds, bx - 1 may be -1)
    int start = 0, cnt = 0;
    if (lane < kSegs) {
        const int plane = lane < 9 ? 0 : lane < 12 ? 1 : 2;
        const int q = plane == 0 ? lane : plane == 1 ? lane - 9 : lane - 12;
        const int kk = bz - 1 + plane;
        const int jj = by - 1 + (plane == 1 ? q : q / 3);
        const int i0 = plane == 1 ? (bx > 0 ? bx - 1 : 0) : bx - 1 + q % 3;
        const int i1 = plane == 1 ? bx + 1 : i0;
        if (jj >= 0 && kk >= 0 && i0 >= 0 && jj < b.m[1] && kk < b.m[2]) {
            const int row = (kk * b.m[1] + jj) * b.m[0];
            start = (int)a.bin_start[row + i0];
            int end = (int)(a.bin_start[row + i1] + a.bin_count[row + i1]);
            const int rz = b.r[2];
            const uint32_t *zo = a.zoff + (size_t)(row + i0) * (rz + 1);
            if (plane == 0) {
                const int d = z - s2 - kk * rz;  // layers of the bin below z - span
                if (d > 0) start += (int)zo[min(d, rz)];
            } else if (plane == 2) {
                const int d = z + s2 + 1 - kk * rz;  // layers of the bin up to z + span
                if (d < rz) end = start + (int)zo[max(d, 0)];
            }
            cnt = max(end - start, 0);
        }
    }
    // segment g holds raw indices [d0, d0 + cnt)
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < kMaxSegs; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const int d0 = incl - cnt;
    const int total = __builtin_amdgcn_readlane(incl, kSegs - 1);
    // lanes from kSegs on are sentinels that no raw index reaches
    const int seg_d = lane < kSegs ? d0 : 0x7fffffff;
    const int seg_o = start - d0;  // raw index -> position in the bin-sorted arrays
    if (lane < kMaxSegs) {
        sh.seg_raw[lane] = seg_d;
        sh.seg_off[lane] = seg_o;
    }
    // byte offset of target 0's entries, less G - 1 entries: the load for
    // the run starts at the entry of its last member (x offset smaller by
    // G - 1); x + w0 y + w01 z modulo 2^32, as bin_lin
    const uint32_t w0 = (uint32_t)a.w[0], w01 = (uint32_t)a.w[0] * (uint32_t)a.w[1];
    const uint32_t tb = a.centre4 - 4u * ((uint32_t)x + w0 * (uint32_t)y + w01 * (uint32_t)z) - 4u * (G - 1);
    // box test: (cx - (x - span0), cy - (y - span1)) modulo 2^16 must not
    // exceed (2 span0 + nt - 1, 2 span1) (offsets of a run's union)
    const u16x2 lo_xy = {(unsigned short)(x - a.span[0]), (unsigned short)(y - a.span[1])};
    const u16x2 box_xy = {(unsigned short)(2 * a.span[0] + nt - 1), (unsigned short)(2 * a.span[1])};
    const uint32_t lane8 = 8u * (uint32_t)lane;
#pragma unroll
    for (int g = 0; g < G; ++g) pos[g] = 0;
    // the segment cursor: cur = offset of the segment holding the next
    // group's first raw index; (dn, on) = start and offset of the next
    // segment not yet entered
    // groups whose loads are in flight together (fewer for long runs, whose
    // rank reads take more registers)
    constexpr int kGatherUnroll = G >= 8 ? 2 : kGatherUnrollMax;
    int nx = 1;
    int cur = __builtin_amdgcn_readlane(seg_o, 0);
    int dn = __builtin_amdgcn_readlane(seg_d, 1), on = __builtin_amdgcn_readlane(seg_o, 1);
    for (int base = 0; base < total; base += 64 * kGatherUnroll) {
        uint32_t boff[kGatherUnroll];
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) {
            const int gb = base + u * 64;
            while (dn <= gb) {
                cur = on;
                ++nx;
                dn = __builtin_amdgcn_readlane(seg_d, nx);
                on = __builtin_amdgcn_readlane(seg_o, nx);
            }
            int off = cur;
            while (dn < gb + 64) {  // segments that start inside this group
                off = gb + lane >= dn ? on : off;
                cur = on;
                ++nx;
                dn = __builtin_amdgcn_readlane(seg_d, nx);
                on = __builtin_amdgcn_readlane(seg_o, nx);
            }
            // position off + j of the bin-sorted arrays, in bytes; lanes past
            // the last candidate read at most kClusterLinPad - 1 entries past
            // the end
            boff[u] = 8u * (uint32_t)(off + gb) + lane8;
        }
        uint2 lin[kGatherUnroll];
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u)
            lin[u] = *reinterpret_cast<const uint2 *>(reinterpret_cast<const char *>(a.bin_lin) + boff[u]);
        // both words in registers once all the groups' loads are issued
        // (left alone, the compiler splits the load and defers the first
        // word behind a branch on the box test)
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) asm volatile("" : "+v"(lin[u].x), "+v"(lin[u].y));
        RankVec<G> rv[kGatherUnroll];
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) {
            const u16x2 cxy = __builtin_bit_cast(u16x2, lin[u].y);
            const u16x2 d = cxy - lo_xy;
            const bool in_box =
                __builtin_bit_cast(uint32_t, __builtin_elementwise_min(d, box_xy)) == __builtin_bit_cast(uint32_t, d);
            // lanes past the last candidate or outside the box read the zero
            // entries before the table
            const bool valid = base + u * 64 + lane < total && in_box;
            rv[u] = load_ranks<G>(a.rank, valid ? lin[u].x + tb : 0u);
        }
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) {
            const uint32_t j = (uint32_t)(base + u * 64 + lane);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g >= nt) continue;
                // the table holds the ranks shifted into place
                const uint32_t rk = rv[u].v[G - 1 - g];
                const uint64_t mk = ballot(rk != 0);
                const int pp = bits_below(mk, pos[g]);
                const uint32_t ev = rk | j;
                if (SPILLED) {
                    if (rk != 0 && pp < a.scratch_per_wave) ge[pp] = ev;
                } else {
                    // every lane stores to LDS: lanes that keep nothing, or
                    // past the capacity (their list is spilled), store to
                    // their own dead slot, so the store needs no branch; the
                    // scratch stores stay masked (fewer lanes, fewer lines)
                    const bool keep = rk != 0 && pp < cap;
                    if (g == 0) {
                        uint32_t *dead = reinterpret_cast<uint32_t *>(sh.pairs) + 128 + lane;
                        *(keep ? e0 + pp : dead) = ev;
                    } else if (keep) {
                        // 32-bit byte offsets from the wavefront's scratch base
                        *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(gs) +
                                                      (((uint32_t)((g - 1) * cap) + (uint32_t)pp) << 2)) = ev;
                    }
                }
                pos[g] += __popcll(mk);
            }
        }
    }
}

// W wavefronts per workgroup; the wavefronts loop over the target list.
//
// SPILLED = false: as many wavefronts as are resident at once take chunks
// of per_wave list entries from the work counters, gather each run of
// targets that share their 27 bins (same y, z and x bin, consecutive x)
// together, then sort the runs' lists one by one; a list longer than a.cap
// is not sorted but its target appended to spill_list.  SPILLED = true
// (G = 1): the wavefronts stride over the spill_count targets of
// spill_list, sorting all their lists in their global scratch.
//
// WAVES8: the register budget of eight wavefronts per SIMD (a few SGPRs
// move to VGPR lanes), used when the LDS capacity allows 32 wavefronts per
// CU; otherwise the budget of seven, which LDS already limits to.
//
// W wavefronts per workgroup take consecutive chunks (neighbouring targets
// along x): their rank-table reads, offsets shifted by G, share the CU's L1
// lines.
template <bool SPILLED, bool WAVES8, int W, int G>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(WAVES8 ? 8 : 1, 8))) void k_cluster(
    ClusterArgs a) {
    static_assert(!SPILLED || G == 1, "the spill pass sorts one target per wavefront");
    __shared__ WaveLds sh_all[W];
    extern __shared__ __attribute__((aligned(16))) uint32_t lds_all[];
    // the spill pass gathers straight into its global scratch
    const int cap = SPILLED ? 0 : a.cap;
    const int w = W == 1 ? 0 : (int)(threadIdx.x >> 6);
    WaveLds &sh = sh_all[w];
    uint32_t *const e0 = lds_all + (size_t)w * cap;
    const int wave_id = blockIdx.x * W + w;
    // the later members' lists of a run; a wavefront without a scratch slot
    // takes its targets one at a time
    uint32_t *const gs = !SPILLED && G > 1 && wave_id < a.run_slots
                             ? a.run_scratch + (size_t)wave_id * cluster_run_words(cap)
                             : nullptr;
    const int lane = threadIdx.x & 63;
    const int X = a.X, Y = a.Y;
    const Keyed key{a.local_bits};
    // global scratch of this wave: the elements of lists longer than a.cap
    uint32_t *ge = SPILLED ? a.scratch + (size_t)wave_id * cluster_scratch_words(a.scratch_per_wave) : nullptr;
    unsigned long long cand_acc = 0, spill_acc = 0;
    const int li_end = SPILLED ? *a.spill_count : a.n_list;
    // list entries [ent0, ent0 + ne), ne <= G
    auto entries = [&](int ent0, int ne) {
        // lane g < ne: entry ent0 + g's target and cell
        int my_t = 0, my_node = 0;
        if (lane < ne) {
            const int li = ent0 + lane;
            my_t = SPILLED ? a.spill_list[li] : a.list ? a.list[li] : li;
            my_node = a.targets[my_t];
        }
        for (int r0 = 0; r0 < ne;) {
            const int node0 = __builtin_amdgcn_readlane(my_node, r0);
            const int x = node0 % X, y = node0 % (X * Y) / X, z = node0 / (X * Y);
            const int bx = x / a.bins.r[0];
            int nt = 1;
            while (gs && r0 + nt < ne) {
                const int xn = x + nt;
                if (__builtin_amdgcn_readlane(my_node, r0 + nt) != node0 + nt || xn >= X || xn / a.bins.r[0] != bx)
                    break;
                ++nt;
            }
            int pos[G];
            if (G > 1 && nt > 1)
                gather_run_<G, SPILLED>(a, sh, x, y, z, nt, e0, cap, ge, gs, pos);
            else
                gather_run_<1, SPILLED>(a, sh, x, y, z, 1, e0, cap, ge, gs, pos);
            wave_sync();
            for (int g = 0; g < nt; ++g) {
                const int t = __builtin_amdgcn_readlane(my_t, r0 + g);
                int total = pos[0];
#pragma unroll
                for (int k = 1; k < G; ++k)
                    if (k == g) total = pos[k];
                if (g > 0 && total <= cap) {
                    // a later member's list from the run scratch into LDS
                    // four elements per lane and instruction: both lists start
                    // 16-byte aligned and hold cap (a multiple of 64) words,
                    // so the last quadruple stays inside them
                    const uint32_t *src = gs + (size_t)(g - 1) * cap;
                    for (int i = 4 * lane; i < total; i += 256)
                        *reinterpret_cast<uint4 *>(e0 + i) = *reinterpret_cast<const uint4 *>(src + i);
                    wave_sync();
                }
                uint32_t *const eg = e0;
                int *out = a.nbr + (size_t)t * a.max_n;
                bool sane = true;
                bool spilled = false;
                if (total <= cap) {
                    sane = topk_sort_(eg, sh, total, a.max_n, key);
                    if (sane) finish_(eg, total, a.max_n, out, sh, a.bin_cells, key);
                } else if (!SPILLED && a.spill_list) {
                    if (lane == 0) a.spill_list[atomicAdd(a.spill_count, 1)] = t;
                    spilled = true;  // counted by the spill pass
                } else if (SPILLED && total <= a.scratch_per_wave) {
                    spill_acc += 1;
                    wave_sync();
                    sane = topk_sort_(ge, sh, total, a.max_n, key);
                    if (sane) finish_(ge, total, a.max_n, out, sh, a.bin_cells, key);
                } else {
                    if (lane == 0) atomicOr(a.overflow, 1u);
                    sane = false;
                }
                if (lane == 0 && !spilled) {
                    cand_acc += (unsigned long long)total;
                    a.cnt[t] = sane ? min(total, a.max_n) : 0;
                    if (!sane) atomicOr(a.overflow, 2u);
                }
                wave_sync();
            }
            r0 += nt;
        }
    };
    if constexpr (SPILLED) {
        for (int li = wave_id; li < li_end; li += gridDim.x * W) entries(li, 1);
    } else {
        // chunks of per_wave entries (a multiple of G) from the work
        // counters: part p holds chunks [n_ch p / P, n_ch (p + 1) / P); a
        // wavefront starts at part blockIdx mod P and moves to the next part
        // when one is exhausted, so all finish within about one chunk
        const int chunk = a.per_wave;
        const int n_ch = (li_end + chunk - 1) / chunk;
        int tries = 0;
        for (;;) {
            int ch = -1;
            while (tries < kClusterWorkParts) {
                const int p = ((int)blockIdx.x + tries) % kClusterWorkParts;
                const int c0 = (int)((long long)n_ch * p / kClusterWorkParts);
                const int c1 = (int)((long long)n_ch * (p + 1) / kClusterWorkParts);
                int c = 0;
                if (lane == 0) c = atomicAdd(a.work + p * kClusterWorkStride, 1);
                c = __builtin_amdgcn_readfirstlane(c);
                if (c < c1 - c0) {
                    ch = c0 + c;
                    break;
                }
                ++tries;
            }
            if (ch < 0) break;
            const int end = min(li_end, (ch + 1) * chunk);
            for (int ent0 = ch * chunk; ent0 < end; ent0 += G) entries(ent0, min(G, end - ent0));
        }
    }
    if (lane == 0 && a.cand_total && cand_acc != 0) atomicAdd(counter_line(a.cand_total), cand_acc);
    if (lane == 0 && a.cand_total && spill_acc != 0) atomicAdd(counter_line(a.cand_total) + 1, spill_acc);
}

}  // namespace

// zoff[b (rz + 1) + d] = cells of bin b with z below the bin's first layer
// + d, d = 0 .. rz (a bin's cells ascend by cell index, hence by z).
static __global__ void k_bin_zoff(const uint32_t *bin_start, const uint32_t *bin_count, const int *bin_cells, int XY,
                           BinArgs b, long long n, uint32_t *zoff) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int rz = b.r[2];
    const long long bin = i / (rz + 1);
    const int d = (int)(i - bin * (rz + 1));
    const int zlim = (int)(bin / ((long long)b.m[0] * b.m[1])) * rz + d;
    const uint32_t s = bin_start[bin];
    uint32_t lo = 0, hi = bin_count[bin];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bin_cells[s + mid] / XY < zlim)
            lo = mid + 1;
        else
            hi = mid;
    }
    zoff[i] = lo;
}

void launch_bin_zoff(const uint32_t *bin_start, const uint32_t *bin_count, const int *bin_cells, const BinArgs &b,
                     long long nbins, uint32_t *zoff, hipStream_t s) {
    const long long n = nbins * (b.r[2] + 1);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bin_zoff, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bin_start, bin_count, bin_cells,
                       b.X * b.Y, b, n, zoff);
}

int cluster_capacity(long long expected_max) {
    const long long c = (expected_max + 63) / 64 * 64;
    return (int)std::min<long long>(kMaxCap, std::max<long long>(kMinCap, c));
}

// Wavefronts per CU that the LDS holds with one list of cap elements each.
static int waves_per_cu(int cap) { return (int)(kLdsPerCu / ((size_t)cap * sizeof(uint32_t) + sizeof(WaveLds))); }

// The main pass's kernel and dynamic LDS for these arguments: runs of up to
// kTargetsPerRun targets, the eight-wavefront register budget when 32
// wavefronts fit the LDS.
struct MainPass {
    const void *kern;
    size_t lds;
    bool w8;
};
static MainPass main_pass(const ClusterArgs &a) {
    MainPass m;
    m.w8 = waves_per_cu(a.cap) >= 32;
    m.lds = kWavesPerGroup * (size_t)a.cap * sizeof(uint32_t);
    m.kern = m.w8 ? reinterpret_cast<const void *>(k_cluster<false, true, kWavesPerGroup, kTargetsPerRun>)
                  : reinterpret_cast<const void *>(k_cluster<false, false, kWavesPerGroup, kTargetsPerRun>);
    return m;
}

// The grid the device holds at once, and at most one workgroup per chunk
// of the list.
static long long main_grid(const ClusterArgs &a, const MainPass &m) {
    constexpr int W = kWavesPerGroup;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m.kern, 64 * W, m.lds) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        throw std::runtime_error("cluster search: occupancy query failed");
    const long long chunks = ((long long)a.n_list + a.per_wave - 1) / a.per_wave;
    return std::max(1ll, std::min((long long)std::max(per_cu, 1) * cus, (chunks + W - 1) / W));
}

long long cluster_main_waves(const ClusterArgs &a) { return main_grid(a, main_pass(a)) * kWavesPerGroup; }

void launch_cluster_search(const ClusterArgs &a, hipStream_t s) {
    if (a.n_list <= 0) return;
    if (a.per_wave < 1 || a.cap < kMinCap || a.cap > kMaxCap || a.cap % 64 != 0 || !a.bin_lin || !a.work ||
        a.per_wave % kTargetsPerRun != 0 || (a.run_slots > 0 && !a.run_scratch) ||
        (a.spill_list && (a.n_spill_waves < 1 || !a.spill_count || !a.scratch)))
        throw std::runtime_error("cluster search: invalid launch arguments");
    const MainPass m = main_pass(a);
    const long long grid = main_grid(a, m);
    constexpr int W = kWavesPerGroup, G = kTargetsPerRun;
    if (m.w8)
        hipLaunchKernelGGL((k_cluster<false, true, W, G>), dim3((unsigned)grid), dim3(64 * W), m.lds, s, a);
    else
        hipLaunchKernelGGL((k_cluster<false, false, W, G>), dim3((unsigned)grid), dim3(64 * W), m.lds, s, a);
    if (a.spill_list)
        hipLaunchKernelGGL((k_cluster<true, false, 1, 1>), dim3(a.n_spill_waves), dim3(64), 0, s, a);
}

}  // namespace hpgl_amd
