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
//    (RankTable), so comparisons are those of the reference;
//  * the introsort is replayed step by step, one wavefront per target, with
//    each __unguarded_partition done in parallel.  In that partition the
//    k-th left stop (first elements with key <= pivot, scanning up) is
//    swapped with the k-th right stop (key >= pivot, scanning down) while
//    the left stop is below the right stop, and the cut is
//    min(l_{s+1}, r_s) for s swaps.  Left/right stop positions come from
//    ballot prefix counts, s from one comparison pass, and the s swaps are
//    disjoint so they run in parallel;
//  * ranges that start at or beyond max_n are never partitioned: they
//    cannot change the first max_n positions, which after the final
//    insertion sort are the stable sort of the partition blocks they fall
//    in.  The first min(n, max_n + 16) positions are therefore finished by a
//    stable rank sort;
//  * a range whose depth budget (2 floor(log2 n)) runs out is heap-sorted
//    by one lane with libstdc++'s make_heap/sort_heap steps.
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

constexpr int kCap = 2048;      // candidates held in LDS per wavefront
constexpr int kStack = 64;

__device__ inline uint32_t key_of(uint64_t e) { return (uint32_t)(e >> 32); }

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

struct WaveLds {
    uint64_t e[kCap];
    uint16_t L[kCap];
    uint16_t R[kCap];
    int stk_lo[kStack], stk_hi[kStack], stk_d[kStack];
};

// libstdc++ heap steps (stl_heap.h) on e[first, last) with
// comp(a, b) = key(a) > key(b).  Run by a single lane.
template <typename E>
__device__ void push_heap_(E *e, long hole, long top, uint64_t value) {
    long parent = (hole - 1) / 2;
    while (hole > top && key_of(e[parent]) > key_of(value)) {
        e[hole] = e[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    e[hole] = value;
}
template <typename E>
__device__ void adjust_heap_(E *e, long hole, long len, uint64_t value) {
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (key_of(e[second]) > key_of(e[second - 1])) second--;
        e[hole] = e[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        e[hole] = e[second - 1];
        hole = second - 1;
    }
    push_heap_(e, hole, top, value);
}
template <typename E>
__device__ void heap_sort_(E *e, long len) {
    if (len >= 2) {  // __make_heap
        long parent = (len - 2) / 2;
        while (true) {
            uint64_t v = e[parent];
            adjust_heap_(e, parent, len, v);
            if (parent == 0) break;
            parent--;
        }
    }
    long last = len;  // __sort_heap
    while (last > 1) {
        --last;
        uint64_t v = e[last];
        e[last] = e[0];
        adjust_heap_(e, 0, last, v);
    }
}

// One __unguarded_partition_pivot on e[lo, hi); returns the cut.
template <typename E, typename P>
__device__ int partition_(E *e, P *Lp, P *Rp, int lo, int hi) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    // __move_median_to_first(lo, lo+1, mid, hi-1)
    {
        const int a = lo + 1, b = lo + (hi - lo) / 2, c = hi - 1;
        const uint32_t ka = key_of(e[a]), kb = key_of(e[b]), kc = key_of(e[c]);
        int sw;
        if (ka > kb) {
            if (kb > kc)
                sw = b;
            else if (ka > kc)
                sw = c;
            else
                sw = a;
        } else if (ka > kc) {
            sw = a;
        } else if (kb > kc) {
            sw = c;
        } else {
            sw = b;
        }
        __syncthreads();
        if (lane == 0) {
            const uint64_t t0 = e[lo];
            e[lo] = e[sw];
            e[sw] = t0;
        }
        __syncthreads();
    }
    const uint32_t kp = key_of(e[lo]);
    // totals of left stops (key <= kp) and right stops (key >= kp)
    int nL = 0, nR = 0;
    for (int base = lo + 1; base < hi; base += 64) {
        const int p = base + lane;
        const bool v = p < hi;
        const uint32_t k = v ? key_of(e[p]) : 0u;
        nL += __popcll(__ballot(v && k <= kp));
        nR += __popcll(__ballot(v && k >= kp));
    }
    // stop positions by rank: Lp[k] = l_{k+1}, Rp[k] = r_{k+1}
    int cL = 0, cR = 0;
    for (int base = lo + 1; base < hi; base += 64) {
        const int p = base + lane;
        const bool v = p < hi;
        const uint32_t k = v ? key_of(e[p]) : 0u;
        const bool lf = v && k <= kp, rf = v && k >= kp;
        const uint64_t mL = __ballot(lf), mR = __ballot(rf);
        if (lf) Lp[cL + __popcll(mL & lt)] = (P)p;
        if (rf) Rp[nR - cR - __popcll(mR & lt) - 1] = (P)p;
        cL += __popcll(mL);
        cR += __popcll(mR);
    }
    __syncthreads();
    // s = number of k with l_k < r_k (a prefix of k, since l rises and r falls)
    const int m = min(nL, nR);
    int s = 0;
    for (int base = 0; base < m; base += 64) {
        const int k = base + lane;
        const bool c = k < m && (int)Lp[k] < (int)Rp[k];
        s += __popcll(__ballot(c));
    }
    const int a = s < nL ? (int)Lp[s] : 0x7fffffff;
    const int b = s >= 1 ? (int)Rp[s - 1] : 0x7fffffff;
    const int cut = min(a, b);
    for (int base = 0; base < s; base += 64) {
        const int k = base + lane;
        if (k < s) {
            const int pl = Lp[k], pr = Rp[k];
            const uint64_t tl = e[pl];
            e[pl] = e[pr];
            e[pr] = tl;
        }
    }
    __syncthreads();
    return cut;
}

// Returns false if the replay reached a state libstdc++'s algorithm cannot
// reach (a cut outside the range or a deeper pending stack than max_n
// ranges); the caller then reports an internal error instead of looping.
template <typename E, typename P>
__device__ bool topk_sort_(E *e, P *Lp, P *Rp, int *slo, int *shi, int *sd, int n, int K) {
    const int lane = threadIdx.x & 63;
    if (n <= 1) return true;
    int lg = 31 - __clz(n);
    int top = 0;
    if (lane == 0) {
        slo[0] = 0;
        shi[0] = n;
        sd[0] = 2 * lg;
    }
    top = 1;
    __syncthreads();
    while (top > 0) {
        --top;
        int lo = slo[top], hi = shi[top], d = sd[top];
        __syncthreads();
        while (hi - lo > 16) {
            if (d == 0) {
                if (lane == 0) heap_sort_(e + lo, hi - lo);
                __syncthreads();
                break;
            }
            --d;
            const int cut = partition_(e, Lp, Rp, lo, hi);
            if (cut <= lo || cut >= hi || top >= kStack) return false;
            if (cut < K) {
                if (lane == 0) {
                    slo[top] = cut;
                    shi[top] = hi;
                    sd[top] = d;
                }
                ++top;
                __syncthreads();
            }
            hi = cut;
        }
    }
    return true;
}

// Stable rank sort of e[0, P) by descending key; writes the first K cells.
template <typename E>
__device__ void finish_(const E *e, int n, int K, int *out) {
    const int lane = threadIdx.x & 63;
    const int P = min(n, K + 16);
    for (int base = 0; base < P; base += 64) {
        const int p = base + lane;
        if (p < P) {
            const uint64_t ep = e[p];
            const uint32_t kp = key_of(ep);
            int r = 0;
            for (int q = 0; q < P; ++q) {
                const uint32_t kq = key_of(e[q]);
                r += (kq > kp) || (kq == kp && q < p);
            }
            if (r < K) out[r] = (int)(uint32_t)ep;
        }
    }
}

__global__ __launch_bounds__(64) void k_cluster(ClusterArgs a) {
    __shared__ WaveLds sh;
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int X = a.X, Y = a.Y;
    const BinArgs &b = a.bins;
    for (int li = blockIdx.x; li < a.n_list; li += gridDim.x) {
        const int t = a.list ? a.list[li] : li;
        const int node = a.targets[t];
        const int x = node % X, y = node % (X * Y) / X, z = node / (X * Y);
        const int bx = x / b.r[0], by = y / b.r[1], bz = z / b.r[2];
        // pass 1: count candidates with rank > 0
        int total = 0;
        for (int kk = bz - 1; kk <= bz + 1; ++kk)
            for (int jj = by - 1; jj <= by + 1; ++jj)
                for (int ii = bx - 1; ii <= bx + 1; ++ii) {
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= b.m[0] || jj >= b.m[1] || kk >= b.m[2]) continue;
                    const int bin = (kk * b.m[1] + jj) * b.m[0] + ii;
                    const int start = (int)a.bin_start[bin], cnt = (int)a.bin_count[bin];
                    for (int base = 0; base < cnt; base += 64) {
                        const int j = base + lane;
                        bool keep = false;
                        if (j < cnt) {
                            const int c = a.bin_cells[start + j];
                            const int dx = c % X - x, dy = c % (X * Y) / X - y, dz = c / (X * Y) - z;
                            if (abs(dx) <= a.span[0] && abs(dy) <= a.span[1] && abs(dz) <= a.span[2])
                                keep = a.rank[((size_t)(dz + a.span[2]) * a.rd[1] + (dy + a.span[1])) * a.rd[0] +
                                              (dx + a.span[0])] != 0;
                        }
                        total += __popcll(__ballot(keep));
                    }
                }
        const bool in_lds = total <= kCap;
        uint64_t *ge = nullptr;
        int *gL = nullptr, *gR = nullptr;
        if (!in_lds) {
            if (total > a.scratch_per_wave || (int)blockIdx.x >= a.n_waves_scratch) {
                if (lane == 0) {
                    atomicOr(a.overflow, 1u);
                    a.cnt[t] = 0;
                }
                continue;
            }
            ge = a.scratch + (size_t)blockIdx.x * a.scratch_per_wave * 2;
            gL = (int *)(ge + a.scratch_per_wave);
            gR = gL + a.scratch_per_wave;
        }
        // pass 2: write candidates in gather order
        int pos = 0;
        for (int kk = bz - 1; kk <= bz + 1; ++kk)
            for (int jj = by - 1; jj <= by + 1; ++jj)
                for (int ii = bx - 1; ii <= bx + 1; ++ii) {
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= b.m[0] || jj >= b.m[1] || kk >= b.m[2]) continue;
                    const int bin = (kk * b.m[1] + jj) * b.m[0] + ii;
                    const int start = (int)a.bin_start[bin], cnt = (int)a.bin_count[bin];
                    for (int base = 0; base < cnt; base += 64) {
                        const int j = base + lane;
                        uint32_t rk = 0;
                        int c = 0;
                        if (j < cnt) {
                            c = a.bin_cells[start + j];
                            const int dx = c % X - x, dy = c % (X * Y) / X - y, dz = c / (X * Y) - z;
                            if (abs(dx) <= a.span[0] && abs(dy) <= a.span[1] && abs(dz) <= a.span[2])
                                rk = a.rank[((size_t)(dz + a.span[2]) * a.rd[1] + (dy + a.span[1])) * a.rd[0] +
                                            (dx + a.span[0])];
                        }
                        const uint64_t m = __ballot(rk != 0);
                        if (rk != 0) {
                            const uint64_t ev = ((uint64_t)rk << 32) | (uint32_t)c;
                            const int p = pos + __popcll(m & lt);
                            if (in_lds)
                                sh.e[p] = ev;
                            else
                                ge[p] = ev;
                        }
                        pos += __popcll(m);
                    }
                }
        __syncthreads();
        int *out = a.nbr + (size_t)t * a.max_n;
        bool sane;
        if (in_lds) {
            sane = topk_sort_(sh.e, sh.L, sh.R, sh.stk_lo, sh.stk_hi, sh.stk_d, total, a.max_n);
            if (sane) finish_(sh.e, total, a.max_n, out);
        } else {
            sane = topk_sort_(ge, gL, gR, sh.stk_lo, sh.stk_hi, sh.stk_d, total, a.max_n);
            if (sane) finish_(ge, total, a.max_n, out);
        }
        if (lane == 0) {
            if (a.cand_total) atomicAdd(a.cand_total, (unsigned long long)total);
            a.cnt[t] = sane ? min(total, a.max_n) : 0;
            if (!sane) atomicOr(a.overflow, 2u);
        }
        __syncthreads();
    }
}

}  // namespace

int cluster_lds_capacity() { return kCap; }

void launch_cluster_search(const ClusterArgs &a, hipStream_t s) {
    if (a.n_list <= 0) return;
    const int grid = a.n_list < a.n_waves_scratch ? a.n_list : a.n_waves_scratch;
    hipLaunchKernelGGL(k_cluster, dim3(grid), dim3(64), 0, s, a);
}

}  // namespace hpgl_amd
