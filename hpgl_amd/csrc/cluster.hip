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
//    min(l_{s+1}, r_s) for s swaps.  One ballot pass records the stop
//    positions, a second pass over k performs the (disjoint) swaps;
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

namespace hpgl_amd {

namespace {

constexpr int kCap = 2048;  // candidates held in LDS per wavefront
constexpr int kStack = 64;

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

struct WaveLds {
    uint32_t e[kCap];
    uint16_t L[kCap];
    uint16_t R[kCap];
    int stk_lo[kStack], stk_hi[kStack], stk_d[kStack];
    int bin_start[27];
    int bin_cnt[27];
    int bin_pref[28];
};

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

// One __unguarded_partition_pivot on e[lo, hi); returns the cut.
// Lp[k] = position of the (k+1)-th left stop counted from lo+1 upwards;
// Rq[k] = position of the (k+1)-th right stop counted from lo+1 upwards,
// so the (k+1)-th right stop counted from hi-1 downwards is Rq[nR-1-k].
template <typename E, typename P>
__device__ int partition_(E *e, P *Lp, P *Rq, int lo, int hi, Keyed key) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    {  // __move_median_to_first(lo, lo+1, mid, hi-1)
        const int a = lo + 1, b = lo + (hi - lo) / 2, c = hi - 1;
        const uint32_t ka = key(e[a]), kb = key(e[b]), kc = key(e[c]);
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
            const E t0 = e[lo];
            e[lo] = e[sw];
            e[sw] = t0;
        }
        __syncthreads();
    }
    const uint32_t kp = key(e[lo]);
    int nL = 0, nR = 0;
    int base = lo + 1;
    for (; base + 64 <= hi; base += 64) {
        const uint32_t k = key(e[base + lane]);
        const bool lf = k <= kp, rf = k >= kp;
        const uint64_t mL = __ballot(lf), mR = __ballot(rf);
        if (lf) Lp[nL + __popcll(mL & lt)] = (P)(base + lane);
        if (rf) Rq[nR + __popcll(mR & lt)] = (P)(base + lane);
        nL += __popcll(mL);
        nR += __popcll(mR);
    }
    if (base < hi) {
        const int p = base + lane;
        const bool v = p < hi;
        const uint32_t k = v ? key(e[p]) : 0u;
        const bool lf = v && k <= kp, rf = v && k >= kp;
        const uint64_t mL = __ballot(lf), mR = __ballot(rf);
        if (lf) Lp[nL + __popcll(mL & lt)] = (P)p;
        if (rf) Rq[nR + __popcll(mR & lt)] = (P)p;
        nL += __popcll(mL);
        nR += __popcll(mR);
    }
    __syncthreads();
    // swap l_k <-> r_k while l_k < r_k; s = number of swaps (a prefix of k)
    const int m = min(nL, nR);
    int s = 0;
    for (int kb = 0; kb < m; kb += 64) {
        const int k = kb + lane;
        bool c = false;
        int pl = 0, pr = 0;
        if (k < m) {
            pl = (int)Lp[k];
            pr = (int)Rq[nR - 1 - k];
            c = pl < pr;
        }
        if (c) {
            const E tl = e[pl];
            e[pl] = e[pr];
            e[pr] = tl;
        }
        const int cnt = __popcll(__ballot(c));
        s += cnt;
        if (cnt < 64) break;
    }
    const int a = s < nL ? (int)Lp[s] : 0x7fffffff;
    const int b = s >= 1 ? (int)Rq[nR - s] : 0x7fffffff;
    __syncthreads();
    return min(a, b);
}

// Returns false if the replay reached a state libstdc++'s algorithm cannot
// reach (a cut outside the range or a deeper pending stack than max_n
// ranges); the caller then reports an internal error instead of looping.
template <typename E, typename P>
__device__ bool topk_sort_(E *e, P *Lp, P *Rq, WaveLds &sh, int n, int K, Keyed key) {
    const int lane = threadIdx.x & 63;
    if (n <= 1) return true;
    const int lg = 31 - __clz(n);
    if (lane == 0) {
        sh.stk_lo[0] = 0;
        sh.stk_hi[0] = n;
        sh.stk_d[0] = 2 * lg;
    }
    int top = 1;
    __syncthreads();
    while (top > 0) {
        --top;
        int lo = sh.stk_lo[top], hi = sh.stk_hi[top], d = sh.stk_d[top];
        __syncthreads();
        while (hi - lo > 16) {
            if (d == 0) {
                if (lane == 0) heap_sort_(e + lo, hi - lo, key);
                __syncthreads();
                break;
            }
            --d;
            const int cut = partition_(e, Lp, Rq, lo, hi, key);
            if (cut <= lo || cut >= hi || top >= kStack) return false;
            if (cut < K) {
                if (lane == 0) {
                    sh.stk_lo[top] = cut;
                    sh.stk_hi[top] = hi;
                    sh.stk_d[top] = d;
                }
                ++top;
                __syncthreads();
            }
            hi = cut;
        }
    }
    return true;
}

// Stable rank sort of e[0, P) by descending key; writes the first K cells,
// decoding each kept element's raw candidate index back to its cell.
template <typename E>
__device__ void finish_(const E *e, int n, int K, int *out, const WaveLds &sh, const int *bin_cells, Keyed key) {
    const int lane = threadIdx.x & 63;
    const int P = min(n, K + 16);
    const uint32_t lmask = (1u << key.lb) - 1u;
    for (int base = 0; base < P; base += 64) {
        const int p = base + lane;
        if (p < P) {
            const uint32_t ep = e[p];
            const uint32_t kp = key(ep);
            int r = 0;
            for (int q = 0; q < P; ++q) {
                const uint32_t kq = key(e[q]);
                r += (kq > kp) || (kq == kp && q < p);
            }
            if (r < K) {
                const int j = (int)(ep & lmask);
                int b = 0;
                while (sh.bin_pref[b + 1] <= j) ++b;
                out[r] = bin_cells[sh.bin_start[b] + (j - sh.bin_pref[b])];
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_cluster(ClusterArgs a) {
    __shared__ WaveLds sh;
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int X = a.X, Y = a.Y;
    const BinArgs &b = a.bins;
    const Keyed key{a.local_bits};
    uint32_t *ge = a.scratch ? a.scratch + (size_t)blockIdx.x * a.scratch_per_wave * 3 : nullptr;
    unsigned long long cand_acc = 0;
    for (int li = blockIdx.x; li < a.n_list; li += gridDim.x) {
        const int t = a.list ? a.list[li] : li;
        const int node = a.targets[t];
        const int x = node % X, y = node % (X * Y) / X, z = node / (X * Y);
        const int bx = x / b.r[0], by = y / b.r[1], bz = z / b.r[2];
        // the 27 bins in gather order (k, j, i ascending) and their offsets
        // in the raw concatenation; bins outside the bin box hold 0 cells
        if (lane < 27) {
            const int ii = bx - 1 + lane % 3, jj = by - 1 + (lane / 3) % 3, kk = bz - 1 + lane / 9;
            int start = 0, cnt = 0;
            if (ii >= 0 && jj >= 0 && kk >= 0 && ii < b.m[0] && jj < b.m[1] && kk < b.m[2]) {
                const int bin = (kk * b.m[1] + jj) * b.m[0] + ii;
                start = (int)a.bin_start[bin];
                cnt = (int)a.bin_count[bin];
            }
            sh.bin_start[lane] = start;
            sh.bin_cnt[lane] = cnt;
        }
        __syncthreads();
        if (lane == 0) {
            int acc = 0;
            for (int q = 0; q < 27; ++q) {
                sh.bin_pref[q] = acc;
                acc += sh.bin_cnt[q];
            }
            sh.bin_pref[27] = acc;
        }
        __syncthreads();
        // gather: keep candidates with rank > 0, in raw order
        int pos = 0;
        for (int q = 0; q < 27; ++q) {
            const int start = sh.bin_start[q], cnt = sh.bin_cnt[q], j0 = sh.bin_pref[q];
            for (int jb = 0; jb < cnt; jb += 64) {
                const int j = jb + lane;
                uint32_t rk = 0;
                if (j < cnt) {
                    int dx, dy, dz;
                    if (a.bin_xyz) {
                        const uint32_t v = a.bin_xyz[start + j];
                        dx = (int)(v & 2047u) - x;
                        dy = (int)((v >> 11) & 2047u) - y;
                        dz = (int)(v >> 22) - z;
                    } else {
                        const int c = a.bin_cells[start + j];
                        dx = c % X - x;
                        dy = c % (X * Y) / X - y;
                        dz = c / (X * Y) - z;
                    }
                    if (abs(dx) <= a.span[0] && abs(dy) <= a.span[1] && abs(dz) <= a.span[2])
                        rk = a.rank[((size_t)(dz + a.span[2]) * a.rd[1] + (dy + a.span[1])) * a.rd[0] +
                                    (dx + a.span[0])];
                }
                const uint64_t mk = __ballot(rk != 0);
                if (rk != 0) {
                    const int p = pos + __popcll(mk & lt);
                    const uint32_t ev = (rk << key.lb) | (uint32_t)(j0 + j);
                    if (p < kCap)
                        sh.e[p] = ev;
                    else if (ge && p < a.scratch_per_wave)
                        ge[p] = ev;
                }
                pos += __popcll(mk);
            }
        }
        const int total = pos;
        __syncthreads();
        int *out = a.nbr + (size_t)t * a.max_n;
        bool sane = true;
        if (total <= kCap) {
            sane = topk_sort_(sh.e, sh.L, sh.R, sh, total, a.max_n, key);
            if (sane) finish_(sh.e, total, a.max_n, out, sh, a.bin_cells, key);
        } else if (ge && total <= a.scratch_per_wave) {
            uint32_t *gL = ge + a.scratch_per_wave;
            uint32_t *gR = gL + a.scratch_per_wave;
            for (int p = lane; p < kCap; p += 64) ge[p] = sh.e[p];  // the head went to LDS
            __syncthreads();
            sane = topk_sort_(ge, gL, gR, sh, total, a.max_n, key);
            if (sane) finish_(ge, total, a.max_n, out, sh, a.bin_cells, key);
        } else {
            if (lane == 0) atomicOr(a.overflow, 1u);
            sane = false;
        }
        if (lane == 0) {
            cand_acc += (unsigned long long)total;
            a.cnt[t] = sane ? min(total, a.max_n) : 0;
            if (!sane) atomicOr(a.overflow, 2u);
        }
        __syncthreads();
    }
    if (lane == 0 && a.cand_total && cand_acc != 0) atomicAdd(counter_line(a.cand_total), cand_acc);
}

}  // namespace

int cluster_lds_capacity() { return kCap; }

void launch_cluster_search(const ClusterArgs &a, hipStream_t s) {
    if (a.n_list <= 0) return;
    const int grid = a.n_list < a.n_waves ? a.n_list : a.n_waves;
    hipLaunchKernelGGL(k_cluster, dim3(grid), dim3(64), 0, s, a);
}

}  // namespace hpgl_amd
