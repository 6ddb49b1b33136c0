// topk_replay.cpp - checks, on the CPU, the formulation the cell-binned
// search kernel (hpgl_amd/csrc/cluster.hip) uses to reproduce the first K
// elements of libstdc++'s std::sort by descending key:
//
//   * each __unguarded_partition pairs the k-th left stop (key <= pivot)
//     with the k-th right stop from the end (key >= pivot) and swaps them
//     while the former is below the latter; the cut is min(L_s, R_{s-1});
//   * ranges longer than 64: two cursors walk 64-position chunks from both
//     ends of the range, each taking a chunk's stop mask when it enters the
//     chunk (on the array as earlier rounds left it) and keeping the
//     unpaired stops; up to 64 pairs are formed per round and swapped
//     together;
//   * a range of at most 64 elements finishes its whole subtree with the
//     elements held one per lane: stops found by ballot, compacted into lane
//     order by a permutation (left stops first, then the other lanes; the
//     same for the right stops from the top), each swapped lane taking its
//     partner's element;
//   * only ranges that start below K are partitioned;
//   * a range whose depth budget runs out is heap-sorted with libstdc++'s
//     make_heap/sort_heap steps;
//   * the first min(n, K+16) positions are finished by a stable sort.
//
// The program replays that formulation sequentially and compares the first
// K (key, payload) pairs with std::sort on random inputs with heavy ties,
// on spatially structured keys and on adversarial inputs that exhaust the
// depth budget.  Exit status 0 on agreement; prints the first mismatch
// otherwise.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

struct E {
    uint32_t key;
    uint32_t id;
};

static bool comp(const E &a, const E &b) { return a.key > b.key; }

constexpr int kInf = 0x7fffffff;

// __move_median_to_first(lo, lo+1, mid, hi-1) on keys k[]; returns the
// position moved to lo.
template <typename KeyAt>
static int median_pos(int lo, int hi, KeyAt key) {
    const int a = lo + 1, b = lo + (hi - lo) / 2, c = hi - 1;
    const uint32_t ka = key(a), kb = key(b), kc = key(c);
    if (ka > kb) {
        if (kb > kc) return b;
        if (ka > kc) return c;
        return a;
    }
    if (ka > kc) return a;
    if (kb > kc) return c;
    return b;
}

static int popc(uint64_t m) { return __builtin_popcountll(m); }

// Ranges longer than 64: the two-cursor partition.
static int partition_stream(std::vector<E> &e, int lo, int hi) {
    const int sw = median_pos(lo, hi, [&](int p) { return e[p].key; });
    std::swap(e[lo], e[sw]);
    const uint32_t kp = e[lo].key;
    const int first = lo + 1;
    const int nch = (hi - first + 63) >> 6;
    auto chunk_mask = [&](int c, bool left) {
        uint64_t m = 0;
        for (int b = 0; b < 64; ++b) {
            const int p = first + (c << 6) + b;
            if (p < hi && (left ? e[p].key <= kp : e[p].key >= kp)) m |= 1ull << b;
        }
        return m;
    };
    int cl = -1, cr = nch;
    uint64_t lrem = 0, rrem = 0;
    int r_prev = kInf;
    for (;;) {
        int lb[64], rb[64], fl = 0, fr = 0;
        for (;;) {  // consume the lowest unpaired stops, then enter the next chunk
            while (lrem && fl < 64) {
                lb[fl++] = first + (cl << 6) + __builtin_ctzll(lrem);
                lrem &= lrem - 1;
            }
            if (fl == 64 || cl + 1 >= nch) break;
            lrem = chunk_mask(++cl, true);
        }
        for (;;) {  // the highest unpaired stops, then the chunk below
            while (rrem && fr < 64) {
                const int b = 63 - __builtin_clzll(rrem);
                rb[fr++] = first + (cr << 6) + b;
                rrem &= ~(1ull << b);
            }
            if (fr == 64 || cr <= 0) break;
            rrem = chunk_mask(--cr, false);
        }
        const int m = std::min(fl, fr);
        int cnt = 0;
        while (cnt < m && lb[cnt] < rb[cnt]) ++cnt;
        for (int k = 0; k < cnt; ++k) std::swap(e[lb[k]], e[rb[k]]);
        if (cnt < 64) {
            const int a = cnt < fl ? lb[cnt] : kInf;
            const int b = cnt >= 1 ? rb[cnt - 1] : r_prev;
            return std::min(a, b);
        }
        r_prev = rb[63];
    }
}

static long g_heap_ranges = 0, g_small_ranges = 0, g_stream_ranges = 0;

static void heap_sort_range(std::vector<E> &e, int lo, int hi) {
    ++g_heap_ranges;
    std::make_heap(e.begin() + lo, e.begin() + hi, comp);
    std::sort_heap(e.begin() + lo, e.begin() + hi, comp);
}

struct R3 {
    int lo, hi, d;
};

// A range of at most 64 elements and its subtree, lane i holding e[lo + i].
static bool small_subtree(std::vector<E> &e, int lo, int hi, int d, int K) {
    const int n = hi - lo;
    E v[64];
    auto load = [&] {
        for (int i = 0; i < 64; ++i) v[i] = i < n ? e[lo + i] : E{0, 0};
    };
    auto store = [&] {
        for (int i = 0; i < n; ++i) e[lo + i] = v[i];
    };
    load();
    std::vector<R3> stk{{0, n, d}};
    while (!stk.empty()) {
        R3 r = stk.back();
        stk.pop_back();
        int l = r.lo, h = r.hi, dd = r.d;
        while (h - l > 16) {
            if (dd == 0) {
                store();
                heap_sort_range(e, lo + l, lo + h);
                load();
                break;
            }
            --dd;
            ++g_small_ranges;
            const int sw = median_pos(l, h, [&](int p) { return v[p].key; });
            std::swap(v[l], v[sw]);
            const uint32_t kp = v[l].key;
            uint64_t lm = 0, rm = 0;
            for (int i = l + 1; i < h; ++i) {
                if (v[i].key <= kp) lm |= 1ull << i;
                if (v[i].key >= kp) rm |= 1ull << i;
            }
            const int nl = popc(lm), nr = popc(rm);
            int l_at[64], r_at[64], il[64], jr[64];
            bool in_l[64], in_r[64];
            for (int i = 0; i < 64; ++i) {
                const uint64_t below = i ? (~0ull >> (64 - i)) : 0;
                const int bl = popc(lm & below), br = popc(rm & below);
                in_l[i] = (lm >> i) & 1;
                in_r[i] = (rm >> i) & 1;
                il[i] = bl;
                jr[i] = nr - 1 - br;
                l_at[in_l[i] ? il[i] : nl + i - bl] = i;  // a permutation of the lanes
                r_at[in_r[i] ? jr[i] : nr + i - br] = i;
            }
            int sc = 0;
            for (int j = 0; j < std::min(nl, nr); ++j) sc += l_at[j] < r_at[j];
            E nv[64];
            for (int i = 0; i < 64; ++i) {
                int src = i;
                if (in_l[i] && il[i] < sc)
                    src = r_at[il[i]];
                else if (in_r[i] && jr[i] < sc)
                    src = l_at[jr[i] & 63];
                nv[i] = v[src];
            }
            std::copy(nv, nv + 64, v);
            const int cut = std::min(sc < nl ? l_at[sc] : kInf, sc >= 1 ? r_at[sc - 1] : kInf);
            if (cut <= l || cut >= h) return false;
            if (lo + cut < K) stk.push_back({cut, h, dd});
            h = cut;
        }
    }
    store();
    return true;
}

static std::vector<E> topk_replay(std::vector<E> e, int K) {
    const int n = (int)e.size();
    if (n > 1) {
        int lg = 0;
        while ((2 << lg) <= n) ++lg;
        std::vector<R3> stk{{0, n, 2 * lg}};
        while (!stk.empty()) {
            R3 r = stk.back();
            stk.pop_back();
            int lo = r.lo, hi = r.hi, d = r.d;
            while (hi - lo > 16) {
                if (hi - lo <= 64) {
                    if (!small_subtree(e, lo, hi, d, K)) {
                        std::fprintf(stderr, "invalid cut (small range)\n");
                        std::exit(2);
                    }
                    break;
                }
                if (d == 0) {
                    heap_sort_range(e, lo, hi);
                    break;
                }
                --d;
                ++g_stream_ranges;
                const int cut = partition_stream(e, lo, hi);
                if (cut <= lo || cut >= hi) {
                    std::fprintf(stderr, "invalid cut\n");
                    std::exit(2);
                }
                if (cut < K) stk.push_back({cut, hi, d});
                hi = cut;
            }
        }
    }
    const int P = std::min(n, K + 16);
    std::stable_sort(e.begin(), e.begin() + P, comp);
    e.resize(std::min(n, K));
    return e;
}

// An input that drives libstdc++'s median-of-3 introsort to its depth
// limit (keys chosen against the pivot rule), to exercise the heap path.
static std::vector<E> adversarial(int n, std::mt19937 &rng) {
    std::vector<E> e(n);
    for (int i = 0; i < n; ++i) e[i] = {(uint32_t)((i % 2) ? i : n - i), (uint32_t)i};
    std::shuffle(e.begin(), e.begin() + n / 8, rng);
    return e;
}

// Keys that vary smoothly with position, like ranks of candidates gathered
// bin by bin (so the median-of-3 pivots are skewed).
static std::vector<E> structured(int n, int distinct, std::mt19937 &rng) {
    std::vector<E> e(n);
    const double f = 3.0 + (double)(rng() % 8);
    for (int i = 0; i < n; ++i) {
        const double s = 0.5 + 0.5 * std::sin(f * i / (double)n);
        e[i] = {(uint32_t)(1 + (uint32_t)(s * (distinct - 1)) + rng() % 3), (uint32_t)i};
    }
    return e;
}

int main() {
    std::mt19937 rng(12345);
    long cases = 0;
    for (int trial = 0; trial < 40000; ++trial) {
        const int n = 1 + (int)(rng() % (trial % 10 == 0 ? 4000 : 1200));
        const int K = 1 + (int)(rng() % 64);
        const int distinct = 1 + (int)(rng() % (trial % 3 == 0 ? 8 : 400));
        std::vector<E> e(n);
        if (trial % 97 == 0) {
            e = adversarial(n, rng);
        } else if (trial % 4 == 1) {
            e = structured(n, distinct, rng);
        } else {
            for (int i = 0; i < n; ++i) e[i] = {(uint32_t)(1 + rng() % distinct), (uint32_t)i};
        }
        std::vector<E> ref = e;
        std::sort(ref.begin(), ref.end(), comp);
        ref.resize(std::min(n, K));
        const std::vector<E> got = topk_replay(e, K);
        for (size_t i = 0; i < ref.size(); ++i) {
            if (got[i].key != ref[i].key || got[i].id != ref[i].id) {
                std::printf("mismatch trial %d n %d K %d pos %zu: got (%u,%u) want (%u,%u)\n", trial, n, K, i,
                            got[i].key, got[i].id, ref[i].key, ref[i].id);
                return 1;
            }
        }
        ++cases;
    }
    std::printf("ok %ld cases, %ld heap-sorted ranges, %ld two-cursor partitions, %ld register partitions\n", cases,
                g_heap_ranges, g_stream_ranges, g_small_ranges);
    return 0;
}
