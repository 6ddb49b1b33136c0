// topk_replay.cpp - checks, on the CPU, the formulation the cell-binned
// search kernel (hpgl_amd/csrc/cluster.hip) uses to reproduce the first K
// elements of libstdc++'s std::sort by descending key:
//
//   * each __unguarded_partition is computed from the positions of the
//     left stops (key <= pivot) and right stops (key >= pivot): the k-th
//     left stop is swapped with the k-th right stop from the end while the
//     former is below the latter, and the cut is min(l_{s+1}, r_s);
//   * only ranges that start below K are partitioned;
//   * a range whose depth budget runs out is heap-sorted with libstdc++'s
//     make_heap/sort_heap steps;
//   * the first min(n, K+16) positions are finished by a stable sort.
//
// The program replays that formulation sequentially and compares the first
// K (key, payload) pairs with std::sort on random inputs with heavy ties and
// on adversarial inputs that exhaust the depth budget.  Exit status 0 on
// agreement; prints the first mismatch otherwise.
#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

struct E {
    uint32_t key;
    uint32_t id;
};

static bool comp(const E &a, const E &b) { return a.key > b.key; }

static int partition_replay(std::vector<E> &e, int lo, int hi) {
    // __move_median_to_first(lo, lo+1, mid, hi-1)
    const int a = lo + 1, b = lo + (hi - lo) / 2, c = hi - 1;
    int sw;
    if (comp(e[a], e[b])) {
        if (comp(e[b], e[c]))
            sw = b;
        else if (comp(e[a], e[c]))
            sw = c;
        else
            sw = a;
    } else if (comp(e[a], e[c])) {
        sw = a;
    } else if (comp(e[b], e[c])) {
        sw = c;
    } else {
        sw = b;
    }
    std::swap(e[lo], e[sw]);
    const uint32_t kp = e[lo].key;
    std::vector<int> L, R;
    for (int p = lo + 1; p < hi; ++p) {
        if (e[p].key <= kp) L.push_back(p);
        if (e[p].key >= kp) R.push_back(p);
    }
    const int nL = (int)L.size(), nR = (int)R.size();
    int s = 0;
    for (int k = 0; k < std::min(nL, nR); ++k) {
        if (L[k] < R[nR - 1 - k]) {
            std::swap(e[L[k]], e[R[nR - 1 - k]]);
            ++s;
        } else {
            break;
        }
    }
    const int x = s < nL ? L[s] : 0x7fffffff;
    const int y = s >= 1 ? R[nR - s] : 0x7fffffff;
    return std::min(x, y);
}

static long g_heap_ranges = 0;

static void heap_sort_range(std::vector<E> &e, int lo, int hi) {
    ++g_heap_ranges;
    std::make_heap(e.begin() + lo, e.begin() + hi, comp);
    std::sort_heap(e.begin() + lo, e.begin() + hi, comp);
}

static std::vector<E> topk_replay(std::vector<E> e, int K) {
    const int n = (int)e.size();
    if (n > 1) {
        struct R3 {
            int lo, hi, d;
        };
        int lg = 0;
        while ((2 << lg) <= n) ++lg;
        std::vector<R3> stk{{0, n, 2 * lg}};
        while (!stk.empty()) {
            R3 r = stk.back();
            stk.pop_back();
            int lo = r.lo, hi = r.hi, d = r.d;
            while (hi - lo > 16) {
                if (d == 0) {
                    heap_sort_range(e, lo, hi);
                    break;
                }
                --d;
                const int cut = partition_replay(e, lo, hi);
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
    std::printf("ok %ld cases, %ld heap-sorted ranges\n", cases, g_heap_ranges);
    return 0;
}
