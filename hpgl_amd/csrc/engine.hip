// engine.hip - one kriging job on the device: host set-up, cell bins,
// search, solve, epilogue, statistics.
#include "engine.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "kernels.h"

namespace hpgl_amd {


void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Device memory cache: freed blocks are kept per (device, size class), the
// classes powers of two from 256 B, and reused, so repeated jobs do not call
// hipMalloc.  Frees are stream-ordered: a block freed while kernels that use
// it may still be queued on the freeing thread's stream (PoolStream) carries
// an event recorded on that stream at the free.  It is handed out again only
// once the event has completed, or to a caller with a stream that is then
// made to wait for the event (the same stream, or any stream when memory is
// short); a block freed with no stream set waits for the whole device first.
// ---------------------------------------------------------------------------
namespace {
struct FreeBlock {
    void *p;
    hipStream_t s;    // stream the free was ordered on (nullptr: none)
    hipEvent_t done;  // completes when the freeing stream's earlier work has (nullptr: free now)
};
std::mutex g_pool_mu;
std::multimap<std::pair<int, size_t>, FreeBlock> g_free;  // (device, size class) -> block
std::map<void *, std::pair<int, size_t>> g_live;
std::vector<hipEvent_t> g_spare_events;
thread_local hipStream_t t_pool_stream = nullptr;
thread_local bool t_pool_stream_set = false;

size_t size_class(size_t bytes) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
}

int current_device() {
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    return d;
}

// hipEventQuery reports a pending event as hipErrorNotReady, which HIP also
// latches as the thread's last error; clear it unless an earlier error is
// waiting to be read.
bool event_done(hipEvent_t e) {
    const hipError_t before = hipPeekAtLastError();
    const hipError_t r = hipEventQuery(e);
    if (r == hipErrorNotReady && before == hipSuccess) (void)hipGetLastError();
    return r == hipSuccess;
}

// Under g_pool_mu.
void recycle_event(hipEvent_t e) {
    if (e) g_spare_events.push_back(e);
}

// Under g_pool_mu: a free block of the class usable on the calling thread's
// stream, or nullptr.  wait_any: take a block whose event is still pending
// and order the calling stream after it.
void *take_block(int dev, size_t c, bool wait_any) {
    auto range = g_free.equal_range({dev, c});
    auto pick = g_free.end();
    bool need_wait = false;
    for (auto it = range.first; it != range.second; ++it) {
        const FreeBlock &b = it->second;
        if (!b.done || event_done(b.done)) {
            pick = it;
            need_wait = false;
            break;
        }
        if (t_pool_stream_set && (wait_any || b.s == t_pool_stream) && pick == g_free.end()) {
            pick = it;
            need_wait = true;
        }
    }
    if (pick == g_free.end()) return nullptr;
    FreeBlock b = pick->second;
    if (need_wait) hip_check(hipStreamWaitEvent(t_pool_stream, b.done, 0), "hipStreamWaitEvent");
    recycle_event(b.done);
    g_free.erase(pick);
    g_live[b.p] = {dev, c};
    return b.p;
}
}  // namespace

PoolStream::PoolStream(hipStream_t s) : prev_(t_pool_stream), prev_set_(t_pool_stream_set) {
    t_pool_stream = s;
    t_pool_stream_set = true;
}

PoolStream::~PoolStream() {
    t_pool_stream = prev_;
    t_pool_stream_set = prev_set_;
}

void *pool_alloc(size_t bytes) {
    if (bytes == 0) bytes = 1;
    const size_t c = size_class(bytes);
    const int dev = current_device();
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (void *p = take_block(dev, c, false)) return p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // memory is short: reuse a pending block of the class behind its
        // event, else return the cached blocks of this device and retry
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (void *q = take_block(dev, c, true)) return q;
        for (auto it = g_free.begin(); it != g_free.end();) {
            if (it->first.first != dev) {
                ++it;
                continue;
            }
            if (it->second.done) (void)hipEventSynchronize(it->second.done);
            (void)hipFree(it->second.p);
            recycle_event(it->second.done);
            it = g_free.erase(it);
        }
        e = hipMalloc(&p, c);
    }
    hip_check(e, "hipMalloc");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_live[p] = {dev, c};
    return p;
}

void pool_free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return;
    FreeBlock b{p, nullptr, nullptr};
    if (t_pool_stream_set) {
        hipEvent_t ev = nullptr;
        if (!g_spare_events.empty()) {
            ev = g_spare_events.back();
            g_spare_events.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
        }
        if (ev && hipEventRecord(ev, t_pool_stream) == hipSuccess) {
            b.s = t_pool_stream;
            b.done = ev;
        } else {
            recycle_event(ev);
            (void)hipDeviceSynchronize();
        }
    } else {
        (void)hipDeviceSynchronize();
    }
    g_free.emplace(it->second, b);
    g_live.erase(it);
}

namespace {
// Writes v to p[0] after spinning for about `ticks` ticks of the 100 MHz
// constant clock (bounded: at most 2^26 polls).
__global__ void k_pool_probe_spin(int *p, int v, long long ticks) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long t0 = wall_clock64();
    for (int i = 0; i < (1 << 26) && wall_clock64() - t0 < ticks; ++i) __builtin_amdgcn_s_sleep(8);
    p[0] = v;
}
}  // namespace

// A block freed on stream A while a kernel that writes it is still queued
// on A must not be handed to stream B before that kernel has completed.
// Returns 0 when that holds, 1 when the spin ended before the check could
// be made (inconclusive), and a negative code when the pool broke the
// ordering or the kernel's write was lost.
int pool_selftest() {
    hipStream_t a = nullptr, b = nullptr;
    hip_check(hipStreamCreateWithFlags(&a, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&b, hipStreamNonBlocking), "hipStreamCreate");
    int rc = 0;
    const size_t bytes = (size_t)3 << 20;  // a size class the library's jobs rarely use
    void *x = nullptr, *y = nullptr;
    {
        PoolStream ps(a);
        x = pool_alloc(bytes);
        hipLaunchKernelGGL(k_pool_probe_spin, dim3(1), dim3(64), 0, a, (int *)x, 7, 5000000ll);  // ~50 ms
        hip_check(hipGetLastError(), "spin kernel");
        pool_free(x);
    }
    {
        PoolStream ps(b);
        const bool pending = pool_pending_blocks() > 0;
        y = pool_alloc(bytes);
        const hipError_t q = hipStreamQuery(a);
        if (q == hipErrorNotReady) (void)hipGetLastError();
        const bool still_running = q == hipErrorNotReady;
        if (!pending || !still_running) rc = 1;
        else if (y == x) rc = -1;  // handed to another stream while pending
        pool_free(y);
    }
    hip_check(hipStreamSynchronize(b), "sync");
    hip_check(hipStreamSynchronize(a), "sync");
    int v = 0;
    hip_check(hipMemcpy(&v, x, sizeof(int), hipMemcpyDeviceToHost), "copy");
    if (v != 7 && rc >= 0) rc = -3;
    (void)hipStreamDestroy(a);
    (void)hipStreamDestroy(b);
    return rc;
}

size_t pool_pending_blocks() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    size_t n = 0;
    for (auto &kv : g_free)
        if (kv.second.done && !event_done(kv.second.done)) ++n;
    return n;
}

namespace {

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        n = count;
        p = (T *)pool_alloc(std::max<size_t>(count, 1) * sizeof(T));
    }
    void release() {
        pool_free(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Events {
    hipEvent_t e[8];
    Events() {
        for (auto &x : e) hip_check(hipEventCreate(&x), "hipEventCreate");
    }
    ~Events() {
        for (auto &x : e) (void)hipEventDestroy(x);
    }
    float ms(int a, int b) {
        float v = 0;
        hip_check(hipEventElapsedTime(&v, e[a], e[b]), "hipEventElapsedTime");
        return v;
    }
};

// Device time of one launch sequence on stream s.
struct KTimer {
    hipEvent_t a, b;
    hipStream_t s;
    explicit KTimer(hipStream_t st) : s(st) {
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        hip_check(hipEventRecord(a, s), "hipEventRecord");
    }
    double stop() {
        hip_check(hipEventRecord(b, s), "hipEventRecord");
        hip_check(hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        return ms;
    }
    ~KTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// Device times of launches on stream s without a host wait per launch: each
// start/stop pair records two events, read by collect() once the stream has
// been synchronised past them for another reason.
class KernelClock {
  public:
    explicit KernelClock(hipStream_t s) : s_(s) {}
    KernelClock(const KernelClock &) = delete;
    KernelClock &operator=(const KernelClock &) = delete;
    ~KernelClock() {
        for (const Mark &m : marks_) {
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
        }
    }
    void start() {
        Mark m{};
        hip_check(hipEventCreate(&m.a), "hipEventCreate");
        hip_check(hipEventCreate(&m.b), "hipEventCreate");
        marks_.push_back(m);
        hip_check(hipEventRecord(m.a, s_), "hipEventRecord");
    }
    // ends the last start(); its time goes to ms and the launch to count
    void stop(double &ms, long long &count) {
        Mark &m = marks_.back();
        hip_check(hipEventRecord(m.b, s_), "hipEventRecord");
        m.ms = &ms;
        m.count = &count;
    }
    // the stream must have passed every recorded event
    void collect() {
        for (const Mark &m : marks_) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, m.a, m.b), "hipEventElapsedTime");
            if (m.ms) *m.ms += ms;
            if (m.count) *m.count += 1;
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
        }
        marks_.clear();
    }

  private:
    struct Mark {
        hipEvent_t a, b;
        double *ms;
        long long *count;
    };
    hipStream_t s_;
    std::vector<Mark> marks_;
};

struct IsZero {
    const uint8_t *m;
    __host__ __device__ bool operator()(const int &i) const { return m[i] == 0; }
};
struct IsNonZero {
    const uint8_t *m;
    __host__ __device__ bool operator()(const int &i) const { return m[i] != 0; }
};
struct FlagEq {
    const uint8_t *f;
    uint8_t v;
    __host__ __device__ bool operator()(const int &i) const { return f[i] == v; }
};

// Ascending list of i in [0, n) with pred(i) into out, its length into
// *d_num on the device (no host wait).
template <typename Pred>
void select_indices_async(Pred pred, int n, DBuf<int> &out, int *d_num, hipStream_t s) {
    out.alloc(n);
    hipcub::CountingInputIterator<int> it(0);
    size_t bytes = 0;
    hip_check(hipcub::DeviceSelect::If(nullptr, bytes, it, out.p, d_num, n, pred, s), "select");
    DBuf<uint8_t> tmp(bytes);
    hip_check(hipcub::DeviceSelect::If(tmp.p, bytes, it, out.p, d_num, n, pred, s), "select");
}

// The same, returning the length.
template <typename Pred>
int select_indices(Pred pred, int n, DBuf<int> &out, hipStream_t s) {
    DBuf<int> d_num(1);
    select_indices_async(pred, n, out, d_num.p, s);
    int num = 0;
    hip_check(hipMemcpyAsync(&num, d_num.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy count");
    hip_check(hipStreamSynchronize(s), "sync");
    return num;
}

__global__ void k_iota(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Adds the number of non-zero bytes of m[0..n) to *out: a grid-stride count
// per lane, summed over the wavefront, one atomic per wavefront (a few
// thousand in all, so the shared counter does not serialise the launch).
__global__ void k_count_nonzero(const uint8_t *m, long long n, unsigned long long *out) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned int c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c += m[i] != 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, (unsigned long long)c);
}

// Sort key of path position i (target cell targets[i]) for the template
// scans: the cell, below the position's density bucket
// b = floor(log2((H + i) / H) / 4) (at most 15), H the hard-data cells,
// informed before the path starts.  A target at position i scans about max_n N / (H + i)
// template offsets, so the lanes of a wavefront in one bucket need scan
// lengths within a factor of 16, while cell order inside the bucket keeps
// their probes close together.  Measured on 64 SGS realisations of a
// 200x200x100 grid with 1% hard data: no buckets 702 ms of scans, factor-2
// buckets 484, factor 4 416, factor 8 387, factor 16 386.
__global__ void k_scan_order_keys(const int *targets, int T, long long H, int cell_bits, uint32_t *keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const unsigned long long q = (unsigned long long)(H + i) / (unsigned long long)H;
    const int b = min((63 - __clzll(q)) >> 2, 15);
    keys[i] = ((uint32_t)b << cell_bits) | (uint32_t)targets[i];
}

template <typename T>
void upload(DBuf<T> &d, const std::vector<T> &h, hipStream_t s) {
    d.alloc(h.size());
    if (!h.empty()) hip_check(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

// Largest covariance table built for the analytic model (entries, 256 MiB
// of doubles); beyond it the solves evaluate the model on the device.
constexpr long long kMaxPairTable = 1ll << 25;

// Cell-binned search launch: list entries per wavefront of the main pass
// (a grid many times the resident set lets the hardware balance the uneven
// per-target work; 8192 resident-sized wavefronts ran 20-25% slower), and
// wavefronts (each with its own global scratch) of the pass that sorts the
// lists longer than the LDS capacity.
constexpr int kClusterPerWave = 32;
constexpr int kClusterSpillWaves = 2048;

int solve_group(int max_n) {
    int g = 8;
    while (g < max_n) g <<= 1;
    return g;
}

DevCov make_devcov(const CovModel &m, const BoxTable *table, const double *d_table, const double *d_fold = nullptr) {
    DevCov c;
    std::memset(&c, 0, sizeof(c));
    c.table = table ? d_table : nullptr;
    c.fold = table ? d_fold : nullptr;
    for (int i = 0; i < 3; ++i) {
        c.tr[i] = table ? table->r[i] : 0;
        c.td[i] = table ? table->d[i] : 1;
    }
    for (int i = 0; i < 9; ++i) c.t[i] = m.t[i];
    c.range0 = m.s.ranges[0];
    c.c0 = m.c0();
    c.type = m.s.type;
    return c;
}

void validate_search(const int grid[3], const int radii[3], int max_n) {
    long long n = 1;
    for (int i = 0; i < 3; ++i) {
        if (grid[i] <= 0) throw std::runtime_error("grid dimensions must be positive");
        if (radii[i] <= 0) throw std::runtime_error("search radiuses must be positive");
        n *= grid[i];
    }
    if (n >= (1ll << 31)) throw std::runtime_error("grid has more than 2^31-1 cells");
    long long box = 1;
    for (int i = 0; i < 3; ++i) box *= 2ll * radii[i] + 1;
    if (box >= (1ll << 31)) throw std::runtime_error("search radiuses too large for the covariance table");
    if (max_n < 0 || max_n > kMaxNeighbours)
        throw std::runtime_error("max_neighbours must be in [0, " + std::to_string(kMaxNeighbours) + "]");
}

// Everything the search and the solves of one covariance need, built on
// the host and uploaded once, then kept resident: jobs that repeat a
// covariance model and radii (a serving loop, the categories of repeated IK
// jobs) reuse them.  Keyed by device, model, radii and table use; the
// cluster-path rank table and the OK pair table are built on first use.
struct CovAssets {
    CovModel model;
    bool use_table = false;
    int radii[3] = {0, 0, 0};
    BoxTable table;
    Template tmpl;
    DBuf<double> d_table;
    std::vector<double> fold;  // the table's octant when it is symmetric (BoxTable::fold)
    DBuf<double> d_fold;
    DBuf<int32_t> d_tmpl;
    std::mutex lazy;  // guards the members below
    bool have_rank = false;
    RankTable rank;
    long long rank_positive = 0;  // offsets with rank > 0
    uint32_t max_rank = 0;
    // the device table's ranks are stored shifted left by the sort elements'
    // index bits (fixed by the template size); -1 when they do not fit
    int rank_lb = -1;
    // half-widths of the device table: the box, widened in x by
    // kClusterRunSlack zero columns on each side
    int wh[3] = {0, 0, 0};
    DBuf<uint32_t> d_rank;
    // host copy of the device rank table, built ahead by prefetch_assets
    // (rank_built) and freed once uploaded
    bool rank_built = false;
    std::vector<uint32_t> rank_wide;
    bool have_pair = false;
    BoxTable pair;
    DBuf<double> d_pair;
    DBuf<double> d_pair_fold;
    explicit CovAssets(const CovSpec &c) : model(c) {}
    // Jobs on other streams may have read these buffers after the last
    // holder's stream did: wait for the device before they are freed.
    ~CovAssets() { (void)hipDeviceSynchronize(); }
};

constexpr size_t kAssetCacheEntries = 32;

struct AssetCache {
    std::mutex mu;
    std::list<std::pair<std::string, std::shared_ptr<CovAssets>>> lru;  // most recent first
};
AssetCache &asset_cache() {
    // never destroyed: its buffers must not be freed after the HIP runtime
    // has shut down at process exit
    static AssetCache *c = new AssetCache;
    return *c;
}

// Assets of (cov, radii, use_table) on the current device; builds them on a
// miss (host time counted in ms_setup_host) and waits for the upload before
// other streams may use them.
std::string asset_key(const CovSpec &cov, const int radii[3], bool use_table) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::string key(reinterpret_cast<const char *>(&dev), sizeof(dev));
    key.append(reinterpret_cast<const char *>(&cov.type), sizeof(cov.type));
    key.append(reinterpret_cast<const char *>(cov.ranges), sizeof(cov.ranges));
    key.append(reinterpret_cast<const char *>(cov.angles), sizeof(cov.angles));
    key.append(reinterpret_cast<const char *>(&cov.sill), sizeof(cov.sill));
    key.append(reinterpret_cast<const char *>(&cov.nugget), sizeof(cov.nugget));
    key.append(reinterpret_cast<const char *>(radii), 3 * sizeof(int));
    key.push_back(use_table ? 1 : 0);
    return key;
}

std::shared_ptr<CovAssets> find_assets(const std::string &key) {
    AssetCache &c = asset_cache();
    std::lock_guard<std::mutex> g(c.mu);
    for (auto it = c.lru.begin(); it != c.lru.end(); ++it)
        if (it->first == key) {
            auto e = it->second;
            c.lru.splice(c.lru.begin(), c.lru, it);
            return e;
        }
    return nullptr;
}

void insert_assets(const std::string &key, const std::shared_ptr<CovAssets> &e) {
    AssetCache &c = asset_cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.lru.emplace_front(key, e);
    while (c.lru.size() > kAssetCacheEntries) c.lru.pop_back();
}

// Host part of the assets (no HIP calls): box table, template, octant.
void build_assets_host(CovAssets &e) {
    if (e.use_table) e.table.build(e.model, e.radii);
    build_template(e.model, e.use_table ? &e.table : nullptr, e.radii, e.tmpl);
    if (!(e.use_table && e.table.fold(e.fold))) e.fold.clear();
}

void upload_assets(CovAssets &e, hipStream_t s) {
    if (e.use_table) {
        upload(e.d_table, e.table.v, s);
        if (!e.fold.empty()) upload(e.d_fold, e.fold, s);
    }
    upload(e.d_tmpl, e.tmpl.xyz, s);
    hip_check(hipStreamSynchronize(s), "sync");
}

std::shared_ptr<CovAssets> get_assets(const CovSpec &cov, const int radii[3], bool use_table, hipStream_t s,
                                      Stats &st) {
    const std::string key = asset_key(cov, radii, use_table);
    if (auto e = find_assets(key)) return e;
    auto e = std::make_shared<CovAssets>(cov);
    e->use_table = use_table;
    for (int i = 0; i < 3; ++i) e->radii[i] = radii[i];
    const double t0 = now_ms();
    build_assets_host(*e);
    st.ms_setup_host += now_ms() - t0;
    upload_assets(*e, s);
    insert_assets(key, e);
    return e;
}

bool build_rank_host(CovAssets &a);

// Builds the host parts of the assets of several covariance models at once,
// one thread per model not yet cached (the box tables, the template sorts
// and, for tables, the cell-binned search's rank tables are independent),
// then uploads them and enters them in the cache, so the jobs' calls of
// get_assets find them.  A rank table that cannot be built here is left to
// assets_rank, which reports the error if the search needs it.
void prefetch_assets(const std::vector<std::pair<CovSpec, std::array<int, 3>>> &specs, bool use_table,
                     hipStream_t s, Stats &st) {
    std::vector<std::string> keys;
    std::vector<std::shared_ptr<CovAssets>> todo;
    for (const auto &sp : specs) {
        const std::string key = asset_key(sp.first, sp.second.data(), use_table);
        if (std::find(keys.begin(), keys.end(), key) != keys.end() || find_assets(key)) continue;
        auto e = std::make_shared<CovAssets>(sp.first);
        e->use_table = use_table;
        for (int i = 0; i < 3; ++i) e->radii[i] = sp.second[i];
        keys.push_back(key);
        todo.push_back(e);
    }
    if (todo.size() < 2) return;  // get_assets builds a single model as fast
    const double t0 = now_ms();
    std::vector<std::string> errors(todo.size());
    {
        // at most 16 threads, each taking the next model
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t k = next++; k < todo.size(); k = next++) {
                try {
                    build_assets_host(*todo[k]);
                } catch (const std::exception &ex) {
                    errors[k] = ex.what();
                    continue;
                }
                try {
                    if (use_table) build_rank_host(*todo[k]);
                } catch (const std::exception &) {
                    todo[k]->rank_built = false;
                }
            }
        };
        // the calling thread works too, so a thread that cannot be started
        // only costs parallelism
        std::vector<std::thread> th;
        const size_t n_threads = std::min<size_t>(todo.size(), 16) - 1;
        try {
            for (size_t t = 0; t < n_threads; ++t) th.emplace_back(work);
        } catch (const std::system_error &) {
        }
        work();
        for (auto &t : th) t.join();
    }
    st.ms_setup_host += now_ms() - t0;
    for (const auto &m : errors)
        if (!m.empty()) throw std::runtime_error(m);
    for (size_t k = 0; k < todo.size(); ++k) {
        upload_assets(*todo[k], s);
        insert_assets(keys[k], todo[k]);
    }
}

// Host part of the rank table: the ranks, their packing shift and the
// device layout in a.rank_wide (no HIP calls).  Throws when the radii are
// too large for the cell-binned search.
bool build_rank_host(CovAssets &a) {
    int span[3];
    for (int i = 0; i < 3; ++i) span[i] = std::max(a.use_table ? a.radii[i] : 2 * a.radii[i] - 1, 0);
    // the search's box test takes offsets modulo 2^16
    if (span[0] + kClusterRunSlack >= (1 << 14) || span[1] >= (1 << 14))
        throw std::runtime_error("search radiuses too large for the cell-binned search");
    int wd[3];
    for (int i = 0; i < 3; ++i) {
        a.wh[i] = span[i] + (i == 0 ? kClusterRunSlack : 0);
        wd[i] = 2 * a.wh[i] + 1;
    }
    // the search kernel indexes the table with 32-bit byte offsets
    if ((long long)wd[0] * wd[1] * wd[2] + 2ll * kClusterRankPad >= (1ll << 30))
        throw std::runtime_error("search radiuses too large for the cell-binned search");
    a.rank.build(a.model, a.use_table ? &a.table : nullptr, span);
    a.rank_positive = 0;
    a.max_rank = 0;
    for (uint32_t r : a.rank.rank) {
        a.rank_positive += r != 0;
        a.max_rank = std::max(a.max_rank, r);
    }
    // elements of the search's sort pack (rank << lb) | raw index, lb the
    // bits of the raw indices of bins of at most template/250 cells
    const long long max_raw = cluster_raw_bound((int)(a.tmpl.size() / 250));
    int lb = 1;
    while ((1ll << lb) < max_raw) ++lb;
    a.rank_lb = (lb < 32 && (unsigned long long)a.max_rank < (1ull << (32 - lb))) ? lb : -1;
    const int shift = std::max(a.rank_lb, 0);
    a.rank_wide.assign((size_t)wd[0] * wd[1] * wd[2] + 2 * kClusterRankPad, 0u);
    const RankTable &rt = a.rank;
    for (int z = 0; z < rt.d[2]; ++z)
        for (int y = 0; y < rt.d[1]; ++y) {
            const size_t src = ((size_t)z * rt.d[1] + y) * rt.d[0];
            const size_t dst = kClusterRankPad + ((size_t)(z - span[2] + a.wh[2]) * wd[1] + (y - span[1] + a.wh[1])) *
                                                     wd[0] + (size_t)(a.wh[0] - span[0]);
            for (int x = 0; x < rt.d[0]; ++x) a.rank_wide[dst + x] = rt.rank[src + x] << shift;
        }
    a.rank_built = true;
    return true;
}

// Rank table of the cluster path over [-span, span]^3: span = r with a
// table, 2r - 1 with the analytic model (candidates lie in the 27 bins).
// The device copy has kClusterRunSlack zero columns on either side in x, so
// a run of neighbouring targets reads its members' entries in one load.
const CovAssets &assets_rank(CovAssets &a, hipStream_t s, Stats &st) {
    std::lock_guard<std::mutex> g(a.lazy);
    if (a.have_rank) return a;
    if (!a.rank_built) {
        const double t0 = now_ms();
        build_rank_host(a);
        st.ms_setup_host += now_ms() - t0;
    }
    upload(a.d_rank, a.rank_wide, s);
    hip_check(hipStreamSynchronize(s), "sync");
    std::vector<uint32_t>().swap(a.rank_wide);
    a.have_rank = true;
    return a;
}

// Neighbour search for targets[0..T) over the informed cells inf[0..I).
// Fills nbr[T][max_n] and cnt[T]; returns (template-path, cluster-path) counts.
void search(CovAssets &as, const int grid[3], const int radii[3], int max_n, const uint8_t *d_mask,
            const int *d_targets, int T, const int *d_inf, int I, int *d_nbr, int *d_cnt, hipStream_t s, Stats &st) {
    if (T == 0) return;
    // the cluster path's rank table and its gather trim follow these radii
    for (int i = 0; i < 3; ++i)
        if (radii[i] != as.radii[i]) throw std::runtime_error("search: radiuses differ from the covariance assets'");
    struct {
        BinGeometry bins;
    } ss;
    ss.bins.init(grid, radii, as.tmpl.size());

    BinArgs ba;
    ba.X = grid[0];
    ba.Y = grid[1];
    ba.Z = grid[2];
    for (int i = 0; i < 3; ++i) {
        ba.r[i] = ss.bins.r[i];
        ba.m[i] = ss.bins.m[i];
    }
    ba.limit = ss.bins.limit;
    DBuf<uint32_t> bin_count(ss.bins.nbins);
    DBuf<int> bin_of(std::max(I, 1));
    launch_fill_u32(bin_count.p, 0, ss.bins.nbins, s);
    launch_bin_ids(d_inf, I, ba, bin_of.p, bin_count.p, s);
    DBuf<uint8_t> flag(T);
    launch_path_flags(d_targets, T, ba, bin_count.p, flag.p, s);
    hip_check(hipGetLastError(), "bin kernels");

    // both paths' target lists, their lengths read with one host wait
    DBuf<int> tlist, clist, d_nums(2);
    select_indices_async(FlagEq{flag.p, 0}, T, tlist, d_nums.p, s);
    select_indices_async(FlagEq{flag.p, 1}, T, clist, d_nums.p + 1, s);
    int nums[2] = {0, 0};
    hip_check(hipMemcpyAsync(nums, d_nums.p, sizeof(nums), hipMemcpyDeviceToHost, s), "copy counts");
    hip_check(hipStreamSynchronize(s), "sync");
    const int n_tmpl = nums[0];
    const int n_clu = nums[1];
    if (n_tmpl + n_clu != T) throw std::runtime_error("search: path lists do not add up");
    KernelClock clk(s);
    st.nodes_template_path += n_tmpl;
    st.nodes_cluster_path += n_clu;

    ScanArgs sa;
    sa.mask = d_mask;
    sa.key = nullptr;
    sa.X = grid[0];
    sa.Y = grid[1];
    sa.Z = grid[2];
    sa.tmpl = (const int4 *)as.d_tmpl.p;
    sa.tsize = as.tmpl.size();
    sa.max_n = max_n;
    sa.targets = d_targets;
    sa.list = tlist.p;
    sa.n_list = n_tmpl;
    sa.nbr = d_nbr;
    sa.cnt = d_cnt;
    sa.simbits = nullptr;
    if (n_tmpl > 0) {
        clk.start();
        launch_template_scan(sa, 0, s);
        hip_check(hipGetLastError(), "template scan");
        clk.stop(st.ms_k_template, st.n_k_template);
    }

    // neighbour total, read with the search's counters
    DBuf<int> total(1);
    unsigned int ov = 0;
    if (n_clu > 0) {
        // bin members: informed cells sorted by (bin, cell); radix sort is stable
        DBuf<int> keys_out(std::max(I, 1)), cells_out(std::max(I, 1));
        size_t bytes = 0;
        int end_bit = 1;
        while ((1ll << end_bit) < ss.bins.nbins) ++end_bit;
        hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, bin_of.p, keys_out.p, d_inf, cells_out.p, I, 0,
                                                     end_bit, s),
                  "radix sort");
        DBuf<uint8_t> tmp(bytes);
        hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, bin_of.p, keys_out.p, d_inf, cells_out.p, I, 0,
                                                     end_bit, s),
                  "radix sort");
        DBuf<uint32_t> bin_start(ss.bins.nbins);
        bytes = 0;
        hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, bin_count.p, bin_start.p, (int)ss.bins.nbins, s),
                  "scan");
        DBuf<uint8_t> tmp2(bytes);
        hip_check(hipcub::DeviceScan::ExclusiveSum(tmp2.p, bytes, bin_count.p, bin_start.p, (int)ss.bins.nbins, s),
                  "scan");
        // sort keys as ranks over the offsets a candidate can have
        const CovAssets &ra = assets_rank(as, s, st);
        const RankTable &rt = ra.rank;

        ClusterArgs ca;
        ca.X = grid[0];
        ca.Y = grid[1];
        ca.Z = grid[2];
        ca.bins = ba;
        ca.bin_start = bin_start.p;
        ca.bin_count = bin_count.p;
        ca.bin_cells = cells_out.p;
        const long long n_zoff = ss.bins.nbins * (ss.bins.r[2] + 1ll);
        if (n_zoff >= (1ll << 31)) throw std::runtime_error("cluster search: too many bins");
        DBuf<uint32_t> zoff(n_zoff);
        launch_bin_zoff(bin_start.p, bin_count.p, cells_out.p, ba, ss.bins.nbins, zoff.p, s);
        ca.zoff = zoff.p;
        ca.rank = ra.d_rank.p;
        for (int i = 0; i < 3; ++i) {
            ca.span[i] = rt.span[i];
            ca.w[i] = 2 * ra.wh[i] + 1;
        }
        ca.centre4 = 4u * (uint32_t)(kClusterRankPad + ra.wh[0] + ca.w[0] * (ra.wh[1] + (long long)ca.w[1] * ra.wh[2]));
        DBuf<uint2> lin((size_t)I + kClusterLinPad);
        launch_pack_lin(cells_out.p, I, grid[0], grid[1], (uint32_t)ca.w[0], (uint32_t)ca.w[0] * (uint32_t)ca.w[1],
                        lin.p, s);
        ca.bin_lin = lin.p;
        ca.max_n = max_n;
        ca.targets = d_targets;
        ca.list = clist.p;
        ca.n_list = n_clu;
        ca.nbr = d_nbr;
        ca.cnt = d_cnt;
        // candidates per target are at most 27 bins x limit cells; an element
        // packs (rank << lb) | raw index into 32 bits
        const long long max_cand = 27ll * std::max(ss.bins.limit, 0);
        const long long max_raw = cluster_raw_bound(std::max(ss.bins.limit, 0));
        int lb = 1;
        while ((1ll << lb) < max_raw) ++lb;
        const uint32_t max_rank = ra.max_rank;
        if (lb >= 32 || (unsigned long long)max_rank >= (1ull << (32 - lb)))
            throw std::runtime_error("cluster search: covariance has too many distinct values for 32-bit sort keys");
        if (lb != ra.rank_lb) throw std::runtime_error("cluster search: rank table shifted for other bins");
        ca.local_bits = lb;
        // expected list length: informed density x offsets with rank > 0,
        // plus six standard deviations of a Poisson count
        const long long n_pos = ra.rank_positive;
        const double expect = (double)I / ((double)grid[0] * grid[1] * grid[2]) * (double)n_pos;
        ca.cap = cluster_capacity((long long)(expect + 6.0 * std::sqrt(expect) + 64.0));
        ca.per_wave = kClusterPerWave;
        DBuf<int> work(kClusterWorkParts * kClusterWorkStride);
        launch_fill_u32(reinterpret_cast<uint32_t *>(work.p), 0u, kClusterWorkParts * kClusterWorkStride, s);
        ca.work = work.p;
        // the later members' lists of runs of neighbouring targets
        ca.run_slots = 0;
        ca.run_scratch = nullptr;
        const long long slots = cluster_main_waves(ca);
        DBuf<uint32_t> run_scratch((size_t)slots * cluster_run_words(ca.cap));
        ca.run_scratch = run_scratch.p;
        ca.run_slots = (int)slots;
        DBuf<uint32_t> scratch;
        DBuf<int> spill_list, spill_count;
        if (max_cand > ca.cap) {
            ca.scratch_per_wave = (int)max_cand;
            ca.n_spill_waves = (int)std::min<long long>(kClusterSpillWaves, n_clu);
            scratch.alloc((size_t)ca.n_spill_waves * cluster_scratch_words((int)max_cand));
            spill_list.alloc(n_clu);
            spill_count.alloc(1);
            hip_check(hipMemsetAsync(spill_count.p, 0, sizeof(int), s), "memset");
            ca.scratch = scratch.p;
            ca.spill_list = spill_list.p;
            ca.spill_count = spill_count.p;
        } else {
            ca.scratch_per_wave = 0;
            ca.scratch = nullptr;
            ca.spill_list = nullptr;
            ca.spill_count = nullptr;
            ca.n_spill_waves = 0;
        }
        DBuf<unsigned int> overflow(1);
        DBuf<unsigned long long> cand(kCounterWords);
        hip_check(hipMemsetAsync(overflow.p, 0, sizeof(unsigned int), s), "memset");
        hip_check(hipMemsetAsync(cand.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
        ca.overflow = overflow.p;
        ca.cand_total = cand.p;
        clk.start();
        launch_cluster_search(ca, s);
        hip_check(hipGetLastError(), "cluster search");
        clk.stop(st.ms_k_cluster, st.n_k_cluster);
        std::vector<unsigned long long> hc(kCounterWords);
        {
            size_t bytes = 0;
            hip_check(hipcub::DeviceReduce::Sum(nullptr, bytes, d_cnt, total.p, T, s), "reduce");
            DBuf<uint8_t> tmp(bytes);
            hip_check(hipcub::DeviceReduce::Sum(tmp.p, bytes, d_cnt, total.p, T, s), "reduce");
        }
        hip_check(hipMemcpyAsync(hc.data(), cand.p, kCounterWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
                  "copy cand");
        hip_check(hipMemcpyAsync(&ov, overflow.p, sizeof(ov), hipMemcpyDeviceToHost, s), "copy overflow");
        int h = 0;
        hip_check(hipMemcpyAsync(&h, total.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        hip_check(hipStreamSynchronize(s), "sync");
        st.neighbours_total += h;
        for (int l = 0; l < kCounterSlots; ++l) {
            st.cluster_candidates += (long long)hc[(size_t)l * kCounterStride];
            st.cluster_lists_spilled += (long long)hc[(size_t)l * kCounterStride + 1];
        }
        if (ov & 1u) throw std::runtime_error("cluster search: candidate list exceeded the scratch capacity");
        if (ov & 2u) throw std::runtime_error("cluster search: internal error in the sort replay");
    } else {
        size_t bytes = 0;
        hip_check(hipcub::DeviceReduce::Sum(nullptr, bytes, d_cnt, total.p, T, s), "reduce");
        DBuf<uint8_t> tmp(bytes);
        hip_check(hipcub::DeviceReduce::Sum(tmp.p, bytes, d_cnt, total.p, T, s), "reduce");
        int h = 0;
        hip_check(hipMemcpyAsync(&h, total.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
        hip_check(hipStreamSynchronize(s), "sync");
        st.neighbours_total += h;
    }
    clk.collect();
}

void write_dump(const Dump *dump, long long cell_offset, const int *d_targets, int T, const int *d_nbr,
                const int *d_cnt, int max_n, hipStream_t s) {
    if (!dump || !dump->indices || T == 0) return;
    std::vector<int> tg(T), cnt(T), nbr((size_t)T * max_n);
    hip_check(hipMemcpyAsync(tg.data(), d_targets, T * sizeof(int), hipMemcpyDeviceToHost, s), "dump");
    hip_check(hipMemcpyAsync(cnt.data(), d_cnt, T * sizeof(int), hipMemcpyDeviceToHost, s), "dump");
    if (max_n > 0)
        hip_check(hipMemcpyAsync(nbr.data(), d_nbr, nbr.size() * sizeof(int), hipMemcpyDeviceToHost, s), "dump");
    hip_check(hipStreamSynchronize(s), "dump");
    for (int t = 0; t < T; ++t) {
        const long long cell = cell_offset + tg[t];
        if (cell < 0 || cell >= dump->n_cells) continue;
        dump->counts[cell] = cnt[t];
        for (int k = 0; k < cnt[t] && k < dump->max_n; ++k) dump->indices[cell * dump->max_n + k] = nbr[(size_t)t * max_n + k];
    }
}

// Bucket key per target (0..7: quad buckets of order 4, 8, ..., 32; 8:
// wider systems) and the buckets' sizes.  A bounded grid strides over the
// targets; each wavefront counts a bucket's lanes with one ballot into the
// workgroup's LDS counts, and each workgroup adds those with one global
// atomic per non-empty bucket (most targets share one bucket, so atomics
// per lane or per small workgroup would serialise on one address).
constexpr int kBuckets = 9;
constexpr int kBucketGrid = 1024;
__global__ void k_bucket_keys(const int *cnt, int T, uint8_t *key, int *count) {
    __shared__ int h[kBuckets];
    if (threadIdx.x < kBuckets) h[threadIdx.x] = 0;
    __syncthreads();
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < T; base += stride) {
        const int t = base + threadIdx.x;
        int k = -1;
        if (t < T) {
            const int n = cnt[t];
            k = n > kQuadMaxMixed ? 8 : (n <= 4 ? 0 : (n + 3) / 4 - 1);
            key[t] = (uint8_t)k;
        }
#pragma unroll
        for (int b = 0; b < kBuckets; ++b) {
            const unsigned long long m = __ballot(k == b);
            if ((threadIdx.x & 63) == 0 && m != 0) atomicAdd(&h[b], (int)__popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < kBuckets && h[threadIdx.x] != 0) atomicAdd(&count[threadIdx.x], h[threadIdx.x]);
}

// Systems per bucket solved first in mixed precision to decide whether the
// bucket's model is too ill-conditioned for FP32 (Gaussian covariance with a
// tiny nugget, near-duplicate samples): when more than half of them are
// handed to FP64, the rest of the bucket goes to FP64 directly.
constexpr int kMixedProbe = 16384;

// FP64 solves of the targets sa.subset[0 .. n_subset): the quad kernel up
// to order 16, lane-per-row kernels above.
void launch_fp64_subset(const SolveArgs &b, int order, hipStream_t s, KernelClock &clk, Stats &st) {
    clk.start();
    if (order <= kQuadMaxFp64 && b.cov.table && !b.correlogram)
        launch_solve_quad(b, order, true, s);
    else
        launch_solve_fp64(b, order <= 32 ? 32 : 64, s);
    hip_check(hipGetLastError(), "solve fp64");
    clk.stop(st.ms_k_solve_fp64, st.n_k_solve_fp64);
}

void launch_mixed_subset(const SolveArgs &b, int order, hipStream_t s, KernelClock &clk, Stats &st) {
    clk.start();
    launch_solve_quad(b, order, false, s);
    hip_check(hipGetLastError(), "solve mixed");
    clk.stop(st.ms_k_solve_mixed, st.n_k_solve_mixed);
}

// Orders below this go to the FP64 quad kernel in mixed mode too: there it
// is both exact and faster than the FP32 factor with FP64 refinement
// (490,000 ordinary-kriging systems of order 4 / 8 / 12 / 16 on MI355X:
// 0.047 / 0.112 / 0.217 / 0.446 ms in FP64 against 0.063 / 0.157 / 0.290 /
// 0.573 ms mixed); from order 20 on the mixed kernels are faster (0.82
// against 3.1 ms at 20, 2.8 against 4.6 ms at 32).
constexpr int kMixedMinOrder = 20;

// Solve all targets, bucketed by neighbour count.  The targets are sorted
// by bucket key once (stable, so each bucket lists its targets in ascending
// order).  Mixed mode: the quad kernels (orders 20..32 with a covariance
// table) in FP32 + FP64 refinement, appending the systems they hand over to
// a per-bucket list; those go through the FP64 kernels after every bucket
// has been launched.  Smaller and wider systems and the analytic-covariance
// fallback go to the FP64 kernels directly.  FP64 mode: every system
// through the FP64 kernels.  Correlogram systems always take the FP64
// kernels.  The host waits for the stream three times plus once per large
// mixed bucket (bucket sizes, each bucket's probe, the fallback counts,
// the end), not once per launch.
void solve(SolveArgs sa, int epi, int precision, DBuf<uint8_t> &status, hipStream_t s, Stats &st) {
    sa.status = status.p;
    sa.subset = nullptr;
    sa.n_subset = 0;
    sa.epi = epi;
    sa.fb_list = nullptr;
    sa.fb_count = nullptr;
    if (sa.T <= 0) return;
    KernelClock clk(s);
    DBuf<uint8_t> keys(sa.T), keys_sorted(sa.T);
    DBuf<int> counts(2 * kBuckets), iota(sa.T), order(sa.T);
    launch_fill_u32(reinterpret_cast<uint32_t *>(counts.p), 0u, 2 * kBuckets, s);
    hipLaunchKernelGGL(k_bucket_keys, dim3(std::min(kBucketGrid, (sa.T + 255) / 256)), dim3(256), 0, s, sa.cnt, sa.T,
                       keys.p, counts.p);
    hipLaunchKernelGGL(k_iota, dim3((sa.T + 255) / 256), dim3(256), 0, s, iota.p, sa.T);
    hip_check(hipGetLastError(), "bucket keys");
    {
        size_t bytes = 0;
        hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.p, keys_sorted.p, iota.p, order.p, sa.T, 0, 4,
                                                     s),
                  "bucket sort");
        DBuf<uint8_t> tmp(bytes);
        hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys.p, keys_sorted.p, iota.p, order.p, sa.T, 0, 4,
                                                     s),
                  "bucket sort");
    }
    int n_k[kBuckets];
    hip_check(hipMemcpyAsync(n_k, counts.p, sizeof(n_k), hipMemcpyDeviceToHost, s), "copy bucket sizes");
    hip_check(hipStreamSynchronize(s), "sync");
    int off[kBuckets + 1];
    off[0] = 0;
    for (int k = 0; k < kBuckets; ++k) off[k + 1] = off[k] + n_k[k];
    if (off[kBuckets] != sa.T) throw std::runtime_error("solve: bucket sizes do not add up");

    // per-bucket fallback lists share the bucket layout of `order`
    DBuf<int> fb_list(sa.T);
    int *fb_count = counts.p + kBuckets;  // zeroed above
    bool any_mixed = false;
    for (int k = 0; k < kBuckets; ++k) {
        const int n = n_k[k];
        if (n == 0) continue;
        const int ord = k == 8 ? 64 : 4 * (k + 1);
        SolveArgs b = sa;
        b.subset = order.p + off[k];
        const bool mixed = precision == 0 && ord >= kMixedMinOrder && k < 8 && sa.cov.table && !sa.correlogram;
        if (!mixed) {
            b.n_subset = n;
            launch_fp64_subset(b, ord, s, clk, st);
            continue;
        }
        any_mixed = true;
        b.fb_list = fb_list.p + off[k];
        b.fb_count = fb_count + k;
        const int probe = n > 4 * kMixedProbe ? kMixedProbe : n;
        b.n_subset = probe;
        launch_mixed_subset(b, ord, s, clk, st);
        if (probe < n) {
            int fb = 0;
            hip_check(hipMemcpyAsync(&fb, fb_count + k, sizeof(int), hipMemcpyDeviceToHost, s), "copy fallbacks");
            hip_check(hipStreamSynchronize(s), "sync");
            b.subset = order.p + off[k] + probe;
            b.n_subset = n - probe;
            if (2 * fb > probe) {
                b.fb_list = nullptr;
                b.fb_count = nullptr;
                launch_fp64_subset(b, ord, s, clk, st);
            } else {
                launch_mixed_subset(b, ord, s, clk, st);
            }
        }
    }
    if (any_mixed) {
        int fbn[kBuckets];
        hip_check(hipMemcpyAsync(fbn, fb_count, sizeof(fbn), hipMemcpyDeviceToHost, s), "copy fallbacks");
        hip_check(hipStreamSynchronize(s), "sync");
        for (int k = 0; k < kBuckets; ++k) {
            if (fbn[k] == 0) continue;
            if (fbn[k] < 0 || fbn[k] > n_k[k]) throw std::runtime_error("solve: fallback count out of range");
            SolveArgs b = sa;
            b.subset = fb_list.p + off[k];
            b.n_subset = fbn[k];
            launch_fp64_subset(b, k == 8 ? 64 : 4 * (k + 1), s, clk, st);
        }
    }
    hip_check(hipStreamSynchronize(s), "sync");
    clk.collect();
}

void collect_counters(const unsigned long long *d_counters, hipStream_t s, Stats &st) {
    std::vector<unsigned long long> h(kCounterWords);
    hip_check(hipMemcpyAsync(h.data(), d_counters, kCounterWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
              "counters");
    hip_check(hipStreamSynchronize(s), "sync");
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < kCounterSlots; ++l)
        for (int w = 0; w < 5; ++w) c[w] += h[(size_t)l * kCounterStride + w];
    st.solves_fp32 += c[0];
    st.solves_fp64 += c[1];
    st.solves_fp32_two_steps += c[4];
    st.nodes_kriged += c[0] + c[1];
    st.nodes_singular += c[2];
    st.nodes_no_neighbours += c[3];
}

}  // namespace

// calc_mean (calc_mean.cpp:6-25): sum of informed values in double, in
// cell order, divided by their count (0 when there are none).
static double device_calc_mean(const float *d_in, const uint8_t *d_mask, long long N, hipStream_t s) {
    std::vector<float> v(N);
    std::vector<uint8_t> m(N);
    hip_check(hipMemcpyAsync(v.data(), d_in, N * sizeof(float), hipMemcpyDeviceToHost, s), "mean");
    hip_check(hipMemcpyAsync(m.data(), d_mask, N, hipMemcpyDeviceToHost, s), "mean");
    hip_check(hipStreamSynchronize(s), "sync");
    double sum = 0;
    int count = 0;
    for (long long i = 0; i < N; ++i)
        if (m[i] != 0) {
            sum += v[i];
            count += 1;
        }
    return count == 0 ? 0.0 : sum / count;
}

// Analytic covariances (ordinary_kriging.cpp, simple_cokriging_markI.cpp)
// as the solve kernels read them: from a table of the model over every
// offset a neighbour pair can take (neighbours lie within 2r-1 of the
// target per axis on the cell-binned path, r on the template path, so pairs
// within 4r-2), built on first use; evaluated on the device when that table
// would be too large.
DevCov analytic_cov(CovAssets &as, const int radii[3], hipStream_t s, Stats &st) {
    int pr[3];
    long long entries = 1;
    for (int i = 0; i < 3; ++i) {
        pr[i] = std::max(2 * radii[i], 4 * radii[i] - 2);
        entries *= 2 * pr[i] + 1;
    }
    if (entries > kMaxPairTable) return make_devcov(as.model, nullptr, nullptr);
    std::lock_guard<std::mutex> g(as.lazy);
    if (!as.have_pair) {
        const double t0 = now_ms();
        as.pair.build(as.model, pr);
        std::vector<double> oct;
        const bool folds = as.pair.fold(oct);
        st.ms_setup_host += now_ms() - t0;
        upload(as.d_pair, as.pair.v, s);
        if (folds) upload(as.d_pair_fold, oct, s);
        hip_check(hipStreamSynchronize(s), "sync");
        as.have_pair = true;
    }
    return make_devcov(as.model, &as.pair, as.d_pair.p, as.d_pair_fold.p);
}

void run_cont(const ContJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st) {
    PoolStream pool_stream(s);
    const double w0 = now_ms();
    validate_search(job.grid, job.radii, job.max_n);
    validate_cov(job.cov);
    const long long N = (long long)job.grid[0] * job.grid[1] * job.grid[2];
    Events ev;
    hip_check(hipEventRecord(ev.e[0], s), "event");

    double mean = job.mean_single;
    if (job.mean_kind == 1 && job.mean_auto) mean = device_calc_mean(job.d_in, job.d_in_mask, N, s);

    DBuf<int> targets, inf;
    const int T = select_indices(IsZero{job.d_in_mask}, (int)N, targets, s);
    const int I = select_indices(IsNonZero{job.d_in_mask}, (int)N, inf, s);
    st.nodes_target += T;

    std::shared_ptr<CovAssets> as = get_assets(job.cov, job.radii, !job.ordinary, s, st);
    DBuf<int> nbr((size_t)T * std::max(job.max_n, 1)), cnt(T);
    hip_check(hipEventRecord(ev.e[1], s), "event");
    search(*as, job.grid, job.radii, job.max_n, job.d_in_mask, targets.p, T, inf.p, I, nbr.p, cnt.p, s, st);
    hip_check(hipEventRecord(ev.e[2], s), "event");
    write_dump(dump, 0, targets.p, T, nbr.p, cnt.p, job.max_n, s);

    DBuf<unsigned long long> counters(kCounterWords);
    hip_check(hipMemsetAsync(counters.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
    DBuf<uint8_t> status(T);
    SolveArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.targets = targets.p;
    sa.T = T;
    sa.nbr = nbr.p;
    sa.cnt = cnt.p;
    sa.max_n = std::max(job.max_n, 1);
    sa.X = job.grid[0];
    sa.Y = job.grid[1];
    sa.Z = job.grid[2];
    if (job.ordinary) {
        sa.cov = analytic_cov(*as, job.radii, s, st);
    } else {
        sa.cov = make_devcov(as->model, &as->table, as->d_table.p, as->d_fold.p);
    }
    sa.ok_system = job.ordinary ? 1 : 0;
    sa.values_f = job.d_in;
    sa.mean_kind = job.mean_kind;
    sa.mean_single = (float)mean;
    sa.mean_cells = job.d_mean;
    sa.mean_on_failure = job.mean_on_failure ? 1 : 0;
    sa.out_f = job.d_out;
    sa.out_mask = job.d_out_mask;
    sa.counters = counters.p;
    hip_check(hipEventRecord(ev.e[3], s), "event");
    if (T > 0) solve(sa, EPI_CONT, precision, status, s, st);
    hip_check(hipEventRecord(ev.e[4], s), "event");
    launch_copy_informed(job.d_in, job.d_in_mask, N, job.d_out, job.d_out_mask, s);
    hip_check(hipEventRecord(ev.e[5], s), "event");
    collect_counters(counters.p, s, st);
    hip_check(hipStreamSynchronize(s), "sync");
    st.ms_search += ev.ms(1, 2);
    st.ms_solve += ev.ms(3, 4);
    st.ms_device_total += ev.ms(0, 5);
    st.ms_wall += now_ms() - w0;
}

void run_ind(const IndJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st) {
    PoolStream pool_stream(s);
    const double w0 = now_ms();
    const long long N = (long long)job.grid[0] * job.grid[1] * job.grid[2];
    const int K = (int)job.cats.size();
    if (K < 1) throw std::runtime_error("indicator kriging needs at least one indicator");
    for (const auto &c : job.cats) {
        validate_search(job.grid, c.radii, c.max_n);
        validate_cov(c.cov);
    }
    Events ev;
    hip_check(hipEventRecord(ev.e[0], s), "event");
    // every cell is estimated, informed ones included (indicator_kriging.h:98-119)
    DBuf<int> targets((size_t)N), inf;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, targets.p, (int)N);
    const int T = (int)N;
    const int I = select_indices(IsNonZero{job.d_in_mask}, (int)N, inf, s);
    st.nodes_target += (long long)T * K;

    DBuf<float> probs(job.median ? 1 : (size_t)K * N);
    DBuf<unsigned long long> counters(kCounterWords);
    hip_check(hipMemsetAsync(counters.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
    DBuf<uint8_t> status(T);
    int maxn_all = 1;
    for (const auto &c : job.cats) maxn_all = std::max(maxn_all, c.max_n);
    {
        // the categories' covariance assets, built on the host side by side
        std::vector<std::pair<CovSpec, std::array<int, 3>>> specs;
        for (const auto &c : job.cats) specs.push_back({c.cov, {c.radii[0], c.radii[1], c.radii[2]}});
        prefetch_assets(specs, true, s, st);
    }
    DBuf<int> nbr((size_t)T * maxn_all), cnt(T);
    for (int k = 0; k < K; ++k) {
        const IndicatorParams &c = job.cats[k];
        std::shared_ptr<CovAssets> as = get_assets(c.cov, c.radii, true, s, st);
        hipEvent_t a0, a1, b0, b1;
        hip_check(hipEventCreate(&a0), "event");
        hip_check(hipEventCreate(&a1), "event");
        hip_check(hipEventCreate(&b0), "event");
        hip_check(hipEventCreate(&b1), "event");
        hip_check(hipEventRecord(a0, s), "event");
        search(*as, job.grid, c.radii, c.max_n, job.d_in_mask, targets.p, T, inf.p, I, nbr.p, cnt.p, s, st);
        hip_check(hipEventRecord(a1, s), "event");
        write_dump(dump, (long long)k * N, targets.p, T, nbr.p, cnt.p, c.max_n, s);
        SolveArgs sa;
        std::memset(&sa, 0, sizeof(sa));
        sa.targets = targets.p;
        sa.T = T;
        sa.nbr = nbr.p;
        sa.cnt = cnt.p;
        sa.max_n = std::max(c.max_n, 1);
        sa.X = job.grid[0];
        sa.Y = job.grid[1];
        sa.Z = job.grid[2];
        sa.cov = make_devcov(as->model, &as->table, as->d_table.p, as->d_fold.p);
        sa.ok_system = 0;
        sa.values_u8 = job.d_in;
        sa.category = job.median ? 1 : k;
        sa.mean_kind = 1;
        // marginal probabilities pass through float (indicator_probability_t)
        sa.mean_single = (float)c.marginal;
        sa.out_f = job.median ? nullptr : probs.p + (size_t)k * N;
        sa.out_u8 = job.d_out;
        sa.out_mask = job.d_out_mask;
        sa.counters = counters.p;
        hip_check(hipEventRecord(b0, s), "event");
        solve(sa, job.median ? EPI_MEDIAN : EPI_IK, precision, status, s, st);
        hip_check(hipEventRecord(b1, s), "event");
        hip_check(hipEventSynchronize(b1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, a0, a1), "elapsed");
        st.ms_search += ms;
        hip_check(hipEventElapsedTime(&ms, b0, b1), "elapsed");
        st.ms_solve += ms;
        (void)hipEventDestroy(a0);
        (void)hipEventDestroy(a1);
        (void)hipEventDestroy(b0);
        (void)hipEventDestroy(b1);
    }
    if (!job.median) launch_ik_argmax(probs.p, K, N, job.d_out, job.d_out_mask, s);
    hip_check(hipEventRecord(ev.e[5], s), "event");
    collect_counters(counters.p, s, st);
    hip_check(hipStreamSynchronize(s), "sync");
    st.ms_device_total += ev.ms(0, 5);
    st.ms_wall += now_ms() - w0;
}

// ---------------------------------------------------------------------------
// Sequential Gaussian simulation (sequential_gaussian_simulation.cpp:18-128,
// sequential_simulation.h:51-110), R realisations per call.
// ---------------------------------------------------------------------------

namespace {

bool is_prime(long long p) {
    if (p < 2) return false;
    for (long long d = 2; d * d <= p; ++d)
        if (p % d == 0) return false;
    return true;
}

// Path LCG of path_random_generator.cpp:21-56 for N cells: A = 1 + the
// product of the distinct primes dividing N (doubled when 4 | N); C = the
// largest prime <= N/3 that does not divide N, else 2.  The path visits
// every cell exactly once iff C does not divide N (Hull-Dobell: A - 1 is a
// multiple of every prime factor of N, and of 4 when 4 | N).
struct Lcg {
    long long A, C;
    bool full;
};
Lcg lcg_for(long long N) {
    long long prod = 1, n = N;
    for (long long p = 2; p * p <= n; ++p)
        if (n % p == 0) {
            prod *= p;
            while (n % p == 0) n /= p;
        }
    if (n > 1) prod *= n;
    if (N % 4 == 0) prod *= 2;
    Lcg l;
    l.A = prod + 1;
    l.C = 2;
    for (long long p = N / 3; p >= 2; --p)
        if (is_prime(p) && N % p != 0) {
            l.C = p;
            break;
        }
    l.full = N == 1 || N % l.C != 0;
    return l;
}

// The reference's loop runs N steps of the LCG and skips cells already
// visited; when the period is short some cells are never visited.  Host
// restatement for that case (grids of a few cells): path[k] = cell at step
// k or -1 for a repeat, rank[cell] = first step or INT_MAX.
void host_path(long long N, const Lcg &l, long long r0, int *path, int *rank) {
    for (long long c = 0; c < N; ++c) rank[c] = 0x7fffffff;
    long long R = r0;
    for (long long k = 0; k < N; ++k) {
        if (rank[R] == 0x7fffffff) {
            rank[R] = (int)k;
            path[k] = (int)R;
        } else {
            path[k] = -1;
        }
        R = (l.A * R + l.C) % N;
    }
}

struct SgsTargetPred {
    const int *key;
    __host__ __device__ bool operator()(const int &cell) const {
        if (cell < 0) return false;
        const int k = key[cell];
        return k >= 0 && k != 0x7fffffff;
    }
};

// Cells of path[0..N) that are simulated, in path order.
int select_path_targets(const int *path, long long N, const int *key, int *out, hipStream_t s) {
    DBuf<int> d_num(1);
    size_t bytes = 0;
    hip_check(hipcub::DeviceSelect::If(nullptr, bytes, path, out, d_num.p, (int)N, SgsTargetPred{key}, s), "select");
    DBuf<uint8_t> tmp(bytes);
    hip_check(hipcub::DeviceSelect::If(tmp.p, bytes, path, out, d_num.p, (int)N, SgsTargetPred{key}, s), "select");
    int num = 0;
    hip_check(hipMemcpyAsync(&num, d_num.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy count");
    hip_check(hipStreamSynchronize(s), "sync");
    return num;
}

__global__ void k_sgs_out_mask(const int *key, long long n, uint8_t *mask) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[i] = key[i] != 0x7fffffff ? 1 : 0;
}

// Path LCGs and mt11213b seeds of all realisations of a call.
struct SimPaths {
    Lcg lcg;
    std::vector<long long> h_R0;
    DBuf<long long> d_A, d_C, d_R0;
    DBuf<uint32_t> d_mt_seed;
};

void sim_paths(long long N, const std::vector<long> &seeds, SimPaths &p, hipStream_t s) {
    const int R = (int)seeds.size();
    p.lcg = lcg_for(N);
    std::vector<long long> h_A(R, p.lcg.A), h_C(R, p.lcg.C);
    std::vector<uint32_t> h_mt_seed(R);
    p.h_R0.resize(R);
    for (int r = 0; r < R; ++r) {
        // PathGen takes the seed through int (sequential_gaussian_simulation.cpp:45,
        // sequential_indicator_simulation.cpp:57)
        const long long r0 = (long long)(int)seeds[r] % N;
        if (r0 < 0) throw std::runtime_error("the seed, taken as int, must not be negative");
        p.h_R0[r] = r0;
        h_mt_seed[r] = (uint32_t)seeds[r];
    }
    upload(p.d_A, h_A, s);
    upload(p.d_C, h_C, s);
    upload(p.d_R0, p.h_R0, s);
    upload(p.d_mt_seed, h_mt_seed, s);
}

// Realisations r0 .. r0 + rc - 1: key of every cell (-1 hard data, path
// rank of a cell to simulate, INT_MAX never informed) and the simulated
// cells of each realisation in path order.
struct SimBatch {
    DBuf<int> key, targets;
    // per realisation, the positions of its targets ordered by cell: the
    // template scans walk the targets in that order so that the probes of a
    // wavefront's 64 lanes fall on neighbouring cells
    DBuf<int> by_cell;
    std::vector<int> h_T;
    std::vector<long long> h_off;
    long long tot = 0;
    int max_T = 0;
};

// Wavefronts of an ordered sweep, all realisations together: enough lanes
// to cover the ready cells of the path (a dependency level of a 4 M-cell
// path holds ~100 k cells) while keeping the flag polls few (more
// wavefronts were measured slower: each polls the L2).
constexpr int kSweepWaves = 1024;

// A second stream of default priority on which the draws are generated
// while the searches and solves run on the job's stream.
struct SideStream {
    hipStream_t s2 = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    SideStream() {
        hip_check(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreate");
    }
    ~SideStream() {
        (void)hipStreamSynchronize(s2);
        (void)hipEventDestroy(ready);
        (void)hipEventDestroy(done);
        (void)hipStreamDestroy(s2);
    }
    SideStream(const SideStream &) = delete;
    SideStream &operator=(const SideStream &) = delete;
};

// mt11213b draws of the batch, one per simulated cell in path order, on the
// side stream after the work queued so far on s; s waits for them at
// sim_draws_wait.
void sim_draws(const SimPaths &p, const SimBatch &b, int r0, DBuf<uint32_t> &draws, DBuf<long long> &d_counts,
               SideStream &side, hipStream_t s) {
    const int rc = (int)b.h_T.size();
    draws.alloc((size_t)rc * std::max(b.max_T, 1));
    std::vector<long long> h_counts(b.h_T.begin(), b.h_T.end());
    upload(d_counts, h_counts, s);
    hip_check(hipEventRecord(side.ready, s), "event");
    hip_check(hipStreamWaitEvent(side.s2, side.ready, 0), "wait event");
    launch_mt11213b(p.d_mt_seed.p + r0, d_counts.p, std::max(b.max_T, 1), rc, draws.p, side.s2);
    hip_check(hipGetLastError(), "mt11213b");
    hip_check(hipEventRecord(side.done, side.s2), "event");
}

void sim_draws_wait(SideStream &side, hipStream_t s) { hip_check(hipStreamWaitEvent(s, side.done, 0), "wait event"); }

void sim_batch(long long N, const SimPaths &p, int r0, int rc, const uint8_t *hard, const uint8_t *sim_mask,
               SimBatch &b, hipStream_t s) {
    DBuf<int> path((size_t)rc * N), rank((size_t)rc * N);
    b.key.alloc((size_t)rc * N);
    if (p.lcg.full) {
        SgsPathArgs pa;
        pa.N = N;
        pa.A = p.d_A.p + r0;
        pa.C = p.d_C.p + r0;
        pa.R0 = p.d_R0.p + r0;
        pa.path = path.p;
        pa.rank = rank.p;
        launch_lcg_path(pa, rc, s);
    } else {
        std::vector<int> hp((size_t)rc * N), hr((size_t)rc * N);
        for (int r = 0; r < rc; ++r)
            host_path(N, p.lcg, p.h_R0[r0 + r], hp.data() + (size_t)r * N, hr.data() + (size_t)r * N);
        hip_check(hipMemcpyAsync(path.p, hp.data(), hp.size() * sizeof(int), hipMemcpyHostToDevice, s), "upload");
        hip_check(hipMemcpyAsync(rank.p, hr.data(), hr.size() * sizeof(int), hipMemcpyHostToDevice, s), "upload");
        hip_check(hipStreamSynchronize(s), "sync");
    }
    launch_sgs_key(hard, sim_mask, rank.p, N, rc, b.key.p, s);
    hip_check(hipGetLastError(), "path kernels");
    rank.release();
    b.h_T.assign(rc, 0);
    b.h_off.assign(rc, 0);
    b.targets.alloc((size_t)rc * N);
    b.tot = 0;
    for (int r = 0; r < rc; ++r) {
        b.h_off[r] = b.tot;
        b.h_T[r] = select_path_targets(path.p + (size_t)r * N, N, b.key.p + (size_t)r * N, b.targets.p + b.tot, s);
        b.tot += b.h_T[r];
    }
    if (b.tot >= (1ll << 31)) throw std::runtime_error("too many simulated cells in one batch");
    b.max_T = 0;
    for (int r = 0; r < rc; ++r) b.max_T = std::max(b.max_T, b.h_T[r]);
    b.by_cell.alloc((size_t)std::max<long long>(b.tot, 1));
    if (b.max_T > 0) {
        // positions ordered by (density bucket, cell); grids of more than
        // 2^28 cells sort by cell alone
        DBuf<int> pos(b.max_T);
        DBuf<uint32_t> keys_in(b.max_T), keys_out(b.max_T);
        hipLaunchKernelGGL(k_iota, dim3((unsigned)((b.max_T + 255) / 256)), dim3(256), 0, s, pos.p, b.max_T);
        int cell_bits = 1;
        while ((1ll << cell_bits) < N) ++cell_bits;
        const bool buckets = cell_bits <= 28;
        const int end_bit = buckets ? cell_bits + 4 : cell_bits;
        size_t bytes = 0;
        hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in.p, keys_out.p, pos.p, b.by_cell.p,
                                                     b.max_T, 0, end_bit, s),
                  "radix sort");
        DBuf<uint8_t> tmp(bytes);
        // hard-data cells (a simulation mask leaves other cells uninformed)
        DBuf<unsigned long long> d_hard(1);
        hip_check(hipMemsetAsync(d_hard.p, 0, sizeof(unsigned long long), s), "memset");
        hipLaunchKernelGGL(k_count_nonzero, dim3((unsigned)std::min<long long>((N + 255) / 256, 1024)), dim3(256), 0,
                           s, hard, N, d_hard.p);
        hip_check(hipGetLastError(), "count hard data");
        unsigned long long h_hard = 0;
        hip_check(hipMemcpyAsync(&h_hard, d_hard.p, sizeof(h_hard), hipMemcpyDeviceToHost, s), "copy count");
        hip_check(hipStreamSynchronize(s), "sync");
        const long long H = std::max<long long>(1, (long long)h_hard);
        for (int r = 0; r < rc; ++r) {
            const int T = b.h_T[r];
            if (T == 0) continue;
            hipLaunchKernelGGL(k_scan_order_keys, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s,
                               b.targets.p + b.h_off[r], T, buckets ? H : (1ll << 62), cell_bits, keys_in.p);
            hip_check(hipGetLastError(), "scan order keys");
            size_t nb = bytes;
            hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.p, nb, keys_in.p, keys_out.p, pos.p,
                                                         b.by_cell.p + b.h_off[r], T, 0, end_bit, s),
                      "radix sort");
        }
    }
}

// Template scan of every target of the batch among the cells informed
// before it (the reference's simulations search with the template only,
// sequential_simulation.h:74, sequential_indicator_simulation.cpp:53).
void sim_scan(const CovAssets &as, const int grid[3], int max_n, const SimBatch &b, int *nbr, int *cnt,
              uint64_t *simbits, hipStream_t s, Stats &st) {
    const long long N = (long long)grid[0] * grid[1] * grid[2];
    const int stride = std::max(max_n, 1);
    for (size_t r = 0; r < b.h_T.size(); ++r) {
        if (b.h_T[r] == 0) continue;
        ScanArgs sa;
        sa.mask = nullptr;
        sa.key = b.key.p + r * N;
        sa.X = grid[0];
        sa.Y = grid[1];
        sa.Z = grid[2];
        sa.tmpl = (const int4 *)as.d_tmpl.p;
        sa.tsize = as.tmpl.size();
        sa.max_n = max_n;
        sa.targets = b.targets.p + b.h_off[r];
        sa.list = b.by_cell.p + b.h_off[r];
        sa.n_list = b.h_T[r];
        sa.nbr = nbr + (size_t)b.h_off[r] * stride;
        sa.cnt = cnt + b.h_off[r];
        sa.simbits = simbits + b.h_off[r];
        launch_template_scan(sa, 1, s);
        hip_check(hipGetLastError(), "template scan");
        st.nodes_template_path += b.h_T[r];
    }
}

// Realisations per batch from free device memory, bytes_per_cell per
// realisation and cell.
int sim_batch_size(int R, size_t bytes_per_cell, long long N) {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t per_real = (size_t)N * bytes_per_cell;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)R, (size_t)(free_b * 0.6) / per_real));
}

}  // namespace

void run_sgs(const SgsJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st) {
    PoolStream pool_stream(s);
    const double w0 = now_ms();
    validate_search(job.grid, job.radii, job.max_n);
    validate_cov(job.cov);
    const int R = (int)job.seeds.size();
    if (R < 1) throw std::runtime_error("sequential simulation needs at least one seed");
    if (job.has_cdf && (job.cdf_values.empty() || job.cdf_values.size() != job.cdf_probs.size()))
        throw std::runtime_error("the non-parametric cdf must have values and probabilities of one positive length");
    const long long N = (long long)job.grid[0] * job.grid[1] * job.grid[2];
    const int max_n = std::max(job.max_n, 1);
    Events ev;
    hip_check(hipEventRecord(ev.e[0], s), "event");
    SimPaths paths;
    sim_paths(N, job.seeds, paths, s);
    SideStream side;

    // normal-score transform of the conditioning data and of the LVM mean
    DBuf<float> cv, cp, data((size_t)N), lvm;
    hip_check(hipMemcpyAsync(data.p, job.d_in, N * sizeof(float), hipMemcpyDeviceToDevice, s), "copy data");
    if (job.d_lvm) {
        lvm.alloc((size_t)N);
        hip_check(hipMemcpyAsync(lvm.p, job.d_lvm, N * sizeof(float), hipMemcpyDeviceToDevice, s), "copy mean");
    }
    if (job.has_cdf) {
        upload(cv, job.cdf_values, s);
        upload(cp, job.cdf_probs, s);
        const long long m = (long long)job.cdf_values.size();
        launch_cdf_forward(data.p, job.d_in_mask, N, cv.p, cp.p, m, s);
        if (job.d_lvm) launch_cdf_forward(lvm.p, nullptr, N, cv.p, cp.p, m, s);
        hip_check(hipGetLastError(), "cdf forward");
    }

    std::shared_ptr<CovAssets> as = get_assets(job.cov, job.radii, true, s, st);
    // per target: lists, weights, counts, draws and the solve's bucket sort
    // and hand-over lists (about 20 bytes)
    const int RC = sim_batch_size(R, 17 + (12 * (size_t)max_n + 52), N);

    DBuf<unsigned long long> counters(kCounterWords);
    hip_check(hipMemsetAsync(counters.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
    DBuf<unsigned int> err(1);

    for (int r0 = 0; r0 < R; r0 += RC) {
        const int rc = std::min(RC, R - r0);
        SimBatch b;
        sim_batch(N, paths, r0, rc, job.d_in_mask, job.d_sim_mask, b, s);
        const long long tot = b.tot;
        st.nodes_target += tot;
        DBuf<uint32_t> draws;
        DBuf<long long> d_counts;
        sim_draws(paths, b, r0, draws, d_counts, side, s);

        // neighbours of every target among the cells informed before it
        DBuf<int> nbr((size_t)std::max<long long>(tot, 1) * max_n), cnt((size_t)std::max<long long>(tot, 1));
        DBuf<uint64_t> simbits((size_t)std::max<long long>(tot, 1));
        hip_check(hipEventRecord(ev.e[1], s), "event");
        sim_scan(*as, job.grid, job.max_n, b, nbr.p, cnt.p, simbits.p, s, st);
        hip_check(hipEventRecord(ev.e[2], s), "event");
        hip_check(hipEventSynchronize(ev.e[2]), "sync");
        st.ms_k_template += ev.ms(1, 2);
        st.ms_search += ev.ms(1, 2);
        st.n_k_template += rc;
        if (dump)
            for (int r = 0; r < rc; ++r)
                write_dump(dump, (long long)(r0 + r) * N, b.targets.p + b.h_off[r], b.h_T[r],
                           nbr.p + (size_t)b.h_off[r] * max_n, cnt.p + b.h_off[r], max_n, s);

        // weights and variances
        DBuf<double> w((size_t)std::max<long long>(tot, 1) * max_n), var((size_t)std::max<long long>(tot, 1));
        DBuf<uint8_t> status((size_t)std::max<long long>(tot, 1));
        if (tot > 0) {
            SolveArgs sa;
            std::memset(&sa, 0, sizeof(sa));
            sa.targets = b.targets.p;
            sa.T = (int)tot;
            sa.nbr = nbr.p;
            sa.cnt = cnt.p;
            sa.max_n = max_n;
            sa.X = job.grid[0];
            sa.Y = job.grid[1];
            sa.Z = job.grid[2];
            sa.cov = make_devcov(as->model, &as->table, as->d_table.p, as->d_fold.p);
            sa.ok_system = job.ordinary ? 1 : 0;
            sa.mean_kind = 0;
            sa.out_var = var.p;
            sa.out_w = w.p;
            sa.counters = counters.p;
            hip_check(hipEventRecord(ev.e[3], s), "event");
            solve(sa, EPI_SGS, precision, status, s, st);
            hip_check(hipEventRecord(ev.e[4], s), "event");
            hip_check(hipEventSynchronize(ev.e[4]), "sync");
            st.ms_solve += ev.ms(3, 4);
        }

        const int max_T = b.max_T;
        // values: conditioning data everywhere, then the ordered sweep
        float *vals = job.d_out + (size_t)r0 * N;
        for (int r = 0; r < rc; ++r)
            hip_check(hipMemcpyAsync(vals + (size_t)r * N, data.p, N * sizeof(float), hipMemcpyDeviceToDevice, s),
                      "copy data");
        DBuf<int> done((size_t)rc * N);
        hip_check(hipMemsetAsync(done.p, 0, (size_t)rc * N * sizeof(int), s), "memset");
        DBuf<unsigned int> ticket(rc);
        hip_check(hipMemsetAsync(ticket.p, 0, rc * sizeof(unsigned int), s), "memset");
        hip_check(hipMemsetAsync(err.p, 0, sizeof(unsigned int), s), "memset");
        DBuf<int> d_T;
        DBuf<long long> d_off;
        upload(d_T, b.h_T, s);
        upload(d_off, b.h_off, s);
        sim_draws_wait(side, s);
        if (tot > 0) {
            SgsSweepArgs a;
            a.N = N;
            a.max_n = max_n;
            a.T = d_T.p;
            a.t_off = d_off.p;
            a.targets = b.targets.p;
            a.nbr = nbr.p;
            a.cnt = cnt.p;
            a.w = w.p;
            a.var = var.p;
            a.status = status.p;
            a.draws = draws.p;
            a.draw_stride = std::max(max_T, 1);
            a.simbits = simbits.p;
            a.lvm = job.d_lvm ? lvm.p : nullptr;
            a.values = vals;
            a.done = done.p;
            a.ticket = ticket.p;
            a.error = err.p;
            a.spin_limit = 1ll << 26;
            const int waves = (int)std::max<long long>(1, std::min<long long>((max_T + 63) / 64, kSweepWaves / rc));
            launch_sgs_sweep(a, rc, waves, s);
            hip_check(hipGetLastError(), "sweep");
        }
        hipLaunchKernelGGL(k_sgs_out_mask, dim3((unsigned)(((long long)rc * N + 255) / 256)), dim3(256), 0, s,
                           b.key.p, (long long)rc * N, job.d_out_mask + (size_t)r0 * N);
        if (job.has_cdf)
            launch_cdf_back(vals, job.d_out_mask + (size_t)r0 * N, (long long)rc * N, cv.p, cp.p,
                            (long long)job.cdf_values.size(), s);
        unsigned int h_err = 0;
        hip_check(hipMemcpyAsync(&h_err, err.p, sizeof(h_err), hipMemcpyDeviceToHost, s), "copy error");
        hip_check(hipStreamSynchronize(s), "sync");
        if (h_err) throw std::runtime_error("sequential simulation: a cell waited on its neighbours past the limit");
    }
    hip_check(hipEventRecord(ev.e[5], s), "event");
    collect_counters(counters.p, s, st);
    hip_check(hipStreamSynchronize(s), "sync");
    st.ms_device_total += ev.ms(0, 5);
    st.ms_wall += now_ms() - w0;
}

// ---------------------------------------------------------------------------
// Sequential indicator simulation (sequential_indicator_simulation.cpp:33-187),
// R realisations per call.  The path, keys, draws and template scans are
// those of the Gaussian simulation; each kriged category has its own
// covariance table, template, neighbour lists and simple-kriging weights,
// and k_sis_sweep samples the categories in path order.
// ---------------------------------------------------------------------------

namespace {

struct SisCategoryBuffers {
    DBuf<int> nbr, cnt;
    DBuf<uint64_t> simbits;
    DBuf<double> w, var;
    DBuf<uint8_t> status;
};

}  // namespace

void run_sis(const SisJob &job, hipStream_t s, int precision, Stats &st) {
    PoolStream pool_stream(s);
    const double w0 = now_ms();
    const int K = (int)job.cats.size();
    if (K < 2) throw std::runtime_error("sequential indicator simulation needs at least two categories");
    if (K > kSisMaxCategories) throw std::runtime_error("sequential indicator simulation supports at most 256 categories");
    for (const auto &c : job.cats) {
        validate_search(job.grid, c.radii, c.max_n);
        validate_cov(c.cov);
    }
    const bool lvm = !job.d_lvm.empty();
    if (lvm && (int)job.d_lvm.size() != K) throw std::runtime_error("one local mean per category is needed");
    if (job.correlogram && !lvm) throw std::runtime_error("correlogram weights need local means");
    const int R = (int)job.seeds.size();
    if (R < 1) throw std::runtime_error("sequential simulation needs at least one seed");
    const long long N = (long long)job.grid[0] * job.grid[1] * job.grid[2];
    // median SIS: only the first category is kriged (sequential_indicator_simulation.cpp:80-100)
    const int kriged = K == 2 ? 1 : K;
    Events ev;
    hip_check(hipEventRecord(ev.e[0], s), "event");
    SimPaths paths;
    sim_paths(N, job.seeds, paths, s);
    SideStream side;

    std::vector<std::shared_ptr<CovAssets>> as(kriged);
    {
        std::vector<std::pair<CovSpec, std::array<int, 3>>> specs;
        for (int c = 0; c < kriged; ++c)
            specs.push_back({job.cats[c].cov, {job.cats[c].radii[0], job.cats[c].radii[1], job.cats[c].radii[2]}});
        prefetch_assets(specs, true, s, st);
    }
    size_t per_cell = 21 + 20;  // and the solve's bucket sort and hand-over lists
    for (int c = 0; c < kriged; ++c) {
        as[c] = get_assets(job.cats[c].cov, job.cats[c].radii, true, s, st);
        per_cell += 12 * (size_t)std::max(job.cats[c].max_n, 1) + 32;
    }
    const int RC = sim_batch_size(R, per_cell, N);

    DBuf<unsigned long long> counters(kCounterWords);
    hip_check(hipMemsetAsync(counters.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
    DBuf<unsigned int> err(1);

    for (int r0 = 0; r0 < R; r0 += RC) {
        const int rc = std::min(RC, R - r0);
        SimBatch b;
        sim_batch(N, paths, r0, rc, job.d_in_mask, job.d_sim_mask, b, s);
        const long long tot = b.tot;
        const size_t tot1 = (size_t)std::max<long long>(tot, 1);
        st.nodes_target += tot;

        DBuf<uint32_t> draws;
        DBuf<long long> d_counts;
        sim_draws(paths, b, r0, draws, d_counts, side, s);

        std::vector<std::unique_ptr<SisCategoryBuffers>> cb(kriged);
        std::vector<SisCategory> h_cats(kriged);
        for (int c = 0; c < kriged; ++c) {
            const IndicatorParams &p = job.cats[c];
            const int max_n = std::max(p.max_n, 1);
            cb[c].reset(new SisCategoryBuffers);
            SisCategoryBuffers &B = *cb[c];
            B.nbr.alloc(tot1 * max_n);
            B.cnt.alloc(tot1);
            B.simbits.alloc(tot1);
            B.w.alloc(tot1 * max_n);
            B.var.alloc(tot1);
            B.status.alloc(tot1);
            hip_check(hipEventRecord(ev.e[1], s), "event");
            sim_scan(*as[c], job.grid, p.max_n, b, B.nbr.p, B.cnt.p, B.simbits.p, s, st);
            hip_check(hipEventRecord(ev.e[2], s), "event");
            hip_check(hipEventSynchronize(ev.e[2]), "sync");
            st.ms_k_template += ev.ms(1, 2);
            st.ms_search += ev.ms(1, 2);
            st.n_k_template += rc;
            if (tot > 0) {
                // simple-kriging weights (sk_weight_calculator_t or
                // corellogram_weight_calculator_t, kriging_interpolation.h:19-200)
                SolveArgs sa;
                std::memset(&sa, 0, sizeof(sa));
                sa.targets = b.targets.p;
                sa.T = (int)tot;
                sa.nbr = B.nbr.p;
                sa.cnt = B.cnt.p;
                sa.max_n = max_n;
                sa.X = job.grid[0];
                sa.Y = job.grid[1];
                sa.Z = job.grid[2];
                sa.cov = make_devcov(as[c]->model, &as[c]->table, as[c]->d_table.p, as[c]->d_fold.p);
                sa.ok_system = 0;
                sa.mean_kind = lvm ? 2 : 0;
                sa.mean_cells = lvm ? job.d_lvm[c] : nullptr;
                sa.correlogram = job.correlogram ? 1 : 0;
                sa.out_var = B.var.p;
                sa.out_w = B.w.p;
                sa.counters = counters.p;
                hip_check(hipEventRecord(ev.e[3], s), "event");
                solve(sa, EPI_SGS, precision, B.status, s, st);
                hip_check(hipEventRecord(ev.e[4], s), "event");
                hip_check(hipEventSynchronize(ev.e[4]), "sync");
                st.ms_solve += ev.ms(3, 4);
            }
            SisCategory &h = h_cats[c];
            h.nbr = B.nbr.p;
            h.cnt = B.cnt.p;
            h.w = B.w.p;
            h.status = B.status.p;
            h.simbits = B.simbits.p;
            h.lvm = lvm ? job.d_lvm[c] : nullptr;
            // marginal probabilities pass through float (ik_params.h:23)
            h.marginal = (float)p.marginal;
            h.max_n = max_n;
        }
        DBuf<SisCategory> d_cats;
        upload(d_cats, h_cats, s);

        const int max_T = b.max_T;
        // categories: conditioning data everywhere, then the ordered sweep
        uint8_t *vals = job.d_out + (size_t)r0 * N;
        for (int r = 0; r < rc; ++r)
            hip_check(hipMemcpyAsync(vals + (size_t)r * N, job.d_in, N, hipMemcpyDeviceToDevice, s), "copy data");
        DBuf<int> done((size_t)rc * N);
        hip_check(hipMemsetAsync(done.p, 0, (size_t)rc * N * sizeof(int), s), "memset");
        DBuf<unsigned int> ticket(rc);
        hip_check(hipMemsetAsync(ticket.p, 0, rc * sizeof(unsigned int), s), "memset");
        hip_check(hipMemsetAsync(err.p, 0, sizeof(unsigned int), s), "memset");
        DBuf<int> d_T;
        DBuf<long long> d_off;
        upload(d_T, b.h_T, s);
        upload(d_off, b.h_off, s);
        sim_draws_wait(side, s);
        if (tot > 0) {
            SisSweepArgs a;
            a.N = N;
            a.kriged = kriged;
            a.K = K;
            a.cats = d_cats.p;
            a.T = d_T.p;
            a.t_off = d_off.p;
            a.targets = b.targets.p;
            a.draws = draws.p;
            a.draw_stride = std::max(max_T, 1);
            a.values = vals;
            a.done = done.p;
            a.ticket = ticket.p;
            a.error = err.p;
            a.spin_limit = 1ll << 26;
            const int waves = (int)std::max<long long>(1, std::min<long long>((max_T + 63) / 64, kSweepWaves / rc));
            launch_sis_sweep(a, rc, waves, s);
            hip_check(hipGetLastError(), "sweep");
        }
        hipLaunchKernelGGL(k_sgs_out_mask, dim3((unsigned)(((long long)rc * N + 255) / 256)), dim3(256), 0, s,
                           b.key.p, (long long)rc * N, job.d_out_mask + (size_t)r0 * N);
        unsigned int h_err = 0;
        hip_check(hipMemcpyAsync(&h_err, err.p, sizeof(h_err), hipMemcpyDeviceToHost, s), "copy error");
        hip_check(hipStreamSynchronize(s), "sync");
        if (h_err) throw std::runtime_error("sequential simulation: a cell waited on its neighbours past the limit");
    }
    hip_check(hipEventRecord(ev.e[5], s), "event");
    collect_counters(counters.p, s, st);
    hip_check(hipStreamSynchronize(s), "sync");
    st.ms_device_total += ev.ms(0, 5);
    st.ms_wall += now_ms() - w0;
}

// ---------------------------------------------------------------------------
// Simple colocated cokriging, Markov models I and II
// (simple_cokriging_markI.cpp:185-305): template search with the analytic
// primary model (sugarbox_neighbour_lookup.h:36-61, no cell-binned path),
// one (n+1)-system per uninformed cell solved by gauss_solve in FP64.
// ---------------------------------------------------------------------------
void run_cokriging(const CokrigingJob &job, hipStream_t s, const Dump *dump, Stats &st) {
    PoolStream pool_stream(s);
    const double w0 = now_ms();
    validate_search(job.grid, job.radii, job.max_n);
    validate_cov(job.primary);
    if (job.mark2) validate_cov(job.secondary);
    const long long N = (long long)job.grid[0] * job.grid[1] * job.grid[2];
    const int max_n = std::max(job.max_n, 1);
    Events ev;
    hip_check(hipEventRecord(ev.e[0], s), "event");
    DBuf<int> targets;
    const int T = select_indices(IsZero{job.d_in_mask}, (int)N, targets, s);
    st.nodes_target += T;
    std::shared_ptr<CovAssets> as = get_assets(job.primary, job.radii, false, s, st);
    DBuf<int> nbr((size_t)std::max(T, 1) * max_n), cnt((size_t)std::max(T, 1));
    hip_check(hipEventRecord(ev.e[1], s), "event");
    if (T > 0) {
        ScanArgs sa;
        sa.mask = job.d_in_mask;
        sa.key = nullptr;
        sa.X = job.grid[0];
        sa.Y = job.grid[1];
        sa.Z = job.grid[2];
        sa.tmpl = (const int4 *)as->d_tmpl.p;
        sa.tsize = as->tmpl.size();
        sa.max_n = job.max_n;
        sa.targets = targets.p;
        sa.list = nullptr;
        sa.n_list = T;
        sa.nbr = nbr.p;
        sa.cnt = cnt.p;
        sa.simbits = nullptr;
        KTimer kt(s);
        launch_template_scan(sa, 0, s);
        hip_check(hipGetLastError(), "template scan");
        st.ms_k_template += kt.stop();
        st.n_k_template += 1;
        st.nodes_template_path += T;
    }
    hip_check(hipEventRecord(ev.e[2], s), "event");
    write_dump(dump, 0, targets.p, T, nbr.p, cnt.p, job.max_n, s);

    DBuf<unsigned long long> counters(kCounterWords);
    hip_check(hipMemsetAsync(counters.p, 0, kCounterWords * sizeof(unsigned long long), s), "memset");
    CokrigingArgs a;
    std::memset(&a, 0, sizeof(a));
    a.targets = targets.p;
    a.T = T;
    a.nbr = nbr.p;
    a.cnt = cnt.p;
    a.max_n = max_n;
    a.X = job.grid[0];
    a.Y = job.grid[1];
    a.Z = job.grid[2];
    a.cov = analytic_cov(*as, job.radii, s, st);
    // C(0) of a model is sill - nugget (cov_model.h:144-181 at h = 0)
    const double c1 = as->model.c0();
    std::shared_ptr<CovAssets> as2;
    if (job.mark2) {
        // cross_cov_model_mark_ii_t (simple_cokriging_markI.cpp:162-183):
        // the secondary model's values at target-neighbour offsets, from
        // its box table over the search radii
        as2 = get_assets(job.secondary, job.radii, true, s, st);
        const double c2 = as2->model.c0();
        a.cross = make_devcov(as2->model, &as2->table, as2->d_table.p, as2->d_fold.p);
        a.sec_var = c2;
        a.coef = job.correlation * sqrt(c2 / c1);
        a.cross0 = a.coef * c2;
    } else {
        // cross_cov_model_mark_i_t (simple_cokriging_markI.cpp:143-160)
        a.cross = a.cov;
        a.sec_var = job.secondary_variance;
        a.coef = job.correlation * sqrt(job.secondary_variance) / sqrt(c1);
        a.cross0 = a.coef * c1;
    }
    a.primary_mean = job.primary_mean;
    a.secondary_mean = job.secondary_mean;
    a.values = job.d_in;
    a.sec = job.d_sec;
    a.sec_mask = job.d_sec_mask;
    a.out = job.d_out;
    a.out_mask = job.d_out_mask;
    a.counters = counters.p;
    hip_check(hipEventRecord(ev.e[3], s), "event");
    {
        KTimer kt(s);
        launch_cokriging(a, s);
        hip_check(hipGetLastError(), "cokriging");
        st.ms_k_solve_fp64 += kt.stop();
        st.n_k_solve_fp64 += T > 0 ? 1 : 0;
    }
    hip_check(hipEventRecord(ev.e[4], s), "event");
    launch_copy_informed(job.d_in, job.d_in_mask, N, job.d_out, job.d_out_mask, s);
    hip_check(hipEventRecord(ev.e[5], s), "event");
    collect_counters(counters.p, s, st);
    hip_check(hipStreamSynchronize(s), "sync");
    st.ms_search += ev.ms(1, 2);
    st.ms_solve += ev.ms(3, 4);
    st.ms_device_total += ev.ms(0, 5);
    st.ms_wall += now_ms() - w0;
}

}  // namespace hpgl_amd
