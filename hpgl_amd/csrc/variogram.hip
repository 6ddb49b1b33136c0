// variogram.hip - experimental semivariograms (the reference's _cvariogram
// module: variograms.cpp, ellipsoid.cpp, stack_layers.h), built as
// hpgl_amd/_cvariogram.so.
//
// calc_variograms (variograms.cpp:362-446) visits every integer offset of
// the search window that lies in the tunnel (is_in_tunnel, :26-40) and
// every informed cell whose partner at that offset is inside the grid and
// informed, adding (v1 - v2)^2 to each lag whose [distance - width/2,
// distance + width/2] holds |offset| (update_lags, :331-360).  Here the
// offsets and their lag ranges are listed on the host in the reference's
// loop order; on the device one thread takes one informed cell and a batch
// of 32 offsets, keeps a double sum and a count per offset in registers,
// and each workgroup reduces them in a fixed order into per-chunk partials,
// which are summed in chunk order.  The sums are deterministic; they are
// associated differently from the reference's single running sum, so a lag
// agrees with it to ~1e-15 relative before the final cast to float.
//
// calc_variograms_from_point_set (:448-502): one thread per first point
// scans all second points, keeping per-lag sums in a per-point row of a
// global array; the rows are summed in point order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cvariogram.h"

#define CVAR_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string t_error;
thread_local long long t_offsets = 0, t_pairs = 0;
thread_local double t_device_ms = 0;

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    explicit DevBuf(size_t n) { hip_ok(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    ~DevBuf() { (void)hipFree(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Device buffers, stream and events of the gridded variogram for the
// calling thread on one device, kept between calls: a buffer grows when a
// call needs more and is otherwise reused; all of it is released when the
// thread exits.
struct Grow {
    void *p = nullptr;
    size_t n = 0;
    template <typename T>
    T *get(size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (bytes > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            hip_ok(hipMalloc(&p, bytes), "hipMalloc");
            n = bytes;
        }
        return static_cast<T *>(p);
    }
};
struct Workspace {
    Grow data, mask, vals, cells, offs, psum, pcnt, osum, ocnt;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0;
    Workspace() = default;
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    ~Workspace() {
        if (!s) return;
        int cur = 0;
        const bool switch_dev = hipGetDevice(&cur) == hipSuccess && cur != dev;
        if (switch_dev) (void)hipSetDevice(dev);
        (void)hipStreamSynchronize(s);
        for (Grow *g : {&data, &mask, &vals, &cells, &offs, &psum, &pcnt, &osum, &ocnt})
            if (g->p) (void)hipFree(g->p);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        if (switch_dev) (void)hipSetDevice(cur);
    }
};
Workspace &workspace() {
    static thread_local std::map<int, Workspace> per_device;
    int dev = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    Workspace &w = per_device[dev];
    if (!w.s) {
        w.dev = dev;
        hip_ok(hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking), "hipStreamCreate");
        hip_ok(hipEventCreate(&w.e0), "hipEventCreate");
        hip_ok(hipEventCreate(&w.e1), "hipEventCreate");
    }
    return w;
}

// ---- host restatements of the geometry (variograms.cpp:9-149) ----

double dot3(const vector_t &a, const vector_t &b) {
    double r = 0.0;
    for (int i = 0; i < 3; ++i) r += a.m_data[i] * b.m_data[i];
    return r;
}

bool in_tunnel(const variogram_search_template_t &t, const vector_t &v) {
    const double ss1 = dot3(v, t.m_ellipsoid.m_direction1);
    const double ss2 = dot3(v, t.m_ellipsoid.m_direction2);
    const double ss3 = dot3(v, t.m_ellipsoid.m_direction3);
    const double s2 = ss2 / t.m_ellipsoid.m_R2;
    const double s3 = ss3 / t.m_ellipsoid.m_R3;
    const double dist = sqrt(s2 * s2 + s3 * s3);
    return dist <= 1.0 && t.m_tol_distance * dist <= ss1;
}

struct Window {
    double min[3], max[3];
};

// calc_search_template_window (:84-128): bounding box of the tunnel's
// corners for i in {0, 1}, j, k in {-1, 1}
Window search_window(const variogram_search_template_t &t) {
    Window w;
    for (int c = 0; c < 3; ++c) {
        w.min[c] = 1e10;
        w.max[c] = -1e10;
    }
    const ellipsoid_t &e = t.m_ellipsoid;
    for (int i = 0; i < 2; ++i)
        for (int j = -1; j < 3; j += 2)
            for (int k = -1; k < 3; k += 2) {
                double v[3];
                for (int c = 0; c < 3; ++c) {
                    const double di = ((e.m_direction1.m_data[c] * t.m_lag_separation) * t.m_num_lags) * i;
                    const double dj = (e.m_direction2.m_data[c] * e.m_R2) * j;
                    const double dk = (e.m_direction3.m_data[c] * e.m_R3) * k;
                    v[c] = (di + dj) + dk;
                }
                for (int c = 0; c < 3; ++c) {
                    if (v[c] < w.min[c]) w.min[c] = v[c];
                    if (v[c] > w.max[c]) w.max[c] = v[c];
                }
            }
    return w;
}

struct Offset {
    int d[3];
    int lag_min, lag_max;  // lags the offset's distance falls in (update_lags)
};

// Offsets of the window in the tunnel, in the reference's loop order
// (variograms.cpp:386-398: i2, j2, k2 start at the window minimum converted
// to int and run while <= the window maximum), with their lag ranges
// (update_lags, :331-360; offsets with no lag are dropped).
std::vector<Offset> tunnel_offsets(const variogram_search_template_t &t, int lag_count) {
    const Window w = search_window(t);
    std::vector<Offset> out;
    for (int i2 = (int)w.min[0]; i2 <= w.max[0]; ++i2)
        for (int j2 = (int)w.min[1]; j2 <= w.max[1]; ++j2)
            for (int k2 = (int)w.min[2]; k2 <= w.max[2]; ++k2) {
                vector_t v = {{(double)i2, (double)j2, (double)k2}};
                if (!in_tunnel(t, v)) continue;
                double r = 0;
                for (int c = 0; c < 3; ++c) r += v.m_data[c] * v.m_data[c];
                const double dist = sqrt(r);
                int lag_min = (int)ceil((dist - (t.m_lag_width / 2) - t.m_first_lag_distance) / t.m_lag_separation);
                if (lag_min >= lag_count) continue;
                int lag_max = (int)floor((dist + (t.m_lag_width / 2) - t.m_first_lag_distance) / t.m_lag_separation);
                if (lag_max < 0) continue;
                if (lag_max >= lag_count) lag_max = lag_count - 1;
                if (lag_min < 0) lag_min = 0;
                if (lag_min > lag_max) continue;
                out.push_back(Offset{{i2, j2, k2}, lag_min, lag_max});
            }
    return out;
}

// ---- device: gridded variogram ----

constexpr int kOffsetBatch = 32;
constexpr int kCellsPerBlock = 256;

struct GridArgs {
    const float *data;
    const unsigned char *mask;
    int dshape[3], dstride[3], mshape[3], mstride[3];
    const int *cells;  // informed cells as (x, y, z) triples
    int n_cells;
    const int *offs;   // offsets (dx, dy, dz) triples
    int n_offs;
    int o_base;        // index of offs[0] among all offsets (pair sampling)
    int percent;
    double *psum;      // [n_chunks][n_offs]
    unsigned long long *pcnt;
};

// Deterministic 32-bit mix of (offset, cell) for the pair sampling.
__device__ inline unsigned int pair_hash(unsigned int o, unsigned int c) {
    unsigned int h = o * 0x9e3779b1u ^ (c + 0x7f4a7c15u) * 0x85ebca6bu;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

__global__ __launch_bounds__(kCellsPerBlock) void k_grid_pairs(GridArgs a) {
    __shared__ double red_s[kCellsPerBlock / 64];
    __shared__ unsigned long long red_c[kCellsPerBlock / 64];
    const int chunk = blockIdx.x;
    const int o0 = blockIdx.y * kOffsetBatch;
    const int i = chunk * kCellsPerBlock + threadIdx.x;
    const bool have = i < a.n_cells;
    int x = 0, y = 0, z = 0;
    double v1 = 0.0;
    if (have) {
        x = a.cells[3 * i];
        y = a.cells[3 * i + 1];
        z = a.cells[3 * i + 2];
        v1 = a.data[a.dstride[0] * x + a.dstride[1] * y + a.dstride[2] * z];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < kOffsetBatch; ++q) {
        const int o = o0 + q;
        if (o >= a.n_offs) break;
        double s = 0.0;
        unsigned long long c = 0;
        if (have) {
            const int xx = x + a.offs[3 * o], yy = y + a.offs[3 * o + 1], zz = z + a.offs[3 * o + 2];
            if (xx >= 0 && yy >= 0 && zz >= 0 && xx < a.mshape[0] && yy < a.mshape[1] && zz < a.mshape[2] &&
                a.mask[a.mstride[0] * xx + a.mstride[1] * yy + a.mstride[2] * zz] != 0 &&
                (a.percent >= 100 || (int)(pair_hash((unsigned)(a.o_base + o), (unsigned)i) % 100u) < a.percent)) {
                const double v2 = a.data[a.dstride[0] * xx + a.dstride[1] * yy + a.dstride[2] * zz];
                const double d = v1 - v2;
                s = d * d;
                c = 1;
            }
        }
        // fixed-order reduction: wavefront butterfly, then the four wavefronts in order
        for (int m = 32; m >= 1; m >>= 1) {
            s += __shfl_xor(s, m);
            c += __shfl_xor(c, m);
        }
        if (lane == 0) {
            red_s[wave] = s;
            red_c[wave] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double ts = 0.0;
            unsigned long long tc = 0;
            for (int w = 0; w < kCellsPerBlock / 64; ++w) {
                ts += red_s[w];
                tc += red_c[w];
            }
            a.psum[(size_t)chunk * a.n_offs + o] = ts;
            a.pcnt[(size_t)chunk * a.n_offs + o] = tc;
        }
        __syncthreads();
    }
}

// The dense device grid from the informed cells: mask 1 and the value at
// each listed cell (the mask was cleared before).
__global__ void k_scatter_cells(const int *cells, const float *vals, int n, int X, int Y, float *data,
                                unsigned char *mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long c = cells[3 * i] + (long long)X * (cells[3 * i + 1] + (long long)Y * cells[3 * i + 2]);
    data[c] = vals[i];
    mask[c] = 1;
}

// Partials of each offset summed over chunks in chunk order.
__global__ void k_sum_chunks(const double *psum, const unsigned long long *pcnt, int n_chunks, int n_offs,
                             double *osum, unsigned long long *ocnt) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_offs) return;
    double s = 0.0;
    unsigned long long c = 0;
    for (int k = 0; k < n_chunks; ++k) {
        s += psum[(size_t)k * n_offs + o];
        c += pcnt[(size_t)k * n_offs + o];
    }
    osum[o] = s;
    ocnt[o] = c;
}

// ---- device: point-set variogram ----

struct PointArgs {
    const float *xs, *ys, *zs, *vs;
    int n;
    double dir[3][3];
    double R2, R3, tol;
    const double *lag_start, *lag_end;
    int lags;
    double *rsum;  // [n][lags]
    unsigned long long *rcnt;
};

__global__ void k_point_pairs(PointArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    double *rs = a.rsum + (size_t)p * a.lags;
    unsigned long long *rc = a.rcnt + (size_t)p * a.lags;
    for (int l = 0; l < a.lags; ++l) {
        rs[l] = 0.0;
        rc[l] = 0;
    }
    const float x1 = a.xs[p], y1 = a.ys[p], z1 = a.zs[p], v1 = a.vs[p];
    for (int q = 0; q < a.n; ++q) {
        // the differences are float (variograms.cpp:473-475)
        const double v[3] = {(double)(x1 - a.xs[q]), (double)(y1 - a.ys[q]), (double)(z1 - a.zs[q])};
        double ss[3];
        for (int d = 0; d < 3; ++d) {
            double r = 0.0;
            for (int c = 0; c < 3; ++c) r += v[c] * a.dir[d][c];
            ss[d] = r;
        }
        const double s2 = ss[1] / a.R2, s3 = ss[2] / a.R3;
        const double td = sqrt(s2 * s2 + s3 * s3);
        if (!(td <= 1.0 && a.tol * td <= ss[0])) continue;
        double r = 0.0;
        for (int c = 0; c < 3; ++c) r += v[c] * v[c];
        const double dist = sqrt(r);
        const double dv = (double)(v1 - a.vs[q]);
        const double var = dv * dv;
        for (int l = 0; l < a.lags; ++l)
            if (a.lag_start[l] <= dist && dist < a.lag_end[l]) {
                rs[l] += var;
                rc[l] += 1;
            }
    }
}

}  // namespace

#define GUARD_BEGIN \
    t_error.clear(); \
    try {
#define GUARD_END                         \
    }                                     \
    catch (const std::exception &e) {     \
        t_error = e.what();               \
        if (t_error.empty()) t_error = "error"; \
    }

CVAR_API void fill_ellipsoid_directions(ellipsoid_t *ell, double azimuth, double dip, double rotation) {
    GUARD_BEGIN
    // ellipsoid.cpp:39-71: M = (A B) C with degrees scaled by 3.14159265358 / 180
    const double coef = 3.14159265358 / 180.0;
    const double ra = azimuth * coef, rd = dip * coef, rr = rotation * coef;
    const double A[9] = {cos(ra), -sin(ra), 0, sin(ra), cos(ra), 0, 0, 0, 1};
    const double B[9] = {cos(rd), 0, -sin(rd), 0, 1, 0, sin(rd), 0, cos(rd)};
    const double C[9] = {1, 0, 0, 0, cos(rr), -sin(rr), 0, sin(rr), cos(rr)};
    auto mul = [](const double *m1, const double *m2, double *m3) {
        double r[9];
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) {
                double s = 0.0;
                for (int i = 0; i < 3; ++i) s += m1[row * 3 + i] * m2[i * 3 + col];
                r[row * 3 + col] = s;
            }
        std::memcpy(m3, r, sizeof(r));
    };
    double M[9] = {0};
    mul(A, B, M);
    mul(M, C, M);
    for (int i = 0; i < 3; ++i) {
        ell->m_direction1.m_data[i] = M[i * 3 + 0];
        ell->m_direction2.m_data[i] = M[i * 3 + 1];
        ell->m_direction3.m_data[i] = M[i * 3 + 2];
    }
    GUARD_END
}

// Informed cells of `data` with z in [z0, z1) as (x, y, z) triples in the
// reference's z, y, x order (variograms.cpp:403-413) and their values.
// Rows with unit mask stride skip 32 empty mask bytes at a time.
static void informed_cells_z(const hard_data_t &d, int z0, int z1, std::vector<int> &cells,
                             std::vector<float> &vals) {
    const int *ds = d.m_data_shape;
    const long long ms0 = d.m_mask_strides[0], ms1 = d.m_mask_strides[1], ms2 = d.m_mask_strides[2];
    const long long ds0 = d.m_data_strides[0], ds1 = d.m_data_strides[1], ds2 = d.m_data_strides[2];
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < ds[1]; ++y) {
            const unsigned char *row = d.m_mask + ms1 * y + ms2 * z;
            const float *vrow = d.m_data + ds1 * y + ds2 * z;
            int x = 0;
            while (x < ds[0]) {
                if (ms0 == 1 && x + 32 <= ds[0]) {
                    uint64_t w[4];
                    std::memcpy(w, row + x, 32);
                    if ((w[0] | w[1] | w[2] | w[3]) == 0) {
                        x += 32;
                        continue;
                    }
                }
                const int xe = ms0 == 1 ? std::min(x + 32, ds[0]) : x + 1;
                for (; x < xe; ++x)
                    if (row[ms0 * x] != 0) {
                        cells.push_back(x);
                        cells.push_back(y);
                        cells.push_back(z);
                        vals.push_back(vrow[ds0 * x]);
                    }
            }
        }
}

// All informed cells, the z layers split among host threads for large
// grids; the parts are joined in z order, so the order is that of a single
// scan.
static void informed_cells(const hard_data_t &d, std::vector<int> &cells, std::vector<float> &vals) {
    const int *ds = d.m_data_shape;
    const long long n = (long long)ds[0] * ds[1] * ds[2];
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::min<long long>({(long long)std::min(hw, 16u), (long long)ds[2], n / (1ll << 20) + 1});
    if (T <= 1) {
        informed_cells_z(d, 0, ds[2], cells, vals);
        return;
    }
    std::vector<std::vector<int>> pc(T);
    std::vector<std::vector<float>> pv(T);
    std::vector<std::thread> th;
    th.reserve(T - 1);
    auto range = [&](int t) { return std::make_pair((int)((long long)ds[2] * t / T), (int)((long long)ds[2] * (t + 1) / T)); };
    for (int t = 1; t < T; ++t) {
        try {
            th.emplace_back([&, t] { informed_cells_z(d, range(t).first, range(t).second, pc[t], pv[t]); });
        } catch (...) {
            break;  // no more threads: the remaining parts are scanned here
        }
    }
    for (int t = 0; t < T; ++t)
        if (t == 0 || t > (int)th.size()) informed_cells_z(d, range(t).first, range(t).second, pc[t], pv[t]);
    for (auto &x : th) x.join();
    size_t nc = 0;
    for (int t = 0; t < T; ++t) nc += pv[t].size();
    cells.reserve(3 * nc);
    vals.reserve(nc);
    for (int t = 0; t < T; ++t) {
        cells.insert(cells.end(), pc[t].begin(), pc[t].end());
        vals.insert(vals.end(), pv[t].begin(), pv[t].end());
    }
}

CVAR_API void calc_variograms(variogram_search_template_t *templ, hard_data_t *data, float *result, int result_length,
                              int percentToUse) {
    GUARD_BEGIN
    if (!templ || !data || !result) throw std::runtime_error("calc_variograms: null argument");
    const int lag_count = templ->m_num_lags <= result_length ? templ->m_num_lags : result_length;
    if (lag_count <= 0) return;
    const std::vector<Offset> offs = tunnel_offsets(*templ, lag_count);
    // informed cells, in the reference's z, y, x order (variograms.cpp:403-413)
    std::vector<int> cells;
    const int *ds = data->m_data_shape;
    const long long n_all = (long long)ds[0] * ds[1] * ds[2];
    long long extent_d = 0, extent_m = 0;
    for (int c = 0; c < 3; ++c) {
        // the reference reads partners inside the mask's shape from the data
        // array: with different shapes it would read outside one of them
        if (ds[c] != data->m_mask_shape[c]) throw std::runtime_error("calc_variograms: data and mask shapes differ");
        if (ds[c] < 0) throw std::runtime_error("calc_variograms: negative shape");
        if (ds[c] > 0) {
            extent_d += (long long)(ds[c] - 1) * data->m_data_strides[c];
            extent_m += (long long)(ds[c] - 1) * data->m_mask_strides[c];
        }
    }
    if (n_all == 0) {
        for (int l = 0; l < lag_count; ++l) result[l] = NAN;
        return;
    }
    // the device grid is indexed with 32-bit cell offsets (k_grid_pairs)
    if (n_all >= (1ll << 31)) throw std::runtime_error("calc_variograms: grid has more than 2^31-1 cells");
    std::vector<float> cell_vals;
    informed_cells(*data, cells, cell_vals);
    const int n_cells = (int)(cells.size() / 3);
    const int n_offs = (int)offs.size();
    std::vector<double> lag_sum(lag_count, 0.0);
    std::vector<long long> lag_cnt(lag_count, 0);
    t_offsets = n_offs;
    t_pairs = (long long)n_offs * n_cells;
    t_device_ms = 0;
    if (n_cells > 0 && n_offs > 0) {
        // the device holds a dense copy of the grid rebuilt from the informed
        // cells (mask 1 and the value there, mask 0 elsewhere; values of
        // uninformed cells are never read), so only the cell list crosses
        // PCIe, not the caller's whole data and mask arrays
        (void)extent_d;
        (void)extent_m;
        Workspace &ws = workspace();
        const int n_chunks = (n_cells + kCellsPerBlock - 1) / kCellsPerBlock;
        float *const d_data = ws.data.get<float>((size_t)n_all);
        float *const d_vals = ws.vals.get<float>((size_t)n_cells);
        unsigned char *const d_mask = ws.mask.get<unsigned char>((size_t)n_all);
        int *const d_cells = ws.cells.get<int>(cells.size());
        int *const d_offs = ws.offs.get<int>(3 * (size_t)n_offs);
        double *const psum = ws.psum.get<double>((size_t)n_chunks * n_offs);
        double *const osum = ws.osum.get<double>((size_t)n_offs);
        unsigned long long *const pcnt = ws.pcnt.get<unsigned long long>((size_t)n_chunks * n_offs);
        unsigned long long *const ocnt = ws.ocnt.get<unsigned long long>((size_t)n_offs);
        const hipStream_t s = ws.s;
        const hipEvent_t e0 = ws.e0, e1 = ws.e1;
        std::vector<int> h_offs(3 * (size_t)n_offs);
        for (int o = 0; o < n_offs; ++o)
            for (int c = 0; c < 3; ++c) h_offs[3 * o + c] = offs[o].d[c];
        try {
            hip_ok(hipMemcpyAsync(d_cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice, s),
                   "copy");
            hip_ok(hipMemcpyAsync(d_vals, cell_vals.data(), cell_vals.size() * sizeof(float), hipMemcpyHostToDevice,
                                  s),
                   "copy");
            hip_ok(hipMemsetAsync(d_mask, 0, (size_t)n_all, s), "memset");
            hipLaunchKernelGGL(k_scatter_cells, dim3((n_cells + 255) / 256), dim3(256), 0, s, d_cells, d_vals,
                               n_cells, ds[0], ds[1], d_data, d_mask);
            hip_ok(hipGetLastError(), "k_scatter_cells");
            hip_ok(hipMemcpyAsync(d_offs, h_offs.data(), h_offs.size() * sizeof(int), hipMemcpyHostToDevice, s),
                   "copy");
            GridArgs a;
            a.data = d_data;
            a.mask = d_mask;
            const int dense[3] = {1, ds[0], ds[0] * ds[1]};
            for (int c = 0; c < 3; ++c) {
                a.dshape[c] = ds[c];
                a.dstride[c] = dense[c];
                a.mshape[c] = ds[c];
                a.mstride[c] = dense[c];
            }
            a.cells = d_cells;
            a.n_cells = n_cells;
            a.offs = d_offs;
            a.n_offs = n_offs;
            a.percent = percentToUse;
            a.o_base = 0;
            a.psum = psum;
            a.pcnt = pcnt;
            hip_ok(hipEventRecord(e0, s), "event");
            const int batches = (n_offs + kOffsetBatch - 1) / kOffsetBatch;
            for (int b0 = 0; b0 < batches; b0 += 65535) {
                // grid y is limited to 65535: shift the offset arrays per launch
                GridArgs ab = a;
                const int nb = std::min(65535, batches - b0);
                ab.offs = a.offs + 3 * (size_t)b0 * kOffsetBatch;
                ab.n_offs = std::min(n_offs - b0 * kOffsetBatch, nb * kOffsetBatch);
                ab.o_base = b0 * kOffsetBatch;
                ab.psum = psum;  // chunk-major with the launch's own offset count
                ab.pcnt = pcnt;
                hipLaunchKernelGGL(k_grid_pairs, dim3(n_chunks, nb), dim3(kCellsPerBlock), 0, s, ab);
                hip_ok(hipGetLastError(), "k_grid_pairs");
                hipLaunchKernelGGL(k_sum_chunks, dim3((ab.n_offs + 255) / 256), dim3(256), 0, s, psum, pcnt,
                                   n_chunks, ab.n_offs, osum + (size_t)b0 * kOffsetBatch,
                                   ocnt + (size_t)b0 * kOffsetBatch);
                hip_ok(hipGetLastError(), "k_sum_chunks");
            }
            hip_ok(hipEventRecord(e1, s), "event");
            std::vector<double> h_sum(n_offs);
            std::vector<unsigned long long> h_cnt(n_offs);
            hip_ok(hipMemcpyAsync(h_sum.data(), osum, n_offs * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
            hip_ok(hipMemcpyAsync(h_cnt.data(), ocnt, n_offs * sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
                   "copy");
            hip_ok(hipStreamSynchronize(s), "sync");
            float ms = 0;
            hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            t_device_ms = ms;
            // each offset's pairs go to every lag of its range, in offset order
            for (int o = 0; o < n_offs; ++o)
                for (int l = offs[o].lag_min; l <= offs[o].lag_max; ++l) {
                    lag_sum[l] += h_sum[o];
                    lag_cnt[l] += (long long)h_cnt[o];
                }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            throw;
        }
    }
    // variograms.cpp:439-442 (the count is an int there; 0 pairs give NaN)
    for (int l = 0; l < lag_count; ++l) result[l] = (float)(lag_sum[l] / (double)lag_cnt[l] / 2);
    GUARD_END
}

CVAR_API void calc_variograms_from_point_set(variogram_search_template_t *templ, cont_point_set_t *ps, float *result,
                                             int result_length) {
    GUARD_BEGIN
    if (!templ || !ps || !result) throw std::runtime_error("calc_variograms_from_point_set: null argument");
    const int lag_count = templ->m_num_lags <= result_length ? templ->m_num_lags : result_length;
    if (lag_count <= 0) return;
    // init_lag_list (variograms.cpp:138-149)
    std::vector<double> lag_start(lag_count), lag_end(lag_count);
    for (int i = 0; i < lag_count; ++i) {
        const double d = i * templ->m_lag_separation + templ->m_first_lag_distance;
        lag_start[i] = d - templ->m_lag_width / 2;
        lag_end[i] = d + templ->m_lag_width / 2;
    }
    std::vector<double> lag_sum(lag_count, 0.0);
    std::vector<long long> lag_cnt(lag_count, 0);
    const int n = ps->size;
    if (n > 0) {
        DevBuf<float> xs(n), ys(n), zs(n), vs(n);
        DevBuf<double> d_start(lag_count), d_end(lag_count), rsum((size_t)n * lag_count);
        DevBuf<unsigned long long> rcnt((size_t)n * lag_count);
        hipStream_t s;
        hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        try {
            hip_ok(hipMemcpyAsync(xs.p, ps->xs, n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
            hip_ok(hipMemcpyAsync(ys.p, ps->ys, n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
            hip_ok(hipMemcpyAsync(zs.p, ps->zs, n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
            hip_ok(hipMemcpyAsync(vs.p, ps->values, n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
            hip_ok(hipMemcpyAsync(d_start.p, lag_start.data(), lag_count * sizeof(double), hipMemcpyHostToDevice, s),
                   "copy");
            hip_ok(hipMemcpyAsync(d_end.p, lag_end.data(), lag_count * sizeof(double), hipMemcpyHostToDevice, s),
                   "copy");
            PointArgs a;
            a.xs = xs.p;
            a.ys = ys.p;
            a.zs = zs.p;
            a.vs = vs.p;
            a.n = n;
            const ellipsoid_t &e = templ->m_ellipsoid;
            for (int c = 0; c < 3; ++c) {
                a.dir[0][c] = e.m_direction1.m_data[c];
                a.dir[1][c] = e.m_direction2.m_data[c];
                a.dir[2][c] = e.m_direction3.m_data[c];
            }
            a.R2 = e.m_R2;
            a.R3 = e.m_R3;
            a.tol = templ->m_tol_distance;
            a.lag_start = d_start.p;
            a.lag_end = d_end.p;
            a.lags = lag_count;
            a.rsum = rsum.p;
            a.rcnt = rcnt.p;
            hipLaunchKernelGGL(k_point_pairs, dim3((n + 63) / 64), dim3(64), 0, s, a);
            hip_ok(hipGetLastError(), "k_point_pairs");
            std::vector<double> h_sum((size_t)n * lag_count);
            std::vector<unsigned long long> h_cnt((size_t)n * lag_count);
            hip_ok(hipMemcpyAsync(h_sum.data(), rsum.p, h_sum.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                   "copy");
            hip_ok(hipMemcpyAsync(h_cnt.data(), rcnt.p, h_cnt.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s),
                   "copy");
            hip_ok(hipStreamSynchronize(s), "sync");
            for (int p = 0; p < n; ++p)
                for (int l = 0; l < lag_count; ++l) {
                    lag_sum[l] += h_sum[(size_t)p * lag_count + l];
                    lag_cnt[l] += (long long)h_cnt[(size_t)p * lag_count + l];
                }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    }
    for (int l = 0; l < lag_count; ++l) {
        result[l] = (float)(lag_sum[l] / (double)lag_cnt[l] / 2);
        if (lag_cnt[l] == 0) result[l] = 0;
    }
    GUARD_END
}

CVAR_API void cvar_stack_layers(float_data_t *thick_layers, int *layer_markers, int layers_count, int nz,
                                float scalez, int blank_value, float_data_t *result) {
    GUARD_BEGIN
    if (!thick_layers || !layer_markers || !result || layers_count <= 0)
        throw std::runtime_error("cvar_stack_layers: no layers");
    // stack_layers.h:6-63: per column, the running depth in layer units
    // (double) fills the cells each positive layer covers with its marker;
    // a negative layer blanks from its (shifted) top down to nz.  Where the
    // running depth has gone below 0 the reference writes below the cube
    // (k < 0); those cells are skipped here.
    const int nx = thick_layers[0].m_data_shape[0];
    const int ny = thick_layers[0].m_data_shape[1];
    std::vector<double> cum((size_t)nx * ny, 0.0);
    for (int layer = 0; layer < layers_count; ++layer) {
        const float_data_t &L = thick_layers[layer];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                const int map_index = L.m_data_strides[0] * i + L.m_data_strides[1] * j;
                if (L.m_data[map_index] > 0) {
                    const float old_k = (float)cum[map_index];
                    cum[map_index] += L.m_data[map_index] / scalez;
                    const float new_k = (float)cum[map_index];
                    int k_next = (int)floor(new_k) + 1;
                    if (k_next > nz) k_next = nz;
                    for (int k = std::max((int)ceil(old_k), 0); k < k_next; ++k)
                        result->m_data[result->m_data_strides[0] * i + result->m_data_strides[1] * j +
                                       result->m_data_strides[2] * k] = (float)layer_markers[layer];
                } else {
                    const int k_start = (int)ceil(cum[map_index] + (L.m_data[map_index] / scalez));
                    for (int k = std::max(k_start, 0); k < nz; ++k)
                        result->m_data[result->m_data_strides[0] * i + result->m_data_strides[1] * j +
                                       result->m_data_strides[2] * k] = (float)blank_value;
                    cum[map_index] += (L.m_data[map_index] / scalez);
                }
            }
    }
    GUARD_END
}

CVAR_API const char *cvar_amd_last_error(void) { return t_error.c_str(); }

CVAR_API void cvar_amd_last_stats(long long *offsets, long long *pairs_tested, double *device_ms) {
    if (offsets) *offsets = t_offsets;
    if (pairs_tested) *pairs_tested = t_pairs;
    if (device_ms) *device_ms = t_device_ms;
}
