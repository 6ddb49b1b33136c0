// api.hip - the extern "C" boundary (include/hpgl.h).
//
// Host-pointer entry points copy the caller's arrays to HBM, run the job on
// the calling thread's HIP stream and copy the outputs back; the *_dev
// variants take device pointers and a stream.  Every exception is caught
// here and reported through hpgl_get_last_exception_message().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hpgl.h"
#include "engine.h"
#include "kernels.h"
#include "propio.h"

using namespace hpgl_amd;

namespace {

thread_local std::string t_error;
thread_local int t_failed = 0;
thread_local Stats t_stats;
thread_local Dump t_dump;
thread_local hipStream_t t_stream = nullptr;
int g_precision = HPGL_AMD_PRECISION_MIXED;
int g_threads = 0;
int (*g_output_handler)(char *, void *) = nullptr;
void *g_output_param = nullptr;
int (*g_progress_handler)(char *, int, void *) = nullptr;
void *g_progress_param = nullptr;

hipStream_t thread_stream() {
    if (!t_stream) hip_check(hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking), "hipStreamCreate");
    return t_stream;
}

void begin_call() {
    t_error.clear();
    t_failed = 0;
    t_stats = Stats();
}

void fail(const char *msg) {
    t_error = msg;
    t_failed = 1;
}

Dump *take_dump(Dump &local) {
    if (!t_dump.indices) return nullptr;
    local = t_dump;
    t_dump = Dump();
    return &local;
}

void progress(const char *stage, int pct) {
    if (g_progress_handler) g_progress_handler(const_cast<char *>(stage), pct, g_progress_param);
}

CovSpec cov_of(int type, const double *ranges, const double *angles, double sill, double nugget) {
    CovSpec c;
    c.type = type;
    for (int i = 0; i < 3; ++i) {
        c.ranges[i] = ranges[i];
        c.angles[i] = angles[i];
    }
    c.sill = sill;
    c.nugget = nugget;
    return c;
}

long long volume(const hpgl_shape_t *s) { return (long long)s->m_data[0] * s->m_data[1] * s->m_data[2]; }

template <typename T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray(size_t count) : n(count) { p = (T *)pool_alloc(std::max<size_t>(count, 1) * sizeof(T)); }
    ~DevArray() { pool_free(p); }
    void put(const T *h, hipStream_t s) {
        if (n) hip_check(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
    }
    void get(T *h, hipStream_t s) {
        if (n) hip_check(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, s), "D2H");
    }
};

// One SK system on real coordinates (simple_kriging.cpp:65-85), analytic
// covariance, dpotf2/dpotrs order; run by one lane: the call is a utility
// on a handful of points, not a grid job.
__global__ void k_sk_weights_real(const float *center, const float *xs, const float *ys, const float *zs, int n,
                                  DevCov c, double *A, double *w, int *info_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    auto C = [&](double x1, double y1, double z1, double x2, double y2, double z2) {
        const double v0 = x1 - x2, v1 = y1 - y2, v2 = z1 - z2;
        double r = 0.0;
        for (int i = 0; i < 3; ++i) {
            double d = 0.0;
            d += v0 * c.t[3 * i + 0];
            d += v1 * c.t[3 * i + 1];
            d += v2 * c.t[3 * i + 2];
            r += d * d;
        }
        const double h = sqrt(r);
        if (h < 0.0001) return c.c0;
        const double a = c.range0;
        if (c.type == 0) {
            if (h > a) return 0.0;
            return c.c0 * (1 - 1.5 * h / a + 0.5 * pow(h / a, 3.0));
        }
        if (c.type == 1) return c.c0 * exp(-3 * h / a);
        return c.c0 * exp(-3 * pow(h / a, 2.0));
    };
    for (int i = 0; i < n; ++i) {
        for (int j = i; j < n; ++j) {
            A[(size_t)i * n + j] = C(xs[i], ys[i], zs[i], xs[j], ys[j], zs[j]);
            A[(size_t)j * n + i] = A[(size_t)i * n + j];
        }
        w[i] = C(xs[i], ys[i], zs[i], center[0], center[1], center[2]);
    }
    int info = 0;
    for (int j = 0; j < n && info == 0; ++j) {  // dpotf2 'U', column-major
        double dot = 0;
        for (int k = 0; k < j; ++k) dot = dot + A[k + j * n] * A[k + j * n];
        double ajj = A[j + j * n] - dot;
        if (ajj <= 0 || ajj != ajj) {
            info = j + 1;
            break;
        }
        ajj = sqrt(ajj);
        A[j + j * n] = ajj;
        const double r = 1.0 / ajj;
        for (int col = j + 1; col < n; ++col) {
            double temp = 0;
            for (int k = 0; k < j; ++k) temp = temp + A[k + col * n] * A[k + j * n];
            A[j + col * n] = r * (A[j + col * n] + (-1.0) * temp);
        }
    }
    if (info == 0) {  // dpotrs
        for (int i = 0; i < n; ++i) {
            double temp = w[i];
            for (int k = 0; k < i; ++k) temp = temp - A[k + i * n] * w[k];
            w[i] = temp / A[i + i * n];
        }
        for (int k = n - 1; k >= 0; --k)
            if (w[k] != 0) {
                w[k] = w[k] / A[k + k * n];
                for (int i = 0; i < k; ++i) w[i] = w[i] - w[k] * A[i + k * n];
            }
    }
    *info_out = info;
}

}  // namespace

#define API extern "C" __attribute__((visibility("default")))
#define GUARD_BEGIN \
    begin_call();   \
    try {
#define GUARD_END                          \
    }                                      \
    catch (const std::exception &e) {      \
        fail(e.what());                    \
    }                                      \
    catch (...) {                          \
        fail("unknown error in hpgl.so");  \
    }

API char *hpgl_get_last_exception_message(void) { return const_cast<char *>(t_error.c_str()); }

API void hpgl_set_output_handler(int (*handler)(char *, void *), void *param) {
    g_output_handler = handler;
    g_output_param = param;
}

API void hpgl_set_progress_handler(int (*handler)(char *, int, void *), void *param) {
    g_progress_handler = handler;
    g_progress_param = param;
}

API void hpgl_set_thread_num(int n) { g_threads = n; }
API int hpgl_get_thread_num(void) { return g_threads; }

// Property files (read_inc_file.cpp, property_writer.cpp, api.cpp:54-193):
// host code in propio.cpp.  Sizes are the shape volumes, as in the reference.
namespace {
long long volume(const hpgl_shape_t &s) { return (long long)s.m_data[0] * s.m_data[1] * s.m_data[2]; }
template <typename F>
void io_call(F f) {
    begin_call();
    try {
        f();
    } catch (const std::exception &e) {
        fail(e.what());
    }
}
}  // namespace

API void hpgl_read_inc_file_float(char *filename, float undefined_value, int size, float *data, unsigned char *mask) {
    io_call([&] { read_inc_float(filename, undefined_value, size, data, mask); });
}

API void hpgl_read_inc_file_byte(char *filename, int undefined_value, int size, unsigned char *data,
                                 unsigned char *mask, unsigned char *values, int values_count) {
    io_call([&] {
        read_inc_byte(filename, undefined_value, size, data, mask);
        // informed bytes equal to values[j] become j (api.cpp:85-98)
        for (int i = 0; i < size; ++i)
            if (mask[i] != 0)
                for (int j = 0; j < values_count; ++j)
                    if (data[i] == values[j]) {
                        data[i] = (unsigned char)j;
                        break;
                    }
    });
}

API void hpgl_write_inc_file_float(char *filename, hpgl_cont_masked_array_t *arr, float undefined_value, char *name) {
    io_call([&] {
        if (!arr) throw std::runtime_error("hpgl_write_inc_file_float: null array");
        write_inc_float(filename, name, arr->m_data, arr->m_mask, volume(arr->m_shape), (double)undefined_value);
    });
}

API void hpgl_write_inc_file_byte(char *filename, hpgl_ind_masked_array_t *arr, int undefined_value, char *name,
                                  unsigned char *values, int values_count) {
    io_call([&] {
        if (!arr) throw std::runtime_error("hpgl_write_inc_file_byte: null array");
        write_inc_byte(filename, name, arr->m_data, arr->m_mask, volume(arr->m_shape), (uint8_t)undefined_value,
                       remap_table(values, values_count, arr->m_indicator_count));
    });
}

API void hpgl_write_gslib_cont_property(hpgl_cont_masked_array_t *data, const char *filename, const char *name,
                                        double undefined_value) {
    io_call([&] {
        if (!data) throw std::runtime_error("hpgl_write_gslib_cont_property: null array");
        write_gslib_float(filename, name, data->m_data, data->m_mask, volume(data->m_shape), undefined_value);
    });
}

API void hpgl_write_gslib_byte_property(hpgl_ind_masked_array_t *data, const char *filename, const char *name,
                                        double undefined_value, unsigned char *values, int values_count) {
    io_call([&] {
        if (!data) throw std::runtime_error("hpgl_write_gslib_byte_property: null array");
        write_gslib_byte(filename, name, data->m_data, data->m_mask, volume(data->m_shape),
                         (uint8_t)(int)undefined_value,
                         remap_table(values, values_count, data->m_indicator_count));
    });
}

API void hpgl_amd_set_precision(int mode) {
    g_precision = mode == HPGL_AMD_PRECISION_FP64 ? HPGL_AMD_PRECISION_FP64 : HPGL_AMD_PRECISION_MIXED;
}
API int hpgl_amd_get_precision(void) { return g_precision; }
API int hpgl_amd_get_last_error(void) { return t_failed; }

API void hpgl_amd_get_last_stats(hpgl_amd_stats_t *o) {
    if (!o) return;
    o->nodes_target = t_stats.nodes_target;
    o->nodes_kriged = t_stats.nodes_kriged;
    o->nodes_no_neighbours = t_stats.nodes_no_neighbours;
    o->nodes_singular = t_stats.nodes_singular;
    o->solves_fp32 = t_stats.solves_fp32;
    o->solves_fp64 = t_stats.solves_fp64;
    o->nodes_template_path = t_stats.nodes_template_path;
    o->nodes_cluster_path = t_stats.nodes_cluster_path;
    o->ms_setup_host = t_stats.ms_setup_host;
    o->ms_search = t_stats.ms_search;
    o->ms_solve = t_stats.ms_solve;
    o->ms_device_total = t_stats.ms_device_total;
    o->ms_wall = t_stats.ms_wall;
    o->ms_k_template = t_stats.ms_k_template;
    o->ms_k_cluster = t_stats.ms_k_cluster;
    o->ms_k_solve_mixed = t_stats.ms_k_solve_mixed;
    o->ms_k_solve_fp64 = t_stats.ms_k_solve_fp64;
    o->n_k_template = t_stats.n_k_template;
    o->n_k_cluster = t_stats.n_k_cluster;
    o->n_k_solve_mixed = t_stats.n_k_solve_mixed;
    o->n_k_solve_fp64 = t_stats.n_k_solve_fp64;
    o->cluster_candidates = t_stats.cluster_candidates;
    o->neighbours_total = t_stats.neighbours_total;
    o->cluster_lists_spilled = t_stats.cluster_lists_spilled;
    o->solves_fp32_two_steps = t_stats.solves_fp32_two_steps;
}

API void hpgl_amd_set_neighbour_dump(int *indices, int *counts, long long n_cells, int max_n) {
    t_dump.indices = indices;
    t_dump.counts = counts;
    t_dump.n_cells = n_cells;
    t_dump.max_n = max_n;
}

API int hpgl_amd_pool_selftest(void) {
    int rc = -100;
    GUARD_BEGIN
    rc = pool_selftest();
    GUARD_END
    return rc;
}

API int hpgl_amd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ---- device-pointer variants ----

static void ok_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                   const hpgl_ok_params_t *p, float *out, unsigned char *out_mask, hipStream_t s) {
    ContJob j;
    std::memset(&j, 0, sizeof(j));
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    j.cov = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    j.max_n = p->m_max_neighbours;
    j.ordinary = true;
    j.mean_kind = 0;
    j.mean_on_failure = false;
    j.d_in = in;
    j.d_in_mask = in_mask;
    j.d_out = out;
    j.d_out_mask = out_mask;
    Dump d;
    progress("ordinary kriging", 0);
    run_cont(j, s, g_precision, take_dump(d), t_stats);
    progress("ordinary kriging", 100);
}

static void sk_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                   const hpgl_sk_params_t *p, float *out, unsigned char *out_mask, hipStream_t s) {
    ContJob j;
    std::memset(&j, 0, sizeof(j));
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    j.cov = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    j.max_n = p->m_max_neighbours;
    j.ordinary = false;
    j.mean_kind = 1;
    j.mean_auto = p->m_automatic_mean != 0;
    j.mean_single = p->m_mean;
    j.mean_on_failure = true;
    j.d_in = in;
    j.d_in_mask = in_mask;
    j.d_out = out;
    j.d_out_mask = out_mask;
    Dump d;
    progress("simple kriging", 0);
    run_cont(j, s, g_precision, take_dump(d), t_stats);
    progress("simple kriging", 100);
}

static void lvm_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape, const float *mean,
                    const hpgl_ok_params_t *p, float *out, unsigned char *out_mask, hipStream_t s) {
    ContJob j;
    std::memset(&j, 0, sizeof(j));
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    j.cov = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    j.max_n = p->m_max_neighbours;
    j.ordinary = false;
    j.mean_kind = 2;
    j.d_mean = mean;
    j.mean_on_failure = true;
    j.d_in = in;
    j.d_in_mask = in_mask;
    j.d_out = out;
    j.d_out_mask = out_mask;
    Dump d;
    progress("lvm kriging", 0);
    run_cont(j, s, g_precision, take_dump(d), t_stats);
    progress("lvm kriging", 100);
}

static void ik_dev(const unsigned char *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                   const hpgl_ik_params_t *params, int K, unsigned char *out, unsigned char *out_mask,
                   hipStream_t s) {
    IndJob j;
    for (int i = 0; i < 3; ++i) j.grid[i] = shape->m_data[i];
    for (int k = 0; k < K; ++k) {
        const hpgl_ik_params_t &p = params[k];
        IndicatorParams c;
        c.cov = cov_of(p.m_covariance_type, p.m_ranges, p.m_angles, p.m_sill, p.m_nugget);
        for (int i = 0; i < 3; ++i) c.radii[i] = p.m_radiuses[i];
        c.max_n = p.m_max_neighbours;
        c.marginal = (double)(float)p.m_marginal_prob;  // ik_params_t keeps it as float
        j.cats.push_back(c);
    }
    j.median = false;
    j.d_in = in;
    j.d_in_mask = in_mask;
    j.d_out = out;
    j.d_out_mask = out_mask;
    Dump d;
    progress("indicator kriging", 0);
    run_ind(j, s, g_precision, take_dump(d), t_stats);
    progress("indicator kriging", 100);
}

API void hpgl_amd_ordinary_kriging_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                                       const hpgl_ok_params_t *p, float *out, unsigned char *out_mask, void *stream) {
    GUARD_BEGIN
    ok_dev(in, in_mask, shape, p, out, out_mask, (hipStream_t)stream);
    GUARD_END
}

API void hpgl_amd_simple_kriging_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                                     const hpgl_sk_params_t *p, float *out, unsigned char *out_mask, void *stream) {
    GUARD_BEGIN
    sk_dev(in, in_mask, shape, p, out, out_mask, (hipStream_t)stream);
    GUARD_END
}

API void hpgl_amd_lvm_kriging_dev(const float *in, const unsigned char *in_mask, const hpgl_shape_t *shape,
                                  const float *mean, const hpgl_ok_params_t *p, float *out, unsigned char *out_mask,
                                  void *stream) {
    GUARD_BEGIN
    lvm_dev(in, in_mask, shape, mean, p, out, out_mask, (hipStream_t)stream);
    GUARD_END
}

API void hpgl_amd_indicator_kriging_dev(const unsigned char *in, const unsigned char *in_mask,
                                        const hpgl_shape_t *shape, const hpgl_ik_params_t *params, int K,
                                        unsigned char *out, unsigned char *out_mask, void *stream) {
    GUARD_BEGIN
    ik_dev(in, in_mask, shape, params, K, out, out_mask, (hipStream_t)stream);
    GUARD_END
}

// ---- host-pointer drop-in entry points ----

API void hpgl_ordinary_kriging(hpgl_cont_masked_array_t *in, hpgl_ok_params_t *p, hpgl_cont_masked_array_t *out) {
    GUARD_BEGIN
    const long long N = volume(&in->m_shape);
    if (volume(&out->m_shape) != N) throw std::runtime_error("input and output sizes differ");
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> di(N), dout(N);
    DevArray<unsigned char> dm(N), dom(N);
    di.put(in->m_data, s);
    dm.put(in->m_mask, s);
    dout.put(out->m_data, s);
    dom.put(out->m_mask, s);
    ok_dev(di.p, dm.p, &in->m_shape, p, dout.p, dom.p, s);
    dout.get(out->m_data, s);
    dom.get(out->m_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    GUARD_END
}

API void hpgl_simple_kriging(float *input_data, unsigned char *input_mask, hpgl_shape_t *in_shape,
                             hpgl_sk_params_t *p, float *output_data, unsigned char *output_mask,
                             hpgl_shape_t *out_shape) {
    GUARD_BEGIN
    const long long N = volume(in_shape);
    if (volume(out_shape) != N) throw std::runtime_error("input and output sizes differ");
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> di(N), dout(N);
    DevArray<unsigned char> dm(N), dom(N);
    di.put(input_data, s);
    dm.put(input_mask, s);
    dout.put(output_data, s);
    dom.put(output_mask, s);
    sk_dev(di.p, dm.p, in_shape, p, dout.p, dom.p, s);
    dout.get(output_data, s);
    dom.get(output_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    GUARD_END
}

API void hpgl_lvm_kriging(float *input_data, unsigned char *input_mask, hpgl_shape_t *in_shape, float *mean_data,
                          hpgl_shape_t *mean_shape, hpgl_ok_params_t *p, float *output_data,
                          unsigned char *output_mask, hpgl_shape_t *out_shape) {
    GUARD_BEGIN
    const long long N = volume(in_shape);
    if (volume(out_shape) != N || volume(mean_shape) != N) throw std::runtime_error("array sizes differ");
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> di(N), dmean(N), dout(N);
    DevArray<unsigned char> dm(N), dom(N);
    di.put(input_data, s);
    dm.put(input_mask, s);
    dmean.put(mean_data, s);
    dout.put(output_data, s);
    dom.put(output_mask, s);
    lvm_dev(di.p, dm.p, in_shape, dmean.p, p, dout.p, dom.p, s);
    dout.get(output_data, s);
    dom.get(output_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    GUARD_END
}

API void hpgl_indicator_kriging(hpgl_ind_masked_array_t *in, hpgl_ind_masked_array_t *out, hpgl_ik_params_t *params,
                                int K) {
    GUARD_BEGIN
    const long long N = volume(&in->m_shape);
    if (volume(&out->m_shape) != N) throw std::runtime_error("input and output sizes differ");
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<unsigned char> di(N), dm(N), dout(N), dom(N);
    di.put(in->m_data, s);
    dm.put(in->m_mask, s);
    dout.put(out->m_data, s);
    dom.put(out->m_mask, s);
    ik_dev(di.p, dm.p, &in->m_shape, params, K, dout.p, dom.p, s);
    dout.get(out->m_data, s);
    dom.get(out->m_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    GUARD_END
}

API void hpgl_median_ik(hpgl_ind_masked_array_t *in, hpgl_median_ik_params_t *p, hpgl_ind_masked_array_t *out) {
    GUARD_BEGIN
    const long long N = volume(&in->m_shape);
    if (volume(&out->m_shape) != N)
        throw std::runtime_error("median_ik_for_two_indicators: Input property size is not equal output property size");
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<unsigned char> di(N), dm(N), dout(N), dom(N);
    di.put(in->m_data, s);
    dm.put(in->m_mask, s);
    dout.put(out->m_data, s);
    dom.put(out->m_mask, s);
    IndJob j;
    for (int i = 0; i < 3; ++i) j.grid[i] = in->m_shape.m_data[i];
    IndicatorParams c;
    c.cov = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    for (int i = 0; i < 3; ++i) c.radii[i] = p->m_radiuses[i];
    c.max_n = p->m_max_neighbours;
    c.marginal = p->m_marginal_probs[1];
    j.cats.push_back(c);
    j.median = true;
    j.d_in = di.p;
    j.d_in_mask = dm.p;
    j.d_out = dout.p;
    j.d_out_mask = dom.p;
    Dump d;
    progress("median ik", 0);
    run_ind(j, s, g_precision, take_dump(d), t_stats);
    progress("median ik", 100);
    dout.get(out->m_data, s);
    dom.get(out->m_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    GUARD_END
}

API int hpgl_simple_kriging_weights(float *center, float *xs, float *ys, float *zs, int n, hpgl_cov_params_t *p,
                                    float *weights) {
    GUARD_BEGIN
    if (n <= 0) return 0;
    CovModel m(cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget));
    DevCov c;
    std::memset(&c, 0, sizeof(c));
    for (int i = 0; i < 9; ++i) c.t[i] = m.t[i];
    c.range0 = m.s.ranges[0];
    c.c0 = m.c0();
    c.type = m.s.type;
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> dc(3), dx(n), dy(n), dz(n);
    DevArray<double> A((size_t)n * n), w(n);
    DevArray<int> info(1);
    dc.put(center, s);
    dx.put(xs, s);
    dy.put(ys, s);
    dz.put(zs, s);
    hipLaunchKernelGGL(k_sk_weights_real, dim3(1), dim3(64), 0, s, dc.p, dx.p, dy.p, dz.p, n, c, A.p, w.p, info.p);
    hip_check(hipGetLastError(), "sk weights");
    std::vector<double> hw(n);
    int hinfo = 0;
    w.get(hw.data(), s);
    info.get(&hinfo, s);
    hip_check(hipStreamSynchronize(s), "sync");
    if (hinfo != 0) throw std::runtime_error("simple_kriging_weights: singular system");
    for (int i = 0; i < n; ++i) weights[i] = (float)hw[i];
    GUARD_END
    return 0;
}

// ---- simple_kriging_weights without allocation ----
//
// Device-pointer variant of hpgl_simple_kriging_weights for callers that
// cannot allocate (code under graph capture, serving loops): the system
// matrix, the double weights and the factorisation status live in the
// file-scope array `scratch` (about 2 MiB per device, held until the
// process ends).  Its device address is looked up once per device with
// hipGetSymbolAddress; host code never passes the array's name itself.
// Every call on a device shares that device's array, so calls that overlap
// on one GPU race.
namespace {
constexpr int kNoAllocMaxPoints = 512;
constexpr size_t kScratchMatrix = (size_t)kNoAllocMaxPoints * kNoAllocMaxPoints * sizeof(double);
constexpr size_t kScratchWeights = (size_t)kNoAllocMaxPoints * sizeof(double);
constexpr size_t kScratchBytes = kScratchMatrix + kScratchWeights + 256;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of its `scratch`

unsigned char *scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    return static_cast<unsigned char *>(p);
}

// out[i] = (float)w[i] when the factorisation succeeded; untouched otherwise.
__global__ void k_weights_out(const double *w, const int *info, int n, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *info == 0) out[i] = (float)w[i];
}
}  // namespace

API int hpgl_amd_simple_kriging_weights_noalloc(const float *center, const float *xs, const float *ys,
                                                const float *zs, int n, const hpgl_cov_params_t *p, float *weights,
                                                void *stream) {
    GUARD_BEGIN
    if (n <= 0) return 0;
    if (n > kNoAllocMaxPoints)
        throw std::runtime_error("simple_kriging_weights_noalloc: at most " + std::to_string(kNoAllocMaxPoints) +
                                 " points");
    CovModel m(cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget));
    DevCov c;
    std::memset(&c, 0, sizeof(c));
    for (int i = 0; i < 9; ++i) c.t[i] = m.t[i];
    c.range0 = m.s.ranges[0];
    c.c0 = m.c0();
    c.type = m.s.type;
    const hipStream_t s = (hipStream_t)stream;
    unsigned char *base = scratch_address();
    double *A = reinterpret_cast<double *>(base);
    double *w = reinterpret_cast<double *>(base + kScratchMatrix);
    int *info = reinterpret_cast<int *>(base + kScratchMatrix + kScratchWeights);
    hipLaunchKernelGGL(k_sk_weights_real, dim3(1), dim3(64), 0, s, center, xs, ys, zs, n, c, A, w, info);
    hip_check(hipGetLastError(), "sk weights");
    hipLaunchKernelGGL(k_weights_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, info, n, weights);
    hip_check(hipGetLastError(), "sk weights out");
    int hinfo = 0;
    hip_check(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s), "copy status");
    hip_check(hipStreamSynchronize(s), "sync");
    if (hinfo != 0) throw std::runtime_error("simple_kriging_weights: singular system");
    GUARD_END
    return 0;
}

// ---- sequential Gaussian simulation ----

static SgsJob sgs_job(const hpgl_shape_t *shape, const hpgl_sgs_params_t *p, const hpgl_non_parametric_cdf_t *cdf,
                      bool lvm) {
    SgsJob j;
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    j.cov = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    j.max_n = p->m_max_neighbours;
    // kriging kind 0 is OK (sgs_params.h:10); the LVM variant always uses SK
    // with the local mean (sequential_gaussian_simulation.cpp:77-128)
    j.ordinary = !lvm && p->m_kriging_kind == 0;
    if (cdf) {
        if (cdf->m_size <= 0 || !cdf->m_values || !cdf->m_probs)
            throw std::runtime_error("the non-parametric cdf is empty");
        j.has_cdf = true;
        j.cdf_values.assign(cdf->m_values, cdf->m_values + cdf->m_size);
        j.cdf_probs.assign(cdf->m_probs, cdf->m_probs + cdf->m_size);
    }
    return j;
}

// Runs the job with conditioning data from host arrays and writes R
// realisations to host arrays out[R][N], out_mask[R][N].
static void sgs_host(SgsJob &j, const float *data, const unsigned char *mask, const float *lvm_mean,
                     const unsigned char *sim_mask, float *out, unsigned char *out_mask) {
    const long long N = (long long)j.grid[0] * j.grid[1] * j.grid[2];
    const int R = (int)j.seeds.size();
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> di(N), dlvm(lvm_mean ? N : 0), dout((size_t)R * N);
    DevArray<unsigned char> dm(N), dsim(sim_mask ? N : 0), dom((size_t)R * N);
    di.put(data, s);
    dm.put(mask, s);
    if (lvm_mean) dlvm.put(lvm_mean, s);
    if (sim_mask) dsim.put(sim_mask, s);
    j.d_in = di.p;
    j.d_in_mask = dm.p;
    j.d_lvm = lvm_mean ? dlvm.p : nullptr;
    j.d_sim_mask = sim_mask ? dsim.p : nullptr;
    j.d_out = dout.p;
    j.d_out_mask = dom.p;
    Dump d;
    progress("sgs", 0);
    run_sgs(j, s, g_precision, take_dump(d), t_stats);
    dout.get(out, s);
    dom.get(out_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    progress("sgs", 100);
}

static void check_same_size(const hpgl_shape_t *a, long long n, const char *what) {
    if (volume(a) != n) throw std::runtime_error(std::string(what) + " has a different size from the property");
}

API void hpgl_sgs_simulation(hpgl_cont_masked_array_t *prop, hpgl_sgs_params_t *params, hpgl_non_parametric_cdf_t *cdf,
                             double *mean, hpgl_ubyte_array_t *simulation_mask) {
    GUARD_BEGIN
    (void)mean;  // only logged by the reference (sequential_gaussian_simulation.cpp:39 uses mean 0)
    const long long N = volume(&prop->m_shape);
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SgsJob j = sgs_job(&prop->m_shape, params, cdf, false);
    j.seeds = {params->m_seed};
    sgs_host(j, prop->m_data, prop->m_mask, nullptr, simulation_mask ? simulation_mask->m_data : nullptr,
             prop->m_data, prop->m_mask);
    GUARD_END
}

API void hpgl_sgs_lvm_simulation(hpgl_cont_masked_array_t *data, hpgl_sgs_params_t *params,
                                 hpgl_non_parametric_cdf_t *cdf, hpgl_float_array_t *means,
                                 hpgl_ubyte_array_t *simulation_mask) {
    GUARD_BEGIN
    const long long N = volume(&data->m_shape);
    if (!means || !means->m_data) throw std::runtime_error("the local mean is missing");
    check_same_size(&means->m_shape, N, "the local mean");
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SgsJob j = sgs_job(&data->m_shape, params, cdf, true);
    j.seeds = {params->m_seed};
    sgs_host(j, data->m_data, data->m_mask, means->m_data, simulation_mask ? simulation_mask->m_data : nullptr,
             data->m_data, data->m_mask);
    GUARD_END
}

API void hpgl_amd_sgs_simulation_batch(const hpgl_cont_masked_array_t *prop, const hpgl_sgs_params_t *params,
                                       const hpgl_non_parametric_cdf_t *cdf, const hpgl_float_array_t *lvm_means,
                                       const hpgl_ubyte_array_t *simulation_mask, const long *seeds, int R,
                                       float *out, unsigned char *out_mask) {
    GUARD_BEGIN
    const long long N = volume(&prop->m_shape);
    if (R < 1 || !seeds) throw std::runtime_error("at least one seed is needed");
    if (lvm_means) check_same_size(&lvm_means->m_shape, N, "the local mean");
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SgsJob j = sgs_job(&prop->m_shape, params, cdf, lvm_means != nullptr);
    j.seeds.assign(seeds, seeds + R);
    sgs_host(j, prop->m_data, prop->m_mask, lvm_means ? lvm_means->m_data : nullptr,
             simulation_mask ? simulation_mask->m_data : nullptr, out, out_mask);
    GUARD_END
}

API void hpgl_amd_sgs_simulation_batch_dev(const float *d_in, const unsigned char *d_in_mask,
                                           const hpgl_shape_t *shape, const hpgl_sgs_params_t *params,
                                           const hpgl_non_parametric_cdf_t *cdf, const float *d_lvm_mean,
                                           const unsigned char *d_sim_mask, const long *seeds, int R, float *d_out,
                                           unsigned char *d_out_mask, void *stream) {
    GUARD_BEGIN
    if (R < 1 || !seeds) throw std::runtime_error("at least one seed is needed");
    SgsJob j = sgs_job(shape, params, cdf, d_lvm_mean != nullptr);
    j.seeds.assign(seeds, seeds + R);
    j.d_in = d_in;
    j.d_in_mask = d_in_mask;
    j.d_lvm = d_lvm_mean;
    j.d_sim_mask = d_sim_mask;
    j.d_out = d_out;
    j.d_out_mask = d_out_mask;
    Dump d;
    run_sgs(j, (hipStream_t)stream, g_precision, take_dump(d), t_stats);
    GUARD_END
}

// ---- sequential indicator simulation ----

static SisJob sis_job(const hpgl_shape_t *shape, const hpgl_ik_params_t *params, int K) {
    if (!params) throw std::runtime_error("the indicator parameters are missing");
    if (K < 2) throw std::runtime_error("sequential indicator simulation needs at least two categories");
    SisJob j;
    for (int i = 0; i < 3; ++i) j.grid[i] = shape->m_data[i];
    // init_sis_params (api_helpers.cpp:70-112)
    for (int k = 0; k < K; ++k) {
        const hpgl_ik_params_t &p = params[k];
        IndicatorParams c;
        c.cov = cov_of(p.m_covariance_type, p.m_ranges, p.m_angles, p.m_sill, p.m_nugget);
        for (int i = 0; i < 3; ++i) c.radii[i] = p.m_radiuses[i];
        c.max_n = p.m_max_neighbours;
        c.marginal = p.m_marginal_prob;
        j.cats.push_back(c);
    }
    return j;
}

// Runs the job with conditioning data from host arrays and writes R
// realisations to host arrays out[R][N], out_mask[R][N].  lvm: K host
// arrays of N means, or nullptr.
static void sis_host(SisJob &j, const unsigned char *data, const unsigned char *mask, const float *const *lvm,
                     const unsigned char *sim_mask, unsigned char *out, unsigned char *out_mask) {
    const long long N = (long long)j.grid[0] * j.grid[1] * j.grid[2];
    const int R = (int)j.seeds.size();
    const int K = (int)j.cats.size();
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<unsigned char> di(N), dm(N), dsim(sim_mask ? N : 0), dout((size_t)R * N), dom((size_t)R * N);
    DevArray<float> dlvm(lvm ? (size_t)K * N : 0);
    di.put(data, s);
    dm.put(mask, s);
    if (sim_mask) dsim.put(sim_mask, s);
    j.d_lvm.clear();
    if (lvm)
        for (int k = 0; k < K; ++k) {
            hip_check(hipMemcpyAsync(dlvm.p + (size_t)k * N, lvm[k], N * sizeof(float), hipMemcpyHostToDevice, s),
                      "upload mean");
            j.d_lvm.push_back(dlvm.p + (size_t)k * N);
        }
    j.d_in = di.p;
    j.d_in_mask = dm.p;
    j.d_sim_mask = sim_mask ? dsim.p : nullptr;
    j.d_out = dout.p;
    j.d_out_mask = dom.p;
    progress("sis", 0);
    run_sis(j, s, g_precision, t_stats);
    dout.get(out, s);
    dom.get(out_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
    progress("sis", 100);
}

static std::vector<const float *> sis_means(const hpgl_float_array_t *mean_data, int K, long long N) {
    if (!mean_data) throw std::runtime_error("the local means are missing");
    std::vector<const float *> m(K);
    for (int k = 0; k < K; ++k) {
        if (!mean_data[k].m_data) throw std::runtime_error("a local mean is missing");
        check_same_size(&mean_data[k].m_shape, N, "a local mean");
        m[k] = mean_data[k].m_data;
    }
    return m;
}

API void hpgl_sis_simulation(hpgl_ind_masked_array_t *data, hpgl_ik_params_t *params, int indicator_count, int seed,
                             hpgl_ubyte_array_t *simulation_mask) {
    GUARD_BEGIN
    const long long N = volume(&data->m_shape);
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SisJob j = sis_job(&data->m_shape, params, indicator_count);
    j.seeds = {(long)seed};
    sis_host(j, data->m_data, data->m_mask, nullptr, simulation_mask ? simulation_mask->m_data : nullptr,
             data->m_data, data->m_mask);
    GUARD_END
}

API void hpgl_sis_simulation_lvm(hpgl_ind_masked_array_t *data, hpgl_ik_params_t *params,
                                 hpgl_float_array_t *mean_data, int indicator_count, int seed,
                                 hpgl_ubyte_array_t *simulation_mask, int use_correlograms) {
    GUARD_BEGIN
    const long long N = volume(&data->m_shape);
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SisJob j = sis_job(&data->m_shape, params, indicator_count);
    const std::vector<const float *> m = sis_means(mean_data, indicator_count, N);
    j.correlogram = use_correlograms != 0;
    j.seeds = {(long)seed};
    sis_host(j, data->m_data, data->m_mask, m.data(), simulation_mask ? simulation_mask->m_data : nullptr,
             data->m_data, data->m_mask);
    GUARD_END
}

API void hpgl_amd_sis_simulation_batch(const hpgl_ind_masked_array_t *data, const hpgl_ik_params_t *params,
                                       int indicator_count, const hpgl_float_array_t *mean_data,
                                       int use_correlograms, const hpgl_ubyte_array_t *simulation_mask,
                                       const int *seeds, int R, unsigned char *out, unsigned char *out_mask) {
    GUARD_BEGIN
    const long long N = volume(&data->m_shape);
    if (R < 1 || !seeds) throw std::runtime_error("at least one seed is needed");
    if (simulation_mask) check_same_size(&simulation_mask->m_shape, N, "the simulation mask");
    SisJob j = sis_job(&data->m_shape, params, indicator_count);
    std::vector<const float *> m;
    if (mean_data) {
        m = sis_means(mean_data, indicator_count, N);
        j.correlogram = use_correlograms != 0;
    }
    for (int r = 0; r < R; ++r) j.seeds.push_back((long)seeds[r]);
    sis_host(j, data->m_data, data->m_mask, mean_data ? m.data() : nullptr,
             simulation_mask ? simulation_mask->m_data : nullptr, out, out_mask);
    GUARD_END
}

API void hpgl_amd_sis_simulation_batch_dev(const unsigned char *d_in, const unsigned char *d_in_mask,
                                           const hpgl_shape_t *shape, const hpgl_ik_params_t *params,
                                           int indicator_count, const float *d_lvm, int use_correlograms,
                                           const unsigned char *d_sim_mask, const int *seeds, int R,
                                           unsigned char *d_out, unsigned char *d_out_mask, void *stream) {
    GUARD_BEGIN
    if (R < 1 || !seeds) throw std::runtime_error("at least one seed is needed");
    SisJob j = sis_job(shape, params, indicator_count);
    const long long N = volume(shape);
    if (d_lvm) {
        for (int k = 0; k < indicator_count; ++k) j.d_lvm.push_back(d_lvm + (size_t)k * N);
        j.correlogram = use_correlograms != 0;
    }
    for (int r = 0; r < R; ++r) j.seeds.push_back((long)seeds[r]);
    j.d_in = d_in;
    j.d_in_mask = d_in_mask;
    j.d_sim_mask = d_sim_mask;
    j.d_out = d_out;
    j.d_out_mask = d_out_mask;
    run_sis(j, (hipStream_t)stream, g_precision, t_stats);
    GUARD_END
}

// ---- simple colocated cokriging ----

static CokrigingJob cokriging_m1_job(const hpgl_shape_t *shape, const hpgl_cokriging_m1_params_t *p) {
    CokrigingJob j;
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    j.primary = cov_of(p->m_covariance_type, p->m_ranges, p->m_angles, p->m_sill, p->m_nugget);
    j.mark2 = false;
    j.secondary_variance = p->m_secondary_variance;
    j.correlation = p->m_correlation_coef;
    j.primary_mean = (float)p->m_primary_mean;
    j.secondary_mean = (float)p->m_secondary_mean;
    j.max_n = p->m_max_neighbours;
    return j;
}

static CokrigingJob cokriging_m2_job(const hpgl_shape_t *shape, const hpgl_cokriging_m2_params_t *p) {
    CokrigingJob j;
    for (int i = 0; i < 3; ++i) {
        j.grid[i] = shape->m_data[i];
        j.radii[i] = p->m_radiuses[i];
    }
    const hpgl_cov_params_t &a = p->m_primary_cov_params, &b = p->m_secondary_cov_params;
    j.primary = cov_of(a.m_covariance_type, a.m_ranges, a.m_angles, a.m_sill, a.m_nugget);
    j.secondary = cov_of(b.m_covariance_type, b.m_ranges, b.m_angles, b.m_sill, b.m_nugget);
    j.mark2 = true;
    j.correlation = p->m_correlation_coef;
    j.primary_mean = (float)p->m_primary_mean;
    j.secondary_mean = (float)p->m_secondary_mean;
    j.max_n = p->m_max_neighbours;
    return j;
}

// Sizes as api.cpp:595-611 checks them.
static void cokriging_sizes(const hpgl_cont_masked_array_t *in, const hpgl_cont_masked_array_t *sec,
                            const hpgl_cont_masked_array_t *out) {
    const long long n = volume(&in->m_shape), n2 = volume(&sec->m_shape), n3 = volume(&out->m_shape);
    if (n != n2)
        throw std::runtime_error("hpgl_simple_cokriging: Size of secondary data (" + std::to_string(n2) +
                                 ") is different from size of primary data (" + std::to_string(n) + ")");
    if (n != n3)
        throw std::runtime_error("hpgl_simple_cokriging: Size of output data (" + std::to_string(n3) +
                                 ") is different from size of primary data (" + std::to_string(n) + ")");
}

static void cokriging_host(CokrigingJob &j, const hpgl_cont_masked_array_t *in, const hpgl_cont_masked_array_t *sec,
                           hpgl_cont_masked_array_t *out) {
    const long long N = volume(&in->m_shape);
    hipStream_t s = thread_stream();
    PoolStream pool_stream(s);
    DevArray<float> di(N), ds(N), dout(N);
    DevArray<unsigned char> dm(N), dsm(N), dom(N);
    di.put(in->m_data, s);
    dm.put(in->m_mask, s);
    ds.put(sec->m_data, s);
    dsm.put(sec->m_mask, s);
    dout.put(out->m_data, s);
    dom.put(out->m_mask, s);
    j.d_in = di.p;
    j.d_in_mask = dm.p;
    j.d_sec = ds.p;
    j.d_sec_mask = dsm.p;
    j.d_out = dout.p;
    j.d_out_mask = dom.p;
    Dump d;
    progress("cokriging", 0);
    run_cokriging(j, s, take_dump(d), t_stats);
    progress("cokriging", 100);
    dout.get(out->m_data, s);
    dom.get(out->m_mask, s);
    hip_check(hipStreamSynchronize(s), "sync");
}

API void hpgl_simple_cokriging_mark1(hpgl_cont_masked_array_t *input_data, hpgl_cont_masked_array_t *secondary_data,
                                     hpgl_cokriging_m1_params_t *params, hpgl_cont_masked_array_t *output_data) {
    GUARD_BEGIN
    cokriging_sizes(input_data, secondary_data, output_data);
    CokrigingJob j = cokriging_m1_job(&input_data->m_shape, params);
    cokriging_host(j, input_data, secondary_data, output_data);
    GUARD_END
}

API void hpgl_simple_cokriging_mark2(hpgl_cont_masked_array_t *primary_data, hpgl_cont_masked_array_t *secondary_data,
                                     hpgl_cokriging_m2_params_t *params, hpgl_cont_masked_array_t *output_data) {
    GUARD_BEGIN
    cokriging_sizes(primary_data, secondary_data, output_data);
    CokrigingJob j = cokriging_m2_job(&primary_data->m_shape, params);
    cokriging_host(j, primary_data, secondary_data, output_data);
    GUARD_END
}

API void hpgl_amd_simple_cokriging_mark1_dev(const float *d_in, const unsigned char *d_in_mask, const float *d_sec,
                                             const unsigned char *d_sec_mask, const hpgl_shape_t *shape,
                                             const hpgl_cokriging_m1_params_t *params, float *d_out,
                                             unsigned char *d_out_mask, void *stream) {
    GUARD_BEGIN
    CokrigingJob j = cokriging_m1_job(shape, params);
    j.d_in = d_in;
    j.d_in_mask = d_in_mask;
    j.d_sec = d_sec;
    j.d_sec_mask = d_sec_mask;
    j.d_out = d_out;
    j.d_out_mask = d_out_mask;
    Dump d;
    run_cokriging(j, (hipStream_t)stream, take_dump(d), t_stats);
    GUARD_END
}

API void hpgl_amd_simple_cokriging_mark2_dev(const float *d_in, const unsigned char *d_in_mask, const float *d_sec,
                                             const unsigned char *d_sec_mask, const hpgl_shape_t *shape,
                                             const hpgl_cokriging_m2_params_t *params, float *d_out,
                                             unsigned char *d_out_mask, void *stream) {
    GUARD_BEGIN
    CokrigingJob j = cokriging_m2_job(shape, params);
    j.d_in = d_in;
    j.d_in_mask = d_in_mask;
    j.d_sec = d_sec;
    j.d_sec_mask = d_sec_mask;
    j.d_out = d_out;
    j.d_out_mask = d_out_mask;
    Dump d;
    run_cokriging(j, (hipStream_t)stream, take_dump(d), t_stats);
    GUARD_END
}

// Failure of a library component outside this file (the request loop):
// sets the calling thread's last-error state as a failed entry point does.
void hpgl_amd::report_error(const char *msg) {
    begin_call();
    fail(msg);
}
