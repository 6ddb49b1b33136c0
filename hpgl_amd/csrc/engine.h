// engine.h - orchestration of one kriging job on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "model.h"

namespace hpgl_amd {

struct Stats {
    long long nodes_target = 0;
    long long nodes_kriged = 0;
    long long nodes_no_neighbours = 0;
    long long nodes_singular = 0;
    long long solves_fp32 = 0;
    long long solves_fp64 = 0;
    long long solves_fp32_two_steps = 0;
    long long nodes_template_path = 0;
    long long nodes_cluster_path = 0;
    double ms_setup_host = 0;
    double ms_search = 0;
    double ms_solve = 0;
    double ms_device_total = 0;
    double ms_wall = 0;
    // per-kernel device time (HIP events around each launch) and launch counts
    double ms_k_template = 0;
    double ms_k_cluster = 0;
    double ms_k_solve_mixed = 0;
    double ms_k_solve_fp64 = 0;
    long long n_k_template = 0;
    long long n_k_cluster = 0;
    long long n_k_solve_mixed = 0;
    long long n_k_solve_fp64 = 0;
    long long cluster_candidates = 0;  // candidate cells gathered by the cluster path
    long long neighbours_total = 0;    // sum of neighbour counts over solved systems
    long long cluster_lists_spilled = 0;  // cell-binned lists longer than the LDS capacity
};

// Neighbour dump requested by the caller (host buffers).
struct Dump {
    int *indices = nullptr;
    int *counts = nullptr;
    long long n_cells = 0;
    int max_n = 0;
};

// A continuous-variable kriging job (OK, SK, LVM) on device buffers.
struct ContJob {
    int grid[3];
    CovSpec cov;
    int radii[3];
    int max_n;
    bool ordinary;          // OK system and analytic covariance
    int mean_kind;          // 0 none, 1 single, 2 per-cell
    double mean_single;     // mean_kind 1 (SK: given or computed from the data)
    bool mean_auto;         // SK automatic mean (calc_mean.cpp:6-25)
    const float *d_mean = nullptr;
    bool mean_on_failure;
    const float *d_in;
    const uint8_t *d_in_mask;
    float *d_out;
    uint8_t *d_out_mask;
};

struct IndicatorParams {
    CovSpec cov;
    int radii[3];
    int max_n;
    double marginal;  // IK: marginal_prob of the category; median IK: marginal_probs[1]
};

// Indicator kriging (K >= 1 categories) or median IK (median = true, one
// parameter set, indicator value 1).
struct IndJob {
    int grid[3];
    std::vector<IndicatorParams> cats;
    bool median;
    const uint8_t *d_in;
    const uint8_t *d_in_mask;
    uint8_t *d_out;
    uint8_t *d_out_mask;
};

// R realisations of sequential Gaussian simulation on device buffers.
struct SgsJob {
    int grid[3];
    CovSpec cov;
    int radii[3];
    int max_n;
    bool ordinary;              // OK system (kriging kind 0, no LVM); SK with mean 0 otherwise
    std::vector<long> seeds;    // one per realisation
    const float *d_in;          // conditioning values (original units)
    const uint8_t *d_in_mask;
    const float *d_lvm = nullptr;       // per-cell mean (original units) or nullptr
    const uint8_t *d_sim_mask = nullptr;  // cells with value 1 are simulated; nullptr = all
    bool has_cdf = false;
    std::vector<float> cdf_values, cdf_probs;
    float *d_out;               // [R][N]
    uint8_t *d_out_mask;        // [R][N]
};

// R realisations of sequential indicator simulation (K >= 2 categories;
// K == 2 is median SIS) on device buffers.
struct SisJob {
    int grid[3];
    std::vector<IndicatorParams> cats;  // marginal: the category's marginal probability
    std::vector<long> seeds;            // one per realisation
    const uint8_t *d_in;                // conditioning categories
    const uint8_t *d_in_mask;
    std::vector<const float *> d_lvm;   // per category, per-cell mean; empty: marginal probabilities
    bool correlogram = false;           // LVM only: correlogram weights (corellogramed_weights_3)
    const uint8_t *d_sim_mask = nullptr;  // cells with value 1 are simulated; nullptr = all
    uint8_t *d_out;                     // [R][N]
    uint8_t *d_out_mask;                // [R][N]
};

// Simple colocated cokriging (simple_cokriging_markI.cpp:185-305).  Model
// I: the cross covariance is the primary model scaled by
// rho sqrt(var2) / sqrt(C1(0)); model II: the secondary model scaled by
// rho sqrt(C2(0) / C1(0)), var2 = C2(0).
struct CokrigingJob {
    int grid[3];
    CovSpec primary;
    bool mark2 = false;
    CovSpec secondary;          // model II
    double secondary_variance;  // model I
    double correlation;
    float primary_mean;         // mean_t (float) in the reference
    float secondary_mean;
    int radii[3];
    int max_n;
    const float *d_in;
    const uint8_t *d_in_mask;
    const float *d_sec;
    const uint8_t *d_sec_mask;
    float *d_out;
    uint8_t *d_out_mask;
};

void run_cont(const ContJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st);
void run_ind(const IndJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st);
void run_sgs(const SgsJob &job, hipStream_t s, int precision, const Dump *dump, Stats &st);
void run_sis(const SisJob &job, hipStream_t s, int precision, Stats &st);
void run_cokriging(const CokrigingJob &job, hipStream_t s, const Dump *dump, Stats &st);

// Device memory held by the library between calls (size-class cache).
// Frees are ordered on the stream named by the innermost PoolStream of the
// calling thread: a freed block is reused only after the work queued on that
// stream before the free has completed, or by a stream made to wait for it.
void *pool_alloc(size_t bytes);
void pool_free(void *p);
// Blocks on the free list whose freeing stream has not yet reached the free.
size_t pool_pending_blocks();
// Checks the stream ordering of pool frees on the current device (0 holds,
// 1 inconclusive, negative broken).
int pool_selftest();

// Names the stream that the calling thread's pool frees are ordered on
// until the scope ends.
class PoolStream {
  public:
    explicit PoolStream(hipStream_t s);
    ~PoolStream();
    PoolStream(const PoolStream &) = delete;
    PoolStream &operator=(const PoolStream &) = delete;

  private:
    hipStream_t prev_;
    bool prev_set_;
};

void hip_check(hipError_t e, const char *what);

// Sets the calling thread's last-error state (api.hip).
void report_error(const char *msg);

}  // namespace hpgl_amd
