// kernels.h - launch interface of the HIP kernels (kernels.hip, search.hip).
//
// Layouts in HBM (all cell arrays are dense Fortran order, cell = x + X*y + X*Y*z):
//   mask      uint8[N]                 informed flag
//   values    float[N] / uint8[N]      property (continuous / categorical)
//   targets   int32[T]                 cells to estimate, ascending
//   nbr       int32[T][max_n]          neighbour cells of each target, search order
//   cnt       int32[T]                 neighbour count (0..max_n)
//   status    uint8[T]                 0 solved, 1 no neighbours, 2 singular,
//                                      3 handed to the FP64 kernel
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hpgl_amd {

constexpr int kMaxNeighbours = 64;

enum SolveStatus : uint8_t { ST_OK = 0, ST_NO_NEIGHBOURS = 1, ST_SINGULAR = 2, ST_FALLBACK = 3 };

// Covariance as the solve kernels see it: either a box table (SK, LVM,
// IK, median IK, SGS: precalculated_covariance.h) or the analytic model
// (OK: cov_model.h) evaluated on the device.
struct DevCov {
    const double *table;  // nullptr => analytic
    int tr[3];            // table radii
    int td[3];            // table extents 2r+1
    // the table's non-negative octant with a layer of zeros appended in each
    // axis (BoxTable::fold), entry (ax, ay, az) at (az (tr[1] + 2) + ay)
    // (tr[0] + 2) + ax for ax = min(|dx|, tr[0] + 1) and so on, when every
    // entry of the table equals the octant's bit for bit (models without
    // rotation); nullptr otherwise
    const double *fold;
    double t[9];          // transform (row-major)
    double range0;
    double c0;            // sill - nugget
    int type;
};

// What the solve epilogue does with the weights.
enum EpilogueKind : int {
    EPI_CONT = 0,    // OK/SK/LVM: write estimate (or failure policy) to out
    EPI_IK = 1,      // indicator kriging: write the category probability
    EPI_MEDIAN = 2,  // median IK: double accumulation, threshold at 0.5
    EPI_SGS = 3,     // SGS: write the weights and the kriging variance
};

struct SolveArgs {
    // system
    const int *targets;
    int T;
    const int *nbr;
    const int *cnt;
    int max_n;
    int X, Y, Z;
    DevCov cov;
    int ok_system;  // 1: ordinary kriging (2 RHS, Lagrange), 0: simple kriging
    // values and means
    const float *values_f;    // EPI_CONT, EPI_SGS
    const uint8_t *values_u8;  // EPI_IK, EPI_MEDIAN
    int category;              // indicator value tested against values_u8
    int mean_kind;             // 0 none, 1 single, 2 per-cell float
    float mean_single;
    const float *mean_cells;
    int mean_on_failure;       // EPI_CONT
    int correlogram;           // scale A and b by indicator sigmas of mean_cells (FP64 kernels only)
    // outputs
    float *out_f;              // EPI_CONT estimate per cell; EPI_IK prob per target; EPI_SGS mean per target
    uint8_t *out_mask;         // EPI_CONT, EPI_MEDIAN
    uint8_t *out_u8;           // EPI_MEDIAN
    double *out_var;           // EPI_SGS variance per target (-1 on failure)
    double *out_w;             // EPI_SGS weights per target [T][max_n]
    uint8_t *status;
    unsigned long long *counters;  // [0] fp32 accepted, [1] fp64 solved, [2] singular, [3] no neighbours
    // list of target positions to solve (nullptr = all T)
    const int *subset;
    int n_subset;
    int epi;  // EpilogueKind
    // mixed kernels: targets handed to FP64 (ST_FALLBACK) are appended to
    // fb_list at positions taken from *fb_count (nullptr = not recorded)
    int *fb_list;
    int *fb_count;
};

// FP64 solves (solve.hip): group <= 32 two systems per wavefront with rows
// in registers, otherwise one system per wavefront with rows in LDS.
void launch_solve_fp64(const SolveArgs &a, int group, hipStream_t s);
// Four lanes per system (solve_quad.hip), order = bucket size: a multiple
// of 4 up to kQuadMaxMixed (mixed) or kQuadMaxFp64 (FP64).
constexpr int kQuadMaxMixed = 32;
constexpr int kQuadMaxFp64 = 16;
void launch_solve_quad(const SolveArgs &a, int order, bool fp64, hipStream_t s);

// Template scan (sugarbox_neighbour_lookup.h:36-61) over targets[list[i]].
// pred_kind 0: informed iff mask[cell] != 0.
// pred_kind 1 (SGS): informed iff key[cell] < key_of_target[i].
// pred_kind 0: cell informed iff mask[cell] != 0.
// pred_kind 1: cell informed iff key[cell] < key[target] (the target's own
// key is its path rank in sequential simulation).
struct ScanArgs {
    const uint8_t *mask;
    const int32_t *key;
    int X, Y, Z;
    const int4 *tmpl;
    int tsize;
    int max_n;
    const int *targets;
    const int *list;  // positions into targets; nullptr = all
    int n_list;
    int *nbr;
    int *cnt;
    // pred_kind 1 only, may be nullptr: bit i set when neighbour i of the
    // target is a simulated cell (key >= 0) rather than conditioning data
    uint64_t *simbits;
};
void launch_template_scan(const ScanArgs &a, int pred_kind, hipStream_t s);

// Cell bins (clusterizer.cpp): per-bin counts of informed cells and the
// decision per target whether any of its 27 bins is over the limit.
struct BinArgs {
    int X, Y, Z;
    int r[3];
    int m[3];
    int limit;
};
void launch_bin_ids(const int *cells, int n, const BinArgs &b, int *bin_of, uint32_t *bin_count, hipStream_t s);
void launch_path_flags(const int *targets, int T, const BinArgs &b, const uint32_t *bin_count, uint8_t *cluster_flag,
                       hipStream_t s);

// Cluster path (sugarbox_indexed_neighbour_lookup.h:66-123) over
// targets[list[i]]: gather the 27 bins' cells, keep rank > 0, order exactly
// as libstdc++ std::sort by descending key, keep max_n.
struct ClusterArgs {
    int X, Y, Z;
    BinArgs bins;
    const uint32_t *bin_start;
    const uint32_t *bin_count;
    const int *bin_cells;     // informed cells sorted by (bin, cell)
    // per cell of bin_cells: x = its rank-table coordinate
    // 4 (x + w[0] y + w[0] w[1] z) mod 2^32, y = x mod 2^16 | (y mod 2^16) << 16
    // (launch_pack_lin); then kClusterLinPad zero entries
    const uint2 *bin_lin;
    const uint32_t *zoff;     // per bin, cells below each z layer (launch_bin_zoff)
    // rank table (ranks shifted left by local_bits) over the covariance box
    // [-span, span]^3, widened in x by
    // kClusterRunSlack zero entries on each side, after kClusterRankPad zero
    // entries: w = (2 span[0] + 1 + 2 kClusterRunSlack, 2 span[1] + 1,
    // 2 span[2] + 1) entries, offset (dx, dy, dz) at byte
    // centre4 + 4 (dx + w[0] dy + w[0] w[1] dz)
    const uint32_t *rank;
    int span[3];
    int w[3];
    uint32_t centre4;
    int max_n;
    const int *targets;
    const int *list;
    int n_list;
    int *nbr;
    int *cnt;
    int local_bits;           // element = (rank << local_bits) | raw index (< cluster_raw_bound)
    int cap;                  // LDS list capacity, a value returned by cluster_capacity
    // main pass: wavefronts resident for the whole launch take chunks of
    // per_wave list entries from kClusterWorkParts counters (work, one
    // 64-byte line each, zero before the launch), each over a contiguous
    // part of the list
    int per_wave;
    int *work;
    // runs of neighbouring targets: the first member's list is gathered
    // into LDS, the others' into run_scratch, cluster_run_words(cap) words
    // per main-pass wavefront, run_slots wavefronts (cluster_main_waves)
    uint32_t *run_scratch;
    int run_slots;
    // lists longer than cap: the main pass appends their targets to
    // spill_list (spill_count entries; nullptr: an overflow error) and a
    // second pass of n_spill_waves wavefronts sorts them in global scratch,
    // cluster_scratch_words(scratch_per_wave) words per wavefront
    int *spill_list;
    int *spill_count;
    int n_spill_waves;
    uint32_t *scratch;
    int scratch_per_wave;
    unsigned int *overflow;   // bit 0: a list exceeded scratch capacity; bit 1: sort replay error
    unsigned long long *cand_total;  // counter lines: word 0 candidates gathered, word 1 lists sorted in scratch
};
void launch_cluster_search(const ClusterArgs &a, hipStream_t s);
// Wavefronts of the cell-binned search's main pass for these arguments
// (the grid the device holds at once, at most one workgroup per chunk).
long long cluster_main_waves(const ClusterArgs &a);
// Zero entries before and after the rank table in its device allocation:
// the search's lanes with no entry to read take the run's entries from the
// start of the allocation.
constexpr int kClusterRankPad = 8;
// Work counters of the cell-binned search's main pass (kClusterWorkStride
// ints apart).
constexpr int kClusterWorkParts = 8;
constexpr int kClusterWorkStride = 16;
// Zero rank-table columns on either side of the box in x: a run of up to
// kClusterRunSlack + 1 targets reads its members' entries together.
constexpr int kClusterRunSlack = 3;
// Run scratch words per wavefront: the later members' lists.
__host__ __device__ inline size_t cluster_run_words(int cap) { return (size_t)kClusterRunSlack * cap; }
// Entries after the bin-sorted rank-table coordinates: a gather pass reads
// up to 255 positions past the last candidate of a target.
constexpr int kClusterLinPad = 256;
// out[i] = (4 (x + w0 y + w01 z) mod 2^32, x mod 2^16 | (y mod 2^16) << 16)
// for the cell cells[i], i < n, and (0, 0) for the kClusterLinPad entries
// after them.
void launch_pack_lin(const int *cells, int n, int X, int Y, uint32_t w0, uint32_t w01, uint2 *out, hipStream_t s);
// zoff[b (r[2] + 1) + d]: cells of bin b whose z lies below the bin's first
// layer + d, for d = 0 .. r[2].
void launch_bin_zoff(const uint32_t *bin_start, const uint32_t *bin_count, const int *bin_cells, const BinArgs &b,
                     long long nbins, uint32_t *zoff, hipStream_t s);
// Raw candidate indices (positions in the concatenation of a target's 27
// bins) of a target whose bins hold at most limit cells each lie below
// this bound.
inline long long cluster_raw_bound(int limit) { return 27ll * limit + 1; }
// Scratch words per wave for lists of up to n candidates (the elements,
// rounded up to a multiple of 64 words).
__host__ __device__ inline long long cluster_scratch_words(int n) { return ((long long)n + 63) / 64 * 64; }
// LDS list capacity of the cell-binned search kernel for lists of up to
// expected_max candidates (the largest built capacity when none suffices).
int cluster_capacity(long long expected_max);

// Sequential Gaussian simulation (sgs.hip), R realisations at once.
struct SgsPathArgs {
    long long N;
    const long long *A, *C, *R0;  // LCG per realisation (path_random_generator.cpp:21-56)
    int *path;                    // [R][N] cell at each path position
    int *rank;                    // [R][N] path position of each cell
};
void launch_lcg_path(const SgsPathArgs &p, int R, hipStream_t s);
void launch_sgs_key(const uint8_t *hard, const uint8_t *sim_mask, const int *rank, long long N, int R, int *key,
                    hipStream_t s);
void launch_mt11213b(const uint32_t *seeds, const long long *counts, long long stride, int R, uint32_t *out,
                     hipStream_t s);
void launch_cdf_forward(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp, long long m,
                        hipStream_t s);
void launch_cdf_back(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp, long long m,
                     hipStream_t s);
struct SgsSweepArgs {
    long long N;
    int max_n;
    const int *T;          // [R] targets per realisation
    const long long *t_off;  // [R] offset of each realisation's targets
    const int *targets;    // cells in path order
    const int *nbr;
    const int *cnt;
    const double *w;
    const double *var;
    const uint8_t *status;  // ST_OK when the system was solved
    const uint32_t *draws;  // [R][draw_stride]
    long long draw_stride;
    const uint64_t *simbits;  // per target: neighbours that are simulated cells
    const float *lvm;       // per-cell mean (transformed) or nullptr
    float *values;          // [R][N]
    int *done;              // [R][N], 1 once the cell's value is written
    unsigned int *ticket;   // [R]
    unsigned int *error;
    long long spin_limit;
};
void launch_sgs_sweep(const SgsSweepArgs &a, int R, int waves_per_realisation, hipStream_t s);

// Sequential indicator simulation (sgs.hip): what the sweep reads of each
// kriged category, indexed by the target's position in the batch.
struct SisCategory {
    const int *nbr;           // [targets][max_n]
    const int *cnt;
    const double *w;          // [targets][max_n] simple-kriging weights
    const uint8_t *status;    // ST_OK when the system was solved
    const uint64_t *simbits;  // neighbours that are simulated cells
    const float *lvm;         // per-cell mean of the category, or nullptr
    float marginal;           // the mean when lvm is nullptr
    int max_n;
};
struct SisSweepArgs {
    long long N;
    int kriged;                // categories kriged: K, or 1 for median SIS (K == 2)
    int K;                     // categories sampled
    const SisCategory *cats;   // [kriged], device memory
    const int *T;              // [R] targets per realisation
    const long long *t_off;    // [R] offset of each realisation's targets
    const int *targets;        // cells in path order
    const uint32_t *draws;     // [R][draw_stride]
    long long draw_stride;
    uint8_t *values;           // [R][N] categories
    int *done;                 // [R][N], 1 once the cell's category is written
    unsigned int *ticket;      // [R]
    unsigned int *error;
    long long spin_limit;
};
constexpr int kSisMaxCategories = 256;
void launch_sis_sweep(const SisSweepArgs &a, int R, int waves_per_realisation, hipStream_t s);

// Simple colocated cokriging, Markov models I and II (cokriging.hip): one
// (n+1)-system per target, solved by gauss_solve in the reference's order.
struct CokrigingArgs {
    const int *targets;
    int T;
    const int *nbr;
    const int *cnt;
    int max_n;
    int X, Y, Z;
    DevCov cov;           // primary covariance between neighbours and to the target
    DevCov cross;         // covariance scaled by coef for the cross terms
    double coef;          // cross-covariance factor
    double sec_var;       // A(n, n): the secondary variance
    double cross0;        // b(n) = coef * cross(0)
    float primary_mean;
    float secondary_mean;
    const float *values;  // primary values
    const float *sec;     // secondary values and mask
    const uint8_t *sec_mask;
    float *out;
    uint8_t *out_mask;
    unsigned long long *counters;  // [1] solved, [2] singular, [3] no neighbours
};
void launch_cokriging(const CokrigingArgs &a, hipStream_t s);

// Statistics counters (counters.h): kCounterSlots lines of kCounterStride
// words; word w of every line accumulates counter w.
constexpr int kCounterSlots = 1024;
constexpr int kCounterStride = 8;
constexpr int kCounterWords = kCounterSlots * kCounterStride;

// Small helpers.
void launch_copy_informed(const float *in, const uint8_t *in_mask, long long n, float *out, uint8_t *out_mask,
                          hipStream_t s);
void launch_ik_argmax(const float *probs, int K, long long n, uint8_t *out, uint8_t *out_mask, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t s);

}  // namespace hpgl_amd
