// sgs.hip - device kernels of sequential Gaussian simulation
// (sequential_gaussian_simulation.cpp:18-128, sequential_simulation.h:51-110).
//
// The reference visits the cells along a random path (a full-period LCG
// over all cells, path_random_generator.cpp:21-94) and, at each cell,
// searches neighbours among the cells informed so far, solves the SK/OK
// system with its variance, draws u from mt11213b and writes
// mean + sqrt(var) * Phi^-1(u).  Which cells are informed when a cell is
// visited depends only on the path: hard data, plus simulated cells of
// smaller path rank.  So every part except the value itself runs in
// parallel over all cells of all realisations:
//   * k_lcg_path: path rank of every cell (LCG jump-ahead per chunk);
//   * k_sgs_key: key = -1 (hard data), rank (cell to simulate) or INT_MAX
//     (never informed); cell c is informed for the target of rank p iff
//     key[c] < p, which the template scan tests;
//   * k_mt11213b: the draws, one workgroup per realisation, 351-word
//     twists in three parallel phases;
//   * weights and variances by the solve kernels (EPI_SGS);
//   * k_sgs_sweep: the values, in path order.  Each wavefront claims 64
//     consecutive path positions from a per-realisation ticket and writes
//     a cell once all its simulated neighbours (all of smaller rank,
//     claimed earlier by running wavefronts) are marked done.  The wait is
//     bounded: past the limit the kernel sets an error flag and exits.
// Sequential indicator simulation uses the same path, keys, draws and
// template scans (one per category) and its own sweep, k_sis_sweep.
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

// (a, c) o (a, c): x -> a (a x + c) + c, modulo m
__device__ inline void affine_sq(unsigned long long &a, unsigned long long &c, unsigned long long m) {
    c = (a * c + c) % m;
    a = (a * a) % m;
}

__global__ void k_lcg_path(SgsPathArgs p) {
    const long long chunk = 256;
    const long long start = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * chunk;
    const int r = blockIdx.y;
    if (start >= p.N) return;
    const unsigned long long M = (unsigned long long)p.N;
    const unsigned long long A = (unsigned long long)p.A[r] % M, C = (unsigned long long)p.C[r] % M;
    // R_start = f^start(R_0) by squaring the affine map
    unsigned long long R = (unsigned long long)p.R0[r];
    unsigned long long a = A, c = C;
    for (long long e = start; e > 0; e >>= 1) {
        if (e & 1) R = (a * R + c) % M;
        affine_sq(a, c, M);
    }
    int *path = p.path + (size_t)r * p.N;
    int *rank = p.rank + (size_t)r * p.N;
    const long long end = start + chunk < p.N ? start + chunk : p.N;
    for (long long k = start; k < end; ++k) {
        path[k] = (int)R;
        rank[R] = (int)k;
        R = (A * R + C) % M;
    }
}

__global__ void k_sgs_key(const uint8_t *hard, const uint8_t *sim_mask, const int *rank, long long N, int R,
                          int *key) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * R) return;
    const long long c = i % N;
    int k;
    if (hard[c] != 0)
        k = -1;
    else if (sim_mask && sim_mask[c] != 1)
        k = 0x7fffffff;
    else
        k = rank[i];
    key[i] = k;
}

// std::mersenne_twister_engine<uint32_t, 32, 351, 175, 19, 0xccab8ee7, 11,
// 0xffffffff, 7, 0x31b6ab00, 15, 0xffe50000, 17, 1812433253> (mt11213b)
constexpr int kMtN = 351, kMtM = 175;
constexpr uint32_t kMtA = 0xccab8ee7u, kUpper = 0xfff80000u, kLower = 0x0007ffffu;

__device__ inline uint32_t mt_mix(uint32_t hi_word, uint32_t lo_word) {
    const uint32_t y = (hi_word & kUpper) | (lo_word & kLower);
    return (y >> 1) ^ ((y & 1u) ? kMtA : 0u);
}

// One workgroup per realisation.  Each twist writes the new state into the
// other of two LDS buffers: the first n - m words read only the old state,
// the rest (and the last word, which needs new words 0 and m - 1) read new
// words of the first part, so a twist takes two barriers; every thread
// tempers and stores its own new word as soon as it has it.
__global__ __launch_bounds__(384) void k_mt11213b(const uint32_t *seeds, const long long *counts, long long stride,
                                                 uint32_t *out) {
    __shared__ uint32_t buf[2][kMtN];
    const int r = blockIdx.x;
    const int k = threadIdx.x;
    const long long count = counts[r];
    uint32_t *o = out + (size_t)r * stride;
    if (k == 0) {
        uint32_t v = seeds[r];
        buf[0][0] = v;
        for (int i = 1; i < kMtN; ++i) {
            v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
            buf[0][i] = v;
        }
    }
    __syncthreads();
    int cur = 0;
    for (long long base = 0; base < count; base += kMtN) {
        const uint32_t *x = buf[cur];
        uint32_t *y = buf[cur ^ 1];
        uint32_t nv = 0;
        // words k < n - m: old state only
        if (k < kMtN - kMtM) {
            nv = x[k + kMtM] ^ mt_mix(x[k], x[k + 1]);
            y[k] = nv;
        }
        __syncthreads();
        // words n - m <= k < n - 1 read new word k + m - n; the last word
        // reads new words m - 1 and 0
        if (k >= kMtN - kMtM && k < kMtN - 1) {
            nv = y[k + kMtM - kMtN] ^ mt_mix(x[k], x[k + 1]);
            y[k] = nv;
        } else if (k == kMtN - 1) {
            nv = y[kMtM - 1] ^ mt_mix(x[kMtN - 1], y[0]);
            y[k] = nv;
        }
        if (k < kMtN && base + k < count) {
            uint32_t z = nv;
            z ^= (z >> 11) & 0xffffffffu;
            z ^= (z << 7) & 0x31b6ab00u;
            z ^= (z << 15) & 0xffe50000u;
            z ^= (z >> 17);
            o[base + k] = z;
        }
        __syncthreads();
        cur ^= 1;
    }
}

// gaussian_cdf_t (gaussian_distribution.h:26-63)
__device__ inline double gauss_prob(double x, double mean, double var) {
    const double p = 0.2316419;
    const double b0 = 0.319381530, b1 = -0.356563782, b2 = 1.781477937, b3 = -1.821255978, b4 = 1.330274429;
    double sx = (x - mean) / sqrt(var);
    const bool neg = sx < 0;
    if (neg) sx = -sx;
    const double t = 1 / (1 + p * sx);
    const double zx = 1 / sqrt(2 * 3.14159265358) * exp(-sx * sx / 2);
    double result = 1 - zx * ((((b4 * t + b3) * t + b2) * t + b1) * t + b0) * t;
    if (neg) result = 1 - result;
    return result;
}
__device__ inline double gauss_inverse(double p, double mean, double var) {
    const double p0 = -0.322232431088, p1 = -1.0, p2 = -0.342242088547, p3 = -0.0204231210245,
                 p4 = -0.453642210148e-4;
    const double q0 = 0.0993484626060, q1 = 0.588581570495, q2 = 0.531103462366, q3 = 0.103537752850,
                 q4 = 0.38560700634e-2;
    const double pp = (p > 0.5) ? (1 - p) : p;
    const double y = sqrt(log(1 / (pp * pp)));
    double xp = y + ((((p4 * y + p3) * y + p2) * y + p1) * y + p0) / ((((q4 * y + q3) * y + q2) * y + q1) * y + q0);
    if (p < 0.5) xp = -xp;
    return xp * sqrt(var) + mean;
}

// non_parametric_cdf_2_t (non_parametric_cdf.h:181-262), float arithmetic
__device__ inline int lower_bound_f(const float *a, long long n, float v) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (int)lo;
}
__device__ inline float np_prob(const float *values, const float *probs, long long n, float value) {
    const int it = lower_bound_f(values, n, value);
    if (it == n) return probs[n - 1];
    if (it == 0) return probs[0];
    const float x1 = values[it - 1], x2 = values[it], y1 = probs[it - 1], y2 = probs[it];
    return y1 + (y2 - y1) / (x2 - x1) * (value - x1);
}
__device__ inline float np_inverse(const float *values, const float *probs, long long n, float prob) {
    if (n <= 0) return 0;
    const int it = lower_bound_f(probs, n, prob);
    if (it == 0) return values[0];
    if (it == n) return values[n - 1];
    const float x1 = values[it - 1], x2 = values[it], y1 = probs[it - 1], y2 = probs[it];
    return x1 + (x2 - x1) / (y2 - y1) * (prob - y1);
}

// transform_cdf_p (cdf_utils.h:14-24) forward: v -> Phi^-1(F_np(v)) on informed cells
__global__ void k_cdf_forward(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp,
                              long long m) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || mask[i] == 0) return;
    const double P = np_prob(cv, cp, m, v[i]);
    v[i] = (float)gauss_inverse(P, 0.0, 1.0);
}
// forward transform of every cell (the LVM mean, sequential_gaussian_simulation.cpp:104)
__global__ void k_cdf_forward_all(float *v, long long n, const float *cv, const float *cp, long long m) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double P = np_prob(cv, cp, m, v[i]);
    v[i] = (float)gauss_inverse(P, 0.0, 1.0);
}
// back-transform of informed cells: v -> F_np^-1((float) Phi(v))
__global__ void k_cdf_back(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp,
                           long long m) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || mask[i] == 0) return;
    const double P = gauss_prob(v[i], 0.0, 1.0);
    v[i] = np_inverse(cv, cp, m, (float)P);
}

// Sequential indicator simulation (sequential_indicator_simulation.cpp:67-121)
// over the same ticketed sweep: a target waits for the simulated neighbours
// of each kriged category in turn (a flag once set stays set), then forms
// each category's probability and samples a category.
//
// Probability of category c (kriging_interpolation.h:239-272, combine
// <uchar, double>, combiner.h:10-24): the mean m_c of the target (LVM
// value or marginal probability), plus w_i ((float)[v_i == c] - m_i) over
// the neighbours in search order, in double; the mean alone when the
// system failed or had no neighbours.  Stored as float.  Median SIS
// (K == 2) krige category 0 only and use {p, 1 - p}.
//
// sample (sample.cpp:16-29): s = u * sum of the K probabilities (double,
// in order), u = draw / (2^32 - 1); the first category whose running sum
// exceeds s, else the last.  The probabilities stay in LDS, [K][64].
__device__ inline void sis_value(const SisSweepArgs &a, long long gk, int k_draw, const uint32_t *draws,
                                 uint8_t *val, float *probs, int lane, int cell) {
    double sum = 0.0;
    for (int c = 0; c < a.kriged; ++c) {
        const SisCategory &C = a.cats[c];
        const float mc = C.lvm ? C.lvm[cell] : C.marginal;
        double prob = mc;
        if (C.status[gk] == ST_OK) {
            const int n = C.cnt[gk];
            const int *row = C.nbr + (size_t)gk * C.max_n;
            const double *w = C.w + (size_t)gk * C.max_n;
            double r = mc;
            // eight neighbours' loads in flight together, then the terms in
            // neighbour order
            for (int i0 = 0; i0 < n; i0 += 8) {
                int nb[8];
                double wv[8];
                float vv[8], mv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) nb[j] = i0 + j < n ? row[i0 + j] : 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const bool in = i0 + j < n;
                    wv[j] = in ? w[i0 + j] : 0.0;
                    vv[j] = (in && val[nb[j]] == (uint8_t)c) ? 1.0f : 0.0f;
                    mv[j] = (in && C.lvm) ? C.lvm[nb[j]] : mc;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (i0 + j < n) r += wv[j] * (double)(vv[j] - mv[j]);
            }
            prob = r;
        }
        const float p0 = (float)prob;
        probs[c * 64 + lane] = p0;
        sum += (double)p0;
        if (a.kriged == 1) {
            const float p1 = (float)(1 - prob);
            probs[64 + lane] = p1;
            sum += (double)p1;
        }
    }
    const double u = (double)draws[k_draw] / 4294967295.0;
    const double s = u * sum;
    double p = 0.0;
    int pick = a.K - 1;
    for (int c = 0; c < a.K; ++c) {
        p += (double)probs[c * 64 + lane];
        if (s < p) {
            pick = c;
            break;
        }
    }
    val[cell] = (uint8_t)pick;
}

__device__ inline int lanes_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The ordered sweep shared by both simulations.  Each lane holds one path
// position at a time and claims the next free one (one atomic per
// wavefront for all lanes that need work) as soon as its cell is written,
// so no lane idles while others of its wavefront wait.  A position waits
// only for positions claimed before it, all held by running lanes, and the
// smallest unfinished position has nothing left to wait for, so the sweep
// always progresses.  The wait is bounded by spin_limit passes without
// progress in the wavefront; past it the error flag is set and every
// wavefront leaves.
//
// Job supplies, per lane: start(k) (the waitset of simulated neighbours of
// path position k and its neighbour row), next_set() (moves to the next
// neighbour list once the current one is done; false when there is none)
// and write(k) (the value of position k; returns its cell).
template <typename Job>
__device__ inline void ordered_sweep(Job &job, int T, unsigned int *ticket, int *done, unsigned int *error,
                                     long long spin_limit) {
    const int lane = threadIdx.x;
    bool pending = false;
    bool drained = false;  // the ticket has passed T: no more claims
    int k = 0;
    long long spins = 0;
    while (true) {
        if (!drained) {
            const uint64_t need = __ballot(!pending);
            if (need) {
                const int leader = __ffsll((unsigned long long)need) - 1;
                int base = 0;
                if (lane == leader) base = (int)atomicAdd(ticket, (unsigned)__popcll(need));
                base = __shfl(base, leader);
                if (base + __popcll(need) >= T) drained = true;
                if (!pending) {
                    const int kn = base + lanes_below(need);
                    if (kn < T) {
                        k = kn;
                        pending = true;
                        job.start(k);
                    }
                }
            }
        }
        if (!__any(pending)) break;
        bool progress = false;
        if (pending) {
            while (job.waitset == 0 && job.next_set()) {
            }
            // poll up to eight outstanding flags per pass, loads in flight together
            int idx[8];
            uint64_t ws = job.waitset;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                idx[q] = ws ? __ffsll((unsigned long long)ws) - 1 : -1;
                if (ws) ws &= ws - 1;
            }
            int fl[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                fl[q] = idx[q] >= 0
                            ? __hip_atomic_load(done + job.row[idx[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : 0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (idx[q] >= 0 && fl[q] != 0) job.waitset &= ~(1ull << idx[q]);
            if (job.waitset == 0 && !job.next_set()) {
                // every neighbour value is published: order the reads after the flags
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                const int cell = job.write(k);
                __hip_atomic_store(done + cell, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                pending = false;
                progress = true;
            }
        }
        if (__any(progress)) {
            spins = 0;
        } else {
            if (++spins > spin_limit) {
                atomicOr(error, 1u);
                break;
            }
            if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// Gaussian simulation: the value mean + sqrt(var) Phi^-1(u) of combine
// (combiner.h:10-24) in neighbour order, N(0, 1) when the system failed.
struct SgsJobLane {
    const SgsSweepArgs &a;
    long long off;
    const uint32_t *draws;
    float *val;
    uint64_t waitset = 0;
    const int *row = nullptr;
    long long gk = 0;
    __device__ SgsJobLane(const SgsSweepArgs &args, long long o, const uint32_t *d, float *v)
        : a(args), off(o), draws(d), val(v) {}
    __device__ void start(int k) {
        gk = off + k;
        waitset = a.simbits[gk];
        row = a.nbr + (size_t)gk * a.max_n;
    }
    __device__ bool next_set() { return false; }
    __device__ int write(int k) {
        const int cell = a.targets[gk];
        const int n = a.cnt[gk];
        const bool ok = a.status[gk] == ST_OK;
        const double v0 = a.var[gk];
        const float m0 = a.lvm ? a.lvm[cell] : 0.0f;
        float mean = m0;
        if (ok) {
            const double *w = a.w + (size_t)gk * a.max_n;
            // eight neighbours' loads in flight together, then the terms in
            // neighbour order
            for (int i0 = 0; i0 < n; i0 += 8) {
                int nb[8];
                double wv[8];
                float vv[8], mv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) nb[j] = i0 + j < n ? row[i0 + j] : 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const bool in = i0 + j < n;
                    wv[j] = in ? w[i0 + j] : 0.0;
                    vv[j] = in ? val[nb[j]] : 0.0f;
                    mv[j] = (in && a.lvm) ? a.lvm[nb[j]] : 0.0f;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (i0 + j < n) mean += (float)(wv[j] * (double)(vv[j] - mv[j]));
            }
        }
        // u = x / (2^32 - 1) (bs_random_generator.cpp:25-28)
        const double u = (double)draws[k] / 4294967295.0;
        const double value = ok ? gauss_inverse(u, (double)mean, v0) : gauss_inverse(u, 0.0, 1.0);
        val[cell] = (float)value;
        return cell;
    }
};

__global__ __launch_bounds__(64) void k_sgs_sweep(SgsSweepArgs a) {
    const int r = blockIdx.y;
    SgsJobLane job(a, a.t_off[r], a.draws + (size_t)r * a.draw_stride, a.values + (size_t)r * a.N);
    ordered_sweep(job, a.T[r], a.ticket + r, a.done + (size_t)r * a.N, a.error, a.spin_limit);
}

// Indicator simulation: waits for the simulated neighbours of each kriged
// category in turn (a flag once set stays set), then sis_value.
struct SisJobLane {
    const SisSweepArgs &a;
    long long off;
    const uint32_t *draws;
    uint8_t *val;
    float *probs;
    uint64_t waitset = 0;
    const int *row = nullptr;
    long long gk = 0;
    int wc = 0;
    __device__ SisJobLane(const SisSweepArgs &args, long long o, const uint32_t *d, uint8_t *v, float *p)
        : a(args), off(o), draws(d), val(v), probs(p) {}
    __device__ void start(int k) {
        gk = off + k;
        wc = 0;
        waitset = a.cats[0].simbits[gk];
        row = a.cats[0].nbr + (size_t)gk * a.cats[0].max_n;
    }
    __device__ bool next_set() {
        if (wc + 1 >= a.kriged) return false;
        ++wc;
        waitset = a.cats[wc].simbits[gk];
        row = a.cats[wc].nbr + (size_t)gk * a.cats[wc].max_n;
        return true;
    }
    __device__ int write(int k) {
        const int cell = a.targets[gk];
        sis_value(a, gk, k, draws, val, probs, threadIdx.x, cell);
        return cell;
    }
};

__global__ __launch_bounds__(64) void k_sis_sweep(SisSweepArgs a) {
    extern __shared__ float sis_probs[];
    const int r = blockIdx.y;
    SisJobLane job(a, a.t_off[r], a.draws + (size_t)r * a.draw_stride, a.values + (size_t)r * a.N, sis_probs);
    ordered_sweep(job, a.T[r], a.ticket + r, a.done + (size_t)r * a.N, a.error, a.spin_limit);
}

inline unsigned blocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_lcg_path(const SgsPathArgs &p, int R, hipStream_t s) {
    const long long threads = (p.N + 255) / 256;
    hipLaunchKernelGGL(k_lcg_path, dim3(blocks(threads, 64), R), dim3(64), 0, s, p);
}

void launch_sgs_key(const uint8_t *hard, const uint8_t *sim_mask, const int *rank, long long N, int R, int *key,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_sgs_key, dim3(blocks(N * R, 256)), dim3(256), 0, s, hard, sim_mask, rank, N, R, key);
}

void launch_mt11213b(const uint32_t *seeds, const long long *counts, long long stride, int R, uint32_t *out,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_mt11213b, dim3(R), dim3(384), 0, s, seeds, counts, stride, out);
}

void launch_cdf_forward(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp, long long m,
                        hipStream_t s) {
    if (n <= 0) return;
    if (mask)
        hipLaunchKernelGGL(k_cdf_forward, dim3(blocks(n, 256)), dim3(256), 0, s, v, mask, n, cv, cp, m);
    else
        hipLaunchKernelGGL(k_cdf_forward_all, dim3(blocks(n, 256)), dim3(256), 0, s, v, n, cv, cp, m);
}

void launch_cdf_back(float *v, const uint8_t *mask, long long n, const float *cv, const float *cp, long long m,
                     hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cdf_back, dim3(blocks(n, 256)), dim3(256), 0, s, v, mask, n, cv, cp, m);
}

void launch_sgs_sweep(const SgsSweepArgs &a, int R, int waves_per_realisation, hipStream_t s) {
    hipLaunchKernelGGL(k_sgs_sweep, dim3(waves_per_realisation, R), dim3(64), 0, s, a);
}

void launch_sis_sweep(const SisSweepArgs &a, int R, int waves_per_realisation, hipStream_t s) {
    const size_t lds = (size_t)a.K * 64 * sizeof(float);
    hipLaunchKernelGGL(k_sis_sweep, dim3(waves_per_realisation, R), dim3(64), lds, s, a);
}

}  // namespace hpgl_amd
