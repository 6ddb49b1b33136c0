// solve.hip - per-node kriging systems in FP64 in the reference's operation
// order: the plain-FP64 path, the fallback of the mixed-precision quad
// kernels (solve_quad.hip) and the path for systems wider than 32.
//
// k_solve_fp64<G> (G = 32): two systems per wavefront, lane i of a group
//   holds row i of the factor in registers; row j and the solution
//   components are broadcast with v_readlane.
// k_solve_fp64_lds (orders 33..64): one system per wavefront, rows of the
//   factor in LDS (stride 65), row j read as a broadcast.
// Both follow dpotf2 + dpotrs (left-looking sequential dot products,
// reciprocal scaling, division in the triangular solves, sequential
// subtraction) with FMA contraction off: with the same A and b they return
// the reference's weights bit for bit, and the same success/failure
// decision.  Rows n..G-1 of a group are padded with the identity.
//
// The combine (combiner.h:10-24) is evaluated exactly as the reference:
// each term (float)(w_i * (double)(v_i - m_i)) is formed per lane, then
// summed in float in neighbour order by the group's first lane.
#include "counters.h"
#include "devcov.h"
#include "divrn.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

template <int G>
struct Shared64 {
    static constexpr int NG = 64 / G;
    double y1[NG][G];
    double y2[NG][G];
    float term[NG][G];
    double dterm[NG][G];
    double sig[NG][G];
    int xyz[NG][3][G];
};

// Group-wide reductions over the G lanes of a group (xor shuffles stay
// inside an aligned power-of-two segment).
template <int G>
__device__ inline double group_max(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, G));
    return v;
}
template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

// Loads the group's neighbour coordinates into LDS and the target's own
// coordinates into registers; returns this lane's neighbour cell (-1 past n).
template <int G>
__device__ inline int stage_coords(const SolveArgs &a, int t, int n, int g, int i, int (*xyz)[3][G], int &cx, int &cy,
                                   int &cz, int &xi, int &yi, int &zi) {
    const int node = a.targets[t];
    cell_xyz(node, a.X, a.Y, cx, cy, cz);
    int nb = -1;
    xi = yi = zi = 0;
    if (i < n) {
        nb = a.nbr[(size_t)t * a.max_n + i];
        cell_xyz(nb, a.X, a.Y, xi, yi, zi);
    }
    xyz[g][0][i] = xi;
    xyz[g][1][i] = yi;
    xyz[g][2][i] = zi;
    return nb;
}

// Mean of neighbour nb.
__device__ inline float nb_mean(const SolveArgs &a, int nb) {
    if (a.mean_kind == 2) return a.mean_cells[nb];
    if (a.mean_kind == 1) return a.mean_single;
    return 0.0f;
}

// Sigma of an indicator with mean m, m moved off 0 and 1 by 1e-5
// (corellogramed_weights_3, my_kriging_weights.h:275-305).
__device__ inline double indicator_sigma(double m) {
    const double delta = 0.00001;
    if (m == 0) m += delta;
    if (m == 1) m -= delta;
    return sqrt(m * (1 - m));
}

// Correlogram systems (corellogramed_weights_3): A(i,j) = C(x_i, x_j)
// (s_i s_j) and b_i = C(x_i, x_0) (s_i s_0) with the sigmas of the per-cell
// means; stores this lane's sigma for the group and returns s_i s_0 (1 for
// plain systems, whose entries are left as they are).
template <int G>
__device__ inline double stage_sigma(const SolveArgs &a, int t, int n, int g, int i, int nb, double (*sig)[G]) {
    if (!a.correlogram) return 1.0;
    const double s0 = indicator_sigma((double)a.mean_cells[a.targets[t]]);
    const double si = i < n ? indicator_sigma((double)a.mean_cells[nb]) : 1.0;
    sig[g][i] = si;
    return si * s0;
}

// Writes the result of one system.  `ok` false means failure (singular or
// no neighbours); w is this lane's weight, term buffers are the group's.
// Every lane of the wave calls this (it holds a barrier); `write` is false
// for groups without a target or whose system went to the FP64 kernel.
template <int G>
__device__ inline void epilogue(const SolveArgs &a, bool write, int t, int n, int g, int i, int nb, bool ok, double w,
                                double variance, float (*term)[G], double (*dterm)[G]) {
    const int node = a.targets[t];
    const int EPI = a.epi;
    if (EPI == EPI_SGS) {
        if (write && ok && i < n) a.out_w[(size_t)t * a.max_n + i] = w;
        if (write && i == 0) a.out_var[t] = ok ? variance : -1.0;
        __syncthreads();
        return;
    }
    float m0;
    if (a.mean_kind == 2)
        m0 = a.mean_cells[node];
    else if (a.mean_kind == 1)
        m0 = a.mean_single;
    else
        m0 = 0.0f;
    if (write && ok && i < n) {
        if (EPI == EPI_CONT) {
            const float v = a.values_f[nb];
            const float m = nb_mean(a, nb);
            term[g][i] = (float)(w * (double)(v - m));
        } else {
            const float v = (a.values_u8[nb] == (uint8_t)a.category) ? 1.0f : 0.0f;
            const float m = nb_mean(a, nb);
            if (EPI == EPI_MEDIAN)
                dterm[g][i] = (double)(w * (double)(v - m));
            else
                term[g][i] = (float)(w * (double)(v - m));
        }
    }
    __syncthreads();
    if (!write || i != 0) return;
    if (EPI == EPI_CONT) {
        if (ok) {
            float r = m0;
            for (int k = 0; k < n; ++k) r += term[g][k];
            a.out_f[node] = r;
            a.out_mask[node] = 1;
        } else if (a.mean_on_failure) {
            a.out_f[node] = m0;
            a.out_mask[node] = 1;
        }
    } else if (EPI == EPI_IK) {
        float r = m0;
        if (ok)
            for (int k = 0; k < n; ++k) r += term[g][k];
        a.out_f[t] = r;
    } else if (EPI == EPI_MEDIAN) {
        double r = (double)m0;
        if (ok)
            for (int k = 0; k < n; ++k) r += dterm[g][k];
        a.out_u8[node] = r <= 0.5 ? 0 : 1;
        a.out_mask[node] = 1;
    }
}

// ---------------------------------------------------------------------------
// FP64 kernel in the reference's operation order (contraction is off for
// the whole translation unit; every product and sum rounds separately).
// ---------------------------------------------------------------------------
// Value of v held by lane j of this lane's group (j uniform), through
// v_readlane: no LDS and no barrier.
template <int G>
__device__ inline double group_bcast(double v, int j, int g) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo0 = __builtin_amdgcn_readlane((uint32_t)b, j);
    const uint32_t hi0 = __builtin_amdgcn_readlane((uint32_t)(b >> 32), j);
    if (G == 64) return __longlong_as_double(((uint64_t)hi0 << 32) | lo0);
    const uint32_t lo1 = __builtin_amdgcn_readlane((uint32_t)b, G + j);
    const uint32_t hi1 = __builtin_amdgcn_readlane((uint32_t)(b >> 32), G + j);
    return g == 0 ? __longlong_as_double(((uint64_t)hi0 << 32) | lo0)
                  : __longlong_as_double(((uint64_t)hi1 << 32) | lo1);
}

// FP64 in the reference's dpotf2 / dpotrs operation order, one system per
// G lanes, lane i holding row i of the factor in registers.  Row j and the
// solution components are broadcast with v_readlane, so the factorisation
// and both triangular solves use no LDS and no barrier.
template <int G>
__global__ __launch_bounds__(64) void k_solve_fp64(SolveArgs a) {
    const int EPI = a.epi;
    constexpr int NG = 64 / G;
    __shared__ Shared64<G> sh;
    const int lane = threadIdx.x;
    const int g = lane / G, i = lane % G;
    const int pos = blockIdx.x * NG + g;
    const int n_items = a.subset ? a.n_subset : a.T;
    const bool valid = pos < n_items;
    const int t = valid ? (a.subset ? a.subset[pos] : pos) : 0;
    const int n = valid ? min(a.cnt[t], G) : 0;

    int cx, cy, cz, xi, yi, zi;
    const int nb = stage_coords<G>(a, t, n, g, i, sh.xyz, cx, cy, cz, xi, yi, zi);
    const double sb = stage_sigma<G>(a, t, n, g, i, nb, sh.sig);
    __syncthreads();

    // A(i,j) = C(x_i, x_j) (my_kriging_weights.h:156-165), identity padding
    double u[G];
    double bi;
    if (a.cov.table) {
        bi = (i < n) ? cov_tab(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const double e = cov_tab(a.cov, sh.xyz[g][0][j] - xi, sh.xyz[g][1][j] - yi, sh.xyz[g][2][j] - zi);
            u[j] = (i < n && j < n) ? e : ((i == j) ? 1.0 : 0.0);
        }
    } else {
        bi = (i < n) ? cov_eval(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (i < n && j < n)
                u[j] = cov_eval(a.cov, sh.xyz[g][0][j] - xi, sh.xyz[g][1][j] - yi, sh.xyz[g][2][j] - zi);
            else
                u[j] = (i == j) ? 1.0 : 0.0;
        }
    }
    if (a.correlogram) {
        const double si = sh.sig[g][i];
        bi = bi * sb;
#pragma unroll
        for (int j = 0; j < G; ++j)
            if (i < n && j < n) u[j] = u[j] * (si * sh.sig[g][j]);
    }
    // the largest system of the wave bounds the unrolled loops
    int nmax = n;
#pragma unroll
    for (int o = 32; o >= G; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    nmax = __builtin_amdgcn_readfirstlane(nmax);

    // dpotf2 'U' as rows of L = U^T: lane i holds L(i, 0..i).  Step j:
    // lane j forms the pivot from its own row; every lane below forms
    // (A_ij + (-1) * sum_k L_ik L_jk) * (1 / L_jj) with row j broadcast.
    int info = 0;
    double diag = 1.0;  // L(i, i) of this lane's row
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (j < nmax) {
            double temp = 0.0;
#pragma unroll
            for (int k = 0; k < j; ++k) {
                const double ljk = group_bcast<G>(u[k], j, g);
                temp = temp + u[k] * ljk;
            }
            // pivot on lane j, whose temp is the sequential sum of its squares
            double ajj = u[j] - temp;
            const bool fail = i == j && info == 0 && j < n && (ajj <= 0.0 || ajj != ajj);
            if (!(ajj > 0.0)) ajj = 1.0;  // failed or padding: keep arithmetic finite
            const double sq = sqrt(ajj);
            const double rcp_l = 1.0 / sq;
            if (i == j) {
                u[j] = sq;
                diag = sq;
            }
            const double rcp = group_bcast<G>(rcp_l, j, g);
            const uint64_t fm = __ballot(fail);
            const bool group_fail = ((fm >> (g * G)) & ((G == 64) ? ~0ull : ((1ull << G) - 1))) != 0;
            if (info == 0 && group_fail) info = j + 1;
            if (i > j) u[j] = rcp * (u[j] + (-1.0) * temp);
        }
    }

    const bool two = a.ok_system != 0;
    // dpotrs: dtrsm('L','U','T','N') (forward with L) then
    // dtrsm('L','U','N','N') (backward with L^T), both right-hand sides;
    // the divisions by L(i, i) go through its reciprocal (div_rn, exact)
    const double dinv = 1.0 / diag;
    double r1 = bi, r2 = two ? ((i < n) ? 1.0 : 0.0) : 0.0;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k < nmax) {
            if (i == k) {
                r1 = div_rn(r1, diag, dinv);
                if (two) r2 = div_rn(r2, diag, dinv);
            }
            const double y1 = group_bcast<G>(r1, k, g);
            const double y2 = group_bcast<G>(r2, k, g);
            if (i > k) {
                r1 = r1 - u[k] * y1;
                r2 = r2 - u[k] * y2;
            }
        }
    }
#pragma unroll
    for (int k = G - 1; k >= 0; --k) {
        if (k < nmax) {
            if (i == k) {
                if (r1 != 0.0) r1 = div_rn(r1, diag, dinv);
                if (r2 != 0.0) r2 = div_rn(r2, diag, dinv);
            }
            const double y1 = group_bcast<G>(r1, k, g);
            const double y2 = group_bcast<G>(r2, k, g);
            // L(k, i): element i of row k, broadcast from lane k
            double lki = 0.0;
#pragma unroll
            for (int c = 0; c < k; ++c) {
                const double v = group_bcast<G>(u[c], k, g);
                if (i == c) lki = v;
            }
            if (i < k) {
                if (y1 != 0.0) r1 = r1 - y1 * lki;
                if (y2 != 0.0) r2 = r2 - y2 * lki;
            }
        }
    }
    const double x1 = r1, x2 = r2;

    double w = x1;
    double variance = 0.0;
    // Sequential sums in neighbour order (my_kriging_weights.h:217-231, 110-127).
    sh.y1[g][i] = x1;
    sh.y2[g][i] = x2;
    sh.dterm[g][i] = bi;
    __syncthreads();
    if (two) {
        double s_sk = 0, s_one = 0;
        for (int k = 0; k < n; ++k) {
            s_sk += sh.y1[g][k];
            s_one += sh.y2[g][k];
        }
        const double mu = (s_sk - 1) / s_one;
        w = x1 - mu * x2;
    }
    __syncthreads();
    if (EPI == EPI_SGS) {
        sh.y1[g][i] = w;
        __syncthreads();
        double v = a.cov.c0;
        for (int k = 0; k < n; ++k) v -= sh.y1[g][k] * sh.dterm[g][k];
        if (two && n > 0) v -= sh.y1[g][n - 1];
        variance = v;
        __syncthreads();
    }

    const bool ok = n > 0 && info == 0;
    if (valid && i == 0) a.status[t] = n == 0 ? ST_NO_NEIGHBOURS : (ok ? ST_OK : ST_SINGULAR);
    wave_count(a.counters, 3, valid && i == 0 && n == 0);
    wave_count(a.counters, 1, valid && i == 0 && ok);
    wave_count(a.counters, 2, valid && i == 0 && n > 0 && !ok);
    epilogue<G>(a, valid, t, n, g, i, nb, ok, w, variance, sh.term, sh.dterm);
}

// k_solve_fp64 for orders 33..64: one system per wavefront with the rows of
// the factor in LDS (row i at stride 65, so a lane's own row reads are
// conflict-free and row j reads are broadcasts); same operation order.
struct ShLds64 {
    double U[64][65];
    double y1[1][64];
    double y2[1][64];
    float term[1][64];
    double dterm[1][64];
    double sig[1][64];
    int xyz[1][3][64];
};

__device__ inline void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64) void k_solve_fp64_lds(SolveArgs a) {
    constexpr int G = 64;
    const int EPI = a.epi;
    __shared__ ShLds64 sh;
    const int i = threadIdx.x;
    const int pos = blockIdx.x;
    const int n_items = a.subset ? a.n_subset : a.T;
    const bool valid = pos < n_items;
    const int t = valid ? (a.subset ? a.subset[pos] : pos) : 0;
    const int n = valid ? min(a.cnt[t], G) : 0;

    int cx, cy, cz, xi, yi, zi;
    const int nb = stage_coords<G>(a, t, n, 0, i, sh.xyz, cx, cy, cz, xi, yi, zi);
    const double sb = stage_sigma<G>(a, t, n, 0, i, nb, sh.sig);
    __syncthreads();

    // lower triangle of A (my_kriging_weights.h:156-165), identity padding
    double bi;
    if (a.cov.table) {
        bi = (i < n) ? cov_tab(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
        for (int j = 0; j <= i; ++j) {
            const double e = cov_tab(a.cov, sh.xyz[0][0][j] - xi, sh.xyz[0][1][j] - yi, sh.xyz[0][2][j] - zi);
            sh.U[i][j] = (i < n && j < n) ? e : ((i == j) ? 1.0 : 0.0);
        }
    } else {
        bi = (i < n) ? cov_eval(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
        for (int j = 0; j <= i; ++j)
            sh.U[i][j] = (i < n && j < n)
                             ? cov_eval(a.cov, sh.xyz[0][0][j] - xi, sh.xyz[0][1][j] - yi, sh.xyz[0][2][j] - zi)
                             : ((i == j) ? 1.0 : 0.0);
    }
    if (a.correlogram) {
        const double si = sh.sig[0][i];
        bi = bi * sb;
        if (i < n)
            for (int j = 0; j <= i; ++j) sh.U[i][j] = sh.U[i][j] * (si * sh.sig[0][j]);
    }
    wave_fence();

    // dpotf2 'U' as rows of L = U^T
    int info = 0;
    double diag = 1.0;
    for (int j = 0; j < n; ++j) {
        double temp = 0.0;
        if (i >= j) {
#pragma unroll 4
            for (int k = 0; k < j; ++k) temp = temp + sh.U[i][k] * sh.U[j][k];
        }
        double ajj = sh.U[i][i] - temp;  // used on lane j only
        const bool fail = i == j && info == 0 && (ajj <= 0.0 || ajj != ajj);
        if (!(ajj > 0.0)) ajj = 1.0;
        const double sq = sqrt(ajj);
        const double rcp = group_bcast<G>(1.0 / sq, j, 0);
        if (info == 0 && __ballot(fail) != 0) info = j + 1;
        if (i == j) {
            sh.U[j][j] = sq;
            diag = sq;
        }
        if (i > j) sh.U[i][j] = rcp * (sh.U[i][j] + (-1.0) * temp);
        wave_fence();
    }

    const bool two = a.ok_system != 0;
    // dpotrs: forward with L, backward with L^T, both right-hand sides; the
    // divisions by L(i, i) go through its reciprocal (div_rn, exact)
    const double dinv = 1.0 / diag;
    double r1 = bi, r2 = two ? ((i < n) ? 1.0 : 0.0) : 0.0;
    for (int k = 0; k < n; ++k) {
        if (i == k) {
            r1 = div_rn(r1, diag, dinv);
            if (two) r2 = div_rn(r2, diag, dinv);
        }
        const double y1 = group_bcast<G>(r1, k, 0);
        const double y2 = group_bcast<G>(r2, k, 0);
        if (i > k) {
            const double lik = sh.U[i][k];
            r1 = r1 - lik * y1;
            r2 = r2 - lik * y2;
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        if (i == k) {
            if (r1 != 0.0) r1 = div_rn(r1, diag, dinv);
            if (r2 != 0.0) r2 = div_rn(r2, diag, dinv);
        }
        const double y1 = group_bcast<G>(r1, k, 0);
        const double y2 = group_bcast<G>(r2, k, 0);
        if (i < k) {
            const double lki = sh.U[k][i];
            if (y1 != 0.0) r1 = r1 - y1 * lki;
            if (y2 != 0.0) r2 = r2 - y2 * lki;
        }
    }
    const double x1 = r1, x2 = r2;

    double w = x1;
    double variance = 0.0;
    sh.y1[0][i] = x1;
    sh.y2[0][i] = x2;
    sh.dterm[0][i] = bi;
    __syncthreads();
    if (two) {
        double s_sk = 0, s_one = 0;
        for (int k = 0; k < n; ++k) {
            s_sk += sh.y1[0][k];
            s_one += sh.y2[0][k];
        }
        const double mu = (s_sk - 1) / s_one;
        w = x1 - mu * x2;
    }
    __syncthreads();
    if (EPI == EPI_SGS) {
        sh.y1[0][i] = w;
        __syncthreads();
        double v = a.cov.c0;
        for (int k = 0; k < n; ++k) v -= sh.y1[0][k] * sh.dterm[0][k];
        if (two && n > 0) v -= sh.y1[0][n - 1];
        variance = v;
        __syncthreads();
    }

    const bool ok = n > 0 && info == 0;
    if (valid && i == 0) a.status[t] = n == 0 ? ST_NO_NEIGHBOURS : (ok ? ST_OK : ST_SINGULAR);
    wave_count(a.counters, 3, valid && i == 0 && n == 0);
    wave_count(a.counters, 1, valid && i == 0 && ok);
    wave_count(a.counters, 2, valid && i == 0 && n > 0 && !ok);
    epilogue<G>(a, valid, t, n, 0, i, nb, ok, w, variance, sh.term, sh.dterm);
}

}  // namespace

void launch_solve_fp64(const SolveArgs &a, int group, hipStream_t s) {
    const int items = a.subset ? a.n_subset : a.T;
    if (items <= 0) return;
    const int g = group <= 32 ? 32 : 64;
    const int ng = 64 / g;
    dim3 grid((items + ng - 1) / ng);
    if (g == 32)
        hipLaunchKernelGGL(k_solve_fp64<32>, grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_solve_fp64_lds, grid, dim3(64), 0, s, a);
}

}  // namespace hpgl_amd
