// solve.hip - per-node kriging systems on CDNA4: build, factor, solve,
// combine, in one kernel per precision.
//
// Mapping: one wavefront (64 lanes, one workgroup) holds 64/G systems of
// order n <= G; lane i of a group owns row i of A (and of L) in VGPRs.
// Rows n..G-1 are padded with the identity so every group runs the same
// unrolled code.  Cross-lane traffic (pivot column, right-hand sides,
// transposed factor) goes through LDS; a workgroup is a single wavefront,
// so __syncthreads() is only an LDS ordering point.
//
// k_solve_mixed: FP32 Cholesky (right-looking, explicit fmaf), two
//   refinement steps whose residual b - A x is formed in FP64 from the FP64
//   row of A kept in registers.  A system goes to the FP64 kernel when the
//   FP32 factorisation breaks down, when the pivot-based condition estimate
//   max(A_kk) / min(pivot) exceeds 1e4 (then FP32 rounding could decide
//   the sign of a pivot the FP64 factorisation sees), or when the second
//   correction is larger than 1e-9 of the solution (refinement did not
//   converge to FP64 accuracy).
// k_solve_fp64: the reference's dpotf2 + dpotrs operation order
//   (left-looking sequential dot products, reciprocal scaling, division in
//   the triangular solves, sequential subtraction), compiled with FMA
//   contraction off: with the same A and b it returns the reference's
//   weights bit for bit, and the same success/failure decision.
//
// The combine (combiner.h:10-24) is evaluated exactly as the reference:
// each term (float)(w_i * (double)(v_i - m_i)) is formed per lane, then
// summed in float in neighbour order by the group's first lane.
#include "counters.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

__device__ inline double cov_eval(const DevCov &c, int dx, int dy, int dz) {
    if (c.table) {
        if (dx < -c.tr[0] || dx > c.tr[0] || dy < -c.tr[1] || dy > c.tr[1] || dz < -c.tr[2] || dz > c.tr[2])
            return 0.0;
        return c.table[((size_t)(dz + c.tr[2]) * c.td[1] + (dy + c.tr[1])) * c.td[0] + (dx + c.tr[0])];
    }
    // cov_model.h:63-77 transfrom_and_norm then :144-181
    const double v0 = dx, v1 = dy, v2 = dz;
    double r = 0.0;
    {
        double d = 0.0;
        d += v0 * c.t[0];
        d += v1 * c.t[1];
        d += v2 * c.t[2];
        r += d * d;
    }
    {
        double d = 0.0;
        d += v0 * c.t[3];
        d += v1 * c.t[4];
        d += v2 * c.t[5];
        r += d * d;
    }
    {
        double d = 0.0;
        d += v0 * c.t[6];
        d += v1 * c.t[7];
        d += v2 * c.t[8];
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
}

__device__ inline void cell_xyz(int cell, int X, int Y, int &x, int &y, int &z) {
    x = cell % X;
    y = cell % (X * Y) / X;
    z = cell / (X * Y);
}

template <int G>
struct Shared {
    static constexpr int NG = 64 / G;
    float col[NG][G + 1];
    float Lt[NG][G * (G + 1)];
    double xb[NG][G];
    float y1[NG][G];
    float y2[NG][G];
    float term[NG][G];
    double dterm[NG][G];
    int xyz[NG][3][G];
    double red[NG][4];
};

template <int G>
struct Shared64 {
    static constexpr int NG = 64 / G;
    double row[NG][G];
    double Ut[NG][G * (G + 1)];
    double y1[NG][G];
    double y2[NG][G];
    int nz1[NG][G];
    int nz2[NG][G];
    double piv[NG][2];
    int fail[NG];
    float term[NG][G];
    double dterm[NG][G];
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
// Mixed precision kernel.
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(64) void k_solve_mixed(SolveArgs a) {
    const int EPI = a.epi;
    constexpr int NG = 64 / G;
    __shared__ Shared<G> sh;
    const int lane = threadIdx.x;
    const int g = lane / G, i = lane % G;
    const int pos = blockIdx.x * NG + g;
    const int n_items = a.subset ? a.n_subset : a.T;
    const bool valid = pos < n_items;
    const int t = valid ? (a.subset ? a.subset[pos] : pos) : 0;
    const int n = valid ? min(a.cnt[t], G) : 0;

    int cx, cy, cz, xi, yi, zi;
    const int nb = stage_coords<G>(a, t, n, g, i, sh.xyz, cx, cy, cz, xi, yi, zi);
    __syncthreads();

    // Row i of A (FP64, kept for the residual) and b_i; identity padding.
    double arow[G];
    const double bi = (i < n) ? cov_eval(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (i < n && j < n)
            arow[j] = cov_eval(a.cov, sh.xyz[g][0][j] - xi, sh.xyz[g][1][j] - yi, sh.xyz[g][2][j] - zi);
        else
            arow[j] = (i == j) ? 1.0 : 0.0;
    }

    // n of the largest system in the wave bounds the unrolled loops.
    int nmax = n;
#pragma unroll
    for (int o = 32; o >= G; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));

    // FP32 right-looking Cholesky, L overwrites l[] (row i).
    float l[G];
#pragma unroll
    for (int j = 0; j < G; ++j) l[j] = (float)arow[j];
    const double maxdiag = group_max<G>(i < n ? arow[i] : 0.0);
    float minpiv = 3.0e38f;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        if (k < nmax) {
            sh.col[g][i] = l[k];
            __syncthreads();
            const float d = sh.col[g][k];
            if (k < n) minpiv = fminf(minpiv, d);
            const float s = sqrtf(fmaxf(d, 1.0e-30f));
            const float inv = 1.0f / s;
            const float lik = (i > k) ? l[k] * inv : 0.0f;
            if (i == k) l[k] = s;
            if (i > k) l[k] = lik;
#pragma unroll
            for (int j = k + 1; j < G; ++j) l[j] = fmaf(-lik, sh.col[g][j] * inv, l[j]);
            __syncthreads();
        }
    }
    // minpiv is group-uniform (each group read the same pivots).
    bool bad = !(minpiv > 1.0e-4f * (float)maxdiag);

    // Transposed factor in LDS: Lt[k*(G+1)+i] = L[k][i] (padded stride, no bank conflicts).
#pragma unroll
    for (int j = 0; j < G; ++j)
        if (j <= i) sh.Lt[g][i * (G + 1) + j] = l[j];
    __syncthreads();

    // Two right-hand sides at once: b and (OK only) the vector of ones.
    const bool two = a.ok_system != 0;
    auto solve2 = [&](float r1, float r2, float &o1, float &o2) {
        // forward: L y = r
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k < nmax) {
                if (i == k) {
                    r1 = r1 / l[k];
                    r2 = r2 / l[k];
                    sh.y1[g][k] = r1;
                    sh.y2[g][k] = r2;
                }
                __syncthreads();
                if (i > k) {
                    r1 = fmaf(-l[k], sh.y1[g][k], r1);
                    r2 = fmaf(-l[k], sh.y2[g][k], r2);
                }
            }
        }
        __syncthreads();
        // backward: L^T x = y
#pragma unroll
        for (int k = G - 1; k >= 0; --k) {
            if (k < nmax) {
                if (i == k) {
                    r1 = r1 / l[k];
                    r2 = r2 / l[k];
                    sh.y1[g][k] = r1;
                    sh.y2[g][k] = r2;
                }
                __syncthreads();
                if (i < k) {
                    const float lki = sh.Lt[g][k * (G + 1) + i];
                    r1 = fmaf(-lki, sh.y1[g][k], r1);
                    r2 = fmaf(-lki, sh.y2[g][k], r2);
                }
            }
        }
        __syncthreads();
        o1 = r1;
        o2 = r2;
    };

    const double ones_i = (i < n) ? 1.0 : 0.0;
    float f1, f2;
    solve2((float)bi, two ? (float)ones_i : 0.0f, f1, f2);
    double x1 = f1, x2 = f2;
    double rel = 0.0;
    for (int step = 0; step < 2; ++step) {
        // residual in FP64
        sh.xb[g][i] = x1;
        __syncthreads();
        double r1 = bi;
#pragma unroll
        for (int j = 0; j < G; ++j) r1 = fma(-arow[j], sh.xb[g][j], r1);
        __syncthreads();
        double r2 = 0.0;
        if (two) {
            sh.xb[g][i] = x2;
            __syncthreads();
            r2 = ones_i;
#pragma unroll
            for (int j = 0; j < G; ++j) r2 = fma(-arow[j], sh.xb[g][j], r2);
            __syncthreads();
        }
        float d1, d2;
        solve2((float)r1, (float)r2, d1, d2);
        x1 += (double)d1;
        x2 += (double)d2;
        if (step == 1) {
            const double num = group_max<G>(fmax(fabs((double)d1), fabs((double)d2)));
            const double den = group_max<G>(fmax(fabs(x1), fabs(x2)));
            rel = num / (den > 0 ? den : 1.0);
        }
    }
    if (!(rel <= 1.0e-9)) bad = true;

    // Weights.  OK: w = x - mu * y with mu = (sum x - 1) / sum y.
    double w = x1;
    double variance = 0.0;
    if (two) {
        const double sx = group_sum<G>(i < n ? x1 : 0.0);
        const double sy = group_sum<G>(i < n ? x2 : 0.0);
        const double mu = (sx - 1.0) / sy;
        w = x1 - mu * x2;
    }
    if (EPI == EPI_SGS) {
        // variance = C(0) - sum w_i b_i (- w_{n-1} for OK, my_kriging_weights.h:246)
        variance = a.cov.c0 - group_sum<G>(i < n ? w * bi : 0.0);
        if (two) {
            sh.xb[g][i] = w;
            __syncthreads();
            variance -= sh.xb[g][n > 0 ? n - 1 : 0];
            __syncthreads();
        }
    }

    const uint8_t st = n == 0 ? ST_NO_NEIGHBOURS : (bad ? ST_FALLBACK : ST_OK);
    if (valid && i == 0) a.status[t] = st;
    wave_count(a.counters, 3, valid && i == 0 && st == ST_NO_NEIGHBOURS);
    wave_count(a.counters, 0, valid && i == 0 && st == ST_OK);
    epilogue<G>(a, valid && (n == 0 || !bad), t, n, g, i, nb, n > 0, w, variance, sh.term, sh.dterm);
}

// ---------------------------------------------------------------------------
// FP64 kernel in the reference's operation order (contraction is off for
// the whole translation unit; every product and sum rounds separately).
// ---------------------------------------------------------------------------
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
    __syncthreads();

    // A(i,j) = C(x_i, x_j) (my_kriging_weights.h:156-165), identity padding
    double u[G];
    const double bi = (i < n) ? cov_eval(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (i < n && j < n)
            u[j] = cov_eval(a.cov, sh.xyz[g][0][j] - xi, sh.xyz[g][1][j] - yi, sh.xyz[g][2][j] - zi);
        else
            u[j] = (i == j) ? 1.0 : 0.0;
    }
    int nmax = n;
#pragma unroll
    for (int o = 32; o >= G; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));

    // dpotf2 'U' as rows of L = U^T: lane i holds L(i, 0..i).
    int info = 0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (j < nmax) {
            if (i == j) {
#pragma unroll
                for (int k = 0; k < j; ++k) sh.row[g][k] = u[k];
            }
            __syncthreads();
            if (i == j) {
                double dot = 0.0;
#pragma unroll
                for (int k = 0; k < j; ++k) dot = dot + u[k] * u[k];
                double ajj = u[j] - dot;
                if (info == 0 && j < n && (ajj <= 0.0 || ajj != ajj)) {
                    sh.fail[g] = j + 1;
                } else {
                    if (!(ajj > 0.0)) ajj = 1.0;  // failed group or padding: keep arithmetic finite
                    ajj = sqrt(ajj);
                    u[j] = ajj;
                    sh.piv[g][0] = 1.0 / ajj;
                    sh.fail[g] = 0;
                }
            }
            __syncthreads();
            if (info == 0 && sh.fail[g] != 0) info = sh.fail[g];
            const double rcp = sh.piv[g][0];
            if (i > j) {
                double temp = 0.0;
#pragma unroll
                for (int k = 0; k < j; ++k) temp = temp + u[k] * sh.row[g][k];
                u[j] = rcp * (u[j] + (-1.0) * temp);
            }
            __syncthreads();
        }
    }

    // U(i,k) = L(k,i) for the back substitution, through LDS.
#pragma unroll
    for (int j = 0; j < G; ++j)
        if (j <= i) sh.Ut[g][i * (G + 1) + j] = u[j];
    __syncthreads();
    double ut[G];
#pragma unroll
    for (int k = 0; k < G; ++k) ut[k] = (k >= i) ? sh.Ut[g][k * (G + 1) + i] : 0.0;

    const bool two = a.ok_system != 0;
    // dpotrs: dtrsm('L','U','T','N') then dtrsm('L','U','N','N')
    auto dpotrs = [&](double r1, double r2, double &o1, double &o2) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (k < nmax) {
                if (i == k) {
                    r1 = r1 / u[k];
                    r2 = r2 / u[k];
                    sh.y1[g][k] = r1;
                    sh.y2[g][k] = r2;
                }
                __syncthreads();
                if (i > k) {
                    r1 = r1 - u[k] * sh.y1[g][k];
                    r2 = r2 - u[k] * sh.y2[g][k];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = G - 1; k >= 0; --k) {
            if (k < nmax) {
                if (i == k) {
                    sh.nz1[g][k] = r1 != 0.0;
                    sh.nz2[g][k] = r2 != 0.0;
                    if (r1 != 0.0) r1 = r1 / u[k];
                    if (r2 != 0.0) r2 = r2 / u[k];
                    sh.y1[g][k] = r1;
                    sh.y2[g][k] = r2;
                }
                __syncthreads();
                if (i < k) {
                    if (sh.nz1[g][k]) r1 = r1 - sh.y1[g][k] * ut[k];
                    if (sh.nz2[g][k]) r2 = r2 - sh.y2[g][k] * ut[k];
                }
            }
        }
        __syncthreads();
        o1 = r1;
        o2 = r2;
    };
    double x1, x2;
    dpotrs(bi, two ? ((i < n) ? 1.0 : 0.0) : 0.0, x1, x2);

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

}  // namespace

void launch_solve_mixed(const SolveArgs &a, int group, hipStream_t s) {
    const int items = a.subset ? a.n_subset : a.T;
    if (items <= 0) return;
    const int g = group <= 32 ? 32 : 64;
    const int ng = 64 / g;
    dim3 grid((items + ng - 1) / ng);
    if (g == 32)
        hipLaunchKernelGGL(k_solve_mixed<32>, grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_solve_mixed<64>, grid, dim3(64), 0, s, a);
}

void launch_solve_fp64(const SolveArgs &a, int group, hipStream_t s) {
    const int items = a.subset ? a.n_subset : a.T;
    if (items <= 0) return;
    const int g = group <= 32 ? 32 : 64;
    const int ng = 64 / g;
    dim3 grid((items + ng - 1) / ng);
    if (g == 32)
        hipLaunchKernelGGL(k_solve_fp64<32>, grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_solve_fp64<64>, grid, dim3(64), 0, s, a);
}

}  // namespace hpgl_amd
