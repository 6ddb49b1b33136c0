// solve_quad.hip - per-node kriging systems of order n <= 32, four lanes
// per system.
//
// A wavefront holds 16 systems; the four lanes of a quad share one system,
// lane q owning rows i = 4m + q (m = 0 .. N/4-1) of A and of its Cholesky
// factor L, in VGPRs.  Every value a row needs from another row of the
// same system moves inside the quad with one DPP quad_perm (a full-rate
// VALU op, no LDS, no barrier).  N is the bucket order (a multiple of 4);
// rows n..N-1 are padded with the identity.
//
// k_solve_quad_mixed: FP32 right-looking Cholesky, forward and backward
//   substitution, one or two refinement steps whose residual b - A x is
//   formed in FP64 with A re-evaluated from the covariance table.  A system
//   is handed to the FP64 kernel when a pivot is not positive, when the
//   pivot-based condition estimate max(A_kk) / min(pivot) exceeds 1e4 (then
//   FP32 rounding could decide the sign of a pivot the FP64 factorisation
//   sees), or when its solution is certified by neither test below.
//   One step is enough when the error bound of iterative refinement
//   certifies it: each step contracts the error by rho <= c n u kappa(A)
//   (u = 2^-24; Higham, Accuracy and Stability of Numerical Algorithms,
//   2nd ed., sections 10.1 and 12.1), so after the first correction d0,
//   |x - x1| <= rho / (1 - rho) |d0|.  The kernel estimates kappa as
//   ||A||_F ||A^-1||_2 (the second from the FP32 factor, below), takes
//   rho = n u kappa and accepts x1 when rho <= 1/2 and 2 rho |d0| / |x1| <=
//   5e-8.  It does so only where the outputs are categories (indicator
//   kriging's argmax, median IK's threshold): a continuous estimate is stored
//   as float32 from a float sum of w_i (v_i - m_i) terms, and weights that
//   differ from the FP64 path's in their twelfth digit flip the rounding of
//   single terms (on 8.3 M simple-kriging nodes: 150 floats differ from the
//   FP64 path's after one step, 1 after two).  A wavefront with a system
//   that is neither certified nor corrected below 1e-9 of its solution runs
//   a second step for all its systems, and a system is accepted after it
//   when the second correction is at most 1e-9 of the solution.
// k_solve_quad_fp64: the reference's dpotf2 + dpotrs operation order
//   (sequential dot products, reciprocal scaling, division in the
//   triangular solves, no fused multiply-add; the translation unit is
//   compiled with contraction off), so with the same A and b the weights
//   and the success/failure decision are the reference's bit for bit.
#include "counters.h"
#include "divrn.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

__device__ inline int dpp_bcast_i(int v, int src) {
    switch (src) {
        case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false);
        case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false);
        case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false);
        default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false);
    }
}
// value of lane `src` of this lane's quad
__device__ inline float qb(float v, int src) { return __int_as_float(dpp_bcast_i(__float_as_int(v), src)); }
__device__ inline double qb(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_bcast_i((int)(b & 0xffffffffll), src);
    const int hi = dpp_bcast_i((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline int qb(int v, int src) { return dpp_bcast_i(v, src); }

// quad-wide reductions (the quad holds one system)
__device__ inline float qswap1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ inline float qswap2(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
}
__device__ inline double qswap1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double qswap2(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0x4E, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x4E, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double qsum(double v) {
    v += qswap1(v);
    v += qswap2(v);
    return v;
}
__device__ inline double qmax(double v) {
    v = fmax(v, qswap1(v));
    v = fmax(v, qswap2(v));
    return v;
}
// Reduce-scatter within the quad: lane q receives the sum over the quad's
// four lanes of their t[q], in the order (t_q + t_{q^1}) + (t_{q^2} +
// t_{q^3}) (two DPP exchange steps).
__device__ inline double quad_reduce_scatter(double t0, double t1, double t2, double t3) {
    const int q = threadIdx.x & 3;
    const bool odd = q & 1, high = q & 2;
    // step 1, with lane q^1: keep the columns of my parity, q&1 and q&1 + 2
    const double ka = (odd ? t1 : t0) + qswap1(odd ? t0 : t1);
    const double kb = (odd ? t3 : t2) + qswap1(odd ? t2 : t3);
    // step 2, with lane q^2: keep column q
    return (high ? kb : ka) + qswap2(high ? ka : kb);
}

__device__ inline int qmaxi(int v) {
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));
    return v;
}

// Covariance from the box table (every job that reaches these kernels has
// one; the analytic OK model is tabulated on the host over the offsets its
// neighbour pairs can take).  Branch-free: out-of-box offsets read entry 0
// and are replaced by 0, so the unrolled build loops issue their loads
// back to back.
// FOLD: the table's octant (DevCov::fold), an eighth of the lines, so more
// of a wavefront's reads hit the vector L1.  Its coordinates arrive scaled
// by the octant's strides (q_setup), so the entry of the offset between two
// points is the sum of their clamped absolute coordinate differences (an
// offset outside the box clamps into the octant's zero layer): no test, no
// multiply.
// (v_sad_u32 gives |a - b| in one instruction; the compiler forms it only
// from a sum with a third operand)
__device__ inline unsigned absdiff_clamp(unsigned a, unsigned b, unsigned lim) {
    unsigned d;
    asm("v_sad_u32 %0, %1, %2, 0" : "=v"(d) : "v"(a), "v"(b));
    return min(d, lim);
}
// covariance of the offset from point i to point j
template <bool FOLD>
__device__ inline double cov_at(const DevCov &c, int xi, int yi, int zi, int xj, int yj, int zj) {
    if constexpr (FOLD) {
        const unsigned sy = c.tr[0] + 2, sz = sy * (c.tr[1] + 2);
        const unsigned idx = absdiff_clamp(xi, xj, c.tr[0] + 1) + absdiff_clamp(yi, yj, (c.tr[1] + 1) * sy) +
                             absdiff_clamp(zi, zj, (c.tr[2] + 1) * sz);
        return c.fold[idx];
    } else {
        const int dx = xj - xi, dy = yj - yi, dz = zj - zi;
        const bool in = dx >= -c.tr[0] && dx <= c.tr[0] && dy >= -c.tr[1] && dy <= c.tr[1] && dz >= -c.tr[2] &&
                        dz <= c.tr[2];
        const int idx = ((dz + c.tr[2]) * c.td[1] + (dy + c.tr[1])) * c.td[0] + (dx + c.tr[0]);
        const double v = c.table[in ? idx : 0];
        return in ? v : 0.0;
    }
}

constexpr int off(int m) { return 2 * m * (m + 1); }  // start of my m-th row (row length 4m+4)

// Per-wavefront LDS of the quad kernels.  The epilogue's terms reuse the
// refinement's solution vector and the right-hand side, both dead by then
// (the SGS variance reads b before the epilogue, which writes no terms for
// SGS).  The second right-hand side's solution of ordinary kriging has its
// own array in the mixed kernel.
template <int N>
struct QShared {
    int xyz[16][3][N];
    union {
        double x1[16][N];
        float tf[16][N];
    };
    union {
        double b[16][N];  // right-hand side
        double td[16][N];
    };
    float fro2[16][4];  // the quad's sum of squares (mixed kernel)
};

// Common set-up of one quad: target, neighbour count, coordinates in LDS.
template <int N>
struct QCtx {
    int t, n, node, q, slot;
    bool valid;
    int cx, cy, cz;
};

// FOLD: y and z are stored scaled by the octant's strides, tr[0] + 2 and
// (tr[0] + 2)(tr[1] + 2) (cov_at; the launchers check that they fit).
template <int N, bool FOLD>
__device__ inline void q_setup(const SolveArgs &a, QShared<N> &sh, QCtx<N> &c) {
    const int sy = FOLD ? a.cov.tr[0] + 2 : 1, sz = FOLD ? sy * (a.cov.tr[1] + 2) : 1;
    const int lane = threadIdx.x & 63;
    c.q = lane & 3;
    c.slot = lane >> 2;
    const int pos = blockIdx.x * 16 + c.slot;
    const int items = a.subset ? a.n_subset : a.T;
    c.valid = pos < items;
    c.t = c.valid ? (a.subset ? a.subset[pos] : pos) : 0;
    c.n = c.valid ? min(a.cnt[c.t], N) : 0;
    c.node = a.targets[c.t];
    c.cx = c.node % a.X;
    c.cy = c.node % (a.X * a.Y) / a.X * sy;
    c.cz = c.node / (a.X * a.Y) * sz;
#pragma unroll
    for (int m = 0; m < N / 4; ++m) {
        const int i = 4 * m + c.q;
        int x = 0, y = 0, z = 0, cell = -1;
        if (i < c.n) {
            cell = a.nbr[(size_t)c.t * a.max_n + i];
            x = cell % a.X;
            y = cell % (a.X * a.Y) / a.X * sy;
            z = cell / (a.X * a.Y) * sz;
        }
        sh.xyz[c.slot][0][i] = x;
        sh.xyz[c.slot][1][i] = y;
        sh.xyz[c.slot][2][i] = z;
    }
    __syncthreads();
}

// A(i, j) of the padded system.  Branch-free: padding rows hold
// coordinates 0 in LDS, so the table read is always in bounds and the
// result is selected afterwards.
template <int N, bool FOLD>
__device__ inline double q_aij(const DevCov &cv, const QShared<N> &sh, int slot, int n, int i, int j) {
    const double v = cov_at<FOLD>(cv, sh.xyz[slot][0][i], sh.xyz[slot][1][i], sh.xyz[slot][2][i], sh.xyz[slot][0][j],
                                  sh.xyz[slot][1][j], sh.xyz[slot][2][j]);
    return (i < n && j < n) ? v : (i == j ? 1.0 : 0.0);
}

template <int N, bool FOLD>
__device__ inline double q_bi(const SolveArgs &a, const QShared<N> &sh, const QCtx<N> &c, int i) {
    const double v = cov_at<FOLD>(a.cov, sh.xyz[c.slot][0][i], sh.xyz[c.slot][1][i], sh.xyz[c.slot][2][i], c.cx, c.cy,
                                  c.cz);
    return i < c.n ? v : 0.0;
}

// Writes one system's result (combine.h order) from the quad's weights.
// w[m] is the weight of my row 4m+q; `write` false skips the output.
template <int N>
__device__ inline void q_epilogue(const SolveArgs &a, QShared<N> *shp, const QCtx<N> &c, bool write, bool ok,
                                  const double *w, double variance) {
    constexpr int M = N / 4;
    const int EPI = a.epi;
    float m0;
    if (a.mean_kind == 2)
        m0 = a.mean_cells[c.node];
    else if (a.mean_kind == 1)
        m0 = a.mean_single;
    else
        m0 = 0.0f;
    if (EPI == EPI_SGS) {
        // the values of simulated neighbours are not known yet: keep the
        // weights and the variance for the ordered sweep
        if (write) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int i = 4 * m + c.q;
                if (ok && i < c.n) a.out_w[(size_t)c.t * a.max_n + i] = w[m];
            }
            if (c.q == 0) a.out_var[c.t] = ok ? variance : -1.0;
        }
        return;
    }
    float tf[M];
    double td[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int i = 4 * m + c.q;
        tf[m] = 0.0f;
        td[m] = 0.0;
        if (ok && i < c.n) {
            const int nb = a.nbr[(size_t)c.t * a.max_n + i];
            const float mi = a.mean_kind == 2 ? a.mean_cells[nb] : (a.mean_kind == 1 ? a.mean_single : 0.0f);
            float v;
            if (EPI == EPI_CONT)
                v = a.values_f[nb];
            else
                v = (a.values_u8[nb] == (uint8_t)a.category) ? 1.0f : 0.0f;
            if (EPI == EPI_MEDIAN)
                td[m] = (double)(w[m] * (double)(v - mi));
            else
                tf[m] = (float)(w[m] * (double)(v - mi));
        }
    }
    // sequential sum in neighbour order by the quad's first lane
    QShared<N> &sh = *shp;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        sh.tf[c.slot][4 * m + c.q] = tf[m];
        sh.td[c.slot][4 * m + c.q] = td[m];
    }
    __syncthreads();
    float rf = m0;
    double rd = (double)m0;
    if (ok && c.q == 0) {
        for (int i = 0; i < c.n; ++i) {
            if (EPI == EPI_MEDIAN)
                rd += sh.td[c.slot][i];
            else
                rf += sh.tf[c.slot][i];
        }
    }
    if (!write || c.q != 0) return;
    if (EPI == EPI_CONT) {
        if (ok) {
            a.out_f[c.node] = rf;
            a.out_mask[c.node] = 1;
        } else if (a.mean_on_failure) {
            a.out_f[c.node] = m0;
            a.out_mask[c.node] = 1;
        }
    } else if (EPI == EPI_IK) {
        a.out_f[c.t] = rf;
    } else if (EPI == EPI_MEDIAN) {
        a.out_u8[c.node] = rd <= 0.5 ? 0 : 1;
        a.out_mask[c.node] = 1;
    }
}

// ---------------------------------------------------------------------------
// Mixed precision.
// ---------------------------------------------------------------------------
// Dinv[m]: reciprocal of the diagonal entry of my row 4m+q (the FP32 factor
// only seeds the FP64 refinement, so products by reciprocals replace the
// divisions).
template <int N>
__device__ inline void q_solve32(const float *L, const float *Dinv, int q, float *v) {
    constexpr int M = N / 4;
    // forward: L y = v (column sweep, y_k broadcast from its row owner)
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int mk = k >> 2, ko = k & 3;
        if (q == ko) v[mk] = v[mk] * Dinv[mk];
        const float yk = qb(v[mk], ko);
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 > k) {
                if (4 * m + q > k) v[m] = fmaf(-L[off(m) + k], yk, v[m]);
            }
    }
    // backward: L^T x = y (row sweep, partial sums reduced in the quad)
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        const int mi = i >> 2, io = i & 3;
        float p = 0.0f;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 > i) {
                if (4 * m + q > i) p = fmaf(L[off(m) + i], v[m], p);
            }
        p += qswap1(p);
        p += qswap2(p);
        if (q == io) v[mi] = (v[mi] - p) * Dinv[mi];
    }
}

// Estimate of ||A^-1||_2 from the FP32 factor (the idea of LINPACK's
// dpoco): L y = e with e_k = +-1 chosen during the forward sweep so that
// |y_k| grows, then L^T z = y; ||z|| / ||e|| (a lower bound that the sign
// choice makes close: y leans towards the dominant singular direction of
// L^-1, which L^-T then amplifies again).  Rows n..N-1 take e_k = 0.
template <int N>
__device__ inline double q_inv_norm_est(const float *L, const float *Dinv, int q, int n) {
    constexpr int M = N / 4;
    float v[M];
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = 0.0f;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int mk = k >> 2, ko = k & 3;
        if (q == ko) {
            const float s = v[mk];
            const float e = k < n ? (s >= 0.0f ? 1.0f : -1.0f) : 0.0f;
            v[mk] = (s + e) * Dinv[mk];
        }
        const float yk = qb(v[mk], ko);
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 > k) {
                if (4 * m + q > k) v[m] = fmaf(-L[off(m) + k], yk, v[m]);
            }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        const int mi = i >> 2, io = i & 3;
        float p = 0.0f;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 > i) {
                if (4 * m + q > i) p = fmaf(L[off(m) + i], v[m], p);
            }
        p += qswap1(p);
        p += qswap2(p);
        if (q == io) v[mi] = (v[mi] - p) * Dinv[mi];
    }
    double zz = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) zz += (double)v[m] * (double)v[m];
    zz = qsum(zz);
    return n > 0 ? sqrt(zz / (double)n) : 0.0;
}

// Tolerance of the one-step certificate: the bound on |x - x1| / |x1|,
// twenty times below the 1e-6 relative error the mixed path is held to.
constexpr double kOneStepTol = 5.0e-8;

// Wavefronts per SIMD that the register budget of k_solve_quad_mixed<N,
// TWO> must allow.  Left alone, orders 28 and 32 and the two-right-hand-side
// kernels from order 20 take more than 256 registers (one wavefront per
// SIMD); capped at two wavefronts they spill a few hundred bytes per lane
// and still run faster (order 32, one right-hand side: 20.9 -> 15.2 ms per
// SGS realisation; two right-hand sides, 490,000 systems: order 20 1.31 ->
// 0.82 ms, 24 1.82 -> 1.18 ms, 32 3.40 -> 2.79 ms).  The order-24 single
// right-hand-side kernel is held to two as well (at three it spills: 332
// bytes per lane and 123 against 89 ms per headline step with an earlier
// build, 196 bytes and 61.0 against 52.1 ms with the right-hand side in
// LDS).
template <int N, bool TWO>
constexpr int kQuadMixedWaves = (N >= 24 || (N >= 20 && TWO)) ? 2 : 1;

template <int N, bool TWO, bool FOLD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kQuadMixedWaves<N, TWO>, 8))) void k_solve_quad_mixed(SolveArgs a) {
    const int EPI = a.epi;
    constexpr int M = N / 4;
    __shared__ QShared<N> sh;
    __shared__ double sh_x2[TWO ? 16 : 1][N];
    QCtx<N> c;
    q_setup<N, FOLD>(a, sh, c);
    const int q = c.q;

    // build L <- A (FP32), b (FP64, kept in LDS: read back for the first
    // solve, each residual and the SGS variance, not held in registers
    // through the factorisation)
    float L[off(M)];
    double maxdiag = 0.0;
    float low2 = 0.0f;  // sum of squares of my rows' entries j <= i
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int i = 4 * m + q;
        sh.b[c.slot][i] = q_bi<N, FOLD>(a, sh, c, i);
#pragma unroll
        for (int j = 0; j < 4 * m + 4; ++j) {
            // lower triangle of the padded system, selected in FP32: A(i, j)
            // on rows i < n (then j <= i < n too), the identity on padding
            // rows, 0 above the diagonal
            const double e = cov_at<FOLD>(a.cov, sh.xyz[c.slot][0][i], sh.xyz[c.slot][1][i], sh.xyz[c.slot][2][i],
                                          sh.xyz[c.slot][0][j], sh.xyz[c.slot][1][j], sh.xyz[c.slot][2][j]);
            const float l = j > i ? 0.0f : (i < c.n ? (float)e : (i == j ? 1.0f : 0.0f));
            L[off(m) + j] = l;
            low2 = fmaf(l, l, low2);
            if (j == i && i < c.n) maxdiag = fmax(maxdiag, e);
        }
    }
    maxdiag = qmax(maxdiag);
    // the quad's sum, parked in LDS until the first refinement step
    sh.fro2[c.slot][q] = (float)qsum((double)low2);

    // right-looking Cholesky in place (hardware square root and reciprocal,
    // within an ulp: the refinement works from this factor in FP64)
    float minpiv = 3.0e38f;
    float Dinv[M];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int mk = k >> 2, ko = k & 3;
        const float d = qb(L[off(mk) + k], ko);
        if (k < c.n) minpiv = fminf(minpiv, d);
        const float s = __builtin_amdgcn_sqrtf(fmaxf(d, 1.0e-30f));
        const float inv = __builtin_amdgcn_rcpf(s);
        if (q == ko) Dinv[mk] = inv;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 >= k) {
                const int i = 4 * m + q;
                const float v = L[off(m) + k];
                L[off(m) + k] = (i > k) ? v * inv : (i == k ? s : v);
            }
#pragma unroll
        for (int j = k + 1; j < N; ++j) {
            const float cj = qb(L[off(j >> 2) + k], j & 3);
#pragma unroll
            for (int m = 0; m < M; ++m)
                if (4 * m + 3 >= j) L[off(m) + j] = fmaf(-L[off(m) + k], cj, L[off(m) + j]);
        }
    }
    bool bad = !(minpiv > 1.0e-4f * (float)maxdiag);

    constexpr bool two = TWO;
    float v1[M], v2[M];
    double x1[M], x2[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int i = 4 * m + q;
        v1[m] = (float)sh.b[c.slot][i];
        v2[m] = (two && i < c.n) ? 1.0f : 0.0f;
    }
    q_solve32<N>(L, Dinv, q, v1);
    if (two) q_solve32<N>(L, Dinv, q, v2);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        x1[m] = v1[m];
        x2[m] = v2[m];
    }
    // Refinement: a second step runs unless every system of the wave has a
    // first correction below 1e-9 of its solution or, where the outputs are
    // categories, is certified after the first (header comment).
    const bool categorical = EPI == EPI_IK || EPI == EPI_MEDIAN;
    bool conv = false, twice = false;
    for (int step = 0; step < 2; ++step) {
        // the table reads of a step are not kept for the next one (that
        // would hold A in registers): the pointer is opaque per step
        DevCov cv = a.cov;
        asm volatile("" : "+s"(cv.table), "+s"(cv.fold));
        // r = b - A x in FP64 for my rows; x goes through LDS
#pragma unroll
        for (int m = 0; m < M; ++m) {
            sh.x1[c.slot][4 * m + q] = x1[m];
            if (TWO) sh_x2[c.slot][4 * m + q] = x2[m];
        }
        __syncthreads();
        double r1[M], r2[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            r1[m] = sh.b[c.slot][4 * m + q];
            r2[m] = (TWO && 4 * m + q < c.n) ? 1.0 : 0.0;
        }
        // A is symmetric: lane q reads the 4x4 blocks (m, mb), mb <= m, of
        // its rows i = 4m+q only; an off-diagonal block also gives the rows
        // 4mb+cc their terms A(4mb+cc, i) x_i, summed over the quad.  The
        // column-block loop stays rolled (an unrolled one keeps too many
        // values live).  Entries of padding rows and columns are read as
        // the covariance of their zero coordinates, unselected: x is zero
        // there, and the padding rows' residuals are not used.
#pragma unroll 1
        for (int mb = 0; mb < M; ++mb) {
            double xc1[4], xc2[4];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                xc1[cc] = sh.x1[c.slot][4 * mb + cc];
                xc2[cc] = TWO ? sh_x2[c.slot][4 * mb + cc] : 0.0;
            }
            double s1 = 0.0, s2 = 0.0;  // terms of row 4mb+q from the blocks below
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if (m < mb) continue;
                const int i = 4 * m + q;
                double e[4];
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    const int j = 4 * mb + cc;
                    e[cc] = cov_at<FOLD>(cv, sh.xyz[c.slot][0][i], sh.xyz[c.slot][1][i], sh.xyz[c.slot][2][i],
                                         sh.xyz[c.slot][0][j], sh.xyz[c.slot][1][j], sh.xyz[c.slot][2][j]);
                }
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    r1[m] = fma(-e[cc], xc1[cc], r1[m]);
                    if (TWO) r2[m] = fma(-e[cc], xc2[cc], r2[m]);
                }
                if (m > mb) {
                    s1 += quad_reduce_scatter(e[0] * x1[m], e[1] * x1[m], e[2] * x1[m], e[3] * x1[m]);
                    if (TWO) s2 += quad_reduce_scatter(e[0] * x2[m], e[1] * x2[m], e[2] * x2[m], e[3] * x2[m]);
                }
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                r1[m] -= m == mb ? s1 : 0.0;
                if (TWO) r2[m] -= m == mb ? s2 : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < M; ++m) {
            v1[m] = (float)r1[m];
            v2[m] = (float)r2[m];
        }
        q_solve32<N>(L, Dinv, q, v1);
        if (two) q_solve32<N>(L, Dinv, q, v2);
        double dmax = 0.0, xmax = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            // padding rows keep x = 0 (their residuals are not formed)
            const bool row = 4 * m + q < c.n;
            const double d1 = row ? (double)v1[m] : 0.0, d2 = row ? (double)v2[m] : 0.0;
            x1[m] += d1;
            x2[m] += d2;
            dmax = fmax(dmax, fmax(fabs(d1), fabs(d2)));
            xmax = fmax(xmax, fmax(fabs(x1[m]), fabs(x2[m])));
        }
        dmax = qmax(dmax);
        xmax = qmax(xmax);
        const double rel = dmax / (xmax > 0 ? xmax : 1.0);
        if (step == 0) {
            conv = rel <= 1.0e-9;
            if (categorical) {
                // contraction bound of a refinement step, n u ||A||_F
                // ||A^-1||_2, with ||A||_F^2 <= 2 (sum of squares over the
                // lower triangle) less the padding rows' unit diagonal
                const double fro = sqrt(fmax(0.0, 2.0 * ((double)sh.fro2[c.slot][q] - (double)(N - c.n))));
                const double rho = (double)c.n * 0x1p-24 * fro * q_inv_norm_est<N>(L, Dinv, q, c.n);
                conv = conv || (rho <= 0.5 && 2.0 * rho * rel <= kOneStepTol);
            }
            if (__all(conv || !c.valid)) break;
        } else {
            twice = true;
            conv = conv || rel <= 1.0e-9;
        }
    }
    if (!conv) bad = true;

    // weights
    double w[M];
    double variance = 0.0;
    if (two) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + q < c.n) {
                s1 += x1[m];
                s2 += x2[m];
            }
        s1 = qsum(s1);
        s2 = qsum(s2);
        const double mu = (s1 - 1.0) / s2;
#pragma unroll
        for (int m = 0; m < M; ++m) w[m] = x1[m] - mu * x2[m];
    } else {
#pragma unroll
        for (int m = 0; m < M; ++m) w[m] = x1[m];
    }
    if (EPI == EPI_SGS) {
        double sv = 0.0, wl = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int i = 4 * m + q;
            if (i < c.n) sv += w[m] * sh.b[c.slot][i];
            if (i == c.n - 1) wl = w[m];
        }
        variance = a.cov.c0 - qsum(sv);
        if (two) variance -= qsum(wl);  // my_kriging_weights.h:246
    }

    const uint8_t st = c.n == 0 ? ST_NO_NEIGHBOURS : (bad ? ST_FALLBACK : ST_OK);
    if (c.valid && q == 0) a.status[c.t] = st;
    wave_count(a.counters, 3, c.valid && q == 0 && st == ST_NO_NEIGHBOURS);
    wave_count(a.counters, 0, c.valid && q == 0 && st == ST_OK);
    wave_count(a.counters, 4, c.valid && q == 0 && st == ST_OK && twice);
    if (a.fb_list) {
        // one atomic per wavefront, then each handed-over target at its rank
        const bool fb = c.valid && q == 0 && st == ST_FALLBACK;
        const unsigned long long m = __ballot(fb);
        int base = 0;
        if ((threadIdx.x & 63) == 0 && m) base = atomicAdd(a.fb_count, (int)__popcll(m));
        base = __shfl(base, 0, 64);
        if (fb) a.fb_list[base + (int)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull))] = c.t;
    }
    q_epilogue<N>(a, &sh, c, c.valid && (c.n == 0 || !bad), c.n > 0, w, variance);
}

// ---------------------------------------------------------------------------
// FP64 in the reference's operation order.
// ---------------------------------------------------------------------------
template <int N, bool FOLD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(N >= 16 ? 2 : 1, 8))) void k_solve_quad_fp64(SolveArgs a) {
    const int EPI = a.epi;
    constexpr int M = N / 4;
    __shared__ QShared<N> sh;
    QCtx<N> c;
    q_setup<N, FOLD>(a, sh, c);
    const int q = c.q;

    double L[off(M)];  // b is kept in LDS through the factorisation
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int i = 4 * m + q;
        sh.b[c.slot][i] = q_bi<N, FOLD>(a, sh, c, i);
#pragma unroll
        for (int j = 0; j < 4 * m + 4; ++j) {
            const double e = q_aij<N, FOLD>(a.cov, sh, c.slot, c.n, i, j);
            L[off(m) + j] = (j <= i) ? e : 0.0;
        }
    }

    // dpotf2 'U' on U = L^T: for each j, row j of L (entries < j) is
    // broadcast, the pivot is A_jj - sum_k L_jk^2 (sequential), and every
    // row below gets (A_ij + (-1) * sum_k L_ik L_jk) * (1 / L_jj).
    int info = 0;
    double dinv[M];  // 1 / L(i, i) of my rows, for the triangular solves (div_rn)
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int mj = j >> 2, jo = j & 3;
        // every accumulator sums over k = 0..j-1 in order, as ddot/dgemv do
        double dot = 0.0;
        double temp[M];
#pragma unroll
        for (int m = 0; m < M; ++m) temp[m] = 0.0;
#pragma unroll
        for (int k = 0; k < j; ++k) {
            const double rjk = qb(L[off(mj) + k], jo);
            dot = dot + rjk * rjk;
#pragma unroll
            for (int m = 0; m < M; ++m)
                if (4 * m + 3 > j) temp[m] = temp[m] + L[off(m) + k] * rjk;
        }
        const double ajj0 = qb(L[off(mj) + j], jo);
        double ajj = ajj0 - dot;
        if (info == 0 && j < c.n && (ajj <= 0.0 || ajj != ajj)) info = j + 1;
        if (!(ajj > 0.0)) ajj = 1.0;  // failed system or padding: keep the arithmetic finite
        ajj = sqrt(ajj);
        const double rcp = 1.0 / ajj;
        if (q == jo) dinv[mj] = rcp;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 >= j) {
                const int i = 4 * m + q;
                if (i == j) L[off(m) + j] = ajj;
                if (i > j) L[off(m) + j] = rcp * (L[off(m) + j] + (-1.0) * temp[m]);
            }
    }

    const bool two = a.ok_system != 0;
    double y1[M], y2[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int i = 4 * m + q;
        y1[m] = sh.b[c.slot][i];
        y2[m] = (two && i < c.n) ? 1.0 : 0.0;
    }
    // dtrsm('L','U','T','N'): temp = b_i; temp -= U(k,i) b_k for k < i; temp / U(i,i)
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int mk = k >> 2, ko = k & 3;
        if (q == ko) {
            y1[mk] = div_rn(y1[mk], L[off(mk) + k], dinv[mk]);
            if (two) y2[mk] = div_rn(y2[mk], L[off(mk) + k], dinv[mk]);
        }
        const double a1 = qb(y1[mk], ko), a2 = qb(y2[mk], ko);
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (4 * m + 3 > k && 4 * m + q > k) {
                y1[m] = y1[m] - L[off(m) + k] * a1;
                y2[m] = y2[m] - L[off(m) + k] * a2;
            }
    }
    // dtrsm('L','U','N','N'): for k descending, if b_k != 0: b_k /= U(k,k);
    // b_i -= b_k U(i,k) for i < k, with U(i,k) = L(k,i) from row k's owner.
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        const int mk = k >> 2, ko = k & 3;
        int nz1 = 0, nz2 = 0;
        if (q == ko) {
            nz1 = y1[mk] != 0.0;
            nz2 = y2[mk] != 0.0;
            if (nz1) y1[mk] = div_rn(y1[mk], L[off(mk) + k], dinv[mk]);
            if (nz2) y2[mk] = div_rn(y2[mk], L[off(mk) + k], dinv[mk]);
        }
        nz1 = qb(nz1, ko);
        nz2 = qb(nz2, ko);
        const double xk1 = qb(y1[mk], ko), xk2 = qb(y2[mk], ko);
#pragma unroll
        for (int i = 0; i < k; ++i) {
            const double lki = qb(L[off(mk) + i], ko);  // row k, column i
            if ((i & 3) == q) {
                const int m = i >> 2;
                if (nz1) y1[m] = y1[m] - xk1 * lki;
                if (nz2) y2[m] = y2[m] - xk2 * lki;
            }
        }
    }

    // weights and variance, sums in neighbour order (my_kriging_weights.h:217-252)
    double w[M];
    double variance = 0.0;
    {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double t1 = qb(y1[i >> 2], i & 3), t2 = qb(y2[i >> 2], i & 3);
            if (i < c.n) {
                s1 += t1;
                s2 += t2;
            }
        }
        if (two) {
            const double mu = (s1 - 1) / s2;
#pragma unroll
            for (int m = 0; m < M; ++m) w[m] = y1[m] - mu * y2[m];
        } else {
#pragma unroll
            for (int m = 0; m < M; ++m) w[m] = y1[m];
        }
    }
    if (EPI == EPI_SGS) {
        double v = a.cov.c0;
        double wl = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double wi = qb(w[i >> 2], i & 3), bi = sh.b[c.slot][i];
            if (i < c.n) v -= wi * bi;
            if (i == c.n - 1) wl = wi;
        }
        if (two && c.n > 0) v -= wl;
        variance = v;
    }

    const bool ok = c.n > 0 && info == 0;
    if (c.valid && q == 0) a.status[c.t] = c.n == 0 ? ST_NO_NEIGHBOURS : (ok ? ST_OK : ST_SINGULAR);
    wave_count(a.counters, 3, c.valid && q == 0 && c.n == 0);
    wave_count(a.counters, 1, c.valid && q == 0 && ok);
    wave_count(a.counters, 2, c.valid && q == 0 && c.n > 0 && !ok);
    q_epilogue<N>(a, &sh, c, c.valid, ok, w, variance);
}

}  // namespace

// FOLD when the covariance has a symmetric octant table (DevCov::fold)
#define HPGL_Q(K, NN)                                                      \
    if (fold)                                                              \
        hipLaunchKernelGGL((K<NN, true>), grid, dim3(64), 0, s, a);        \
    else                                                                   \
        hipLaunchKernelGGL((K<NN, false>), grid, dim3(64), 0, s, a)
#define HPGL_QM2(NN, F)                                                                    \
    if (a.ok_system)                                                                       \
        hipLaunchKernelGGL((k_solve_quad_mixed<NN, true, F>), grid, dim3(64), 0, s, a);   \
    else                                                                                   \
        hipLaunchKernelGGL((k_solve_quad_mixed<NN, false, F>), grid, dim3(64), 0, s, a)
#define HPGL_QM(NN)            \
    if (fold) {                \
        HPGL_QM2(NN, true);    \
    } else {                   \
        HPGL_QM2(NN, false);   \
    }

void launch_solve_quad(const SolveArgs &a, int order, bool fp64, hipStream_t s) {
    const int items = a.subset ? a.n_subset : a.T;
    if (items <= 0) return;
    dim3 grid((items + 15) / 16);
    // the octant path stores y and z scaled by its strides in 32-bit ints
    const long long sy = a.cov.tr[0] + 2, sz = sy * (a.cov.tr[1] + 2);
    const bool fold = a.cov.fold && (long long)(a.Y - 1) * sy <= INT32_MAX && (long long)(a.Z - 1) * sz <= INT32_MAX;
    if (fp64) {
        switch (order) {
            case 4: HPGL_Q(k_solve_quad_fp64, 4); break;
            case 8: HPGL_Q(k_solve_quad_fp64, 8); break;
            case 12: HPGL_Q(k_solve_quad_fp64, 12); break;
            default: HPGL_Q(k_solve_quad_fp64, 16); break;
        }
    } else {
        switch (order) {
            case 4: HPGL_QM(4); break;
            case 8: HPGL_QM(8); break;
            case 12: HPGL_QM(12); break;
            case 16: HPGL_QM(16); break;
            case 20: HPGL_QM(20); break;
            case 24: HPGL_QM(24); break;
            case 28: HPGL_QM(28); break;
            default: HPGL_QM(32); break;
        }
    }
}

}  // namespace hpgl_amd
