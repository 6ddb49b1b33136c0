// solve.hip - per-node kriging systems in FP64 in the reference's operation
// order: the plain-FP64 path, the fallback of the mixed-precision quad
// kernels (solve_quad.hip) and the path for systems wider than the quad
// kernels take.
//
// k_solve_fp64_bc<NB> (orders up to 8 NB, NB = 4 or 8): one system per
//   wavefront, the factor formed right-looking on an 8 x 8 lane grid with
//   per-element accumulators in registers and the finished columns in LDS
//   as a packed triangle; the forward substitution is fused into the
//   factor steps, the backward substitution runs lane per row.
// It follows dpotf2 + dpotrs (sequential dot products, reciprocal scaling,
// division in the triangular solves, sequential subtraction) with FMA
// contraction off: with the same A and b it returns the reference's weights
// bit for bit, and the same success/failure decision.  Rows n..M-1 are
// padded with the identity.
//
// The combine (combiner.h:10-24) is evaluated exactly as the reference:
// each term (float)(w_i * (double)(v_i - m_i)) is formed per lane, then
// summed in float in neighbour order by the first lane.
#include "counters.h"
#include "devcov.h"
#include "divrn.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

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

// k_solve_fp64_bc<NB> (orders up to M = 8 NB): one system per wavefront.
// The factor is formed right-looking on an 8 x 8 lane grid: lane (r, c)
// (lane = r + 8 c) keeps, for every block pair a >= b, the running dot
// product of element (r + 8a, c + 8b) in a register, and the matrix and the
// finished columns live in LDS as a packed lower triangle (column j holds
// rows j..M-1).  Step k finishes column k on the lanes c == k mod 8, reads
// the column back as broadcasts (rows r + 8a and c + 8b) and adds
// L(i,k) L(j,k) into every trailing accumulator.  Each accumulator starts at
// 0 and receives its products in k order, rounded separately, so element
// (i, j) sees exactly dpotf2's sum: AJJ = A(j,j) - sum, L(i,j) =
// (1 / L(j,j)) (A(i,j) + (-1) sum).  The forward substitution is fused into
// the steps and the backward substitution runs lane per row, both in
// dtrsm's order.  Accumulators of finished columns and of the upper half of
// the diagonal blocks take updates that are never read.
__host__ __device__ constexpr int tri_col(int m, int j) { return m * j - j * (j - 1) / 2; }

template <int NB>
struct ShBC {
    static constexpr int M = 8 * NB;
    static constexpr int TRI = M * (M + 1) / 2;
    union {
        double L[TRI];  // (i, j), i >= j, at tri_col(M, j) + i - j
        struct {        // after the triangular solves
            double y1[1][64];
            double y2[1][64];
            double dterm[1][64];
            float term[1][64];
        } post;
    };
    double sig[1][64];
    int xyz[1][3][64];
};

__device__ inline void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Per-lane state of k_solve_fp64_bc: the accumulators of the lane's
// elements (a, b), a >= b, at a (a + 1) / 2 + b; column k+1 of A at rows
// r + 8a; the right-hand sides (row l on lane l) and the last forward value.
// Forward substitution state: the running right-hand sides (row l on lane
// l, updated on every lane; only rows past the current column matter), the
// current forward values, and the finished values (lane l keeps y_l).
struct BcState_rhs {
    double r1, r2, y1, y2, f1, f2;
};
constexpr int kNoFailure = 1 << 30;
template <int NB>
struct BcState {
    double acc[NB * (NB + 1) / 2];
    double av[NB];
    BcState_rhs rhs;
    int info;
};

// Pivot and column kn (block KBN): L(kn, kn) = sqrt(A - sum) on lane
// (kr, kr), its reciprocal broadcast, the failure flag, and
// L(i, kn) = (1 / L(kn, kn)) (A(i, kn) + (-1) sum) written to LDS by the
// lanes c == kr.  Lanes that own no element write to their own dead slot
// (the sigmas, not read after the matrix is built) instead of branching,
// so the chain can be scheduled together with independent updates.  Rows
// past n hold zeros of the padding and stay zero.  Returns this lane's
// square root (meaningful on the pivot lane) and the broadcast reciprocal.
template <int NB, int KBN>
__device__ __forceinline__ void bc_column(ShBC<NB> &sh, BcState<NB> &s, int kn, int l, int r, int c, double &sq,
                                          double &rcp) {
    constexpr int M = 8 * NB;
    const int kr = kn - 8 * KBN;
    const int col = tri_col(M, kn);
    double ajj = s.av[KBN] - s.acc[KBN * (KBN + 1) / 2 + KBN];
    const bool piv = l == 9 * kr;
    const bool fail = piv && (ajj <= 0.0 || ajj != ajj);
    if (!(ajj > 0.0)) ajj = 1.0;  // failed or not the pivot lane: keep arithmetic finite
    sq = sqrt(ajj);
    rcp = group_bcast<64>(1.0 / sq, 9 * kr, 0);
    // first failing column, kept without a branch (info is 1 + that column)
    s.info = min(s.info, __ballot(fail) != 0 ? kn + 1 : kNoFailure);
    double *const dead = &sh.sig[0][l];
    *(piv ? &sh.L[col] : dead) = sq;
    const bool own = c == kr;
    *((own && r > kr) ? &sh.L[col + r - kr] : dead) =
        rcp * (s.av[KBN] + (-1.0) * s.acc[KBN * (KBN + 1) / 2 + KBN]);
#pragma unroll
    for (int ab = KBN + 1; ab < NB; ++ab)
        *(own ? &sh.L[col + r + 8 * ab - kn] : dead) = rcp * (s.av[ab] + (-1.0) * s.acc[ab * (ab + 1) / 2 + KBN]);
}

// Forward substitution (dtrsm 'L','U','T','N') value of column kn:
// r_kn is final, y_kn = r_kn / L(kn, kn) through the reciprocal (div_rn,
// exact); lane kn keeps it.
__device__ __forceinline__ void bc_forward(BcState_rhs &s, int kn, int kr, int l, double sq, double rcp, bool two) {
    const double lkk = group_bcast<64>(sq, 9 * kr, 0);
    s.y1 = div_rn(group_bcast<64>(s.r1, kn, 0), lkk, rcp);
    s.y2 = two ? div_rn(group_bcast<64>(s.r2, kn, 0), lkk, rcp) : 0.0;
    if (l == kn) {
        s.f1 = s.y1;
        s.f2 = s.y2;
    }
}

// Step k (block KB) with one column of lookahead: read column k, apply it
// to the right-hand sides and to block column KBN (which holds column k+1),
// finish column k+1, apply column k to the remaining live block columns
// (independent of the pivot chain, which they overlap), then form y_{k+1}.
template <int NB, int KB, int KBN>
__device__ __forceinline__ void bc_step(ShBC<NB> &sh, BcState<NB> &s, int k, int l, int r, int c, int n, bool two) {
    constexpr int M = 8 * NB;
    const int col = tri_col(M, k);
    wave_fence();
    // column k at rows r + 8a and c + 8b (rows above k read stale entries
    // of column k-1 and feed only dead accumulators)
    double lr[NB], lc[NB];
#pragma unroll
    for (int ab = KB; ab < NB; ++ab) {
        lr[ab] = sh.L[col + r + 8 * ab - k];
        lc[ab] = sh.L[col + c + 8 * ab - k];
    }
    {
        // rows at or above k (and lanes past M) read other entries and
        // update dead values
        const double lik = sh.L[col + (l < M ? l : M - 1) - k];
        s.rhs.r1 = s.rhs.r1 - lik * s.rhs.y1;
        s.rhs.r2 = s.rhs.r2 - lik * s.rhs.y2;
    }
    if (k + 1 >= n) return;
    const int kn = k + 1;
    const int coln = tri_col(M, kn);
#pragma unroll
    for (int ab = KBN; ab < NB; ++ab) s.av[ab] = sh.L[coln + r + 8 * ab - kn];
#pragma unroll
    for (int ab = KBN; ab < NB; ++ab) {
        const int e = ab * (ab + 1) / 2 + KBN;
        s.acc[e] = s.acc[e] + lr[ab] * lc[KBN];
    }
    double sq, rcp;
    bc_column<NB, KBN>(sh, s, kn, l, r, c, sq, rcp);
    // block column KB is finished once KBN moves past it
#pragma unroll
    for (int bb = KB; bb < NB; ++bb) {
        if (bb <= KBN) continue;
#pragma unroll
        for (int ab = bb; ab < NB; ++ab) {
            const int e = ab * (ab + 1) / 2 + bb;
            s.acc[e] = s.acc[e] + lr[ab] * lc[bb];
        }
    }
    bc_forward(s.rhs, kn, kn - 8 * KBN, l, sq, rcp, two);
}

template <int NB, int KB>
struct BcBlock {
    __device__ __forceinline__ static void run(ShBC<NB> &sh, BcState<NB> &s, int l, int r, int c, int n, bool two) {
        if (8 * KB >= n) return;
        const int kend = min(8, n - 8 * KB);
        const int kmain = min(7, kend);
        for (int kr = 0; kr < kmain; ++kr) bc_step<NB, KB, KB>(sh, s, 8 * KB + kr, l, r, c, n, two);
        if (kend == 8) bc_step<NB, KB, (KB + 1 < NB ? KB + 1 : KB)>(sh, s, 8 * KB + 7, l, r, c, n, two);
        if constexpr (KB + 1 < NB) BcBlock<NB, KB + 1>::run(sh, s, l, r, c, n, two);
    }
};

// FOLD: table reads from the octant (DevCov::fold), an eighth of the lines.
template <int NB, bool FOLD>
__global__ __launch_bounds__(64) void k_solve_fp64_bc(SolveArgs a) {
    constexpr int G = 64;
    constexpr int M = 8 * NB;
    const int EPI = a.epi;
    __shared__ ShBC<NB> sh;
    const int l = threadIdx.x;
    const int r = l & 7, c = l >> 3;
    const int pos = blockIdx.x;
    const int n_items = a.subset ? a.n_subset : a.T;
    const bool valid = pos < n_items;
    const int t = valid ? (a.subset ? a.subset[pos] : pos) : 0;
    const int n = valid ? min(a.cnt[t], M) : 0;

    int cx, cy, cz, xi, yi, zi;
    const int nb = stage_coords<G>(a, t, n, 0, l, sh.xyz, cx, cy, cz, xi, yi, zi);
    const double sb = stage_sigma<G>(a, t, n, 0, l, nb, sh.sig);
    __syncthreads();

    // b (my_kriging_weights.h:156-165) on lane i = l
    double bi;
    if (FOLD)
        bi = (l < n) ? cov_fold(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
    else if (a.cov.table)
        bi = (l < n) ? cov_tab(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
    else
        bi = (l < n) ? cov_eval(a.cov, cx - xi, cy - yi, cz - zi) : 0.0;
    if (a.correlogram) bi = bi * sb;

    // lower triangle of A, A(i, j) = C(x_j - x_i), identity padding; lane
    // (r, c) writes the elements it will own.  Table reads are issued all
    // at once (the table is larger than L2); the analytic model is
    // evaluated in a loop.
    if (a.cov.table) {
        int xr[3][NB], xc[3][NB];
#pragma unroll
        for (int ab = 0; ab < NB; ++ab)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                xr[d][ab] = sh.xyz[0][d][r + 8 * ab];
                xc[d][ab] = sh.xyz[0][d][c + 8 * ab];
            }
        double e[NB * (NB + 1) / 2];
#pragma unroll
        for (int ab = 0; ab < NB; ++ab)
#pragma unroll
            for (int bb = 0; bb <= ab; ++bb) {
                const int i = r + 8 * ab, j = c + 8 * bb;
                const int q = ab * (ab + 1) / 2 + bb;
                if (i >= j && i < n && j < n)
                    e[q] = FOLD ? cov_fold(a.cov, xc[0][bb] - xr[0][ab], xc[1][bb] - xr[1][ab], xc[2][bb] - xr[2][ab])
                                : cov_tab(a.cov, xc[0][bb] - xr[0][ab], xc[1][bb] - xr[1][ab], xc[2][bb] - xr[2][ab]);
                else
                    e[q] = (i == j) ? 1.0 : 0.0;
            }
#pragma unroll
        for (int ab = 0; ab < NB; ++ab)
#pragma unroll
            for (int bb = 0; bb <= ab; ++bb) {
                const int i = r + 8 * ab, j = c + 8 * bb;
                if (i >= j) {
                    double v = e[ab * (ab + 1) / 2 + bb];
                    if (a.correlogram && i < n && j < n) v = v * (sh.sig[0][i] * sh.sig[0][j]);
                    sh.L[tri_col(M, j) + i - j] = v;
                }
            }
    } else {
#pragma unroll 1
        for (int ab = 0; ab < NB; ++ab) {
#pragma unroll 1
            for (int bb = 0; bb <= ab; ++bb) {
                const int i = r + 8 * ab, j = c + 8 * bb;
                if (i >= j) {
                    double v;
                    if (i < n && j < n) {
                        const int dx = sh.xyz[0][0][j] - sh.xyz[0][0][i];
                        const int dy = sh.xyz[0][1][j] - sh.xyz[0][1][i];
                        const int dz = sh.xyz[0][2][j] - sh.xyz[0][2][i];
                        v = cov_eval(a.cov, dx, dy, dz);
                        if (a.correlogram) v = v * (sh.sig[0][i] * sh.sig[0][j]);
                    } else {
                        v = (i == j) ? 1.0 : 0.0;
                    }
                    sh.L[tri_col(M, j) + i - j] = v;
                }
            }
        }
    }
    wave_fence();

    // dpotf2 'U' as columns of L = U^T with the forward substitution
    // fused in; column 0 first, then steps 0..n-1
    const bool two = a.ok_system != 0;
    BcState<NB> s;
#pragma unroll
    for (int e = 0; e < NB * (NB + 1) / 2; ++e) s.acc[e] = 0.0;
#pragma unroll
    for (int ab = 0; ab < NB; ++ab) s.av[ab] = sh.L[r + 8 * ab];
    s.info = kNoFailure;
    s.rhs.r1 = bi;
    s.rhs.r2 = two ? ((l < n) ? 1.0 : 0.0) : 0.0;
    s.rhs.y1 = s.rhs.y2 = 0.0;
    s.rhs.f1 = s.rhs.r1;
    s.rhs.f2 = s.rhs.r2;
    if (n > 0) {
        double sq, rcp;
        bc_column<NB, 0>(sh, s, 0, l, r, c, sq, rcp);
        bc_forward(s.rhs, 0, 0, l, sq, rcp, two);
        BcBlock<NB, 0>::run(sh, s, l, r, c, n, two);
    }
    wave_fence();
    const int info = s.info == kNoFailure ? 0 : s.info;
    double r1 = s.rhs.f1, r2 = s.rhs.f2;

    // backward substitution (dtrsm 'L','U','N','N') with L^T, both
    // right-hand sides; L(k, i) of the next step is read one step ahead
    const int i = l;
    const double diag = (i < n) ? sh.L[tri_col(M, i)] : 1.0;
    const double dinv = 1.0 / diag;
    double lki_next = (i < n - 1) ? sh.L[tri_col(M, i) + (n - 1) - i] : 0.0;
    for (int k = n - 1; k >= 0; --k) {
        const double lki = lki_next;
        if (i < k - 1) lki_next = sh.L[tri_col(M, i) + (k - 1) - i];
        if (i == k) {
            if (r1 != 0.0) r1 = div_rn(r1, diag, dinv);
            if (r2 != 0.0) r2 = div_rn(r2, diag, dinv);
        }
        const double y1 = group_bcast<G>(r1, k, 0);
        const double y2 = group_bcast<G>(r2, k, 0);
        if (i < k) {
            if (y1 != 0.0) r1 = r1 - y1 * lki;
            if (y2 != 0.0) r2 = r2 - y2 * lki;
        }
    }
    const double x1 = r1, x2 = r2;
    wave_fence();  // the sums below reuse the factor's LDS

    double w = x1;
    double variance = 0.0;
    sh.post.y1[0][i] = x1;
    sh.post.y2[0][i] = x2;
    sh.post.dterm[0][i] = bi;
    __syncthreads();
    if (two) {
        double s_sk = 0, s_one = 0;
        for (int k = 0; k < n; ++k) {
            s_sk += sh.post.y1[0][k];
            s_one += sh.post.y2[0][k];
        }
        const double mu = (s_sk - 1) / s_one;
        w = x1 - mu * x2;
    }
    __syncthreads();
    if (EPI == EPI_SGS) {
        sh.post.y1[0][i] = w;
        __syncthreads();
        double v = a.cov.c0;
        for (int k = 0; k < n; ++k) v -= sh.post.y1[0][k] * sh.post.dterm[0][k];
        if (two && n > 0) v -= sh.post.y1[0][n - 1];
        variance = v;
        __syncthreads();
    }

    const bool ok = n > 0 && info == 0;
    if (valid && i == 0) a.status[t] = n == 0 ? ST_NO_NEIGHBOURS : (ok ? ST_OK : ST_SINGULAR);
    wave_count(a.counters, 3, valid && i == 0 && n == 0);
    wave_count(a.counters, 1, valid && i == 0 && ok);
    wave_count(a.counters, 2, valid && i == 0 && n > 0 && !ok);
    epilogue<G>(a, valid, t, n, 0, i, nb, ok, w, variance, sh.post.term, sh.post.dterm);
}

}  // namespace

void launch_solve_fp64(const SolveArgs &a, int group, hipStream_t s) {
    const int items = a.subset ? a.n_subset : a.T;
    if (items <= 0) return;
    // one system per wavefront
    const dim3 grid(items);
    const bool fold = a.cov.table && a.cov.fold;
    if (group <= 32) {
        if (fold)
            hipLaunchKernelGGL((k_solve_fp64_bc<4, true>), grid, dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_solve_fp64_bc<4, false>), grid, dim3(64), 0, s, a);
    } else {
        if (fold)
            hipLaunchKernelGGL((k_solve_fp64_bc<8, true>), grid, dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_solve_fp64_bc<8, false>), grid, dim3(64), 0, s, a);
    }
}

}  // namespace hpgl_amd
