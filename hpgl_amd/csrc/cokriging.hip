// cokriging.hip - simple colocated cokriging, Markov models I and II
// (simple_cokriging_markI.cpp:19-305), one system per wavefront.
//
// The system of n neighbours and the colocated secondary value
// (build_system, :19-55) is (n+1) x (n+1):
//   A(i,j) = C(x_i, x_j), A(i,n) = A(n,i) = k C_x(x_0, x_i), A(n,n) = var2,
//   b(i) = C(x_i, x_0), b(n) = k C_x(0),
// with C the primary covariance, C_x the primary (model I) or secondary
// (model II) covariance and k the cross-covariance factor.  It is solved by
// the reference's gauss_solve (gauss_solver.cpp:5-65) in its operation
// order: for each column the first unused row with a non-zero entry is the
// pivot, the pivot row is divided by it and every unused row has it
// subtracted; back substitution starts each unknown X[i] from B[i] (the
// right-hand side of row i, not of the pivot row of column i: the two agree
// whenever the pivots come in row order, as for covariance matrices) and
// subtracts A[pivot row][j] X[j] for j from n down to i+1.  Rows live in LDS
// sized for max_n + 1 unknowns (stride max_n + 2 doubles), so small systems
// leave room for many wavefronts per CU; lane i owns rows i and i + 64.
// Contraction is off, so with the same covariances the weights are the
// reference's bit for bit.  combine (:82-104) then rounds to float after
// each double addition.
#include "counters.h"
#include "devcov.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

// The wavefront's LDS for systems of up to S = max_n + 1 unknowns.
struct CoLds {
    double *A;  // [S][S + 1]
    double *B;  // [S]
    double *X;  // [S]
    int *order;  // [S]
    int *used;   // [S]
    int *xyz;    // [3][max_n]
    int stride;
    __device__ CoLds(double *base, int max_n) {
        const int S = max_n + 1;
        stride = S + 1;
        A = base;
        B = A + (size_t)S * stride;
        X = B + S;
        order = reinterpret_cast<int *>(X + S);
        used = order + S;
        xyz = used + S;
    }
    __device__ double &a(int r, int c) { return A[r * stride + c]; }
};

__device__ inline double co_cov(const DevCov &c, int dx, int dy, int dz) {
    return c.table ? cov_tab(c, dx, dy, dz) : cov_eval(c, dx, dy, dz);
}

__global__ __launch_bounds__(64) void k_cokriging(CokrigingArgs a) {
    extern __shared__ double co_lds[];
    CoLds sh(co_lds, a.max_n);
    const int mn = a.max_n;
    const int lane = threadIdx.x;
    const int t = blockIdx.x;
    const int n = min(a.cnt[t], a.max_n);
    const int size = n + 1;
    const int node = a.targets[t];
    int cx, cy, cz;
    cell_xyz(node, a.X, a.Y, cx, cy, cz);
    int xi = 0, yi = 0, zi = 0, nb = -1;
    if (lane < n) {
        nb = a.nbr[(size_t)t * a.max_n + lane];
        cell_xyz(nb, a.X, a.Y, xi, yi, zi);
        sh.xyz[lane] = xi;
        sh.xyz[mn + lane] = yi;
        sh.xyz[2 * mn + lane] = zi;
    }
    __syncthreads();

    // build_system (simple_cokriging_markI.cpp:19-55)
    if (lane < n) {
        for (int j = 0; j < n; ++j)
            sh.a(lane, j) = co_cov(a.cov, sh.xyz[j] - xi, sh.xyz[mn + j] - yi, sh.xyz[2 * mn + j] - zi);
        const double cr = a.coef * co_cov(a.cross, xi - cx, yi - cy, zi - cz);
        sh.a(lane, n) = cr;
        sh.a(n, lane) = cr;
        sh.B[lane] = co_cov(a.cov, cx - xi, cy - yi, cz - zi);
    }
    if (lane == 0) {
        sh.B[n] = a.cross0;
        sh.a(n, n) = a.sec_var;
    }
    for (int r = lane; r < size; r += 64) sh.used[r] = 0;
    __syncthreads();

    // gauss_solve (gauss_solver.cpp:5-52)
    bool ok = n > 0;
    for (int i = 0; ok && i < size; ++i) {
        const uint64_t lo = __ballot(lane < size && sh.used[lane] == 0 && sh.a(lane, i) != 0.0);
        int row = lo ? __ffsll((unsigned long long)lo) - 1 : -1;
        if (row < 0 && size > 64 && sh.used[64] == 0 && sh.a(64, i) != 0.0) row = 64;
        if (row < 0) {
            ok = false;  // the matrix is singular
            break;
        }
        const double coef = sh.a(row, i);
        __syncthreads();
        if (lane == 0) {
            sh.used[row] = 1;
            sh.order[i] = row;
            sh.B[row] = sh.B[row] / coef;
        }
        for (int j = i + lane; j < size; j += 64) sh.a(row, j) = sh.a(row, j) / coef;
        __syncthreads();
        for (int r = lane; r < size; r += 64) {
            if (sh.used[r] != 0) continue;
            const double c = sh.a(r, i);
            for (int k = i; k < size; ++k) sh.a(r, k) = sh.a(r, k) - c * sh.a(row, k);
            sh.B[r] = sh.B[r] - c * sh.B[row];
        }
        __syncthreads();
    }

    // back substitution (gauss_solver.cpp:54-62): X[i] = B[i] - sum over
    // j = n .. i+1 of A[order[i]][j] X[j], in that order
    if (ok) {
        double acc0 = lane < size ? sh.B[lane] : 0.0;
        double acc1 = lane + 64 < size ? sh.B[lane + 64] : 0.0;
        const int row0 = lane < size ? sh.order[lane] : 0;
        const int row1 = lane + 64 < size ? sh.order[lane + 64] : 0;
        for (int j = size - 1; j >= 0; --j) {
            if (lane == (j & 63)) sh.X[j] = j < 64 ? acc0 : acc1;
            __syncthreads();
            const double xj = sh.X[j];
            if (lane < j) acc0 = acc0 - sh.a(row0, j) * xj;
            if (lane + 64 < j) acc1 = acc1 - sh.a(row1, j) * xj;
        }
    }

    wave_count(a.counters, 1, lane == 0 && ok);
    wave_count(a.counters, 2, lane == 0 && n > 0 && !ok);
    wave_count(a.counters, 3, lane == 0 && n == 0);
    if (lane != 0) return;
    // combine (simple_cokriging_markI.cpp:82-104); the failure value :235
    const float pm = a.primary_mean, sm = a.secondary_mean;
    const float sv = a.sec_mask[node] != 0 ? a.sec[node] : sm;
    float result;
    if (ok) {
        result = pm;
        for (int i = 0; i < n; ++i) {
            const int c = a.nbr[(size_t)t * a.max_n + i];
            result = (float)((double)result + (double)(a.values[c] - pm) * sh.X[i]);
        }
        result = (float)((double)result + sh.X[n] * (double)(sv - sm));
    } else {
        result = pm + sv - sm;
    }
    a.out[node] = result;
    a.out_mask[node] = 1;
}

// Systems of at most 16 unknowns (max_neighbours <= 15): four per
// wavefront, 16 lanes each, lane i of a group owning row i.  Same
// operation order as k_cokriging; a group whose system is singular (or has
// no neighbours) keeps taking part in the wavefront's barriers.
constexpr int kGroupRows = 16;

struct CoGroupLds {
    double A[4][kGroupRows][kGroupRows + 1];
    double B[4][kGroupRows];
    double X[4][kGroupRows];
    int order[4][kGroupRows];
    int used[4][kGroupRows];
    int xyz[4][3][kGroupRows];
};

__global__ __launch_bounds__(64) void k_cokriging_g16(CokrigingArgs a) {
    __shared__ CoGroupLds sh;
    const int lane = threadIdx.x;
    const int g = lane >> 4, i = lane & 15;
    const int t = blockIdx.x * 4 + g;
    const bool valid = t < a.T;
    const int n = valid ? min(a.cnt[t], a.max_n) : 0;
    const int size = n + 1;
    const int node = valid ? a.targets[t] : 0;
    int cx, cy, cz;
    cell_xyz(node, a.X, a.Y, cx, cy, cz);
    int xi = 0, yi = 0, zi = 0;
    if (i < n) {
        const int nb = a.nbr[(size_t)t * a.max_n + i];
        cell_xyz(nb, a.X, a.Y, xi, yi, zi);
        sh.xyz[g][0][i] = xi;
        sh.xyz[g][1][i] = yi;
        sh.xyz[g][2][i] = zi;
    }
    __syncthreads();
    // build_system (simple_cokriging_markI.cpp:19-55)
    if (i < n) {
        for (int j = 0; j < n; ++j)
            sh.A[g][i][j] = co_cov(a.cov, sh.xyz[g][0][j] - xi, sh.xyz[g][1][j] - yi, sh.xyz[g][2][j] - zi);
        const double cr = a.coef * co_cov(a.cross, xi - cx, yi - cy, zi - cz);
        sh.A[g][i][n] = cr;
        sh.A[g][n][i] = cr;
        sh.B[g][i] = co_cov(a.cov, cx - xi, cy - yi, cz - zi);
    }
    if (i == 0 && valid) {
        sh.B[g][n] = a.cross0;
        sh.A[g][n][n] = a.sec_var;
    }
    sh.used[g][i] = 0;
    __syncthreads();
    // the largest system of the wavefront bounds the column loop
    int smax = valid ? size : 0;
    for (int o = 16; o <= 32; o <<= 1) smax = max(smax, __shfl_xor(smax, o));
    // gauss_solve (gauss_solver.cpp:5-52)
    bool ok = valid && n > 0;
    for (int c = 0; c < smax; ++c) {
        const bool act = ok && c < size;
        const uint64_t m = __ballot(act && i < size && sh.used[g][i] == 0 && sh.A[g][i][c] != 0.0);
        const uint32_t gm = (uint32_t)(m >> (16 * g)) & 0xffffu;
        const int row = gm ? __ffs(gm) - 1 : -1;
        if (act && row < 0) ok = false;  // the matrix is singular
        const bool step = act && row >= 0;
        const double coef = step ? sh.A[g][row][c] : 1.0;
        __syncthreads();
        if (step) {
            if (i == 0) {
                sh.used[g][row] = 1;
                sh.order[g][c] = row;
                sh.B[g][row] = sh.B[g][row] / coef;
            }
            if (i >= c && i < size) sh.A[g][row][i] = sh.A[g][row][i] / coef;
        }
        __syncthreads();
        if (step && i < size && i != row && sh.used[g][i] == 0) {
            const double f = sh.A[g][i][c];
            for (int k = c; k < size; ++k) sh.A[g][i][k] = sh.A[g][i][k] - f * sh.A[g][row][k];
            sh.B[g][i] = sh.B[g][i] - f * sh.B[g][row];
        }
        __syncthreads();
    }
    // back substitution (gauss_solver.cpp:54-62)
    double acc = ok && i < size ? sh.B[g][i] : 0.0;
    const int rowi = ok && i < size ? sh.order[g][i] : 0;
    for (int j = smax - 1; j >= 0; --j) {
        if (ok && i == j && j < size) sh.X[g][j] = acc;
        __syncthreads();
        if (ok && j < size && i < j) acc = acc - sh.A[g][rowi][j] * sh.X[g][j];
    }
    wave_count(a.counters, 1, i == 0 && ok);
    wave_count(a.counters, 2, i == 0 && valid && n > 0 && !ok);
    wave_count(a.counters, 3, i == 0 && valid && n == 0);
    __syncthreads();
    if (i != 0 || !valid) return;
    // combine (simple_cokriging_markI.cpp:82-104); the failure value :235
    const float pm = a.primary_mean, sm = a.secondary_mean;
    const float sv = a.sec_mask[node] != 0 ? a.sec[node] : sm;
    float result;
    if (ok) {
        result = pm;
        for (int k = 0; k < n; ++k) {
            const int cc = a.nbr[(size_t)t * a.max_n + k];
            result = (float)((double)result + (double)(a.values[cc] - pm) * sh.X[g][k]);
        }
        result = (float)((double)result + sh.X[g][n] * (double)(sv - sm));
    } else {
        result = pm + sv - sm;
    }
    a.out[node] = result;
    a.out_mask[node] = 1;
}

}  // namespace

void launch_cokriging(const CokrigingArgs &a, hipStream_t s) {
    if (a.T <= 0) return;
    if (a.max_n < kGroupRows) {
        hipLaunchKernelGGL(k_cokriging_g16, dim3((a.T + 3) / 4), dim3(64), 0, s, a);
        return;
    }
    const size_t S = (size_t)a.max_n + 1;
    const size_t lds = sizeof(double) * (S * (S + 1) + 2 * S) + sizeof(int) * (2 * S + 3 * (size_t)a.max_n);
    hipLaunchKernelGGL(k_cokriging, dim3(a.T), dim3(64), lds, s, a);
}

}  // namespace hpgl_amd
