// search.hip - neighbour search kernels and small element-wise helpers.
//
// Template scan (sugarbox_neighbour_lookup.h:36-61): one lane per target
// cell; the lanes of a wavefront hold neighbouring cells and walk the same
// template offsets in lockstep, so the 64 mask probes of one offset hit one
// or two cache lines.  A lane stops when it has max_n neighbours.  Once at
// most kTailLanes lanes of a wavefront are still searching, those targets
// are handed, with their count and next offset, to a tail pass that gives
// each one a whole wavefront (64 offsets per round), so no wavefront idles
// behind one lane in a sparse neighbourhood.  Both passes take the finds in
// template order, so the lists are those of the sequential scan.
//
// Cell bins (clusterizer.cpp:70-184): counts per bin of informed cells and,
// per target, whether any of its 27 bins holds more than `limit` cells (the
// reference then uses the template scan, sugarbox_indexed_neighbour_lookup.h:68).
#include "engine.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace hpgl_amd {

namespace {

// Template offsets a wavefront probes per round: their loads are issued
// together, then the results are taken in template order, so a lane keeps
// exactly the offsets the sequential scan keeps and ignores those past its
// max_n-th find.
constexpr int kScanUnroll = 8;

// Hand-over to the tail pass: (list position, next template offset) per
// target, and their count.
struct ScanTail {
    int2 *cont;
    unsigned *n_cont;
};
// Lanes still searching at which a wavefront hands them to the tail pass.
constexpr int kTailLanes = 8;
// Largest max_n whose finds the first pass stages in LDS (256 x max_n
// words per workgroup) and writes out as whole rows at its end; above it
// each find is stored as it is taken.
constexpr int kScanStageMax = 32;

// STAGE: find c of thread i is kept at s_found[c * 256 + i] (consecutive
// lanes, consecutive banks) until the wavefront leaves the walk; then the
// wavefront writes its 64 lists in row-major order, 64 consecutive list
// words per store, instead of one scattered word per lane and find.
template <int PRED, bool STAGE>
__global__ __launch_bounds__(256) void k_template_scan(ScanArgs a, ScanTail tl) {
    extern __shared__ int s_found[];
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = li < a.n_list;
    const int t = active ? (a.list ? a.list[li] : li) : 0;
    const int node = active ? a.targets[t] : 0;
    const int X = a.X, Y = a.Y, Z = a.Z;
    const int x = node % X, y = node % (X * Y) / X, z = node / (X * Y);
    const int32_t my_key = (PRED == 1 && active) ? a.key[node] : 0;
    int count = 0;
    int *out = a.nbr + (size_t)t * a.max_n;
    const int max_n = a.max_n;
    bool want = active && count < max_n;
    uint64_t sim = 0;
    const int lane = threadIdx.x & 63;
    for (int k0 = 0; k0 < a.tsize; k0 += kScanUnroll) {
        const uint64_t wm = __ballot(want);
        if (wm == 0) break;
        if (__popcll(wm) <= kTailLanes) {
            // the lanes still searching continue in the tail pass from k0
            unsigned base = 0;
            if (lane == 0) base = atomicAdd(tl.n_cont, (unsigned)__popcll(wm));
            base = __shfl(base, 0);
            if (want) tl.cont[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(wm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)wm, 0u))] = make_int2(li, k0);
            break;
        }
        // the round's template entries (wave-uniform) and every lane's
        // probe; lanes that do not probe read their own cell instead, so
        // all loads are issued before the first is used
        int4 o[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) o[u] = a.tmpl[__builtin_amdgcn_readfirstlane(min(k0 + u, a.tsize - 1))];
        int idx[kScanUnroll];
        bool probe[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
            const int px = x + o[u].x, py = y + o[u].y, pz = z + o[u].z;
            probe[u] = want & (k0 + u < a.tsize) & ((unsigned)px < (unsigned)X) & ((unsigned)py < (unsigned)Y) &
                       ((unsigned)pz < (unsigned)Z);
            idx[u] = probe[u] ? (pz * Y + py) * X + px : node;
        }
        int32_t v[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) v[u] = PRED == 0 ? (int32_t)a.mask[idx[u]] : a.key[idx[u]];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) {
            bool inf;
            if (PRED == 0) {
                inf = v[u] != 0;
            } else {
                inf = v[u] < my_key;
                if (want && probe[u] && inf && v[u] >= 0) sim |= 1ull << count;
            }
            if (want && probe[u] && inf) {
                if (STAGE)
                    s_found[count * 256 + threadIdx.x] = idx[u];
                else
                    out[count] = idx[u];
                ++count;
                want = count < max_n;
            }
        }
    }
    if (STAGE) {
        // every exit of the walk above is wave-uniform, so the whole
        // wavefront is here; it reads back only its own lanes' finds
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int wbase = threadIdx.x & ~63;
        const float inv = 1.0f / (float)max_n;
        for (int e = lane; e < 64 * max_n; e += 64) {
            // row r = e / max_n (e < 4096 is exact in float; one step corrects the rounding)
            int r = (int)((float)e * inv);
            if (r * max_n > e)
                --r;
            else if ((r + 1) * max_n <= e)
                ++r;
            const int col = e - r * max_n;
            const int tr = __shfl(t, r, 64);
            const int cr = __shfl(count, r, 64);
            if (col < cr) a.nbr[(size_t)tr * max_n + col] = s_found[col * 256 + wbase + r];
        }
    }
    if (active) a.cnt[t] = count;
    if (PRED == 1 && active && a.simbits) a.simbits[t] = sim;
}

// Tail pass: one wavefront per handed-over target, 64 consecutive template
// offsets per round.  The round's finds are taken in offset order (lane
// order) up to max_n, so the list continues exactly as the sequential scan
// would; the count and, for the key predicate, the simulated-neighbour
// bits continue from what the first pass stored.
template <int PRED>
__global__ __launch_bounds__(256) void k_template_scan_tail(ScanArgs a, ScanTail tl) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int n_waves = (gridDim.x * blockDim.x) >> 6;
    const int n_cont = (int)*tl.n_cont;
    const int X = a.X, Y = a.Y, Z = a.Z;
    const int max_n = a.max_n;
    for (int e = wave; e < n_cont; e += n_waves) {
        const int2 c = tl.cont[e];
        const int t = a.list ? a.list[c.x] : c.x;
        const int node = a.targets[t];
        const int x = node % X, y = node % (X * Y) / X, z = node / (X * Y);
        const int32_t my_key = PRED == 1 ? a.key[node] : 0;
        int count = a.cnt[t];
        uint64_t sim = (PRED == 1 && a.simbits) ? a.simbits[t] : 0ull;
        int *out = a.nbr + (size_t)t * max_n;
        for (int k0 = c.y; k0 < a.tsize && count < max_n; k0 += 64) {
            const int k = k0 + lane;
            const int4 o = a.tmpl[k < a.tsize ? k : 0];
            const int px = x + o.x, py = y + o.y, pz = z + o.z;
            const bool probe = (k < a.tsize) & ((unsigned)px < (unsigned)X) & ((unsigned)py < (unsigned)Y) &
                               ((unsigned)pz < (unsigned)Z);
            const int idx = probe ? (pz * Y + py) * X + px : node;
            const int32_t v = PRED == 0 ? (int32_t)a.mask[idx] : a.key[idx];
            const bool inf = probe && (PRED == 0 ? v != 0 : v < my_key);
            const uint64_t m = __ballot(inf);
            const int pos = count + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const bool take = inf && pos < max_n;
            if (take) out[pos] = idx;
            if (PRED == 1) {
                uint64_t bit = (take && v >= 0) ? (1ull << pos) : 0ull;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) bit |= __shfl_xor(bit, d, 64);
                sim |= bit;
            }
            count = min(max_n, count + (int)__popcll(m));
        }
        if (lane == 0) {
            a.cnt[t] = count;
            if (PRED == 1 && a.simbits) a.simbits[t] = sim;
        }
    }
}

__global__ void k_bin_ids(const int *cells, int n, BinArgs b, int *bin_of, uint32_t *bin_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cells[i];
    const int x = c % b.X, y = c % (b.X * b.Y) / b.X, z = c / (b.X * b.Y);
    const int bx = x / b.r[0], by = y / b.r[1], bz = z / b.r[2];
    const int bin = (bz * b.m[1] + by) * b.m[0] + bx;
    bin_of[i] = bin;
    atomicAdd(bin_count + bin, 1u);
}

__global__ void k_path_flags(const int *targets, int T, BinArgs b, const uint32_t *bin_count, uint8_t *flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int c = targets[t];
    const int x = c % b.X, y = c % (b.X * b.Y) / b.X, z = c / (b.X * b.Y);
    const int bx = x / b.r[0], by = y / b.r[1], bz = z / b.r[2];
    bool exceeded = false;
    for (int k = bz - 1; k <= bz + 1; ++k)
        for (int j = by - 1; j <= by + 1; ++j)
            for (int i = bx - 1; i <= bx + 1; ++i) {
                // rect_3d_t::has is inclusive at both ends (geometry.h:21-24);
                // i <= m-1 always holds for cells inside the grid.
                if (i < 0 || j < 0 || k < 0 || i >= b.m[0] || j >= b.m[1] || k >= b.m[2]) continue;
                if (bin_count[(k * b.m[1] + j) * b.m[0] + i] > (uint32_t)b.limit) exceeded = true;
            }
    flag[t] = exceeded ? 0 : 1;  // 1 = cluster path
}

__global__ void k_copy_informed(const float *in, const uint8_t *in_mask, long long n, float *out, uint8_t *out_mask) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (in_mask[i] != 0) {
        out[i] = in[i];
        out_mask[i] = 1;
    }
}

// most_probable_category (cdf_utils.cpp:9-24): first index of the maximum,
// compared in double.
__global__ void k_ik_argmax(const float *probs, int K, long long n, uint8_t *out, uint8_t *out_mask) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double best = probs[i];
    int idx = 0;
    for (int k = 1; k < K; ++k) {
        const double p = probs[(long long)k * n + i];
        if (p > best) {
            best = p;
            idx = k;
        }
    }
    out[i] = (uint8_t)idx;
    out_mask[i] = 1;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Differences of the first words are offsets in the rank table's strides
// (modulo 2^32, exact for offsets inside the table); the second word holds
// x and y modulo 2^16 for the search's box test.
__global__ void k_pack_lin(const int *cells, int n, int X, int Y, uint32_t w0, uint32_t w01, uint2 *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n + kClusterLinPad) return;
    if (i >= n) {
        out[i] = make_uint2(0u, 0u);
        return;
    }
    const int c = cells[i];
    const uint32_t x = c % X, y = c % (X * Y) / X, z = c / (X * Y);
    out[i] = make_uint2(4u * (x + w0 * y + w01 * z), (x & 0xffffu) | (y << 16));
}

inline unsigned blocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_template_scan(const ScanArgs &a, int pred_kind, hipStream_t s) {
    if (a.n_list <= 0) return;
    // the hand-over list (at most one entry per target) and its count, in
    // stream-ordered pool memory
    ScanTail tl;
    tl.cont = (int2 *)pool_alloc((size_t)a.n_list * sizeof(int2));
    tl.n_cont = (unsigned *)pool_alloc(sizeof(unsigned));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)tl.n_cont, 0u, 1ll);
    const unsigned tail_blocks = std::min<unsigned>(blocks(a.n_list, 256), 512u);
    const bool stage = a.max_n >= 1 && a.max_n <= kScanStageMax;
    const size_t lds = stage ? (size_t)256 * a.max_n * sizeof(int) : 0;
    const dim3 grid(blocks(a.n_list, 256));
    if (pred_kind == 0) {
        if (stage)
            hipLaunchKernelGGL((k_template_scan<0, true>), grid, dim3(256), lds, s, a, tl);
        else
            hipLaunchKernelGGL((k_template_scan<0, false>), grid, dim3(256), 0, s, a, tl);
        hipLaunchKernelGGL(k_template_scan_tail<0>, dim3(tail_blocks), dim3(256), 0, s, a, tl);
    } else {
        if (stage)
            hipLaunchKernelGGL((k_template_scan<1, true>), grid, dim3(256), lds, s, a, tl);
        else
            hipLaunchKernelGGL((k_template_scan<1, false>), grid, dim3(256), 0, s, a, tl);
        hipLaunchKernelGGL(k_template_scan_tail<1>, dim3(tail_blocks), dim3(256), 0, s, a, tl);
    }
    pool_free(tl.cont);
    pool_free(tl.n_cont);
}

void launch_bin_ids(const int *cells, int n, const BinArgs &b, int *bin_of, uint32_t *bin_count, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bin_ids, dim3(blocks(n, 256)), dim3(256), 0, s, cells, n, b, bin_of, bin_count);
}

void launch_path_flags(const int *targets, int T, const BinArgs &b, const uint32_t *bin_count, uint8_t *flag,
                       hipStream_t s) {
    if (T <= 0) return;
    hipLaunchKernelGGL(k_path_flags, dim3(blocks(T, 256)), dim3(256), 0, s, targets, T, b, bin_count, flag);
}

void launch_copy_informed(const float *in, const uint8_t *in_mask, long long n, float *out, uint8_t *out_mask,
                          hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_informed, dim3(blocks(n, 256)), dim3(256), 0, s, in, in_mask, n, out, out_mask);
}

void launch_ik_argmax(const float *probs, int K, long long n, uint8_t *out, uint8_t *out_mask, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ik_argmax, dim3(blocks(n, 256)), dim3(256), 0, s, probs, K, n, out, out_mask);
}

void launch_pack_lin(const int *cells, int n, int X, int Y, uint32_t w0, uint32_t w01, uint2 *out, hipStream_t s) {
    if (n < 0) return;
    hipLaunchKernelGGL(k_pack_lin, dim3(blocks((long long)n + kClusterLinPad, 256)), dim3(256), 0, s, cells, n, X, Y,
                       w0, w01, out);
}

void launch_fill_u32(uint32_t *p, uint32_t v, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
}

}  // namespace hpgl_amd
