// depth_kernels.hip - gfx950 kernels of the chiaro-fusion depth hot path
// that do not evaluate ZNCC patches (those are in depth_zncc.hip).
//
// Stage map (reference lines in crates/chiaro-fusion/src/depth.rs):
//   k_sgm_paths        accumulate_path x 8 families :1279-1383
//   k_sgm_select       path sum + select_depths     :1385-1488
//   k_complete         complete_depth_field         :1490-1564
//   k_reject_isolated  reject_isolated_direct_depths :881-924
//   k_cc_*             reject_small_direct_components :933-989
//   k_fit_planes       fit_local_depth_planes       :995-1060
//   k_consensus        enforce_warp_consensus       :582-618
//   k_boundaries       suppress_warp_boundaries     :620-672
#include <hip/hip_runtime.h>

#include "depth_device.hpp"
#include "depth_internal.hpp"

namespace chd {

// -------------------------------------------------------------------- SGM

constexpr uint32_t kSgmMaxChunks = 8;  // labels <= 512

// DPP controls of the gfx9 family (whole-wave shifts and row broadcasts).
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138,
              kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ float dpp(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), Ctrl,
                                                      RowMask, 0xf, false));
}

// Minimum over the wavefront, returned in a scalar register. min is exact
// and the path values are never -0.0 or NaN, so the reduction order does not
// matter.
__device__ __forceinline__ float wave_min_uniform(float v) {
    const float inf = __int_as_float(0x7f800000);
    v = fminf(v, dpp<kDppRowShr1>(inf, v));
    v = fminf(v, dpp<kDppRowShr2>(inf, v));
    v = fminf(v, dpp<kDppRowShr4>(inf, v));
    v = fminf(v, dpp<kDppRowShr8>(inf, v));
    v = fminf(v, dpp<kDppRowBcast15, 0xa>(inf, v));
    v = fminf(v, dpp<kDppRowBcast31, 0xc>(inf, v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

// A wavefront walks one path (d, p) of one of the 8 families at a time,
// lanes over labels (label = lane + 64 j), writing that family's L_r volume. Family order is
// H->, H<-, V-down, V-up, DR->, DR<-, DL->, DL<- (depth.rs:1287-1309).
// The path geometry is wave-uniform (scalar registers); the costs of the
// next kSgmPrefetch nodes are loaded ahead so that the recurrence waits on
// its own arithmetic, not on memory. Lanes past the last label hold +inf,
// which the shifts turn into the reference's missing neighbours.
constexpr uint32_t kSgmPrefetch = 8;

// Label set of a node: every label, or with band windows (lo, hi non-null)
// label 0 and lo[node] .. hi[node]. The costs outside the set are +inf, so
// every path value there is +inf too. With windows the path sweep reads,
// computes and writes only the 64-label chunks that meet a node's set (the
// others hold +inf in registers, exactly what reading and computing them
// would give); the select reads the set only and takes +inf elsewhere, so
// path values outside the sets are never read.
struct LabelSet {
    uint32_t lo, hi;  // lo = 1, hi = L - 1 without windows
    __device__ __forceinline__ bool has(uint32_t l) const { return l == 0 || (l >= lo && l <= hi); }
};
__device__ __forceinline__ LabelSet label_set(const uint16_t *lo, const uint16_t *hi, size_t node,
                                              uint32_t L) {
    LabelSet w;
    w.lo = lo ? lo[node] : 1u;
    w.hi = lo ? hi[node] : L - 1;
    return w;
}

__host__ __device__ __forceinline__ uint32_t sgm_path_count(uint32_t columns, uint32_t rows) {
    return 2 * rows + 2 * columns + 4 * (columns + rows - 1);
}

// Path of rank `rank` in the order of decreasing length: the horizontal
// paths (length C), the vertical ones (R), then the diagonals of the four
// diagonal families from the longest (min(C, R), |C - R| + 1 of them per
// family) down to length 1 (two per family and length). Returns the family
// d (0..7, the reference's order) and the path index p within it, as
// diagonal_paths numbers them (depth.rs:1313-1338). A bijection onto the
// 2R + 2C + 4(C + R - 1) paths.
__device__ __forceinline__ void sgm_path_by_rank(uint32_t rank, uint32_t C, uint32_t R,
                                                 uint32_t &d, uint32_t &p) {
    if (rank < 2 * R) {
        d = rank / R;
        p = rank % R;
        return;
    }
    rank -= 2 * R;
    if (rank < 2 * C) {
        d = 2 + rank / C;
        p = rank % C;
        return;
    }
    rank -= 2 * C;
    const uint32_t m = C < R ? C : R, M = C < R ? R : C, full = M - m + 1;
    uint32_t f, q;  // family among the diagonal four, rank within it
    if (rank < 4 * full) {
        f = rank % 4;
        q = rank / 4;
    } else {
        const uint32_t g = rank - 4 * full;
        f = (g % 8) / 2;
        q = full + 2 * (g / 8) + g % 2;
    }
    d = 4 + f;
    const bool mirrored = f >= 2;
    if (q < full) {  // full-length diagonals
        if (C >= R)
            p = mirrored ? R - 1 + q : q;
        else
            p = q == 0 ? (mirrored ? C - 1 : 0) : C + q - 1;
        return;
    }
    // the two diagonals of length l: one starting on the first row, one on
    // the first (mirrored: last) column
    const uint32_t k = q - full, l = m - 1 - k / 2;
    if (k % 2 == 0)
        p = mirrored ? l - 1 : C - l;
    else
        p = C + R - 1 - l;
}

// Chunks of 64 labels that meet the label set of `node` (bit j = labels
// 64j .. 64j + 63): chunk 0 (label 0) and the chunks of lo .. hi.
// Wave-uniform.
__device__ __forceinline__ uint32_t chunk_mask(const uint16_t *__restrict__ win_lo,
                                               const uint16_t *__restrict__ win_hi,
                                               int64_t node) {
    const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)win_lo[node]) >> 6;
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)win_hi[node]) >> 6;
    return 1u | (((2u << b) - 1u) & ~((1u << a) - 1u));
}

// Win: band windows are given (win_lo/hi non-null; a kernel of its own, so
// that the plain sweep is compiled without the chunk tests).
template <uint32_t J, bool Win>
__device__ __forceinline__ void sgm_path(const float *__restrict__ costs,
                                         const float *__restrict__ guidance, uint32_t C,
                                         uint32_t R, uint32_t L, float *__restrict__ paths,
                                         uint32_t d, uint32_t p,
                                         const uint16_t *__restrict__ win_lo,
                                         const uint16_t *__restrict__ win_hi) {
    const uint32_t lane = threadIdx.x & 63u;
    int64_t start, stride;
    uint32_t len;
    switch (d) {
    case 0: start = (int64_t)p * C; stride = 1; len = C; break;
    case 1: start = (int64_t)p * C + C - 1; stride = -1; len = C; break;
    case 2: start = p; stride = C; len = R; break;
    case 3: start = (int64_t)(R - 1) * C + p; stride = -(int64_t)C; len = R; break;
    default: {
        // diagonal_paths (depth.rs:1313-1338)
        const bool mirrored = d >= 6;
        uint32_t c0, r0;
        if (p < C) {
            c0 = p;
            r0 = 0;
        } else {
            c0 = mirrored ? C - 1 : 0;
            r0 = p - C + 1;
        }
        len = mirrored ? min(c0 + 1, R - r0) : min(C - c0, R - r0);
        int64_t s = mirrored ? (int64_t)C - 1 : (int64_t)C + 1;
        start = (int64_t)r0 * C + c0;
        stride = s;
        if (d == 5 || d == 7) {
            start += (int64_t)(len - 1) * s;
            stride = -s;
        }
        break;
    }
    }
    const size_t volume = (size_t)C * R * L;
    float *out = paths + (size_t)d * volume;
    const float inf = __int_as_float(0x7f800000);

    float prev[J];
    uint32_t prev_mask = ~0u;  // chunks of the previous node's set (Win)
    if constexpr (Win) prev_mask = chunk_mask(win_lo, win_hi, start);
#pragma unroll
    for (uint32_t j = 0; j < J; j++) {
        const uint32_t l = lane + 64 * j;
        prev[j] = inf;
        if (l < L && (!Win || ((prev_mask >> j) & 1u))) {
            prev[j] = costs[(size_t)start * L + l];
            out[(size_t)start * L + l] = prev[j];
        }
    }
    float g_prev = guidance[start];
    // ring of prefetched costs, guidance (and chunk masks) for steps i .. i +
    // kSgmPrefetch - 1
    float ahead[kSgmPrefetch][J];
    float g_ahead[kSgmPrefetch];
    uint32_t m_ahead[kSgmPrefetch];
#pragma unroll
    for (uint32_t k = 0; k < kSgmPrefetch; k++) {
        const uint32_t i = 1 + k;
        const int64_t node = start + (int64_t)i * stride;
        g_ahead[k] = i < len ? guidance[node] : 0.0f;
        uint32_t mk = ~0u;
        if constexpr (Win) {
            mk = i < len ? chunk_mask(win_lo, win_hi, node) : 0u;
            m_ahead[k] = mk;
        }
#pragma unroll
        for (uint32_t j = 0; j < J; j++) {
            const uint32_t l = lane + 64 * j;
            ahead[k][j] =
                i < len && l < L && (!Win || ((mk >> j) & 1u)) ? costs[(size_t)node * L + l] : inf;
        }
    }
    for (uint32_t base = 1; base < len; base += kSgmPrefetch) {
#pragma unroll
        for (uint32_t k = 0; k < kSgmPrefetch; k++) {
            const uint32_t i = base + k;
            if (i >= len) break;
            const int64_t node = start + (int64_t)i * stride;
            float cost[J];
#pragma unroll
            for (uint32_t j = 0; j < J; j++) cost[j] = ahead[k][j];
            const float g = g_ahead[k];
            uint32_t mask = ~0u;
            if constexpr (Win) mask = m_ahead[k];
            {
                const uint32_t ip = i + kSgmPrefetch;
                const int64_t pnode = start + (int64_t)ip * stride;
                if (ip < len) {
                    g_ahead[k] = guidance[pnode];
                    uint32_t pm = ~0u;
                    if constexpr (Win) {
                        pm = chunk_mask(win_lo, win_hi, pnode);
                        m_ahead[k] = pm;
                    }
#pragma unroll
                    for (uint32_t j = 0; j < J; j++) {
                        const uint32_t l = lane + 64 * j;
                        ahead[k][j] =
                            l < L && (!Win || ((pm >> j) & 1u)) ? costs[(size_t)pnode * L + l] : inf;
                    }
                }
            }
            const float edge = fabsf(g - g_prev);
            g_prev = g;
            const float edge_scale = 1.0f / (1.0f + 4.0f * edge);
            const float p1 = kSgmSmallPenalty * rmaxf(edge_scale, 0.35f);
            const float p2 = kSgmLargePenalty * rmaxf(edge_scale, 0.12f);
            float m = prev[0];
#pragma unroll
            for (uint32_t j = 1; j < J; j++)
                if (!Win || ((prev_mask >> j) & 1u)) m = fminf(m, prev[j]);
            m = wave_min_uniform(m);
            const float jump = m + p2;
            float cur[J];
#pragma unroll
            for (uint32_t j = 0; j < J; j++) {
                if constexpr (Win) {
                    if (!((mask >> j) & 1u)) {  // +inf costs: +inf values
                        cur[j] = inf;
                        continue;
                    }
                }
                // neighbours across chunk edges come from lane 63 of the
                // previous chunk and lane 0 of the next
                const float carry_up =
                    j > 0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(prev[j > 0 ? j - 1 : 0]), 63)) : inf;
                const float carry_dn =
                    j + 1 < J ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(prev[j + 1 < J ? j + 1 : j]), 0)) : inf;
                const float up = dpp<kDppWaveShr1>(carry_up, prev[j]);
                const float dn = dpp<kDppWaveShl1>(carry_dn, prev[j]);
                const float mn = fminf(fminf(fminf(prev[j], up + p1), dn + p1), jump);
                cur[j] = cost[j] + mn - m;
            }
#pragma unroll
            for (uint32_t j = 0; j < J; j++) {
                const uint32_t l = lane + 64 * j;
                if (l < L && (!Win || ((mask >> j) & 1u))) out[(size_t)node * L + l] = cur[j];
                prev[j] = l < L ? cur[j] : inf;
            }
            if constexpr (Win) prev_mask = mask;
        }
    }
}

// Waves take paths by rank in a snake order over the rounds (wave w takes
// rank w in even rounds and W - 1 - w in odd ones), so that with the ranks
// sorted by decreasing length every wave, and so every SIMD, gets about the
// same number of steps. (One wave per path in launch order gave the most
// loaded SIMD a third more steps than the average.)
__device__ __forceinline__ uint32_t snake_rank(uint32_t round, uint32_t wave, uint32_t waves) {
    return round * waves + (round % 2 == 0 ? wave : waves - 1 - wave);
}

template <uint32_t J, bool Win>
__global__ void __launch_bounds__(256) k_sgm_paths(const float *__restrict__ costs,
                                                   const float *__restrict__ guidance, uint32_t C,
                                                   uint32_t R, uint32_t L,
                                                   float *__restrict__ paths, uint32_t waves,
                                                   const uint16_t *__restrict__ win_lo,
                                                   const uint16_t *__restrict__ win_hi) {
    const uint32_t wave =
        __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (wave >= waves) return;
    const uint32_t total = sgm_path_count(C, R);
    for (uint32_t round = 0; round * waves < total; round++) {
        const uint32_t rank = snake_rank(round, wave, waves);
        if (rank >= total) continue;
        uint32_t d, p;
        sgm_path_by_rank(rank, C, R, d, p);
        sgm_path<J, Win>(costs, guidance, C, R, L, paths, d, p, win_lo, win_hi);
    }
}

// All path sweeps of several captures in one launch; first[] is the prefix
// of each job's path count (njobs + 1 entries). With `interleave` (every job
// on the same grid) global rank g is rank g / njobs of job g % njobs, so the
// longest paths of every capture come first.
template <uint32_t J, bool Win>
__global__ void __launch_bounds__(256) k_sgm_paths_batch(const SgmJob *__restrict__ jobs,
                                                         const uint32_t *__restrict__ first,
                                                         uint32_t njobs, uint32_t L,
                                                         uint32_t waves, uint32_t interleave) {
    const uint32_t wave =
        __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (wave >= waves) return;
    const uint32_t total = first[njobs];
    for (uint32_t round = 0; round * waves < total; round++) {
        const uint32_t g = snake_rank(round, wave, waves);
        if (g >= total) continue;
        uint32_t j = 0, local;
        if (interleave) {
            j = g % njobs;
            local = g / njobs;
        } else {
            while (j + 1 < njobs && first[j + 1] <= g) j++;
            local = g - first[j];
        }
        const SgmJob job = jobs[j];
        uint32_t d, p;
        sgm_path_by_rank(local, job.columns, job.rows, d, p);
        sgm_path<J, Win>(job.costs, job.guidance, job.columns, job.rows, L, job.paths, d, p,
                         job.win_lo, job.win_hi);
    }
}

uint32_t sgm_waves(uint32_t columns, uint32_t rows) { return sgm_path_count(columns, rows); }

// Path waves of one launch: four per SIMD of the device (CUs read once per
// device), at most one per path.
static uint32_t sgm_launch_waves(uint32_t paths) {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    int n = dev >= 0 && dev < 64 ? cus[dev] : 0;
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        if (dev >= 0 && dev < 64) cus[dev] = n;
    }
    const uint32_t w = (uint32_t)n * 16u;
    return paths < w ? paths : w;
}

hipError_t launch_sgm_batch(hipStream_t stream, const SgmJob *jobs, const uint32_t *first,
                            uint32_t njobs, uint32_t total_paths, uint32_t labels,
                            uint32_t paths_per_block, bool same_grids, bool windows) {
    const uint32_t J = (labels + 63) / 64;
    const uint32_t wpb = paths_per_block == 1 || paths_per_block == 2 ? paths_per_block : 4;
    const uint32_t waves = sgm_launch_waves(total_paths);
    const dim3 blocks((waves + wpb - 1) / wpb), threads(64 * wpb);
    const uint32_t interleave = same_grids ? 1u : 0u;
    switch (J) {
#define CHD_SGM_CASE(n)                                                                        \
    case n:                                                                                    \
        if (windows)                                                                           \
            hipLaunchKernelGGL((k_sgm_paths_batch<n, true>), blocks, threads, 0, stream, jobs, \
                               first, njobs, labels, waves, interleave);                       \
        else                                                                                   \
            hipLaunchKernelGGL((k_sgm_paths_batch<n, false>), blocks, threads, 0, stream, jobs,\
                               first, njobs, labels, waves, interleave);                       \
        break;
        CHD_SGM_CASE(1)
        CHD_SGM_CASE(2)
        CHD_SGM_CASE(3)
        CHD_SGM_CASE(4)
        CHD_SGM_CASE(5)
        CHD_SGM_CASE(6)
        CHD_SGM_CASE(7)
        CHD_SGM_CASE(8)
#undef CHD_SGM_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sgm(hipStream_t stream, const float *costs, const float *guidance,
                      uint32_t columns, uint32_t rows, uint32_t labels, float *paths,
                      uint32_t paths_per_block, const uint16_t *window_lo,
                      const uint16_t *window_hi) {
    const uint32_t J = (labels + 63) / 64;
    const uint32_t waves = sgm_launch_waves(sgm_path_count(columns, rows));
    const uint32_t wpb = paths_per_block == 1 || paths_per_block == 2 ? paths_per_block : 4;
    const dim3 blocks((waves + wpb - 1) / wpb), threads(64 * wpb);
    switch (J) {
#define CHD_SGM_CASE(n)                                                                        \
    case n:                                                                                    \
        if (window_lo)                                                                         \
            hipLaunchKernelGGL((k_sgm_paths<n, true>), blocks, threads, 0, stream, costs,      \
                               guidance, columns, rows, labels, paths, waves, window_lo,       \
                               window_hi);                                                     \
        else                                                                                   \
            hipLaunchKernelGGL((k_sgm_paths<n, false>), blocks, threads, 0, stream, costs,     \
                               guidance, columns, rows, labels, paths, waves, nullptr,         \
                               nullptr);                                                       \
        break;
        CHD_SGM_CASE(1)
        CHD_SGM_CASE(2)
        CHD_SGM_CASE(3)
        CHD_SGM_CASE(4)
        CHD_SGM_CASE(5)
        CHD_SGM_CASE(6)
        CHD_SGM_CASE(7)
        CHD_SGM_CASE(8)
#undef CHD_SGM_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Value of label `l` (wave-uniform) held as chunk l/64 of lane l%64.
template <uint32_t J>
__device__ __forceinline__ float label_value(const float (&v)[J], uint32_t l) {
    float out = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < J; j++)
        if (j == (l >> 6))
            out = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[j]), (int)(l & 63u)));
    return out;
}

// Minimum of an int over the wavefront, in a scalar register (the order of
// an integer minimum does not matter).
__device__ __forceinline__ int32_t wave_min_i32_uniform(int32_t v) {
    const int32_t big = 0x7fffffff;
    auto step = [&](int32_t shifted) { v = shifted < v ? shifted : v; };
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowShr1, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowShr2, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowShr4, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowShr8, 0xf, 0xf, false));
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowBcast15, 0xa, 0xf, false));
    step(__builtin_amdgcn_update_dpp(big, v, kDppRowBcast31, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// select_depths (depth.rs:1385-1467) for one node whose regularised costs
// are held as chunks in the wave; lane 0 writes the node.
template <uint32_t J>
__device__ void select_node(const float (&s)[J], uint32_t L, uint32_t lane, size_t node,
                            const float *scores, const uint8_t *tested, const double *labels,
                            const chd_options &o, chd_node_depth *field, uint8_t *fillable,
                            uint16_t *best_label_out) {
    chd_node_depth result;
    result.depth = 0.0;
    result.confidence = 0.0f;
    result.improvement = 0.0f;
    result.valid = 0;
    result.regularized = 0;
    for (int k = 0; k < 6; k++) result.pad[k] = 0;
    uint8_t fill = 0;
    uint16_t best_out = 0;
    if (tested[node] && L >= 2) {
        // WTA: first minimum over labels 1..L-1 by total_cmp: the smallest
        // key over the wave, then the lowest label holding it (chunks in
        // label order, lanes by the ballot's lowest bit)
        int32_t key[J];
        int32_t bk = 0x7fffffff;
#pragma unroll
        for (uint32_t j = 0; j < J; j++) {
            const uint32_t l = lane + 64 * j;
            key[j] = tkey32(s[j]);
            if (l >= 1 && l < L && key[j] < bk) bk = key[j];
        }
        bk = wave_min_i32_uniform(bk);
        uint32_t best = 0;
#pragma unroll
        for (uint32_t j = J; j-- > 0;) {
            const uint32_t l = lane + 64 * j;
            const uint64_t hit = __ballot(l >= 1 && l < L && key[j] == bk);
            if (hit) best = 64 * j + (uint32_t)__builtin_ctzll(hit);
        }
        best_out = (uint16_t)best;
        const float best_cost = label_value(s, best);
        const float left = label_value(s, best >= 1 ? best - 1 : 0);
        const float right = label_value(s, best + 1 < L ? best + 1 : best);
        // competing minimum over labels more than two away
        int32_t ck = 0x7fffffff;
        bool candidate = false;
#pragma unroll
        for (uint32_t j = 0; j < J; j++) {
            const uint32_t l = lane + 64 * j;
            const uint32_t diff = l > best ? l - best : best - l;
            if (l >= 1 && l < L && diff > 2) {
                candidate = true;
                if (key[j] < ck) ck = key[j];
            }
        }
        const int have = __ballot(candidate) != 0;
        ck = wave_min_i32_uniform(ck);
        if (lane == 0) {
            // sublabel_depth (depth.rs:1469-1488)
            double depth = labels[best];
            if (!(best <= 1 || best + 1 >= L)) {
                float curvature = left - 2.0f * best_cost + right;
                if (finite_f(curvature) && !(curvature <= 1.0e-6f)) {
                    double offset =
                        (double)rclampf(0.5f * (left - right) / curvature, -0.75f, 0.75f);
                    double lower_inverse = 1.0 / labels[best - 1];
                    double upper_inverse = 1.0 / labels[best + 1];
                    double inverse_step = (upper_inverse - lower_inverse) * 0.5;
                    depth = 1.0 / (1.0 / labels[best] + offset * inverse_step);
                }
            }
            const size_t offset = node * L;
            const float score = scores[offset + best];
            if (!finite_f(score) || score < kMinimumRegularizedScore) {
                fill = 1;
            } else {
                // key -> value: invert the total_cmp transform
                float competing = best_cost;
                if (have) {
                    int32_t raw = ck ^ (int32_t)(((uint32_t)(ck >> 31)) >> 1);
                    competing = __int_as_float(raw);
                }
                const float margin = rmaxf((competing - best_cost) / kSgmDirections, 0.0f);
                const float baseline = scores[offset];
                const float improvement = finite_f(baseline) ? score - baseline : 0.0f;
                if (!(finite_f(baseline) && improvement < -kMaximumRegularizedBaselineLoss)) {
                    const bool measured =
                        score >= o.minimum_score && improvement >= o.minimum_improvement * 0.5f &&
                        (margin >= o.minimum_margin || improvement >= o.minimum_improvement);
                    const float score_confidence = rclampf(
                        (score - o.minimum_score) / rmaxf(1.0f - o.minimum_score, 1.0e-3f), 0.0f,
                        1.0f);
                    const float margin_confidence =
                        rclampf(margin / rmaxf(o.minimum_margin, 1.0e-3f), 0.0f, 1.0f);
                    const float improvement_confidence =
                        rclampf(improvement / rmaxf(o.minimum_improvement, 1.0e-3f), 0.0f, 1.0f);
                    result.valid = 1;
                    result.depth = depth;
                    result.confidence =
                        measured ? rclampf(0.30f + 0.45f * score_confidence +
                                               0.25f * margin_confidence,
                                           0.0f, 1.0f)
                                 : rclampf(0.10f + 0.25f * score_confidence +
                                               0.15f * margin_confidence +
                                               0.10f * improvement_confidence,
                                           0.10f, 0.55f);
                    result.improvement = improvement;
                    result.regularized = measured ? 0 : 1;
                }
            }
        }
    }
    if (lane == 0) {
        field[node] = result;
        fillable[node] = fill;
        if (best_label_out) best_label_out[node] = best_out;
    }
}

// Sum of the eight L_r volumes in the reference's order, fused with
// select_depths. One wavefront per node.
template <uint32_t J>
__global__ void __launch_bounds__(256) k_sgm_select(const float *__restrict__ paths,
                                                    const float *__restrict__ scores,
                                                    const uint8_t *__restrict__ tested,
                                                    const double *__restrict__ labels,
                                                    uint32_t nodes, uint32_t L,
                                                    chd_options o, float *aggregated,
                                                    chd_node_depth *field, uint8_t *fillable,
                                                    uint16_t *best_label,
                                                    const uint16_t *__restrict__ win_lo,
                                                    const uint16_t *__restrict__ win_hi) {
    const uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (node >= nodes) return;
    const size_t volume = (size_t)nodes * L;
    const LabelSet w = label_set(win_lo, win_hi, node, L);
    float s[J];
#pragma unroll
    for (uint32_t j = 0; j < J; j++) {
        s[j] = 0.0f;
        const uint32_t l = lane + 64 * j;
        if (l < L) {
            const size_t at = (size_t)node * L + l;
            float acc = __int_as_float(0x7f800000);  // outside the set: 8 x +inf
            if (w.has(l)) {
                acc = 0.0f;
#pragma unroll
                for (uint32_t d = 0; d < 8; d++) acc += paths[d * volume + at];
            }
            s[j] = acc;
            if (aggregated) aggregated[at] = acc;
        }
    }
    select_node<J>(s, L, lane, node, scores, tested, labels, o, field, fillable, best_label);
}

hipError_t launch_sgm_select(hipStream_t stream, const float *paths, const float *scores,
                             const uint8_t *tested, const double *labels, uint32_t columns,
                             uint32_t rows, uint32_t label_count, const chd_options &options,
                             float *aggregated, chd_node_depth *field, uint8_t *fillable,
                             uint16_t *best_label, const uint16_t *window_lo,
                             const uint16_t *window_hi) {
    const uint32_t J = (label_count + 63) / 64;
    if (J > kSgmMaxChunks) return hipErrorInvalidValue;
    const uint32_t nodes = columns * rows;
    dim3 blocks((nodes + 3) / 4);
    switch (J) {
#define CHD_SELECT_CASE(n)                                                                     \
    case n:                                                                                    \
        hipLaunchKernelGGL(k_sgm_select<n>, blocks, dim3(256), 0, stream, paths, scores,       \
                           tested, labels, nodes, label_count, options, aggregated, field,     \
                           fillable, best_label, window_lo, window_hi);                        \
        break;
        CHD_SELECT_CASE(1)
        CHD_SELECT_CASE(2)
        CHD_SELECT_CASE(3)
        CHD_SELECT_CASE(4)
        CHD_SELECT_CASE(5)
        CHD_SELECT_CASE(6)
        CHD_SELECT_CASE(7)
        CHD_SELECT_CASE(8)
#undef CHD_SELECT_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// select_depths alone, from a given regularised volume.
template <uint32_t J>
__global__ void __launch_bounds__(256) k_select_depths(const float *__restrict__ regularised,
                                                       const float *__restrict__ scores,
                                                       const uint8_t *__restrict__ tested,
                                                       const double *__restrict__ labels,
                                                       uint32_t nodes, uint32_t L,
                                                       chd_options o, chd_node_depth *field,
                                                       uint8_t *fillable, uint16_t *best_label) {
    const uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (node >= nodes) return;
    float s[J];
#pragma unroll
    for (uint32_t j = 0; j < J; j++) {
        const uint32_t l = lane + 64 * j;
        s[j] = l < L ? regularised[(size_t)node * L + l] : 0.0f;
    }
    select_node<J>(s, L, lane, node, scores, tested, labels, o, field, fillable, best_label);
}

hipError_t launch_select_depths(hipStream_t stream, const float *regularised,
                                const float *scores, const uint8_t *tested,
                                const double *labels, uint32_t nodes, uint32_t label_count,
                                const chd_options &options, chd_node_depth *field,
                                uint8_t *fillable, uint16_t *best_label) {
    const uint32_t J = (label_count + 63) / 64;
    if (J > kSgmMaxChunks) return hipErrorInvalidValue;
    dim3 blocks((nodes + 3) / 4);
    switch (J) {
#define CHD_SELECT_CASE(n)                                                                     \
    case n:                                                                                    \
        hipLaunchKernelGGL(k_select_depths<n>, blocks, dim3(256), 0, stream, regularised,      \
                           scores, tested, labels, nodes, label_count, options, field,         \
                           fillable, best_label);                                              \
        break;
        CHD_SELECT_CASE(1)
        CHD_SELECT_CASE(2)
        CHD_SELECT_CASE(3)
        CHD_SELECT_CASE(4)
        CHD_SELECT_CASE(5)
        CHD_SELECT_CASE(6)
        CHD_SELECT_CASE(7)
        CHD_SELECT_CASE(8)
#undef CHD_SELECT_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// complete_depth_field, one iteration (depth.rs:1498-1563). Iterations past
// the reference's early exit reproduce their input, so running all of them
// equals the reference's loop.
__global__ void k_complete(const chd_node_depth *previous, chd_node_depth *field,
                           const float *guidance, const uint8_t *fillable, uint32_t C,
                           uint32_t R, uint32_t minimum_support) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= C * R) return;
    chd_node_depth out = previous[index];
    const uint32_t row = index / C, column = index % C;
    if (!previous[index].valid && fillable[index]) {
        double inv[8];
        float w[8];
        uint32_t m = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                if (dx == 0 && dy == 0) continue;
                int x = (int)column + dx, y = (int)row + dy;
                if (x < 0 || y < 0 || x >= (int)C || y >= (int)R) continue;
                const uint32_t ni = (uint32_t)y * C + (uint32_t)x;
                const chd_node_depth nb = previous[ni];
                if (!nb.valid) continue;
                const float edge = fabsf(guidance[index] - guidance[ni]);
                if (edge > 0.35f) continue;
                const float spatial = (dx != 0 && dy != 0) ? 0.707f : 1.0f;
                const float weight = nb.confidence * spatial * libm_expf(-3.0f * edge);
                if (weight >= 0.05f) {
                    const double v = 1.0 / nb.depth;
                    uint32_t j = m;
                    while (j > 0 && tkey64(inv[j - 1]) > tkey64(v)) {
                        inv[j] = inv[j - 1];
                        w[j] = w[j - 1];
                        j--;
                    }
                    inv[j] = v;
                    w[j] = weight;
                    m++;
                }
            }
        if (m >= minimum_support) {
            float total = -0.0f;
            for (uint32_t i = 0; i < m; i++) total += w[i];
            float accumulated = 0.0f;
            double inverse_depth = inv[m / 2];
            for (uint32_t i = 0; i < m; i++) {
                accumulated += w[i];
                if (accumulated >= total * 0.5f) {
                    inverse_depth = inv[i];
                    break;
                }
            }
            float maxw = 0.0f;
            for (uint32_t i = 0; i < m; i++) maxw = rmaxf(maxw, w[i]);
            out.valid = 1;
            out.depth = 1.0 / inverse_depth;
            out.confidence = rclampf(maxw * 0.92f, 0.15f, 0.75f);
            out.improvement = 0.0f;
            out.regularized = 1;
        }
    }
    field[index] = out;
}

hipError_t launch_complete_depth_field(hipStream_t stream, const chd_node_depth *previous,
                                       chd_node_depth *field, const float *guidance,
                                       const uint8_t *fillable, uint32_t columns,
                                       uint32_t rows, uint32_t minimum_neighbour_support) {
    const uint32_t n = columns * rows;
    hipLaunchKernelGGL(k_complete, dim3((n + 255) / 256), dim3(256), 0, stream, previous, field,
                       guidance, fillable, columns, rows, minimum_neighbour_support);
    return hipGetLastError();
}

// ---------------------------------------------------------------- filters

__global__ void k_reject_isolated(const chd_node_depth *source, chd_node_depth *field,
                                  const float *guidance, uint32_t C, uint32_t R, double tol) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= C * R) return;
    chd_node_depth node = source[index];
    if (node.valid) {
        const uint32_t row = index / C, column = index % C;
        const double inverse = 1.0 / node.depth;
        uint32_t agreeing = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                if (dx == 0 && dy == 0) continue;
                const int x = (int)column + dx, y = (int)row + dy;
                if (x < 0 || y < 0 || x >= (int)C || y >= (int)R) continue;
                const uint32_t ni = (uint32_t)y * C + (uint32_t)x;
                const chd_node_depth nbr = source[ni];
                if (!nbr.valid) continue;
                const float image_edge = fabsf(guidance[index] - guidance[ni]);
                if (image_edge <= 0.6f && fabs(1.0 / nbr.depth - inverse) <= tol) agreeing++;
            }
        const uint32_t minimum = node.confidence >= 0.75f ? 1u : 2u;
        if (agreeing < minimum) clear_node(node);
    }
    field[index] = node;
}

hipError_t launch_reject_isolated(hipStream_t stream, const chd_node_depth *source,
                                  chd_node_depth *field, const float *guidance,
                                  uint32_t columns, uint32_t rows, double inverse_tolerance) {
    const uint32_t n = columns * rows;
    hipLaunchKernelGGL(k_reject_isolated, dim3((n + 255) / 256), dim3(256), 0, stream, source,
                       field, guidance, columns, rows, inverse_tolerance);
    return hipGetLastError();
}

// Connected components of the symmetric relation of
// reject_small_direct_components by concurrent union-find: roots link
// larger index under smaller with compare-and-swap.
// Parents always point to a smaller index of the same component, so a walk
// ends at the root; while walking, each visited node is re-pointed to its
// grandparent (a shorter path to the same root; concurrent writers only
// ever store ancestors).
__device__ __forceinline__ uint32_t cc_find(uint32_t *parent, uint32_t x) {
    uint32_t curr = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (curr != x) {
        uint32_t prev = x, next;
        while (curr > (next = __hip_atomic_load(&parent[curr], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(&parent[prev], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            prev = curr;
            curr = next;
        }
    }
    return curr;
}

__device__ __forceinline__ bool cc_edge(const chd_node_depth *source, const float *guidance,
                                        uint32_t i, uint32_t j, double tol) {
    const float image_edge = fabsf(guidance[i] - guidance[j]);
    const double inverse_difference = fabs(1.0 / source[j].depth - 1.0 / source[i].depth);
    return image_edge <= 0.6f && inverse_difference <= tol;
}

// Components are first joined inside tiles of kCcTileW x kCcTileH nodes:
// each block runs a union-find over its tile's edges in LDS (parents by
// local index, always a smaller one), then stores every node's parent as its
// local root's global index. Row-major order within a tile is the global
// order, so parents still point to smaller indices of the same component,
// and k_cc_union only has the edges that cross tile borders left to join in
// global memory (a global union-find over every edge, each node first under
// its smallest related neighbour, took 0.83-1.1 ms for 813 k fine nodes).
constexpr uint32_t kCcTileW = 64, kCcTileH = 16;

__device__ __forceinline__ uint32_t cc_local_find(uint32_t *lp, uint32_t x) {
    uint32_t p = __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return x;
}

__device__ __forceinline__ bool cc_related(const double *inverse, const float *guide, uint32_t a,
                                           uint32_t b, double tol) {
    // cc_edge with the reciprocal depths computed once per node
    return fabsf(guide[a] - guide[b]) <= 0.6f && fabs(inverse[b] - inverse[a]) <= tol;
}

__global__ void __launch_bounds__(256) k_cc_local(const chd_node_depth *source,
                                                  const float *guidance, uint32_t C, uint32_t R,
                                                  double tol, uint32_t *parent, uint32_t *sizes) {
    constexpr uint32_t N = kCcTileW * kCcTileH;
    __shared__ uint32_t lp[N];
    __shared__ double inverse[N];
    __shared__ float guide[N];
    __shared__ uint8_t valid[N];
    const uint32_t tiles_x = (C + kCcTileW - 1) / kCcTileW;
    const uint32_t x0 = (blockIdx.x % tiles_x) * kCcTileW, y0 = (blockIdx.x / tiles_x) * kCcTileH;
    const uint32_t w = min(kCcTileW, C - x0), h = min(kCcTileH, R - y0);
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
        const uint32_t lx = k % kCcTileW, ly = k / kCcTileW;
        lp[k] = k;
        valid[k] = 0;
        if (lx < w && ly < h) {
            const uint32_t g = (y0 + ly) * C + x0 + lx;
            sizes[g] = 0;
            const chd_node_depth d = source[g];
            valid[k] = d.valid ? 1 : 0;
            inverse[k] = 1.0 / d.depth;
            guide[k] = guidance[g];
        }
    }
    __syncthreads();
    // forward edges inside the tile: right, down-left, down, down-right
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
        if (!valid[k]) continue;
        const int lx = (int)(k % kCcTileW), ly = (int)(k / kCcTileW);
        const int nx[4] = {lx + 1, lx - 1, lx, lx + 1};
        const int ny[4] = {ly, ly + 1, ly + 1, ly + 1};
        for (int e = 0; e < 4; e++) {
            if (nx[e] < 0 || nx[e] >= (int)w || ny[e] >= (int)h) continue;
            const uint32_t j = (uint32_t)ny[e] * kCcTileW + (uint32_t)nx[e];
            if (!valid[j] || !cc_related(inverse, guide, k, j, tol)) continue;
            uint32_t a = k, b = j;
            while (true) {
                a = cc_local_find(lp, a);
                b = cc_local_find(lp, b);
                if (a == b) break;
                if (a > b) {
                    const uint32_t s = a;
                    a = b;
                    b = s;
                }
                uint32_t expected = b;
                if (__hip_atomic_compare_exchange_strong(&lp[b], &expected, a, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP))
                    break;
                b = expected;
            }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
        const uint32_t lx = k % kCcTileW, ly = k / kCcTileW;
        if (lx >= w || ly >= h) continue;
        const uint32_t r = cc_local_find(lp, k);
        parent[(y0 + ly) * C + x0 + lx] = (y0 + r / kCcTileW) * C + x0 + r % kCcTileW;
    }
}

__global__ void k_cc_union(const chd_node_depth *source, const float *guidance, uint32_t C,
                           uint32_t R, double tol, uint32_t *parent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * R || !source[i].valid) return;
    const int row = (int)(i / C), column = (int)(i % C);
    const int nx[4] = {column + 1, column - 1, column, column + 1};
    const int ny[4] = {row, row + 1, row + 1, row + 1};
    for (int k = 0; k < 4; k++) {
        if (nx[k] < 0 || nx[k] >= (int)C || ny[k] >= (int)R) continue;
        // edges inside a tile were joined by k_cc_local
        if ((uint32_t)nx[k] / kCcTileW == (uint32_t)column / kCcTileW &&
            (uint32_t)ny[k] / kCcTileH == (uint32_t)row / kCcTileH)
            continue;
        const uint32_t j = (uint32_t)ny[k] * C + (uint32_t)nx[k];
        if (!source[j].valid || !cc_edge(source, guidance, i, j, tol)) continue;
        uint32_t a = i, b = j;
        while (true) {
            a = cc_find(parent, a);
            b = cc_find(parent, b);
            if (a == b) break;
            if (a > b) {
                uint32_t t = a;
                a = b;
                b = t;
            }
            uint32_t expected = b;
            if (__hip_atomic_compare_exchange_strong(&parent[b], &expected, a, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
            b = expected;
        }
    }
}

// Component sizes: neighbouring lanes mostly share a root, so each run of
// equal roots in a wavefront adds its length with one atomic.
__global__ void k_cc_count(const chd_node_depth *source, uint32_t *parent, uint32_t *sizes,
                           uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool valid = i < n && source[i].valid;
    const uint32_t root = valid ? cc_find(parent, i) : 0xffffffffu;
    const uint32_t before = __shfl_up(root, 1, 64);
    const bool head = valid && (lane == 0 || before != root);
    const bool boundary = !valid || lane == 0 || before != root;
    const unsigned long long ends = __ballot(boundary);
    if (head) {
        const unsigned long long later = lane == 63 ? 0ull : ends >> (lane + 1);
        const uint32_t next = later ? lane + 1 + (uint32_t)__builtin_ctzll(later) : 64u;
        atomicAdd(&sizes[root], next - lane);
    }
}

__global__ void k_cc_apply(const chd_node_depth *source, chd_node_depth *field, uint32_t *parent,
                           const uint32_t *sizes, uint32_t n, uint32_t minimum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    chd_node_depth node = source[i];
    if (node.valid && sizes[cc_find(parent, i)] < minimum) clear_node(node);
    field[i] = node;
}

hipError_t launch_reject_components(hipStream_t stream, const chd_node_depth *source,
                                    chd_node_depth *field, const float *guidance,
                                    uint32_t columns, uint32_t rows, double inverse_tolerance,
                                    uint32_t minimum_nodes, uint32_t *parent, uint32_t *sizes) {
    const uint32_t n = columns * rows;
    const dim3 grid((n + 255) / 256), block(256);
    if (minimum_nodes <= 1) {
        hipError_t e = hipMemcpyAsync(field, source, (size_t)n * sizeof(chd_node_depth),
                                      hipMemcpyDeviceToDevice, stream);
        return e;
    }
    const uint32_t tiles = ((columns + kCcTileW - 1) / kCcTileW) * ((rows + kCcTileH - 1) / kCcTileH);
    hipLaunchKernelGGL(k_cc_local, dim3(tiles), block, 0, stream, source, guidance, columns, rows,
                       inverse_tolerance, parent, sizes);
    hipLaunchKernelGGL(k_cc_union, grid, block, 0, stream, source, guidance, columns, rows,
                       inverse_tolerance, parent);
    hipLaunchKernelGGL(k_cc_count, grid, block, 0, stream, source, parent, sizes, n);
    hipLaunchKernelGGL(k_cc_apply, grid, block, 0, stream, source, field, parent, sizes, n,
                       minimum_nodes);
    return hipGetLastError();
}

__global__ void k_fit_planes(const chd_node_depth *source, chd_node_depth *field,
                             const float *guidance, uint32_t C, uint32_t R, double tol) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= C * R) return;
    chd_node_depth node = source[index];
    if (node.valid) {
        const uint32_t row = index / C, column = index % C;
        const double centre_inverse = 1.0 / node.depth;
        double normal[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double rhs[3] = {0, 0, 0};
        uint32_t support = 0;
        for (int dy = -2; dy <= 2; dy++)
            for (int dx = -2; dx <= 2; dx++) {
                const int x = (int)column + dx, y = (int)row + dy;
                if (x < 0 || y < 0 || x >= (int)C || y >= (int)R) continue;
                const uint32_t ni = (uint32_t)y * C + (uint32_t)x;
                const chd_node_depth nbr = source[ni];
                if (!nbr.valid) continue;
                const double inverse = 1.0 / nbr.depth;
                const float image_edge = fabsf(guidance[index] - guidance[ni]);
                if (image_edge > 0.6f || fabs(inverse - centre_inverse) > tol) continue;
                const double spatial = libm_exp((double)(-(dx * dx + dy * dy)) / 6.0);
                const double weight =
                    (double)nbr.confidence * spatial * libm_exp(-2.5 * (double)image_edge);
                const double basis[3] = {1.0, (double)dx, (double)dy};
                for (int r = 0; r < 3; r++) {
                    rhs[r] += weight * basis[r] * inverse;
                    for (int c = 0; c < 3; c++) normal[r * 3 + c] += weight * basis[r] * basis[c];
                }
                support++;
            }
        if (support >= 6) {
            double inv[9];
            if (inverse3(normal, inv)) {
                double inverse = inv[0] * rhs[0] + inv[1] * rhs[1] + inv[2] * rhs[2];
                if (finite_d(inverse) && inverse > 0.0) {
                    const double maximum_adjustment = tol * 0.25;
                    inverse = rclamp(inverse, centre_inverse - maximum_adjustment,
                                     centre_inverse + maximum_adjustment);
                    node.depth = 1.0 / inverse;
                }
            }
        }
    }
    field[index] = node;
}

hipError_t launch_fit_planes(hipStream_t stream, const chd_node_depth *source,
                             chd_node_depth *field, const float *guidance, uint32_t columns,
                             uint32_t rows, double inverse_tolerance) {
    const uint32_t n = columns * rows;
    hipLaunchKernelGGL(k_fit_planes, dim3((n + 255) / 256), dim3(256), 0, stream, source, field,
                       guidance, columns, rows, inverse_tolerance);
    return hipGetLastError();
}

__device__ __forceinline__ bool warp_mapping(const chd_node_warp &d, float2 &g, float2 &p) {
    if (d.code == CHD_WARP_GLOBAL) {
        g = p = make_float2(d.global[0], d.global[1]);
        return true;
    }
    if (d.code == CHD_WARP_REFINED) {
        g = make_float2(d.global[0], d.global[1]);
        p = make_float2(d.point[0], d.point[1]);
        return true;
    }
    return false;
}

__device__ __forceinline__ bool decisions_agree(uint32_t kind, const chd_node_warp &c,
                                                const chd_node_warp &n) {
    if (kind == 0 && c.code == CHD_WARP_REFINED && n.code == CHD_WARP_REFINED) {
        const float cdx = c.point[0] - c.global[0], cdy = c.point[1] - c.global[1];
        const float ndx = n.point[0] - n.global[0], ndy = n.point[1] - n.global[1];
        const float dx = cdx - ndx, dy = cdy - ndy;
        return dx * dx + dy * dy <= 9.0f;
    }
    return kind == 1 && c.code == CHD_WARP_OCCLUDED && n.code == CHD_WARP_OCCLUDED;
}

__global__ void k_consensus(const chd_node_warp *source, chd_node_warp *decisions, uint32_t C,
                            uint32_t R) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= C * R) return;
    chd_node_warp d = source[index];
    uint32_t kind, minimum;
    if (d.code == CHD_WARP_REFINED) {
        kind = 0;
        minimum = d.measured ? 2 : 4;
    } else if (d.code == CHD_WARP_OCCLUDED) {
        kind = 1;
        minimum = d.measured ? 3 : 5;
    } else {
        decisions[index] = d;
        return;
    }
    const int row = (int)(index / C), column = (int)(index % C);
    uint32_t neighbours = 0;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            if (dx == 0 && dy == 0) continue;
            const int x = column + dx, y = row + dy;
            if (x < 0 || y < 0 || x >= (int)C || y >= (int)R) continue;
            neighbours += decisions_agree(kind, d, source[(uint32_t)y * C + (uint32_t)x]) ? 1u : 0u;
        }
    if (neighbours < minimum) {
        chd_node_warp g;
        g.code = CHD_WARP_GLOBAL;
        g.global[0] = d.global[0];
        g.global[1] = d.global[1];
        g.point[0] = g.point[1] = 0.0f;
        g.confidence = 0.0f;
        g.measured = 0;
        d = g;
    }
    decisions[index] = d;
}

hipError_t launch_consensus(hipStream_t stream, const chd_node_warp *source,
                            chd_node_warp *decisions, uint32_t columns, uint32_t rows) {
    const uint32_t n = columns * rows;
    hipLaunchKernelGGL(k_consensus, dim3((n + 255) / 256), dim3(256), 0, stream, source,
                       decisions, columns, rows);
    return hipGetLastError();
}

// A node is a boundary when any of its four-neighbour pairs (the pairs the
// reference visits as (i, right) and (i, down)) has both a mapping jump and
// an image edge; the pair test is symmetric.
__global__ void k_boundaries(const chd_node_warp *source, chd_node_warp *decisions,
                             const float *guidance, uint32_t C, uint32_t R) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= C * R) return;
    chd_node_warp d = source[index];
    float2 g, p;
    if (warp_mapping(d, g, p)) {
        const int row = (int)(index / C), column = (int)(index % C);
        const int nx[4] = {column + 1, column, column - 1, column};
        const int ny[4] = {row, row + 1, row, row - 1};
        bool boundary = false;
        for (int k = 0; k < 4 && !boundary; k++) {
            if (nx[k] < 0 || ny[k] < 0 || nx[k] >= (int)C || ny[k] >= (int)R) continue;
            const uint32_t ni = (uint32_t)ny[k] * C + (uint32_t)nx[k];
            float2 ng, np;
            if (!warp_mapping(source[ni], ng, np)) continue;
            // evaluate as the reference does with (first, second) in visit order
            const bool forward = k < 2;
            const float2 ag = forward ? g : ng, ap = forward ? p : np;
            const float2 bg = forward ? ng : g, bpnt = forward ? np : p;
            const uint32_t ia = forward ? index : ni, ib = forward ? ni : index;
            const float dax = ap.x - ag.x, day = ap.y - ag.y;
            const float dbx = bpnt.x - bg.x, dby = bpnt.y - bg.y;
            const float diffx = dax - dbx, diffy = day - dby;
            const bool mapping_edge =
                diffx * diffx + diffy * diffy > kWarpBoundaryDeltaPx * kWarpBoundaryDeltaPx;
            const bool image_edge = fabsf(guidance[ia] - guidance[ib]) > kWarpBoundaryContrast;
            boundary = mapping_edge && image_edge;
        }
        if (boundary) {
            chd_node_warp b;
            b.code = CHD_WARP_BOUNDARY;
            b.global[0] = g.x;
            b.global[1] = g.y;
            b.point[0] = b.point[1] = 0.0f;
            b.confidence = 0.0f;
            b.measured = 0;
            d = b;
        }
    }
    decisions[index] = d;
}

hipError_t launch_boundaries(hipStream_t stream, const chd_node_warp *source,
                             chd_node_warp *decisions, const float *guidance, uint32_t columns,
                             uint32_t rows) {
    const uint32_t n = columns * rows;
    hipLaunchKernelGGL(k_boundaries, dim3((n + 255) / 256), dim3(256), 0, stream, source,
                       decisions, guidance, columns, rows);
    return hipGetLastError();
}

__global__ void k_target_outputs(const chd_node_warp *decisions, uint32_t n, float *points,
                                 float *confidence, uint8_t *codes, unsigned long long *counts) {
    __shared__ unsigned int local[3];
    if (threadIdx.x < 3) local[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const chd_node_warp d = decisions[i];
        float px = __int_as_float(0x7fc00000), py = px, conf = 0.0f;
        switch (d.code) {
        case CHD_WARP_GLOBAL: px = d.global[0]; py = d.global[1]; conf = 1.0f; break;
        case CHD_WARP_REFINED:
            px = d.point[0]; py = d.point[1]; conf = d.confidence;
            atomicAdd(&local[0], 1u);
            break;
        case CHD_WARP_OCCLUDED: px = d.global[0]; py = d.global[1]; atomicAdd(&local[1], 1u); break;
        case CHD_WARP_BOUNDARY: px = d.global[0]; py = d.global[1]; atomicAdd(&local[2], 1u); break;
        default: break;
        }
        points[2 * (size_t)i] = px;
        points[2 * (size_t)i + 1] = py;
        confidence[i] = conf;
        codes[i] = (uint8_t)d.code;
    }
    __syncthreads();
    if (threadIdx.x < 3 && local[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)local[threadIdx.x]);
}

hipError_t launch_target_outputs(hipStream_t stream, const chd_node_warp *decisions,
                                 uint32_t nodes, float *points, float *confidence,
                                 uint8_t *codes, unsigned long long *counts) {
    hipLaunchKernelGGL(k_target_outputs, dim3((nodes + 255) / 256), dim3(256), 0, stream,
                       decisions, nodes, points, confidence, codes, counts);
    return hipGetLastError();
}

// Report inputs (depth.rs:394-418). Counts of the fine field (tested, measured, regularised nodes) and the keys
// whose medians the report takes (depth of every valid node, improvement of
// every measured node; the medians sort them, so their order is free). Each
// block takes a run of kStatsNodes nodes in two passes: the first counts and
// reserves the block's key ranges with one atomic per counter, the second
// writes the keys in node order. (One atomic per wave and counter on five
// shared addresses took 0.31 ms for 813 k nodes, serialised in L2.)
constexpr uint32_t kStatsNodes = 2048;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (uint32_t m = 32; m > 0; m >>= 1) v += __shfl_xor(v, (int)m, 64);
    return v;
}

__global__ void __launch_bounds__(256) k_field_stats(const chd_node_depth *field,
                                                     const uint8_t *tested, uint32_t n,
                                                     unsigned long long *counts,
                                                     unsigned long long *depth_keys,
                                                     unsigned int *improvement_keys,
                                                     uint32_t *key_counts) {
    __shared__ uint32_t partial[4][4];  // [wave][tested, measured, regularised, valid]
    __shared__ uint32_t base[2];        // the block's first depth / improvement key slot
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t first = blockIdx.x * kStatsNodes;
    const uint32_t last = min(n, first + kStatsNodes);
    uint32_t c[4] = {0, 0, 0, 0};
    for (uint32_t i = first + threadIdx.x; i < last; i += blockDim.x) {
        const chd_node_depth d = field[i];
        c[0] += tested[i] ? 1u : 0u;
        c[1] += d.valid && !d.regularized ? 1u : 0u;
        c[2] += d.valid && d.regularized ? 1u : 0u;
        c[3] += d.valid ? 1u : 0u;
    }
    for (int k = 0; k < 4; k++) {
        const uint32_t s = wave_sum(c[k]);
        if (lane == 0) partial[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot[4] = {0, 0, 0, 0};
        for (uint32_t w = 0; w < blockDim.x / 64; w++)
            for (int k = 0; k < 4; k++) tot[k] += partial[w][k];
        for (int k = 0; k < 3; k++)
            if (tot[k]) atomicAdd(&counts[k], (unsigned long long)tot[k]);
        base[0] = tot[3] ? atomicAdd(&key_counts[0], tot[3]) : 0u;
        base[1] = tot[1] ? atomicAdd(&key_counts[1], tot[1]) : 0u;
    }
    __syncthreads();
    uint32_t run_v = base[0], run_m = base[1];
    for (uint32_t i0 = first; i0 < last; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        chd_node_depth d;
        clear_node(d);
        if (i < last) d = field[i];
        const bool valid = d.valid, measured = d.valid && !d.regularized;
        const unsigned long long mv = __ballot(valid), mm = __ballot(measured);
        __syncthreads();  // the previous round's partial[] reads are done
        if (lane == 0) {
            partial[wave][0] = (uint32_t)__popcll(mv);
            partial[wave][1] = (uint32_t)__popcll(mm);
        }
        __syncthreads();
        uint32_t off_v = 0, off_m = 0, all_v = 0, all_m = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) {
            if (w < wave) {
                off_v += partial[w][0];
                off_m += partial[w][1];
            }
            all_v += partial[w][0];
            all_m += partial[w][1];
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        if (valid)
            depth_keys[run_v + off_v + (uint32_t)__popcll(mv & below)] =
                (unsigned long long)tkey64(d.depth) ^ 0x8000000000000000ull;
        if (measured)
            improvement_keys[run_m + off_m + (uint32_t)__popcll(mm & below)] =
                (unsigned int)tkey32(d.improvement) ^ 0x80000000u;
        run_v += all_v;
        run_m += all_m;
    }
}

hipError_t launch_field_stats(hipStream_t stream, const chd_node_depth *field,
                              const uint8_t *tested, uint32_t nodes,
                              unsigned long long *counts, unsigned long long *depth_keys,
                              unsigned int *improvement_keys, uint32_t *key_counts) {
    if (nodes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_field_stats, dim3((nodes + kStatsNodes - 1) / kStatsNodes), dim3(256), 0,
                       stream, field, tested, nodes, counts, depth_keys, improvement_keys,
                       key_counts);
    return hipGetLastError();
}

__global__ void k_final_map(const chd_node_depth *field, const uint8_t *tested, uint32_t n,
                            double *depth, float *confidence, uint8_t *provenance) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const chd_node_depth d = field[i];
    if (d.valid) {
        depth[i] = d.depth;
        confidence[i] = d.confidence;
        provenance[i] = d.regularized ? CHD_PROVENANCE_REGULARIZED : CHD_PROVENANCE_MEASURED;
    } else {
        depth[i] = __longlong_as_double(0x7ff8000000000000LL);
        confidence[i] = tested[i] ? 1.0f : 0.0f;
        provenance[i] = tested[i] ? CHD_PROVENANCE_GLOBAL : CHD_PROVENANCE_UNSUPPORTED;
    }
}

hipError_t launch_final_map(hipStream_t stream, const chd_node_depth *field,
                            const uint8_t *tested, uint32_t nodes, double *depth,
                            float *confidence, uint8_t *provenance) {
    hipLaunchKernelGGL(k_final_map, dim3((nodes + 255) / 256), dim3(256), 0, stream, field,
                       tested, nodes, depth, confidence, provenance);
    return hipGetLastError();
}

// ----------------------------------------------------------------- small

__global__ void k_pad_plane(float *plane, uint32_t width, uint32_t height) {
    const uint32_t pitch = width + 1;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < height) plane[(size_t)i * pitch + width] = plane[(size_t)i * pitch + width - 1];
    if (i < pitch) {
        const uint32_t x = i < width ? i : width - 1;
        plane[(size_t)height * pitch + i] = plane[(size_t)(height - 1) * pitch + x];
    }
}

// A plane of pitch `width` into a padded one (pitch width + 1, height + 1
// rows), padding included: one thread per stored element.
__global__ void k_pitch_copy(const float *src, uint32_t width, uint32_t height, float *dst) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x > width || y > height) return;
    const uint32_t sx = x < width ? x : width - 1, sy = y < height ? y : height - 1;
    dst[(size_t)y * (width + 1) + x] = src[(size_t)sy * width + sx];
}

hipError_t launch_pitch_copy(hipStream_t stream, const float *src, uint32_t width, uint32_t height,
                             float *dst) {
    hipLaunchKernelGGL(k_pitch_copy, dim3((width + 1 + 255) / 256, height + 1), dim3(256), 0, stream,
                       src, width, height, dst);
    return hipGetLastError();
}

hipError_t launch_pad_plane(hipStream_t stream, float *plane, uint32_t width, uint32_t height) {
    const uint32_t n = (width + 1 > height ? width + 1 : height);
    hipLaunchKernelGGL(k_pad_plane, dim3((n + 255) / 256), dim3(256), 0, stream, plane, width,
                       height);
    return hipGetLastError();
}

// Mosaic::luminance_half (image.rs:401-426) with a unit flat field: the
// f32 sum of a 2x2 CFA cell is an exact integer below 2^18, so the rest of
// the per-pixel formula, (sum * 0.25 - black) / range clamped at 0 and then
// ln(1 + 1000 x), is a function of that integer, tabulated by the host with
// the platform libm. With `headroom`, reserve_highlight_headroom
// (image.rs:69-75) first maps each u16 sample s to (s + 1) / 2 in wrapping
// u16 arithmetic. One thread per plane pixel, one 4-byte load per cell row
// when the width is even (cells then start 4-byte aligned in a 4-byte
// aligned mosaic), 2-byte loads otherwise.
__global__ void __launch_bounds__(256) k_luminance_half(const uint16_t *__restrict__ samples,
                                                        uint32_t width, uint32_t plane_width,
                                                        uint32_t plane_height,
                                                        const float *__restrict__ table,
                                                        int headroom, float *__restrict__ plane) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= plane_width || j >= plane_height) return;
    const uint16_t *r0 = samples + (size_t)(2 * j) * width + 2 * i;
    uint32_t a, b, c, d;
    if ((width & 1u) == 0) {
        const uint32_t top = *(const uint32_t *)r0, bottom = *(const uint32_t *)(r0 + width);
        a = top & 0xffffu;
        b = top >> 16;
        c = bottom & 0xffffu;
        d = bottom >> 16;
    } else {
        a = r0[0];
        b = r0[1];
        c = r0[width];
        d = r0[width + 1];
    }
    if (headroom) {
        a = ((a + 1u) & 0xffffu) >> 1;
        b = ((b + 1u) & 0xffffu) >> 1;
        c = ((c + 1u) & 0xffffu) >> 1;
        d = ((d + 1u) & 0xffffu) >> 1;
    }
    plane[(size_t)j * plane_width + i] = table[a + b + c + d];
}

hipError_t launch_luminance_half(hipStream_t stream, const uint16_t *samples, uint32_t width,
                                 uint32_t height, const float *table, int headroom,
                                 float *plane) {
    const uint32_t pw = width / 2, ph = height / 2;
    if (pw == 0 || ph == 0) return hipSuccess;
    hipLaunchKernelGGL(k_luminance_half, dim3((pw + 255) / 256, ph), dim3(256), 0, stream,
                       samples, width, pw, ph, table, headroom, plane);
    return hipGetLastError();
}

// 32 x 32 tiles through LDS; both sides coalesced.
__global__ void __launch_bounds__(256) k_transpose_plane(const float *rows, float *columns,
                                                         uint32_t width, uint32_t height) {
    __shared__ float tile[32][33];
    const uint32_t w = width + 1, h = height + 1;  // padded extents
    const uint32_t x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {
        const uint32_t x = x0 + threadIdx.x, y = y0 + k;
        if (x < w && y < h) tile[k][threadIdx.x] = rows[(size_t)y * w + x];
    }
    __syncthreads();
    for (uint32_t k = threadIdx.y; k < 32; k += 8) {
        const uint32_t y = y0 + threadIdx.x, x = x0 + k;
        if (x < w && y < h) columns[(size_t)x * h + y] = tile[threadIdx.x][k];
    }
}

// Element (line, along) of the quad array: the pair at (line, along) and
// the pair on the next line (the last stored line and column repeat the
// padding, which no in-bounds sample reads).
__global__ void __launch_bounds__(256) k_quad_plane(const float *__restrict__ plane,
                                                    uint32_t pitch, uint32_t lines,
                                                    float4 *__restrict__ quad) {
    const uint32_t along = blockIdx.x * blockDim.x + threadIdx.x, line = blockIdx.y;
    if (along >= pitch) return;
    const uint32_t a1 = along + 1 < pitch ? along + 1 : along;
    const uint32_t l1 = line + 1 < lines ? line + 1 : line;
    const float *r0 = plane + (size_t)line * pitch, *r1 = plane + (size_t)l1 * pitch;
    quad[(size_t)line * pitch + along] = make_float4(r0[along], r0[a1], r1[along], r1[a1]);
}

hipError_t launch_quad_plane(hipStream_t stream, const float *plane, uint32_t pitch,
                             uint32_t lines, float4 *quad) {
    hipLaunchKernelGGL(k_quad_plane, dim3((pitch + 255) / 256, lines), dim3(256), 0, stream, plane,
                       pitch, lines, quad);
    return hipGetLastError();
}

hipError_t launch_transpose_plane(hipStream_t stream, const float *rows, float *columns,
                                  uint32_t width, uint32_t height) {
    const dim3 grid((width + 1 + 31) / 32, (height + 1 + 31) / 32);
    hipLaunchKernelGGL(k_transpose_plane, grid, dim3(32, 8), 0, stream, rows, columns, width,
                       height);
    return hipGetLastError();
}

__global__ void k_epipolar_axes(const ViewDev *views, const uint32_t *vlist, uint32_t count,
                                uint32_t reference, double near_depth, double far_depth,
                                uint32_t *vertical) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const chd_camera &r = *views[reference].camera;
    const chd_camera &t = *views[vlist[k]].camera;
    d3 o, d;
    pixel_to_ray(r, (double)(r.width / 2), (double)(r.height / 2), o, d);
    const d3 pn{o.x + d.x * near_depth, o.y + d.y * near_depth, o.z + d.z * near_depth};
    const d3 pf{o.x + d.x * far_depth, o.y + d.y * far_depth, o.z + d.z * far_depth};
    double xn, yn, xf, yf;
    uint32_t v = 0;
    if (project(t, pn, xn, yn) && project(t, pf, xf, yf)) v = fabs(yn - yf) > fabs(xn - xf);
    vertical[k] = v;
}

hipError_t launch_epipolar_axes(hipStream_t stream, const ViewDev *views, const uint32_t *vlist,
                                uint32_t count, uint32_t reference, double near_depth,
                                double far_depth, uint32_t *vertical) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_epipolar_axes, dim3((count + 63) / 64), dim3(64), 0, stream, views,
                       vlist, count, reference, near_depth, far_depth, vertical);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t *data, uint32_t value, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] = value;
}
__global__ void k_fill_u64(unsigned long long *data, unsigned long long value, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] = value;
}

__global__ void k_fill_u8(uint8_t *data, uint8_t value, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] = value;
}

hipError_t launch_fill_u8(hipStream_t stream, uint8_t *data, uint8_t value, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, data,
                       value, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t stream, uint32_t *data, uint32_t value, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, data,
                       value, n);
    return hipGetLastError();
}
hipError_t launch_fill_u64(hipStream_t stream, unsigned long long *data,
                           unsigned long long value, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, data,
                       value, n);
    return hipGetLastError();
}

__global__ void k_expf(const float *x, float *y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = libm_expf(x[i]);
}
__global__ void k_exp(const double *x, double *y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = libm_exp(x[i]);
}

hipError_t launch_expf(hipStream_t stream, const float *x, float *y, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_expf, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, x, y, n);
    return hipGetLastError();
}
hipError_t launch_exp(hipStream_t stream, const double *x, double *y, size_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_exp, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, x, y, n);
    return hipGetLastError();
}

}  // namespace chd
