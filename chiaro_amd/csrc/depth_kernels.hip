// depth_kernels.hip - gfx950 kernels of the chiaro-fusion depth hot path.
//
// Stage map (reference lines in crates/chiaro-fusion/src/depth.rs):
//   k_cost_volume      build_cost_volume            :1187-1277
//   k_sgm_paths        accumulate_path x 8 families :1279-1383
//   k_sgm_select       path sum + select_depths     :1385-1488
//   k_complete         complete_depth_field         :1490-1564
//   k_direct           measure_direct_depths        :715-876
//   k_direct_wide      its r = 8 fallback           :803-846
//   k_reject_isolated  reject_isolated_direct_depths :881-924
//   k_cc_*             reject_small_direct_components :933-989
//   k_fit_planes       fit_local_depth_planes       :995-1060
//   k_targets          per-target decisions         :420-487, 1634-1736
//   k_consensus        enforce_warp_consensus       :582-618
//   k_boundaries       suppress_warp_boundaries     :620-672
//
// The ZNCC kernels share one block-level "patch context" in LDS: per node the
// reference-tap values and bilateral weights (label- and view-independent),
// per (node, view) the global-warp tap positions, the label-independent half
// of map_at_depth (the corrected reference ray), and the min/max tap position
// that decides Plane::sample's bounds test for every tap at once (the
// f64 -> f32 tap coordinate is monotone in the warp coordinate). A work item
// is then one (node, hypothesis, view) ZNCC over the taps in the reference's
// dy-major order, with each of the five sums accumulated in its own register
// exactly as the reference does.
#include <hip/hip_runtime.h>

#include "depth_device.hpp"
#include "depth_internal.hpp"

namespace chd {

// ---------------------------------------------------------- patch context

struct PatchLayout {
    uint32_t cams, pos, ray, ext, gc, tap, rv, centre, sums, spatial, node_ok, nv_ok, ray_ok,
        total;
};

__host__ __device__ inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

__host__ __device__ inline PatchLayout patch_layout(uint32_t nb, uint32_t v, uint32_t t) {
    PatchLayout l;
    uint32_t o = 0;
    l.cams = o; o = align16(o + v * (uint32_t)sizeof(chd_camera));
    l.pos = o; o = align16(o + nb * 16u);
    l.ray = o; o = align16(o + nb * v * 6u * 8u);
    l.ext = o; o = align16(o + nb * v * 16u);
    l.gc = o; o = align16(o + nb * v * 8u);
    l.tap = o; o = align16(o + nb * v * t * 16u);
    l.rv = o; o = align16(o + nb * t * 4u);
    l.centre = o; o = align16(o + nb * 4u);
    l.sums = o; o = align16(o + nb * 3u * 4u);
    l.spatial = o; o = align16(o + t * 4u);
    l.node_ok = o; o = align16(o + nb * 4u);
    l.nv_ok = o; o = align16(o + nb * v * 4u);
    l.ray_ok = o; o = align16(o + nb * v * 4u);
    l.total = o;
    return l;
}

struct PatchLds {
    chd_camera *cams;  // the scored views' cameras, in view-slot order
    double2 *pos;
    double *ray;
    float4 *ext;
    float2 *gc;
    float4 *tap;  // [node][view][tap]: weight, weight * reference, warp x, warp y
    float *rv;
    float *centre;
    float *sums;
    float *spatial;
    int *node_ok;
    int *nv_ok;
    int *ray_ok;
};

__device__ inline PatchLds bind_patch(unsigned char *base, const PatchLayout &l) {
    PatchLds p;
    p.cams = (chd_camera *)(base + l.cams);
    p.pos = (double2 *)(base + l.pos);
    p.ray = (double *)(base + l.ray);
    p.ext = (float4 *)(base + l.ext);
    p.gc = (float2 *)(base + l.gc);
    p.tap = (float4 *)(base + l.tap);
    p.rv = (float *)(base + l.rv);
    p.centre = (float *)(base + l.centre);
    p.sums = (float *)(base + l.sums);
    p.spatial = (float *)(base + l.spatial);
    p.node_ok = (int *)(base + l.node_ok);
    p.nv_ok = (int *)(base + l.nv_ok);
    p.ray_ok = (int *)(base + l.ray_ok);
    return p;
}

__device__ __forceinline__ DevPlane view_plane(const ViewDev &v) {
    DevPlane p;
    p.data = v.plane;
    p.width = v.plane_width;
    p.height = v.plane_height;
    p.pitch = v.plane_pitch;
    return p;
}
__device__ __forceinline__ DevWarp view_warp(const ViewDev &v) {
    DevWarp w;
    w.points = v.warp;
    w.step = v.warp_step;
    w.columns = v.warp_columns;
    w.rows = v.warp_rows;
    return w;
}

// Fills the patch context for `nb` nodes. On entry P.pos[n] holds each
// node's raster position and P.node_ok[n] whether the node exists; both
// must be visible (a __syncthreads() between writing them and the call).
// On exit (after a barrier) every array is valid. node_ok becomes 0 when
// any reference sample of the patch is outside the plane (the ZNCC is then
// None for every hypothesis and view); nv_ok[n][a] becomes 0 when the
// global warp is undefined at the centre or any tap (score_one_view /
// warped_patch_zncc return None); ray_ok[n][a] becomes 0 when the
// label-independent half of map_at_depth fails.
__device__ void patch_prologue(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t T,
                               float *centre_value) {
    const uint32_t V = S.view_count, tid = threadIdx.x, nth = blockDim.x;
    const int r = (int)S.radius;
    const uint32_t width = 2u * S.radius + 1u;
    const ViewDev &R = S.views[S.reference];
    const DevPlane rp = view_plane(R);
    const float sigma = rmaxf((float)S.radius * 0.75f, 1.0f);

    for (uint32_t i = tid; i < nb * V; i += nth) {
        P.nv_ok[i] = 1;
        P.ray_ok[i] = 1;
    }
    for (uint32_t n = tid; n < nb; n += nth) centre_value[n] = 0.0f;
    {
        constexpr uint32_t words = sizeof(chd_camera) / 4;
        for (uint32_t i = tid; i < V * words; i += nth) {
            const uint32_t *src = (const uint32_t *)S.views[S.vlist[i / words]].camera;
            ((uint32_t *)P.cams)[i] = src[i % words];
        }
    }
    __syncthreads();

    // phase 1: spatial weights, reference samples, warp taps, corrected rays
    for (uint32_t t = tid; t < T; t += nth) {
        int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
        float distance_sq = (float)(dx * dx + dy * dy);
        P.spatial[t] = libm_expf(-distance_sq / (2.0f * sigma * sigma));
    }
    for (uint32_t n = tid; n < nb; n += nth) {
        double2 p = P.pos[n];
        float c;
        if (plane_sample(rp, to_plane(p.x), to_plane(p.y), c)) {
            P.centre[n] = c;
            centre_value[n] = c;
        } else {
            P.node_ok[n] = 0;
        }
    }
    for (uint32_t i = tid; i < nb * T; i += nth) {
        uint32_t n = i / T, t = i % T;
        double2 p = P.pos[n];
        int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
        double px = p.x + (double)dx * 2.0, py = p.y + (double)dy * 2.0;
        float v;
        if (!plane_sample(rp, to_plane(px), to_plane(py), v)) {
            P.node_ok[n] = 0;
            v = 0.0f;
        }
        P.rv[i] = v;
    }
    for (uint32_t i = tid; i < nb * V * T; i += nth) {
        uint32_t n = i / (V * T), a = (i / T) % V, t = i % T;
        double2 p = P.pos[n];
        int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
        double px = p.x + (double)dx * 2.0, py = p.y + (double)dy * 2.0;
        float2 q;
        if (!warp_map(view_warp(S.views[S.vlist[a]]), (float)px, (float)py, q)) {
            P.nv_ok[n * V + a] = 0;
            q = make_float2(0.0f, 0.0f);
        }
        P.tap[i].z = q.x;
        P.tap[i].w = q.y;
    }
    for (uint32_t i = tid; i < nb * V; i += nth) {
        uint32_t n = i / V, a = i % V;
        double2 p = P.pos[n];
        const ViewDev &tv = S.views[S.vlist[a]];
        float2 gc;
        if (!warp_map(view_warp(tv), (float)p.x, (float)p.y, gc)) {
            P.nv_ok[i] = 0;
            P.ray_ok[i] = 0;
            continue;
        }
        P.gc[i] = gc;
        // corrected_reference = reference.map_from(target, gc, 1e8)
        // (depth.rs:1752-1756), then the reference ray through it.
        d3 o, d;
        pixel_to_ray(P.cams[a], (double)gc.x, (double)gc.y, o, d);
        d3 far{o.x + d.x * kInfinityDepth, o.y + d.y * kInfinityDepth, o.z + d.z * kInfinityDepth};
        double cx, cy;
        if (!project(*R.camera, far, cx, cy)) {
            P.ray_ok[i] = 0;
            continue;
        }
        pixel_to_ray(*R.camera, cx, cy, o, d);
        double *ray = P.ray + (size_t)i * 6;
        ray[0] = o.x; ray[1] = o.y; ray[2] = o.z;
        ray[3] = d.x; ray[4] = d.y; ray[5] = d.z;
    }
    __syncthreads();

    // phase 2: bilateral weights (copied into every view's tap record), tap
    // extremes
    for (uint32_t i = tid; i < nb * T; i += nth) {
        uint32_t n = i / T, t = i % T;
        if (!P.node_ok[n]) continue;
        float rv = P.rv[i];
        float range = libm_expf(-1.2f * fabsf(rv - P.centre[n]));
        float w = P.spatial[t] * range;
        float wr = w * rv;
        for (uint32_t a = 0; a < V; a++) {
            float4 *e = P.tap + ((size_t)n * V + a) * T + t;
            e->x = w;
            e->y = wr;
        }
    }
    for (uint32_t i = tid; i < nb * V; i += nth) {
        if (!P.nv_ok[i]) continue;
        const float4 *g = P.tap + (size_t)i * T;
        float x0 = g[0].z, x1 = g[0].z, y0 = g[0].w, y1 = g[0].w;
        for (uint32_t t = 1; t < T; t++) {
            x0 = fminf(x0, g[t].z);
            x1 = fmaxf(x1, g[t].z);
            y0 = fminf(y0, g[t].w);
            y1 = fmaxf(y1, g[t].w);
        }
        P.ext[i] = make_float4(x0, x1, y0, y1);
    }
    __syncthreads();

    // phase 3: label- and view-independent sums, in tap order
    for (uint32_t n = tid; n < nb; n += nth) {
        if (!P.node_ok[n] || V == 0) continue;
        float count = 0.0f, sum_r = 0.0f, sum_rr = 0.0f;
        for (uint32_t t = 0; t < T; t++) {
            const float4 w = P.tap[(size_t)n * V * T + t];
            count += w.x;
            sum_r += w.y;
            sum_rr += w.y * P.rv[n * T + t];
        }
        P.sums[n * 3 + 0] = count;
        P.sums[n * 3 + 1] = sum_r;
        P.sums[n * 3 + 2] = sum_rr;
    }
    __syncthreads();
}

// warped_patch_zncc (depth.rs:1772-1822) for node n, view slot a and the
// raster displacement (dx, dy).
__device__ __forceinline__ bool patch_zncc(const PatchLds &P, uint32_t n, uint32_t a, uint32_t V,
                                           uint32_t T, const DevPlane &tp, double dx, double dy,
                                           float &score, unsigned int *evaluations) {
    const float4 e = P.ext[n * V + a];
    const float xmin = to_plane((double)e.x + dx), xmax = to_plane((double)e.y + dx);
    const float ymin = to_plane((double)e.z + dy), ymax = to_plane((double)e.w + dy);
    if (!(xmin >= 0.0f && ymin >= 0.0f && xmax <= (float)(tp.width - 1) &&
          ymax <= (float)(tp.height - 1)))
        return false;
    atomicAdd(evaluations, 1u);
    const float4 *taps = P.tap + (size_t)(n * V + a) * T;
    const __amdgpu_buffer_rsrc_t rsrc = plane_rsrc(tp);
    float st = 0.0f, stt = 0.0f, sp = 0.0f;
#pragma unroll 9
    for (uint32_t t = 0; t < T; t++) {
        const float4 e = taps[t];
        const float x = to_plane((double)e.z + dx);
        const float y = to_plane((double)e.w + dy);
        const float tv = plane_sample_inside(rsrc, tp.pitch, x, y);
        const float wv = e.x * tv;
        st += wv;
        stt += wv * tv;
        sp += e.y * tv;
    }
    const float count = P.sums[n * 3 + 0], sr = P.sums[n * 3 + 1], srr = P.sums[n * 3 + 2];
    const float covariance = sp - sr * st / count;
    const float reference_energy = srr - sr * sr / count;
    const float target_energy = stt - st * st / count;
    const float denominator = sqrtf(rmaxf(reference_energy, 0.0f) * rmaxf(target_energy, 0.0f));
    if (!(denominator > 1.0e-6f)) return false;
    score = rclampf(covariance / denominator, -1.0f, 1.0f);
    return true;
}

// map_at_depth (depth.rs:1738-1759) from the cached corrected ray, then the
// 128 px parallax gate of score_one_view (depth.rs:1604-1614).
__device__ __forceinline__ bool delta_at_depth(const PatchLds &P, uint32_t n, uint32_t a,
                                               uint32_t V, const chd_camera &target,
                                               double depth, double &dx, double &dy) {
    const uint32_t i = n * V + a;
    if (!P.ray_ok[i]) return false;
    const double *r = P.ray + (size_t)i * 6;
    d3 point{r[0] + r[3] * depth, r[1] + r[4] * depth, r[2] + r[5] * depth};
    double qx, qy;
    if (!project(target, point, qx, qy)) return false;
    if (!(finite_d(qx) && finite_d(qy))) return false;
    const float2 gc = P.gc[i];
    dx = qx - (double)gc.x;
    dy = qy - (double)gc.y;
    return !(fabs(dx) > 128.0 || fabs(dy) > 128.0);
}

// aggregate (depth.rs:1761-1770) of V view scores at stride `stride` (NaN =
// the view returned None): mean of the best best_view_count in descending
// total_cmp order, summed from -0.0.
__device__ __forceinline__ bool aggregate_scores(const float *res, uint32_t stride, uint32_t V,
                                                 uint32_t minimum_support,
                                                 uint32_t best_view_count, float &out) {
    uint32_t count = 0;
    for (uint32_t a = 0; a < V; a++) count += res[a * stride] == res[a * stride] ? 1u : 0u;
    if (count < minimum_support) return false;
    const uint32_t k = best_view_count < count ? best_view_count : count;
    if (k < minimum_support) return false;
    uint32_t used = 0;
    float sum = -0.0f;
    for (uint32_t i = 0; i < k; i++) {
        int best = -1;
        int32_t best_key = 0;
        for (uint32_t a = 0; a < V; a++) {
            float s = res[a * stride];
            if ((used >> a) & 1u || s != s) continue;
            int32_t key = tkey32(s);
            if (best < 0 || key > best_key) {
                best = (int)a;
                best_key = key;
            }
        }
        used |= 1u << best;
        sum += res[best * stride];
    }
    out = sum / (float)k;
    return true;
}

// ------------------------------------------------------------ cost volume

__global__ void __launch_bounds__(256) k_cost_volume(PatchSpec S, GridSpec G, const double *labels,
                                                     uint32_t L, uint32_t nb, uint32_t T,
                                                     uint32_t minimum_support,
                                                     uint32_t best_view_count, float *scores,
                                                     float *costs, float *guidance,
                                                     uint8_t *tested, unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    unsigned int *evaluations = &evaluation_count;
    const uint32_t V = S.view_count;
    const PatchLayout lay = patch_layout(nb, V, T);
    const PatchLds P = bind_patch(lds, lay);
    float *res = (float *)(lds + lay.total);                 // [nb][V][L]
    float *centre_value = res + (size_t)nb * V * L;          // [nb]
    int *node_tested = (int *)(centre_value + nb);           // [nb]
    const uint32_t nodes = G.columns * G.rows;
    const uint32_t node0 = blockIdx.x * nb;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;

    for (uint32_t n = tid; n < nb; n += nth) {
        uint32_t g = node0 + n;
        uint32_t row = g / G.columns, column = g % G.columns;
        P.pos[n] = make_double2((double)((size_t)column * G.step), (double)((size_t)row * G.step));
        P.node_ok[n] = g < nodes ? 1 : 0;
        node_tested[n] = 0;
    }
    if (tid == 0) evaluation_count = 0;
    __syncthreads();
    patch_prologue(S, P, nb, T, centre_value);

    // The view loop is outermost so that the target plane (and its buffer
    // resource) and camera are uniform across the block.
    for (uint32_t a = 0; a < V; a++) {
        const ViewDev &view = S.views[S.vlist[a]];
        const DevPlane tp = view_plane(view);
        for (uint32_t i = tid; i < nb * L; i += nth) {
            const uint32_t n = i / L, l = i % L;
            float s = __int_as_float(0x7fc00000);
            if (P.node_ok[n] && P.nv_ok[n * V + a]) {
                double dx = 0.0, dy = 0.0;
                bool ok = true;
                if (l > 0) ok = delta_at_depth(P, n, a, V, P.cams[a], labels[l], dx, dy);
                float sc;
                if (ok && patch_zncc(P, n, a, V, T, tp, dx, dy, sc, evaluations)) s = sc;
            }
            res[((size_t)n * V + a) * L + l] = s;
        }
    }
    __syncthreads();

    for (uint32_t i = tid; i < nb * L; i += nth) {
        const uint32_t n = i / L, l = i % L, g = node0 + n;
        if (g >= nodes) continue;
        float score;
        if (!aggregate_scores(res + (size_t)n * V * L + l, L, V, minimum_support, best_view_count,
                              score))
            score = __int_as_float(0x7fc00000);
        float cost;
        if (finite_f(score)) {
            node_tested[n] = 1;
            float near_prior =
                l == 0 ? 0.0f : kNearDepthPrior * (float)(l - 1) / (float)(L - 2);
            cost = rclampf(1.0f - score, 0.0f, 2.0f) + near_prior;
        } else {
            cost = kMissingCost;
        }
        scores[(size_t)g * L + l] = score;
        costs[(size_t)g * L + l] = cost;
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += nth) {
        const uint32_t g = node0 + n;
        if (g >= nodes) continue;
        guidance[g] = centre_value[n];  // reference_guidance (depth.rs:1824-1832)
        tested[g] = (uint8_t)node_tested[n];
    }
    if (tid == 0 && work && evaluation_count) atomicAdd(work, (unsigned long long)evaluation_count);
}

size_t cost_lds_bytes(uint32_t nb, uint32_t views, uint32_t taps, uint32_t label_count) {
    PatchLayout l = patch_layout(nb, views, taps);
    return (size_t)l.total + (size_t)nb * views * label_count * 4 + (size_t)nb * 8;
}

hipError_t launch_cost_volume(hipStream_t stream, const PatchSpec &patch, const GridSpec &grid,
                              const double *labels, uint32_t label_count,
                              uint32_t minimum_support, uint32_t best_view_count,
                              float *scores, float *costs, float *guidance, uint8_t *tested,
                              unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t V = patch.view_count > 0 ? patch.view_count : 1;
    uint32_t nb = 1024 / label_count;
    if (nb < 1) nb = 1;
    if (nb > 8) nb = 8;
    // 32 KB of LDS per 256-thread block keeps 5 blocks (20 waves) per CU
    while (nb > 1 && cost_lds_bytes(nb, V, T, label_count) > 32 * 1024) nb--;
    size_t lds = cost_lds_bytes(nb, V, T, label_count);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const uint32_t nodes = grid.columns * grid.rows;
    dim3 blocks((nodes + nb - 1) / nb);
    PatchSpec s = patch;
    s.view_count = patch.view_count;
    hipLaunchKernelGGL(k_cost_volume, blocks, dim3(256), lds, stream, s, grid, labels, label_count,
                       nb, T, minimum_support, best_view_count, scores, costs, guidance, tested,
                       work);
    return hipGetLastError();
}

// -------------------------------------------------------------------- SGM

constexpr uint32_t kSgmMaxChunks = 8;  // labels <= 512

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

// One wavefront walks one path of one of the 8 families, lanes over labels
// (label = lane + 64 j), writing that family's L_r volume. Family order is
// H->, H<-, V-down, V-up, DR->, DR<-, DL->, DL<- (depth.rs:1287-1309).
__global__ void __launch_bounds__(256) k_sgm_paths(const float *__restrict__ costs,
                                                   const float *__restrict__ guidance, uint32_t C,
                                                   uint32_t R, uint32_t L, uint32_t J,
                                                   float *__restrict__ paths) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nH = R, nV = C, nD = C + R - 1;
    uint32_t d, p;
    if (wave < 2 * nH) {
        d = wave / nH;
        p = wave % nH;
    } else if (wave < 2 * nH + 2 * nV) {
        uint32_t w = wave - 2 * nH;
        d = 2 + w / nV;
        p = w % nV;
    } else if (wave < 2 * nH + 2 * nV + 4 * nD) {
        uint32_t w = wave - 2 * nH - 2 * nV;
        d = 4 + w / nD;
        p = w % nD;
    } else {
        return;
    }
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

    float prev[kSgmMaxChunks];
    int64_t node = start;
#pragma unroll
    for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
        prev[j] = inf;
        if (j < J) {
            uint32_t l = lane + 64 * j;
            if (l < L) {
                prev[j] = costs[(size_t)node * L + l];
                out[(size_t)node * L + l] = prev[j];
            }
        }
    }
    for (uint32_t i = 1; i < len; i++) {
        const int64_t previous_node = node;
        node += stride;
        const float edge = fabsf(guidance[node] - guidance[previous_node]);
        const float edge_scale = 1.0f / (1.0f + 4.0f * edge);
        const float p1 = kSgmSmallPenalty * rmaxf(edge_scale, 0.35f);
        const float p2 = kSgmLargePenalty * rmaxf(edge_scale, 0.12f);
        float m = inf;
#pragma unroll
        for (uint32_t j = 0; j < kSgmMaxChunks; j++)
            if (j < J) m = fminf(m, prev[j]);
        m = wave_min(m);
        const float jump = m + p2;
        float cur[kSgmMaxChunks];
#pragma unroll
        for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
            cur[j] = inf;
            if (j < J) {
                float up = __shfl_up(prev[j], 1, 64);
                float dn = __shfl_down(prev[j], 1, 64);
                float carry_up = j > 0 ? __shfl(prev[j > 0 ? j - 1 : 0], 63, 64) : inf;
                float carry_dn = j + 1 < J ? __shfl(prev[j + 1 < kSgmMaxChunks ? j + 1 : j], 0, 64) : inf;
                if (lane == 0) up = carry_up;
                if (lane == 63) dn = carry_dn;
                const uint32_t l = lane + 64 * j;
                if (l < L) {
                    const float same = prev[j];
                    const float lower = l > 0 ? up + p1 : inf;
                    const float higher = l + 1 < L ? dn + p1 : inf;
                    const float mn = fminf(fminf(fminf(same, lower), higher), jump);
                    cur[j] = costs[(size_t)node * L + l] + mn - m;
                    out[(size_t)node * L + l] = cur[j];
                }
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kSgmMaxChunks; j++) prev[j] = cur[j];
    }
}

hipError_t launch_sgm(hipStream_t stream, const float *costs, const float *guidance,
                      uint32_t columns, uint32_t rows, uint32_t labels, float *paths) {
    const uint32_t J = (labels + 63) / 64;
    if (J > kSgmMaxChunks) return hipErrorInvalidValue;
    const uint64_t waves = 2ull * rows + 2ull * columns + 4ull * (columns + rows - 1);
    dim3 blocks((uint32_t)((waves + 3) / 4));
    hipLaunchKernelGGL(k_sgm_paths, blocks, dim3(256), 0, stream, costs, guidance, columns, rows,
                       labels, J, paths);
    return hipGetLastError();
}

// Broadcast value of label `l` held as chunk l/64 of lane l%64.
__device__ __forceinline__ float label_value(const float *v, uint32_t J, uint32_t l) {
    float out = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
        if (j < J) {
            float x = __shfl(v[j], (int)(l & 63u), 64);
            if (j == (l >> 6)) out = x;
        }
    }
    return out;
}

// select_depths (depth.rs:1385-1467) for one node whose regularised costs
// are held as chunks in the wave; lane 0 writes the node.
__device__ void select_node(const float *s, uint32_t J, uint32_t L, uint32_t lane, size_t node,
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
        // WTA: first minimum over labels 1..L-1 by total_cmp
        int32_t bk = 0x7fffffff;
        uint32_t bl = 0xffffffffu;
#pragma unroll
        for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
            if (j < J) {
                uint32_t l = lane + 64 * j;
                if (l >= 1 && l < L) {
                    int32_t key = tkey32(s[j]);
                    if (key < bk) {
                        bk = key;
                        bl = l;
                    }
                }
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            int32_t ok = __shfl_xor(bk, off, 64);
            uint32_t ol = __shfl_xor(bl, off, 64);
            if (ok < bk || (ok == bk && ol < bl)) {
                bk = ok;
                bl = ol;
            }
        }
        const uint32_t best = bl;
        best_out = (uint16_t)best;
        const float best_cost = label_value(s, J, best);
        const float left = label_value(s, J, best >= 1 ? best - 1 : 0);
        const float right = label_value(s, J, best + 1 < L ? best + 1 : best);
        // competing minimum over labels more than two away
        int32_t ck = 0x7fffffff;
        int have = 0;
#pragma unroll
        for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
            if (j < J) {
                uint32_t l = lane + 64 * j;
                uint32_t diff = l > best ? l - best : best - l;
                if (l >= 1 && l < L && diff > 2) {
                    int32_t key = tkey32(s[j]);
                    if (!have || key < ck) {
                        ck = key;
                        have = 1;
                    }
                }
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            int32_t ok = __shfl_xor(ck, off, 64);
            int oh = __shfl_xor(have, off, 64);
            if (oh && (!have || ok < ck)) {
                ck = ok;
                have = 1;
            }
        }
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
__global__ void __launch_bounds__(256) k_sgm_select(const float *__restrict__ paths,
                                                    const float *__restrict__ scores,
                                                    const uint8_t *__restrict__ tested,
                                                    const double *__restrict__ labels,
                                                    uint32_t nodes, uint32_t L, uint32_t J,
                                                    chd_options o, float *aggregated,
                                                    chd_node_depth *field, uint8_t *fillable,
                                                    uint16_t *best_label) {
    const uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (node >= nodes) return;
    const size_t volume = (size_t)nodes * L;
    float s[kSgmMaxChunks];
#pragma unroll
    for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
        s[j] = 0.0f;
        if (j < J) {
            const uint32_t l = lane + 64 * j;
            if (l < L) {
                const size_t at = (size_t)node * L + l;
                float acc = 0.0f;
#pragma unroll
                for (uint32_t d = 0; d < 8; d++) acc += paths[d * volume + at];
                s[j] = acc;
                if (aggregated) aggregated[at] = acc;
            }
        }
    }
    select_node(s, J, L, lane, node, scores, tested, labels, o, field, fillable, best_label);
}

hipError_t launch_sgm_select(hipStream_t stream, const float *paths, const float *scores,
                             const uint8_t *tested, const double *labels, uint32_t columns,
                             uint32_t rows, uint32_t label_count, const chd_options &options,
                             float *aggregated, chd_node_depth *field, uint8_t *fillable,
                             uint16_t *best_label) {
    const uint32_t J = (label_count + 63) / 64;
    if (J > kSgmMaxChunks) return hipErrorInvalidValue;
    const uint32_t nodes = columns * rows;
    dim3 blocks((nodes + 3) / 4);
    hipLaunchKernelGGL(k_sgm_select, blocks, dim3(256), 0, stream, paths, scores, tested, labels,
                       nodes, label_count, J, options, aggregated, field, fillable, best_label);
    return hipGetLastError();
}

// select_depths alone, from a given regularised volume.
__global__ void __launch_bounds__(256) k_select_depths(const float *__restrict__ regularised,
                                                       const float *__restrict__ scores,
                                                       const uint8_t *__restrict__ tested,
                                                       const double *__restrict__ labels,
                                                       uint32_t nodes, uint32_t L, uint32_t J,
                                                       chd_options o, chd_node_depth *field,
                                                       uint8_t *fillable, uint16_t *best_label) {
    const uint32_t node = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (node >= nodes) return;
    float s[kSgmMaxChunks];
#pragma unroll
    for (uint32_t j = 0; j < kSgmMaxChunks; j++) {
        s[j] = 0.0f;
        if (j < J) {
            const uint32_t l = lane + 64 * j;
            if (l < L) s[j] = regularised[(size_t)node * L + l];
        }
    }
    select_node(s, J, L, lane, node, scores, tested, labels, o, field, fillable, best_label);
}

hipError_t launch_select_depths(hipStream_t stream, const float *regularised,
                                const float *scores, const uint8_t *tested,
                                const double *labels, uint32_t nodes, uint32_t label_count,
                                const chd_options &options, chd_node_depth *field,
                                uint8_t *fillable, uint16_t *best_label) {
    const uint32_t J = (label_count + 63) / 64;
    if (J > kSgmMaxChunks) return hipErrorInvalidValue;
    dim3 blocks((nodes + 3) / 4);
    hipLaunchKernelGGL(k_select_depths, blocks, dim3(256), 0, stream, regularised, scores, tested,
                       labels, nodes, label_count, J, options, field, fillable, best_label);
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

// ------------------------------------------------------------ direct pass

constexpr uint32_t kDirectNodes = 8;     // fine nodes per block (at most)
constexpr uint32_t kMaxCandidates = 32;  // direct_depth_candidates bound

// select_direct_depth (depth.rs:1109-1165) over `count` candidates whose
// aggregated scores are `score[k]` (NaN = None).
__device__ bool select_direct_depth(const double *depths, const float *score, uint32_t count,
                                    bool has_baseline, float baseline, bool seeded,
                                    const chd_options &o, chd_node_depth &out) {
    const double far_inverse = 1.0 / o.far_depth;
    const double inverse_range = 1.0 / o.near_depth - far_inverse;
    uint32_t idx[kMaxCandidates];
    float obj[kMaxCandidates];
    uint32_t n = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (score[i] != score[i]) continue;
        const double near_fraction = rclamp((1.0 / depths[i] - far_inverse) / inverse_range, 0.0, 1.0);
        const float objective = score[i] - kNearDepthPrior * (float)near_fraction;
        uint32_t j = n;  // stable, descending by total_cmp
        while (j > 0 && tkey32(obj[j - 1]) < tkey32(objective)) {
            obj[j] = obj[j - 1];
            idx[j] = idx[j - 1];
            j--;
        }
        obj[j] = objective;
        idx[j] = i;
        n++;
    }
    if (n == 0) return false;
    const uint32_t best_index = idx[0];
    const float best_score = score[best_index];
    const float best_objective = obj[0];
    float competing = best_objective;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t d = idx[k] > best_index ? idx[k] - best_index : best_index - idx[k];
        if (d > 1) {
            competing = obj[k];
            break;
        }
    }
    const float margin = rmaxf(best_objective - competing, 0.0f);
    const float improvement = has_baseline ? best_score - baseline : 0.0f;
    const float score_floor = rmaxf(o.minimum_score - 0.03f, kMinimumRegularizedScore);
    const float required_margin = seeded ? o.minimum_margin * 0.6f : o.minimum_margin;
    const bool improves_global = !has_baseline || improvement >= o.minimum_improvement * 0.5f;
    const bool supported = best_score >= score_floor && improves_global &&
                           (margin >= required_margin || improvement >= o.minimum_improvement);
    if (!supported) return false;
    const float score_confidence =
        rclampf((best_score - score_floor) / rmaxf(1.0f - score_floor, 1.0e-3f), 0.0f, 1.0f);
    const float margin_confidence = rclampf(margin / rmaxf(required_margin, 1.0e-3f), 0.0f, 1.0f);
    out.valid = 1;
    out.depth = depths[best_index];
    out.confidence =
        rclampf(0.30f + 0.45f * score_confidence + 0.25f * margin_confidence, 0.0f, 1.0f);
    out.improvement = improvement;
    out.regularized = 0;
    return true;
}

// best_depth_candidate (depth.rs:1167-1185): last maximum.
__device__ int best_depth_candidate(const double *depths, const float *score, uint32_t count,
                                    const chd_options &o) {
    const double far_inverse = 1.0 / o.far_depth;
    const double inverse_range = 1.0 / o.near_depth - far_inverse;
    int best = -1;
    int32_t best_key = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (score[i] != score[i]) continue;
        const double near_fraction = rclamp((1.0 / depths[i] - far_inverse) / inverse_range, 0.0, 1.0);
        const int32_t key = tkey32(score[i] - kNearDepthPrior * (float)near_fraction);
        if (best < 0 || key >= best_key) {
            best = (int)i;
            best_key = key;
        }
    }
    return best;
}

__device__ __forceinline__ void clear_node(chd_node_depth &n) {
    n.depth = 0.0;
    n.confidence = 0.0f;
    n.improvement = 0.0f;
    n.valid = 0;
    n.regularized = 0;
    for (int k = 0; k < 6; k++) n.pad[k] = 0;
}

// Candidate scoring shared by the direct and wide kernels: for each view,
// items are (node, slot) with slot 0 the global-warp baseline and slot k the
// depth cand[node][k - 1] (base[] is the per-view prefix of 1 + ncand);
// results land in res[node][view][slot].
__device__ void score_slots(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t T,
                            const double *cand, const uint32_t *ncand, const uint32_t *base,
                            uint32_t total, float *res, unsigned int *evaluations) {
    const uint32_t V = S.view_count;
    for (uint32_t a = 0; a < V; a++) {
        const ViewDev &view = S.views[S.vlist[a]];
        const DevPlane tp = view_plane(view);
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            uint32_t n = 0;
            while (n + 1 < nb && base[n + 1] <= i) n++;
            const uint32_t slot = i - base[n];
            float s = __int_as_float(0x7fc00000);
            if (P.node_ok[n] && P.nv_ok[n * V + a]) {
                double dx = 0.0, dy = 0.0;
                bool ok = true;
                if (slot > 0)
                    ok = delta_at_depth(P, n, a, V, P.cams[a], cand[n * kMaxCandidates + slot - 1],
                                        dx, dy);
                float sc;
                if (ok && patch_zncc(P, n, a, V, T, tp, dx, dy, sc, evaluations)) s = sc;
            }
            res[(size_t)n * V * (kMaxCandidates + 1) + (size_t)a * (kMaxCandidates + 1) + slot] = s;
        }
    }
}

__global__ void __launch_bounds__(256) k_direct(PatchSpec S, GridSpec F, GridSpec Cg,
                                                const chd_node_depth *__restrict__ coarse,
                                                double inverse_step, chd_options o, uint32_t T,
                                                uint32_t nb, chd_node_depth *field, float *guidance,
                                                uint8_t *tested, WideEntry *wide,
                                                uint32_t *wide_count, unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    unsigned int *evaluations = &evaluation_count;
    const uint32_t V = S.view_count;
    const PatchLayout lay = patch_layout(nb, V, T);
    const PatchLds P = bind_patch(lds, lay);
    double *cand = (double *)(lds + lay.total);                 // [nb][32]
    float *res = (float *)(cand + nb * kMaxCandidates);         // [nb][V][33]
    float *agg = res + (size_t)nb * V * (kMaxCandidates + 1);   // [nb][33]
    float *centre_value = agg + nb * (kMaxCandidates + 1);      // [nb]
    uint32_t *ncand = (uint32_t *)(centre_value + nb);          // [nb]
    uint32_t *base = ncand + nb;                                // [nb + 1]
    uint32_t *seeded = base + nb + 1;                           // [nb]
    const uint32_t nodes = F.columns * F.rows, node0 = blockIdx.x * nb;
    const uint32_t tid = threadIdx.x;
    const double far_inverse = 1.0 / o.far_depth, near_inverse = 1.0 / o.near_depth;

    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        const uint32_t g = node0 + n;
        const uint32_t row = g / F.columns, column = g % F.columns;
        const double px = (double)((size_t)column * F.step), py = (double)((size_t)row * F.step);
        P.pos[n] = make_double2(px, py);
        P.node_ok[n] = g < nodes ? 1 : 0;
        uint32_t count = 0;
        if (g < nodes) {
            // nearest_coarse_depth (depth.rs:1062-1076)
            const uint32_t cc = sat_u32d(rclamp(round(px / (double)Cg.step), 0.0, (double)(Cg.columns - 1)));
            const uint32_t cr = sat_u32d(rclamp(round(py / (double)Cg.step), 0.0, (double)(Cg.rows - 1)));
            const chd_node_depth seed = coarse[(size_t)cr * Cg.columns + cc];
            seeded[n] = seed.valid;
            double *c = cand + n * kMaxCandidates;
            // direct_depth_candidates (depth.rs:1078-1107)
            if (seed.valid) {
                for (int offset = -6; offset <= 6; offset++) {
                    const double inverse = rclamp(1.0 / seed.depth + (double)offset * inverse_step * 0.5,
                                                  far_inverse, near_inverse);
                    const double depth = 1.0 / inverse;
                    if (count == 0 || fabs(1.0 / c[count - 1] - 1.0 / depth) > inverse_step * 0.1)
                        c[count++] = depth;
                }
            } else {
                // inverse_depth_samples(near, far, 32), reversed
                for (uint32_t k = 0; k < 32; k++) {
                    const uint32_t index = 31 - k;
                    const double w = (double)index / 31.0;
                    c[k] = 1.0 / (near_inverse * (1.0 - w) + far_inverse * w);
                }
                count = 32;
            }
        } else {
            seeded[n] = 0;
        }
        ncand[n] = count;
    }
    if (tid == 0) evaluation_count = 0;
    __syncthreads();
    if (tid == 0) {
        uint32_t b = 0;
        for (uint32_t n = 0; n < nb; n++) {
            base[n] = b;
            b += (node0 + n < nodes) ? 1 + ncand[n] : 0;
        }
        base[nb] = b;
    }
    patch_prologue(S, P, nb, T, centre_value);
    score_slots(S, P, nb, T, cand, ncand, base, base[nb], res, evaluations);
    __syncthreads();
    // aggregate each (node, slot) over the views
    for (uint32_t i = tid; i < nb * (kMaxCandidates + 1); i += blockDim.x) {
        const uint32_t n = i / (kMaxCandidates + 1), slot = i % (kMaxCandidates + 1);
        if (node0 + n >= nodes || slot > ncand[n]) continue;
        float a;
        if (!aggregate_scores(res + (size_t)n * V * (kMaxCandidates + 1) + slot, kMaxCandidates + 1,
                              V, o.minimum_support, o.best_view_count, a))
            a = __int_as_float(0x7fc00000);
        agg[i] = a;
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        const uint32_t g = node0 + n;
        if (g >= nodes) continue;
        guidance[g] = centre_value[n];
        const float *sc = agg + n * (kMaxCandidates + 1);
        const float baseline = sc[0];
        const bool has_baseline = baseline == baseline;
        bool any = false;
        for (uint32_t k = 0; k < ncand[n]; k++) any |= sc[1 + k] == sc[1 + k];
        tested[g] = (uint8_t)(has_baseline || any);
        const double *c = cand + n * kMaxCandidates;
        chd_node_depth selected;
        clear_node(selected);
        const bool ok = select_direct_depth(c, sc + 1, ncand[n], has_baseline, baseline,
                                            seeded[n] != 0, o, selected);
        if (!ok) {
            const int best = best_depth_candidate(c, sc + 1, ncand[n], o);
            if (best >= 0 && sc[1 + best] >= kMinimumRegularizedScore) {
                const uint32_t first = best >= 3 ? (uint32_t)best - 3 : 0;
                const uint32_t last = (uint32_t)best + 3 < ncand[n] - 1 ? (uint32_t)best + 3 : ncand[n] - 1;
                const uint32_t slot = atomicAdd(wide_count, 1u);
                WideEntry e;
                e.node = g;
                e.count = last - first + 1;
                e.seeded = seeded[n];
                e.pad = 0;
                for (uint32_t k = 0; k < 7; k++) e.depth[k] = k < e.count ? c[first + k] : 0.0;
                wide[slot] = e;
            }
        }
        field[g] = selected;
    }
    if (tid == 0 && work && evaluation_count) atomicAdd(work, (unsigned long long)evaluation_count);
}

hipError_t launch_direct(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                         const GridSpec &coarse, const chd_node_depth *coarse_field,
                         double inverse_step, const chd_options &direct_options,
                         chd_node_depth *field, float *guidance, uint8_t *tested,
                         WideEntry *wide, uint32_t *wide_count, unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t V = patch.view_count > 0 ? patch.view_count : 1;
    auto bytes = [&](uint32_t nb) {
        PatchLayout l = patch_layout(nb, V, T);
        return (size_t)l.total + nb * kMaxCandidates * 8 +
               (size_t)nb * V * (kMaxCandidates + 1) * 4 + nb * (kMaxCandidates + 1) * 4 +
               nb * 4 + nb * 4 + (nb + 1) * 4 + nb * 4;
    };
    uint32_t nb = kDirectNodes;
    while (nb > 1 && bytes(nb) > 64 * 1024) nb--;
    const size_t lds = bytes(nb);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const uint32_t nodes = fine.columns * fine.rows;
    hipLaunchKernelGGL(k_direct, dim3((nodes + nb - 1) / nb), dim3(256), lds, stream, patch, fine,
                       coarse, coarse_field, inverse_step, direct_options, T, nb, field, guidance,
                       tested, wide, wide_count, work);
    return hipGetLastError();
}

constexpr uint32_t kWideNodes = 4;  // wide-fallback nodes per block (at most)

// The r = max(2r, 8) re-measurement of nodes whose first pass produced no
// supported depth but a candidate scoring at least 0.30 (depth.rs:803-846).
// A fixed grid walks the entry list.
__global__ void __launch_bounds__(256) k_direct_wide(PatchSpec S, GridSpec F, chd_options o,
                                                     uint32_t T, uint32_t nb,
                                                     const WideEntry *wide,
                                                     const uint32_t *wide_count,
                                                     chd_node_depth *field,
                                                     unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    unsigned int *evaluations = &evaluation_count;
    if (threadIdx.x == 0) evaluation_count = 0;
    const uint32_t V = S.view_count;
    const PatchLayout lay = patch_layout(nb, V, T);
    const PatchLds P = bind_patch(lds, lay);
    double *cand = (double *)(lds + lay.total);                 // [nb][32]
    float *res = (float *)(cand + nb * kMaxCandidates);         // [nb][V][33]
    float *agg = res + (size_t)nb * V * (kMaxCandidates + 1);   // [nb][33]
    float *centre_value = agg + nb * (kMaxCandidates + 1);      // [nb]
    uint32_t *ncand = (uint32_t *)(centre_value + nb);          // [nb]
    uint32_t *base = ncand + nb;                                // [nb + 1]
    const uint32_t tid = threadIdx.x;
    const uint32_t count = *wide_count;
    for (uint32_t first = blockIdx.x * nb; first < count; first += gridDim.x * nb) {
        __syncthreads();
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t e = first + n;
            if (e < count) {
                const WideEntry w = wide[e];
                const uint32_t row = w.node / F.columns, column = w.node % F.columns;
                P.pos[n] = make_double2((double)((size_t)column * F.step), (double)((size_t)row * F.step));
                P.node_ok[n] = 1;
                ncand[n] = w.count;
                for (uint32_t k = 0; k < 7; k++) cand[n * kMaxCandidates + k] = w.depth[k];
            } else {
                P.pos[n] = make_double2(0.0, 0.0);
                P.node_ok[n] = 0;
                ncand[n] = 0;
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t b = 0;
            for (uint32_t n = 0; n < nb; n++) {
                base[n] = b;
                b += (first + n < count) ? 1 + ncand[n] : 0;
            }
            base[nb] = b;
        }
        patch_prologue(S, P, nb, T, centre_value);
        score_slots(S, P, nb, T, cand, ncand, base, base[nb], res, evaluations);
        __syncthreads();
        for (uint32_t i = tid; i < nb * 8; i += blockDim.x) {
            const uint32_t n = i / 8, slot = i % 8;
            if (first + n >= count || slot > ncand[n]) continue;
            float a;
            if (!aggregate_scores(res + (size_t)n * V * (kMaxCandidates + 1) + slot,
                                  kMaxCandidates + 1, V, o.minimum_support, o.best_view_count, a))
                a = __int_as_float(0x7fc00000);
            agg[n * (kMaxCandidates + 1) + slot] = a;
        }
        __syncthreads();
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t e = first + n;
            if (e >= count) continue;
            const WideEntry w = wide[e];
            const float *sc = agg + n * (kMaxCandidates + 1);
            chd_node_depth selected;
            clear_node(selected);
            select_direct_depth(cand + n * kMaxCandidates, sc + 1, ncand[n], sc[0] == sc[0], sc[0],
                                w.seeded != 0, o, selected);
            field[w.node] = selected;
        }
    }
    __syncthreads();
    if (tid == 0 && work && evaluation_count) atomicAdd(work, (unsigned long long)evaluation_count);
}

hipError_t launch_direct_wide(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                              const chd_options &wide_options, const WideEntry *wide,
                              const uint32_t *wide_count, uint32_t wide_capacity,
                              chd_node_depth *field, unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t V = patch.view_count > 0 ? patch.view_count : 1;
    auto bytes = [&](uint32_t nb) {
        PatchLayout l = patch_layout(nb, V, T);
        return (size_t)l.total + nb * kMaxCandidates * 8 +
               (size_t)nb * V * (kMaxCandidates + 1) * 4 + nb * (kMaxCandidates + 1) * 4 +
               nb * 4 + nb * 4 + (nb + 1) * 4;
    };
    uint32_t nb = kWideNodes;
    while (nb > 1 && bytes(nb) > 96 * 1024) nb--;
    const size_t lds = bytes(nb);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    uint32_t blocks = (wide_capacity + nb - 1) / nb;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_direct_wide, dim3(blocks), dim3(256), lds, stream, patch, fine,
                       wide_options, T, nb, wide, wide_count, field, work);
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

// Each node starts under its smallest-index neighbour it is related to (or
// itself), which already links most of a surface into short chains.
__global__ void k_cc_init(const chd_node_depth *source, const float *guidance, uint32_t C,
                          uint32_t R, double tol, uint32_t *parent, uint32_t *sizes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C * R) return;
    sizes[i] = 0;
    uint32_t p = i;
    if (source[i].valid) {
        const int row = (int)(i / C), column = (int)(i % C);
        const int nx[4] = {column - 1, column, column + 1, column - 1};
        const int ny[4] = {row - 1, row - 1, row - 1, row};
        for (int k = 0; k < 4; k++) {  // neighbours of smaller index, ascending
            if (nx[k] < 0 || ny[k] < 0 || nx[k] >= (int)C) continue;
            const uint32_t j = (uint32_t)ny[k] * C + (uint32_t)nx[k];
            if (source[j].valid && cc_edge(source, guidance, i, j, tol)) {
                p = j;
                break;
            }
        }
    }
    parent[i] = p;
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
    hipLaunchKernelGGL(k_cc_init, grid, block, 0, stream, source, guidance, columns, rows,
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

// ------------------------------------------------------ per-target stage

constexpr uint32_t kTargetNodes = 16;

// refine_one_view (depth.rs:1634-1708) and the decision block
// (depth.rs:429-486) for one target view (S.vlist[0]) over a tile of fine
// nodes. Phase 1 scores the five shared-depth offsets and the global
// baseline, phase 2 the eight +-1.5 px residuals around the best depth.
__global__ void __launch_bounds__(128) k_targets(PatchSpec S, GridSpec F,
                                                 const chd_node_depth *__restrict__ field,
                                                 double inverse_step, chd_options o, uint32_t T,
                                                 chd_node_warp *decisions,
                                                 unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluations;
    const uint32_t nb = kTargetNodes;
    const PatchLayout lay = patch_layout(nb, 1, T);
    const PatchLds P = bind_patch(lds, lay);
    float2 *points = (float2 *)(lds + lay.total);       // [nb][5]
    float *phase = (float *)(points + nb * 5);          // [nb][8]
    float *centre_value = phase + nb * 8;               // [nb]
    int *state = (int *)(centre_value + nb);            // [nb] 0 skip, 1 refine, -1 failed
    float2 *gp = (float2 *)(state + nb + (nb & 1));     // [nb] global point
    float2 *best_point = gp + nb;                       // [nb]
    float *best_score = (float *)(best_point + nb);     // [nb]
    float *baseline = best_score + nb;                  // [nb]
    const uint32_t nodes = F.columns * F.rows, node0 = blockIdx.x * nb, tid = threadIdx.x;
    const ViewDev &view = S.views[S.vlist[0]];
    const DevWarp gw = view_warp(view);
    const DevPlane tp = view_plane(view);
    const double min_inverse = 1.0 / o.far_depth, max_inverse = 1.0 / o.near_depth;

    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        const uint32_t g = node0 + n;
        const uint32_t row = g / F.columns, column = g % F.columns;
        const double px = (double)((size_t)column * F.step), py = (double)((size_t)row * F.step);
        P.pos[n] = make_double2(px, py);
        P.node_ok[n] = g < nodes ? 1 : 0;
        state[n] = 0;
        if (g < nodes) {
            chd_node_warp d;
            d.code = CHD_WARP_UNDEFINED;
            d.global[0] = d.global[1] = 0.0f;
            d.point[0] = d.point[1] = 0.0f;
            d.confidence = 0.0f;
            d.measured = 0;
            float2 q;
            if (warp_map(gw, (float)px, (float)py, q)) {
                d.global[0] = q.x;
                d.global[1] = q.y;
                gp[n] = q;
                d.code = CHD_WARP_GLOBAL;
                if (field[g].valid) state[n] = 1;
            }
            decisions[g] = d;
        }
    }
    if (tid == 0) evaluations = 0;
    __syncthreads();
    patch_prologue(S, P, nb, T, centre_value);
    // the five depth-offset points (map_at_depth); any failure ends the
    // refinement (the `?` of depth.rs:1654-1660)
    for (uint32_t i = tid; i < nb * 5; i += blockDim.x) {
        const uint32_t n = i / 5, k = i % 5;
        if (state[n] != 1) continue;
        const double offsets[5] = {-1.0, -0.5, 0.0, 0.5, 1.0};
        const double shared_inverse = 1.0 / field[node0 + n].depth;
        const double inverse =
            rclamp(shared_inverse + offsets[k] * inverse_step, min_inverse, max_inverse);
        const double depth = 1.0 / inverse;
        bool ok = P.ray_ok[n] != 0;
        double qx = 0.0, qy = 0.0;
        if (ok) {
            const double *r = P.ray + (size_t)n * 6;
            d3 point{r[0] + r[3] * depth, r[1] + r[4] * depth, r[2] + r[5] * depth};
            ok = project(P.cams[0], point, qx, qy) && finite_d(qx) && finite_d(qy);
        }
        points[i] = make_float2((float)qx, (float)qy);
        if (!ok) state[n] = -1;
    }
    __syncthreads();
    // phase 1: five offsets + baseline
    for (uint32_t i = tid; i < nb * 6; i += blockDim.x) {
        const uint32_t n = i / 6, k = i % 6;
        if (state[n] == 0) continue;
        float s = __int_as_float(0x7fc00000);
        if (k == 5) {
            // score_one_view(.., None, ..): delta 0
            float sc;
            if (P.node_ok[n] && P.nv_ok[n] && patch_zncc(P, n, 0, 1, T, tp, 0.0, 0.0, sc, &evaluations)) s = sc;
            baseline[n] = s;
        } else if (state[n] == 1) {
            // score_one_view_at_point (depth.rs:1710-1736)
            const float2 q = points[n * 5 + k];
            const double dx = (double)(q.x - gp[n].x), dy = (double)(q.y - gp[n].y);
            float sc;
            if (!(fabs(dx) > 128.0 || fabs(dy) > 128.0) && P.node_ok[n] && P.nv_ok[n] &&
                patch_zncc(P, n, 0, 1, T, tp, dx, dy, sc, &evaluations))
                s = sc;
            phase[n * 8 + k] = s;
        }
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        if (state[n] != 1) continue;
        const float offsets[5] = {1.0f, 0.5f, 0.0f, 0.5f, 1.0f};
        bool have = false, failed = false;
        float bo = 0.0f, bs = 0.0f;
        float2 bp = make_float2(0.0f, 0.0f);
        for (uint32_t k = 0; k < 5; k++) {
            const float s = phase[n * 8 + k];
            if (s != s) {
                failed = true;
                break;
            }
            const float objective = s - offsets[k] * 0.002f;
            if (!have || objective > bo) {
                have = true;
                bo = objective;
                bs = s;
                bp = points[n * 5 + k];
            }
        }
        if (failed) {
            state[n] = -1;
            continue;
        }
        best_score[n] = bs;
        best_point[n] = bp;
    }
    __syncthreads();
    // phase 2: residual offsets, dy-major
    for (uint32_t i = tid; i < nb * 8; i += blockDim.x) {
        const uint32_t n = i / 8, k = i % 8;
        if (state[n] != 1) continue;
        const float steps[3] = {-1.5f, 0.0f, 1.5f};
        const uint32_t cell = k < 4 ? k : k + 1;
        const float dyo = steps[cell / 3], dxo = steps[cell % 3];
        const float2 point = make_float2(best_point[n].x + dxo, best_point[n].y + dyo);
        const double dx = (double)(point.x - gp[n].x), dy = (double)(point.y - gp[n].y);
        float s = __int_as_float(0x7fc00000), sc;
        if (!(fabs(dx) > 128.0 || fabs(dy) > 128.0) && P.node_ok[n] && P.nv_ok[n] &&
            patch_zncc(P, n, 0, 1, T, tp, dx, dy, sc, &evaluations))
            s = sc;
        phase[n * 8 + k] = s;
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        const uint32_t g = node0 + n;
        if (state[n] == 0 || g >= nodes) continue;
        bool has_selected = false;
        float selected_score = 0.0f;
        float2 selected_point = make_float2(0.0f, 0.0f);
        if (state[n] == 1) {
            const float steps[3] = {-1.5f, 0.0f, 1.5f};
            float b_obj = best_score[n], b_score = best_score[n];
            float2 b_point = best_point[n];
            bool failed = false;
            for (uint32_t k = 0; k < 8; k++) {
                const float s = phase[n * 8 + k];
                if (s != s) {
                    failed = true;
                    break;
                }
                const uint32_t cell = k < 4 ? k : k + 1;
                const float dyo = steps[cell / 3], dxo = steps[cell % 3];
                const float residual_sq = dxo * dxo + dyo * dyo;
                const float objective = s - residual_sq * 0.002f;
                if (objective > b_obj) {
                    b_obj = objective;
                    b_score = s;
                    b_point = make_float2(best_point[n].x + dxo, best_point[n].y + dyo);
                }
            }
            if (!failed) {
                has_selected = true;
                selected_score = b_score;
                selected_point = b_point;
            }
        }
        const chd_node_depth node = field[g];
        const float base = baseline[n];
        const bool has_base = base == base;
        const bool supported = has_selected && selected_score >= o.minimum_score &&
                               (!has_base || selected_score + 0.01f >= base);
        const bool contradicted = has_selected && has_base && selected_score + 0.12f < base;
        chd_node_warp d = decisions[g];
        if (contradicted) {
            d.code = CHD_WARP_OCCLUDED;
            d.measured = node.regularized ? 0u : 1u;
        } else if (supported) {
            d.code = CHD_WARP_REFINED;
            d.point[0] = selected_point.x;
            d.point[1] = selected_point.y;
            d.confidence = node.regularized ? node.confidence * 0.6f : 0.5f + 0.5f * node.confidence;
            d.measured = node.regularized ? 0u : 1u;
        }
        decisions[g] = d;
    }
    if (tid == 0 && work && evaluations) atomicAdd(work, (unsigned long long)evaluations);
}

hipError_t launch_targets(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                          const chd_node_depth *field, double inverse_step,
                          const chd_options &options, chd_node_warp *decisions,
                          unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t nb = kTargetNodes;
    PatchLayout l = patch_layout(nb, 1, T);
    size_t lds = (size_t)l.total + nb * 5 * 8 + nb * 8 * 4 + nb * 4 + (nb + (nb & 1)) * 4 +
                 nb * 8 + nb * 8 + nb * 4 + nb * 4;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const uint32_t nodes = fine.columns * fine.rows;
    hipLaunchKernelGGL(k_targets, dim3((nodes + nb - 1) / nb), dim3(128), lds, stream, patch,
                       fine, field, inverse_step, options, T, decisions, work);
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

// Report inputs (depth.rs:394-418): counts and total_cmp sort keys of the
// selected depths and of the measured nodes' improvements.
// Wavefront-aggregated: one atomic per wave for each count and key list.
__device__ __forceinline__ uint32_t wave_append(bool take, uint32_t *counter) {
    const unsigned long long mask = __ballot(take);
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __builtin_ctzll(mask | (1ull << 63));
    uint32_t base = 0;
    if (mask && (int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__global__ void k_field_stats(const chd_node_depth *field, const uint8_t *tested, uint32_t n,
                              unsigned long long *counts, unsigned long long *depth_keys,
                              unsigned int *improvement_keys, uint32_t *key_counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool in = i < n;
    chd_node_depth d;
    clear_node(d);
    if (in) d = field[i];
    const bool t = in && tested[i];
    const bool measured = d.valid && !d.regularized, regularized = d.valid && d.regularized;
    const unsigned long long mt = __ballot(t), mm = __ballot(measured), mr = __ballot(regularized);
    if (lane == 0) {
        if (mt) atomicAdd(&counts[0], (unsigned long long)__popcll(mt));
        if (mm) atomicAdd(&counts[1], (unsigned long long)__popcll(mm));
        if (mr) atomicAdd(&counts[2], (unsigned long long)__popcll(mr));
    }
    const uint32_t slot = wave_append(d.valid, &key_counts[0]);
    if (d.valid) depth_keys[slot] = (unsigned long long)tkey64(d.depth) ^ 0x8000000000000000ull;
    const uint32_t s2 = wave_append(measured, &key_counts[1]);
    if (measured) improvement_keys[s2] = (unsigned int)tkey32(d.improvement) ^ 0x80000000u;
}

hipError_t launch_field_stats(hipStream_t stream, const chd_node_depth *field,
                              const uint8_t *tested, uint32_t nodes,
                              unsigned long long *counts, unsigned long long *depth_keys,
                              unsigned int *improvement_keys, uint32_t *key_counts) {
    hipLaunchKernelGGL(k_field_stats, dim3((nodes + 255) / 256), dim3(256), 0, stream, field,
                       tested, nodes, counts, depth_keys, improvement_keys, key_counts);
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

hipError_t launch_pad_plane(hipStream_t stream, float *plane, uint32_t width, uint32_t height) {
    const uint32_t n = (width + 1 > height ? width + 1 : height);
    hipLaunchKernelGGL(k_pad_plane, dim3((n + 255) / 256), dim3(256), 0, stream, plane, width,
                       height);
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
