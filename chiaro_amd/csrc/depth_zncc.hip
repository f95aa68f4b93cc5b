// depth_zncc.hip - the ZNCC kernels of the depth hot path (gfx950).
//
// Stage map (reference lines in crates/chiaro-fusion/src/depth.rs):
//   k_cost_volume      build_cost_volume              :1187-1277
//   k_direct_score     measure_direct_depths, scoring :752-802
//   k_direct_select    select_direct_depth + fallback trigger :796-810, 1109-1185
//   k_direct_wide      the r = 8 fallback             :803-846
//   k_targets          per-target refinement and decision :420-487, 1634-1736
//
// All of them evaluate warped_patch_zncc (depth.rs:1772-1822) through one
// block-level patch context in LDS, built in two parts:
//  * per node (patch_nodes): the reference-tap samples, bilateral weights w
//    and w * r, and the label- and view-independent sums sum(w), sum(w r),
//    sum(w r^2) accumulated in tap order (the reference keeps each of its
//    five sums in its own accumulator, so hoisting these three is exact);
//  * per view: the centre's global point and the label-independent half of
//    map_at_depth (the corrected reference ray), computed for a group of up
//    to kRayViews views at once (one thread per node and view, so the long
//    f64 chains run side by side), and, rebuilt inside the view loop
//    (patch_view), the global-warp tap positions with their min/max, which
//    decides Plane::sample's bounds test for all taps at once because
//    ((f64(g) + delta) - 0.5) * 0.5 -> f32 is monotone in g. A group of
//    lanes per node computes the taps and reduces the min/max (exact in any
//    order) with lane shuffles.
// A work item is one (node, hypothesis) ZNCC for the current view, looping
// the taps in the reference's dy-major order. With the view loop outermost
// the target plane, its buffer resource and the camera are uniform across
// the block.
#include <hip/hip_runtime.h>

#include "depth_device.hpp"
#include "depth_internal.hpp"

namespace chd {

constexpr uint32_t kMaxCandidates = 32;  // direct_depth_candidates bound (depth.rs:1101-1105)

// ---------------------------------------------------------- patch context

struct PatchLayout {
    uint32_t cams, pos, ray, ext, gc, gq, wr, sums, centre, spatial, node_ok, nv_ok, gc_ok,
        ray_ok, loff, lat, total;
    uint32_t ray_views;  // views per ray group: the per-view arrays hold [ray_views][nb]
};

// Shared tap lattice of a block whose nodes lie on one grid row, one grid
// step apart, with an even step: every tap of every node is a point of one
// lattice of 2-px spacing, (nb - 1) * step / 2 + 2r + 1 points wide and
// 2r + 1 high, so the reference samples and the global-warp tap positions
// are computed once per lattice point instead of once per (node, tap) (a
// strip of 16 fine nodes at a 4 px step has 351 points against 1,296 taps).
// Each point is computed with the same operations as the per-node form
// (raster position p.x + 2 dx is the same exact integer either way), so the
// contexts are identical.
__host__ __device__ inline uint32_t lattice_width(uint32_t nb, uint32_t step, uint32_t radius) {
    return (nb - 1) * (step / 2) + 2 * radius + 1;
}

// Views whose centre points and corrected rays are computed together.
constexpr uint32_t kRayViews = 4;

__host__ __device__ inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

// f64_taps: the per-view tap positions are kept as f64 (16 B per tap, two
// conversions fewer per tap in the loop) rather than f32 (8 B, half the LDS).
// lat_points: points of the block's shared tap lattice (0: no lattice).
__host__ __device__ inline PatchLayout patch_layout(uint32_t nb, uint32_t v, uint32_t t,
                                                    bool f64_taps, uint32_t lat_points = 0) {
    PatchLayout l;
    const uint32_t rv = v < kRayViews ? (v > 0 ? v : 1u) : kRayViews;
    l.ray_views = rv;
    uint32_t o = 0;
    l.cams = o; o = align16(o + (uint32_t)sizeof(chd_camera));  // the current view's
    l.pos = o; o = align16(o + nb * 16u);
    l.ray = o; o = align16(o + rv * nb * 6u * 8u);
    l.ext = o; o = align16(o + nb * 16u);
    l.gc = o; o = align16(o + rv * nb * 8u);
    // first holds the reference samples; a lattice replaces both
    l.gq = o; o = align16(o + (lat_points ? 0u : nb * t * (f64_taps ? 16u : 8u)));
    l.wr = o; o = align16(o + nb * t * 8u);
    l.sums = o; o = align16(o + nb * 3u * 4u);
    l.centre = o; o = align16(o + nb * 4u);
    l.spatial = o; o = align16(o + t * 4u);
    l.node_ok = o; o = align16(o + nb * 4u);
    l.nv_ok = o; o = align16(o + rv * nb * 4u);
    l.gc_ok = o; o = align16(o + rv * nb * 4u);
    l.ray_ok = o; o = align16(o + rv * nb * 4u);
    l.loff = o; o = align16(o + (lat_points ? t * 2u : 0u));
    l.lat = o; o = align16(o + lat_points * 8u);
    l.total = o;
    return l;
}

struct PatchLds {
    chd_camera *cam;   // camera of the current view (copied by patch_view)
    double2 *pos;      // raster position of each node
    // ray, gc, nv_ok, gc_ok, ray_ok: [ray group slot][node] in the layout;
    // patch_view returns a copy pointing at the current view's slot
    double *ray;       // [node][6]: corrected reference ray (current view)
    float4 *ext;       // [node]: min x, max x, min y, max y warp tap (current view)
    float2 *gc;        // [node]: global point of the centre (current view)
    double2 *gq;       // [node][tap]: global-warp tap positions (current view), f32
                       // values held as f64 (f64_taps layouts)
    float2 *gqf;       // the same positions as f32 (other layouts; aliases gq)
    float *rv;         // [node][tap]: reference samples (aliases gq until the views)
    bool f64_taps;
    float2 *wr;        // [node][tap]: w, w * r
    float *sums;       // [node][3]: sum w, sum w r, sum w r^2
    float *centre;
    float *spatial;    // [tap]
    int *node_ok;      // every reference sample of the patch exists
    int *nv_ok;        // current view: global warp defined at the centre and every tap
    int *gc_ok;        // current view: global warp defined at the centre
    int *ray_ok;       // current view: corrected reference ray exists
    uint32_t ray_views;
    // shared tap lattice (lat_w = 0: none): points [2r + 1][lat_w], node n's
    // taps start at column n * lat_cols; loff[t] = (t / (2r + 1)) * lat_w +
    // t % (2r + 1). Holds the reference samples (x) and their validity (y)
    // in patch_nodes, then each view's global-warp positions (NaN where the
    // warp is undefined; a defined warp point is never NaN).
    float2 *lat;
    uint16_t *loff;
    uint32_t lat_w, lat_cols;
    uint32_t tap_width;  // 2r + 1, for the row-wise tap loop (set by with_rows / with_lattice)
};

__device__ inline PatchLds bind_patch(unsigned char *base, const PatchLayout &l, bool f64_taps) {
    PatchLds p;
    p.f64_taps = f64_taps;
    p.gqf = (float2 *)(base + l.gq);
    p.cam = (chd_camera *)(base + l.cams);
    p.pos = (double2 *)(base + l.pos);
    p.ray = (double *)(base + l.ray);
    p.ext = (float4 *)(base + l.ext);
    p.gc = (float2 *)(base + l.gc);
    p.gq = (double2 *)(base + l.gq);
    p.rv = (float *)(base + l.gq);
    p.wr = (float2 *)(base + l.wr);
    p.sums = (float *)(base + l.sums);
    p.centre = (float *)(base + l.centre);
    p.spatial = (float *)(base + l.spatial);
    p.node_ok = (int *)(base + l.node_ok);
    p.nv_ok = (int *)(base + l.nv_ok);
    p.gc_ok = (int *)(base + l.gc_ok);
    p.ray_ok = (int *)(base + l.ray_ok);
    p.ray_views = l.ray_views;
    p.lat = (float2 *)(base + l.lat);
    p.loff = (uint16_t *)(base + l.loff);
    p.lat_w = 0;
    p.lat_cols = 0;
    p.tap_width = 0;
    return p;
}

// A context whose ZNCCs run the row-wise tap loop (tap positions of row dy
// at a stride: the node's own T taps, or a shared lattice).
__device__ inline PatchLds with_rows(PatchLds p, uint32_t radius) {
    p.tap_width = 2 * radius + 1;
    return p;
}

// The lattice form of a patch context for nodes 0 .. nb - 1 at raster
// columns x0 + n * step of one row (step even); bind_patch's layout must
// hold lattice_width(nb, step, radius) * (2 radius + 1) points.
__device__ inline PatchLds with_lattice(PatchLds p, uint32_t nb, uint32_t step, uint32_t radius) {
    p.lat_w = lattice_width(nb, step, radius);
    p.lat_cols = step / 2;
    p.tap_width = 2 * radius + 1;
    return p;
}

// The row-major plane, or with prefer_columns its column-major copy where
// the view has one. Lanes run over depth hypotheses of a node (and, in the
// direct and target kernels, over a few neighbouring nodes), whose samples
// follow the epipolar line: the copy keeps them in fewer cache lines when
// that line is mostly vertical (measured faster in all three kernels).
__device__ __forceinline__ DevPlane view_plane(const ViewDev &v, bool prefer_columns) {
    DevPlane p;
    p.width = v.plane_width;
    p.height = v.plane_height;
    // a target view's quad array is built from the plane this picks
    p.quad = prefer_columns ? v.quad : nullptr;
    if (prefer_columns && v.columns) {
        p.data = v.columns;
        p.pitch = v.plane_height + 1;
        p.lines = v.plane_width + 1;
        p.transposed = 1;
    } else {
        p.data = v.plane;
        p.pitch = v.plane_pitch;
        p.lines = v.plane_height + 1;
        p.transposed = 0;
    }
    return p;
}

__device__ __forceinline__ DevWarp view_warp(const ViewDev &v) {
    DevWarp w;
    w.points = v.warp;
    w.step = v.warp_step;
    w.columns = v.warp_columns;
    w.rows = v.warp_rows;
    w.inverse_step = w.step != 0 && (w.step & (w.step - 1)) == 0 ? 1.0f / (float)w.step : 0.0f;
    return w;
}

// Centre global point and corrected reference ray of nodes 0 .. nb-1 for the
// views first .. first + count - 1 (count <= ray_views), one (node, view)
// per thread (the block's last threads, which have the fewest of the
// accompanying sample items), into ray-group slots 0 .. count - 1; nv_ok
// starts as gc_ok (patch_view clears it where a tap's warp is undefined).
// Reads the cameras from global memory (the LDS copy may still be in flight).
__device__ void patch_rays(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t first,
                           uint32_t count, uint32_t n_begin = 0, uint32_t n_end = ~0u) {
    const chd_camera &reference = *S.views[S.reference].camera;
    if (n_end > nb) n_end = nb;
    const uint32_t m = n_end > n_begin ? n_end - n_begin : 0;
    for (uint32_t i = blockDim.x - 1 - threadIdx.x; i < m * count; i += blockDim.x) {
        const uint32_t slot = i / m, n = n_begin + i % m, k = slot * nb + n;
        const ViewDev &view = S.views[S.vlist[first + slot]];
        const double2 p = P.pos[n];
        float2 gc;
        if (!warp_map(view_warp(view), (float)p.x, (float)p.y, gc)) {
            P.nv_ok[k] = 0;
            P.gc_ok[k] = 0;
            P.ray_ok[k] = 0;
            continue;
        }
        P.nv_ok[k] = 1;
        P.gc_ok[k] = 1;
        P.gc[k] = gc;
        // corrected_reference = reference.map_from(target, gc, 1e8)
        // (depth.rs:1752-1756), then the reference ray through it
        d3 o, d;
        pixel_to_ray(*view.camera, (double)gc.x, (double)gc.y, o, d);
        const d3 far{o.x + d.x * kInfinityDepth, o.y + d.y * kInfinityDepth,
                     o.z + d.z * kInfinityDepth};
        double cx, cy;
        if (!project(reference, far, cx, cy)) {
            P.ray_ok[k] = 0;
            continue;
        }
        P.ray_ok[k] = 1;
        pixel_to_ray(reference, cx, cy, o, d);
        double *ray = P.ray + (size_t)k * 6;
        ray[0] = o.x;
        ray[1] = o.y;
        ray[2] = o.z;
        ray[3] = d.x;
        ray[4] = d.y;
        ray[5] = d.z;
    }
}

// Reference sample of tap t of node n (patch_nodes has run or is past its
// sampling step): from the lattice, or the per-node array.
__device__ __forceinline__ float ref_sample(const PatchLds &P, uint32_t n, uint32_t t, uint32_t T) {
    return P.lat_w ? P.lat[P.loff[t] + n * P.lat_cols].x : P.rv[n * T + t];
}

// Node half of the patch context. On entry P.pos[n] and P.node_ok[n] (node
// exists) are set and visible. On exit node_ok[n] is 0 when the centre or any
// tap of the reference patch falls outside the reference plane (every ZNCC of
// the node is then None), centre_value[n] holds reference_guidance
// (depth.rs:1824-1832), and the first ray group (views 0 .. ray_views - 1)
// is computed.
__device__ void patch_nodes(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t T,
                            float *centre_value) {
    const uint32_t V = S.view_count, tid = threadIdx.x, nth = blockDim.x;
    const int r = (int)S.radius;
    const uint32_t width = 2u * S.radius + 1u;
    const DevPlane rp = view_plane(S.views[S.reference], false);
    const float sigma = rmaxf((float)S.radius * 0.75f, 1.0f);
    if (V > 0) patch_rays(S, P, nb, 0, V < P.ray_views ? V : P.ray_views);
    for (uint32_t t = tid; t < T; t += nth) {
        const int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
        const float distance_sq = (float)(dx * dx + dy * dy);
        P.spatial[t] = libm_expf(-distance_sq / (2.0f * sigma * sigma));
    }
    for (uint32_t n = tid; n < nb; n += nth) {
        const double2 p = P.pos[n];
        float c;
        if (plane_sample(rp, to_plane(p.x), to_plane(p.y), c)) {
            P.centre[n] = c;
            centre_value[n] = c;
        } else {
            centre_value[n] = 0.0f;
            P.node_ok[n] = 0;
        }
    }
    if (P.lat_w) {
        // the reference samples once per lattice point, then per (node, tap)
        const uint32_t W = P.lat_w;
        const double2 p0 = P.pos[0];
        for (uint32_t t = tid; t < T; t += nth)
            P.loff[t] = (uint16_t)((t / width) * W + t % width);
        for (uint32_t i = tid; i < W * width; i += nth) {
            const uint32_t col = i % W, row = i / W;
            float v;
            const bool ok = plane_sample(rp, to_plane(p0.x + (double)((int)col - r) * 2.0),
                                         to_plane(p0.y + (double)((int)row - r) * 2.0), v);
            P.lat[i] = make_float2(ok ? v : 0.0f, ok ? 1.0f : 0.0f);
        }
        __syncthreads();
        for (uint32_t i = tid; i < nb * T; i += nth) {
            const uint32_t n = i / T, t = i % T;
            if (P.lat[P.loff[t] + n * P.lat_cols].y == 0.0f) P.node_ok[n] = 0;
        }
    } else {
#pragma unroll 3
        for (uint32_t i = tid; i < nb * T; i += nth) {
            const uint32_t n = i / T, t = i % T;
            const double2 p = P.pos[n];
            const int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
            float v;
            if (!plane_sample(rp, to_plane(p.x + (double)dx * 2.0),
                              to_plane(p.y + (double)dy * 2.0), v)) {
                P.node_ok[n] = 0;
                v = 0.0f;
            }
            P.rv[i] = v;
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < nb * T; i += nth) {
        const uint32_t n = i / T, t = i % T;
        if (!P.node_ok[n]) continue;
        const float rv = ref_sample(P, n, t, T);
        const float range = libm_expf(-1.2f * fabsf(rv - P.centre[n]));
        const float w = P.spatial[t] * range;
        P.wr[i] = make_float2(w, w * rv);
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += nth) {
        if (!P.node_ok[n]) continue;
        float count = 0.0f, sum_r = 0.0f, sum_rr = 0.0f;
        for (uint32_t t = 0; t < T; t++) {
            const float2 w = P.wr[n * T + t];
            count += w.x;
            sum_r += w.y;
            sum_rr += w.y * ref_sample(P, n, t, T);
        }
        P.sums[n * 3 + 0] = count;
        P.sums[n * 3 + 1] = sum_r;
        P.sums[n * 3 + 2] = sum_rr;
    }
    __syncthreads();
}

// View half of the patch context for view slot `a`; returns the context with
// the per-view pointers at a's ray-group slot. Must be entered by the whole
// block after every read of the previous view's context has finished. The
// first ray group comes from patch_nodes unless first_group_ready is false
// (a second pass over the views). With `live`, the block first checks
// whether any node of n_begin .. n_end - 1 has its reference patch and the
// view's global warp at its centre; if none has, *live is false, every ZNCC
// of the view is None for these nodes and the tap context is not built (the
// narrow modules of a rig see a small part of the reference). `wanted`
// (optional, per node) narrows the check to the nodes a caller will score.
__device__ PatchLds patch_view(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t T,
                               uint32_t a, bool first_group_ready = true, uint32_t n_begin = 0,
                               uint32_t n_end = ~0u, bool *live = nullptr,
                               const int *wanted = nullptr) {
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const int r = (int)S.radius;
    const uint32_t width = 2u * S.radius + 1u;
    const uint32_t V = S.view_count, R = P.ray_views, slot = a % R;
    if (n_end > nb) n_end = nb;
    if (slot == 0 && (a > 0 || !first_group_ready)) {
        patch_rays(S, P, nb, a, V - a < R ? V - a : R, n_begin, n_end);
        __syncthreads();
    }
    PatchLds Q = P;
    Q.ray = P.ray + (size_t)slot * nb * 6;
    Q.gc = P.gc + slot * nb;
    Q.nv_ok = P.nv_ok + slot * nb;
    Q.gc_ok = P.gc_ok + slot * nb;
    Q.ray_ok = P.ray_ok + slot * nb;
    if (live) {
        int any = 0;
        for (uint32_t n = n_begin + tid; n < n_end; n += nth)
            any |= P.node_ok[n] && Q.nv_ok[n] && (!wanted || wanted[n]) ? 1 : 0;
        *live = __syncthreads_or(any) != 0;
        if (!*live) return Q;
    }
    const DevWarp gw = view_warp(S.views[S.vlist[a]]);
    {  // the view's camera into LDS (the previous view's readers are done)
        constexpr uint32_t words = sizeof(chd_camera) / 4;
        const uint32_t *src = (const uint32_t *)S.views[S.vlist[a]].camera;
        for (uint32_t i = tid; i < words; i += nth) ((uint32_t *)P.cam)[i] = src[i];
    }
    // G lanes per node (a power of two dividing the wavefront) compute the
    // node's tap positions and reduce their min/max; only nodes n_begin ..
    // n_end - 1 (the others' contexts are not read until rebuilt)
    const uint32_t m = n_end > n_begin ? n_end - n_begin : 0;
    const uint32_t G = m <= 4 ? 64u : m <= 8 ? 32u : 16u;
    const uint32_t per_pass = nth / G, lane = tid % G;
    if (P.lat_w) {
        // the global warp once per lattice point (NaN where undefined)
        const uint32_t W = P.lat_w;
        const double2 p0 = P.pos[0];
        const float nan = __int_as_float(0x7fc00000);
        for (uint32_t i = tid; i < W * width; i += nth) {
            const uint32_t col = i % W, row = i / W;
            float2 q;
            if (!warp_map(gw, (float)(p0.x + (double)((int)col - r) * 2.0),
                          (float)(p0.y + (double)((int)row - r) * 2.0), q))
                q = make_float2(nan, nan);
            P.lat[i] = q;
        }
        __syncthreads();
    }
    for (uint32_t n0 = n_begin; n0 < n_end; n0 += per_pass) {
        const uint32_t n = n0 + tid / G;
        float x0 = __int_as_float(0x7f800000), y0 = x0;
        float x1 = -x0, y1 = -x0;
        if (n < n_end) {
            const double2 p = P.pos[n];
#pragma unroll 3
            for (uint32_t t = lane; t < T; t += G) {
                float2 q;
                if (P.lat_w) {
                    q = P.lat[P.loff[t] + n * P.lat_cols];
                    if (q.x != q.x) {
                        Q.nv_ok[n] = 0;
                        q = make_float2(0.0f, 0.0f);
                    }
                } else {
                    const int dx = (int)(t % width) - r, dy = (int)(t / width) - r;
                    if (!warp_map(gw, (float)(p.x + (double)dx * 2.0),
                                  (float)(p.y + (double)dy * 2.0), q)) {
                        Q.nv_ok[n] = 0;
                        q = make_float2(0.0f, 0.0f);
                    }
                    // the tap loop reads a lattice in place
                    const uint32_t i = n * T + t;
                    if (P.f64_taps)
                        P.gq[i] = make_double2((double)q.x, (double)q.y);
                    else
                        P.gqf[i] = q;
                }
                x0 = fminf(x0, q.x);
                x1 = fmaxf(x1, q.x);
                y0 = fminf(y0, q.y);
                y1 = fmaxf(y1, q.y);
            }
        }
        for (uint32_t m = G >> 1; m > 0; m >>= 1) {
            x0 = fminf(x0, __shfl_xor(x0, (int)m, (int)G));
            x1 = fmaxf(x1, __shfl_xor(x1, (int)m, (int)G));
            y0 = fminf(y0, __shfl_xor(y0, (int)m, (int)G));
            y1 = fmaxf(y1, __shfl_xor(y1, (int)m, (int)G));
        }
        if (n < n_end && lane == 0) P.ext[n] = make_float4(x0, x1, y0, y1);
    }
    __syncthreads();
    return Q;
}

// One batch of kTapBatch taps from t0 of the three target sums of
// warped_patch_zncc: all nine quad loads are issued before the first
// interpolation, then the sums run in tap order. Tail batches (t0 + kTapBatch
// > T) re-read the last tap for the missing ones and drop them; full batches
// need no clamp, so their LDS reads take constant offsets from one address.
constexpr uint32_t kTapBatch = 9;
template <bool Transposed, bool F64Taps, bool Tail>
__device__ __forceinline__ void tap_batch(const float2 *wr, const double2 *gq, const float2 *gqf,
                                          uint32_t t0, uint32_t T, uint32_t pitch,
                                          __amdgpu_buffer_rsrc_t rsrc,
                                          double dx, double dy, float &st, float &stt, float &sp) {
    PendingSample ps[kTapBatch];
#pragma unroll
    for (uint32_t k = 0; k < kTapBatch; k++) {
        const uint32_t t = Tail ? (t0 + k < T ? t0 + k : T - 1) : t0 + k;
        double qx, qy;
        if constexpr (F64Taps) {
            const double2 q = gq[t];
            qx = q.x;
            qy = q.y;
        } else {
            const float2 q = gqf[t];
            qx = (double)q.x;
            qy = (double)q.y;
        }
        ps[k] = plane_sample_issue<Transposed>(rsrc, pitch, to_plane(qx + dx),
                                               to_plane(qy + dy));
    }
    // keep the batch's loads ahead of its interpolations (the scheduler
    // would otherwise sink each load next to its use)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (uint32_t k = 0; k < kTapBatch; k++) {
        if (Tail && t0 + k >= T) break;
        const float2 w = wr[t0 + k];
        const float tv = plane_sample_finish<Transposed>(ps[k]);
        const float wv = w.x * tv;
        st += wv;
        stt += wv * tv;
        sp += w.y * tv;
    }
}

// The three target sums of warped_patch_zncc over the taps in the
// reference's dy-major order, each in its own f32 accumulator.
// Rows (f32 tap positions only): the taps row by row, tap positions of
// consecutive rows at the lattice's stride (or the node's own, 2r + 1);
// the same taps in the same order.
template <bool Transposed, bool F64Taps, bool Rows = false>
__device__ __forceinline__ void tap_sums(const PatchLds &P, uint32_t n, uint32_t T, uint32_t pitch,
                                         __amdgpu_buffer_rsrc_t rsrc,
                                         double dx, double dy, float &st, float &stt, float &sp) {
    const float2 *wr = P.wr + (size_t)n * T;
    if constexpr (Rows) {
        static_assert(!F64Taps, "row-wise taps are f32");
        const uint32_t width = P.tap_width;
        const float2 *g = P.lat_w ? P.lat + n * P.lat_cols : P.gqf + (size_t)n * T;
        const uint32_t gs = P.lat_w ? P.lat_w : width;
        for (uint32_t row = 0; row < width; row++, wr += width, g += gs) {
            uint32_t t0 = 0;
            for (; t0 + kTapBatch <= width; t0 += kTapBatch)
                tap_batch<Transposed, false, false>(wr, nullptr, g, t0, width, pitch, rsrc,
                                                    dx, dy, st, stt, sp);
            if (t0 < width)
                tap_batch<Transposed, false, true>(wr, nullptr, g, t0, width, pitch, rsrc,
                                                   dx, dy, st, stt, sp);
        }
        return;
    }
    const double2 *gq = P.gq + (size_t)n * T;
    const float2 *gqf = P.gqf + (size_t)n * T;
    uint32_t t0 = 0;
    for (; t0 + kTapBatch <= T; t0 += kTapBatch)
        tap_batch<Transposed, F64Taps, false>(wr, gq, gqf, t0, T, pitch, rsrc, dx, dy,
                                              st, stt, sp);
    if (t0 < T)
        tap_batch<Transposed, F64Taps, true>(wr, gq, gqf, t0, T, pitch, rsrc, dx, dy,
                                             st, stt, sp);
}

// warped_patch_zncc (depth.rs:1772-1822) in two halves. The bounds half:
// every tap's bilinear footprint lies inside the plane (the extremes of the
// warped taps decide it, since to_plane is monotone).
__device__ __forceinline__ bool zncc_inside(const PatchLds &P, uint32_t n, const DevPlane &tp,
                                            double dx, double dy) {
    const float4 e = P.ext[n];
    const float xmin = to_plane((double)e.x + dx), xmax = to_plane((double)e.y + dx);
    const float ymin = to_plane((double)e.z + dy), ymax = to_plane((double)e.w + dy);
    return xmin >= 0.0f && ymin >= 0.0f && xmax <= (float)(tp.width - 1) &&
           ymax <= (float)(tp.height - 1);
}

// The sums half, for a displacement zncc_inside accepted.
template <bool F64Taps, bool Rows = false>
__device__ __forceinline__ bool zncc_taps(const PatchLds &P, uint32_t n, uint32_t T,
                                          const DevPlane &tp, __amdgpu_buffer_rsrc_t rsrc,
                                          double dx, double dy,
                                          float &score) {
    float st = 0.0f, stt = 0.0f, sp = 0.0f;
    if (tp.transposed)
        tap_sums<true, F64Taps, Rows>(P, n, T, tp.pitch, rsrc, dx, dy, st, stt, sp);
    else
        tap_sums<false, F64Taps, Rows>(P, n, T, tp.pitch, rsrc, dx, dy, st, stt, sp);
    const float count = P.sums[n * 3 + 0], sr = P.sums[n * 3 + 1], srr = P.sums[n * 3 + 2];
    const float covariance = sp - sr * st / count;
    const float reference_energy = srr - sr * sr / count;
    const float target_energy = stt - st * st / count;
    const float denominator = sqrtf(rmaxf(reference_energy, 0.0f) * rmaxf(target_energy, 0.0f));
    if (!(denominator > 1.0e-6f)) return false;
    score = rclampf(covariance / denominator, -1.0f, 1.0f);
    return true;
}

// warped_patch_zncc (depth.rs:1772-1822) of node n against the current view
// displaced by the raster delta (dx, dy); false is the reference's None.
// Evaluations count the displacements that pass the bounds check.
// F64Taps must match the layout's f64_taps.
template <bool F64Taps, bool Rows = false>
__device__ __forceinline__ bool patch_zncc(const PatchLds &P, uint32_t n, uint32_t T,
                                           const DevPlane &tp, __amdgpu_buffer_rsrc_t rsrc,
                                           double dx, double dy, float &score,
                                           unsigned int *evaluations) {
    if (!zncc_inside(P, n, tp, dx, dy)) return false;
    // one LDS atomic per wavefront: the first active lane adds the count of
    // the lanes that passed (per-lane atomics on one address serialise)
    const uint64_t passed = __ballot(1);
    if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)passed) - 1))
        atomicAdd(evaluations, (unsigned int)__popcll(passed));
    return zncc_taps<F64Taps, Rows>(P, n, T, tp, rsrc, dx, dy, score);
}

// map_at_depth (depth.rs:1738-1759) from the cached corrected ray of node n,
// then score_one_view's 128 px parallax gate (depth.rs:1604-1614).
__device__ __forceinline__ bool delta_at_depth(const PatchLds &P, uint32_t n,
                                               const chd_camera &target, double depth, double &dx,
                                               double &dy) {
    if (!P.ray_ok[n]) return false;
    const double *r = P.ray + (size_t)n * 6;
    const d3 point{r[0] + r[3] * depth, r[1] + r[4] * depth, r[2] + r[5] * depth};
    double qx, qy;
    if (!project(target, point, qx, qy)) return false;
    if (!(finite_d(qx) && finite_d(qy))) return false;
    const float2 gc = P.gc[n];
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
            const float s = res[a * stride];
            if ((used >> a) & 1u || s != s) continue;
            const int32_t key = tkey32(s);
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

// aggregate over the views as a running top-K, for kernels whose view loop
// is outermost: slot[0] holds the count of scored views (as bits), slot[1..K]
// the K largest scores so far in descending total_cmp order (an equal score
// stays behind the earlier view's, as the reference's first-maximum pick
// has it; equal keys are equal bits, so the order among them cannot change
// the sum). With K = min(best_view_count, V) the sum below takes the same
// values in the same order as aggregate_scores.
__device__ __forceinline__ void topk_insert(float *slot, uint32_t K, float s) {
    if (s != s) return;
    const uint32_t c = __float_as_uint(slot[0]);
    const int32_t key = tkey32(s);
    float *v = slot + 1;
    uint32_t j = c < K ? c : K;
    while (j > 0 && tkey32(v[j - 1]) < key) {
        if (j < K) v[j] = v[j - 1];
        j--;
    }
    if (j < K) v[j] = s;
    slot[0] = __uint_as_float(c + 1);
}

__device__ __forceinline__ bool topk_aggregate(const float *slot, uint32_t minimum_support,
                                               uint32_t best_view_count, float &out) {
    const uint32_t count = __float_as_uint(slot[0]);
    if (count < minimum_support) return false;
    const uint32_t k = best_view_count < count ? best_view_count : count;
    if (k < minimum_support) return false;
    float sum = -0.0f;
    for (uint32_t i = 0; i < k; i++) sum += slot[1 + i];
    out = sum / (float)k;
    return true;
}

__host__ __device__ __forceinline__ uint32_t topk_slots(uint32_t best_view_count, uint32_t V) {
    const uint32_t k = best_view_count < V ? best_view_count : V;
    return 1 + (k > 0 ? k : 1);
}

__device__ __forceinline__ float nan_f() { return __int_as_float(0x7fc00000); }

// Row-aligned node tiles: block b takes `count` consecutive nodes of one
// grid row, from node `first` (rows * row_tiles(G, nb) blocks; a row's
// tiles differ by at most one node and hold at most nb), so that the
// block's taps share one lattice.
struct NodeTile {
    uint32_t first, count;
};
__host__ __device__ __forceinline__ uint32_t row_tiles(const GridSpec &G, uint32_t nb) {
    return (G.columns + nb - 1) / nb;
}
__device__ __forceinline__ NodeTile row_tile(const GridSpec &G, uint32_t nb, uint32_t b) {
    const uint32_t per_row = row_tiles(G, nb);
    const uint32_t row = b / per_row, k = b % per_row;
    const uint32_t c0 = (uint32_t)((uint64_t)k * G.columns / per_row);
    const uint32_t c1 = (uint32_t)((uint64_t)(k + 1) * G.columns / per_row);
    return NodeTile{row * G.columns + c0, c1 - c0};
}
// Lattice points of a row tile's context (0 for odd steps: no lattice).
__host__ __device__ __forceinline__ uint32_t tile_lattice_points(uint32_t nb, uint32_t step,
                                                                 uint32_t radius) {
    return step % 2 == 0 ? lattice_width(nb, step, radius) * (2 * radius + 1) : 0u;
}

__device__ __forceinline__ double2 node_position(const GridSpec &G, uint32_t g) {
    const uint32_t row = g / G.columns, column = g % G.columns;
    return make_double2((double)((size_t)column * G.step), (double)((size_t)row * G.step));
}

// ------------------------------------------------------------ cost volume

// build_cost_volume (depth.rs:1187-1277) for one work item of job J: the
// nodes of tile `tile` (J.nb nodes) and the labels of slab `slab` (J.slab
// labels). Every score is computed as in the single launch; a slab only
// limits which labels this item writes, so any tiling and slab choice gives
// the same volume. `tested` must be cleared beforehand: an item sets it to 1
// for the nodes with a finite score among its labels.
// TopK: the views' scores of an item are kept as a running top-K (when K <
// V this takes less LDS); else all V are kept and aggregate_scores picks.
template <bool TopK>
__device__ void cost_tile(const CostJob &J, uint32_t tile, uint32_t slab, unsigned char *lds,
                          unsigned int *evaluation_count) {
    const PatchSpec &S = J.S;
    const uint32_t V = S.view_count, T = J.T, L = J.L, nb = J.nb;
    const uint32_t l0 = slab * J.slab, Ls = min(J.slab, L - l0);
    const PatchLayout lay = patch_layout(nb, V, T, true);
    const PatchLds P = bind_patch(lds, lay, true);
    const uint32_t KS = TopK ? topk_slots(J.best_view_count, V) : V;
    float *res = (float *)(lds + lay.total);  // [nb * Ls][KS] running top-K, or [V][nb * Ls]
    float *centre_value = res + (size_t)nb * Ls * KS;  // [nb]
    int *node_tested = (int *)(centre_value + nb);    // [nb]
    const uint32_t nodes = J.G.columns * J.G.rows, node0 = tile * nb;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;

    for (uint32_t n = tid; n < nb; n += nth) {
        P.pos[n] = node_position(J.G, node0 + n);
        P.node_ok[n] = node0 + n < nodes ? 1 : 0;
        node_tested[n] = 0;
    }
    if (tid == 0) *evaluation_count = 0;
    // each item's slot is only touched by the thread that owns the item
    if (TopK)
        for (uint32_t i = tid; i < nb * Ls; i += nth) res[(size_t)i * KS] = __uint_as_float(0u);
    __syncthreads();
    patch_nodes(S, P, nb, T, centre_value);

    for (uint32_t a = 0; a < V; a++) {
        bool live;
        const PatchLds Q = patch_view(S, P, nb, T, a, true, 0, ~0u, &live);
        if (!live) {
            // every score of the view is None: the top-K is unchanged
            for (uint32_t i = tid; i < nb * Ls; i += nth) {
                const uint32_t n = i / Ls, l = l0 + i % Ls;
                if (!TopK) res[(size_t)a * nb * Ls + i] = nan_f();
                if (J.view_scores && node0 + n < nodes)
                    J.view_scores[((size_t)a * nodes + node0 + n) * L + l] = nan_f();
            }
            continue;
        }
        const DevPlane tp = view_plane(S.views[S.vlist[a]], true);
        const __amdgpu_buffer_rsrc_t rsrc = plane_rsrc(tp);
        for (uint32_t i = tid; i < nb * Ls; i += nth) {
            const uint32_t n = i / Ls, l = l0 + i % Ls;
            float s = nan_f();
            if (Q.node_ok[n] && Q.nv_ok[n]) {
                double dx = 0.0, dy = 0.0;
                bool ok = true;
                if (l > 0) ok = delta_at_depth(Q, n, *Q.cam, J.labels[l], dx, dy);
                float sc;
                if (ok && patch_zncc<true>(Q, n, T, tp, rsrc, dx, dy, sc,
                                           evaluation_count))
                    s = sc;
            }
            if (TopK)
                topk_insert(res + (size_t)i * KS, KS - 1, s);
            else
                res[(size_t)a * nb * Ls + i] = s;
            if (J.view_scores && node0 + n < nodes)
                J.view_scores[((size_t)a * nodes + node0 + n) * L + l] = s;
        }
        __syncthreads();
    }

    for (uint32_t i = tid; i < nb * Ls; i += nth) {
        const uint32_t n = i / Ls, l = l0 + i % Ls, g = node0 + n;
        if (g >= nodes) continue;
        float score;
        const bool scored =
            TopK ? topk_aggregate(res + (size_t)i * KS, J.minimum_support, J.best_view_count, score)
                 : aggregate_scores(res + i, nb * Ls, V, J.minimum_support, J.best_view_count,
                                    score);
        if (!scored) score = nan_f();
        float cost;
        if (finite_f(score)) {
            node_tested[n] = 1;
            const float near_prior =
                l == 0 ? 0.0f : kNearDepthPrior * (float)(l - 1) / (float)(L - 2);
            cost = rclampf(1.0f - score, 0.0f, 2.0f) + near_prior;
        } else {
            cost = kMissingCost;
        }
        J.scores[(size_t)g * L + l] = score;
        J.costs[(size_t)g * L + l] = cost;
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += nth) {
        const uint32_t g = node0 + n;
        if (g >= nodes) continue;
        J.guidance[g] = centre_value[n];  // every slab writes the same value
        if (node_tested[n]) J.tested[g] = 1;
    }
    if (tid == 0 && J.work && *evaluation_count)
        atomicAdd(J.work, (unsigned long long)*evaluation_count);
}

// The contiguous run of `total` items that XCD x takes in XCD order: runs
// of total / 8 items, the first total % 8 of them one longer.
__device__ __forceinline__ void xcd_run(uint32_t x, uint32_t total, uint32_t &start,
                                        uint32_t &count) {
    const uint32_t q = total / kXcds, r = total % kXcds;
    start = x * q + min(x, r);
    count = q + (x < r ? 1u : 0u);
}

// The band form of cost_tile: each node sweeps label 0 (the global warp)
// and its window lo .. hi only (window_lo/hi; a full window is 1 .. L - 1),
// as proposed by the learned band proposal. Scores and costs of the labels
// outside the window are written as NaN and +inf, so the SGM that follows
// never selects them; inside the window every value is the full sweep's.
// One slab; items are (node, label) over the nodes' windows. The tap
// positions are held as f64, as in the full sweep (two conversions fewer
// per tap than f32 positions: 3-5 % less time per ZNCC on config 4 bands;
// a tile lattice shared by the nodes, as in the direct pass, was slower
// here, where nodes are 8 px apart).
template <bool TopK>
__device__ void cost_tile_band(const CostJob &J, uint32_t tile, unsigned char *lds,
                               unsigned int *evaluation_count) {
    const PatchSpec &S = J.S;
    const uint32_t V = S.view_count, T = J.T, L = J.L, nb = J.nb;
    const PatchLayout lay = patch_layout(nb, V, T, true);
    const PatchLds P = bind_patch(lds, lay, true);
    const uint32_t KS = TopK ? topk_slots(J.best_view_count, V) : V;
    const uint32_t C = J.slab;                        // items per pass over the views
    float *res = (float *)(lds + lay.total);          // [C][KS] running top-K, or [V][C]
    float *centre_value = res + (size_t)C * KS;       // [nb]
    int *node_tested = (int *)(centre_value + nb);    // [nb]
    uint32_t *base = (uint32_t *)(node_tested + nb);  // [nb + 1]
    uint32_t *lo = base + nb + 1;                     // [nb]
    uint8_t *item_node = (uint8_t *)(lo + nb);        // [nb * L]
    const uint32_t nodes = J.G.columns * J.G.rows, node0 = tile * nb;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;

    for (uint32_t n = tid; n < nb; n += nth) {
        const uint32_t g = node0 + n;
        P.pos[n] = node_position(J.G, g);
        P.node_ok[n] = g < nodes ? 1 : 0;
        node_tested[n] = 0;
        lo[n] = g < nodes ? J.window_lo[g] : 1;
    }
    if (tid < 64) {
        // wavefront 0: exclusive prefix of the nodes' item counts (nb <= 64)
        const uint32_t g = node0 + tid;
        const uint32_t c = tid < nb && g < nodes ? 2u + J.window_hi[g] - J.window_lo[g] : 0u;
        uint32_t incl = c;
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, 64);
            if (tid >= o) incl += v;
        }
        if (tid < nb) base[tid] = incl - c;
        if (tid + 1 == nb) base[nb] = incl;
        if (tid == 0) *evaluation_count = 0;
    }
    __syncthreads();
    patch_nodes(S, P, nb, T, centre_value);
    for (uint32_t n = tid; n < nb; n += nth)
        for (uint32_t k = base[n]; k < base[n + 1]; k++) item_node[k] = (uint8_t)n;
    __syncthreads();
    const uint32_t items = base[nb];

    // passes of at most C items (one for banded tiles; wide or fallback
    // windows take several, each rebuilding the view contexts). A slot is
    // only touched by the thread that owns its item.
    for (uint32_t c0 = 0; c0 < items; c0 += C) {
        const uint32_t cn = min(C, items - c0);
        if (TopK)
            for (uint32_t i = tid; i < cn; i += nth) res[(size_t)i * KS] = __uint_as_float(0u);
        for (uint32_t a = 0; a < V; a++) {
            bool live;
            const PatchLds Q = patch_view(S, P, nb, T, a, c0 == 0, item_node[c0],
                                          item_node[c0 + cn - 1] + 1u, &live);
            if (!live) {
                // every score of the view is None: the top-K is unchanged
                if (!TopK)
                    for (uint32_t i = tid; i < cn; i += nth) res[(size_t)a * C + i] = nan_f();
                continue;
            }
            const DevPlane tp = view_plane(S.views[S.vlist[a]], true);
            const __amdgpu_buffer_rsrc_t rsrc = plane_rsrc(tp);
            for (uint32_t i = tid; i < cn; i += nth) {
                const uint32_t it = c0 + i, n = item_node[it], j = it - base[n];
                const uint32_t l = j == 0 ? 0 : lo[n] + j - 1;
                float s = nan_f();
                if (Q.node_ok[n] && Q.nv_ok[n]) {
                    double dx = 0.0, dy = 0.0;
                    bool ok = true;
                    if (l > 0) ok = delta_at_depth(Q, n, *Q.cam, J.labels[l], dx, dy);
                    float sc;
                    if (ok && patch_zncc<true>(Q, n, T, tp, rsrc, dx, dy, sc,
                                              evaluation_count))
                        s = sc;
                }
                if (TopK)
                    topk_insert(res + (size_t)i * KS, KS - 1, s);
                else
                    res[(size_t)a * C + i] = s;
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < cn; i += nth) {
            const uint32_t it = c0 + i, n = item_node[it], j = it - base[n], g = node0 + n;
            const uint32_t l = j == 0 ? 0 : lo[n] + j - 1;
            float score;
            const bool scored =
                TopK ? topk_aggregate(res + (size_t)i * KS, J.minimum_support, J.best_view_count,
                                      score)
                     : aggregate_scores(res + i, C, V, J.minimum_support, J.best_view_count,
                                        score);
            if (!scored) score = nan_f();
            float cost;
            if (finite_f(score)) {
                node_tested[n] = 1;
                const float near_prior =
                    l == 0 ? 0.0f : kNearDepthPrior * (float)(l - 1) / (float)(L - 2);
                cost = rclampf(1.0f - score, 0.0f, 2.0f) + near_prior;
            } else {
                cost = kMissingCost;
            }
            J.scores[(size_t)g * L + l] = score;
            J.costs[(size_t)g * L + l] = cost;
        }
    }
    // labels outside the windows: score None, cost +inf
    for (uint32_t i = tid; i < nb * L; i += nth) {
        const uint32_t n = i / L, l = i % L, g = node0 + n;
        if (g >= nodes || l == 0 || (l >= lo[n] && l < lo[n] + (base[n + 1] - base[n]) - 1))
            continue;
        J.scores[(size_t)g * L + l] = nan_f();
        J.costs[(size_t)g * L + l] = __int_as_float(0x7f800000);
    }
    __syncthreads();
    for (uint32_t n = tid; n < nb; n += nth) {
        const uint32_t g = node0 + n;
        if (g >= nodes) continue;
        J.guidance[g] = centre_value[n];
        if (node_tested[n]) J.tested[g] = 1;
    }
    if (tid == 0 && J.work && *evaluation_count)
        atomicAdd(J.work, (unsigned long long)*evaluation_count);
}

// One launch over one job: block b is item b (tile item / slabs, slab item %
// slabs), or with XCD order item k of XCD b % 8's contiguous run, where k =
// b / 8 (the hardware hands workgroups to the XCDs round robin, so each L2
// then serves neighbouring tiles).
// At most four waves per SIMD fit anyway (LDS and the 256-thread blocks), so
// the tap loop may spend up to 128 VGPRs keeping a step's 18 plane loads in
// flight (with a higher occupancy target the compiler serialises them).
template <bool TopK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_cost_volume(CostJob J) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    uint32_t item = blockIdx.x;
    if (J.xcd_order == 1) {
        uint32_t start, count;
        xcd_run(blockIdx.x % kXcds, gridDim.x, start, count);
        item = start + blockIdx.x / kXcds;
    }
    cost_tile<TopK>(J, item / J.slabs, item % J.slabs, lds, &evaluation_count);
}

// The band sweep's launch (its own kernel, so that the full sweep's register
// allocation does not carry the band's bookkeeping).
template <bool TopK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_cost_volume_band(CostJob J) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    uint32_t item = blockIdx.x;
    if (J.xcd_order == 1) {
        uint32_t start, count;
        xcd_run(blockIdx.x % kXcds, gridDim.x, start, count);
        item = start + blockIdx.x / kXcds;
    }
    cost_tile_band<TopK>(J, item, lds, &evaluation_count);
}

// Persistent form over several jobs (captures): each block takes the next
// (job, tile, slab) item from a queue counter until the queue is empty, so
// captures of uneven cost share the chip without a tail per capture. `first`
// holds each job's first item index (njobs + 1 entries); `counter` must be 0
// at launch.
// The queue is split into 8 sequences, one per XCD, each with its own
// counter: by default XCD x takes items x, x + 8, x + 16, ..., the order in
// which the hardware hands a plain launch's workgroups to the XCDs (measured
// faster than one shared counter); with xcd_order XCD x takes its contiguous
// run instead. A block drains its own XCD's sequence, then the others' in
// turn.
// Mode: the kind of every job (0 full, 1 full with top-K, 2 band, 3 band
// with top-K), or kQueueMixed to decide per job (one kernel then carries the
// register needs of all four).
constexpr int kQueueMixed = 4;
template <int Mode>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_cost_volume_queue(const CostJob *__restrict__ jobs, const uint32_t *__restrict__ first,
                    uint32_t njobs, uint32_t *counters, uint32_t xcd_order) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    __shared__ uint32_t next;
    const uint32_t total = first[njobs];
    const uint32_t own = blockIdx.x % kXcds;
    uint32_t run = 0;  // sequences tried so far
    uint32_t round = 0;
    for (;;) {
        __syncthreads();  // the previous item has finished with `next` and LDS
        if (threadIdx.x == 0 && xcd_order == 2) {
            next = blockIdx.x + round * gridDim.x;  // static: no counter
            round++;
        } else if (threadIdx.x == 0) {
            next = total;
            while (run < kXcds) {
                const uint32_t x = (own + run) % kXcds;
                uint32_t start, count;
                if (xcd_order) {
                    xcd_run(x, total, start, count);
                } else {
                    start = x;
                    count = total > x ? (total - x + kXcds - 1) / kXcds : 0;
                }
                const uint32_t k = atomicAdd(counters + x, 1u);
                if (k < count) {
                    next = xcd_order ? start + k : start + k * kXcds;
                    break;
                }
                run++;
            }
        }
        __syncthreads();
        const uint32_t item = next;
        if (item >= total) break;
        uint32_t j = 0;
        while (j + 1 < njobs && first[j + 1] <= item) j++;
        const CostJob J = jobs[j];
        const uint32_t local = item - first[j];
        const int mode = Mode == kQueueMixed ? (J.window_lo ? 2 : 0) + (J.topk ? 1 : 0) : Mode;
        if ((Mode == kQueueMixed || Mode == 3) && mode == 3)
            cost_tile_band<true>(J, local, lds, &evaluation_count);
        else if ((Mode == kQueueMixed || Mode == 2) && mode == 2)
            cost_tile_band<false>(J, local, lds, &evaluation_count);
        else if ((Mode == kQueueMixed || Mode == 1) && mode == 1)
            cost_tile<true>(J, local / J.slabs, local % J.slabs, lds, &evaluation_count);
        else if (Mode == kQueueMixed || Mode == 0)
            cost_tile<false>(J, local / J.slabs, local % J.slabs, lds, &evaluation_count);
    }
}

// slab: labels per item, or for band items the item capacity of one pass
// (band items also hold base, lo and item_node for up to L labels a node)
size_t cost_lds_bytes(uint32_t nb, uint32_t views, uint32_t taps, uint32_t slab,
                      uint32_t best_view_count, bool band = false, uint32_t labels = 0) {
    const PatchLayout l = patch_layout(nb, views, taps, true);
    const uint32_t k = topk_slots(best_view_count, views);
    const size_t slots = (size_t)(band ? slab : nb * slab) * (k < views ? k : views);
    size_t bytes = (size_t)l.total + slots * 4 + (size_t)nb * 8;
    if (band) bytes += (size_t)(2 * nb + 1) * 4 + (size_t)nb * labels;
    return bytes;
}

// Dynamic LDS per block for four blocks per CU (the occupancy the tap loop's
// registers allow): a quarter of the CU's 160 KB less room for the kernels'
// static variables, which share the block's allocation (272 B in the cost
// and direct kernels, 784 B with the target kernel's item list). A block a
// few bytes over the quarter runs three to a CU: the band sweep at 40,704 B
// of dynamic LDS held 11.5 waves per CU against the full sweep's 15.5 and
// took 18 % longer for its items.
constexpr size_t kQuarterLds = 40 * 1024 - 1024;

// Default shape: about 1024 (node, label) units per block, within a quarter
// of the CU's LDS, one slab.
void cost_job_shape(CostJob &J, uint32_t nodes_per_block, uint32_t label_slab) {
    const uint32_t V = J.S.view_count;
    J.T = (2 * J.S.radius + 1) * (2 * J.S.radius + 1);
    J.slab = label_slab == 0 || label_slab > J.L ? J.L : label_slab;
    J.topk = topk_slots(J.best_view_count, V) < V ? 1u : 0u;
    uint32_t nb = nodes_per_block;
    if (J.window_lo) {
        // band windows hold 18 .. 66 labels: about 512 items per block (two
        // rounds of the 256 threads) at the widest window, 4 .. 14 nodes
        // (measured on config 4: 14 nodes best at +-8 and +-16, 6 at +-32 and
        // for fallback tiles); within a quarter of the CU's LDS (four blocks
        // per CU) with room for +-16 windows in one pass, the rest of the
        // quarter holding more items per pass (fallback tiles take several
        // passes)
        const uint32_t KS = J.topk ? topk_slots(J.best_view_count, V) : V;
        if (nb == 0) {
            nb = J.window_max ? 512 / J.window_max : 14;
            nb = nb < 4 ? 4 : nb > 14 ? 14 : nb;
            while (nb > 1 &&
                   cost_lds_bytes(nb, V, J.T, nb * 34, J.best_view_count, true, J.L) > kQuarterLds)
                nb--;
        }
        const size_t fixed = cost_lds_bytes(nb, V, J.T, 0, J.best_view_count, true, J.L);
        size_t cap = fixed < kQuarterLds ? (kQuarterLds - fixed) / ((size_t)KS * 4) : 0;
        if (cap > (size_t)nb * J.L) cap = (size_t)nb * J.L;
        J.slab = cap > 64 ? (uint32_t)cap : 64u;
        J.nb = nb;
        J.tiles = (J.G.columns * J.G.rows + nb - 1) / nb;
        J.slabs = 1;
        return;
    } else if (nb == 0) {
        nb = 1024 / J.slab;
        if (nb < 1) nb = 1;
        if (nb > 16) nb = 16;
        while (nb > 1 && cost_lds_bytes(nb, V, J.T, J.slab, J.best_view_count) > kQuarterLds) nb--;
    }
    J.nb = nb;
    const uint32_t nodes = J.G.columns * J.G.rows;
    J.tiles = (nodes + nb - 1) / nb;
    J.slabs = (J.L + J.slab - 1) / J.slab;
}

size_t cost_job_lds(const CostJob &J) {
    return cost_lds_bytes(J.nb, J.S.view_count, J.T, J.slab, J.best_view_count,
                          J.window_lo != nullptr, J.L);
}

hipError_t launch_cost_volume(hipStream_t stream, const CostJob &job) {
    const size_t lds = cost_job_lds(job);
    if (lds > 160 * 1024 || job.nb == 0) return hipErrorInvalidValue;
    if (job.window_lo && job.topk)
        hipLaunchKernelGGL(k_cost_volume_band<true>, dim3(job.tiles), dim3(256), lds, stream, job);
    else if (job.window_lo)
        hipLaunchKernelGGL(k_cost_volume_band<false>, dim3(job.tiles), dim3(256), lds, stream, job);
    else if (job.topk)
        hipLaunchKernelGGL(k_cost_volume<true>, dim3(job.tiles * job.slabs), dim3(256), lds, stream,
                           job);
    else
        hipLaunchKernelGGL(k_cost_volume<false>, dim3(job.tiles * job.slabs), dim3(256), lds,
                           stream, job);
    return hipGetLastError();
}

int cost_job_mode(const CostJob &job) { return (job.window_lo ? 2 : 0) + (job.topk ? 1 : 0); }

hipError_t launch_cost_volume_queue(hipStream_t stream, const CostJob *jobs,
                                    const uint32_t *first, uint32_t njobs, size_t lds,
                                    uint32_t blocks, uint32_t *counters, uint32_t xcd_order,
                                    int mode) {
    if (lds > 160 * 1024 || njobs == 0) return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(stream, counters, 0u, kXcds);
    if (e != hipSuccess) return e;
    switch (mode) {
#define CHD_QUEUE(m)                                                                           \
    case m:                                                                                    \
        hipLaunchKernelGGL(k_cost_volume_queue<m>, dim3(blocks), dim3(256), lds, stream, jobs,  \
                           first, njobs, counters, xcd_order);                                 \
        break;
        CHD_QUEUE(0)
        CHD_QUEUE(1)
        CHD_QUEUE(2)
        CHD_QUEUE(3)
        CHD_QUEUE(kQueueMixed)
#undef CHD_QUEUE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ direct pass

// nearest_coarse_depth (depth.rs:1062-1076) and direct_depth_candidates
// (depth.rs:1078-1107); returns the candidate count.
__device__ uint32_t direct_candidates(const GridSpec &Cg, const chd_node_depth *coarse, double px,
                                      double py, double inverse_step, const chd_options &o,
                                      double *c, bool &seeded) {
    const double far_inverse = 1.0 / o.far_depth, near_inverse = 1.0 / o.near_depth;
    const uint32_t cc = sat_u32d(rclamp(round(px / (double)Cg.step), 0.0, (double)(Cg.columns - 1)));
    const uint32_t cr = sat_u32d(rclamp(round(py / (double)Cg.step), 0.0, (double)(Cg.rows - 1)));
    const chd_node_depth seed = coarse[(size_t)cr * Cg.columns + cc];
    seeded = seed.valid != 0;
    uint32_t count = 0;
    if (seeded) {
        for (int offset = -6; offset <= 6; offset++) {
            const double inverse = rclamp(1.0 / seed.depth + (double)offset * inverse_step * 0.5,
                                          far_inverse, near_inverse);
            const double depth = 1.0 / inverse;
            if (count == 0 || fabs(1.0 / c[count - 1] - 1.0 / depth) > inverse_step * 0.1)
                c[count++] = depth;
        }
    } else {
        // inverse_depth_samples(near, far, 32), reversed (depth.rs:1834-1843)
        for (uint32_t k = 0; k < 32; k++) {
            const double w = (double)(31 - k) / 31.0;
            c[k] = 1.0 / (near_inverse * (1.0 - w) + far_inverse * w);
        }
        count = 32;
    }
    return count;
}

// select_direct_depth (depth.rs:1109-1165) over `count` candidates whose
// aggregated scores are score[k] (NaN = None).
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
        const uint32_t d = idx[k] > best_index ? idx[k] - best_index : best_index - idx[k];
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

// best_depth_candidate (depth.rs:1167-1185): the last maximum.
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

// Candidate scoring shared by the direct and wide kernels for view slot a:
// items are (node, slot), slot 0 the global-warp baseline and slot k the
// depth cand[node * stride + k - 1]; base[] is the prefix of 1 + ncand over
// the block's nodes and item_node[] maps an item to its node. Results land
// in res[(node * V + view) * (stride + 1) + slot].
// KS = 0: res is [node][view][stride + 1] (every view's score kept);
// KS > 0: res is [node][stride + 1][KS], a running top-K of the views'
// scores per (node, slot) (topk_insert; the views come one at a time).
template <bool Rows>
__device__ void score_slots(const PatchSpec &S, const PatchLds &P, uint32_t nb, uint32_t T,
                            uint32_t a, const double *cand, const uint32_t *base,
                            const uint8_t *item_node, uint32_t stride, float *res,
                            unsigned int *evaluations, uint32_t KS = 0) {
    const uint32_t V = S.view_count, total = base[nb];
    const DevPlane tp = view_plane(S.views[S.vlist[a]], true);
    const __amdgpu_buffer_rsrc_t rsrc = plane_rsrc(tp);
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const uint32_t n = item_node[i];
        const uint32_t slot = i - base[n];
        float s = nan_f();
        if (P.node_ok[n] && P.nv_ok[n]) {
            double dx = 0.0, dy = 0.0;
            bool ok = true;
            if (slot > 0)
                ok = delta_at_depth(P, n, *P.cam, cand[n * stride + slot - 1], dx, dy);
            float sc;
            if (ok && patch_zncc<false, Rows>(P, n, T, tp, rsrc, dx, dy, sc,
                                              evaluations))
                s = sc;
        }
        if (KS)
            topk_insert(res + ((size_t)n * (stride + 1) + slot) * KS, KS - 1, s);
        else
            res[((size_t)n * V + a) * (stride + 1) + slot] = s;
    }
}

// The scoring half of measure_direct_depths (depth.rs:752-795): for every
// fine node the global baseline and each candidate depth, aggregated over
// the views, go to agg[node][slot] (NaN = None); also guidance and tested.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_direct_score(PatchSpec S, GridSpec F, GridSpec Cg,
                                                      const chd_node_depth *__restrict__ coarse,
                                                      double inverse_step, chd_options o,
                                                      uint32_t T, uint32_t nb, float *agg,
                                                      float *guidance, uint8_t *tested,
                                                      unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int evaluation_count;
    const uint32_t V = S.view_count;
    const uint32_t lat_points = tile_lattice_points(nb, F.step, S.radius);
    const PatchLayout lay = patch_layout(nb, V, T, false, lat_points);
    PatchLds P = bind_patch(lds, lay, false);
    P = lat_points ? with_lattice(P, nb, F.step, S.radius) : with_rows(P, S.radius);
    double *cand = (double *)(lds + lay.total);                 // [nb][32]
    // the views' scores per (node, slot): all V, or a running top-K when
    // best_view_count < V (KS slots: count + the K largest)
    const uint32_t KS = topk_slots(o.best_view_count, V) < V ? topk_slots(o.best_view_count, V) : 0;
    float *res = (float *)(cand + nb * kMaxCandidates);  // [nb][V][33] or [nb][33][KS]
    float *centre_value = res + (size_t)nb * (KS ? KS : V) * kDirectSlots;  // [nb]
    uint32_t *ncand = (uint32_t *)(centre_value + nb);          // [nb]
    uint32_t *node_tested = ncand + nb;                         // [nb]
    uint32_t *base = node_tested + nb;                          // [nb + 1]
    uint8_t *item_node = (uint8_t *)(base + nb + 1);            // [nb * 33]
    const NodeTile tile = row_tile(F, nb, blockIdx.x);
    const uint32_t node0 = tile.first, count = tile.count, tid = threadIdx.x;

    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        // slots past the tile continue the row (the lattice spans them;
        // they are never scored)
        const double2 p0 = node_position(F, node0);
        const double2 p = make_double2(p0.x + (double)(n * F.step), p0.y);
        P.pos[n] = p;
        P.node_ok[n] = n < count ? 1 : 0;
        node_tested[n] = 0;
        bool seeded;
        ncand[n] = n < count ? direct_candidates(Cg, coarse, p.x, p.y, inverse_step, o,
                                                 cand + n * kMaxCandidates, seeded)
                             : 0;
    }
    if (tid == 0) evaluation_count = 0;
    __syncthreads();
    if (tid == 0) {
        uint32_t b = 0;
        for (uint32_t n = 0; n < nb; n++) {
            base[n] = b;
            b += n < count ? 1 + ncand[n] : 0;
        }
        base[nb] = b;
    }
    patch_nodes(S, P, nb, T, centre_value);
    for (uint32_t n = tid; n < nb; n += blockDim.x)
        for (uint32_t k = 0; n < count && k <= ncand[n]; k++) item_node[base[n] + k] = (uint8_t)n;
    if (KS)  // no view scored yet (a slot is only touched by its item's thread)
        for (uint32_t i = tid; i < nb * kDirectSlots; i += blockDim.x)
            res[(size_t)i * KS] = __uint_as_float(0u);
    __syncthreads();
    for (uint32_t a = 0; a < V; a++) {
        // a view no node of the tile sees builds no tap context: its items
        // fail the node / warp test in score_slots and record None
        bool live;
        const PatchLds Q = patch_view(S, P, nb, T, a, true, 0, count, &live);
        score_slots<true>(S, Q, nb, T, a, cand, base, item_node, kMaxCandidates, res,
                          &evaluation_count, KS);
        __syncthreads();
    }
    for (uint32_t i = tid; i < nb * kDirectSlots; i += blockDim.x) {
        const uint32_t n = i / kDirectSlots, slot = i % kDirectSlots;
        const uint32_t g = node0 + n;
        if (n >= count || slot > ncand[n]) continue;
        float a;
        const bool scored =
            KS ? topk_aggregate(res + ((size_t)n * kDirectSlots + slot) * KS, o.minimum_support,
                                o.best_view_count, a)
               : aggregate_scores(res + (size_t)n * V * kDirectSlots + slot, kDirectSlots, V,
                                  o.minimum_support, o.best_view_count, a);
        if (!scored)
            a = nan_f();
        else
            node_tested[n] = 1;  // baseline or any candidate measured (depth.rs:794-795)
        agg[(size_t)g * kDirectSlots + slot] = a;
    }
    __syncthreads();
    for (uint32_t n = tid; n < count; n += blockDim.x) {
        const uint32_t g = node0 + n;
        guidance[g] = centre_value[n];
        tested[g] = (uint8_t)node_tested[n];
    }
    if (tid == 0 && work && evaluation_count) atomicAdd(work, (unsigned long long)evaluation_count);
}

// The selection half (depth.rs:796-810): select_direct_depth, or an entry
// of the wide-fallback list when the best candidate scores at least 0.30.
// One thread per fine node; the candidates are recomputed exactly as the
// scoring kernel made them.
__global__ void __launch_bounds__(256) k_direct_select(GridSpec F, GridSpec Cg,
                                                       const chd_node_depth *__restrict__ coarse,
                                                       double inverse_step, chd_options o,
                                                       const float *__restrict__ agg,
                                                       chd_node_depth *field, WideEntry *wide,
                                                       uint32_t *wide_count) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= F.columns * F.rows) return;
    const double2 p = node_position(F, g);
    double c[kMaxCandidates];
    bool seeded;
    const uint32_t count = direct_candidates(Cg, coarse, p.x, p.y, inverse_step, o, c, seeded);
    const float *sc = agg + (size_t)g * kDirectSlots;
    const float baseline = sc[0];
    chd_node_depth selected;
    clear_node(selected);
    if (!select_direct_depth(c, sc + 1, count, baseline == baseline, baseline, seeded, o,
                             selected)) {
        const int best = best_depth_candidate(c, sc + 1, count, o);
        if (best >= 0 && sc[1 + best] >= kMinimumRegularizedScore) {
            const uint32_t first = best >= 3 ? (uint32_t)best - 3 : 0;
            const uint32_t last = (uint32_t)best + 3 < count - 1 ? (uint32_t)best + 3 : count - 1;
            WideEntry e;
            e.node = g;
            e.count = last - first + 1;
            e.seeded = seeded ? 1u : 0u;
            e.pad = 0;
            for (uint32_t k = 0; k < 7; k++) e.depth[k] = k < e.count ? c[first + k] : 0.0;
            wide[atomicAdd(wide_count, 1u)] = e;
        }
    }
    field[g] = selected;
}

size_t direct_lds_bytes(uint32_t nb, uint32_t V, uint32_t T, uint32_t step, uint32_t radius,
                        uint32_t best_view_count) {
    const PatchLayout l = patch_layout(nb, V, T, false, tile_lattice_points(nb, step, radius));
    const uint32_t ks = topk_slots(best_view_count, V), per = ks < V ? ks : V;
    return (size_t)l.total + (size_t)nb * kMaxCandidates * 8 +
           (size_t)nb * per * kDirectSlots * 4 + (size_t)nb * 4 * 3 + (nb + 1) * 4 +
           (size_t)nb * kDirectSlots;
}

// Nodes per k_direct_score block and its dynamic LDS: 18 nodes give 252
// (seeded) to 594 (unseeded) items per view for 256 threads, within the LDS
// quarter that keeps four blocks per CU.
uint32_t direct_block_nodes(uint32_t V, uint32_t T, uint32_t step, uint32_t r,
                            uint32_t best_view_count, size_t &lds) {
    uint32_t nb = 18;
    while (nb > 1 && direct_lds_bytes(nb, V, T, step, r, best_view_count) > kQuarterLds) nb--;
    lds = direct_lds_bytes(nb, V, T, step, r, best_view_count);
    return nb;
}

hipError_t launch_direct(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                         const GridSpec &coarse, const chd_node_depth *coarse_field,
                         double inverse_step, const chd_options &direct_options,
                         chd_node_depth *field, float *guidance, uint8_t *tested,
                         WideEntry *wide, uint32_t *wide_count, float *agg,
                         unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t V = patch.view_count;
    size_t lds = 0;
    const uint32_t nb =
        direct_block_nodes(V, T, fine.step, patch.radius, direct_options.best_view_count, lds);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const uint32_t nodes = fine.columns * fine.rows;
    hipLaunchKernelGGL(k_direct_score, dim3(fine.rows * row_tiles(fine, nb)), dim3(256), lds,
                       stream, patch, fine, coarse, coarse_field, inverse_step, direct_options, T,
                       nb, agg, guidance, tested, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_direct_select, dim3((nodes + 255) / 256), dim3(256), 0, stream, fine,
                       coarse, coarse_field, inverse_step, direct_options, agg, field, wide,
                       wide_count);
    return hipGetLastError();
}

constexpr uint32_t kWideNodes = 8;  // wide-fallback nodes per block (at most)
constexpr uint32_t kWideSlots = 8;  // baseline + at most 7 candidates

size_t wide_lds_bytes(uint32_t nb, uint32_t V, uint32_t T) {
    const PatchLayout l = patch_layout(nb, V, T, false);
    return (size_t)l.total + (size_t)nb * kWideSlots * 8 + (size_t)nb * V * kWideSlots * 4 +
           (size_t)nb * kWideSlots * 4 + (size_t)nb * 4 * 2 + (nb + 1) * 4 + (size_t)nb * kWideSlots;
}

// The max(2r, 8) re-measurement of nodes whose first pass produced no
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
    const uint32_t V = S.view_count;
    const PatchLayout lay = patch_layout(nb, V, T, false);
    const PatchLds P = bind_patch(lds, lay, false);
    double *cand = (double *)(lds + lay.total);                   // [nb][7]
    float *res = (float *)(cand + nb * kWideSlots);               // [nb][V][8]
    float *agg = res + (size_t)nb * V * kWideSlots;               // [nb][8]
    float *centre_value = agg + nb * kWideSlots;                  // [nb]
    uint32_t *ncand = (uint32_t *)(centre_value + nb);            // [nb]
    uint32_t *base = ncand + nb;                                  // [nb + 1]
    uint8_t *item_node = (uint8_t *)(base + nb + 1);              // [nb * 8]
    const uint32_t tid = threadIdx.x;
    if (tid == 0) evaluation_count = 0;
    const uint32_t count = *wide_count;
    for (uint32_t first = blockIdx.x * nb; first < count; first += gridDim.x * nb) {
        __syncthreads();
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t e = first + n;
            if (e < count) {
                const WideEntry w = wide[e];
                P.pos[n] = node_position(F, w.node);
                P.node_ok[n] = 1;
                ncand[n] = w.count;
                for (uint32_t k = 0; k < 7; k++) cand[n * (kWideSlots - 1) + k] = w.depth[k];
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
                b += first + n < count ? 1 + ncand[n] : 0;
            }
            base[nb] = b;
        }
        patch_nodes(S, P, nb, T, centre_value);
        for (uint32_t n = tid; n < nb; n += blockDim.x)
            for (uint32_t k = 0; first + n < count && k <= ncand[n]; k++)
                item_node[base[n] + k] = (uint8_t)n;
        __syncthreads();
        for (uint32_t a = 0; a < V; a++) {
            const PatchLds Q = patch_view(S, P, nb, T, a);
            score_slots<false>(S, Q, nb, T, a, cand, base, item_node, kWideSlots - 1, res,
                               &evaluation_count);
            __syncthreads();
        }
        for (uint32_t i = tid; i < nb * kWideSlots; i += blockDim.x) {
            const uint32_t n = i / kWideSlots, slot = i % kWideSlots;
            float a = nan_f();
            if (first + n < count && slot <= ncand[n] &&
                !aggregate_scores(res + (size_t)n * V * kWideSlots + slot, kWideSlots, V,
                                  o.minimum_support, o.best_view_count, a))
                a = nan_f();
            agg[i] = a;
        }
        __syncthreads();
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t e = first + n;
            if (e >= count) continue;
            const float *sc = agg + n * kWideSlots;
            chd_node_depth selected;
            clear_node(selected);
            select_direct_depth(cand + n * (kWideSlots - 1), sc + 1, ncand[n], sc[0] == sc[0], sc[0],
                                wide[e].seeded != 0, o, selected);
            field[wide[e].node] = selected;
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
    const uint32_t V = patch.view_count;
    uint32_t nb = kWideNodes;
    while (nb > 1 && wide_lds_bytes(nb, V, T) > 64 * 1024) nb--;
    const size_t lds = wide_lds_bytes(nb, V, T);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    uint32_t blocks = (wide_capacity + nb - 1) / nb;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_direct_wide, dim3(blocks), dim3(256), lds, stream, patch, fine,
                       wide_options, T, nb, wide, wide_count, field, work);
    return hipGetLastError();
}

// ------------------------------------------------------ per-target stage

constexpr uint32_t kTargetNodes = 32;

// refine_one_view (depth.rs:1634-1708) and the decision block
// (depth.rs:429-486) for every target (S.vlist) over a tile of fine nodes,
// writing decisions[target slot][node]. Per target, phase A runs one item
// per (node, shared-depth offset): map_at_depth and its ZNCC (any None among
// the five makes the refinement None, the reference's `?`s; running all five
// first changes nothing since they have no side effects), plus the global
// baseline; phase B the eight +-1.5 px residuals around the best depth.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_targets(PatchSpec S, GridSpec F,
                                                 const chd_node_depth *__restrict__ field,
                                                 double inverse_step, chd_options o, uint32_t T,
                                                 uint32_t nb, chd_node_warp *decisions,
                                                 unsigned long long *work) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ unsigned int list_count;
    __shared__ uint16_t list[kTargetNodes * 8];  // items of the phase that pass every gate
    const uint32_t V = S.view_count;
    const uint32_t lat_points = tile_lattice_points(nb, F.step, S.radius);
    const PatchLayout lay = patch_layout(nb, V, T, false, lat_points);
    PatchLds P = bind_patch(lds, lay, false);
    P = lat_points ? with_lattice(P, nb, F.step, S.radius) : with_rows(P, S.radius);
    float2 *points = (float2 *)(lds + lay.total);  // [nb][5]
    float2 *best_point = points + nb * 5;          // [nb]
    float *phase = (float *)(best_point + nb);     // [nb][8]
    float *centre_value = phase + nb * 8;          // [nb]
    float *best_score = centre_value + nb;         // [nb]
    float *baseline = best_score + nb;             // [nb]
    int *state = (int *)(baseline + nb);           // [nb] 0 skip, 1 refine, -1 None
    const NodeTile tile = row_tile(F, nb, blockIdx.x);
    const uint32_t nodes = F.columns * F.rows, node0 = tile.first, count = tile.count;
    const uint32_t tid = threadIdx.x;
    const double min_inverse = 1.0 / o.far_depth, max_inverse = 1.0 / o.near_depth;

    int *measured = state + nb;                    // [nb] the node has a depth
    for (uint32_t n = tid; n < nb; n += blockDim.x) {
        // slots past the tile continue the row (the lattice spans them;
        // they are never scored)
        const double2 p0 = node_position(F, node0);
        P.pos[n] = make_double2(p0.x + (double)(n * F.step), p0.y);
        P.node_ok[n] = n < count ? 1 : 0;
        measured[n] = n < count && field[node0 + n].valid ? 1 : 0;
    }
    if (tid == 0) list_count = 0;
    unsigned int evaluations = 0;  // thread 0's tally of ZNCCs past the bounds check
    __syncthreads();
    patch_nodes(S, P, nb, T, centre_value);

    for (uint32_t a = 0; a < V; a++) {
        bool live;
        const PatchLds Q = patch_view(S, P, nb, T, a, true, 0, count, &live, measured);
        chd_node_warp *out = decisions + (size_t)a * nodes;
        if (!live) {
            // no node with a depth has its reference patch and this view's
            // global warp: every ZNCC of the view is None, so every node is
            // Global where the warp is defined at its centre and Undefined
            // elsewhere (depth.rs:433-440, 458-485)
            for (uint32_t n = tid; n < nb; n += blockDim.x) {
                const uint32_t g = node0 + n;
                if (n >= count) continue;
                chd_node_warp d;
                d.code = Q.gc_ok[n] ? CHD_WARP_GLOBAL : CHD_WARP_UNDEFINED;
                d.global[0] = Q.gc_ok[n] ? Q.gc[n].x : 0.0f;
                d.global[1] = Q.gc_ok[n] ? Q.gc[n].y : 0.0f;
                d.point[0] = d.point[1] = 0.0f;
                d.confidence = 0.0f;
                d.measured = 0;
                out[g] = d;
            }
            __syncthreads();
            continue;
        }
        const DevPlane tp = view_plane(S.views[S.vlist[a]], true);
        const __amdgpu_buffer_rsrc_t rsrc = plane_rsrc(tp);
        // Undefined / Global / to refine (depth.rs:433-440)
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t g = node0 + n;
            state[n] = n < count && Q.gc_ok[n] && field[g].valid ? 1 : 0;
        }
        __syncthreads();
        // phase A: offsets k < 5 (score_one_view_at_point, depth.rs:1710-1736)
        // and the baseline k = 5 (score_one_view with no depth: delta 0).
        // Every item first records None and its point; the items that pass
        // the gates (projection, parallax, patch and bounds) join a list whose
        // ZNCCs then run on consecutive threads, so no lane idles in the taps.
        for (uint32_t i = tid; i < nb * 6; i += blockDim.x) {
            const uint32_t n = i / 6, k = i % 6;
            if (state[n] == 0) continue;
            bool run = false;
            if (k == 5) {
                baseline[n] = nan_f();
                run = Q.node_ok[n] && Q.nv_ok[n] && zncc_inside(Q, n, tp, 0.0, 0.0);
            } else {
                const double offsets[5] = {-1.0, -0.5, 0.0, 0.5, 1.0};
                const double shared_inverse = 1.0 / field[node0 + n].depth;
                const double inverse =
                    rclamp(shared_inverse + offsets[k] * inverse_step, min_inverse, max_inverse);
                const double depth = 1.0 / inverse;
                double qx = 0.0, qy = 0.0;
                bool ok = Q.ray_ok[n] != 0;
                if (ok) {
                    const double *r = Q.ray + (size_t)n * 6;
                    const d3 point{r[0] + r[3] * depth, r[1] + r[4] * depth, r[2] + r[5] * depth};
                    ok = project(*Q.cam, point, qx, qy) && finite_d(qx) && finite_d(qy);
                }
                const float2 q = make_float2((float)qx, (float)qy);
                points[n * 5 + k] = q;
                phase[n * 8 + k] = nan_f();
                if (ok) {
                    const double dx = (double)(q.x - Q.gc[n].x), dy = (double)(q.y - Q.gc[n].y);
                    run = !(fabs(dx) > 128.0 || fabs(dy) > 128.0) && Q.node_ok[n] &&
                          Q.nv_ok[n] && zncc_inside(Q, n, tp, dx, dy);
                }
            }
            if (run) list[atomicAdd(&list_count, 1u)] = (uint16_t)i;
        }
        __syncthreads();
        {
            const uint32_t count = list_count;
            if (tid == 0) evaluations += count;
            for (uint32_t j = tid; j < count; j += blockDim.x) {
                const uint32_t i = list[j], n = i / 6, k = i % 6;
                double dx = 0.0, dy = 0.0;
                if (k < 5) {
                    const float2 q = points[n * 5 + k];
                    dx = (double)(q.x - Q.gc[n].x);
                    dy = (double)(q.y - Q.gc[n].y);
                }
                float sc;
                if (zncc_taps<false, true>(Q, n, T, tp, rsrc, dx, dy, sc)) {
                    if (k == 5)
                        baseline[n] = sc;
                    else
                        phase[n * 8 + k] = sc;
                }
            }
        }
        __syncthreads();
        // the best shared-depth offset: first maximum of score - |offset| * 0.002
        if (tid == 0) list_count = 0;
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            if (state[n] != 1) continue;
            const float magnitude[5] = {1.0f, 0.5f, 0.0f, 0.5f, 1.0f};
            bool failed = false;
            float bo = 0.0f, bs = 0.0f;
            float2 bp = make_float2(0.0f, 0.0f);
            for (uint32_t k = 0; k < 5; k++) {
                const float s = phase[n * 8 + k];
                if (s != s) {
                    failed = true;
                    break;
                }
                const float objective = s - magnitude[k] * 0.002f;
                if (k == 0 || objective > bo) {
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
        // phase B: residual offsets, dy-major, gated and listed as in phase A
        const float steps[3] = {-1.5f, 0.0f, 1.5f};
        for (uint32_t i = tid; i < nb * 8; i += blockDim.x) {
            const uint32_t n = i / 8, k = i % 8;
            if (state[n] != 1) continue;
            const uint32_t cell = k < 4 ? k : k + 1;
            const float2 point = make_float2(best_point[n].x + steps[cell % 3],
                                             best_point[n].y + steps[cell / 3]);
            const double dx = (double)(point.x - Q.gc[n].x), dy = (double)(point.y - Q.gc[n].y);
            phase[n * 8 + k] = nan_f();
            if (!(fabs(dx) > 128.0 || fabs(dy) > 128.0) && Q.node_ok[n] && Q.nv_ok[n] &&
                zncc_inside(Q, n, tp, dx, dy))
                list[atomicAdd(&list_count, 1u)] = (uint16_t)i;
        }
        __syncthreads();
        {
            const uint32_t count = list_count;
            if (tid == 0) evaluations += count;
            for (uint32_t j = tid; j < count; j += blockDim.x) {
                const uint32_t i = list[j], n = i / 8, k = i % 8;
                const uint32_t cell = k < 4 ? k : k + 1;
                const float2 point = make_float2(best_point[n].x + steps[cell % 3],
                                                 best_point[n].y + steps[cell / 3]);
                const double dx = (double)(point.x - Q.gc[n].x);
                const double dy = (double)(point.y - Q.gc[n].y);
                float sc;
                if (zncc_taps<false, true>(Q, n, T, tp, rsrc, dx, dy, sc))
                    phase[n * 8 + k] = sc;
            }
        }
        __syncthreads();
        if (tid == 0) list_count = 0;
        for (uint32_t n = tid; n < nb; n += blockDim.x) {
            const uint32_t g = node0 + n;
            if (n >= count) continue;
            chd_node_warp d;
            d.code = CHD_WARP_UNDEFINED;
            d.global[0] = d.global[1] = 0.0f;
            d.point[0] = d.point[1] = 0.0f;
            d.confidence = 0.0f;
            d.measured = 0;
            if (Q.gc_ok[n]) {
                d.code = CHD_WARP_GLOBAL;
                d.global[0] = Q.gc[n].x;
                d.global[1] = Q.gc[n].y;
            }
            if (state[n] != 0) {
                bool has_selected = false;
                float selected_score = 0.0f;
                float2 selected_point = make_float2(0.0f, 0.0f);
                if (state[n] == 1) {
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
                if (contradicted) {
                    d.code = CHD_WARP_OCCLUDED;
                    d.measured = node.regularized ? 0u : 1u;
                } else if (supported) {
                    d.code = CHD_WARP_REFINED;
                    d.point[0] = selected_point.x;
                    d.point[1] = selected_point.y;
                    d.confidence =
                        node.regularized ? node.confidence * 0.6f : 0.5f + 0.5f * node.confidence;
                    d.measured = node.regularized ? 0u : 1u;
                }
            }
            out[g] = d;
        }
        __syncthreads();
    }
    if (tid == 0 && work && evaluations) atomicAdd(work, (unsigned long long)evaluations);
}

size_t targets_lds_bytes(uint32_t nb, uint32_t V, uint32_t T, uint32_t step, uint32_t radius) {
    const PatchLayout l = patch_layout(nb, V, T, false, tile_lattice_points(nb, step, radius));
    return (size_t)l.total + nb * 5 * 8 + nb * 8 + nb * 8 * 4 + nb * 4 * 5;
}

// Nodes per k_targets block and its dynamic LDS: 32 nodes give 192 phase-A
// and 256 phase-B items for 256 threads; with the shared lattice a quarter
// of the CU's LDS holds them (four blocks per CU).
uint32_t targets_block_nodes(uint32_t V, uint32_t T, uint32_t step, uint32_t r, size_t &lds) {
    uint32_t nb = kTargetNodes;
    while (nb > 1 && targets_lds_bytes(nb, V, T, step, r) > kQuarterLds) nb--;
    lds = targets_lds_bytes(nb, V, T, step, r);
    return nb;
}

hipError_t zncc_occupancy(const CostJob &cost, const PatchSpec &direct, const PatchSpec &target,
                          uint32_t fine_step, uint32_t best_view_count, int blocks[3]) {
    const void *cost_fn =
        cost.window_lo ? (cost.topk ? (const void *)k_cost_volume_band<true>
                                    : (const void *)k_cost_volume_band<false>)
                       : (cost.topk ? (const void *)k_cost_volume<true>
                                    : (const void *)k_cost_volume<false>);
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks[0], cost_fn, 256,
                                                                cost_job_lds(cost));
    if (e != hipSuccess) return e;
    size_t lds = 0;
    const uint32_t Td = (2 * direct.radius + 1) * (2 * direct.radius + 1);
    direct_block_nodes(direct.view_count, Td, fine_step, direct.radius, best_view_count, lds);
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks[1], (const void *)k_direct_score,
                                                     256, lds);
    if (e != hipSuccess) return e;
    const uint32_t Tt = (2 * target.radius + 1) * (2 * target.radius + 1);
    targets_block_nodes(target.view_count, Tt, fine_step, target.radius, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks[2], (const void *)k_targets, 256,
                                                        lds);
}

hipError_t launch_targets(hipStream_t stream, const PatchSpec &patch, const GridSpec &fine,
                          const chd_node_depth *field, double inverse_step,
                          const chd_options &options, chd_node_warp *decisions,
                          unsigned long long *work) {
    const uint32_t T = (2 * patch.radius + 1) * (2 * patch.radius + 1);
    const uint32_t V = patch.view_count;
    size_t lds = 0;
    const uint32_t nb = targets_block_nodes(V, T, fine.step, patch.radius, lds);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_targets, dim3(fine.rows * row_tiles(fine, nb)), dim3(256), lds, stream,
                       patch, fine, field, inverse_step, options, T, nb, decisions, work);
    return hipGetLastError();
}

}  // namespace chd
