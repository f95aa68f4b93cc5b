// align_kernels.hip - the device half of align_module's refinement
// (crates/chiaro-fusion/src/align.rs:270-519) on gfx950.
//
//   k_downsample       Plane::downsample, one pyramid level (image.rs:469-484)
//   k_initial_grid     the factory model: map_from(.., 1e8) on an 8 px grid
//                      (align.rs:317-324)
//   k_render_through   the target through initial(correction(p)) at one
//                      level (align.rs:523-547)
//   k_match_patches    match_patch (image.rs:531-640) of many windows, one
//                      block per window: the reference window and its
//                      statistics in LDS, one lane per integer shift
//                      summing the window in the reference's row-major order
//                      (neighbouring lanes read neighbouring target texels),
//                      then the peak and its sub-pixel parabola
//   k_gain_samples     the texel pairs of the photometric gain
//   k_ransac_count     the consensus count of every RANSAC hypothesis
//                      (align.rs:655-686), one block per hypothesis
//   k_render_bandwidth the target through a warp at the reference plane's
//                      sampling, blurred by a Gaussian footprint when the
//                      target is magnified (resolution.rs:254-270, 441-465)
//
// The RANSAC samples, the 4-point and final least-squares fits and the
// residual medians run on the host; the 300 consensus counts over a few
// thousand pairs, which dominate RANSAC, run on the device.
// Planes of the pyramids are stored with one replicated padding column and
// row (pitch width + 1), so Plane::sample's clamp of x0 + 1 at the border is
// a plain read.
#include <hip/hip_runtime.h>

#include "depth_device.hpp"
#include "depth_internal.hpp"

namespace chd {

namespace {

__global__ void k_downsample(const float *src, uint32_t sw, uint32_t sh, uint32_t spitch,
                             float *dst, uint32_t dw, uint32_t dh, uint32_t dpitch) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= dw || j >= dh) return;
    const uint32_t x = 2 * i, y = 2 * j;
    const uint32_t x1 = min(x + 1, sw - 1), y1 = min(y + 1, sh - 1);
    const float a = src[(size_t)y * spitch + x], b = src[(size_t)y * spitch + x1];
    const float c = src[(size_t)y1 * spitch + x], d = src[(size_t)y1 * spitch + x1];
    dst[(size_t)j * dpitch + i] = 0.25f * (a + b + c + d);
}

__global__ void k_initial_grid(const chd_camera *reference, const chd_camera *target,
                               uint32_t columns, uint32_t rows, uint32_t step, float2 *out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= columns || r >= rows) return;
    // target.map_from(reference, p, 1e8) filtered to finite points
    // (geometry.rs:249-252)
    d3 o, d;
    pixel_to_ray(*reference, (double)((size_t)c * step), (double)((size_t)r * step), o, d);
    const d3 far{o.x + d.x * kInfinityDepth, o.y + d.y * kInfinityDepth, o.z + d.z * kInfinityDepth};
    double qx, qy;
    float2 q = make_float2(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000));
    if (project(*target, far, qx, qy) && finite_d(qx) && finite_d(qy))
        q = make_float2((float)qx, (float)qy);
    out[(size_t)r * columns + c] = q;
}

struct Homography {
    double h[9];
};

// apply_homography (math.rs:131-138)
__device__ __forceinline__ bool apply_homography(const Homography &H, double x, double y,
                                                 double &ox, double &oy) {
    const double *h = H.h;
    const double v0 = h[0] * x + h[1] * y + h[2] * 1.0;
    const double v1 = h[3] * x + h[4] * y + h[5] * 1.0;
    const double v2 = h[6] * x + h[7] * y + h[8] * 1.0;
    if (fabs(v2) < 1e-12) return false;
    ox = v0 / v2;
    oy = v1 / v2;
    return true;
}

__global__ void k_render_through(uint32_t width, uint32_t height, uint32_t reference_scale,
                                 DevPlane target, double target_scale, DevWarp initial,
                                 Homography correction, float *out) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= width || y >= height) return;
    const double rs = (double)reference_scale;
    const double px = ((double)(float)x * rs * 2.0) + rs - 0.5;
    const double py = ((double)(float)y * rs * 2.0) + rs - 0.5;
    float value = __int_as_float(0x7fc00000);
    double cx, cy;
    float2 q;
    float v;
    if (apply_homography(correction, px, py, cx, cy) && warp_map(initial, (float)cx, (float)cy, q) &&
        plane_sample(target, (float)(((double)q.x + 0.5) / (2.0 * target_scale) - 0.5),
                     (float)(((double)q.y + 0.5) / (2.0 * target_scale) - 0.5), v))
        value = v;
    out[(size_t)y * width + x] = value;
}

__device__ __forceinline__ int32_t total_key(float f) {
    const int32_t k = __float_as_int(f);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}

__device__ __forceinline__ float refine_peak(float minus, float centre, float plus) {
    const float denominator = minus - 2.0f * centre + plus;
    if (fabsf(denominator) < 1e-9f) return 0.0f;
    const float v = 0.5f * (minus - plus) / denominator;
    return v < -0.5f ? -0.5f : v > 0.5f ? 0.5f : v;  // f32::clamp keeps NaN
}

// One block per window k. LDS: the centred reference window (size^2 f32)
// and the shift scores ((2 radius + 1)^2 f32).
__global__ void __launch_bounds__(256) k_match_patches(const float *reference, const float *target,
                                                       uint32_t width, uint32_t height,
                                                       const uint32_t *rxs, const uint32_t *rys,
                                                       uint32_t size, uint32_t radius,
                                                       int coverage, float *shift_out,
                                                       float *score_out, uint8_t *valid_out) {
    extern __shared__ __align__(16) float lds_f[];
    __shared__ float stats[2];  // mean, std
    __shared__ int usable;
    const uint32_t k = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const uint32_t rx = rxs[k], ry = rys[k];
    const uint32_t area = size * size, span = 2 * radius + 1, shifts = span * span;
    float *window = lds_f;           // [size * size]
    float *scores = lds_f + area;    // [span * span]
    // match_patch's bounds (image.rs:540-550), both planes width x height
    const bool inside = rx + size <= width && ry + size <= height && rx >= radius && ry >= radius &&
                        rx + size + radius <= width && ry + size + radius <= height;
    if (!inside) {
        if (tid == 0) valid_out[k] = 0;
        return;
    }
    int nan_seen = 0;
    for (uint32_t i = tid; i < area; i += nth) {
        const uint32_t y = i / size, x = i % size;
        window[i] = reference[(size_t)(ry + y) * width + rx + x];
        if (coverage == kCoverWindowNan) {
            const float t = target[(size_t)(ry + y) * width + rx + x];
            nan_seen |= t != t;
        }
    }
    if (coverage == kCoverSearchFinite) {
        // every target pixel the search can read is finite; the bounds
        // above keep [rx - radius, rx + size + radius) inside the plane
        const uint32_t side = size + 2 * radius;
        for (uint32_t i = tid; i < side * side; i += nth) {
            const uint32_t y = i / side, x = i % side;
            const float t = target[(size_t)(ry - radius + y) * width + rx - radius + x];
            nan_seen |= !(t - t == 0.0f);
        }
    }
    const int uncovered = __syncthreads_or(nan_seen);
    if (tid == 0) {
        // Iterator::sum in order from -0.0, then the centred second moment
        float sum = -0.0f;
        for (uint32_t i = 0; i < area; i++) sum += window[i];
        const float n = (float)area;
        const float mean = sum / n;
        float sq = -0.0f;
        for (uint32_t i = 0; i < area; i++) {
            const float dv = window[i] - mean;
            sq += dv * dv;
        }
        stats[0] = mean;
        stats[1] = sqrtf(sq / n);
        usable = !uncovered && !(stats[1] < 1e-4f);
    }
    __syncthreads();
    if (!usable) {
        if (tid == 0) valid_out[k] = 0;
        return;
    }
    const float mean = stats[0], ref_std = stats[1], n = (float)area;
    for (uint32_t i = tid; i < area; i += nth) window[i] = window[i] - mean;
    __syncthreads();
    for (uint32_t s = tid; s < shifts; s += nth) {
        const uint32_t dy = s / span, dx = s % span;
        const uint32_t tx = rx + dx - radius, ty = ry + dy - radius;
        float sum = 0.0f, sum_sq = 0.0f, cross = 0.0f;
        for (uint32_t y = 0; y < size; y++) {
            const float *row = target + (size_t)(ty + y) * width + tx;
            const float *rrow = window + y * size;
            for (uint32_t x = 0; x < size; x++) {
                const float t = row[x];
                sum += t;
                sum_sq += t * t;
                cross += t * rrow[x];
            }
        }
        const float m = sum / n;
        const float var = fmaxf(sum_sq / n - m * m, 0.0f);
        float score = -__int_as_float(0x7f800000);
        if (!(var < 1e-8f)) score = (cross / n) / (sqrtf(var) * ref_std);
        scores[s] = score;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t best = 0;  // max_by(total_cmp): the last maximum
        int32_t best_key = total_key(scores[0]);
        for (uint32_t s = 1; s < shifts; s++) {
            const int32_t key = total_key(scores[s]);
            if (key >= best_key) {
                best = s;
                best_key = key;
            }
        }
        const float score = scores[best];
        if (!(score - score == 0.0f)) {
            valid_out[k] = 0;
            return;
        }
        const uint32_t bx = best % span, by = best / span;
        const float sub_x = bx > 0 && bx + 1 < span
                                ? refine_peak(scores[by * span + bx - 1], score, scores[by * span + bx + 1])
                                : 0.0f;
        const float sub_y = by > 0 && by + 1 < span
                                ? refine_peak(scores[(by - 1) * span + bx], score, scores[(by + 1) * span + bx])
                                : 0.0f;
        shift_out[2 * k] = ((float)bx - (float)radius) + sub_x;
        shift_out[2 * k + 1] = ((float)by - (float)radius) + sub_y;
        score_out[k] = score;
        valid_out[k] = 1;
    }
}

// The samples of photometric_gain (align.rs:795-824): at every 16th
// reference plane pixel (x, y) = (16 (i + 1), 16 (j + 1)) the reference
// texel and the target sampled through the warp at ((q - 0.5) / 2); the
// ratios, their sort and median stay on the host (the platform expf).
__global__ void k_gain_samples(DevPlane reference, DevPlane target, DevWarp warp, uint32_t gx,
                               uint32_t gy, float2 *samples, uint8_t *ok) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= gx || j >= gy) return;
    const uint32_t x = 16 * (i + 1), y = 16 * (j + 1);
    const size_t k = (size_t)j * gx + i;
    float2 q;
    float t;
    const bool defined = warp_map(warp, (float)x * 2.0f + 0.5f, (float)y * 2.0f + 0.5f, q) &&
                         plane_sample(target, (q.x - 0.5f) / 2.0f, (q.y - 0.5f) / 2.0f, t);
    ok[k] = defined ? 1 : 0;
    samples[k] = make_float2(reference.data[(size_t)y * reference.pitch + x], defined ? t : 0.0f);
}

// RANSAC's consensus counts (align.rs:672-683): one block per hypothesis
// (9 f64 each, row-major), threads over the pairs (a.x, a.y, b.x, b.y); a
// pair counts when its residual under the hypothesis is at most the
// threshold. Counts are integers, so the order of the block's sum is free.
__global__ void __launch_bounds__(256) k_ransac_count(const double4 *pairs, uint32_t count,
                                                      const double *hypotheses, double threshold,
                                                      uint32_t *counts) {
    __shared__ uint32_t total;
    Homography H;
    for (int i = 0; i < 9; i++) H.h[i] = hypotheses[(size_t)blockIdx.x * 9 + i];
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    uint32_t n = 0;
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const double4 p = pairs[i];
        double px, py;
        if (apply_homography(H, p.x, p.y, px, py)) {
            const double dx = px - p.z, dy = py - p.w;
            n += sqrt(dx * dx + dy * dy) <= threshold ? 1u : 0u;
        }
    }
    atomicAdd(&total, n);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

__global__ void k_render_bandwidth(uint32_t width, uint32_t height, DevPlane target, DevWarp base,
                                   BandwidthKernel kernel, float *out) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= width || y >= height) return;
    float value = __int_as_float(0x7fc00000);
    float2 q;
    if (warp_map(base, (float)x * 2.0f + 0.5f, (float)y * 2.0f + 0.5f, q)) {
        const float u = (q.x + 0.5f) * 0.5f - 0.5f, v = (q.y + 0.5f) * 0.5f - 0.5f;
        float s;
        if (kernel.taps == 0) {
            if (plane_sample(target, u, v, s)) value = s;
        } else {
            // the weighted mean over the footprint's defined samples, summed
            // row by row from (-taps, -taps)
            const int r = (int)kernel.taps;
            float sum = 0.0f, weight = 0.0f;
            int k = 0;
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++, k++)
                    if (plane_sample(target, u + (float)dx, v + (float)dy, s)) {
                        sum += kernel.weight[k] * s;
                        weight += kernel.weight[k];
                    }
            if (weight > 1.0e-8f) value = sum / weight;
        }
    }
    out[(size_t)y * width + x] = value;
}

}  // namespace

hipError_t launch_gain_samples(hipStream_t stream, const DevPlane &reference,
                              const DevPlane &target, const DevWarp &warp, uint32_t gx,
                              uint32_t gy, float2 *samples, uint8_t *ok) {
    if (gx == 0 || gy == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gain_samples, dim3((gx + 63) / 64, gy), dim3(64), 0, stream, reference,
                       target, warp, gx, gy, samples, ok);
    return hipGetLastError();
}

hipError_t launch_ransac_count(hipStream_t stream, const double *pairs, uint32_t count,
                              const double *hypotheses, uint32_t hypothesis_count,
                              double threshold, uint32_t *counts) {
    if (hypothesis_count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ransac_count, dim3(hypothesis_count), dim3(256), 0, stream,
                       (const double4 *)pairs, count, hypotheses, threshold, counts);
    return hipGetLastError();
}

hipError_t launch_render_bandwidth(hipStream_t stream, uint32_t width, uint32_t height,
                                   const DevPlane &target, const DevWarp &base,
                                   const BandwidthKernel &kernel, float *out) {
    const dim3 block(256), grid((width + 255) / 256, height);
    hipLaunchKernelGGL(k_render_bandwidth, grid, block, 0, stream, width, height, target, base,
                       kernel, out);
    return hipGetLastError();
}

hipError_t launch_downsample(hipStream_t stream, const float *src, uint32_t sw, uint32_t sh,
                             uint32_t spitch, float *dst, uint32_t dw, uint32_t dh,
                             uint32_t dpitch) {
    const dim3 block(256), grid((dw + 255) / 256, dh);
    hipLaunchKernelGGL(k_downsample, grid, block, 0, stream, src, sw, sh, spitch, dst, dw, dh,
                       dpitch);
    return hipGetLastError();
}

hipError_t launch_initial_grid(hipStream_t stream, const chd_camera *reference,
                               const chd_camera *target, uint32_t columns, uint32_t rows,
                               uint32_t step, float2 *out) {
    const dim3 block(64), grid((columns + 63) / 64, rows);
    hipLaunchKernelGGL(k_initial_grid, grid, block, 0, stream, reference, target, columns, rows,
                       step, out);
    return hipGetLastError();
}

hipError_t launch_render_through(hipStream_t stream, uint32_t width, uint32_t height,
                                 uint32_t reference_scale, const DevPlane &target,
                                 double target_scale, const DevWarp &initial,
                                 const double correction[9], float *out) {
    Homography h;
    for (int i = 0; i < 9; i++) h.h[i] = correction[i];
    const dim3 block(256), grid((width + 255) / 256, height);
    hipLaunchKernelGGL(k_render_through, grid, block, 0, stream, width, height, reference_scale,
                       target, target_scale, initial, h, out);
    return hipGetLastError();
}

size_t match_patches_lds(uint32_t size, uint32_t radius) {
    return ((size_t)size * size + (size_t)(2 * radius + 1) * (2 * radius + 1)) * sizeof(float);
}

hipError_t launch_match_patches(hipStream_t stream, const float *reference, const float *target,
                                uint32_t width, uint32_t height, const uint32_t *rx,
                                const uint32_t *ry, uint32_t count, uint32_t size, uint32_t radius,
                                int coverage, float *shift, float *score, uint8_t *valid) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_match_patches, dim3(count), dim3(256), match_patches_lds(size, radius),
                       stream, reference, target, width, height, rx, ry, size, radius, coverage,
                       shift, score, valid);
    return hipGetLastError();
}

}  // namespace chd
