// synth_kernels.hip - the synthesis consumer's depth and warp stage on
// gfx950: for every canvas pixel, the nearest depth node
// (DenseDepthMap::sample_nearest, depth.rs:152-159) and, per source, the
// warp point, border feather, local warp confidence, base weight and
// close-side focus weight that synthesize (synth.rs:633-725) derives before
// it samples the mosaics.
//
// One thread per canvas pixel; the sources are looped in order and each
// output is a [source][row][column] plane, so a wavefront's stores of one
// source are 64 consecutive words. The 64-row band filter
// (band_touches_module, synth.rs:1688-1721) is evaluated on the host, a
// handful of warp lookups per band and source, and read here as a flag.
#include <hip/hip_runtime.h>

#include "depth_device.hpp"
#include "depth_internal.hpp"

namespace chd {

namespace {

__device__ __forceinline__ float keep_nan_clamp(float v, float lo, float hi) {
    return v < lo ? lo : v > hi ? hi : v;
}

// Warp::confidence (align.rs:112-131); 0 without a confidence array
__device__ __forceinline__ float warp_confidence(const SynthSourceDev &s, float x, float y) {
    const DevWarp &w = s.warp;
    const bool scaled = w.inverse_step != 0.0f;
    const float fx = scaled ? x * w.inverse_step : x / (float)w.step;
    const float fy = scaled ? y * w.inverse_step : y / (float)w.step;
    if (fx < 0.0f || fy < 0.0f || !s.confidence) return 0.0f;
    const uint32_t c0 = min(sat_u32(floorf(fx)), w.columns - 1);
    const uint32_t r0 = min(sat_u32(floorf(fy)), w.rows - 1);
    const uint32_t c1 = min(c0 + 1, w.columns - 1), r1 = min(r0 + 1, w.rows - 1);
    const float tx = fx - (float)c0, ty = fy - (float)r0;
    const float *v = s.confidence;
    const float top = v[(size_t)r0 * w.columns + c0] * (1.0f - tx) + v[(size_t)r0 * w.columns + c1] * tx;
    const float bottom = v[(size_t)r1 * w.columns + c0] * (1.0f - tx) + v[(size_t)r1 * w.columns + c1] * tx;
    return keep_nan_clamp(top * (1.0f - ty) + bottom * ty, 0.0f, 1.0f);
}

// focus_consistency_weight (synth.rs:1469-1491), powi(4) as two squarings
__device__ __forceinline__ float focus_consistency(bool has_depth, double depth, float confidence,
                                                   const SynthSourceDev &s) {
    if (s.reference || s.magnification <= 1.05f) return 1.0f;
    if (!has_depth || !s.has_focus_distance) return 1.0f;
    const double focus = s.focus_distance;
    if (!finite_d(depth) || !finite_d(focus) || depth <= 0.0 || focus <= 0.0) return 1.0f;
    const float mismatch = (float)(focus / depth) * (s.magnification - 1.0f);
    const float relative = mismatch / 0.75f;
    const float square = relative * relative;
    const float supported = 1.0f / (1.0f + square * square);
    return 1.0f - keep_nan_clamp(confidence, 0.0f, 1.0f) * (1.0f - supported);
}

__device__ void synthesis_pixel(const SynthCanvasDev &cv, const DepthMapDev &map,
                                const SynthSourceDev *sources, uint32_t count,
                                const uint8_t *touch, uint32_t u, uint32_t v, float2 *points,
                                float *feather, float *local_out, float *base_out,
                                float *focus_out);

__global__ void __launch_bounds__(256) k_synthesis_geometry(SynthCanvasDev cv, DepthMapDev map,
                                                            const SynthSourceDev *sources,
                                                            uint32_t count, const uint8_t *touch,
                                                            float2 *points, float *feather,
                                                            float *local_out, float *base_out,
                                                            float *focus_out) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= cv.width) return;
    // rows beyond the grid's y extent (65,535) are taken in strides
    for (uint32_t v = blockIdx.y; v < cv.height; v += gridDim.y)
        synthesis_pixel(cv, map, sources, count, touch, u, v, points, feather, local_out, base_out,
                        focus_out);
}

__device__ void synthesis_pixel(const SynthCanvasDev &cv, const DepthMapDev &map,
                                const SynthSourceDev *sources, uint32_t count,
                                const uint8_t *touch, uint32_t u, uint32_t v, float2 *points,
                                float *feather, float *local_out, float *base_out,
                                float *focus_out) {
    const float rx = cv.crop_x + ((float)u + 0.5f) / cv.scale - 0.5f;
    const float ry = cv.crop_y + ((float)v + 0.5f) / cv.scale - 0.5f;
    // sample_nearest, then node.depth (NaN = None) with its confidence
    bool has_depth = false;
    double depth = 0.0;
    float depth_confidence = 0.0f;
    if (map.depth && !(rx < 0.0f || ry < 0.0f || map.step == 0)) {
        const uint32_t column = sat_u32(roundf(rx / (float)map.step));
        const uint32_t row = sat_u32(roundf(ry / (float)map.step));
        if (column < map.columns && row < map.rows) {
            const size_t node = (size_t)row * map.columns + column;
            const double d = map.depth[node];
            if (d == d) {
                has_depth = true;
                depth = d;
                depth_confidence = map.confidence[node];
            }
        }
    }
    const size_t plane = (size_t)cv.width * cv.height, pixel = (size_t)v * cv.width + u;
    const uint8_t *band_touch = touch + (size_t)(v / kSynthBandRows) * count;
    for (uint32_t s = 0; s < count; s++) {
        const SynthSourceDev src = sources[s];
        float2 q = make_float2(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000));
        float f = 0.0f, local = 0.0f, base = 0.0f, focus = 0.0f;
        float2 m;
        if (band_touch[s] && warp_map(src.warp, rx, ry, m) && m.x >= 0.0f && m.y >= 0.0f &&
            m.x <= (float)(src.mosaic_width - 1) && m.y <= (float)(src.mosaic_height - 1)) {
            float border = fminf(m.x, m.y);
            border = fminf(border, (float)(src.mosaic_width - 1) - m.x);
            border = fminf(border, (float)(src.mosaic_height - 1) - m.y);
            float t = keep_nan_clamp(border / cv.feather_scale, 0.0f, 1.0f);
            const float fe = t * t * (3.0f - 2.0f * t);
            const float lc = warp_confidence(src, rx, ry);
            if (!(fe <= 0.0f || lc <= 0.0f)) {
                q = m;
                f = fe;
                local = lc;
                base = fe * lc * src.magnification * src.magnification * src.source_confidence;
                focus = focus_consistency(has_depth, depth, depth_confidence, src);
            }
        }
        const size_t o = (size_t)s * plane + pixel;
        if (points) points[o] = q;
        if (feather) feather[o] = f;
        if (local_out) local_out[o] = local;
        if (base_out) base_out[o] = base;
        if (focus_out) focus_out[o] = focus;
    }
}

}  // namespace

hipError_t launch_synthesis_geometry(hipStream_t stream, const SynthCanvasDev &canvas,
                                     const DepthMapDev &map, const SynthSourceDev *sources,
                                     uint32_t count, const uint8_t *touch, float2 *points,
                                     float *feather, float *local_confidence, float *base_weight,
                                     float *focus_weight) {
    const dim3 block(256), grid((canvas.width + 255) / 256, canvas.height < 65535u ? canvas.height : 65535u);
    hipLaunchKernelGGL(k_synthesis_geometry, grid, block, 0, stream, canvas, map, sources, count,
                       touch, points, feather, local_confidence, base_weight, focus_weight);
    return hipGetLastError();
}

}  // namespace chd
