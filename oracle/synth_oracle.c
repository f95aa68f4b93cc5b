/* synth_oracle.c - CPU restatement of the depth- and warp-driven part of
 * synthesize (crates/chiaro-fusion/src/synth.rs:591-725) and of
 * DenseDepthMap::sample_nearest (crates/chiaro-fusion/src/depth.rs:152-159).
 *
 * Test infrastructure only: the checker of chd_synthesis_geometry. Built
 * with the rest of the oracle (no FMA contraction, Rust f32 semantics).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "depth_oracle.h"

#define SYNTH_BAND_ROWS 64 /* png16.rs:29, parallel.rs:26-43: 64-row bands */

/* `as usize` of an f32: NaN and negatives to 0, saturating above */
static size_t to_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return SIZE_MAX;
    return (size_t)f;
}

static float keep_nan_clamp(float v, float lo, float hi) { return v < lo ? lo : v > hi ? hi : v; }

/* smoothstep (synth.rs:1723-1726) */
static float smooth(float t) {
    t = keep_nan_clamp(t, 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* DenseDepthMap::sample_nearest (depth.rs:152-159) followed by node():
 * the node index, or 0 for None. */
int or_sample_nearest(const chd_depth_map *m, float x, float y, size_t *index) {
    if (x < 0.0f || y < 0.0f || m->step == 0) return 0;
    const size_t column = to_usize(roundf(x / (float)m->step));
    const size_t row = to_usize(roundf(y / (float)m->step));
    if (column >= m->columns || row >= m->rows) return 0;
    *index = row * m->columns + column;
    return 1;
}

/* Warp::confidence (align.rs:112-131); 0 without a confidence array */
static float warp_confidence_at(const chd_warp *w, float x, float y) {
    const float fx = x / (float)w->step, fy = y / (float)w->step;
    if (fx < 0.0f || fy < 0.0f || !w->confidence) return 0.0f;
    size_t c0 = to_usize(floorf(fx)), r0 = to_usize(floorf(fy));
    if (c0 > w->columns - 1) c0 = w->columns - 1;
    if (r0 > w->rows - 1) r0 = w->rows - 1;
    const size_t c1 = c0 + 1 < w->columns - 1 ? c0 + 1 : w->columns - 1;
    const size_t r1 = r0 + 1 < w->rows - 1 ? r0 + 1 : w->rows - 1;
    const float tx = fx - (float)c0, ty = fy - (float)r0;
    const float *v = w->confidence;
    const float top = v[r0 * w->columns + c0] * (1.0f - tx) + v[r0 * w->columns + c1] * tx;
    const float bottom = v[r1 * w->columns + c0] * (1.0f - tx) + v[r1 * w->columns + c1] * tx;
    return keep_nan_clamp(top * (1.0f - ty) + bottom * ty, 0.0f, 1.0f);
}

/* focus_consistency_weight (synth.rs:1469-1491). relative.powi(4) is
 * evaluated as two squarings: LLVM lowers a constant exponent by repeated
 * squaring, and compiler-rt's __powisf2 (binary exponentiation) rounds the
 * same two products for exponent 4. */
static float focus_consistency(int has_depth, double depth, float confidence,
                               const chd_synth_source *s) {
    if (s->reference || s->magnification <= 1.05f) return 1.0f;
    if (!has_depth || !s->has_focus_distance) return 1.0f;
    const double focus = s->focus_distance;
    if (!isfinite(depth) || !isfinite(focus) || depth <= 0.0 || focus <= 0.0) return 1.0f;
    const float mismatch = (float)(focus / depth) * (s->magnification - 1.0f);
    const float relative = mismatch / 0.75f;
    const float square = relative * relative;
    const float supported = 1.0f / (1.0f + square * square);
    return 1.0f - keep_nan_clamp(confidence, 0.0f, 1.0f) * (1.0f - supported);
}

/* band_touches_module (synth.rs:1688-1721) for canvas rows start..end */
int or_band_touches(const chd_synth_source *s, size_t start, size_t end, const chd_canvas *cv) {
    const chd_warp *w = &s->warp;
    const float margin = (float)(w->step * 2) * fmaxf(s->magnification, 1.0f);
    const size_t rows[3] = {start, (start + end) / 2, end > 0 ? end - 1 : 0};
    for (int k = 0; k < 3; k++) {
        const float ry = cv->crop_y + ((float)rows[k] + 0.5f) / cv->scale - 0.5f;
        size_t samples = to_usize(cv->crop_width / (float)w->step);
        if (samples < 8) samples = 8;
        samples += 1;
        for (size_t i = 0; i <= samples; i++) {
            const float rx = cv->crop_x + ((float)i / (float)samples) * (cv->crop_width - 1.0f);
            float q[2];
            if (or_warp_map(w, rx, ry, q) && q[0] >= -margin && q[1] >= -margin &&
                q[0] <= (float)s->mosaic_width + margin && q[1] <= (float)s->mosaic_height + margin)
                return 1;
        }
    }
    return 0;
}

/* synth.rs:591-620: the canvas geometry checks and size */
int or_canvas_size(const chd_canvas *cv, uint32_t *width, uint32_t *height) {
    if (cv->scale != cv->scale || cv->scale <= 0.0f || cv->crop_width < 1.0f ||
        cv->crop_height < 1.0f)
        return CHD_ERR_ARGUMENT;
    const float w = fmaxf(roundf(cv->crop_width * cv->scale), 1.0f);
    const float h = fmaxf(roundf(cv->crop_height * cv->scale), 1.0f);
    *width = (uint32_t)to_usize(w);
    *height = (uint32_t)to_usize(h);
    return CHD_OK;
}

/* The per-pixel source loop of synthesize up to its weights
 * (synth.rs:633-725). */
int or_synthesis_geometry(const chd_depth_map *map, const chd_synth_source *sources,
                          uint32_t count, const chd_canvas *cv, float *points, float *feather,
                          float *local_confidence, float *base_weight, float *focus_weight) {
    if (count == 0) return CHD_ERR_ARGUMENT;
    uint32_t width, height;
    if (or_canvas_size(cv, &width, &height) != CHD_OK) return CHD_ERR_ARGUMENT;
    const size_t plane = (size_t)width * height;
    const uint32_t bands = (height + SYNTH_BAND_ROWS - 1) / SYNTH_BAND_ROWS;
    uint8_t *touch = (uint8_t *)malloc((size_t)bands * count);
    for (uint32_t b = 0; b < bands; b++) {
        const size_t start = (size_t)b * SYNTH_BAND_ROWS;
        const size_t end = start + SYNTH_BAND_ROWS < height ? start + SYNTH_BAND_ROWS : height;
        for (uint32_t s = 0; s < count; s++)
            touch[(size_t)b * count + s] = (uint8_t)or_band_touches(&sources[s], start, end, cv);
    }
    const float feather_scale = fmaxf(cv->feather_px, 1.0f);
    for (uint32_t v = 0; v < height; v++) {
        const float ry = cv->crop_y + ((float)v + 0.5f) / cv->scale - 0.5f;
        const uint32_t band = v / SYNTH_BAND_ROWS;
        for (uint32_t u = 0; u < width; u++) {
            const float rx = cv->crop_x + ((float)u + 0.5f) / cv->scale - 0.5f;
            int has_depth = 0;
            double depth = 0.0;
            float depth_confidence = 0.0f;
            size_t node;
            if (map && or_sample_nearest(map, rx, ry, &node) && map->depth[node] == map->depth[node]) {
                has_depth = 1;
                depth = map->depth[node];
                depth_confidence = map->confidence[node];
            }
            for (uint32_t s = 0; s < count; s++) {
                const chd_synth_source *src = &sources[s];
                const size_t o = (size_t)s * plane + (size_t)v * width + u;
                float q[2] = {NAN, NAN}, f = 0.0f, local = 0.0f, base = 0.0f, focus = 0.0f;
                float m[2];
                if (touch[(size_t)band * count + s] && or_warp_map(&src->warp, rx, ry, m) &&
                    m[0] >= 0.0f && m[1] >= 0.0f && m[0] <= (float)(src->mosaic_width - 1) &&
                    m[1] <= (float)(src->mosaic_height - 1)) {
                    float border = fminf(m[0], m[1]);
                    border = fminf(border, (float)(src->mosaic_width - 1) - m[0]);
                    border = fminf(border, (float)(src->mosaic_height - 1) - m[1]);
                    const float fe = smooth(border / feather_scale);
                    const float lc = warp_confidence_at(&src->warp, rx, ry);
                    if (!(fe <= 0.0f || lc <= 0.0f)) {
                        q[0] = m[0];
                        q[1] = m[1];
                        f = fe;
                        local = lc;
                        base = fe * lc * src->magnification * src->magnification * src->confidence;
                        focus = focus_consistency(has_depth, depth, depth_confidence, src);
                    }
                }
                if (points) {
                    points[2 * o] = q[0];
                    points[2 * o + 1] = q[1];
                }
                if (feather) feather[o] = f;
                if (local_confidence) local_confidence[o] = local;
                if (base_weight) base_weight[o] = base;
                if (focus_weight) focus_weight[o] = focus;
            }
        }
    }
    free(touch);
    return CHD_OK;
}
