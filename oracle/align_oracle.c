/*
 * align_oracle.c - CPU restatement of align_module's refinement
 * (crates/chiaro-fusion/src/align.rs:270-519, 644-793) and match_patch
 * (crates/chiaro-fusion/src/image.rs:531-640), with Plane::downsample and
 * Plane::pyramid (image.rs:469-498) and apply_homography (math.rs:131-138).
 *
 * Test infrastructure only: the checker of the GPU alignment
 * (chd_align_module) and of chd_match_patches. Built like the rest of the
 * oracle: no FMA contraction, no fast-math; Rust's float semantics kept
 * (sequential sums starting from -0.0 for Iterator::sum, max_by returning
 * the last maximum, total_cmp sorts, saturating casts).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "depth_oracle.h"

typedef struct {
    float *data;
    uint32_t width, height;
} plane_t;

static float at(const plane_t *p, uint32_t x, uint32_t y) { return p->data[(size_t)y * p->width + x]; }

/* Plane::downsample (image.rs:469-484) */
static plane_t downsample(const plane_t *p) {
    plane_t out;
    out.width = p->width / 2 > 1 ? p->width / 2 : 1;
    out.height = p->height / 2 > 1 ? p->height / 2 : 1;
    out.data = (float *)malloc((size_t)out.width * out.height * sizeof(float));
    for (uint32_t j = 0; j < out.height; j++)
        for (uint32_t i = 0; i < out.width; i++) {
            uint32_t x = 2 * i, y = 2 * j;
            uint32_t x1 = x + 1 < p->width - 1 ? x + 1 : p->width - 1;
            uint32_t y1 = y + 1 < p->height - 1 ? y + 1 : p->height - 1;
            out.data[(size_t)j * out.width + i] =
                0.25f * (at(p, x, y) + at(p, x1, y) + at(p, x, y1) + at(p, x1, y1));
        }
    return out;
}

/* Plane::pyramid (image.rs:487-498); levels[0] aliases the input. */
static uint32_t pyramid(const plane_t *base, uint32_t min_size, plane_t *levels, uint32_t cap) {
    uint32_t n = 0;
    levels[n++] = *base;
    uint32_t floor_size = min_size > 8 ? min_size : 8;
    while (n < cap) {
        const plane_t *last = &levels[n - 1];
        uint32_t short_side = last->width < last->height ? last->width : last->height;
        if (short_side / 2 < floor_size) break;
        levels[n] = downsample(last);
        n++;
    }
    return n;
}

uint32_t or_pyramid_levels(uint32_t width, uint32_t height, uint32_t min_size) {
    uint32_t n = 1, floor_size = min_size > 8 ? min_size : 8;
    while (n < 32) {
        uint32_t s = width < height ? width : height;
        if (s / 2 < floor_size) break;
        width = width / 2 > 1 ? width / 2 : 1;
        height = height / 2 > 1 ? height / 2 : 1;
        n++;
    }
    return n;
}

/* Plane::sample (image.rs:452-466) */
static int sample(const plane_t *p, float x, float y, float *out) {
    return or_plane_sample(p->data, p->width, p->height, x, y, out);
}

static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
/* f32::total_cmp as an integer key */
static inline int64_t tkey(float f) {
    int32_t k = (int32_t)f2u(f);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}

/* match_patch (image.rs:531-640) */
int or_match_patch(const float *reference, uint32_t rw, uint32_t rh, const float *target,
                   uint32_t tw, uint32_t th, uint32_t rx, uint32_t ry, uint32_t size,
                   uint32_t radius, float shift[2], float *score_out) {
    if (rx + size > rw || ry + size > rh || rx < radius || ry < radius ||
        rx + size + radius > tw || ry + size + radius > th || rx + size > tw || ry + size > th)
        return 0;
    const float n = (float)(size * size);
    float *ref_centered = (float *)malloc((size_t)size * size * sizeof(float));
    float sum = -0.0f;
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++) sum += reference[(size_t)(ry + y) * rw + rx + x];
    const float ref_mean = sum / n;
    float sq = -0.0f;
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++) {
            float d = reference[(size_t)(ry + y) * rw + rx + x] - ref_mean;
            sq += d * d;
        }
    const float ref_std = sqrtf(sq / n);
    if (ref_std < 1e-4f) {
        free(ref_centered);
        return 0;
    }
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++)
            ref_centered[y * size + x] = reference[(size_t)(ry + y) * rw + rx + x] - ref_mean;
    const uint32_t span = 2 * radius + 1;
    float *scores = (float *)malloc((size_t)span * span * sizeof(float));
    for (uint32_t i = 0; i < span * span; i++) scores[i] = -INFINITY;
    for (uint32_t dy = 0; dy < span; dy++)
        for (uint32_t dx = 0; dx < span; dx++) {
            uint32_t tx = rx + dx - radius, ty = ry + dy - radius;
            float s = 0.0f, s2 = 0.0f, cross = 0.0f;
            for (uint32_t y = 0; y < size; y++) {
                const float *row = target + (size_t)(ty + y) * tw + tx;
                const float *rrow = ref_centered + (size_t)y * size;
                for (uint32_t x = 0; x < size; x++) {
                    float t = row[x];
                    s += t;
                    s2 += t * t;
                    cross += t * rrow[x];
                }
            }
            float mean = s / n;
            float var = fmaxf(s2 / n - mean * mean, 0.0f); /* f32::max: NaN gives 0 */
            if (var < 1e-8f) continue;
            float covariance = cross / n;
            scores[dy * span + dx] = covariance / (sqrtf(var) * ref_std);
        }
    /* max_by(total_cmp): the last maximum */
    uint32_t best = 0;
    for (uint32_t i = 1; i < span * span; i++)
        if (tkey(scores[i]) >= tkey(scores[best])) best = i;
    float score = scores[best];
    int ok = isfinite(score);
    if (ok) {
        uint32_t bx = best % span, by = best / span;
        float sub_x = 0.0f, sub_y = 0.0f;
#define REFINE(m, c, p, out)                                             \
    do {                                                                 \
        float den = (m) - 2.0f * (c) + (p);                              \
        if (fabsf(den) < 1e-9f) out = 0.0f;                              \
        else {                                                           \
            float v = 0.5f * ((m) - (p)) / den;                          \
            out = v < -0.5f ? -0.5f : v > 0.5f ? 0.5f : v;               \
        }                                                                \
    } while (0)
        if (bx > 0 && bx + 1 < span) REFINE(scores[by * span + bx - 1], score, scores[by * span + bx + 1], sub_x);
        if (by > 0 && by + 1 < span) REFINE(scores[(by - 1) * span + bx], score, scores[(by + 1) * span + bx], sub_y);
#undef REFINE
        shift[0] = ((float)bx - (float)radius) + sub_x;
        shift[1] = ((float)by - (float)radius) + sub_y;
        *score_out = score;
    }
    free(scores);
    free(ref_centered);
    return ok;
}

/* ------------------------------------------------------- homographies */

static void mat_mul(const double *a, const double *b, double *out) {
    double r[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] +
                           a[i * 3 + 2] * b[2 * 3 + j];
    memcpy(out, r, sizeof r);
}

/* apply_homography (math.rs:131-138) */
int or_apply_homography(const double *h, const double p[2], double out[2]) {
    double v[3] = {h[0] * p[0] + h[1] * p[1] + h[2] * 1.0, h[3] * p[0] + h[4] * p[1] + h[5] * 1.0,
                   h[6] * p[0] + h[7] * p[1] + h[8] * 1.0};
    if (fabs(v[2]) < 1e-12) return 0;
    out[0] = v[0] / v[2];
    out[1] = v[1] / v[2];
    return 1;
}

static inline int64_t tkey64(double d) {
    int64_t k;
    memcpy(&k, &d, 8);
    return k ^ (int64_t)(((uint64_t)(k >> 63)) >> 1);
}

/* solve_linear (align.rs:765-793) */
static int solve_linear(double a[8][8], double b[8], double x[8]) {
    for (int column = 0; column < 8; column++) {
        int pivot = column; /* max_by over |a[i][column]|: the last maximum */
        for (int i = column + 1; i < 8; i++)
            if (tkey64(fabs(a[i][column])) >= tkey64(fabs(a[pivot][column]))) pivot = i;
        if (fabs(a[pivot][column]) < 1e-12) return 0;
        for (int k = 0; k < 8; k++) {
            double t = a[column][k];
            a[column][k] = a[pivot][k];
            a[pivot][k] = t;
        }
        double t = b[column];
        b[column] = b[pivot];
        b[pivot] = t;
        double pivot_row[8];
        memcpy(pivot_row, a[column], sizeof pivot_row);
        for (int row = column + 1; row < 8; row++) {
            double factor = a[row][column] / pivot_row[column];
            for (int k = column; k < 8; k++) a[row][k] -= factor * pivot_row[k];
            b[row] -= factor * b[column];
        }
    }
    for (int row = 7; row >= 0; row--) {
        double sum = b[row];
        for (int k = row + 1; k < 8; k++) sum -= a[row][k] * x[k];
        x[row] = sum / a[row][row];
    }
    return 1;
}

/* fit_homography_least_squares (align.rs:713-762); pairs as
 * (ax, ay, bx, by, score) */
int or_fit_homography_least_squares(const double *pairs, const float *scores, uint32_t count,
                                    double h[9]) {
    if (count < 4) return 0;
    const double n = (double)count;
    double s_ax = -0.0, s_ay = -0.0, s_bx = -0.0, s_by = -0.0;
    for (uint32_t i = 0; i < count; i++) s_ax += pairs[i * 4 + 0];
    for (uint32_t i = 0; i < count; i++) s_ay += pairs[i * 4 + 1];
    for (uint32_t i = 0; i < count; i++) s_bx += pairs[i * 4 + 2];
    for (uint32_t i = 0; i < count; i++) s_by += pairs[i * 4 + 3];
    const double ax = s_ax / n, ay = s_ay / n, bx = s_bx / n, by = s_by / n;
    double s_a = -0.0, s_b = -0.0;
    for (uint32_t i = 0; i < count; i++) {
        double dx = pairs[i * 4 + 0] - ax, dy = pairs[i * 4 + 1] - ay;
        s_a += sqrt(dx * dx + dy * dy);
    }
    for (uint32_t i = 0; i < count; i++) {
        double dx = pairs[i * 4 + 2] - bx, dy = pairs[i * 4 + 3] - by;
        s_b += sqrt(dx * dx + dy * dy);
    }
    double sa = s_a / n, sb = s_b / n;
    sa = fmax(sa, 1e-9); /* f64::max: NaN gives the other operand */
    sb = fmax(sb, 1e-9);
    double ata[8][8], atb[8];
    memset(ata, 0, sizeof ata);
    memset(atb, 0, sizeof atb);
    for (uint32_t i = 0; i < count; i++) {
        double w = fmax((double)scores[i], 0.05);
        double x = (pairs[i * 4 + 0] - ax) / sa, y = (pairs[i * 4 + 1] - ay) / sa;
        double u = (pairs[i * 4 + 2] - bx) / sb, v = (pairs[i * 4 + 3] - by) / sb;
        double rows[2][8] = {{x, y, 1.0, 0.0, 0.0, 0.0, -u * x, -u * y},
                             {0.0, 0.0, 0.0, x, y, 1.0, -v * x, -v * y}};
        double rhs[2] = {u, v};
        for (int r = 0; r < 2; r++)
            for (int a = 0; a < 8; a++) {
                atb[a] += w * rows[r][a] * rhs[r];
                for (int b = 0; b < 8; b++) ata[a][b] += w * rows[r][a] * rows[r][b];
            }
    }
    double sol[8];
    if (!solve_linear(ata, atb, sol)) return 0;
    const double normalized[9] = {sol[0], sol[1], sol[2], sol[3], sol[4], sol[5], sol[6], sol[7], 1.0};
    const double t_a[9] = {1.0 / sa, 0.0, -ax / sa, 0.0, 1.0 / sa, -ay / sa, 0.0, 0.0, 1.0};
    const double t_b_inv[9] = {sb, 0.0, bx, 0.0, sb, by, 0.0, 0.0, 1.0};
    double tmp[9], hh[9];
    mat_mul(t_b_inv, normalized, tmp);
    mat_mul(tmp, t_a, hh);
    double scale = hh[8];
    if (fabs(scale) < 1e-12 || !isfinite(scale)) return 0;
    for (int k = 0; k < 9; k++) h[k] = hh[k] / scale;
    return 1;
}

static int cmp_f32_total(const void *a, const void *b) {
    int64_t x = tkey(*(const float *)a), y = tkey(*(const float *)b);
    return x < y ? -1 : x > y;
}

/* fit_homography_ransac (align.rs:644-710); residuals receives the sorted
 * inlier residuals (count of them in *inliers). */
int or_fit_homography_ransac(const double *pairs, const float *scores, uint32_t count,
                             float threshold, double h[9], uint32_t *inliers, float *residuals) {
    if (count < 6) return 0;
    uint64_t state = 0x9E3779B97F4A7C15ull ^ (uint64_t)count;
    uint32_t *best = (uint32_t *)malloc(count * sizeof(uint32_t));
    uint32_t *current = (uint32_t *)malloc(count * sizeof(uint32_t));
    double *subset = (double *)malloc(count * 4 * sizeof(double));
    float *subset_scores = (float *)malloc(count * sizeof(float));
    uint32_t best_count = 0;
    for (int iteration = 0; iteration < 300; iteration++) {
        uint32_t sample[4];
        for (int slot = 0; slot < 4; slot++) {
            for (;;) {
                state = state * 6364136223846793005ull + 1442695040888963407ull;
                uint32_t candidate = (uint32_t)((state >> 33) % (uint64_t)count);
                int seen = 0;
                for (int k = 0; k < slot; k++) seen |= sample[k] == candidate;
                if (!seen) {
                    sample[slot] = candidate;
                    break;
                }
            }
        }
        for (int k = 0; k < 4; k++) {
            memcpy(subset + k * 4, pairs + (size_t)sample[k] * 4, 4 * sizeof(double));
            subset_scores[k] = scores[sample[k]];
        }
        double hs[9];
        if (!or_fit_homography_least_squares(subset, subset_scores, 4, hs)) continue;
        uint32_t c = 0;
        for (uint32_t i = 0; i < count; i++) {
            double p[2];
            if (or_apply_homography(hs, pairs + (size_t)i * 4, p)) {
                double dx = p[0] - pairs[i * 4 + 2], dy = p[1] - pairs[i * 4 + 3];
                if (sqrt(dx * dx + dy * dy) <= (double)threshold) current[c++] = i;
            }
        }
        if (c > best_count) {
            memcpy(best, current, c * sizeof(uint32_t));
            best_count = c;
        }
    }
    int ok = 0;
    if (best_count >= 6) {
        for (uint32_t k = 0; k < best_count; k++) {
            memcpy(subset + k * 4, pairs + (size_t)best[k] * 4, 4 * sizeof(double));
            subset_scores[k] = scores[best[k]];
        }
        double hf[9];
        if (or_fit_homography_least_squares(subset, subset_scores, best_count, hf)) {
            uint32_t c = 0;
            for (uint32_t i = 0; i < count; i++) {
                double p[2];
                if (or_apply_homography(hf, pairs + (size_t)i * 4, p)) {
                    double dx = p[0] - pairs[i * 4 + 2], dy = p[1] - pairs[i * 4 + 3];
                    float r = (float)sqrt(dx * dx + dy * dy);
                    if (r <= threshold) residuals[c++] = r;
                }
            }
            if (c >= 6) {
                qsort(residuals, c, sizeof(float), cmp_f32_total);
                memcpy(h, hf, sizeof hf);
                *inliers = c;
                ok = 1;
            }
        }
    }
    free(best);
    free(current);
    free(subset);
    free(subset_scores);
    return ok;
}

/* ------------------------------------------------------- align_module */

static int initial_map(const chd_warp *grid, double px, double py, double out[2]) {
    float q[2];
    if (!or_warp_map(grid, (float)px, (float)py, q)) return 0;
    out[0] = (double)q[0];
    out[1] = (double)q[1];
    return 1;
}

/* the initial mapping through the correction: initial(correction(p)) */
static int corrected_map(const chd_warp *grid, const double *correction, double px, double py,
                         double out[2]) {
    double p[2] = {px, py}, c[2];
    if (!or_apply_homography(correction, p, c)) return 0;
    return initial_map(grid, c[0], c[1], out);
}

/* render_through (align.rs:523-547) */
static void render_through(const plane_t *reference_plane, uint32_t reference_scale,
                           const plane_t *target_plane, double target_scale, const chd_warp *grid,
                           const double *correction, float *out) {
    for (uint32_t y = 0; y < reference_plane->height; y++)
        for (uint32_t x = 0; x < reference_plane->width; x++) {
            double p[2] = {((double)(float)x * (double)reference_scale * 2.0) +
                               (double)reference_scale - 0.5,
                           ((double)(float)y * (double)reference_scale * 2.0) +
                               (double)reference_scale - 0.5};
            float value = NAN, v;
            double q[2];
            if (corrected_map(grid, correction, p[0], p[1], q) &&
                sample(target_plane, (float)((q[0] + 0.5) / (2.0 * target_scale) - 0.5),
                       (float)((q[1] + 0.5) / (2.0 * target_scale) - 0.5), &v))
                value = v;
            out[(size_t)y * reference_plane->width + x] = value;
        }
}

static int covered(const float *rendered, uint32_t width, uint32_t x, uint32_t y, uint32_t size) {
    for (uint32_t row = y; row < y + size; row++)
        for (uint32_t col = x; col < x + size; col++)
            if (rendered[(size_t)row * width + col] != rendered[(size_t)row * width + col]) return 0;
    return 1;
}

void or_default_align_options(chd_align_options *o) {
    o->grid_step = 32;
    o->patch = 32;
    o->coarse_radius = 12;
    o->min_score = 0.5f;
    o->inlier_px = 3.0f;
    o->min_inlier_ratio = 0.45f;
    o->refine = 1;
}

/* photometric_gain (align.rs:795-826) */
static int photometric_gain(const plane_t *reference, const plane_t *target, const chd_warp *warp,
                            float *gain) {
    size_t cap = ((size_t)reference->width / 16 + 1) * ((size_t)reference->height / 16 + 1);
    float *ratios = (float *)malloc(cap * sizeof(float));
    size_t n = 0;
    for (uint32_t y = 16; y + 16 < reference->height; y += 16)
        for (uint32_t x = 16; x + 16 < reference->width; x += 16) {
            float rx = (float)x * 2.0f + 0.5f, ry = (float)y * 2.0f + 0.5f, q[2], t;
            if (!or_warp_map(warp, rx, ry, q)) continue;
            if (!sample(target, (q[0] - 0.5f) / 2.0f, (q[1] - 0.5f) / 2.0f, &t)) continue;
            float r = at(reference, x, y);
            float lr = (expf(r) - 1.0f) / 1000.0f, lt = (expf(t) - 1.0f) / 1000.0f;
            if (lr > 0.01f && lt > 0.01f && lr < 0.95f && lt < 0.95f) ratios[n++] = lr / lt;
        }
    int ok = n >= 16;
    if (ok) {
        qsort(ratios, n, sizeof(float), cmp_f32_total);
        float g = ratios[n / 2];
        *gain = g < 0.25f ? 0.25f : g > 4.0f ? 4.0f : g;
    }
    free(ratios);
    return ok;
}

static void warp_fill(uint32_t width, uint32_t height, uint32_t step, chd_warp *w,
                      const chd_warp *grid, const double *correction) {
    w->step = step;
    w->columns = (width + step - 1) / step + 1;
    w->rows = (height + step - 1) / step + 1;
    for (uint32_t r = 0; r < w->rows; r++)
        for (uint32_t c = 0; c < w->columns; c++) {
            size_t i = (size_t)r * w->columns + c;
            double q[2];
            int ok = grid ? corrected_map(grid, correction, (double)((size_t)c * step),
                                          (double)((size_t)r * step), q)
                          : (q[0] = (double)((size_t)c * step), q[1] = (double)((size_t)r * step), 1);
            w->points[2 * i] = ok ? (float)q[0] : NAN;
            w->points[2 * i + 1] = ok ? (float)q[1] : NAN;
            w->confidence[i] = ok ? 1.0f : 0.0f;
        }
}

int or_align_module(const chd_view *reference, const chd_view *target,
                    const chd_align_options *o, chd_warp *warp, chd_alignment_report *report) {
    memset(report, 0, sizeof *report);
    const uint32_t width = reference->width, height = reference->height;
    if (reference == target) { /* align.rs:291-309 */
        warp_fill(width, height, o->grid_step, warp, NULL, NULL);
        report->initialised_from = CHD_ALIGN_FROM_REFERENCE;
        report->coverage = 1.0f;
        report->inlier_ratio = 1.0f;
        report->accepted = 1;
        report->status = CHD_ALIGN_STATUS_REFERENCE;
        report->gain = 1.0f;
        return CHD_OK;
    }
    if (!reference->camera || !target->camera) return CHD_ERR_ARGUMENT;
    report->initialised_from = CHD_ALIGN_FROM_CALIBRATION;
    /* Warp::from_fn at 8 px of map_from(.., 1e8) (align.rs:317-324) */
    chd_warp grid;
    grid.step = 8;
    grid.columns = (width + 7) / 8 + 1;
    grid.rows = (height + 7) / 8 + 1;
    grid.points = (float *)malloc((size_t)grid.columns * grid.rows * 2 * sizeof(float));
    grid.confidence = (float *)malloc((size_t)grid.columns * grid.rows * sizeof(float));
    for (uint32_t r = 0; r < grid.rows; r++)
        for (uint32_t c = 0; c < grid.columns; c++) {
            size_t i = (size_t)r * grid.columns + c;
            double p[2] = {(double)((size_t)c * 8), (double)((size_t)r * 8)}, q[2];
            int ok = or_map_from(target->camera, reference->camera, p, 1.0e8, q) &&
                     isfinite(q[0]) && isfinite(q[1]);
            grid.points[2 * i] = ok ? (float)q[0] : NAN;
            grid.points[2 * i + 1] = ok ? (float)q[1] : NAN;
            grid.confidence[i] = ok ? 1.0f : 0.0f;
        }
    double correction[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float *finest = NULL;
    uint32_t finest_count = 0;
    plane_t ref_base = {(float *)reference->luminance, reference->plane_width, reference->plane_height};
    plane_t tgt_base = {(float *)target->luminance, target->plane_width, target->plane_height};
    if (o->refine) {
        plane_t rp[32], tp[32];
        uint32_t rn = pyramid(&ref_base, 96, rp, 32), tn = pyramid(&tgt_base, 48, tp, 32);
        double magnification;
        {
            double c[2] = {(double)width / 2.0, (double)height / 2.0}, hh = 64.0, a[2], b[2], d[2];
            if (initial_map(&grid, c[0], c[1], a) && initial_map(&grid, c[0] + hh, c[1], b) &&
                initial_map(&grid, c[0], c[1] + hh, d)) {
                double dx = sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / hh;
                double dy = sqrt((d[0] - a[0]) * (d[0] - a[0]) + (d[1] - a[1]) * (d[1] - a[1])) / hh;
                magnification = sqrt(dx * dy);
            } else {
                magnification = target->camera->focal_px / reference->camera->focal_px;
            }
        }
        for (int level = (int)rn - 1; level >= 0; level--) {
            uint32_t scale = 1u << level;
            const plane_t *reference_plane = &rp[level];
            /* log2().round().max(0.0) as usize, then min(len - 1) */
            double l2 = fmax(round(log2((double)scale * magnification)), 0.0);
            uint32_t target_level = l2 >= (double)(tn - 1) ? tn - 1 : (uint32_t)l2;
            const plane_t *target_plane = &tp[target_level];
            double target_scale = (double)(1u << target_level);
            float *rendered = (float *)malloc((size_t)reference_plane->width *
                                              reference_plane->height * sizeof(float));
            render_through(reference_plane, scale, target_plane, target_scale, &grid, correction,
                           rendered);
            uint32_t radius = level + 1 == (int)rn ? o->coarse_radius : 3;
            uint32_t patch = o->patch < reference_plane->width / 4 ? o->patch : reference_plane->width / 4;
            if (patch < 8) patch = 8;
            uint32_t stride = patch / 2;
            size_t cap = 16;
            size_t count = 0;
            double *pairs = (double *)malloc(cap * 4 * sizeof(double));
            float *scores = (float *)malloc(cap * sizeof(float));
            for (uint32_t y = radius; y + patch + radius <= reference_plane->height; y += stride)
                for (uint32_t x = radius; x + patch + radius <= reference_plane->width; x += stride) {
                    float shift[2], score;
                    if (!covered(rendered, reference_plane->width, x, y, patch)) continue;
                    if (!or_match_patch(reference_plane->data, reference_plane->width,
                                        reference_plane->height, rendered, reference_plane->width,
                                        reference_plane->height, x, y, patch, radius, shift, &score))
                        continue;
                    if (!(score >= o->min_score)) continue;
                    if (count == cap) {
                        cap *= 2;
                        pairs = (double *)realloc(pairs, cap * 4 * sizeof(double));
                        scores = (float *)realloc(scores, cap * sizeof(float));
                    }
                    float cx = (float)x + (float)patch / 2.0f, cy = (float)y + (float)patch / 2.0f;
                    float sx = cx + shift[0], sy = cy + shift[1];
                    double s = (double)scale;
                    pairs[count * 4 + 0] = (double)cx * s * 2.0 + s - 0.5;
                    pairs[count * 4 + 1] = (double)cy * s * 2.0 + s - 0.5;
                    pairs[count * 4 + 2] = (double)sx * s * 2.0 + s - 0.5;
                    pairs[count * 4 + 3] = (double)sy * s * 2.0 + s - 0.5;
                    scores[count] = score;
                    count++;
                }
            free(rendered);
            float threshold = o->inlier_px * (float)scale * 2.0f;
            if (!(threshold >= o->inlier_px)) threshold = o->inlier_px;
            double update[9];
            uint32_t inliers = 0;
            float *residuals = (float *)malloc((count + 1) * sizeof(float));
            chd_level_report *lr = report->level_count < CHD_MAX_ALIGN_LEVELS
                                       ? &report->levels[report->level_count++]
                                       : NULL;
            if (or_fit_homography_ransac(pairs, scores, (uint32_t)count, threshold, update, &inliers,
                                         residuals)) {
                mat_mul(correction, update, correction);
                if (lr) {
                    lr->scale = scale * 2;
                    lr->patches = (uint32_t)count;
                    lr->inliers = inliers;
                    lr->median_residual_px = residuals[inliers / 2];
                }
                if (level == 0) {
                    free(finest);
                    finest = residuals;
                    residuals = NULL;
                    finest_count = inliers;
                    report->patches = (uint32_t)count;
                    report->inliers = inliers;
                }
            } else if (lr) {
                lr->scale = scale * 2;
                lr->patches = (uint32_t)count;
                lr->inliers = 0;
                lr->median_residual_px = NAN;
            }
            free(residuals);
            free(pairs);
            free(scores);
        }
        for (uint32_t k = 1; k < rn; k++) free(rp[k].data);
        for (uint32_t k = 1; k < tn; k++) free(tp[k].data);
    }
    warp_fill(width, height, o->grid_step, warp, &grid, correction);
    size_t covered_count = 0, total = (size_t)warp->columns * warp->rows;
    for (size_t i = 0; i < total; i++) {
        float x = warp->points[2 * i], y = warp->points[2 * i + 1];
        if (isfinite(x) && x >= 0.0f && y >= 0.0f && x <= (float)(target->width - 1) &&
            y <= (float)(target->height - 1))
            covered_count++;
    }
    report->coverage = (float)covered_count / (float)total;
    if (finest_count > 0) {
        report->residual_median_px = finest[finest_count / 2];
        size_t p90 = (size_t)finest_count * 9 / 10;
        if (p90 > finest_count - 1) p90 = finest_count - 1;
        report->residual_p90_px = finest[p90];
    }
    {
        double centre[2] = {(double)(width / 2), (double)(height / 2)}, moved[2];
        if (!or_apply_homography(correction, centre, moved)) {
            moved[0] = centre[0];
            moved[1] = centre[1];
        }
        report->correction_median_px[0] = (float)(moved[0] - centre[0]);
        report->correction_median_px[1] = (float)(moved[1] - centre[1]);
    }
    report->inlier_ratio = report->patches == 0 ? 0.0f : (float)report->inliers / (float)report->patches;
    report->accepted = !o->refine || report->inlier_ratio >= o->min_inlier_ratio;
    if (!report->accepted) report->status = CHD_ALIGN_STATUS_REJECTED;
    else if (!o->refine) report->status = CHD_ALIGN_STATUS_FACTORY_MODEL;
    else if (report->inliers >= 12 && report->residual_median_px < o->inlier_px)
        report->status = CHD_ALIGN_STATUS_REFINED;
    else if (report->inliers > 0) report->status = CHD_ALIGN_STATUS_WEAK;
    else report->status = CHD_ALIGN_STATUS_NO_SUPPORT;
    float gain;
    report->gain = photometric_gain(&ref_base, &tgt_base, warp, &gain) ? gain : 1.0f;
    free(finest);
    free(grid.points);
    free(grid.confidence);
    return CHD_OK;
}

/* -------------------------------------------- refine_resolution_warp */

#define LOCAL_WARP_STEP 32
#define LOCAL_MATCH_PATCH 12
#define LOCAL_MATCH_RADIUS 2
#define MAGNIFIED_MATCH_RADIUS 5
#define LOCAL_MINIMUM_SCORE 0.72f

static float clamp01(float v) { return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v; }

/* smoothstep (resolution.rs:485-489) */
static float smoothstep(float v) {
    v = clamp01(v);
    return v * v * (3.0f - 2.0f * v);
}

/* Warp::magnification (align.rs:135-143) */
static int warp_magnification(const chd_warp *w, float x, float y, float *out) {
    float h = (float)w->step, a[2], b[2], c[2];
    if (!or_warp_map(w, x, y, a) || !or_warp_map(w, x + h, y, b) || !or_warp_map(w, x, y + h, c))
        return 0;
    float dx = sqrtf((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / h;
    float dy = sqrtf((c[0] - a[0]) * (c[0] - a[0]) + (c[1] - a[1]) * (c[1] - a[1])) / h;
    *out = fmaxf(sqrtf(dx * dy), 1e-6f);
    return 1;
}

/* Warp::confidence (align.rs:112-131) */
static float warp_confidence(const chd_warp *w, float x, float y) {
    float fx = x / (float)w->step, fy = y / (float)w->step;
    if (fx < 0.0f || fy < 0.0f || !w->confidence) return 0.0f;
    float flx = floorf(fx), fly = floorf(fy);
    uint32_t c0 = !(flx > 0.0f) ? 0 : (flx >= 4294967296.0f ? UINT32_MAX : (uint32_t)flx);
    uint32_t r0 = !(fly > 0.0f) ? 0 : (fly >= 4294967296.0f ? UINT32_MAX : (uint32_t)fly);
    if (c0 > w->columns - 1) c0 = w->columns - 1;
    if (r0 > w->rows - 1) r0 = w->rows - 1;
    uint32_t c1 = c0 + 1 < w->columns - 1 ? c0 + 1 : w->columns - 1;
    uint32_t r1 = r0 + 1 < w->rows - 1 ? r0 + 1 : w->rows - 1;
    float tx = fx - (float)c0, ty = fy - (float)r0;
#define V(c, r) w->confidence[(size_t)(r) * w->columns + (c)]
    float top = V(c0, r0) * (1.0f - tx) + V(c1, r0) * tx;
    float bottom = V(c0, r1) * (1.0f - tx) + V(c1, r1) * tx;
#undef V
    float v = top * (1.0f - ty) + bottom * ty;
    return v < 0.0f ? 0.0f : v > 1.0f ? 1.0f : v;
}

/* sample_at_reference_bandwidth (resolution.rs:441-465) */
static int sample_at_bandwidth(const plane_t *t, float x, float y, float magnification, float *out) {
    if (magnification <= 1.15f) return sample(t, x, y, out);
    float sigma = fmaxf(0.42f * magnification, 0.55f);
    float rr = ceilf(2.0f * sigma);
    rr = rr < 1.0f ? 1.0f : rr > 4.0f ? 4.0f : rr;
    int radius = (int)rr;
    float sum = 0.0f, weight = 0.0f;
    for (int dy = -radius; dy <= radius; dy++)
        for (int dx = -radius; dx <= radius; dx++) {
            float distance_squared = (float)(dx * dx + dy * dy);
            float kernel = expf(-0.5f * distance_squared / (sigma * sigma));
            float v;
            if (sample(t, x + (float)dx, y + (float)dy, &v)) {
                sum += kernel * v;
                weight += kernel;
            }
        }
    if (!(weight > 1.0e-8f)) return 0;
    *out = sum / weight;
    return 1;
}

static int window_finite(const float *r, uint32_t w, size_t x0, size_t y0, size_t x1, size_t y1) {
    for (size_t row = y0; row < y1; row++)
        for (size_t col = x0; col < x1; col++)
            if (!isfinite(r[row * w + col])) return 0;
    return 1;
}

/* f32::round: half away from zero */
static float round_half_away(float v) { return roundf(v); }

/* refine_resolution_warp (resolution.rs:232-432) */
int or_refine_resolution_warp(const float *reference, uint32_t rw, uint32_t rh,
                              const float *target, uint32_t tw, uint32_t th, const chd_warp *base,
                              uint32_t width, uint32_t height, chd_warp *out,
                              chd_resolution_report *report) {
    plane_t tp = {(float *)target, tw, th};
    float centre[2] = {(float)width * 0.5f, (float)height * 0.5f}, mag;
    if (!warp_magnification(base, centre[0], centre[1], &mag)) mag = 1.0f;
    mag = mag < 0.5f ? 0.5f : mag > 4.0f ? 4.0f : mag;
    uint32_t radius = mag > 1.35f ? MAGNIFIED_MATCH_RADIUS : LOCAL_MATCH_RADIUS;
    float *rendered = (float *)malloc((size_t)rw * rh * sizeof(float));
    for (uint32_t y = 0; y < rh; y++)
        for (uint32_t x = 0; x < rw; x++) {
            float rx = (float)x * 2.0f + 0.5f, ry = (float)y * 2.0f + 0.5f, q[2], v = NAN;
            if (or_warp_map(base, rx, ry, q)) {
                float s;
                if (sample_at_bandwidth(&tp, (q[0] + 0.5f) * 0.5f - 0.5f, (q[1] + 0.5f) * 0.5f - 0.5f,
                                        mag, &s))
                    v = s;
            }
            rendered[(size_t)y * rw + x] = v;
        }
    uint32_t columns = (width + LOCAL_WARP_STEP - 1) / LOCAL_WARP_STEP + 1;
    uint32_t rows = (height + LOCAL_WARP_STEP - 1) / LOCAL_WARP_STEP + 1;
    size_t nodes = (size_t)columns * rows;
    uint8_t *has = (uint8_t *)calloc(nodes, 1);
    float *mshift = (float *)calloc(nodes * 2, sizeof(float));
    float *mscore = (float *)calloc(nodes, sizeof(float));
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            float rx = (float)(column * LOCAL_WARP_STEP), ry = (float)(row * LOCAL_WARP_STEP);
            if (rx >= (float)width || ry >= (float)height) continue;
            long plane_x = (long)round_half_away((rx - 0.5f) * 0.5f);
            long plane_y = (long)round_half_away((ry - 0.5f) * 0.5f);
            long x = plane_x - LOCAL_MATCH_PATCH / 2, y = plane_y - LOCAL_MATCH_PATCH / 2;
            if (x < (long)radius || y < (long)radius ||
                x + LOCAL_MATCH_PATCH + (long)radius > (long)rw ||
                y + LOCAL_MATCH_PATCH + (long)radius > (long)rh)
                continue;
            if (!window_finite(rendered, rw, (size_t)(x - radius), (size_t)(y - radius),
                               (size_t)(x + LOCAL_MATCH_PATCH + radius),
                               (size_t)(y + LOCAL_MATCH_PATCH + radius)))
                continue;
            float shift[2], score;
            if (!or_match_patch(reference, rw, rh, rendered, rw, rh, (uint32_t)x, (uint32_t)y,
                                LOCAL_MATCH_PATCH, radius, shift, &score))
                continue;
            if (!(score >= LOCAL_MINIMUM_SCORE)) continue;
            size_t i = (size_t)row * columns + column;
            has[i] = 1;
            mshift[2 * i] = shift[0] * 2.0f;
            mshift[2 * i + 1] = shift[1] * 2.0f;
            mscore[i] = score;
        }
    free(rendered);
    uint8_t *cgrid = (uint8_t *)calloc(nodes, 1);
    float *correction = (float *)calloc(nodes * 2, sizeof(float));
    float *corrections = (float *)malloc((nodes + 1) * sizeof(float));
    size_t ncorr = 0;
    float confidence_sum = 0.0f;
    out->step = LOCAL_WARP_STEP;
    out->columns = columns;
    out->rows = rows;
    for (uint32_t row = 0; row < rows; row++)
        for (uint32_t column = 0; column < columns; column++) {
            float rx = (float)(column * LOCAL_WARP_STEP), ry = (float)(row * LOCAL_WARP_STEP);
            size_t index = (size_t)row * columns + column;
            int refined = 0;
            float point[2], local_conf = 0.0f, med[2];
            if (has[index]) {
                float xs[9], ys[9];
                int n = 0;
                uint32_t r0 = row > 0 ? row - 1 : 0, r1 = row + 1 < rows - 1 ? row + 1 : rows - 1;
                uint32_t c0 = column > 0 ? column - 1 : 0, c1 = column + 1 < columns - 1 ? column + 1 : columns - 1;
                for (uint32_t nr = r0; nr <= r1; nr++)
                    for (uint32_t nc = c0; nc <= c1; nc++) {
                        size_t j = (size_t)nr * columns + nc;
                        if (has[j]) {
                            xs[n] = mshift[2 * j];
                            ys[n] = mshift[2 * j + 1];
                            n++;
                        }
                    }
                if (n >= 3) {
                    qsort(xs, n, sizeof(float), cmp_f32_total);
                    qsort(ys, n, sizeof(float), cmp_f32_total);
                    med[0] = xs[n / 2];
                    med[1] = ys[n / 2];
                    if (!(hypotf(mshift[2 * index] - med[0], mshift[2 * index + 1] - med[1]) > 1.5f) &&
                        or_warp_map(base, rx + med[0], ry + med[1], point)) {
                        float sc = smoothstep((mscore[index] - LOCAL_MINIMUM_SCORE) / 0.20f);
                        local_conf = sc * warp_confidence(base, rx, ry);
                        refined = 1;
                    }
                }
            }
            if (refined) {
                out->points[2 * index] = point[0];
                out->points[2 * index + 1] = point[1];
                out->confidence[index] = local_conf;
                cgrid[index] = 1;
                correction[2 * index] = med[0];
                correction[2 * index + 1] = med[1];
                corrections[ncorr++] = hypotf(med[0], med[1]);
                confidence_sum += local_conf;
            } else {
                float q[2];
                if (!or_warp_map(base, rx, ry, q)) q[0] = q[1] = NAN;
                out->points[2 * index] = q[0];
                out->points[2 * index + 1] = q[1];
                out->confidence[index] = 0.0f;
            }
        }
    /* two completion passes (resolution.rs:362-412) */
    uint8_t *csnap = (uint8_t *)malloc(nodes);
    float *corr_snap = (float *)malloc(nodes * 2 * sizeof(float));
    float *conf_snap = (float *)malloc(nodes * sizeof(float));
    for (int pass = 0; pass < 2; pass++) {
        memcpy(csnap, cgrid, nodes);
        memcpy(corr_snap, correction, nodes * 2 * sizeof(float));
        memcpy(conf_snap, out->confidence, nodes * sizeof(float));
        for (uint32_t row = 0; row < rows; row++)
            for (uint32_t column = 0; column < columns; column++) {
                size_t index = (size_t)row * columns + column;
                if (csnap[index]) continue;
                float xs[9], ys[9], csum = 0.0f;
                int n = 0;
                uint32_t r0 = row > 0 ? row - 1 : 0, r1 = row + 1 < rows - 1 ? row + 1 : rows - 1;
                uint32_t c0 = column > 0 ? column - 1 : 0, c1 = column + 1 < columns - 1 ? column + 1 : columns - 1;
                for (uint32_t nr = r0; nr <= r1; nr++)
                    for (uint32_t nc = c0; nc <= c1; nc++) {
                        size_t j = (size_t)nr * columns + nc;
                        if (csnap[j]) {
                            xs[n] = corr_snap[2 * j];
                            ys[n] = corr_snap[2 * j + 1];
                            csum += conf_snap[j];
                            n++;
                        }
                    }
                if (n < 3) continue;
                float sx[9], sy[9];
                memcpy(sx, xs, n * sizeof(float));
                memcpy(sy, ys, n * sizeof(float));
                qsort(sx, n, sizeof(float), cmp_f32_total);
                qsort(sy, n, sizeof(float), cmp_f32_total);
                float c[2] = {sx[n / 2], sy[n / 2]};
                int coherent = 1;
                for (int k = 0; k < n; k++)
                    if (!(hypotf(sx[k] - c[0], sy[k] - c[1]) <= 1.5f)) coherent = 0;
                if (!coherent) continue;
                float rx = (float)(column * LOCAL_WARP_STEP), ry = (float)(row * LOCAL_WARP_STEP), p[2];
                if (or_warp_map(base, rx + c[0], ry + c[1], p)) {
                    out->points[2 * index] = p[0];
                    out->points[2 * index + 1] = p[1];
                    out->confidence[index] = 0.45f * csum / (float)n;
                    cgrid[index] = 1;
                    correction[2 * index] = c[0];
                    correction[2 * index + 1] = c[1];
                }
            }
    }
    qsort(corrections, ncorr, sizeof(float), cmp_f32_total);
    size_t supported = 0;
    for (size_t i = 0; i < nodes; i++) supported += out->confidence[i] > 0.0f;
    float denom = (float)(nodes > 0 ? nodes : 1);
    report->verified_fraction = (float)ncorr / denom;
    report->supported_fraction = (float)supported / denom;
    report->median_correction_px = ncorr > 0 ? corrections[ncorr / 2] : 0.0f;
    report->mean_confidence = confidence_sum / (float)(ncorr > 0 ? ncorr : 1);
    free(has);
    free(mshift);
    free(mscore);
    free(cgrid);
    free(correction);
    free(corrections);
    free(csnap);
    free(corr_snap);
    free(conf_snap);
    return CHD_OK;
}
