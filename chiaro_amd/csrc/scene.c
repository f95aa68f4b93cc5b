/*
 * scene.c - seeded synthetic multi-view captures for the depth path.
 *
 * The reference has no scene generator (its only fakes are gradient mock
 * LRIs, crates/chiaro/src/mock.rs:62), so the workloads are defined here:
 *
 *  - scene: a Voronoi partition of the reference image into fronto-parallel
 *    planes (in the reference camera frame) with inverse depths drawn
 *    uniformly from [inverse_depth_min, inverse_depth_max]; the farthest
 *    plane doubles as an unbounded backdrop;
 *  - texture: a sum of 4 octaves of value noise attached to reference-raster
 *    coordinates, rendered on a canvas with a margin around the reference;
 *  - views: the reference is the clean canvas; every other module is an
 *    inverse warp: each target pixel's ray meets the planes nearest first,
 *    the first whose Voronoi cell contains the reference projection is the
 *    visible one, and the canvas is sampled bilinearly there;
 *  - sensor: f32 in [0, 1] plus Gaussian noise, or 12-bit RGGB Bayer codes
 *    with black level, Poisson shot noise and Gaussian read noise;
 *  - luminance_half: the reference's input preparation
 *    (crates/chiaro-fusion/src/image.rs:401-426) with a unit flat field,
 *    evaluated with the platform logf as the reference does;
 *  - global warps: align_module(refine = false), i.e. the 8 px infinity-map
 *    grid resampled to the 32 px warp grid (align.rs:317-324, 456-459).
 *
 * Every random quantity is a pure function of (seed, stream, index) through
 * SplitMix64, so output bytes do not depend on the thread count. Built with
 * -ffp-contract=off so the camera arithmetic matches the reference's.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/chiaro_depth.h"

typedef struct chs_module {
    double focal_px;
    uint32_t width;
    uint32_t height;
    double center[3];      /* optical centre in the world (reference) frame, mm */
    double rotation_wc[9]; /* world-to-camera rotation, row-major */
} chs_module;

typedef struct chs_scene {
    uint64_t seed;
    uint32_t plane_count;
    double inverse_depth_min; /* 1/mm */
    double inverse_depth_max;
    uint32_t texture_cell_px; /* lattice spacing of the coarsest octave */
} chs_scene;

typedef struct chs_bayer {
    float black_level;   /* DN */
    float white_level;   /* DN */
    float exposure;      /* fraction of (white - black) reached by radiance 1 */
    float electrons_per_dn;
    float read_noise_dn;
} chs_bayer;

/* ------------------------------------------------------------ randomness */

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
static inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t index) {
    return splitmix64(splitmix64(seed ^ splitmix64(stream * 0x632be59bd9b4e019ULL)) ^ index);
}
static inline double unit(uint64_t h) { return (double)(h >> 11) * 0x1p-53; }
static inline double gaussian(uint64_t seed, uint64_t stream, uint64_t index) {
    double u1 = unit(hash3(seed, stream, 2 * index));
    double u2 = unit(hash3(seed, stream, 2 * index + 1));
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ----------------------------------------------------------- the scene */

typedef struct {
    uint32_t n;
    double site_x[64], site_y[64];
    double inverse[64]; /* 1/mm */
    uint32_t order[64]; /* planes sorted nearest first */
    double canvas_x0, canvas_y0;
    uint32_t canvas_w, canvas_h;
    float *canvas;
    uint8_t *cells; /* Voronoi cell of each integer canvas point (or NULL) */
} scene_state;

static uint32_t voronoi_cell(const scene_state *s, double x, double y) {
    uint32_t best = 0;
    double bd = INFINITY;
    for (uint32_t i = 0; i < s->n; i++) {
        double dx = x - s->site_x[i], dy = y - s->site_y[i];
        double d = dx * dx + dy * dy;
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

/* voronoi_cell through the canvas's cell map: a point whose four enclosing
 * grid points share one cell lies in that cell (cells are convex), else the
 * sites are searched. Same result as voronoi_cell, much cheaper. */
static uint32_t cell_at(const scene_state *s, double x, double y) {
    const double cx = x - s->canvas_x0, cy = y - s->canvas_y0;
    if (s->cells && cx >= 0.0 && cy >= 0.0 && cx < (double)(s->canvas_w - 1) &&
        cy < (double)(s->canvas_h - 1)) {
        const size_t i = (size_t)cx, j = (size_t)cy, w = s->canvas_w;
        const uint8_t c = s->cells[j * w + i];
        if (s->cells[j * w + i + 1] == c && s->cells[(j + 1) * w + i] == c &&
            s->cells[(j + 1) * w + i + 1] == c)
            return c;
    }
    return voronoi_cell(s, x, y);
}

static double value_noise(uint64_t seed, uint32_t octave, double x, double y, double cell) {
    double fx = x / cell, fy = y / cell;
    double ix = floor(fx), iy = floor(fy);
    double tx = fx - ix, ty = fy - iy;
    tx = tx * tx * (3.0 - 2.0 * tx);
    ty = ty * ty * (3.0 - 2.0 * ty);
    int64_t cx = (int64_t)ix, cy = (int64_t)iy;
    double v[4];
    for (int k = 0; k < 4; k++) {
        int64_t gx = cx + (k & 1), gy = cy + (k >> 1);
        uint64_t key = ((uint64_t)(gx + (1 << 20)) << 24) ^ (uint64_t)(gy + (1 << 20));
        v[k] = unit(hash3(seed, 1000 + octave, key));
    }
    double top = v[0] + (v[1] - v[0]) * tx, bottom = v[2] + (v[3] - v[2]) * tx;
    return top + (bottom - top) * ty;
}

static double texture(const chs_scene *sc, double x, double y) {
    double sum = 0.0, amplitude = 1.0, norm = 0.0, cell = (double)sc->texture_cell_px;
    for (uint32_t o = 0; o < 4; o++) {
        sum += amplitude * value_noise(sc->seed, o, x, y, cell);
        norm += amplitude;
        amplitude *= 0.5;
        cell *= 0.5;
    }
    return sum / norm;
}

static int build_scene(const chs_scene *sc, const chs_module *ref, uint32_t margin,
                       scene_state *s) {
    if (sc->plane_count == 0 || sc->plane_count > 64) return 0;
    s->n = sc->plane_count;
    for (uint32_t i = 0; i < s->n; i++) {
        s->site_x[i] = unit(hash3(sc->seed, 1, i)) * (double)ref->width;
        s->site_y[i] = unit(hash3(sc->seed, 2, i)) * (double)ref->height;
        s->inverse[i] = sc->inverse_depth_min +
                        (sc->inverse_depth_max - sc->inverse_depth_min) * unit(hash3(sc->seed, 3, i));
        s->order[i] = i;
    }
    /* nearest (largest inverse depth) first; ties by index */
    for (uint32_t i = 1; i < s->n; i++) {
        uint32_t v = s->order[i], j = i;
        while (j > 0 && s->inverse[s->order[j - 1]] < s->inverse[v]) {
            s->order[j] = s->order[j - 1];
            j--;
        }
        s->order[j] = v;
    }
    s->canvas_x0 = -(double)margin;
    s->canvas_y0 = -(double)margin;
    s->canvas_w = ref->width + 2 * margin;
    s->canvas_h = ref->height + 2 * margin;
    s->canvas = (float *)malloc((size_t)s->canvas_w * s->canvas_h * sizeof(float));
    s->cells = NULL;
    return s->canvas != NULL;
}

/* bilinear canvas sample at reference-raster (x, y) */
static float canvas_sample(const scene_state *s, double x, double y) {
    double cx = x - s->canvas_x0, cy = y - s->canvas_y0;
    if (cx < 0) cx = 0;
    if (cy < 0) cy = 0;
    if (cx > s->canvas_w - 1) cx = s->canvas_w - 1;
    if (cy > s->canvas_h - 1) cy = s->canvas_h - 1;
    uint32_t x0 = (uint32_t)cx, y0 = (uint32_t)cy;
    uint32_t x1 = x0 + 1 < s->canvas_w ? x0 + 1 : x0, y1 = y0 + 1 < s->canvas_h ? y0 + 1 : y0;
    double tx = cx - x0, ty = cy - y0;
    const float *c = s->canvas;
    double top = c[(size_t)y0 * s->canvas_w + x0] * (1 - tx) + c[(size_t)y0 * s->canvas_w + x1] * tx;
    double bot = c[(size_t)y1 * s->canvas_w + x0] * (1 - tx) + c[(size_t)y1 * s->canvas_w + x1] * tx;
    return (float)(top * (1 - ty) + bot * ty);
}

/* ---------------------------------------------------------- rendering */

typedef struct {
    const chs_scene *sc;
    const scene_state *s;
    const chs_module *modules;
    uint32_t ref, view, row_begin, row_end;
    int canvas_pass;
    float noise_sigma;
    const chs_bayer *bayer;
    float *image;
    uint16_t *raw;
} render_job;

static float radiance(const render_job *j, uint32_t view, uint32_t px, uint32_t py) {
    const chs_module *m = &j->modules[view];
    const chs_module *r = &j->modules[j->ref];
    const scene_state *s = j->s;
    if (view == j->ref) return canvas_sample(s, px, py);
    /* camera ray in the world frame: d = R^T K^-1 q (pinhole, principal
     * point at the image centre) */
    double cx = ((double)m->width - 1.0) * 0.5, cy = ((double)m->height - 1.0) * 0.5;
    double v[3] = {((double)px - cx) / m->focal_px, ((double)py - cy) / m->focal_px, 1.0};
    const double *R = m->rotation_wc;
    double d[3] = {R[0] * v[0] + R[3] * v[1] + R[6] * v[2], R[1] * v[0] + R[4] * v[1] + R[7] * v[2],
                   R[2] * v[0] + R[5] * v[1] + R[8] * v[2]};
    double rcx = ((double)r->width - 1.0) * 0.5, rcy = ((double)r->height - 1.0) * 0.5;
    double last_x = 0, last_y = 0;
    for (uint32_t k = 0; k < s->n; k++) {
        uint32_t i = s->order[k];
        double z = 1.0 / s->inverse[i];
        if (d[2] <= 0) break;
        double t = (z - m->center[2]) / d[2];
        double X = m->center[0] + t * d[0], Y = m->center[1] + t * d[1];
        /* reference camera at the origin with R = I */
        double rx = r->focal_px * X / z + rcx, ry = r->focal_px * Y / z + rcy;
        last_x = rx;
        last_y = ry;
        if (cell_at(s, rx, ry) == i) return canvas_sample(s, rx, ry);
    }
    return canvas_sample(s, last_x, last_y); /* farthest plane as backdrop */
}

static void *render_worker(void *arg) {
    render_job *j = (render_job *)arg;
    if (j->canvas_pass) {
        const scene_state *s = j->s;
        for (uint32_t y = j->row_begin; y < j->row_end; y++)
            for (uint32_t x = 0; x < s->canvas_w; x++) {
                const size_t at = (size_t)y * s->canvas_w + x;
                s->canvas[at] = (float)texture(j->sc, s->canvas_x0 + x, s->canvas_y0 + y);
                if (s->cells)
                    s->cells[at] = (uint8_t)voronoi_cell(s, s->canvas_x0 + x, s->canvas_y0 + y);
            }
        return NULL;
    }
    const chs_module *m = &j->modules[j->view];
    for (uint32_t y = j->row_begin; y < j->row_end; y++)
        for (uint32_t x = 0; x < m->width; x++) {
            size_t idx = (size_t)y * m->width + x;
            double value = radiance(j, j->view, x, y);
            if (j->bayer) {
                const chs_bayer *b = j->bayer;
                double signal_dn = value * b->exposure * (b->white_level - b->black_level);
                double electrons = signal_dn * b->electrons_per_dn;
                double shot;
                if (electrons < 20.0) {
                    /* Poisson by inversion */
                    double u = unit(hash3(j->sc->seed, 5000 + j->view, idx));
                    double p = exp(-electrons), cdf = p;
                    uint32_t k = 0;
                    while (u > cdf && k < 200) {
                        k++;
                        p *= electrons / k;
                        cdf += p;
                    }
                    shot = k;
                } else {
                    /* normal approximation of Poisson(lambda >= 20) */
                    shot = floor(electrons + sqrt(electrons) *
                                                 gaussian(j->sc->seed, 5000 + j->view, idx) + 0.5);
                    if (shot < 0) shot = 0;
                }
                double dn = b->black_level + shot / b->electrons_per_dn +
                            b->read_noise_dn * gaussian(j->sc->seed, 6000 + j->view, idx);
                dn = floor(dn + 0.5);
                if (dn < 0) dn = 0;
                if (dn > b->white_level) dn = b->white_level;
                j->raw[idx] = (uint16_t)dn;
            } else {
                double v = value + (double)j->noise_sigma * gaussian(j->sc->seed, 4000 + j->view, idx);
                if (v < 0) v = 0;
                if (v > 1) v = 1;
                j->image[idx] = (float)v;
            }
        }
    return NULL;
}

static void run_rows(render_job *proto, uint32_t rows, uint32_t threads) {
    if (threads == 0) threads = 1;
    if (threads > 64) threads = 64;
    if (threads > rows) threads = rows;
    render_job jobs[64];
    pthread_t tid[64];
    uint32_t per = (rows + threads - 1) / threads, n = 0;
    for (uint32_t first = 0; first < rows; first += per) {
        jobs[n] = *proto;
        jobs[n].row_begin = first;
        jobs[n].row_end = first + per < rows ? first + per : rows;
        pthread_create(&tid[n], NULL, render_worker, &jobs[n]);
        n++;
    }
    for (uint32_t i = 0; i < n; i++) pthread_join(tid[i], NULL);
}

static uint32_t rig_margin(const chs_scene *sc, const chs_module *modules, uint32_t n,
                           uint32_t ref) {
    double worst = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        double b = sqrt(modules[i].center[0] * modules[i].center[0] +
                        modules[i].center[1] * modules[i].center[1]);
        double d = modules[ref].focal_px * b * sc->inverse_depth_max;
        if (d > worst) worst = d;
    }
    return (uint32_t)ceil(worst) + 64;
}

int chs_render(const chs_scene *sc, const chs_module *modules, uint32_t count, uint32_t ref,
               float noise_sigma, const chs_bayer *bayer, float **images, uint16_t **raws,
               uint32_t threads) {
    if (!sc || !modules || ref >= count) return -1;
    scene_state s;
    if (!build_scene(sc, &modules[ref], rig_margin(sc, modules, count, ref), &s)) return -1;
    s.cells = (uint8_t *)malloc((size_t)s.canvas_w * s.canvas_h);
    render_job proto;
    memset(&proto, 0, sizeof proto);
    proto.sc = sc;
    proto.s = &s;
    proto.modules = modules;
    proto.ref = ref;
    proto.canvas_pass = 1;
    run_rows(&proto, s.canvas_h, threads);
    proto.canvas_pass = 0;
    proto.noise_sigma = noise_sigma;
    proto.bayer = bayer;
    for (uint32_t v = 0; v < count; v++) {
        proto.view = v;
        proto.image = images ? images[v] : NULL;
        proto.raw = raws ? raws[v] : NULL;
        if ((bayer && !proto.raw) || (!bayer && !proto.image)) {
            free(s.canvas);
            free(s.cells);
            return -1;
        }
        run_rows(&proto, modules[v].height, threads);
    }
    free(s.canvas);
    free(s.cells);
    return 0;
}

/* Ground truth for the reference view: inverse of the distance along each
 * raster pixel's ray (the quantity the plane sweep hypothesises), 1/mm. */
int chs_truth_inverse_distance(const chs_scene *sc, const chs_module *ref_module,
                               uint32_t step, float *out) {
    scene_state s;
    if (!build_scene(sc, ref_module, 0, &s)) return -1;
    free(s.canvas);
    uint32_t columns = (ref_module->width + step - 1) / step + 1;
    uint32_t rows = (ref_module->height + step - 1) / step + 1;
    double cx = ((double)ref_module->width - 1.0) * 0.5, cy = ((double)ref_module->height - 1.0) * 0.5;
    for (uint32_t r = 0; r < rows; r++)
        for (uint32_t c = 0; c < columns; c++) {
            double x = (double)c * step, y = (double)r * step;
            uint32_t cell = voronoi_cell(&s, x, y);
            double vx = (x - cx) / ref_module->focal_px, vy = (y - cy) / ref_module->focal_px;
            double norm = sqrt(vx * vx + vy * vy + 1.0);
            out[(size_t)r * columns + c] = (float)(s.inverse[cell] / norm);
        }
    return 0;
}

/* ------------------------------------------------------ luminance_half */

/* Mosaic::luminance_half (image.rs:401-426) on f32 raster samples with a
 * unit flat field. */
int chs_luminance_half_f32(const float *raster, uint32_t width, uint32_t height,
                           float black, float white, float *plane) {
    uint32_t w = width / 2, h = height / 2;
    float range = white - black > 1.0f ? white - black : 1.0f;
    for (uint32_t j = 0; j < h; j++) {
        const float *row0 = raster + (size_t)(2 * j) * width;
        const float *row1 = raster + (size_t)(2 * j + 1) * width;
        for (uint32_t i = 0; i < w; i++) {
            float sum = row0[2 * i] + row0[2 * i + 1] + row1[2 * i] + row1[2 * i + 1];
            float linear = (sum * 0.25f - black) / range;
            if (!(linear > 0.0f)) linear = linear != linear ? linear : 0.0f;
            linear = linear * 1.0f;
            plane[(size_t)j * w + i] = logf(1.0f + 1000.0f * linear);
        }
    }
    return 0;
}

/* The same on u16 codes shifted left by `shift` bits (the pipeline's Q6
 * step, pipeline.rs:1026-1030); black/white are given in shifted units. */
int chs_luminance_half_u16(const uint16_t *raster, uint32_t width, uint32_t height,
                           uint32_t shift, float black, float white, float *plane) {
    uint32_t w = width / 2, h = height / 2;
    float range = white - black > 1.0f ? white - black : 1.0f;
    for (uint32_t j = 0; j < h; j++) {
        const uint16_t *row0 = raster + (size_t)(2 * j) * width;
        const uint16_t *row1 = raster + (size_t)(2 * j + 1) * width;
        for (uint32_t i = 0; i < w; i++) {
            float sum = (float)(uint16_t)(row0[2 * i] << shift) +
                        (float)(uint16_t)(row0[2 * i + 1] << shift) +
                        (float)(uint16_t)(row1[2 * i] << shift) +
                        (float)(uint16_t)(row1[2 * i + 1] << shift);
            float linear = (sum * 0.25f - black) / range;
            if (!(linear > 0.0f)) linear = linear != linear ? linear : 0.0f;
            linear = linear * 1.0f;
            plane[(size_t)j * w + i] = logf(1.0f + 1000.0f * linear);
        }
    }
    return 0;
}

/* ------------------------------------------------------------ cameras */

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void mul_vec3(const double *m, const double *v, double *o) {
    double r0 = dot3(m, v), r1 = dot3(m + 3, v), r2 = dot3(m + 6, v);
    o[0] = r0; o[1] = r1; o[2] = r2;
}
static void normalize3(const double *a, double *o) {
    double n = sqrt(dot3(a, a));
    if (n == 0.0) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; return; }
    double s = 1.0 / n;
    o[0] = a[0] * s; o[1] = a[1] * s; o[2] = a[2] * s;
}

/* ResolvedCamera::new for a canonical pose (geometry.rs:64-133) with
 * K = [[f, 0, (w-1)/2], [0, f, (h-1)/2], [0, 0, 1]] and t = -R c. */
int chs_camera(const chs_module *m, chd_camera *out) {
    memset(out, 0, sizeof *out);
    out->width = m->width;
    out->height = m->height;
    double k[9] = {m->focal_px, 0, ((double)m->width - 1.0) * 0.5,
                   0, m->focal_px, ((double)m->height - 1.0) * 0.5,
                   0, 0, 1};
    memcpy(out->k, k, sizeof k);
    double det = k[0] * (k[4] * k[8] - k[5] * k[7]) - k[1] * (k[3] * k[8] - k[5] * k[6]) +
                 k[2] * (k[3] * k[7] - k[4] * k[6]);
    if (fabs(det) < 1e-300) return -1;
    double inv = 1.0 / det;
#define C(r0, c0, r1, c1) (k[(r0)*3 + (c0)] * k[(r1)*3 + (c1)] - k[(r0)*3 + (c1)] * k[(r1)*3 + (c0)])
    double ki[9] = {C(1, 1, 2, 2) * inv, -C(0, 1, 2, 2) * inv, C(0, 1, 1, 2) * inv,
                    -C(1, 0, 2, 2) * inv, C(0, 0, 2, 2) * inv, -C(0, 0, 1, 2) * inv,
                    C(1, 0, 2, 1) * inv, -C(0, 0, 2, 1) * inv, C(0, 0, 1, 1) * inv};
#undef C
    memcpy(out->k_inverse, ki, sizeof ki);
    out->flip_around_x = -1;
    out->pose_kind = 0;
    memcpy(out->rotation_wc, m->rotation_wc, sizeof out->rotation_wc);
    const double *R = m->rotation_wc;
    double rcw[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
    memcpy(out->rotation_cw, rcw, sizeof rcw);
    /* translation t = -R c, then center_world = -(R^T t) */
    double rc[3];
    mul_vec3(R, m->center, rc);
    double t[3] = {-rc[0], -rc[1], -rc[2]};
    memcpy(out->translation_wc, t, sizeof t);
    double c[3];
    mul_vec3(rcw, t, c);
    out->center[0] = c[0] * -1.0;
    out->center[1] = c[1] * -1.0;
    out->center[2] = c[2] * -1.0;
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(out->orientation_correction, I, sizeof I);
    out->focal_px = m->focal_px;
    return 0;
}

/* pinhole map_from (geometry.rs:152-252 without distortion, mirror or
 * orientation correction, which synthetic rigs do not use) */
static int map_from(const chd_camera *self, const chd_camera *reference, const double pixel[2],
                    double depth, double out[2]) {
    double h[3] = {pixel[0], pixel[1], 1.0}, t[3], d[3], world[3], n1[3], dir[3];
    mul_vec3(reference->k_inverse, h, t);
    normalize3(t, d);
    mul_vec3(reference->rotation_cw, d, world);
    normalize3(world, n1);
    double n2[3];
    mul_vec3(reference->orientation_correction, n1, n2);
    normalize3(n2, dir);
    double p[3] = {reference->center[0] + dir[0] * depth, reference->center[1] + dir[1] * depth,
                   reference->center[2] + dir[2] * depth};
    double r[3];
    mul_vec3(self->rotation_wc, p, r);
    double cam[3] = {r[0] + self->translation_wc[0], r[1] + self->translation_wc[1],
                     r[2] + self->translation_wc[2]};
    if (cam[2] <= 0.0) return 0;
    double n[2] = {cam[0] / cam[2], cam[1] / cam[2]};
    const double *k = self->k;
    double ideal[2] = {k[0] * n[0] + k[1] * n[1] + k[2], k[4] * n[1] + k[5]};
    double mx = (double)self->width * 0.5, my = (double)self->height * 0.5;
    if (ideal[0] < -mx || ideal[1] < -my || ideal[0] > (double)self->width + mx ||
        ideal[1] > (double)self->height + my)
        return 0;
    out[0] = ideal[0];
    out[1] = ideal[1];
    return 1;
}

static inline size_t f32_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return SIZE_MAX;
    return (size_t)v;
}

static int warp_map(const float *points, uint32_t step, uint32_t columns, uint32_t rows, float x,
                    float y, float out[2]) {
    float fx = x / (float)step, fy = y / (float)step;
    if (fx < 0.0f || fy < 0.0f) return 0;
    size_t c0 = f32_usize(floorf(fx)), r0 = f32_usize(floorf(fy));
    if (c0 > columns - 1) c0 = columns - 1;
    if (r0 > rows - 1) r0 = rows - 1;
    size_t c1 = c0 + 1 < columns - 1 ? c0 + 1 : columns - 1;
    size_t r1 = r0 + 1 < rows - 1 ? r0 + 1 : rows - 1;
    float tx = fx - (float)c0, ty = fy - (float)r0;
    const float *a = points + 2 * (r0 * columns + c0), *b = points + 2 * (r0 * columns + c1);
    const float *c = points + 2 * (r1 * columns + c0), *d = points + 2 * (r1 * columns + c1);
    for (int k = 0; k < 2; k++) {
        float top = a[k] * (1.0f - tx) + b[k] * tx;
        float bottom = c[k] * (1.0f - tx) + d[k] * tx;
        out[k] = top * (1.0f - ty) + bottom * ty;
    }
    return !(out[0] != out[0] || out[1] != out[1]);
}

/* align_module(refine = false) on a calibrated pair: Warp::from_fn at 8 px
 * over map_from(.., 1e8) filtered to finite values, resampled to grid_step
 * through the identity correction (align.rs:317-324, 456-459). */
int chs_global_warp(const chd_camera *reference, const chd_camera *target, uint32_t width,
                    uint32_t height, uint32_t grid_step, float *points, float *confidence) {
    uint32_t ic = (width + 7) / 8 + 1, ir = (height + 7) / 8 + 1;
    float *initial = (float *)malloc((size_t)ic * ir * 2 * sizeof(float));
    if (!initial) return -1;
    for (uint32_t r = 0; r < ir; r++)
        for (uint32_t c = 0; c < ic; c++) {
            double p[2] = {(double)((size_t)c * 8), (double)((size_t)r * 8)}, q[2];
            size_t i = (size_t)r * ic + c;
            if (map_from(target, reference, p, 1.0e8, q) && isfinite(q[0]) && isfinite(q[1])) {
                initial[2 * i] = (float)q[0];
                initial[2 * i + 1] = (float)q[1];
            } else {
                initial[2 * i] = NAN;
                initial[2 * i + 1] = NAN;
            }
        }
    uint32_t oc = (width + grid_step - 1) / grid_step + 1, orows = (height + grid_step - 1) / grid_step + 1;
    for (uint32_t r = 0; r < orows; r++)
        for (uint32_t c = 0; c < oc; c++) {
            size_t i = (size_t)r * oc + c;
            double px = (double)((size_t)c * grid_step), py = (double)((size_t)r * grid_step);
            /* identity homography: (px*1 + py*0 + 0) / 1 */
            double hx = px * 1.0 + py * 0.0 + 1.0 * 0.0, hy = px * 0.0 + py * 1.0 + 1.0 * 0.0;
            double hw = px * 0.0 + py * 0.0 + 1.0 * 1.0;
            float q[2];
            if (warp_map(initial, 8, ic, ir, (float)(hx / hw), (float)(hy / hw), q)) {
                points[2 * i] = q[0];
                points[2 * i + 1] = q[1];
                if (confidence) confidence[i] = 1.0f;
            } else {
                points[2 * i] = NAN;
                points[2 * i + 1] = NAN;
                if (confidence) confidence[i] = 0.0f;
            }
        }
    free(initial);
    return 0;
}
