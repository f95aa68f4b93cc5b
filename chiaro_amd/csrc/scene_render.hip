// scene_render.hip - the seeded synthetic capture generator on the GPU.
//
// The same model as the host generator (scene.c): a Voronoi partition of
// the reference raster into fronto-parallel planes, a 4-octave value-noise
// texture on a canvas around the reference, every other module rendered by
// inverse warping (the nearest plane whose cell contains the reference
// projection is the visible one), and 12-bit Bayer codes with black level,
// Poisson shot noise and Gaussian read noise. Every random quantity is a
// pure function of (seed, stream, index) through SplitMix64, and every
// operation is the host generator's in the same order in f64 (built without
// contraction), so the device writes the host generator's codes; the tests
// compare them byte for byte. The codes are stored shifted left by `shift`
// bits, as the pipeline feeds luminance_half (pipeline.rs:1026-1030).
//
// One thread per canvas point (texture value and Voronoi cell), then one
// thread per module pixel; a module's pixels are independent, so the launch
// is a plain 2-D grid over the image, rows of 64-wide wavefronts reading
// neighbouring canvas texels.
#include <hip/hip_runtime.h>

#include "depth_internal.hpp"

namespace chd {

namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t index) {
    return splitmix64(splitmix64(seed ^ splitmix64(stream * 0x632be59bd9b4e019ULL)) ^ index);
}
__device__ __forceinline__ double unit(uint64_t h) { return (double)(h >> 11) * 0x1p-53; }
__device__ __forceinline__ double gaussian(uint64_t seed, uint64_t stream, uint64_t index) {
    double u1 = unit(hash3(seed, stream, 2 * index));
    double u2 = unit(hash3(seed, stream, 2 * index + 1));
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ uint32_t voronoi_cell(const SceneDev &s, double x, double y) {
    uint32_t best = 0;
    double bd = INFINITY;
    for (uint32_t i = 0; i < s.n; i++) {
        double dx = x - s.site_x[i], dy = y - s.site_y[i];
        double d = dx * dx + dy * dy;
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

__device__ double value_noise(uint64_t seed, uint32_t octave, double x, double y, double cell) {
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

__device__ double texture(const SceneDev &s, double x, double y) {
    double sum = 0.0, amplitude = 1.0, norm = 0.0, cell = (double)s.texture_cell_px;
    for (uint32_t o = 0; o < 4; o++) {
        sum += amplitude * value_noise(s.seed, o, x, y, cell);
        norm += amplitude;
        amplitude *= 0.5;
        cell *= 0.5;
    }
    return sum / norm;
}

__global__ void k_scene_canvas(const SceneDev *scene, float *canvas, uint8_t *cells) {
    const SceneDev &s = *scene;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= s.canvas_w) return;
    const size_t at = (size_t)y * s.canvas_w + x;
    const double cx = s.canvas_x0 + x, cy = s.canvas_y0 + y;
    canvas[at] = (float)texture(s, cx, cy);
    cells[at] = (uint8_t)voronoi_cell(s, cx, cy);
}

__device__ uint32_t cell_at(const SceneDev &s, const uint8_t *cells, double x, double y) {
    const double cx = x - s.canvas_x0, cy = y - s.canvas_y0;
    if (cx >= 0.0 && cy >= 0.0 && cx < (double)(s.canvas_w - 1) && cy < (double)(s.canvas_h - 1)) {
        const size_t i = (size_t)cx, j = (size_t)cy, w = s.canvas_w;
        const uint8_t c = cells[j * w + i];
        if (cells[j * w + i + 1] == c && cells[(j + 1) * w + i] == c &&
            cells[(j + 1) * w + i + 1] == c)
            return c;
    }
    return voronoi_cell(s, x, y);
}

__device__ float canvas_sample(const SceneDev &s, const float *canvas, double x, double y) {
    double cx = x - s.canvas_x0, cy = y - s.canvas_y0;
    if (cx < 0) cx = 0;
    if (cy < 0) cy = 0;
    if (cx > s.canvas_w - 1) cx = s.canvas_w - 1;
    if (cy > s.canvas_h - 1) cy = s.canvas_h - 1;
    uint32_t x0 = (uint32_t)cx, y0 = (uint32_t)cy;
    uint32_t x1 = x0 + 1 < s.canvas_w ? x0 + 1 : x0, y1 = y0 + 1 < s.canvas_h ? y0 + 1 : y0;
    double tx = cx - x0, ty = cy - y0;
    const float *c = canvas;
    double top = c[(size_t)y0 * s.canvas_w + x0] * (1 - tx) + c[(size_t)y0 * s.canvas_w + x1] * tx;
    double bot = c[(size_t)y1 * s.canvas_w + x0] * (1 - tx) + c[(size_t)y1 * s.canvas_w + x1] * tx;
    return (float)(top * (1 - ty) + bot * ty);
}

__device__ double radiance(const SceneDev &s, const float *canvas, const uint8_t *cells,
                           const RenderModule &m, const RenderModule &r, bool is_reference,
                           uint32_t px, uint32_t py) {
    if (is_reference) return canvas_sample(s, canvas, px, py);
    double cx = ((double)m.width - 1.0) * 0.5, cy = ((double)m.height - 1.0) * 0.5;
    double v[3] = {((double)px - cx) / m.focal_px, ((double)py - cy) / m.focal_px, 1.0};
    const double *R = m.rotation_wc;
    double d[3] = {R[0] * v[0] + R[3] * v[1] + R[6] * v[2], R[1] * v[0] + R[4] * v[1] + R[7] * v[2],
                   R[2] * v[0] + R[5] * v[1] + R[8] * v[2]};
    double rcx = ((double)r.width - 1.0) * 0.5, rcy = ((double)r.height - 1.0) * 0.5;
    double last_x = 0, last_y = 0;
    for (uint32_t k = 0; k < s.n; k++) {
        uint32_t i = s.order[k];
        double z = 1.0 / s.inverse[i];
        if (d[2] <= 0) break;
        double t = (z - m.center[2]) / d[2];
        double X = m.center[0] + t * d[0], Y = m.center[1] + t * d[1];
        double rx = r.focal_px * X / z + rcx, ry = r.focal_px * Y / z + rcy;
        last_x = rx;
        last_y = ry;
        if (cell_at(s, cells, rx, ry) == i) return canvas_sample(s, canvas, rx, ry);
    }
    return canvas_sample(s, canvas, last_x, last_y);
}

__global__ void k_render_bayer(const SceneDev *scene, const float *canvas, const uint8_t *cells,
                               RenderModule m, RenderModule r, uint32_t view, int is_reference,
                               BayerSpec b, uint32_t shift, uint16_t *out) {
    const SceneDev &s = *scene;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= m.width) return;
    const size_t idx = (size_t)y * m.width + x;
    const double value = radiance(s, canvas, cells, m, r, is_reference != 0, x, y);
    const double signal_dn = value * b.exposure * (b.white_level - b.black_level);
    const double electrons = signal_dn * b.electrons_per_dn;
    double shot;
    if (electrons < 20.0) {
        // Poisson by inversion
        const double u = unit(hash3(s.seed, 5000 + view, idx));
        double p = exp(-electrons), cdf = p;
        uint32_t k = 0;
        while (u > cdf && k < 200) {
            k++;
            p *= electrons / k;
            cdf += p;
        }
        shot = k;
    } else {
        // normal approximation of Poisson(lambda >= 20)
        shot = floor(electrons + sqrt(electrons) * gaussian(s.seed, 5000 + view, idx) + 0.5);
        if (shot < 0) shot = 0;
    }
    double dn = b.black_level + shot / b.electrons_per_dn +
                b.read_noise_dn * gaussian(s.seed, 6000 + view, idx);
    dn = floor(dn + 0.5);
    if (dn < 0) dn = 0;
    if (dn > b.white_level) dn = b.white_level;
    out[idx] = (uint16_t)((uint32_t)(uint16_t)dn << shift);
}

}  // namespace

hipError_t launch_scene_canvas(hipStream_t stream, const SceneDev *scene, uint32_t canvas_w,
                               uint32_t canvas_h, float *canvas, uint8_t *cells) {
    const dim3 block(256), grid((canvas_w + 255) / 256, canvas_h);
    hipLaunchKernelGGL(k_scene_canvas, grid, block, 0, stream, scene, canvas, cells);
    return hipGetLastError();
}

hipError_t launch_render_bayer(hipStream_t stream, const SceneDev *scene, const float *canvas,
                               const uint8_t *cells, const RenderModule &module,
                               const RenderModule &reference, uint32_t view, bool is_reference,
                               const BayerSpec &bayer, uint32_t shift, uint16_t *out) {
    const dim3 block(256), grid((module.width + 255) / 256, module.height);
    hipLaunchKernelGGL(k_render_bayer, grid, block, 0, stream, scene, canvas, cells, module,
                       reference, view, is_reference ? 1 : 0, bayer, shift, out);
    return hipGetLastError();
}

}  // namespace chd
