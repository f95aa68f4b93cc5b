// depth_device.hpp - device-side arithmetic of the depth path (gfx950).
//
// Everything here reproduces the reference's f32/f64 arithmetic bit for bit:
// the translation units that include it are compiled with -ffp-contract=off
// (the reference, Rust, never contracts a*b+c), IEEE division and square root,
// no fast-math, and explicit fma() only where the platform libm's algorithm
// fuses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/chiaro_depth.h"

namespace chd {

// depth.rs:23-37
constexpr double kInfinityDepth = 1.0e8;
constexpr float kMissingCost = 2.25f;
constexpr float kSgmSmallPenalty = 0.06f;
constexpr float kSgmLargePenalty = 0.45f;
constexpr float kSgmDirections = 8.0f;
constexpr float kMinimumRegularizedScore = 0.30f;
constexpr float kMaximumRegularizedBaselineLoss = 0.04f;
constexpr float kNearDepthPrior = 0.10f;
constexpr float kWarpBoundaryDeltaPx = 2.0f;
constexpr float kWarpBoundaryContrast = 0.30f;
constexpr uint32_t kMinimumDirectComponentNodes = 24;

// ------------------------------------------------------------ Rust helpers

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ uint64_t d2u(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ double u2d(uint64_t u) { return __longlong_as_double((long long)u); }

// f32::total_cmp key: order-preserving signed integer.
__device__ __forceinline__ int32_t tkey32(float f) {
    int32_t k = (int32_t)f2u(f);
    return k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1);
}
__device__ __forceinline__ int64_t tkey64(double d) {
    int64_t k = (int64_t)d2u(d);
    return k ^ (int64_t)(((uint64_t)(k >> 63)) >> 1);
}
// f32::max / f32::min (maxNum / minNum on non-NaN inputs as written).
__device__ __forceinline__ float rmaxf(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
__device__ __forceinline__ float rminf(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}
__device__ __forceinline__ double rmax(double a, double b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
__device__ __forceinline__ float rclampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
__device__ __forceinline__ double rclamp(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// `as usize` on floats saturates (NaN, negative -> 0); indices here are
// bounded by the grid, so 32 bits suffice after the clamp to the grid.
__device__ __forceinline__ uint32_t sat_u32(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}
__device__ __forceinline__ uint32_t sat_u32d(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}
__device__ __forceinline__ bool finite_d(double v) { return v - v == 0.0; }
__device__ __forceinline__ bool finite_f(float v) { return v - v == 0.0f; }

// ---------------------------------------------------- transcendental funcs

// expf as computed by the platform libm the reference links (glibc >= 2.28,
// ARM optimized-routines algorithm): 2^(k/32) from a 32-entry table of
// correctly rounded doubles (less i << 47) times a cubic in the remainder,
// fused as in its FMA build.
__constant__ const uint64_t kExp2fTable[32] = {
    0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
    0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
    0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
    0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
    0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
    0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
    0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
    0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL};

__device__ __forceinline__ float libm_expf(float x) {
    const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
    const double shift = 0x1.8p+52;
    const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0;
    const double c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0;
    const double c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
    uint32_t abstop = (f2u(x) >> 20) & 0x7ffu;
    if (abstop >= 0x42bu) {  // |x| >= 88 or NaN
        if (f2u(x) == 0xff800000u) return 0.0f;
        if (abstop >= 0x7f8u) return x + x;
        if (x > 0x1.62e42ep6f) return __int_as_float(0x7f800000);
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    double xd = (double)x;
    double kd = __builtin_fma(inv_ln2_n, xd, shift);
    uint64_t ki = d2u(kd);
    kd -= shift;
    double r = __builtin_fma(inv_ln2_n, xd, -kd);
    uint64_t t = kExp2fTable[ki % 32];
    t += ki << (52 - 5);
    double s = u2d(t);
    double z = __builtin_fma(c0, r, c1);
    double r2 = r * r;
    double y = __builtin_fma(c2, r, 1.0);
    y = __builtin_fma(z, r2, y);
    y = y * s;
    return (float)y;
}

// f64 exp as the platform libm computes it for |x| < 512: 128-entry table
// 2^(k/128) = H[k] (1 + T[k]) and a degree-5 polynomial (ARM
// optimized-routines exp), fused as in its FMA build. The path only calls it
// with x in [-1.5, 0].
__constant__ const uint64_t kExpTable[256] = {
#include "exp_table.inc"
};

__device__ __forceinline__ double libm_exp(double x) {
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128.0;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double shift = 0x1.8p52;
    const double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(d2u(x) >> 52) & 0x7ffu;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;
        // |x| >= 512 never occurs on the path (arguments are -2.5 * |dg|
        // and -d^2/6); saturate like the libm's overflow/underflow ends.
        if (x != x) return x;
        return x > 0 ? __longlong_as_double(0x7ff0000000000000LL) : 0.0;
    }
    double kd = __builtin_fma(inv_ln2_n, x, shift);
    uint64_t ki = d2u(kd);
    kd -= shift;
    double r = __builtin_fma(kd, neg_ln2_lo_n, __builtin_fma(kd, neg_ln2_hi_n, x));
    uint64_t idx = 2 * (ki % 128);
    uint64_t top = ki << (52 - 7);
    double tail = u2d(kExpTable[idx]);
    uint64_t sbits = kExpTable[idx + 1] + top;
    double r2 = r * r;
    double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, c5, c4),
                               __builtin_fma(r2, __builtin_fma(r, c3, c2), tail + r));
    double scale = u2d(sbits);
    return __builtin_fma(scale, tmp, scale);
}

// Option<NodeDepth>::None
__device__ __forceinline__ void clear_node(chd_node_depth &n) {
    n.depth = 0.0;
    n.confidence = 0.0f;
    n.improvement = 0.0f;
    n.valid = 0;
    n.regularized = 0;
    for (int k = 0; k < 6; k++) n.pad[k] = 0;
}

// ------------------------------------------------------------- math.rs

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ double dot3(const double *a, d3 b) {
    return a[0] * b.x + a[1] * b.y + a[2] * b.z;  // math.rs:11-13
}
__device__ __forceinline__ d3 mul_vec3(const double *m, d3 v) {  // math.rs:48-50
    return d3{dot3(m, v), dot3(m + 3, v), dot3(m + 6, v)};
}
__device__ __forceinline__ d3 mul_vec3_t(const double *m, d3 v) {  // transpose(m) * v
    return d3{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
__device__ __forceinline__ d3 normalize3(d3 a) {  // math.rs:39-46
    double n = __builtin_sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (n == 0.0) return a;
    double s = 1.0 / n;
    return d3{a.x * s, a.y * s, a.z * s};
}
__device__ __forceinline__ bool inverse3(const double *m, double *out) {  // math.rs:71-97
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                 m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (fabs(det) < 1e-300) return false;
    double inv = 1.0 / det;
#define CHD_C(r0, c0, r1, c1) \
    (m[(r0)*3 + (c0)] * m[(r1)*3 + (c1)] - m[(r0)*3 + (c1)] * m[(r1)*3 + (c0)])
    out[0] = CHD_C(1, 1, 2, 2) * inv;
    out[1] = -CHD_C(0, 1, 2, 2) * inv;
    out[2] = CHD_C(0, 1, 1, 2) * inv;
    out[3] = -CHD_C(1, 0, 2, 2) * inv;
    out[4] = CHD_C(0, 0, 2, 2) * inv;
    out[5] = -CHD_C(0, 0, 1, 2) * inv;
    out[6] = CHD_C(1, 0, 2, 1) * inv;
    out[7] = -CHD_C(0, 0, 2, 1) * inv;
    out[8] = CHD_C(0, 0, 1, 1) * inv;
#undef CHD_C
    return true;
}

// ---------------------------------------------------------- geometry.rs

__device__ __forceinline__ double dist_coeff(const chd_distortion &d, uint32_t i) {
    return i < d.coeff_count ? d.coeffs[i] : 0.0;
}

__device__ inline void distort_normalized(const chd_distortion &d, double x, double y,
                                          double &ox, double &oy) {  // geometry.rs:315-326
    double k1 = dist_coeff(d, 0), k2 = dist_coeff(d, 1), p1 = dist_coeff(d, 2);
    double p2 = dist_coeff(d, 3), k3 = dist_coeff(d, 4);
    double r2 = x * x + y * y;
    double radial = 1.0 + k1 * r2 + k2 * r2 * r2 + k3 * r2 * r2 * r2;
    ox = x * radial + 2.0 * p1 * x * y + p2 * (r2 + 2.0 * x * x);
    oy = y * radial + p1 * (r2 + 2.0 * y * y) + 2.0 * p2 * x * y;
}

__device__ inline void distort(const chd_distortion &d, double ix, double iy, double &ox,
                               double &oy) {  // geometry.rs:328-333
    double nx = (ix - d.center[0]) / d.normalization[0];
    double ny = (iy - d.center[1]) / d.normalization[1];
    double dx, dy;
    distort_normalized(d, nx, ny, dx, dy);
    ox = dx * d.normalization[0] + d.center[0];
    oy = dy * d.normalization[1] + d.center[1];
}

__device__ inline void undistort(const chd_distortion &d, double px, double py, double &ox,
                                 double &oy) {  // geometry.rs:335-356
    double obx = (px - d.center[0]) / d.normalization[0];
    double oby = (py - d.center[1]) / d.normalization[1];
    double ex = obx, ey = oby;
    for (int i = 0; i < 20; i++) {
        double dx, dy;
        distort_normalized(d, ex, ey, dx, dy);
        double nx = ex + (obx - dx), ny = ey + (oby - dy);
        double delta = rmax(fabs(nx - ex), fabs(ny - ey));
        ex = nx;
        ey = ny;
        if (delta < 1e-12) break;
    }
    ox = ex * d.normalization[0] + d.center[0];
    oy = ey * d.normalization[1] + d.center[1];
}

// reflect_image_axis (geometry.rs:143-150), written as two value selects
// (a branch that negates one of two referenced locals makes the compiler
// address them through a stack slot).
__device__ __forceinline__ void reflect_axis(int flip, double &x, double &y) {
    const double rx = flip == 0 ? -x : x;
    const double ry = flip == 1 ? -y : y;
    x = rx;
    y = ry;
}

// pixel_to_ray (geometry.rs:152-181)
__device__ inline void pixel_to_ray(const chd_camera &c, double px, double py, d3 &origin,
                                    d3 &direction) {
    double ix = px, iy = py;
    if (c.has_distortion) undistort(c.distortion, px, py, ix, iy);
    d3 d = normalize3(mul_vec3(c.k_inverse, d3{ix, iy, 1.0}));
    if (c.flip_around_x >= 0) {
        double x = d.x, y = d.y;
        reflect_axis(c.flip_around_x, x, y);
        d = normalize3(d3{x, y, d.z});
    }
    d3 world;
    if (c.pose_kind == 0) {
        origin = d3{c.center[0], c.center[1], c.center[2]};
        world = mul_vec3(c.rotation_cw, d);
    } else {
        origin = d3{c.virtual_center[0], c.virtual_center[1], c.virtual_center[2]};
        world = mul_vec3(c.reflect, mul_vec3(c.real_cw, d));
    }
    direction = normalize3(mul_vec3(c.orientation_correction, normalize3(world)));
}

__device__ __forceinline__ bool is_identity(const double *m) {
    return m[0] == 1.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 0.0 && m[4] == 1.0 &&
           m[5] == 0.0 && m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0;
}

// project with the field-of-view guard (geometry.rs:196-243)
__device__ inline bool project(const chd_camera &c, d3 point, double &ox, double &oy) {
    if (!is_identity(c.orientation_correction)) {
        const double *ctr = c.pose_kind == 0 ? c.center : c.virtual_center;
        d3 s{point.x - ctr[0], point.y - ctr[1], point.z - ctr[2]};
        d3 r = mul_vec3_t(c.orientation_correction, s);
        point = d3{ctr[0] + r.x, ctr[1] + r.y, ctr[2] + r.z};
    }
    d3 cam;
    if (c.pose_kind == 0) {
        d3 r = mul_vec3(c.rotation_wc, point);
        cam = d3{r.x + c.translation_wc[0], r.y + c.translation_wc[1], r.z + c.translation_wc[2]};
    } else {
        d3 s{point.x - c.virtual_center[0], point.y - c.virtual_center[1],
             point.z - c.virtual_center[2]};
        cam = mul_vec3_t(c.real_cw, mul_vec3(c.reflect, s));
    }
    if (cam.z <= 0.0) return false;
    double nx = cam.x / cam.z, ny = cam.y / cam.z;
    reflect_axis(c.flip_around_x, nx, ny);
    const double *k = c.k;
    double ix = k[0] * nx + k[1] * ny + k[2];
    double iy = k[4] * ny + k[5];
    double mx = (double)c.width * 0.5, my = (double)c.height * 0.5;
    if (ix < -mx || iy < -my || ix > (double)c.width + mx || iy > (double)c.height + my)
        return false;
    if (c.has_distortion) {
        distort(c.distortion, ix, iy, ox, oy);
    } else {
        ox = ix;
        oy = iy;
    }
    return true;
}

// ----------------------------------------------------- image.rs / align.rs

// A luminance plane stored with one replicated extra column and row
// (pitch = width + 1, height + 1 rows): Plane::sample reads
// at(min(x0 + 1, w - 1)), and at x0 = w - 1 the padding holds that same
// value, so the pair (x0, x0 + 1) can always be loaded together.
// A transposed plane holds the same padded samples column by column (pitch
// = height + 1, width + 1 stored columns), so that the pair (y0, y0 + 1) is
// adjacent instead; target planes whose epipolar lines run mostly vertically
// get such a copy, which keeps the samples of neighbouring depth hypotheses
// in few cache lines.
struct DevPlane {
    const float *data;
    const float4 *quad;      // 2 x 2 neighbourhood per stored element, or null
    uint32_t width, height;  // logical size of the plane
    uint32_t pitch;          // f32 between stored rows (columns when transposed)
    uint32_t lines;          // stored rows (columns when transposed), padding included
    uint32_t transposed;
};

struct DevWarp {
    const float2 *points;
    uint32_t step, columns, rows;
    float inverse_step;  // 1 / step when step is a power of two (then exact), else 0
};

// The synthesis stage's inputs on the device: one source (SynthSource's
// fields that stage reads, the warp's confidence beside its points), the
// depth map's node arrays and the canvas.
constexpr uint32_t kSynthBandRows = 64;  // png16.rs:29: rows per output band
struct SynthSourceDev {
    DevWarp warp;
    const float *confidence;  // warp confidence per node; null = none
    uint32_t mosaic_width, mosaic_height;
    float magnification, source_confidence;
    int32_t reference, has_focus_distance;
    double focus_distance;
};
struct DepthMapDev {
    const double *depth;  // NaN = None; null = no map
    const float *confidence;
    uint32_t columns, rows, step;
};
struct SynthCanvasDev {
    float crop_x, crop_y, scale, feather_scale;  // feather_scale = max(feather_px, 1)
    uint32_t width, height;
};

// Plane::sample (image.rs:451-466) on a padded row-major plane.
__device__ __forceinline__ bool plane_sample(const DevPlane &p, float x, float y, float &out) {
    if (!(x >= 0.0f && y >= 0.0f && x <= (float)(p.width - 1) && y <= (float)(p.height - 1)))
        return false;
    float fx = floorf(x), fy = floorf(y);
    uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
    float tx = x - fx, ty = y - fy;
    const float *r0 = p.data + (size_t)y0 * p.pitch + x0;
    const float *r1 = r0 + p.pitch;
    float top = r0[0] * (1.0f - tx) + r0[1] * tx;
    float bottom = r1[0] * (1.0f - tx) + r1[1] * tx;
    out = top * (1.0f - ty) + bottom * ty;
    return true;
}

// Buffer resource over a plane's quad array (gfx9-family raw buffer, 32-bit
// data format, 16 bytes per stored element); reads past the array return 0
// instead of faulting. Callers pass a plane that is uniform across the
// wavefront; reading the address through readfirstlane keeps the resource
// in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const DevPlane &p) {
    const uint64_t a = (uint64_t)p.quad;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const uint32_t bytes = __builtin_amdgcn_readfirstlane(p.pitch * p.lines * 16u);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), (short)0,
                                             (int)bytes, 0x00020000);
}

// Plane::sample without the bounds test, for points proven inside
// [0, w-1] x [0, h-1], in two halves for callers that issue the loads of
// several samples before interpolating any of them. One 16-byte load of the
// quad array fetches (x0, x0 + 1) of rows y0 and y0 + 1, or, on a
// transposed plane, (y0, y0 + 1) of columns x0 and x0 + 1 (the texture path
// takes about the same time per load instruction for 8 and 16 bytes a lane,
// so this halves its work against two 8-byte loads of the plane). The
// interpolation is the same either way. For x >= 0, truncation is floor and
// x - floor(x) is exact, which is what the hardware fract computes. Stored
// lines and pitches stay below 2^24, so the line product fits the full-rate
// 24-bit multiply.
struct PendingSample {
    float2 a, b;  // the pair on line y0 (x0) and the pair on the next line
    float tx, ty;
};
template <bool Transposed>
__device__ __forceinline__ PendingSample plane_sample_issue(__amdgpu_buffer_rsrc_t rsrc,
                                                           uint32_t pitch, float x, float y) {
    const uint32_t line = Transposed ? (uint32_t)x : (uint32_t)y;
    const uint32_t along = Transposed ? (uint32_t)y : (uint32_t)x;
    const uint32_t offset = (__umul24(line, pitch) + along) * 16u;
    const float4 q =
        __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, offset, 0, 0));
    PendingSample p;
    p.a = make_float2(q.x, q.y);
    p.b = make_float2(q.z, q.w);
    p.tx = __builtin_amdgcn_fractf(x);
    p.ty = __builtin_amdgcn_fractf(y);
    return p;
}
template <bool Transposed>
__device__ __forceinline__ float plane_sample_finish(const PendingSample &p) {
    const float p00 = p.a.x, p10 = Transposed ? p.b.x : p.a.y;
    const float p01 = Transposed ? p.a.y : p.b.x, p11 = p.b.y;
    const float top = p00 * (1.0f - p.tx) + p10 * p.tx;
    const float bottom = p01 * (1.0f - p.tx) + p11 * p.tx;
    return top * (1.0f - p.ty) + bottom * p.ty;
}

// Warp::map (align.rs:84-111)
__device__ __forceinline__ bool warp_map(const DevWarp &w, float x, float y, float2 &out) {
    // x / step; a power-of-two step divides exactly as a multiplication by
    // its reciprocal (both round the same exact quotient)
    const bool scaled = w.inverse_step != 0.0f;
    float fx = scaled ? x * w.inverse_step : x / (float)w.step;
    float fy = scaled ? y * w.inverse_step : y / (float)w.step;
    if (fx < 0.0f || fy < 0.0f) return false;
    uint32_t c0 = min(sat_u32(floorf(fx)), w.columns - 1);
    uint32_t r0 = min(sat_u32(floorf(fy)), w.rows - 1);
    uint32_t c1 = min(c0 + 1, w.columns - 1), r1 = min(r0 + 1, w.rows - 1);
    float tx = fx - (float)c0, ty = fy - (float)r0;
    float2 a = w.points[(size_t)r0 * w.columns + c0];
    float2 b = w.points[(size_t)r0 * w.columns + c1];
    float2 c = w.points[(size_t)r1 * w.columns + c0];
    float2 d = w.points[(size_t)r1 * w.columns + c1];
    float topx = a.x * (1.0f - tx) + b.x * tx;
    float botx = c.x * (1.0f - tx) + d.x * tx;
    float ox = topx * (1.0f - ty) + botx * ty;
    float topy = a.y * (1.0f - tx) + b.y * tx;
    float boty = c.y * (1.0f - tx) + d.y * tx;
    float oy = topy * (1.0f - ty) + boty * ty;
    if (ox != ox || oy != oy) return false;
    out = make_float2(ox, oy);
    return true;
}

// plane coordinate of a raster position: ((p - 0.5) * 0.5) as f32
// (depth.rs:1774-1800). Halving is exact and commutes with rounding, so
// round(p - 0.5) * 0.5 == round(p * 0.5 - 0.25), which one fused
// multiply-add with the exact product p * 0.5 computes.
__device__ __forceinline__ float to_plane(double p) {
    return (float)__builtin_fma(p, 0.5, -0.25);
}

}  // namespace chd
