// proposal.hip - the learned hypothesis-band proposal (gfx950).
//
// Build-only component with no counterpart in the reference (the reference
// sweeps every hypothesis on the CPU): a 7-layer 3x3 convolutional
// encoder-decoder at quarter resolution reads the reference and two target
// planes and predicts, per pixel, a centre label, a log band half-width and
// a confidence. The cost volume then sweeps only +-8 labels around the
// centre where the tile's confidence is high and falls back to the full
// sweep elsewhere (the band windows below feed cost_tile).
//
// Arithmetic: BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_16x16x16_bf16). Every convolution is an implicit GEMM:
// rows are output pixels, columns output channels, and the reduction runs
// over (tap, input channel) with k = tap * cin_p + c. Operand layout of the
// 16x16x16 instruction, lane l = 16 * q + r (q = l / 16, r = l % 16):
//   A: row r, k = 4q .. 4q + 3;  B: column r, k = 4q .. 4q + 3;
//   D: rows 4q .. 4q + 3 of column r.
// Backward: dX in gather form (each input pixel collects dY W^T over the
// output pixels that read it, written once), dW = X_col^T dY from operands
// staged transposed in LDS with the bias gradient alongside, accumulated in
// registers per block and added with one f32 atomic per element and block;
// Adam on f32 master weights, then BF16 copies in both layouts.
#include <hip/hip_runtime.h>

#include "proposal.hpp"

namespace chd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Input pixel of output pixel (oy, ox) at tap (ky, kx); false in the zero
// padding.
__device__ __forceinline__ bool conv_source(const ConvLayer &L, uint32_t oy, uint32_t ox,
                                            uint32_t ky, uint32_t kx, uint32_t &iy,
                                            uint32_t &ix) {
    if (L.upsample) {
        const int uy = (int)oy + (int)ky - 1, ux = (int)ox + (int)kx - 1;
        if (uy < 0 || ux < 0 || uy >= (int)(2 * L.hin) || ux >= (int)(2 * L.win)) return false;
        iy = (uint32_t)uy >> 1;
        ix = (uint32_t)ux >> 1;
        return true;
    }
    const int y = (int)(oy * L.stride) + (int)ky - 1, x = (int)(ox * L.stride) + (int)kx - 1;
    if (y < 0 || x < 0 || y >= (int)L.hin || x >= (int)L.win) return false;
    iy = (uint32_t)y;
    ix = (uint32_t)x;
    return true;
}

// ------------------------------------------------------------- forward

// One wavefront per 16 output pixels; all cout_p / 16 column tiles.
// Chunks = cin_p / 16 and Tiles = cout_p / 16 are template parameters, so
// the tap and chunk loops unroll and every operand load of the wave is in
// flight before the first product; the products run in (tap, chunk) order.
template <bool OutF32, uint32_t Chunks, uint32_t Tiles>
__global__ void __launch_bounds__(256) k_conv_fwd(ConvLayer L, uint32_t batch,
                                                  const uint16_t *__restrict__ in,
                                                  const uint16_t *__restrict__ wt,
                                                  const float *__restrict__ bias, void *out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t plane = L.hout * L.wout, M = batch * plane;
    const uint32_t m0 = wave * 16;
    if (m0 >= M) return;
    const uint32_t q = lane >> 4, r = lane & 15u;
    const uint32_t m = m0 + r;
    const bool mv = m < M;
    const uint32_t n = mv ? m / plane : 0, rem = mv ? m % plane : 0;
    const uint32_t oy = rem / L.wout, ox = rem % L.wout;
    constexpr uint32_t cin_p = Chunks * 16, K = 9 * cin_p;
    bf16x4 a[9][Chunks];
#pragma unroll
    for (uint32_t tap = 0; tap < 9; tap++) {
        uint32_t iy = 0, ix = 0;
        const bool valid = mv && conv_source(L, oy, ox, tap / 3, tap % 3, iy, ix);
        const uint16_t *px = in + (((size_t)n * L.hin + iy) * L.win + ix) * cin_p;
#pragma unroll
        for (uint32_t c = 0; c < Chunks; c++) {
            a[tap][c] = bf16x4{0, 0, 0, 0};
            if (valid) a[tap][c] = *(const bf16x4 *)(px + c * 16 + 4 * q);
        }
    }
    f32x4 acc[Tiles];
#pragma unroll
    for (uint32_t t = 0; t < Tiles; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t tap = 0; tap < 9; tap++) {
#pragma unroll
        for (uint32_t c = 0; c < Chunks; c++) {
            const uint32_t k0 = tap * cin_p + c * 16;
#pragma unroll
            for (uint32_t t = 0; t < Tiles; t++) {
                const bf16x4 b = *(const bf16x4 *)(wt + (size_t)(t * 16 + r) * K + k0 + 4 * q);
                acc[t] = mfma16(a[tap][c], b, acc[t]);
            }
        }
    }
#pragma unroll
    for (uint32_t t = 0; t < Tiles; t++) {
        const uint32_t ch = t * 16 + r;
        for (uint32_t v = 0; v < 4; v++) {
            const uint32_t p = m0 + 4 * q + v;
            if (p >= M) continue;
            float y = 0.0f;
            if (ch < L.cout) {
                y = acc[t][v] + bias[ch];
                if (L.relu) y = fmaxf(y, 0.0f);
            }
            if (OutF32)
                ((float *)out)[(size_t)p * (Tiles * 16) + ch] = y;
            else
                ((uint16_t *)out)[(size_t)p * (Tiles * 16) + ch] = f2bf(y);
        }
    }
}

// ------------------------------------------------------------ backward

// dy (f32) -> BF16, zero where the layer's ReLU output was zero.
__global__ void k_grad_bf16(const float *__restrict__ dy, const uint16_t *__restrict__ y,
                            size_t n, int relu, uint16_t *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = relu && !(bf2f(y[i]) > 0.0f) ? 0.0f : dy[i];
    out[i] = f2bf(g);
}

// dX in gather form: one wavefront per 16 input pixels x 16 input channels,
// dX[p][c] = sum over (tap, output pixel o that reads p at that tap, co) of
// dY[o][co] W[tap][c][co], as an implicit GEMM over (tap, candidate, co).
// A stride-1 or stride-2 layer has at most one candidate output pixel per
// tap; an upsampling layer (nearest 2x2) up to four. Every element of dX
// is written once: no atomics, no clearing.
template <uint32_t CoutChunks, bool Upsample>
__global__ void __launch_bounds__(256) k_conv_dx_gather(ConvLayer L, uint32_t batch,
                                                        const uint16_t *__restrict__ dy,
                                                        const uint16_t *__restrict__ wk,
                                                        float *__restrict__ dx) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t plane_in = L.hin * L.win, M = batch * plane_in;
    const uint32_t ctiles = L.cin_p / 16, mtiles = (M + 15) / 16;
    if (wave >= mtiles * ctiles) return;
    const uint32_t m0 = (wave / ctiles) * 16, c0 = (wave % ctiles) * 16;
    const uint32_t q = lane >> 4, r = lane & 15u;
    constexpr uint32_t cout_p = CoutChunks * 16;
    const uint32_t m = m0 + r;
    const bool mv = m < M;
    const uint32_t n = mv ? m / plane_in : 0, rem = mv ? m % plane_in : 0;
    const int iy = (int)(rem / L.win), ix = (int)(rem % L.win);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t tap = 0; tap < 9; tap++) {
        const int ky = (int)(tap / 3), kx = (int)(tap % 3);
        constexpr uint32_t kCandidates = Upsample ? 4 : 1;
#pragma unroll
        for (uint32_t cand = 0; cand < kCandidates; cand++) {
            int oy, ox;
            bool valid = mv;
            if (Upsample) {
                // the upsampled grid rows 2 iy and 2 iy + 1 repeat this row
                oy = 2 * iy + (int)(cand >> 1) - ky + 1;
                ox = 2 * ix + (int)(cand & 1u) - kx + 1;
            } else {
                const int ty = iy - ky + 1, tx = ix - kx + 1;
                const int s = (int)L.stride;
                valid = valid && ty >= 0 && tx >= 0 && ty % s == 0 && tx % s == 0;
                oy = ty / s;
                ox = tx / s;
            }
            valid = valid && oy >= 0 && ox >= 0 && oy < (int)L.hout && ox < (int)L.wout;
            const uint16_t *row =
                dy + (((size_t)n * L.hout + (valid ? oy : 0)) * L.wout + (valid ? ox : 0)) * cout_p;
            const uint16_t *wrow = wk + (size_t)(tap * L.cin_p + c0 + r) * cout_p;
#pragma unroll
            for (uint32_t cc = 0; cc < CoutChunks; cc++) {
                bf16x4 a = {0, 0, 0, 0};
                if (valid) a = *(const bf16x4 *)(row + cc * 16 + 4 * q);
                const bf16x4 b = *(const bf16x4 *)(wrow + cc * 16 + 4 * q);
                acc = mfma16(a, b, acc);
            }
        }
    }
    for (uint32_t v = 0; v < 4; v++) {
        const uint32_t p = m0 + 4 * q + v;
        if (p < M) dx[(size_t)p * L.cin_p + c0 + r] = acc[v];
    }
}

// dW = X_col^T dY and db = sum dY, staged through LDS: a block takes
// kDwSteps runs of kDwPixels (= its thread count) output pixels; per run
// each thread loads its pixel's dY row and, per tap, its gathered input row
// with 16-byte loads and stores them transposed ([channel][pixel]), so that
// each lane of the products reads four consecutive pixels of one channel
// as one 8-byte LDS load. The (tap, c tile, co tile) output tiles are dealt
// to the four wavefronts and accumulate in registers over the block's
// pixels; one f32 atomic per element and block at the end.
constexpr uint32_t kDwPixels = 256;
constexpr uint32_t kDwSteps = 16;
constexpr uint32_t kDwPitch = kDwPixels + 4;  // bf16 per LDS row (8-byte aligned rows)
template <uint32_t CinTiles, uint32_t CoutTiles>
__global__ void __launch_bounds__(256) k_conv_dw_staged(ConvLayer L, uint32_t batch,
                                                        const uint16_t *__restrict__ in,
                                                        const uint16_t *__restrict__ dy,
                                                        float *__restrict__ dw,
                                                        float *__restrict__ db) {
    constexpr uint32_t cin_p = CinTiles * 16, cout_p = CoutTiles * 16;
    constexpr uint32_t kTiles = 9 * CinTiles * CoutTiles, kPerWave = (kTiles + 3) / 4;
    __shared__ __align__(16) uint16_t dyT[cout_p * kDwPitch];
    __shared__ __align__(16) uint16_t xT[cin_p * kDwPitch];
    __shared__ float bias_sum[cout_p];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t q = lane >> 4, r = lane & 15u;
    const uint32_t plane = L.hout * L.wout, M = batch * plane;
    f32x4 acc[kPerWave];
#pragma unroll
    for (uint32_t u = 0; u < kPerWave; u++) acc[u] = f32x4{0.f, 0.f, 0.f, 0.f};
    // bias partial of channel tid % cout_p over pixel segment tid / cout_p
    constexpr uint32_t kSegments = 256 / cout_p, kSegment = kDwPixels / kSegments;
    float bias = 0.0f;
    if (tid < cout_p) bias_sum[tid] = 0.0f;
    for (uint32_t step = 0; step < kDwSteps; step++) {
        const uint32_t p0 = (blockIdx.x * kDwSteps + step) * kDwPixels;
        if (p0 >= M) break;
        const uint32_t g = p0 + tid;  // this thread's pixel
        const bool live = g < M;
        const uint32_t nn = live ? g / plane : 0, rem = live ? g % plane : 0;
        const uint32_t oy = rem / L.wout, ox = rem % L.wout;
        __syncthreads();
        {
            const uint4 *src = (const uint4 *)(dy + (size_t)(live ? g : 0) * cout_p);
#pragma unroll
            for (uint32_t j = 0; j < cout_p / 8; j++) {
                uint4 v = live ? src[j] : make_uint4(0u, 0u, 0u, 0u);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (uint32_t h = 0; h < 8; h++)
                    dyT[(j * 8 + h) * kDwPitch + tid] = (uint16_t)(w[h >> 1] >> (16 * (h & 1u)));
            }
        }
        __syncthreads();
        {
            const uint32_t co = tid % cout_p, seg = tid / cout_p;
            for (uint32_t p = seg * kSegment; p < (seg + 1) * kSegment; p++)
                bias += bf2f(dyT[co * kDwPitch + p]);
        }
#pragma unroll
        for (uint32_t tap = 0; tap < 9; tap++) {
            if (tap > 0) __syncthreads();
            {
                uint32_t iy = 0, ix = 0;
                const bool valid = live && conv_source(L, oy, ox, tap / 3, tap % 3, iy, ix);
                const uint4 *src =
                    (const uint4 *)(in + (((size_t)nn * L.hin + iy) * L.win + ix) * cin_p);
#pragma unroll
                for (uint32_t j = 0; j < cin_p / 8; j++) {
                    uint4 v = valid ? src[j] : make_uint4(0u, 0u, 0u, 0u);
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (uint32_t h = 0; h < 8; h++)
                        xT[(j * 8 + h) * kDwPitch + tid] = (uint16_t)(w[h >> 1] >> (16 * (h & 1u)));
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t t = 0; t < CinTiles * CoutTiles; t++) {
                const uint32_t j = tap * CinTiles * CoutTiles + t;
                if (j % 4 != wave) continue;
                const uint32_t ct = t / CoutTiles, ot = t % CoutTiles;
                const uint16_t *arow = xT + (ct * 16 + r) * kDwPitch + 4 * q;
                const uint16_t *brow = dyT + (ot * 16 + r) * kDwPitch + 4 * q;
                f32x4 a_acc = acc[j / 4];
#pragma unroll 4
                for (uint32_t s = 0; s < kDwPixels; s += 16)
                    a_acc = mfma16(*(const bf16x4 *)(arow + s), *(const bf16x4 *)(brow + s), a_acc);
                acc[j / 4] = a_acc;
            }
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < kPerWave; u++) {
        const uint32_t j = u * 4 + wave;
        if (j >= kTiles) continue;
        const uint32_t tap = j / (CinTiles * CoutTiles), t = j % (CinTiles * CoutTiles);
        const uint32_t ct = t / CoutTiles, ot = t % CoutTiles;
        for (uint32_t v = 0; v < 4; v++)
            atomicAdd(dw + (size_t)(tap * cin_p + ct * 16 + 4 * q + v) * cout_p + ot * 16 + r,
                      acc[u][v]);
    }
    __syncthreads();
    atomicAdd(&bias_sum[tid % cout_p], bias);
    __syncthreads();
    if (tid < cout_p) atomicAdd(db + tid, bias_sum[tid]);
}

// ------------------------------------------------------------ loss

// Per pixel of the last layer: outputs o0 (centre, label units = half + k o0
// with half = (L - 1) / 2 and k = (L - 1) / 8 for L labels: 64 + 16 o0 at 129),
// o1 (log band half-width s), o2 (confidence logit z); label t (0 = no
// label). Loss = |t - c| / 8 (the centre) + a e^-s + s with a = |t - c|
// held constant (the Laplace scale of the centre's error, which does not
// pull on the centre) + BCE(z, |t - c| <= 6): the confidence learns whether
// a band of +-8 labels around the centre covers the full sweep's label with
// a margin of 2. Divided by `norm` pixels; writes d loss / d outputs.
__global__ void k_proposal_loss(const float *__restrict__ out, const uint16_t *__restrict__ labels,
                                uint32_t M, uint32_t cout_p, float norm, float half, float k,
                                float *__restrict__ dout, float *__restrict__ totals) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.0f, valid = 0.0f;
    if (p < M) {
        const float *o = out + (size_t)p * cout_p;
        float *d = dout + (size_t)p * cout_p;
        for (uint32_t c = 0; c < cout_p; c++) d[c] = 0.0f;
        const uint32_t t = labels[p];
        if (t > 0) {
            const float centre = half + k * o[0];
            const float s = fminf(fmaxf(o[1], -3.0f), 5.0f);
            const float z = o[2];
            const float e = __expf(-s), diff = centre - (float)t, ad = fabsf(diff);
            // the scale term sees the error capped at 32 labels, so one far
            // miss cannot swamp the width gradient
            const float al = fminf(ad, 32.0f);
            const float y = ad <= 6.0f ? 1.0f : 0.0f;
            const float sig = 1.0f / (1.0f + __expf(-z));
            loss = ad * 0.125f + al * e + s + (fmaxf(z, 0.0f) - z * y + __logf(1.0f + __expf(-fabsf(z))));
            d[0] = (diff > 0.0f ? 1.0f : diff < 0.0f ? -1.0f : 0.0f) * (k * 0.125f) / norm;
            d[1] = (o[1] > -3.0f && o[1] < 5.0f) ? (1.0f - al * e) / norm : 0.0f;
            d[2] = (sig - y) / norm;
            valid = 1.0f;
        }
    }
    // block sums of loss and labelled pixels
    __shared__ float sl[256], sv[256];
    sl[threadIdx.x] = loss;
    sv[threadIdx.x] = valid;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sl[threadIdx.x] += sl[threadIdx.x + s];
            sv[threadIdx.x] += sv[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(totals, sl[0]);
        atomicAdd(totals + 1, sv[0]);
    }
}

// ------------------------------------------------------------ Adam

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint32_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 copies of the f32 master weights [k][cout_p]: wk = [k][cout_p] and
// wt = [cout_p][k]; padded input channels and output channels stay zero.
__global__ void k_weights_bf16(ConvLayer L, const float *__restrict__ w, uint16_t *__restrict__ wk,
                               uint16_t *__restrict__ wt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.k * L.cout_p) return;
    const uint32_t k = i / L.cout_p, co = i % L.cout_p;
    const bool live = (k % L.cin_p) < L.cin && co < L.cout;
    const uint16_t h = live ? f2bf(w[i]) : 0;
    wk[i] = h;
    wt[(size_t)co * L.k + k] = h;
}

// ------------------------------------------------------------ inputs

// Quarter-resolution network input from the half-resolution planes of up
// to three views (reference first): the mean of each 2x2 block, scaled by
// 1/8 (log-luminance lies in [0, ln 1001]), as BF16 channels 0..2 of a
// 16-channel pixel.
__global__ void k_quarter_input(const float *__restrict__ p0, const float *__restrict__ p1,
                                const float *__restrict__ p2, uint32_t pitch0, uint32_t pitch1,
                                uint32_t pitch2, uint32_t plane_w, uint32_t plane_h, uint32_t w,
                                uint32_t h, uint32_t factor, float offset, float scale,
                                uint16_t *__restrict__ out) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const float *planes[3] = {p0, p1, p2};
    const uint32_t pitches[3] = {pitch0, pitch1, pitch2};
    uint16_t *o = out + ((size_t)y * w + x) * 16;
    for (uint32_t c = 0; c < 16; c++) {
        float v = 0.0f;
        if (c < 3 && planes[c]) {
            // factor x factor cell mean, summed row by row (clamped at the edges)
            const float *p = planes[c];
            const uint32_t pt = pitches[c];
            float sum = 0.0f;
            for (uint32_t dy = 0; dy < factor; dy++) {
                const uint32_t yy = min(factor * y + dy, plane_h - 1);
                for (uint32_t dx = 0; dx < factor; dx++)
                    sum += p[(size_t)yy * pt + min(factor * x + dx, plane_w - 1)];
            }
            const float mean = (1.0f / (float)(factor * factor)) * sum;
            v = (mean - offset) * scale;
        }
        o[c] = f2bf(v);
    }
}

// ------------------------------------------------------------ band windows

// One block per tile x tile coarse nodes: the tile's mean confidence decides
// between the band and the full sweep (lo = 1, hi = L - 1). A banded node
// sweeps centre +- w labels, clamped to 1 .. L-1 (label 0, the global warp,
// is always swept), with w = ceil(4 b) for the predicted Laplace scale b =
// exp(log half-width) - the +-4b interval holds 98 % of a Laplace error -
// clamped to [half, max_half]. Node (c, r) reads the network pixel
// (min(c, w - 1), min(r, h - 1)).
__global__ void __launch_bounds__(256) k_band_windows(const float *__restrict__ out,
                                                      uint32_t cout_p, uint32_t w, uint32_t h,
                                                      uint32_t columns, uint32_t rows,
                                                      uint32_t L, float centre_half,
                                                      float centre_k, float threshold,
                                                      uint32_t half, uint32_t max_half, uint32_t tile,
                                                      uint32_t fallback_half,
                                                      uint16_t *__restrict__ lo,
                                                      uint16_t *__restrict__ hi,
                                                      uint32_t *__restrict__ full_tiles) {
    __shared__ float sum[256];
    __shared__ float count[256];
    const uint32_t tiles_x = (columns + tile - 1) / tile;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    float s = 0.0f, cnt = 0.0f;
    for (uint32_t i = threadIdx.x; i < tile * tile; i += blockDim.x) {
        const uint32_t c = tx * tile + i % tile, r = ty * tile + i / tile;
        if (c >= columns || r >= rows) continue;
        const float *o = out + ((size_t)min(r, h - 1) * w + min(c, w - 1)) * cout_p;
        s += 1.0f / (1.0f + __expf(-o[2]));
        cnt += 1.0f;
    }
    sum[threadIdx.x] = s;
    count[threadIdx.x] = cnt;
    __syncthreads();
    for (uint32_t k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            sum[threadIdx.x] += sum[threadIdx.x + k];
            count[threadIdx.x] += count[threadIdx.x + k];
        }
        __syncthreads();
    }
    const bool full = !(sum[0] >= threshold * count[0]);
    if (threadIdx.x == 0 && full) atomicAdd(full_tiles, 1u);
    for (uint32_t i = threadIdx.x; i < tile * tile; i += blockDim.x) {
        const uint32_t c = tx * tile + i % tile, r = ty * tile + i / tile;
        if (c >= columns || r >= rows) continue;
        const size_t node = (size_t)r * columns + c;
        if (full && fallback_half == 0) {
            lo[node] = 1;
            hi[node] = (uint16_t)(L - 1);
            continue;
        }
        const float *o = out + ((size_t)min(r, h - 1) * w + min(c, w - 1)) * cout_p;
        const float centre = rintf(centre_half + centre_k * o[0]);
        const int ctr = (int)fminf(fmaxf(centre, 1.0f), (float)(L - 1));
        if (full) {  // a wide window instead of every label
            lo[node] = (uint16_t)max(1, ctr - (int)fallback_half);
            hi[node] = (uint16_t)min((int)L - 1, ctr + (int)fallback_half);
            continue;
        }
        const float b = __expf(fminf(fmaxf(o[1], -3.0f), 5.0f));
        const int w = (int)fminf(fmaxf(ceilf(4.0f * b), (float)half), (float)max_half);
        lo[node] = (uint16_t)max(1, ctr - w);
        hi[node] = (uint16_t)min((int)L - 1, ctr + w);
    }
}

// ------------------------------------------------------------ launchers

hipError_t launch_conv_fwd(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *in,
                           const uint16_t *wt, const float *bias, void *out, bool out_f32) {
    const uint32_t M = batch * L.hout * L.wout;
    const dim3 grid((M + 63) / 64);
    const uint32_t chunks = L.cin_p / 16, tiles = L.cout_p / 16;
#define CHD_CONV(F, C, T)                                                                     \
    if (out_f32 == F && chunks == C && tiles == T) {                                          \
        hipLaunchKernelGGL((k_conv_fwd<F, C, T>), grid, dim3(256), 0, s, L, batch, in, wt, bias, \
                           out);                                                              \
        return hipGetLastError();                                                             \
    }
    CHD_CONV(false, 1, 1) CHD_CONV(false, 1, 2) CHD_CONV(false, 2, 1) CHD_CONV(false, 2, 2)
    CHD_CONV(true, 1, 1) CHD_CONV(true, 1, 2) CHD_CONV(true, 2, 1) CHD_CONV(true, 2, 2)
#undef CHD_CONV
    return hipErrorInvalidValue;  // channel counts above 32
}

hipError_t launch_grad_bf16(hipStream_t s, const float *dy, const uint16_t *y, size_t n, bool relu,
                            uint16_t *out) {
    hipLaunchKernelGGL(k_grad_bf16, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, dy, y, n,
                       relu ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_conv_dx(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *dy,
                          const uint16_t *wk, float *dx) {
    const uint32_t M = batch * L.hin * L.win;
    const uint32_t waves = ((M + 15) / 16) * (L.cin_p / 16);
    const dim3 grid((waves + 3) / 4), block(256);
    if (L.cout_p == 16 && L.upsample)
        hipLaunchKernelGGL((k_conv_dx_gather<1, true>), grid, block, 0, s, L, batch, dy, wk, dx);
    else if (L.cout_p == 16)
        hipLaunchKernelGGL((k_conv_dx_gather<1, false>), grid, block, 0, s, L, batch, dy, wk, dx);
    else if (L.cout_p == 32 && L.upsample)
        hipLaunchKernelGGL((k_conv_dx_gather<2, true>), grid, block, 0, s, L, batch, dy, wk, dx);
    else if (L.cout_p == 32)
        hipLaunchKernelGGL((k_conv_dx_gather<2, false>), grid, block, 0, s, L, batch, dy, wk, dx);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_conv_dw(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *in,
                          const uint16_t *dy, float *dw, float *db) {
    const uint32_t M = batch * L.hout * L.wout, per_block = kDwPixels * kDwSteps;
    const dim3 grid((M + per_block - 1) / per_block), block(256);
    if (L.cin_p == 16 && L.cout_p == 16)
        hipLaunchKernelGGL((k_conv_dw_staged<1, 1>), grid, block, 0, s, L, batch, in, dy, dw, db);
    else if (L.cin_p == 16 && L.cout_p == 32)
        hipLaunchKernelGGL((k_conv_dw_staged<1, 2>), grid, block, 0, s, L, batch, in, dy, dw, db);
    else if (L.cin_p == 32 && L.cout_p == 16)
        hipLaunchKernelGGL((k_conv_dw_staged<2, 1>), grid, block, 0, s, L, batch, in, dy, dw, db);
    else if (L.cin_p == 32 && L.cout_p == 32)
        hipLaunchKernelGGL((k_conv_dw_staged<2, 2>), grid, block, 0, s, L, batch, in, dy, dw, db);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_proposal_loss(hipStream_t s, const float *out, const uint16_t *labels,
                                uint32_t M, uint32_t cout_p, uint32_t label_count, float norm,
                                float *dout, float *totals) {
    const float half = (float)(label_count - 1) * 0.5f, k = (float)(label_count - 1) * 0.125f;
    hipLaunchKernelGGL(k_proposal_loss, dim3((M + 255) / 256), dim3(256), 0, s, out, labels, M,
                       cout_p, norm, half, k, dout, totals);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, uint32_t n,
                       float lr, float b1, float b2, float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2,
                       eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_weights_bf16(hipStream_t s, const ConvLayer &L, const float *w, uint16_t *wk,
                               uint16_t *wt) {
    const uint32_t n = L.k * L.cout_p;
    hipLaunchKernelGGL(k_weights_bf16, dim3((n + 255) / 256), dim3(256), 0, s, L, w, wk, wt);
    return hipGetLastError();
}

hipError_t launch_quarter_input(hipStream_t s, const float *const planes[3],
                                const uint32_t pitches[3], uint32_t plane_w, uint32_t plane_h,
                                uint32_t w, uint32_t h, uint32_t factor, float offset,
                                float scale, uint16_t *out) {
    hipLaunchKernelGGL(k_quarter_input, dim3((w + 127) / 128, h), dim3(128), 0, s, planes[0],
                       planes[1], planes[2], pitches[0], pitches[1], pitches[2], plane_w, plane_h,
                       w, h, factor, offset, scale, out);
    return hipGetLastError();
}

hipError_t launch_band_windows(hipStream_t s, const float *out, uint32_t cout_p, uint32_t w,
                               uint32_t h, uint32_t columns, uint32_t rows, uint32_t L,
                               float threshold, uint32_t half, uint32_t max_half, uint32_t tile,
                               uint32_t fallback_half, uint16_t *lo, uint16_t *hi,
                               uint32_t *full_tiles) {
    const float centre_half = (float)(L - 1) * 0.5f, centre_k = (float)(L - 1) * 0.125f;
    const uint32_t tiles = ((columns + tile - 1) / tile) * ((rows + tile - 1) / tile);
    hipLaunchKernelGGL(k_band_windows, dim3(tiles), dim3(256), 0, s, out, cout_p, w, h, columns,
                       rows, L, centre_half, centre_k, threshold, half, max_half, tile,
                       fallback_half, lo, hi, full_tiles);
    return hipGetLastError();
}

}  // namespace chd
