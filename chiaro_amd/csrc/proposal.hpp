// proposal.hpp - layer descriptor and launchers of the learned band
// proposal (proposal.hip), shared with the C ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace chd {

// One 3x3 convolution (pad 1) of the proposal network. Activations are
// [batch][h][w][c_p] BF16 with channels padded to a multiple of 16; k =
// 9 * cin_p. With upsample the input is first repeated 2x2 (nearest), so
// the output grid is at most (2 hin) x (2 win).
struct ConvLayer {
    uint32_t cin, cin_p, cout, cout_p;
    uint32_t hin, win, hout, wout;
    uint32_t stride, upsample, relu;
    uint32_t k;
};

hipError_t launch_conv_fwd(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *in,
                           const uint16_t *wt, const float *bias, void *out, bool out_f32);
hipError_t launch_grad_bf16(hipStream_t s, const float *dy, const uint16_t *y, size_t n, bool relu,
                            uint16_t *out);
// dX of one layer (f32, every element written; channels padded to 16 or 32).
hipError_t launch_conv_dx(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *dy,
                          const uint16_t *wk, float *dx);
// dW and db of one layer, added into dw and db (cleared beforehand).
hipError_t launch_conv_dw(hipStream_t s, const ConvLayer &L, uint32_t batch, const uint16_t *in,
                          const uint16_t *dy, float *dw, float *db);
// Labels 1 .. label_count - 1 (0 = unlabelled); centre = (L-1)/2 + (L-1)/8 o0.
hipError_t launch_proposal_loss(hipStream_t s, const float *out, const uint16_t *labels,
                                uint32_t M, uint32_t cout_p, uint32_t label_count, float norm,
                                float *dout, float *totals);
hipError_t launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, uint32_t n,
                       float lr, float b1, float b2, float eps, float bc1, float bc2);
hipError_t launch_weights_bf16(hipStream_t s, const ConvLayer &L, const float *w, uint16_t *wk,
                               uint16_t *wt);
hipError_t launch_quarter_input(hipStream_t s, const float *const planes[3],
                                const uint32_t pitches[3], uint32_t plane_w, uint32_t plane_h,
                                uint32_t w, uint32_t h, uint32_t factor, float offset, float scale,
                                uint16_t *out);
hipError_t launch_band_windows(hipStream_t s, const float *out, uint32_t cout_p, uint32_t w,
                               uint32_t h, uint32_t columns, uint32_t rows, uint32_t L,
                               float threshold, uint32_t half, uint32_t max_half, uint32_t tile,
                               uint32_t fallback_half, uint16_t *lo, uint16_t *hi,
                               uint32_t *full_tiles);

}  // namespace chd
