/*
 * nectar_denoise.h — C ABI of the learned denoiser in libnectar_hip.so.
 *
 * Replaces the OIDN stage of the reference's readback chain
 * (OIDNDenoiser::setup/denoise, csrc/data/include/data/denoise.h:25-83, fed
 * with beauty, albedo := DIFFUSE AOV and normal := WORLD_NORMAL AOV by
 * RenderEngine::RenderEngine, csrc/engine/src/engine.cpp:25-28).  OIDN's
 * network is not available; this is a U-Net of its own, trained on this
 * library's path-traced images, with BF16 MFMA convolutions (FP32
 * accumulation) for inference and training.  Besides OIDN's inputs it takes
 * the DEPTH AOV (render_layers.h:11-14; distance to the first hit, 0 where
 * the camera ray misses).
 *
 * Network: inputs log1p(beauty), log1p(albedo), the normal clamped to
 * [-1, 1] and log1p(depth) (10 channels, padded to 16; NaN reads as 0);
 * five levels of 32/48/64/96/128 channels, two 3x3 convolutions + ReLU per
 * level and side, 2x2 average pooling, nearest 2x upsampling, skip
 * concatenation [encoder, upsampled]; a 3x3 output convolution adds a
 * residual to log1p(beauty); the result is expm1(max(y, 0)).
 *
 * Parameters are one fp32 vector: for every convolution in the order
 * enc0a enc0b enc1a enc1b enc2a enc2b enc3a enc3b enc4a enc4b
 * dec3a dec3b dec2a dec2b dec1a dec1b dec0a dec0b out,
 * weights [cout][3*3][cin] then bias [cout].
 *
 * Images are planar fp32 on the device: N x 3 x H x W per input.
 */
#ifndef NECTAR_DENOISE_H
#define NECTAR_DENOISE_H

#include <stddef.h>
#include <stdint.h>

#include "nectar_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nr_unet nr_unet;

/* Number of fp32 parameters of the network. */
size_t nr_unet_param_count(void);

/* Create the network on HIP device `device` with seeded He-normal weights. NULL on error. */
nr_unet* nr_unet_create(int device, uint32_t seed);
void nr_unet_destroy(nr_unet* net);

/* Copy parameters in / out (host fp32, nr_unet_param_count() values). */
int nr_unet_set_params(nr_unet* net, const float* params, size_t n);
int nr_unet_get_params(nr_unet* net, float* params, size_t n);

/* Denoise N images of H x W (any size; padded internally to multiples of 16).
 * d_beauty, d_albedo, d_normal, d_out: device planar fp32, N x 3 x H x W;
 * d_depth: device fp32, N x 1 x H x W. */
int nr_unet_forward(nr_unet* net, const float* d_beauty, const float* d_albedo, const float* d_normal,
                    const float* d_depth, int32_t N, int32_t H, int32_t W, float* d_out);

/* One training step on N crops of H x W (multiples of 16): forward, L1 loss
 * on log1p radiance against d_target, backward, Adam (beta 0.9/0.999, eps
 * 1e-8) with learning rate lr.  *loss receives the mean loss before the update. */
int nr_unet_train_step(nr_unet* net, const float* d_beauty, const float* d_albedo, const float* d_normal,
                       const float* d_depth, const float* d_target, int32_t N, int32_t H, int32_t W, float lr,
                       float* loss);

/* Loss and its gradient with respect to every parameter (host fp32,
 * nr_unet_param_count() values), without updating: the check entry point. */
int nr_unet_gradients(nr_unet* net, const float* d_beauty, const float* d_albedo, const float* d_normal,
                      const float* d_depth, const float* d_target, int32_t N, int32_t H, int32_t W, float* grad,
                      float* loss);

/* Denoise a view's BEAUTY layer using its DIFFUSE, WORLD_NORMAL and DEPTH
 * layers into d_out (device planar fp32, 3 x H x W).  Waits for the view's render. */
int nr_denoise(nr_ctx* ctx, int view, nr_unet* net, float* d_out);

/* GPU time (ms) of the last nr_unet_forward / nr_denoise and the FLOPs it
 * executed (2 x multiply-adds of the convolutions at the padded size). */
int nr_unet_last_stats(nr_unet* net, float* ms, double* flops);

#ifdef __cplusplus
}
#endif

#endif /* NECTAR_DENOISE_H */
