/* Learned warm start: a 3D U-Net trained and run by the library's own HIP
 * kernels (bfloat16 MFMA convolutions with binary32 accumulation).
 *
 * No counterpart in the reference: HemeLB always starts from equilibrium
 * (Code/lb/InitialCondition.hpp:39-52) or a checkpoint.  The network maps a
 * crop of the signed-distance wall field plus iolet-parameter channels to
 * velocity and pressure; the solver then starts from the equilibrium of the
 * predicted field (hlb_set_distributions).
 *
 * Layout: inputs [N][S][S][S][in_channels] and outputs
 * [N][S][S][S][out_channels], binary32, channels-last (z fastest among the
 * spatial indices).  Parameters are one binary32 vector in this order: for
 * each 3x3x3 convolution e1a, e1b, e2a, e2b, e3a, e3b, e4a, e4b, d3a, d3b,
 * d2a, d2b, d1a, d1b its weights [27][Cout][Cin] (tap 9*kd + 3*kh + kw for
 * offsets kd-1, kh-1, kw-1) then its bias [Cout]; then the 1x1x1 head
 * weights [out_channels][C1] and bias.  Encoder level k has chans[k]
 * channels; the decoder concatenates [upsampled, skip] channels.  The
 * input is zero-padded to 16 channels, so e1a has Cin = 16. */
#ifndef HEMELB_UNET_H
#define HEMELB_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hlbn_net hlbn_net;

/* side: multiple of 8 (three 2x poolings); chans: multiples of 16;
 * in_channels <= 16; out_channels <= 16. */
int hlbn_create(int32_t side, int32_t batch, int32_t in_channels, int32_t out_channels,
                const int32_t chans[4], int32_t device, hlbn_net** out);
void hlbn_destroy(hlbn_net* net);
int64_t hlbn_param_count(const hlbn_net* net);
int hlbn_set_params(hlbn_net* net, const float* params);
int hlbn_get_params(hlbn_net* net, float* params);
/* Gradients of the last hlbn_backward, parameter order. */
int hlbn_get_grads(hlbn_net* net, float* grads);

/* Forward pass of a full batch.  `device` != 0: input/output are device
 * pointers (output may be NULL); otherwise host pointers. */
int hlbn_forward(hlbn_net* net, const float* input, float* output, int32_t device);
/* Loss of the last forward pass against `target` (same layout as the
 * output) over the voxels with mask[v] != 0 ([N][S][S][S]):
 * sum_c w_c (out_c - target_c)^2 / (mask_count * out_channels), and the
 * gradients of every parameter.  `device` as above. */
int hlbn_backward(hlbn_net* net, const float* target, const uint8_t* mask,
                  const float* channel_weights, double mask_count, double* loss,
                  int32_t device);
/* One Adam step (decoupled weight decay) with the gradients of the last
 * hlbn_backward. */
int hlbn_adam(hlbn_net* net, float lr, float beta1, float beta2, float eps,
              float weight_decay);
/* Device bytes held by the network's buffers. */
int64_t hlbn_device_bytes(const hlbn_net* net);
/* Milliseconds of the last forward and backward passes (HIP events). */
int hlbn_last_times(const hlbn_net* net, double* forward_ms, double* backward_ms);
void* hlbn_stream(const hlbn_net* net);
const char* hlbn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
