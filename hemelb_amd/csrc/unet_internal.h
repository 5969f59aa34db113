// Launches of the warm-start U-Net kernels (unet_kernels.hip) for the host
// orchestration (unet.cpp).  Plain pointers; every function returns 0 or a
// HIP error code.  bfloat16 tensors are uint16 bit patterns, channels-last.
#pragma once

#include <cstdint>

namespace hlbn {

struct ConvGeom {
  int N, D, H, W;
};

// y[p][yoff + o] = (relu)(sum_t,c x[p + off_t][c] w[t][o][c] + bias[o]) into
// a binary32 and/or bfloat16 tensor with yc channels.  ci, co multiples of 16.
// ws: conv_fwd_ws(g, ci, co) floats of workspace for the K slices of coarse
// levels (null when that is 0).
int conv_fwd(const ConvGeom& g, const uint16_t* x, int ci, const uint16_t* w,
             const float* bias, int co, float* y32, uint16_t* y16, int yc, int yoff, int relu,
             float* ws, void* stream);
int64_t conv_fwd_ws(const ConvGeom& g, int ci, int co);
// Position chunk of the backward-weight partial sums.
int64_t wgrad_chunk(const ConvGeom& g);
// dw[t][o][c] = sum_p dy[p][o] x[p + off_t][c]; part holds
// wgrad_parts(g, ci, co) floats (partial sums over runs of positions, added
// in a fixed order).
int conv_wgrad(const ConvGeom& g, const uint16_t* x, int ci, const uint16_t* dy, int co,
               float* part, float* dw, void* stream);
int64_t wgrad_parts(const ConvGeom& g, int ci, int co);
// out[c] = sum_p a[p][c] (chunked, fixed order); part holds ceil(P/chunk)*C.
int channel_sum(const float* a, int64_t P, int C, float* part, int64_t chunk, float* out,
                void* stream);
// bfloat16 copies of w [27][co][ci] and of its flipped transpose
// wt[26 - t][ci][co] (the weights of the backward-data convolution).
int flip_transpose(const float* w, int co, int ci, uint16_t* w16, uint16_t* wt16,
                   void* stream);
int relu_bwd(const float* dpost, const uint16_t* post, int64_t n, uint16_t* dpre16,
             float* dpre32, void* stream);
int pool(const ConvGeom& fine, int C, const uint16_t* x, uint16_t* y, void* stream);
int pool_bwd_add(const ConvGeom& fine, int C, const float* dy, float* dx, void* stream);
int up(const ConvGeom& coarse, int C, const uint16_t* x, uint16_t* y, int yc, int yoff,
       void* stream);
int up_bwd(const ConvGeom& coarse, int C, const float* dy, int dyc, int dyoff, float* dx,
           void* stream);
int copy_channels(int64_t P, int C, const uint16_t* x, uint16_t* y, int yc, int yoff,
                  void* stream);
int take_channels(int64_t P, int C, const float* dy, int dyc, int dyoff, float* dx, int add,
                  void* stream);
int pack_input(int64_t P, int cin, int cpad, const float* in, uint16_t* x, void* stream);
int head_fwd(int64_t P, int K, int O, const uint16_t* h, const float* W, const float* b,
             float* out, void* stream);
int head_bwd(int64_t P, int K, int O, const float* dout, const uint16_t* h, const float* W,
             float* dh, float* part, int64_t chunk, float* dW, float* db, void* stream);
int mse(int64_t P, int O, const float* out, const float* target, const uint8_t* mask,
        const float* wc, float inv_count, float* dout, float* lterm, void* stream);
int adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
         float eps, float wd, float c1, float c2, void* stream);

}  // namespace hlbn
