// Host orchestration of the warm-start 3D U-Net: parameters, activation and
// gradient buffers in HBM, the forward and backward passes as sequences of
// kernel launches on one stream, and Adam.
#include "hemelb_unet.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "unet_internal.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return 1;
}

#define HIPN(expr)                                                       \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define KERN(expr)                                                       \
  do {                                                                   \
    int e_ = (expr);                                                     \
    if (e_) return fail(std::string("kernel launch failed: ") + #expr);  \
  } while (0)

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) {
    n = count;
    return hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1));
  }
};

struct Conv {
  const char* name;
  int ci, co;
  int64_t woff, boff;  // offsets into the parameter vector
  int level;           // resolution level 0..3
  Buf<uint16_t> w16, wt16;
};

}  // namespace

struct hlbn_net {
  int S, N, cin, cinp, cout, C[4];
  int device;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  float fwd_ms = 0, bwd_ms = 0;
  Conv conv[14];  // layers in forward order
  int64_t nparam = 0, head_w = 0, head_b = 0;
  Buf<float> params, grads, adam_m, adam_v;
  int64_t adam_t = 0;
  hlbn::ConvGeom geo[4];
  int64_t P[4];
  // activations (bfloat16)
  Buf<uint16_t> xin, a[14], pooled[3], cat[3];
  Buf<float> out, in32, tgt32, lterm, dout;
  Buf<uint8_t> mask8;
  // gradients (binary32) of each activation and of the concat buffers
  Buf<float> g[14], gcat[3], gpool[3], genc[3];
  Buf<uint16_t> dpre16;
  Buf<float> dpre32, part;
  Buf<float> wc;
  int64_t head_chunk = 0;

  ~hlbn_net() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

enum { E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3A, D3B, D2A, D2B, D1A, D1B };

int refresh_weights(hlbn_net* n) {
  for (auto& c : n->conv)
    KERN(hlbn::flip_transpose(n->params.p + c.woff, c.co, c.ci, c.w16.p, c.wt16.p, n->stream));
  return 0;
}

// One 3x3x3 convolution + ReLU layer of the forward pass.
int conv_layer(hlbn_net* n, int L, const uint16_t* x, uint16_t* y) {
  const Conv& c = n->conv[L];
  KERN(hlbn::conv_fwd(n->geo[c.level], x, c.ci, c.w16.p, n->params.p + c.boff, c.co, nullptr,
                      y, c.co, 0, 1, n->part.p, n->stream));
  return 0;
}

// Backward of layer L: dpost = gradient of its (post-ReLU) output; writes
// its parameter gradients and, when dx != nullptr, the input gradient.
int conv_layer_back(hlbn_net* n, int L, const uint16_t* x, const float* dpost, float* dx) {
  const Conv& c = n->conv[L];
  const hlbn::ConvGeom& G = n->geo[c.level];
  const int64_t P = n->P[c.level];
  KERN(hlbn::relu_bwd(dpost, n->a[L].p, P * c.co, n->dpre16.p, n->dpre32.p, n->stream));
  KERN(hlbn::conv_wgrad(G, x, c.ci, n->dpre16.p, c.co, n->part.p, n->grads.p + c.woff,
                        n->stream));
  KERN(hlbn::channel_sum(n->dpre32.p, P, c.co, n->part.p, hlbn::wgrad_chunk(G),
                         n->grads.p + c.boff, n->stream));
  if (dx)
    KERN(hlbn::conv_fwd(G, n->dpre16.p, c.co, c.wt16.p, nullptr, c.ci, dx, nullptr, c.ci, 0, 0,
                        n->part.p, n->stream));
  return 0;
}

int forward_impl(hlbn_net* n) {
  const int C1 = n->C[0], C2 = n->C[1], C3 = n->C[2], C4 = n->C[3];
  if (conv_layer(n, E1A, n->xin.p, n->a[E1A].p) || conv_layer(n, E1B, n->a[E1A].p, n->a[E1B].p))
    return 1;
  KERN(hlbn::pool(n->geo[0], C1, n->a[E1B].p, n->pooled[0].p, n->stream));
  if (conv_layer(n, E2A, n->pooled[0].p, n->a[E2A].p) ||
      conv_layer(n, E2B, n->a[E2A].p, n->a[E2B].p))
    return 1;
  KERN(hlbn::pool(n->geo[1], C2, n->a[E2B].p, n->pooled[1].p, n->stream));
  if (conv_layer(n, E3A, n->pooled[1].p, n->a[E3A].p) ||
      conv_layer(n, E3B, n->a[E3A].p, n->a[E3B].p))
    return 1;
  KERN(hlbn::pool(n->geo[2], C3, n->a[E3B].p, n->pooled[2].p, n->stream));
  if (conv_layer(n, E4A, n->pooled[2].p, n->a[E4A].p) ||
      conv_layer(n, E4B, n->a[E4A].p, n->a[E4B].p))
    return 1;
  // decoder: cat[k] = [up(coarser), skip] at level k
  KERN(hlbn::up(n->geo[3], C4, n->a[E4B].p, n->cat[2].p, C4 + C3, 0, n->stream));
  KERN(hlbn::copy_channels(n->P[2], C3, n->a[E3B].p, n->cat[2].p, C4 + C3, C4, n->stream));
  if (conv_layer(n, D3A, n->cat[2].p, n->a[D3A].p) ||
      conv_layer(n, D3B, n->a[D3A].p, n->a[D3B].p))
    return 1;
  KERN(hlbn::up(n->geo[2], C3, n->a[D3B].p, n->cat[1].p, C3 + C2, 0, n->stream));
  KERN(hlbn::copy_channels(n->P[1], C2, n->a[E2B].p, n->cat[1].p, C3 + C2, C3, n->stream));
  if (conv_layer(n, D2A, n->cat[1].p, n->a[D2A].p) ||
      conv_layer(n, D2B, n->a[D2A].p, n->a[D2B].p))
    return 1;
  KERN(hlbn::up(n->geo[1], C2, n->a[D2B].p, n->cat[0].p, C2 + C1, 0, n->stream));
  KERN(hlbn::copy_channels(n->P[0], C1, n->a[E1B].p, n->cat[0].p, C2 + C1, C2, n->stream));
  if (conv_layer(n, D1A, n->cat[0].p, n->a[D1A].p) ||
      conv_layer(n, D1B, n->a[D1A].p, n->a[D1B].p))
    return 1;
  KERN(hlbn::head_fwd(n->P[0], C1, n->cout, n->a[D1B].p, n->params.p + n->head_w,
                      n->params.p + n->head_b, n->out.p, n->stream));
  return 0;
}

int backward_impl(hlbn_net* n) {
  const int C1 = n->C[0], C2 = n->C[1], C3 = n->C[2], C4 = n->C[3];
  hipStream_t s = n->stream;
  KERN(hlbn::head_bwd(n->P[0], C1, n->cout, n->dout.p, n->a[D1B].p, n->params.p + n->head_w,
                      n->g[D1B].p, n->part.p, n->head_chunk, n->grads.p + n->head_w,
                      n->grads.p + n->head_b, s));
  // decoder level 0
  if (conv_layer_back(n, D1B, n->a[D1A].p, n->g[D1B].p, n->g[D1A].p) ||
      conv_layer_back(n, D1A, n->cat[0].p, n->g[D1A].p, n->gcat[0].p))
    return 1;
  KERN(hlbn::up_bwd(n->geo[1], C2, n->gcat[0].p, C2 + C1, 0, n->g[D2B].p, s));
  KERN(hlbn::take_channels(n->P[0], C1, n->gcat[0].p, C2 + C1, C2, n->genc[0].p, 0, s));
  // decoder level 1
  if (conv_layer_back(n, D2B, n->a[D2A].p, n->g[D2B].p, n->g[D2A].p) ||
      conv_layer_back(n, D2A, n->cat[1].p, n->g[D2A].p, n->gcat[1].p))
    return 1;
  KERN(hlbn::up_bwd(n->geo[2], C3, n->gcat[1].p, C3 + C2, 0, n->g[D3B].p, s));
  KERN(hlbn::take_channels(n->P[1], C2, n->gcat[1].p, C3 + C2, C3, n->genc[1].p, 0, s));
  // decoder level 2
  if (conv_layer_back(n, D3B, n->a[D3A].p, n->g[D3B].p, n->g[D3A].p) ||
      conv_layer_back(n, D3A, n->cat[2].p, n->g[D3A].p, n->gcat[2].p))
    return 1;
  KERN(hlbn::up_bwd(n->geo[3], C4, n->gcat[2].p, C4 + C3, 0, n->g[E4B].p, s));
  KERN(hlbn::take_channels(n->P[2], C3, n->gcat[2].p, C4 + C3, C4, n->genc[2].p, 0, s));
  // bottleneck and encoder
  if (conv_layer_back(n, E4B, n->a[E4A].p, n->g[E4B].p, n->g[E4A].p) ||
      conv_layer_back(n, E4A, n->pooled[2].p, n->g[E4A].p, n->gpool[2].p))
    return 1;
  KERN(hlbn::pool_bwd_add(n->geo[2], C3, n->gpool[2].p, n->genc[2].p, s));
  if (conv_layer_back(n, E3B, n->a[E3A].p, n->genc[2].p, n->g[E3A].p) ||
      conv_layer_back(n, E3A, n->pooled[1].p, n->g[E3A].p, n->gpool[1].p))
    return 1;
  KERN(hlbn::pool_bwd_add(n->geo[1], C2, n->gpool[1].p, n->genc[1].p, s));
  if (conv_layer_back(n, E2B, n->a[E2A].p, n->genc[1].p, n->g[E2A].p) ||
      conv_layer_back(n, E2A, n->pooled[0].p, n->g[E2A].p, n->gpool[0].p))
    return 1;
  KERN(hlbn::pool_bwd_add(n->geo[0], C1, n->gpool[0].p, n->genc[0].p, s));
  if (conv_layer_back(n, E1B, n->a[E1A].p, n->genc[0].p, n->g[E1A].p) ||
      conv_layer_back(n, E1A, n->xin.p, n->g[E1A].p, nullptr))
    return 1;
  return 0;
}

}  // namespace

extern "C" {

const char* hlbn_last_error(void) { return g_err.c_str(); }

int hlbn_create(int32_t side, int32_t batch, int32_t in_channels, int32_t out_channels,
                const int32_t chans[4], int32_t device, hlbn_net** out) {
  if (!out || !chans) return fail("null argument");
  *out = nullptr;
  if (side < 8 || side % 8 || batch < 1 || in_channels < 1 || in_channels > 16 ||
      out_channels < 1 || out_channels > 16)
    return fail("side must be a positive multiple of 8; 1..16 input and output channels");
  for (int k = 0; k < 4; ++k)
    if (chans[k] < 16 || chans[k] % 16) return fail("channel counts must be multiples of 16");
  // the backward-weight kernel indexes activations with 32-bit offsets
  const int64_t widest = std::max<int64_t>({chans[1] + chans[0], (chans[2] + chans[1]) / 8,
                                            (chans[3] + chans[2]) / 64, chans[3] / 512});
  if ((int64_t)batch * side * side * side * widest >= ((int64_t)1 << 31))
    return fail("batch * side^3 * channels must stay below 2^31");
  auto n = std::make_unique<hlbn_net>();
  n->S = side;
  n->N = batch;
  n->cin = in_channels;
  n->cinp = 16;
  n->cout = out_channels;
  for (int k = 0; k < 4; ++k) n->C[k] = chans[k];
  n->device = device;
  const int C1 = chans[0], C2 = chans[1], C3 = chans[2], C4 = chans[3];
  struct Spec { const char* name; int ci, co, level; };
  const Spec spec[14] = {{"e1a", n->cinp, C1, 0}, {"e1b", C1, C1, 0}, {"e2a", C1, C2, 1},
                         {"e2b", C2, C2, 1},      {"e3a", C2, C3, 2}, {"e3b", C3, C3, 2},
                         {"e4a", C3, C4, 3},      {"e4b", C4, C4, 3}, {"d3a", C4 + C3, C3, 2},
                         {"d3b", C3, C3, 2},      {"d2a", C3 + C2, C2, 1}, {"d2b", C2, C2, 1},
                         {"d1a", C2 + C1, C1, 0}, {"d1b", C1, C1, 0}};
  int64_t off = 0;
  for (int L = 0; L < 14; ++L) {
    Conv& c = n->conv[L];
    c.name = spec[L].name;
    // the first layer's weights cover the padded input channels; the
    // parameter vector holds only the real ones (zero padding)
    c.ci = spec[L].ci;
    c.co = spec[L].co;
    c.level = spec[L].level;
    c.woff = off;
    off += (int64_t)27 * c.co * c.ci;
    c.boff = off;
    off += c.co;
  }
  n->head_w = off;
  off += (int64_t)out_channels * C1;
  n->head_b = off;
  off += out_channels;
  n->nparam = off;
  for (int k = 0; k < 4; ++k) {
    const int s = side >> k;
    n->geo[k] = hlbn::ConvGeom{batch, s, s, s};
    n->P[k] = (int64_t)batch * s * s * s;
  }
  HIPN(hipSetDevice(device));
  HIPN(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  for (auto& e : n->ev) HIPN(hipEventCreate(&e));
  HIPN(n->params.alloc(n->nparam));
  HIPN(n->grads.alloc(n->nparam));
  HIPN(n->adam_m.alloc(n->nparam));
  HIPN(n->adam_v.alloc(n->nparam));
  HIPN(hipMemset(n->params.p, 0, sizeof(float) * n->nparam));
  HIPN(hipMemset(n->grads.p, 0, sizeof(float) * n->nparam));
  HIPN(hipMemset(n->adam_m.p, 0, sizeof(float) * n->nparam));
  HIPN(hipMemset(n->adam_v.p, 0, sizeof(float) * n->nparam));
  int64_t max_part = 0, max_pc = 0;
  for (auto& c : n->conv) {
    HIPN(c.w16.alloc((size_t)27 * c.co * c.ci));
    HIPN(c.wt16.alloc((size_t)27 * c.co * c.ci));
    const int64_t P = n->P[c.level];
    const int64_t ch = hlbn::wgrad_chunk(n->geo[c.level]);
    max_part = std::max(max_part, hlbn::wgrad_parts(n->geo[c.level], c.ci, c.co));
    // K-slice workspaces of the forward and backward-data convolutions
    max_part = std::max(max_part, hlbn::conv_fwd_ws(n->geo[c.level], c.ci, c.co));
    max_part = std::max(max_part, hlbn::conv_fwd_ws(n->geo[c.level], c.co, c.ci));
    max_part = std::max(max_part, (P + ch - 1) / ch * c.co);
    max_pc = std::max(max_pc, P * c.co);
  }
  // positions per partial sum of the head's weight gradient and of the
  // per-channel sums at full resolution: 16384 chunks of 64 at a 64^3 x 4
  // batch, one thread each (1024 chunks of 1024 took 175 us)
  n->head_chunk = std::max<int64_t>(64, (n->P[0] + 16383) / 16384);
  max_part = std::max(max_part, (n->P[0] + n->head_chunk - 1) / n->head_chunk *
                                    (int64_t)out_channels * C1);
  HIPN(n->part.alloc(max_part));
  HIPN(n->dpre16.alloc(max_pc));
  HIPN(n->dpre32.alloc(max_pc));
  const int64_t P0 = n->P[0];
  HIPN(n->xin.alloc(P0 * n->cinp));
  HIPN(n->in32.alloc(P0 * in_channels));
  for (int L = 0; L < 14; ++L) {
    HIPN(n->a[L].alloc(n->P[n->conv[L].level] * n->conv[L].co));
    HIPN(n->g[L].alloc(n->P[n->conv[L].level] * n->conv[L].co));
  }
  for (int k = 0; k < 3; ++k) {
    HIPN(n->pooled[k].alloc(n->P[k + 1] * chans[k]));
    HIPN(n->gpool[k].alloc(n->P[k + 1] * chans[k]));
    HIPN(n->cat[k].alloc(n->P[k] * (chans[k + 1] + chans[k])));
    HIPN(n->gcat[k].alloc(n->P[k] * (chans[k + 1] + chans[k])));
    HIPN(n->genc[k].alloc(n->P[k] * chans[k]));
  }
  HIPN(n->out.alloc(P0 * out_channels));
  HIPN(n->dout.alloc(P0 * out_channels));
  HIPN(n->lterm.alloc(P0 * out_channels));
  HIPN(n->tgt32.alloc(P0 * out_channels));
  HIPN(n->mask8.alloc(P0));
  HIPN(n->wc.alloc(out_channels));
  if (refresh_weights(n.get())) return 1;
  HIPN(hipStreamSynchronize(n->stream));
  *out = n.release();
  return 0;
}

void hlbn_destroy(hlbn_net* net) { delete net; }

int64_t hlbn_param_count(const hlbn_net* n) { return n ? n->nparam : 0; }

int64_t hlbn_device_bytes(const hlbn_net* n) {
  if (!n) return 0;
  int64_t b = 4 * 4 * n->nparam;
  for (auto& c : n->conv) b += 2 * 2 * (int64_t)c.w16.n;
  b += 2 * (int64_t)n->xin.n + 4 * (int64_t)n->in32.n;
  for (int L = 0; L < 14; ++L) b += 2 * (int64_t)n->a[L].n + 4 * (int64_t)n->g[L].n;
  for (int k = 0; k < 3; ++k)
    b += 2 * (int64_t)(n->pooled[k].n + n->cat[k].n) +
         4 * (int64_t)(n->gpool[k].n + n->gcat[k].n + n->genc[k].n);
  b += 4 * (int64_t)(n->out.n + n->dout.n + n->lterm.n + n->tgt32.n + n->part.n + n->dpre32.n);
  b += 2 * (int64_t)n->dpre16.n + (int64_t)n->mask8.n;
  return b;
}

int hlbn_set_params(hlbn_net* n, const float* params) {
  if (!n || !params) return fail("null argument");
  HIPN(hipMemcpy(n->params.p, params, sizeof(float) * n->nparam, hipMemcpyHostToDevice));
  n->adam_t = 0;
  HIPN(hipMemset(n->adam_m.p, 0, sizeof(float) * n->nparam));
  HIPN(hipMemset(n->adam_v.p, 0, sizeof(float) * n->nparam));
  if (refresh_weights(n)) return 1;
  HIPN(hipStreamSynchronize(n->stream));
  return 0;
}

int hlbn_get_params(hlbn_net* n, float* params) {
  if (!n || !params) return fail("null argument");
  HIPN(hipStreamSynchronize(n->stream));
  HIPN(hipMemcpy(params, n->params.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
  return 0;
}

int hlbn_get_grads(hlbn_net* n, float* grads) {
  if (!n || !grads) return fail("null argument");
  HIPN(hipStreamSynchronize(n->stream));
  HIPN(hipMemcpy(grads, n->grads.p, sizeof(float) * n->nparam, hipMemcpyDeviceToHost));
  return 0;
}

int hlbn_forward(hlbn_net* n, const float* input, float* output, int32_t device) {
  if (!n || !input) return fail("null argument");
  const int64_t P0 = n->P[0];
  const float* in = input;
  if (!device) {
    HIPN(hipMemcpyAsync(n->in32.p, input, sizeof(float) * P0 * n->cin, hipMemcpyHostToDevice,
                        n->stream));
    in = n->in32.p;
  }
  HIPN(hipEventRecord(n->ev[0], n->stream));
  KERN(hlbn::pack_input(P0, n->cin, n->cinp, in, n->xin.p, n->stream));
  if (forward_impl(n)) return 1;
  HIPN(hipEventRecord(n->ev[1], n->stream));
  if (output) {
    HIPN(hipMemcpyAsync(output, n->out.p, sizeof(float) * P0 * n->cout,
                        device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, n->stream));
  }
  HIPN(hipStreamSynchronize(n->stream));
  HIPN(hipEventElapsedTime(&n->fwd_ms, n->ev[0], n->ev[1]));
  return 0;
}

int hlbn_backward(hlbn_net* n, const float* target, const uint8_t* mask,
                  const float* channel_weights, double mask_count, double* loss,
                  int32_t device) {
  if (!n || !target || !mask || !channel_weights) return fail("null argument");
  const int64_t P0 = n->P[0];
  const float* t = target;
  const uint8_t* m = mask;
  if (!device) {
    HIPN(hipMemcpyAsync(n->tgt32.p, target, sizeof(float) * P0 * n->cout,
                        hipMemcpyHostToDevice, n->stream));
    HIPN(hipMemcpyAsync(n->mask8.p, mask, P0, hipMemcpyHostToDevice, n->stream));
    t = n->tgt32.p;
    m = n->mask8.p;
  }
  HIPN(hipMemcpyAsync(n->wc.p, channel_weights, sizeof(float) * n->cout, hipMemcpyHostToDevice,
                      n->stream));
  HIPN(hipEventRecord(n->ev[1], n->stream));
  const float inv = (float)(1.0 / (std::max(mask_count, 1.0) * n->cout));
  KERN(hlbn::mse(P0, n->cout, n->out.p, t, m, n->wc.p, inv, n->dout.p, n->lterm.p, n->stream));
  if (backward_impl(n)) return 1;
  HIPN(hipEventRecord(n->ev[2], n->stream));
  if (loss) {
    // loss terms summed per channel in a fixed order, then on the host
    Buf<float> sums;
    HIPN(sums.alloc(n->cout));
    KERN(hlbn::channel_sum(n->lterm.p, P0, n->cout, n->part.p, n->head_chunk, sums.p,
                           n->stream));
    std::vector<float> h(n->cout);
    HIPN(hipMemcpyAsync(h.data(), sums.p, sizeof(float) * n->cout, hipMemcpyDeviceToHost,
                        n->stream));
    HIPN(hipStreamSynchronize(n->stream));
    double l = 0;
    for (float v : h) l += v;
    *loss = l;
  }
  HIPN(hipStreamSynchronize(n->stream));
  HIPN(hipEventElapsedTime(&n->bwd_ms, n->ev[1], n->ev[2]));
  return 0;
}

int hlbn_adam(hlbn_net* n, float lr, float b1, float b2, float eps, float wd) {
  if (!n) return fail("null argument");
  n->adam_t += 1;
  const float c1 = 1.0f - (float)std::pow((double)b1, (double)n->adam_t);
  const float c2 = 1.0f - (float)std::pow((double)b2, (double)n->adam_t);
  KERN(hlbn::adam(n->nparam, n->params.p, n->grads.p, n->adam_m.p, n->adam_v.p, lr, b1, b2, eps,
                  wd, c1, c2, n->stream));
  if (refresh_weights(n)) return 1;
  HIPN(hipStreamSynchronize(n->stream));
  return 0;
}

int hlbn_last_times(const hlbn_net* n, double* fwd, double* bwd) {
  if (!n) return fail("null argument");
  if (fwd) *fwd = n->fwd_ms;
  if (bwd) *bwd = n->bwd_ms;
  return 0;
}

void* hlbn_stream(const hlbn_net* n) { return n ? (void*)n->stream : nullptr; }

}  // extern "C"
