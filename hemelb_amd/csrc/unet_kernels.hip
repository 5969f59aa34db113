// HIP kernels of the warm-start 3D U-Net (gfx950): 3x3x3 convolutions as
// implicit GEMMs on bfloat16 MFMA with binary32 accumulation (forward,
// backward-data through the same kernel with flipped, transposed weights,
// backward-weight with fixed-order partial sums), and the elementwise
// pieces of training: ReLU, 2x2x2 mean pooling, 2x nearest upsampling,
// channel concatenation, masked MSE, bias and head reductions, Adam.
//
// Tensors are channels-last: element (p, c) of a [P][C] tensor, where p
// enumerates (n, d, h, w) with w fastest.  bfloat16 values are stored as
// uint16 bit patterns.  Convolution weights are [27][Cout][Cin], tap
// t = 9*kd + 3*kh + kw for offsets (kd-1, kh-1, kw-1).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "unet_internal.h"

namespace hlbn {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}
// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct Grid3 {
  int N, D, H, W;
  __host__ __device__ __forceinline__ int64_t P() const { return (int64_t)N * D * H * W; }
};

// Coordinates (w, h, d) of position p inside its sample (P < 2^31).
struct Coord3 {
  int w, h, d;
};
__device__ __forceinline__ Coord3 coord_of(const Grid3& g, int p) {
  const int r1 = p / g.W, r2 = r1 / g.H;
  return Coord3{p - r1 * g.W, r1 - r2 * g.H, r2 % g.D};
}
// The coordinates of position p + step from those of p (no division).
__device__ __forceinline__ void advance(const Grid3& g, Coord3& c, int step) {
  c.w += step;
  while (c.w >= g.W) {
    c.w -= g.W;
    if (++c.h == g.H) {
      c.h = 0;
      if (++c.d == g.D) c.d = 0;
    }
  }
}

// --------------------------------------------------------------------------
// Forward 3x3x3 convolution (also backward-data).  One wave computes 16
// consecutive positions x CT tiles of 16 output channels with
// v_mfma_f32_16x16x16_bf16: A = input patches (16 positions x 16 channels of
// one tap), B = weights (16 channels x 16 outputs).  Lane l supplies A row
// l%16 and B column l%16, K elements 4*(l/16)..+3; it receives D rows
// 4*(l/16)..+3 of column l%16.
// --------------------------------------------------------------------------
template <int CT>
__global__ __launch_bounds__(256) void k_conv_fwd(Grid3 g, const uint16_t* __restrict__ x,
                                                  int ci, const uint16_t* __restrict__ w,
                                                  const float* __restrict__ bias, int co,
                                                  int tiles_per_wave,
                                                  float* __restrict__ y32,
                                                  uint16_t* __restrict__ y16, int yc,
                                                  int yoff, int relu) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int ngroups = (co / 16 + tiles_per_wave - 1) / tiles_per_wave;
  const int64_t pgroup = wave / ngroups;
  const int tile0 = (int)(wave % ngroups) * tiles_per_wave;
  const int64_t P = g.P();
  const int64_t p0 = pgroup * 16;
  if (p0 >= P) return;
  const int row = lane & 15, kq = lane >> 4;
  const int64_t p = p0 + row;
  const bool pvalid = p < P;
  float4_t acc[CT];
#pragma unroll
  for (int j = 0; j < CT; ++j) acc[j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const int ntile = min(CT, co / 16 - tile0);
  // the lane's position decomposed once; each tap checks its neighbour
  // against the volume (zero padding outside)
  const Coord3 pc = pvalid ? coord_of(g, (int)p) : Coord3{0, 0, 0};
  for (int t = 0; t < 27; ++t) {
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const bool inside = pvalid && (unsigned)(pc.d + dz) < (unsigned)g.D &&
                        (unsigned)(pc.h + dy) < (unsigned)g.H &&
                        (unsigned)(pc.w + dx) < (unsigned)g.W;
    const int64_t q = inside ? p + ((int64_t)dz * g.H + dy) * g.W + dx : -1;
    for (int c0 = 0; c0 < ci; c0 += 16) {
      short4_t a = short4_t{0, 0, 0, 0};
      if (q >= 0) a = *reinterpret_cast<const short4_t*>(x + q * ci + c0 + 4 * kq);
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        if (j < ntile) {
          const int oc = (tile0 + j) * 16 + row;
          const short4_t b =
              *reinterpret_cast<const short4_t*>(w + ((int64_t)t * co + oc) * ci + c0 + 4 * kq);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CT; ++j) {
    if (j >= ntile) continue;
    const int oc = (tile0 + j) * 16 + row;
    const float bv = bias ? bias[oc] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t pp = p0 + 4 * kq + r;
      if (pp >= P) continue;
      float v = acc[j][r] + bv;
      if (relu) v = v > 0.f ? v : 0.f;
      const int64_t o = pp * yc + yoff + oc;
      if (y32) y32[o] = v;
      if (y16) y16[o] = f2bf(v);
    }
  }
}

// --------------------------------------------------------------------------
// Backward-weight: dW[t][co][ci] = sum_p dy[p][co] * x[p + off_t][ci].  Per
// wave: one chunk of positions, TG taps (one kd plane: kh, kw in 0..2), one
// 16 x 16 (co x ci) tile; MFMA with A = dy^T (co x 16 positions), loaded once
// per K step for all TG taps, and B = the shifted input (16 positions x ci).
// Partial sums of chunk k go to part[k][27][co][ci]; a second kernel adds the
// chunks in order.
// --------------------------------------------------------------------------
template <int TG>
__global__ __launch_bounds__(256) void k_conv_wgrad(Grid3 g, const uint16_t* __restrict__ x,
                                                    int ci, const uint16_t* __restrict__ dy,
                                                    int co, int64_t chunk64,
                                                    float* __restrict__ part) {
  // 32-bit positions and element offsets (the host keeps P * channels < 2^31)
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int mt = co / 16, nt = ci / 16, ngr = 27 / TG;
  const int per_chunk = ngr * mt * nt;
  const int k = wave / per_chunk;
  const int P = (int)g.P();
  const int chunk = (int)chunk64;
  const int nchunks = (P + chunk - 1) / chunk;
  if (k >= nchunks) return;
  int rem = wave % per_chunk;
  const int tg = rem / (mt * nt);
  rem %= (mt * nt);
  const int m0 = rem / nt, n0 = rem % nt;
  const int row = lane & 15, kq = lane >> 4;
  const int oc = m0 * 16 + row, ic = n0 * 16 + row;
  float4_t acc[TG];
#pragma unroll
  for (int i = 0; i < TG; ++i) acc[i] = float4_t{0.f, 0.f, 0.f, 0.f};
  const int pbeg = k * chunk, pend = min(P, pbeg + chunk);
  const int kd = tg * TG / 9 - 1;  // TG = 9: one kd plane per tap group
  // Coordinates of the lane's four positions, advanced by 16 per step
  // instead of divided out again.
  Coord3 pc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) pc[j] = coord_of(g, min(pbeg + 4 * kq + j, P - 1));
  for (int pk = pbeg; pk < pend; pk += 16) {
    short4_t a = short4_t{0, 0, 0, 0};
    int pos[4];
    unsigned okd = 0;  // bit j: position j valid and its kd-plane neighbour inside
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      pos[j] = pk + 4 * kq + j;
      if (pos[j] < pend) {
        a[j] = (short)dy[pos[j] * co + oc];
        if ((unsigned)(pc[j].d + kd) < (unsigned)g.D) okd |= 1u << j;
      }
    }
#pragma unroll
    for (int tt = 0; tt < TG; ++tt) {
      const int dyy = tt / 3 - 1, dx = tt % 3 - 1;
      const int off = (kd * g.H + dyy) * g.W + dx;
      short4_t b = short4_t{0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if ((okd >> j) & 1u) {
          if ((unsigned)(pc[j].h + dyy) < (unsigned)g.H && (unsigned)(pc[j].w + dx) < (unsigned)g.W)
            b[j] = (short)x[(pos[j] + off) * ci + ic];
        }
      }
      acc[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tt], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) advance(g, pc[j], 16);
  }
#pragma unroll
  for (int tt = 0; tt < TG; ++tt) {
    const int t = tg * TG + tt;
    float* out = part + ((int64_t)(k * 27 + t) * co) * ci;
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(m0 * 16 + 4 * kq + r) * ci + ic] = acc[tt][r];
  }
}

// out[i] = sum over k of part[k][i] in a fixed order (deterministic): a
// 256-thread block covers kOut consecutive outputs; each of its 256 / kOut
// phases adds the chunks k = phase, phase + phases, ... in order, and the
// phase sums are then added in phase order.
template <int kOut>
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part,
                                                    int64_t nchunks, int64_t n,
                                                    float* __restrict__ out) {
  constexpr int kPhases = 256 / kOut;
  __shared__ float sh[256];
  const int j = (int)threadIdx.x % kOut, q = (int)threadIdx.x / kOut;
  const int64_t i = (int64_t)blockIdx.x * kOut + j;
  float s = 0.f;
  if (i < n)
    for (int64_t k = q; k < nchunks; k += kPhases) s += part[k * n + i];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (q == 0 && i < n) {
    float t = sh[j];
    for (int r = 1; r < kPhases; ++r) t += sh[r * kOut + j];
    out[i] = t;
  }
}

// part[k][c] = sum of a[p][c] over the positions p of chunk k, for C
// dividing 256: one block per chunk reads the chunk's [positions][C] slab
// with consecutive threads on consecutive elements (thread t always on
// channel t % C), then adds the threads of each channel in a fixed tree.
__global__ __launch_bounds__(256) void k_channel_sum_block(const float* __restrict__ a,
                                                           int64_t P, int C, int64_t chunk,
                                                           float* __restrict__ part) {
  __shared__ float sh[256];
  const int64_t k = blockIdx.x;
  const int64_t e0 = k * chunk * C, e1 = min(P, (k + 1) * chunk) * C;
  float s = 0.f;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) s += a[e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int stride = 128; stride >= C; stride >>= 1) {
    if ((int)threadIdx.x < stride) sh[threadIdx.x] += sh[threadIdx.x + stride];
    __syncthreads();
  }
  if ((int)threadIdx.x < C) part[k * C + threadIdx.x] = sh[threadIdx.x];
}

// Per-channel sums for any C: one thread per (chunk, channel).
__global__ void k_channel_sum_part(const float* __restrict__ a, int64_t P, int C,
                                   int64_t chunk, float* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nchunks = (P + chunk - 1) / chunk;
  if (i >= nchunks * C) return;
  const int64_t k = i / C;
  const int c = (int)(i % C);
  const int64_t end = min(P, (k + 1) * chunk);
  float s = 0.f;
  for (int64_t p = k * chunk; p < end; ++p) s += a[p * C + c];
  part[k * C + c] = s;
}

// Weights for backward-data: wt[26 - t][ci][co] = w[t][co][ci] (bf16).
__global__ void k_flip_transpose(const float* __restrict__ w, int co, int ci,
                                 uint16_t* __restrict__ w16, uint16_t* __restrict__ wt16) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)27 * co * ci;
  if (i >= n) return;
  const int t = (int)(i / ((int64_t)co * ci));
  const int o = (int)((i / ci) % co), c = (int)(i % ci);
  const uint16_t b = f2bf(w[i]);
  w16[i] = b;
  wt16[((int64_t)(26 - t) * ci + c) * co + o] = b;
}

// dpre_bf16[i] = dpost[i] if post[i] > 0 else 0  (ReLU backward)
__global__ void k_relu_bwd(const float* __restrict__ dpost, const uint16_t* __restrict__ post,
                           int64_t n, uint16_t* __restrict__ dpre16,
                           float* __restrict__ dpre32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = bf2f(post[i]) > 0.f ? dpost[i] : 0.f;
  dpre16[i] = f2bf(v);
  if (dpre32) dpre32[i] = v;
}

// 2x2x2 mean pooling of bf16 [N,D,H,W,C] -> bf16 [N,D/2,H/2,W/2,C].
__global__ void k_pool(Grid3 g, int C, const uint16_t* __restrict__ x,
                       uint16_t* __restrict__ y) {
  const Grid3 h{g.N, g.D / 2, g.H / 2, g.W / 2};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= h.P() * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % h.W), hh = (int)((q / h.W) % h.H), d = (int)((q / h.W / h.H) % h.D);
  const int64_t n = q / ((int64_t)h.W * h.H * h.D);
  float s = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int e = 0; e < 2; ++e) {
        const int64_t p = ((n * g.D + 2 * d + a) * g.H + 2 * hh + b) * g.W + 2 * w + e;
        s += bf2f(x[p * C + c]);
      }
  y[i] = f2bf(s * 0.125f);
}

// dx[p] += dy[parent(p)] / 8 (float), x at full resolution.
__global__ void k_pool_bwd_add(Grid3 g, int C, const float* __restrict__ dy,
                               float* __restrict__ dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.P() * C) return;
  const int c = (int)(i % C);
  const int64_t p = i / C;
  const int w = (int)(p % g.W), h = (int)((p / g.W) % g.H), d = (int)((p / g.W / g.H) % g.D);
  const int64_t n = p / ((int64_t)g.W * g.H * g.D);
  const int64_t q = ((n * (g.D / 2) + d / 2) * (g.H / 2) + h / 2) * (g.W / 2) + w / 2;
  dx[i] += dy[q * C + c] * 0.125f;
}

// Nearest 2x upsampling of bf16 [N,D,H,W,C] into channels [off, off+C) of
// the bf16 tensor y [N,2D,2H,2W,yc].
__global__ void k_up(Grid3 g, int C, const uint16_t* __restrict__ x, uint16_t* __restrict__ y,
                     int yc, int yoff) {
  const Grid3 f{g.N, 2 * g.D, 2 * g.H, 2 * g.W};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.P() * C) return;
  const int c = (int)(i % C);
  const int64_t p = i / C;
  const int w = (int)(p % f.W), h = (int)((p / f.W) % f.H), d = (int)((p / f.W / f.H) % f.D);
  const int64_t n = p / ((int64_t)f.W * f.H * f.D);
  const int64_t q = ((n * g.D + d / 2) * g.H + h / 2) * g.W + w / 2;
  y[p * yc + yoff + c] = x[q * C + c];
}

// dx[q][c] = sum of the 8 children of q in channels [off, off+C) of dy.
__global__ void k_up_bwd(Grid3 g, int C, const float* __restrict__ dy, int dyc, int dyoff,
                         float* __restrict__ dx) {
  const Grid3 f{g.N, 2 * g.D, 2 * g.H, 2 * g.W};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.P() * C) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % g.W), h = (int)((q / g.W) % g.H), d = (int)((q / g.W / g.H) % g.D);
  const int64_t n = q / ((int64_t)g.W * g.H * g.D);
  float s = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int e = 0; e < 2; ++e) {
        const int64_t p = ((n * f.D + 2 * d + a) * f.H + 2 * h + b) * f.W + 2 * w + e;
        s += dy[p * dyc + dyoff + c];
      }
  dx[i] = s;
}

// Copies bf16 channels: y[p][yoff + c] = x[p][c], c < C.
__global__ void k_copy_channels(int64_t P, int C, const uint16_t* __restrict__ x,
                                uint16_t* __restrict__ y, int yc, int yoff) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  y[(i / C) * yc + yoff + (i % C)] = x[i];
}

// dx[p][c] (+)= dy[p][dyoff + c], c < C.
__global__ void k_take_channels(int64_t P, int C, const float* __restrict__ dy, int dyc,
                                int dyoff, float* __restrict__ dx, int add) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * C) return;
  const float v = dy[(i / C) * dyc + dyoff + (i % C)];
  dx[i] = add ? dx[i] + v : v;
}

// Input packing: float [P][cin] -> bf16 [P][cpad] with zero padding.
__global__ void k_pack_input(int64_t P, int cin, int cpad, const float* __restrict__ in,
                             uint16_t* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * cpad) return;
  const int c = (int)(i % cpad);
  x[i] = c < cin ? f2bf(in[(i / cpad) * cin + c]) : (uint16_t)0;
}

// 1x1x1 head: out[p][o] = sum_k h[p][k] W[o][k] + b[o]   (binary32)
__global__ void k_head_fwd(int64_t P, int K, int O, const uint16_t* __restrict__ h,
                           const float* __restrict__ W, const float* __restrict__ b,
                           float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * O) return;
  const int o = (int)(i % O);
  const int64_t p = i / O;
  float s = b[o];
  for (int k = 0; k < K; ++k) s += bf2f(h[p * K + k]) * W[o * K + k];
  out[i] = s;
}

// dh[p][k] = sum_o dout[p][o] W[o][k]
__global__ void k_head_bwd_data(int64_t P, int K, int O, const float* __restrict__ dout,
                                const float* __restrict__ W, float* __restrict__ dh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * K) return;
  const int k = (int)(i % K);
  const int64_t p = i / K;
  float s = 0.f;
  for (int o = 0; o < O; ++o) s += dout[p * O + o] * W[o * K + k];
  dh[i] = s;
}

// part[chunk][o*K + k] = sum over the chunk of dout[p][o] h[p][k]
__global__ void k_head_wgrad_part(int64_t P, int K, int O, const float* __restrict__ dout,
                                  const uint16_t* __restrict__ h, int64_t chunk,
                                  float* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nchunks = (P + chunk - 1) / chunk;
  if (i >= nchunks * O * K) return;
  const int64_t c = i / ((int64_t)O * K);
  const int ok = (int)(i % ((int64_t)O * K));
  const int o = ok / K, k = ok % K;
  const int64_t end = min(P, (c + 1) * chunk);
  float s = 0.f;
  for (int64_t p = c * chunk; p < end; ++p) s += dout[p * O + o] * bf2f(h[p * K + k]);
  part[i] = s;
}

// Masked weighted MSE: per-position loss terms and the gradient
// dout = 2 w_c (out - t) m / (count * O).
__global__ void k_mse(int64_t P, int O, const float* __restrict__ out,
                      const float* __restrict__ target, const uint8_t* __restrict__ mask,
                      const float* __restrict__ wc, float inv_count,
                      float* __restrict__ dout, float* __restrict__ lterm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * O) return;
  const int o = (int)(i % O);
  const float m = mask[i / O] ? 1.f : 0.f;
  const float d = (out[i] - target[i]) * m;
  dout[i] = 2.f * wc[o] * d * inv_count;
  lterm[i] = wc[o] * d * d * inv_count;
}

// Adam (decoupled weight decay) on binary32 master parameters.
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                       float b2, float eps, float wd, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

inline dim3 blocks_for(int64_t n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

// out[i] = sum over k of part[k * n + i] (fixed order): 16 outputs per
// block in 16 chunk phases.
void sum_chunks(const float* part, int64_t nchunks, int64_t n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_chunks<16>, blocks_for(n, 16), dim3(256), 0, s, part, nchunks, n, out);
}

}  // namespace

int conv_fwd(const ConvGeom& cg, const uint16_t* x, int ci, const uint16_t* w,
             const float* bias, int co, float* y32, uint16_t* y16, int yc, int yoff, int relu,
             void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  const int64_t P = (int64_t)cg.N * cg.D * cg.H * cg.W;
  const int tiles = co / 16;
  int tpw = tiles >= 8 ? 8 : tiles;  // tiles per wave
  // Coarse levels have few positions: fewer tiles per wave, so that the grid
  // still has about four waves per SIMD (the input patches are then loaded
  // once per tile group).
  while (tpw > 1 && ((P + 15) / 16) * ((tiles + tpw - 1) / tpw) < 4096) tpw = (tpw + 1) / 2;
  const int ngroups = (tiles + tpw - 1) / tpw;
  const int64_t waves = ((P + 15) / 16) * ngroups;
  const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
  hipStream_t s = (hipStream_t)stream;
  switch (tpw) {
    case 1: hipLaunchKernelGGL(k_conv_fwd<1>, grid, block, 0, s, g, x, ci, w, bias, co, tpw, y32, y16, yc, yoff, relu); break;
    case 2: hipLaunchKernelGGL(k_conv_fwd<2>, grid, block, 0, s, g, x, ci, w, bias, co, tpw, y32, y16, yc, yoff, relu); break;
    case 3: hipLaunchKernelGGL(k_conv_fwd<4>, grid, block, 0, s, g, x, ci, w, bias, co, tpw, y32, y16, yc, yoff, relu); break;
    case 4: hipLaunchKernelGGL(k_conv_fwd<4>, grid, block, 0, s, g, x, ci, w, bias, co, tpw, y32, y16, yc, yoff, relu); break;
    default: hipLaunchKernelGGL(k_conv_fwd<8>, grid, block, 0, s, g, x, ci, w, bias, co, tpw, y32, y16, yc, yoff, relu); break;
  }
  return (int)hipGetLastError();
}

int64_t wgrad_chunk(const ConvGeom& cg) {
  // up to 2048 chunks: with one wave per (chunk, tap plane, 16x16 tile) the
  // finest level still has thousands of waves
  const int64_t P = (int64_t)cg.N * cg.D * cg.H * cg.W;
  int64_t chunk = (P + 2047) / 2048;
  chunk = (chunk + 15) / 16 * 16;
  return chunk < 256 ? 256 : chunk;
}

int conv_wgrad(const ConvGeom& cg, const uint16_t* x, int ci, const uint16_t* dy, int co,
               float* part, float* dw, void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  const int64_t P = g.P();
  const int64_t chunk = wgrad_chunk(cg);
  const int64_t nchunks = (P + chunk - 1) / chunk;
  constexpr int TG = 9;
  const int64_t waves = nchunks * (27 / TG) * (co / 16) * (ci / 16);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL((k_conv_wgrad<TG>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, g,
                     x, ci, dy, co, chunk, part);
  const int64_t n = (int64_t)27 * co * ci;
  sum_chunks(part, nchunks, n, dw, s);
  return (int)hipGetLastError();
}

int channel_sum(const float* a, int64_t P, int C, float* part, int64_t chunk, float* out,
                void* stream) {
  const int64_t nchunks = (P + chunk - 1) / chunk;
  hipStream_t s = (hipStream_t)stream;
  if (C > 0 && 256 % C == 0)
    hipLaunchKernelGGL(k_channel_sum_block, dim3((unsigned)nchunks), dim3(256), 0, s, a, P, C,
                       chunk, part);
  else
    hipLaunchKernelGGL(k_channel_sum_part, blocks_for(nchunks * C), dim3(256), 0, s, a, P, C,
                       chunk, part);
  sum_chunks(part, nchunks, (int64_t)C, out, s);
  return (int)hipGetLastError();
}

int flip_transpose(const float* w, int co, int ci, uint16_t* w16, uint16_t* wt16, void* stream) {
  hipLaunchKernelGGL(k_flip_transpose, blocks_for((int64_t)27 * co * ci), dim3(256), 0,
                     (hipStream_t)stream, w, co, ci, w16, wt16);
  return (int)hipGetLastError();
}

int relu_bwd(const float* dpost, const uint16_t* post, int64_t n, uint16_t* dpre16,
             float* dpre32, void* stream) {
  hipLaunchKernelGGL(k_relu_bwd, blocks_for(n), dim3(256), 0, (hipStream_t)stream, dpost, post,
                     n, dpre16, dpre32);
  return (int)hipGetLastError();
}

int pool(const ConvGeom& cg, int C, const uint16_t* x, uint16_t* y, void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  hipLaunchKernelGGL(k_pool, blocks_for(g.P() / 8 * C), dim3(256), 0, (hipStream_t)stream, g, C,
                     x, y);
  return (int)hipGetLastError();
}

int pool_bwd_add(const ConvGeom& cg, int C, const float* dy, float* dx, void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  hipLaunchKernelGGL(k_pool_bwd_add, blocks_for(g.P() * C), dim3(256), 0, (hipStream_t)stream,
                     g, C, dy, dx);
  return (int)hipGetLastError();
}

int up(const ConvGeom& coarse, int C, const uint16_t* x, uint16_t* y, int yc, int yoff,
       void* stream) {
  const Grid3 g{coarse.N, coarse.D, coarse.H, coarse.W};
  hipLaunchKernelGGL(k_up, blocks_for(g.P() * 8 * C), dim3(256), 0, (hipStream_t)stream, g, C, x,
                     y, yc, yoff);
  return (int)hipGetLastError();
}

int up_bwd(const ConvGeom& coarse, int C, const float* dy, int dyc, int dyoff, float* dx,
           void* stream) {
  const Grid3 g{coarse.N, coarse.D, coarse.H, coarse.W};
  hipLaunchKernelGGL(k_up_bwd, blocks_for(g.P() * C), dim3(256), 0, (hipStream_t)stream, g, C,
                     dy, dyc, dyoff, dx);
  return (int)hipGetLastError();
}

int copy_channels(int64_t P, int C, const uint16_t* x, uint16_t* y, int yc, int yoff,
                  void* stream) {
  hipLaunchKernelGGL(k_copy_channels, blocks_for(P * C), dim3(256), 0, (hipStream_t)stream, P, C,
                     x, y, yc, yoff);
  return (int)hipGetLastError();
}

int take_channels(int64_t P, int C, const float* dy, int dyc, int dyoff, float* dx, int add,
                  void* stream) {
  hipLaunchKernelGGL(k_take_channels, blocks_for(P * C), dim3(256), 0, (hipStream_t)stream, P, C,
                     dy, dyc, dyoff, dx, add);
  return (int)hipGetLastError();
}

int pack_input(int64_t P, int cin, int cpad, const float* in, uint16_t* x, void* stream) {
  hipLaunchKernelGGL(k_pack_input, blocks_for(P * cpad), dim3(256), 0, (hipStream_t)stream, P,
                     cin, cpad, in, x);
  return (int)hipGetLastError();
}

int head_fwd(int64_t P, int K, int O, const uint16_t* h, const float* W, const float* b,
             float* out, void* stream) {
  hipLaunchKernelGGL(k_head_fwd, blocks_for(P * O), dim3(256), 0, (hipStream_t)stream, P, K, O,
                     h, W, b, out);
  return (int)hipGetLastError();
}

int head_bwd(int64_t P, int K, int O, const float* dout, const uint16_t* h, const float* W,
             float* dh, float* part, int64_t chunk, float* dW, float* db, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_head_bwd_data, blocks_for(P * K), dim3(256), 0, s, P, K, O, dout, W, dh);
  const int64_t nchunks = (P + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_head_wgrad_part, blocks_for(nchunks * O * K), dim3(256), 0, s, P, K, O,
                     dout, h, chunk, part);
  sum_chunks(part, nchunks, (int64_t)O * K, dW, s);
  return channel_sum(dout, P, O, part, chunk, db, stream);
}

int mse(int64_t P, int O, const float* out, const float* target, const uint8_t* mask,
        const float* wc, float inv_count, float* dout, float* lterm, void* stream) {
  hipLaunchKernelGGL(k_mse, blocks_for(P * O), dim3(256), 0, (hipStream_t)stream, P, O, out,
                     target, mask, wc, inv_count, dout, lterm);
  return (int)hipGetLastError();
}

int adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
         float eps, float wd, float c1, float c2, void* stream) {
  hipLaunchKernelGGL(k_adam, blocks_for(n), dim3(256), 0, (hipStream_t)stream, n, p, g, m, v, lr,
                     b1, b2, eps, wd, c1, c2);
  return (int)hipGetLastError();
}

}  // namespace hlbn
