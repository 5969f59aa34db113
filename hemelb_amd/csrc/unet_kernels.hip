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

#include <algorithm>
#include <cstdint>

#include "unet_internal.h"

namespace hlbn {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

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
// Forward 3x3x3 convolution (also backward-data): an implicit GEMM on
// gfx950's v_mfma_f32_16x16x32_bf16 with M = positions, N = output channels
// and K = 27 taps x input channels.  One wave computes MT tiles of 16
// consecutive positions x NT tiles of 16 output channels.  K advances in
// steps of 32: two half-steps of 16 (tap, 16-channel group) pairs, tap-major,
// the 28th tap of an odd count padded with zeros.  Lane l supplies row l%16
// of A (its position) and column l%16 of B (its output channel) with the 8
// K-values 8*(l/16)..+7, each one 16-byte load of 8 consecutive channels
// (channels-last activations, [27][Cout][Cin] weights); it receives rows
// 4*(l/16)..+3 of column l%16.  An A fragment serves NT MFMAs, a B fragment
// MT.  Coarse levels, with few positions and long K, split K over
// gridDim.y slices of `hpw` half-steps: each slice writes its raw sums to
// ws[slice][position][co] and k_conv_reduce adds them in slice order.
// --------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8_t load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(p));
}

template <int MT, int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(Grid3 g, const uint16_t* __restrict__ x,
                                                  int ci, const uint16_t* __restrict__ w,
                                                  const float* __restrict__ bias, int co,
                                                  float* __restrict__ y32,
                                                  uint16_t* __restrict__ y16, int yc,
                                                  int yoff, int relu, int hpw,
                                                  float* __restrict__ ws) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int ngroups = (co / 16 + NT - 1) / NT;
  const int64_t pgroup = wave / ngroups;
  const int tile0 = (int)(wave % ngroups) * NT;
  const int64_t P = g.P();
  const int64_t p0 = pgroup * 16 * MT;
  if (p0 >= P) return;
  const int row = lane & 15, q = lane >> 4;
  const int half = q >> 1, sub = q & 1;
  int64_t pp[MT];
  bool pv[MT];
  Coord3 pc[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    pp[m] = p0 + 16 * m + row;
    pv[m] = pp[m] < P;
    pc[m] = pv[m] ? coord_of(g, (int)pp[m]) : Coord3{0, 0, 0};
  }
  f32x4_t acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[m][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int cpt = ci / 16;       // 16-channel groups per tap
  const int split = blockIdx.y;
  const int hbeg = split * hpw;
  const int nhalf = min(27 * cpt, hbeg + hpw);  // half-steps of 16 K-values
  const bf16x8_t zero = __builtin_bit_cast(bf16x8_t, u32x4_t{0u, 0u, 0u, 0u});
  for (int h0 = hbeg; h0 < nhalf; h0 += 2) {
    const int h = h0 + half;
    const bool hv = h < nhalf;
    const int t = hv ? h / cpt : 0;
    const int c = (h - t * cpt) * 16 + 8 * sub;
    const int dz = t / 9 - 1, dyy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const int64_t off = ((int64_t)dz * g.H + dyy) * g.W + dx;
    bf16x8_t a[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const bool inside = hv && pv[m] && (unsigned)(pc[m].d + dz) < (unsigned)g.D &&
                          (unsigned)(pc[m].h + dyy) < (unsigned)g.H &&
                          (unsigned)(pc[m].w + dx) < (unsigned)g.W;
      a[m] = inside ? load8(x + (pp[m] + off) * ci + c) : zero;
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int oc = (tile0 + j) * 16 + row;
      const bf16x8_t b = (hv && oc < co) ? load8(w + ((int64_t)t * co + oc) * ci + c) : zero;
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m][j], 0, 0, 0);
    }
  }
  if (ws) {
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int oc = (tile0 + j) * 16 + row;
      if (oc >= co) continue;
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t pos = p0 + 16 * m + 4 * q + r;
          if (pos < P) ws[((int64_t)split * P + pos) * co + oc] = acc[m][j][r];
        }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int oc = (tile0 + j) * 16 + row;
    if (oc >= co) continue;
    const float bv = bias ? bias[oc] : 0.f;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t pos = p0 + 16 * m + 4 * q + r;
        if (pos >= P) continue;
        float v = acc[m][j][r] + bv;
        if (relu) v = v > 0.f ? v : 0.f;
        const int64_t o = pos * yc + yoff + oc;
        if (y32) y32[o] = v;
        if (y16) y16[o] = f2bf(v);
      }
  }
}

// y = (relu)(sum over the K slices of ws, in slice order, + bias).
__global__ void k_conv_reduce(int64_t P, int co, int nsplit, const float* __restrict__ ws,
                              const float* __restrict__ bias, int relu,
                              float* __restrict__ y32, uint16_t* __restrict__ y16, int yc,
                              int yoff) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * co) return;
  const int oc = (int)(i % co);
  const int64_t pos = i / co;
  float v = 0.f;
  for (int k = 0; k < nsplit; ++k) v += ws[(int64_t)k * P * co + i];
  v += bias ? bias[oc] : 0.f;
  if (relu) v = v > 0.f ? v : 0.f;
  const int64_t o = pos * yc + yoff + oc;
  if (y32) y32[o] = v;
  if (y16) y16[o] = f2bf(v);
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

// --------------------------------------------------------------------------
// Backward-weight on v_mfma_f32_16x16x32_bf16 with the operands staged in
// LDS: dW[t][co][ci] = sum_p dy[p][co] x[p + off_t][ci] is, per tap, a GEMM
// with M = 16 output channels, N = 16 input channels and K = positions.  A
// workgroup takes one (co, ci) tile pair and a run of `bpw` boxes of
// BD x BH x BW = 256 positions (BW = 16 or 8).  Per box it stages
//  * sA: dy of the box transposed to [co][position] (one 16-byte row
//    segment per lane and K step), and
//  * sB: x of the box and its one-voxel halo transposed to
//    [dx][ci][d][h][w], three copies shifted by dx = -1, 0, +1, so that the
//    eight positions a lane needs for any tap are one aligned 16-byte read;
// then wave w accumulates taps w, w+4, ... over the box's 8 K steps of 32
// positions (octets of 8 consecutive w).  Rows are padded by 8 elements
// against LDS bank conflicts.  Each run of boxes writes one partial sum.
// --------------------------------------------------------------------------
template <int BW>
struct WBox {
  static constexpr int BD = 4, BH = 256 / (4 * BW);
  static constexpr int HD = BD + 2, HH = BH + 2;
  static constexpr int ROW_A = 256 + 8;             // [co][position] row
  static constexpr int ROW_B = HD * HH * BW + 8;    // [ci][d][h][w] plane
};

template <int BW>
__global__ __launch_bounds__(256) void k_conv_wgrad_box(Grid3 g, const uint16_t* __restrict__ x,
                                                        int ci, const uint16_t* __restrict__ dy,
                                                        int co, int bpw,
                                                        float* __restrict__ part) {
  using B = WBox<BW>;
  __shared__ __attribute__((aligned(16))) uint16_t sA[16 * B::ROW_A];
  __shared__ __attribute__((aligned(16))) uint16_t sB[3 * 16 * B::ROW_B];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int row = lane & 15, q = lane >> 4;
  const int nct = co / 16, nit = ci / 16;
  const int tile = blockIdx.x % (nct * nit);
  const int chunk = blockIdx.x / (nct * nit);
  const int co0 = (tile / nit) * 16, ci0 = (tile % nit) * 16;
  const int nbd = g.D / B::BD, nbh = g.H / B::BH, nbw = g.W / BW;
  const int nboxes = g.N * nbd * nbh * nbw;
  f32x4_t acc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) acc[k] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int b_end = min(nboxes, (chunk + 1) * bpw);
  for (int b = chunk * bpw; b < b_end; ++b) {
    const int bw0 = (b % nbw) * BW, bh0 = ((b / nbw) % nbh) * B::BH;
    const int bd0 = ((b / (nbw * nbh)) % nbd) * B::BD, n = b / (nbw * nbh * nbd);
    __syncthreads();  // the previous box's reads are done
    {
      // dy: thread = (octet o of 8 positions, co pair j)
      const int o = tid >> 3, j = tid & 7;
      const int od = o / (B::BH * BW / 8), oh = (o / (BW / 8)) % B::BH, ow = (o % (BW / 8)) * 8;
      const int64_t p = (((int64_t)n * g.D + bd0 + od) * g.H + bh0 + oh) * g.W + bw0 + ow;
      uint32_t v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        v[i] = *reinterpret_cast<const uint32_t*>(dy + (p + i) * co + co0 + 2 * j);
      u32x4_t lo, hi;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        lo[i] = (v[2 * i] & 0xFFFFu) | (v[2 * i + 1] << 16);
        hi[i] = (v[2 * i] >> 16) | (v[2 * i + 1] & 0xFFFF0000u);
      }
      *reinterpret_cast<u32x4_t*>(&sA[(2 * j) * B::ROW_A + o * 8]) = lo;
      *reinterpret_cast<u32x4_t*>(&sA[(2 * j + 1) * B::ROW_A + o * 8]) = hi;
    }
    // x halo: task = (halo row (hd, hh), w octet, ci pair)
    constexpr int ntask = B::HD * B::HH * (BW / 8) * 8;
    for (int task = tid; task < ntask; task += 256) {
      const int j = task & 7, rest = task >> 3;
      const int wo = rest % (BW / 8), r2 = rest / (BW / 8);
      const int hh = r2 % B::HH, hd = r2 / B::HH;
      const int d = bd0 + hd - 1, h = bh0 + hh - 1;
      const bool rowin = (unsigned)d < (unsigned)g.D && (unsigned)h < (unsigned)g.H;
      const int64_t rbase = (((int64_t)n * g.D + d) * g.H + h) * g.W;
      uint32_t v[10];
#pragma unroll
      for (int i = 0; i < 10; ++i) {
        const int w = bw0 + wo * 8 + i - 1;
        v[i] = (rowin && (unsigned)w < (unsigned)g.W)
                   ? *reinterpret_cast<const uint32_t*>(x + (rbase + w) * ci + ci0 + 2 * j)
                   : 0u;
      }
#pragma unroll
      for (int sh = 0; sh < 3; ++sh) {
        u32x4_t lo, hi;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t a = v[sh + 2 * i], c = v[sh + 2 * i + 1];
          lo[i] = (a & 0xFFFFu) | (c << 16);
          hi[i] = (a >> 16) | (c & 0xFFFF0000u);
        }
        const int base = ((hd * B::HH + hh) * BW) + wo * 8;
        *reinterpret_cast<u32x4_t*>(&sB[(sh * 16 + 2 * j) * B::ROW_B + base]) = lo;
        *reinterpret_cast<u32x4_t*>(&sB[(sh * 16 + 2 * j + 1) * B::ROW_B + base]) = hi;
      }
    }
    __syncthreads();
#pragma unroll 2
    for (int ks = 0; ks < 8; ++ks) {
      const int o = 4 * ks + q;
      const int od = o / (B::BH * BW / 8), oh = (o / (BW / 8)) % B::BH, ow = (o % (BW / 8)) * 8;
      const bf16x8_t a = __builtin_bit_cast(
          bf16x8_t, *reinterpret_cast<const u32x4_t*>(&sA[row * B::ROW_A + o * 8]));
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int t = wv + 4 * k;
        if (t < 27) {
          const int dz = t / 9, dyy = (t / 3) % 3, dx = t % 3;  // offsets + 1
          const bf16x8_t bb = __builtin_bit_cast(
              bf16x8_t, *reinterpret_cast<const u32x4_t*>(
                            &sB[(dx * 16 + row) * B::ROW_B + ((od + dz) * B::HH + oh + dyy) * BW +
                                ow]));
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[k], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int t = wv + 4 * k;
    if (t >= 27) continue;
    float* out = part + ((int64_t)chunk * 27 + t) * co * ci;
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(int64_t)(co0 + 4 * q + r) * ci + ci0 + row] = acc[k][r];
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
// block in 16 chunk phases, or, for few outputs (the biases, the head),
// 4 outputs in 64 phases or one output in 256 phases, so that the grid
// still has thousands of threads.
void sum_chunks(const float* part, int64_t nchunks, int64_t n, float* out, hipStream_t s) {
  if (n >= 2048)
    hipLaunchKernelGGL(k_sum_chunks<16>, blocks_for(n, 16), dim3(256), 0, s, part, nchunks, n,
                       out);
  else if (n >= 256)
    hipLaunchKernelGGL(k_sum_chunks<4>, blocks_for(n, 4), dim3(256), 0, s, part, nchunks, n,
                       out);
  else
    hipLaunchKernelGGL(k_sum_chunks<1>, blocks_for(n, 1), dim3(256), 0, s, part, nchunks, n,
                       out);
}

}  // namespace

namespace {
// Launch plan of a forward convolution: NT output-channel tiles per wave
// (an A fragment feeds NT MFMAs), MT position tiles (a B fragment feeds MT)
// and the number of K slices.  Fine levels keep MT = 4 or 2; coarse levels,
// with few positions, first take fewer position tiles per wave and then
// split K so that the grid has about four waves per SIMD and each wave a
// short chain of K steps.
struct FwdPlan {
  int mt, nt, nsplit, hpw;
  int64_t waves;
};
FwdPlan fwd_plan(const ConvGeom& cg, int ci, int co) {
  const int64_t P = (int64_t)cg.N * cg.D * cg.H * cg.W;
  const int tiles = co / 16;
  FwdPlan f;
  f.nt = tiles >= 8 ? 8 : (tiles >= 3 ? 4 : tiles);
  f.mt = f.nt >= 8 ? 2 : 4;
  const int ngroups = (tiles + f.nt - 1) / f.nt;
  constexpr int64_t kTarget = 4096;
  while (f.mt > 1 && ((P + 16 * f.mt - 1) / (16 * f.mt)) * ngroups < kTarget) f.mt /= 2;
  f.waves = ((P + 16 * f.mt - 1) / (16 * f.mt)) * ngroups;
  const int nhalf = 27 * (ci / 16);
  const int ksteps = (nhalf + 1) / 2;
  int ns = (int)std::min<int64_t>(ksteps, (kTarget + f.waves - 1) / f.waves);
  int kps = (ksteps + ns - 1) / ns;  // K steps per slice
  f.hpw = 2 * kps;
  f.nsplit = (nhalf + f.hpw - 1) / f.hpw;
  if (f.nsplit <= 1) {
    f.nsplit = 1;
    f.hpw = nhalf + 1;
  }
  return f;
}
}  // namespace

int64_t conv_fwd_ws(const ConvGeom& cg, int ci, int co) {
  const FwdPlan f = fwd_plan(cg, ci, co);
  return f.nsplit > 1 ? (int64_t)f.nsplit * cg.N * cg.D * cg.H * cg.W * co : 0;
}

int conv_fwd(const ConvGeom& cg, const uint16_t* x, int ci, const uint16_t* w,
             const float* bias, int co, float* y32, uint16_t* y16, int yc, int yoff, int relu,
             float* ws, void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  const int64_t P = (int64_t)cg.N * cg.D * cg.H * cg.W;
  const FwdPlan f = fwd_plan(cg, ci, co);
  if (f.nsplit > 1 && !ws) return (int)hipErrorInvalidValue;
  float* slices = f.nsplit > 1 ? ws : nullptr;
  const dim3 grid((unsigned)((f.waves + 3) / 4), (unsigned)f.nsplit), block(256);
  hipStream_t s = (hipStream_t)stream;
  bool launched = false;
#define HLBN_FWD(M_, N_)                                                                   \
  if (!launched && f.mt == M_ && f.nt == N_) {                                             \
    hipLaunchKernelGGL((k_conv_fwd<M_, N_>), grid, block, 0, s, g, x, ci, w, bias, co, y32,  \
                       y16, yc, yoff, relu, f.hpw, slices);                                \
    launched = true;                                                                       \
  }
  HLBN_FWD(4, 1) HLBN_FWD(2, 1) HLBN_FWD(1, 1)
  HLBN_FWD(4, 2) HLBN_FWD(2, 2) HLBN_FWD(1, 2)
  HLBN_FWD(4, 4) HLBN_FWD(2, 4) HLBN_FWD(1, 4)
  HLBN_FWD(2, 8) HLBN_FWD(1, 8)
#undef HLBN_FWD
  if (!launched) return (int)hipErrorInvalidValue;
  if (slices)
    hipLaunchKernelGGL(k_conv_reduce, blocks_for(P * co), dim3(256), 0, s, P, co, f.nsplit,
                       slices, bias, relu, y32, y16, yc, yoff);
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

namespace {
// Box width of the LDS-staged weight gradient for this level (0: the level
// does not tile into 4 x BH x BW boxes and takes the direct kernel).
int wgrad_box_width(const ConvGeom& cg) {
  if (cg.D % 4) return 0;
  if (cg.W % 16 == 0 && cg.H % 4 == 0) return 16;
  if (cg.W % 8 == 0 && cg.H % 8 == 0) return 8;
  return 0;
}
// Boxes per workgroup: at least about 1024 workgroups over the tile pairs.
int64_t wgrad_boxes_per_group(const ConvGeom& cg, int ci, int co) {
  const int64_t nboxes = (int64_t)cg.N * cg.D * cg.H * cg.W / 256;
  const int64_t tiles = (int64_t)(co / 16) * (ci / 16);
  return std::max<int64_t>(1, nboxes * tiles / 1024);
}
}  // namespace

int64_t wgrad_parts(const ConvGeom& cg, int ci, int co) {
  const int64_t P = (int64_t)cg.N * cg.D * cg.H * cg.W;
  int64_t nchunks;
  if (wgrad_box_width(cg)) {
    const int64_t bpw = wgrad_boxes_per_group(cg, ci, co);
    nchunks = (P / 256 + bpw - 1) / bpw;
  } else {
    const int64_t chunk = wgrad_chunk(cg);
    nchunks = (P + chunk - 1) / chunk;
  }
  return nchunks * 27 * co * ci;
}

int conv_wgrad(const ConvGeom& cg, const uint16_t* x, int ci, const uint16_t* dy, int co,
               float* part, float* dw, void* stream) {
  const Grid3 g{cg.N, cg.D, cg.H, cg.W};
  const int64_t P = g.P();
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)27 * co * ci;
  const int bw = wgrad_box_width(cg);
  if (bw) {
    const int64_t bpw = wgrad_boxes_per_group(cg, ci, co);
    const int64_t nchunks = (P / 256 + bpw - 1) / bpw;
    const dim3 grid((unsigned)(nchunks * (co / 16) * (ci / 16))), block(256);
    if (bw == 16)
      hipLaunchKernelGGL(k_conv_wgrad_box<16>, grid, block, 0, s, g, x, ci, dy, co, (int)bpw,
                         part);
    else
      hipLaunchKernelGGL(k_conv_wgrad_box<8>, grid, block, 0, s, g, x, ci, dy, co, (int)bpw,
                         part);
    sum_chunks(part, nchunks, n, dw, s);
    return (int)hipGetLastError();
  }
  const int64_t chunk = wgrad_chunk(cg);
  const int64_t nchunks = (P + chunk - 1) / chunk;
  constexpr int TG = 9;
  const int64_t waves = nchunks * (27 / TG) * (co / 16) * (ci / 16);
  hipLaunchKernelGGL((k_conv_wgrad<TG>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, g,
                     x, ci, dy, co, chunk, part);
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
