// Learned warm start: a small BF16 U-Net that maps a (template, reference) pair to a displacement
// field, with its forward pass, backward pass and Adam update written for gfx950 (no counterpart in
// the reference; the gradient flow then refines its prediction).
//
// Network (S scales, base width C): per scale two 3x3 convolutions with ReLU (C, 2C, 4C, ... channels),
// 2x2 average pooling between encoder scales, nearest-neighbour upsampling whose output is
// concatenated with the encoder's skip tensor before the decoder's two convolutions, and a 1x1 head to
// 2 channels (the displacement in pixels). The input's 2 channels are padded to 16.
//
// Layout: activations NHWC in BF16 (channels innermost, so a convolution's reduction dimension is
// contiguous); gradients of activations NHWC in FP32; weights FP32 masters [9][Cout][Cin] with BF16
// copies in the forward layout [tap][Cout][Cin] and the data-gradient layout [tap][Cin][Cout].
// A 3x3 convolution is an implicit GEMM on MFMA (v_mfma_f32_32x32x16_bf16): rows = output channels
// (weights), columns = 32 consecutive pixels, reduction = 9 taps x input channels in steps of 16;
// FP32 accumulation, bias and ReLU fused into the epilogue. The upsampling and the concatenation are
// not materialised: the convolution reads channel c < C1 from the coarse tensor at (y/2, x/2) and the
// rest from the skip tensor. The data gradient is the same kernel over the pre-activation gradient with
// flipped taps and transposed weights; the weight gradient reduces over pixels per chunk on MFMA and
// sums the chunks in a fixed order (deterministic).
#include "ms_gpu.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace msg {
int set_error(int code, const std::string &m);  // the calling thread's msg_last_error() (ms_gpu.hip)
}

namespace {

int ufail(int code, const std::string &m) { return msg::set_error(code, "msg_unet: " + m); }
#define UHIP(x)                                                                                    \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return ufail(MSG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define UCHK(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != MSG_OK) return r_; \
  } while (0)

typedef unsigned short bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// FP32 -> BF16, round to nearest even (NaN stays NaN)
__host__ __device__ inline bf16 f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
__device__ inline bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

inline int cdivu(long a, long b) { return (int)((a + b - 1) / b); }

// The input of a convolution at H x W: channels [0, c1) from x1 (H/2 x W/2 and upsampled when up1),
// channels [c1, c1 + c2) from x2 (H x W). c1 and c2 are multiples of 16 (c2 may be 0).
struct ConvIn {
  const bf16 *x1;
  const bf16 *x2;
  int c1, c2, up1;
};

// ------------------------------------------------------------------------------------------------
// 3x3 convolution, forward (or data gradient with flip = 1 and the transposed weights):
// y[p][co] = act(bias[co] + sum_tap sum_ci w[tap'][co][ci] in[p + off(tap)][ci]), tap' = tap or 8 - tap.
// One wave: 32 output channels x NT times and 32 x NP consecutive pixels of one image row (W % (32 NP)
// == 0); four waves per workgroup along the pixels. A weight operand loaded for an MFMA serves the NP
// pixel tiles, an input operand the NT channel tiles. Output FP32 or BF16, NHWC with `cout` channels.
// ------------------------------------------------------------------------------------------------
template <int NT, int NP, typename OutT>
__global__ void __launch_bounds__(256)
k_conv3x3(ConvIn in, const bf16 *__restrict__ w, const float *__restrict__ bias, OutT *__restrict__ y, int B, int H,
          int W, int cout, int relu, int flip) {
  const int cin = in.c1 + in.c2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long M = (long)B * H * W;
  const long m0 = ((long)blockIdx.x * 4 + wave) * 32 * NP;
  if (m0 >= M) return;  // whole waves only (M % (32 NP) == 0)
  const int n0 = blockIdx.y * 32 * NT;
  const int px = lane & 31, kh = lane >> 5;
  // the NP tiles lie in one image row: same row and batch, x advancing by 32
  const int x0 = (int)((m0 + px) % W);
  const long t0 = (m0 + px) / W;
  const int yy = (int)(t0 % H), b = (int)(t0 / H);
  f32x16 acc[NT][NP];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[t][q][r] = 0.f;
  const int Hh = H >> 1, Wh = W >> 1;
  for (int tap = 0; tap < 9; ++tap) {
    const int sy = yy + tap / 3 - 1;
    const int wt = flip ? 8 - tap : tap;
    const bf16 *p1[NP], *p2[NP];
    bool inb[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int sx = x0 + 32 * q + tap % 3 - 1;
      inb[q] = sy >= 0 && sy < H && sx >= 0 && sx < W;
      p1[q] = in.x1;
      p2[q] = in.x2;
      if (inb[q]) {
        p1[q] += (in.up1 ? (((long)b * Hh + (sy >> 1)) * Wh + (sx >> 1)) : (((long)b * H + sy) * W + sx)) * in.c1;
        if (in.c2) p2[q] += (((long)b * H + sy) * W + sx) * in.c2;
      }
    }
    const bf16 *wrow = w + ((long)wt * cout + n0 + px) * cin + 8 * kh;
    for (int c0 = 0; c0 < cin; c0 += 16) {
      const int c = c0 + 8 * kh;
      u16x8 bv[NP];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        bv[q] = (u16x8)0;
        if (inb[q]) bv[q] = *(const u16x8 *)(c < in.c1 ? p1[q] + c : p2[q] + (c - in.c1));
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const u16x8 av = *(const u16x8 *)(wrow + (long)t * 32 * cin + c0);
#pragma unroll
        for (int q = 0; q < NP; ++q)
          acc[t][q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(av), as_bf16x8(bv[q]), acc[t][q], 0, 0, 0);
      }
    }
  }
  // lane: pixel m0 + 32 q + px (column px of tile q); rows = output channels (r/4)*8 + kh*4 + r%4
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const long m = m0 + 32 * q + px;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = n0 + t * 32 + j * 8 + kh * 4;
        OutT *dst = y + m * cout + co;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float v = acc[t][q][j * 4 + e] + (bias ? bias[co + e] : 0.f);
          if (relu) v = v > 0.f ? v : 0.f;
          if constexpr (sizeof(OutT) == 2)
            dst[e] = f2bf(v);
          else
            dst[e] = v;
        }
      }
  }
}

// ------------------------------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution over one chunk of pixels:
// part[chunk][tap][co][ci] = sum_{p in chunk} dz[p][co] in[p + off(tap)][ci].
// Grid: x = chunk, y = tap * (cout / 32) + co block, z = ci block of 32. Each wave reduces a quarter of
// the chunk with MFMA (rows co, columns ci, reduction over 16 pixels per instruction). Both operands
// are channel-innermost while the reduction runs over pixels, so a wave stages each step's 16 pixels
// x 32 channels of dz and of the input through its own LDS block with one 16-byte load per lane and
// operand, then reads them across pixels; the four waves are summed in LDS in a fixed order.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_conv3x3_wgrad(ConvIn in, const bf16 *__restrict__ dz, float *__restrict__ part, int B, int H, int W, int cout,
                int chunk) {
  const int cin = in.c1 + in.c2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cob = cout / 32;
  const int tap = blockIdx.y / cob, co0 = (blockIdx.y % cob) * 32, ci0 = blockIdx.z * 32;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const long M = (long)B * H * W;
  const long pbeg = (long)blockIdx.x * chunk, pend = pbeg + chunk < M ? pbeg + chunk : M;
  const int px = lane & 31, kh = lane >> 5;
  // staging: lane -> pixel q = lane / 4 of the step, channels 8 g .. 8 g + 7 of the 32-block
  const int q = lane >> 2, grp = lane & 3;
  const int cs = ci0 + 8 * grp;  // first input channel this lane stages
  const bool cs_ok = cs < cin;
  const bool cs1 = cs < in.c1;
  const int Hh = H >> 1, Wh = W >> 1;
  __shared__ bf16 sa[4][16 * 32], sb[4][16 * 32];  // [wave][pixel][channel]
  bf16 *la = sa[wave], *lb = sb[wave];
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const long quarter = (chunk / 4 + 15) / 16 * 16;
  const long wbeg = pbeg + wave * quarter;
  const long wend = wbeg + quarter < pend ? wbeg + quarter : pend;
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  for (long p0 = wbeg; p0 < wend; p0 += 16) {
    // 16 consecutive pixels from a multiple of 16 lie in one image row (W % 32 == 0)
    const int row = (int)(p0 / W), x0 = (int)(p0 - (long)row * W);
    const int yy = row % H, b = row / H;
    const int sy = yy + dy;
    const bool pin = p0 + q < wend;
    u16x8 av = (u16x8)0, bv = (u16x8)0;
    if (pin) av = *(const u16x8 *)(dz + (p0 + q) * cout + co0 + 8 * grp);
    const int sx = x0 + q + dx;
    if (pin && cs_ok && sy >= 0 && sy < H && sx >= 0 && sx < W) {
      const bf16 *src;
      if (cs1)
        src = in.x1 + (in.up1 ? (((long)b * Hh + (sy >> 1)) * Wh + (sx >> 1)) : (((long)b * H + sy) * W + sx)) * in.c1 + cs;
      else
        src = in.x2 + (((long)b * H + sy) * W + sx) * in.c2 + (cs - in.c1);
      bv = *(const u16x8 *)src;
    }
    wave_sync();  // the previous step's reads are done
    *(u16x8 *)(la + q * 32 + 8 * grp) = av;
    *(u16x8 *)(lb + q * 32 + 8 * grp) = bv;
    wave_sync();
    u16x8 a, bb;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = la[(8 * kh + k) * 32 + px];
      bb[k] = lb[(8 * kh + k) * 32 + px];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), as_bf16x8(bb), acc, 0, 0, 0);
  }
  __shared__ float red[4][16][64];
#pragma unroll
  for (int r = 0; r < 16; ++r) red[wave][r][lane] = acc[r];
  __syncthreads();
  if (wave == 0) {
    const int ci = ci0 + px;
    float *dst = part + (((long)blockIdx.x * 9 + tap) * cout) * cin;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float v = ((red[0][r][lane] + red[1][r][lane]) + red[2][r][lane]) + red[3][r][lane];
      const int orow = co0 + (r >> 2) * 8 + kh * 4 + (r & 3);
      if (ci < cin) dst[(long)orow * cin + ci] = v;
    }
  }
}

// dw[i] (+)= sum over chunks in order of part[chunk][i]
__global__ void k_sum_chunks(const float *__restrict__ part, int nchunks, long n, float *__restrict__ dw) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a = 0.f;
  for (int c = 0; c < nchunks; ++c) a += part[(long)c * n + i];
  dw[i] = a;
}

// dz = bf16(dy * [y > 0]) and per-block partial sums of dz (in FP32) per channel for the bias gradient.
// With C dividing 256 a block of 256 threads covers RELU_CHUNKS x 256 consecutive values and every
// thread keeps its channel; with C a multiple of 256 a block covers 256 channels of one pixel.
constexpr int RELU_CHUNKS = 16;
__global__ void k_relu_grad(const float *__restrict__ dy, const bf16 *__restrict__ y, bf16 *__restrict__ dz, long n,
                            int C, float *__restrict__ bpart) {
  const int chunks = C <= 256 ? RELU_CHUNKS : 1;
  float v = 0.f;
  for (int k = 0; k < chunks; ++k) {
    const long i = ((long)blockIdx.x * chunks + k) * 256 + threadIdx.x;
    if (i < n) {
      const float u = bf2f(y[i]) > 0.f ? dy[i] : 0.f;
      dz[i] = f2bf(u);
      v += u;
    }
  }
  __shared__ float s[256];
  s[threadIdx.x] = v;
  __syncthreads();
  // channels repeat every C threads (C divides 256 or 256 divides C)
  if (C <= 256) {
    if ((int)threadIdx.x < C) {
      float a = 0.f;
      for (int k = threadIdx.x; k < 256; k += C) a += s[k];
      bpart[(long)blockIdx.x * C + threadIdx.x] = a;
    }
  } else {
    bpart[(long)blockIdx.x * 256 + threadIdx.x] = v;  // C > 256: block covers part of one pixel
  }
}

// out[c] = sum over rows r of part[r * stride + off(c)], one workgroup per column c: each thread sums
// the rows r = t, t + 256, ... in order, then a fixed tree over the 256 threads (deterministic).
// Bias gradients (rows: the blocks of k_relu_grad; C > 256: a pixel's channels span C / 256 blocks)
// and the head's partials use it.
__global__ void __launch_bounds__(256)
k_col_sum(const float *__restrict__ part, int nrows, long stride, int C, float *__restrict__ out) {
  const int c = blockIdx.x;
  float a = 0.f;
  if (C <= 256 || stride != 256) {
    for (int r = threadIdx.x; r < nrows; r += 256) a += part[(long)r * stride + c];
  } else {
    const int per = C / 256;  // blocks per pixel
    for (int r = c / 256 + per * (int)threadIdx.x; r < nrows; r += per * 256) a += part[(long)r * 256 + (c % 256)];
  }
  __shared__ float s[256];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = s[0];
}

// 2x2 average pooling, NHWC bf16: y[b][yy][x][c] = mean of x[b][2yy..][2x..][c]
__global__ void k_pool(const bf16 *__restrict__ x, bf16 *__restrict__ y, int B, int H, int W, int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  const long n = (long)B * Ho * Wo * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long t = i / C;
  const int xo = (int)(t % Wo);
  t /= Wo;
  const int yo = (int)(t % Ho), b = (int)(t / Ho);
  const long r0 = (((long)b * H + 2 * yo) * W + 2 * xo) * C + c, r1 = r0 + (long)W * C;
  const float s = (bf2f(x[r0]) + bf2f(x[r0 + C])) + (bf2f(x[r1]) + bf2f(x[r1 + C]));
  y[i] = f2bf(0.25f * s);
}
// dx[b][y][x][c] += 0.25 dy[b][y/2][x/2][c]
__global__ void k_pool_grad(const float *__restrict__ dy, float *__restrict__ dx, int B, int H, int W, int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)B * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long t = i / C;
  const int x = (int)(t % W);
  t /= W;
  const int yy = (int)(t % H), b = (int)(t / H);
  dx[i] += 0.25f * dy[(((long)b * (H / 2) + yy / 2) * (W / 2) + x / 2) * C + c];
}

// The data gradient of the virtual input (FP32, H x W x (c1 + c2)) to its two sources: the coarse
// tensor (summed over each 2x2 block when upsampled) and the skip tensor, accumulated.
__global__ void k_split_grad1(const float *__restrict__ dv, float *__restrict__ d1, int B, int H, int W, int c1,
                              int c2, int up1) {
  const int cin = c1 + c2;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Hs = up1 ? H / 2 : H, Ws = up1 ? W / 2 : W;
  const long n = (long)B * Hs * Ws * c1;
  if (i >= n) return;
  const int c = (int)(i % c1);
  long t = i / c1;
  const int x = (int)(t % Ws);
  t /= Ws;
  const int yy = (int)(t % Hs), b = (int)(t / Hs);
  float a;
  if (up1) {
    const long r0 = (((long)b * H + 2 * yy) * W + 2 * x) * cin + c, r1 = r0 + (long)W * cin;
    a = (dv[r0] + dv[r0 + cin]) + (dv[r1] + dv[r1 + cin]);
  } else {
    a = dv[(((long)b * H + yy) * W + x) * cin + c];
  }
  d1[i] += a;
}
__global__ void k_split_grad2(const float *__restrict__ dv, float *__restrict__ d2, long npix, int c1, int c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * c2) return;
  const long p = i / c2;
  const int c = (int)(i % c2);
  d2[i] += dv[p * (c1 + c2) + c1 + c];
}

// Input pairs [B][2][H][W] FP32 (template, reference) -> NHWC bf16 with 16 channels (14 zero).
__global__ void k_pack_input(const float *__restrict__ in, bf16 *__restrict__ x, int B, int H, int W) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W, n = (long)B * hw;
  if (i >= n) return;
  const long b = i / hw, q = i % hw;
  bf16 *dst = x + i * 16;
  dst[0] = f2bf(in[(b * 2) * hw + q]);
  dst[1] = f2bf(in[(b * 2 + 1) * hw + q]);
#pragma unroll
  for (int c = 2; c < 16; ++c) dst[c] = 0;
}

// 1x1 head to 2 channels: out[b][o][y][x] = bh[o] + sum_c wh[o][c] x[p][c] (FP32, planar)
__global__ void k_head(const bf16 *__restrict__ x, const float *__restrict__ wh, const float *__restrict__ bh,
                       float *__restrict__ out, int B, int H, int W, int C) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long hw = (long)H * W, n = (long)B * hw;
  if (p >= n) return;
  float a0 = bh[0], a1 = bh[1];
  const bf16 *src = x + p * C;
  for (int c = 0; c < C; ++c) {
    const float v = bf2f(src[c]);
    a0 += wh[c] * v;
    a1 += wh[C + c] * v;
  }
  const long b = p / hw, q = p % hw;
  out[(b * 2) * hw + q] = a0;
  out[(b * 2 + 1) * hw + q] = a1;
}

// Loss 1/(2 n) sum (out - target)^2 (n = all output values) and its gradient d = (out - target) / n;
// per-block partial sums of the loss in FP64.
__global__ void k_mse(const float *__restrict__ out, const float *__restrict__ tgt, float *__restrict__ d, long n,
                      double *__restrict__ lpart) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0;
  if (i < n) {
    const float e = out[i] - tgt[i];
    d[i] = e / (float)n;
    l = 0.5 * (double)e * (double)e / (double)n;
  }
  __shared__ double s[256];
  s[threadIdx.x] = l;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = s[0];
}

// Head backward over 256 pixels per workgroup: dx[p][c] += sum_o wh[o][c] dout[o][p], and per-block
// partials of dwh[o][c] = sum_p dout[o][p] x[p][c] (2C) and dbh[o] = sum_p dout[o][p] (2). Thread t
// takes channel t % C of the pixels t / C, t / C + 256 / C, ... (C divides 256), so the loads and
// stores of a wave are consecutive channels; the 256 / C partial sums of a channel are added in a
// fixed order.
constexpr int HEAD_CMAX = 256;
__global__ void __launch_bounds__(256)
k_head_grad(const bf16 *__restrict__ x, const float *__restrict__ wh, const float *__restrict__ dout,
            float *__restrict__ dx, int B, int H, int W, int C, float *__restrict__ part) {
  const long hw = (long)H * W, n = (long)B * hw;
  const long p0 = (long)blockIdx.x * 256;
  const int t = threadIdx.x, c = t % C, G = 256 / C, g = t / C;
  const float w0 = wh[c], w1 = wh[C + c];
  float a0 = 0.f, a1 = 0.f, s0 = 0.f, s1 = 0.f;
  for (int q = g; q < 256; q += G) {
    const long p = p0 + q;
    if (p >= n) break;
    const long b = p / hw, r = p % hw;
    const float d0 = dout[(b * 2) * hw + r], d1 = dout[(b * 2 + 1) * hw + r];
    const float v = bf2f(x[p * C + c]);
    dx[p * C + c] += w0 * d0 + w1 * d1;
    a0 += d0 * v;
    a1 += d1 * v;
    s0 += d0;
    s1 += d1;
  }
  __shared__ float r0[256], r1[256], q0[256], q1[256];
  r0[t] = a0;
  r1[t] = a1;
  q0[t] = s0;
  q1[t] = s1;
  __syncthreads();
  float *dst = part + (long)blockIdx.x * (2 * C + 2);
  if (t < C) {
    float u0 = 0.f, u1 = 0.f;
    for (int k = 0; k < G; ++k) {
      u0 += r0[k * C + t];
      u1 += r1[k * C + t];
    }
    dst[t] = u0;
    dst[C + t] = u1;
  }
  if (t == 0) {  // every pixel once: the threads of channel 0
    float u0 = 0.f, u1 = 0.f;
    for (int k = 0; k < G; ++k) {
      u0 += q0[k * C];
      u1 += q1[k * C];
    }
    dst[2 * C] = u0;
    dst[2 * C + 1] = u1;
  }
}

// Adam (Kingma & Ba 2015, Algorithm 1) on FP32 masters.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi, v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// BF16 copies of a layer's weights: wf[tap][co][ci] (forward) and wd[tap][ci][co] (data gradient)
__global__ void k_weights_bf16(const float *__restrict__ w, bf16 *__restrict__ wf, bf16 *__restrict__ wd, int cout,
                               int cin) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = 9L * cout * cin;
  if (i >= n) return;
  const int ci = (int)(i % cin);
  const long t = i / cin;
  const int co = (int)(t % cout), tap = (int)(t / cout);
  const bf16 v = f2bf(w[i]);
  wf[i] = v;
  if (wd) wd[((long)tap * cin + ci) * cout + co] = v;
}

// ------------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------------
struct Dev {
  void *p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (n >= bytes) return MSG_OK;
    if (p) UHIP(hipFree(p));
    p = nullptr;
    n = 0;
    UHIP(hipMalloc(&p, bytes ? bytes : 16));
    n = bytes;
    return MSG_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// A tensor of the network: NHWC at scale s (H >> s, W >> s) with C channels.
struct Act {
  int scale, C;
  Dev y;   // bf16 values
  Dev dy;  // fp32 gradient
};

struct Conv {
  int in1, in2;   // source activations (in2 < 0: none)
  int up1;        // in1 is upsampled
  int out;        // destination activation
  int cin1, cin2, cout, cin_real;
  long woff, boff;  // offsets in the flat parameter vector
};

}  // namespace

struct msg_unet {
  int scales = 4, base = 32;
  std::vector<Act> acts;
  std::vector<Conv> convs;
  std::vector<std::pair<int, int>> pools;  // (from, to)
  int head_in = -1;
  long whoff = 0, bhoff = 0, nparams = 0;
  Dev w, g, m1, m2;          // FP32 masters, gradient, Adam moments (flat)
  std::vector<Dev> wf, wd;   // per conv: bf16 forward / data-gradient weights
  Dev in_packed, out, dout, dv, dz, part, bpart, hpart, lpart;
  int B = 0, H = 0, Wd = 0;
  long step = 0;
  hipStream_t s = nullptr;
};

namespace {

int build_graph(msg_unet *u) {
  const int S = u->scales, C = u->base;
  auto act = [&](int scale, int c) {
    u->acts.push_back(Act{scale, c, Dev(), Dev()});
    return (int)u->acts.size() - 1;
  };
  const int input = act(0, 16);
  std::vector<int> skip(S);
  int cur = input, ch = 16;
  for (int s = 0; s < S; ++s) {
    const int co = C << s;
    if (s > 0) {
      const int p = act(s, ch);
      u->pools.push_back({cur, p});
      cur = p;
    }
    const int a = act(s, co);
    u->convs.push_back(Conv{cur, -1, 0, a, ch, 0, co, s == 0 ? 2 : ch, 0, 0});
    const int b = act(s, co);
    u->convs.push_back(Conv{a, -1, 0, b, co, 0, co, co, 0, 0});
    skip[s] = b;
    cur = b;
    ch = co;
  }
  for (int s = S - 2; s >= 0; --s) {
    const int co = C << s;
    const int a = act(s, co);
    u->convs.push_back(Conv{cur, skip[s], 1, a, ch, co, co, ch + co, 0, 0});
    const int b = act(s, co);
    u->convs.push_back(Conv{a, -1, 0, b, co, 0, co, co, 0, 0});
    cur = b;
    ch = co;
  }
  u->head_in = cur;
  long off = 0;
  for (auto &c : u->convs) {
    c.woff = off;
    off += 9L * c.cout * (c.cin1 + c.cin2);
    c.boff = off;
    off += c.cout;
  }
  u->whoff = off;
  off += 2L * C;
  u->bhoff = off;
  off += 2;
  u->nparams = off;
  return MSG_OK;
}

int refresh_bf16(msg_unet *u) {
  for (size_t k = 0; k < u->convs.size(); ++k) {
    const Conv &c = u->convs[k];
    const int cin = c.cin1 + c.cin2;
    const long n = 9L * c.cout * cin;
    hipLaunchKernelGGL(k_weights_bf16, dim3(cdivu(n, 256)), dim3(256), 0, u->s, u->w.as<float>() + c.woff,
                       u->wf[k].as<bf16>(), k == 0 ? nullptr : u->wd[k].as<bf16>(), c.cout, cin);
    UHIP(hipGetLastError());
  }
  return MSG_OK;
}

int ensure_acts(msg_unet *u, int B, int H, int W, bool grads) {
  if (H % (32 << (u->scales - 1)) != 0 || W % (32 << (u->scales - 1)) != 0)
    return ufail(MSG_ERR_ARG, "msg_unet: H and W must be multiples of 32 * 2^(scales-1)");
  if ((long)B * H * W * 16 >= (1L << 31)) return ufail(MSG_ERR_ARG, "batch x H x W too large");
  for (size_t k = 1; k < u->acts.size(); ++k) {  // activation 0 is the packed input
    Act &a = u->acts[k];
    const size_t px = (size_t)B * (H >> a.scale) * (W >> a.scale);
    UCHK(a.y.ensure(px * a.C * 2));
    if (grads) UCHK(a.dy.ensure(px * a.C * 4));
  }
  const size_t px0 = (size_t)B * H * W;
  UCHK(u->in_packed.ensure(px0 * 16 * 2));
  u->acts[0].y.p = u->in_packed.p;  // not owned (freed through in_packed)
  u->acts[0].y.n = 0;
  UCHK(u->out.ensure(px0 * 2 * 4));
  if (grads) {
    UCHK(u->dout.ensure(px0 * 2 * 4));
    size_t dvmax = 0, dzmax = 0, partmax = 0, bpmax = 0;
    for (auto &c : u->convs) {
      const int s = u->acts[c.out].scale;
      const size_t px = (size_t)B * (H >> s) * (W >> s);
      const int cin = c.cin1 + c.cin2;
      dvmax = std::max(dvmax, px * cin * 4);
      dzmax = std::max(dzmax, px * c.cout * 2);
      const size_t chunks = (px + 4095) / 4096;
      partmax = std::max(partmax, chunks * 9 * c.cout * cin * 4);
      bpmax = std::max(bpmax, (px * c.cout + 255) / 256 * 256 * 4);
    }
    UCHK(u->dv.ensure(dvmax));
    UCHK(u->dz.ensure(dzmax));
    UCHK(u->part.ensure(partmax));
    UCHK(u->bpart.ensure(bpmax));
    UCHK(u->hpart.ensure((px0 + 255) / 256 * (2 * u->base + 2) * 4));
    UCHK(u->lpart.ensure((px0 * 2 + 255) / 256 * 8));
  }
  u->B = B;
  u->H = H;
  u->Wd = W;
  return MSG_OK;
}

int launch_conv(msg_unet *u, const Conv &c, const bf16 *w, const float *bias, void *y, bool out_f32, int cout,
                int relu, int flip, ConvIn in, int H, int W) {
  const long M = (long)u->B * H * W;
  const int nt = (cout % 64 == 0) ? 2 : 1;
  const int np = (W % 64 == 0) ? 2 : 1;
  dim3 grid(cdivu(M, 128 * np), cout / (32 * nt));
  if (out_f32) {
    float *yo = (float *)y;
    if (nt == 2 && np == 2)
      hipLaunchKernelGGL((k_conv3x3<2, 2, float>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else if (nt == 2)
      hipLaunchKernelGGL((k_conv3x3<2, 1, float>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else if (np == 2)
      hipLaunchKernelGGL((k_conv3x3<1, 2, float>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else
      hipLaunchKernelGGL((k_conv3x3<1, 1, float>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
  } else {
    bf16 *yo = (bf16 *)y;
    if (nt == 2 && np == 2)
      hipLaunchKernelGGL((k_conv3x3<2, 2, bf16>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else if (nt == 2)
      hipLaunchKernelGGL((k_conv3x3<2, 1, bf16>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else if (np == 2)
      hipLaunchKernelGGL((k_conv3x3<1, 2, bf16>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
    else
      hipLaunchKernelGGL((k_conv3x3<1, 1, bf16>), grid, dim3(256), 0, u->s, in, w, bias, yo, u->B, H, W, cout, relu, flip);
  }
  UHIP(hipGetLastError());
  (void)c;
  return MSG_OK;
}

ConvIn conv_input(msg_unet *u, const Conv &c) {
  ConvIn in;
  in.x1 = u->acts[c.in1].y.as<bf16>();
  in.x2 = c.in2 >= 0 ? u->acts[c.in2].y.as<bf16>() : nullptr;
  in.c1 = c.cin1;
  in.c2 = c.cin2;
  in.up1 = c.up1;
  return in;
}

int forward(msg_unet *u, const float *input) {
  const int B = u->B, H = u->H, W = u->Wd;
  const long px0 = (long)B * H * W;
  hipLaunchKernelGGL(k_pack_input, dim3(cdivu(px0, 256)), dim3(256), 0, u->s, input, u->in_packed.as<bf16>(), B, H, W);
  UHIP(hipGetLastError());
  size_t pi = 0;
  for (size_t k = 0; k < u->convs.size(); ++k) {
    const Conv &c = u->convs[k];
    // the pool feeding this convolution runs first
    while (pi < u->pools.size() && u->pools[pi].second == c.in1) {
      const Act &a = u->acts[u->pools[pi].first];
      const int Hs = H >> a.scale, Ws = W >> a.scale;
      const long n = (long)B * (Hs / 2) * (Ws / 2) * a.C;
      hipLaunchKernelGGL(k_pool, dim3(cdivu(n, 256)), dim3(256), 0, u->s, a.y.as<bf16>(),
                         u->acts[u->pools[pi].second].y.as<bf16>(), B, Hs, Ws, a.C);
      UHIP(hipGetLastError());
      ++pi;
    }
    const int s = u->acts[c.out].scale;
    UCHK(launch_conv(u, c, u->wf[k].as<bf16>(), u->w.as<float>() + c.boff, u->acts[c.out].y.p, false, c.cout, 1, 0,
                     conv_input(u, c), H >> s, W >> s));
  }
  hipLaunchKernelGGL(k_head, dim3(cdivu(px0, 256)), dim3(256), 0, u->s, u->acts[u->head_in].y.as<bf16>(),
                     u->w.as<float>() + u->whoff, u->w.as<float>() + u->bhoff, u->out.as<float>(), B, H, W,
                     u->acts[u->head_in].C);
  UHIP(hipGetLastError());
  return MSG_OK;
}

int backward(msg_unet *u, const float *target, double *loss) {
  const int B = u->B, H = u->H, W = u->Wd;
  const long px0 = (long)B * H * W;
  float *g = u->g.as<float>();
  for (auto &a : u->acts)
    if (a.dy.p) UHIP(hipMemsetAsync(a.dy.p, 0, a.dy.n, u->s));
  const long nout = px0 * 2;
  const int nlb = cdivu(nout, 256);
  hipLaunchKernelGGL(k_mse, dim3(nlb), dim3(256), 0, u->s, u->out.as<float>(), target, u->dout.as<float>(), nout,
                     u->lpart.as<double>());
  UHIP(hipGetLastError());
  // head
  const int C0 = u->acts[u->head_in].C;
  const int nhb = cdivu(px0, 256);
  hipLaunchKernelGGL(k_head_grad, dim3(nhb), dim3(256), 0, u->s, u->acts[u->head_in].y.as<bf16>(),
                     u->w.as<float>() + u->whoff, u->dout.as<float>(), u->acts[u->head_in].dy.as<float>(), B, H, W,
                     C0, u->hpart.as<float>());
  UHIP(hipGetLastError());
  hipLaunchKernelGGL(k_col_sum, dim3(2 * C0 + 2), dim3(256), 0, u->s, u->hpart.as<float>(), nhb,
                     (long)(2 * C0 + 2), 2 * C0 + 2, g + u->whoff);  // dwh then dbh are contiguous in the flat vector
  UHIP(hipGetLastError());
  size_t pi = u->pools.size();
  for (int k = (int)u->convs.size() - 1; k >= 0; --k) {
    const Conv &c = u->convs[k];
    const Act &o = u->acts[c.out];
    const int Hs = H >> o.scale, Ws = W >> o.scale;
    const long px = (long)B * Hs * Ws;
    const long n = px * c.cout;
    const int nb = cdivu(n, 256L * (c.cout <= 256 ? RELU_CHUNKS : 1));
    hipLaunchKernelGGL(k_relu_grad, dim3(nb), dim3(256), 0, u->s, o.dy.as<float>(), o.y.as<bf16>(), u->dz.as<bf16>(),
                       n, c.cout, u->bpart.as<float>());
    UHIP(hipGetLastError());
    hipLaunchKernelGGL(k_col_sum, dim3(c.cout), dim3(256), 0, u->s, u->bpart.as<float>(), nb,
                       (long)(c.cout <= 256 ? c.cout : 256), c.cout, g + c.boff);
    UHIP(hipGetLastError());
    const ConvIn in = conv_input(u, c);
    const int cin = c.cin1 + c.cin2;
    const int chunk = 4096;
    const int nch = cdivu(px, chunk);
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(nch, 9 * (c.cout / 32), cdivu(cin, 32)), dim3(256), 0, u->s, in,
                       u->dz.as<bf16>(), u->part.as<float>(), B, Hs, Ws, c.cout, chunk);
    UHIP(hipGetLastError());
    const long nw = 9L * c.cout * cin;
    hipLaunchKernelGGL(k_sum_chunks, dim3(cdivu(nw, 256)), dim3(256), 0, u->s, u->part.as<float>(), nch, nw,
                       g + c.woff);
    UHIP(hipGetLastError());
    if (k == 0) continue;  // no gradient for the input
    // data gradient of the virtual input (FP32), then to its sources
    ConvIn din;
    din.x1 = u->dz.as<bf16>();
    din.x2 = nullptr;
    din.c1 = c.cout;
    din.c2 = 0;
    din.up1 = 0;
    UCHK(launch_conv(u, c, u->wd[k].as<bf16>(), nullptr, u->dv.p, true, cin, 0, 1, din, Hs, Ws));
    const Act &a1 = u->acts[c.in1];
    const long n1 = (long)B * (H >> a1.scale) * (W >> a1.scale) * c.cin1;
    hipLaunchKernelGGL(k_split_grad1, dim3(cdivu(n1, 256)), dim3(256), 0, u->s, u->dv.as<float>(), a1.dy.as<float>(),
                       B, Hs, Ws, c.cin1, c.cin2, c.up1);
    UHIP(hipGetLastError());
    if (c.in2 >= 0) {
      hipLaunchKernelGGL(k_split_grad2, dim3(cdivu(px * c.cin2, 256)), dim3(256), 0, u->s, u->dv.as<float>(),
                         u->acts[c.in2].dy.as<float>(), px, c.cin1, c.cin2);
      UHIP(hipGetLastError());
    }
    // a pooled activation passes its gradient on to the pool's source
    while (pi > 0 && u->pools[pi - 1].second == c.in1) {
      --pi;
      const Act &src = u->acts[u->pools[pi].first];
      const int Hp = H >> src.scale, Wp = W >> src.scale;
      const long np = (long)B * Hp * Wp * src.C;
      hipLaunchKernelGGL(k_pool_grad, dim3(cdivu(np, 256)), dim3(256), 0, u->s, a1.dy.as<float>(),
                         src.dy.as<float>(), B, Hp, Wp, src.C);
      UHIP(hipGetLastError());
    }
  }
  if (loss) {
    std::vector<double> lp(nlb);
    UHIP(hipMemcpyAsync(lp.data(), u->lpart.p, sizeof(double) * nlb, hipMemcpyDeviceToHost, u->s));
    UHIP(hipStreamSynchronize(u->s));
    double a = 0;
    for (double v : lp) a += v;
    *loss = a;
  }
  return MSG_OK;
}

}  // namespace

extern "C" {

int msg_unet_create(msg_unet **net, int scales, int base_channels, unsigned long long seed) {
  // powers of two: every layer's channel count then divides 256 or is a multiple of it, which the
  // per-channel reductions assume
  if (!net || scales < 1 || scales > 6 || base_channels < 32 || base_channels > HEAD_CMAX ||
      (base_channels & (base_channels - 1)) != 0)
    return ufail(MSG_ERR_ARG, "msg_unet_create: scales in 1..6, base_channels 32, 64, 128 or 256");
  msg_unet *u = new msg_unet();
  u->scales = scales;
  u->base = base_channels;
  build_graph(u);
  int rc = MSG_OK;
  auto bail = [&](int r) {
    rc = r;
    return r;
  };
  if (bail(u->w.ensure(u->nparams * 4)) || bail(u->g.ensure(u->nparams * 4)) || bail(u->m1.ensure(u->nparams * 4)) ||
      bail(u->m2.ensure(u->nparams * 4))) {
    delete u;
    return rc;
  }
  u->wf.resize(u->convs.size());
  u->wd.resize(u->convs.size());
  for (size_t k = 0; k < u->convs.size(); ++k) {
    const Conv &c = u->convs[k];
    const size_t n = 9L * c.cout * (c.cin1 + c.cin2) * 2;
    if (u->wf[k].ensure(n) || u->wd[k].ensure(n)) {
      delete u;
      return MSG_ERR_HIP;
    }
  }
  if (hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking) != hipSuccess) {
    delete u;
    return ufail(MSG_ERR_HIP, "msg_unet_create: stream");
  }
  // He initialisation from the seed (the padded input channels stay zero), zero biases
  std::vector<float> h(u->nparams, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (auto &c : u->convs) {
    const int cin = c.cin1 + c.cin2;
    const double sd = std::sqrt(2.0 / (9.0 * c.cin_real));
    for (int t = 0; t < 9; ++t)
      for (int o = 0; o < c.cout; ++o)
        for (int i = 0; i < cin; ++i)
          h[c.woff + ((long)t * c.cout + o) * cin + i] = (c.in1 != 0 || i < c.cin_real) ? (float)(sd * nd(rng)) : 0.f;
  }
  const double sdh = std::sqrt(1.0 / u->base);
  for (int k = 0; k < 2 * u->base; ++k) h[u->whoff + k] = (float)(sdh * nd(rng));
  *net = u;
  return msg_unet_set_params(u, h.data());
}

int msg_unet_destroy(msg_unet *u) {
  if (!u) return MSG_OK;
  if (u->s) (void)hipStreamSynchronize(u->s);
  for (size_t k = 0; k < u->acts.size(); ++k) {
    if (k == 0) u->acts[0].y.p = nullptr;  // the packed input buffer, owned by in_packed
    u->acts[k].y.release();
    u->acts[k].dy.release();
  }
  for (Dev *d : {&u->w, &u->g, &u->m1, &u->m2, &u->in_packed, &u->out, &u->dout, &u->dv, &u->dz, &u->part,
                 &u->bpart, &u->hpart, &u->lpart})
    d->release();
  for (auto &d : u->wf) d.release();
  for (auto &d : u->wd) d.release();
  if (u->s) (void)hipStreamDestroy(u->s);
  delete u;
  return MSG_OK;
}

int msg_unet_param_count(msg_unet *u, long long *count) {
  if (!u || !count) return ufail(MSG_ERR_ARG, "invalid argument");
  *count = u->nparams;
  return MSG_OK;
}

int msg_unet_get_params(msg_unet *u, float *host) {
  if (!u || !host) return ufail(MSG_ERR_ARG, "invalid argument");
  UHIP(hipMemcpyAsync(host, u->w.p, u->nparams * 4, hipMemcpyDeviceToHost, u->s));
  UHIP(hipStreamSynchronize(u->s));
  return MSG_OK;
}

int msg_unet_set_params(msg_unet *u, const float *host) {
  if (!u || !host) return ufail(MSG_ERR_ARG, "invalid argument");
  UHIP(hipMemcpyAsync(u->w.p, host, u->nparams * 4, hipMemcpyHostToDevice, u->s));
  UHIP(hipMemsetAsync(u->m1.p, 0, u->nparams * 4, u->s));
  UHIP(hipMemsetAsync(u->m2.p, 0, u->nparams * 4, u->s));
  u->step = 0;
  UCHK(refresh_bf16(u));
  UHIP(hipStreamSynchronize(u->s));
  return MSG_OK;
}

int msg_unet_forward(msg_unet *u, const float *in, float *out, int batch, int H, int W, void *stream) {
  if (!u || !in || !out || batch < 1) return ufail(MSG_ERR_ARG, "invalid argument");
  if (stream) UHIP(hipStreamSynchronize((hipStream_t)stream));
  UCHK(ensure_acts(u, batch, H, W, false));
  UCHK(forward(u, in));
  UHIP(hipMemcpyAsync(out, u->out.p, (size_t)batch * 2 * H * W * 4, hipMemcpyDeviceToDevice, u->s));
  UHIP(hipStreamSynchronize(u->s));
  return MSG_OK;
}

int msg_unet_gradient(msg_unet *u, const float *in, const float *target, int batch, int H, int W, double *loss,
                      float *grad_host, void *stream) {
  if (!u || !in || !target || batch < 1) return ufail(MSG_ERR_ARG, "invalid argument");
  if (stream) UHIP(hipStreamSynchronize((hipStream_t)stream));
  UCHK(ensure_acts(u, batch, H, W, true));
  UCHK(forward(u, in));
  UCHK(backward(u, target, loss));
  if (grad_host) UHIP(hipMemcpyAsync(grad_host, u->g.p, u->nparams * 4, hipMemcpyDeviceToHost, u->s));
  UHIP(hipStreamSynchronize(u->s));
  return MSG_OK;
}

int msg_unet_adam_step(msg_unet *u, float lr) {
  if (!u || !(lr > 0)) return ufail(MSG_ERR_ARG, "invalid argument");
  ++u->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
  hipLaunchKernelGGL(k_adam, dim3(cdivu(u->nparams, 256)), dim3(256), 0, u->s, u->w.as<float>(), u->g.as<float>(),
                     u->m1.as<float>(), u->m2.as<float>(), u->nparams, lr, b1, b2, 1e-8f, c1, c2);
  UHIP(hipGetLastError());
  UCHK(refresh_bf16(u));
  UHIP(hipStreamSynchronize(u->s));
  return MSG_OK;
}

}  // extern "C"
