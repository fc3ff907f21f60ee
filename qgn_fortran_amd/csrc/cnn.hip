// A 4-layer 3x3 convolutional network on doubly periodic grids, written for
// the gfx950 matrix cores: BF16 weights and activations, FP32 accumulation
// (v_mfma_f32_16x16x16bf16_1k), FP32 master weights trained with Adam.  It is
// the learned subgrid PV closure of the coarse model.
//
// Layers: cin -> 32 -> 32 -> 32 -> cout, ReLU after the first three; the
// padding is periodic (every tap wraps around the grid).  Tensors are
// (n, H, W, C) with the channel fastest ("NHWC").
//
// Parameter vector (FP32, host and device): for each layer l, W_l[co][tap][ci]
// then b_l[co], tap = 3 (dy + 1) + (dx + 1) for the input offset (dy, dx).
//
// Convolution as an implicit GEMM: per wave, D (16 output channels x 16
// pixels) += A (16 output channels x 16 input channels of one tap) *
// B (16 input channels x 16 pixels).  MFMA operand layouts on gfx950:
// A[i][k] in lane l at i = l % 16, k = 4 (l / 16) + r; B[k][j] at j = l % 16,
// k = 4 (l / 16) + r; D[i][j] at i = 4 (l / 16) + r, j = l % 16 (r = 0..3).
// With channels fastest, each lane's A and B operands are 8 contiguous bytes,
// and each lane's 4 results are 4 consecutive output channels of one pixel.
// The backward pass for the data is the same convolution with the weights
// transposed (ci <-> co) and the taps mirrored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cnn.hpp"
#include "qgn.h"

struct qgn_cnn;

namespace qcnn {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int kLayers = 4;
constexpr int kHidden = 32;

__host__ __device__ constexpr int pad16(int c) { return (c + 15) / 16 * 16; }

// round to nearest even; NaN stays NaN
__device__ __forceinline__ u16 f2bf(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u16)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(((unsigned)h) << 16); }

// MODE 0: bf16 out = relu(acc + bias)          (hidden layers, forward)
// MODE 1: fp32 out = acc + bias                (last layer, forward)
// MODE 2: bf16 out = acc * (mask > 0)          (backward data; mask = layer input)
// in: (n, H, W, CI) bf16; wpk: (COP, 9, CI) bf16 with rows >= CO zero;
// out: (n, H, W, CO).  Block: 4 waves; each wave takes TP tiles of 16
// consecutive pixels of one row (wave w of block b: tiles TP (4 b + w) ..),
// so every weight operand read from LDS feeds TP independent MFMA chains.
// Needs W % 16 == 0 and n H W % 16 == 0.
template <int CI, int CO, int MODE, int TP>
__global__ void __launch_bounds__(256) k_conv3x3(const u16* __restrict__ in, const u16* __restrict__ wpk,
                                                 const float* __restrict__ bias, const u16* __restrict__ mask,
                                                 void* __restrict__ out, int H, int W, long ntiles) {
  constexpr int COP = pad16(CO);
  constexpr int NW = COP * 9 * CI;
  __shared__ __align__(16) u16 ws[NW];
  for (int e = threadIdx.x * 8; e < NW; e += 256 * 8)
    *reinterpret_cast<uint4*>(&ws[e]) = *reinterpret_cast<const uint4*>(&wpk[e]);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const long t0 = ((long)blockIdx.x * 4 + w) * TP;
  long img[TP];
  int y[TP], x0[TP];
  bool on[TP];
#pragma unroll
  for (int u = 0; u < TP; ++u) {
    on[u] = t0 + u < ntiles;
    const long p0 = on[u] ? (t0 + u) * 16 : 0;
    const long row = p0 / W;                // n H + y
    x0[u] = (int)(p0 - row * W);
    y[u] = (int)(row % H);
    img[u] = (row - y[u]) * W;               // first pixel of image n
  }
  f4v acc[TP][COP / 16];
#pragma unroll
  for (int u = 0; u < TP; ++u)
#pragma unroll
    for (int mt = 0; mt < COP / 16; ++mt) acc[u][mt] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const u16* src[TP];
#pragma unroll
    for (int u = 0; u < TP; ++u) {
      int yy = y[u] + dy;
      yy += yy < 0 ? H : 0;
      yy -= yy >= H ? H : 0;
      int xx = x0[u] + j + dx;
      xx += xx < 0 ? W : 0;
      xx -= xx >= W ? W : 0;
      src[u] = in + (img[u] + (long)yy * W + xx) * CI + 4 * g;
    }
#pragma unroll
    for (int kc = 0; kc < CI / 16; ++kc) {
      s4v b[TP];
#pragma unroll
      for (int u = 0; u < TP; ++u) b[u] = *reinterpret_cast<const s4v*>(src[u] + kc * 16);
#pragma unroll
      for (int mt = 0; mt < COP / 16; ++mt) {
        const s4v a = *reinterpret_cast<const s4v*>(&ws[((mt * 16 + j) * 9 + tap) * CI + kc * 16 + 4 * g]);
#pragma unroll
        for (int u = 0; u < TP; ++u) acc[u][mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[u], acc[u][mt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < TP; ++u) {
    if (!on[u]) continue;
    const long p = (t0 + u) * 16 + j;
#pragma unroll
    for (int mt = 0; mt < COP / 16; ++mt) {
      const int co = mt * 16 + 4 * g;
      if (co >= CO) continue;               // CO is a multiple of 4
      if constexpr (MODE == 1) {
        float4 v;
        v.x = acc[u][mt][0] + bias[co];
        v.y = acc[u][mt][1] + bias[co + 1];
        v.z = acc[u][mt][2] + bias[co + 2];
        v.w = acc[u][mt][3] + bias[co + 3];
        *reinterpret_cast<float4*>(reinterpret_cast<float*>(out) + p * CO + co) = v;
      } else {
        float v[4];
        if constexpr (MODE == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = fmaxf(acc[u][mt][r] + bias[co + r], 0.f);
        } else {
          const ushort4 m = *reinterpret_cast<const ushort4*>(mask + p * CO + co);
          v[0] = bf2f(m.x) > 0.f ? acc[u][mt][0] : 0.f;
          v[1] = bf2f(m.y) > 0.f ? acc[u][mt][1] : 0.f;
          v[2] = bf2f(m.z) > 0.f ? acc[u][mt][2] : 0.f;
          v[3] = bf2f(m.w) > 0.f ? acc[u][mt][3] : 0.f;
        }
        ushort4 o;
        o.x = f2bf(v[0]);
        o.y = f2bf(v[1]);
        o.z = f2bf(v[2]);
        o.w = f2bf(v[3]);
        *reinterpret_cast<ushort4*>(reinterpret_cast<u16*>(out) + p * CO + co) = o;
      }
    }
  }
}

// The same convolution on 2-D tiles with the input window staged in LDS.
// A workgroup computes a tile of TH x TW pixels (TH TW = 64 TP) of one
// image, each wave TP rows of 16 pixels; the (TH + 2) x (TW + 2) input
// window with its periodic halo is loaded once with 16-byte loads, all of a
// thread's loads in flight together, so each input pixel is read about
// (TH + 2)(TW + 2) / (TH TW) times instead of 3.  The MFMA B operands are
// read from LDS at a per-lane base plus a constant offset per tap.  A staged
// pixel holds CI + 4 values, so the 16 pixels of an operand fall on
// distinct banks.  Needs TH | H and TW | W.
template <int CI, int CO, int MODE, int TP, int TW>
__global__ void __launch_bounds__(256) k_conv3x3_tile(const u16* __restrict__ in, const u16* __restrict__ wpk,
                                                      const float* __restrict__ bias, const u16* __restrict__ mask,
                                                      void* __restrict__ out, int H, int W) {
  constexpr int COP = pad16(CO);
  constexpr int NW = COP * 9 * CI;
  // LDS weight rows of 9 CI values padded by 8 (16 bytes): the 16 output
  // channels of an A operand read fall in distinct banks
  constexpr int WS = 9 * CI + 8;
  constexpr int TH = 64 * TP / TW;
  constexpr int WW = TW + 2, NR = TH + 2;
  constexpr int PS = CI + 4;            // bf16 values per staged pixel
  constexpr int CH = CI / 8;            // 16-byte chunks per pixel
  constexpr int NCH = NR * WW * CH;
  constexpr int MAXC = (NCH + 255) / 256;
  static_assert(TW % 16 == 0 && (64 * TP) % TW == 0, "tile of whole 16-pixel rows");
  __shared__ __align__(16) u16 ws[COP * WS];
  __shared__ __align__(16) u16 xs[NR * WW * PS];
  const int tx_n = W / TW, ty_n = H / TH;
  const int img = blockIdx.x / (tx_n * ty_n);
  const int trem = blockIdx.x - img * tx_n * ty_n;
  const int y0 = (trem / tx_n) * TH, x0 = (trem % tx_n) * TW;
  const long imgp = (long)img * H * W;   // first pixel of the image
  // the image as a raw buffer: lanes past the window load zero, no branch
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<u16*>(in + imgp * CI), (short)0, H * W * CI * 2, 0x00020000);
  typedef int i4v __attribute__((ext_vector_type(4)));
  i4v v[MAXC];
#pragma unroll
  for (int u = 0; u < MAXC; ++u) {
    const int c = threadIdx.x + u * 256;
    const int pix = c / CH, sub = c - pix * CH;
    const int r = pix / WW, cc = pix - r * WW;
    int yy = y0 - 1 + r;
    yy += yy < 0 ? H : 0;
    yy -= yy >= H ? H : 0;
    int xx = x0 - 1 + cc;
    xx += xx < 0 ? W : 0;
    xx -= xx >= W ? W : 0;
    const int off = c < NCH ? ((yy * W + xx) * CI + sub * 8) * 2 : 0x7ffffff0;
    v[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
  }
  for (int e = threadIdx.x * 8; e < NW; e += 256 * 8) {
    const int co = e / (9 * CI);   // 9 CI is a multiple of 8: a chunk stays in its row
    *reinterpret_cast<uint4*>(&ws[co * WS + e - co * 9 * CI]) = *reinterpret_cast<const uint4*>(&wpk[e]);
  }
#pragma unroll
  for (int u = 0; u < MAXC; ++u) {
    const int c = threadIdx.x + u * 256;
    if (c >= NCH) continue;
    const int pix = c / CH, sub = c - pix * CH;
    uint2* d = reinterpret_cast<uint2*>(xs + pix * PS + sub * 8);   // 8-byte aligned
    d[0] = make_uint2((unsigned)v[u].x, (unsigned)v[u].y);
    d[1] = make_uint2((unsigned)v[u].z, (unsigned)v[u].w);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  // tile k = w TP + u: row k / (TW / 16) of the workgroup tile, 16 pixels
  // from column 16 (k % (TW / 16)); window position (row + 1, column + 1)
  const u16* base[TP];
#pragma unroll
  for (int u = 0; u < TP; ++u) {
    const int k = w * TP + u;
    const int tr = k / (TW / 16), tc = (k % (TW / 16)) * 16 + j;
    base[u] = xs + ((tr + 1) * WW + tc + 1) * PS + 4 * g;
  }
  f4v acc[TP][COP / 16];
#pragma unroll
  for (int u = 0; u < TP; ++u)
#pragma unroll
    for (int mt = 0; mt < COP / 16; ++mt) acc[u][mt] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int toff = ((tap / 3 - 1) * WW + (tap % 3 - 1)) * PS;
#pragma unroll
    for (int kc = 0; kc < CI / 16; ++kc) {
      s4v b[TP];
#pragma unroll
      for (int u = 0; u < TP; ++u) b[u] = *reinterpret_cast<const s4v*>(base[u] + toff + kc * 16);
#pragma unroll
      for (int mt = 0; mt < COP / 16; ++mt) {
        const s4v a = *reinterpret_cast<const s4v*>(&ws[(mt * 16 + j) * WS + tap * CI + kc * 16 + 4 * g]);
#pragma unroll
        for (int u = 0; u < TP; ++u) acc[u][mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[u], acc[u][mt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < TP; ++u) {
    const int k = w * TP + u;
    const int tr = k / (TW / 16), tc = (k % (TW / 16)) * 16 + j;
    const long p = imgp + (long)(y0 + tr) * W + x0 + tc;
#pragma unroll
    for (int mt = 0; mt < COP / 16; ++mt) {
      const int co = mt * 16 + 4 * g;
      if (co >= CO) continue;               // CO is a multiple of 4
      if constexpr (MODE == 1) {
        float4 o;
        o.x = acc[u][mt][0] + bias[co];
        o.y = acc[u][mt][1] + bias[co + 1];
        o.z = acc[u][mt][2] + bias[co + 2];
        o.w = acc[u][mt][3] + bias[co + 3];
        *reinterpret_cast<float4*>(reinterpret_cast<float*>(out) + p * CO + co) = o;
      } else {
        float f[4];
        if constexpr (MODE == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) f[r] = fmaxf(acc[u][mt][r] + bias[co + r], 0.f);
        } else {
          const ushort4 m = *reinterpret_cast<const ushort4*>(mask + p * CO + co);
          f[0] = bf2f(m.x) > 0.f ? acc[u][mt][0] : 0.f;
          f[1] = bf2f(m.y) > 0.f ? acc[u][mt][1] : 0.f;
          f[2] = bf2f(m.z) > 0.f ? acc[u][mt][2] : 0.f;
          f[3] = bf2f(m.w) > 0.f ? acc[u][mt][3] : 0.f;
        }
        // two packed conversions (round to nearest even, as f2bf for finite values)
        typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
        typedef float f2v __attribute__((ext_vector_type(2)));
        const bf2v lo = __builtin_convertvector((f2v){f[0], f[1]}, bf2v);
        const bf2v hi = __builtin_convertvector((f2v){f[2], f[3]}, bf2v);
        uint2 o;
        __builtin_memcpy(&o.x, &lo, 4);
        __builtin_memcpy(&o.y, &hi, 4);
        *reinterpret_cast<uint2*>(reinterpret_cast<u16*>(out) + p * CO + co) = o;
      }
    }
  }
}

// Weight and bias gradients of one layer, per workgroup:
//   dW[co][tap][ci] = sum_p dY[p][co] X[p + off(tap)][ci],  db[co] = sum_p dY[p][co]
// over the workgroup's pixels, into its slice of `part` (CO 9 CI weights,
// then CO biases); k_wgrad_reduce_chunks/_final add the slices in a fixed order, so a
// training step is reproducible.  X: (n, H, W, CI) bf16 (layer input), dY:
// (n, H, W, COP) bf16 (gradient of the layer's pre-activation, channels >=
// CO zero).  Block: `rows` image rows of image blockIdx.y.
//
// The reduction runs over pixels, so both MFMA operands are staged in LDS
// channel-major (each channel's pixels contiguous): a lane's four pixels of
// A (dY, 16 output channels x 16 pixels) are one 8-byte read, and those of
// B (X, 16 pixels x 16 input channels) for the three horizontal taps of a
// row come from one 8-byte and one 4-byte read (pixels p0-1 .. p0+4; the
// middle tap takes its words across a 2-byte boundary).  The X rows carry
// their periodic neighbours (pixel -1 and W) and live in a 3-row ring, so
// each image row is staged once; row strides of W + 4 values keep the 16
// channels of a read in distinct LDS banks.  Each wave owns one 16 x 16
// (output, input) channel tile for all 9 taps (9 independent accumulation
// chains); with 2 such tiles, two waves split the 16-pixel chunks of each
// row and add their sums in wave order at the end.
template <int CI, int CO>
__global__ void __launch_bounds__(256) k_conv_wgrad(const u16* __restrict__ x, const u16* __restrict__ dy,
                                                    float* __restrict__ part, int H, int W, int rows) {
  constexpr int COP = pad16(CO);
  constexpr int NMT = COP / 16, NNT = CI / 16, NP = NMT * NNT;
  static_assert(NP == 2 || NP == 4, "channel tiles");
  constexpr int WPP = 4 / NP;            // waves per channel tile
  extern __shared__ __align__(16) u16 lds[];
  const int S = W + 4;                   // row stride (values) of the channel-major rows
  u16* xs = lds;                         // [3 ring slots][CI][S]: pixel p at p + 1, p = -1 .. W
  u16* ds = lds + 3 * CI * S;            // [COP][S]: pixel p at p
  __shared__ float bpart[8 * 256];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int n = blockIdx.y;
  const int y0 = blockIdx.x * rows;
  const int y1 = y0 + rows < H ? y0 + rows : H;
  const int pair = w / WPP, sub = w % WPP;
  const int mt = pair / NNT, nt = pair % NNT;
  f4v acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
  // bias partial sums: thread tid's staging loads always hold channels
  // 8 (tid % (COP / 8)) .. + 7 (256 is a multiple of COP / 8)
  float bacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  part += (long)(blockIdx.y * gridDim.x + blockIdx.x) * (CO * 9 * CI + CO);
  // image row yy (periodic) of X into ring slot yy mod 3, channel-major with neighbours
  auto stage_x = [&](int yy) {
    int yr = yy;
    yr += yr < 0 ? H : 0;
    yr -= yr >= H ? H : 0;
    u16* dst = xs + ((yy + 3) % 3) * CI * S;
    const u16* src = x + ((long)n * H + yr) * W * CI;
    for (int e = tid; e < (W + 2) * (CI / 8); e += 256) {
      const int pp = e / (CI / 8), c8 = 8 * (e - pp * (CI / 8));   // padded pixel pp = p + 1
      int px = pp - 1;
      px += px < 0 ? W : 0;
      px -= px >= W ? W : 0;
      const uint4 v = *reinterpret_cast<const uint4*>(src + (long)px * CI + c8);
      const u16* h = reinterpret_cast<const u16*>(&v);
#pragma unroll
      for (int c = 0; c < 8; ++c) dst[(c8 + c) * S + pp] = h[c];
    }
  };
  for (int yy = y0 - 1; yy <= y0; ++yy) stage_x(yy);
  for (int y = y0; y < y1; ++y) {
    stage_x(y + 1);   // slot (y + 1) mod 3 held row y - 2, which the previous row's chunks no longer read
    {
      const u16* src = dy + ((long)n * H + y) * W * COP;
      for (int e = tid; e < W * (COP / 8); e += 256) {
        const int px = e / (COP / 8), c8 = 8 * (e - px * (COP / 8));
        const uint4 v = *reinterpret_cast<const uint4*>(src + (long)px * COP + c8);
        const u16* h = reinterpret_cast<const u16*>(&v);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          ds[(c8 + c) * S + px] = h[c];
          bacc[c] += bf2f(h[c]);
        }
      }
    }
    __syncthreads();
    const u16* arow = ds + (mt * 16 + j) * S + 4 * g;
    for (int ks = sub; ks < W / 16; ks += WPP) {
      const s4v a = *reinterpret_cast<const s4v*>(arow + 16 * ks);
#pragma unroll
      for (int ddy = 0; ddy < 3; ++ddy) {
        const u16* brow = xs + ((y + ddy - 1 + 3) % 3) * CI * S + (nt * 16 + j) * S + 16 * ks + 4 * g;
        const uint2 w01 = *reinterpret_cast<const uint2*>(brow);       // pixels p0-1 .. p0+2
        const unsigned w2 = *reinterpret_cast<const unsigned*>(brow + 4);   // pixels p0+3, p0+4
        uint2 m;
        m.x = __builtin_amdgcn_alignbyte(w01.y, w01.x, 2);               // pixels p0, p0+1
        m.y = __builtin_amdgcn_alignbyte(w2, w01.y, 2);                  // pixels p0+2, p0+3
        const uint2 hi = make_uint2(w01.y, w2);                          // pixels p0+1 .. p0+4
        s4v b0, b1, b2;
        __builtin_memcpy(&b0, &w01, 8);
        __builtin_memcpy(&b1, &m, 8);
        __builtin_memcpy(&b2, &hi, 8);
        acc[3 * ddy + 0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc[3 * ddy + 0], 0, 0, 0);
        acc[3 * ddy + 1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc[3 * ddy + 1], 0, 0, 0);
        acc[3 * ddy + 2] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b2, acc[3 * ddy + 2], 0, 0, 0);
      }
    }
    __syncthreads();   // ds and the ring slot of row y - 1 are rewritten next
  }
  // the staging buffers are free: combine the waves of one channel tile in
  // wave order, then the bias partial sums in thread order
  float* red = reinterpret_cast<float*>(lds);
  if (WPP == 2 && sub == 1) {
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[((pair * 9 + t) * 4 + r) * 64 + lane] = acc[t][r];
  }
  __syncthreads();
  if (sub == 0) {
#pragma unroll
    for (int t = 0; t < 9; ++t) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[t][r];
        if (WPP == 2) v += red[((pair * 9 + t) * 4 + r) * 64 + lane];
        const int co = mt * 16 + 4 * g + r;
        if (co < CO) part[(co * 9 + t) * CI + nt * 16 + j] = v;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) bpart[c * 256 + tid] = bacc[c];
  __syncthreads();
  if (tid < CO) {
    // channel tid = 8 c8 + c: threads k with k % (COP / 8) == c8, in order
    const int c8 = tid / 8, c = tid % 8;
    float sb = 0.f;
    for (int k = c8; k < 256; k += COP / 8) sb += bpart[c * 256 + k];
    part[CO * 9 * CI + tid] = sb;
  }
}

// grad[i] += sum over the nblk workgroup slices of part[b * psz + i]; i <
// CO 9 CI are weights (dw), the rest biases (db).  Two passes, each in a
// fixed order: k_wgrad_reduce_chunks sums chunk c of the slices ([c per,
// (c + 1) per)) for 64 entries per workgroup (wave q takes slices q, q + 4,
// ... of the chunk in order, then the four wave sums are added in wave
// order) into red[c psz + i]; k_wgrad_reduce_final adds the chunk sums in
// chunk order.  Chunks spread the slices over many workgroups (a training
// batch of 10 images of 256^2 has 640 slices).
constexpr int kWgradChunks = 16;
__global__ void __launch_bounds__(256) k_wgrad_reduce_chunks(const float* __restrict__ part, int nblk, int psz,
                                                             float* __restrict__ red) {
  __shared__ float ws[4][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int c = blockIdx.y;
  const int per = (nblk + kWgradChunks - 1) / kWgradChunks;
  const int b0 = c * per, b1 = b0 + per < nblk ? b0 + per : nblk;
  float sacc = 0.f;
  if (i < psz)
    for (int b = b0 + q; b < b1; b += 4) sacc += part[(long)b * psz + i];
  ws[q][lane] = sacc;
  __syncthreads();
  if (q == 0 && i < psz) red[(long)c * psz + i] = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
}
__global__ void __launch_bounds__(256) k_wgrad_reduce_final(const float* __restrict__ red, int psz, int nw,
                                                            float* __restrict__ dw, float* __restrict__ db) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= psz) return;
  float t = 0.f;
#pragma unroll
  for (int c = 0; c < kWgradChunks; ++c) t += red[(long)c * psz + i];
  if (i < nw) dw[i] += t;
  else db[i - nw] += t;
}

// Mean-squared error of pred against target over n_pix * CO values; writes
// dL/dpred as bf16 with COP channels (the padding channels zero, COP a power
// of two) and each workgroup's sum of squares to sq[blockIdx.x] (the host
// adds the kMseBlocks partial sums in order, so the loss is reproducible).
constexpr int kMseBlocks = 1024;
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ pred, const float* __restrict__ tgt,
                                             long npix, int CO, int COP, float scale, u16* __restrict__ dpred,
                                             double* __restrict__ sq) {
  __shared__ double red[4];
  double s = 0.0;
  const long n = npix * COP;
  const int lg = __ffs(COP) - 1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long p = e >> lg;
    const int c = (int)e & (COP - 1);
    float d = 0.f;
    if (c < CO) {
      d = pred[p * CO + c] - tgt[p * CO + c];
      s += (double)d * d;
    }
    dpred[e] = f2bf(scale * d);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) sq[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Adam on the FP32 parameter vector; c1 = 1/(1 - b1^t), c2 = 1/(1 - b2^t).
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// bf16 weights of one layer from the FP32 W[co][tap][ci]: forward layout
// wf[COP][9][CI] and backward-data layout wb[CI][9][COP] with mirrored taps.
__global__ void k_pack_w(const float* __restrict__ W, int CI, int CO, int COP, u16* __restrict__ wf,
                         u16* __restrict__ wb) {
  const int n = COP * 9 * CI;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int co = e / (9 * CI);
  const int rem = e - co * 9 * CI;
  const int tap = rem / CI, ci = rem - tap * CI;
  const float v = co < CO ? W[(co * 9 + tap) * CI + ci] : 0.f;
  const u16 h = f2bf(v);
  wf[e] = h;
  wb[(ci * 9 + (8 - tap)) * COP + co] = h;
}

struct Layer {
  int ci, co, cop;
  long woff, boff;   // offsets of W and b in the parameter vector
};

int allow_lds(const void* f, size_t bytes) {
  if (bytes <= 65536) return 0;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess ? 0 : -1;
}

}  // namespace qcnn

using namespace qcnn;

struct qgn_cnn {
  int cin = 16, cout = 8;
  Layer L[kLayers];
  int np = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 parameters, gradients, Adam moments
  u16 *wf[kLayers] = {}, *wb[kLayers] = {};
  int t = 0;                   // Adam step
  // activations of the last forward: act[0] = input (not owned), act[1..3] hidden
  const u16* act[kLayers] = {};
  u16* hid[3] = {};
  u16 *dpred = nullptr, *dbuf[2] = {};
  float* pred = nullptr;       // last forward output (not owned)
  long cap = 0;                // pixels the buffers hold
  int n = 0, H = 0, W = 0;
  double* sq = nullptr;
  double* h_sq = nullptr;
  float* part = nullptr;       // weight-gradient workgroup slices
  long part_cap = 0;           // floats part holds
  hipStream_t stream = nullptr;
  // Models (as their closure) and ensembles (in their captured graphs) that
  // use the activation buffers and packed weights: while any holds the
  // network, those buffers may neither move nor be freed.
  int holds = 0;
  bool closure_fwd = false;    // the last forward ran inside a model's RHS
};

namespace {

thread_local std::string g_cnn_err;
int cfail(int code, const std::string& m) {
  g_cnn_err = m;
  return code;
}
#define CCHK(x)                                                                  \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) return cfail(QGN_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define CRET(x)          \
  do {                   \
    int r_ = (x);        \
    if (r_) return r_;   \
  } while (0)

int repack(qgn_cnn* h, hipStream_t s) {
  for (int l = 0; l < kLayers; ++l) {
    const Layer& L = h->L[l];
    const int n = L.cop * 9 * L.ci;
    hipLaunchKernelGGL(k_pack_w, dim3((n + 255) / 256), dim3(256), 0, s, h->w + L.woff, L.ci, L.co, L.cop, h->wf[l],
                       h->wb[l]);
  }
  CCHK(hipGetLastError());
  return 0;
}

int ensure(qgn_cnn* h, long npix) {
  if (npix <= h->cap) return 0;
  if (h->holds > 0)
    return cfail(QGN_EINVAL, "cnn: the network is attached to a model or captured by an ensemble, which use its "
                             "activation buffers for " + std::to_string(h->cap) + " pixels; a pass on " +
                             std::to_string(npix) + " pixels would reallocate them (detach it first)");
  for (u16* p : h->hid)
    if (p) (void)hipFree(p);
  for (u16* p : h->dbuf)
    if (p) (void)hipFree(p);
  if (h->dpred) (void)hipFree(h->dpred);
  for (int i = 0; i < 3; ++i) CCHK(hipMalloc((void**)&h->hid[i], npix * kHidden * sizeof(u16)));
  for (int i = 0; i < 2; ++i) CCHK(hipMalloc((void**)&h->dbuf[i], npix * kHidden * sizeof(u16)));
  CCHK(hipMalloc((void**)&h->dpred, npix * h->L[kLayers - 1].cop * sizeof(u16)));
  h->cap = npix;
  return 0;
}

template <int CI, int CO, int MODE>
int conv(const u16* in, const u16* w, const float* b, const u16* mask, void* out, int n, int H, int W, hipStream_t s) {
  constexpr int TP = 4;   // pixel tiles per wave
  // 2-D tiles with the input window in LDS: 16 x 16 pixels where they
  // divide the image, else 8 x 32
  if (H % 16 == 0 && W % 16 == 0 && (long)H * W * CI * 2 < 0x7ffffff0L) {
    hipLaunchKernelGGL((k_conv3x3_tile<CI, CO, MODE, TP, 16>), dim3((unsigned)((long)n * (H / 16) * (W / 16))),
                       dim3(256), 0, s, in, w, b, mask, out, H, W);
    CCHK(hipGetLastError());
    return 0;
  }
  if (H % 8 == 0 && W % 32 == 0 && (long)H * W * CI * 2 < 0x7ffffff0L) {
    hipLaunchKernelGGL((k_conv3x3_tile<CI, CO, MODE, TP, 32>), dim3((unsigned)((long)n * (H / 8) * (W / 32))),
                       dim3(256), 0, s, in, w, b, mask, out, H, W);
    CCHK(hipGetLastError());
    return 0;
  }
  const long ntiles = (long)n * H * W / 16;
  const long blocks = (ntiles + 4 * TP - 1) / (4 * TP);
  hipLaunchKernelGGL((k_conv3x3<CI, CO, MODE, TP>), dim3((unsigned)blocks), dim3(256), 0, s, in, w, b, mask, out, H,
                     W, ntiles);
  CCHK(hipGetLastError());
  return 0;
}

// image rows per weight-gradient workgroup: 640 workgroups for 10 images of
// 256^2 (2.5 per CU), 9248 floats of partial sums each
constexpr int kWgradRows = 4;

// workgroup slices of the weight-gradient scratch for a batch of n images of
// H rows, and floats per slice (the largest layer: 32 x 9 x 32 + 32)
long wgrad_slices(int n, int H) { return (long)n * ((H + kWgradRows - 1) / kWgradRows); }
constexpr long kWgradSlice = 32L * 9 * 32 + 32;

template <int CI, int CO>
int wgrad(const u16* x, const u16* dy, float* dw, float* db, int n, int H, int W, float* part, hipStream_t s) {
  constexpr int COP = pad16(CO);
  constexpr int PSZ = CO * 9 * CI + CO;
  static_assert(PSZ <= kWgradSlice, "slice");
  const size_t lds = std::max((size_t)(3 * CI + COP) * (W + 4) * sizeof(u16), (size_t)9 * 4 * 64 * 4 * 2);
  auto kern = k_conv_wgrad<CI, CO>;
  if (allow_lds((const void*)kern, lds)) return cfail(QGN_EINVAL, "wgrad: grid row too wide for LDS");
  const dim3 grid((H + kWgradRows - 1) / kWgradRows, n);
  hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, x, dy, part, H, W, kWgradRows);
  CCHK(hipGetLastError());
  // the chunk sums follow the slices in the same scratch buffer
  const int nblk = (int)(grid.x * grid.y);
  float* red = part + (long)nblk * kWgradSlice;
  hipLaunchKernelGGL(k_wgrad_reduce_chunks, dim3((PSZ + 63) / 64, kWgradChunks), dim3(256), 0, s, part, nblk, PSZ,
                     red);
  CCHK(hipGetLastError());
  hipLaunchKernelGGL(k_wgrad_reduce_final, dim3((PSZ + 255) / 256), dim3(256), 0, s, red, PSZ, CO * 9 * CI, dw, db);
  CCHK(hipGetLastError());
  return 0;
}

int check_shape(const qgn_cnn* h, int n, int H, int W) {
  if (n < 1 || H < 3 || W < 16 || W % 16 || ((long)n * H * W) % 64)
    return cfail(QGN_EINVAL, "cnn: needs W % 16 == 0, n*H*W % 64 == 0, H >= 3");
  // the weight gradient's LDS rows: 3 input rows and one gradient row of 32 channels, W + 4 values each
  if ((3L * 32 + 32) * (W + 4) * 2 > 160 * 1024) return cfail(QGN_EINVAL, "cnn: W too large");
  return 0;
}

// layer l forward: input act[l] -> hid[l] or y
int layer_fwd(qgn_cnn* h, int l, const u16* in, void* out, int n, int H, int W, hipStream_t s) {
  const Layer& L = h->L[l];
  const float* b = h->w + L.boff;
  if (l < kLayers - 1) {
    if (L.ci == 16) return conv<16, 32, 0>(in, h->wf[l], b, nullptr, out, n, H, W, s);
    return conv<32, 32, 0>(in, h->wf[l], b, nullptr, out, n, H, W, s);
  }
  if (L.co == 8) return conv<32, 8, 1>(in, h->wf[l], b, nullptr, out, n, H, W, s);
  return conv<32, 16, 1>(in, h->wf[l], b, nullptr, out, n, H, W, s);
}

}  // namespace

namespace qcnn {

int forward(qgn_cnn* h, const u16* x, int n, int H, int W, float* y, hipStream_t s, bool closure) {
  CRET(check_shape(h, n, H, W));
  CRET(ensure(h, (long)n * H * W));
  h->closure_fwd = closure;
  h->act[0] = x;
  for (int l = 0; l < kLayers; ++l) {
    void* out = l < kLayers - 1 ? (void*)h->hid[l] : (void*)y;
    CRET(layer_fwd(h, l, h->act[l], out, n, H, W, s));
    if (l < kLayers - 1) h->act[l + 1] = h->hid[l];
  }
  h->pred = y;
  h->n = n;
  h->H = H;
  h->W = W;
  return 0;
}

int reserve(qgn_cnn* h, long npix) { return ensure(h, npix); }

void hold(qgn_cnn* h, int delta) { h->holds += delta; }

const char* last_error() { return g_cnn_err.c_str(); }

int channels(const qgn_cnn* h, int* cin, int* cout) {
  *cin = h->cin;
  *cout = h->cout;
  return 0;
}

}  // namespace qcnn

extern "C" {

const char* qgn_cnn_last_error(void) { return g_cnn_err.c_str(); }

int qgn_cnn_destroy(qgn_cnn* h) {
  if (!h) return 0;
  if (h->holds > 0)
    return cfail(QGN_EINVAL, "cnn: the network is attached to a model or captured by an ensemble; detach it "
                             "(qgn_set_closure(ctx, NULL, ...)) and destroy the ensemble first");
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* ptrs[] = {h->w,       h->g,       h->m,       h->v,       h->dpred,  h->sq,
                  h->hid[0],  h->hid[1],  h->hid[2],  h->dbuf[0], h->dbuf[1], h->part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int l = 0; l < kLayers; ++l) {
    if (h->wf[l]) (void)hipFree(h->wf[l]);
    if (h->wb[l]) (void)hipFree(h->wb[l]);
  }
  if (h->h_sq) (void)hipHostFree(h->h_sq);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

static int cnn_create_impl(qgn_cnn* h, int cin, int cout, const float* params) {
  if (cin != 16 && cin != 32) return cfail(QGN_EINVAL, "cnn: cin must be 16 or 32");
  if (cout != 8 && cout != 16) return cfail(QGN_EINVAL, "cnn: cout must be 8 or 16");
  h->cin = cin;
  h->cout = cout;
  const int ci[kLayers] = {cin, kHidden, kHidden, kHidden};
  const int co[kLayers] = {kHidden, kHidden, kHidden, cout};
  long off = 0;
  for (int l = 0; l < kLayers; ++l) {
    h->L[l] = Layer{ci[l], co[l], pad16(co[l]), off, off + (long)co[l] * 9 * ci[l]};
    off = h->L[l].boff + co[l];
  }
  h->np = (int)off;
  const size_t pb = (size_t)h->np * sizeof(float);
  CCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  CCHK(hipMalloc((void**)&h->w, pb));
  CCHK(hipMalloc((void**)&h->g, pb));
  CCHK(hipMalloc((void**)&h->m, pb));
  CCHK(hipMalloc((void**)&h->v, pb));
  CCHK(hipMemset(h->m, 0, pb));
  CCHK(hipMemset(h->v, 0, pb));
  CCHK(hipMemset(h->g, 0, pb));
  CCHK(hipMemcpy(h->w, params, pb, hipMemcpyHostToDevice));
  for (int l = 0; l < kLayers; ++l) {
    const size_t nb = (size_t)h->L[l].cop * 9 * h->L[l].ci * sizeof(u16);
    CCHK(hipMalloc((void**)&h->wf[l], nb));
    CCHK(hipMalloc((void**)&h->wb[l], nb));
    CCHK(hipMemset(h->wb[l], 0, nb));
  }
  CCHK(hipMalloc((void**)&h->sq, kMseBlocks * sizeof(double)));
  CCHK(hipHostMalloc((void**)&h->h_sq, kMseBlocks * sizeof(double), hipHostMallocDefault));
  CRET(repack(h, h->stream));
  CCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qgn_cnn_create(qgn_cnn** out, int cin, int cout, const float* params) {
  if (!out || !params) return cfail(QGN_EINVAL, "null argument");
  *out = nullptr;
  qgn_cnn* h = new qgn_cnn();
  const int rc = cnn_create_impl(h, cin, cout, params);
  if (rc) {
    std::string m = g_cnn_err;
    qgn_cnn_destroy(h);
    g_cnn_err = m;
    return rc;
  }
  *out = h;
  return 0;
}

int qgn_cnn_nparams(qgn_cnn* h, int* n) {
  if (!h || !n) return cfail(QGN_EINVAL, "null argument");
  *n = h->np;
  return 0;
}

int qgn_cnn_get_params(qgn_cnn* h, float* params, float* grads) {
  if (!h) return cfail(QGN_EINVAL, "null argument");
  CCHK(hipStreamSynchronize(h->stream));
  if (params) CCHK(hipMemcpy(params, h->w, (size_t)h->np * sizeof(float), hipMemcpyDeviceToHost));
  if (grads) CCHK(hipMemcpy(grads, h->g, (size_t)h->np * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int qgn_cnn_set_params(qgn_cnn* h, const float* params) {
  if (!h || !params) return cfail(QGN_EINVAL, "null argument");
  CCHK(hipStreamSynchronize(h->stream));
  CCHK(hipMemcpy(h->w, params, (size_t)h->np * sizeof(float), hipMemcpyHostToDevice));
  CRET(repack(h, h->stream));
  CCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qgn_cnn_forward(qgn_cnn* h, const void* x, int n, int H, int W, float* y) {
  if (!h || !x || !y) return cfail(QGN_EINVAL, "null argument");
  CRET(qcnn::forward(h, reinterpret_cast<const u16*>(x), n, H, W, y, h->stream, false));
  CCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// Backward of the last forward against `target`: gradients (overwritten)
// and the mean-squared error.
static int backward_impl(qgn_cnn* h, const float* target, double* loss) {
  if (!h->pred) return cfail(QGN_EINVAL, "cnn: backward needs a forward first");
  if (h->closure_fwd)
    return cfail(QGN_EINVAL, "cnn: the last forward pass was the closure's, inside a model's RHS; call "
                             "qgn_cnn_forward before qgn_cnn_backward");
  hipStream_t s = h->stream;
  const int n = h->n, H = h->H, W = h->W;
  const long npix = (long)n * H * W;
  const Layer& L4 = h->L[kLayers - 1];
  const long need = (wgrad_slices(n, H) + kWgradChunks) * kWgradSlice;   // slices, then chunk sums
  if (need > h->part_cap) {
    CCHK(hipStreamSynchronize(s));
    if (h->part) (void)hipFree(h->part);
    h->part = nullptr;
    h->part_cap = 0;
    CCHK(hipMalloc((void**)&h->part, (size_t)need * sizeof(float)));
    h->part_cap = need;
  }
  float* part = h->part;
  CCHK(hipMemsetAsync(h->g, 0, (size_t)h->np * sizeof(float), s));
  const float scale = (float)(2.0 / ((double)npix * L4.co));
  hipLaunchKernelGGL(k_mse, dim3(kMseBlocks), dim3(256), 0, s, h->pred, target, npix, L4.co, L4.cop, scale,
                     h->dpred, h->sq);
  CCHK(hipGetLastError());
  // layer 4 (no activation): dY = dpred
  float* g = h->g;
  if (L4.co == 8) CRET((wgrad<32, 8>(h->act[3], h->dpred, g + L4.woff, g + L4.boff, n, H, W, part, s)));
  else CRET((wgrad<32, 16>(h->act[3], h->dpred, g + L4.woff, g + L4.boff, n, H, W, part, s)));
  const u16* dy = h->dpred;
  int nb = 0;
  for (int l = kLayers - 1; l >= 1; --l) {
    // dX of layer l masked by the ReLU of layer l-1: the pre-activation gradient of layer l-1
    u16* dx = h->dbuf[nb];
    nb ^= 1;
    if (l == kLayers - 1) {
      if (L4.cop == 16) CRET((conv<16, 32, 2>(dy, h->wb[l], nullptr, h->act[l], dx, n, H, W, s)));
    } else {
      CRET((conv<32, 32, 2>(dy, h->wb[l], nullptr, h->act[l], dx, n, H, W, s)));
    }
    const Layer& Lp = h->L[l - 1];
    if (Lp.ci == 16) CRET((wgrad<16, 32>(h->act[l - 1], dx, g + Lp.woff, g + Lp.boff, n, H, W, part, s)));
    else CRET((wgrad<32, 32>(h->act[l - 1], dx, g + Lp.woff, g + Lp.boff, n, H, W, part, s)));
    dy = dx;
  }
  CCHK(hipMemcpyAsync(h->h_sq, h->sq, kMseBlocks * sizeof(double), hipMemcpyDeviceToHost, s));
  CCHK(hipStreamSynchronize(s));
  double sum = 0.0;
  for (int b = 0; b < kMseBlocks; ++b) sum += h->h_sq[b];
  if (loss) *loss = sum / ((double)npix * L4.co);
  return 0;
}

int qgn_cnn_backward(qgn_cnn* h, const float* target, double* loss) {
  if (!h || !target) return cfail(QGN_EINVAL, "null argument");
  return backward_impl(h, target, loss);
}

int qgn_cnn_adam(qgn_cnn* h, float lr, float b1, float b2, float eps) {
  if (!h) return cfail(QGN_EINVAL, "null argument");
  h->t += 1;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, h->t)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, h->t)));
  hipLaunchKernelGGL(k_adam, dim3((h->np + 255) / 256), dim3(256), 0, h->stream, h->w, h->g, h->m, h->v, h->np, lr,
                     b1, b2, eps, c1, c2);
  CCHK(hipGetLastError());
  CRET(repack(h, h->stream));
  CCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int qgn_cnn_train_step(qgn_cnn* h, const void* x, const float* target, int n, int H, int W, float* y, float lr,
                       double* loss) {
  if (!h || !x || !target || !y) return cfail(QGN_EINVAL, "null argument");
  CRET(qcnn::forward(h, reinterpret_cast<const u16*>(x), n, H, W, y, h->stream, false));
  CRET(backward_impl(h, target, loss));
  return qgn_cnn_adam(h, lr, 0.9f, 0.999f, 1e-8f);
}

}  // extern "C"
