// Kernels of the learned grid-1 predictor (see cnn.h): implicit-GEMM 3x3
// convolutions on v_mfma_f32_16x16x32_bf16, forward and backward, plus the
// loss, the Adam step, weight packing, input downsampling and the outlier
// gate of the fallback.
//
// MFMA operand layout (16x16x32, D = A * B, A 16x32, B 32x16): lane l feeds
// row l % 16 of A and column l % 16 of B with the eight K values
// 8 * (l / 16) + 0..7, and receives D rows 4 * (l / 16) + 0..3 of column
// l % 16.  Any fixed assignment of K values to (lane group, slot) gives the
// same product as long as A and B use the same one.
#include "cnn.h"

#include <cmath>

namespace pivx {

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(16)));
typedef float f32x4 __attribute__((__vector_size__(16)));
typedef unsigned int u32x4 __attribute__((__vector_size__(16)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (unsigned int)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ bf16x8 pack8(const uint16_t* v) {
  u32x4 u;
  u[0] = v[0] | ((unsigned int)v[1] << 16);
  u[1] = v[2] | ((unsigned int)v[3] << 16);
  u[2] = v[4] | ((unsigned int)v[5] << 16);
  u[3] = v[6] | ((unsigned int)v[7] << 16);
  return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// gradient through the leaky ReLU, from the stored (bf16) layer output
__device__ __forceinline__ uint16_t act_grad(uint16_t g, uint16_t post) {
  return bf2f(post) > 0.f ? g : f2bf(bf2f(g) * kCnnLeak);
}

// Eight K values of the input patch of output pixel (oy, ox): k = tap * CIN + ci.
template <int CIN, int S>
__device__ __forceinline__ bf16x8 patch8(const uint16_t* in, int hin, int win, int oy, int ox, int k) {
  if constexpr (CIN % 8 == 0) {
    if (k >= 9 * CIN) return zero8();
    const int tap = k / CIN, ci = k - tap * CIN;
    const int ky = tap / 3, kx = tap - 3 * ky;
    const int iy = oy * S + ky - 1, ix = ox * S + kx - 1;
    if ((unsigned)iy >= (unsigned)hin || (unsigned)ix >= (unsigned)win) return zero8();
    return ld8(in + ((long long)iy * win + ix) * CIN + ci);
  } else {
    uint16_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int kk = k + i;
      uint16_t x = 0;
      if (kk < 9 * CIN) {
        const int tap = kk / CIN, ci = kk - tap * CIN;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int iy = oy * S + ky - 1, ix = ox * S + kx - 1;
        if ((unsigned)iy < (unsigned)hin && (unsigned)ix < (unsigned)win) x = in[((long long)iy * win + ix) * CIN + ci];
      }
      v[i] = x;
    }
    return pack8(v);
  }
}

// Eight K values (k = tap * COUT + co) of the output-gradient patch that
// reaches input pixel (iy, ix): output pixel o with o * S + tap - 1 = i.
template <int COUT, int S, bool ACT>
__device__ __forceinline__ bf16x8 gpatch8(const uint16_t* gout, const uint16_t* post, int hout, int wout, int iy,
                                          int ix, int k) {
  uint16_t v[8];
  if constexpr (COUT % 8 == 0) {
    if (k >= 9 * COUT) return zero8();
    const int tap = k / COUT, co = k - tap * COUT;
    const int ky = tap / 3, kx = tap - 3 * ky;
    int oy = iy + 1 - ky, ox = ix + 1 - kx;
    if (oy < 0 || ox < 0) return zero8();
    if constexpr (S == 2) {
      if ((oy & 1) || (ox & 1)) return zero8();
      oy >>= 1;
      ox >>= 1;
    }
    if (oy >= hout || ox >= wout) return zero8();
    const long long o = ((long long)oy * wout + ox) * COUT + co;
    if constexpr (!ACT) {
      return ld8(gout + o);
    } else {
      const u32x4 g = *reinterpret_cast<const u32x4*>(gout + o);
      const u32x4 p = *reinterpret_cast<const u32x4*>(post + o);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        v[2 * i] = act_grad((uint16_t)(g[i] & 0xffffu), (uint16_t)(p[i] & 0xffffu));
        v[2 * i + 1] = act_grad((uint16_t)(g[i] >> 16), (uint16_t)(p[i] >> 16));
      }
      return pack8(v);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int kk = k + i;
      uint16_t x = 0;
      if (kk < 9 * COUT) {
        const int tap = kk / COUT, co = kk - tap * COUT;
        const int ky = tap / 3, kx = tap - 3 * ky;
        int oy = iy + 1 - ky, ox = ix + 1 - kx;
        bool ok = oy >= 0 && ox >= 0;
        if (S == 2) {
          ok = ok && !(oy & 1) && !(ox & 1);
          oy >>= 1;
          ox >>= 1;
        }
        ok = ok && oy < hout && ox < wout;
        if (ok) {
          const long long o = ((long long)oy * wout + ox) * COUT + co;
          x = ACT ? act_grad(gout[o], post[o]) : gout[o];
        }
      }
      v[i] = x;
    }
    return pack8(v);
  }
}

// Network input value at downsampled pixel (y, x): the normalised mean of a
// 4x4 block of frame 1 (rows 4-byte aligned) or frame 2 (any alignment).
__device__ __forceinline__ uint16_t down_f1(const uint8_t* f1, long long p1, int y, int x) {
  unsigned int s1 = 0;
  const uint8_t* q = f1 + (long long)y * kCnnDown * p1 + x * kCnnDown;
#pragma unroll
  for (int r = 0; r < kCnnDown; ++r, q += p1)  // the four bytes' sum, one v_sad_u8 against 0
    s1 = __builtin_amdgcn_sad_u8(*reinterpret_cast<const unsigned int*>(q), 0u, s1);
  const float inv = 1.0f / (kCnnDown * kCnnDown);
  return f2bf(((float)s1 * inv - kCnnMu) * kCnnInvSigma);
}
__device__ __forceinline__ uint16_t down_f2(const uint8_t* f2, long long p2, int y, int x) {
  unsigned int s2 = 0;
  const uint8_t* q = f2 + (long long)y * kCnnDown * p2 + x * kCnnDown;
#pragma unroll
  for (int r = 0; r < kCnnDown; ++r, q += p2) s2 += q[0] + q[1] + q[2] + q[3];
  const float inv = 1.0f / (kCnnDown * kCnnDown);
  return f2bf(((float)s2 * inv - kCnnMu) * kCnnInvSigma);
}

// Field vector (vx, vy) of output cell (oy, ox) of the last layer: the FP32
// planes at base + oy * wout + ox, the validity flag (optional) and the
// border-replicated extension of the field (optional, the extend stage).
__device__ __forceinline__ void store_field(float* out_x, float* out_y, uint8_t* out_v, float* ext_x, float* ext_y,
                                            int ext_b, int hout, int wout, long long base, int oy, int ox, float vx,
                                            float vy) {
  const long long o = base + (long long)oy * wout + ox;
  out_x[o] = vx;
  out_y[o] = vy;
  if (out_v) out_v[o] = 1;
  if (ext_x) {
    const int B = ext_b, nxe = wout + 2 * B;
    const int i0 = ox == 0 ? 0 : ox + B, i1 = ox == wout - 1 ? nxe - 1 : ox + B;
    const int j0 = oy == 0 ? 0 : oy + B, j1 = oy == hout - 1 ? hout + 2 * B - 1 : oy + B;
    for (int j = j0; j <= j1; ++j)
      for (int i = i0; i <= i1; ++i) {
        ext_x[j * nxe + i] = vx;
        ext_y[j * nxe + i] = vy;
      }
  }
}

// ---------------------------------------------------------------- forward
// One output tile: MT output-channel tiles (of 16) of PT * 16 consecutive
// output pixels of one row; the COUT / 16 / MT channel blocks of a pixel
// group are consecutive tiles.  D[co][pixel] = W[co][k] * X[k][pixel].  With
// KS = 4 the four waves of a workgroup split the K steps of one tile and sum
// their partial tiles through LDS (short dependent chains for the small,
// deep layers); with KS = 1 each wave owns a tile.  The K loop is unrolled
// by three so that the loads of later steps are in flight while the matrix
// cores work.
template <int CIN, int COUT, int S, bool ACT, bool FINAL, int PT, int MT, int KS>
__global__ void __launch_bounds__(256) k_conv_fwd(ConvFwdArgs a) {
  constexpr int KP = cnn_kpad(CIN);
  constexpr int NCB = cnn_pad16(COUT) / 16 / MT;  // channel blocks per pixel group
  constexpr int KSTEPS = KP / 32;
  constexpr int KPER = (KSTEPS + KS - 1) / KS;   // K steps per wave
  __shared__ f32x4 part[KS > 1 ? KS - 1 : 1][MT * PT][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int gpr = a.wout / (16 * PT);
  const long long gw = KS > 1 ? (long long)blockIdx.x : (long long)blockIdx.x * 4 + wave;
  const int ks = KS > 1 ? wave : 0;
  const long long g = gw / NCB;
  const int cb = (int)(gw - g * NCB);
  if (g >= (long long)a.batch * a.hout * gpr) return;  // whole workgroup (KS > 1) or whole wave
  const int gx = (int)(g % gpr);
  const long long r = g / gpr;
  const int oy = (int)(r % a.hout), b = (int)(r / a.hout);
  const int ox0 = gx * 16 * PT;
  const uint16_t* wbase = a.w + (long long)cb * MT * 16 * KP;
  const uint16_t* in = a.in + (long long)(a.batch_index ? a.batch_index[b] : b) * a.in_bstride;
  f32x4 acc[MT][PT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int pt = 0; pt < PT; ++pt) acc[mt][pt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kbeg = ks * KPER * 32, kend = min(KP, (ks + 1) * KPER * 32);
#pragma unroll 3
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const int k = k0 + 8 * kg;
    bf16x8 wa[MT], xb[PT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) wa[mt] = ld8(wbase + (long long)(mt * 16 + col) * KP + k);
#pragma unroll
    for (int pt = 0; pt < PT; ++pt) xb[pt] = patch8<CIN, S>(in, a.hin, a.win, oy, ox0 + pt * 16 + col, k);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int pt = 0; pt < PT; ++pt) acc[mt][pt] = mfma(wa[mt], xb[pt], acc[mt][pt]);
  }
  if constexpr (KS > 1) {
    // partial tiles of waves 1..KS-1 to wave 0, summed in a fixed order
    if (ks > 0)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int pt = 0; pt < PT; ++pt) part[ks - 1][mt * PT + pt][lane] = acc[mt][pt];
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int pt = 0; pt < PT; ++pt) acc[mt][pt] += part[q][mt * PT + pt][lane];
  }
#pragma unroll
  for (int pt = 0; pt < PT; ++pt) {
    const int ox = ox0 + pt * 16 + col;
    if constexpr (FINAL) {
      if (kg == 0)
        store_field(a.out_x, a.out_y, a.out_v, a.ext_x, a.ext_y, a.ext_b, a.hout, a.wout,
                    (long long)b * a.out_bstride, oy, ox, acc[0][pt][0] + a.bias[0], acc[0][pt][1] + a.bias[1]);
    } else {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const int co0 = (cb * MT + mt) * 16 + 4 * kg;
        if (co0 >= COUT) continue;
        uint16_t h[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = acc[mt][pt][j] + a.bias[co0 + j];
          if (ACT) v = v > 0.f ? v : v * kCnnLeak;
          h[j] = f2bf(v);
        }
        uint2 u;
        u.x = h[0] | ((unsigned int)h[1] << 16);
        u.y = h[2] | ((unsigned int)h[3] << 16);
        *reinterpret_cast<uint2*>(a.out + (((long long)b * a.hout + oy) * a.wout + ox) * COUT + co0) = u;
      }
    }
  }
}


// ---------------------------------------------------------------- forward, fused layers
// For one pair (batch 1) the network runs as two launches: k_conv_head
// (layers 0-2, from the frames or the input plane) and k_conv_tail (layers
// 3-5, to the field).  A workgroup computes a tile of its last layer and
// recomputes the layers before it on the tile plus the halo each 3x3 layer
// needs (cells outside a layer's plane are set to 0, the zero padding of the
// next layer), with every intermediate region in LDS.  Each product runs its
// K steps in the ranges k_conv_fwd splits them into (KS = 4 for the deep
// layers), summed in the same order, and rounds to bf16 at the same points,
// so the field is bit-identical to the layer-by-layer forward pass.

// K steps of a 3x3 layer with CIN input channels and k_conv_fwd's split of
// them over KS waves (KPER steps each, the last range possibly empty).
template <int CIN>
struct KPlan {
  static constexpr int steps = cnn_kpad(CIN) / 32;
  static constexpr int ks = steps >= 8 ? 4 : 1;
  static constexpr int per = (steps + ks - 1) / ks;
};

// LDS elements per cell of a region of C channels (16 bytes of padding).
template <int C>
constexpr int region_pix() { return C + 8; }

// This lane's weight slices of output channels 16 m .. 16 m + 15 (all K
// steps) from the lane-ordered layout: one contiguous 1 KiB load per step.
template <int CIN>
__device__ __forceinline__ void region_weights(const uint16_t* wl, int m, int lane,
                                               bf16x8 (&wreg)[KPlan<CIN>::steps]) {
  const uint16_t* p = wl + ((long long)m * KPlan<CIN>::steps * 64 + lane) * 8;
#pragma unroll
  for (int st = 0; st < KPlan<CIN>::steps; ++st) wreg[st] = ld8(p + st * 64 * 8);
}

// One 16-channel x 16-cell product of a 3x3 layer (CIN a multiple of 8) from
// LDS region src (side RS) to cells p0 .. p0 + 15 of a region of side RD at
// stride S: cell (r, c), tap (ky, kx) reads src cell (S r + ky, S c + kx).
template <int CIN, int RS, int RD, int S>
__device__ __forceinline__ f32x4 region_product(const uint16_t* src, const bf16x8 (&wreg)[KPlan<CIN>::steps], int p0,
                                                int col, int kg) {
  using P = KPlan<CIN>;
  constexpr int PIX = region_pix<CIN>();
  int p = p0 + col;
  p = p < RD * RD ? p : RD * RD - 1;  // cells past the region: any valid cell, not stored
  const int r = p / RD, c = p - (p / RD) * RD;
  f32x4 part[P::ks];
#pragma unroll
  for (int q = 0; q < P::ks; ++q) part[q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < P::per; ++j)
#pragma unroll
    for (int q = 0; q < P::ks; ++q) {
      const int st = q * P::per + j;
      if (st < P::steps) {
        const int k = st * 32 + 8 * kg;
        const int tap = k / CIN, ci = k - tap * CIN;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const bf16x8 xb = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u32x4*>(src + ((S * r + ky) * RS + (S * c + kx)) * PIX + ci));
        part[q] = mfma(wreg[st], xb, part[q]);
      }
    }
  f32x4 acc = part[0];
#pragma unroll
  for (int q = 1; q < P::ks; ++q) acc += part[q];
  return acc;
}

// Bias, leaky ReLU and bf16 of the four channels co0 .. co0 + 3 of region
// cell p (side RD, origin (y0, x0) in a plane of h x w cells; 0 outside it).
template <int COUT, int RD>
__device__ __forceinline__ void region_store(uint16_t* dst, const f32x4& acc, const float (&bb)[4], int p, int co0,
                                             int y0, int x0, int h, int w) {
  if (p >= RD * RD) return;
  const int r = p / RD, c = p - (p / RD) * RD;
  const bool inside = (unsigned)(y0 + r) < (unsigned)h && (unsigned)(x0 + c) < (unsigned)w;
  uint16_t hv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float v = acc[j] + bb[j];
    v = v > 0.f ? v : v * kCnnLeak;
    hv[j] = inside ? f2bf(v) : (uint16_t)0;
  }
  uint2 u;
  u.x = hv[0] | ((unsigned int)hv[1] << 16);
  u.y = hv[2] | ((unsigned int)hv[3] << 16);
  *reinterpret_cast<uint2*>(dst + p * region_pix<COUT>() + co0) = u;
}

// Workgroups of the fused network kernels: eight waves, two per SIMD (a
// handful of workgroups per pair, so the waves of one workgroup are what
// hides the LDS and matrix-core latency).
constexpr int kNetThreads = 512;

// Hidden layer CIN -> 64 from region src (side RS) into region dst (side
// RD, stride S, origin (y0, x0) in its plane of h x w cells): wave w computes
// output channels 16 (w mod 4) .. + 15, with their weights (lane-ordered, w)
// held in registers, for every other block of 16 cells (w / 4 picks which).
template <int CIN, int RS, int RD, int S>
__device__ __forceinline__ void region_hidden64(const uint16_t* src, uint16_t* dst, const uint16_t* w,
                                                const float* bias, int y0, int x0, int h, int wd, int wave, int col,
                                                int kg) {
  const int m = wave & 3;
  bf16x8 wreg[KPlan<CIN>::steps];
  region_weights<CIN>(w, m, kg * 16 + col, wreg);
  const int co0 = m * 16 + 4 * kg;
  float bb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bb[j] = bias[co0 + j];
  for (int p0 = 16 * (wave >> 2); p0 < RD * RD; p0 += 32) {
    const f32x4 acc = region_product<CIN, RS, RD, S>(src, wreg, p0, col, kg);
    region_store<64, RD>(dst, acc, bb, p0 + col, co0, y0, x0, h, wd);
  }
}

constexpr int kTailTile = 8;  // field cells per tile side
constexpr int kHeadTile = 4;  // layer-2 cells per tile side

__global__ void __launch_bounds__(kNetThreads) k_conv_tail(ConvTailArgs a) {
  constexpr int T = kTailTile, RI = T + 6, RA = T + 4, RB = T + 2;
  constexpr int PIX = region_pix<64>();
  // region I (layer-2 output, side RI) and later region B (side RB) share
  // the first buffer; region A (side RA) has its own
  __shared__ __attribute__((aligned(16))) uint16_t bufI[RI * RI * PIX];
  __shared__ __attribute__((aligned(16))) uint16_t bufA[RA * RA * PIX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int tiles_x = (a.w + T - 1) / T;
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
  const int oy0 = ty * T, ox0 = tx * T;
  // stage region I: cells (oy0 - 3 + r, ox0 - 3 + c), 8 channels per thread
  // and step; all of a thread's loads are issued before its first LDS store
  {
    constexpr int N = (RI * RI * 8 + kNetThreads - 1) / kNetThreads;
    u32x4 v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int i = (int)threadIdx.x + kNetThreads * j;
      const int cell = i >> 3, c8 = i & 7;
      const int r = cell / RI, c = cell - (cell / RI) * RI;
      const int y = oy0 - 3 + r, x = ox0 - 3 + c;
      v[j] = u32x4{0u, 0u, 0u, 0u};
      if (i < RI * RI * 8 && (unsigned)y < (unsigned)a.h && (unsigned)x < (unsigned)a.w)
        v[j] = *reinterpret_cast<const u32x4*>(a.in + ((long long)y * a.w + x) * 64 + c8 * 8);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int i = (int)threadIdx.x + kNetThreads * j;
      if (i < RI * RI * 8) *reinterpret_cast<u32x4*>(bufI + (i >> 3) * PIX + (i & 7) * 8) = v[j];
    }
  }
  __syncthreads();
  region_hidden64<64, RI, RA, 1>(bufI, bufA, a.w3, a.b3, oy0 - 2, ox0 - 2, a.h, a.w, wave, col, kg);
  __syncthreads();  // region I is read: its buffer takes region B
  region_hidden64<64, RA, RB, 1>(bufA, bufI, a.w4, a.b4, oy0 - 1, ox0 - 1, a.h, a.w, wave, col, kg);
  __syncthreads();
  // the field: one 16-cell product for each of the first T * T / 16 waves
  if (wave * 16 >= T * T) return;
  bf16x8 wreg[KPlan<64>::steps];
  region_weights<64>(a.w5, 0, lane, wreg);
  const f32x4 acc = region_product<64, RB, T, 1>(bufI, wreg, wave * 16, col, kg);
  const int p = wave * 16 + col;
  const int oy = oy0 + p / T, ox = ox0 + (p - (p / T) * T);
  if (kg == 0 && p < T * T && oy < a.h && ox < a.w)
    store_field(a.out_x, a.out_y, a.out_v, a.ext_x, a.ext_y, a.ext_b, a.h, a.w, 0, oy, ox, acc[0] + a.b5[0],
                acc[1] + a.b5[1]);
}

// Layers 0-2 of one pair: a kHeadTile^2 tile of the layer-2 plane (ph x pw
// cells), layer 1 on 9 x 9 cells, layer 0 on 19 x 19, from the network input
// on 39 x 39 cells (read from the plane `in`, or formed from the frames).
__global__ void __launch_bounds__(kNetThreads) k_conv_head(ConvHeadArgs a) {
  constexpr int T = kHeadTile, R1 = 2 * T + 1, R0 = 2 * R1 + 1, RX = 2 * R0 + 1;
  constexpr int PIX0 = region_pix<32>(), PIX1 = region_pix<64>();
  __shared__ __attribute__((aligned(16))) uint16_t bufX[RX * RX * 2];  // input cells, 2 channels
  __shared__ __attribute__((aligned(16))) uint16_t buf0[R0 * R0 * PIX0];
  __shared__ __attribute__((aligned(16))) uint16_t buf1[R1 * R1 * PIX1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int h0 = 2 * a.ph * 2, w0 = 2 * a.pw * 2;  // layer-0 plane (input / 2)
  const int h1 = 2 * a.ph, w1 = 2 * a.pw;          // layer-1 plane
  const int hx = 2 * h0, wx = 2 * w0;              // network input plane
  const int tiles_x = (a.pw + T - 1) / T;
  const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - (blockIdx.x / tiles_x) * tiles_x;
  const int oy0 = ty * T, ox0 = tx * T;                  // layer-2 origin
  const int y1 = 2 * oy0 - 1, x1 = 2 * ox0 - 1;          // layer-1 region origin
  const int y0 = 2 * y1 - 1, x0 = 2 * x1 - 1;            // layer-0 region origin
  const int yx = 2 * y0 - 1, xx = 2 * x0 - 1;            // input region origin
  const bool f2_aligned = ((a.p2 & 3) | (reinterpret_cast<uintptr_t>(a.f2) & 3)) == 0;
  // stage the input region (0 outside the plane); all of a thread's loads
  // are issued before its first LDS store
  {
    constexpr int N = (RX * RX + kNetThreads - 1) / kNetThreads;
    unsigned int v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int cell = (int)threadIdx.x + kNetThreads * j;
      const int r = cell / RX, c = cell - (cell / RX) * RX;
      const int y = yx + r, x = xx + c;
      v[j] = 0u;
      if (cell < RX * RX && (unsigned)y < (unsigned)hx && (unsigned)x < (unsigned)wx) {
        if (a.f1) {
          // frame 2 rows 4-byte aligned (the engine's bordered frame): 4-byte loads as for frame 1
          const uint16_t b = f2_aligned ? down_f1(a.f2, a.p2, y, x) : down_f2(a.f2, a.p2, y, x);
          v[j] = down_f1(a.f1, a.p1, y, x) | ((unsigned int)b << 16);
        } else {
          v[j] = *reinterpret_cast<const unsigned int*>(a.in + ((long long)y * wx + x) * 2);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int cell = (int)threadIdx.x + kNetThreads * j;
      if (cell < RX * RX) *reinterpret_cast<unsigned int*>(bufX + cell * 2) = v[j];
    }
  }
  __syncthreads();
  // layer 0 (2 -> 32, stride 2, one K step): one cell block per task, both
  // channel blocks from one operand.  A lane's 8 K values are both channels
  // of the input cells at taps 4 kg .. 4 kg + 3, i.e. four 32-bit words of
  // bufX (taps past 8 are the K padding, 0).
  {
    constexpr int NB = (R0 * R0 + 15) / 16;
    // weights and biases of both channel blocks, loaded once
    const bf16x8 wa0 = ld8(a.w0 + lane * 8);
    const bf16x8 wa1 = ld8(a.w0 + (64 + lane) * 8);
    float bb0[4], bb1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bb0[j] = a.b0[4 * kg + j];
      bb1[j] = a.b0[16 + 4 * kg + j];
    }
    int toff[4];
    bool tok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int tap = 4 * kg + j;
      tok[j] = tap < 9;
      toff[j] = tok[j] ? (tap / 3) * RX + tap % 3 : 0;
    }
    const uint32_t* const bx = reinterpret_cast<const uint32_t*>(bufX);
    for (int nb = wave; nb < NB; nb += kNetThreads / 64) {
      int p = nb * 16 + col;
      p = p < R0 * R0 ? p : R0 * R0 - 1;
      const int r = p / R0, c = p - (p / R0) * R0;
      const int base = 2 * r * RX + 2 * c;
      u32x4 xw;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t v = bx[base + toff[j]];
        xw[j] = tok[j] ? v : 0u;
      }
      const bf16x8 xb = __builtin_bit_cast(bf16x8, xw);
      const f32x4 acc0 = mfma(wa0, xb, f32x4{0.f, 0.f, 0.f, 0.f});
      const f32x4 acc1 = mfma(wa1, xb, f32x4{0.f, 0.f, 0.f, 0.f});
      region_store<32, R0>(buf0, acc0, bb0, nb * 16 + col, 4 * kg, y0, x0, h0, w0);
      region_store<32, R0>(buf0, acc1, bb1, nb * 16 + col, 16 + 4 * kg, y0, x0, h0, w0);
    }
  }
  __syncthreads();
  region_hidden64<32, R0, R1, 2>(buf0, buf1, a.w1, a.b1, y1, x1, h1, w1, wave, col, kg);
  __syncthreads();
  // layer 2 (64 -> 64, stride 2): one 16-cell product for each of waves 0-3
  // (channel block = wave), to the plane
  if (wave >= 4) return;
  bf16x8 wreg[KPlan<64>::steps];
  region_weights<64>(a.w2, wave, lane, wreg);
  const f32x4 acc = region_product<64, R1, T, 2>(buf1, wreg, 0, col, kg);
  const int co0 = wave * 16 + 4 * kg;
  const int oy = oy0 + col / T, ox = ox0 + (col - (col / T) * T);
  if (oy < a.ph && ox < a.pw) {
    uint16_t hv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float v = acc[j] + a.b2[co0 + j];
      v = v > 0.f ? v : v * kCnnLeak;
      hv[j] = f2bf(v);
    }
    uint2 u;
    u.x = hv[0] | ((unsigned int)hv[1] << 16);
    u.y = hv[2] | ((unsigned int)hv[3] << 16);
    *reinterpret_cast<uint2*>(a.out + ((long long)oy * a.pw + ox) * 64 + co0) = u;
  }
}

// ---------------------------------------------------------------- backward, data
// D[ci][input pixel] = WT[ci][k] * G[k][input pixel], k = tap * COUT + co.
template <int CIN, int COUT, int S, bool ACT, int PT>
__global__ void __launch_bounds__(256) k_conv_bwd_data(ConvBwdDataArgs a) {
  constexpr int KP = cnn_kpad(COUT);
  constexpr int MT = cnn_pad16(CIN) / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int gpr = a.win / (16 * PT);
  const long long g = (long long)blockIdx.x * 4 + wave;
  if (g >= (long long)a.batch * a.hin * gpr) return;
  const int gx = (int)(g % gpr);
  const long long r = g / gpr;
  const int iy = (int)(r % a.hin), b = (int)(r / a.hin);
  // pixel of lane column c in tile pt: at stride 1 the run ix0 + pt * 16 + c;
  // at stride 2 the wave takes the 16 pixels of one parity of a 32-pixel
  // span, ix0 + 2 (pt * 16 + c) + par, so that a tap reaches either all or
  // none of them and the taps that reach none are skipped (their K steps
  // would add exact zeros)
  const int par = S == 2 ? gx & 1 : 0;
  const int ix0 = S == 2 ? (gx >> 1) * 32 * PT + par : gx * 16 * PT;
  auto pix = [&](int pt, int c) { return S == 2 ? ix0 + 2 * (pt * 16 + c) : ix0 + pt * 16 + c; };
  const long long so = (long long)b * a.hout * a.wout * COUT;
  const uint16_t* gout = a.gout + so;
  const uint16_t* post = ACT ? a.post + so : nullptr;
  f32x4 acc[MT][PT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int pt = 0; pt < PT; ++pt) acc[mt][pt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < KP; k0 += 32) {
    if constexpr (S == 2 && COUT % 32 == 0) {
      // the step's tap (COUT a multiple of 32: one tap per step) reaches this
      // wave's row and column parity, or the step is all zeros
      const int tap = k0 / COUT, ky = tap / 3, kx = tap - 3 * (tap / 3);
      if (((iy + 1 - ky) & 1) || ((par + 1 - kx) & 1)) continue;
    }
    const int k = k0 + 8 * kg;
    bf16x8 wa[MT], xb[PT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) wa[mt] = ld8(a.wt + (long long)(mt * 16 + col) * KP + k);
#pragma unroll
    for (int pt = 0; pt < PT; ++pt) xb[pt] = gpatch8<COUT, S, ACT>(gout, post, a.hout, a.wout, iy, pix(pt, col), k);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int pt = 0; pt < PT; ++pt) acc[mt][pt] = mfma(wa[mt], xb[pt], acc[mt][pt]);
  }
#pragma unroll
  for (int pt = 0; pt < PT; ++pt) {
    const int ix = pix(pt, col);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int ci0 = mt * 16 + 4 * kg;
      if (ci0 >= CIN) continue;
      uint2 u;
      u.x = f2bf(acc[mt][pt][0]) | ((unsigned int)f2bf(acc[mt][pt][1]) << 16);
      u.y = f2bf(acc[mt][pt][2]) | ((unsigned int)f2bf(acc[mt][pt][3]) << 16);
      *reinterpret_cast<uint2*>(a.gin + (((long long)b * a.hin + iy) * a.win + ix) * CIN + ci0) = u;
    }
  }
}

// ---------------------------------------------------------------- backward, weights
// D[co][n] = Gpre[co][o] * X[o][n], n = tap * CIN + ci, K = the output pixels
// o of one chunk (linear over [batch][hout][wout]).  Both operands are
// gathered one pixel at a time (a lane's eight K values are eight pixels).
// Grid: (chunks, groups of NTG column tiles, groups of MTG row tiles + 1),
// one wave per workgroup; the last z slice sums the bias gradient.
template <int CIN, int COUT, int S, bool ACT, int NTG, int MTG>
__global__ void __launch_bounds__(64) k_conv_bwd_w(ConvBwdWArgs a) {
  constexpr int KPI = cnn_kpad(CIN);
  const int mt0 = blockIdx.z * MTG;  // this workgroup's row tiles mt0 .. mt0 + MTG - 1
  const long long npx = (long long)a.batch * a.hout * a.wout;
  if (blockIdx.z == gridDim.z - 1) {
    // the last z slice: the bias partial sums of the chunk (the same bf16
    // gradient values), one lane per channel, in pixel order; the loads of a
    // batch of pixels are issued before its sums
    if (blockIdx.y != 0 || threadIdx.x >= COUT) return;
    const long long ob = (long long)blockIdx.x * a.chunk_px;
    const long long oe = ob + a.chunk_px < npx ? ob + a.chunk_px : npx;
    constexpr int B = 64;
    float s = 0.f;
    for (long long o0 = ob; o0 < oe; o0 += B) {
      float v[B];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const long long q = ((o0 + j < oe) ? o0 + j : ob) * COUT + threadIdx.x;
        v[j] = bf2f(ACT ? act_grad(a.gout[q], a.post[q]) : a.gout[q]);
      }
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (o0 + j < oe) s += v[j];
    }
    a.partial_b[(long long)blockIdx.x * COUT + threadIdx.x] = s;
    return;
  }
  const int lane = threadIdx.x, col = lane & 15, kg = lane >> 4;
  const long long ob = (long long)blockIdx.x * a.chunk_px;
  const long long oe = ob + a.chunk_px < npx ? ob + a.chunk_px : npx;
  const int nt0 = blockIdx.y * NTG;
  f32x4 acc[MTG][NTG];
#pragma unroll
  for (int mt = 0; mt < MTG; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTG; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int hw = a.hout * a.wout;
  // this lane's columns n = tap * CIN + ci of the NTG tiles
  int ntap_ky[NTG], ntap_kx[NTG], nci[NTG];
  bool nok[NTG];
#pragma unroll
  for (int nt = 0; nt < NTG; ++nt) {
    const int n = (nt0 + nt) * 16 + col;
    const int tap = n / CIN;
    nci[nt] = n - tap * CIN;
    ntap_ky[nt] = tap / 3;
    ntap_kx[nt] = tap - 3 * (tap / 3);
    nok[nt] = n < 9 * CIN;
  }
  // one 32-pixel step per iteration (two or four steps gathered ahead of
  // their products were measured slower)
  for (long long o0 = ob; o0 < oe; o0 += 32) {
    const long long ok = o0 + 8 * kg;
    // (sample, row, column) of the lane's 8 consecutive output pixels: one
    // division for the first, then steps along the row
    int pb[8], py[8], px[8];
    {
      const int o = (int)ok;  // the host keeps batch x hout x wout below 2^31
      int b = o / hw;
      const int rem = o - b * hw;
      int oy = rem / a.wout, ox = rem - (rem / a.wout) * a.wout;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        pb[i] = b;
        py[i] = oy;
        px[i] = ox;
        if (++ox == a.wout) {
          ox = 0;
          if (++oy == a.hout) {
            oy = 0;
            ++b;
          }
        }
      }
    }
    // every gather load is unconditional (from a valid address), the value
    // selected afterwards, so that all of a step's loads are in flight at once
    bf16x8 ga[MTG], xb[NTG];
#pragma unroll
    for (int mt = 0; mt < MTG; ++mt) {
      const int co = (mt0 + mt) * 16 + col;
      uint16_t v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const long long o = ok + i;
        const bool ov = o < oe && co < COUT;
        const long long q = (ov ? o : ob) * COUT + (co < COUT ? co : 0);
        const uint16_t g = a.gout[q];
        const uint16_t x = ACT ? act_grad(g, a.post[q]) : g;
        v[i] = ov ? x : (uint16_t)0;
      }
      ga[mt] = pack8(v);
    }
    long long sb[8];  // sample offsets of the 8 pixels
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int b = pb[i] < a.batch ? pb[i] : a.batch - 1;
      sb[i] = (long long)(a.batch_index ? a.batch_index[b] : b) * a.in_bstride;
    }
#pragma unroll
    for (int nt = 0; nt < NTG; ++nt) {
      uint16_t v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int iy = py[i] * S + ntap_ky[nt] - 1, ix = px[i] * S + ntap_kx[nt] - 1;
        const bool iv = ok + i < oe && nok[nt] && (unsigned)iy < (unsigned)a.hin && (unsigned)ix < (unsigned)a.win;
        const uint16_t x = a.in[sb[i] + ((long long)(iv ? iy : 0) * a.win + (iv ? ix : 0)) * CIN + (iv ? nci[nt] : 0)];
        v[i] = iv ? x : (uint16_t)0;
      }
      xb[nt] = pack8(v);
    }
#pragma unroll
    for (int mt = 0; mt < MTG; ++mt)
#pragma unroll
      for (int nt = 0; nt < NTG; ++nt) acc[mt][nt] = mfma(ga[mt], xb[nt], acc[mt][nt]);
  }
  float* part = a.partial + (long long)blockIdx.x * cnn_pad16(COUT) * KPI;
#pragma unroll
  for (int mt = 0; mt < MTG; ++mt)
#pragma unroll
    for (int nt = 0; nt < NTG; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = (mt0 + mt) * 16 + 4 * kg + j, n = (nt0 + nt) * 16 + col;
        part[(long long)co * KPI + n] = acc[mt][nt][j];
      }
}

// Layer 0 (2 -> 32, stride 2) of k_conv_bwd_w, the same products and bias
// sums in the same order from operands staged in LDS: each block of 256
// output pixels of a chunk is read once with row-contiguous loads, the
// activation gradient as [channel][pixel] and the input patches as
// [k][pixel] (k = tap * 2 + ci, rows 18-31 the zero padding), and every MFMA
// operand is one 16-byte LDS read.  Waves 0-3 compute output tile
// (w / 2, w mod 2); wave 4 adds the staged gradients of each channel to the
// bias sum in pixel order.  (The gathering form issued about 24,000 VALU
// instructions per wave for 43 MFMAs; its bias sums, read pixel by pixel from
// global memory, took 386 of its 437 us.)
template <bool ACT>
__global__ void __launch_bounds__(320) k_conv_bwd_w0(ConvBwdWArgs a) {
  constexpr int COUT = 32, KPI = cnn_kpad(2), PB = 256, LP = PB + 8;
  __shared__ __attribute__((aligned(16))) uint16_t gA[COUT][LP];  // activation gradient [co][pixel]
  __shared__ __attribute__((aligned(16))) uint16_t gB[32][LP];    // input patches [k][pixel]
  const long long npx = (long long)a.batch * a.hout * a.wout;
  const long long ob = (long long)blockIdx.x * a.chunk_px;
  const long long oe = ob + a.chunk_px < npx ? ob + a.chunk_px : npx;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int col = lane & 15, kg = lane >> 4, mt = (wave >> 1) & 1, nt = wave & 1;
  if (t < PB)
    for (int k = 18; k < 32; ++k) gB[k][t] = 0;
  const int hw = a.hout * a.wout;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (long long p0 = ob; p0 < oe; p0 += PB) {
    __syncthreads();  // the previous block's operands are read
    const long long o = p0 + t;
    const bool ov = o < oe;
    if (t < PB) {
      // pixel t: 32 channels of the gradient, four 16-byte loads of each plane
      const long long q = (ov ? o : ob) * COUT;
      u32x4 g[4], pp[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        g[j] = *reinterpret_cast<const u32x4*>(a.gout + q + 8 * j);
        if (ACT) pp[j] = *reinterpret_cast<const u32x4*>(a.post + q + 8 * j);
      }
      // pixel t: its 3x3 patch of both input channels (0 outside the plane)
      const int oi = (int)(ov ? o : ob);
      const int b = oi / hw, rem = oi - b * hw;
      const int oy = rem / a.wout, ox = rem - (rem / a.wout) * a.wout;
      const uint32_t* in = reinterpret_cast<const uint32_t*>(
          a.in + (long long)(a.batch_index ? a.batch_index[b] : b) * a.in_bstride);
      uint32_t w[9];
#pragma unroll
      for (int ky = 0; ky < 3; ++ky)
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
          const int iy = 2 * oy + ky - 1, ix = 2 * ox + kx - 1;
          const bool iv = ov && (unsigned)iy < (unsigned)a.hin && (unsigned)ix < (unsigned)a.win;
          const uint32_t v = in[(long long)(iv ? iy : 0) * a.win + (iv ? ix : 0)];  // both channels of the cell
          w[3 * ky + kx] = iv ? v : 0u;
        }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint16_t gv = (uint16_t)(g[j][e / 2] >> (16 * (e & 1)));
          const uint16_t x = ACT ? act_grad(gv, (uint16_t)(pp[j][e / 2] >> (16 * (e & 1)))) : gv;
          gA[8 * j + e][t] = ov ? x : (uint16_t)0;
        }
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        gB[2 * tap][t] = (uint16_t)(w[tap] & 0xffffu);
        gB[2 * tap + 1][t] = (uint16_t)(w[tap] >> 16);
      }
    }
    __syncthreads();
    if (wave < 4) {
#pragma unroll
      for (int st = 0; st < PB / 32; ++st) {
        if (p0 + 32 * st >= oe) break;
        const bf16x8 ga =
            __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&gA[mt * 16 + col][32 * st + 8 * kg]));
        const bf16x8 xb =
            __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&gB[nt * 16 + col][32 * st + 8 * kg]));
        acc = mfma(ga, xb, acc);
      }
    } else if (lane < COUT) {
      // bias: channel `lane`, the block's pixels in order
      const int n = oe - p0 < PB ? (int)(oe - p0) : PB;
      for (int p8 = 0; p8 < n; p8 += 8) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(&gA[lane][p8]);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (p8 + e < n) bsum += bf2f((uint16_t)(v[e / 2] >> (16 * (e & 1))));
      }
    }
  }
  if (wave < 4) {
    float* part = a.partial + (long long)blockIdx.x * COUT * KPI;
#pragma unroll
    for (int j = 0; j < 4; ++j) part[(long long)(mt * 16 + 4 * kg + j) * KPI + nt * 16 + col] = acc[j];
  } else if (lane < COUT) {
    a.partial_b[(long long)blockIdx.x * COUT + lane] = bsum;
  }
}

// The hidden layers (COUT = 64, CIN = 32 or 64) of k_conv_bwd_w from operands
// staged in LDS, as k_conv_bwd_w0 stages layer 0: blocks of 64 output
// pixels of one output row (the host checks wout and the chunk size), the
// activation gradient as [channel][pixel] and this workgroup's 288 K columns
// of the input patches (n = tap * CIN + ci, from n0 = 288 blockIdx.y) as
// [n][pixel], both read with 16-byte loads.  Wave w of four computes the 18
// output tiles of channel block w; wave 4 adds the staged gradients to the
// bias sums (workgroups with blockIdx.y = 0).  Same products and sums, in the
// same order, as k_conv_bwd_w.
template <int CIN, int S, bool ACT>
__global__ void __launch_bounds__(320) k_conv_bwd_wh(ConvBwdWArgs a) {
  constexpr int COUT = 64, KPI = cnn_kpad(CIN), PB = 64, LP = PB + 8, NG = 288, NT = NG / 16;
  static_assert(9 * CIN % NG == 0 && KPI == 9 * CIN, "K columns in groups of 288");
  __shared__ __attribute__((aligned(16))) uint16_t gA[COUT][LP];  // activation gradient [co][pixel]
  __shared__ __attribute__((aligned(16))) uint16_t gB[NG][LP];    // input patches [n - n0][pixel]
  const long long npx = (long long)a.batch * a.hout * a.wout;
  const long long ob = (long long)blockIdx.x * a.chunk_px;
  const long long oe = ob + a.chunk_px < npx ? ob + a.chunk_px : npx;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const int n0 = blockIdx.y * NG;
  const int hw = a.hout * a.wout;
  f32x4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (long long p0 = ob; p0 < oe; p0 += PB) {
    __syncthreads();  // the previous block's operands are read
    if (t < 256) {
      // the block's pixels p0 .. p0 + 63 lie in one output row
      const int o0 = (int)p0;  // the host keeps batch x hout x wout below 2^31
      const int b = o0 / hw, rem = o0 - b * hw;
      const int oy = rem / a.wout, ox0 = rem - (rem / a.wout) * a.wout;
      {
        // gradient: pixel t / 4, channels 16 (t mod 4) .. + 15
        const int p = t >> 2, cg = (t & 3) * 16;
        const long long o = p0 + p;
        const bool ov = o < oe;
        const long long q = (ov ? o : ob) * COUT + cg;
        u32x4 g[2], pp[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          g[j] = *reinterpret_cast<const u32x4*>(a.gout + q + 8 * j);
          if (ACT) pp[j] = *reinterpret_cast<const u32x4*>(a.post + q + 8 * j);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const uint16_t gv = (uint16_t)(g[j][e / 2] >> (16 * (e & 1)));
            const uint16_t x = ACT ? act_grad(gv, (uint16_t)(pp[j][e / 2] >> (16 * (e & 1)))) : gv;
            gA[cg + 8 * j + e][p] = ov ? x : (uint16_t)0;
          }
      }
      // patches: 36 chunks of 8 K columns per pixel, consecutive threads
      // taking consecutive chunks of one pixel
      const uint16_t* in = a.in + (long long)(a.batch_index ? a.batch_index[b] : b) * a.in_bstride;
      for (int i = t; i < PB * (NG / 8); i += 256) {
        const int p = i / (NG / 8), ch = i - p * (NG / 8);
        const int n = n0 + 8 * ch, tap = n / CIN, ci = n - tap * CIN;
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        const int iy = oy * S + ky - 1, ix = (ox0 + p) * S + kx - 1;
        const bool iv = p0 + p < oe && (unsigned)iy < (unsigned)a.hin && (unsigned)ix < (unsigned)a.win;
        const u32x4 v = *reinterpret_cast<const u32x4*>(in + ((long long)(iv ? iy : 0) * a.win + (iv ? ix : 0)) * CIN +
                                                         (iv ? ci : 0));
#pragma unroll
        for (int e = 0; e < 8; ++e) gB[8 * ch + e][p] = iv ? (uint16_t)(v[e / 2] >> (16 * (e & 1))) : (uint16_t)0;
      }
    }
    __syncthreads();
    if (wave < 4) {
#pragma unroll
      for (int st = 0; st < PB / 32; ++st) {
        if (p0 + 32 * st >= oe) break;
        const bf16x8 ga =
            __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&gA[wave * 16 + col][32 * st + 8 * kg]));
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const bf16x8 xb =
              __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&gB[nt * 16 + col][32 * st + 8 * kg]));
          acc[nt] = mfma(ga, xb, acc[nt]);
        }
      }
    } else if (blockIdx.y == 0) {
      // bias: channel `lane`, the block's pixels in order
      const int n = oe - p0 < PB ? (int)(oe - p0) : PB;
      for (int p8 = 0; p8 < n; p8 += 8) {
        const u32x4 v = *reinterpret_cast<const u32x4*>(&gA[lane][p8]);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (p8 + e < n) bsum += bf2f((uint16_t)(v[e / 2] >> (16 * (e & 1))));
      }
    }
  }
  if (wave < 4) {
    float* part = a.partial + (long long)blockIdx.x * COUT * KPI;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) part[(long long)(wave * 16 + 4 * kg + j) * KPI + n0 + nt * 16 + col] = acc[nt][j];
  } else if (blockIdx.y == 0) {
    a.partial_b[(long long)blockIdx.x * COUT + lane] = bsum;
  }
}

// grad_w [cout][9][cin] and grad_b [cout] of every layer from the partial
// sums, one launch for all layers (each layer's partial sums in a buffer of
// their own): thread t takes element t - first of the layer whose range holds
// it and sums it in chunk order, the loads of a batch of chunks issued first.
__global__ void k_wgrad_reduce_all(WgradReduceAll r) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int l = 0;
  while (l + 1 < kCnnLayers && t >= r.layer[l + 1].first) ++l;
  const WgradReduceLayer& L = r.layer[l];
  const int e = t - L.first;
  const int nw = L.cout * 9 * L.cin;
  if (e < 0 || e >= nw + L.cout) return;
  auto chunk_sum = [&](const float* p, long long stride) {
    constexpr int B = 16;
    float s = 0.f;
    for (int c0 = 0; c0 < L.nchunks; c0 += B) {
      float v[B];
#pragma unroll
      for (int j = 0; j < B; ++j) v[j] = c0 + j < L.nchunks ? p[(c0 + j) * stride] : 0.f;
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (c0 + j < L.nchunks) s += v[j];
    }
    return s;
  };
  if (e < nw) {
    const int co = e / (9 * L.cin), n = e - co * 9 * L.cin;  // n = tap * cin + ci
    L.grad[e] = chunk_sum(L.partial + (long long)co * L.kpi + n, (long long)L.mrows * L.kpi);
  } else {
    const int co = e - nw;
    L.grad[e] = chunk_sum(L.partial_b + co, L.cout);
  }
}

// ---------------------------------------------------------------- loss, optimizer
constexpr int kLossBlock = 256;

__global__ void __launch_bounds__(kLossBlock) k_mse(const float* out, long long out_bstride, const float* labels,
                                                    const int* batch_index, int batch, int hw, uint16_t* gout,
                                                    float* loss_part) {
  __shared__ float red[kLossBlock];
  const long long n = (long long)batch * 2 * hw;
  const long long t = (long long)blockIdx.x * kLossBlock + threadIdx.x;
  const float inv = 2.0f / (float)n;
  float s = 0.f;
  if (t < n) {
    const int b = (int)(t / (2 * hw));
    const int rem = (int)(t - (long long)b * 2 * hw);
    const int c = rem / hw, p = rem - c * hw;
    const float o = out[(long long)b * out_bstride + (long long)c * hw + p];
    const long long li = (long long)(batch_index ? batch_index[b] : b) * 2 * hw + (long long)c * hw + p;
    const float d = o - labels[li];
    gout[((long long)b * hw + p) * 2 + c] = f2bf(d * inv);
    s = d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kLossBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_loss_final(const float* part, int n, double count, float* loss) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = (float)(red[0] / count);
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

__global__ void k_pack_layer(const float* w, int cin, int cout, uint16_t* wf, uint16_t* wt, uint16_t* wl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cout * 9 * cin) return;
  const int co = t / (9 * cin), r = t - co * 9 * cin, tap = r / cin, ci = r - tap * cin;
  const uint16_t h = f2bf(w[t]);
  wf[(long long)co * cnn_kpad(cin) + tap * cin + ci] = h;
  wt[(long long)ci * cnn_kpad(cout) + tap * cout + co] = h;
  wl[cnn_lane_index(co, r, cin)] = h;
}


// ---------------------------------------------------------------- input, gate
__global__ void k_downsample(const uint8_t* f1, long long p1, const uint8_t* f2, long long p2, int w, int h,
                             uint16_t* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= w * h) return;
  const int y = t / w, x = t - (t / w) * w;
  const uint16_t a = down_f1(f1, p1, y, x), b = down_f2(f2, p2, y, x);
  reinterpret_cast<unsigned int*>(out)[t] = a | ((unsigned int)b << 16);
}

__global__ void __launch_bounds__(256) k_gate(const int8_t* accepted, int n, int max_outliers, int* gate, int* count) {
  __shared__ int red[256];
  int c = 0;
  for (int i = threadIdx.x; i < n; i += 256) c += accepted[i] < 0;
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *count = red[0];
    *gate = red[0] > max_outliers ? 1 : 0;
  }
}

__global__ void k_fill_u8(uint8_t* p, int n, uint8_t v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

inline int blocks(long long n, int b) { return (int)((n + b - 1) / b); }

template <int CIN, int COUT, int S, bool ACT, bool FINAL>
hipError_t fwd(const ConvFwdArgs& a, hipStream_t s) {
  constexpr int PT = 1;
  constexpr int MT = cnn_pad16(COUT) / 16;  // every channel block of a pixel group in one wave (one input operand per K step)
  constexpr int NCB = cnn_pad16(COUT) / 16 / MT;
  constexpr int KS = cnn_kpad(CIN) / 32 >= 8 ? 4 : 1;  // split the K steps of the deep layers over 4 waves
  if (a.wout % (16 * PT) != 0) return hipErrorInvalidValue;
  const long long tiles = (long long)a.batch * a.hout * (a.wout / (16 * PT)) * NCB;
  const long long wgs = KS > 1 ? tiles : blocks(tiles, 4);
  hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, S, ACT, FINAL, PT, MT, KS>), dim3((unsigned)wgs), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int CIN, int COUT, int S, bool ACT, int PT>
hipError_t bwd_data_pt(const ConvBwdDataArgs& a, hipStream_t s) {
  if (a.win % (16 * PT * S) != 0) return hipErrorInvalidValue;
  const long long groups = (long long)a.batch * a.hin * (a.win / (16 * PT));
  hipLaunchKernelGGL((k_conv_bwd_data<CIN, COUT, S, ACT, PT>), dim3(blocks(groups, 4)), dim3(256), 0, s, a);
  return hipGetLastError();
}

// Stride 2 skips the taps that reach none of a wave's pixels, which leaves
// the waves short: up to four 16-pixel tiles per wave where the row allows.
template <int CIN, int COUT, int S, bool ACT>
hipError_t bwd_data(const ConvBwdDataArgs& a, hipStream_t s) {
  if constexpr (S == 2) {
    if (a.win % (16 * 4 * S) == 0) return bwd_data_pt<CIN, COUT, S, ACT, 4>(a, s);
    if (a.win % (16 * 2 * S) == 0) return bwd_data_pt<CIN, COUT, S, ACT, 2>(a, s);
  }
  return bwd_data_pt<CIN, COUT, S, ACT, 1>(a, s);
}

template <int CIN, int COUT, int S, bool ACT>
hipError_t bwd_w(const ConvBwdWArgs& a, int nchunks, hipStream_t s) {
  constexpr int MT = cnn_pad16(COUT) / 16;
  constexpr int ntiles = cnn_kpad(CIN) / 16;
  // a small layer (layer 0: 2 x 2 output tiles) takes one tile per
  // workgroup, so that its launch has enough waves to hide the gathers'
  // latency; every tile sums its pixels in the same order either way
  constexpr bool split = MT * ntiles <= 4;
  constexpr int NTG = split ? 1 : 2, MTG = split ? 1 : MT;
  static_assert(ntiles % NTG == 0, "column tiles per group");
  if (a.chunk_px % 32 != 0 || (long long)a.batch * a.hout * a.wout >= (1LL << 31) - 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_conv_bwd_w<CIN, COUT, S, ACT, NTG, MTG>), dim3(nchunks, ntiles / NTG, MT / MTG + 1), dim3(64), 0,
                     s, a);
  return hipGetLastError();
}

}  // namespace

int cnn_param_count() {
  int n = 0;
  for (const CnnLayer& L : kCnnArch) n += L.cout * 9 * L.cin + L.cout;
  return n;
}

int cnn_weight_offset(int layer) {
  int n = 0;
  for (int l = 0; l < layer; ++l) n += kCnnArch[l].cout * 9 * kCnnArch[l].cin + kCnnArch[l].cout;
  return n;
}

hipError_t launch_conv_fwd(int layer, const ConvFwdArgs& a, hipStream_t s) {
  switch (layer) {
    case 0: return fwd<2, 32, 2, true, false>(a, s);
    case 1: return fwd<32, 64, 2, true, false>(a, s);
    case 2: return fwd<64, 64, 2, true, false>(a, s);
    case 3: return fwd<64, 64, 1, true, false>(a, s);
    case 4: return fwd<64, 64, 1, true, false>(a, s);
    case 5: return fwd<64, 2, 1, false, true>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_conv_head(const ConvHeadArgs& a, hipStream_t s) {
  if (a.ph < 1 || a.pw < 1 || (!a.f1 && !a.in) || (a.f1 && (!a.f2 || (a.p1 & 3) != 0 ||
                                                          (reinterpret_cast<uintptr_t>(a.f1) & 3) != 0)))
    return hipErrorInvalidValue;
  const int tiles = ((a.ph + kHeadTile - 1) / kHeadTile) * ((a.pw + kHeadTile - 1) / kHeadTile);
  hipLaunchKernelGGL(k_conv_head, dim3(tiles), dim3(kNetThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_conv_tail(const ConvTailArgs& a, hipStream_t s) {
  if (a.h < 1 || a.w < 1) return hipErrorInvalidValue;
  const int tiles = ((a.h + kTailTile - 1) / kTailTile) * ((a.w + kTailTile - 1) / kTailTile);
  hipLaunchKernelGGL(k_conv_tail, dim3(tiles), dim3(kNetThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_conv_bwd_data(int layer, const ConvBwdDataArgs& a, hipStream_t s) {
  switch (layer) {
    case 1: return bwd_data<32, 64, 2, true>(a, s);
    case 2: return bwd_data<64, 64, 2, true>(a, s);
    case 3: return bwd_data<64, 64, 1, true>(a, s);
    case 4: return bwd_data<64, 64, 1, true>(a, s);
    case 5: return bwd_data<64, 2, 1, false>(a, s);
    default: return hipErrorInvalidValue;  // no gradient is needed for the network input
  }
}

// The staged hidden-layer kernel takes blocks of 64 pixels of one output row.
bool staged_wh(const ConvBwdWArgs& a) {
  return a.wout % 64 == 0 && a.chunk_px % 64 == 0 && (long long)a.batch * a.hout * a.wout < (1LL << 31) - 64;
}

template <int CIN, int S>
hipError_t bwd_wh(const ConvBwdWArgs& a, int nchunks, hipStream_t s) {
  hipLaunchKernelGGL((k_conv_bwd_wh<CIN, S, true>), dim3(nchunks, 9 * CIN / 288), dim3(320), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_conv_bwd_w(int layer, const ConvBwdWArgs& a, int nchunks, hipStream_t s) {
  switch (layer) {
    case 0:
      if (a.chunk_px % 32 != 0 || (long long)a.batch * a.hout * a.wout >= (1LL << 31) - 256) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_conv_bwd_w0<true>, dim3(nchunks), dim3(320), 0, s, a);
      return hipGetLastError();
    case 1: return staged_wh(a) ? bwd_wh<32, 2>(a, nchunks, s) : bwd_w<32, 64, 2, true>(a, nchunks, s);
    case 2: return staged_wh(a) ? bwd_wh<64, 2>(a, nchunks, s) : bwd_w<64, 64, 2, true>(a, nchunks, s);
    case 3: return staged_wh(a) ? bwd_wh<64, 1>(a, nchunks, s) : bwd_w<64, 64, 1, true>(a, nchunks, s);
    case 4: return staged_wh(a) ? bwd_wh<64, 1>(a, nchunks, s) : bwd_w<64, 64, 1, true>(a, nchunks, s);
    case 5: return bwd_w<64, 2, 1, false>(a, nchunks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_wgrad_reduce_all(const float* const* partial, const float* const* partial_b, const int* nchunks,
                                   float* grad, hipStream_t s) {
  WgradReduceAll r{};
  int first = 0;
  for (int l = 0; l < kCnnLayers; ++l) {
    const CnnLayer& L = kCnnArch[l];
    WgradReduceLayer& d = r.layer[l];
    d.partial = partial[l];
    d.partial_b = partial_b[l];
    d.nchunks = nchunks[l];
    d.cin = L.cin;
    d.cout = L.cout;
    d.kpi = cnn_kpad(L.cin);
    d.mrows = cnn_pad16(L.cout);
    d.grad = grad + cnn_weight_offset(l);
    d.first = first;
    first += L.cout * 9 * L.cin + L.cout;
  }
  hipLaunchKernelGGL(k_wgrad_reduce_all, dim3(blocks(first, 256)), dim3(256), 0, s, r);
  return hipGetLastError();
}

int cnn_loss_blocks(int batch, int hw) { return blocks((long long)batch * 2 * hw, kLossBlock); }

hipError_t launch_mse(const float* out, long long out_bstride, const float* labels, const int* batch_index, int batch,
                      int hw, uint16_t* gout, float* loss_part, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(cnn_loss_blocks(batch, hw)), dim3(kLossBlock), 0, s, out, out_bstride, labels,
                     batch_index, batch, hw, gout, loss_part);
  return hipGetLastError();
}

hipError_t launch_loss_final(const float* loss_part, int n, double count, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, s, loss_part, n, count, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float* master, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, long long t, hipStream_t s) {
  const float c1 = 1.f - (float)std::pow((double)b1, (double)t);
  const float c2 = 1.f - (float)std::pow((double)b2, (double)t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, s, master, grad, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_pack(const float* master, uint16_t* const* wf, uint16_t* const* wt, uint16_t* const* wl,
                       hipStream_t s) {
  for (int l = 0; l < kCnnLayers; ++l) {
    const CnnLayer& L = kCnnArch[l];
    const int n = L.cout * 9 * L.cin;
    hipLaunchKernelGGL(k_pack_layer, dim3(blocks(n, 256)), dim3(256), 0, s, master + cnn_weight_offset(l), L.cin,
                       L.cout, wf[l], wt[l], wl[l]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_downsample(const uint8_t* f1, long long p1, const uint8_t* f2, long long p2, int nx, int ny,
                             uint16_t* out, hipStream_t s) {
  const int w = nx / kCnnDown, h = ny / kCnnDown;
  if ((p1 & 3) != 0 || (reinterpret_cast<uintptr_t>(f1) & 3) != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_downsample, dim3(blocks((long long)w * h, 256)), dim3(256), 0, s, f1, p1, f2, p2, w, h, out);
  return hipGetLastError();
}

hipError_t launch_gate(const int8_t* accepted, int n, int max_outliers, int* gate, int* count, hipStream_t s) {
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(256), 0, s, accepted, n, max_outliers, gate, count);
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, int n, uint8_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks(n, 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace pivx
