// Hand-written gfx950 kernels of the multi-grid PIV pipeline.
//
// Reference stages and their replacements (xcorr2d_V1.0/xcorr2d_multi_grid.cu):
//   organize_data_fft2d_multi(_deformed) cu:497-517/654-673, the three batched
//   cuFFT C2C transforms cu:1334-1350, multiply_conjugate cu:519-534,
//   complex2amp_xcorr2d cu:536-545, fftshift2d_xcorr2d_row/col cu:547-583,
//   find_local_max_2d cu:585-606, reduce_max_index_general/final cu:4-172,
//   replace_local_max_2d cu:298-307, peak_subpixel_4peak_2d cu:309-350 and
//   kernel_remove_nan cu:457-471
//     -> one fused launch per pass, register FFTs: k_xcorr_pair<WX,WY> (two
//        windows per wave, shapes up to 32x32), k_xcorr_single<WX,WY> (one
//        window per wave, up to 64x64), k_xcorr_fast<WX,WY,WPB> (16x16), and
//        k_xcorr_generic (any legal size, LDS or global planes).
//   Kernel_Window_Deformation_2d cu:608-652           -> k_deform, or inside
//     k_xcorr_single (deform_window) for windows that do not overlap (and in
//     the learned pipeline's gated grid-1 fallback), with the term tables of
//     k_deform_terms
//   extend_filtered_vecfield_p1/p2 cu:1547-1618       -> k_extend
//   Upsampling_vecfield_kernel_2d cu:675-704          -> k_upsample
//   median_filter_data_prep, organize_median_filter_data, median_validity,
//   Calc_median, Calc_devi, median_filtering cu:174-296,484-495 -> k_prep,
//     k_median2<WX,WY> (3x3, 5x5, 7x7), k_median<NS> / k_median_generic
//   fill_blank_vecfield(_update) + host loop cu:352-397,1514-1528 -> k_fill_list,
//     k_fill
//   organize_smooth_filter_data + smoothing_tophat cu:399-443 -> k_smooth
//   MHE enhancement (MHE7_4_gpu.cu)                  -> k_mhe_hist, k_mhe_map
#include "kernels.h"
#include "fft_reg.h"
#include "sortnet.h"
#include <math.h>
#include <type_traits>
#include <algorithm>

namespace pivx {

// ============================================================== shared helpers

// 4 * conj(F1[k]) * F2[k] from the packed spectrum Z = FFT(f1 + i*f2):
// F1 = (Z[k] + conj(Z[-k]))/2, F2 = (Z[k] - conj(Z[-k]))/(2i).
// The reference's product order is conj(F1)*F2 (cu:530-531).
__device__ __forceinline__ float2 xprod(float2 A, float2 Am) {
  float f1r = A.x + Am.x, f1i = A.y - Am.y;   // 2*F1
  float f2r = A.y + Am.y, f2i = Am.x - A.x;   // 2*F2
  return make_float2(f1r * f2r + f1i * f2i, f1r * f2i - f1i * f2r);
}

// Order used by the top-N rounds: larger value first, ties to the lower
// row-major pixel index (the reference keeps the lower slot on ties, cu:7,51;
// its slots are filled in atomic order, here fixed to raster order).
__device__ __forceinline__ bool peak_before(float v, int i, float bv, int bi) {
  return v > bv || (v == bv && i < bi);
}

// The same order as one unsigned 64-bit key: a peak value (>= 0, whose bits
// order as unsigned integers) above the complemented pixel index, so that the
// larger key is the earlier peak.  Key 0 stands for "no peak".
__device__ __forceinline__ uint64_t peak_key(float v, int i) {
  return ((uint64_t)__float_as_uint(v) << 32) | (uint32_t)~i;
}
__device__ __forceinline__ float key_value(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ int key_index(uint64_t k) { return (int)~(uint32_t)k; }

// Max of a key over the W lanes of each aligned W-lane group.
template <int W>
__device__ __forceinline__ uint64_t group_max_key(uint64_t k) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)k, off, W), hi = __shfl_xor((uint32_t)(k >> 32), off, W);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    k = o > k ? o : k;
  }
  return k;
}

// Rows lo..hi of a column as a bit mask, where lo = max(by, 1) and
// hi = min(WY - 1 - by, WY - 2): the rows of the 4-neighbour walk inside the
// search border (cu:596).
template <typename Bits, int WY>
__device__ __forceinline__ Bits search_rows(int by) {
  const int lo = max(by, 1), hi = min(WY - 1 - by, WY - 2);
  if (hi < lo) return 0;
  return (((Bits)2 << hi) - 1) & ~(((Bits)1 << lo) - 1);
}

// 2 * w + (cond ? 1 : 0) as one add with carry whose carry-in is the lane
// mask of cond.
__device__ __forceinline__ uint32_t shift_in(uint32_t w, bool cond) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(cond);
  uint32_t r;
  uint64_t carry_out;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(w), "s"(m));
  return r;
}

// Rows 1..WY-2 of column cl of a correlation plane cc (row stride SX) that are
// 4-neighbour maxima with a positive value, as a bit mask (bit j = row j).
// "c0 > 0 and c0 >= each neighbour" (cu:597) is c0 >= max(up, dn, lv, rv,
// smallest positive float) with the max taken over the bit patterns: a
// window's plane holds values >= 0 (which order as their bits), or NaN
// throughout when a frame-2 value is NaN, where neither form finds a maximum.
// The walk runs upwards and shifts each row's flag in at bit 0, rows from 32
// on into the high word.
template <int WY, int SX, typename Bits>
__device__ __forceinline__ Bits column_maxima(const float* cc, int cl, int lft, int rgt) {
  static_assert(WY <= 8 * (int)sizeof(Bits), "one bit per row");
  uint32_t lo = 0, hi = 0;
  float dn = cc[(WY - 1) * SX + cl];
  float cur = cc[(WY - 2) * SX + cl];
  static_for<WY - 2>([&](auto jc) {
    constexpr int j = WY - 2 - jc;  // WY-2 down to 1
    const float up = cc[(j - 1) * SX + cl];
    const float lv = cc[j * SX + lft], rv = cc[j * SX + rgt];
    const uint32_t m = max(max(max(__float_as_uint(up), __float_as_uint(dn)), __float_as_uint(lv)),
                           max(__float_as_uint(rv), 1u));
    const bool is_max = cur >= __uint_as_float(m);
    if constexpr (j >= 32) {
      hi = shift_in(hi, is_max);  // row j at bit j - 32
    } else {
      lo = shift_in(lo, is_max);  // row j at bit j - 1
    }
    dn = cur;
    cur = up;
  });
  lo <<= 1;
  if constexpr (sizeof(Bits) == 8) {
    return ((Bits)hi << 32) | lo;
  } else {
    return lo;
  }
}

// The two largest peak keys (k1 > k2, 0 = none) over the rows in `bits` of
// column cl of a correlation plane cc (row stride SX, window width WX).
template <typename Bits, int WX, int SX>
__device__ __forceinline__ void best_two(const float* cc, Bits bits, int cl, uint64_t& k1, uint64_t& k2) {
  k1 = 0;
  k2 = 0;
  for (Bits rb = bits; rb; rb &= rb - 1) {
    int j;
    if constexpr (sizeof(Bits) == 8) {
      j = __builtin_ctzll(rb);
    } else {
      j = __builtin_ctz(rb);
    }
    const uint64_t kj = peak_key(cc[j * SX + cl], j * WX + cl);
    const bool gt1 = kj > k1;
    k2 = gt1 ? k1 : (kj > k2 ? kj : k2);
    k1 = gt1 ? kj : k1;
  }
}

__device__ __forceinline__ void wave_argmax(float& bv, int& bi) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float ov = __shfl_xor(bv, off, 64);
    int oi = __shfl_xor(bi, off, 64);
    if (peak_before(ov, oi, bv, bi)) { bv = ov; bi = oi; }
  }
}

// |v| * 0.25 / N (the packed product carries a factor 4): one multiply when N
// is a power of two (both forms are exact scalings), else the two-step form.
template <int N>
__device__ __forceinline__ float cc_scale(float v) {
  if constexpr ((N & (N - 1)) == 0) return fabsf(v) * (0.25f / (float)N);
  else return (fabsf(v) * 0.25f) / (float)N;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumber them so that every XCD walks one contiguous range of windows
// (neighbouring windows share most of their pixels).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  constexpr int NXCD = 8;
  const int q = nb / NXCD, r = nb % NXCD;
  const int x = b % NXCD, k = b / NXCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// 3-point Gaussian fit of one peak (cu:319-347) plus the NaN/Inf removal of
// kernel_remove_nan (cu:457-471).  cc is the fftshifted plane with row stride
// sx; (i, j) the peak pixel.  Returns validity; writes the displacement.
__device__ __forceinline__ bool subpixel_fit(const float* cc, int sx, int wsx, int wsy, int i, int j, float v,
                                             float px, float py, float lx1, float lx2, float ly1, float ly2,
                                             float& dx, float& dy) {
#pragma clang fp contract(off)
  dx = 0.f;
  dy = 0.f;
  if (i == 0 || i == wsx - 1 || j == 0 || j == wsy - 1 || !(v > 0.f)) return false;
  float f1 = logf(v);
  float f0 = logf(cc[j * sx + i - 1]);
  float f2 = logf(cc[j * sx + i + 1]);
  float x = (float)i + (f0 - f2) / (f0 + f2 - 2 * f1) / 2 - (float)wsx / 2 + px;
  f0 = logf(cc[(j - 1) * sx + i]);
  f2 = logf(cc[(j + 1) * sx + i]);
  float y = (float)j + (f0 - f2) / (f0 + f2 - 2 * f1) / 2 - (float)wsy / 2 + py;
  if (x < lx1 || x > lx2 || y < ly1 || y > ly2) return false;
  if (isnan(x) || isnan(y) || isinf(x) || isinf(y)) return false;
  dx = x;
  dy = y;
  return true;
}

// subpixel_fit for the precision form: an axis whose FP32 log-ratio
// denominator f0 + f2 - 2 f1 is below 1/16 in magnitude (a nearly flat peak,
// where the ratio magnifies every rounding of the logarithms) is recomputed
// with FP64 logarithms and FP64 ratio.
__device__ __forceinline__ float fit_axis_fp64near(float v, float fm, float fp) {
#pragma clang fp contract(off)
  const float f1 = logf(v), f0 = logf(fm), f2 = logf(fp);
  const float den = f0 + f2 - 2 * f1;
  if (fabsf(den) >= 0.0625f) return (f0 - f2) / den / 2;
  const double d1 = log((double)v), d0 = log((double)fm), d2 = log((double)fp);
  return (float)((d0 - d2) / (d0 + d2 - 2.0 * d1) / 2.0);
}

__device__ __forceinline__ bool subpixel_fit_fp64near(const float* cc, int sx, int wsx, int wsy, int i, int j, float v,
                                                      float px, float py, float lx1, float lx2, float ly1, float ly2,
                                                      float& dx, float& dy) {
#pragma clang fp contract(off)
  dx = 0.f;
  dy = 0.f;
  if (i == 0 || i == wsx - 1 || j == 0 || j == wsy - 1 || !(v > 0.f)) return false;
  const float x = (float)i + fit_axis_fp64near(v, cc[j * sx + i - 1], cc[j * sx + i + 1]) - (float)wsx / 2 + px;
  const float y = (float)j + fit_axis_fp64near(v, cc[(j - 1) * sx + i], cc[(j + 1) * sx + i]) - (float)wsy / 2 + py;
  if (x < lx1 || x > lx2 || y < ly1 || y > ly2) return false;
  if (isnan(x) || isnan(y) || isinf(x) || isinf(y)) return false;
  dx = x;
  dy = y;
  return true;
}

// subpixel_fit with its two axes on a lane pair: the even lane of the pair
// fits x, the odd lane y (same operations as above), the pair exchanges the
// coordinates and both apply the tests.  Lanes 2k and 2k + 1 of a W-lane
// group hold the same peak (v at pixel idx, -1 / 0 for none); returns its
// validity and writes the displacement on both lanes.
template <int W, bool FIT64 = false>
__device__ __forceinline__ bool subpixel_fit_pair(const float* cc, int sx, int wsx, int wsy, int idx, float v,
                                                  float px, float py, float lx1, float lx2, float ly1, float ly2,
                                                  bool yaxis, float& dx, float& dy) {
#pragma clang fp contract(off)
  const int i = idx % wsx, j = idx / wsx;
  const bool fit = v > 0.f && i != 0 && i != wsx - 1 && j != 0 && j != wsy - 1;
  float d = 0.f;
  if (fit) {
    const int c = j * sx + i, step = yaxis ? sx : 1;
    float r;
    if constexpr (FIT64) {
      r = fit_axis_fp64near(v, cc[c - step], cc[c + step]);  // FP64 near flat peaks (precision form)
    } else {
      const float f1 = logf(v);
      const float f0 = logf(cc[c - step]);
      const float f2 = logf(cc[c + step]);
      r = (f0 - f2) / (f0 + f2 - 2 * f1) / 2;
    }
    d = (float)(yaxis ? j : i) + r - (float)(yaxis ? wsy : wsx) / 2 + (yaxis ? py : px);
  }
  const float od = __shfl_xor(d, 1, W);
  const float x = yaxis ? od : d, y = yaxis ? d : od;
  dx = 0.f;
  dy = 0.f;
  if (!fit) return false;
  if (x < lx1 || x > lx2 || y < ly1 || y > ly2) return false;
  if (isnan(x) || isnan(y) || isinf(x) || isinf(y)) return false;
  dx = x;
  dy = y;
  return true;
}

// ============================================================== deformation per pixel
// Kernel_Window_Deformation_2d (cu:608-652): for every aug pixel, bilinear
// displacement from the extended filtered field, then bilinear sample of frame
// 2 at the displaced position.  Out-of-frame taps read 0 (the reference's zero
// padding); taps beyond the reference's padded buffer are counted.  The sums
// x + c1*e1 + ... + c4*e4 (cu:631-632) and c1*v1 + ... + c4*v4 (cu:649) are the
// fused multiply-add chains a CUDA build forms from that source by default
// (--fmad=true): fma(c4, e4, fma(c3, e3, fma(c2, e2, fma(c1, e1, x)))) and
// fma(c4, v4, fma(c3, v3, fma(c1, v1, c2*v2))).  Used by k_deform and by the
// gather of k_xcorr_single when the deformation is fused into it.

// Interpolation terms of aug coordinate v along one axis (cu:615-624): a =
// (v - st) / sp, the cell floor(a) clamped to [0, n - 2] (a guard: the
// extended field covers every aug pixel), the fraction a - floor(a) of the
// unclamped cell.  Returns (fraction, 1 - fraction, cell as int bits, v).
__device__ __forceinline__ float4 deform_terms(int v, float st, float sp, int n) {
#pragma clang fp contract(off)
  const float fv = (float)v;
  float t = (fv - st) / sp;
  int i = (int)floorf(t);
  t = t - (float)i;
  i = min(max(i, 0), n - 2);
  return make_float4(t, 1 - t, __int_as_float(i), fv);
}

// The four field samples of a cell.  A pixel's displacement is a convex
// combination of them, so while every sample is below hfws - 2 in magnitude
// no tap can leave the reference's padded buffer and the per-pixel test is
// skipped (may_oob false).
struct DefCell {
  float ex1, ex2, ex3, ex4, ey1, ey2, ey3, ey4;
  bool may_oob;
};

__device__ __forceinline__ DefCell load_cell(const float* ex_x, const float* ex_y, int nxe, int ixf, int iyf,
                                             float oob_guard) {
  DefCell e;
  const int i1 = iyf * nxe + ixf, i3 = i1 + nxe;
  e.ex1 = ex_x[i1];
  e.ex2 = ex_x[i1 + 1];
  e.ex3 = ex_x[i3];
  e.ex4 = ex_x[i3 + 1];
  e.ey1 = ex_y[i1];
  e.ey2 = ex_y[i1 + 1];
  e.ey3 = ex_y[i3];
  e.ey4 = ex_y[i3 + 1];
  const float em = fmaxf(fmaxf(fmaxf(fabsf(e.ex1), fabsf(e.ex2)), fmaxf(fabsf(e.ex3), fabsf(e.ex4))),
                         fmaxf(fmaxf(fabsf(e.ey1), fabsf(e.ey2)), fmaxf(fabsf(e.ey3), fabsf(e.ey4))));
  e.may_oob = !(em < oob_guard);  // NaN samples keep the exact test
  return e;
}

// Frame 2 for the taps.  It is read through a buffer resource that starts
// one byte before the frame and ends with it: a tap outside the frame gets an
// offset beyond that span, which the hardware returns as 0 without a memory
// access, so the four taps are unconditional loads and a pixel waits on one
// round trip.  The range check covers the vector offset only, so taps 2 and 4
// (x + 1) share the vector offset of taps 1 and 3 and add 1 as the scalar
// offset; with the base one byte early every in-frame tap's vector offset is
// >= 0.
// Cells that cannot leave the reference's padded buffer read their taps
// through a second resource, f2f, whose byte (jy + hfws_y) * pitch + jx is
// the frame byte at aug (jx, jy): the frame sits in a zero border that covers
// that buffer, so those taps need no test (jy + hfws_y >= 1 and jx > -pitch
// keep the offset positive).
struct DefFrame {
  __amdgpu_buffer_rsrc_t f2r, f2f;
  unsigned pitch;
  int nxm, nym, left_x, left_y, hfws_x, hfws_y;
  unsigned span_x, span_y;  // the reference's aug_ex buffer: [-hfws, n_aug + hfws - 1)
};

// an offset beyond every frame (frames hold < 2^31 bytes), also after the +1
// of taps 2 and 4; the bits of -2.0f, an inline constant of the select
constexpr unsigned DEF_TAP_OUT = 0xc0000000u;

__device__ __forceinline__ DefFrame def_frame(const uint8_t* f2, long long pitch, int nx, int ny, int left_x, int left_y,
                                              int hfws_x, int hfws_y, int nx_aug, int ny_aug) {
  DefFrame F;
  F.pitch = (unsigned)pitch;
  // byte b of the resource is frame byte b - 1 (byte 0 is never read)
  F.f2r = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(f2) - 1, (short)0, (int)(F.pitch * (unsigned)ny + 1u),
                                            0x00020000);
  // aug (0, -hfws_y); rows -hfws_y .. ny_aug + hfws_y of the zero border
  F.f2f = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(f2) - (long long)(left_y + hfws_y) * pitch - left_x,
                                            (short)0, (int)(F.pitch * (unsigned)(ny_aug + 2 * hfws_y + 1)), 0x00020000);
  F.nxm = nx - 1;
  F.nym = ny - 1;
  F.left_x = left_x;
  F.left_y = left_y;
  F.hfws_x = hfws_x;
  F.hfws_y = hfws_y;
  F.span_x = (unsigned)(nx_aug + 2 * hfws_x - 1);
  F.span_y = (unsigned)(ny_aug + 2 * hfws_y - 1);
  return F;
}

// The deformed value of the aug pixel with column terms (ax, ax1, fidx) and
// row terms (ay, ay1, fidy) in cell e; out-of-buffer taps are added to oob.
// COUNT false: the caller has checked that the cell cannot produce any
// (e.may_oob false), and the taps are read from the zero border unchecked.
template <bool COUNT = true>
__device__ __forceinline__ float deform_px(const DefFrame& F, const DefCell& e, float ax, float ax1, float fidx, float ay,
                                           float ay1, float fidy, unsigned& oob) {
#pragma clang fp contract(off)
  float ce1 = ax1 * ay1, ce2 = ax * ay1, ce3 = ax1 * ay, ce4 = ax * ay;
  float sx = fmaf(ce4, e.ex4, fmaf(ce3, e.ex3, fmaf(ce2, e.ex2, fmaf(ce1, e.ex1, fidx))));
  float sy = fmaf(ce4, e.ey4, fmaf(ce3, e.ey3, fmaf(ce2, e.ey2, fmaf(ce1, e.ey1, fidy))));
  // (float)(int)floor(s) == floor(s) for |s| < 2^31, so the fraction
  // s - (float)(int)floor(s) of cu:637-638 is s - floor(s)
  const float fx = floorf(sx), fy = floorf(sy);
  const int jx = (int)fx, jy = (int)fy;
  sx = sx - fx;
  sy = sy - fy;
  ce1 = (1 - sx) * (1 - sy);
  ce2 = sx * (1 - sy);
  ce3 = (1 - sx) * sy;
  ce4 = sx * sy;
  if constexpr (!COUNT) {
    // |displacement| < hfws - 2: every tap is inside the zero border
    const unsigned o1 = __umul24((unsigned)(jy + F.hfws_y), F.pitch) + (unsigned)jx, o3 = o1 + F.pitch;
    const float t1 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2f, o1, 0, 0);
    const float t2 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2f, o1, 1, 0);
    const float t3 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2f, o3, 0, 0);
    const float t4 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2f, o3, 1, 0);
    return fmaf(ce4, t4, fmaf(ce3, t3, fmaf(ce1, t1, ce2 * t2)));
  }
  const int xi = jx - F.left_x, yi = jy - F.left_y;
  const bool x0 = (unsigned)xi <= (unsigned)F.nxm, x1 = (unsigned)(xi + 1) <= (unsigned)F.nxm;
  const bool yok0 = (unsigned)yi <= (unsigned)F.nym, yok1 = (unsigned)(yi + 1) <= (unsigned)F.nym;
  // resource offsets (frame offset + 1) of taps 3 and 1; exact whenever the
  // tap or its x + 1 neighbour is inside the frame (24-bit product: rows and
  // row counts are below 2^24)
  const unsigned o3 = __umul24((unsigned)(yi + 1), F.pitch) + (unsigned)(xi + 1), o1 = o3 - F.pitch;
  const unsigned q1 = (yok0 && x0) ? o1 : DEF_TAP_OUT, q2 = (yok0 && x1) ? o1 : DEF_TAP_OUT;
  const unsigned q3 = (yok1 && x0) ? o3 : DEF_TAP_OUT, q4 = (yok1 && x1) ? o3 : DEF_TAP_OUT;
  const float t1 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2r, q1, 0, 0);
  const float t2 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2r, q2, 1, 0);
  const float t3 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2r, q3, 0, 0);
  const float t4 = (float)__builtin_amdgcn_raw_buffer_load_b8(F.f2r, q4, 1, 0);
  if (__builtin_expect(e.may_oob, 0))
    oob += ((unsigned)(jx + F.hfws_x) >= F.span_x || (unsigned)(jy + F.hfws_y) >= F.span_y) ? 1u : 0u;
  return fmaf(ce4, t4, fmaf(ce3, t3, fmaf(ce1, t1, ce2 * t2)));
}

// Gather of one window pixel: frame-1 value and frame-2 value at aug
// coordinates (ax, ay).  Image pixels outside the frame read as 0 (the
// reference's zero-padded aug_ex buffers, cu:951-954, 1271-1282).
__device__ __forceinline__ float2 gather_px(const XcArgs& a, int ax, int ay) {
  int xi = ax - a.left_x, yi = ay - a.left_y;
  bool in = (unsigned)xi < (unsigned)a.nx && (unsigned)yi < (unsigned)a.ny;
  float v1 = 0.f, v2 = 0.f;
  if (in) v1 = (float)a.f1[(long long)yi * a.pitch1 + xi];
  if (a.f2def) v2 = a.f2def[(long long)ay * a.def_pitch + ax];
  else if (in) v2 = (float)a.f2u8[(long long)yi * a.pitch2 + xi];
  return make_float2(v1, v2);
}

// ============================================================== fast fused kernel
//
// One workgroup = WPB horizontally consecutive windows of one vector row.
// LDS holds one padded complex plane per window (row stride WX+1 so that both
// row-per-lane and column-per-lane accesses are bank-conflict free).
//   1 gather  z = f1 + i*f2                       (coalesced along window rows)
//   2 row FFTs, one row per thread, in registers
//   3 column FFTs, one column per thread
//   4 Hermitian split -> conj(F1)*F2 for windows a=2p, b=2p+1, packed as
//     Q = Pa + i*Pb, inverse column FFTs
//   5 inverse row FFTs: Re -> cc_a, Im -> cc_b, scaled by 1/N and written
//     fftshifted as float planes (into window b's LDS plane)
//   6 per window, one wave: 4-neighbour local maxima inside the search border,
//     Npeak_save rounds of wave64 arg-max, Gaussian fit of each peak.
template <int WX, int WY, int WPB>
struct FastCfg {
  static constexpr int LPW = WX > WY ? WX : WY;
  static constexpr int NT = ((WPB * LPW + 63) / 64) * 64;
  static constexpr int SX = WX + 1;
};

// Synchronisation of the threads that share the LDS planes: a single-wave
// group only needs the compiler not to move LDS accesses across this point
// (a wave's LDS operations are performed in issue order); larger groups use
// the workgroup barrier.
template <int NT>
__device__ __forceinline__ void group_sync() {
  if constexpr (NT <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <int WX, int WY, int WPB>
__global__ void __launch_bounds__((FastCfg<WX, WY, WPB>::NT)) k_xcorr_fast(XcArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  static_assert(WPB % 2 == 0, "windows are processed in pairs");
  constexpr int NT = FastCfg<WX, WY, WPB>::NT;
  constexpr int SX = FastCfg<WX, WY, WPB>::SX;
  constexpr int NPIX = WX * WY;
  __shared__ float2 pl[WPB][WY][SX];

  const int tid = threadIdx.x;
  const int iy = blockIdx.x / a.chunks_x;
  const int ix0 = (blockIdx.x - iy * a.chunks_x) * WPB;

  // 1. gather: all of a thread's loads are issued before the LDS stores
  {
    constexpr int NG = (WPB * NPIX + NT - 1) / NT;
    float2 z[NG];
    static_for<NG>([&](auto gc) {
      constexpr int g = gc;
      const int e = tid + g * NT;
      const int wl = e / NPIX;
      const int rem = e - wl * NPIX;
      const int r = rem / WX, c = rem - (rem / WX) * WX;
      const int ix = ix0 + wl;
      z[g] = make_float2(0.f, 0.f);
      if ((WPB * NPIX % NT == 0 || e < WPB * NPIX) && ix < a.nx_vec)
        z[g] = gather_px(a, ix * a.sp_x + c, iy * a.sp_y + r);
    });
    static_for<NG>([&](auto gc) {
      constexpr int g = gc;
      const int e = tid + g * NT;
      if (WPB * NPIX % NT == 0 || e < WPB * NPIX) {
        const int wl = e / NPIX;
        const int rem = e - wl * NPIX;
        pl[wl][rem / WX][rem - (rem / WX) * WX] = z[g];
      }
    });
  }
  group_sync<NT>();

  // 2. forward row FFTs
  if (tid < WPB * WY) {
    const int wl = tid / WY, r = tid - (tid / WY) * WY;
    float2 x[WX];
    static_for<WX>([&](auto c) { x[c] = pl[wl][r][c]; });
    fft_reg<WX, false>(x);
    static_for<WX>([&](auto c) { pl[wl][r][c] = x[c]; });
  }
  group_sync<NT>();

  // 3. forward column FFTs
  if (tid < WPB * WX) {
    const int wl = tid / WX, c = tid - (tid / WX) * WX;
    float2 x[WY];
    static_for<WY>([&](auto r) { x[r] = pl[wl][r][c]; });
    fft_reg<WY, false>(x);
    static_for<WY>([&](auto r) { pl[wl][r][c] = x[r]; });
  }
  group_sync<NT>();

  // 4. split + conj-multiply + pair packing, inverse column FFTs
  {
    const bool act = tid < (WPB / 2) * WX;
    const int p = tid / WX, c = tid - (tid / WX) * WX;
    const int cm = (WX - c) % WX;
    float2 q[WY];
    if (act) {
      static_for<WY>([&](auto rc) {
        constexpr int r = rc;
        constexpr int rm = (WY - r) % WY;
        float2 pa = xprod(pl[2 * p][r][c], pl[2 * p][rm][cm]);
        float2 pb = xprod(pl[2 * p + 1][r][c], pl[2 * p + 1][rm][cm]);
        q[r] = make_float2(pa.x - pb.y, pa.y + pb.x);
      });
    }
    group_sync<NT>();
    if (act) {
      fft_reg<WY, true>(q);
      static_for<WY>([&](auto r) { pl[2 * p][r][c] = q[r]; });
    }
  }
  group_sync<NT>();

  // 5. inverse row FFTs -> shifted correlation planes
  if (tid < (WPB / 2) * WY) {
    const int p = tid / WY, r = tid - (tid / WY) * WY;
    float2 x[WX];
    static_for<WX>([&](auto c) { x[c] = pl[2 * p][r][c]; });
    fft_reg<WX, true>(x);
    float* cca = reinterpret_cast<float*>(&pl[2 * p + 1][0][0]);
    float* ccb = cca + WY * SX;
    const int rs = (r + WY / 2) % WY;
    static_for<WX>([&](auto cc) {
      constexpr int c = cc;
      constexpr int cs = (c + WX / 2) % WX;
      // |IFFT(4 conj(F1) F2)| * 0.25 / N == |IFFT(conj(F1) F2)| / N  (cu:542)
      cca[rs * SX + cs] = cc_scale<NPIX>(x[c].x);
      ccb[rs * SX + cs] = cc_scale<NPIX>(x[c].y);
    });
  }
  group_sync<NT>();

  // 6. peaks, one wave per window
  constexpr int NW = NT / 64;
  constexpr int PPL = (NPIX + 63) / 64;
  const int lane = tid & 63, wave = tid >> 6;
  for (int wl = wave; wl < WPB; wl += NW) {
    const int ix = ix0 + wl;
    if (ix >= a.nx_vec) break;
    const float* cc = reinterpret_cast<const float*>(&pl[wl | 1][0][0]) + (wl & 1) * WY * SX;
    const int w = iy * a.nx_vec + ix;
    float v[PPL];
    int cnt = 0;
    static_for<PPL>([&](auto mc) {
      constexpr int m = mc;
      const int pidx = lane + 64 * m;
      const int i = pidx % WX, j = pidx / WX;
      float val = 0.f;
      if (pidx < NPIX && i >= a.bx && i <= WX - 1 - a.bx && j >= a.by && j <= WY - 1 - a.by) {
        const float c0 = cc[j * SX + i];
        if (c0 > 0.f && c0 >= cc[j * SX + i - 1] && c0 >= cc[j * SX + i + 1] && c0 >= cc[(j - 1) * SX + i] &&
            c0 >= cc[(j + 1) * SX + i]) {
          val = c0;
          ++cnt;
        }
      }
      v[m] = val;
    });
    cnt = wave_sum(cnt);
    const float px = a.pred_x ? a.pred_x[w] : 0.f;
    const float py = a.pred_y ? a.pred_y[w] : 0.f;
    if (lane == 0) {
      a.ncand[w] = cnt;
      if (a.zero_x) a.zero_x[w] = a.zero_y[w] = 0.f;
    }
    // Npeak_save rounds of arg-max; round k's winner is kept by lane k % 64
    // and all fits of a group of 64 rounds run in parallel afterwards.
    float pv = INFINITY;
    int pi = -1;
    for (int k0 = 0; k0 < a.K; k0 += 64) {
      const int kn = min(64, a.K - k0);
      float my_v = 0.f;
      int my_i = -1;
      for (int k = 0; k < kn; ++k) {
        float bv = 0.f;
        int bi = 0x7fffffff;
        static_for<PPL>([&](auto mc) {
          constexpr int m = mc;
          const int pidx = lane + 64 * m;
          const float val = v[m];
          const bool after = val < pv || (val == pv && pidx > pi);
          if (val > 0.f && after && peak_before(val, pidx, bv, bi)) {
            bv = val;
            bi = pidx;
          }
        });
        wave_argmax(bv, bi);
        pv = bv;
        pi = bi;
        if (lane == k) {
          my_v = bv;
          my_i = bv > 0.f ? bi : -1;
        }
      }
      if (lane < kn) {
        const long long o = (long long)w * a.K + k0 + lane;
        float dx = 0.f, dy = 0.f;
        bool ok = false;
        if (my_v > 0.f) ok = subpixel_fit(cc, SX, WX, WY, my_i % WX, my_i / WX, my_v, px, py, a.lx1, a.lx2, a.ly1,
                                          a.ly2, dx, dy);
        a.disp_x[o] = dx;
        a.disp_y[o] = dy;
        a.valid[o] = ok ? 1 : 0;
        a.pk_val[o] = my_v;
        a.pk_idx[o] = my_i;
      }
    }
  }
}

// ============================================================== wave-per-pair kernel
//
// Windows up to 32x32: one 64-lane wave owns two horizontally adjacent
// windows a (lanes 0-31) and b (lanes 32-63); no workgroup barriers.
//   1 each lane loads one window row of frame 1 and frame 2 straight into
//     registers (wide loads; per-pixel zero fill only at the image border)
//   2 row FFT in registers, rows to LDS
//   3 column FFT, one column per lane, both windows at once
//   4 Hermitian split of both windows, packed as Q = Pa + i*Pb; each column
//     IFFT is split over the lane pair (l, l+32) by one radix-2 step (each
//     half transforms the even or the odd rows)
//   5 row IFFTs split the same way, halves exchanged through LDS; Re is
//     cc_a, Im is cc_b (scaled, fftshifted)
//   6 peaks: each half-wave walks its window by columns (3 LDS reads per
//     pixel), Npeak_save rounds of 32-lane arg-max, winners removed in place,
//     Gaussian fits in parallel.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


__device__ __forceinline__ int half_sum(int v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
  return v;
}

// WX consecutive u8 pixels as floats: 16-byte chunk loads (4-byte chunks for
// a remainder of WX % 16), any alignment.
template <int WX>
__device__ __forceinline__ void load_u8_row(const uint8_t* p, float (&f)[WX]) {
  auto put = [&](int q, uint32_t w) {
    f[4 * q + 0] = (float)(w & 255u);
    f[4 * q + 1] = (float)((w >> 8) & 255u);
    f[4 * q + 2] = (float)((w >> 16) & 255u);
    f[4 * q + 3] = (float)(w >> 24);
  };
  static_for<WX / 16>([&](auto qc) {
    constexpr int q = qc;
    uint4 w;
    __builtin_memcpy(&w, p + 16 * q, 16);
    put(4 * q + 0, w.x);
    put(4 * q + 1, w.y);
    put(4 * q + 2, w.z);
    put(4 * q + 3, w.w);
  });
  static_for<(WX % 16) / 4>([&](auto qc) {
    constexpr int q = (WX / 16) * 4 + qc;
    uint32_t w;
    __builtin_memcpy(&w, p + 4 * q, 4);
    put(q, w);
  });
}

// Pixels s..s+3 of an image row (nx >= 4 pixels) as a packed u8 word, 0 for
// pixels outside [0, nx): one 4-byte load at the clamped start, shifted into
// place and masked.
__device__ __forceinline__ uint32_t load_u8x4_edge(const uint8_t* row, int s, int nx) {
  const int cs = min(max(s, 0), nx - 4);
  uint32_t w;
  __builtin_memcpy(&w, row + cs, 4);
  const int sh = s - cs;  // pixel s + b is byte b + sh of w
  const uint64_t wide = (uint64_t)w << 32;
  const int rs = min(max(32 + 8 * sh, 0), 64);
  const uint32_t moved = rs >= 64 ? 0u : (uint32_t)(wide >> rs);
  const int lo = min(max(-s, 0), 4), hi = min(max(nx - s, 0), 4);  // inside bytes [lo, hi)
  const uint64_t m = ((1ull << (8 * hi)) - 1ull) & ~((1ull << (8 * lo)) - 1ull);
  return lo < hi ? moved & (uint32_t)m : 0u;
}

// A window row of u8 pixels starting at image column xi0 of image row yi, 0
// outside the frame (rows crossing the frame edge).
template <int WX>
__device__ __forceinline__ void load_u8_row_edge(const uint8_t* img, long long pitch, int nx, int ny, int xi0, int yi,
                                                 float (&f)[WX]) {
  const bool row_in = (unsigned)yi < (unsigned)ny;
  const uint8_t* row = img + (long long)min(max(yi, 0), ny - 1) * pitch;
  static_for<WX / 4>([&](auto qc) {
    constexpr int q = qc;
    const uint32_t w = row_in ? load_u8x4_edge(row, xi0 + 4 * q, nx) : 0u;
    f[4 * q + 0] = (float)(w & 255u);
    f[4 * q + 1] = (float)((w >> 8) & 255u);
    f[4 * q + 2] = (float)((w >> 16) & 255u);
    f[4 * q + 3] = (float)(w >> 24);
  });
}

// One window row of WX pixels starting at aug coordinates (ax0, ay):
// x[c] = (frame 1, frame 2).  Rows fully inside the image use wide loads;
// rows crossing the frame edge use masked 4-byte loads with 0 outside.
template <int WX, class A>
__device__ __forceinline__ void load_row(const A& a, int ax0, int ay, float2 (&x)[WX]) {
  const int xi0 = ax0 - a.left_x, yi = ay - a.left_y;
  const bool row_in = (unsigned)yi < (unsigned)a.ny;
  const bool inside = row_in && xi0 >= 0 && xi0 + WX <= a.nx;
  float f1[WX], f2[WX];
  if (inside) {
    load_u8_row<WX>(a.f1 + (long long)yi * a.pitch1 + xi0, f1);
  } else {
    load_u8_row_edge<WX>(a.f1, a.pitch1, a.nx, a.ny, xi0, yi, f1);
  }
  if (a.fdef.col) {
    // frame 2 comes from the deformed window plane (deform_window)
    static_for<WX>([&](auto c) { x[c].x = f1[c]; });
    return;
  }
  if (a.f2def) {
    const float* q = a.f2def + (long long)ay * a.def_pitch + ax0;
    static_for<WX / 4>([&](auto qc) {
      constexpr int k = qc;
      float4 t;
      __builtin_memcpy(&t, q + 4 * k, 16);
      f2[4 * k + 0] = t.x;
      f2[4 * k + 1] = t.y;
      f2[4 * k + 2] = t.z;
      f2[4 * k + 3] = t.w;
    });
  } else if (inside) {
    load_u8_row<WX>(a.f2u8 + (long long)yi * a.pitch2 + xi0, f2);
  } else {
    load_u8_row_edge<WX>(a.f2u8, a.pitch2, a.nx, a.ny, xi0, yi, f2);
  }
  static_for<WX>([&](auto c) { x[c] = make_float2(f1[c], f2[c]); });
}

// Frame 2 of the window at aug (ax0, ay0) deformed into the plane
// p[r * SX + c] (FusedDef): lane = window column c, rows r_lo..r_hi - 1 (a
// multiple of 8 each) in order; the row terms are wave-uniform and come
// through scalar loads, and the cell changes with the row, as in k_deform.
template <int WX, int WY, int SX, class A>
__device__ __forceinline__ void deform_window(const A& a, int ax0, int ay0, float* p, int lane, int r_lo = 0,
                                              int r_hi = WY) {
  const auto& d = a.fdef;
  if (lane >= WX) return;
  const float4 ct = d.col[ax0 + lane];
  const int ixf = __float_as_int(ct.z);
  const DefFrame F = def_frame(a.f2u8, a.pitch2, a.nx, a.ny, a.left_x, a.left_y, d.hfws_x, d.hfws_y, d.nx_aug, d.ny_aug);
  const float oob_guard = (float)(min(d.hfws_x, d.hfws_y) - 2);
  // the row terms through the constant address space: scalar loads
  typedef const __attribute__((address_space(4))) float ConstF;
  ConstF* const rows = (ConstF*)d.row + 4 * ay0;
  static_assert(WY % 8 == 0, "rows in groups of 8");
  unsigned oob = 0;
  int cell = -1;
  DefCell e{};
  // rows in groups of 8: when a group lies in one cell whose samples stay
  // below the out-of-buffer guard (the common case) the group has no branch
  // and its 32 frame taps are in flight together
  for (int r0 = r_lo; r0 < r_hi; r0 += 8) {
    float ay[8], ay1[8], fidy[8];
    int iyf[8];
    static_for<8>([&](auto k) {
      ay[k] = rows[4 * (r0 + k)];
      ay1[k] = rows[4 * (r0 + k) + 1];
      iyf[k] = __float_as_int(rows[4 * (r0 + k) + 2]);
      fidy[k] = rows[4 * (r0 + k) + 3];
    });
    if (iyf[0] == iyf[7]) {  // cells are monotone in the row
      if (iyf[0] != cell) {
        cell = iyf[0];
        e = load_cell(d.ex_x, d.ex_y, d.nxe, ixf, cell, oob_guard);
      }
      if (__builtin_expect(e.may_oob, 0)) {
        static_for<8>([&](auto k) {
          p[(r0 + k) * SX + lane] = deform_px<true>(F, e, ct.x, ct.y, ct.w, ay[k], ay1[k], fidy[k], oob);
        });
      } else {
        static_for<8>([&](auto k) {
          p[(r0 + k) * SX + lane] = deform_px<false>(F, e, ct.x, ct.y, ct.w, ay[k], ay1[k], fidy[k], oob);
        });
      }
    } else {
      static_for<8>([&](auto k) {
        if (iyf[k] != cell) {
          cell = iyf[k];
          e = load_cell(d.ex_x, d.ex_y, d.nxe, ixf, cell, oob_guard);
        }
        p[(r0 + k) * SX + lane] = e.may_oob ? deform_px<true>(F, e, ct.x, ct.y, ct.w, ay[k], ay1[k], fidy[k], oob)
                                            : deform_px<false>(F, e, ct.x, ct.y, ct.w, ay[k], ay1[k], fidy[k], oob);
      });
    }
  }
  if (oob && d.oob) atomicAdd(d.oob, (unsigned long long)oob);
}

// LDS of one wave-pair work item: one complex plane (= both windows' real
// planes), 8.4 KB at 32x32.
template <int WX, int WY>
constexpr int pair_lds_floats() { return 2 * WY * (WX + 1); }

// Work item `bid` (window pair bid % chunks_x of vector row bid / chunks_x)
// of the wave-pair kernel; lds holds pair_lds_floats<WX, WY>() floats.
template <int WX, int WY, class A>
__device__ __forceinline__ void xcorr_pair_body(const A& a, int bid, float* lds) {
  static_assert(WX <= 32 && WY <= 32 && WX % 4 == 0 && WY % 2 == 0, "pair kernel covers windows up to 32x32");
  constexpr int SX = WX + 1;
  constexpr int NPIX = WX * WY;
  constexpr int HX = WX / 2, HY = WY / 2;

  const int lane = threadIdx.x;
  const int h = lane >> 5, l = lane & 31;
  const int iy = bid / a.chunks_x;
  const int ix = (bid - iy * a.chunks_x) * 2 + h;
  const bool win_ok = ix < a.nx_vec;
  float* const pr = lds + h * WY * SX;  // this half's real plane

  // 1-2. row gather + forward row FFT, then the transpose (Re, then Im)
  float2 col[WY];
  {
    float2 x[WX];
    if (l < WY) {
      // the unused second window of an odd row loads its neighbour (its
      // results are never stored)
      load_row<WX>(a, (win_ok ? ix : ix - 1) * a.sp_x, iy * a.sp_y + l, x);
      fft_reg<WX, false>(x);
      static_for<WX>([&](auto c) { pr[l * SX + c] = x[c].x; });
    }
    wave_sync();
    if (l < WX) static_for<WY>([&](auto r) { col[r].x = pr[r * SX + l]; });
    wave_sync();
    if (l < WY) static_for<WX>([&](auto c) { pr[l * SX + c] = x[c].y; });
    wave_sync();
    if (l < WX) static_for<WY>([&](auto r) { col[r].y = pr[r * SX + l]; });
  }

  // 3. forward column FFTs (both windows)
  if (l < WX) fft_reg<WY, false>(col);
  wave_sync();

  // 4. split + column IFFTs.  Q = Pa + i*Pb with P = 4 conj(F1) F2.  Lane
  //    (h, l) holds column l of window h and takes the mirror column -l from
  //    lane (h, -l) by lane permutation (no LDS), forming its window's P for
  //    all rows in place; one permlane32 swap per component then gives lane
  //    (h, l) the rows ky = 2t + h of both windows.  Every column IFFT of
  //    length WY is shared by the lane pair (l, l + 32): half h transforms
  //    the rows ky = 2t + h (length WY/2), then X[y] = E[y] + w^y O[y] and
  //    X[y + WY/2] = E[y] - w^y O[y] (w = e^{2 pi i/WY}).
  float2* const pl = reinterpret_cast<float2*>(lds);
  float2 q[HY];
  if (l < WX) {
    const int src = h * 32 + (WX - l) % WX;
    auto mirror = [&](float2 v) { return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64)); };
    // rows 0..WY/2 from Z and the mirror column (the partner's rows
    // WY-ky are still Z: only rows 0..WY/2 are overwritten, each after its
    // own mirror read) ...
    static_for<HY + 1>([&](auto kc) {
      constexpr int ky = kc;
      constexpr int km = (WY - ky) % WY;
      col[ky] = xprod(col[ky], mirror(col[km]));
    });
    // ... and rows WY/2+1..WY-1 as conjugates of the partner's rows: the
    // correlation is real, so P[-ky][-kx] = conj P[ky][kx], and xprod with
    // its operands swapped is the exact conjugate (same rounded real part,
    // negated imaginary part)
    static_for<HY - 1>([&](auto kc) {
      constexpr int ky = HY + 1 + kc;
      const float2 pm = mirror(col[WY - ky]);
      col[ky] = make_float2(pm.x, -pm.y);
    });
    // lanes 32-63 of the even rows <-> lanes 0-31 of the odd rows
    auto swap32 = [](float& lo, float& hi) {
      const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
      lo = __uint_as_float(r[0]);
      hi = __uint_as_float(r[1]);
    };
    static_for<HY>([&](auto tc) {
      constexpr int t = tc;
      float2 pa = col[2 * t], pb = col[2 * t + 1];
      swap32(pa.x, pb.x);
      swap32(pa.y, pb.y);
      q[t] = make_float2(pa.x - pb.y, pa.y + pb.x);
    });
    fft_reg<HY, true>(q);
    if (h) static_for<HY>([&](auto kc) { q[kc] = twiddle<kc, WY, true>(q[kc]); });
  }
  // plane as eo[h][y][kx], y < WY/2: E (h = 0) and w^y O (h = 1)
  float2* const eo = pl;
  if (l < WX) static_for<HY>([&](auto k) { eo[(h * HY + k) * SX + l] = q[k]; });
  wave_sync();

  // 5. row IFFTs, shared the same way: lane (h, y) transforms the columns
  //    kx = 2t + h of row y of X, the pair exchanges halves through LDS and
  //    lane h writes output columns [h*WX/2, (h+1)*WX/2) of both correlation
  //    planes (Re -> window a, Im -> window b), scaled and fftshifted.
  float2 r[HX];
  const int yy = l < HY ? l : l - HY;
  if (l < WY) {
    const float sg = l < HY ? 1.f : -1.f;
    static_for<HX>([&](auto tc) {
      constexpr int t = tc;
      const float2 e = eo[yy * SX + 2 * t + h], o = eo[(HY + yy) * SX + 2 * t + h];
      r[t] = make_float2(fmaf(sg, o.x, e.x), fmaf(sg, o.y, e.y));
    });
    fft_reg<HX, true>(r);
    if (h) static_for<HX>([&](auto kc) { r[kc] = twiddle<kc, WX, true>(r[kc]); });
  }
  wave_sync();  // every read of eo is done before the plane is reused
  // plane as xb[h][k][y]: E' (h = 0) and w^k O' (h = 1) of row y
  float2* const xb = pl;
  float* const cc0 = lds;
  if (l < WY) {
    static_for<HX>([&](auto k) { xb[(h * HX + k) * WY + l] = r[k]; });
    wave_sync();
    float2 p[HX];
    static_for<HX>([&](auto k) { p[k] = xb[((1 - h) * HX + k) * WY + l]; });
    wave_sync();  // the correlation planes overwrite xb
    const float sh = h ? -1.f : 1.f;
    // output row (l + WY/2) % WY, columns from h ? 0 : WX/2
    float* const orow = cc0 + ((l + HY) % WY) * SX + (h ? 0 : HX);
    static_for<HX>([&](auto kc) {
      constexpr int k = kc;
      // h = 0: E' + w^k O';  h = 1: E' - w^k O'
      const float re = fmaf(sh, r[k].x, p[k].x), im = fmaf(sh, r[k].y, p[k].y);
      orow[k] = cc_scale<NPIX>(re);
      orow[WY * SX + k] = cc_scale<NPIX>(im);
    });
  }
  wave_sync();
  // 6. peaks: half-wave h walks window h column l, keeping a bit mask of the
  //    rows of its column that are local maxima; Npeak_save rounds of 32-lane
  //    arg-max over the lanes' best maxima
  const float* cc = cc0 + h * WY * SX;
  const bool col_ok = l < WX && l >= a.bx && l <= WX - 1 - a.bx;
  const int cl = l < WX ? l : WX - 1;
  const int lft = cl > 0 ? cl - 1 : 0, rgt = cl < WX - 1 ? cl + 1 : WX - 1;
  // the search border (rows from a uniform mask, then the lane's column)
  const uint32_t rows = search_rows<uint32_t, WY>(a.by);
  const uint32_t keep = col_ok ? rows : 0u;
  uint32_t bits = column_maxima<WY, SX, uint32_t>(cc, cl, lft, rgt) & keep;
  const int n = half_sum(__popc(bits));
  const int w = iy * a.nx_vec + ix;
  if (win_ok && l == 0) {
    a.ncand[w] = n;
    if (a.zero_x) a.zero_x[w] = a.zero_y[w] = 0.f;
  }
  const float px = (win_ok && a.pred_x) ? a.pred_x[w] : 0.f;
  const float py = (win_ok && a.pred_y) ? a.pred_y[w] : 0.f;
  // the lane's best two maxima of its column as peak keys (every maximum is
  // > 0, so key 0 is "none"); the owner of a round's winner clears that row
  // and promotes its second, and a lane that wins again rescans its column
  uint64_t mk, mk2;
  best_two<uint32_t, WX, SX>(cc, bits, cl, mk, mk2);
  bool have2 = true;
  // rounds in groups of 16: round k's winner is kept by lanes 2k and 2k + 1,
  // which fit its x and its y
  for (int k0 = 0; k0 < a.K; k0 += 16) {
    const int kn = min(16, a.K - k0);
    float my_v = 0.f;
    int my_i = -1;
    for (int k = 0; k < kn; ++k) {
      const uint64_t best = group_max_key<32>(mk);
      if (best != 0 && best == mk) {
        bits &= ~((uint32_t)1 << (key_index(mk) / WX));
        if (have2) {
          mk = mk2;
          have2 = false;
        } else {
          best_two<uint32_t, WX, SX>(cc, bits, cl, mk, mk2);
          have2 = true;
        }
      }
      if ((l >> 1) == k && best != 0) {
        my_v = key_value(best);
        my_i = key_index(best);
      }
    }
    if (win_ok && l < 2 * kn) {
      const long long o = (long long)w * a.K + k0 + (l >> 1);
      const bool ya = (l & 1) != 0;
      float dx, dy;
      const bool ok = subpixel_fit_pair<32>(cc, SX, WX, WY, my_i, my_v, px, py, a.lx1, a.lx2, a.ly1, a.ly2, ya, dx, dy);
      if (ya) {
        a.disp_y[o] = dy;
        a.pk_idx[o] = my_i;
      } else {
        a.disp_x[o] = dx;
        a.valid[o] = ok ? 1 : 0;
        a.pk_val[o] = my_v;
      }
    }
  }
}

template <int WX, int WY>
__global__ void __launch_bounds__(64) k_xcorr_pair(XcArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  __shared__ float lds[pair_lds_floats<WX, WY>()];
  xcorr_pair_body<WX, WY>(a, xcd_remap(blockIdx.x, gridDim.x), lds);
}

// ============================================================== wave-per-window kernel
//
// Windows up to 64x64: one 64-lane wave owns one window, no workgroup
// barriers.  The window's LDS footprint is about one real plane
// (WY x (WX + 2) floats, 16.9 KB at 64x64) so that two windows share every
// SIMD: the transpose between the forward row and column FFTs moves the real
// parts, then the imaginary parts, through one float plane.
//
// The product P = 4 conj(F1) F2 is Hermitian (the correlation is real), so the
// inverse runs on half the spectrum, every line shared by a lane pair (radix-2
// decimation in time: half h transforms the odd or even samples, the odd half
// is twiddled, the pair exchanges halves through LDS and combines):
//   * column IFFTs for kx = 0..WX/2, where columns 0 and WX/2 (real after the
//     transform) are packed into one: P[:, 0] + i P[:, WX/2];
//   * real-output row IFFTs two rows at a time, H = G[y] + i G[y + WY/2] with
//     G[y][-kx] = conj(G[y][kx]).
// LDS of one wave-per-window work item (floats, then WY float2 of pnyq).
template <int WX, int WY>
constexpr int single_lds_floats() {
  constexpr int SX = WX + 1, HX = WX / 2, HY = WY / 2, GS = HX + 1;
  return std::max(std::max(WY * SX, 2 * WY * GS), 4 * HX * HY);
}

// Peak stage of a wave-per-window plane (fftshifted correlation cc, row
// stride SX floats, in LDS): lane l walks column l of the window, keeping a
// bit mask of the rows that are local maxima; Npeak_save rounds of 64-lane
// arg-max; Gaussian fits of the winners.
template <int WX, int WY, int SX, bool FIT64 = false, class A>
__device__ __forceinline__ void single_peaks(const A& a, int w, const float* cc, int lane) {
  using Bits = typename std::conditional<(WY <= 32), uint32_t, uint64_t>::type;
  const bool col_ok = lane < WX && lane >= a.bx && lane <= WX - 1 - a.bx;
  const int cl = lane < WX ? lane : WX - 1;
  const int lft = cl > 0 ? cl - 1 : 0, rgt = cl < WX - 1 ? cl + 1 : WX - 1;
  // the search border (rows from a uniform mask, then the lane's column)
  const Bits rows = search_rows<Bits, WY>(a.by);
  const Bits keep = col_ok ? rows : (Bits)0;
  Bits bits = column_maxima<WY, SX, Bits>(cc, cl, lft, rgt) & keep;
  const int n = wave_sum(__popcll((unsigned long long)bits));
  if (lane == 0) {
    a.ncand[w] = n;
    if (a.zero_x) a.zero_x[w] = a.zero_y[w] = 0.f;
  }
  const float px = a.pred_x ? a.pred_x[w] : 0.f;
  const float py = a.pred_y ? a.pred_y[w] : 0.f;
  // the lane's best two maxima of its column as peak keys (every maximum is
  // > 0, so key 0 is "none"); the owner of a round's winner clears that row
  // and promotes its second, and a lane that wins again rescans its column
  uint64_t mk, mk2;
  best_two<Bits, WX, SX>(cc, bits, cl, mk, mk2);
  bool have2 = true;
  // rounds in groups of 32: round k's winner is kept by lanes 2k and 2k + 1,
  // which fit its x and its y
  for (int k0 = 0; k0 < a.K; k0 += 32) {
    const int kn = min(32, a.K - k0);
    float my_v = 0.f;
    int my_i = -1;
    for (int k = 0; k < kn; ++k) {
      const uint64_t best = group_max_key<64>(mk);
      if (best != 0 && best == mk) {
        bits &= ~((Bits)1 << (key_index(mk) / WX));
        if (have2) {
          mk = mk2;
          have2 = false;
        } else {
          best_two<Bits, WX, SX>(cc, bits, cl, mk, mk2);
          have2 = true;
        }
      }
      if ((lane >> 1) == k && best != 0) {
        my_v = key_value(best);
        my_i = key_index(best);
      }
    }
    if (lane < 2 * kn) {
      const long long o = (long long)w * a.K + k0 + (lane >> 1);
      const bool ya = (lane & 1) != 0;
      float dx, dy;
      const bool ok =
          subpixel_fit_pair<64, FIT64>(cc, SX, WX, WY, my_i, my_v, px, py, a.lx1, a.lx2, a.ly1, a.ly2, ya, dx, dy);
      if (ya) {
        a.disp_y[o] = dy;
        a.pk_idx[o] = my_i;
      } else {
        a.disp_x[o] = dx;
        a.valid[o] = ok ? 1 : 0;
        a.pk_val[o] = my_v;
      }
    }
  }
}

// Window w of the wave-per-window kernel; lds holds single_lds_floats floats,
// pnyq WY float2.
template <int WX, int WY, class A>
__device__ __forceinline__ void xcorr_single_body(const A& a, int w, float* lds, float2* pnyq) {
  static_assert(WX <= 64 && WY <= 64 && WX % 4 == 0 && WY % 4 == 0, "single-wave kernel covers windows up to 64x64");
  constexpr int SX = WX + 1;   // real plane row stride (floats)
  constexpr int NPIX = WX * WY;
  constexpr int HX = WX / 2, HY = WY / 2;
  constexpr int GS = HX + 1;   // G plane row stride (complex)

  const int lane = threadIdx.x;
  const int iy = w / a.nx_vec, ix = w - iy * a.nx_vec;
  float* const sp = lds;

  // 0. fused deformation: frame 2 of the window into the plane
  if (a.fdef.col) {
    deform_window<WX, WY, SX>(a, ix * a.sp_x, iy * a.sp_y, sp, lane);
    wave_sync();
  }
  // 1-2. row gather + forward row FFT, then the transpose (Re, then Im)
  float2 col[WY];
  {
    float2 x[WX];
    if (lane < WY) {
      load_row<WX>(a, ix * a.sp_x, iy * a.sp_y + lane, x);
      // each lane reads and then overwrites only its own row of the plane
      if (a.fdef.col) static_for<WX>([&](auto c) { x[c].y = sp[lane * SX + c]; });
      fft_reg<WX, false>(x);
      static_for<WX>([&](auto c) { sp[lane * SX + c] = x[c].x; });
    }
    wave_sync();
    if (lane < WX) static_for<WY>([&](auto r) { col[r].x = sp[r * SX + lane]; });
    wave_sync();
    if (lane < WY) static_for<WX>([&](auto c) { sp[lane * SX + c] = x[c].y; });
    wave_sync();
    if (lane < WX) static_for<WY>([&](auto r) { col[r].y = sp[r * SX + lane]; });
  }

  // 3. forward column FFTs (lane = column kx)
  if (lane < WX) fft_reg<WY, false>(col);
  wave_sync();

  // 4. product and column IFFTs.  Lane (h4, t4) gathers Z[ky][t4] and the
  //    mirror Z[-ky][-t4] for its rows ky = 2m + h4 (again Re, then Im),
  //    lanes ky < WY the column WX/2 pair for pnyq.
  const int h4 = lane / HX, t4 = lane - (lane / HX) * HX;  // valid for lane < WX
  const int tm = (WX - t4) % WX;
  const int ky0 = lane < WY ? lane : 0;
  const int kym = (WY - ky0) % WY;
  float2 own[HY], mir[HY];
  float2 n0, n1;
  if (lane < WX) static_for<WY>([&](auto r) { sp[r * SX + lane] = col[r].x; });
  wave_sync();
  if (lane < WX) {
    static_for<HY>([&](auto mc) {
      constexpr int m = mc;
      const int ky = 2 * m + h4;
      own[m].x = sp[ky * SX + t4];
      mir[m].x = sp[((WY - ky) % WY) * SX + tm];
    });
  }
  if (lane < WY) {
    n0.x = sp[ky0 * SX + HX];
    n1.x = sp[kym * SX + HX];
  }
  wave_sync();
  if (lane < WX) static_for<WY>([&](auto r) { sp[r * SX + lane] = col[r].y; });
  wave_sync();
  if (lane < WX) {
    static_for<HY>([&](auto mc) {
      constexpr int m = mc;
      const int ky = 2 * m + h4;
      own[m].y = sp[ky * SX + t4];
      mir[m].y = sp[((WY - ky) % WY) * SX + tm];
    });
  }
  if (lane < WY) {
    n0.y = sp[ky0 * SX + HX];
    n1.y = sp[kym * SX + HX];
    pnyq[lane] = xprod(n0, n1);
  }
  wave_sync();
  float2* const xb = reinterpret_cast<float2*>(lds);
  if (lane < WX) {
    // lanes t4 == 0 add i * P[ky][WX/2] (the broadcast read is harmless elsewhere)
    const float f = t4 == 0 ? 1.f : 0.f;
    float2 q[HY];
    static_for<HY>([&](auto mc) {
      constexpr int m = mc;
      const float2 p = xprod(own[m], mir[m]);
      const float2 pn = pnyq[2 * m + h4];
      q[m] = make_float2(fmaf(-f, pn.y, p.x), fmaf(f, pn.x, p.y));
    });
    fft_reg<HY, true>(q);
    if (h4) static_for<HY>([&](auto kc) { q[kc] = twiddle<kc, WY, true>(q[kc]); });
    // xb[h][k][t]: E (h = 0) and w^k O (h = 1) of column t
    static_for<HY>([&](auto k) { xb[(h4 * HY + k) * HX + t4] = q[k]; });
    wave_sync();
    float2 p[HY];
    static_for<HY>([&](auto k) { p[k] = xb[((1 - h4) * HY + k) * HX + t4]; });
    wave_sync();
    // G rows k + h4 * HY of column t4: E + w^k O (h4 = 0), E - w^k O (h4 = 1)
    const float sh = h4 ? -1.f : 1.f;
    float2* const gs = reinterpret_cast<float2*>(lds);
    static_for<HY>([&](auto k) {
      gs[(h4 * HY + k) * GS + t4] = make_float2(fmaf(sh, q[k].x, p[k].x), fmaf(sh, q[k].y, p[k].y));
    });
  }
  wave_sync();

  // 5. real-output row IFFTs: lane (h5, y5) transforms the samples kx = 2m + h5
  //    of H = G[y5] + i G[y5 + WY/2]; the pair exchanges halves and lane h5
  //    writes output columns [h5 * WX/2, (h5 + 1) * WX/2) of rows y5 (Im) and
  //    y5 + WY/2 (Re), scaled and fftshifted.
  const float2* const gs = reinterpret_cast<const float2*>(lds);
  float* const ccp = lds;
  const int h5 = lane / HY, y5 = lane - (lane / HY) * HY;  // valid for lane < WY
  if (lane < WY) {
    float2 hr[HX];
    const float2* const g0r = gs + y5 * GS;
    const float2* const g1r = gs + (y5 + HY) * GS;
    static_for<HX>([&](auto mc) {
      constexpr int m = mc;
      if constexpr (m == 0) {
        // kx = 0 (h5 = 0, real column 0) or kx = 1 (h5 = 1, direct)
        const float2 g0 = g0r[h5], g1 = g1r[h5];
        hr[m] = h5 ? make_float2(g0.x - g1.y, g0.y + g1.x) : make_float2(g0.x, g1.x);
      } else if constexpr (2 * m < HX) {
        const float2 g0 = g0r[2 * m + h5], g1 = g1r[2 * m + h5];
        hr[m] = make_float2(g0.x - g1.y, g0.y + g1.x);
      } else if constexpr (2 * m == HX) {
        // kx = WX/2 (h5 = 0: imaginary part of the packed column 0) or
        // kx = WX/2 + 1 (h5 = 1: conjugate of column WX/2 - 1)
        const int c = h5 ? HX - 1 : 0;
        const float2 g0 = g0r[c], g1 = g1r[c];
        hr[m] = h5 ? make_float2(g0.x + g1.y, g1.x - g0.y) : make_float2(g0.y, g1.y);
      } else {
        const float2 g0 = g0r[WX - 2 * m - h5], g1 = g1r[WX - 2 * m - h5];
        hr[m] = make_float2(g0.x + g1.y, g1.x - g0.y);
      }
    });
    fft_reg<HX, true>(hr);
    if (h5) static_for<HX>([&](auto kc) { hr[kc] = twiddle<kc, WX, true>(hr[kc]); });
    wave_sync();  // all of G is read before the exchange overwrites it
    float2* const xb2 = reinterpret_cast<float2*>(lds);
    static_for<HX>([&](auto k) { xb2[(h5 * HX + k) * HY + y5] = hr[k]; });
    wave_sync();
    float2 p[HX];
    static_for<HX>([&](auto k) { p[k] = xb2[((1 - h5) * HX + k) * HY + y5]; });
    wave_sync();
    const float sh = h5 ? -1.f : 1.f;
    // output rows y5 (Im) and y5 + WY/2 (Re), columns from h5 ? 0 : WX/2
    float* const orow = ccp + y5 * SX + (h5 ? 0 : HX);
    static_for<HX>([&](auto kc) {
      constexpr int k = kc;
      const float re = fmaf(sh, hr[k].x, p[k].x), im = fmaf(sh, hr[k].y, p[k].y);
      orow[HY * SX + k] = cc_scale<NPIX>(re);
      orow[k] = cc_scale<NPIX>(im);
    });
  }
  wave_sync();
  // 6. peaks
  single_peaks<WX, WY, SX>(a, w, ccp, lane);
}

template <int WX, int WY>
__global__ void __launch_bounds__(64) k_xcorr_single(XcArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  __shared__ float lds[single_lds_floats<WX, WY>()];
  __shared__ float2 pnyq[WY];  // column WX/2 of P
  xcorr_single_body<WX, WY>(a, blockIdx.x, lds, pnyq);
}

// ============================================================== batched work queue
// The queue's tables are read through the constant address space: with a
// uniform index these are scalar loads at the point of use, as for the
// per-pair kernels' kernel arguments (read through a generic pointer they
// would be flat loads into VGPRs, reloaded after every store; copied out
// whole they would hold 60 SGPRs for the whole pass).
__device__ __forceinline__ int batch_start(const BatchQueue& q, int p) {
  return ((const __attribute__((address_space(4))) int*)q.item_start)[p];
}
typedef __attribute__((address_space(4))) XcArgs ConstXcArgs;
// Claims of an XCD range [lo, hi): claim n covers `chunk` items interleaved
// with the claims of the range's other g workgroups, items
// lo + (n / g) g chunk + n % g + j g for j < chunk.  The workgroups of a
// range then work on neighbouring items at any moment (windows sharing rows
// meet in the XCD's L2) however many items a claim holds, and the first
// items of successive claims only increase: a claim whose first item is past
// the range ends the walk.  Every item of the range belongs to exactly one
// claim.
__device__ __forceinline__ long long batch_claim_first(int n, int lo, int g, int chunk) {
  return (long long)lo + (long long)(n / g) * g * chunk + n % g;
}
__device__ __forceinline__ const ConstXcArgs& batch_args(const BatchQueue& q, int p) {
  return ((const ConstXcArgs*)q.args)[p];
}

// ============================================================== two-wave 64x64 kernel
//
// One 128-thread workgroup (two waves) per 64x64 window.  The window's
// registers and LDS are shared by two waves (about half the VGPRs per lane
// and half the LDS per wave of the wave-per-window kernel), so that twice as
// many waves share every SIMD to hide the LDS and memory latency of the
// transforms.  Every length-64 line is split in two by one radix-2
// decimation-in-time step whose butterfly is left to the reader of the next
// phase: with E and O the length-32 transforms of the even and odd samples
// and T[k] = w^k O[k], X[k] = E[k] + T[k] and X[k + 32] = E[k] - T[k].
//   1. row FFTs: wave h transforms the samples 2n + h of row `lane` (wave 1
//      then twiddles); E and T of the row go to the plane as columns k and
//      32 + k;
//   2. column FFTs: wave h reads the rows 2n + h of column `lane`, forming
//      X = E +- T as it reads, and transforms them (wave 1 twiddles); E' and
//      T' go to the plane as rows k and 32 + k;
//   3. the product P = 4 conj(F1) F2 on the half spectrum kx = 0..31, Z and
//      its mirror Z[-ky][-kx] formed as E' +- T' from the plane, columns 0
//      and 32 packed as P[:, 0] + i P[:, 32] (both transform to real
//      columns); each column IFFT over ky is split four ways by a radix-4
//      decimation-in-time step (thread q transforms the rows ky = 4m + q,
//      length 16, and twiddles by w^(q k)) and combined in place in LDS, the
//      four inputs of a radix-4 butterfly being exactly the slots of its
//      four outputs;
//   4. real-output row IFFTs of H = G[y] + i G[y + 32] (y < 32), G[y][-kx] =
//      conj(G[y][kx]), split four ways the same way over kx = 4m + q; the
//      outputs are scaled and written fftshifted (Re: row y + 32, Im: row y);
//   5. peaks in wave 0 (single_peaks); wave 1 is done.
// The planes move through LDS as real parts, then imaginary parts; the
// phases are separated by workgroup barriers.
#ifndef PIVX_SINGLE64_TWO_WAVES
#define PIVX_SINGLE64_TWO_WAVES 1
#endif
constexpr bool kSingle64TwoWaves = PIVX_SINGLE64_TWO_WAVES != 0;  // 64x64 windows take this kernel
constexpr int kW2SX = 65;  // real plane row stride (floats)
constexpr int kW2GS = 33;  // complex row stride of the G / exchange plane
constexpr int kW2LdsFloats = 64 * kW2GS * 2;

// Pixels of parity P (columns 2n + P) of a row of u8 pixels, any alignment:
// 16-byte chunk loads, bytes P and P + 2 of every word.
template <int WX, int P>
__device__ __forceinline__ void load_u8_row_par(const uint8_t* p, float (&f)[WX / 2]) {
  static_assert(WX % 16 == 0, "16-byte chunks");
  static_for<WX / 16>([&](auto qc) {
    constexpr int q = qc;
    uint4 w;
    __builtin_memcpy(&w, p + 16 * q, 16);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    static_for<4>([&](auto jc) {
      constexpr int j = jc;
      f[8 * q + 2 * j + 0] = (float)((ws[j] >> (8 * P)) & 255u);
      f[8 * q + 2 * j + 1] = (float)((ws[j] >> (8 * P + 16)) & 255u);
    });
  });
}

// The same for a row crossing the frame edge (0 outside the frame).
template <int WX, int P>
__device__ __forceinline__ void load_u8_row_edge_par(const uint8_t* img, long long pitch, int nx, int ny, int xi0,
                                                     int yi, float (&f)[WX / 2]) {
  const bool row_in = (unsigned)yi < (unsigned)ny;
  const uint8_t* row = img + (long long)min(max(yi, 0), ny - 1) * pitch;
  static_for<WX / 4>([&](auto qc) {
    constexpr int q = qc;
    const uint32_t w = row_in ? load_u8x4_edge(row, xi0 + 4 * q, nx) : 0u;
    f[2 * q + 0] = (float)((w >> (8 * P)) & 255u);
    f[2 * q + 1] = (float)((w >> (8 * P + 16)) & 255u);
  });
}

// Samples 2n + P of one 64-pixel window row at aug (ax0, ay) as load_row
// reads them: x[n] = (frame 1, frame 2); with the fused deformation frame 2
// comes from the plane (the caller).
template <int P, class A>
__device__ __forceinline__ void load_row_par64(const A& a, int ax0, int ay, float2 (&x)[32]) {
  const int xi0 = ax0 - a.left_x, yi = ay - a.left_y;
  const bool row_in = (unsigned)yi < (unsigned)a.ny;
  const bool inside = row_in && xi0 >= 0 && xi0 + 64 <= a.nx;
  float f1[32], f2[32];
  if (inside) load_u8_row_par<64, P>(a.f1 + (long long)yi * a.pitch1 + xi0, f1);
  else load_u8_row_edge_par<64, P>(a.f1, a.pitch1, a.nx, a.ny, xi0, yi, f1);
  if (a.fdef.col) {
    static_for<32>([&](auto c) { x[c].x = f1[c]; });
    return;
  }
  if (a.f2def) {
    const float* q = a.f2def + (long long)ay * a.def_pitch + ax0;
    static_for<16>([&](auto qc) {
      constexpr int k = qc;
      float4 t;
      __builtin_memcpy(&t, q + 4 * k, 16);
      f2[2 * k + 0] = P ? t.y : t.x;
      f2[2 * k + 1] = P ? t.w : t.z;
    });
  } else if (inside) {
    load_u8_row_par<64, P>(a.f2u8 + (long long)yi * a.pitch2 + xi0, f2);
  } else {
    load_u8_row_edge_par<64, P>(a.f2u8, a.pitch2, a.nx, a.ny, xi0, yi, f2);
  }
  static_for<32>([&](auto c) { x[c] = make_float2(f1[c], f2[c]); });
}

// w^(q k) (w = e^(2 pi i / 64), inverse sign) on the 16 values of thread
// quarter q, q in 0..3 (uniform within each half-wave).
__device__ __forceinline__ void quarter_twiddle(float2 (&f)[16], int q) {
  if (q == 1) static_for<16>([&](auto k) { f[k] = twiddle<k, 64, true>(f[k]); });
  else if (q == 2) static_for<16>([&](auto k) { f[k] = twiddle<2 * k, 64, true>(f[k]); });
  else if (q == 3) static_for<16>([&](auto k) { f[k] = twiddle<3 * k, 64, true>(f[k]); });
}

// In-place inverse radix-4 butterflies over the exchange plane x4[(16 q + k)
// * kW2GS + c] (q = input quarter), for k = 4 kb .. 4 kb + 3 of column c:
// output p lands in slot (16 p + k), the line's element k + 16 p.
__device__ __forceinline__ void radix4_inplace(float2* x4, int kb, int c, float2 (&out)[16]) {
  static_for<4>([&](auto jc) {
    constexpr int j = jc;
    const int k = 4 * kb + j;
    const float2 t0 = x4[(k)*kW2GS + c], t1 = x4[(16 + k) * kW2GS + c];
    const float2 t2 = x4[(32 + k) * kW2GS + c], t3 = x4[(48 + k) * kW2GS + c];
    const float2 a0 = cadd(t0, t2), a1 = csub(t0, t2), b0 = cadd(t1, t3), b1 = csub(t1, t3);
    out[4 * j + 0] = cadd(a0, b0);                                   // p = 0
    out[4 * j + 1] = make_float2(a1.x - b1.y, a1.y + b1.x);          // p = 1: a1 + i b1
    out[4 * j + 2] = csub(a0, b0);                                   // p = 2
    out[4 * j + 3] = make_float2(a1.x + b1.y, a1.y - b1.x);          // p = 3: a1 - i b1
  });
}

template <class A>
__device__ __forceinline__ void xcorr_single64x2_body(const A& a, int w, float* lds, float2* pnyq) {
  constexpr int N = 64, H = 32, SX = kW2SX, GS = kW2GS;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int iy = w / a.nx_vec, ix = w - iy * a.nx_vec;
  const int ax0 = ix * a.sp_x, ay0 = iy * a.sp_y;
  float* const sp = lds;

  // 0. fused deformation: wave h deforms frame-2 rows 32h..32h+31 into the plane
  if (a.fdef.col) {
    deform_window<N, N, SX>(a, ax0, ay0, sp, lane, H * wv, H * wv + H);
    __syncthreads();
  }
  // 1. row FFTs: wave = sample parity, lane = row
  float2 v[H];
  if (wv == 0) load_row_par64<0>(a, ax0, ay0 + lane, v);
  else load_row_par64<1>(a, ax0, ay0 + lane, v);
  if (a.fdef.col) {
    static_for<H>([&](auto n) { v[n].y = sp[lane * SX + 2 * n + wv]; });
    __syncthreads();  // the deformed plane is read before it is rewritten
  }
  fft_reg<H, false>(v);
  if (wv) static_for<H>([&](auto k) { v[k] = twiddle<k, N, false>(v[k]); });
  // E (wave 0) / T (wave 1) of row `lane` as plane columns 32 wv + k; the
  // column phase reads X[r][c] = E[r][c % 32] +- T[r][c % 32] (+ for c < 32)
  float2 cz[H];
  const int cm = lane & (H - 1);
  const float cs = lane < H ? 1.f : -1.f;
  static_for<H>([&](auto k) { sp[lane * SX + H * wv + k] = v[k].x; });
  __syncthreads();
  static_for<H>([&](auto nc) {
    const int r = 2 * nc + wv;
    cz[nc].x = fmaf(cs, sp[r * SX + H + cm], sp[r * SX + cm]);
  });
  __syncthreads();
  static_for<H>([&](auto k) { sp[lane * SX + H * wv + k] = v[k].y; });
  __syncthreads();
  static_for<H>([&](auto nc) {
    const int r = 2 * nc + wv;
    cz[nc].y = fmaf(cs, sp[r * SX + H + cm], sp[r * SX + cm]);
  });
  // 2. column FFTs: wave = row parity, lane = column
  fft_reg<H, false>(cz);
  if (wv) static_for<H>([&](auto k) { cz[k] = twiddle<k, N, false>(cz[k]); });
  __syncthreads();  // every read of the row plane is done

  // 3. product on the half spectrum and column IFFTs split four ways.
  //    E' (wave 0) / T' (wave 1) of column `lane` as plane rows 32 wv + k:
  //    Z[ky][c] = E'[ky % 32][c] +- T'[ky % 32][c] (+ for ky < 32).
  //    Thread (q, t): column t < 32, rows ky = 4m + q.
  const int q = tid >> 5, t = tid & 31;
  const int tm = (N - t) & (N - 1);
  float2 zo[16], zm[16];
  float2 n0, n1;  // column 32 at row ky = tid and its mirror (wave 0)
  const int kyn = lane, kynm = (N - lane) & (N - 1);
  auto zread = [&](int ky, int c) {
    const int kk = ky & (H - 1);
    const float sg = ky < H ? 1.f : -1.f;
    return fmaf(sg, sp[(H + kk) * SX + c], sp[kk * SX + c]);
  };
  static_for<H>([&](auto k) { sp[(H * wv + k) * SX + lane] = cz[k].x; });
  __syncthreads();
  static_for<16>([&](auto mc) {
    constexpr int m = mc;
    const int ky = 4 * m + q;
    zo[m].x = zread(ky, t);
    zm[m].x = zread((N - ky) & (N - 1), tm);
  });
  if (wv == 0) {
    n0.x = zread(kyn, H);
    n1.x = zread(kynm, H);
  }
  __syncthreads();
  static_for<H>([&](auto k) { sp[(H * wv + k) * SX + lane] = cz[k].y; });
  __syncthreads();
  static_for<16>([&](auto mc) {
    constexpr int m = mc;
    const int ky = 4 * m + q;
    zo[m].y = zread(ky, t);
    zm[m].y = zread((N - ky) & (N - 1), tm);
  });
  if (wv == 0) {
    n0.y = zread(kyn, H);
    n1.y = zread(kynm, H);
    pnyq[lane] = xprod(n0, n1);
  }
  __syncthreads();  // pnyq written; every read of the column plane is done
  float2* const x4 = reinterpret_cast<float2*>(lds);
  {
    // threads t == 0 add i * P[ky][32] (the broadcast read is harmless elsewhere)
    const float f = t == 0 ? 1.f : 0.f;
    float2 qv[16];
    static_for<16>([&](auto mc) {
      constexpr int m = mc;
      const float2 p = xprod(zo[m], zm[m]);
      const float2 pn = pnyq[4 * m + q];
      qv[m] = make_float2(fmaf(-f, pn.y, p.x), fmaf(f, pn.x, p.y));
    });
    fft_reg<16, true>(qv);
    quarter_twiddle(qv, q);
    static_for<16>([&](auto k) { x4[(16 * q + k) * GS + t] = qv[k]; });
  }
  __syncthreads();
  {
    float2 g[16];
    radix4_inplace(x4, q, t, g);  // kb = q: k = 4q .. 4q + 3 of column t
    static_for<4>([&](auto jc) {
      static_for<4>([&](auto pc) {
        constexpr int j = jc, pp = pc;
        x4[(16 * pp + 4 * q + j) * GS + t] = g[4 * j + pp];
      });
    });
  }
  __syncthreads();

  // 4. row IFFTs: thread (q, y), y < 32, samples kx = 4m + q of H_y
  const int y = t;
  {
    const float2* const g0 = x4 + y * GS;
    const float2* const g1 = x4 + (y + H) * GS;
    float2 hv[16];
    static_for<16>([&](auto mc) {
      constexpr int m = mc;
      if constexpr (m == 0) {
        // kx = q: the packed real column 0 when q == 0
        const float2 a0 = g0[q], a1 = g1[q];
        hv[m] = q == 0 ? make_float2(a0.x, a1.x) : make_float2(a0.x - a1.y, a0.y + a1.x);
      } else if constexpr (m < 8) {
        const float2 a0 = g0[4 * m + q], a1 = g1[4 * m + q];
        hv[m] = make_float2(a0.x - a1.y, a0.y + a1.x);
      } else if constexpr (m == 8) {
        // kx = 32 + q: the packed column's imaginary part when q == 0, else
        // the conjugate of column 32 - q
        const int c = q == 0 ? 0 : H - q;
        const float2 a0 = g0[c], a1 = g1[c];
        hv[m] = q == 0 ? make_float2(a0.y, a1.y) : make_float2(a0.x + a1.y, a1.x - a0.y);
      } else {
        const int c = N - 4 * m - q;
        const float2 a0 = g0[c], a1 = g1[c];
        hv[m] = make_float2(a0.x + a1.y, a1.x - a0.y);
      }
    });
    fft_reg<16, true>(hv);
    quarter_twiddle(hv, q);
    __syncthreads();  // every read of G is done
    static_for<16>([&](auto k) { x4[(16 * q + k) * GS + y] = hv[k]; });
  }
  __syncthreads();
  {
    float2 o[16];
    radix4_inplace(x4, q, y, o);  // outputs x = 4q + j + 16 p of row pair y
    __syncthreads();              // every read of the exchange plane is done
    static_for<4>([&](auto jc) {
      static_for<4>([&](auto pc) {
        constexpr int j = jc, pp = pc;
        const int xo = 4 * q + j + 16 * pp;
        const int cx = (xo + H) & (N - 1);
        const float2 r = o[4 * j + pp];
        sp[(y + H) * SX + cx] = cc_scale<N * N>(r.x);
        sp[y * SX + cx] = cc_scale<N * N>(r.y);
      });
    });
  }
  __syncthreads();
  // 5. peaks (wave 0; wave 1 is done with this window)
  if (wv == 0) single_peaks<N, N, SX>(a, w, sp, lane);
}

__global__ void __launch_bounds__(128) k_xcorr_single64x2(XcArgs a) {
  __shared__ __align__(16) float lds[kW2LdsFloats];
  __shared__ float2 pnyq[64];
  xcorr_single64x2_body(a, blockIdx.x, lds, pnyq);
}

// The same behind the learned pipeline's fallback gate: launched with fewer
// workgroups, which walk the windows, so that the launch costs little when
// the gate is closed (the usual case).
__global__ void __launch_bounds__(128) k_xcorr_single64x2_gated(XcArgs a) {
  if (*a.gate == 0) return;
  __shared__ __align__(16) float lds[kW2LdsFloats];
  __shared__ float2 pnyq[64];
  const int nt = a.nx_vec * a.ny_vec;
  for (int w = blockIdx.x; w < nt; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();  // the last window's peak stage is done
    xcorr_single64x2_body(a, w, lds, pnyq);
  }
}

// The persistent batched form of the two-wave 64x64 kernel (same queue and
// claim order as k_xcorr_batch; `workgroups` counts two-wave workgroups).
__global__ void __launch_bounds__(128) k_xcorr_batch64x2(BatchQueue q) {
  constexpr int NXCD = 8;
  __shared__ __align__(16) float lds[kW2LdsFloats];
  __shared__ float2 pnyq[64];
  __shared__ int claim;
  const int home = blockIdx.x % NXCD;
  const int g = max(1, (int)gridDim.x / NXCD);  // workgroups per XCD range
  for (int r = 0; r < NXCD; ++r) {
    const int x = (home + r) % NXCD;
    const int lo = (int)((long long)q.total * x / NXCD), hi = (int)((long long)q.total * (x + 1) / NXCD);
    int p0 = 0;
    for (;;) {
      __syncthreads();  // the last item's planes and the last claim are read
      if (threadIdx.x == 0) claim = atomicAdd(q.counter + x, 1);
      __syncthreads();
      const long long first = batch_claim_first(__builtin_amdgcn_readfirstlane(claim), lo, g, q.chunk);
      if (first >= hi) break;  // this range is drained (both waves see the same claim)
      while (first >= batch_start(q, p0 + 1)) ++p0;
      int p = p0;
      for (int j = 0; j < q.chunk; ++j) {
        const long long it = first + (long long)j * g;
        if (it >= hi) break;
        const int item = (int)it;
        while (item >= batch_start(q, p + 1)) ++p;  // the items of a claim increase
        if (j > 0) __syncthreads();                 // the previous item's peak stage is done
        const ConstXcArgs& a = batch_args(q, p);
        xcorr_single64x2_body(a, item - batch_start(q, p), lds, pnyq);
      }
    }
  }
}

// ============================================================== generic fused kernel
//
// Any legal window (4..1024, multiple of 4, radices from the host plan).  One
// workgroup per window; two complex planes either in dynamic LDS or in a
// global scratch slice.  Batched Stockham stages with runtime radix over all
// lines of a plane, then the same product / inverse / peak logic as above,
// with the exact magnitude sqrt(re^2 + im^2) of cu:542.
struct GenPlan {
  int nrx, nry;
  int rx[16], ry[16];
};

__device__ float2* fft_lines(float2* src, float2* dst, int N, int lines, int es, int ls, const int* rad, int nrad,
                             const float2* tw, bool inv) {
  int Ns = 1;
  for (int s = 0; s < nrad; ++s) {
    const int R = rad[s];
    const int NR = N / R;
    const int tstep = N / (Ns * R);
    for (int item = threadIdx.x; item < lines * NR; item += blockDim.x) {
      const int line = item / NR, j = item - (item / NR) * NR;
      const int k = j % Ns;
      const float2* in = src + (long long)line * ls;
      float2* out = dst + (long long)line * ls;
      const int idxD = (j / Ns) * Ns * R + k;
      if (R == 2 || R == 4) {
        float2 v[4];
        for (int r = 0; r < R; ++r) {
          float2 x = in[(long long)(j + r * NR) * es];
          float2 w = tw[(r * k * tstep) % N];
          if (inv) w.y = -w.y;
          v[r] = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        }
        if (R == 2) {
          out[(long long)idxD * es] = cadd(v[0], v[1]);
          out[(long long)(idxD + Ns) * es] = csub(v[0], v[1]);
        } else {
          float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
          float2 t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
          float2 t3 = inv ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
          out[(long long)idxD * es] = cadd(t0, t2);
          out[(long long)(idxD + Ns) * es] = cadd(t1, t3);
          out[(long long)(idxD + 2 * Ns) * es] = csub(t0, t2);
          out[(long long)(idxD + 3 * Ns) * es] = csub(t1, t3);
        }
      } else {
        // direct length-R DFT, one output at a time (odd prime radices)
        for (int qo = 0; qo < R; ++qo) {
          float2 acc = make_float2(0.f, 0.f);
          for (int r = 0; r < R; ++r) {
            float2 x = in[(long long)(j + r * NR) * es];
            float2 w = tw[((long long)r * k * tstep + (long long)qo * r * NR) % N];
            if (inv) w.y = -w.y;
            acc.x += x.x * w.x - x.y * w.y;
            acc.y += x.x * w.y + x.y * w.x;
          }
          out[(long long)(idxD + qo * Ns) * es] = acc;
        }
      }
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

__device__ void block_argmax(float& bv, int& bi, float* sv, int* si) {
  wave_argmax(bv, bi);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
    sv[wave] = bv;
    si[wave] = bi;
  }
  __syncthreads();
  bv = sv[0];
  bi = si[0];
  for (int q = 1; q < nw; ++q)
    if (peak_before(sv[q], si[q], bv, bi)) {
      bv = sv[q];
      bi = si[q];
    }
}

// Peak stage over a shifted correlation plane cc [WY][WX] in LDS, run by the
// whole workgroup: 4-neighbour local maxima inside the search border
// (cu:585-606) into cand, the uncapped count, Npeak_save rounds of arg-max
// (value desc, raster index asc) and the Gaussian fit of each peak.  fit64:
// the FP64 fit of nearly flat peaks (subpixel_fit_fp64near).
template <int TWX = 0, int TWY = 0>  // compile-time window sides, or 0: the run-time rwx, rwy
__device__ void plane_peaks(const XcArgs& a, int w, float* cc, float* cand, int rwx, int rwy, float* red_v,
                            int* red_i, bool fit64) {
  const int WX = TWX ? TWX : rwx, WY = TWY ? TWY : rwy;
  const int N = WX * WY;
  int cnt = 0;
  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    const int j = e / WX, i = e - (e / WX) * WX;
    float val = 0.f;
    if (i >= a.bx && i <= WX - 1 - a.bx && j >= a.by && j <= WY - 1 - a.by) {
      const float c0 = cc[e];
      if (c0 > 0.f && c0 >= cc[e - 1] && c0 >= cc[e + 1] && c0 >= cc[e - WX] && c0 >= cc[e + WX]) {
        val = c0;
        ++cnt;
      }
    }
    cand[e] = val;
  }
  cnt = wave_sum(cnt);
  __shared__ int cnt_s[8];
  if ((threadIdx.x & 63) == 0) cnt_s[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int q = 0; q < (int)((blockDim.x + 63) >> 6); ++q) t += cnt_s[q];
    a.ncand[w] = t;
    if (a.zero_x) a.zero_x[w] = a.zero_y[w] = 0.f;
  }
  const float px = a.pred_x ? a.pred_x[w] : 0.f;
  const float py = a.pred_y ? a.pred_y[w] : 0.f;
  float pv = INFINITY;
  int pi = -1;
  for (int k = 0; k < a.K; ++k) {
    float bv = 0.f;
    int bi = 0x7fffffff;
    for (int e = threadIdx.x; e < N; e += blockDim.x) {
      const float val = cand[e];
      const bool after = val < pv || (val == pv && e > pi);
      if (val > 0.f && after && peak_before(val, e, bv, bi)) {
        bv = val;
        bi = e;
      }
    }
    block_argmax(bv, bi, red_v, red_i);
    pv = bv;
    pi = bi;
    if (threadIdx.x == 0) {
      const long long o = (long long)w * a.K + k;
      float dx = 0.f, dy = 0.f;
      bool ok = false;
      if (bv > 0.f)
        ok = fit64 ? subpixel_fit_fp64near(cc, WX, WX, WY, bi % WX, bi / WX, bv, px, py, a.lx1, a.lx2, a.ly1, a.ly2, dx, dy)
                   : subpixel_fit(cc, WX, WX, WY, bi % WX, bi / WX, bv, px, py, a.lx1, a.lx2, a.ly1, a.ly2, dx, dy);
      a.disp_x[o] = dx;
      a.disp_y[o] = dy;
      a.valid[o] = ok ? 1 : 0;
      a.pk_val[o] = bv;
      a.pk_idx[o] = bv > 0.f ? bi : -1;
    }
  }
}

__global__ void __launch_bounds__(256) k_xcorr_generic(XcArgs a, int WX, int WY, GenPlan plan, const float2* twx,
                                                       const float2* twy, float2* gscratch, int win0) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  extern __shared__ float2 dyn[];
  __shared__ float red_v[8];
  __shared__ int red_i[8];
  const int N = WX * WY;
  const int w = win0 + blockIdx.x;
  if (w >= a.nx_vec * a.ny_vec) return;
  const int iy = w / a.nx_vec, ix = w - (w / a.nx_vec) * a.nx_vec;
  float2* A = gscratch ? gscratch + (long long)blockIdx.x * 2 * N : dyn;
  float2* B = A + N;

  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    const int r = e / WX, c = e - (e / WX) * WX;
    A[e] = gather_px(a, ix * a.sp_x + c, iy * a.sp_y + r);
  }
  __syncthreads();
  float2* Z = fft_lines(A, B, WX, WY, 1, WX, plan.rx, plan.nrx, twx, false);
  float2* O = (Z == A) ? B : A;
  Z = fft_lines(Z, O, WY, WX, WX, 1, plan.ry, plan.nry, twy, false);
  O = (Z == A) ? B : A;
  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    const int r = e / WX, c = e - (e / WX) * WX;
    const int rm = (WY - r) % WY, cm = (WX - c) % WX;
    O[e] = xprod(Z[e], Z[rm * WX + cm]);
  }
  __syncthreads();
  Z = fft_lines(O, Z, WX, WY, 1, WX, plan.rx, plan.nrx, twx, true);
  O = (Z == A) ? B : A;
  Z = fft_lines(Z, O, WY, WX, WX, 1, plan.ry, plan.nry, twy, true);
  O = (Z == A) ? B : A;
  float* cc = reinterpret_cast<float*>(O);
  float* cand = cc + N;
  for (int e = threadIdx.x; e < N; e += blockDim.x) {
    const int r = e / WX, c = e - (e / WX) * WX;
    const int rs = (r + WY / 2) % WY, cs = (c + WX / 2) % WX;
    const float2 q = Z[e];
    cc[rs * WX + cs] = (sqrtf(q.x * q.x + q.y * q.y) * 0.25f) / (float)N;
  }
  __syncthreads();
  plane_peaks(a, w, cc, cand, WX, WY, red_v, red_i, false);
}

// ============================================================== batched passes
// One persistent launch runs a correlation pass over many pairs of uneven
// size: every pair contributes its work items (window pairs of the wave-pair
// kernel, windows of the wave-per-window kernel) to one queue, and each
// workgroup claims `chunk` items at a time (interleaved with the other
// workgroups' claims, batch_claim_first) from an atomic counter until the
// queue is empty.  Each item runs the same body as the per-pair
// kernel with that pair's arguments, so results are bit-identical.
// The queue is split into one contiguous range per XCD (workgroups are
// dispatched to the 8 XCDs round-robin by workgroup index), so that
// neighbouring windows, which share most of their pixels, are claimed by
// workgroups of one XCD and meet in its L2; a workgroup whose range is empty
// moves on to the next XCD's range, so every range drains.
// waves per SIMD the per-pair kernels of a shape reach (their register use),
// kept by the persistent form
template <int WX, int WY, bool PAIR>
constexpr int batch_waves_per_eu() { return PAIR ? 4 : (WX * WY <= 48 * 24 ? 3 : 2); }  // larger: 3 spills

template <int WX, int WY, bool PAIR>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(batch_waves_per_eu<WX, WY, PAIR>())))
k_xcorr_batch(BatchQueue q) {
  constexpr int NL = PAIR ? pair_lds_floats<WX, WY>() : single_lds_floats<WX, WY>();
  constexpr int NXCD = 8;
  __shared__ float lds[NL];
  __shared__ float2 pnyq[WY];
  const int home = blockIdx.x % NXCD;
  const int g = max(1, (int)gridDim.x / NXCD);  // workgroups per XCD range
  for (int r = 0; r < NXCD; ++r) {
    const int x = (home + r) % NXCD;
    const int lo = (int)((long long)q.total * x / NXCD), hi = (int)((long long)q.total * (x + 1) / NXCD);
    int p0 = 0;
    for (;;) {
      int n = 0;
      if (threadIdx.x == 0) n = atomicAdd(q.counter + x, 1);
      const long long first = batch_claim_first(__builtin_amdgcn_readfirstlane(__shfl(n, 0, 64)), lo, g, q.chunk);
      if (first >= hi) break;  // this range is drained (every wave gets here for every range)
      while (first >= batch_start(q, p0 + 1)) ++p0;
      int p = p0;
      for (int j = 0; j < q.chunk; ++j) {
        const long long it = first + (long long)j * g;
        if (it >= hi) break;
        const int item = (int)it;
        while (item >= batch_start(q, p + 1)) ++p;  // the items of a claim increase
        const ConstXcArgs& a = batch_args(q, p);
        if constexpr (PAIR) xcorr_pair_body<WX, WY>(a, item - batch_start(q, p), lds);
        else xcorr_single_body<WX, WY>(a, item - batch_start(q, p), lds, pnyq);
        wave_sync();  // the next item rewrites the LDS planes
      }
    }
  }
}

// ============================================================== precision form
// Correlation as DFT-by-matrix contractions on the matrix cores (FP16 or
// BF16 operands, FP32 accumulation; with SPLIT every operand x is carried as
// hi = round(x) plus lo = round(x - hi) and each product as hi*hi + hi*lo +
// lo*hi, which brings the products back to about FP32 accuracy), replacing
// the batched cuFFT plans of cu:922-931, 1334-1350 for windows up to 64x64.
// With F = C - iS (C = cos, S = sin of 2 pi jk / n) and z = f1 + i f2:
//   T = F_H z        (Tr = [C  S][zr; zi],  Ti = [-S C][zr; zi])
//   Z = T F_W        (Zr = [Tr Ti][C; S],   Zi = [Tr Ti][-S; C])
//   P = conj(F1) F2  from Z and its mirror (Hermitian split), scaled by a
//                    power of two so that FP16 operands cannot overflow
//   U = F_H^H P      (Ur = [C -S][Pr; Pi],  Ui = [S C][Pr; Pi])
//   cc = Re(U F_W^H) = [Ur Ui][C; -S]
// Each step is a set of 16x16 output tiles of v_mfma_f32_16x16x32_{f16,bf16}
// with K padded per half to a multiple of 16; the DFT matrices are constant
// tables (DftTables), the data operands live in LDS in the layout the next
// step reads (A row-major with K contiguous, B with K contiguous per
// column).  The plane then goes through plane_peaks with the FP64 fit of
// flat peaks.
typedef _Float16 f16x8_t __attribute__((__vector_size__(16)));
typedef __bf16 bf16x8_t __attribute__((__vector_size__(16)));
typedef float f32x4_t __attribute__((__vector_size__(16)));
typedef unsigned int u32x4_t __attribute__((__vector_size__(16)));

template <bool BF>
struct DftTy;
template <>
struct DftTy<false> {
  static __device__ __forceinline__ uint16_t cvt(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
  static __device__ __forceinline__ float back(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }
  static __device__ __forceinline__ f32x4_t mma(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b), c,
                                                  0, 0, 0);
  }
};
template <>
struct DftTy<true> {
  static __device__ __forceinline__ uint16_t cvt(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
  static __device__ __forceinline__ float back(uint16_t b) { return __builtin_bit_cast(float, (unsigned int)b << 16); }
  static __device__ __forceinline__ f32x4_t mma(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
  }
};

__device__ __forceinline__ u32x4_t ld16(const uint16_t* p) { return *reinterpret_cast<const u32x4_t*>(p); }

// one 16x16 output tile: D[m0 + .][n0 + .] = sum_k A[m][k] B[n][k]
template <bool BF, bool SPLIT>
__device__ __forceinline__ f32x4_t dft_tile(const uint16_t* Ah, const uint16_t* Al, int lda, const uint16_t* Bh,
                                            const uint16_t* Bl, int ldb, int m0, int n0, int kdim, int lane) {
  f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, kg = lane >> 4;
  const uint16_t* ah = Ah + (long long)(m0 + r) * lda + 8 * kg;
  const uint16_t* bh = Bh + (long long)(n0 + r) * ldb + 8 * kg;
  const uint16_t* al = Al + (long long)(m0 + r) * lda + 8 * kg;
  const uint16_t* bl = Bl + (long long)(n0 + r) * ldb + 8 * kg;
  for (int k0 = 0; k0 < kdim; k0 += 32) {
    const u32x4_t a = ld16(ah + k0), b = ld16(bh + k0);
    acc = DftTy<BF>::mma(a, b, acc);
    if constexpr (SPLIT) {
      acc = DftTy<BF>::mma(a, ld16(bl + k0), acc);
      acc = DftTy<BF>::mma(ld16(al + k0), b, acc);
    }
  }
  return acc;
}

template <bool BF>
__device__ __forceinline__ void dft_put(uint16_t* hi, uint16_t* lo, int i, float v) {
  const uint16_t h = DftTy<BF>::cvt(v);
  hi[i] = h;
  lo[i] = DftTy<BF>::cvt(v - DftTy<BF>::back(h));
}

// One workgroup (4 waves) per window of W x H (multiples of 8, at most 64).
template <int W, int H, bool BF>
__global__ void __launch_bounds__(256) k_xcorr_dft(XcArgs a, DftTables tb, int split) {
  if (a.gate && *a.gate == 0) return;
  constexpr int PW = (W + 15) / 16 * 16, PH = (H + 15) / 16 * 16;
  constexpr int KH = 2 * PH, KW = 2 * PW, LB = KH + 8, LA = KW + 8;
  constexpr int N = W * H;
  constexpr int tiles = (PH / 16) * (PW / 16);
  constexpr int PER = (N + 255) / 256;  // plane elements per thread
  __shared__ __align__(16) uint16_t Bh[PW * LB], Bl[PW * LB];  // B-layout data: z, then P
  __shared__ __align__(16) uint16_t Ah[PH * LA], Al[PH * LA];  // A-layout data: T, then U
  __shared__ __align__(16) float Zf[N * 2];                    // Z (complex), then cc + cand
  __shared__ float red_v[8];
  __shared__ int red_i[8];
  __shared__ float smax[4];
  const int w = blockIdx.x;
  if (w >= a.nx_vec * a.ny_vec) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int iy = w / a.nx_vec, ix = w - (w / a.nx_vec) * a.nx_vec;
  // 0. z into the B layout: column c holds [zr[0..H) | pad | zi[0..H) | pad]
  if constexpr (PW != W || PH != H) {
    for (int e = tid; e < PW * LB; e += 256) Bh[e] = Bl[e] = 0;
    __syncthreads();
  }
  for (int e = tid; e < N; e += 256) {
    const int r = e / W, c = e - (e / W) * W;
    const float2 z = gather_px(a, ix * a.sp_x + c, iy * a.sp_y + r);
    dft_put<BF>(Bh, Bl, c * LB + r, z.x);
    dft_put<BF>(Bh, Bl, c * LB + PH + r, z.y);
  }
  __syncthreads();
  // 1. [Tr | Ti] = A1[o] * z  -> A layout, row k = [Tr[k][0..PW) | Ti[k][0..PW)]
  for (int t = wave; t < 2 * tiles; t += 4) {
    const int o = t / tiles, q = t - o * tiles, mt = q / (PW / 16), nt = q - mt * (PW / 16);
    const f32x4_t d = split ? dft_tile<BF, true>(tb.a1h[o], tb.a1l[o], KH, Bh, Bl, LB, mt * 16, nt * 16, KH, lane)
                            : dft_tile<BF, false>(tb.a1h[o], tb.a1l[o], KH, Bh, Bl, LB, mt * 16, nt * 16, KH, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = mt * 16 + 4 * (lane >> 4) + j, c = nt * 16 + (lane & 15);
      dft_put<BF>(Ah, Al, k * LA + o * PW + c, d[j]);
    }
  }
  __syncthreads();
  // 2. [Zr, Zi] = T * B2[o] -> Zf[(k * W + l) * 2 + o]
  for (int t = wave; t < 2 * tiles; t += 4) {
    const int o = t / tiles, q = t - o * tiles, mt = q / (PW / 16), nt = q - mt * (PW / 16);
    const f32x4_t d = split ? dft_tile<BF, true>(Ah, Al, LA, tb.b2h[o], tb.b2l[o], KW, mt * 16, nt * 16, KW, lane)
                            : dft_tile<BF, false>(Ah, Al, LA, tb.b2h[o], tb.b2l[o], KW, mt * 16, nt * 16, KW, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = mt * 16 + 4 * (lane >> 4) + j, l = nt * 16 + (lane & 15);
      if (k < H && l < W) Zf[(k * W + l) * 2 + o] = d[j];
    }
  }
  __syncthreads();
  // 3. P = conj(F1) F2 (4x, folded into the final scale), scaled by 2^-s so
  //    that |P| <= 256: the next products stay inside the FP16 range
  const float2* Z = reinterpret_cast<const float2*>(Zf);
  float mx = 0.f;
  float2 Pv[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q;
    if (e < N) {
      const int r = e / W, c = e - (e / W) * W;
      const int rm = r ? H - r : 0, cm = c ? W - c : 0;
      Pv[q] = xprod(Z[e], Z[rm * W + cm]);
      mx = fmaxf(mx, fmaxf(fabsf(Pv[q].x), fabsf(Pv[q].y)));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) smax[wave] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
  int ex = 0;
  if (mx > 0.f) (void)frexpf(mx, &ex);  // mx < 2^ex
  const int sh = 8 - ex;                 // |P| * 2^sh < 256
  __syncthreads();  // every thread has read Z before the B layout is rewritten
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q;
    if (e < N) {
      const int r = e / W, c = e - (e / W) * W;
      dft_put<BF>(Bh, Bl, c * LB + r, ldexpf(Pv[q].x, sh));
      dft_put<BF>(Bh, Bl, c * LB + PH + r, ldexpf(Pv[q].y, sh));
    }
  }
  __syncthreads();
  // 4. [Ur | Ui] = A3[o] * P -> A layout
  for (int t = wave; t < 2 * tiles; t += 4) {
    const int o = t / tiles, q = t - o * tiles, mt = q / (PW / 16), nt = q - mt * (PW / 16);
    const f32x4_t d = split ? dft_tile<BF, true>(tb.a3h[o], tb.a3l[o], KH, Bh, Bl, LB, mt * 16, nt * 16, KH, lane)
                            : dft_tile<BF, false>(tb.a3h[o], tb.a3l[o], KH, Bh, Bl, LB, mt * 16, nt * 16, KH, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = mt * 16 + 4 * (lane >> 4) + j, c = nt * 16 + (lane & 15);
      dft_put<BF>(Ah, Al, k * LA + o * PW + c, d[j]);
    }
  }
  __syncthreads();
  // 5. cc = Re(U F_W^H), |.| / N (the exact correlation is real and >= 0),
  //    written fftshifted
  float* cc = Zf;
  float* cand = Zf + N;
  const float scale = ldexpf(0.25f / (float)N, -sh);
  for (int t = wave; t < tiles; t += 4) {
    const int mt = t / (PW / 16), nt = t - mt * (PW / 16);
    const f32x4_t d = split ? dft_tile<BF, true>(Ah, Al, LA, tb.b4h, tb.b4l, KW, mt * 16, nt * 16, KW, lane)
                            : dft_tile<BF, false>(Ah, Al, LA, tb.b4h, tb.b4l, KW, mt * 16, nt * 16, KW, lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = mt * 16 + 4 * (lane >> 4) + j, c = nt * 16 + (lane & 15);
      if (r < H && c < W) {
        const int rs = r < H / 2 ? r + H / 2 : r - H / 2, cs = c < W / 2 ? c + W / 2 : c - W / 2;
        cc[rs * W + cs] = fabsf(d[j]) * scale;
      }
    }
  }
  __syncthreads();
  plane_peaks<W, H>(a, w, cc, cand, W, H, red_v, red_i, true);
}

// ============================================================== precision form, 32x32, one wave per window
// The precision form of 32x32 windows with one wave per window and the DFT
// matrices in registers.  F = C - iS; every product is one 16x16x32 MFMA over
// the whole K = 32 of a matrix (2 x 2 output tiles).  Lane l = 16 g + r
// supplies row r of an A tile (K values 8g..8g+7) and column r of a B tile
// (the same K values) and receives rows 4g..4g+3 of column r of the output
// tile.  C and S are symmetric, so one register copy of each (hi, lo) serves
// as A operand (rows) and as B operand (columns).
//   T  = z F:         Tr = zr C + zi S          Ti = zi C - zr S
//   Z  = F T:         Zr = C Tr + S Ti          Zi = C Ti - S Tr
//   P  = 4 conj(F1) F2 from Z and its mirror, scaled by 2^sh (|P| < 256)
//   U  = F^H P:       Ur = C Pr - S Pi          Ui = C Pi + S Pr
//   cc = Re(U F^H) =  Ur C - Ui S
// Between products the FP32 results pass through an LDS plane pair (real and
// imaginary parts, 32 x 33 floats each) and are read back in the next
// operand layout, then rounded to FP16 or BF16 (hi, and lo = x - hi when
// split).  The plane then goes through the peak stage of the wave-per-window
// kernel, with the FP64 fit of flat peaks.
constexpr int kD32S = 33;  // plane row stride (floats)

typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
// two values rounded to nearest (one packed conversion), and back to FP32
template <bool BF>
__device__ __forceinline__ uint32_t pack2(float x, float y) {
  const f32x2_t v = {x, y};
  if constexpr (BF) return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
  else return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2_t));
}
template <bool BF>
__device__ __forceinline__ f32x2_t unpack2(uint32_t u) {
  if constexpr (BF) return f32x2_t{__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u)};
  else return __builtin_convertvector(__builtin_bit_cast(f16x2_t, u), f32x2_t);
}
// eight values as an MFMA operand: hi, and lo = v - hi (split)
template <bool BF, bool SPLIT>
__device__ __forceinline__ void operand8(const float (&v)[8], u32x4_t& hi, u32x4_t& lo) {
#pragma clang fp contract(off)
  static_for<4>([&](auto kc) {
    constexpr int k = kc;
    hi[k] = pack2<BF>(v[2 * k], v[2 * k + 1]);
    if constexpr (SPLIT) {
      const f32x2_t h = unpack2<BF>(hi[k]);
      lo[k] = pack2<BF>(v[2 * k] - h.x, v[2 * k + 1] - h.y);
    }
  });
}
__device__ __forceinline__ u32x4_t neg8(u32x4_t v) {  // sign of every half
  const uint32_t m = 0x80008000u;
  return u32x4_t{v[0] ^ m, v[1] ^ m, v[2] ^ m, v[3] ^ m};
}
// acc + A B with the split products hi*hi + hi*lo + lo*hi
template <bool BF, bool SPLIT>
__device__ __forceinline__ f32x4_t mm(f32x4_t acc, u32x4_t ah, u32x4_t al, u32x4_t bh, u32x4_t bl) {
  acc = DftTy<BF>::mma(ah, bh, acc);
  if constexpr (SPLIT) {
    acc = DftTy<BF>::mma(ah, bl, acc);
    acc = DftTy<BF>::mma(al, bh, acc);
  }
  return acc;
}

template <bool BF, bool SPLIT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_xcorr_dft32w(XcArgs a, const uint16_t* lanetab) {
#pragma clang fp contract(off)
  if (a.gate && *a.gate == 0) return;
  constexpr int N = 32, SX = kD32S;
  __shared__ __align__(16) float pr[N * SX], pi[N * SX];
  const int w = blockIdx.x;
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const int iy = w / a.nx_vec, ix = w - iy * a.nx_vec;
  const int ax0 = ix * a.sp_x, ay0 = iy * a.sp_y;
  // C (m = 0) and S (m = 1) in the lane layout: [m][t][hl] = M[16t + r][8g .. 8g+7]
  u32x4_t cst[2][2][2];
  {
    const u32x4_t* tab = reinterpret_cast<const u32x4_t*>(lanetab);
    static_for<8>([&](auto qc) {
      constexpr int q = qc;
      cst[q >> 2][(q >> 1) & 1][q & 1] = tab[q * 64 + lane];
    });
  }
  const u32x4_t z4 = {0u, 0u, 0u, 0u};
  auto C = [&](int t, int hl) { return cst[0][t][hl]; };
  auto S = [&](int t, int hl) { return cst[1][t][hl]; };
  // ---- 1. T = z F.  A operands: z rows 16 mt + r, columns 8g .. 8g+7
  u32x4_t zr[2][2], zi[2][2], nzr[2][2];
  static_for<2>([&](auto mc) {
    constexpr int mt = mc;
    float2 x[8];
    load_row<8>(a, ax0 + 8 * g, ay0 + 16 * mt + r, x);
    float vr[8], vi[8];
    static_for<8>([&](auto jc) { vr[jc] = x[jc].x; vi[jc] = x[jc].y; });
    operand8<BF, SPLIT>(vr, zr[mt][0], zr[mt][1]);
    operand8<BF, SPLIT>(vi, zi[mt][0], zi[mt][1]);
    nzr[mt][0] = neg8(zr[mt][0]);
    nzr[mt][1] = SPLIT ? neg8(zr[mt][1]) : z4;
  });
  // tiles (mt, nt) of a product, written to the planes in the output layout
  auto store = [&](const f32x4_t (&re)[2][2], const f32x4_t (&im)[2][2]) {
    static_for<2>([&](auto mc) {
      static_for<2>([&](auto nc) {
        constexpr int mt = mc, nt = nc;
        static_for<4>([&](auto ic) {
          constexpr int i = ic;
          const int e = (16 * mt + 4 * g + i) * SX + 16 * nt + r;
          pr[e] = re[mt][nt][i];
          pi[e] = im[mt][nt][i];
        });
      });
    });
  };
  {
    f32x4_t tr[2][2], ti[2][2];
    static_for<2>([&](auto mc) {
      static_for<2>([&](auto nc) {
        constexpr int mt = mc, nt = nc;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, zr[mt][0], zr[mt][1], C(nt, 0), C(nt, 1));
        tr[mt][nt] = mm<BF, SPLIT>(acc, zi[mt][0], zi[mt][1], S(nt, 0), S(nt, 1));
        acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, zi[mt][0], zi[mt][1], C(nt, 0), C(nt, 1));
        ti[mt][nt] = mm<BF, SPLIT>(acc, nzr[mt][0], nzr[mt][1], S(nt, 0), S(nt, 1));
      });
    });
    store(tr, ti);
  }
  wave_sync();
  // ---- 2. Z = F T.  B operands: T rows 8g .. 8g+7 of column 16 nt + r
  {
    u32x4_t br[2][2], bi[2][2], nbr[2][2];
    static_for<2>([&](auto nc) {
      constexpr int nt = nc;
      float vr[8], vi[8];
      static_for<8>([&](auto jc) {
        constexpr int j = jc;
        vr[j] = pr[(8 * g + j) * SX + 16 * nt + r];
        vi[j] = pi[(8 * g + j) * SX + 16 * nt + r];
      });
      operand8<BF, SPLIT>(vr, br[nt][0], br[nt][1]);
      operand8<BF, SPLIT>(vi, bi[nt][0], bi[nt][1]);
      nbr[nt][0] = neg8(br[nt][0]);
      nbr[nt][1] = SPLIT ? neg8(br[nt][1]) : z4;
    });
    wave_sync();  // the planes are read before they are rewritten
    f32x4_t zre[2][2], zim[2][2];
    static_for<2>([&](auto mc) {
      static_for<2>([&](auto nc) {
        constexpr int mt = mc, nt = nc;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, C(mt, 0), C(mt, 1), br[nt][0], br[nt][1]);
        zre[mt][nt] = mm<BF, SPLIT>(acc, S(mt, 0), S(mt, 1), bi[nt][0], bi[nt][1]);
        acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, C(mt, 0), C(mt, 1), bi[nt][0], bi[nt][1]);
        zim[mt][nt] = mm<BF, SPLIT>(acc, S(mt, 0), S(mt, 1), nbr[nt][0], nbr[nt][1]);
      });
    });
    store(zre, zim);
  }
  wave_sync();
  // ---- 3. P = 4 conj(F1) F2 at the B positions of step 4 (rows 8g + j,
  //      columns 16 nt + r) from Z and its mirror Z[-ky][-kx]; scaled by
  //      2^sh so that |P| < 256 (FP16 range for the next products)
  float2 P[2][8];
  float mx = 0.f;
  static_for<2>([&](auto nc) {
    constexpr int nt = nc;
    const int kx = 16 * nt + r, kxm = (N - kx) & (N - 1);
    static_for<8>([&](auto jc) {
      constexpr int j = jc;
      const int ky = 8 * g + j, kym = (N - ky) & (N - 1);
      const float2 zo = make_float2(pr[ky * SX + kx], pi[ky * SX + kx]);
      const float2 zm = make_float2(pr[kym * SX + kxm], pi[kym * SX + kxm]);
      P[nt][j] = xprod(zo, zm);
      mx = fmaxf(mx, fmaxf(fabsf(P[nt][j].x), fabsf(P[nt][j].y)));
    });
  });
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  int ex = 0;
  if (mx > 0.f) (void)frexpf(mx, &ex);  // mx < 2^ex
  const int sh = 8 - ex;
  u32x4_t ppr[2][2], ppi[2][2], npi[2][2];
  static_for<2>([&](auto nc) {
    constexpr int nt = nc;
    float vr[8], vi[8];
    static_for<8>([&](auto jc) {
      vr[jc] = ldexpf(P[nt][jc].x, sh);
      vi[jc] = ldexpf(P[nt][jc].y, sh);
    });
    operand8<BF, SPLIT>(vr, ppr[nt][0], ppr[nt][1]);
    operand8<BF, SPLIT>(vi, ppi[nt][0], ppi[nt][1]);
    npi[nt][0] = neg8(ppi[nt][0]);
    npi[nt][1] = SPLIT ? neg8(ppi[nt][1]) : z4;
  });
  wave_sync();  // every read of Z is done before U is stored
  // ---- 4. U = F^H P
  {
    f32x4_t ure[2][2], uim[2][2];
    static_for<2>([&](auto mc) {
      static_for<2>([&](auto nc) {
        constexpr int mt = mc, nt = nc;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, C(mt, 0), C(mt, 1), ppr[nt][0], ppr[nt][1]);
        ure[mt][nt] = mm<BF, SPLIT>(acc, S(mt, 0), S(mt, 1), npi[nt][0], npi[nt][1]);
        acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, C(mt, 0), C(mt, 1), ppi[nt][0], ppi[nt][1]);
        uim[mt][nt] = mm<BF, SPLIT>(acc, S(mt, 0), S(mt, 1), ppr[nt][0], ppr[nt][1]);
      });
    });
    store(ure, uim);
  }
  wave_sync();
  // ---- 5. cc = Re(U F^H) = Ur C - Ui S.  A operands: U rows 16 mt + r,
  //      columns 8g .. 8g+7
  f32x4_t cc4[2][2];
  {
    u32x4_t ar[2][2], nai[2][2];
    static_for<2>([&](auto mc) {
      constexpr int mt = mc;
      float vr[8], vi[8];
      static_for<8>([&](auto jc) {
        constexpr int j = jc;
        vr[j] = pr[(16 * mt + r) * SX + 8 * g + j];
        vi[j] = -pi[(16 * mt + r) * SX + 8 * g + j];
      });
      operand8<BF, SPLIT>(vr, ar[mt][0], ar[mt][1]);
      operand8<BF, SPLIT>(vi, nai[mt][0], nai[mt][1]);
    });
    wave_sync();  // the planes are read before the correlation plane is written
    static_for<2>([&](auto mc) {
      static_for<2>([&](auto nc) {
        constexpr int mt = mc, nt = nc;
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
        acc = mm<BF, SPLIT>(acc, ar[mt][0], ar[mt][1], C(nt, 0), C(nt, 1));
        cc4[mt][nt] = mm<BF, SPLIT>(acc, nai[mt][0], nai[mt][1], S(nt, 0), S(nt, 1));
      });
    });
  }
  // |.| / N (the exact correlation is real and >= 0), written fftshifted
  const float scale = ldexpf(0.25f / (float)(N * N), -sh);
  static_for<2>([&](auto mc) {
    static_for<2>([&](auto nc) {
      constexpr int mt = mc, nt = nc;
      static_for<4>([&](auto ic) {
        constexpr int i = ic;
        const int row = (16 * mt + 4 * g + i + N / 2) & (N - 1), col = (16 * nt + r + N / 2) & (N - 1);
        pr[row * SX + col] = fabsf(cc4[mt][nt][i]) * scale;
      });
    });
  });
  wave_sync();
  single_peaks<N, N, SX, true>(a, w, pr, lane);
}

// The same for 64x64 windows with one 256-thread workgroup (four waves) per
// window: K = 64 takes two 16x16x32 MFMAs per tile product, the output has
// 4 x 4 tiles, and the waves split each product by tile rows (steps with a
// constant B operand: wave w computes rows 16w..16w+15) or by tile columns
// (steps with a constant A operand: columns 16w..16w+15), so that every wave
// reads or converts only its own band of the data operand.  The DFT matrices
// are read in their lane layout straight from the table (L1/L2-resident):
// [C, S][hi, lo][tile t][k-step ks][lane][8] = M[16t + lane % 16][32 ks + 8 (lane / 16) + j].
template <bool BF, bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_xcorr_dft64w(XcArgs a, const uint16_t* lanetab) {
#pragma clang fp contract(off)
  if (a.gate && *a.gate == 0) return;
  constexpr int N = 64, SX = N + 1;
  __shared__ __align__(16) float pr[N * SX], pi[N * SX];
  __shared__ float smax[4];
  const int w = blockIdx.x;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, r = lane & 15, g = lane >> 4;
  const int iy = w / a.nx_vec, ix = w - iy * a.nx_vec;
  const int ax0 = ix * a.sp_x, ay0 = iy * a.sp_y;
  const u32x4_t* const tab = reinterpret_cast<const u32x4_t*>(lanetab);
  auto cst = [&](int m, int hl, int t, int ks) { return tab[(((m * 2 + hl) * 4 + t) * 2 + ks) * 64 + lane]; };
  const u32x4_t z4 = {0u, 0u, 0u, 0u};
  auto lo_or0 = [&](u32x4_t v) { return SPLIT ? v : z4; };
  auto put = [&](int mt, int nt, f32x4_t re, f32x4_t im) {
    static_for<4>([&](auto ic) {
      constexpr int i = ic;
      const int e = (16 * mt + 4 * g + i) * SX + 16 * nt + r;
      pr[e] = re[i];
      pi[e] = im[i];
    });
  };
  // ---- 1. T = z F, tile rows 16 wv + r: A = z (two k-steps), B = F columns
  {
    u32x4_t zr[2][2], zi[2][2], nzr[2][2];
    static_for<2>([&](auto kc) {
      constexpr int ks = kc;
      float2 x[8];
      load_row<8>(a, ax0 + 32 * ks + 8 * g, ay0 + 16 * wv + r, x);
      float vr[8], vi[8];
      static_for<8>([&](auto jc) { vr[jc] = x[jc].x; vi[jc] = x[jc].y; });
      operand8<BF, SPLIT>(vr, zr[ks][0], zr[ks][1]);
      operand8<BF, SPLIT>(vi, zi[ks][0], zi[ks][1]);
      nzr[ks][0] = neg8(zr[ks][0]);
      nzr[ks][1] = lo_or0(neg8(zr[ks][1]));
    });
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t tr = {0.f, 0.f, 0.f, 0.f}, ti = {0.f, 0.f, 0.f, 0.f};
      static_for<2>([&](auto kc) {
        constexpr int ks = kc;
        const u32x4_t ch = cst(0, 0, nt, ks), cl = lo_or0(cst(0, 1, nt, ks));
        const u32x4_t sh_ = cst(1, 0, nt, ks), sl = lo_or0(cst(1, 1, nt, ks));
        tr = mm<BF, SPLIT>(tr, zr[ks][0], zr[ks][1], ch, cl);
        tr = mm<BF, SPLIT>(tr, zi[ks][0], zi[ks][1], sh_, sl);
        ti = mm<BF, SPLIT>(ti, zi[ks][0], zi[ks][1], ch, cl);
        ti = mm<BF, SPLIT>(ti, nzr[ks][0], nzr[ks][1], sh_, sl);
      });
      put(wv, nt, tr, ti);
    }
  }
  __syncthreads();
  // ---- 2. Z = F T, tile columns 16 wv + r: A = F rows, B = T (rows 32 ks + 8g + j)
  {
    u32x4_t br[2][2], bi[2][2], nbr[2][2];
    static_for<2>([&](auto kc) {
      constexpr int ks = kc;
      float vr[8], vi[8];
      static_for<8>([&](auto jc) {
        constexpr int j = jc;
        vr[j] = pr[(32 * ks + 8 * g + j) * SX + 16 * wv + r];
        vi[j] = pi[(32 * ks + 8 * g + j) * SX + 16 * wv + r];
      });
      operand8<BF, SPLIT>(vr, br[ks][0], br[ks][1]);
      operand8<BF, SPLIT>(vi, bi[ks][0], bi[ks][1]);
      nbr[ks][0] = neg8(br[ks][0]);
      nbr[ks][1] = lo_or0(neg8(br[ks][1]));
    });
    __syncthreads();  // the T planes are read before Z is stored
    for (int mt = 0; mt < 4; ++mt) {
      f32x4_t zre = {0.f, 0.f, 0.f, 0.f}, zim = {0.f, 0.f, 0.f, 0.f};
      static_for<2>([&](auto kc) {
        constexpr int ks = kc;
        const u32x4_t ch = cst(0, 0, mt, ks), cl = lo_or0(cst(0, 1, mt, ks));
        const u32x4_t sh_ = cst(1, 0, mt, ks), sl = lo_or0(cst(1, 1, mt, ks));
        zre = mm<BF, SPLIT>(zre, ch, cl, br[ks][0], br[ks][1]);
        zre = mm<BF, SPLIT>(zre, sh_, sl, bi[ks][0], bi[ks][1]);
        zim = mm<BF, SPLIT>(zim, ch, cl, bi[ks][0], bi[ks][1]);
        zim = mm<BF, SPLIT>(zim, sh_, sl, nbr[ks][0], nbr[ks][1]);
      });
      put(mt, wv, zre, zim);
    }
  }
  __syncthreads();
  // ---- 3. P = 4 conj(F1) F2 at rows 32 ks + 8g + j of column 16 wv + r,
  //      scaled by 2^sh so that |P| < 256 over the whole window
  float2 P[2][8];
  float mx = 0.f;
  {
    const int kx = 16 * wv + r, kxm = (N - kx) & (N - 1);
    static_for<2>([&](auto kc) {
      constexpr int ks = kc;
      static_for<8>([&](auto jc) {
        constexpr int j = jc;
        const int ky = 32 * ks + 8 * g + j, kym = (N - ky) & (N - 1);
        const float2 zo = make_float2(pr[ky * SX + kx], pi[ky * SX + kx]);
        const float2 zm = make_float2(pr[kym * SX + kxm], pi[kym * SX + kxm]);
        P[ks][j] = xprod(zo, zm);
        mx = fmaxf(mx, fmaxf(fabsf(P[ks][j].x), fabsf(P[ks][j].y)));
      });
    });
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) smax[wv] = mx;
  __syncthreads();  // also: every read of Z is done before U is stored
  mx = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
  int ex = 0;
  if (mx > 0.f) (void)frexpf(mx, &ex);  // mx < 2^ex
  const int sh = 8 - ex;
  // ---- 4. U = F^H P, tile columns 16 wv + r
  {
    u32x4_t ppr[2][2], ppi[2][2], npi[2][2];
    static_for<2>([&](auto kc) {
      constexpr int ks = kc;
      float vr[8], vi[8];
      static_for<8>([&](auto jc) {
        vr[jc] = ldexpf(P[ks][jc].x, sh);
        vi[jc] = ldexpf(P[ks][jc].y, sh);
      });
      operand8<BF, SPLIT>(vr, ppr[ks][0], ppr[ks][1]);
      operand8<BF, SPLIT>(vi, ppi[ks][0], ppi[ks][1]);
      npi[ks][0] = neg8(ppi[ks][0]);
      npi[ks][1] = lo_or0(neg8(ppi[ks][1]));
    });
    for (int mt = 0; mt < 4; ++mt) {
      f32x4_t ure = {0.f, 0.f, 0.f, 0.f}, uim = {0.f, 0.f, 0.f, 0.f};
      static_for<2>([&](auto kc) {
        constexpr int ks = kc;
        const u32x4_t ch = cst(0, 0, mt, ks), cl = lo_or0(cst(0, 1, mt, ks));
        const u32x4_t sh_ = cst(1, 0, mt, ks), sl = lo_or0(cst(1, 1, mt, ks));
        ure = mm<BF, SPLIT>(ure, ch, cl, ppr[ks][0], ppr[ks][1]);
        ure = mm<BF, SPLIT>(ure, sh_, sl, npi[ks][0], npi[ks][1]);
        uim = mm<BF, SPLIT>(uim, ch, cl, ppi[ks][0], ppi[ks][1]);
        uim = mm<BF, SPLIT>(uim, sh_, sl, ppr[ks][0], ppr[ks][1]);
      });
      put(mt, wv, ure, uim);
    }
  }
  __syncthreads();
  // ---- 5. cc = Re(U F^H) = Ur C - Ui S, tile rows 16 wv + r: A = U rows
  {
    u32x4_t ar[2][2], nai[2][2];
    static_for<2>([&](auto kc) {
      constexpr int ks = kc;
      float vr[8], vi[8];
      static_for<8>([&](auto jc) {
        constexpr int j = jc;
        vr[j] = pr[(16 * wv + r) * SX + 32 * ks + 8 * g + j];
        vi[j] = -pi[(16 * wv + r) * SX + 32 * ks + 8 * g + j];
      });
      operand8<BF, SPLIT>(vr, ar[ks][0], ar[ks][1]);
      operand8<BF, SPLIT>(vi, nai[ks][0], nai[ks][1]);
    });
    __syncthreads();  // the U planes are read before the correlation plane is written
    const float scale = ldexpf(0.25f / (float)(N * N), -sh);
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t cc = {0.f, 0.f, 0.f, 0.f};
      static_for<2>([&](auto kc) {
        constexpr int ks = kc;
        cc = mm<BF, SPLIT>(cc, ar[ks][0], ar[ks][1], cst(0, 0, nt, ks), lo_or0(cst(0, 1, nt, ks)));
        cc = mm<BF, SPLIT>(cc, nai[ks][0], nai[ks][1], cst(1, 0, nt, ks), lo_or0(cst(1, 1, nt, ks)));
      });
      // |.| / N^2 (the exact correlation is real and >= 0), written fftshifted
      static_for<4>([&](auto ic) {
        constexpr int i = ic;
        const int row = (16 * wv + 4 * g + i + N / 2) & (N - 1), col = (16 * nt + r + N / 2) & (N - 1);
        pr[row * SX + col] = fabsf(cc[i]) * scale;
      });
    }
  }
  __syncthreads();
  if (wv == 0) single_peaks<N, N, SX, true>(a, w, pr, lane);
}

// ============================================================== deformation
// One thread per aug column, DEF_ROWS consecutive rows.  The row terms are
// computed once per workgroup into LDS and the column terms once per thread
// (deform_terms); the four field samples of a cell are reloaded only when
// the row crosses into the next cell; each pixel is deform_px.
constexpr int DEF_ROWS = 16;
// Grid caps of the kernels the learned pipeline launches behind its fallback
// gate: an empty launch then dispatches a few hundred workgroups instead of
// one per window or band.
constexpr int kGatedBands = 32;
constexpr int kGatedWindows = 1024;

// One block: aug rows y0 .. y0 + DEF_ROWS - 1, columns 256 bx .. 256 bx + 255.
__device__ __forceinline__ void deform_block(const DefArgs& a, int y0, int bx, float4* rowt) {
  const int nrow = min(DEF_ROWS, a.ny_aug - y0);
  if (threadIdx.x < nrow) rowt[threadIdx.x] = deform_terms(y0 + (int)threadIdx.x, a.y_st, a.spy, a.nye);
  __syncthreads();
  const int idx = bx * 256 + (int)threadIdx.x;
  if (idx >= a.nx_aug) return;
  const float4 ct = deform_terms(idx, a.x_st, a.spx, a.nxe);
  const int ixf = __float_as_int(ct.z);
  const DefFrame F = def_frame(a.f2, a.pitch, a.nx, a.ny, a.left_x, a.left_y, a.hfws_x, a.hfws_y, a.nx_aug, a.ny_aug);
  const float oob_guard = (float)(min(a.hfws_x, a.hfws_y) - 2);
  unsigned oob = 0;
  int cell = -1;
  DefCell e{};
  // this block's output rows as a buffer resource: row r of pixel idx at
  // byte r * nx_aug * 4 (scalar offset) + idx * 4
  const unsigned row_bytes = (unsigned)a.nx_aug * 4u;
  // stores carry the non-temporal (streaming) cache policy: the 117 MB plane
  // of a grid-2 pass is written once and read once, by the next kernel
  constexpr int kDefStoreNT = 2;
  const __amdgpu_buffer_rsrc_t outr = __builtin_amdgcn_make_buffer_rsrc(a.out + (long long)y0 * a.nx_aug, (short)0,
                                                                        (int)(row_bytes * (unsigned)nrow), 0x00020000);
  // row terms read at constant offsets from one base register
  int rbase = 0;
  asm volatile("" : "+v"(rbase));
  const float4* const rowp = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(rowt) + rbase);
  auto row = [&](int r) {
    const float4 rt = rowp[r];
    const int iyf = __builtin_amdgcn_readfirstlane(__float_as_int(rt.z));
    if (iyf != cell) {
      cell = iyf;
      e = load_cell(a.ex_x, a.ex_y, a.nxe, ixf, iyf, oob_guard);
    }
    const float v = e.may_oob ? deform_px<true>(F, e, ct.x, ct.y, ct.w, rt.x, rt.y, rt.w, oob)
                              : deform_px<false>(F, e, ct.x, ct.y, ct.w, rt.x, rt.y, rt.w, oob);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), outr, (int)(idx * 4), (int)(r * row_bytes), kDefStoreNT);
  };
  // full blocks unrolled (row terms at constant LDS offsets), in groups of 8
  // rows: a group inside one cell whose samples stay below the guard has no
  // branch, and its 32 frame taps per lane are in flight together
  if (nrow == DEF_ROWS) {
    static_for<DEF_ROWS / 8>([&](auto gc) {
      constexpr int r0 = 8 * gc;
      float4 rt[8];
      static_for<8>([&](auto k) { rt[k] = rowp[r0 + k]; });
      const int iy0 = __builtin_amdgcn_readfirstlane(__float_as_int(rt[0].z));
      const int iy7 = __builtin_amdgcn_readfirstlane(__float_as_int(rt[7].z));
      if (iy0 == iy7) {  // cells are monotone in the row
        if (iy0 != cell) {
          cell = iy0;
          e = load_cell(a.ex_x, a.ex_y, a.nxe, ixf, iy0, oob_guard);
        }
        float v[8];
        if (__builtin_expect(e.may_oob, 0)) {
          static_for<8>([&](auto k) { v[k] = deform_px<true>(F, e, ct.x, ct.y, ct.w, rt[k].x, rt[k].y, rt[k].w, oob); });
        } else {
          static_for<8>([&](auto k) { v[k] = deform_px<false>(F, e, ct.x, ct.y, ct.w, rt[k].x, rt[k].y, rt[k].w, oob); });
        }
        static_for<8>([&](auto k) {
          __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[k]), outr, (int)(idx * 4), (int)((r0 + k) * row_bytes), kDefStoreNT);
        });
      } else {
        static_for<8>([&](auto k) { row(r0 + k); });
      }
    });
  } else {
    for (int r = 0; r < nrow; ++r) row(r);
  }
  if (oob) atomicAdd(a.oob, (unsigned long long)oob);
}

// Blocks in row bands: gridDim.y is the band count, or fewer when the launch
// is gated (the learned pipeline's fallback, which mostly does nothing: a
// small grid keeps its empty launches cheap) and the bands are walked.
__global__ void __launch_bounds__(256) k_deform(DefArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  __shared__ float4 rowt[DEF_ROWS];
  const int nby = (a.ny_aug + DEF_ROWS - 1) / DEF_ROWS;
  for (int by = blockIdx.y; by < nby; by += gridDim.y) {
    if (by != (int)blockIdx.y) __syncthreads();  // the last band's row terms are read
    deform_block(a, by * DEF_ROWS, blockIdx.x, rowt);
  }
}

// The per-column and per-row terms of k_deform as tables (FusedDef).
__global__ void k_deform_terms(DefArgs a, float4* col, float4* row) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.nx_aug) col[t] = deform_terms(t, a.x_st, a.spx, a.nxe);
  if (t < a.ny_aug) row[t] = deform_terms(t, a.y_st, a.spy, a.nye);
}

// ============================================================== field extension
// extend_filtered_vecfield_p1/p2 (cu:1547-1618): copy into an (nx+2b)x(ny+2b)
// field with the border replicated b times (b = 3, cuh:190-191).
__global__ void k_extend(const float* fx, const float* fy, float* ex, float* ey, int nx, int ny, int b,
                         const int* gate) {
  if (gate && *gate == 0) return;
  const int nxe = nx + 2 * b, nye = ny + 2 * b;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nxe * nye; t += gridDim.x * blockDim.x) {
    const int j = t / nxe, i = t - (t / nxe) * nxe;
    const int si = min(max(i - b, 0), nx - 1), sj = min(max(j - b, 0), ny - 1);
    ex[t] = fx[sj * nx + si];
    ey[t] = fy[sj * nx + si];
  }
}

// The extension entries (i, j) that replicate vector (vx, vy) — the same
// values k_extend writes, from the thread that produced the vector.
__device__ __forceinline__ void write_ext(float* ex, float* ey, int nx, int ny, int b, int vx, int vy, float x,
                                          float y) {
  const int nxe = nx + 2 * b;
  const int i0 = vx == 0 ? 0 : vx + b, i1 = vx == nx - 1 ? nxe - 1 : vx + b;
  const int j0 = vy == 0 ? 0 : vy + b, j1 = vy == ny - 1 ? ny + 2 * b - 1 : vy + b;
  for (int j = j0; j <= j1; ++j)
    for (int i = i0; i <= i1; ++i) {
      ex[j * nxe + i] = x;
      ey[j * nxe + i] = y;
    }
}

// ============================================================== upsampling
// Upsampling_vecfield_kernel_2d (cu:675-704): grid-2 predictor at image
// position (k*sp2 + sp2/2) from the extended grid-1 field (origin x_st1).  The
// bilinear sums (cu:700-701) are fma(c4, e4, fma(c3, e3, fma(c1, e1, c2*e2))),
// the chain a CUDA build forms from that source by default (--fmad=true).
__global__ void k_upsample(float* dx2, float* dy2, const float* ex, const float* ey, int nxe, int nye, float x_st1,
                           float y_st1, float sp1x, float sp1y, int sp2x, int sp2y, int nx2, int ny2, float* ex2x,
                           float* ex2y, int border) {
#pragma clang fp contract(off)
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nx2 * ny2; t += gridDim.x * blockDim.x) {
    const int vy = t / nx2, vx = t - (t / nx2) * nx2;
    const int idx = vx * sp2x + sp2x / 2, idy = vy * sp2y + sp2y / 2;
    float ax = ((float)idx - x_st1) / sp1x;
    float ay = ((float)idy - y_st1) / sp1y;
    int ix = (int)floorf(ax), iy = (int)floorf(ay);
    ax = ax - (float)ix;
    ay = ay - (float)iy;
    const float ce1 = (1 - ax) * (1 - ay), ce2 = ax * (1 - ay), ce3 = (1 - ax) * ay, ce4 = ax * ay;
    ix = min(max(ix, 0), nxe - 2);
    iy = min(max(iy, 0), nye - 2);
    const int i1 = iy * nxe + ix, i2 = i1 + 1, i3 = i1 + nxe, i4 = i3 + 1;
    const float ux = fmaf(ce4, ex[i4], fmaf(ce3, ex[i3], fmaf(ce1, ex[i1], ce2 * ex[i2])));
    const float uy = fmaf(ce4, ey[i4], fmaf(ce3, ey[i3], fmaf(ce1, ey[i1], ce2 * ey[i2])));
    dx2[t] = ux;
    dy2[t] = uy;
    if (ex2x) write_ext(ex2x, ex2y, nx2, ny2, border, vx, vy, ux, uy);
  }
}

// ============================================================== median test
// One thread per vector.  Neighbours (valid entries of the current field in
// the wx*wy window, skipping slot (wx*wy-1)/2, cu:255-283) are held in a
// fixed slot array padded with +inf, sorted by a bitonic network; the exact
// median is the same order statistic the reference's partial bubble sort
// produces (cu:174-207).  Deviations |v - median| (cu:209-223), their median,
// then the candidate test of median_filtering (cu:225-253).
template <int NS>
__device__ __forceinline__ void bitonic_sort(float (&v)[NS]) {
#pragma unroll
  for (int k = 2; k <= NS; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const float lo = fminf(v[i], v[l]), hi = fmaxf(v[i], v[l]);
          if ((i & k) == 0) {
            v[i] = lo;
            v[l] = hi;
          } else {
            v[i] = hi;
            v[l] = lo;
          }
        }
      }
    }
  }
}

template <int NS>
__device__ __forceinline__ float select_sorted(const float (&s)[NS], int idx) {
  float r = 0.f;
#pragma unroll
  for (int i = 0; i < NS; ++i)
    if (i == idx) r = s[i];
  return r;
}

template <int NS>
__device__ __forceinline__ float median_of(float (&s)[NS], int n) {
#pragma clang fp contract(off)
  bitonic_sort<NS>(s);
  if (n <= 0) return 0.f;  // reference reads outside the list here (cu:199); defined as 0
  if (n % 2 == 0) return (select_sorted<NS>(s, n / 2) + select_sorted<NS>(s, n / 2 - 1)) / 2;
  return select_sorted<NS>(s, (n - 1) / 2);
}

// The candidate test of median_filtering (cu:237-248): both normalised
// residuals |d - median| / (residual median + eps) below the threshold.  With
// a.fp64 (the precision form) the normalisation is evaluated in FP64.
__device__ __forceinline__ bool median_accepts(const MedArgs& a, float cx, float cy, float mx, float my, float rx,
                                               float ry) {
#pragma clang fp contract(off)
  if (a.fp64) {
    const double dx = fabs((double)cx - (double)mx) / ((double)rx + (double)a.eps);
    const double dy = fabs((double)cy - (double)my) / ((double)ry + (double)a.eps);
    return dx < (double)a.thr && dy < (double)a.thr;
  }
  const float devix = fabsf(cx - mx) / (rx + a.eps);
  const float deviy = fabsf(cy - my) / (ry + a.eps);
  return devix < a.thr && deviy < a.thr;
}

__device__ __forceinline__ void median_decide(const MedArgs& a, int t, bool vmed, float mx, float my, float rx,
                                              float ry) {
#pragma clang fp contract(off)
  float ox = 0.f, oy = 0.f;
  bool ok = false;
  int acc = -1;
  if (vmed) {
    for (int k = 0; k < a.K; ++k) {
      const long long o = (long long)t * a.K + k;
      if (!a.cv[o]) continue;
      if (!median_accepts(a, a.cx[o], a.cy[o], mx, my, rx, ry)) continue;
      {
        ok = true;
        ox = a.cx[o];
        oy = a.cy[o];
        acc = k;
        break;
      }
    }
  }
  a.dx[t] = ox;
  a.dy[t] = oy;
  a.dv[t] = ok ? 1 : 0;
  if (a.dx2) {
    a.dx2[t] = ox;
    a.dy2[t] = oy;
    a.dv2[t] = ok ? 1 : 0;
  }
  if (a.accepted) a.accepted[t] = (int8_t)acc;
  if (a.hole_list && !ok) {
    const int slot = atomicAdd(a.nholes, 1);
    if (slot < a.hole_cap) a.hole_list[slot] = t;
  }
}

template <int NS>
__global__ void __launch_bounds__(128) k_median(MedArgs a) {
#pragma clang fp contract(off)
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  const int nt = a.nx_vec * a.ny_vec;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.reset_count && t == 0) *a.reset_count = 0;
  if (t >= nt) return;
  const int vy = t / a.nx_vec, vx = t - (t / a.nx_vec) * a.nx_vec;
  const int ntm = a.wx * a.wy;
  const int skip = (ntm - 1) / 2;
  float X[NS], Y[NS];
  int n = 0;
  int col = 0, row = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    float xv = INFINITY, yv = INFINITY;
    if (s < ntm && s != skip) {
      const int qx = vx - a.hfwx + col, qy = vy - a.hfwy + row;
      if (qx >= 0 && qx < a.nx_vec && qy >= 0 && qy < a.ny_vec) {
        const long long q = (long long)(qy * a.nx_vec + qx) * a.sstride;
        if (a.sv[q]) {
          xv = a.sx[q];
          yv = a.sy[q];
          ++n;
        }
      }
    }
    X[s] = xv;
    Y[s] = yv;
    if (++col == a.wx) {
      col = 0;
      ++row;
    }
  }
  const bool vmed = n >= a.min_nneigh;
  float mx = 0.f, my = 0.f, rx = 0.f, ry = 0.f;
  if (vmed) {
    mx = median_of<NS>(X, n);
    my = median_of<NS>(Y, n);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      X[s] = s < n ? fabsf(X[s] - mx) : INFINITY;
      Y[s] = s < n ? fabsf(Y[s] - my) : INFINITY;
    }
    rx = median_of<NS>(X, n);
    ry = median_of<NS>(Y, n);
  }
  median_decide(a, t, vmed, mx, my, rx, ry);
}

// Compile-time median windows (3x3, 5x5, 7x7): two lanes per vector, lane
// c = 0 ranks the x components and c = 1 the y components.  The M = wx*wy-1
// neighbour slots (the window without its centre slot, in raster order) hold
// valid values or +inf; a pruned odd-even merge network leaves the order
// statistics 0..M/2 in place, which contain the median of any count of valid
// neighbours.  Deviations |v - median| of the finite slots go through the same
// network.  The lanes swap their two medians and both run the candidate test
// of median_filtering (cu:225-253) on identical inputs.
template <int M>
__device__ __forceinline__ void select_net(float (&v)[MedianNet<M>::NS]) {
  static_for<MedianNet<M>::net.n>([&](auto cc) {
    constexpr int x = MedianNet<M>::net.a[cc], y = MedianNet<M>::net.b[cc];
    const float lo = fminf(v[x], v[y]), hi = fmaxf(v[x], v[y]);
    v[x] = lo;
    v[y] = hi;
  });
}

template <int M>
__device__ __forceinline__ float median_low(const float (&v)[MedianNet<M>::NS], int n) {
#pragma clang fp contract(off)
  // order statistic idx of the sorted prefix 0..M/2
  auto at = [&](int idx) {
    float r = 0.f;
    static_for<M / 2 + 1>([&](auto ic) { r = (ic == idx) ? v[ic] : r; });
    return r;
  };
  if (n <= 0) return 0.f;  // reference reads outside the list here (cu:199); defined as 0
  if (n % 2 == 0) return (at(n / 2) + at(n / 2 - 1)) / 2;
  return at((n - 1) / 2);
}

// Lane g of the pass: vector g / 2, component g % 2 (both lanes of a vector
// sit in one wave, as neighbours).
template <int WXM, int WYM>
__device__ __forceinline__ bool median2_lane(const MedArgs& a, int g) {
#pragma clang fp contract(off)
  constexpr int M = WXM * WYM - 1;
  constexpr int NS = MedianNet<M>::NS;
  constexpr int SKIP = (WXM * WYM - 1) / 2;
  constexpr int HX = (WXM - 1) / 2, HY = (WYM - 1) / 2;
  const int nt = a.nx_vec * a.ny_vec;
  const int t = g >> 1, c = g & 1;
  const bool live = t < nt;
  const int tt = live ? t : nt - 1;  // both lanes of a pair stay active for the swap
  const int vy = tt / a.nx_vec, vx = tt - vy * a.nx_vec;
  const float* src = c ? a.sy : a.sx;
  float v[NS];
  int n = 0;
  static_for<NS>([&](auto kc) {
    constexpr int k = kc;
    if constexpr (k >= M) {
      v[k] = INFINITY;
    } else {
      constexpr int s = k < SKIP ? k : k + 1;
      const int qx = vx - HX + s % WXM, qy = vy - HY + s / WXM;
      const bool inb = (unsigned)qx < (unsigned)a.nx_vec && (unsigned)qy < (unsigned)a.ny_vec;
      const long long q = (long long)(inb ? qy * a.nx_vec + qx : 0) * a.sstride;
      const bool ok = inb && a.sv[q];
      const float val = src[q];
      v[k] = ok ? val : INFINITY;
      n += ok ? 1 : 0;
    }
  });
  float med = 0.f, dev = 0.f;
  const bool vmed = n >= a.min_nneigh;
  if (vmed) {
    select_net<M>(v);
    med = median_low<M>(v, n);
    static_for<M>([&](auto kc) { v[kc] = v[kc] < INFINITY ? fabsf(v[kc] - med) : INFINITY; });
    select_net<M>(v);
    dev = median_low<M>(v, n);
  }
  const float omed = __shfl_xor(med, 1), odev = __shfl_xor(dev, 1);
  if (!live) return false;
  const float mx = c ? omed : med, my = c ? med : omed;
  const float rx = c ? odev : dev, ry = c ? dev : odev;
  float ox = 0.f, oy = 0.f;
  bool okc = false;
  int acc = -1;
  if (vmed) {
    for (int k = 0; k < a.K; ++k) {
      const long long o = (long long)t * a.K + k;
      const bool cvk = a.cv[o] != 0;
      const float cxk = a.cx[o], cyk = a.cy[o];
      if (!cvk) continue;
      if (median_accepts(a, cxk, cyk, mx, my, rx, ry)) {
        okc = true;
        ox = cxk;
        oy = cyk;
        acc = k;
        break;
      }
    }
  }
  if (c == 0) {
    a.dx[t] = ox;
    a.dv[t] = okc ? 1 : 0;
    if (a.dx2) {
      a.dx2[t] = ox;
      a.dv2[t] = okc ? 1 : 0;
    }
    if (a.accepted) a.accepted[t] = (int8_t)acc;
    if (a.hole_list && !okc) {
      const int slot = atomicAdd(a.nholes, 1);
      if (slot < a.hole_cap) a.hole_list[slot] = t;
    }
  } else {
    a.dy[t] = oy;
    if (a.dy2) a.dy2[t] = oy;
  }
  return c == 0 && !okc;  // this lane's vector is left invalid
}

template <int WXM, int WYM>
__global__ void __launch_bounds__(256) k_median2(MedArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  if (a.reset_count && blockIdx.x == 0 && threadIdx.x == 0) *a.reset_count = 0;
  // neighbouring blocks read mostly the same rows: keep them on one XCD's L2
  const bool rejected = median2_lane<WXM, WYM>(a, xcd_remap(blockIdx.x, gridDim.x) * 256 + (int)threadIdx.x);
  if (a.gate_out) {
    // outlier gate: block counts into *gate_acc; the last block to finish
    // (ticket) reads the total, sets the gate and clears both counters
    const int n = __syncthreads_count(rejected);
    if (threadIdx.x == 0) {
      atomicAdd(a.gate_acc, n);
      __threadfence();
      if (atomicAdd(a.gate_ticket, 1u) == gridDim.x - 1) {
        const int total = atomicExch(a.gate_acc, 0);
        *a.gate_count = total;
        *a.gate_out = total > a.gate_max ? 1 : 0;
        atomicExch(a.gate_ticket, 0u);
      }
    }
  }
}

// Windows larger than the register path: per-vector lists in global scratch,
// insertion sort (any wx*wy the reference accepts).
__device__ float median_scratch(float* s, int n) {
#pragma clang fp contract(off)
  for (int i = 1; i < n; ++i) {
    const float v = s[i];
    int j = i - 1;
    while (j >= 0 && s[j] > v) {
      s[j + 1] = s[j];
      --j;
    }
    s[j + 1] = v;
  }
  if (n <= 0) return 0.f;
  if (n % 2 == 0) return (s[n / 2] + s[n / 2 - 1]) / 2;
  return s[(n - 1) / 2];
}

__global__ void __launch_bounds__(128) k_median_generic(MedArgs a) {
#pragma clang fp contract(off)
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  const int nt = a.nx_vec * a.ny_vec;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (a.reset_count && t == 0) *a.reset_count = 0;
  if (t >= nt) return;
  const int vy = t / a.nx_vec, vx = t - (t / a.nx_vec) * a.nx_vec;
  const int ntm = a.wx * a.wy;
  float* X = a.scratch + (long long)t * 2 * ntm;
  float* Y = X + ntm;
  int n = 0;
  for (int s = 0; s < ntm; ++s) {
    if (s == (ntm - 1) / 2) continue;
    const int qx = vx - a.hfwx + s % a.wx, qy = vy - a.hfwy + s / a.wx;
    if (qx >= 0 && qx < a.nx_vec && qy >= 0 && qy < a.ny_vec) {
      const long long q = (long long)(qy * a.nx_vec + qx) * a.sstride;
      if (a.sv[q]) {
        X[n] = a.sx[q];
        Y[n] = a.sy[q];
        ++n;
      }
    }
  }
  const bool vmed = n >= a.min_nneigh;
  float mx = 0.f, my = 0.f, rx = 0.f, ry = 0.f;
  if (vmed) {
    mx = median_scratch(X, n);
    my = median_scratch(Y, n);
    for (int s = 0; s < n; ++s) {
      X[s] = fabsf(X[s] - mx);
      Y[s] = fabsf(Y[s] - my);
    }
    rx = median_scratch(X, n);
    ry = median_scratch(Y, n);
  }
  median_decide(a, t, vmed, mx, my, rx, ry);
}

// ============================================================== hole filling
// Jacobi sweeps of fill_blank_vecfield (cu:352-397) until a sweep fills
// nothing, run by one workgroup over a compacted hole list so that no host
// round trip is needed per sweep (the reference copies a counter back per
// sweep, cu:1527).  Each sweep first computes every hole's 3x3 mean of valid
// neighbours (raster sum order), then applies all fills at once.
// Hole list of the field (any order: each sweep's result is independent of
// the order holes are visited in).  *nholes is 0 on entry (the launch ahead of
// the listing one zeroes it); k_fill resets it as well.  A count beyond
// list_cap cannot arise from a pass that starts at 0 (list_cap >= every
// grid's vector count); k_fill adds any excess to *drops, which the host
// reports as an error.
__device__ __forceinline__ void fill_list_range(const FillArgs& a, int first, int stride) {
  const int nt = a.nx_vec * a.ny_vec;
  for (int t = first; t < nt; t += stride)
    if (!a.dv[t]) {
      const int slot = atomicAdd(a.nholes, 1);
      if (slot < a.list_cap) a.list_a[slot] = t;
    }
}

__global__ void __launch_bounds__(256) k_fill_list(FillArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  fill_list_range(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// 256 threads (four waves): a single larger workgroup waited up to 175 us for
// a CU with room beside concurrent correlation kernels of other handles; at
// four waves it is placed at once (longest of 750 launches 8 us).
// All sweeps, by every thread of the workgroup; sh is workgroup-shared.
struct FillShared {
  int n_holes, n_next, n_filled;
};
__device__ void fill_sweeps(const FillArgs& a, FillShared& sh) {
#pragma clang fp contract(off)
  int& n_holes = sh.n_holes;
  int& n_next = sh.n_next;
  int& n_filled = sh.n_filled;
  if (threadIdx.x == 0) {
    const int counted = *a.nholes;
    n_holes = min(counted, a.list_cap);
    if (counted > a.list_cap && a.drops) atomicAdd(a.drops, (unsigned int)(counted - a.list_cap));
    n_filled = 0;
  }
  __syncthreads();
  int* cur = a.list_a;
  int* nxt = a.list_b;
  int sweeps = 0;
  while (true) {
    const int nh = n_holes;
    ++sweeps;  // the reference runs one more (empty) sweep before its loop ends
    if (nh == 0) break;
    for (int h = threadIdx.x; h < nh; h += blockDim.x) {
      const int t = cur[h];
      const int vy = t / a.nx_vec, vx = t - (t / a.nx_vec) * a.nx_vec;
      int cnt = 0;
      float sx = 0.f, sy = 0.f;
      for (int j = max(0, vy - 1); j <= min(a.ny_vec - 1, vy + 1); ++j)
        for (int k = max(0, vx - 1); k <= min(a.nx_vec - 1, vx + 1); ++k) {
          const int q = j * a.nx_vec + k;
          if (a.dv[q]) {
            ++cnt;
            sx += a.dx[q];
            sy += a.dy[q];
          }
        }
      a.tmp_f[h] = cnt > 0;
      if (cnt > 0) {
        a.tmp_x[h] = sx / (float)cnt;
        a.tmp_y[h] = sy / (float)cnt;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) n_next = 0;
    __syncthreads();
    int filled = 0;
    for (int h = threadIdx.x; h < nh; h += blockDim.x) {
      const int t = cur[h];
      if (a.tmp_f[h]) {
        a.dx[t] = a.tmp_x[h];
        a.dy[t] = a.tmp_y[h];
        a.dv[t] = 1;
        ++filled;
      } else {
        nxt[atomicAdd(&n_next, 1)] = t;
      }
    }
    if (filled) atomicAdd(&n_filled, filled);
    __syncthreads();
    const bool progress = n_filled > 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      n_holes = n_next;
      n_filled = 0;
    }
    __syncthreads();
    if (!progress) break;
    int* tt = cur;
    cur = nxt;
    nxt = tt;
  }
  if (threadIdx.x == 0) {
    if (a.sweeps) *a.sweeps = sweeps;
    *a.nholes = 0;
  }
}

__global__ void __launch_bounds__(256) k_fill(FillArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  __shared__ FillShared sh;
  fill_sweeps(a, sh);
}

// ============================================================== smoothing
// organize_smooth_filter_data + smoothing_tophat (cu:399-443): mean over the
// in-bounds wx*wy window including the centre, ignoring validity; raster sum.
__device__ __forceinline__ void smooth_range(const SmoothArgs& a, int first, int stride) {
#pragma clang fp contract(off)
  const int nt = a.nx_vec * a.ny_vec;
  for (int t = first; t < nt; t += stride) {
    const int vy = t / a.nx_vec, vx = t - (t / a.nx_vec) * a.nx_vec;
    float sx = 0.f, sy = 0.f;
    int n = 0;
    for (int r = 0; r < a.wy; ++r) {
      const int qy = vy - a.hfwy + r;
      if (qy < 0 || qy >= a.ny_vec) continue;
      for (int c = 0; c < a.wx; ++c) {
        const int qx = vx - a.hfwx + c;
        if (qx < 0 || qx >= a.nx_vec) continue;
        sx += a.sx[qy * a.nx_vec + qx];
        sy += a.sy[qy * a.nx_vec + qx];
        ++n;
      }
    }
    const float ox = sx / (float)n, oy = sy / (float)n;
    a.dx[t] = ox;
    a.dy[t] = oy;
    if (a.ex) write_ext(a.ex, a.ey, a.nx_vec, a.ny_vec, a.border, vx, vy, ox, oy);
  }
}

__global__ void k_smooth(SmoothArgs a) {
  if (a.gate && *a.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  smooth_range(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__device__ __forceinline__ void prep_range(const PrepArgs& p, int first, int stride) {
  for (int t = first; t < p.nt; t += stride) {
    const long long o = (long long)t * p.K;
    p.dv[t] = p.cv[o];
    p.dx[t] = p.cx[o];
    p.dy[t] = p.cy[o];
  }
}

__global__ void k_prep(PrepArgs p, const int* gate) {
  if (gate && *gate == 0) return;
  if (p.reset_count && blockIdx.x == 0 && threadIdx.x == 0) *p.reset_count = 0;
  prep_range(p, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The steps of a filter sequence in one workgroup (FilterSeq), a barrier
// after each; __syncthreads makes each step's global writes visible to the
// whole workgroup before the next step reads them.
constexpr int kFilterSeqThreads = 1024;
template <int WXM, int WYM>
__device__ void filter_seq_run(const FilterSeq& q, FillShared& sh, int nthreads) {
  const int tid = (int)threadIdx.x;
  for (int i = 0; i < q.nsteps; ++i) {
    const int kind = q.step[i].kind;
    if (kind == FS_PREP) {
      const PrepArgs p = q.step[i].prep;
      if (p.reset_count && tid == 0) *p.reset_count = 0;
      prep_range(p, tid, kFilterSeqThreads);
    } else if (kind == FS_ZERO) {
      if (tid == 0) *q.step[i].zero = 0;
    } else if (kind == FS_MEDIAN) {
      const MedArgs m = q.step[i].med;
      if (m.reset_count && tid == 0) *m.reset_count = 0;
      __syncthreads();  // the reset lands before any append of this pass
      const int lanes = 2 * m.nx_vec * m.ny_vec;
      // every lane of a wave takes part in each round (the two lanes of a
      // vector exchange their medians); lanes past the end write nothing
      for (int g0 = 0; g0 < lanes; g0 += nthreads) median2_lane<WXM, WYM>(m, g0 + tid);
    } else if (kind == FS_FILL_LIST) {
      fill_list_range(q.step[i].fill, tid, nthreads);
    } else if (kind == FS_FILL) {
      fill_sweeps(q.step[i].fill, sh);
    } else if (kind == FS_SMOOTH) {
      smooth_range(q.step[i].smooth, tid, nthreads);
    }
    __syncthreads();
  }
}

template <int WXM, int WYM>
__global__ void __launch_bounds__(kFilterSeqThreads) k_filter_seq(FilterSeq q) {
  if (q.gate && *q.gate == 0) return;  // learned pipeline: grid 1 runs only as the fallback
  __shared__ FillShared sh;
  filter_seq_run<WXM, WYM>(q, sh, kFilterSeqThreads);
}

// The learned pipeline's gated grid-1 pass in one launch: the two-wave 64x64
// correlation walking the windows (k_xcorr_single64x2_gated), then, in the
// workgroup that finishes last, the pass's filter sequence.  Every thread
// releases its results before the workgroup takes its ticket, and the last
// workgroup acquires before it reads the others'; it clears the ticket for
// the next launch (all tickets are taken by then).  With the gate closed no
// workgroup touches the ticket.
template <int WXM, int WYM>
__global__ void __launch_bounds__(128) k_xcorr_single64x2_gated_filter(XcArgs a, FilterSeq q, unsigned int* ticket) {
  if (*a.gate == 0) return;
  __shared__ __align__(16) float lds[kW2LdsFloats];
  __shared__ float2 pnyq[64];
  __shared__ FillShared sh;
  __shared__ int last;
  const int nt = a.nx_vec * a.ny_vec;
  for (int w = blockIdx.x; w < nt; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();  // the last window's peak stage is done
    xcorr_single64x2_body(a, w, lds, pnyq);
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1 ? 1 : 0;
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
  filter_seq_run<WXM, WYM>(q, sh, 128);
}
__global__ void k_zero_ints(int* p, int n, const int* gate) {
  if (gate && *gate == 0) return;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = 0;
}

// Streaming copy used to measure the achievable HBM bandwidth and to calibrate
// the FETCH_SIZE / WRITE_SIZE counters against a known byte count.
__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     size_t n4) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += (size_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

// ============================================================== MHE enhancement
// One wave per tile, four tiles per 256-thread workgroup; no workgroup
// barriers.  Tile rows are read 4 pixels per load when ws % 4 == 0, with up
// to MHE_U loads of a lane in flight before their use.
constexpr int MHE_SUB = 4;  // sub-histograms per wave (lanes l, l+4, ... share one)
constexpr int MHE_U = 8;

__device__ __forceinline__ int mhe_neigh_count(int tx, int ty, int mx, int my) {
  // Calc_cnt_neigh_kernel (mhe:45-64)
  if (tx >= 1 && tx <= mx - 2 && ty >= 1 && ty <= my - 2) return 9;
  if ((tx == 0 && ty == 0) || (tx == 0 && ty == my - 1) || (tx == mx - 1 && ty == 0) || (tx == mx - 1 && ty == my - 1))
    return 4;
  return 6;
}

// MakeHist_kernel (mhe:66-83): the 256-bin histogram of the wave's tile,
// written into the tile's slot of the bordered histogram array.
// (16 sub-histograms of packed 16-bit counts, so that fewer lanes share one,
// were measured slower: 25.6 against 21.9 us per 6600x4400 frame.)
__global__ void __launch_bounds__(256) k_mhe_hist(MheArgs a) {
  __shared__ uint32_t hs[4][MHE_SUB][257];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * 4 + wv;  // tile rows banded per XCD
  if (tile >= a.tiles_x * a.tiles_y) return;
  uint32_t (*h)[257] = hs[wv];
  const int sub = lane & (MHE_SUB - 1);
#pragma unroll
  for (int q = 0; q < MHE_SUB; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) h[q][lane + 64 * j] = 0;
  wave_sync();
  auto bump = [&](uint32_t v) { atomicAdd(&h[sub][v], 1u); };
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const uint8_t* base = a.in + (long long)(a.ys + ty * a.ws) * a.nx + a.xs + tx * a.ws;
  if ((a.ws & 3) == 0) {
    const int q4 = a.ws >> 2, nd = a.ws * q4;
    auto word = [&](int d) {
      const int r = d / q4, c = d - r * q4;
      uint32_t w;
      __builtin_memcpy(&w, base + (long long)r * a.nx + 4 * c, 4);
      return w;
    };
    auto count = [&](uint32_t w) {
      bump(w & 255u);
      bump((w >> 8) & 255u);
      bump((w >> 16) & 255u);
      bump(w >> 24);
    };
    // batches of MHE_U words per lane, every load of a batch issued before
    // the first count (a 40x40 tile is one batch of at most 7 words)
    for (int d0 = lane; d0 < nd; d0 += 64 * MHE_U) {
      uint32_t w[MHE_U];
#pragma unroll
      for (int j = 0; j < MHE_U; ++j) w[j] = d0 + 64 * j < nd ? word(d0 + 64 * j) : 0u;
#pragma unroll
      for (int j = 0; j < MHE_U; ++j)
        if (d0 + 64 * j < nd) count(w[j]);
    }
  } else {
    const int npix = a.ws * a.ws;
    for (int p0 = lane; p0 < npix; p0 += 64 * MHE_U) {
      uint32_t v[MHE_U];
#pragma unroll
      for (int j = 0; j < MHE_U; ++j) {
        const int p = p0 + 64 * j, r = p / a.ws, c = p - r * a.ws;
        v[j] = p < npix ? base[(long long)r * a.nx + c] : 0u;
      }
#pragma unroll
      for (int j = 0; j < MHE_U; ++j)
        if (p0 + 64 * j < npix) bump(v[j]);
    }
  }
  wave_sync();
  // lane l: levels 4l .. 4l + 3, summed over the sub-histograms
  uint32_t o[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int q = 0; q < MHE_SUB; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] += h[q][4 * lane + j];
  reinterpret_cast<uint4*>(a.hist + ((long long)(ty + 1) * (a.tiles_x + 2) + tx + 1) * 256)[lane] =
      make_uint4(o[0], o[1], o[2], o[3]);
}

// CreateTransfer_kernel + prescan_kernel + Calc_Enhanced_value_kernel +
// Enhanced_kernel (mhe:85-137, 5-43), one wave per tile, lane l holding grey
// levels 4l..4l+3: the 3x3 tile neighbourhood's histogram count at each level,
// the inclusive prefix sum over the levels, then
//   v = min(255, max(cum - n*Ntrash, 0) * 255 / (n * Nnontrash)), v(0) = 0, v(255) = 255
// with n the tile's neighbour count.  The wave then maps its tile's pixels
// and zeroes the pixels outside the tiled region that lie beside it (edge
// tiles own the margins), so every frame pixel is written exactly once.
__global__ void __launch_bounds__(256) k_mhe_map(MheArgs a) {
  __shared__ uint8_t luts[4][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * 4 + wv;  // tile rows banded per XCD
  if (tile >= a.tiles_x * a.tiles_y) return;
  uint8_t* lut = luts[wv];
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int row = a.tiles_x + 2;
  const uint4* hc = reinterpret_cast<const uint4*>(a.hist + ((long long)(ty + 1) * row + tx + 1) * 256) + lane;
  uint4 hv[9];
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) hv[(dy + 1) * 3 + dx + 1] = hc[((long long)dy * row + dx) * 64];
  int v[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    v[0] += (int)hv[q].x;
    v[1] += (int)hv[q].y;
    v[2] += (int)hv[q].z;
    v[3] += (int)hv[q].w;
  }
  v[1] += v[0];
  v[2] += v[1];
  v[3] += v[2];
  int run = v[3];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(run, off, 64);
    if (lane >= off) run += o;
  }
  const int before = run - v[3];
  const int n = mhe_neigh_count(tx, ty, a.tiles_x, a.tiles_y);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int p = 4 * lane + j;
    long long c = (long long)(before + v[j]) - (long long)n * a.ntrash;
    c = c >= 0 ? c : 0;
    c = c * 255 / ((long long)n * a.nnontrash);
    int e = c < 255 ? (int)c : 255;
    if (p == 0) e = 0;
    if (p == 255) e = 255;
    lut[p] = (uint8_t)e;
  }
  wave_sync();
  // rectangle owned by this tile: the tile, widened to the frame edge on edge tiles
  const int mx = a.tiles_x, my = a.tiles_y;
  const int x0t = a.xs + tx * a.ws, y0t = a.ys + ty * a.ws;
  const int xlo = tx == 0 ? 0 : x0t, xhi = tx == mx - 1 ? a.nx : x0t + a.ws;
  const int ylo = ty == 0 ? 0 : y0t, yhi = ty == my - 1 ? a.ny : y0t + a.ws;
  const int w = xhi - xlo;
  if (w == a.ws && yhi - ylo == a.ws && (a.ws & 3) == 0) {
    // the rectangle is the tile itself: 4 pixels per load and store
    const int q4 = a.ws >> 2, nd = a.ws * q4;
    const long long ob = (long long)y0t * a.nx + x0t, obo = (long long)y0t * a.out_pitch + x0t;
    auto word = [&](int d) {
      const int r = d / q4, c = d - r * q4;
      uint32_t x;
      __builtin_memcpy(&x, a.in + ob + (long long)r * a.nx + 4 * c, 4);
      return x;
    };
    auto put = [&](int d, uint32_t x) {
      const int r = d / q4, c = d - r * q4;
      const uint32_t m = (uint32_t)lut[x & 255u] | ((uint32_t)lut[(x >> 8) & 255u] << 8) |
                         ((uint32_t)lut[(x >> 16) & 255u] << 16) | ((uint32_t)lut[x >> 24] << 24);
      __builtin_memcpy(a.out + obo + (long long)r * a.out_pitch + 4 * c, &m, 4);
    };
    int d0 = lane;
    for (; d0 + 64 * (MHE_U - 1) < nd; d0 += 64 * MHE_U) {
      uint32_t x[MHE_U];
#pragma unroll
      for (int j = 0; j < MHE_U; ++j) x[j] = word(d0 + 64 * j);
#pragma unroll
      for (int j = 0; j < MHE_U; ++j) put(d0 + 64 * j, x[j]);
    }
    for (; d0 < nd; d0 += 64) put(d0, word(d0));
    return;
  }
  const int npix = w * (yhi - ylo);
  for (int q = lane; q < npix; q += 64) {
    const int r = q / w, y = ylo + r, x = xlo + (q - r * w);
    const long long o = (long long)y * a.nx + x;
    const bool in_tile = x >= x0t && x < x0t + a.ws && y >= y0t && y < y0t + a.ws;
    a.out[(long long)y * a.out_pitch + x] = in_tile ? lut[a.in[o]] : (uint8_t)0;
  }
}

// ============================================================== launchers

static inline int grid_for(long long n, int block, int cap = 65536) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <int WX, int WY>
static hipError_t launch_pair(const XcArgs& a0, hipStream_t s) {
  XcArgs a = a0;
  a.chunks_x = (a.nx_vec + 1) / 2;
  hipLaunchKernelGGL((k_xcorr_pair<WX, WY>), dim3(a.chunks_x * a.ny_vec), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int WX, int WY>
static hipError_t launch_single(const XcArgs& a, hipStream_t s) {
  if constexpr (WX == 64 && WY == 64 && kSingle64TwoWaves) {
    if (a.gate)
      hipLaunchKernelGGL(k_xcorr_single64x2_gated, dim3(std::min(a.nx_vec * a.ny_vec, kGatedWindows)), dim3(128), 0,
                         s, a);
    else
      hipLaunchKernelGGL(k_xcorr_single64x2, dim3(a.nx_vec * a.ny_vec), dim3(128), 0, s, a);
  }
  else
    hipLaunchKernelGGL((k_xcorr_single<WX, WY>), dim3(a.nx_vec * a.ny_vec), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int WX, int WY, int WPB>
static hipError_t launch_fast(const XcArgs& a0, hipStream_t s) {
  if constexpr (WX <= 32 && WY <= 32 && WX + WY > 32) {
    return launch_pair<WX, WY>(a0, s);
  } else if constexpr (WX <= 64 && WY <= 64 && (WX > 32 || WY > 32)) {
    return launch_single<WX, WY>(a0, s);
  } else {
    XcArgs a = a0;
    a.chunks_x = (a.nx_vec + WPB - 1) / WPB;
    const int blocks = a.chunks_x * a.ny_vec;
    hipLaunchKernelGGL((k_xcorr_fast<WX, WY, WPB>), dim3(blocks), dim3(FastCfg<WX, WY, WPB>::NT), 0, s, a);
    return hipGetLastError();
  }
}

// (WX, WY, windows per workgroup).  Shapes up to 32 wide run one 64-lane wave
// per workgroup (wave-level synchronisation only); larger ones use barriers.
#define PIVX_FAST_SIZES(X) \
  X(32, 32, 2)             \
  X(64, 64, 2)             \
  X(16, 16, 4)             \
  X(24, 24, 2)             \
  X(48, 48, 4)             \
  X(64, 32, 4)             \
  X(32, 64, 4)             \
  X(32, 16, 2)             \
  X(16, 32, 2)             \
  X(48, 24, 8)

bool xcorr_fuses_deform(int ws_x, int ws_y) {
  return xcorr_fast_supported(ws_x, ws_y) && ws_x <= 64 && ws_y <= 64 && (ws_x > 32 || ws_y > 32);
}

hipError_t launch_deform_terms(const DefArgs& a, float4* col, float4* row, hipStream_t s) {
  const int n = a.nx_aug > a.ny_aug ? a.nx_aug : a.ny_aug;
  hipLaunchKernelGGL(k_deform_terms, dim3(grid_for(n, 256)), dim3(256), 0, s, a, col, row);
  return hipGetLastError();
}

bool xcorr_fast_supported(int ws_x, int ws_y) {
#define PIVX_CASE(WX, WY, WPB) \
  if (ws_x == WX && ws_y == WY) return true;
  PIVX_FAST_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return false;
}

static void factor_radices(int n, int* rad, int& nr) {
  nr = 0;
  while (n % 4 == 0) { rad[nr++] = 4; n /= 4; }
  while (n % 2 == 0) { rad[nr++] = 2; n /= 2; }
  for (int p = 3; n > 1; p += 2)
    while (n % p == 0) { rad[nr++] = p; n /= p; }
}

hipError_t launch_xcorr(const XcArgs& a, int ws_x, int ws_y, float2* gscratch, size_t gscratch_windows,
                        const float2* twx, const float2* twy, hipStream_t s) {
  if (twx == nullptr) {
#define PIVX_CASE(WX, WY, WPB) \
  if (ws_x == WX && ws_y == WY) return launch_fast<WX, WY, WPB>(a, s);
    PIVX_FAST_SIZES(PIVX_CASE)
#undef PIVX_CASE
    return hipErrorInvalidValue;
  }
  GenPlan plan;
  factor_radices(ws_x, plan.rx, plan.nrx);
  factor_radices(ws_y, plan.ry, plan.nry);
  const long long nwin = (long long)a.nx_vec * a.ny_vec;
  const size_t plane_bytes = (size_t)ws_x * ws_y * sizeof(float2);
  if (2 * plane_bytes <= 64 * 1024) {
    hipLaunchKernelGGL(k_xcorr_generic, dim3((unsigned)nwin), dim3(256), 2 * plane_bytes, s, a, ws_x, ws_y, plan, twx,
                       twy, (float2*)nullptr, 0);
    return hipGetLastError();
  }
  if (gscratch == nullptr || gscratch_windows == 0) return hipErrorInvalidValue;
  for (long long w0 = 0; w0 < nwin; w0 += (long long)gscratch_windows) {
    const long long nb = (nwin - w0) < (long long)gscratch_windows ? (nwin - w0) : (long long)gscratch_windows;
    hipLaunchKernelGGL(k_xcorr_generic, dim3((unsigned)nb), dim3(256), 0, s, a, ws_x, ws_y, plan, twx, twy, gscratch,
                       (int)w0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <int W, int H>
static hipError_t launch_dft_shape(const XcArgs& a, int mode, const DftTables& t, hipStream_t s) {
  const dim3 g((unsigned)(a.nx_vec * a.ny_vec)), b(256);
  const int split = mode >= 3 ? 1 : 0;
  if constexpr (W == 64 && H == 64) {
    if (t.lane64) {  // four waves per window, matrices from the lane-layout table
      if (mode == 1) hipLaunchKernelGGL((k_xcorr_dft64w<false, false>), g, b, 0, s, a, t.lane64);
      else if (mode == 2) hipLaunchKernelGGL((k_xcorr_dft64w<true, false>), g, b, 0, s, a, t.lane64);
      else if (mode == 3) hipLaunchKernelGGL((k_xcorr_dft64w<false, true>), g, b, 0, s, a, t.lane64);
      else hipLaunchKernelGGL((k_xcorr_dft64w<true, true>), g, b, 0, s, a, t.lane64);
      return hipGetLastError();
    }
  }
  if constexpr (W == 32 && H == 32) {
    if (t.lane32) {  // one wave per window, matrices in registers
      const dim3 b64(64);
      if (mode == 1) hipLaunchKernelGGL((k_xcorr_dft32w<false, false>), g, b64, 0, s, a, t.lane32);
      else if (mode == 2) hipLaunchKernelGGL((k_xcorr_dft32w<true, false>), g, b64, 0, s, a, t.lane32);
      else if (mode == 3) hipLaunchKernelGGL((k_xcorr_dft32w<false, true>), g, b64, 0, s, a, t.lane32);
      else hipLaunchKernelGGL((k_xcorr_dft32w<true, true>), g, b64, 0, s, a, t.lane32);
      return hipGetLastError();
    }
  }
  if (mode == 2 || mode == 4) hipLaunchKernelGGL((k_xcorr_dft<W, H, true>), g, b, 0, s, a, t, split);
  else hipLaunchKernelGGL((k_xcorr_dft<W, H, false>), g, b, 0, s, a, t, split);
  return hipGetLastError();
}

// window shapes of the precision form (sides multiples of 8 up to 64)
#define PIVX_DFT_SIZES(X) X(32, 32) X(64, 64) X(16, 16) X(24, 24) X(48, 48) X(64, 32) X(32, 64) X(32, 16) X(16, 32) X(48, 24)

bool xcorr_dft_supported(int ws_x, int ws_y) {
#define PIVX_CASE(WX, WY) \
  if (ws_x == WX && ws_y == WY) return true;
  PIVX_DFT_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return false;
}

hipError_t launch_xcorr_dft(const XcArgs& a, int ws_x, int ws_y, int mode, const DftTables& t, hipStream_t s) {
  if (mode < 1 || mode > 4) return hipErrorInvalidValue;
#define PIVX_CASE(WX, WY) \
  if (ws_x == WX && ws_y == WY) return launch_dft_shape<WX, WY>(a, mode, t, s);
  PIVX_DFT_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return hipErrorInvalidValue;
}

// Shapes with a batched (persistent) correlation kernel: the wave-pair and
// wave-per-window shapes of the register-FFT set.
#define PIVX_BATCH_SIZES(X) \
  X(32, 32, true) X(24, 24, true) X(32, 16, true) X(16, 32, true) X(64, 64, false) X(48, 48, false) \
  X(64, 32, false) X(32, 64, false) X(48, 24, false)

bool xcorr_batch_supported(int ws_x, int ws_y) {
#define PIVX_CASE(WX, WY, PAIR) \
  if (ws_x == WX && ws_y == WY) return true;
  PIVX_BATCH_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return false;
}

long long xcorr_batch_items(int ws_x, int ws_y, int nx_vec, int ny_vec, int* chunks_x) {
  *chunks_x = 0;
#define PIVX_CASE(WX, WY, PAIR)                                        \
  if (ws_x == WX && ws_y == WY) {                                      \
    if (PAIR) {                                                        \
      *chunks_x = (nx_vec + 1) / 2;                                    \
      return (long long)*chunks_x * ny_vec;                            \
    }                                                                  \
    return (long long)nx_vec * ny_vec;                                 \
  }
  PIVX_BATCH_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return -1;
}

hipError_t launch_xcorr_batch(int ws_x, int ws_y, const BatchQueue& q, int workgroups, hipStream_t s) {
  if (q.chunk < 1 || workgroups < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, q.counter, 8, (const int*)nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
#define PIVX_CASE(WX, WY, PAIR)                                                                     \
  if (ws_x == WX && ws_y == WY) {                                                                   \
    if (WX == 64 && WY == 64 && kSingle64TwoWaves) /* two waves per workgroup */                  \
      hipLaunchKernelGGL(k_xcorr_batch64x2, dim3(std::max(1, workgroups / 2)), dim3(128), 0, s, q);  \
    else                                                                                            \
      hipLaunchKernelGGL((k_xcorr_batch<WX, WY, PAIR>), dim3(workgroups), dim3(64), 0, s, q);     \
    return hipGetLastError();                                                                       \
  }
  PIVX_BATCH_SIZES(PIVX_CASE)
#undef PIVX_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_deform(const DefArgs& a, hipStream_t s) {
  const int nby = (a.ny_aug + DEF_ROWS - 1) / DEF_ROWS;
  const dim3 grid((a.nx_aug + 255) / 256, a.gate ? std::min(nby, kGatedBands) : nby);
  hipLaunchKernelGGL(k_deform, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_extend(const float* fx, const float* fy, float* ex, float* ey, int nx, int ny, int border,
                         hipStream_t s, const int* gate) {
  const long long n = (long long)(nx + 2 * border) * (ny + 2 * border);
  hipLaunchKernelGGL(k_extend, dim3(grid_for(n, 256)), dim3(256), 0, s, fx, fy, ex, ey, nx, ny, border, gate);
  return hipGetLastError();
}

hipError_t launch_upsample(float* dx2, float* dy2, const float* ex1x, const float* ex1y, int nxe1, int nye1,
                           float x_st1, float y_st1, float sp1x, float sp1y, int sp2x, int sp2y, int nx2, int ny2,
                           float* ex2x, float* ex2y, int border, hipStream_t s) {
  hipLaunchKernelGGL(k_upsample, dim3(grid_for((long long)nx2 * ny2, 256)), dim3(256), 0, s, dx2, dy2, ex1x, ex1y,
                     nxe1, nye1, x_st1, y_st1, sp1x, sp1y, sp2x, sp2y, nx2, ny2, ex2x, ex2y, border);
  return hipGetLastError();
}

hipError_t launch_median(const MedArgs& a, hipStream_t s) {
  const int nt = a.nx_vec * a.ny_vec;
  const int ntm = a.wx * a.wy;
  dim3 g(grid_for(nt, 128, 1 << 30)), b(128);
  const dim3 g2(grid_for(2LL * nt, 256, 1 << 30)), b2(256);
  if (a.wx == 5 && a.wy == 5) hipLaunchKernelGGL((k_median2<5, 5>), g2, b2, 0, s, a);
  else if (a.wx == 3 && a.wy == 3) hipLaunchKernelGGL((k_median2<3, 3>), g2, b2, 0, s, a);
  else if (a.wx == 7 && a.wy == 7) hipLaunchKernelGGL((k_median2<7, 7>), g2, b2, 0, s, a);
  else if (ntm <= 8) hipLaunchKernelGGL(k_median<8>, g, b, 0, s, a);
  else if (ntm <= 16) hipLaunchKernelGGL(k_median<16>, g, b, 0, s, a);
  else if (ntm <= 32) hipLaunchKernelGGL(k_median<32>, g, b, 0, s, a);
  else if (ntm <= 64) hipLaunchKernelGGL(k_median<64>, g, b, 0, s, a);
  else {
    if (a.scratch == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_median_generic, g, b, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_fill(const FillArgs& a, hipStream_t s) {
  if (!a.prelisted) {
    hipLaunchKernelGGL(k_fill_list, dim3(grid_for((long long)a.nx_vec * a.ny_vec, 256, 2048)), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_smooth(const SmoothArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_smooth, dim3(grid_for((long long)a.nx_vec * a.ny_vec, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_prep(const float* cx, const float* cy, const uint8_t* cv, int K, float* dx, float* dy, uint8_t* dv,
                       int nt, int* reset_count, hipStream_t s, const int* gate) {
  const PrepArgs p{cx, cy, cv, K, dx, dy, dv, nt, reset_count};
  hipLaunchKernelGGL(k_prep, dim3(grid_for(nt, 256)), dim3(256), 0, s, p, gate);
  return hipGetLastError();
}

bool median_gate_supported(int wxm, int wym) { return filter_seq_supported(wxm, wym); }

bool xcorr_gated_filter_supported(int ws_x, int ws_y, int wxm, int wym) {
  return ws_x == 64 && ws_y == 64 && kSingle64TwoWaves && filter_seq_supported(wxm, wym);
}

hipError_t launch_xcorr_gated_filter(const XcArgs& a, int ws_x, int ws_y, const FilterSeq& q, unsigned int* ticket,
                                     hipStream_t s) {
  if (!a.gate || !ticket || !xcorr_gated_filter_supported(ws_x, ws_y, q.wxm, q.wym) || q.nsteps < 0 ||
      q.nsteps > kFilterSeqSteps)
    return hipErrorInvalidValue;
  const dim3 g(std::min(a.nx_vec * a.ny_vec, kGatedWindows)), b(128);
  if (q.wxm == 5) hipLaunchKernelGGL((k_xcorr_single64x2_gated_filter<5, 5>), g, b, 0, s, a, q, ticket);
  else if (q.wxm == 3) hipLaunchKernelGGL((k_xcorr_single64x2_gated_filter<3, 3>), g, b, 0, s, a, q, ticket);
  else hipLaunchKernelGGL((k_xcorr_single64x2_gated_filter<7, 7>), g, b, 0, s, a, q, ticket);
  return hipGetLastError();
}

bool filter_seq_supported(int wxm, int wym) {
  return (wxm == 5 && wym == 5) || (wxm == 3 && wym == 3) || (wxm == 7 && wym == 7);
}

hipError_t launch_filter_seq(const FilterSeq& q, hipStream_t s) {
  if (q.nsteps < 0 || q.nsteps > kFilterSeqSteps) return hipErrorInvalidValue;
  const dim3 g(1), b(kFilterSeqThreads);
  if (q.wxm == 5 && q.wym == 5) hipLaunchKernelGGL((k_filter_seq<5, 5>), g, b, 0, s, q);
  else if (q.wxm == 3 && q.wym == 3) hipLaunchKernelGGL((k_filter_seq<3, 3>), g, b, 0, s, q);
  else if (q.wxm == 7 && q.wym == 7) hipLaunchKernelGGL((k_filter_seq<7, 7>), g, b, 0, s, q);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t s, const int* gate) {
  hipLaunchKernelGGL(k_zero_ints, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, gate);
  return hipGetLastError();
}
hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n4 = bytes / 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(grid_for((long long)n4, 256, 256 * 64)), dim3(256), 0, s,
                     (const float4*)src, (float4*)dst, n4);
  return hipGetLastError();
}
hipError_t launch_mhe(const MheArgs& a, hipStream_t s) {
  const int tiles = a.tiles_x * a.tiles_y;
  if (tiles < 1) return hipErrorInvalidValue;
  const dim3 g((tiles + 3) / 4), b(256);
  hipLaunchKernelGGL(k_mhe_hist, g, b, 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mhe_map, g, b, 0, s, a);
  return hipGetLastError();
}

}  // namespace pivx
