// oracle.cpp -- CPU restatement of the reference's multi-grid PIV stages.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the GPU path, and
// the timed scalar CPU baseline.  Never linked into the product.
//
// Each stage restates the reference text it cites (xcorr2d_V1.0/
// xcorr2d_multi_grid.cu = cu, main.cpp = main) in plain loops, templated on the
// arithmetic type T (double: the FP64 oracle; float: the FP32 transcription).
// Racy reference behaviour is fixed deterministically and documented:
//   * local-maximum candidates are kept uncapped and ranked by (value desc,
//     raster index asc); the reference appends them in atomic order into at
//     most ws_x*ws_y/16 slots (cu:598-601) and breaks ties to the lower slot;
//   * neighbour lists are gathered in raster order (cu:275, 417 use atomics);
//     the median is order-independent, the smoothing sum uses raster order;
//   * deformation taps outside the frame read 0 (cu:635-649 reads the zero
//     padding, or past its buffer for |disp| > ws/2).
//
// How each function is pinned (tests/test_oracle.py):
//   orc_fft1d            -- direct O(n^2) DFT in long double for n up to 1024
//   orc_correlate_window -- FFT path vs the direct circular sum; integer exact
//                           for u8 input; delta/constant/shift known answers
//   orc_xcorr_pass       -- brute-force local maxima and sorting on tiny planes;
//                           sampled-Gaussian peaks recover their sub-pixel
//                           centre exactly (3-point Gaussian estimator)
//   orc_deform           -- zero field = identity; integer field = exact shift;
//                           linear field checked against the closed form
//   orc_extend           -- closed form (clamp indexing)
//   orc_upsample         -- constant and linear fields reproduced exactly
//   orc_median_pass      -- numpy median on random fields; hand-built
//                           outlier / candidate-replacement / too-few-neighbour
//                           cases
//   orc_fill             -- isolated hole = mean of valid 3x3 neighbours;
//                           numpy Jacobi restatement on random masks
//   orc_smooth           -- linear field unchanged in the interior; numpy box
//                           mean
//   orc_geometry         -- reference result files (I=206,J=137 / I=412,J=275
//                           for 6600x4400 at spacing 32/16)
//   orc_mhe              -- numpy brute force (per-tile np.bincount, 3x3 sums,
//                           np.cumsum) on random frames; constant frames map
//                           to 255; margins outside the tiles are 0; corner /
//                           edge / interior neighbour counts on a 3x3 tiling
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <deque>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kExBorder = 3;  // cuh:190-191

// ---------------------------------------------------------------- FFT
// Recursive mixed-radix decimation in time over a per-length plan holding the
// twiddle table exp(-2*pi*i*j/n) (evaluated in long double, rounded to T once)
// and scratch buffers; one plan cache per thread.
template <class T>
struct FFT {
  using C = std::complex<T>;
  struct Plan {
    int n = 0;
    std::vector<C> tw, a, b, c;
  };
  static int small_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
      if (n % p == 0) return p;
    return n;
  }
  static Plan& plan(int n) {
    thread_local std::deque<Plan> cache;  // deque: references stay valid
    for (auto& p : cache)
      if (p.n == n) return p;
    cache.emplace_back();
    Plan& p = cache.back();
    p.n = n;
    p.tw.resize(n);
    for (int j = 0; j < n; ++j) {
      long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)n;
      p.tw[j] = C((T)std::cos(a), (T)std::sin(a));
    }
    p.a.resize(n);
    p.b.resize(n);
    p.c.resize(n);
    return p;
  }
  // out[k] = sum_j in[j*stride] * exp(sign 2 pi i j k / n); tw is the table
  // of the top-level length N = n * f.
  static void run(const C* in, int n, int stride, C* out, int sign, C* scratch, const C* tw, int f) {
    if (n == 1) {
      out[0] = in[0];
      return;
    }
    const int p = small_factor(n);
    const int m = n / p;
    for (int q = 0; q < p; ++q) run(in + (long)q * stride, m, stride * p, scratch + (long)q * m, sign, out, tw, f * p);
    if (p == 2 || p == 4) {
      // t_q = s_q[k] * w_n^{q k}; then a length-p DFT of the t_q
      for (int k = 0; k < m; ++k) {
        C t[4];
        t[0] = scratch[k];
        for (int q = 1; q < p; ++q) {
          C w = tw[(long)q * k * f];
          if (sign > 0) w = std::conj(w);
          t[q] = scratch[(long)q * m + k] * w;
        }
        if (p == 2) {
          out[k] = t[0] + t[1];
          out[k + m] = t[0] - t[1];
        } else {
          const C a = t[0] + t[2], b = t[0] - t[2], c = t[1] + t[3], d0 = t[1] - t[3];
          const C d = sign < 0 ? C(d0.imag(), -d0.real()) : C(-d0.imag(), d0.real());  // d0 * (-+i)
          out[k] = a + c;
          out[k + m] = b + d;
          out[k + 2 * m] = a - c;
          out[k + 3 * m] = b - d;
        }
      }
      return;
    }
    for (int k = 0; k < m; ++k)
      for (int s = 0; s < p; ++s) {
        C acc(0, 0);
        for (int q = 0; q < p; ++q) {
          C w = tw[(long)((long)q * (k + (long)s * m) % n) * f];
          if (sign > 0) w = std::conj(w);
          acc += scratch[(long)q * m + k] * w;
        }
        out[k + (long)s * m] = acc;
      }
  }
  // Power-of-two lengths: iterative radix-4/2 Stockham autosort passes.
  static void stockham_pow2(C* x, int n, int sign, Plan& P) {
    C* src = x;
    C* dst = P.a.data();
    for (int ns = 1; ns < n;) {
      const int R = ((n / ns) % 4 == 0) ? 4 : 2;
      const int nr = n / R, step = n / (ns * R);
      for (int j = 0; j < nr; ++j) {
        const int k = j & (ns - 1);
        C v[4];
        for (int r = 0; r < R; ++r) {
          C w = P.tw[(long)r * k * step];
          if (sign > 0) w = std::conj(w);
          v[r] = src[j + r * nr] * w;
        }
        const int d = (j - k) * R + k;
        if (R == 2) {
          dst[d] = v[0] + v[1];
          dst[d + ns] = v[0] - v[1];
        } else {
          const C a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], e0 = v[1] - v[3];
          const C e = sign < 0 ? C(e0.imag(), -e0.real()) : C(-e0.imag(), e0.real());
          dst[d] = a + c;
          dst[d + ns] = b + e;
          dst[d + 2 * ns] = a - c;
          dst[d + 3 * ns] = b - e;
        }
      }
      std::swap(src, dst);
      ns *= R;
    }
    if (src != x) std::copy(src, src + n, x);
  }
  static void transform(C* x, int n, int sign) {
    Plan& P = plan(n);
    if ((n & (n - 1)) == 0) {
      stockham_pow2(x, n, sign, P);
      return;
    }
    std::copy(x, x + n, P.a.begin());
    run(P.a.data(), n, 1, P.b.data(), sign, P.c.data(), P.tw.data(), 1);
    std::copy(P.b.begin(), P.b.end(), x);
  }
  // 2-D transform of a row-major wsy x wsx array: rows then columns.
  static void transform2d(C* x, int wsx, int wsy, int sign) {
    thread_local std::vector<C> line;
    line.resize(std::max(wsx, wsy));
    for (int r = 0; r < wsy; ++r) transform(x + (long)r * wsx, wsx, sign);
    for (int c = 0; c < wsx; ++c) {
      for (int r = 0; r < wsy; ++r) line[r] = x[(long)r * wsx + c];
      transform(line.data(), wsy, sign);
      for (int r = 0; r < wsy; ++r) x[(long)r * wsx + c] = line[r];
    }
  }
};

// ---------------------------------------------------------------- correlation (A5-A8)
// organize -> FFT f1, FFT f2 (cu:1334-1341) -> conj(F1)*F2 (cu:530-531) ->
// inverse (cu:1347) -> |.|/N (cu:542) -> fftshift rows then columns (cu:547-583).
template <class T>
void correlate(const T* f1w, const T* f2w, int wsx, int wsy, int method, T* cc) {
  const int N = wsx * wsy;
  std::vector<T> un(N);
  if (method == 0) {
    using C = std::complex<T>;
    std::vector<C> a(N), b(N);
    for (int i = 0; i < N; ++i) {
      a[i] = C(f1w[i], 0);
      b[i] = C(f2w[i], 0);
    }
    FFT<T>::transform2d(a.data(), wsx, wsy, -1);
    FFT<T>::transform2d(b.data(), wsx, wsy, -1);
    for (int i = 0; i < N; ++i) {
      const T x1 = a[i].real(), y1 = a[i].imag(), x2 = b[i].real(), y2 = b[i].imag();
      a[i] = C(x1 * x2 + y1 * y2, x1 * y2 - x2 * y1);
    }
    FFT<T>::transform2d(a.data(), wsx, wsy, +1);
    for (int i = 0; i < N; ++i) un[i] = std::sqrt(a[i].real() * a[i].real() + a[i].imag() * a[i].imag()) / (T)N;
  } else {
    // direct circular correlation, unshifted index s = (sx, sy)
    for (int sy = 0; sy < wsy; ++sy)
      for (int sx = 0; sx < wsx; ++sx) {
        T acc = 0;
        for (int y = 0; y < wsy; ++y)
          for (int x = 0; x < wsx; ++x)
            acc += f1w[y * wsx + x] * f2w[((y + sy) % wsy) * wsx + (x + sx) % wsx];
        un[sy * wsx + sx] = std::fabs(acc);
      }
  }
  // shifted (i, j) <- unshifted ((i + wsx/2) % wsx, (j + wsy/2) % wsy)
  for (int j = 0; j < wsy; ++j)
    for (int i = 0; i < wsx; ++i) cc[j * wsx + i] = un[((j + wsy / 2) % wsy) * wsx + (i + wsx / 2) % wsx];
}

// ---------------------------------------------------------------- peaks + sub-pixel (A10-A15)
struct Cand {
  double v;
  int i;
};

// Largest change of the 3-point fit d = (L0 - L2) / (L0 + L2 - 2 L1) / 2
// (L = log of the three plane values f0, f1, f2 along one axis) when each
// value moves by at most delta: first-order bound sum_i |dd/dL_i| * delta / f_i.
inline double fit_sensitivity(double f0, double f1, double f2, double delta) {
  if (!(f0 > 0 && f1 > 0 && f2 > 0)) return 0.0;  // log of a non-positive value: NaN, the vector is invalid
  const double L0 = std::log(f0), L1 = std::log(f1), L2 = std::log(f2);
  const double N = L0 - L2, D = L0 + L2 - 2 * L1;
  if (D == 0) return INFINITY;
  const double d0 = (D - N) / (2 * D * D), d2 = (-D - N) / (2 * D * D), d1 = N / (D * D);
  return std::fabs(d0) * delta / f0 + std::fabs(d1) * delta / f1 + std::fabs(d2) * delta / f2;
}

template <class T>
void peaks_and_fit(const T* cc, int wsx, int wsy, int bx, int by, int K, T px, T py, T lx1, T lx2, T ly1, T ly2,
                   double tie_rel, double fit_rel, double fit_tol, double* dxo, double* dyo, uint8_t* vo, double* pvo, int* pio,
                   int* nc, uint8_t* tie) {
  std::vector<Cand> c;
  for (int j = by; j <= wsy - 1 - by; ++j)
    for (int i = bx; i <= wsx - 1 - bx; ++i) {
      const T v = cc[j * wsx + i];
      // cu:596-597
      if (v > 0 && v >= cc[j * wsx + i - 1] && v >= cc[j * wsx + i + 1] && v >= cc[(j - 1) * wsx + i] &&
          v >= cc[(j + 1) * wsx + i])
        c.push_back({(double)v, j * wsx + i});
    }
  std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.v > b.v || (a.v == b.v && a.i < b.i); });
  if (nc) *nc = (int)c.size();
  for (int k = 0; k < K; ++k) {
    double dx = 0, dy = 0;
    bool ok = false;
    double pv = 0;
    int pi = -1;
    if (k < (int)c.size()) {
      pv = c[k].v;
      pi = c[k].i;
      const int i = pi % wsx, j = pi / wsx;
      // cu:327-347 and kernel_remove_nan cu:464-468
      if (!(i == 0 || i == wsx - 1 || j == 0 || j == wsy - 1 || pv <= 0)) {
        const T f1 = std::log(cc[j * wsx + i]);
        T f0 = std::log(cc[j * wsx + i - 1]);
        T f2 = std::log(cc[j * wsx + i + 1]);
        const T x = (T)i + (f0 - f2) / (f0 + f2 - 2 * f1) / 2 - (T)wsx / 2 + px;
        f0 = std::log(cc[(j - 1) * wsx + i]);
        f2 = std::log(cc[(j + 1) * wsx + i]);
        const T y = (T)j + (f0 - f2) / (f0 + f2 - 2 * f1) / 2 - (T)wsy / 2 + py;
        ok = !(x < lx1 || x > lx2 || y < ly1 || y > ly2) && std::isfinite(x) && std::isfinite(y);
        if (ok) {
          dx = (double)x;
          dy = (double)y;
        }
      }
    }
    if (dxo) dxo[k] = dx;
    if (dyo) dyo[k] = dy;
    if (vo) vo[k] = ok ? 1 : 0;
    if (pvo) pvo[k] = pv;
    if (pio) pio[k] = pi;
  }
  if (tie) {
    bool t = false;
    const double top = c.empty() ? 0.0 : c[0].v;
    const double delta = tie_rel * top;
    for (int k = 0; k < K && k + 1 < (int)c.size(); ++k)
      if (c[k].v - c[k + 1].v < delta) t = true;
    const double vth = ((int)c.size() >= K) ? c[K - 1].v - delta : -1.0;
    if (!t && delta > 0) {
      for (int j = by; j <= wsy - 1 - by && !t; ++j)
        for (int i = bx; i <= wsx - 1 - bx && !t; ++i) {
          const double v = (double)cc[j * wsx + i];
          if (v < vth) continue;
          const int nb[4] = {j * wsx + i - 1, j * wsx + i + 1, (j - 1) * wsx + i, (j + 1) * wsx + i};
          for (int q = 0; q < 4; ++q)
            if (std::fabs(v - (double)cc[nb[q]]) < delta) t = true;
          if (std::fabs(v) < delta) t = true;
        }
    }
    // an ill-conditioned fit (a nearly flat peak): values within fit_rel *
    // peak move the sub-pixel displacement of a valid candidate by more than
    // fit_tol
    bool flat = false;
    const double fdelta = fit_rel * top;
    if (!t && fdelta > 0 && fit_tol > 0) {
      for (int k = 0; k < K && k < (int)c.size() && !flat; ++k) {
        if (vo && !vo[k]) continue;
        const int i = c[k].i % wsx, j = c[k].i / wsx;
        if (i == 0 || i == wsx - 1 || j == 0 || j == wsy - 1) continue;
        const double f1 = (double)cc[j * wsx + i];
        const double sx = fit_sensitivity((double)cc[j * wsx + i - 1], f1, (double)cc[j * wsx + i + 1], fdelta);
        const double sy = fit_sensitivity((double)cc[(j - 1) * wsx + i], f1, (double)cc[(j + 1) * wsx + i], fdelta);
        if (std::max(sx, sy) > fit_tol) flat = true;
      }
    }
    *tie = t ? 1 : (flat ? 2 : 0);
  }
}

// Gather of window w (organize_data_fft2d_multi cu:497-517 and _deformed
// cu:654-673): pixel (c, r) at aug coordinates (ix*spx + c, iy*spy + r),
// image coordinates = aug - left, 0 outside the image.
template <class T>
void gather(const orc_pass* P, int w, T* f1w, T* f2w) {
  const int ix = w % P->nx_vec, iy = w / P->nx_vec;
  for (int r = 0; r < P->wsy; ++r)
    for (int c = 0; c < P->wsx; ++c) {
      const int ax = ix * P->spx + c, ay = iy * P->spy + r;
      const int xi = ax - P->left_x, yi = ay - P->left_y;
      const bool in = xi >= 0 && xi < P->nx && yi >= 0 && yi < P->ny;
      f1w[r * P->wsx + c] = in ? (T)P->f1[(long)yi * P->nx + xi] : (T)0;
      if (P->f2def) f2w[r * P->wsx + c] = (T)P->f2def[(long)ay * P->nx_aug + ax];
      else f2w[r * P->wsx + c] = in ? (T)P->f2u8[(long)yi * P->nx + xi] : (T)0;
    }
}

template <class T>
int xcorr_pass(const orc_pass* P, double* dx, double* dy, uint8_t* vv, double* pv, int* pi, int* nc, uint8_t* tie) {
  const int nt = P->nx_vec * P->ny_vec;
  const int K = P->K;
  const int b = P->win_begin > 0 ? P->win_begin : 0;
  const int e = P->win_end > 0 ? std::min(P->win_end, nt) : nt;
  const int N = P->wsx * P->wsy;
  int threads = P->threads > 0 ? P->threads : 1;
#ifndef _OPENMP
  threads = 1;
#endif
#pragma omp parallel num_threads(threads)
  {
    std::vector<T> f1w(N), f2w(N), cc(N);
#pragma omp for schedule(static)
    for (int w = b; w < e; ++w) {
      gather<T>(P, w, f1w.data(), f2w.data());
      correlate<T>(f1w.data(), f2w.data(), P->wsx, P->wsy, P->method, cc.data());
      const T px = P->pred_x ? (T)P->pred_x[w] : (T)0, py = P->pred_y ? (T)P->pred_y[w] : (T)0;
      const long o = (long)w * K;
      peaks_and_fit<T>(cc.data(), P->wsx, P->wsy, P->bx, P->by, K, px, py, (T)P->lx1, (T)P->lx2, (T)P->ly1,
                       (T)P->ly2, P->tie_rel, P->fit_rel, P->fit_tol, dx ? dx + o : nullptr, dy ? dy + o : nullptr, vv ? vv + o : nullptr,
                       pv ? pv + o : nullptr, pi ? pi + o : nullptr, nc ? nc + w : nullptr, tie ? tie + w : nullptr);
    }
  }
  return 0;
}

// ---------------------------------------------------------------- bilinear sums
// The reference's bilinear sums as the fused multiply-add chains a CUDA build
// forms from their source by default (--fmad=true; the rest of the oracle is
// compiled with contraction off):
//   x + c1*e1 + c2*e2 + c3*e3 + c4*e4 -> fma(c4,e4, fma(c3,e3, fma(c2,e2, fma(c1,e1, x))))
//   c1*e1 + c2*e2 + c3*e3 + c4*e4     -> fma(c4,e4, fma(c3,e3, fma(c1,e1, c2*e2)))
template <class T>
inline T fma4(T c1, T e1, T c2, T e2, T c3, T e3, T c4, T e4, T x) {
  return std::fma(c4, e4, std::fma(c3, e3, std::fma(c2, e2, std::fma(c1, e1, x))));
}
template <class T>
inline T sum4(T c1, T e1, T c2, T e2, T c3, T e3, T c4, T e4) {
  return std::fma(c4, e4, std::fma(c3, e3, std::fma(c1, e1, c2 * e2)));
}

// ---------------------------------------------------------------- deformation (A16)
template <class T>
void deform(const uint8_t* f2, int nx, int ny, int left_x, int left_y, int nx_aug, int ny_aug, const double* exx,
            const double* exy, int nxe, int nye, T x_st, T y_st, T spx, T spy, double* out) {
  for (int idy = 0; idy < ny_aug; ++idy)
    for (int idx = 0; idx < nx_aug; ++idx) {
      T ax = ((T)idx - x_st) / spx;
      T ay = ((T)idy - y_st) / spy;
      int ix = (int)std::floor(ax), iy = (int)std::floor(ay);
      ax = ax - (T)ix;
      ay = ay - (T)iy;
      T ce1 = (1 - ax) * (1 - ay), ce2 = ax * (1 - ay), ce3 = (1 - ax) * ay, ce4 = ax * ay;
      ix = std::min(std::max(ix, 0), nxe - 2);
      iy = std::min(std::max(iy, 0), nye - 2);
      const int i1 = iy * nxe + ix, i2 = i1 + 1, i3 = i1 + nxe, i4 = i3 + 1;
      T sx = fma4(ce1, (T)exx[i1], ce2, (T)exx[i2], ce3, (T)exx[i3], ce4, (T)exx[i4], (T)idx);
      T sy = fma4(ce1, (T)exy[i1], ce2, (T)exy[i2], ce3, (T)exy[i3], ce4, (T)exy[i4], (T)idy);
      const int jx = (int)std::floor(sx), jy = (int)std::floor(sy);
      sx = sx - (T)jx;
      sy = sy - (T)jy;
      ce1 = (1 - sx) * (1 - sy);
      ce2 = sx * (1 - sy);
      ce3 = (1 - sx) * sy;
      ce4 = sx * sy;
      auto px = [&](int x, int y) -> T {
        const int xi = x - left_x, yi = y - left_y;
        return (xi >= 0 && xi < nx && yi >= 0 && yi < ny) ? (T)f2[(long)yi * nx + xi] : (T)0;
      };
      out[(long)idy * nx_aug + idx] =
          (double)sum4(ce1, px(jx, jy), ce2, px(jx + 1, jy), ce3, px(jx, jy + 1), ce4, px(jx + 1, jy + 1));
    }
}

// ---------------------------------------------------------------- upsampling (A18)
template <class T>
void upsample(const double* ex, const double* ey, int nxe, int nye, T x_st1, T y_st1, T sp1x, T sp1y, int sp2x,
              int sp2y, int nx2, int ny2, double* dx2, double* dy2) {
  for (int vy = 0; vy < ny2; ++vy)
    for (int vx = 0; vx < nx2; ++vx) {
      const int idx = vx * sp2x + sp2x / 2, idy = vy * sp2y + sp2y / 2;
      T ax = ((T)idx - x_st1) / sp1x;
      T ay = ((T)idy - y_st1) / sp1y;
      int ix = (int)std::floor(ax), iy = (int)std::floor(ay);
      ax = ax - (T)ix;
      ay = ay - (T)iy;
      const T ce1 = (1 - ax) * (1 - ay), ce2 = ax * (1 - ay), ce3 = (1 - ax) * ay, ce4 = ax * ay;
      ix = std::min(std::max(ix, 0), nxe - 2);
      iy = std::min(std::max(iy, 0), nye - 2);
      const int i1 = iy * nxe + ix, i2 = i1 + 1, i3 = i1 + nxe, i4 = i3 + 1;
      const long t = (long)vy * nx2 + vx;
      dx2[t] = (double)sum4(ce1, (T)ex[i1], ce2, (T)ex[i2], ce3, (T)ex[i3], ce4, (T)ex[i4]);
      dy2[t] = (double)sum4(ce1, (T)ey[i1], ce2, (T)ey[i2], ce3, (T)ey[i3], ce4, (T)ey[i4]);
    }
}

// ---------------------------------------------------------------- validation (A19-A25)
template <class T>
T median_of(std::vector<T>& v) {  // exact order-statistic median (cu:184-203)
  const int n = (int)v.size();
  if (n == 0) return 0;
  std::sort(v.begin(), v.end());
  if (n % 2 == 0) return (v[n / 2] + v[n / 2 - 1]) / 2;
  return v[(n - 1) / 2];
}

template <class T>
void median_pass(const double* sx, const double* sy, const uint8_t* sv, int sstride, const double* cx, const double* cy,
                 const uint8_t* cv, int K, int nxv, int nyv, int wx, int wy, int min_nneigh, T thr, T eps, double* dx,
                 double* dy, uint8_t* dv, double* margin) {
  const int ntm = wx * wy, hx = (wx - 1) / 2, hy = (wy - 1) / 2;
  std::vector<T> X, Y;
  for (int t = 0; t < nxv * nyv; ++t) {
    const int vy = t / nxv, vx = t % nxv;
    X.clear();
    Y.clear();
    for (int s = 0; s < ntm; ++s) {  // cu:262-279
      if (s == (ntm - 1) / 2) continue;
      const int qx = vx - hx + s % wx, qy = vy - hy + s / wx;
      if (qx < 0 || qx >= nxv || qy < 0 || qy >= nyv) continue;
      const long q = (long)(qy * nxv + qx) * sstride;
      if (!sv[q]) continue;
      X.push_back((T)sx[q]);
      Y.push_back((T)sy[q]);
    }
    const int n = (int)X.size();
    const bool vmed = n >= min_nneigh;  // cu:291
    double ox = 0, oy = 0;
    bool ok = false;
    double mg = INFINITY;
    if (vmed) {
      const T mx = median_of(X), my = median_of(Y);
      std::vector<T> DX(n), DY(n);
      for (int i = 0; i < n; ++i) {  // cu:218
        DX[i] = std::fabs(X[i] - mx);
        DY[i] = std::fabs(Y[i] - my);
      }
      const T rx = median_of(DX), ry = median_of(DY);
      for (int k = 0; k < K; ++k) {  // cu:237-248
        const long o = (long)t * K + k;
        if (!cv[o]) continue;
        const T ex = std::fabs((T)cx[o] - mx) / (rx + eps);
        const T ey = std::fabs((T)cy[o] - my) / (ry + eps);
        mg = std::min(mg, std::min(std::fabs((double)ex - (double)thr), std::fabs((double)ey - (double)thr)) / (double)thr);
        if (ex < thr && ey < thr) {
          ok = true;
          ox = cx[o];
          oy = cy[o];
          break;
        }
      }
    }
    dx[t] = ox;
    dy[t] = oy;
    dv[t] = ok ? 1 : 0;
    if (margin) margin[t] = mg;
  }
}

template <class T>
int fill(double* dx, double* dy, uint8_t* dv, int nxv, int nyv) {
  const int nt = nxv * nyv;
  std::vector<T> nxs(nt), nys(nt);
  std::vector<uint8_t> f(nt);
  int sweeps = 0;
  while (true) {  // cu:1514-1528: repeat while a sweep filled something
    int cnt_fill = 0;
    for (int t = 0; t < nt; ++t) {
      f[t] = 0;
      if (dv[t]) continue;
      const int vy = t / nxv, vx = t % nxv;
      int cnt = 0;
      T sx = 0, sy = 0;
      for (int j = std::max(0, vy - 1); j <= std::min(nyv - 1, vy + 1); ++j)
        for (int k = std::max(0, vx - 1); k <= std::min(nxv - 1, vx + 1); ++k)
          if (dv[j * nxv + k]) {
            ++cnt;
            sx += (T)dx[j * nxv + k];
            sy += (T)dy[j * nxv + k];
          }
      if (cnt > 0) {
        f[t] = 1;
        nxs[t] = sx / (T)cnt;
        nys[t] = sy / (T)cnt;
        ++cnt_fill;
      }
    }
    ++sweeps;
    for (int t = 0; t < nt; ++t)
      if (f[t]) {
        dx[t] = (double)nxs[t];
        dy[t] = (double)nys[t];
        dv[t] = 1;
      }
    if (cnt_fill == 0) break;
  }
  return sweeps;
}

template <class T>
void smooth(const double* sx, const double* sy, int nxv, int nyv, int wx, int wy, double* dx, double* dy) {
  const int hx = (wx - 1) / 2, hy = (wy - 1) / 2;
  for (int t = 0; t < nxv * nyv; ++t) {
    const int vy = t / nxv, vx = t % nxv;
    T ax = 0, ay = 0;
    int n = 0;
    for (int r = 0; r < wy; ++r)
      for (int c = 0; c < wx; ++c) {
        const int qx = vx - hx + c, qy = vy - hy + r;
        if (qx < 0 || qx >= nxv || qy < 0 || qy >= nyv) continue;
        ax += (T)sx[qy * nxv + qx];
        ay += (T)sy[qy * nxv + qx];
        ++n;
      }
    dx[t] = (double)(ax / (T)n);
    dy[t] = (double)(ay / (T)n);
  }
}

// ---------------------------------------------------------------- geometry (A1)
int vec_count(int n, int sp) {  // cu:771-777
  int t1 = (n - sp / 2 - 1) / sp;
  int t2 = n - (sp / 2 + 1 + t1 * sp);
  while (t2 < (sp / 2 - 1)) {
    t1 -= 1;
    t2 += sp;
  }
  return t1 + 1;
}

struct Geo {
  int nx_vec, ny_vec, bx, by, slots, lx, ly, rx, ry, nx_aug, ny_aug, nsec, neach, nlast;
};

Geo geometry(int nx, int ny, int nsec, int wsx, int wsy, int spx, int spy, float pctx, float pcty) {
  Geo g{};
  g.bx = std::max(1, wsx / 2 - (int)roundf((float)wsx * pctx));
  g.by = std::max(1, wsy / 2 - (int)roundf((float)wsy * pcty));
  g.slots = wsx * wsy / 16;
  g.nx_vec = vec_count(nx, spx);
  g.ny_vec = vec_count(ny, spy);
  const int nt = g.nx_vec * g.ny_vec;
  g.nsec = nsec;
  if (nsec > 0 && nt > 0) {
    g.neach = nt / nsec;
    g.nlast = (nt % nsec == 0) ? g.neach : nt - (nsec - 1) * g.neach;
    if (nt % nsec != 0)
      while (g.nlast > g.neach && g.neach > 0) {
        g.nlast -= g.neach;
        g.nsec += 1;
      }
  }
  const int hx = wsx / 2, hy = wsy / 2;
  g.lx = spx / 2 < hx ? hx - spx / 2 : 0;
  g.ly = spy / 2 < hy ? hy - spy / 2 : 0;
  g.rx = (nx - spx / 2 - 1 - (g.nx_vec - 1) * spx) < (hx - 1) ? spx / 2 + 1 + (g.nx_vec - 1) * spx + hx - 1 - nx : 0;
  g.ry = (ny - spy / 2 - 1 - (g.ny_vec - 1) * spy) < (hy - 1) ? spy / 2 + 1 + (g.ny_vec - 1) * spy + hy - 1 - ny : 0;
  g.nx_aug = nx + g.lx + g.rx;
  g.ny_aug = ny + g.ly + g.ry;
  return g;
}

// ---------------------------------------------------------------- whole pair (A26, A28)
template <class T>
struct Pipeline {
  const orc_run* R;
  const uint8_t *f1, *f2;
  Geo g[2];
  std::vector<double> rx[2], ry[2], fx[2], fy[2], ex[2], ey[2];
  std::vector<uint8_t> rv[2], fv[2];

  void extend_field(int gi) {
    const int nxe = g[gi].nx_vec + 2 * kExBorder, nye = g[gi].ny_vec + 2 * kExBorder;
    ex[gi].assign((size_t)nxe * nye, 0.0);
    ey[gi].assign((size_t)nxe * nye, 0.0);
    orc_extend(fx[gi].data(), fy[gi].data(), g[gi].nx_vec, g[gi].ny_vec, kExBorder, ex[gi].data(), ey[gi].data());
  }

  void calc(int gi, bool deformed) {
    const int wsx = gi ? R->ws_x2 : R->ws_x1, wsy = gi ? R->ws_y2 : R->ws_y1;
    const int spx = gi ? R->sp_x2 : R->sp_x1, spy = gi ? R->sp_y2 : R->sp_y1;
    std::vector<double> def;
    if (deformed) {
      def.resize((size_t)g[gi].nx_aug * g[gi].ny_aug);
      deform<T>(f2, R->nx, R->ny, g[gi].lx, g[gi].ly, g[gi].nx_aug, g[gi].ny_aug, ex[gi].data(), ey[gi].data(),
                g[gi].nx_vec + 2 * kExBorder, g[gi].ny_vec + 2 * kExBorder, (T)(wsx / 2 - spx * kExBorder),
                (T)(wsy / 2 - spy * kExBorder), (T)spx, (T)spy, def.data());
    }
    orc_pass P{};
    P.f1 = f1;
    P.f2u8 = deformed ? nullptr : f2;
    P.f2def = deformed ? def.data() : nullptr;
    P.nx = R->nx;
    P.ny = R->ny;
    P.nx_aug = g[gi].nx_aug;
    P.ny_aug = g[gi].ny_aug;
    P.wsx = wsx;
    P.wsy = wsy;
    P.spx = spx;
    P.spy = spy;
    P.left_x = g[gi].lx;
    P.left_y = g[gi].ly;
    P.nx_vec = g[gi].nx_vec;
    P.ny_vec = g[gi].ny_vec;
    P.bx = g[gi].bx;
    P.by = g[gi].by;
    P.K = R->K;
    P.lx1 = R->lx1;  // both grids use the first-grid limit variables (cu:1440, 1755)
    P.lx2 = R->lx2;
    P.ly1 = R->ly1;
    P.ly2 = R->ly2;
    P.pred_x = fx[gi].data();
    P.pred_y = fy[gi].data();
    P.threads = R->threads;
    xcorr_pass<T>(&P, rx[gi].data(), ry[gi].data(), rv[gi].data(), nullptr, nullptr, nullptr, nullptr);
  }

  void filter(int gi, int npm, int minn, double thr, int nps, bool dofill, double* ox, double* oy, uint8_t* ov) {
    const int nxv = g[gi].nx_vec, nyv = g[gi].ny_vec, nt = nxv * nyv, K = R->K;
    const int wxm = gi ? R->wx2_median : R->wx1_median, wym = gi ? R->wy2_median : R->wy1_median;
    const int wxs = gi ? R->wx2_smooth : R->wx1_smooth, wys = gi ? R->wy2_smooth : R->wy1_smooth;
    // median_filter_data_prep (cu:484-495)
    std::vector<double> cx(nt), cy(nt);
    std::vector<uint8_t> cv(nt);
    for (int t = 0; t < nt; ++t) {
      cx[t] = rx[gi][(size_t)t * K];
      cy[t] = ry[gi][(size_t)t * K];
      cv[t] = rv[gi][(size_t)t * K];
    }
    for (int p = 0; p < npm; ++p) {
      std::vector<double> nx_(nt), ny_(nt);
      std::vector<uint8_t> nv(nt);
      median_pass<T>(cx.data(), cy.data(), cv.data(), 1, rx[gi].data(), ry[gi].data(), rv[gi].data(), K, nxv, nyv, wxm,
                     wym, minn, (T)thr, (T)0.1f, nx_.data(), ny_.data(), nv.data(), nullptr);
      cx.swap(nx_);
      cy.swap(ny_);
      cv.swap(nv);
    }
    if (dofill) fill<T>(cx.data(), cy.data(), cv.data(), nxv, nyv);
    for (int s = 0; s < nps; ++s) {
      std::vector<double> sx(nt), sy(nt);
      smooth<T>(cx.data(), cy.data(), nxv, nyv, wxs, wys, sx.data(), sy.data());
      cx.swap(sx);
      cy.swap(sy);
    }
    std::copy(cx.begin(), cx.end(), ox);
    std::copy(cy.begin(), cy.end(), oy);
    std::copy(cv.begin(), cv.end(), ov);
  }
};

template <class T>
int run_pair(const orc_run* R, const uint8_t* f1, const uint8_t* f2, double* r1x, double* r1y, uint8_t* r1v,
             double* o1x, double* o1y, uint8_t* o1v, double* r2x, double* r2y, uint8_t* r2v, double* o2x, double* o2y,
             uint8_t* o2v, double* m2x, double* m2y, uint8_t* m2v) {
  Pipeline<T> S;
  S.R = R;
  S.f1 = f1;
  S.f2 = f2;
  S.g[0] = geometry(R->nx, R->ny, R->nsec, R->ws_x1, R->ws_y1, R->sp_x1, R->sp_y1, (float)R->pct_x1, (float)R->pct_y1);
  S.g[1] = geometry(R->nx, R->ny, R->nsec, R->ws_x2, R->ws_y2, R->sp_x2, R->sp_y2, (float)R->pct_x2, (float)R->pct_y2);
  for (int gi = 0; gi < 2; ++gi) {
    const size_t nt = (size_t)S.g[gi].nx_vec * S.g[gi].ny_vec;
    S.rx[gi].assign(nt * R->K, 0.0);
    S.ry[gi].assign(nt * R->K, 0.0);
    S.rv[gi].assign(nt * R->K, 0);
    S.fx[gi].assign(nt, 0.0);
    S.fy[gi].assign(nt, 0.0);
    S.fv[gi].assign(nt, 0);
    const size_t ne = (size_t)(S.g[gi].nx_vec + 6) * (S.g[gi].ny_vec + 6);
    S.ex[gi].assign(ne, 0.0);
    S.ey[gi].assign(ne, 0.0);
  }
  for (int j = 0; j < R->npass1; ++j) {  // main:289-303
    if (j > 0) S.extend_field(0);
    if (j == 0) {  // cu:1287-1298: predictor zeroed, no deformation
      std::fill(S.fx[0].begin(), S.fx[0].end(), 0.0);
      std::fill(S.fy[0].begin(), S.fy[0].end(), 0.0);
    }
    S.calc(0, j > 0);
    S.filter(0, R->npass1_median, R->min_nneigh1, R->thr1, R->npass1_smooth, true, S.fx[0].data(), S.fy[0].data(),
             S.fv[0].data());
  }
  std::vector<double> mx, my;
  std::vector<uint8_t> mv;
  if (R->npass2 > 0) {  // main:305-326
    S.extend_field(0);
    const Geo& g1 = S.g[0];
    const Geo& g2 = S.g[1];
    upsample<T>(S.ex[0].data(), S.ey[0].data(), g1.nx_vec + 6, g1.ny_vec + 6, (T)(R->sp_x1 / 2 - 3 * R->sp_x1),
                (T)(R->sp_y1 / 2 - 3 * R->sp_y1), (T)R->sp_x1, (T)R->sp_y1, R->sp_x2, R->sp_y2, g2.nx_vec, g2.ny_vec,
                S.fx[1].data(), S.fy[1].data());
    for (int j = 0; j < R->npass2; ++j) {
      S.extend_field(1);
      S.calc(1, true);
      S.filter(1, R->npass2_median, R->min_nneigh2, R->thr2, R->npass2_smooth, true, S.fx[1].data(), S.fy[1].data(),
               S.fv[1].data());
    }
    const size_t nt2 = (size_t)g2.nx_vec * g2.ny_vec;
    mx.resize(nt2);
    my.resize(nt2);
    mv.resize(nt2);
    S.filter(1, R->npass2_median, R->min_nneigh2, R->thr2, 0, false, mx.data(), my.data(), mv.data());  // main:363
  }
  auto put = [](const std::vector<double>& v, double* o) {
    if (o) std::copy(v.begin(), v.end(), o);
  };
  auto putb = [](const std::vector<uint8_t>& v, uint8_t* o) {
    if (o) std::copy(v.begin(), v.end(), o);
  };
  put(S.rx[0], r1x);
  put(S.ry[0], r1y);
  putb(S.rv[0], r1v);
  put(S.fx[0], o1x);
  put(S.fy[0], o1y);
  putb(S.fv[0], o1v);
  put(S.rx[1], r2x);
  put(S.ry[1], r2y);
  putb(S.rv[1], r2v);
  put(S.fx[1], o2x);
  put(S.fy[1], o2y);
  putb(S.fv[1], o2v);
  put(mx, m2x);
  put(my, m2y);
  putb(mv, m2v);
  return 0;
}

}  // namespace

extern "C" {

double orc_fit_sensitivity(double f0, double f1, double f2, double delta) {
  return fit_sensitivity(f0, f1, f2, delta);
}

int orc_geometry(int nx, int ny, int nsec, int wsx, int wsy, int spx, int spy, float pctx, float pcty, int* out) {
  if (spx < 1 || spy < 1 || !out) return 1;
  Geo g = geometry(nx, ny, nsec, wsx, wsy, spx, spy, pctx, pcty);
  const int v[14] = {g.nx_vec, g.ny_vec, g.bx, g.by, g.slots, g.lx, g.ly, g.rx, g.ry, g.nx_aug, g.ny_aug, g.nsec, g.neach, g.nlast};
  std::memcpy(out, v, sizeof(v));
  return 0;
}

int orc_correlate_window(int prec, const double* f1w, const double* f2w, int wsx, int wsy, int method, double* cc) {
  const int N = wsx * wsy;
  if (prec == 32) {
    std::vector<float> a(f1w, f1w + N), b(f2w, f2w + N), c(N);
    correlate<float>(a.data(), b.data(), wsx, wsy, method, c.data());
    for (int i = 0; i < N; ++i) cc[i] = c[i];
  } else {
    correlate<double>(f1w, f2w, wsx, wsy, method, cc);
  }
  return 0;
}

int orc_fft1d(int prec, double* x, int n, int sign) {
  if (prec == 32) {
    std::vector<std::complex<float>> v(n);
    for (int i = 0; i < n; ++i) v[i] = {(float)x[2 * i], (float)x[2 * i + 1]};
    FFT<float>::transform(v.data(), n, sign);
    for (int i = 0; i < n; ++i) {
      x[2 * i] = v[i].real();
      x[2 * i + 1] = v[i].imag();
    }
  } else {
    FFT<double>::transform(reinterpret_cast<std::complex<double>*>(x), n, sign);
  }
  return 0;
}

int orc_xcorr_pass(int prec, const orc_pass* P, double* dx, double* dy, uint8_t* vv, double* pv, int* pi, int* nc,
                   uint8_t* tie) {
  if (!P || !P->f1 || (!P->f2u8 && !P->f2def)) return 1;
  return prec == 32 ? xcorr_pass<float>(P, dx, dy, vv, pv, pi, nc, tie) : xcorr_pass<double>(P, dx, dy, vv, pv, pi, nc, tie);
}

// One correlation pass split into its stages for the CPU baseline (one
// thread, FP32): seconds spent in the window gather (cu:497-517, 654-673), in
// the three transforms with the conjugate product, magnitude and shift
// (cu:519-583, 1334-1350), and in the peak search, top-N ranking and fit
// (cu:585-606, 4-172, 298-350); same arithmetic as orc_xcorr_pass.
int orc_xcorr_stage_seconds(const orc_pass* P, double* t_gather, double* t_corr, double* t_peaks) {
  if (!P || !P->f1 || (!P->f2u8 && !P->f2def) || !t_gather || !t_corr || !t_peaks) return 1;
  using clk = std::chrono::steady_clock;
  const int nt = P->nx_vec * P->ny_vec, N = P->wsx * P->wsy, K = P->K;
  std::vector<float> f1w(N), f2w(N), cc(N);
  std::vector<double> dx(K), dy(K), pv(K);
  std::vector<uint8_t> vv(K);
  std::vector<int> pi(K);
  double tg = 0, tc = 0, tp = 0;
  for (int w = 0; w < nt; ++w) {
    const auto a = clk::now();
    gather<float>(P, w, f1w.data(), f2w.data());
    const auto b = clk::now();
    correlate<float>(f1w.data(), f2w.data(), P->wsx, P->wsy, P->method, cc.data());
    const auto c = clk::now();
    const float px = P->pred_x ? (float)P->pred_x[w] : 0.f, py = P->pred_y ? (float)P->pred_y[w] : 0.f;
    int ncand = 0;
    peaks_and_fit<float>(cc.data(), P->wsx, P->wsy, P->bx, P->by, K, px, py, (float)P->lx1, (float)P->lx2,
                         (float)P->ly1, (float)P->ly2, 0.0, 0.0, 0.0, dx.data(), dy.data(), vv.data(), pv.data(),
                         pi.data(), &ncand, nullptr);
    const auto d = clk::now();
    tg += std::chrono::duration<double>(b - a).count();
    tc += std::chrono::duration<double>(c - b).count();
    tp += std::chrono::duration<double>(d - c).count();
  }
  *t_gather = tg;
  *t_corr = tc;
  *t_peaks = tp;
  return 0;
}

int orc_deform(int prec, const uint8_t* f2, int nx, int ny, int left_x, int left_y, int nx_aug, int ny_aug,
               const double* ex_x, const double* ex_y, int nxe, int nye, double x_st, double y_st, double spx,
               double spy, double* out) {
  if (prec == 32)
    deform<float>(f2, nx, ny, left_x, left_y, nx_aug, ny_aug, ex_x, ex_y, nxe, nye, (float)x_st, (float)y_st,
                  (float)spx, (float)spy, out);
  else
    deform<double>(f2, nx, ny, left_x, left_y, nx_aug, ny_aug, ex_x, ex_y, nxe, nye, x_st, y_st, spx, spy, out);
  return 0;
}

int orc_extend(const double* fx, const double* fy, int nx, int ny, int b, double* ex, double* ey) {
  const int nxe = nx + 2 * b, nye = ny + 2 * b;
  for (int j = 0; j < nye; ++j)
    for (int i = 0; i < nxe; ++i) {
      const int si = std::min(std::max(i - b, 0), nx - 1), sj = std::min(std::max(j - b, 0), ny - 1);
      ex[j * nxe + i] = fx[sj * nx + si];
      ey[j * nxe + i] = fy[sj * nx + si];
    }
  return 0;
}

int orc_upsample(int prec, const double* ex, const double* ey, int nxe, int nye, double x_st1, double y_st1,
                 double sp1x, double sp1y, int sp2x, int sp2y, int nx2, int ny2, double* dx2, double* dy2) {
  if (prec == 32)
    upsample<float>(ex, ey, nxe, nye, (float)x_st1, (float)y_st1, (float)sp1x, (float)sp1y, sp2x, sp2y, nx2, ny2, dx2, dy2);
  else
    upsample<double>(ex, ey, nxe, nye, x_st1, y_st1, sp1x, sp1y, sp2x, sp2y, nx2, ny2, dx2, dy2);
  return 0;
}

int orc_median_pass(int prec, const double* sx, const double* sy, const uint8_t* sv, int sstride, const double* cx,
                    const double* cy, const uint8_t* cv, int K, int nx_vec, int ny_vec, int wx, int wy,
                    int min_nneigh, double thr, double eps, double* dx, double* dy, uint8_t* dv, double* margin) {
  if (prec == 32)
    median_pass<float>(sx, sy, sv, sstride, cx, cy, cv, K, nx_vec, ny_vec, wx, wy, min_nneigh, (float)thr, (float)eps,
                       dx, dy, dv, margin);
  else
    median_pass<double>(sx, sy, sv, sstride, cx, cy, cv, K, nx_vec, ny_vec, wx, wy, min_nneigh, thr, eps, dx, dy, dv,
                        margin);
  return 0;
}

int orc_fill(int prec, double* dx, double* dy, uint8_t* dv, int nx_vec, int ny_vec, int* sweeps) {
  const int s = prec == 32 ? fill<float>(dx, dy, dv, nx_vec, ny_vec) : fill<double>(dx, dy, dv, nx_vec, ny_vec);
  if (sweeps) *sweeps = s;
  return 0;
}

int orc_smooth(int prec, const double* sx, const double* sy, int nx_vec, int ny_vec, int wx, int wy, double* dx,
               double* dy) {
  if (prec == 32) smooth<float>(sx, sy, nx_vec, ny_vec, wx, wy, dx, dy);
  else smooth<double>(sx, sy, nx_vec, ny_vec, wx, wy, dx, dy);
  return 0;
}

int orc_run_pair(int prec, const orc_run* R, const uint8_t* f1, const uint8_t* f2, double* r1x, double* r1y,
                 uint8_t* r1v, double* o1x, double* o1y, uint8_t* o1v, double* r2x, double* r2y, uint8_t* r2v,
                 double* o2x, double* o2y, uint8_t* o2v, double* m2x, double* m2y, uint8_t* m2v) {
  if (!R || !f1 || !f2) return 1;
  if (prec == 32)
    return run_pair<float>(R, f1, f2, r1x, r1y, r1v, o1x, o1y, o1v, r2x, r2y, r2v, o2x, o2y, o2v, m2x, m2y, m2v);
  return run_pair<double>(R, f1, f2, r1x, r1y, r1v, o1x, o1y, o1v, r2x, r2y, r2v, o2x, o2y, o2v, m2x, m2y, m2v);
}

// MHE contrast enhancement (MHE7_4_gpu.cu = mhe), restated step by step:
// init geometry (mhe:149-178), neighbour counts (mhe:45-64), tile histograms
// with a zero border ring (mhe:66-83, 216), 3x3 transfer sums (mhe:85-101),
// inclusive prefix sum over the 256 levels (prescan_kernel writes the
// exclusive scan shifted by one, mhe:5-43), enhanced values (mhe:103-119) and
// the mapping of the tiled region (mhe:121-137); pixels outside it keep the 0
// that d_enhanced is cleared to at init (mhe:197).  Integer arithmetic only;
// the product c*255 is formed in 64 bits.
int orc_mhe(const uint8_t* in, uint8_t* out, int nx, int ny, int ws, float thresh) {
  if (ws < 1 || ws > nx || ws > ny) return 1;
  int width, height, xs, ys;
  if (nx % ws == 0 && ny % ws == 0) {
    width = nx;
    height = ny;
    xs = 0;
    ys = 0;
  } else {
    width = ws * (nx / ws);
    xs = (nx - width) / 2;
    height = ws * (ny / ws);
    ys = (ny - height) / 2;
  }
  const int maxI = height / ws, maxJ = width / ws;
  const int ntrash = (int)roundf((float)(ws * ws) * thresh);
  const int nnontrash = ws * ws - ntrash;
  if (nnontrash < 1) return 1;
  std::vector<int> cnt((size_t)maxI * maxJ);
  for (int j = 0; j < maxI; ++j)
    for (int i = 0; i < maxJ; ++i) {
      int c;
      if (i >= 1 && i <= maxJ - 2 && j >= 1 && j <= maxI - 2) c = 9;
      else if ((i == 0 && j == 0) || (i == 0 && j == maxI - 1) || (i == maxJ - 1 && j == 0) ||
               (i == maxJ - 1 && j == maxI - 1))
        c = 4;
      else c = 6;
      cnt[(size_t)j * maxJ + i] = c;
    }
  std::vector<int> levels((size_t)(maxI + 2) * (maxJ + 2) * 256, 0);
  for (int idy = 0; idy < height; ++idy)
    for (int idx = 0; idx < width; ++idx) {
      const int v = in[(size_t)(idy + ys) * nx + xs + idx];
      const int index = (maxJ + 2) * (idy / ws + 1) + 1 + idx / ws;
      ++levels[(size_t)index * 256 + v];
    }
  std::vector<int> lut((size_t)maxI * maxJ * 256);
  for (int ty = 0; ty < maxI; ++ty)
    for (int tx = 0; tx < maxJ; ++tx) {
      const int index = (ty + 1) * (maxJ + 2) + 1 + tx;
      const int n = cnt[(size_t)ty * maxJ + tx];
      long long cum = 0;
      for (int p = 0; p < 256; ++p) {
        long long t = 0;
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) t += levels[(size_t)(index + dy * (maxJ + 2) + dx) * 256 + p];
        cum += t;
        long long c = cum - (long long)n * ntrash;
        c = c >= 0 ? c : 0;
        c = std::min<long long>(255, c * 255 / ((long long)n * nnontrash));
        if (p == 0) c = 0;
        if (p == 255) c = 255;
        lut[((size_t)ty * maxJ + tx) * 256 + p] = (int)c;
      }
    }
  std::memset(out, 0, (size_t)nx * ny);
  for (int idy = 0; idy < height; ++idy)
    for (int idx = 0; idx < width; ++idx) {
      const size_t o = (size_t)(idy + ys) * nx + xs + idx;
      const int tile = (idy / ws) * maxJ + idx / ws;
      out[o] = (uint8_t)lut[(size_t)tile * 256 + in[o]];
    }
  return 0;
}

}  // extern "C"
