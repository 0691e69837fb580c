// Register-resident mixed-radix FFTs for small compile-time lengths (gfx950).
//
// One thread owns one whole 1-D line (N complex values held in VGPRs) and runs a
// fully unrolled Stockham autosort transform on it: every index and every
// twiddle factor is a compile-time constant, so the generated code is pure VALU
// work with literal operands and no memory traffic.  Trivial twiddles (multiples
// of pi/4) are specialised away.  The transform is unnormalised in both
// directions, matching the batched C2C transforms the reference plans at
// xcorr2d_multi_grid.cu:922-931 and executes at cu:1334-1350 (forward
// exp(-2*pi*i*k*n/N), inverse exp(+2*pi*i*k*n/N)).
//
// The Stockham stage is the textbook one (Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors", SC'08): for stage size Ns
// and radix R, butterfly j reads x[j + r*N/R], twiddles by w^(r*(j mod Ns)),
// runs a length-R DFT and writes y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].
#pragma once
#include <hip/hip_runtime.h>
#include <utility>

#ifndef PIVX_HD
#define PIVX_HD __host__ __device__ __forceinline__
#endif

namespace pivx {

// ---------------------------------------------------------------- constexpr trig
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*m/n with exact octant reduction done in integers.
struct CS { double c, s; };
constexpr CS cossin2pi(long m, long n) {
  long mm = m % n;
  if (mm < 0) mm += n;
  long e = 8 * mm;               // angle in units of pi/(4n)
  long o = e / n;                // octant 0..7
  long r = e - o * n;            // remainder in [0, n)
  double phi = double(r) / double(n) * (kPi / 4);        // in [0, pi/4)
  double psi = double(n - r) / double(n) * (kPi / 4);    // pi/4 - phi, in (0, pi/4]
  double sp = taylor_sin(phi), cp = taylor_cos(phi);
  double sq = taylor_sin(psi), cq = taylor_cos(psi);
  switch (o) {
    case 0: return {cp, sp};
    case 1: return {sq, cq};
    case 2: return {-sp, cp};
    case 3: return {-cq, sq};
    case 4: return {-cp, -sp};
    case 5: return {-sq, -cq};
    case 6: return {sp, -cp};
    default: return {cq, -sq};
  }
}

// ---------------------------------------------------------------- static_for
template <int... Is, class F>
PIVX_HD void static_for_impl(std::integer_sequence<int, Is...>, F&& f) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
PIVX_HD void static_for(F&& f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// ---------------------------------------------------------------- complex helpers
PIVX_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
PIVX_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// a * exp(sgn * 2*pi*i*M/N), sgn = -1 forward, +1 inverse.  M, N compile-time.
template <int M, int N, bool INV>
PIVX_HD float2 twiddle(float2 a) {
  constexpr int m = ((M % N) + N) % N;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (2 * m == N) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * m == N) {
    // forward: * (-i);  inverse: * (+i)
    if constexpr (INV) return make_float2(-a.y, a.x);
    else return make_float2(a.y, -a.x);
  } else if constexpr (4 * m == 3 * N) {
    if constexpr (INV) return make_float2(a.y, -a.x);
    else return make_float2(-a.y, a.x);
  } else if constexpr (8 * m == N || 8 * m == 3 * N || 8 * m == 5 * N || 8 * m == 7 * N) {
    // (+-1 +- i)/sqrt(2): two adds and two multiplies.
    constexpr float h = 0.70710678118654752440f;
    constexpr CS w = cossin2pi(m, N);
    constexpr float wc = w.c > 0 ? h : -h;
    constexpr float ws = (INV ? w.s : -w.s) > 0 ? h : -h;
    // (a.x + i a.y)(wc + i ws) with |wc| = |ws| = h
    if constexpr (wc == ws) {
      return make_float2((a.x - a.y) * wc, (a.x + a.y) * wc);
    } else {
      return make_float2((a.x + a.y) * wc, (a.y - a.x) * wc);
    }
  } else {
    constexpr CS w = cossin2pi(m, N);
    constexpr float c = float(w.c);
    constexpr float s = float(INV ? w.s : -w.s);
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// (a + w*b, a - w*b) with w = exp(sgn * 2*pi*i*M/N), the twiddled radix-2
// butterfly.  For a nontrivial w = c + i*s it is written in tangent form
// (w*b = c*(b.x - (s/c)*b.y, b.y + (s/c)*b.x), or the cotangent form when
// |s| > |c|, so the ratio never exceeds 1): two fused multiply-adds for the
// scaled product and four for the two outputs, six instructions where a
// complex multiply and four additions take eight.
template <int M, int N, bool INV>
PIVX_HD void bfly2(float2 a, float2 b, float2& p, float2& q) {
  constexpr int m = ((M % N) + N) % N;
  if constexpr (m == 0 || 2 * m == N || 4 * m == N || 4 * m == 3 * N) {
    const float2 wb = twiddle<m, N, INV>(b);
    p = cadd(a, wb);
    q = csub(a, wb);
  } else {
    constexpr CS w = cossin2pi(m, N);
    constexpr double c = w.c, sn = INV ? w.s : -w.s;
    if constexpr ((c < 0 ? -c : c) >= (sn < 0 ? -sn : sn)) {
      constexpr float t = float(sn / c), cf = float(c);
      const float u = fmaf(-t, b.y, b.x), v = fmaf(t, b.x, b.y);
      p = make_float2(fmaf(cf, u, a.x), fmaf(cf, v, a.y));
      q = make_float2(fmaf(-cf, u, a.x), fmaf(-cf, v, a.y));
    } else {
      constexpr float t = float(c / sn), sf = float(sn);
      const float u = fmaf(t, b.x, -b.y), v = fmaf(t, b.y, b.x);
      p = make_float2(fmaf(sf, u, a.x), fmaf(sf, v, a.y));
      q = make_float2(fmaf(-sf, u, a.x), fmaf(-sf, v, a.y));
    }
  }
}

// ---------------------------------------------------------------- small DFTs
template <int R, bool INV>
PIVX_HD void dft_small(float2 (&v)[R]) {
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    float2 t3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);  // d * (+-i)
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    constexpr float h = 0.86602540378443864676f;  // sqrt(3)/2
    float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    float2 x0 = cadd(v[0], s);
    float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    // forward: X1 = m - i*h*d, X2 = m + i*h*d
    float2 jd = INV ? make_float2(-h * d.y, h * d.x) : make_float2(h * d.y, -h * d.x);
    v[0] = x0;
    v[1] = cadd(m, jd);
    v[2] = csub(m, jd);
  } else {
    // generic odd radix: direct DFT with compile-time twiddles
    float2 in[R];
    static_for<R>([&](auto rc) { in[rc] = v[rc]; });
    static_for<R>([&](auto qc) {
      constexpr int q = qc;
      float2 acc = in[0];
      static_for<R - 1>([&](auto rc) {
        constexpr int r = rc + 1;
        acc = cadd(acc, twiddle<q * r, R, INV>(in[r]));
      });
      v[q] = acc;
    });
  }
}

constexpr int pick_radix(int rem) {
  if (rem % 4 == 0) return 4;
  if (rem % 2 == 0) return 2;
  for (int p = 3; p <= rem; p += 2)
    if (rem % p == 0) return p;
  return rem;
}

// Radix-2 and radix-4 stages run as twiddled radix-2 butterflies: a radix-4
// butterfly with inputs twiddled by w^0, w^m, w^2m, w^3m is
//   (v0, v2) -> v0 +- w^2m v2 = t0, t1;   (v1, v3) -> v1 +- w^2m v3 = s, d;
//   X0, X2 = t0 +- w^m s;   X1, X3 = t1 +- (-+i) w^m d
// where -+i is the radix-4 rotation, i.e. the twiddle index m + N/4.
template <int N, int R, int Ns, bool INV>
PIVX_HD void stockham_stage(const float2 (&x)[N], float2 (&y)[N]) {
  static_for<N / R>([&](auto jc) {
    constexpr int j = jc;
    constexpr int k = j % Ns;
    constexpr int m = k * (N / (Ns * R));
    constexpr int idxD = (j / Ns) * Ns * R + k;
    if constexpr (R == 2) {
      bfly2<m, N, INV>(x[j], x[j + N / 2], y[idxD], y[idxD + Ns]);
    } else if constexpr (R == 4) {
      float2 t0, t1, sv, dv;
      bfly2<2 * m, N, INV>(x[j], x[j + 2 * (N / 4)], t0, t1);
      bfly2<2 * m, N, INV>(x[j + N / 4], x[j + 3 * (N / 4)], sv, dv);
      bfly2<m, N, INV>(t0, sv, y[idxD], y[idxD + 2 * Ns]);
      bfly2<m + N / 4, N, INV>(t1, dv, y[idxD + Ns], y[idxD + 3 * Ns]);
    } else {
      float2 v[R];
      static_for<R>([&](auto rc) {
        constexpr int r = rc;
        v[r] = twiddle<(r * m) % N, N, INV>(x[j + r * (N / R)]);
      });
      dft_small<R, INV>(v);
      static_for<R>([&](auto rc) {
        constexpr int r = rc;
        y[idxD + r * Ns] = v[r];
      });
    }
  });
}

template <int N, int Ns, bool INV>
PIVX_HD void fft_stages(float2 (&x)[N]) {
  if constexpr (Ns < N) {
    constexpr int R = pick_radix(N / Ns);
    float2 y[N];
    stockham_stage<N, R, Ns, INV>(x, y);
    fft_stages<N, Ns * R, INV>(y);
    static_for<N>([&](auto ic) { x[ic] = y[ic]; });
  }
}

// In-register unnormalised DFT of length N (forward if !INV).
template <int N, bool INV>
PIVX_HD void fft_reg(float2 (&x)[N]) {
  fft_stages<N, 1, INV>(x);
}

}  // namespace pivx
