// CPU restatement of the learned grid-1 predictor (test infrastructure: the
// checker of the GPU network kernels, never part of the product).
//
// The network has no reference counterpart (it replaces the grid-1 passes of
// main.cpp:289-303); this file restates, in FP64 and with plain loops, what
// piv_xcorr2d_v1.0_amd/csrc/cnn.hip computes: six 3x3 convolutions with zero
// padding 1 (strides 2, 2, 2, 1, 1, 1; channels 2-32-64-64-64-64-2), leaky
// ReLU (slope 0.1) after all but the last, a mean-squared-error loss, and its
// gradient by the chain rule.  With bf16 != 0 the values the GPU stores in
// bfloat16 (network input, weights, hidden activations, loss gradient, the
// gradient through each leaky ReLU and each propagated input gradient) are
// rounded to bfloat16 (round to nearest even) at the same points, so that the
// GPU can be compared against it to accumulation-order rounding.
//
// Pinned by tests/test_oracle.py: central finite differences of the FP64 loss
// for sampled parameters of every layer, and closed forms (a single-tap
// kernel reproduces a shifted, strided copy of the input).
#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Layer {
  int cin, cout, stride, act;
};
const Layer kArch[6] = {{2, 32, 2, 1}, {32, 64, 2, 1}, {64, 64, 2, 1}, {64, 64, 1, 1}, {64, 64, 1, 1}, {64, 2, 1, 0}};
const double kLeak = 0.1;

double bf16(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}
inline double rnd(double x, bool on) { return on ? bf16(x) : x; }

size_t weight_offset(int l) {
  size_t n = 0;
  for (int i = 0; i < l; ++i) n += (size_t)kArch[i].cout * 9 * kArch[i].cin + kArch[i].cout;
  return n;
}

struct Net {
  int B, H, W;
  bool emu;
  std::vector<double> w[6], b[6];  // w[co][tap][ci]
  std::vector<double> act[7];      // act[0] = input, act[l + 1] = output of layer l, NHWC per sample
  int h[7], wd[7];
};

void load(Net& n, const double* params, int B, int H, int W, bool emu) {
  n.B = B;
  n.H = H;
  n.W = W;
  n.emu = emu;
  n.h[0] = H;
  n.wd[0] = W;
  for (int l = 0; l < 6; ++l) {
    const Layer& L = kArch[l];
    const size_t o = weight_offset(l), nw = (size_t)L.cout * 9 * L.cin;
    n.w[l].resize(nw);
    n.b[l].resize(L.cout);
    for (size_t i = 0; i < nw; ++i) n.w[l][i] = rnd(params[o + i], emu);
    for (int i = 0; i < L.cout; ++i) n.b[l][i] = emu ? (double)(float)params[o + nw + i] : params[o + nw + i];
    n.h[l + 1] = n.h[l] / L.stride;
    n.wd[l + 1] = n.wd[l] / L.stride;
  }
}

void forward(Net& n, const double* x0) {
  const size_t n0 = (size_t)n.B * n.H * n.W * 2;
  n.act[0].assign(x0, x0 + n0);
  if (n.emu)
    for (double& v : n.act[0]) v = bf16(v);
  for (int l = 0; l < 6; ++l) {
    const Layer& L = kArch[l];
    const int hi = n.h[l], wi = n.wd[l], ho = n.h[l + 1], wo = n.wd[l + 1];
    std::vector<double>& out = n.act[l + 1];
    out.assign((size_t)n.B * ho * wo * L.cout, 0.0);
    for (int bb = 0; bb < n.B; ++bb)
      for (int oy = 0; oy < ho; ++oy)
        for (int ox = 0; ox < wo; ++ox)
          for (int co = 0; co < L.cout; ++co) {
            double s = 0.0;
            for (int tap = 0; tap < 9; ++tap) {
              const int iy = oy * L.stride + tap / 3 - 1, ix = ox * L.stride + tap % 3 - 1;
              if (iy < 0 || iy >= hi || ix < 0 || ix >= wi) continue;
              const double* in = &n.act[l][(((size_t)bb * hi + iy) * wi + ix) * L.cin];
              const double* wr = &n.w[l][((size_t)co * 9 + tap) * L.cin];
              for (int ci = 0; ci < L.cin; ++ci) s += wr[ci] * in[ci];
            }
            double v = s + n.b[l][co];
            if (L.act) v = v > 0 ? v : v * kLeak;
            if (l < 5) v = rnd(v, n.emu);
            out[(((size_t)bb * ho + oy) * wo + ox) * L.cout + co] = v;
          }
  }
}

}  // namespace

extern "C" {

int orc_cnn_param_count(void) { return (int)weight_offset(6); }

int orc_cnn_forward(const double* params, const double* x0, int batch, int h, int w, int bf16_storage, double* out) {
  if (!params || !x0 || !out || batch < 1 || h % 8 || w % 8 || h < 8 || w < 8) return 1;
  Net n;
  load(n, params, batch, h, w, bf16_storage != 0);
  forward(n, x0);
  const int ho = n.h[6], wo = n.wd[6], hw = ho * wo;
  for (int bb = 0; bb < batch; ++bb)
    for (int p = 0; p < hw; ++p)
      for (int c = 0; c < 2; ++c) out[((size_t)bb * 2 + c) * hw + p] = n.act[6][((size_t)bb * hw + p) * 2 + c];
  return 0;
}

int orc_cnn_grad(const double* params, const double* x0, const double* labels, int batch, int h, int w,
                 int bf16_storage, double* loss, double* grad) {
  if (!params || !x0 || !labels || batch < 1 || h % 8 || w % 8 || h < 8 || w < 8) return 1;
  Net n;
  const bool emu = bf16_storage != 0;
  load(n, params, batch, h, w, emu);
  forward(n, x0);
  const int ho = n.h[6], wo = n.wd[6], hw = ho * wo;
  const double cnt = (double)batch * 2 * hw;
  // dL/d(output), NHWC
  std::vector<double> g((size_t)batch * hw * 2);
  double ls = 0.0;
  for (int bb = 0; bb < batch; ++bb)
    for (int c = 0; c < 2; ++c)
      for (int p = 0; p < hw; ++p) {
        const double d = n.act[6][((size_t)bb * hw + p) * 2 + c] - labels[((size_t)bb * 2 + c) * hw + p];
        ls += d * d;
        g[((size_t)bb * hw + p) * 2 + c] = rnd(2.0 * d / cnt, emu);
      }
  if (loss) *loss = ls / cnt;
  if (!grad) return 0;
  for (int l = 5; l >= 0; --l) {
    const Layer& L = kArch[l];
    const int hi = n.h[l], wi = n.wd[l], hh = n.h[l + 1], ww = n.wd[l + 1];
    // gradient through the activation, from the stored output
    std::vector<double> gp(g.size());
    for (size_t i = 0; i < g.size(); ++i) {
      if (!L.act || n.act[l + 1][i] > 0) gp[i] = g[i];
      else gp[i] = rnd(g[i] * kLeak, emu);
    }
    const size_t o = weight_offset(l), nw = (size_t)L.cout * 9 * L.cin;
    double* gw = grad + o;
    double* gb = grad + o + nw;
    std::fill(gw, gw + nw + L.cout, 0.0);
    std::vector<double> gin;
    if (l > 0) gin.assign((size_t)batch * hi * wi * L.cin, 0.0);
    for (int bb = 0; bb < batch; ++bb)
      for (int oy = 0; oy < hh; ++oy)
        for (int ox = 0; ox < ww; ++ox)
          for (int co = 0; co < L.cout; ++co) {
            const double gv = gp[(((size_t)bb * hh + oy) * ww + ox) * L.cout + co];
            gb[co] += gv;
            for (int tap = 0; tap < 9; ++tap) {
              const int iy = oy * L.stride + tap / 3 - 1, ix = ox * L.stride + tap % 3 - 1;
              if (iy < 0 || iy >= hi || ix < 0 || ix >= wi) continue;
              const size_t ii = (((size_t)bb * hi + iy) * wi + ix) * L.cin;
              const size_t wi0 = ((size_t)co * 9 + tap) * L.cin;
              for (int ci = 0; ci < L.cin; ++ci) {
                gw[wi0 + ci] += gv * n.act[l][ii + ci];
                if (l > 0) gin[ii + ci] += gv * n.w[l][wi0 + ci];
              }
            }
          }
    if (l > 0) {
      if (emu)
        for (double& v : gin) v = bf16(v);
      g.swap(gin);
    }
  }
  return 0;
}

}  // extern "C"
