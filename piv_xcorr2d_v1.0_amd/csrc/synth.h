// Seeded synthetic particle-image pairs (no reference counterpart: the
// reference's input images are not available).  The flow model is shared by
// the host generator (hostio.cpp), the device renderer (synth.hip) and the
// analytic ground truth (pivx_synth_flow).
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PIVX_HD __host__ __device__
#else
#define PIVX_HD
#endif

namespace pivx {

// 0: uniform (a0, a1); 1: Lamb-Oseen vortex at (xc, yc), core radius a0,
// circulation gamma, plus shear a2 * (y - yc) in x; 2: sum of 8 Fourier modes
// scaled to a maximum displacement.
struct SynthFlow {
  int kind;
  double a0, a1, a2, xc, yc, gamma;
  double fp[8], fq[8], fa[8], fb[8], fphi[8], fpsi[8], scale;
};

PIVX_HD inline void synth_flow_eval(const SynthFlow& f, double x, double y, double& u, double& v) {
  const double pi = 3.14159265358979323846;
  if (f.kind == 0) {
    u = f.a0;
    v = f.a1;
  } else if (f.kind == 1) {
    const double dx = x - f.xc, dy = y - f.yc, r2 = dx * dx + dy * dy;
    const double ut_over_r = (r2 < 1e-12) ? f.gamma / (2 * pi * f.a0 * f.a0)
                                          : f.gamma / (2 * pi * r2) * (1 - exp(-r2 / (f.a0 * f.a0)));
    u = -ut_over_r * dy + f.a2 * dy;
    v = ut_over_r * dx;
  } else {
    u = v = 0;
    for (int m = 0; m < 8; ++m) {
      const double ph = 2 * pi * (f.fp[m] * x + f.fq[m] * y);
      u += f.fa[m] * sin(ph + f.fphi[m]);
      v += f.fb[m] * cos(ph + f.fpsi[m]);
    }
    u *= f.scale;
    v *= f.scale;
  }
}

// Render parameters of one pair on the device (particle model of
// pivx_synth_params, particle count round(density * (nx + 32) * (ny + 32))
// over [-16, n + 16)^2, counter-based random numbers from `seed`).
struct SynthRender {
  int nx, ny;
  unsigned long long seed;
  long long nparticles;
  float sigma, imin, imax, background, noise_sigma;
  SynthFlow flow;
};

}  // namespace pivx
