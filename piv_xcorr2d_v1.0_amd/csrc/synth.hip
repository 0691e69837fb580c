// Device renderer of seeded synthetic particle-image pairs (no reference
// counterpart; the reference's inputs are not available).  Same particle
// model as the host generator pivx_synth_pair: Gaussian particles of
// standard deviation sigma and peak intensity U[imin, imax] uniformly over
// [-16, n + 16)^2, frame 2 rendering every particle displaced by the flow at
// its frame-1 position, background plus Gaussian noise, rounded and clipped
// to u8.  Random numbers are counter-based (particle and pixel index), and
// intensities accumulate in 32-bit fixed point (1/4096 grey level) with
// integer atomics, so a pair renders identically on every run.
#include "synth.h"
#include "kernels.h"

namespace pivx {

namespace {

constexpr float kFix = 4096.0f;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uni(unsigned long long h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ void splat(unsigned int* acc, int nx, int ny, float X, float Y, float amp, float inv2s2,
                                      int rad) {
  const int cx = (int)floorf(X + 0.5f), cy = (int)floorf(Y + 0.5f);
  for (int y = max(0, cy - rad); y <= min(ny - 1, cy + rad); ++y) {
    const float dy = (float)y - Y;
    for (int x = max(0, cx - rad); x <= min(nx - 1, cx + rad); ++x) {
      const float dx = (float)x - X;
      const unsigned int q = (unsigned int)(amp * __expf(-(dx * dx + dy * dy) * inv2s2) * kFix + 0.5f);
      if (q) atomicAdd(acc + (long long)y * nx + x, q);
    }
  }
}

__global__ void __launch_bounds__(256) k_synth_particles(SynthRender r, unsigned int* acc1, unsigned int* acc2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r.nparticles) return;
  const unsigned long long base = mix64(r.seed) ^ ((unsigned long long)i * 4ULL);
  const double px = -16.0 + uni(mix64(base)) * (r.nx + 32.0);
  const double py = -16.0 + uni(mix64(base + 1)) * (r.ny + 32.0);
  const float amp = (float)(r.imin + uni(mix64(base + 2)) * (r.imax - r.imin));
  double u, v;
  synth_flow_eval(r.flow, px, py, u, v);
  const float inv2s2 = 1.0f / (2.0f * r.sigma * r.sigma);
  const int rad = (int)ceilf(4.0f * r.sigma);
  splat(acc1, r.nx, r.ny, (float)px, (float)py, amp, inv2s2, rad);
  splat(acc2, r.nx, r.ny, (float)(px + u), (float)(py + v), amp, inv2s2, rad);
}

// background + noise + particles -> u8 (rows of pitch p), accumulators cleared
__global__ void __launch_bounds__(256) k_synth_finish(SynthRender r, unsigned int* acc1, unsigned int* acc2,
                                                      uint8_t* f1, long long p1, uint8_t* f2, long long p2) {
  const long long n = (long long)r.nx * r.ny;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = (int)(t / r.nx), x = (int)(t - (long long)y * r.nx);
  const unsigned long long nb = mix64(r.seed ^ 0x6e6f697365ULL);
#pragma unroll
  for (int frame = 0; frame < 2; ++frame) {
    unsigned int* acc = frame ? acc2 : acc1;
    const unsigned long long h = mix64(nb + (unsigned long long)t * 2 + frame);
    double u1 = uni(h), u2 = uni(mix64(h));
    if (u1 < 1e-300) u1 = 1e-300;
    const float g = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    const float val = r.background + r.noise_sigma * g + (float)acc[t] / kFix;
    const float q = fminf(255.f, fmaxf(0.f, rintf(val)));
    (frame ? f2 + (long long)y * p2 : f1 + (long long)y * p1)[x] = (uint8_t)q;
    acc[t] = 0;
  }
}

}  // namespace

// acc1, acc2: nx * ny zeroed accumulators (left zeroed on return)
hipError_t launch_synth(const SynthRender& r, unsigned int* acc1, unsigned int* acc2, uint8_t* f1, long long p1,
                        uint8_t* f2, long long p2, hipStream_t s) {
  if (r.nparticles > 0) {
    hipLaunchKernelGGL(k_synth_particles, dim3((unsigned)((r.nparticles + 255) / 256)), dim3(256), 0, s, r, acc1,
                       acc2);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const long long n = (long long)r.nx * r.ny;
  hipLaunchKernelGGL(k_synth_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, r, acc1, acc2, f1, p1, f2,
                     p2);
  return hipGetLastError();
}

}  // namespace pivx
