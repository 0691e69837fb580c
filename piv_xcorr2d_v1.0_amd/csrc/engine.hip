// Host engine of the MI355X PIV path: geometry, device memory, pass
// sequencing and the C ABI declared in include/piv_xcorr2d.h.
//
// Replaces the host side of class xcorr2d_multi
// (xcorr2d_V1.0/xcorr2d_multi_grid.cu:706-2063, .cuh:12-208).  Differences in
// form, not in results:
//   * frames live once in HBM as plain nx*ny u8 images; the four zero-padded
//     per-grid copies and their 4*Ny row copies (cu:933-954, 1271-1282) are
//     replaced by bounds-checked gathers that read 0 outside the frame;
//   * each correlation pass is one fused kernel (no cuFFT plans, no
//     sectioning: nsec_cc is validated and reported but every pass covers all
//     windows at once);
//   * the hole-filling loop runs on the device (no per-sweep host copy);
//   * everything is ordered on one HIP stream owned by the handle.
#include "../../include/piv_xcorr2d.h"
#include "kernels.h"
#include "cnn.h"
#include "hostio.h"
#include "synth.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace pivx;

namespace {

constexpr int kExBorder = 3;       // vec_ex_border_p1/p2 (cuh:190-191)
constexpr float kEpsilon = 0.1f;   // epsilon (cuh:192)

struct GridBuf {
  pivx_geometry g{};
  bool fast = false;
  float *dx = nullptr, *dy = nullptr;  // raw candidates [nt*K]
  uint8_t* dv = nullptr;
  float* pkv = nullptr;
  int* pki = nullptr;
  int* ncand = nullptr;                // [nt]
  float *fx = nullptr, *fy = nullptr;  // filtered field [nt]
  uint8_t* fv = nullptr;
  float *sx = nullptr, *sy = nullptr;  // ping-pong partner of the filtered field
  uint8_t* sv = nullptr;
  float *ux = nullptr, *uy = nullptr;  // second intermediate set of the median passes
  uint8_t* uv = nullptr;
  float *ex = nullptr, *ey = nullptr;  // extended field (nx+6)*(ny+6)
  float* deform = nullptr;             // deformed frame 2 [ny_aug][nx_aug]
  // Deformation fused into the gather of the correlation kernel (FusedDef):
  // possible when the windows do not overlap and take the wave-per-window
  // kernel; the interpolation terms per aug column and row.
  bool fuse = false;
  float4 *dcol = nullptr, *drow = nullptr;
  float2 *twx = nullptr, *twy = nullptr;   // generic-path twiddles
  float2* gscratch = nullptr;              // generic-path planes for big windows
  size_t gscratch_windows = 0;
  float* med_scratch = nullptr;            // generic median lists
  int8_t* accepted = nullptr;              // [nt] accepted candidate of the last median pass
  bool deformed_once = false;
};

}  // namespace

struct pivx_ctx {
  pivx_init_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  GridBuf grid[2];
  float *mo_x = nullptr, *mo_y = nullptr;  // grid-2 median-only field (run_pair)
  uint8_t* mo_v = nullptr;
  uint8_t *f1 = nullptr, *f2 = nullptr;    // frames: f1 [ny][nx]; f2 rows of pitch2 inside a zero border
  long long pitch2 = 0;
  bool have_frames = false;
  bool in_pair = false;                    // inside pivx_run_pair's sequence
  int fuse_mode = 1;                       // pivx_set_fused_deform
  // MHE enhancement of uploaded frames (MHE7_4_gpu.cu), off unless pivx_set_mhe
  bool mhe_on = false, mhe_ready = false;
  MheArgs mhe{};
  uint8_t* mhe_in = nullptr;    // frame staging (row pitch nx)
  uint8_t* mhe_out = nullptr;   // output of pivx_mhe_enhance
  unsigned long long* counters = nullptr;  // [2]: cumulative out-of-buffer deformation taps per grid
  int* sweeps = nullptr;                   // [2]
  int* nholes = nullptr;                   // hole-list counter of the fill kernels
  unsigned int* drops = nullptr;           // holes counted beyond the list capacity since the last check
  const int* gate = nullptr;               // while set, grid kernels run only if *gate != 0 (learned fallback)
  int prec = 0;                            // correlation precision form (pivx_set_correlation_precision)
  DftTables dtab[2] = {};                  // its DFT-matrix tables per grid
  // learned pipeline (pivx_run_pair_learned): fallback flag and outlier count
  // of the network field, network input and hidden activations (batch 1)
  int* lgate = nullptr;
  int* lcount = nullptr;
  int* lgate_work = nullptr;  // [3]: count accumulator and block ticket of the gated median, ticket of the
                              // fallback's correlation-and-filter launches; 0 between launches
  bool learned_ready = false;  // lgate, lcount and cact are allocated
  uint16_t* cact[5] = {};
  hipGraph_t lgraph = nullptr;
  hipGraphExec_t lgexec = nullptr;
  const pivx_cnn* lnet = nullptr;
  pivx_filter_params lf[2] = {};
  float lfrac = -1.f;
  unsigned int *sacc1 = nullptr, *sacc2 = nullptr;  // accumulators of the device renderer
  long long drops_total = 0;               // host total of *drops, reported by pivx_get_stats
  int *list_a = nullptr, *list_b = nullptr;
  int list_cap = 0;                        // entries of list_a / list_b
  float *tmp_x = nullptr, *tmp_y = nullptr;
  uint8_t* tmp_f = nullptr;
  std::vector<void*> allocs;
  std::string err;
  // Stage timing: event pairs recorded on the stream without host waits,
  // resolved by pivx_get_stage_ms.  Graph replays are timed whole (stage 5).
  struct Rec { int gi, stage, a, b; };
  struct Timeline {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    int start = -1;
    std::vector<Rec> recs;
  };
  bool timing = false;
  Timeline tl;
  // Graph mode: pivx_run_pair replays one captured hipGraph of the whole
  // per-pair sequence (re-captured when the filter parameters change).
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  pivx_filter_params gf[2] = {};
};

namespace {

int fail(pivx_ctx* c, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  set_global_error(buf);
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, PIVX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(pivx_ctx* c, T** p, size_t n) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(n * sizeof(T), 16);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return PIVX_OK;
}

// float -> bfloat16 bits, round to nearest even
uint16_t host_bf16(float f) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Reference vector count along one axis (cu:771-777).
int vec_count(int n, int sp) {
  int t1 = (n - sp / 2 - 1) / sp;
  int t2 = n - (sp / 2 + 1 + t1 * sp);
  while (t2 < (sp / 2 - 1)) {
    t1 -= 1;
    t2 += sp;
  }
  return t1 + 1;
}

// Geometry of one grid (cu:751-894).
pivx_geometry make_geometry(int nx, int ny, int nsec, int wsx, int wsy, int spx, int spy, float pctx, float pcty) {
  pivx_geometry g{};
  g.ws_x = wsx;
  g.ws_y = wsy;
  g.spacing_x = spx;
  g.spacing_y = spy;
  g.border_x = std::max(1, wsx / 2 - (int)roundf((float)wsx * pctx));
  g.border_y = std::max(1, wsy / 2 - (int)roundf((float)wsy * pcty));
  g.npeak_slots = wsx * wsy / 4 / 4;
  g.nx_vec = vec_count(nx, spx);
  g.ny_vec = vec_count(ny, spy);
  g.nt_vec = g.nx_vec * g.ny_vec;
  g.nsec = nsec;
  if (g.nt_vec > 0 && nsec > 0) {
    if (g.nt_vec % nsec == 0) {
      g.neach = g.nt_vec / nsec;
      g.nlast = g.neach;
    } else {
      g.neach = g.nt_vec / nsec;
      g.nlast = g.nt_vec - (nsec - 1) * g.neach;
      while (g.nlast > g.neach && g.neach > 0) {
        g.nlast -= g.neach;
        g.nsec += 1;
      }
    }
  }
  const int hx = wsx / 2, hy = wsy / 2;
  g.left_x = (spx / 2 < hx) ? hx - spx / 2 : 0;
  g.left_y = (spy / 2 < hy) ? hy - spy / 2 : 0;
  g.right_x = ((nx - spx / 2 - 1 - (g.nx_vec - 1) * spx) < (hx - 1)) ? spx / 2 + 1 + (g.nx_vec - 1) * spx + hx - 1 - nx : 0;
  g.right_y = ((ny - spy / 2 - 1 - (g.ny_vec - 1) * spy) < (hy - 1)) ? spy / 2 + 1 + (g.ny_vec - 1) * spy + hy - 1 - ny : 0;
  g.nx_aug = nx + g.left_x + g.right_x;
  g.ny_aug = ny + g.left_y + g.right_y;
  return g;
}

int validate(const pivx_init_params* p) {
  // Reference checks, same order and messages (cu:896-920).
  if (p->ws_x1 < 4 || p->ws_y1 < 4) return fail(nullptr, PIVX_ERR_PARAM, "The interrogation window size (1st grid) is too small!");
  if (p->ws_x1 > 1024 || p->ws_y1 > 1024) return fail(nullptr, PIVX_ERR_PARAM, "The interrogation window size (1st grid) is too big!");
  if (p->ws_x2 < 4 || p->ws_y2 < 4) return fail(nullptr, PIVX_ERR_PARAM, "The interrogation window size (2nd grid) is too small!");
  if (p->ws_x2 > 1024 || p->ws_y2 > 1024) return fail(nullptr, PIVX_ERR_PARAM, "The interrogation window size (2nd grid) is too big!");
  if (p->ws_x1 % 4 != 0 || p->ws_x2 % 4 != 0) return fail(nullptr, PIVX_ERR_PARAM, "The window x-dimension is not divisible by 4!");
  if (p->ws_y1 % 4 != 0 || p->ws_y2 % 4 != 0) return fail(nullptr, PIVX_ERR_PARAM, "The window y-dimension is not divisible by 4!");
  // Inputs on which the reference divides by zero or indexes out of bounds.
  if (p->nx < 1 || p->ny < 1) return fail(nullptr, PIVX_ERR_PARAM, "image size must be positive (nx=%d, ny=%d)", p->nx, p->ny);
  // the reference indexes frames with int (cu:933-954); frame offsets here are 32-bit too
  if ((long long)p->nx * p->ny > 0x7fffffffLL || p->nx >= (1 << 24) || p->ny >= (1 << 24))
    return fail(nullptr, PIVX_ERR_PARAM, "image too large (nx*ny = %lld pixels, at most 2^31-1)", (long long)p->nx * p->ny);
  if (p->nsec_cc < 1) return fail(nullptr, PIVX_ERR_PARAM, "Nsec_cc must be >= 1 (got %d)", p->nsec_cc);
  if (p->spacing_x1 < 1 || p->spacing_y1 < 1 || p->spacing_x2 < 1 || p->spacing_y2 < 1)
    return fail(nullptr, PIVX_ERR_PARAM, "vector spacing must be >= 1");
  if (p->npeak_save < 1) return fail(nullptr, PIVX_ERR_PARAM, "Npeak_save must be >= 1 (got %d)", p->npeak_save);
  if (p->npass1 < 0 || p->npass2 < 0) return fail(nullptr, PIVX_ERR_PARAM, "pass counts must be >= 0");
  if (p->wx1_median < 1 || p->wy1_median < 1 || p->wx2_median < 1 || p->wy2_median < 1 || p->wx1_smooth < 1 ||
      p->wy1_smooth < 1 || p->wx2_smooth < 1 || p->wy2_smooth < 1)
    return fail(nullptr, PIVX_ERR_PARAM, "median/smoothing window sizes must be >= 1");
  return PIVX_OK;
}

void host_twiddles(int n, std::vector<float2>& tw) {
  tw.resize(n);
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)n;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
}

GridBuf* grid_of(pivx_ctx* c, int grid) {
  if (!c || (grid != 1 && grid != 2)) return nullptr;
  return &c->grid[grid - 1];
}

int next_event(pivx_ctx::Timeline& t) {
  if (t.used == t.pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    t.pool.push_back(e);
  }
  return (int)t.used++;
}
void tic(pivx_ctx* c) {
  if (!c || !c->timing) return;
  pivx_ctx::Timeline& t = c->tl;
  t.start = next_event(t);
  if (t.start >= 0) (void)hipEventRecord(t.pool[t.start], c->stream);
}
void toc(pivx_ctx* c, int gi, int stage) {
  if (!c) return;
  pivx_ctx::Timeline& t = c->tl;
  if (!c->timing || t.start < 0) return;
  const int e = next_event(t);
  if (e < 0) return;
  (void)hipEventRecord(t.pool[e], c->stream);
  t.recs.push_back({gi, stage, t.start, e});
  t.start = -1;
}

int check_launch(pivx_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "GPU kernel error at %s: %s", what, hipGetErrorString(e));
  return PIVX_OK;
}

// Deformation arguments of grid gi (cu:1300-1301, 1624-1625).
DefArgs def_args(pivx_ctx* c, int gi) {
  GridBuf& G = c->grid[gi];
  const pivx_geometry& g = G.g;
  DefArgs d{};
  d.out = G.deform;
  d.nx_aug = g.nx_aug;
  d.ny_aug = g.ny_aug;
  d.f2 = c->f2;
  d.pitch = c->pitch2;
  d.nx = c->p.nx;
  d.ny = c->p.ny;
  d.left_x = g.left_x;
  d.left_y = g.left_y;
  d.hfws_x = g.ws_x / 2;
  d.hfws_y = g.ws_y / 2;
  d.ex_x = G.ex;
  d.ex_y = G.ey;
  d.nxe = g.nx_vec + 2 * kExBorder;
  d.nye = g.ny_vec + 2 * kExBorder;
  d.x_st = (float)(g.ws_x / 2 - g.spacing_x * kExBorder);
  d.y_st = (float)(g.ws_y / 2 - g.spacing_y * kExBorder);
  d.spx = (float)g.spacing_x;
  d.spy = (float)g.spacing_y;
  d.oob = c->counters + gi;
  d.gate = c->gate;
  return d;
}

// Deformation of frame 2 by the extended field of grid gi.
int run_deform(pivx_ctx* c, int gi) {
  GridBuf& G = c->grid[gi];
  const DefArgs d = def_args(c, gi);
  tic(c);
  const int rc = check_launch(c, launch_deform(d, c->stream), "Kernel_Window_Deformation_2d");
  toc(c, gi, 0);
  G.deformed_once = true;
  return rc;
}

// With zero_pred set (first grid-1 pass) the predictor is 0 and the kernel
// itself zeroes the filtered field, replacing set_float_zero_kernel
// (cu:1287-1298).
// With fused set, frame 2 is deformed inside the correlation kernel's gather
// (FusedDef) and the deformed plane is not written.
// Arguments of one correlation pass of grid gi (shared by the per-handle
// launch and the batched launch over many handles).
XcArgs make_xcargs(pivx_ctx* c, int gi, bool deformed, bool zero_pred, bool fused) {
  GridBuf& G = c->grid[gi];
  const pivx_geometry& g = G.g;
  XcArgs a{};
  a.f1 = c->f1;
  a.f2u8 = deformed && !fused ? nullptr : c->f2;
  a.f2def = deformed && !fused ? G.deform : nullptr;
  if (fused) {
    a.fdef.col = G.dcol;
    a.fdef.row = G.drow;
    a.fdef.ex_x = G.ex;
    a.fdef.ex_y = G.ey;
    a.fdef.nxe = g.nx_vec + 2 * kExBorder;
    a.fdef.hfws_x = g.ws_x / 2;
    a.fdef.hfws_y = g.ws_y / 2;
    a.fdef.nx_aug = g.nx_aug;
    a.fdef.ny_aug = g.ny_aug;
    // with overlapping windows (the learned pipeline's fallback) a pixel is
    // deformed once per window that holds it: its taps are not counted
    a.fdef.oob = G.fuse ? c->counters + gi : nullptr;
  }
  a.pitch1 = c->p.nx;
  a.pitch2 = c->pitch2;
  a.def_pitch = g.nx_aug;
  a.nx = c->p.nx;
  a.ny = c->p.ny;
  a.left_x = g.left_x;
  a.left_y = g.left_y;
  a.nx_vec = g.nx_vec;
  a.ny_vec = g.ny_vec;
  a.sp_x = g.spacing_x;
  a.sp_y = g.spacing_y;
  a.bx = g.border_x;
  a.by = g.border_y;
  a.K = c->p.npeak_save;
  // Both grids use the first-grid limit variables (cu:1440, 1755).
  a.lx1 = c->p.dispx_lim1;
  a.lx2 = c->p.dispx_lim2;
  a.ly1 = c->p.dispy_lim1;
  a.ly2 = c->p.dispy_lim2;
  a.pred_x = zero_pred ? nullptr : G.fx;
  a.pred_y = zero_pred ? nullptr : G.fy;
  a.zero_x = zero_pred ? G.fx : nullptr;
  a.zero_y = zero_pred ? G.fy : nullptr;
  a.disp_x = G.dx;
  a.disp_y = G.dy;
  a.valid = G.dv;
  a.pk_val = G.pkv;
  a.pk_idx = G.pki;
  a.ncand = G.ncand;
  a.gate = c->gate;
  return a;
}

int run_xcorr(pivx_ctx* c, int gi, bool deformed, bool zero_pred = false, bool fused = false) {
  GridBuf& G = c->grid[gi];
  const pivx_geometry& g = G.g;
  const XcArgs a = make_xcargs(c, gi, deformed, zero_pred, fused);
  tic(c);
  hipError_t e;
  if (c->prec) e = launch_xcorr_dft(a, g.ws_x, g.ws_y, c->prec, c->dtab[gi], c->stream);
  else if (G.fast) e = launch_xcorr(a, g.ws_x, g.ws_y, nullptr, 0, nullptr, nullptr, c->stream);
  else e = launch_xcorr(a, g.ws_x, g.ws_y, G.gscratch, G.gscratch_windows, G.twx, G.twy, c->stream);
  const int rc = check_launch(c, e, "correlation");
  toc(c, gi, 1);
  return rc;
}

// Everything one filter sequence reads and works in: the raw candidates, the
// field and window sizes, and the work buffers (owned by a handle, or carved
// from the static scratch of pivx_filter_field_noalloc).
struct FilterJob {
  const float* cx; const float* cy; const uint8_t* cv; int K;  // raw candidates [nt][K]
  int nx_vec, ny_vec;
  int wxm, wym, wxs, wys;                                        // median and smoothing windows
  float *sx, *sy; uint8_t* sv;                                   // ping-pong partner of the output
  float *ux, *uy; uint8_t* uv;                                   // second intermediate set
  int *list_a, *list_b; float *tmp_x, *tmp_y; uint8_t* tmp_f;    // hole lists and pending fills
  int* nholes; unsigned int* drops; int* sweeps; int list_cap;
  float* med_scratch;                                            // generic median lists (wx*wy > 64) or null
  int8_t* accepted;                                              // accepted candidate per vector or null
  hipStream_t stream;
  const int* gate;                                               // launches run only if *gate != 0 (null: always)
  int fp64;                                                      // FP64 residual normalisation (precision form)
};

// MedianFilter_Smoothing_p1/p2 (cu:1451-1545, 1766-1863) writing into (ox, oy, ov).
// Median passes, hole filling and smoothing of one grid into (ox, oy, ov).
// Buffers are planned so that no pass needs a copy: the last median pass
// writes the valid flags to ov and the components to T0 = (ox, oy) when the
// number of smoothing passes is even, else to the partner (sx, sy), so that
// the smoothing ping-pong ends in (ox, oy); earlier median passes alternate
// between two other buffer sets.  With `ext` set and at least one smoothing
// pass, the last smoothing pass also writes the extended field (ext_x, ext_y)
// and *ext_done is set.  With (mx, my, mv) set, the field after the median
// stage (before filling and smoothing) is also written there.  c may be null
// (no stage timing; errors go to the context-free error text).
// Whether a gated filter sequence runs as one FilterSeq (k_filter_seq).
bool filter_fits_seq(int npass_median, int npass_smooth, int wxm, int wym) {
  const int nsteps = (npass_median > 0 ? 1 + npass_median : 2) + 2 + std::max(0, npass_smooth);
  return (npass_median <= 0 || filter_seq_supported(wxm, wym)) && nsteps <= kFilterSeqSteps;
}

// With defer set (a gated sequence that fits one FilterSeq), the sequence is
// returned in *defer instead of launched.
int enqueue_filter(pivx_ctx* c, int gi, const FilterJob& J, int npass_median, int min_nneigh, float insert_thr,
                   int npass_smooth, int iffill, float* ox, float* oy, uint8_t* ov, float* ext_x, float* ext_y,
                   bool* ext_done, float* mx, float* my, uint8_t* mv, FilterSeq* defer = nullptr) {
  const int nt = J.nx_vec * J.ny_vec;
  const int K = J.K;
  const int wxm = J.wxm, wym = J.wym, wxs = J.wxs, wys = J.wys;
  const hipStream_t stream = J.stream;
  if (ext_done) *ext_done = false;
  const bool smooth_odd = npass_smooth > 0 && npass_smooth % 2 == 1;
  float* t0x = smooth_odd ? J.sx : ox;
  float* t0y = smooth_odd ? J.sy : oy;
  const bool fill = iffill == 1;
  const bool listed = fill && npass_median > 0;  // the last median pass lists the holes
  int rc;
  // The gated grid-1 fallback of the learned pipeline mostly does nothing, yet
  // every launch of it costs its stream a dispatch: it runs the whole sequence
  // as one launch of k_filter_seq (one workgroup, the same per-vector code)
  // when the median window has the register path and the steps fit.
  const bool seq = J.gate != nullptr && filter_fits_seq(npass_median, npass_smooth, wxm, wym);
  if (defer && !seq) return fail(c, PIVX_ERR_STATE, "filter sequence cannot be deferred");
  FilterSeq q{};
  q.wxm = wxm;
  q.wym = wym;
  q.gate = J.gate;
  auto add_step = [&](int kind) -> FilterStep& {
    FilterStep& st = q.step[q.nsteps++];
    st.kind = kind;
    return st;
  };
  auto prep = [&](float* dx, float* dy, uint8_t* dv, int* reset) -> hipError_t {
    if (!seq) return launch_prep(J.cx, J.cy, J.cv, K, dx, dy, dv, nt, reset, stream, J.gate);
    add_step(FS_PREP).prep = PrepArgs{J.cx, J.cy, J.cv, K, dx, dy, dv, nt, reset};
    return hipSuccess;
  };
  auto zero = [&](int* p) -> hipError_t {
    if (!seq) return launch_zero_ints(p, 1, stream, J.gate);
    add_step(FS_ZERO).zero = p;
    return hipSuccess;
  };
  auto median = [&](const MedArgs& m) -> hipError_t {
    if (!seq) return launch_median(m, stream);
    add_step(FS_MEDIAN).med = m;
    return hipSuccess;
  };
  auto fill_holes = [&](const FillArgs& f) -> hipError_t {
    if (!seq) return launch_fill(f, stream);
    if (!f.prelisted) add_step(FS_FILL_LIST).fill = f;
    add_step(FS_FILL).fill = f;
    return hipSuccess;
  };
  auto smooth = [&](const SmoothArgs& sa) -> hipError_t {
    if (!seq) return launch_smooth(sa, stream);
    add_step(FS_SMOOTH).smooth = sa;
    return hipSuccess;
  };
  if (!defer) tic(c);  // a deferred sequence is launched by the caller, untimed
  // The hole counter must be 0 when the holes are listed (by the last median
  // pass, or by k_fill_list without median passes).  k_fill zeroes it after
  // use; the launch ahead of the listing one zeroes it as well, so that no
  // earlier sequence (an interrupted pair, a caller's partial sequence) can
  // leave a stale count behind.
  if (npass_median <= 0) {
    rc = check_launch(c, prep(t0x, t0y, ov, fill ? J.nholes : nullptr), "median_filter_data_prep");
    if (!rc && mx) rc = check_launch(c, prep(mx, my, mv, nullptr), "median_filter_data_prep");
    if (rc) return rc;
  } else {
    if (listed && npass_median == 1 && (rc = check_launch(c, zero(J.nholes), "hole counter"))) return rc;
    // intermediate sets: I1 = (ux, uy, uv), I2 = whichever of (ox, oy) / (sx, sy) is not T0, with sv
    float* i2x = smooth_odd ? ox : J.sx;
    float* i2y = smooth_odd ? oy : J.sy;
    const float* srcx = J.cx;
    const float* srcy = J.cy;
    const uint8_t* srcv = J.cv;
    int sstride = K;
    for (int p = 0; p < npass_median; ++p) {
      const int back = npass_median - 1 - p;  // 0 for the last pass
      float* tx = back == 0 ? t0x : (back % 2 == 1 ? J.ux : i2x);
      float* ty = back == 0 ? t0y : (back % 2 == 1 ? J.uy : i2y);
      uint8_t* tv = back == 0 ? ov : (back % 2 == 1 ? J.uv : J.sv);
      MedArgs m{};
      m.sx = srcx;
      m.sy = srcy;
      m.sv = srcv;
      m.sstride = sstride;
      m.cx = J.cx;
      m.cy = J.cy;
      m.cv = J.cv;
      m.K = K;
      m.dx = tx;
      m.dy = ty;
      m.dv = tv;
      if (back == 0) {
        m.dx2 = mx;
        m.dy2 = my;
        m.dv2 = mv;
      }
      m.nx_vec = J.nx_vec;
      m.ny_vec = J.ny_vec;
      m.wx = wxm;
      m.wy = wym;
      m.hfwx = (wxm - 1) / 2;
      m.hfwy = (wym - 1) / 2;
      m.min_nneigh = min_nneigh;
      m.thr = insert_thr;
      m.eps = kEpsilon;
      m.accepted = back == 0 ? J.accepted : nullptr;
      m.scratch = J.med_scratch;
      m.hole_list = (back == 0 && listed) ? J.list_a : nullptr;
      m.nholes = (back == 0 && listed) ? J.nholes : nullptr;
      m.hole_cap = J.list_cap;
      m.reset_count = (listed && p == 0 && back > 0) ? J.nholes : nullptr;
      m.gate = J.gate;
      m.fp64 = J.fp64;
      rc = check_launch(c, median(m), "median_filtering");
      if (rc) return rc;
      srcx = tx;
      srcy = ty;
      srcv = tv;
      sstride = 1;
    }
  }
  if (!seq) toc(c, gi, 2);
  if (fill) {
    FillArgs f{};
    f.dx = t0x;
    f.dy = t0y;
    f.dv = ov;
    f.nx_vec = J.nx_vec;
    f.ny_vec = J.ny_vec;
    f.list_a = J.list_a;
    f.list_b = J.list_b;
    f.tmp_x = J.tmp_x;
    f.tmp_y = J.tmp_y;
    f.tmp_f = J.tmp_f;
    f.nholes = J.nholes;
    f.list_cap = J.list_cap;
    f.drops = J.drops;
    f.gate = J.gate;
    f.sweeps = J.sweeps;
    f.prelisted = listed ? 1 : 0;
    if (!seq) tic(c);
    rc = check_launch(c, fill_holes(f), "fill_blank_vecfield");
    if (!seq) toc(c, gi, 3);
    if (rc) return rc;
  }
  if (npass_smooth > 0) {
    if (!seq) tic(c);
    float *cx = t0x, *cy = t0y;
    for (int s = 0; s < npass_smooth; ++s) {
      float* nx_ = (cx == ox) ? J.sx : ox;
      float* ny_ = (cy == oy) ? J.sy : oy;
      SmoothArgs sa{};
      sa.gate = J.gate;
      sa.sx = cx;
      sa.sy = cy;
      sa.dx = nx_;
      sa.dy = ny_;
      sa.nx_vec = J.nx_vec;
      sa.ny_vec = J.ny_vec;
      sa.wx = wxs;
      sa.wy = wys;
      sa.hfwx = (wxs - 1) / 2;
      sa.hfwy = (wys - 1) / 2;
      if (s == npass_smooth - 1 && ext_x) {
        sa.ex = ext_x;
        sa.ey = ext_y;
        sa.border = kExBorder;
        if (ext_done) *ext_done = true;
      }
      rc = check_launch(c, smooth(sa), "smoothing_tophat");
      if (rc) return rc;
      cx = nx_;
      cy = ny_;
    }
    if (!seq) toc(c, gi, 4);
  }
  if (seq && defer) {
    *defer = q;
  } else if (seq) {
    rc = check_launch(c, launch_filter_seq(q, stream), "filter sequence");
    toc(c, gi, 2);
    if (rc) return rc;
  }
  return PIVX_OK;
}

int run_filter(pivx_ctx* c, int gi, int npass_median, int min_nneigh, float insert_thr, int npass_smooth, int iffill,
               float* ox, float* oy, uint8_t* ov, float* ext_x = nullptr, float* ext_y = nullptr,
               bool* ext_done = nullptr, float* mx = nullptr, float* my = nullptr, uint8_t* mv = nullptr,
               FilterSeq* defer = nullptr) {
  GridBuf& G = c->grid[gi];
  FilterJob J{};
  J.cx = G.dx;
  J.cy = G.dy;
  J.cv = G.dv;
  J.K = c->p.npeak_save;
  J.nx_vec = G.g.nx_vec;
  J.ny_vec = G.g.ny_vec;
  J.wxm = gi == 0 ? c->p.wx1_median : c->p.wx2_median;
  J.wym = gi == 0 ? c->p.wy1_median : c->p.wy2_median;
  J.wxs = gi == 0 ? c->p.wx1_smooth : c->p.wx2_smooth;
  J.wys = gi == 0 ? c->p.wy1_smooth : c->p.wy2_smooth;
  J.sx = G.sx;
  J.sy = G.sy;
  J.sv = G.sv;
  J.ux = G.ux;
  J.uy = G.uy;
  J.uv = G.uv;
  J.list_a = c->list_a;
  J.list_b = c->list_b;
  J.tmp_x = c->tmp_x;
  J.tmp_y = c->tmp_y;
  J.tmp_f = c->tmp_f;
  J.nholes = c->nholes;
  J.drops = c->drops;
  J.sweeps = c->sweeps + gi;
  J.list_cap = c->list_cap;
  J.med_scratch = G.med_scratch;
  J.accepted = G.accepted;
  J.stream = c->stream;
  J.gate = c->gate;
  J.fp64 = c->prec != 0 ? 1 : 0;
  return enqueue_filter(c, gi, J, npass_median, min_nneigh, insert_thr, npass_smooth, iffill, ox, oy, ov, ext_x, ext_y,
                        ext_done, mx, my, mv, defer);
}

int extend(pivx_ctx* c, int gi) {
  GridBuf& G = c->grid[gi];
  return check_launch(c, launch_extend(G.fx, G.fy, G.ex, G.ey, G.g.nx_vec, G.g.ny_vec, kExBorder, c->stream, c->gate),
                      "extend_filtered_vecfield");
}

int alloc_grid(pivx_ctx* c, int gi) {
  GridBuf& G = c->grid[gi];
  const pivx_geometry& g = G.g;
  const size_t nt = (size_t)g.nt_vec, K = (size_t)c->p.npeak_save;
  int rc = 0;
  rc |= dalloc(c, &G.dx, nt * K);
  rc |= dalloc(c, &G.dy, nt * K);
  rc |= dalloc(c, &G.dv, nt * K);
  rc |= dalloc(c, &G.pkv, nt * K);
  rc |= dalloc(c, &G.pki, nt * K);
  rc |= dalloc(c, &G.ncand, nt);
  rc |= dalloc(c, &G.accepted, nt);
  rc |= dalloc(c, &G.fx, nt);
  rc |= dalloc(c, &G.fy, nt);
  rc |= dalloc(c, &G.fv, nt);
  rc |= dalloc(c, &G.sx, nt);
  rc |= dalloc(c, &G.sy, nt);
  rc |= dalloc(c, &G.sv, nt);
  rc |= dalloc(c, &G.ux, nt);
  rc |= dalloc(c, &G.uy, nt);
  rc |= dalloc(c, &G.uv, nt);
  const size_t ne = (size_t)(g.nx_vec + 2 * kExBorder) * (g.ny_vec + 2 * kExBorder);
  rc |= dalloc(c, &G.ex, ne);
  rc |= dalloc(c, &G.ey, ne);
  rc |= dalloc(c, &G.deform, (size_t)g.nx_aug * g.ny_aug);
  if (rc) return PIVX_ERR_HIP;
  HIPCHK(c, hipMemsetAsync(G.ex, 0, ne * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.ey, 0, ne * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.fx, 0, nt * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.fy, 0, nt * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.fv, 0, nt, c->stream));
  HIPCHK(c, hipMemsetAsync(G.dx, 0, nt * K * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.dy, 0, nt * K * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(G.dv, 0, nt * K, c->stream));
  HIPCHK(c, hipMemsetAsync(G.ncand, 0, nt * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(G.accepted, 0, nt, c->stream));
  G.fast = xcorr_fast_supported(g.ws_x, g.ws_y);
  if (!G.fast) {
    std::vector<float2> tx, ty;
    host_twiddles(g.ws_x, tx);
    host_twiddles(g.ws_y, ty);
    if (dalloc(c, &G.twx, tx.size()) || dalloc(c, &G.twy, ty.size())) return PIVX_ERR_HIP;
    HIPCHK(c, hipMemcpy(G.twx, tx.data(), tx.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(G.twy, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice));
    const size_t plane = (size_t)g.ws_x * g.ws_y * sizeof(float2);
    if (2 * plane > 64 * 1024) {
      // about 1 GiB of plane scratch, at most all windows
      size_t w = std::max<size_t>(1, (size_t(1) << 30) / (2 * plane));
      w = std::min<size_t>(w, nt);
      G.gscratch_windows = w;
      if (dalloc(c, &G.gscratch, w * 2 * (size_t)g.ws_x * g.ws_y)) return PIVX_ERR_HIP;
    }
  }
  const int wxm = gi == 0 ? c->p.wx1_median : c->p.wx2_median;
  const int wym = gi == 0 ? c->p.wy1_median : c->p.wy2_median;
  if ((long long)wxm * wym > 64) {
    if (dalloc(c, &G.med_scratch, nt * 2 * (size_t)wxm * wym)) return PIVX_ERR_HIP;
  }
  // fused deformation: each aug pixel in at most one window (the learned
  // pipeline's gated fallback fuses it with overlapping windows as well)
  G.fuse = xcorr_fuses_deform(g.ws_x, g.ws_y) && g.spacing_x >= g.ws_x && g.spacing_y >= g.ws_y;
  if (xcorr_fuses_deform(g.ws_x, g.ws_y)) {
    if (dalloc(c, &G.dcol, (size_t)g.nx_aug) || dalloc(c, &G.drow, (size_t)g.ny_aug)) return PIVX_ERR_HIP;
    const int rc2 = check_launch(c, launch_deform_terms(def_args(c, gi), G.dcol, G.drow, c->stream), "deformation terms");
    if (rc2) return rc2;
  }
  return PIVX_OK;
}

// Zero border of frame 2 for grid g (pixels left/right, rows top/bottom):
// the reference's aug_ex buffer of the grid, aug [-hfws, n_aug + hfws - 1)
// in each axis (image = aug - left, cu:885-894), plus the row below it that
// the deformation's unchecked tap resource spans.
void frame2_border(const pivx_geometry& g, int nx, int ny, int& bx, int& by) {
  const int hx = g.ws_x / 2, hy = g.ws_y / 2;
  bx = std::max({0, hx + g.left_x, g.nx_aug + hx - 1 - g.left_x - nx});
  by = std::max({0, hy + g.left_y, g.ny_aug + hy + 2 - g.left_y - ny});
}

}  // namespace

// ---------------------------------------------------------------- allocation-free filter
// Work memory of pivx_median_filter_smoothing_noalloc: a file-scope device
// array (each GPU holds its own copy for the life of the process), sized for
// PIVX_NOALLOC_MAX_VECTORS vectors: six float planes (two intermediate field
// sets, pending fills), two int hole lists, three byte planes and a 256-byte
// block of counters.  Host code never uses the array's name as an address:
// the address of each device's copy comes from hipGetSymbolAddress, taken
// once per device with that device current.
constexpr size_t kNaN = PIVX_NOALLOC_MAX_VECTORS;
constexpr size_t kNaBytes = 256 + 8 * 4 * kNaN + 3 * kNaN;
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

int scratch_of_device(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, PIVX_ERR_PARAM, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(nullptr, PIVX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return PIVX_OK;
}
}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* pivx_last_error(const pivx_ctx* ctx) { return ctx ? ctx->err.c_str() : global_error(); }

int pivx_init(const pivx_init_params* p, int device_id, pivx_ctx** out) {
  if (!p || !out) return fail(nullptr, PIVX_ERR_PARAM, "null argument");
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  pivx_geometry g1 = make_geometry(p->nx, p->ny, p->nsec_cc, p->ws_x1, p->ws_y1, p->spacing_x1, p->spacing_y1,
                                   p->max_dispx_percent1, p->max_dispy_percent1);
  pivx_geometry g2 = make_geometry(p->nx, p->ny, p->nsec_cc, p->ws_x2, p->ws_y2, p->spacing_x2, p->spacing_y2,
                                   p->max_dispx_percent2, p->max_dispy_percent2);
  if (g1.nx_vec < 1 || g1.ny_vec < 1 || g2.nx_vec < 1 || g2.ny_vec < 1)
    return fail(nullptr, PIVX_ERR_PARAM, "image %dx%d holds no vector at the given spacing", p->nx, p->ny);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) return fail(nullptr, PIVX_ERR_HIP, "No GPU on this machine!");
  if (device_id < 0 || device_id >= ndev)
    return fail(nullptr, PIVX_ERR_PARAM, "device id %d out of range (%d devices)", device_id, ndev);
  e = hipSetDevice(device_id);
  if (e != hipSuccess) return fail(nullptr, PIVX_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  pivx_ctx* c = new pivx_ctx();
  c->p = *p;
  c->device = device_id;
  c->grid[0].g = g1;
  c->grid[1].g = g2;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(nullptr, PIVX_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  const size_t npix = (size_t)p->nx * p->ny;
  const size_t ntmax = (size_t)std::max(g1.nt_vec, g2.nt_vec);
  rc = 0;
  rc |= dalloc(c, &c->f1, npix);
  // frame 2 inside a zero border wide enough for both grids (the deformation
  // reads taps of in-range cells from it without checks); cleared once here,
  // uploads write the interior only
  int bx1, by1, bx2, by2;
  frame2_border(g1, p->nx, p->ny, bx1, by1);
  frame2_border(g2, p->nx, p->ny, bx2, by2);
  const int bx = (std::max(bx1, bx2) + 3) & ~3, by = std::max(by1, by2);
  c->pitch2 = p->nx + 2LL * bx;
  const size_t f2bytes = (size_t)c->pitch2 * ((size_t)p->ny + 2 * by);
  uint8_t* f2buf = nullptr;
  rc |= dalloc(c, &f2buf, f2bytes);
  if (f2buf) {
    c->f2 = f2buf + (size_t)by * c->pitch2 + bx;
    if (hipMemsetAsync(f2buf, 0, f2bytes, c->stream) != hipSuccess) rc |= fail(c, PIVX_ERR_HIP, "frame border clear failed");
  }
  rc |= dalloc(c, &c->counters, 2);
  rc |= dalloc(c, &c->sweeps, 2);
  rc |= dalloc(c, &c->nholes, 1);
  rc |= dalloc(c, &c->drops, 1);
  c->list_cap = (int)ntmax;
  rc |= dalloc(c, &c->list_a, ntmax);
  rc |= dalloc(c, &c->list_b, ntmax);
  rc |= dalloc(c, &c->tmp_x, ntmax);
  rc |= dalloc(c, &c->tmp_y, ntmax);
  rc |= dalloc(c, &c->tmp_f, ntmax);
  rc |= dalloc(c, &c->mo_x, (size_t)g2.nt_vec);
  rc |= dalloc(c, &c->mo_y, (size_t)g2.nt_vec);
  rc |= dalloc(c, &c->mo_v, (size_t)g2.nt_vec);
  if (rc == 0) rc = alloc_grid(c, 0);
  if (rc == 0) rc = alloc_grid(c, 1);
  if (rc == 0 && hipMemsetAsync(c->counters, 0, 2 * sizeof(unsigned long long), c->stream) != hipSuccess) rc = PIVX_ERR_HIP;
  if (rc == 0 && hipMemsetAsync(c->sweeps, 0, 2 * sizeof(int), c->stream) != hipSuccess) rc = PIVX_ERR_HIP;
  if (rc == 0 && hipMemsetAsync(c->nholes, 0, sizeof(int), c->stream) != hipSuccess) rc = PIVX_ERR_HIP;
  if (rc == 0 && hipMemsetAsync(c->drops, 0, sizeof(unsigned int), c->stream) != hipSuccess) rc = PIVX_ERR_HIP;
  if (rc == 0 && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, PIVX_ERR_HIP, "init synchronisation failed");
  if (rc) {
    std::string msg = c->err;
    pivx_freemem(c);
    set_global_error(msg.c_str());
    return rc;
  }
  *out = c;
  return PIVX_OK;
}

static void drop_graph(pivx_ctx* c);
static void drop_learned_graph(pivx_ctx* c);

int pivx_freemem(pivx_ctx* c) {
  if (!c) return PIVX_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* q : c->allocs) (void)hipFree(q);
  drop_graph(c);
  drop_learned_graph(c);
  for (hipEvent_t e : c->tl.pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PIVX_OK;
}

int pivx_compute_geometry(const pivx_init_params* p, int grid, pivx_geometry* g) {
  if (!p || !g || (grid != 1 && grid != 2)) return fail(nullptr, PIVX_ERR_PARAM, "bad argument");
  const int rc = validate(p);
  if (rc) return rc;
  *g = grid == 1 ? make_geometry(p->nx, p->ny, p->nsec_cc, p->ws_x1, p->ws_y1, p->spacing_x1, p->spacing_y1,
                                 p->max_dispx_percent1, p->max_dispy_percent1)
                 : make_geometry(p->nx, p->ny, p->nsec_cc, p->ws_x2, p->ws_y2, p->spacing_x2, p->spacing_y2,
                                 p->max_dispx_percent2, p->max_dispy_percent2);
  return PIVX_OK;
}

int pivx_get_geometry(const pivx_ctx* c, int grid, pivx_geometry* g) {
  if (!c || !g || (grid != 1 && grid != 2)) return fail(nullptr, PIVX_ERR_PARAM, "bad argument");
  *g = c->grid[grid - 1].g;
  return PIVX_OK;
}

// MHE of one frame already in c->mhe_in into `out` (device, row pitch out_pitch).
static int mhe_run(pivx_ctx* c, uint8_t* out, long long out_pitch) {
  MheArgs m = c->mhe;
  m.in = c->mhe_in;
  m.out = out;
  m.out_pitch = out_pitch;
  tic(c);
  const int rc = check_launch(c, launch_mhe(m, c->stream), "mhe_gpu::processsing");
  toc(c, 0, 6);
  return rc;
}

static int upload(pivx_ctx* c, const uint8_t* a, const uint8_t* b, size_t pitch, hipMemcpyKind kind,
                  bool wait = true) {
  if (!c || !a || !b) return fail(c, PIVX_ERR_PARAM, "null frame");
  if (pitch < (size_t)c->p.nx) return fail(c, PIVX_ERR_PARAM, "pitch %zu < nx %d", pitch, c->p.nx);
  HIPCHK(c, hipSetDevice(c->device));
  if (c->mhe_on) {
    // main.cpp:267-283: each frame is enhanced before prep_img_data
    int rc;
    HIPCHK(c, hipMemcpy2DAsync(c->mhe_in, c->p.nx, a, pitch, c->p.nx, c->p.ny, kind, c->stream));
    if ((rc = mhe_run(c, c->f1, c->p.nx))) return rc;
    HIPCHK(c, hipMemcpy2DAsync(c->mhe_in, c->p.nx, b, pitch, c->p.nx, c->p.ny, kind, c->stream));
    if ((rc = mhe_run(c, c->f2, c->pitch2))) return rc;
  } else {
    HIPCHK(c, hipMemcpy2DAsync(c->f1, c->p.nx, a, pitch, c->p.nx, c->p.ny, kind, c->stream));
    HIPCHK(c, hipMemcpy2DAsync(c->f2, c->pitch2, b, pitch, c->p.nx, c->p.ny, kind, c->stream));
  }
  if (kind == hipMemcpyHostToDevice && wait) HIPCHK(c, hipStreamSynchronize(c->stream));  // caller may reuse its buffers
  c->have_frames = true;
  return PIVX_OK;
}

int pivx_prep_img_data(pivx_ctx* c, const uint8_t* img1, const uint8_t* img2, size_t pitch) {
  return upload(c, img1, img2, pitch, hipMemcpyHostToDevice);
}

int pivx_prep_img_data_device(pivx_ctx* c, const uint8_t* d1, const uint8_t* d2, size_t pitch) {
  return upload(c, d1, d2, pitch, hipMemcpyDeviceToDevice);
}

int pivx_prep_img_data_async(pivx_ctx* c, const uint8_t* img1, const uint8_t* img2, size_t pitch) {
  return upload(c, img1, img2, pitch, hipMemcpyHostToDevice, false);
}

void* pivx_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void pivx_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

// mhe_gpu::init (mhe:149-212) for the handle's frame size: tile grid, central
// region and thresholds; the buffers are allocated once.
int pivx_set_mhe(pivx_ctx* c, int enable, int ws, float thresh) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (!enable) {
    c->mhe_on = false;
    return PIVX_OK;
  }
  const int nx = c->p.nx, ny = c->p.ny;
  if (ws < 1 || ws > nx || ws > ny)
    return fail(c, PIVX_ERR_PARAM, "MHE window %d must lie in [1, min(nx, ny)] = [1, %d]", ws, std::min(nx, ny));
  const int ntrash = (int)roundf((float)(ws * ws) * thresh);
  const int nnontrash = ws * ws - ntrash;
  if (nnontrash < 1)
    return fail(c, PIVX_ERR_PARAM, "MHE threshold %g leaves no counted pixels in a %dx%d tile", thresh, ws, ws);
  MheArgs m{};
  m.nx = nx;
  m.ny = ny;
  m.ws = ws;
  if (nx % ws == 0 && ny % ws == 0) {
    m.xs = 0;
    m.ys = 0;
  } else {
    m.xs = (nx - ws * (nx / ws)) / 2;
    m.ys = (ny - ws * (ny / ws)) / 2;
  }
  m.tiles_x = nx / ws;
  m.tiles_y = ny / ws;
  m.ntrash = ntrash;
  m.nnontrash = nnontrash;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nh = (size_t)(m.tiles_x + 2) * (m.tiles_y + 2) * 256;
  if (c->mhe_ready && c->mhe.tiles_x == m.tiles_x && c->mhe.tiles_y == m.tiles_y) {
    m.hist = c->mhe.hist;
  } else {
    int rc = dalloc(c, &m.hist, nh);
    if (!rc && !c->mhe_in) rc = dalloc(c, &c->mhe_in, (size_t)nx * ny);
    if (!rc && !c->mhe_out) rc = dalloc(c, &c->mhe_out, (size_t)nx * ny);
    if (rc) return rc;
    // the border ring of the histogram array stays 0 (d_levels' zero frame, mhe:216)
    HIPCHK(c, hipMemsetAsync(m.hist, 0, nh * sizeof(uint32_t), c->stream));
  }
  c->mhe = m;
  c->mhe_ready = true;
  c->mhe_on = true;
  return PIVX_OK;
}

// mhe_gpu::processsing + d_enhanced download (mhe:214-253) on one host frame.
int pivx_mhe_enhance(pivx_ctx* c, const uint8_t* img, uint8_t* out, size_t pitch) {
  if (!c || !img || !out) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (!c->mhe_ready) return fail(c, PIVX_ERR_STATE, "MHE not configured: call pivx_set_mhe first");
  if (pitch < (size_t)c->p.nx) return fail(c, PIVX_ERR_PARAM, "pitch %zu < nx %d", pitch, c->p.nx);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpy2DAsync(c->mhe_in, c->p.nx, img, pitch, c->p.nx, c->p.ny, hipMemcpyHostToDevice, c->stream));
  int rc = mhe_run(c, c->mhe_out, c->p.nx);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy2DAsync(out, pitch, c->mhe_out, c->p.nx, c->p.nx, c->p.ny, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PIVX_OK;
}

int pivx_calc_cc_map_p1(pivx_ctx* c, int idx_pass) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (!c->have_frames) return fail(c, PIVX_ERR_STATE, "no frames: call pivx_prep_img_data first");
  if (idx_pass < 0) return fail(c, PIVX_ERR_PARAM, "idx_pass must be >= 0");
  int rc;
  if (idx_pass == 0) return run_xcorr(c, 0, false, true);
  // the deformation of non-overlapping windows runs in the correlation
  // kernel (within pivx_run_pair by default, pivx_set_fused_deform); so does
  // that of the learned pipeline's gated fallback, whatever the overlap: it
  // saves the fallback a launch per pass that mostly does nothing
  const GridBuf& G1 = c->grid[0];
  if (c->prec == 0 && ((G1.fuse && (c->fuse_mode == 2 || (c->fuse_mode == 1 && c->in_pair))) ||
                       (c->gate && c->in_pair && xcorr_fuses_deform(G1.g.ws_x, G1.g.ws_y))))
    return run_xcorr(c, 0, true, false, true);
  rc = run_deform(c, 0);
  if (rc) return rc;
  return run_xcorr(c, 0, true);
}

int pivx_calc_cc_map_p2(pivx_ctx* c) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (!c->have_frames) return fail(c, PIVX_ERR_STATE, "no frames: call pivx_prep_img_data first");
  int rc = run_deform(c, 1);
  if (rc) return rc;
  return run_xcorr(c, 1, true);
}

int pivx_median_filter_smoothing_p1(pivx_ctx* c, int npass_median, int min_nneigh, float residual_thr,
                                    float insert_thr, int npass_smooth) {
  (void)residual_thr;  // read but never used by the reference
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  GridBuf& G = c->grid[0];
  return run_filter(c, 0, npass_median, min_nneigh, insert_thr, npass_smooth, 1, G.fx, G.fy, G.fv);
}

int pivx_median_filter_smoothing_p2(pivx_ctx* c, int npass_median, int min_nneigh, float residual_thr,
                                    float insert_thr, int npass_smooth, int iffill) {
  (void)residual_thr;
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  GridBuf& G = c->grid[1];
  return run_filter(c, 1, npass_median, min_nneigh, insert_thr, npass_smooth, iffill, G.fx, G.fy, G.fv);
}

int pivx_median_filter_smoothing_noalloc(const float* d_cand_x, const float* d_cand_y, const uint8_t* d_cand_valid,
                                         int nx_vec, int ny_vec, int npeak_save, int wx_median, int wy_median,
                                         int npass_median, int min_nneigh, float residual_thr, float insert_thr,
                                         int wx_smooth, int wy_smooth, int npass_smooth, int iffill, float* d_out_x,
                                         float* d_out_y, uint8_t* d_out_valid, void* stream) {
  (void)residual_thr;  // read but never used by the reference
  if (!d_cand_x || !d_cand_y || !d_cand_valid || !d_out_x || !d_out_y || !d_out_valid)
    return fail(nullptr, PIVX_ERR_PARAM, "null buffer");
  if (nx_vec < 1 || ny_vec < 1 || npeak_save < 1) return fail(nullptr, PIVX_ERR_PARAM, "bad field size");
  const long long nt = (long long)nx_vec * ny_vec;
  if (nt > (long long)PIVX_NOALLOC_MAX_VECTORS)
    return fail(nullptr, PIVX_ERR_PARAM, "field of %lld vectors exceeds the allocation-free limit of %d vectors", nt,
                PIVX_NOALLOC_MAX_VECTORS);
  if (wx_median < 1 || wy_median < 1 || wx_smooth < 1 || wy_smooth < 1)
    return fail(nullptr, PIVX_ERR_PARAM, "median/smoothing window sizes must be >= 1");
  if (npass_median > 0 && wx_median * wy_median > 64)
    return fail(nullptr, PIVX_ERR_PARAM,
                "median window %dx%d exceeds the allocation-free limit of 64 vectors (larger windows need the "
                "per-vector lists of pivx_median_filter_smoothing_p2)", wx_median, wy_median);
  int dev = 0;
  HIPCHK(nullptr, hipGetDevice(&dev));
  unsigned char* base = nullptr;
  int rc = scratch_of_device(dev, &base);
  if (rc) return rc;
  FilterJob J{};
  J.cx = d_cand_x;
  J.cy = d_cand_y;
  J.cv = d_cand_valid;
  J.K = npeak_save;
  J.nx_vec = nx_vec;
  J.ny_vec = ny_vec;
  J.wxm = wx_median;
  J.wym = wy_median;
  J.wxs = wx_smooth;
  J.wys = wy_smooth;
  // counters first, then the planes (each a multiple of 256 bytes apart)
  J.nholes = reinterpret_cast<int*>(base);
  J.drops = reinterpret_cast<unsigned int*>(base + 64);
  J.sweeps = reinterpret_cast<int*>(base + 128);
  float* f = reinterpret_cast<float*>(base + 256);
  J.sx = f;
  J.sy = f + kNaN;
  J.ux = f + 2 * kNaN;
  J.uy = f + 3 * kNaN;
  J.tmp_x = f + 4 * kNaN;
  J.tmp_y = f + 5 * kNaN;
  J.list_a = reinterpret_cast<int*>(f + 6 * kNaN);
  J.list_b = reinterpret_cast<int*>(f + 7 * kNaN);
  uint8_t* b8 = reinterpret_cast<uint8_t*>(f + 8 * kNaN);
  J.sv = b8;
  J.uv = b8 + kNaN;
  J.tmp_f = b8 + 2 * kNaN;
  J.list_cap = (int)kNaN;
  J.med_scratch = nullptr;
  J.accepted = nullptr;
  J.stream = static_cast<hipStream_t>(stream);
  return enqueue_filter(nullptr, 1, J, npass_median, min_nneigh, insert_thr, npass_smooth, iffill, d_out_x, d_out_y,
                        d_out_valid, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int pivx_extend_filtered_vecfield_p1(pivx_ctx* c) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  return extend(c, 0);
}

int pivx_extend_filtered_vecfield_p2(pivx_ctx* c) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  return extend(c, 1);
}

static int upsample(pivx_ctx* c, bool write_ext) {
  GridBuf& G1 = c->grid[0];
  GridBuf& G2 = c->grid[1];
  const pivx_geometry& g1 = G1.g;
  const pivx_geometry& g2 = G2.g;
  return check_launch(
      c,
      launch_upsample(G2.fx, G2.fy, G1.ex, G1.ey, g1.nx_vec + 2 * kExBorder, g1.ny_vec + 2 * kExBorder,
                      (float)(g1.spacing_x / 2 - kExBorder * g1.spacing_x),
                      (float)(g1.spacing_y / 2 - kExBorder * g1.spacing_y), (float)g1.spacing_x, (float)g1.spacing_y,
                      g2.spacing_x, g2.spacing_y, g2.nx_vec, g2.ny_vec, write_ext ? G2.ex : nullptr,
                      write_ext ? G2.ey : nullptr, kExBorder, c->stream),
      "Upsampling_vecfield_kernel_2d");
}

int pivx_upsampling_vector_field(pivx_ctx* c) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  return upsample(c, false);
}

// The per-pair sequence of main.cpp:288-375.  The extension of a field is
// written by the kernel that produces the field (last smoothing pass,
// upsampling) whenever it can be; extend() runs only where it could not.
struct InPair {
  pivx_ctx* c;
  explicit InPair(pivx_ctx* x) : c(x) { c->in_pair = true; }
  ~InPair() { c->in_pair = false; }
};

// Grid-1 passes of main.cpp:289-303; *ext_ok tells whether the last
// smoothing pass wrote the extended field.
static int enqueue_grid1(pivx_ctx* c, const pivx_filter_params* f1, bool* ext_ok) {
  int rc;
  GridBuf& G1 = c->grid[0];
  const pivx_geometry& g1 = G1.g;
  *ext_ok = false;
  // The learned pipeline's gated fallback: each pass is one launch, the
  // correlation whose last workgroup runs the pass's filter sequence (the
  // deformation inside the correlation), where the shapes allow it.
  const bool one_launch = c->gate && c->in_pair && c->prec == 0 && c->lgate_work &&
                          xcorr_gated_filter_supported(g1.ws_x, g1.ws_y, c->p.wx1_median, c->p.wy1_median) &&
                          xcorr_fuses_deform(g1.ws_x, g1.ws_y) &&
                          filter_fits_seq(f1->npass_median, f1->npass_smooth, c->p.wx1_median, c->p.wy1_median);
  for (int j = 0; j < c->p.npass1; ++j) {
    if (j > 0 && !*ext_ok && (rc = extend(c, 0))) return rc;
    if (one_launch) {
      FilterSeq q{};
      if ((rc = run_filter(c, 0, f1->npass_median, f1->min_nneigh, f1->insert_thr, f1->npass_smooth, 1, G1.fx, G1.fy,
                           G1.fv, G1.ex, G1.ey, ext_ok, nullptr, nullptr, nullptr, &q)))
        return rc;
      const XcArgs a = make_xcargs(c, 0, j > 0, j == 0, j > 0);
      rc = check_launch(c, launch_xcorr_gated_filter(a, g1.ws_x, g1.ws_y, q,
                                                     reinterpret_cast<unsigned int*>(c->lgate_work + 2), c->stream),
                        "gated correlation and filter");
      if (rc) return rc;
      continue;
    }
    if ((rc = pivx_calc_cc_map_p1(c, j))) return rc;
    if ((rc = run_filter(c, 0, f1->npass_median, f1->min_nneigh, f1->insert_thr, f1->npass_smooth, 1, G1.fx, G1.fy,
                         G1.fv, G1.ex, G1.ey, ext_ok)))
      return rc;
  }
  return PIVX_OK;
}

// Upsampling and grid-2 passes of main.cpp:305-375.
static int enqueue_grid2(pivx_ctx* c, const pivx_filter_params* f2, bool ext_ok) {
  int rc;
  GridBuf& G2 = c->grid[1];
  if (c->p.npass2 > 0) {
    if (!ext_ok && (rc = extend(c, 0))) return rc;
    if ((rc = upsample(c, true))) return rc;
    ext_ok = true;
    // The median-only re-run of main.cpp:363-375 filters the last raw
    // candidates with the same median settings and stops after the median
    // passes: its field is the last pass's field before filling and
    // smoothing, which that pass's last median launch also writes to mo.
    for (int j = 0; j < c->p.npass2; ++j) {
      if (!ext_ok && (rc = extend(c, 1))) return rc;
      if ((rc = pivx_calc_cc_map_p2(c))) return rc;
      const bool last = j == c->p.npass2 - 1;
      if ((rc = run_filter(c, 1, f2->npass_median, f2->min_nneigh, f2->insert_thr, f2->npass_smooth, 1, G2.fx, G2.fy,
                           G2.fv, G2.ex, G2.ey, &ext_ok, last ? c->mo_x : nullptr, last ? c->mo_y : nullptr,
                           last ? c->mo_v : nullptr)))
        return rc;
    }
  }
  return PIVX_OK;
}

static int run_pair_enqueue(pivx_ctx* c, const pivx_filter_params* f1, const pivx_filter_params* f2) {
  InPair in_pair(c);
  bool ext_ok = false;
  const int rc = enqueue_grid1(c, f1, &ext_ok);
  if (rc) return rc;
  return enqueue_grid2(c, f2, ext_ok);
}

static void drop_graph(pivx_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

static bool same_filter(const pivx_filter_params& a, const pivx_filter_params& b) {
  return a.npass_median == b.npass_median && a.min_nneigh == b.min_nneigh && a.insert_thr == b.insert_thr &&
         a.npass_smooth == b.npass_smooth;
}

int pivx_run_pair(pivx_ctx* c, const pivx_filter_params* f1, const pivx_filter_params* f2) {
  if (!c || !f1 || !f2) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (!c->have_frames) return fail(c, PIVX_ERR_STATE, "no frames: call pivx_prep_img_data first");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->use_graph) return run_pair_enqueue(c, f1, f2);
  if (!c->gexec || !same_filter(c->gf[0], *f1) || !same_filter(c->gf[1], *f2)) {
    drop_graph(c);
    // stage events are not captured: a replay is timed as a whole (stage 5)
    const bool timing = c->timing;
    c->timing = false;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = run_pair_enqueue(c, f1, f2);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = timing;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    c->graph = g;
    e = hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(c);
      return fail(c, PIVX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->gf[0] = *f1;
    c->gf[1] = *f2;
  }
  tic(c);
  HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
  toc(c, 0, 5);
  return PIVX_OK;
}

int pivx_set_graph_mode(pivx_ctx* c, int on) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  c->use_graph = on != 0;
  if (!c->use_graph) {
    drop_graph(c);
    drop_learned_graph(c);
  }
  return PIVX_OK;
}

int pivx_set_fused_deform(pivx_ctx* c, int mode) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (mode < 0 || mode > 2) return fail(c, PIVX_ERR_PARAM, "fused deformation mode must be 0, 1 or 2 (got %d)", mode);
  if (mode != c->fuse_mode) {  // a captured pair holds the old choice
    drop_graph(c);
    drop_learned_graph(c);
  }
  c->fuse_mode = mode;
  return PIVX_OK;
}

// Moves the device's dropped-hole count into drops_total; a non-zero count is
// reported once, as an error of the call that finds it.
static int check_drops(pivx_ctx* c) {
  unsigned int d = 0;
  HIPCHK(c, hipMemcpyAsync(&d, c->drops, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (d == 0) return PIVX_OK;
  const unsigned int zero = 0;
  HIPCHK(c, hipMemcpyAsync(c->drops, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->drops_total += d;
  return fail(c, PIVX_ERR_STATE, "hole list overflow: %u invalid vectors were not filled", d);
}

// Host tables of the DFT-by-matrix precision form for a W x H window: the
// matrices k_xcorr_dft contracts with, each entry split into hi = round(x)
// and lo = round(x - hi) in FP16 or BF16.
static uint16_t host_f16(float f) {  // round to nearest even, with subnormals
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const float a = std::fabs(f);
  if (!(a < 65520.f)) return (uint16_t)(sign | 0x7c00u);  // overflow (inputs never reach it)
  if (a < 6.103515625e-05f) {                              // subnormal: multiples of 2^-24
    const float q = std::nearbyint(a * 16777216.f);
    return (uint16_t)(sign | (uint32_t)q);
  }
  uint32_t v = u & 0x7fffffffu;
  const uint32_t lsb = (v >> 13) & 1u;
  v += 0xfffu + lsb;
  const uint32_t e = (v >> 23) - 127 + 15, m = (v >> 13) & 0x3ffu;
  return (uint16_t)(sign | (e << 10) | m);
}
static float host_f16_back(uint16_t h) {
  const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0) v = (float)m * 5.9604644775390625e-08f;
  else {
    const uint32_t u = ((e - 15 + 127) << 23) | (m << 13);
    std::memcpy(&v, &u, 4);
  }
  uint32_t r;
  std::memcpy(&r, &v, 4);
  r |= sign;
  std::memcpy(&v, &r, 4);
  return v;
}

static int build_dft_tables(pivx_ctx* c, int gi, bool bf) {
  const pivx_geometry& g = c->grid[gi].g;
  const int W = g.ws_x, H = g.ws_y, PW = (W + 15) / 16 * 16, PH = (H + 15) / 16 * 16;
  auto cs = [](int n, int j, int k, bool sine) {
    const double ang = 2.0 * M_PI * (double)(((long long)j * k) % n) / (double)n;
    return sine ? std::sin(ang) : std::cos(ang);
  };
  // A tables [PH][2PH]: (sa, ca) signs of the S and C halves: value(k, kk) = first half or second half
  struct Spec { int n, pad; bool rowmajor; int first, second; };  // 1 = C, 2 = S, negative = negated, 0 = none
  const Spec specs[7] = {{H, PH, true, 1, 2},  {H, PH, true, -2, 1},   // a1: [C S], [-S C]
                         {W, PW, false, 1, 2}, {W, PW, false, -2, 1},  // b2: [C; S], [-S; C]
                         {H, PH, true, 1, -2}, {H, PH, true, 2, 1},    // a3: [C -S], [S C]
                         {W, PW, false, 1, -2}};                       // b4: [C; -S]
  size_t total = 0;
  for (const Spec& sp : specs) total += (size_t)sp.pad * 2 * sp.pad * 2;
  std::vector<uint16_t> buf(total, 0);
  size_t off = 0;
  const uint16_t* base = nullptr;
  uint16_t* dev = nullptr;
  if (dalloc(c, &dev, total)) return PIVX_ERR_HIP;
  base = dev;
  const uint16_t* ptr[7][2];
  for (int t = 0; t < 7; ++t) {
    const Spec& sp = specs[t];
    const int K = 2 * sp.pad;
    uint16_t* hi = buf.data() + off;
    uint16_t* lo = hi + (size_t)sp.pad * K;
    for (int row = 0; row < sp.n; ++row)
      for (int kk = 0; kk < K; ++kk) {
        const int half = kk < sp.pad ? 0 : 1, idx = kk - half * sp.pad;
        if (idx >= sp.n) continue;
        const int which = half == 0 ? sp.first : sp.second;
        // row-major tables hold M[row][idx]; column tables hold column `row` of M, entry idx: M[idx][row] (C, S symmetric)
        double v = cs(sp.n, row, idx, std::abs(which) == 2);
        if (which < 0) v = -v;
        const float f = (float)v;
        uint16_t h, l;
        if (bf) {
          h = host_bf16(f);
          float hb;
          const uint32_t hu = (uint32_t)h << 16;
          std::memcpy(&hb, &hu, 4);
          l = host_bf16(f - hb);
        } else {
          h = host_f16(f);
          l = host_f16(f - host_f16_back(h));
        }
        hi[(size_t)row * K + kk] = h;
        lo[(size_t)row * K + kk] = l;
      }
    ptr[t][0] = base + off;
    ptr[t][1] = base + off + (size_t)sp.pad * K;
    off += (size_t)sp.pad * K * 2;
  }
  HIPCHK(c, hipMemcpy(dev, buf.data(), total * sizeof(uint16_t), hipMemcpyHostToDevice));
  DftTables& T = c->dtab[gi];
  T.lane32 = nullptr;
  if (W == 32 && H == 32) {
    // C and S in the MFMA lane layout: [m][t][hl][lane][j] = M[16t + lane % 16][8 (lane / 16) + j]
    std::vector<uint16_t> lt(2 * 2 * 2 * 64 * 8);
    for (int m = 0; m < 2; ++m)
      for (int t = 0; t < 2; ++t)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int row = 16 * t + (lane & 15), k = 8 * (lane >> 4) + j;
            const float f = (float)cs(32, row, k, m == 1);
            uint16_t h, l;
            if (bf) {
              h = host_bf16(f);
              float hb;
              const uint32_t hu = (uint32_t)h << 16;
              std::memcpy(&hb, &hu, 4);
              l = host_bf16(f - hb);
            } else {
              h = host_f16(f);
              l = host_f16(f - host_f16_back(h));
            }
            lt[(((size_t)(m * 2 + t) * 2 + 0) * 64 + lane) * 8 + j] = h;
            lt[(((size_t)(m * 2 + t) * 2 + 1) * 64 + lane) * 8 + j] = l;
          }
    uint16_t* ldev = nullptr;
    if (dalloc(c, &ldev, lt.size())) return PIVX_ERR_HIP;
    HIPCHK(c, hipMemcpy(ldev, lt.data(), lt.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    T.lane32 = ldev;
  }
  T.lane64 = nullptr;
  if (W == 64 && H == 64) {
    std::vector<uint16_t> lt(2 * 2 * 4 * 2 * 64 * 8);
    for (int m = 0; m < 2; ++m)
      for (int t = 0; t < 4; ++t)
        for (int ks = 0; ks < 2; ++ks)
          for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < 8; ++j) {
              const int row = 16 * t + (lane & 15), k = 32 * ks + 8 * (lane >> 4) + j;
              const float f = (float)cs(64, row, k, m == 1);
              uint16_t h, l;
              if (bf) {
                h = host_bf16(f);
                float hb;
                const uint32_t hu = (uint32_t)h << 16;
                std::memcpy(&hb, &hu, 4);
                l = host_bf16(f - hb);
              } else {
                h = host_f16(f);
                l = host_f16(f - host_f16_back(h));
              }
              for (int hl = 0; hl < 2; ++hl)
                lt[((((size_t)(m * 2 + hl) * 4 + t) * 2 + ks) * 64 + lane) * 8 + j] = hl ? l : h;
            }
    uint16_t* ldev = nullptr;
    if (dalloc(c, &ldev, lt.size())) return PIVX_ERR_HIP;
    HIPCHK(c, hipMemcpy(ldev, lt.data(), lt.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    T.lane64 = ldev;
  }
  for (int o = 0; o < 2; ++o) {
    T.a1h[o] = ptr[o][0];
    T.a1l[o] = ptr[o][1];
    T.b2h[o] = ptr[2 + o][0];
    T.b2l[o] = ptr[2 + o][1];
    T.a3h[o] = ptr[4 + o][0];
    T.a3l[o] = ptr[4 + o][1];
  }
  T.b4h = ptr[6][0];
  T.b4l = ptr[6][1];
  return PIVX_OK;
}

int pivx_set_correlation_precision(pivx_ctx* c, int mode) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (mode < 0 || mode > 4) return fail(c, PIVX_ERR_PARAM, "precision mode must be 0..4 (got %d)", mode);
  if (mode == c->prec) return PIVX_OK;
  HIPCHK(c, hipSetDevice(c->device));
  if (mode != 0) {
    for (int gi = 0; gi < 2; ++gi) {
      const pivx_geometry& g = c->grid[gi].g;
      if (!xcorr_dft_supported(g.ws_x, g.ws_y))
        return fail(c, PIVX_ERR_PARAM,
                    "the DFT-by-matrix precision form is compiled for windows 16x16, 24x24, 32x32, 48x48, 64x64, "
                    "64x32, 32x64, 32x16, 16x32 and 48x24 (grid %d: %dx%d)", gi + 1, g.ws_x, g.ws_y);
    }
    for (int gi = 0; gi < 2; ++gi) {
      const int rc = build_dft_tables(c, gi, mode == 2 || mode == 4);
      if (rc) return rc;
    }
  }
  drop_graph(c);
  drop_learned_graph(c);
  c->prec = mode;
  return PIVX_OK;
}

int pivx_synchronize(pivx_ctx* c) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_drops(c);
}

int pivx_get_dims(const pivx_ctx* c, int grid, int* nx_vec, int* ny_vec) {
  if (!c || (grid != 1 && grid != 2)) return fail(nullptr, PIVX_ERR_PARAM, "bad argument");
  if (nx_vec) *nx_vec = c->grid[grid - 1].g.nx_vec;
  if (ny_vec) *ny_vec = c->grid[grid - 1].g.ny_vec;
  return PIVX_OK;
}

int pivx_get_grid_coords(const pivx_ctx* c, int grid, float* x, float* y) {
  if (!c || (grid != 1 && grid != 2)) return fail(nullptr, PIVX_ERR_PARAM, "bad argument");
  const pivx_geometry& g = c->grid[grid - 1].g;
  for (int j = 0; j < g.ny_vec; ++j)
    for (int k = 0; k < g.nx_vec; ++k) {
      if (x) x[j * g.nx_vec + k] = (float)(g.spacing_x / 2 + k * g.spacing_x);
      if (y) y[j * g.nx_vec + k] = (float)(g.spacing_y / 2 + j * g.spacing_y);
    }
  return PIVX_OK;
}

static int d2h(pivx_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!dst) return PIVX_OK;
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return PIVX_OK;
}
static int h2d(pivx_ctx* c, void* dst, const void* src, size_t bytes) {
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return PIVX_OK;
}

int pivx_get_raw(pivx_ctx* c, int grid, float* dx, float* dy, uint8_t* valid) {
  GridBuf* G = grid_of(c, grid);
  if (!G) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const size_t n = (size_t)G->g.nt_vec * c->p.npeak_save;
  int rc = d2h(c, dx, G->dx, n * 4) | d2h(c, dy, G->dy, n * 4) | d2h(c, valid, G->dv, n);
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_get_filtered(pivx_ctx* c, int grid, float* dx, float* dy, uint8_t* valid) {
  if (c && grid == 3) {  // grid-2 median-only field of pivx_run_pair
    const size_t n = (size_t)c->grid[1].g.nt_vec;
    int rc = d2h(c, dx, c->mo_x, n * 4) | d2h(c, dy, c->mo_y, n * 4) | d2h(c, valid, c->mo_v, n);
    if (rc) return rc;
    return pivx_synchronize(c);
  }
  GridBuf* G = grid_of(c, grid);
  if (!G) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const size_t n = (size_t)G->g.nt_vec;
  int rc = d2h(c, dx, G->fx, n * 4) | d2h(c, dy, G->fy, n * 4) | d2h(c, valid, G->fv, n);
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_get_peaks(pivx_ctx* c, int grid, float* value, int* index, int* ncand) {
  GridBuf* G = grid_of(c, grid);
  if (!G) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const size_t n = (size_t)G->g.nt_vec * c->p.npeak_save;
  int rc = d2h(c, value, G->pkv, n * 4) | d2h(c, index, G->pki, n * 4) | d2h(c, ncand, G->ncand, G->g.nt_vec * 4);
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_get_deformed(pivx_ctx* c, int grid, float* out) {
  GridBuf* G = grid_of(c, grid);
  if (!G || !out) return fail(c, PIVX_ERR_PARAM, "bad argument");
  int rc = d2h(c, out, G->deform, (size_t)G->g.nx_aug * G->g.ny_aug * 4);
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_get_stats(pivx_ctx* c, pivx_stats* st) {
  if (!c || !st) return fail(c, PIVX_ERR_PARAM, "bad argument");
  unsigned long long h[2];
  int sw[2];
  int pending = 0;
  int rc = d2h(c, h, c->counters, sizeof(h)) | d2h(c, sw, c->sweeps, sizeof(sw)) |
           d2h(c, &pending, c->nholes, sizeof(pending));
  if (rc) return rc;
  std::vector<int> nc[2];
  std::vector<int8_t> acc[2];
  for (int g = 0; g < 2; ++g) {
    const size_t nt = (size_t)c->grid[g].g.nt_vec;
    nc[g].resize(nt);
    acc[g].resize(nt);
    rc = d2h(c, nc[g].data(), c->grid[g].ncand, nt * sizeof(int)) | d2h(c, acc[g].data(), c->grid[g].accepted, nt);
    if (rc) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)check_drops(c);  // counted into drops_total; the stats carry it instead of an error
  st->hole_list_drops = c->drops_total;
  st->holes_pending = pending;
  for (int g = 0; g < 2; ++g) {
    const int cap = c->grid[g].g.npeak_slots;
    long long ov = 0, out = 0;
    for (int v : nc[g]) ov += v > cap;
    for (int8_t a : acc[g]) out += a != 0;
    st->overflow_windows[g] = ov;
    st->outliers[g] = out;
    st->oob_taps[g] = (long long)h[g];
    st->fill_sweeps[g] = sw[g];
  }
  return PIVX_OK;
}

int pivx_set_filtered(pivx_ctx* c, int grid, const float* dx, const float* dy, const uint8_t* valid) {
  GridBuf* G = grid_of(c, grid);
  if (!G || !dx || !dy) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const size_t n = (size_t)G->g.nt_vec;
  int rc = h2d(c, G->fx, dx, n * 4) | h2d(c, G->fy, dy, n * 4);
  if (rc) return rc;
  if (valid) rc = h2d(c, G->fv, valid, n);
  else HIPCHK(c, hipMemsetAsync(G->fv, 1, n, c->stream));
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_set_raw(pivx_ctx* c, int grid, const float* dx, const float* dy, const uint8_t* valid) {
  GridBuf* G = grid_of(c, grid);
  if (!G || !dx || !dy) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const size_t n = (size_t)G->g.nt_vec * c->p.npeak_save;
  int rc = h2d(c, G->dx, dx, n * 4) | h2d(c, G->dy, dy, n * 4);
  if (rc) return rc;
  if (valid) rc = h2d(c, G->dv, valid, n);
  else HIPCHK(c, hipMemsetAsync(G->dv, 1, n, c->stream));
  if (rc) return rc;
  return pivx_synchronize(c);
}

int pivx_enable_timing(pivx_ctx* c, int enable) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  c->timing = enable != 0;
  c->tl.used = 0;
  c->tl.start = -1;
  c->tl.recs.clear();
  return PIVX_OK;
}

int pivx_get_stage_ms(pivx_ctx* c, int grid, int stage, double* ms, int* launches) {
  if (!c || (grid != 1 && grid != 2) || stage < 0 || stage > 6) return fail(c, PIVX_ERR_PARAM, "bad argument");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double sum = 0.0;
  int n = 0;
  const pivx_ctx::Timeline& tl = c->tl;
  for (const auto& r : tl.recs) {
    if (r.gi != grid - 1 || r.stage != stage) continue;
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, tl.pool[r.a], tl.pool[r.b]));
    sum += t;
    ++n;
  }
  if (ms) *ms = n ? sum / n : 0.0;
  if (launches) *launches = n;
  return PIVX_OK;
}

int pivx_bandwidth_probe(pivx_ctx* c, size_t bytes, int reps, double* gbps) {
  if (!c || bytes < 16 || reps < 1 || !gbps) return fail(c, PIVX_ERR_PARAM, "bad argument");
  bytes &= ~(size_t)15;
  HIPCHK(c, hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  HIPCHK(c, hipMalloc(&a, bytes));
  hipError_t e = hipMalloc(&b, bytes);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(c, PIVX_ERR_HIP, "hipMalloc: %s", hipGetErrorString(e));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, bytes, c->stream);
  if (e == hipSuccess) e = launch_stream_copy(a, b, bytes, c->stream);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stream_copy(a, b, bytes, c->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "bandwidth probe: %s", hipGetErrorString(e));
  *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  return PIVX_OK;
}

// Host snapshot of everything write_vec_data_general reads (cu:1865-1973),
// in page-locked memory so that the fetch is one asynchronous copy train on
// the handle's stream, and the files can be written by host threads while
// the GPU runs the next pair.
struct pivx_results {
  int device = 0;
  int K = 0;
  pivx_geometry g[2]{};
  std::vector<float> gx[2], gy[2];     // grid coordinates (fixed per handle)
  float* raw_x[2] = {nullptr, nullptr};  // [nt][K]
  float* raw_y[2] = {nullptr, nullptr};
  float* fil_x[3] = {nullptr, nullptr, nullptr};  // grid 1, grid 2, grid-2 median-only
  float* fil_y[3] = {nullptr, nullptr, nullptr};
  hipEvent_t done = nullptr;
  bool pending = false;
};

int pivx_results_create(pivx_ctx* c, pivx_results** out) {
  if (!c || !out) return fail(c, PIVX_ERR_PARAM, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  pivx_results* r = new pivx_results();
  r->device = c->device;
  r->K = c->p.npeak_save;
  bool ok = true;
  for (int gi = 0; gi < 2; ++gi) {
    r->g[gi] = c->grid[gi].g;
    const size_t nt = (size_t)r->g[gi].nt_vec;
    r->gx[gi].resize(nt);
    r->gy[gi].resize(nt);
    pivx_get_grid_coords(c, gi + 1, r->gx[gi].data(), r->gy[gi].data());
    ok = ok && hipHostMalloc((void**)&r->raw_x[gi], nt * r->K * 4 + 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&r->raw_y[gi], nt * r->K * 4 + 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&r->fil_x[gi], nt * 4 + 4, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&r->fil_y[gi], nt * 4 + 4, hipHostMallocDefault) == hipSuccess;
  }
  const size_t nt2 = (size_t)r->g[1].nt_vec;
  ok = ok && hipHostMalloc((void**)&r->fil_x[2], nt2 * 4 + 4, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&r->fil_y[2], nt2 * 4 + 4, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&r->done, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    pivx_results_destroy(r);
    return fail(c, PIVX_ERR_HIP, "page-locked result buffers could not be allocated");
  }
  *out = r;
  return PIVX_OK;
}

int pivx_results_fetch(pivx_ctx* c, pivx_results* r) {
  if (!c || !r) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (r->device != c->device || r->K != c->p.npeak_save || r->g[0].nt_vec != c->grid[0].g.nt_vec ||
      r->g[1].nt_vec != c->grid[1].g.nt_vec)
    return fail(c, PIVX_ERR_PARAM, "result snapshot was created for another handle");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = 0;
  for (int gi = 0; gi < 2; ++gi) {
    const GridBuf& G = c->grid[gi];
    const size_t nt = (size_t)G.g.nt_vec;
    rc |= d2h(c, r->raw_x[gi], G.dx, nt * r->K * 4);
    rc |= d2h(c, r->raw_y[gi], G.dy, nt * r->K * 4);
    rc |= d2h(c, r->fil_x[gi], G.fx, nt * 4);
    rc |= d2h(c, r->fil_y[gi], G.fy, nt * 4);
  }
  const size_t nt2 = (size_t)c->grid[1].g.nt_vec;
  rc |= d2h(c, r->fil_x[2], c->mo_x, nt2 * 4);
  rc |= d2h(c, r->fil_y[2], c->mo_y, nt2 * 4);
  if (rc) return rc;
  HIPCHK(c, hipEventRecord(r->done, c->stream));
  r->pending = true;
  return PIVX_OK;
}

int pivx_results_wait(pivx_results* r) {
  if (!r) return fail(nullptr, PIVX_ERR_PARAM, "null argument");
  if (r->pending) {
    const hipError_t e = hipEventSynchronize(r->done);
    if (e != hipSuccess) return fail(nullptr, PIVX_ERR_HIP, "hipEventSynchronize: %s", hipGetErrorString(e));
    r->pending = false;
  }
  return PIVX_OK;
}

int pivx_results_write(const pivx_results* r, const char* path, int selection, int crlf) {
  if (!r || !path) return fail(nullptr, PIVX_ERR_PARAM, "null argument");
  if (r->pending) return fail(nullptr, PIVX_ERR_STATE, "results not ready: call pivx_results_wait first");
  const int gi = (selection == 0 || selection == 1) ? 0 : 1;
  const pivx_geometry& g = r->g[gi];
  int rc;
  if (selection == 0 || selection == 2) {
    rc = write_tecplot(path, g.nx_vec, g.ny_vec, r->gx[gi].data(), r->gy[gi].data(), r->raw_x[gi], r->raw_y[gi], r->K,
                       true, crlf != 0);
  } else if (selection == 1 || selection == 3 || selection == 4) {
    const int f = selection == 4 ? 2 : gi;
    rc = write_tecplot(path, g.nx_vec, g.ny_vec, r->gx[gi].data(), r->gy[gi].data(), r->fil_x[f], r->fil_y[f], 1,
                       false, crlf != 0);
  } else {
    return fail(nullptr, PIVX_ERR_PARAM, "selection must be 0..4");
  }
  if (rc) return fail(nullptr, PIVX_ERR_IO, "Error occurs while the vector file is being written: %s", path);
  return PIVX_OK;
}

void pivx_results_destroy(pivx_results* r) {
  if (!r) return;
  (void)hipSetDevice(r->device);
  if (r->pending) (void)hipEventSynchronize(r->done);
  for (float* p : {r->raw_x[0], r->raw_x[1], r->raw_y[0], r->raw_y[1], r->fil_x[0], r->fil_x[1], r->fil_x[2],
                   r->fil_y[0], r->fil_y[1], r->fil_y[2]})
    if (p) (void)hipHostFree(p);
  if (r->done) (void)hipEventDestroy(r->done);
  delete r;
}

int pivx_write_vec_data_general(pivx_ctx* c, const char* path, int selection, int crlf) {
  if (!c || !path) return fail(c, PIVX_ERR_PARAM, "bad argument");
  const int grid = (selection == 0 || selection == 1) ? 1 : 2;
  const pivx_geometry& g = c->grid[grid - 1].g;
  const int nt = g.nt_vec, K = c->p.npeak_save;
  std::vector<float> gx(nt), gy(nt);
  pivx_get_grid_coords(c, grid, gx.data(), gy.data());
  int rc;
  if (selection == 0 || selection == 2) {
    std::vector<float> dx((size_t)nt * K), dy((size_t)nt * K);
    if ((rc = pivx_get_raw(c, grid, dx.data(), dy.data(), nullptr))) return rc;
    rc = write_tecplot(path, g.nx_vec, g.ny_vec, gx.data(), gy.data(), dx.data(), dy.data(), K, true, crlf != 0);
  } else if (selection == 1 || selection == 3 || selection == 4) {
    std::vector<float> dx(nt), dy(nt);
    if ((rc = pivx_get_filtered(c, selection == 4 ? 3 : grid, dx.data(), dy.data(), nullptr))) return rc;
    rc = write_tecplot(path, g.nx_vec, g.ny_vec, gx.data(), gy.data(), dx.data(), dy.data(), 1, false, crlf != 0);
  } else {
    return fail(c, PIVX_ERR_PARAM, "selection must be 0..4");
  }
  if (rc) return fail(c, PIVX_ERR_IO, "Error occurs while the vector file is being written: %s", path);
  return PIVX_OK;
}

}  // extern "C"

// ====================================================================== learned grid-1 predictor
struct pivx_cnn {
  int device = 0;
  int np = 0;
  hipStream_t stream = nullptr;
  float *master = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t* wf[kCnnLayers] = {};
  uint16_t* wt[kCnnLayers] = {};
  uint16_t* wl[kCnnLayers] = {};  // lane-ordered copy for the fused inference kernels
  long long step = 0;
  // batch workspace (training and the host test entry points)
  int wb = 0, wh = 0, ww = 0;
  uint16_t* act[kCnnLayers - 1] = {};
  float* out = nullptr;
  uint16_t* g[2] = {};
  float *partial = nullptr, *partial_b = nullptr, *loss_part = nullptr, *loss = nullptr;
  size_t pstride = 0;  // floats per layer in partial (one region per layer)
  int* dindex = nullptr;
  uint16_t* xin = nullptr;    // host-entry input batch
  float* lab = nullptr;       // host-entry labels
  std::vector<void*> allocs, ws;
};

namespace {

constexpr int kCnnChunks = 256;  // weight-gradient partial sums per layer

int cnn_fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_global_error(buf);
  return PIVX_ERR_PARAM;
}

template <class T>
int nalloc(std::vector<void*>& list, T** p, size_t n) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(n * sizeof(T), 256);
  const hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    set_global_error((std::string("hipMalloc: ") + hipGetErrorString(e)).c_str());
    return PIVX_ERR_HIP;
  }
  list.push_back(q);
  *p = static_cast<T*>(q);
  return PIVX_OK;
}


// spatial size of layer l's input for a network input h x w
void layer_dims(int h, int w, int l, int& hin, int& win, int& hout, int& wout) {
  hin = h;
  win = w;
  for (int i = 0; i < l; ++i) {
    hin /= kCnnArch[i].stride;
    win /= kCnnArch[i].stride;
  }
  hout = hin / kCnnArch[l].stride;
  wout = win / kCnnArch[l].stride;
}

bool cnn_shape_ok(int h, int w) { return h >= 8 && w >= 128 && h % 8 == 0 && w % 128 == 0; }

// The frames of a pair as the network input (batch 1): layer 0 averages them
// over 4x4 blocks as it reads them, as k_downsample would.
struct FrameInput {
  const uint8_t* f1; long long p1;
  const uint8_t* f2; long long p2;
};

// Forward pass over a batch: x0 sample b at x0 + (bidx ? bidx[b] : b) * xstride;
// hidden activations into act[0..4]; output planes at out_x / out_y + b * ostride.
int cnn_forward_enqueue(const pivx_cnn* net, const uint16_t* x0, long long xstride, const int* bidx, int B, int H,
                        int W, uint16_t* const* act, float* out_x, float* out_y, long long ostride, hipStream_t s,
                        uint8_t* out_v = nullptr, float* ext_x = nullptr, float* ext_y = nullptr, int ext_b = 0,
                        const FrameInput* frames = nullptr, bool inference = false) {
  auto bias = [&](int l) { return net->master + cnn_weight_offset(l) + kCnnArch[l].cout * 9 * kCnnArch[l].cin; };
  if (inference && B == 1 && !bidx) {
    // one pair, no backward pass to follow: two launches, layers 0-2
    // (k_conv_head, from the input plane or the frames) and layers 3-5
    // (k_conv_tail), bit-identical to the layer-by-layer pass below, which
    // also keeps every layer's activations for the backward pass
    static_assert(kCnnLayers == 6 && kCnnArch[0].cin == 2 && kCnnArch[0].cout == 32 && kCnnArch[0].stride == 2 &&
                      kCnnArch[1].cout == 64 && kCnnArch[1].stride == 2 && kCnnArch[2].cout == 64 &&
                      kCnnArch[2].stride == 2 && kCnnArch[3].cout == 64 && kCnnArch[3].stride == 1 &&
                      kCnnArch[4].cout == 64 && kCnnArch[4].stride == 1 && kCnnArch[5].cout == 2 &&
                      kCnnArch[5].stride == 1,
                  "k_conv_head and k_conv_tail compute the layers of this architecture");
    ConvHeadArgs hd{};
    int hin, win;
    layer_dims(H, W, 2, hin, win, hd.ph, hd.pw);
    if (frames) {
      hd.f1 = frames->f1;
      hd.p1 = frames->p1;
      hd.f2 = frames->f2;
      hd.p2 = frames->p2;
    } else {
      hd.in = x0;
    }
    hd.w0 = net->wl[0];
    hd.w1 = net->wl[1];
    hd.w2 = net->wl[2];
    hd.b0 = bias(0);
    hd.b1 = bias(1);
    hd.b2 = bias(2);
    hd.out = act[2];
    hipError_t e = launch_conv_head(hd, s);
    if (e != hipSuccess) return cnn_fail("network layers 0-2: %s", hipGetErrorString(e));
    ConvTailArgs t{};
    layer_dims(H, W, 3, hin, win, t.h, t.w);
    t.in = act[2];
    t.w3 = net->wl[3];
    t.w4 = net->wl[4];
    t.w5 = net->wl[5];
    t.b3 = bias(3);
    t.b4 = bias(4);
    t.b5 = bias(5);
    t.out_x = out_x;
    t.out_y = out_y;
    t.out_v = out_v;
    t.ext_x = ext_x;
    t.ext_y = ext_y;
    t.ext_b = ext_b;
    e = launch_conv_tail(t, s);
    if (e != hipSuccess) return cnn_fail("network layers 3-5: %s", hipGetErrorString(e));
    return PIVX_OK;
  }
  for (int l = 0; l < kCnnLayers; ++l) {
    const CnnLayer& L = kCnnArch[l];
    ConvFwdArgs a{};
    layer_dims(H, W, l, a.hin, a.win, a.hout, a.wout);
    a.in = l == 0 ? x0 : act[l - 1];
    a.in_bstride = l == 0 ? xstride : (long long)a.hin * a.win * L.cin;
    a.batch_index = l == 0 ? bidx : nullptr;
    a.batch = B;
    a.w = net->wf[l];
    a.bias = bias(l);
    if (l < kCnnLayers - 1) {
      a.out = act[l];
    } else {
      a.out_x = out_x;
      a.out_y = out_y;
      a.out_bstride = ostride;
      a.out_v = out_v;
      a.ext_x = ext_x;
      a.ext_y = ext_y;
      a.ext_b = ext_b;
    }
    const hipError_t e = launch_conv_fwd(l, a, s);
    if (e != hipSuccess) return cnn_fail("network layer %d: %s", l, hipGetErrorString(e));
  }
  return PIVX_OK;
}

// Loss gradient in net->g[0] -> parameter gradient in net->grad.
int cnn_backward_enqueue(pivx_cnn* net, const uint16_t* x0, long long xstride, const int* bidx, int B, int H, int W,
                         hipStream_t s) {
  int cur = 0;
  // every layer's partial sums in a buffer of its own, reduced in one launch
  // at the end
  const float* lpart[kCnnLayers];
  const float* lpart_b[kCnnLayers];
  int lchunks[kCnnLayers];
  for (int l = kCnnLayers - 1; l >= 0; --l) {
    const CnnLayer& L = kCnnArch[l];
    int hin, win, hout, wout;
    layer_dims(H, W, l, hin, win, hout, wout);
    const long long npx = (long long)B * hout * wout;
    const int chunk = (int)(((npx + kCnnChunks - 1) / kCnnChunks + 31) / 32 * 32);
    const int nchunks = (int)((npx + chunk - 1) / chunk);
    ConvBwdWArgs w{};
    w.in = l == 0 ? x0 : net->act[l - 1];
    w.in_bstride = l == 0 ? xstride : (long long)hin * win * L.cin;
    w.batch_index = l == 0 ? bidx : nullptr;
    w.gout = net->g[cur];
    w.post = L.act ? net->act[l] : nullptr;
    w.partial = net->partial + (size_t)l * net->pstride;
    w.partial_b = net->partial_b + (size_t)l * 64 * kCnnChunks;
    lpart[l] = w.partial;
    lpart_b[l] = w.partial_b;
    lchunks[l] = nchunks;
    w.batch = B;
    w.hin = hin;
    w.win = win;
    w.hout = hout;
    w.wout = wout;
    w.chunk_px = chunk;
    hipError_t e = launch_conv_bwd_w(l, w, nchunks, s);
    if (e == hipSuccess && l > 0) {
      ConvBwdDataArgs d{};
      d.gout = net->g[cur];
      d.post = L.act ? net->act[l] : nullptr;
      d.wt = net->wt[l];
      d.gin = net->g[cur ^ 1];
      d.batch = B;
      d.hin = hin;
      d.win = win;
      d.hout = hout;
      d.wout = wout;
      e = launch_conv_bwd_data(l, d, s);
      cur ^= 1;
    }
    if (e != hipSuccess) return cnn_fail("network backward, layer %d: %s", l, hipGetErrorString(e));
  }
  const hipError_t e = launch_wgrad_reduce_all(lpart, lpart_b, lchunks, net->grad, s);
  if (e != hipSuccess) return cnn_fail("network gradient reduction: %s", hipGetErrorString(e));
  return PIVX_OK;
}

// Batch workspace for B samples of H x W input (reallocated when larger).
int cnn_workspace(pivx_cnn* net, int B, int H, int W) {
  if (net->wb >= B && net->wh == H && net->ww == W) return PIVX_OK;
  (void)hipStreamSynchronize(net->stream);
  for (void* q : net->ws) (void)hipFree(q);
  net->ws.clear();
  net->wb = net->wh = net->ww = 0;
  int rc = 0;
  size_t gmax = (size_t)B * (H / 8) * (W / 8) * 2;
  size_t pmax = 0;
  for (int l = 0; l < kCnnLayers; ++l) {
    int hin, win, hout, wout;
    layer_dims(H, W, l, hin, win, hout, wout);
    const size_t n = (size_t)B * hout * wout * kCnnArch[l].cout;
    if (l < kCnnLayers - 1) rc |= nalloc(net->ws, &net->act[l], n);
    gmax = std::max(gmax, n);
    gmax = std::max(gmax, (size_t)B * hin * win * kCnnArch[l].cin);
    pmax = std::max(pmax, (size_t)cnn_pad16(kCnnArch[l].cout) * cnn_kpad(kCnnArch[l].cin));
  }
  rc |= nalloc(net->ws, &net->out, (size_t)B * 2 * (H / 8) * (W / 8));
  rc |= nalloc(net->ws, &net->g[0], gmax);
  rc |= nalloc(net->ws, &net->g[1], gmax);
  net->pstride = pmax * kCnnChunks;
  rc |= nalloc(net->ws, &net->partial, net->pstride * kCnnLayers);
  rc |= nalloc(net->ws, &net->partial_b, (size_t)64 * kCnnChunks * kCnnLayers);
  rc |= nalloc(net->ws, &net->loss_part, (size_t)cnn_loss_blocks(B, (H / 8) * (W / 8)));
  rc |= nalloc(net->ws, &net->loss, 1);
  rc |= nalloc(net->ws, &net->xin, (size_t)B * H * W * 2);
  rc |= nalloc(net->ws, &net->lab, (size_t)B * 2 * (H / 8) * (W / 8));
  if (rc) return PIVX_ERR_HIP;
  net->wb = B;
  net->wh = H;
  net->ww = W;
  return PIVX_OK;
}

int cnn_pack(pivx_cnn* net) {
  const hipError_t e = launch_pack(net->master, net->wf, net->wt, net->wl, net->stream);
  if (e != hipSuccess) return cnn_fail("weight packing: %s", hipGetErrorString(e));
  return PIVX_OK;
}

// One training step on the batch in net's workspace: forward, loss, backward, Adam.
int cnn_step_enqueue(pivx_cnn* net, const uint16_t* x0, long long xstride, const int* bidx, const float* labels,
                     int B, int H, int W, float lr, float* loss_slot, hipStream_t s) {
  const int hw = (H / 8) * (W / 8);
  int rc = cnn_forward_enqueue(net, x0, xstride, bidx, B, H, W, net->act, net->out, net->out + hw, 2LL * hw, s);
  if (rc) return rc;
  hipError_t e = launch_mse(net->out, 2LL * hw, labels, bidx, B, hw, net->g[0], net->loss_part, s);
  if (e == hipSuccess)
    e = launch_loss_final(net->loss_part, cnn_loss_blocks(B, hw), (double)B * 2 * hw, loss_slot, s);
  if (e != hipSuccess) return cnn_fail("loss: %s", hipGetErrorString(e));
  if ((rc = cnn_backward_enqueue(net, x0, xstride, bidx, B, H, W, s))) return rc;
  if (lr > 0.f) {
    net->step += 1;
    e = launch_adam(net->master, net->grad, net->m, net->v, net->np, lr, 0.9f, 0.999f, 1e-8f, net->step, s);
    if (e == hipSuccess) e = launch_pack(net->master, net->wf, net->wt, net->wl, s);
    if (e != hipSuccess) return cnn_fail("Adam step: %s", hipGetErrorString(e));
  }
  return PIVX_OK;
}

// The learned pipeline needs grid 1 at spacing 32x32 with one vector per
// 32x32 block of the frame (the network's output cells).
int learned_shape_check(const pivx_ctx* c) {
  const pivx_geometry& g = c->grid[0].g;
  if (g.spacing_x != kCnnDown * kCnnStride || g.spacing_y != kCnnDown * kCnnStride)
    return cnn_fail("the learned grid-1 predictor needs grid-1 spacing %dx%d (got %dx%d)", kCnnDown * kCnnStride,
                    kCnnDown * kCnnStride, g.spacing_x, g.spacing_y);
  if (c->p.nx % 512 != 0 || c->p.ny % 32 != 0)
    return cnn_fail("the learned grid-1 predictor needs nx a multiple of 512 and ny of 32 (got %dx%d)", c->p.nx,
                    c->p.ny);
  if (g.nx_vec != c->p.nx / 32 || g.ny_vec != c->p.ny / 32)
    return cnn_fail("grid-1 vector count %dx%d does not match the network output %dx%d", g.nx_vec, g.ny_vec,
                    c->p.nx / 32, c->p.ny / 32);
  return PIVX_OK;
}

// per-handle buffers of the learned pipeline (allocated once, outside any capture)
int learned_alloc(pivx_ctx* c) {
  if (c->learned_ready) return PIVX_OK;
  const int H = c->p.ny / kCnnDown, W = c->p.nx / kCnnDown;
  int rc = 0;
  if (!c->lgate) rc |= dalloc(c, &c->lgate, 1);
  if (!c->lcount) rc |= dalloc(c, &c->lcount, 1);
  if (!c->lgate_work) rc |= dalloc(c, &c->lgate_work, 3);
  for (int l = 0; l < kCnnLayers - 1; ++l) {
    int hin, win, hout, wout;
    layer_dims(H, W, l, hin, win, hout, wout);
    if (!c->cact[l]) rc |= dalloc(c, &c->cact[l], (size_t)hout * wout * kCnnArch[l].cout);
  }
  if (rc) return PIVX_ERR_HIP;
  HIPCHK(c, hipMemsetAsync(c->lgate, 0, sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(c->lcount, 0, sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(c->lgate_work, 0, 3 * sizeof(int), c->stream));
  c->learned_ready = true;
  return PIVX_OK;
}

int learned_enqueue(pivx_ctx* c, const pivx_cnn* net, const pivx_filter_params* f1, const pivx_filter_params* f2,
                    float max_frac) {
  InPair in_pair(c);
  GridBuf& G1 = c->grid[0];
  const pivx_geometry& g1 = G1.g;
  const int nt1 = g1.nt_vec;
  const int H = c->p.ny / kCnnDown, W = c->p.nx / kCnnDown;
  int rc;
  // network field -> grid-1 field; the first layer reads the frames
  // directly (4x4 block means formed on the fly), the last also sets the
  // field's flags and writes its extension (the fallback rewrites both)
  const FrameInput fin{c->f1, c->p.nx, c->f2, c->pitch2};
  if ((rc = cnn_forward_enqueue(net, nullptr, 0, nullptr, 1, H, W, c->cact, G1.fx, G1.fy, 0, c->stream, G1.fv, G1.ex,
                                G1.ey, kExBorder, &fin, true))) {
    c->err = global_error();
    return rc;
  }
  // universal outlier test of the network field: the grid-1 median test with
  // the network vector as the only candidate; rejected vectors are counted
  MedArgs m{};
  m.sx = G1.fx;
  m.sy = G1.fy;
  m.sv = G1.fv;
  m.sstride = 1;
  m.cx = G1.fx;
  m.cy = G1.fy;
  m.cv = G1.fv;
  m.K = 1;
  m.dx = G1.ux;
  m.dy = G1.uy;
  m.dv = G1.uv;
  m.nx_vec = g1.nx_vec;
  m.ny_vec = g1.ny_vec;
  m.wx = c->p.wx1_median;
  m.wy = c->p.wy1_median;
  m.hfwx = (m.wx - 1) / 2;
  m.hfwy = (m.wy - 1) / 2;
  m.min_nneigh = f1->min_nneigh;
  m.thr = f1->insert_thr;
  m.eps = kEpsilon;
  m.accepted = G1.accepted;
  m.scratch = G1.med_scratch;
  const int max_out = (int)std::floor((double)max_frac * nt1);
  // the median pass sets the gate itself where it can (its last block to
  // finish counts the rejected vectors), else k_gate counts them
  const bool gate_in_median = median_gate_supported(m.wx, m.wy);
  if (gate_in_median) {
    m.gate_out = c->lgate;
    m.gate_count = c->lcount;
    m.gate_max = max_out;
    m.gate_acc = c->lgate_work;
    m.gate_ticket = reinterpret_cast<unsigned int*>(c->lgate_work + 1);
  }
  if ((rc = check_launch(c, launch_median(m, c->stream), "network field median test"))) return rc;
  if (!gate_in_median &&
      (rc = check_launch(c, launch_gate(G1.accepted, nt1, max_out, c->lgate, c->lcount, c->stream), "fallback gate")))
    return rc;
  // plain grid 1, run only when the gate says so; with a fraction of 1 or
  // more the gate can never open and the fallback is not enqueued at all
  bool ext_ok = false;
  if (max_frac < 1.f) {
    c->gate = c->lgate;
    rc = enqueue_grid1(c, f1, &ext_ok);
    c->gate = nullptr;
    if (rc) return rc;
  }
  // The extension now belongs to whichever field grid 1 holds: the network's
  // (written above) or the fallback's (written by its last smoothing pass, or
  // rewritten by the ungated extension in enqueue_grid2 when there is none).
  return enqueue_grid2(c, f2, ext_ok);
}

}  // namespace

static void drop_learned_graph(pivx_ctx* c) {
  if (c->lgexec) (void)hipGraphExecDestroy(c->lgexec);
  if (c->lgraph) (void)hipGraphDestroy(c->lgraph);
  c->lgexec = nullptr;
  c->lgraph = nullptr;
}

extern "C" {

int pivx_synth_pair_device(pivx_ctx* c, const pivx_synth_params* sp) {
  if (!c || !sp) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (sp->nx != c->p.nx || sp->ny != c->p.ny)
    return fail(c, PIVX_ERR_PARAM, "synthetic frame size %dx%d differs from the handle's %dx%d", sp->nx, sp->ny,
                c->p.nx, c->p.ny);
  if (sp->flow < 0 || sp->flow > 2 || !(sp->sigma > 0.f)) return fail(c, PIVX_ERR_PARAM, "bad synthetic parameters");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t npix = (size_t)c->p.nx * c->p.ny;
  if (!c->sacc1) {
    if (dalloc(c, &c->sacc1, npix) || dalloc(c, &c->sacc2, npix)) return PIVX_ERR_HIP;
    HIPCHK(c, hipMemsetAsync(c->sacc1, 0, npix * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->sacc2, 0, npix * 4, c->stream));
  }
  SynthRender r{};
  r.nx = sp->nx;
  r.ny = sp->ny;
  r.seed = sp->seed;
  r.nparticles = (long long)std::llround((double)sp->density * (sp->nx + 32.0) * (sp->ny + 32.0));
  r.sigma = sp->sigma;
  r.imin = sp->imin;
  r.imax = sp->imax;
  r.background = sp->background;
  r.noise_sigma = sp->noise_sigma;
  synth_flow_of(sp, &r.flow);
  const int rc = check_launch(c, launch_synth(r, c->sacc1, c->sacc2, c->f1, c->p.nx, c->f2, c->pitch2, c->stream),
                              "synthetic pair");
  if (rc) return rc;
  c->have_frames = true;
  return PIVX_OK;
}

int pivx_cnn_param_count(void) { return cnn_param_count(); }

int pivx_cnn_destroy(pivx_cnn* net) {
  if (!net) return PIVX_OK;
  (void)hipSetDevice(net->device);
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  for (void* q : net->ws) (void)hipFree(q);
  for (void* q : net->allocs) (void)hipFree(q);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
  return PIVX_OK;
}

int pivx_cnn_create(int device_id, pivx_cnn** out) {
  if (!out) return cnn_fail("null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device_id < 0 || device_id >= ndev)
    return cnn_fail("device id %d out of range", device_id);
  if (hipSetDevice(device_id) != hipSuccess) return cnn_fail("hipSetDevice failed");
  pivx_cnn* net = new pivx_cnn();
  net->device = device_id;
  net->np = cnn_param_count();
  if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess) {
    delete net;
    return cnn_fail("hipStreamCreate failed");
  }
  int rc = 0;
  rc |= nalloc(net->allocs, &net->master, net->np);
  rc |= nalloc(net->allocs, &net->m, net->np);
  rc |= nalloc(net->allocs, &net->v, net->np);
  rc |= nalloc(net->allocs, &net->grad, net->np);
  for (int l = 0; l < kCnnLayers && !rc; ++l) {
    const CnnLayer& L = kCnnArch[l];
    const size_t nf = (size_t)cnn_pad16(L.cout) * cnn_kpad(L.cin), nt = (size_t)cnn_pad16(L.cin) * cnn_kpad(L.cout);
    rc |= nalloc(net->allocs, &net->wf[l], nf);
    rc |= nalloc(net->allocs, &net->wt[l], nt);
    rc |= nalloc(net->allocs, &net->wl[l], nf);
    if (!rc && (hipMemsetAsync(net->wf[l], 0, nf * 2, net->stream) != hipSuccess ||
                hipMemsetAsync(net->wt[l], 0, nt * 2, net->stream) != hipSuccess ||
                hipMemsetAsync(net->wl[l], 0, nf * 2, net->stream) != hipSuccess))
      rc = PIVX_ERR_HIP;
  }
  if (!rc && (hipMemsetAsync(net->master, 0, (size_t)net->np * 4, net->stream) != hipSuccess ||
              hipMemsetAsync(net->m, 0, (size_t)net->np * 4, net->stream) != hipSuccess ||
              hipMemsetAsync(net->v, 0, (size_t)net->np * 4, net->stream) != hipSuccess ||
              hipMemsetAsync(net->grad, 0, (size_t)net->np * 4, net->stream) != hipSuccess ||
              hipStreamSynchronize(net->stream) != hipSuccess))
    rc = PIVX_ERR_HIP;
  if (rc) {
    pivx_cnn_destroy(net);
    return rc;
  }
  *out = net;
  return PIVX_OK;
}

int pivx_cnn_set_params(pivx_cnn* net, const float* params, int n) {
  if (!net || !params || n != net->np) return cnn_fail("parameter count must be %d", net ? net->np : 0);
  (void)hipSetDevice(net->device);
  if (hipMemcpyAsync(net->master, params, (size_t)n * 4, hipMemcpyHostToDevice, net->stream) != hipSuccess ||
      hipMemsetAsync(net->m, 0, (size_t)n * 4, net->stream) != hipSuccess ||
      hipMemsetAsync(net->v, 0, (size_t)n * 4, net->stream) != hipSuccess)
    return cnn_fail("parameter upload failed");
  net->step = 0;
  int rc = cnn_pack(net);
  if (rc) return rc;
  if (hipStreamSynchronize(net->stream) != hipSuccess) return cnn_fail("parameter upload failed");
  return PIVX_OK;
}

int pivx_cnn_get_params(pivx_cnn* net, float* params, int n) {
  if (!net || !params || n != net->np) return cnn_fail("parameter count must be %d", net ? net->np : 0);
  (void)hipSetDevice(net->device);
  if (hipMemcpyAsync(params, net->master, (size_t)n * 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess ||
      hipStreamSynchronize(net->stream) != hipSuccess)
    return cnn_fail("parameter download failed");
  return PIVX_OK;
}

int pivx_cnn_init_params(pivx_cnn* net, unsigned long long seed) {
  if (!net) return cnn_fail("null network");
  std::vector<float> p((size_t)net->np, 0.f);
  unsigned long long st = seed;
  auto next = [&st]() {
    st += 0x9e3779b97f4a7c15ULL;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < kCnnLayers; ++l) {
    const CnnLayer& L = kCnnArch[l];
    const double sd = std::sqrt(2.0 / (9.0 * L.cin)) * (l == kCnnLayers - 1 ? 0.1 : 1.0);
    const int o = cnn_weight_offset(l);
    for (int i = 0; i < L.cout * 9 * L.cin; ++i) {
      double u1 = next(), u2 = next();
      if (u1 < 1e-300) u1 = 1e-300;
      p[o + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
    }
  }
  return pivx_cnn_set_params(net, p.data(), net->np);
}

static int cnn_upload_batch(pivx_cnn* net, const float* x0, const float* labels, int B, int H, int W) {
  int rc = cnn_workspace(net, B, H, W);
  if (rc) return rc;
  std::vector<uint16_t> xb((size_t)B * H * W * 2);
  for (size_t i = 0; i < xb.size(); ++i) xb[i] = host_bf16(x0[i]);
  if (hipMemcpy(net->xin, xb.data(), xb.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
    return cnn_fail("input upload failed");
  if (labels && hipMemcpy(net->lab, labels, (size_t)B * 2 * (H / 8) * (W / 8) * 4, hipMemcpyHostToDevice) != hipSuccess)
    return cnn_fail("label upload failed");
  return PIVX_OK;
}

int pivx_cnn_forward(pivx_cnn* net, const float* x0, int batch, int h, int w, float* out) {
  if (!net || !x0 || !out || batch < 1) return cnn_fail("bad argument");
  if (!cnn_shape_ok(h, w)) return cnn_fail("network input %dx%d: h must be a multiple of 8, w of 128", h, w);
  (void)hipSetDevice(net->device);
  int rc = cnn_upload_batch(net, x0, nullptr, batch, h, w);
  if (rc) return rc;
  const int hw = (h / 8) * (w / 8);
  if ((rc = cnn_forward_enqueue(net, net->xin, (long long)h * w * 2, nullptr, batch, h, w, net->act, net->out,
                                net->out + hw, 2LL * hw, net->stream, nullptr, nullptr, nullptr, 0, nullptr, true)))
    return rc;
  if (hipMemcpyAsync(out, net->out, (size_t)batch * 2 * hw * 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess ||
      hipStreamSynchronize(net->stream) != hipSuccess)
    return cnn_fail("network forward failed");
  return PIVX_OK;
}

int pivx_cnn_gradient(pivx_cnn* net, const float* x0, const float* labels, int batch, int h, int w, float* loss,
                      float* grad) {
  if (!net || !x0 || !labels || batch < 1) return cnn_fail("bad argument");
  if (!cnn_shape_ok(h, w)) return cnn_fail("network input %dx%d: h must be a multiple of 8, w of 128", h, w);
  (void)hipSetDevice(net->device);
  int rc = cnn_upload_batch(net, x0, labels, batch, h, w);
  if (rc) return rc;
  if ((rc = cnn_step_enqueue(net, net->xin, (long long)h * w * 2, nullptr, net->lab, batch, h, w, 0.f, net->loss,
                             net->stream)))
    return rc;
  if (loss && hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess)
    return cnn_fail("loss download failed");
  if (grad && hipMemcpyAsync(grad, net->grad, (size_t)net->np * 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess)
    return cnn_fail("gradient download failed");
  if (hipStreamSynchronize(net->stream) != hipSuccess) return cnn_fail("network gradient failed");
  return PIVX_OK;
}

int pivx_cnn_adam_step(pivx_cnn* net, float lr) {
  if (!net || !(lr > 0.f)) return cnn_fail("bad argument");
  (void)hipSetDevice(net->device);
  net->step += 1;
  hipError_t e = launch_adam(net->master, net->grad, net->m, net->v, net->np, lr, 0.9f, 0.999f, 1e-8f, net->step,
                             net->stream);
  if (e == hipSuccess) e = launch_pack(net->master, net->wf, net->wt, net->wl, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  if (e != hipSuccess) return cnn_fail("Adam step: %s", hipGetErrorString(e));
  return PIVX_OK;
}

int pivx_cnn_train(pivx_ctx* c, pivx_cnn* net, const pivx_cnn_train_params* tp, float* epoch_loss,
                   pivx_cnn_train_report* rep) {
  if (!c || !net || !tp) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (net->device != c->device) return fail(c, PIVX_ERR_PARAM, "network and handle are on different devices");
  if (tp->npairs < 1 || tp->epochs < 1 || tp->batch < 1 || tp->batch > tp->npairs || !(tp->lr > 0.f))
    return fail(c, PIVX_ERR_PARAM, "bad training parameters");
  int rc = learned_shape_check(c);
  if (rc) {
    c->err = global_error();
    return rc;
  }
  if (c->p.npass1 < 1) return fail(c, PIVX_ERR_PARAM, "labels need at least one grid-1 pass");
  HIPCHK(c, hipSetDevice(c->device));
  const int N = tp->npairs, B = tp->batch;
  const int H = c->p.ny / kCnnDown, W = c->p.nx / kCnnDown;
  const int hw = (H / 8) * (W / 8);
  const long long xstride = (long long)H * W * 2;
  if ((rc = cnn_workspace(net, B, H, W))) return rc;
  // the dataset: network inputs and plain grid-1 labels of every pair
  std::vector<void*> data;
  struct Release {
    std::vector<void*>& d;
    ~Release() {
      for (void* q : d) (void)hipFree(q);
    }
  } release{data};
  uint16_t* X = nullptr;
  float* Y = nullptr;
  int* perm = nullptr;
  float* losses = nullptr;
  const int steps = N / B;
  if (nalloc(data, &X, (size_t)N * xstride) || nalloc(data, &Y, (size_t)N * 2 * hw) ||
      nalloc(data, &perm, (size_t)steps * B) || nalloc(data, &losses, (size_t)tp->epochs * steps))
    return fail(c, PIVX_ERR_HIP, "training data allocation failed");
  const auto t0 = std::chrono::steady_clock::now();
  GridBuf& G1 = c->grid[0];
  c->gate = nullptr;
  for (int i = 0; i < N; ++i) {
    pivx_synth_params sp = tp->synth;
    sp.seed = tp->synth.seed + (unsigned long long)i;
    if ((rc = pivx_synth_pair_device(c, &sp))) return rc;
    bool ext_ok = false;
    {
      InPair in_pair(c);
      if ((rc = enqueue_grid1(c, &tp->grid1, &ext_ok))) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(Y + (size_t)i * 2 * hw, G1.fx, (size_t)hw * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(Y + (size_t)i * 2 * hw + hw, G1.fy, (size_t)hw * 4, hipMemcpyDeviceToDevice, c->stream));
    const hipError_t e = launch_downsample(c->f1, c->p.nx, c->f2, c->pitch2, c->p.nx, c->p.ny, X + (size_t)i * xstride,
                                           c->stream);
    if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "network input: %s", hipGetErrorString(e));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const auto t1 = std::chrono::steady_clock::now();
  // minibatch Adam over seeded shuffles, cosine-decayed learning rate
  std::vector<int> order((size_t)N);
  unsigned long long st = tp->shuffle_seed;
  auto rnd = [&st]() {
    st += 0x9e3779b97f4a7c15ULL;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  const long long total = (long long)tp->epochs * steps;
  long long t = 0;
  for (int ep = 0; ep < tp->epochs; ++ep) {
    for (int i = 0; i < N; ++i) order[i] = i;
    for (int i = N - 1; i > 0; --i) std::swap(order[i], order[(size_t)(rnd() % (unsigned long long)(i + 1))]);
    // the previous epoch's steps read perm: wait before overwriting it
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(perm, order.data(), (size_t)steps * B * sizeof(int), hipMemcpyHostToDevice));
    for (int s = 0; s < steps; ++s, ++t) {
      const float lr = (float)(0.5 * tp->lr * (1.0 + std::cos(M_PI * (double)t / (double)total)));
      if ((rc = cnn_step_enqueue(net, X, xstride, perm + (size_t)s * B, Y, B, H, W, lr,
                                 losses + (size_t)ep * steps + s, c->stream))) {
        c->err = global_error();
        return rc;
      }
    }
  }
  std::vector<float> lh((size_t)total);
  HIPCHK(c, hipMemcpyAsync(lh.data(), losses, lh.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const auto t2 = std::chrono::steady_clock::now();
  std::vector<float> el((size_t)tp->epochs, 0.f);
  for (int ep = 0; ep < tp->epochs; ++ep) {
    double s = 0.0;
    for (int k = 0; k < steps; ++k) s += lh[(size_t)ep * steps + k];
    el[ep] = (float)(s / steps);
    if (epoch_loss) epoch_loss[ep] = el[ep];
  }
  if (rep) {
    rep->gen_seconds = std::chrono::duration<double>(t1 - t0).count();
    rep->train_seconds = std::chrono::duration<double>(t2 - t1).count();
    rep->steps = (int)total;
    rep->first_epoch_loss = el.front();
    rep->last_epoch_loss = el.back();
  }
  return PIVX_OK;
}

static bool same_learned(const pivx_ctx* c, const pivx_cnn* net, const pivx_filter_params* f1,
                         const pivx_filter_params* f2, float frac) {
  return c->lnet == net && same_filter(c->lf[0], *f1) && same_filter(c->lf[1], *f2) && c->lfrac == frac;
}

int pivx_run_pair_learned(pivx_ctx* c, const pivx_cnn* net, const pivx_filter_params* f1,
                          const pivx_filter_params* f2, float max_outlier_fraction) {
  if (!c || !net || !f1 || !f2) return fail(c, PIVX_ERR_PARAM, "null argument");
  if (!c->have_frames) return fail(c, PIVX_ERR_STATE, "no frames: call pivx_prep_img_data first");
  if (net->device != c->device) return fail(c, PIVX_ERR_PARAM, "network and handle are on different devices");
  if (!(max_outlier_fraction >= 0.f)) return fail(c, PIVX_ERR_PARAM, "max_outlier_fraction must be >= 0");
  int rc = learned_shape_check(c);
  if (rc) {
    c->err = global_error();
    return rc;
  }
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = learned_alloc(c))) return rc;
  // the network's parameters were written on its own stream
  HIPCHK(c, hipStreamSynchronize(net->stream));
  if (!c->use_graph) return learned_enqueue(c, net, f1, f2, max_outlier_fraction);
  if (!c->lgexec || !same_learned(c, net, f1, f2, max_outlier_fraction)) {
    drop_learned_graph(c);
    const bool timing = c->timing;
    c->timing = false;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = learned_enqueue(c, net, f1, f2, max_outlier_fraction);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = timing;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, PIVX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    c->lgraph = g;
    e = hipGraphInstantiate(&c->lgexec, c->lgraph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_learned_graph(c);
      return fail(c, PIVX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    c->lnet = net;
    c->lf[0] = *f1;
    c->lf[1] = *f2;
    c->lfrac = max_outlier_fraction;
  }
  tic(c);
  HIPCHK(c, hipGraphLaunch(c->lgexec, c->stream));
  toc(c, 0, 5);
  return PIVX_OK;
}

int pivx_get_learned_stats(pivx_ctx* c, int* outliers, int* fallback) {
  if (!c) return fail(nullptr, PIVX_ERR_PARAM, "null handle");
  if (!c->lgate) return fail(c, PIVX_ERR_STATE, "no learned pair has run on this handle");
  int v[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(&v[0], c->lcount, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&v[1], c->lgate, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (outliers) *outliers = v[0];
  if (fallback) *fallback = v[1];
  return PIVX_OK;
}

}  // extern "C"

// ====================================================================== batched passes
struct pivx_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<pivx_ctx*> ctx;
  XcArgs* dargs[3] = {};  // pass kinds: grid-1 pass 0, grid-1 deformed pass, grid-2 pass
  int* dstart[2] = {};    // work-item offsets per grid
  int total[2] = {0, 0};
  int* counter = nullptr;
  int cus = 1;
  int wgs[2] = {0, 0}, chunk[2] = {1, 1};
  double tuned_ms[2] = {0.0, 0.0};
  std::vector<hipEvent_t> ev;  // one per handle
  hipEvent_t evb = nullptr;    // on the batch stream
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  pivx_filter_params gf[2] = {};
  std::vector<void*> allocs;
};

namespace {

int batch_fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_global_error(buf);
  return PIVX_ERR_PARAM;
}

BatchQueue batch_queue(const pivx_batch* b, int kind) {
  BatchQueue q{};
  const int gi = kind == 2 ? 1 : 0;
  q.args = b->dargs[kind];
  q.item_start = b->dstart[gi];
  q.npairs = (int)b->ctx.size();
  q.total = b->total[gi];
  q.counter = b->counter;
  q.chunk = b->chunk[gi];
  return q;
}

int batch_corr(pivx_batch* b, int kind) {
  const int gi = kind == 2 ? 1 : 0;
  const pivx_geometry& g = b->ctx[0]->grid[gi].g;
  const hipError_t e = launch_xcorr_batch(g.ws_x, g.ws_y, batch_queue(b, kind), b->wgs[gi], b->stream);
  if (e != hipSuccess) return batch_fail("batched correlation: %s", hipGetErrorString(e));
  return PIVX_OK;
}

#define BCHK(expr)                                                                        \
  do {                                                                                    \
    const hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return batch_fail("%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

// The per-pair sequence of pivx_run_pair for every handle, with each
// correlation pass one persistent launch over all pairs on the batch stream;
// the per-pair stages (deformation, validation, extension, upsampling) run on
// the handles' own streams, joined to the batch stream by events.
int batch_enqueue(pivx_batch* b, const pivx_filter_params* f1, const pivx_filter_params* f2) {
  const size_t n = b->ctx.size();
  const pivx_init_params& P = b->ctx[0]->p;
  std::vector<char> ext_ok(n, 0);
  int rc;
  BCHK(hipEventRecord(b->evb, b->stream));
  for (pivx_ctx* c : b->ctx) BCHK(hipStreamWaitEvent(c->stream, b->evb, 0));
  auto join_to_batch = [&]() -> int {
    for (size_t i = 0; i < n; ++i) {
      BCHK(hipEventRecord(b->ev[i], b->ctx[i]->stream));
      BCHK(hipStreamWaitEvent(b->stream, b->ev[i], 0));
    }
    return PIVX_OK;
  };
  auto fork_from_batch = [&]() -> int {
    BCHK(hipEventRecord(b->evb, b->stream));
    for (pivx_ctx* c : b->ctx) BCHK(hipStreamWaitEvent(c->stream, b->evb, 0));
    return PIVX_OK;
  };
  for (int j = 0; j < P.npass1; ++j) {
    for (size_t i = 0; i < n && j > 0; ++i) {
      pivx_ctx* c = b->ctx[i];
      if (!ext_ok[i] && (rc = extend(c, 0))) return rc;
      if ((rc = run_deform(c, 0))) return rc;
    }
    if ((rc = join_to_batch())) return rc;
    if ((rc = batch_corr(b, j == 0 ? 0 : 1))) return rc;
    if ((rc = fork_from_batch())) return rc;
    for (size_t i = 0; i < n; ++i) {
      pivx_ctx* c = b->ctx[i];
      GridBuf& G1 = c->grid[0];
      bool ok = false;
      if ((rc = run_filter(c, 0, f1->npass_median, f1->min_nneigh, f1->insert_thr, f1->npass_smooth, 1, G1.fx, G1.fy,
                           G1.fv, G1.ex, G1.ey, &ok)))
        return rc;
      ext_ok[i] = ok;
    }
  }
  if (P.npass2 > 0) {
    for (size_t i = 0; i < n; ++i) {
      pivx_ctx* c = b->ctx[i];
      if (!ext_ok[i] && (rc = extend(c, 0))) return rc;
      if ((rc = upsample(c, true))) return rc;
      ext_ok[i] = 1;
    }
    for (int j = 0; j < P.npass2; ++j) {
      for (size_t i = 0; i < n; ++i) {
        pivx_ctx* c = b->ctx[i];
        if (!ext_ok[i] && (rc = extend(c, 1))) return rc;
        if ((rc = run_deform(c, 1))) return rc;
      }
      if ((rc = join_to_batch())) return rc;
      if ((rc = batch_corr(b, 2))) return rc;
      if ((rc = fork_from_batch())) return rc;
      const bool last = j == P.npass2 - 1;
      for (size_t i = 0; i < n; ++i) {
        pivx_ctx* c = b->ctx[i];
        GridBuf& G2 = c->grid[1];
        bool ok = false;
        if ((rc = run_filter(c, 1, f2->npass_median, f2->min_nneigh, f2->insert_thr, f2->npass_smooth, 1, G2.fx, G2.fy,
                             G2.fv, G2.ex, G2.ey, &ok, last ? c->mo_x : nullptr, last ? c->mo_y : nullptr,
                             last ? c->mo_v : nullptr)))
          return rc;
        ext_ok[i] = ok;
      }
    }
  }
  return join_to_batch();
}

void drop_batch_graph(pivx_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" {

int pivx_batch_destroy(pivx_batch* b) {
  if (!b) return PIVX_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_batch_graph(b);
  for (void* q : b->allocs) (void)hipFree(q);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  if (b->evb) (void)hipEventDestroy(b->evb);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PIVX_OK;
}

int pivx_batch_create(pivx_ctx* const* ctxs, int n, pivx_batch** out) {
  if (!ctxs || n < 1 || !out) return batch_fail("bad argument");
  *out = nullptr;
  const pivx_ctx* c0 = ctxs[0];
  if (!c0) return batch_fail("null handle");
  for (int gi = 0; gi < 2; ++gi) {
    const pivx_geometry& g = c0->grid[gi].g;
    if (!xcorr_batch_supported(g.ws_x, g.ws_y))
      return batch_fail("no batched correlation kernel for %dx%d windows (grid %d)", g.ws_x, g.ws_y, gi + 1);
  }
  for (int i = 0; i < n; ++i) {
    const pivx_ctx* c = ctxs[i];
    if (!c) return batch_fail("null handle %d", i);
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == c) return batch_fail("handle %d appears twice", i);
    if (c->device != c0->device) return batch_fail("handles on different devices");
    if (c->prec != 0) return batch_fail("handle %d uses the precision form; batched passes use the FP32 kernels", i);
    const pivx_init_params &p = c->p, &p0 = c0->p;
    if (p.ws_x1 != p0.ws_x1 || p.ws_y1 != p0.ws_y1 || p.ws_x2 != p0.ws_x2 || p.ws_y2 != p0.ws_y2 ||
        p.npass1 != p0.npass1 || p.npass2 != p0.npass2 || p.npeak_save != p0.npeak_save)
      return batch_fail("handle %d differs in window sizes, pass counts or Npeak_save", i);
  }
  if (hipSetDevice(c0->device) != hipSuccess) return batch_fail("hipSetDevice failed");
  pivx_batch* b = new pivx_batch();
  b->device = c0->device;
  b->ctx.assign(ctxs, ctxs + n);
  int rc = PIVX_OK;
  auto hip = [&](hipError_t e) {
    if (e != hipSuccess && rc == PIVX_OK) rc = batch_fail("batch set-up: %s", hipGetErrorString(e));
  };
  hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  b->ev.assign((size_t)n, nullptr);
  for (int i = 0; i < n && rc == PIVX_OK; ++i) hip(hipEventCreateWithFlags(&b->ev[i], hipEventDisableTiming));
  if (rc == PIVX_OK) hip(hipEventCreateWithFlags(&b->evb, hipEventDisableTiming));
  int cus = 0;
  if (rc == PIVX_OK) hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  b->cus = std::max(1, cus);
  std::vector<XcArgs> args[3];
  std::vector<int> start[2];
  for (int gi = 0; gi < 2 && rc == PIVX_OK; ++gi) {
    start[gi].push_back(0);
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
      const pivx_geometry& g = ctxs[i]->grid[gi].g;
      int chunks_x = 0;
      tot += xcorr_batch_items(g.ws_x, g.ws_y, g.nx_vec, g.ny_vec, &chunks_x);
      if (tot > 0x7fffffffLL) rc = batch_fail("batch holds too many work items");
      start[gi].push_back((int)tot);
      for (int kind = (gi == 0 ? 0 : 2); kind <= (gi == 0 ? 1 : 2); ++kind) {
        XcArgs a = make_xcargs(ctxs[i], gi, kind != 0, kind == 0, false);
        a.chunks_x = chunks_x;
        args[kind].push_back(a);
      }
    }
    b->total[gi] = (int)tot;
  }
  for (int kind = 0; kind < 3 && rc == PIVX_OK; ++kind) {
    void* q = nullptr;
    hip(hipMalloc(&q, args[kind].size() * sizeof(XcArgs)));
    if (q) {
      b->allocs.push_back(q);
      b->dargs[kind] = static_cast<XcArgs*>(q);
      hip(hipMemcpy(q, args[kind].data(), args[kind].size() * sizeof(XcArgs), hipMemcpyHostToDevice));
    }
  }
  for (int gi = 0; gi < 2 && rc == PIVX_OK; ++gi) {
    void* q = nullptr;
    hip(hipMalloc(&q, start[gi].size() * sizeof(int)));
    if (q) {
      b->allocs.push_back(q);
      b->dstart[gi] = static_cast<int*>(q);
      hip(hipMemcpy(q, start[gi].data(), start[gi].size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  if (rc == PIVX_OK) {
    void* q = nullptr;
    hip(hipMalloc(&q, 256));
    if (q) {
      b->allocs.push_back(q);
      b->counter = static_cast<int*>(q);
    }
  }
  // untuned defaults: 8 persistent waves per CU, 4 items per claim
  for (int gi = 0; gi < 2; ++gi) {
    b->wgs[gi] = 8 * b->cus;
    b->chunk[gi] = 4;
  }
  if (rc != PIVX_OK) {
    pivx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PIVX_OK;
}

int pivx_batch_set_graph_mode(pivx_batch* b, int on) {
  if (!b) return batch_fail("null batch");
  b->use_graph = on != 0;
  if (!b->use_graph) drop_batch_graph(b);
  return PIVX_OK;
}

int pivx_batch_run(pivx_batch* b, const pivx_filter_params* f1, const pivx_filter_params* f2) {
  if (!b || !f1 || !f2) return batch_fail("null argument");
  for (pivx_ctx* c : b->ctx)
    if (!c->have_frames) return batch_fail("a handle of the batch has no frames");
  if (hipSetDevice(b->device) != hipSuccess) return batch_fail("hipSetDevice failed");
  if (!b->use_graph) return batch_enqueue(b, f1, f2);
  if (!b->gexec || !same_filter(b->gf[0], *f1) || !same_filter(b->gf[1], *f2)) {
    drop_batch_graph(b);
    BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, f1, f2);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return batch_fail("hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    const hipError_t e2 = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
      drop_batch_graph(b);
      return batch_fail("hipGraphInstantiate: %s", hipGetErrorString(e2));
    }
    b->gf[0] = *f1;
    b->gf[1] = *f2;
  }
  BCHK(hipGraphLaunch(b->gexec, b->stream));
  return PIVX_OK;
}

int pivx_batch_synchronize(pivx_batch* b) {
  if (!b) return batch_fail("null batch");
  BCHK(hipStreamSynchronize(b->stream));
  for (pivx_ctx* c : b->ctx) {
    const int rc = pivx_synchronize(c);
    if (rc) return rc;
  }
  return PIVX_OK;
}

int pivx_batch_autotune(pivx_batch* b, int reps) {
  if (!b || reps < 1) return batch_fail("bad argument");
  for (pivx_ctx* c : b->ctx)
    if (!c->have_frames) return batch_fail("a handle of the batch has no frames");
  if (hipSetDevice(b->device) != hipSuccess) return batch_fail("hipSetDevice failed");
  drop_batch_graph(b);
  for (pivx_ctx* c : b->ctx) BCHK(hipStreamSynchronize(c->stream));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  BCHK(hipEventCreate(&t0));
  BCHK(hipEventCreate(&t1));
  const int waves_per_cu[4] = {4, 8, 16, 32};
  const int chunks[3] = {1, 4, 16};
  int rc = PIVX_OK;
  for (int gi = 0; gi < 2 && rc == PIVX_OK; ++gi) {
    const int kind = gi == 0 ? 0 : 2;
    double best = 1e30;
    int bw = b->wgs[gi], bc = b->chunk[gi];
    for (int wv : waves_per_cu)
      for (int ch : chunks) {
        b->wgs[gi] = wv * b->cus;
        b->chunk[gi] = ch;
        if ((rc = batch_corr(b, kind))) break;  // warm-up
        if (hipEventRecord(t0, b->stream) != hipSuccess) rc = batch_fail("event record failed");
        for (int r = 0; r < reps && rc == PIVX_OK; ++r) rc = batch_corr(b, kind);
        float ms = 0.f;
        if (rc == PIVX_OK && (hipEventRecord(t1, b->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                              hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
          rc = batch_fail("autotune timing failed");
        if (rc) break;
        if (ms / reps < best) {
          best = ms / reps;
          bw = b->wgs[gi];
          bc = b->chunk[gi];
        }
      }
    b->wgs[gi] = bw;
    b->chunk[gi] = bc;
    b->tuned_ms[gi] = best;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return rc;
}

int pivx_batch_get_tuning(const pivx_batch* b, int grid, int* waves_per_cu, int* items_per_claim, double* ms,
                          long long* items) {
  if (!b || (grid != 1 && grid != 2)) return batch_fail("bad argument");
  const int gi = grid - 1;
  if (waves_per_cu) *waves_per_cu = b->wgs[gi] / b->cus;
  if (items_per_claim) *items_per_claim = b->chunk[gi];
  if (ms) *ms = b->tuned_ms[gi];
  if (items) *items = b->total[gi];
  return PIVX_OK;
}

}  // extern "C"
