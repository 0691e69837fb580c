/*
 * oracle.h -- CPU restatement of the reference's PIV stages.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load liboracle.so, and only as the checker (or
 * as the timed CPU baseline); the product (libpivx.so, piv_xcorr2d) never
 * links or calls it.
 *
 * Every stage is templated on the arithmetic type and exported with a `prec`
 * argument: 64 = the FP64 oracle, 32 = an FP32 transcription in the
 * reference's operation order (used for bit-exact checks of the FP32
 * element-wise stages and as the timed scalar CPU baseline).  All array
 * arguments are double (float values are represented exactly).
 *
 * Citations: xcorr2d_V1.0/xcorr2d_multi_grid.cu (cu:N), main.cpp (main:N).
 * Parity pins: see the header of oracle.cpp.
 */
#ifndef PIV_ORACLE_H
#define PIV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry of one grid (cu:751-894).  out[14] = nx_vec, ny_vec, border_x,
 * border_y, npeak_slots, left_x, left_y, right_x, right_y, nx_aug, ny_aug,
 * nsec, neach, nlast. */
int orc_geometry(int nx, int ny, int nsec, int wsx, int wsy, int spx, int spy, float pctx, float pcty, int* out);

/* Correlation plane of one window (A3-A8): cc[j][i] (fftshifted, row-major,
 * wsy x wsx) = |IFFT(conj(FFT f1) * FFT f2)| / N.  method 0: FFT, 1: direct
 * circular sum  cc(s) = sum_x f1(x) f2((x + s) mod ws). */
int orc_correlate_window(int prec, const double* f1w, const double* f2w, int wsx, int wsy, int method, double* cc);

/* Unnormalised 1-D DFT (sign -1 forward, +1 inverse) of n complex values
 * (interleaved re, im), by the oracle's mixed-radix FFT. */
int orc_fft1d(int prec, double* x, int n, int sign);

typedef struct orc_pass {
  const uint8_t* f1;        /* frame 1 [ny][nx] */
  const uint8_t* f2u8;      /* frame 2 (undeformed pass) or NULL */
  const double* f2def;      /* deformed frame 2 [ny_aug][nx_aug] or NULL */
  int nx, ny, nx_aug, ny_aug;
  int wsx, wsy, spx, spy, left_x, left_y, nx_vec, ny_vec, bx, by, K;
  double lx1, lx2, ly1, ly2;
  const double* pred_x;     /* predictor [nt] or NULL */
  const double* pred_y;
  int method;               /* 0 FFT, 1 direct */
  double tie_rel;           /* relative tie margin for the tie flags */
  int threads;              /* OpenMP threads (<= 0: 1) */
  int win_begin, win_end;   /* window range [begin, end); end <= 0 means all */
  double fit_tol;           /* > 0: also flag windows whose sub-pixel fit moves by more than fit_tol px */
  double fit_rel;           /*   when the plane values move by fit_rel * peak (nearly flat peaks) */
} orc_pass;

/* One correlation pass (A3-A15): uncapped 4-neighbour local maxima, top-K in
 * (value desc, raster index asc) order, 3-point Gaussian fit with predictor
 * and limits, NaN/Inf -> invalid.  Outputs [nt][K] (pk_idx -1 when no
 * candidate), ncand [nt], tie [nt] = 1 when the window's top-K order or
 * local-maximum set is within tie_rel * peak of flipping, 2 when instead a
 * valid candidate's fit is ill-conditioned (see fit_tol).  Any output may be
 * NULL. */
int orc_xcorr_pass(int prec, const orc_pass* P, double* disp_x, double* disp_y, uint8_t* valid, double* pk_val,
                   int* pk_idx, int* ncand, uint8_t* tie);

/* First-order bound on how far the 3-point Gaussian fit of one axis
 * (cu:334-340) moves when each of its three plane values f0 (left), f1 (peak),
 * f2 (right) moves by at most delta; used for the ill-conditioned-fit flag. */
double orc_fit_sensitivity(double f0, double f1, double f2, double delta);

/* One FP32 correlation pass, single-threaded, timed per stage (seconds):
 * gather, transforms with product/magnitude/shift, peaks with fit. */
int orc_xcorr_stage_seconds(const orc_pass* P, double* t_gather, double* t_corr, double* t_peaks);

/* Kernel_Window_Deformation_2d (cu:608-652); taps outside the image read 0. */
int orc_deform(int prec, const uint8_t* f2, int nx, int ny, int left_x, int left_y, int nx_aug, int ny_aug,
               const double* ex_x, const double* ex_y, int nxe, int nye, double x_st, double y_st, double spx,
               double spy, double* out);

/* extend_filtered_vecfield (cu:1547-1576): border replicated b times. */
int orc_extend(const double* fx, const double* fy, int nx, int ny, int b, double* ex, double* ey);

/* Upsampling_vecfield_kernel_2d (cu:675-704). */
int orc_upsample(int prec, const double* ex, const double* ey, int nxe, int nye, double x_st1, double y_st1,
                 double sp1x, double sp1y, int sp2x, int sp2y, int nx2, int ny2, double* dx2, double* dy2);

/* One normalised-median pass (cu:174-296): neighbourhood from (sx, sy, sv)
 * with element stride sstride, candidates (cx, cy, cv) [nt][K].  margin (may be
 * NULL) receives per vector the smallest |statistic - thr| / thr over the
 * candidates tested. */
int orc_median_pass(int prec, const double* sx, const double* sy, const uint8_t* sv, int sstride, const double* cx,
                    const double* cy, const uint8_t* cv, int K, int nx_vec, int ny_vec, int wx, int wy,
                    int min_nneigh, double thr, double eps, double* dx, double* dy, uint8_t* dv, double* margin);

/* Hole filling until a sweep fills nothing (cu:352-397, 1514-1528). */
int orc_fill(int prec, double* dx, double* dy, uint8_t* dv, int nx_vec, int ny_vec, int* sweeps);

/* One top-hat smoothing pass (cu:399-443), raster summation order. */
int orc_smooth(int prec, const double* sx, const double* sy, int nx_vec, int ny_vec, int wx, int wy, double* dx,
               double* dy);

/* The whole per-pair sequence of main:288-375. */
typedef struct orc_run {
  int nx, ny, nsec;
  int ws_x1, ws_y1, sp_x1, sp_y1, npass1;
  int ws_x2, ws_y2, sp_x2, sp_y2, npass2;
  int wx1_median, wy1_median, wx2_median, wy2_median;
  int wx1_smooth, wy1_smooth, wx2_smooth, wy2_smooth;
  int K;
  double pct_x1, pct_y1, pct_x2, pct_y2;
  double lx1, lx2, ly1, ly2;
  int npass1_median, min_nneigh1; double thr1; int npass1_smooth;
  int npass2_median, min_nneigh2; double thr2; int npass2_smooth;
  int threads;
} orc_run;

/* Outputs (any may be NULL): raw1 [nt1][K] dx,dy,valid; filt1 [nt1];
 * raw2 [nt2][K]; filt2 [nt2] (filled + smoothed); mo2 [nt2] (median-only). */
int orc_run_pair(int prec, const orc_run* R, const uint8_t* f1, const uint8_t* f2, double* raw1x, double* raw1y,
                 uint8_t* raw1v, double* f1x, double* f1y, uint8_t* f1v, double* raw2x, double* raw2y, uint8_t* raw2v,
                 double* f2x, double* f2y, uint8_t* f2v, double* mo2x, double* mo2y, uint8_t* mo2v);

/* MHE contrast enhancement of one nx*ny u8 frame (MHE7_4_gpu.cu:45-253).
 * Returns 1 for a window or threshold the reference cannot run. */
int orc_mhe(const uint8_t* in, uint8_t* out, int nx, int ny, int ws, float thresh);

/* Learned grid-1 predictor (cnn_oracle.cpp).  params: the master layout of
 * the GPU network, per layer W [cout][9][cin] then b [cout]; x0: network
 * input [batch][h][w][2]; out and labels: [batch][2][h/8 * w/8].  bf16_storage
 * rounds to bfloat16 wherever the GPU stores bfloat16.  grad (optional):
 * dL/dparams of the mean squared error in the params layout. */
int orc_cnn_param_count(void);
int orc_cnn_forward(const double* params, const double* x0, int batch, int h, int w, int bf16_storage, double* out);
int orc_cnn_grad(const double* params, const double* x0, const double* labels, int batch, int h, int w,
                 int bf16_storage, double* loss, double* grad);

#ifdef __cplusplus
}
#endif
#endif
