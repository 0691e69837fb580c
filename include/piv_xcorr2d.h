/*
 * piv_xcorr2d.h -- C ABI of the MI355X-native multi-grid PIV cross-correlation
 * engine (libpivx.so).
 *
 * The reference exposes this path as the C++ class `xcorr2d_multi`
 * (xcorr2d_V1.0/xcorr2d_multi_grid.cuh:12-208) driven by `main`
 * (xcorr2d_V1.0/main.cpp:227-376).  Each entry point below replaces one method
 * of that class (cited per function); the object-with-public-members pattern is
 * replaced by an opaque handle, plain pointers and sizes, and int status codes.
 *
 * Conventions
 *  - Every call returns PIVX_OK (0) or a positive PIVX_ERR_* code; the text of
 *    the last failure is available from pivx_last_error(ctx) (or
 *    pivx_last_error(NULL) for failures of pivx_init and of the host utilities).
 *  - No call exits the process, prints a message box or waits for input.
 *  - A handle owns one HIP stream on the device it was created on; all GPU work
 *    of that handle is ordered on that stream.  Calls on one handle must be
 *    serialised by the caller; distinct handles are independent.
 *  - Buffers passed in are caller-owned and only read/written during the call.
 *  - grid = 1 or 2 selects the coarse or fine grid (the reference's _p1/_p2).
 */
#ifndef PIV_XCORR2D_H
#define PIV_XCORR2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIVX_OK 0
#define PIVX_ERR_PARAM 1   /* a parameter the reference rejects, or an invalid argument */
#define PIVX_ERR_HIP 2     /* a HIP runtime failure (allocation, launch, copy) */
#define PIVX_ERR_STATE 3   /* call out of order (e.g. no frames uploaded yet) */
#define PIVX_ERR_IO 4      /* file could not be opened, read or written */
#define PIVX_ERR_FORMAT 5  /* malformed parameter, index or TIFF file */

typedef struct pivx_ctx pivx_ctx;

/* The 30 scalars of xcorr2d_multi::init (xcorr2d_multi_grid.cu:716-718), in
 * the same order and with the same meaning. */
typedef struct pivx_init_params {
  int nx, ny, nsec_cc;
  int ws_x1, ws_y1, spacing_x1, spacing_y1, npass1;
  int ws_x2, ws_y2, spacing_x2, spacing_y2, npass2;
  int wx1_median, wy1_median, wx2_median, wy2_median;
  int wx1_smooth, wy1_smooth, wx2_smooth, wy2_smooth;
  int npeak_save;
  float max_dispx_percent1, max_dispy_percent1, max_dispx_percent2, max_dispy_percent2;
  float dispx_lim1, dispx_lim2, dispy_lim1, dispy_lim2;
} pivx_init_params;

/* Derived geometry of one grid (xcorr2d_multi_grid.cu:751-894). */
typedef struct pivx_geometry {
  int ws_x, ws_y, spacing_x, spacing_y;
  int nx_vec, ny_vec, nt_vec;
  int border_x, border_y;      /* peak-search border, cu:751-754 */
  int npeak_slots;             /* reference candidate cap ws_x*ws_y/16, cu:768-769 */
  int left_x, left_y, right_x, right_y;   /* zero padding, cu:831-880 */
  int nx_aug, ny_aug;          /* padded ("aug") frame size, cu:882-891 */
  int nsec, neach, nlast;      /* reference memory sections, cu:805-829 */
} pivx_geometry;

/* Validation-stage parameters of one grid, as passed by main.cpp:299-318. */
typedef struct pivx_filter_params {
  int npass_median, min_nneigh;
  float residual_thr;   /* read but unused by the reference (cu:1451-1545) */
  float insert_thr;
  int npass_smooth;
} pivx_filter_params;

/* Diagnostics (no reference counterpart), per grid. */
typedef struct pivx_stats {
  long long overflow_windows[2];  /* last correlation pass: windows whose local-maximum count exceeds the reference's slot cap */
  long long outliers[2];          /* last median pass: vectors whose first candidate was not the one accepted */
  int fill_sweeps[2];             /* last hole filling: sweeps run, counted like the reference's loop */
  long long oob_taps[2];          /* cumulative since init: deformation taps beyond the reference's padded buffer (read as 0) */
  long long hole_list_drops;      /* cumulative since init: invalid vectors the hole filling could not list (0 unless
                                     the hole counter was left stale; pivx_synchronize returns PIVX_ERR_STATE when it
                                     finds new ones) */
  int holes_pending;              /* hole counter between filter calls: 0 whenever the stream is idle */
} pivx_stats;

/* ------------------------------------------------------------------ lifecycle */

/* xcorr2d_multi::init (cu:716-1269) + gpusetup::AssignGpu (gpuinfo.cu:62-66).
 * Applies the reference's acceptance rules exactly (4 <= ws <= 1024, ws % 4 == 0
 * for both grids, cu:896-920), plus rejections for inputs on which the
 * reference divides by zero or reads out of bounds (nsec_cc < 1, spacing < 1,
 * npeak_save < 1, image smaller than one spacing). */
int pivx_init(const pivx_init_params* p, int device_id, pivx_ctx** out);

/* xcorr2d_multi::freemem (cu:1976-2063).  Accepts NULL. */
int pivx_freemem(pivx_ctx* ctx);

/* Text of the last error on ctx (or of the last context-free failure when ctx
 * is NULL).  Never NULL; empty when there was no error. */
const char* pivx_last_error(const pivx_ctx* ctx);

int pivx_get_geometry(const pivx_ctx* ctx, int grid, pivx_geometry* g);

/* The geometry pivx_init derives for `grid` (cu:751-894), computed on the host
 * without a device (no reference counterpart: the reference derives it inside
 * init).  Applies pivx_init's parameter checks and returns their errors. */
int pivx_compute_geometry(const pivx_init_params* p, int grid, pivx_geometry* g);

/* ------------------------------------------------------------------ inputs */

/* xcorr2d_multi::prep_img_data (cu:1271-1282).  The reference takes the frames
 * from its public pinned members h_img1_org/h_img2_org (written by
 * main.cpp:264-266); here the caller passes host pointers to two nx*ny 8-bit
 * frames with row pitch `pitch` bytes (>= nx).  Copies them to the device. */
int pivx_prep_img_data(pivx_ctx* ctx, const uint8_t* img1, const uint8_t* img2, size_t pitch);

/* Same, from frames already resident in device memory of the handle's GPU
 * (device-to-device copy, no host staging). */
int pivx_prep_img_data_device(pivx_ctx* ctx, const uint8_t* d_img1, const uint8_t* d_img2, size_t pitch);

/* MHE contrast enhancement, replacing the mhe_gpu class (MHE7_4_gpu.h:7-40;
 * mhe_gpu::init at MHE7_4_gpu.cu:149-212, processsing at :214-253).  With
 * enable = 1 every frame later passed to pivx_prep_img_data[_device] is
 * enhanced on the GPU before use, as main.cpp:267-283 does when Parameter.dat
 * line 8 (ifmhe) is 1; ws and thresh are lines 9-10.  Pixels of the central
 * region of ws x ws tiles are mapped through their tile's transfer table,
 * pixels outside it become 0.  Rejects ws outside [1, min(nx, ny)] and
 * thresholds that leave no counted pixel per tile (the reference divides by
 * zero there).  enable = 0 switches the enhancement off. */
int pivx_set_mhe(pivx_ctx* ctx, int enable, int ws, float thresh);

/* One frame through the configured MHE (host in, host out, same pitch),
 * without touching the handle's frames: the d_enhanced result of
 * mhe_gpu::processsing (MHE7_4_gpu.cu:214-253) copied back as main.cpp:274
 * does.  Requires a prior pivx_set_mhe(ctx, 1, ...). */
int pivx_mhe_enhance(pivx_ctx* ctx, const uint8_t* img, uint8_t* out, size_t pitch);

/* ------------------------------------------------------------------ pipelined host I/O
 * No reference counterpart: the reference serialises disk reads, row-by-row
 * uploads, compute and file formatting per pair (main.cpp:227-376,
 * cu:1271-1282, 1865-1973).  These calls let a driver read the next pair and
 * write the previous pair's files while the GPU works. */

/* Page-locked host memory (so that uploads and result copies overlap host
 * work); NULL on failure. */
void* pivx_host_alloc(size_t bytes);
void pivx_host_free(void* p);

/* pivx_prep_img_data without the final wait: the uploads (and the MHE, if
 * enabled) are enqueued on the handle's stream.  The frames must stay
 * unchanged until the stream has passed them (pivx_synchronize, or
 * pivx_results_wait on a later fetch). */
int pivx_prep_img_data_async(pivx_ctx* ctx, const uint8_t* img1, const uint8_t* img2, size_t pitch);

/* Host snapshot of the fields write_vec_data_general reads (raw and filtered
 * fields of both grids and the grid-2 median-only field), in page-locked
 * memory.  fetch enqueues the copies on the handle's stream and returns;
 * wait blocks until they have landed; write formats one of the five files
 * (selection as in pivx_write_vec_data_general) from the snapshot only, so it
 * may run on any host thread while the handle is used for the next pair.  A
 * snapshot belongs to the handle that created it. */
typedef struct pivx_results pivx_results;
int pivx_results_create(pivx_ctx* ctx, pivx_results** out);
int pivx_results_fetch(pivx_ctx* ctx, pivx_results* r);
int pivx_results_wait(pivx_results* r);
int pivx_results_write(const pivx_results* r, const char* path, int selection, int crlf);
void pivx_results_destroy(pivx_results* r);

/* ------------------------------------------------------------------ passes */

/* xcorr2d_multi::Calc_cc_map_p1 (cu:1284-1449): one grid-1 correlation pass.
 * idx_pass == 0 zeroes the grid-1 predictor and correlates the undeformed
 * frames; idx_pass > 0 deforms frame 2 by the extended filtered field. */
int pivx_calc_cc_map_p1(pivx_ctx* ctx, int idx_pass);

/* xcorr2d_multi::MedianFilter_Smoothing_p1 (cu:1451-1545): normalised median
 * test with candidate replacement, hole filling (always), top-hat smoothing. */
int pivx_median_filter_smoothing_p1(pivx_ctx* ctx, int npass_median, int min_nneigh,
                                    float residual_thr, float insert_thr, int npass_smooth);

/* xcorr2d_multi::extend_filtered_vecfield_p1 (cu:1547-1576). */
int pivx_extend_filtered_vecfield_p1(pivx_ctx* ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the validation stage of pivx_median_filter_smoothing_p2
 * (median passes, optional hole filling, smoothing) on caller device buffers,
 * with no handle and no allocation.  Inputs: raw candidates
 * [nx_vec*ny_vec][npeak_save] (dx, dy, valid) in device memory of the current
 * device; outputs: the filtered field [nx_vec*ny_vec] in device memory.  The
 * other arguments mean what they mean for pivx_init (median and smoothing
 * windows) and pivx_median_filter_smoothing_p2.  Work memory is a static
 * device array of the library (about 4.6 MB per GPU, held until the process
 * ends) covering at most PIVX_NOALLOC_MAX_VECTORS vectors; larger fields, and
 * median windows of more than 64 vectors, are rejected with PIVX_ERR_PARAM.
 * All work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting.  Every call on one GPU shares that
 * array: calls that overlap in time on one GPU (other streams, other host
 * threads) race and must be serialised by the caller.  Errors are reported
 * through pivx_last_error(NULL). */
#define PIVX_NOALLOC_MAX_VECTORS 131072
int pivx_median_filter_smoothing_noalloc(const float* d_cand_x, const float* d_cand_y, const uint8_t* d_cand_valid,
                                         int nx_vec, int ny_vec, int npeak_save, int wx_median, int wy_median,
                                         int npass_median, int min_nneigh, float residual_thr, float insert_thr,
                                         int wx_smooth, int wy_smooth, int npass_smooth, int iffill, float* d_out_x,
                                         float* d_out_y, uint8_t* d_out_valid, void* stream);

/* xcorr2d_multi::Upsampling_vector_field (cu:1578-1587): bilinear grid-1 ->
 * grid-2 predictor. */
int pivx_upsampling_vector_field(pivx_ctx* ctx);

/* xcorr2d_multi::extend_filtered_vecfield_p2 (cu:1589-1618). */
int pivx_extend_filtered_vecfield_p2(pivx_ctx* ctx);

/* xcorr2d_multi::Calc_cc_map_p2 (cu:1620-1764): one grid-2 pass (always
 * deformed). */
int pivx_calc_cc_map_p2(pivx_ctx* ctx);

/* xcorr2d_multi::MedianFilter_Smoothing_p2 (cu:1766-1863); iffill selects hole
 * filling as in the reference. */
int pivx_median_filter_smoothing_p2(pivx_ctx* ctx, int npass_median, int min_nneigh,
                                    float residual_thr, float insert_thr, int npass_smooth,
                                    int iffill);

/* The per-pair pass sequence of main.cpp:288-375 without file output: grid-1
 * passes, upsampling, grid-2 passes, then the grid-2 "median-only" re-run
 * (written to the separate median-only field, see pivx_get_filtered).  All work
 * is enqueued on the handle's stream; no host synchronisation inside. */
int pivx_run_pair(pivx_ctx* ctx, const pivx_filter_params* grid1, const pivx_filter_params* grid2);

/* Graph mode (no reference counterpart): with on != 0, pivx_run_pair captures
 * the whole per-pair sequence once into a hipGraph on the handle's stream and
 * replays it (re-captured when the filter parameters change).  Results are
 * identical to the eager sequence. */
int pivx_set_graph_mode(pivx_ctx* ctx, int on);

/* Fused deformation (no reference counterpart).  Where the grid-1 windows do
 * not overlap (spacing >= window) and take the wave-per-window correlation
 * kernel (one side above 32, both at most 64), frame 2 can be deformed inside
 * the correlation kernel's gather instead of into the deformed plane
 * (Kernel_Window_Deformation_2d, cu:608-652, then a read of the plane); the
 * correlation results are bit-identical, and the out-of-buffer tap count
 * (pivx_stats) then covers the window pixels only (aug pixels outside every
 * window reach no result and are not deformed).  mode 0:
 * never; 1 (default): inside pivx_run_pair only; 2: also in
 * pivx_calc_cc_map_p1.  A fused pass leaves the deformed plane of grid 1
 * (pivx_get_deformed) as it was. */
int pivx_set_fused_deform(pivx_ctx* ctx, int mode);

/* Precision form of the correlation (no reference counterpart; replaces the
 * FP32 cuFFT transforms of cu:922-931, 1334-1350 for the handle's passes):
 * 0 = FP32 register FFTs (default), or a DFT-by-matrix contraction on the
 * matrix cores (F_H z F_W, conj-multiply, F_H^H P F_W^H) with 1 = FP16 or
 * 2 = BF16 operands, 3 = FP16 or 4 = BF16 operands split into hi + lo parts
 * (three products each, about FP32 accuracy); FP32 accumulation in all.
 * Modes 1-4 also evaluate the Gaussian fit of nearly flat peaks and the
 * median test's normalised residuals in FP64.  Windows of both grids must be
 * one of 16x16, 24x24, 32x32, 48x48, 64x64, 64x32, 32x64, 32x16, 16x32 or
 * 48x24.  The fused grid-1 deformation is not used while a mode other than 0
 * is set. */
int pivx_set_correlation_precision(pivx_ctx* ctx, int mode);

/* Wait for all work of the handle. */
int pivx_synchronize(pivx_ctx* ctx);

/* ------------------------------------------------------------------ outputs */

/* xcorr2d_multi::write_vec_data_general (cu:1865-1973): selection 0 = grid-1
 * raw candidates, 1 = grid-1 filtered, 2 = grid-2 raw, 3 = grid-2 filtered.
 * crlf != 0 writes CRLF line ends (the reference's Windows text mode). */
int pivx_write_vec_data_general(pivx_ctx* ctx, const char* path, int selection, int crlf);

int pivx_get_dims(const pivx_ctx* ctx, int grid, int* nx_vec, int* ny_vec);
/* Vector centres as written by the reference (cu:1237-1254). */
int pivx_get_grid_coords(const pivx_ctx* ctx, int grid, float* x, float* y);
/* Raw candidates [nt_vec][npeak_save] (d_disp_{x,y}_p*, d_valid_p*). */
int pivx_get_raw(pivx_ctx* ctx, int grid, float* dx, float* dy, uint8_t* valid);
/* Filtered field [nt_vec] (d_disp_{x,y}_filter_p*, d_valid_filter_p*). */
int pivx_get_filtered(pivx_ctx* ctx, int grid, float* dx, float* dy, uint8_t* valid);
/* Integer peak data [nt_vec][npeak_save]: correlation value and row-major pixel
 * index j*ws_x+i in the fftshifted window plane, and the uncapped local-maximum
 * count per window [nt_vec]. Any pointer may be NULL. */
int pivx_get_peaks(pivx_ctx* ctx, int grid, float* value, int* index, int* ncand);
/* The deformed frame 2 of the last deformed pass of `grid` that wrote it (a
 * pass with fused deformation does not, see pivx_set_fused_deform),
 * [ny_aug][nx_aug]. */
int pivx_get_deformed(pivx_ctx* ctx, int grid, float* out);
int pivx_get_stats(pivx_ctx* ctx, pivx_stats* st);

/* Stage-test hooks: overwrite the filtered field (the predictor) or the raw
 * candidates of a grid with caller data, so each stage can be fed identical
 * inputs.  valid may be NULL (all valid). */
int pivx_set_filtered(pivx_ctx* ctx, int grid, const float* dx, const float* dy, const uint8_t* valid);
int pivx_set_raw(pivx_ctx* ctx, int grid, const float* dx, const float* dy, const uint8_t* valid);

/* Average device time (ms) per launch of a stage since timing was enabled,
 * measured with HIP events on the handle's stream.
 * stage: 0 deformation, 1 correlation, 2 median, 3 fill, 4 smoothing,
 * 5 one whole pair replayed as a graph, 6 MHE of one frame (for 5 and 6 the
 * grid argument is ignored: use 1). */
int pivx_enable_timing(pivx_ctx* ctx, int enable);
int pivx_get_stage_ms(pivx_ctx* ctx, int grid, int stage, double* ms, int* launches);

/* Measurement helper (no reference counterpart): `reps` streaming copies of
 * `bytes` between two scratch buffers on the handle's stream; returns the
 * achieved read+write bandwidth in GB/s.  Allocates and frees its buffers. */
int pivx_bandwidth_probe(pivx_ctx* ctx, size_t bytes, int reps, double* gbps);

/* ------------------------------------------------------------------ host utilities */

/* Parameter.dat (main.cpp:93-158): 50 positional lines, CR/LF tolerant; numeric
 * lines parsed like _wtoi/_wtof (leading number, 0 on garbage).  Fewer than 50
 * lines -> PIVX_ERR_FORMAT ("Parameter missing"). */
typedef struct pivx_parameter_file {
  char prefix_f1[1024], suffix_f1[256], prefix_f2[1024], suffix_f2[256];
  char dir_index[1024], folder_result[1024];
  int device_id, ifmhe, mhe_ws;
  float mhe_thresh;
  pivx_init_params init;           /* nx .. dispy_lim2 */
  int npass1_median, min_nneigh_p1;
  float residual_threshold_p1, insert_threshold_p1;
  int npass1_smooth;
  int npass2_median, min_nneigh_p2;
  float residual_threshold_p2, insert_threshold_p2;
  int npass2_smooth;
} pivx_parameter_file;
int pivx_parse_parameter_file(const char* path, pivx_parameter_file* out);

/* Pair-index file (main.cpp:170-190): one integer per line until EOF.  Returns
 * the number of indices (may exceed max_n; only max_n are stored), or -1. */
int pivx_read_pair_index(const char* path, int* idx, int max_n);

/* The five output names of main.cpp:327-376 for pair `pair`; which = 0 (grid-1
 * raw), 1 (grid-1 filtered), 2 (grid-2 raw), 3 (grid-2 filtered), 4 (grid-2
 * median-only).  Writes folder + name into out (size cap). */
int pivx_output_name(const pivx_parameter_file* pf, int pair, int which, char* out, size_t cap);

/* The Tecplot text format of write_vec_data_general (cu:1875-1886) for
 * caller arrays (host only): raw != 0 writes x y dx1 dy1 .. dx4 dy4 from
 * [nt][K] candidate arrays (columns for k >= K are written as 0.00), else
 * x y dx dy from [nt] arrays. */
int pivx_write_tecplot(const char* path, int nx_vec, int ny_vec, const float* gx, const float* gy, const float* dx,
                       const float* dy, int K, int raw, int crlf);

/* 8-bit TIFF (ImageIOPkg.cpp:124-429): header fields and strip read. */
int pivx_tiff_info(const char* path, int* width, int* height, int* bits_per_sample);
int pivx_tiff_read8(const char* path, uint8_t* buf, size_t cap);
/* Minimal baseline 8-bit writer (one strip) used to build test inputs. */
int pivx_tiff_write8(const char* path, const uint8_t* buf, int width, int height);

/* Seeded synthetic particle-image pair (no reference counterpart; the
 * reference's inputs are not available).  flow: 0 uniform (a0,a1), 1
 * Lamb-Oseen vortex at the image centre (a0 = core radius, a1 = peak swirl
 * px) plus linear shear a2 px/px, 2 sum of 8 random Fourier modes with maximum
 * displacement a0.  Frame 2 renders every particle displaced by the flow at its
 * frame-1 position. */
typedef struct pivx_synth_params {
  int nx, ny;
  unsigned long long seed;
  float density;       /* particles per pixel */
  float sigma;         /* Gaussian particle sigma, px */
  float imin, imax;    /* peak intensity range */
  float background, noise_sigma;
  int flow;
  float a0, a1, a2;
} pivx_synth_params;
int pivx_synth_pair(const pivx_synth_params* sp, uint8_t* img1, uint8_t* img2);
/* The synthetic flow's displacement at n points. */
int pivx_synth_flow(const pivx_synth_params* sp, const float* x, const float* y, int n,
                    float* u, float* v);

/* Device rendering of a synthetic pair straight into the handle's frames (no
 * reference counterpart): the particle model and flows of pivx_synth_pair
 * with counter-based random numbers, so frames differ from the host
 * generator's but follow the same flow (pivx_synth_flow) and render
 * identically on every call.  sp->nx, sp->ny must equal the handle's frame
 * size.  Enqueued on the handle's stream; the first call allocates two
 * nx*ny 32-bit accumulators. */
int pivx_synth_pair_device(pivx_ctx* ctx, const pivx_synth_params* sp);

/* ------------------------------------------------------------------ learned grid-1 predictor
 * No reference counterpart.  A 6-layer 3x3 convolutional network (channels
 * 2-32-64-64-64-64-2, strides 2,2,2,1,1,1, leaky ReLU 0.1, 131,042
 * parameters, bfloat16 weights and activations with FP32 accumulation on the
 * matrix cores, FP32 master weights) maps the frame pair averaged over 4x4
 * pixel blocks to the grid-1 displacement field, replacing the grid-1 passes
 * of main.cpp:289-303 (Calc_cc_map_p1 / MedianFilter_Smoothing_p1) for a
 * grid-1 spacing of 32x32 px.  It is trained on the GPU, by the project's own
 * forward, backward and Adam kernels, on synthetic pairs labelled by the plain
 * grid-1 passes. */
typedef struct pivx_cnn pivx_cnn;
int pivx_cnn_param_count(void);
int pivx_cnn_create(int device_id, pivx_cnn** out);
int pivx_cnn_destroy(pivx_cnn* net);
/* He-normal weights (last layer scaled by 0.1), zero biases, from a seed. */
int pivx_cnn_init_params(pivx_cnn* net, unsigned long long seed);
/* FP32 master parameters, per layer W [cout][3*3][cin] then b [cout]. */
int pivx_cnn_set_params(pivx_cnn* net, const float* params, int n);
int pivx_cnn_get_params(pivx_cnn* net, float* params, int n);
/* Host-side test entry points on a batch: x0 [batch][h][w][2] network input
 * (rounded to bfloat16 on upload), out and labels [batch][2][(h/8)*(w/8)];
 * w must be a multiple of 128 and h of 8.  pivx_cnn_gradient leaves the
 * gradient on the device for pivx_cnn_adam_step and copies it to grad (if
 * not NULL, pivx_cnn_param_count() floats). */
int pivx_cnn_forward(pivx_cnn* net, const float* x0, int batch, int h, int w, float* out);
int pivx_cnn_gradient(pivx_cnn* net, const float* x0, const float* labels, int batch, int h, int w, float* loss,
                      float* grad);
int pivx_cnn_adam_step(pivx_cnn* net, float lr);

/* Training on synthetic pairs.  Pair i is rendered on the device with seed
 * synth.seed + i into ctx's frames (ctx frame size = synth.nx x synth.ny,
 * grid 1 at spacing 32x32 over nx a multiple of 512 and ny of 32), labelled
 * by ctx's plain grid-1 passes with the `grid1` filter, and stored as network
 * input; then `epochs` passes of minibatch Adam (cosine-decayed learning rate)
 * over a seeded shuffle. */
typedef struct pivx_cnn_train_params {
  int npairs, epochs, batch;
  float lr;
  unsigned long long shuffle_seed;
  pivx_synth_params synth;
  pivx_filter_params grid1;
} pivx_cnn_train_params;
typedef struct pivx_cnn_train_report {
  double gen_seconds, train_seconds;   /* wall time of labelling and of training */
  int steps;
  float first_epoch_loss, last_epoch_loss;   /* mean squared error, px^2 per component */
} pivx_cnn_train_report;
/* epoch_loss (optional): mean loss of every epoch. */
int pivx_cnn_train(pivx_ctx* ctx, pivx_cnn* net, const pivx_cnn_train_params* tp, float* epoch_loss,
                   pivx_cnn_train_report* report);

/* The per-pair sequence of pivx_run_pair with the grid-1 passes replaced by
 * the network: the network field becomes the grid-1 field, the normalised
 * median test of main.cpp's grid-1 filter (its window, insert threshold and
 * minimum neighbour count from `grid1`) counts the vectors it rejects, and
 * when more than max_outlier_fraction of them are rejected the plain grid-1
 * passes run instead (decided on the device; the grid-1 kernels are launched
 * either way and return at once when not needed; with max_outlier_fraction
 * >= 1 the fallback cannot trigger and is not launched).  Then upsampling and the
 * grid-2 passes as in pivx_run_pair.  Graph mode (pivx_set_graph_mode) applies.
 * The network's parameters must not change while pairs using it are queued. */
int pivx_run_pair_learned(pivx_ctx* ctx, const pivx_cnn* net, const pivx_filter_params* grid1,
                          const pivx_filter_params* grid2, float max_outlier_fraction);
/* Of the last learned pair: vectors of the network field rejected by the
 * median test, and 1 when the plain grid-1 passes ran instead. */
int pivx_get_learned_stats(pivx_ctx* ctx, int* outliers, int* fallback);

/* ------------------------------------------------------------------ batched passes
 * No reference counterpart (the reference runs one pair at a time,
 * main.cpp:248-380).  A batch runs the pivx_run_pair sequence of many handles
 * (pairs of uneven frame size) at once: each correlation pass is one
 * persistent work-queue launch over the windows of all pairs, on the batch's
 * own stream, and the per-pair stages run on the handles' streams, joined by
 * events.  Results are bit-identical to pivx_run_pair on each handle.  The
 * handles keep their buffers and getters; all must share the device, window
 * sizes of both grids (a shape with a batched kernel: the wave-pair shapes
 * 32x32, 24x24, 32x16, 16x32 and the wave-per-window shapes 64x64, 48x48,
 * 64x32, 32x64, 48x24), pass counts and Npeak_save, and use the FP32
 * correlation.  While a batch exists its handles must not run other work. */
typedef struct pivx_batch pivx_batch;
int pivx_batch_create(pivx_ctx* const* ctxs, int n, pivx_batch** out);
int pivx_batch_destroy(pivx_batch* b);
/* Graph mode: the whole batch sequence captured once into one hipGraph and
 * replayed (re-captured when the filter parameters change). */
int pivx_batch_set_graph_mode(pivx_batch* b, int on);
int pivx_batch_run(pivx_batch* b, const pivx_filter_params* grid1, const pivx_filter_params* grid2);
int pivx_batch_synchronize(pivx_batch* b);
/* Autotuner: times `reps` correlation passes of each grid for every
 * persistent grid size (4, 8, 16, 32 waves per compute unit) and claim size
 * (1, 4, 16 work items) and keeps the fastest.  Overwrites the handles' raw
 * candidates; run it before the real passes. */
int pivx_batch_autotune(pivx_batch* b, int reps);
int pivx_batch_get_tuning(const pivx_batch* b, int grid, int* waves_per_cu, int* items_per_claim, double* ms,
                          long long* items);

#ifdef __cplusplus
}
#endif
#endif /* PIV_XCORR2D_H */
