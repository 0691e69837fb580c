// Device-kernel argument blocks and launchers of the PIV engine (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pivx {

// Deformation fused into the gather of the wave-per-window kernel, for
// windows that do not overlap (each aug pixel then belongs to one window and
// is deformed once): frame 2 is sampled at the displaced position of every
// window pixel with the arithmetic of k_deform instead of read from a
// materialized plane.  col == nullptr: off.
struct FusedDef {
  const float4* col;      // per aug column: (ax, 1 - ax, ixf as int bits, (float)idx)
  const float4* row;      // per aug row: (ay, 1 - ay, iyf as int bits, (float)idy)
  const float* ex_x; const float* ex_y; int nxe;  // extended field of the grid
  int hfws_x, hfws_y, nx_aug, ny_aug;             // for the out-of-buffer tap count
  unsigned long long* oob;                        // cumulative out-of-buffer tap count (null: not counted)
};

// Fused correlate-and-peak pass over all windows of one grid.
struct XcArgs {
  const uint8_t* f1;      // frame 1, image coordinates, row pitch pitch1
  const uint8_t* f2u8;    // frame 2 (undeformed pass) or nullptr
  const float* f2def;     // deformed frame 2 in padded ("aug") coordinates or nullptr
  long long pitch1, pitch2;
  int def_pitch;          // row length of f2def (nx_aug)
  int nx, ny;             // image size
  int left_x, left_y;     // aug coordinate = image coordinate + left
  int nx_vec, ny_vec, sp_x, sp_y;
  int bx, by, K;          // search border, Npeak_save
  float lx1, lx2, ly1, ly2;
  const float* pred_x;    // predictor (filtered field), nullptr = 0
  const float* pred_y;
  float* disp_x; float* disp_y; uint8_t* valid;   // [nt_vec][K]
  float* pk_val; int* pk_idx;                     // [nt_vec][K]
  int* ncand;                                     // [nt_vec] uncapped local-maximum count
  int chunks_x;           // blocks per vector row (fast path)
  float* zero_x;          // if set: zeroed at every window (first-pass predictor reset)
  float* zero_y;
  FusedDef fdef;          // frame 2 deformed in the gather (wave-per-window kernel only)
  const int* gate;         // if set: the launch does nothing unless *gate != 0
};

// f2 (and XcArgs::f2u8 when frame 2 is deformed in the gather) points at
// pixel (0, 0) of a frame stored inside a zero border: the bytes of aug
// coordinates [-hfws_x, nx_aug + hfws_x) x [-hfws_y, ny_aug + hfws_y + 1)
// (image = aug - left) are readable and 0 outside the image.  Taps of cells
// whose field samples stay below the out-of-buffer guard are read from it
// without checks.
struct DefArgs {
  float* out; int nx_aug, ny_aug;
  const uint8_t* f2; long long pitch; int nx, ny; int left_x, left_y;
  int hfws_x, hfws_y;     // the reference's padded buffer: tap count and the zero border
  const float* ex_x; const float* ex_y; int nxe, nye;
  float x_st, y_st, spx, spy;
  unsigned long long* oob;   // cumulative count of taps beyond the reference's padded buffer
  const int* gate;         // if set: the launch does nothing unless *gate != 0
};

struct MedArgs {
  const float* sx; const float* sy; const uint8_t* sv; int sstride;   // neighbourhood field
  const float* cx; const float* cy; const uint8_t* cv; int K;         // raw candidates
  float* dx; float* dy; uint8_t* dv;                                  // output filtered field
  float* dx2; float* dy2; uint8_t* dv2;  // optional second copy of the output (null: none)
  int nx_vec, ny_vec, wx, wy, hfwx, hfwy, min_nneigh;
  float thr, eps;
  int8_t* accepted;        // [nt] index of the accepted candidate, -1 if none (may be null)
  float* scratch;          // generic path only: 4*wx*wy floats per vector
  int* hole_list;          // if set: indices of vectors left invalid are appended here
  int* nholes;             //   (count in *nholes, which must be 0 before the pass)
  int hole_cap;            //   capacity of hole_list; appends beyond it are not stored (k_fill counts them)
  int* reset_count;        // if set: the launch writes 0 here (the hole counter, ahead of a later listing pass)
  int fp64;                // normalised residuals in FP64 (precision form)
  const int* gate;         // if set: the launch does nothing unless *gate != 0
  // optional outlier gate (k_median2 shapes): the pass counts the vectors it
  // leaves invalid into *gate_count and sets *gate_out = (count > gate_max);
  // gate_acc and gate_ticket are work counters, 0 between launches
  int* gate_out; int* gate_count; int gate_max; int* gate_acc; unsigned int* gate_ticket;
};

struct FillArgs {
  float* dx; float* dy; uint8_t* dv;
  int nx_vec, ny_vec;
  int* list_a; int* list_b; float* tmp_x; float* tmp_y; uint8_t* tmp_f;
  int* nholes;             // hole counter, 0 between fills
  int list_cap;            // capacity of list_a and list_b
  unsigned int* drops;     // holes counted beyond list_cap (never filled) are added here
  int* sweeps;
  int prelisted;           // holes already in list_a / *nholes (appended by the median pass)
  const int* gate;         // if set: the launch does nothing unless *gate != 0
};

struct SmoothArgs {
  const float* sx; const float* sy; float* dx; float* dy;
  int nx_vec, ny_vec, wx, wy, hfwx, hfwy;
  float* ex; float* ey; int border;   // if ex is set: also write the border-replicated extension
  const int* gate;         // if set: the launch does nothing unless *gate != 0
};

struct PrepArgs {
  const float* cx; const float* cy; const uint8_t* cv; int K;
  float* dx; float* dy; uint8_t* dv; int nt;
  int* reset_count;        // if set: 0 is written here
};

// A whole filter sequence (median passes, hole listing and filling,
// smoothing) run by one workgroup in one launch: the steps of the launch
// sequence, in order, with a workgroup barrier between them.  Used for the
// learned pipeline's gated grid-1 fallback, whose launches mostly do nothing:
// one launch instead of one per step.  Each step runs the per-vector code of
// its own kernel, so the results are those of the launch sequence.
enum FilterStepKind : int { FS_PREP = 0, FS_ZERO = 1, FS_MEDIAN = 2, FS_FILL_LIST = 3, FS_FILL = 4, FS_SMOOTH = 5 };
struct FilterStep {
  int kind;
  union {
    PrepArgs prep;
    int* zero;             // FS_ZERO: one int set to 0
    MedArgs med;           // k_median2 window shapes only (3x3, 5x5, 7x7)
    FillArgs fill;
    SmoothArgs smooth;
  };
};
constexpr int kFilterSeqSteps = 12;
struct FilterSeq {
  FilterStep step[kFilterSeqSteps];
  int nsteps;
  int wxm, wym;            // median window of the FS_MEDIAN steps
  const int* gate;         // if set: the launch does nothing unless *gate != 0
};
static_assert(sizeof(FilterSeq) <= 4096, "kernel arguments are limited to 4 KB");

// MHE contrast enhancement of one u8 frame (MHE7_4_gpu.cu).  The central
// tiles_x*ws by tiles_y*ws region (origin xs, ys) is split into ws x ws tiles;
// every pixel there is mapped through its tile's transfer table, every pixel
// outside it becomes 0.
struct MheArgs {
  const uint8_t* in; uint8_t* out; int nx, ny;
  int ws, xs, ys, tiles_x, tiles_y;
  int ntrash, nnontrash;
  long long out_pitch;     // row pitch of out (in has row pitch nx)
  uint32_t* hist;          // [(tiles_y+2)*(tiles_x+2)][256]; the border ring is kept at 0
};

// Launchers (return hipGetLastError() of the launch).
hipError_t launch_xcorr(const XcArgs& a, int ws_x, int ws_y, float2* gscratch, size_t gscratch_windows,
                        const float2* twx, const float2* twy, hipStream_t s);
bool xcorr_fast_supported(int ws_x, int ws_y);
// Precision form of the correlation: DFT-by-matrix on the matrix cores for
// windows up to 64x64 (sides multiples of 8).  Tables of the constant DFT
// matrices (bf16 or fp16 bits, hi and lo parts), laid out as k_xcorr_dft
// reads them: a1 / a3 [pad16(H)][2 pad16(H)] row-major, b2 / b4
// [pad16(W)][2 pad16(W)] (K contiguous per output column).
struct DftTables {
  const uint16_t* a1h[2]; const uint16_t* a1l[2];   // [C S], [-S C]   (step 1: Tr, Ti)
  const uint16_t* b2h[2]; const uint16_t* b2l[2];   // [C; S], [-S; C] (step 2: Zr, Zi)
  const uint16_t* a3h[2]; const uint16_t* a3l[2];   // [C -S], [S C]   (step 3: Ur, Ui)
  const uint16_t* b4h; const uint16_t* b4l;         // [C; -S]         (step 4: cc)
  // 32x32 windows: C and S in the MFMA lane layout of the wave-per-window
  // kernel, [C, S][tile t][hi, lo][lane][8]: M[16t + lane % 16][8 (lane / 16) + j]
  const uint16_t* lane32;
  // 64x64 windows: [C, S][hi, lo][tile t][k-step ks][lane][8]: M[16t + lane % 16][32 ks + 8 (lane / 16) + j]
  const uint16_t* lane64;
};
// mode 1 FP16, 2 BF16, 3 FP16 hi+lo, 4 BF16 hi+lo; shapes: xcorr_dft_supported
bool xcorr_dft_supported(int ws_x, int ws_y);
hipError_t launch_xcorr_dft(const XcArgs& a, int ws_x, int ws_y, int mode, const DftTables& t, hipStream_t s);
// Persistent work-queue correlation pass over many pairs (same window shape).
struct BatchQueue {
  const XcArgs* args;      // [npairs] per-pair arguments of the pass (device); chunks_x set for pair shapes
  const int* item_start;   // [npairs + 1] first work item of every pair; item_start[npairs] = total
  int npairs, total;
  int* counter;            // [8] claims taken per XCD range (zeroed by the launcher's first kernel)
  int chunk;               // items per claim (interleaved with other workgroups' claims)
};
bool xcorr_batch_supported(int ws_x, int ws_y);
// work items of one pair of the given grid size (chunks_x: window pairs per vector row, 0 for single-window shapes)
long long xcorr_batch_items(int ws_x, int ws_y, int nx_vec, int ny_vec, int* chunks_x);
hipError_t launch_xcorr_batch(int ws_x, int ws_y, const BatchQueue& q, int workgroups, hipStream_t s);
hipError_t launch_deform(const DefArgs& a, hipStream_t s);
// The per-column and per-row interpolation terms of the deformation (as
// k_deform computes them) into col[nx_aug] and row[ny_aug], for FusedDef.
hipError_t launch_deform_terms(const DefArgs& a, float4* col, float4* row, hipStream_t s);
// Whether launch_xcorr runs windows of this shape on the wave-per-window
// kernel, the one that accepts FusedDef.
bool xcorr_fuses_deform(int ws_x, int ws_y);
hipError_t launch_extend(const float* fx, const float* fy, float* ex, float* ey, int nx, int ny, int border,
                         hipStream_t s, const int* gate = nullptr);
hipError_t launch_upsample(float* dx2, float* dy2, const float* ex1x, const float* ex1y, int nxe1, int nye1,
                           float x_st1, float y_st1, float sp1x, float sp1y, int sp2x, int sp2y, int nx2, int ny2,
                           float* ex2x, float* ex2y, int border, hipStream_t s);
hipError_t launch_median(const MedArgs& a, hipStream_t s);
hipError_t launch_fill(const FillArgs& a, hipStream_t s);
hipError_t launch_smooth(const SmoothArgs& a, hipStream_t s);
// filter := candidate 0 (median_filter_data_prep, cu:484-495); reset_count,
// if set, is zeroed by the launch.
hipError_t launch_prep(const float* cx, const float* cy, const uint8_t* cv, int K, float* dx, float* dy, uint8_t* dv,
                       int nt, int* reset_count, hipStream_t s, const int* gate = nullptr);
// p[0..n) = 0 by a kernel (no memset node in a captured pair graph).
hipError_t launch_zero_ints(int* p, int n, hipStream_t s, const int* gate = nullptr);
// One workgroup runs the steps of q (median windows 3x3, 5x5 or 7x7 only).
bool filter_seq_supported(int wxm, int wym);
// A gated correlation pass whose last workgroup then runs the filter
// sequence q (the learned pipeline's fallback, 64x64 windows); ticket: one
// counter, 0 between launches.
bool xcorr_gated_filter_supported(int ws_x, int ws_y, int wxm, int wym);
hipError_t launch_xcorr_gated_filter(const XcArgs& a, int ws_x, int ws_y, const FilterSeq& q, unsigned int* ticket,
                                     hipStream_t s);
// Whether launch_median honours MedArgs::gate_out (the k_median2 shapes).
bool median_gate_supported(int wxm, int wym);
hipError_t launch_filter_seq(const FilterSeq& q, hipStream_t s);
hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_mhe(const MheArgs& a, hipStream_t s);
// Device rendering of one synthetic pair (synth.hip) into frames f1, f2 (row
// pitches p1, p2); acc1, acc2 are nx * ny zeroed accumulators, left zeroed.
struct SynthRender;
hipError_t launch_synth(const SynthRender& r, unsigned int* acc1, unsigned int* acc2, uint8_t* f1, long long p1,
                        uint8_t* f2, long long p2, hipStream_t s);

}  // namespace pivx
