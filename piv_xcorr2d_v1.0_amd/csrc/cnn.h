// Learned grid-1 predictor: a 6-layer 3x3 convolutional network that maps a
// 4x-downsampled frame pair to the grid-1 displacement field (64x64 vectors
// for a 2048x2048 pair at spacing 32).  No reference counterpart: it replaces
// the grid-1 passes of the reference's main loop (main.cpp:289-303,
// Calc_cc_map_p1 cu:1284-1449 and MedianFilter_Smoothing_p1 cu:1451-1545).
//
// Numerics: BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_16x16x32_bf16), FP32 master weights and Adam moments.
// Activations are stored NHWC so that 8 consecutive channels of one pixel are
// one 16-byte load, the K-slice one lane feeds to an MFMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pivx {

struct CnnLayer {
  int cin, cout, stride, act;  // act: leaky ReLU (slope kCnnLeak) after the layer
};
constexpr int kCnnLayers = 6;
constexpr CnnLayer kCnnArch[kCnnLayers] = {{2, 32, 2, 1},  {32, 64, 2, 1}, {64, 64, 2, 1},
                                           {64, 64, 1, 1}, {64, 64, 1, 1}, {64, 2, 1, 0}};
constexpr float kCnnLeak = 0.1f;
constexpr int kCnnDown = 4;    // input = frames averaged over 4x4 pixel blocks
constexpr int kCnnStride = 8;  // output cell = 8x8 input pixels = 32x32 frame pixels
// input normalisation: (block mean - kCnnMu) * kCnnInvSigma
constexpr float kCnnMu = 30.0f;
constexpr float kCnnInvSigma = 1.0f / 40.0f;

// GEMM K extent of a layer with c input channels (9 taps x c, padded to the
// 32-wide K step of the MFMA), and an M/N extent padded to 16 rows.
constexpr int cnn_kpad(int c) { return (9 * c + 31) / 32 * 32; }
constexpr int cnn_pad16(int c) { return (c + 15) / 16 * 16; }

// Lane-ordered weight layout of the fused inference kernels: for channel
// block m (16 output channels) and K step st (32 of k = tap * cin + ci), the
// 64 lanes' MFMA operands are one contiguous 1 KiB block, lane
// kg * 16 + col holding channel 16 m + col, k = 32 st + 8 kg .. + 7.  Same
// size as the forward layout (pad16(cout) x kpad(cin)), zero padded.
__host__ __device__ constexpr long long cnn_lane_index(int co, int k, int cin) {
  return (((long long)(co / 16) * (cnn_kpad(cin) / 32) + k / 32) * 64 + ((k % 32) / 8) * 16 + co % 16) * 8 + k % 8;
}

// Forward convolution of one layer over a batch.
struct ConvFwdArgs {
  const uint16_t* in;      // bf16 NHWC [batch][hin][win][cin]
  long long in_bstride;    // elements per sample of `in`
  const int* batch_index;  // if set: sample b reads in + batch_index[b] * in_bstride
  int batch, hin, win, hout, wout;
  const uint16_t* w;       // bf16 [pad16(cout)][kpad(cin)], k = tap * cin + ci, zero padded
  const float* bias;       // [cout]
  uint16_t* out;           // bf16 NHWC [batch][hout][wout][cout] (hidden layers)
  float* out_x;            // last layer: FP32 planes, sample b at out_x + b * out_bstride
  float* out_y;
  long long out_bstride;
  uint8_t* out_v;          // last layer, batch 1 (optional): validity flags of the field, set to 1
  float* ext_x;            // last layer, batch 1 (optional): the field's extension, border ext_b replicated
  float* ext_y;
  int ext_b;
};

// Layers 0..2 of one pair (batch 1) in one launch (k_conv_head): from the
// network input plane, or formed from the frames, to the layer-2 plane.
struct ConvHeadArgs {
  const uint16_t* in;        // network input, bf16 [8 ph][8 pw][2], or null with f1 set
  const uint8_t* f1; long long p1; const uint8_t* f2; long long p2;  // the frames (optional)
  int ph, pw;                // layer-2 plane (input / 8)
  const uint16_t* w0; const uint16_t* w1; const uint16_t* w2;  // lane-ordered weights (cnn_lane_index)
  const float* b0; const float* b1; const float* b2;           // biases
  uint16_t* out;             // layer-2 output, bf16 NHWC [ph][pw][64]
};

// Layers 3..5 of one pair (batch 1) in one launch (k_conv_tail): the two
// 64 -> 64 stride-1 layers and the 64 -> 2 field layer.
struct ConvTailArgs {
  const uint16_t* in;        // layer-2 output, bf16 NHWC [h][w][64]
  int h, w;                  // field size
  const uint16_t* w3; const uint16_t* w4; const uint16_t* w5;  // lane-ordered weights (cnn_lane_index)
  const float* b3; const float* b4; const float* b5;           // biases
  float* out_x; float* out_y;  // the field planes [h][w]
  uint8_t* out_v;            // optional: validity flags, set to 1
  float* ext_x; float* ext_y; int ext_b;  // optional: the field's border-replicated extension
};

// Gradient with respect to the layer input.
struct ConvBwdDataArgs {
  const uint16_t* gout;    // bf16 dL/d(layer output) NHWC [batch][hout][wout][cout]
  const uint16_t* post;    // bf16 layer output (for the leaky ReLU slope) or null
  const uint16_t* wt;      // bf16 [pad16(cin)][kpad(cout)], k = tap * cout + co
  uint16_t* gin;           // bf16 dL/d(layer input) NHWC [batch][hin][win][cin]
  int batch, hin, win, hout, wout;
};

// Gradient with respect to the weights: per chunk of output pixels, partial
// sums [chunk][pad16(cout)][kpad(cin)] (reduced in a fixed order afterwards).
struct ConvBwdWArgs {
  const uint16_t* in; long long in_bstride; const int* batch_index;
  const uint16_t* gout; const uint16_t* post;
  float* partial;          // [nchunks][pad16(cout)][kpad(cin)]
  float* partial_b;        // [nchunks][cout] bias partial sums
  int batch, hin, win, hout, wout;
  int chunk_px;            // output pixels per chunk, multiple of 32
};

int cnn_param_count();
int cnn_weight_offset(int layer);  // master layout per layer: W [cout][9][cin], then b [cout]

hipError_t launch_conv_fwd(int layer, const ConvFwdArgs& a, hipStream_t s);
hipError_t launch_conv_head(const ConvHeadArgs& a, hipStream_t s);
hipError_t launch_conv_tail(const ConvTailArgs& a, hipStream_t s);
hipError_t launch_conv_bwd_data(int layer, const ConvBwdDataArgs& a, hipStream_t s);
hipError_t launch_conv_bwd_w(int layer, const ConvBwdWArgs& a, int nchunks, hipStream_t s);
// grad_w [cout][9][cin] and grad_b [cout] of every layer in one launch, from
// per-layer partial buffers; grad is the whole parameter gradient (layer l at
// cnn_weight_offset(l)).
struct WgradReduceLayer {
  const float* partial; const float* partial_b; float* grad;
  int nchunks, cin, cout, kpi, mrows, first;
};
struct WgradReduceAll {
  WgradReduceLayer layer[kCnnLayers];
};
hipError_t launch_wgrad_reduce_all(const float* const* partial, const float* const* partial_b, const int* nchunks,
                                   float* grad, hipStream_t s);
// Mean squared error of the FP32 output planes against labels [n][2][hw];
// writes dL/d(out) as bf16 NHWC [batch][hw][2] and the per-block partial
// sums of squared errors into loss_part (nblocks = cnn_loss_blocks(...)).
int cnn_loss_blocks(int batch, int hw);
hipError_t launch_mse(const float* out, long long out_bstride, const float* labels, const int* batch_index, int batch,
                      int hw, uint16_t* gout, float* loss_part, hipStream_t s);
// loss = sum(loss_part[0..n)) / count, one workgroup, fixed order.
hipError_t launch_loss_final(const float* loss_part, int n, double count, float* loss, hipStream_t s);
// One Adam step on the FP32 master parameters (t = step number, from 1).
hipError_t launch_adam(float* master, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                       float eps, long long t, hipStream_t s);
// bf16 GEMM layouts of every layer's weights from the FP32 master: forward
// (wf), transposed for the data gradient (wt) and lane-ordered (wl).
hipError_t launch_pack(const float* master, uint16_t* const* wf, uint16_t* const* wt, uint16_t* const* wl,
                       hipStream_t s);
// Frames (u8, row pitches p1, p2) -> bf16 NHWC [ny/4][nx/4][2] network input.
hipError_t launch_downsample(const uint8_t* f1, long long p1, const uint8_t* f2, long long p2, int nx, int ny,
                             uint16_t* out, hipStream_t s);
// gate = 1 when more than max_outliers vectors have accepted < 0 (the
// normalised median test rejected the network's vector), else 0; *count
// receives the number.
hipError_t launch_gate(const int8_t* accepted, int n, int max_outliers, int* gate, int* count, hipStream_t s);
// fv[0..n) = 1
hipError_t launch_fill_u8(uint8_t* p, int n, uint8_t v, hipStream_t s);

}  // namespace pivx
