// Host-side file formats of the PIV path (no GPU code).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/piv_xcorr2d.h"

namespace pivx {

// Process-wide text of the last context-free failure (pivx_last_error(NULL)).
void set_global_error(const char* msg);
const char* global_error();

// Tecplot ASCII vector file of write_vec_data_general (xcorr2d_multi_grid.cu:1865-1973).
// raw: columns x y dx1 dy1 .. dx4 dy4 from [nt][K] arrays (missing candidates
// K < 4 are written as 0.00); otherwise x y dx dy from [nt] arrays.
int write_tecplot(const char* path, int nx_vec, int ny_vec, const float* gx, const float* gy, const float* dx,
                  const float* dy, int K, bool raw, bool crlf);

// The flow of a synthetic pair (kind, modes, scale) as the host generator
// builds it from its parameters.
struct SynthFlow;
void synth_flow_of(const pivx_synth_params* sp, SynthFlow* out);

}  // namespace pivx
