// Undecimated periodic 2D wavelet transform (analysis / synthesis passes).
#pragma once

#include "shx_common.h"

namespace shx {

constexpr int RDWT_MAX_TAPS = 32;

// Per-tap coefficients of the two filters one pass applies.
struct RdwtFilt {
  int lh;
  float f0[RDWT_MAX_TAPS];
  float f1[RDWT_MAX_TAPS];
};

// x (nimg x m x n) -> yl (nimg x m x n), yh (nimg x m x 3nL); f holds the
// analysis taps (a0, a1); tmp: 2*m*n floats per image.
hipError_t rdwt_analysis(const float* x, float* yl, float* yh, float* tmp, int m, int n, int L, const RdwtFilt& f,
                         int nimg, hipStream_t s);
// yl, yh -> x; f holds the synthesis taps (g0, g1).
hipError_t rdwt_synthesis(const float* yl, const float* yh, float* x, float* tmp, int m, int n, int L,
                          const RdwtFilt& f, int nimg, hipStream_t s);

}  // namespace shx
