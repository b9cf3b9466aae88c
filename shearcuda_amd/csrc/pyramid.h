// Launch interface of the a trous pyramid kernels.
#pragma once

#include "shx_common.h"

namespace shx {

// Full FP64 filters, row-major LEN x LEN: which = 0 h0, 1 h1, 2 g0, 3 g1.
const double* maxflat_filter(int which, int* len);
// Largest |F - sum_i u_i f_i^T| over both filter pairs (FP64 factors).
double maxflat_factor_error();
// FP64 joint factors of pair 0 (h0, h1) or 1 (g0, g1) on a 19 x 19 grid:
// u (shared basis), a (first filter), b (second filter minus the centre
// impulse), each rank x len, row-major; F(r, c) = sum_i u_i[r] f_i[c].
bool maxflat_pair_fp64(int pair, int* len, int* rank, const double** u, const double** a, const double** b);
// True when both pairs have the rank the kernels are specialised for.
bool maxflat_ok();
// Masks of the DCT-domain pyramid: nscales + 1 planes of L x L in FP32,
// divided by L^2.  pair 0 (decomposition, h0 / h1) or 1 (reconstruction,
// g0 / g1).  Plane p < nscales: the band-pass of level lev = nscales - 1 - p,
// F1(2^lev) prod_{i<lev} F0(2^i); plane nscales: the lowpass
// prod_{i<nscales} F0(2^i).  F(m)[kr][kc] is the filter's response on the
// DCT-II grid at dilation m (row frequency kr, column frequency kc).
void dct_pyramid_masks(int L, int nscales, int pair, float* out);

// Per-image element strides of the three planes a level touches
// (dec: x, ylow, yhigh; rec: xlow, yhigh, out).  0 for a single image.
struct ImgStrides {
  long long a, b, c;
};

// Threshold applied to the rec level's xlow input: none, one value, or one
// per image in device memory (dev[img * stride]; stride 0: one for all);
// hard or soft (soft != 0).
struct LowThresh {
  int mode;  // 0 none, 1 v, 2 dev
  float v;
  const float* dev;
  int stride;
  int soft;
};

// ShrinkRule of a lowpass read.
__host__ __device__ inline int low_rule(const LowThresh& lt) { return lt.mode == 0 ? 0 : (lt.soft ? 2 : 1); }
__device__ inline float low_value(const LowThresh& lt, int img) {
  return lt.mode == 1 ? lt.v : (lt.mode == 2 ? lt.dev[(size_t)img * lt.stride] : 0.0f);
}

constexpr int PYR_VARIANTS = 3;  // tile shapes the autotuner chooses from

hipError_t launch_atrous_dec(const float* x, float* ylow, float* yhigh, const ImgStrides& st, int nimg, int L, int m,
                             hipStream_t s, int variant = 0);
hipError_t launch_atrous_rec(const float* xlow, const float* yhigh, float* out, const ImgStrides& st, int nimg, int L,
                             int m, const LowThresh& lt, float sx, float sy, hipStream_t s, int variant = 0);

// Adjoints of one level (gradients).  zext: scratch of the extended grid,
// side adjoint_ext_side(L, m), one plane (dec) or two planes (rec) per image,
// zstride floats between images.
//   dec level y_low = C_h0 E x, y_high = C_h1 E x:
//       gx = E^T (C_h0^T glow + C_h1^T ghigh);  strides {glow, ghigh, gx}
//   rec level out = C_g0 E x_low + C_g1 E y:
//       glow = E^T C_g0^T gout, ghigh = E^T C_g1^T gout;  strides {gout, glow, ghigh}
int adjoint_ext_side(int L, int m);
hipError_t launch_atrous_dec_adjoint(const float* glow, const float* ghigh, float* gx, float* zext,
                                     const ImgStrides& st, long long zstride, int nimg, int L, int m, hipStream_t s);
hipError_t launch_atrous_rec_adjoint(const float* gout, float* glow, float* ghigh, float* zext, const ImgStrides& st,
                                     long long zstride, int nimg, int L, int m, hipStream_t s);

}  // namespace shx
