// 3D FFT line passes and Fourier-domain band-pass masks for the 3D shearlet
// path (gfx950).
//
// A real volume of side n is transformed on a grid of side N >= n (zero
// padded) into a half spectrum halved along memory axis 0:
//   spectrum[k0][k1][k2], k0 in [0, N/2], k1, k2 in [0, N), k2 contiguous.
// Every transform is three line passes:
//   pencil  axis 0, real <-> half complex, two real columns per complex FFT
//   cols    axis 1, complex, a tile of adjacent columns per block
//   rows    axis 2, complex, contiguous lines
// Each pass reads n_in valid inputs placed at FFT index o_in + m (the rest are
// zero) and writes the FFT outputs with index k in [o_out, o_out + n_out) to
// position k - o_out.  All passes are unnormalised.
#pragma once

#include "shx_common.h"

namespace shx {

// Lines along the contiguous axis.  Line l of batch b starts at
// in + b*in_bs + l*in_ld (and likewise for out).
struct Rows3 {
  const float2* in;
  long long in_ld, in_bs;
  int n_in, o_in;
  float2* out;
  long long out_ld, out_bs;
  int n_out, o_out;
  int nlines;
};

// Inverse rows of F_g . S * scale for G bands (g = 0..G-1): S is one level
// spectrum per batch, F the band spectra (band stride f_gs), out band g at
// out + g*out_gs + b*out_bs.  Lines have N elements in S and F.
struct Rows3Mul {
  const float2* S;
  long long s_bs;
  const float2* F;
  long long f_gs;
  float scale;
  float2* out;
  long long out_ld, out_gs, out_bs;
  int n_out, o_out;
  int nlines, G;
};

// Fused inverse of one half-spectrum plane for G bands: plane k0 of
// F_g . S * scale is inverse transformed along axis 2 (rows) and axis 1
// (cols) with the output window [o_out, o_out + n_out) on both, giving
// out[g][k0][i1][i2] (i1, i2 < n_out) at out + b*out_bs + g*out_gs +
// (k0*n_out + i1)*n_out + i2.  S and F planes are N x N complex.
struct Plane3Inv {
  const float2* S;
  long long s_bs;
  const float2* F;
  long long f_gs;
  float scale;
  float2* out;
  long long out_gs, out_bs;
  int n_out, o_out, G, H;
};

// Forward rows of G band inputs, multiplied by conj(F_g) * scale and summed
// over the bands into out (lines of N; accumulate = add to out).
struct Rows3Cmac {
  const float2* in;
  long long in_ld, in_gs, in_bs;
  int n_in, o_in;
  const float2* F;
  long long f_gs;
  float scale;
  float2* out;
  long long out_bs;
  int nlines, G, accumulate;
};

// Lines along axis 1 of [P][rows][W] arrays: element (p, m, c) at
// in + p*in_ps + m*W + c, batch b adds b*in_bs.
struct Cols3 {
  const float2* in;
  long long in_ps, in_bs;
  int n_in, o_in;
  float2* out;
  long long out_ps, out_bs;
  int n_out, o_out;
  int W, P;
};

// Axis-0 real -> half complex.  Real input [n_in][E] (element (m, e) at
// in + m*E + e), output [N/2+1][E] complex.  Batch (g, b) adds
// g*in_gs + b*in_bs (and out_gs / out_bs).
struct Pencil3Fwd {
  const float* in;
  long long in_gs, in_bs;
  int n_in, o_in;
  float2* out;
  long long out_gs, out_bs;
  int E;
};

// Axis-0 half complex -> real, window [o_out, o_out + n_out) written as
// [n_out][E].  mode 0: store band g to out + g*out_gs; mode 1: sum over the G
// bands of the hard-thresholded values (keep |v| > th[b*th_bs + th0 + g]);
// mode 2: plain sum over the G bands.  For modes 1 and 2, first = 1 writes
// the sum, first = 0 adds it to out.
struct Pencil3Inv {
  const float2* in;
  long long in_gs, in_bs;
  float* out;
  long long out_gs, out_bs;
  int n_out, o_out, E, G;
  const float* th;
  long long th_bs;
  int th0, mode, first;
};

// 1D passband factor v[|f|], |f| <= P, of the separable lowpass mask
// Lf(k) = v(|f0|) v(|f1|) v(|f2|) on the box max |f_d| <= P (0 outside).
constexpr int BP_MAX_P = 160;
struct BpMask {
  int P;
  double v[BP_MAX_P + 1];
};

bool fft3_supported(int N);
hipError_t launch3_rows(int N, int dir, const Rows3& io, int nb, const float2* tw, hipStream_t s);
hipError_t launch3_rows_mul(int N, const Rows3Mul& io, int nb, const float2* tw, hipStream_t s);
hipError_t launch3_rows_cmac(int N, const Rows3Cmac& io, int nb, const float2* tw, hipStream_t s);
// cw: output columns per block (64 or 32; 0 = default).
bool plane3_supported(int N, int n_out);
hipError_t launch3_plane_inv(int N, const Plane3Inv& io, int nb, int cw, const float2* tw, hipStream_t s);
hipError_t launch3_cols(int N, int dir, const Cols3& io, int nb, const float2* tw, hipStream_t s);
hipError_t launch3_pencil_fwd(int N, const Pencil3Fwd& io, int G, int nb, const float2* tw, hipStream_t s);
hipError_t launch3_pencil_inv(int N, const Pencil3Inv& io, int nb, const float2* tw, hipStream_t s);

// X (half spectrum of side n) -> Hs = X Hf hscale (side n) and
// Lp = crop(X Lf) lscale (side M = n / D).  Batch stride xs / hs / ls.
hipError_t launch3_bp_split(const float2* X, long long xs, float2* Hs, long long hs, float2* Lp, long long ls, int n,
                            int M, const BpMask& m, double hscale, double lscale, int nb, hipStream_t s);
// out = H Hf hscale + ext(Lp) Lf lscale (side n; Lp side M).
hipError_t launch3_bp_merge(const float2* H, long long hs, const float2* Lp, long long ls, float2* out,
                            long long os, int n, int M, const BpMask& m, double hscale, double lscale, int nb,
                            hipStream_t s);

}  // namespace shx
