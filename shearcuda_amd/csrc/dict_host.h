// FP64 host construction of shearing-filter spectra and FFT twiddles.
#pragma once

#include <complex>
#include <vector>

namespace shx {

// Normalised spectra of one scale, D x L x (L/2+1) complex (row u, column v).
void scale_spectra_fp64(int L, int D, int n1, const double* w, int shift, std::vector<std::complex<double>>& H);
// Packed FP32 form used by the column kernels, scaled by 1/L^2:
//   main: D x L x (L/2) complex, column 0 = (H(u,0)+H(u,L/2))/2
//   beta: D x L complex = (H(u,0)-H(u,L/2))/2
void pack_spectra_fp32(int L, int D, const std::vector<std::complex<double>>& H, float* main_out, float* beta_out);
// tw[k] = exp(-2 pi i k / L), k < L, interleaved (re, im).
void twiddles_fp32(int L, float* out);
// dtw[k] = exp(-i pi k / (2L)), k < L, interleaved (re, im): the DCT-II
// pre/post twiddles of an L-point transform computed by an L-point FFT.
void dct_twiddles_fp32(int L, float* out);

}  // namespace shx
