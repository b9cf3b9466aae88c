// Shearing-filter spectra of one scale built on the device (FP64 direct DFT of
// the small-support spatial filters, normalisation, FP32 packing).
#pragma once

#include <hip/hip_runtime.h>

namespace shx {

// w_host: D x n1 x n1 FP64 spatial filters (row-major); writes the packed
// spectra (D x L x L/2 float2, scaled by 1/L^2) and the beta vectors (D x L
// float2) of the scale into device memory.  Synchronous (null stream).
// Returns SHX_OK, or an error code with the message set.
int spectra_device(int L, int D, int n1, const double* w_host, int shift, float2* main_out, float2* beta_out,
                   int scale);
// The same from filters already in device memory (w_dev: D x n1 x n1 FP64).
int spectra_device_from_dev(int L, int D, int n1, const double* w_dev, int shift, float2* main_out,
                            float2* beta_out, int scale);

}  // namespace shx
