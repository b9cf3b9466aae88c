// Shearing-filter spectra built on the device, in FP64.
//
// H_k(u, v) = sum_{r,c} w_k[r][c] e^{-2 pi i (u rr + v cc) / L} with
// rr = (r - s) mod L, cc = (c - s) mod L (s = 0, or n1/2 for the MEX GPU
// path's circular shift, shearcuda/shear_cuda.cu:2303-2305), i.e.
// fft2(w_k, L, L) of shearlet_toolbox_1/nsst_dec2.m:50; then every frequency is
// divided by sqrt(sum_k |H_k|^2) (nsst_dec2.m:54-58, prepareMyerFilters_kernel
// shear_cuda.cu:2399-2423) and the result is packed as the host builder packs
// it (scaled by 1/L^2; column 0 carries alpha = (H(u,0) + H(u,L/2))/2, the beta
// vector (H(u,0) - H(u,L/2))/2).
//
// The spatial filters have small support (n1 x n1, n1 <= L), so the transform
// is evaluated directly: a row DFT of the n1 nonzero rows at the L/2+1 column
// frequencies, then a column DFT of those n1 rows at every row frequency, both
// against an FP64 twiddle table indexed by (frequency * position) mod L.  The
// normalisation needs every direction of a bin, so each thread of the column
// pass owns one bin and evaluates its directions twice (norm, then output)
// rather than holding them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "api_err.h"
#include "dict_dev.h"

namespace shx {
namespace {

constexpr int ST = 256;

__global__ void __launch_bounds__(ST) k_spec_rows(const double* __restrict__ w, int n1, int L, int s,
                                                  const double2* __restrict__ tw, double2* __restrict__ G) {
  const int NC = L / 2 + 1;
  const int v = blockIdx.x * ST + threadIdx.x;
  const int r = blockIdx.y, k = blockIdx.z;
  if (v >= NC) return;
  const double* wr = w + ((size_t)k * n1 + r) * n1;
  double re = 0.0, im = 0.0;
  for (int c = 0; c < n1; ++c) {
    const long long cc = (c - s + L) & (L - 1);
    const double2 t = tw[(v * cc) & (L - 1)];
    const double a = wr[c];
    re += a * t.x;
    im += a * t.y;
  }
  G[((size_t)k * n1 + r) * NC + v] = make_double2(re, im);
}

__device__ __forceinline__ double2 col_dft(const double2* __restrict__ Gk, int n1, int NC, int v, int u, int L, int s,
                                           const double2* __restrict__ tw) {
  double re = 0.0, im = 0.0;
  for (int r = 0; r < n1; ++r) {
    const long long rr = (r - s + L) & (L - 1);
    const double2 t = tw[(u * rr) & (L - 1)];
    const double2 g = Gk[(size_t)r * NC + v];
    re += t.x * g.x - t.y * g.y;
    im += t.x * g.y + t.y * g.x;
  }
  return make_double2(re, im);
}

__global__ void __launch_bounds__(ST) k_spec_cols(const double2* __restrict__ G, int D, int n1, int L, int s,
                                                  const double2* __restrict__ tw, float2* __restrict__ main_out,
                                                  float2* __restrict__ beta_out, int* __restrict__ bad) {
  const int N = L / 2, NC = N + 1;
  const int v = blockIdx.x * ST + threadIdx.x;  // packed slot; slot 0 also carries v = N
  const int u = blockIdx.y;
  if (v >= N) return;
  const size_t gk = (size_t)n1 * NC;
  double d0 = 0.0, dn = 0.0;
  for (int k = 0; k < D; ++k) {
    const double2 h = col_dft(G + k * gk, n1, NC, v, u, L, s, tw);
    d0 += h.x * h.x + h.y * h.y;
    if (v == 0) {
      const double2 g = col_dft(G + k * gk, n1, NC, N, u, L, s, tw);
      dn += g.x * g.x + g.y * g.y;
    }
  }
  d0 = sqrt(d0);
  dn = sqrt(dn);
  const double sc = 1.0 / ((double)L * (double)L);
  bool ok = true;
  for (int k = 0; k < D; ++k) {
    const double2 h = col_dft(G + k * gk, n1, NC, v, u, L, s, tw);
    const double hx = h.x / d0, hy = h.y / d0;
    float2* mk = main_out + (size_t)k * L * N + (size_t)u * N;
    if (v == 0) {
      const double2 g = col_dft(G + k * gk, n1, NC, N, u, L, s, tw);
      const double gx = g.x / dn, gy = g.y / dn;
      const double ax = 0.5 * (hx + gx) * sc, ay = 0.5 * (hy + gy) * sc;
      const double bx = 0.5 * (hx - gx) * sc, by = 0.5 * (hy - gy) * sc;
      ok = ok && isfinite(ax) && isfinite(ay) && isfinite(bx) && isfinite(by);
      mk[0] = make_float2((float)ax, (float)ay);
      beta_out[(size_t)k * L + u] = make_float2((float)bx, (float)by);
    } else {
      const double px = hx * sc, py = hy * sc;
      ok = ok && isfinite(px) && isfinite(py);
      mk[v] = make_float2((float)px, (float)py);
    }
  }
  if (!ok) bad[0] = 1;
}

}  // namespace

namespace {
int spectra_impl(int L, int D, int n1, const double* w_src, bool src_on_device, int shift, float2* main_out,
                 float2* beta_out, int scale);
}  // namespace

int spectra_device(int L, int D, int n1, const double* w_host, int shift, float2* main_out, float2* beta_out,
                   int scale) {
  return spectra_impl(L, D, n1, w_host, false, shift, main_out, beta_out, scale);
}

int spectra_device_from_dev(int L, int D, int n1, const double* w_dev, int shift, float2* main_out,
                            float2* beta_out, int scale) {
  return spectra_impl(L, D, n1, w_dev, true, shift, main_out, beta_out, scale);
}

namespace {
int spectra_impl(int L, int D, int n1, const double* w_src, bool src_on_device, int shift, float2* main_out,
                 float2* beta_out, int scale) {
  const int N = L / 2, NC = N + 1;
  const int s = shift ? n1 / 2 : 0;
  std::vector<double2> tw(L);
  for (int t = 0; t < L; ++t) {
    const std::complex<double> z = std::polar(1.0, -2.0 * M_PI * t / L);
    tw[t] = make_double2(z.real(), z.imag());
  }
  double* w = nullptr;
  double2 *G = nullptr, *twd = nullptr;
  int* bad = nullptr;
  const size_t bw = sizeof(double) * (size_t)D * n1 * n1, bg = sizeof(double2) * (size_t)D * n1 * NC;
  hipError_t e = hipMalloc(&w, bw);
  if (e == hipSuccess) e = hipMalloc(&G, bg);
  if (e == hipSuccess) e = hipMalloc(&twd, sizeof(double2) * L);
  if (e == hipSuccess) e = hipMalloc(&bad, sizeof(int));
  int r = SHX_OK;
  if (e != hipSuccess) r = fail(SHX_ERR_ALLOC, "spectra workspace allocation failed: %s", hipGetErrorString(e));
  int hbad = 0;
  if (r == SHX_OK && ((e = hipMemcpy(w, w_src, bw, src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)) !=
                          hipSuccess ||
                      (e = hipMemcpy(twd, tw.data(), sizeof(double2) * L, hipMemcpyHostToDevice)) != hipSuccess ||
                      (e = hipMemcpy(bad, &hbad, sizeof(int), hipMemcpyHostToDevice)) != hipSuccess))
    r = fail(SHX_ERR_HIP, "spectra upload failed: %s", hipGetErrorString(e));
  if (r == SHX_OK) {
    hipLaunchKernelGGL(k_spec_rows, dim3((NC + ST - 1) / ST, n1, D), dim3(ST), 0, nullptr, w, n1, L, s, twd, G);
    hipLaunchKernelGGL(k_spec_cols, dim3((N + ST - 1) / ST, L), dim3(ST), 0, nullptr, G, D, n1, L, s, twd, main_out,
                       beta_out, bad);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
      r = fail(SHX_ERR_HIP, "spectra kernels failed: %s", hipGetErrorString(e));
    else if (hbad)
      r = fail(SHX_ERR_ARG, "scale %d: filter spectra vanish at some frequency (division by zero)", scale);
  }
  if (w) (void)hipFree(w);
  if (G) (void)hipFree(G);
  if (twd) (void)hipFree(twd);
  if (bad) (void)hipFree(bad);
  return r;
}
}  // namespace

}  // namespace shx
