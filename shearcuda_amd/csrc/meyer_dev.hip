// Spatial Meyer shearing filters of one scale, built on the device in FP64
// (shearlet_toolbox_1/shearing_filters_Myer.m:18-27 =
// shearcuda/shearing_filters_Myer.m:29-40):
//   1. the pseudo-polar grid of an n x n block: 2n lines through the centre
//      of an (n+1)-point grid, rounded to cells (gen_x_y_cordinates.m:22-111),
//      and the number of line samples in each cell (its avg_pol);
//   2. 2^level Meyer windows over the 2n lines (windowing.m:11-27,
//      meyer_wind.m:11-19);
//   3. per direction, the mean window value of the line samples in each cell
//      (rec_from_pol.m:16-35, option 0);
//   4. the centred inverse 2D DFT of that block, real part, / sqrt(n).
// The grid must come out exactly as the FP64 host arithmetic rounds it, so
// contraction of products and sums into FMAs is off in this file.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "api_err.h"
#include "meyer_dev.h"

#pragma clang fp contract(off)

namespace shx {
namespace {

// Half away from zero, as MATLAB's round on these (small, exact-ish) values.
__device__ __forceinline__ double round_half_away(double v) {
  return v < 0.0 ? -floor(-v + 0.5) : floor(v + 0.5);
}

// One thread per (row i of the (n+1)-point MATLAB arrays, sample j < n):
// x1(i,j), y1(i,j), x2(i,j), y2(i,j) placed into the n x n grids
// x1n = flipud(x1(1:n, 1:n)), y1n = y1(1:n, 1:n), x2n = x2(2:n+1, 1:n),
// y2n = y2(2:n+1, 1:n) with y2n(n, 1) = n (1-based values).
__global__ void k_meyer_grid(int n, int* __restrict__ x1n, int* __restrict__ y1n, int* __restrict__ x2n,
                             int* __restrict__ y2n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based sample
  const int i = blockIdx.y;                              // 0-based row, 0..n
  if (j >= n) return;
  const int N = n + 1;
  const double x0 = i + 1, y0 = 1.0, xn = N - (i + 1) + 1, yn = N;
  double xa, ya, xb, yb;
  if (xn == x0) {
    xa = (N - 1) / 2.0 + 1.0;
    ya = j + 1;
    xb = j + 1;
    yb = (N - 1) / 2.0 + 1.0;
  } else {
    const double m1 = (yn - y0) / (xn - x0);
    // linspace(x0, xn, N): x0 + (k * (xn - x0)) / (N - 1), the last point xn
    const double xt = (j == N - 1) ? xn : x0 + ((double)j * (xn - x0)) / (double)(N - 1);
    ya = round_half_away(m1 * (xt - x0) + y0);
    xa = round_half_away(xt);
    xb = ya;
    yb = xa;
  }
  if (i < n) {
    x1n[(n - 1 - i) * n + j] = (int)xa;
    y1n[i * n + j] = (int)ya;
  }
  if (i >= 1) {
    x2n[(i - 1) * n + j] = (int)xb;
    y2n[(i - 1) * n + j] = (i - 1 == n - 1 && j == 0) ? n : (int)yb;
  }
}

__device__ double meyer_wind(double x) {
  if (-1.0 / 3 + 0.5 < x && x < 1.0 / 3 + 0.5) return 1.0;
  if ((1.0 / 3 + 0.5 <= x && x <= 2.0 / 3 + 0.5) || (-2.0 / 3 + 0.5 <= x && x <= 1.0 / 3 + 0.5)) {
    const double w = 3.0 * fabs(x - 0.5) - 1.0;
    const double z = pow(w, 4.0) * (35.0 - 84.0 * w + 70.0 * pow(w, 2.0) - 20.0 * pow(w, 3.0));
    const double c = cos(M_PI / 2.0 * z);
    return c * c;
  }
  return 0.0;
}

// wf[r][band] of windowing(ones(2n), D): band b's window g(index) is written
// at line floor(mod(k + b T, 2n)) for k = -T/2 .. 1.5T-1 (later writes win).
__global__ void k_meyer_windows(int n, int D, double* __restrict__ wf) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int NL = 2 * n;
  if (r >= NL) return;
  const int T = NL / D;
  double val = 0.0;
  int index = 0;
  for (double k = -T / 2.0; k <= 1.5 * T - 1 + 1e-9; k += 1.0, ++index) {
    double m = fmod(k + (double)b * T, (double)NL);
    if (m < 0) m += NL;
    if ((int)floor(m) == r) val = meyer_wind((-T / 2.0 + index) / T);
  }
  wf[(size_t)r * D + b] = val;
}

__device__ __forceinline__ bool in_block(int y, int x, int n) { return y >= 1 && y <= n && x >= 1 && x <= n; }

// Line samples per cell (avg_pol of gen_x_y_cordinates.m:89-111).  A sample
// outside the block (which the grid construction never makes for the sizes
// tested) sets *bad instead of being counted; so does an empty cell.
__global__ void k_meyer_count(int n, const int* __restrict__ x1n, const int* __restrict__ y1n,
                              const int* __restrict__ x2n, const int* __restrict__ y2n, double* __restrict__ F,
                              int* __restrict__ bad) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int b = 0;
  for (int c = 0; c < n * n; ++c) F[c] = 0.0;
  for (int e = 0; e < n * n; ++e) {
    if (in_block(y1n[e], x1n[e], n)) F[(y1n[e] - 1) * n + x1n[e] - 1] += 1.0;
    else b = 1;
  }
  for (int e = 0; e < n * n; ++e) {
    if (in_block(y2n[e], x2n[e], n)) F[(y2n[e] - 1) * n + x2n[e] - 1] += 1.0;
    else b = 1;
  }
  for (int c = 0; c < n * n; ++c)
    if (F[c] == 0.0) b = 1;
  *bad = b;
}

// One thread per direction: the lines of set 1 carry wf[0..n), set 2
// wf[n..2n); every sample adds its line's value to its cell, set 1 first, in
// line order (the order of the host builder), then the cells are divided by
// their sample counts.
__global__ void k_meyer_support(int n, int D, const int* __restrict__ x1n, const int* __restrict__ y1n,
                                const int* __restrict__ x2n, const int* __restrict__ y2n,
                                const double* __restrict__ wf, const double* __restrict__ F,
                                double* __restrict__ C) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= D) return;
  double* Ck = C + (size_t)k * n * n;
  for (int c = 0; c < n * n; ++c) Ck[c] = 0.0;
  for (int e = 0; e < n * n; ++e)
    if (in_block(y1n[e], x1n[e], n)) Ck[(y1n[e] - 1) * n + x1n[e] - 1] += wf[(size_t)(e / n) * D + k];
  for (int e = 0; e < n * n; ++e)
    if (in_block(y2n[e], x2n[e], n)) Ck[(y2n[e] - 1) * n + x2n[e] - 1] += wf[(size_t)(n + e / n) * D + k];
  for (int c = 0; c < n * n; ++c) Ck[c] = Ck[c] / F[c];
}

// Centred inverse DFT, rows: R[k][u][q] = sum_v C[k][u][v] e^{2 pi i (v + h)(q - h) / n},
// h = floor(n/2) (fftshift before and after ifft2); tw[t] = e^{2 pi i t / n}.
__global__ void k_meyer_idft_rows(int n, int D, const double* __restrict__ C, const double2* __restrict__ tw,
                                  double2* __restrict__ R) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int u = blockIdx.y, k = blockIdx.z;
  if (q >= n) return;
  const int h = n / 2;
  const double* row = C + ((size_t)k * n + u) * n;
  double re = 0.0, im = 0.0;
  for (int v = 0; v < n; ++v) {
    int t = ((v + h) * (q - h)) % n;
    if (t < 0) t += n;
    re += row[v] * tw[t].x;
    im += row[v] * tw[t].y;
  }
  R[((size_t)k * n + u) * n + q] = make_double2(re, im);
}

// Columns and output: w[k][p][q] = Re(sum_u R[k][u][q] e^{2 pi i (u + h)(p - h) / n}) / n^2 / sqrt(n).
__global__ void k_meyer_idft_cols(int n, int D, const double2* __restrict__ R, const double2* __restrict__ tw,
                                  double* __restrict__ w) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y, k = blockIdx.z;
  if (q >= n) return;
  const int h = n / 2;
  double re = 0.0;
  for (int u = 0; u < n; ++u) {
    int t = ((u + h) * (p - h)) % n;
    if (t < 0) t += n;
    const double2 r = R[((size_t)k * n + u) * n + q];
    re += r.x * tw[t].x - r.y * tw[t].y;
  }
  w[((size_t)k * n + p) * n + q] = re / ((double)n * (double)n) / sqrt((double)n);
}

// Device scratch of one call, freed on every return path.
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t get(T** out, size_t count) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, sizeof(T) * count);
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T*>(q);
    return e;
  }
};

}  // namespace

int meyer_filters_device(int n, int level, double* w_dev, hipStream_t s) {
  if (n < 2 || n > 1024) return fail(SHX_ERR_ARG, "filter size %d outside [2, 1024]", n);
  if (level < 0 || level > 5) return fail(SHX_ERR_ARG, "direction level %d outside [0, 5]", level);
  const int D = 1 << level;
  if ((2 * n) % D != 0) return fail(SHX_ERR_ARG, "2 * filter size (%d) must be a multiple of 2^level (%d)", 2 * n, D);
  if (!w_dev) return fail(SHX_ERR_ARG, "null output");
  std::vector<double2> tw(n);
  for (int t = 0; t < n; ++t) {
    const std::complex<double> z = std::polar(1.0, 2.0 * M_PI * t / n);
    tw[t] = make_double2(z.real(), z.imag());
  }
  Scratch sc;
  int *x1n, *y1n, *x2n, *y2n;
  double *wf, *F, *C;
  double2 *R, *twd;
  int* bad;
  const size_t nn = (size_t)n * n;
  hipError_t e;
  if ((e = sc.get(&x1n, nn)) != hipSuccess || (e = sc.get(&y1n, nn)) != hipSuccess ||
      (e = sc.get(&x2n, nn)) != hipSuccess || (e = sc.get(&y2n, nn)) != hipSuccess ||
      (e = sc.get(&wf, (size_t)2 * n * D)) != hipSuccess || (e = sc.get(&F, nn)) != hipSuccess ||
      (e = sc.get(&C, nn * D)) != hipSuccess || (e = sc.get(&R, nn * D)) != hipSuccess ||
      (e = sc.get(&twd, (size_t)n)) != hipSuccess || (e = sc.get(&bad, 1)) != hipSuccess)
    return fail(SHX_ERR_ALLOC, "Meyer filter workspace allocation failed: %s", hipGetErrorString(e));
  SHX_HIP(hipMemcpyAsync(twd, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));
  const int B = 64;
  hipLaunchKernelGGL(k_meyer_grid, dim3((n + B - 1) / B, n + 1), dim3(B), 0, s, n, x1n, y1n, x2n, y2n);
  hipLaunchKernelGGL(k_meyer_windows, dim3((2 * n + B - 1) / B, D), dim3(B), 0, s, n, D, wf);
  hipLaunchKernelGGL(k_meyer_count, dim3(1), dim3(1), 0, s, n, x1n, y1n, x2n, y2n, F, bad);
  hipLaunchKernelGGL(k_meyer_support, dim3((D + B - 1) / B), dim3(B), 0, s, n, D, x1n, y1n, x2n, y2n, wf, F, C);
  hipLaunchKernelGGL(k_meyer_idft_rows, dim3((n + B - 1) / B, n, D), dim3(B), 0, s, n, D, C, twd, R);
  hipLaunchKernelGGL(k_meyer_idft_cols, dim3((n + B - 1) / B, n, D), dim3(B), 0, s, n, D, R, twd, w_dev);
  SHX_HIP(hipGetLastError());
  int hbad = 0;
  SHX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  // the scratch is freed when this returns
  SHX_HIP(hipStreamSynchronize(s));
  if (hbad) return fail(SHX_ERR_UNSUPPORTED, "pseudo-polar grid of size %d leaves the block or an empty cell", n);
  return SHX_OK;
}

}  // namespace shx
