// Undecimated (redundant) periodic 2D wavelet transform, Rice Wavelet Toolbox
// semantics, as the reference's mrdwt/mirdwt (shearcuda/src/mrdwt_r.h:70-207,
// mirdwt_r.h:92-218; GPU shearcuda/src_cuda/mrdwt_cuda.cpp:220-253,
// mirdwt_cuda.cpp:210-255).
//
// Row-major m x n images; yh is m x 3nL, level l (1-based) in columns
// [3n(l-1), 3nl) as [lh hl hh] (first letter: filtering along a row).
// Level l uses dilation sf = 2^(l-1) and periodic indexing.
//   analysis  (mrdwt):  out[i] = sum_j in[(i + j*sf) mod n] * a[j]
//             a0[j] = h[j],  a1[j] = h[lh-1-j] * (-1 if lh-1-j is even)
//             rows: A = a0 (x) yl, B = a1 (x) yl; columns: yl = a0 A,
//             lh = a1 A, hl = a0 B, hh = a1 B
//   synthesis (mirdwt): out[i] = sum_t d[(i - t*sf) mod n] * g[t]
//             g0[t] = h[t]/2,  g1[t] = h[lh-1-t]/2 * (-1 if t is odd)
//             columns: X = g0 x + g1 lh, XH = g0 hl + g1 hh; rows:
//             x = g0 X + g1 XH
// The adjoint of the analysis is the synthesis with filters 2g and the adjoint
// of the synthesis is the analysis with filters a/2 (host side picks them).
#include "rdwt.h"

namespace shx {

__device__ __forceinline__ int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// analysis along rows: A, B from yl
__global__ void k_rdwt_rows_an(const float* __restrict__ yl, float* __restrict__ A, float* __restrict__ B, int m,
                               int n, int sf, RdwtFilt f, long long img_stride, long long tmp_stride) {
  const int img = blockIdx.z;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  const float* row = yl + (size_t)img * img_stride + (size_t)r * n;
  float a = 0.f, b = 0.f;
  for (int j = 0; j < f.lh; ++j) {
    const float v = row[wrap(c + j * sf, n)];
    a = fmaf(f.f0[j], v, a);
    b = fmaf(f.f1[j], v, b);
  }
  const size_t o = (size_t)img * tmp_stride + (size_t)r * n + c;
  A[o] = a;
  B[o] = b;
}

// analysis along columns: yl, lh, hl, hh of level `lev` from A, B
__global__ void k_rdwt_cols_an(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ yl,
                               float* __restrict__ yh, int m, int n, int sf, int L, int lev, RdwtFilt f,
                               long long img_stride, long long yh_stride, long long tmp_stride) {
  const int img = blockIdx.z;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  const float* a = A + (size_t)img * tmp_stride + c;
  const float* b = B + (size_t)img * tmp_stride + c;
  float ll = 0.f, lh = 0.f, hl = 0.f, hh = 0.f;
  for (int j = 0; j < f.lh; ++j) {
    const int rr = wrap(r + j * sf, m);
    const float va = a[(size_t)rr * n], vb = b[(size_t)rr * n];
    ll = fmaf(f.f0[j], va, ll);
    lh = fmaf(f.f1[j], va, lh);
    hl = fmaf(f.f0[j], vb, hl);
    hh = fmaf(f.f1[j], vb, hh);
  }
  yl[(size_t)img * img_stride + (size_t)r * n + c] = ll;
  float* yr = yh + (size_t)img * yh_stride + (size_t)r * 3 * n * L + (size_t)3 * n * (lev - 1);
  yr[c] = lh;
  yr[n + c] = hl;
  yr[2 * n + c] = hh;
}

// synthesis along columns: X, XH from x and the level's (lh, hl, hh)
__global__ void k_rdwt_cols_syn(const float* __restrict__ x, const float* __restrict__ yh, float* __restrict__ X,
                                float* __restrict__ XH, int m, int n, int sf, int L, int lev, RdwtFilt f,
                                long long img_stride, long long yh_stride, long long tmp_stride) {
  const int img = blockIdx.z;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  const float* xi = x + (size_t)img * img_stride + c;
  const float* yb = yh + (size_t)img * yh_stride + (size_t)3 * n * (lev - 1) + c;
  const size_t yrow = (size_t)3 * n * L;
  float s0 = 0.f, s1 = 0.f;
  for (int t = 0; t < f.lh; ++t) {
    const int rr = wrap(r - t * sf, m);
    const float* yr = yb + (size_t)rr * yrow;
    s0 = fmaf(f.f0[t], xi[(size_t)rr * n], s0);
    s0 = fmaf(f.f1[t], yr[0], s0);
    s1 = fmaf(f.f0[t], yr[n], s1);
    s1 = fmaf(f.f1[t], yr[2 * n], s1);
  }
  const size_t o = (size_t)img * tmp_stride + (size_t)r * n + c;
  X[o] = s0;
  XH[o] = s1;
}

// synthesis along rows: x from X, XH
__global__ void k_rdwt_rows_syn(const float* __restrict__ X, const float* __restrict__ XH, float* __restrict__ x,
                                int m, int n, int sf, RdwtFilt f, long long img_stride, long long tmp_stride) {
  const int img = blockIdx.z;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= n) return;
  const float* a = X + (size_t)img * tmp_stride + (size_t)r * n;
  const float* b = XH + (size_t)img * tmp_stride + (size_t)r * n;
  float s = 0.f;
  for (int t = 0; t < f.lh; ++t) {
    const int cc = wrap(c - t * sf, n);
    s = fmaf(f.f0[t], a[cc], s);
    s = fmaf(f.f1[t], b[cc], s);
  }
  x[(size_t)img * img_stride + (size_t)r * n + c] = s;
}

hipError_t rdwt_analysis(const float* x, float* yl, float* yh, float* tmp, int m, int n, int L, const RdwtFilt& f,
                         int nimg, hipStream_t s) {
  const long long is = (long long)m * n, ys = (long long)m * 3 * n * L, ts = 2LL * m * n;
  const dim3 blk(256), grid((n + 255) / 256, m, nimg);
  hipError_t e = hipSuccess;
  if (yl != x && (e = hipMemcpyAsync(yl, x, sizeof(float) * is * nimg, hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return e;
  int sf = 1;
  for (int lev = 1; lev <= L; ++lev, sf *= 2) {
    hipLaunchKernelGGL(k_rdwt_rows_an, grid, blk, 0, s, yl, tmp, tmp + is, m, n, sf, f, is, ts);
    hipLaunchKernelGGL(k_rdwt_cols_an, grid, blk, 0, s, tmp, tmp + is, yl, yh, m, n, sf, L, lev, f, is, ys, ts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t rdwt_synthesis(const float* yl, const float* yh, float* x, float* tmp, int m, int n, int L,
                          const RdwtFilt& f, int nimg, hipStream_t s) {
  const long long is = (long long)m * n, ys = (long long)m * 3 * n * L, ts = 2LL * m * n;
  const dim3 blk(256), grid((n + 255) / 256, m, nimg);
  hipError_t e = hipSuccess;
  if (x != yl && (e = hipMemcpyAsync(x, yl, sizeof(float) * is * nimg, hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return e;
  int sf = 1 << (L - 1);
  for (int lev = L; lev >= 1; --lev, sf /= 2) {
    hipLaunchKernelGGL(k_rdwt_cols_syn, grid, blk, 0, s, x, yh, tmp, tmp + is, m, n, sf, L, lev, f, is, ys, ts);
    hipLaunchKernelGGL(k_rdwt_rows_syn, grid, blk, 0, s, tmp, tmp + is, x, m, n, sf, f, is, ts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace shx
