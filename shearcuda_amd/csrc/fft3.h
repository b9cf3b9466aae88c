// Three-pass column FFT of 512 points (8 x 8 x 8) for a tile of CT columns,
// gfx950.  Every thread of a block of CT * 64 threads, (c, tau) with
// c = threadIdx.x % CT and tau = threadIdx.x / CT, holds eight elements of
// column c: rows tau + 64 s, s = 0..7.  The transform keeps that layout: on
// return the thread holds X[tau + 64 s].  A forward transform, a pointwise
// product and an inverse transform therefore chain in registers, with eight
// values per thread (the two-pass split of 512 = 16 x 32 holds sixteen).
//
// With n = a + 64 b and k = p + 8 q (a < 64; b, p < 8; q < 64):
//   stage 1 (thread a):       Y_a[p] = DFT8_b x[a + 64 b],  Z_p[a] = w512^(a p) Y_a[p]
//   stage 2 (tau = p + 8 a0): V[r]   = w64^(a0 r) DFT8_a1 Z_p[a0 + 8 a1]
//   stage 3 (tau = p + 8 r):  X[p + 8 r + 64 s] = DFT8_a0 V_{p,a0}[r]
// with one LDS transpose between stages.  The LDS slot of row e, column c is
// e * CT + c + (e / 64) * CT: the four access patterns (stage-1 stores,
// stage-2 loads and stores, stage-3 loads) are then free of bank conflicts
// for CT = 4 and 8.
#pragma once

#include "fft_lds.h"
#include "fft_reg.h"

namespace shx {

// Twiddle table reads through a buffer descriptor: one 32-bit offset
// register per lane-dependent index instead of a 64-bit address.
struct TwView {
  BufView bv;
  __device__ __forceinline__ TwView(const float2* tw, int n) : bv(tw, (unsigned)(n * sizeof(float2))) {}
  __device__ __forceinline__ float2 operator[](int i) const { return bv.ld2((unsigned)(i * sizeof(float2)), 0u); }
};

template <int CT>
struct Col3 {
  static constexpr int N = 512, TPC = 64, T = CT * TPC;
  static constexpr int LDS = N * CT + (N / 64) * CT;  // float2 slots
  __device__ static __forceinline__ int ix(int e, int c) { return e * CT + c + (e >> 6) * CT; }
};

// Twiddles of one thread (forward sign; the inverse uses their conjugates):
// w1[p] = w512^(tau p) for stage 1, w2[r] = w512^(8 a0 r), a0 = tau / 8, for
// stage 2 (w1[0] = w2[0] = 1 are not stored).  tw[m] = exp(-2 pi i m / 512).
struct Fft3Tw {
  float2 w1[8];
  float2 w2[8];
};

__device__ __forceinline__ Fft3Tw fft3_tw(int tau, const float2* __restrict__ tw_table) {
  const TwView tw(tw_table, 512);
  Fft3Tw t;
  const int a0 = tau >> 3;
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    t.w1[k] = tw[tau * k];
    t.w2[k] = tw[8 * a0 * k];
  }
  return t;
}

template <int DIR>
__device__ __forceinline__ float2 tw_apply(float2 v, float2 w) {
  return DIR < 0 ? cmul(v, w) : cmulc(v, w);
}

// In place on v (layout above); DIR -1 forward, +1 inverse (unnormalised).
// buf may be written at once (the caller synchronises before reusing it),
// and is read until return: a barrier must precede its next write.
template <int CT, int DIR>
__device__ __forceinline__ void fft512_3(float2 (&v)[8], float2* buf, const Fft3Tw& t) {
  using E = Col3<CT>;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  // stage 1
  dft8<DIR>(v);
#pragma unroll
  for (int p = 1; p < 8; ++p) v[p] = tw_apply<DIR>(v[p], t.w1[p]);
#pragma unroll
  for (int p = 0; p < 8; ++p) buf[E::ix(p * 64 + tau, c)] = v[p];
  __syncthreads();
  // stage 2: tau = p + 8 a0
  const int p2 = tau & 7, a0 = tau >> 3;
#pragma unroll
  for (int a1 = 0; a1 < 8; ++a1) v[a1] = buf[E::ix(p2 * 64 + a0 + 8 * a1, c)];
  dft8<DIR>(v);
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = tw_apply<DIR>(v[r], t.w2[r]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[E::ix(p2 * 64 + 8 * r + a0, c)] = v[r];
  __syncthreads();
  // stage 3: tau = p + 8 r (r = tau / 8)
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = buf[E::ix(p2 * 64 + 8 * a0 + k, c)];
  dft8<DIR>(v);
}

// ---- 1024 points = 16 x 8 x 8 on 128 threads per column ---------------------
// Thread tid of CT * 128 threads: column c = (tid / 2) % CT and column-local
// index q = (tid & 1) + 2 (tid / (2 CT)) in [0, 128), so that the two halves
// of each radix-16 DFT of stage 1 sit in adjacent lanes (dft_pair).  With
// n = a + 64 b (a < 64, b < 16) and k = p + 16 q' (p < 16, q' < 64):
//   stage 1 (a = q / 2, lane half h = q & 1, holding b = 2 j + h, j < 8):
//                         Z_p[a] = w1024^(a p) DFT16_b x[a + 64 b], p = k + 8 h
//   stage 2 (p = q & 15, a0 = q / 16):
//                         V_{p,a0}[r] = w64^(a0 r) DFT8_a1 Z_p[a0 + 8 a1]
//   stage 3 (p = q & 15, r = q / 16):
//                         X[p + 16 r + 128 s] = DFT8_a0 V_{p,a0}[r] = X[q + 128 s]
// The forward transform takes the "natural" layout v[j] = x[in_row(q) +
// 128 j], in_row(q) = q / 2 + 64 (q & 1), and returns the spectral layout
// v[s] = X[q + 128 s].  The inverse runs the conjugated stages in reverse
// order (spectral in, natural out; both unnormalised).
template <int CT>
struct Col3m {
  static constexpr int N = 1024, TPC = 128, T = CT * TPC;
  // one pad of CT after every 64 rows and, for CT >= 4, 16 more after row
  // 511 (the stage-1 stores of a lane pair are 512 rows apart)
  static constexpr int PADH = CT >= 4 ? 16 : 0;
  static constexpr int LDS = N * CT + (N / 64) * CT + PADH;
  __device__ static __forceinline__ int ix(int e, int c) { return e * CT + c + (e >> 6) * CT + (e >> 9) * PADH; }
  __device__ static __forceinline__ int col() { return (threadIdx.x >> 1) % CT; }
  __device__ static __forceinline__ int q() { return (threadIdx.x & 1) + 2 * (threadIdx.x / (2 * CT)); }
  __device__ static __forceinline__ int in_row(int q) { return (q >> 1) + 64 * (q & 1); }
};

// Transpose of dft_pair: the thread holds X[k + R h] (k < R) and gets
// z[j] = x[2 j + h] of the length-2R DFT with sign DIR:
//   h = 0: DFT_R(X[k] + X[k + R]);  h = 1: DFT_R(W^k (X[k] - X[k + R])), W = w_2R^DIR.
template <int R, int DIR>
__device__ __forceinline__ void dft_pair_t(float2 (&z)[R], int h) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const float2 o = swap_pair(z[k]);
    z[k] = h ? ctw_mul<2 * R, DIR>(csub(o, z[k]), k) : cadd(z[k], o);
  }
  dft<R, DIR>(z);
}

template <int CT>
__device__ __forceinline__ void fft1024_fwd(float2 (&v)[8], float2* buf, const float2* __restrict__ tw_table) {
  using E = Col3m<CT>;
  const TwView tw(tw_table, 1024);
  const int c = E::col(), q = E::q();
  const int t = q >> 1, hh = q & 1, p = q & 15, a0 = q >> 4;
  float2 w1[8], w2[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w1[k] = tw[t * (k + 8 * hh)];
#pragma unroll
  for (int r = 1; r < 8; ++r) w2[r] = tw[16 * a0 * r];
  // stage 1: radix 16 over the lane pair
  dft_pair<8, -1>(v, hh);
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[E::ix((k + 8 * hh) * 64 + t, c)] = cmul(v[k], w1[k]);
  __syncthreads();
  // stage 2
#pragma unroll
  for (int a1 = 0; a1 < 8; ++a1) v[a1] = buf[E::ix(p * 64 + a0 + 8 * a1, c)];
  dft8<-1>(v);
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = cmul(v[r], w2[r]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[E::ix(p * 64 + 8 * r + a0, c)] = v[r];
  __syncthreads();
  // stage 3 (r = q / 16 = a0)
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = buf[E::ix(p * 64 + 8 * a0 + k, c)];
  dft8<-1>(v);
}

template <int CT>
__device__ __forceinline__ void fft1024_inv(float2 (&v)[8], float2* buf, const float2* __restrict__ tw_table) {
  using E = Col3m<CT>;
  const TwView tw(tw_table, 1024);
  const int c = E::col(), q = E::q();
  const int t = q >> 1, hh = q & 1, p = q & 15, a0 = q >> 4;
  float2 w1[8], w2[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w1[k] = tw[t * (k + 8 * hh)];
#pragma unroll
  for (int r = 1; r < 8; ++r) w2[r] = tw[16 * a0 * r];
  // stage 3, conjugated (over s; r = q / 16)
  dft8<+1>(v);
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[E::ix(p * 64 + 8 * a0 + k, c)] = v[k];
  __syncthreads();
  // stage 2, conjugated: twiddles, then the DFT over r
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[E::ix(p * 64 + 8 * r + a0, c)];
#pragma unroll
  for (int r = 1; r < 8; ++r) v[r] = cmulc(v[r], w2[r]);
  dft8<+1>(v);
  __syncthreads();
#pragma unroll
  for (int a1 = 0; a1 < 8; ++a1) buf[E::ix(p * 64 + a0 + 8 * a1, c)] = v[a1];
  __syncthreads();
  // stage 1, conjugated: twiddles, then the transposed pair DFT
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = cmulc(buf[E::ix((k + 8 * hh) * 64 + t, c)], w1[k]);
  dft_pair_t<8, +1>(v, hh);
}

// ---- 1024 points = 16 x 8 x 8 on 64 threads per column (16 values each) ----
// Thread (c, tau) of a block of CT * 64 threads, c = threadIdx.x % CT, holds
// v[t] = x[tau + 64 t], t < 16, and gets X[tau + 64 t] back (same layout).
// With n = a + 64 b (a < 64, b < 16) and k = p + 16 q (p < 16, q = r + 8 s):
//   stage 1 (thread a = tau):     Z_p[a] = w1024^(a p) DFT16_b x[a + 64 b]
//   stage 2 (p = tau & 15; a0 = tau >> 4 and a0 + 4):
//                                 V_{p,a0}[r] = w64^(a0 r) DFT8_a1 Z_p[a0 + 8 a1]
//   stage 3 (p = tau & 15; r = tau >> 4 and r + 4):
//                                 X[p + 16 r + 128 s] = DFT8_a0 V_{p,a0}[r]
// (p + 16 r = tau: the two stage-3 DFTs give rows tau + 64 (2 s) and
// tau + 64 (2 s + 1)).  LDS slot of row e, column c: e * CT + c + (e / 64) * CT,
// the layout of the 512-point engine: every access pattern spreads a wave's
// lanes evenly over the banks for CT = 4 and 8.
template <int CT>
struct Col4 {
  static constexpr int N = 1024, TPC = 64, T = CT * TPC;
  static constexpr int LDS = N * CT + (N / 64) * CT;
  __device__ static __forceinline__ int ix(int e, int c) { return e * CT + c + (e >> 6) * CT; }
};

// w1[p] = w1024^(tau p), w2[r] = w64^(a0 r) (a0 = tau >> 4; the second
// stage-2 DFT, a0 + 4, takes w2[r] w16^r).  tw[m] = exp(-2 pi i m / 1024).
struct Fft4Tw {
  float2 w1[16];
  float2 w2[8];
};

__device__ __forceinline__ Fft4Tw fft4_tw(int tau, const float2* __restrict__ tw_table) {
  const TwView tw(tw_table, 1024);
  Fft4Tw t;
  const int a0 = tau >> 4;
#pragma unroll
  for (int p = 1; p < 16; ++p) t.w1[p] = tw[tau * p];
#pragma unroll
  for (int r = 1; r < 8; ++r) t.w2[r] = tw[16 * a0 * r];
  return t;
}

// In place on v (layout above); DIR -1 forward, +1 inverse (unnormalised).
// buf is written at once (the caller makes sure earlier reads of it are done)
// and read until return.
template <int CT, int DIR>
__device__ __forceinline__ void fft1024_4(float2 (&v)[16], float2* buf, const Fft4Tw& t) {
  using E = Col4<CT>;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  // stage 1
  dft<16, DIR>(v);
#pragma unroll
  for (int p = 1; p < 16; ++p) v[p] = tw_apply<DIR>(v[p], t.w1[p]);
#pragma unroll
  for (int p = 0; p < 16; ++p) buf[E::ix(p * 64 + tau, c)] = v[p];
  __syncthreads();
  // stage 2
  const int p = tau & 15, h = tau >> 4;
#pragma unroll
  for (int a1 = 0; a1 < 8; ++a1) {
    v[a1] = buf[E::ix(p * 64 + h + 8 * a1, c)];
    v[8 + a1] = buf[E::ix(p * 64 + h + 4 + 8 * a1, c)];
  }
  dft8<DIR>(v);
  dft8<DIR>(v + 8);
#pragma unroll
  for (int r = 1; r < 8; ++r) {
    v[r] = tw_apply<DIR>(v[r], t.w2[r]);
    v[8 + r] = tw_apply<DIR>(ctw_mul<16, DIR>(v[8 + r], r), t.w2[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    buf[E::ix(p * 64 + 8 * r + h, c)] = v[r];
    buf[E::ix(p * 64 + 8 * r + h + 4, c)] = v[8 + r];
  }
  __syncthreads();
  // stage 3 (r = h and h + 4)
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = buf[E::ix(p * 64 + 8 * h + k, c)];
    v[8 + k] = buf[E::ix(p * 64 + 8 * (h + 4) + k, c)];
  }
  dft8<DIR>(v);
  dft8<DIR>(v + 8);
  float2 o[16];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    o[2 * s] = v[s];
    o[2 * s + 1] = v[8 + s];
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = o[k];
}

// The engines with 64 threads per column (rows tau + 64 s, NV values per
// thread), by transform length, for kernels written once for both.
template <int L, int CT>
struct ColEng;
template <int CT>
struct ColEng<512, CT> {
  static constexpr int NV = 8;
  using E = Col3<CT>;
  using Tw = Fft3Tw;
  __device__ static __forceinline__ Tw twiddles(int tau, const float2* __restrict__ tw) { return fft3_tw(tau, tw); }
  template <int DIR>
  __device__ static __forceinline__ void fft(float2 (&v)[NV], float2* buf, const Tw& t) {
    fft512_3<CT, DIR>(v, buf, t);
  }
};
template <int CT>
struct ColEng<1024, CT> {
  static constexpr int NV = 16;
  using E = Col4<CT>;
  using Tw = Fft4Tw;
  __device__ static __forceinline__ Tw twiddles(int tau, const float2* __restrict__ tw) { return fft4_tw(tau, tw); }
  template <int DIR>
  __device__ static __forceinline__ void fft(float2 (&v)[NV], float2* buf, const Tw& t) {
    fft1024_4<CT, DIR>(v, buf, t);
  }
};

}  // namespace shx
