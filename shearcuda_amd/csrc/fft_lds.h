// Batched mixed-radix (8/4/2) Stockham FFT staged through LDS, for gfx950.
//
// A workgroup holds B transforms of N complex points in LDS; element (b, i)
// lives at buf[b * bs + i * es].  Rows use es = 1, bs = N (+pad); column tiles
// use es = tile width, bs = 1.  Element index e maps to LDS slot
// lds_ix<PADSH>(e).  Each stage: every thread loads the R inputs of
// its butterflies into registers, applies the twiddles and the radix-R DFT,
// waits at a barrier, and writes the R outputs to their autosorted places, so
// the transform is in place in LDS and its output is in natural order.
//
// Stage with radix R and span Ns (product of the earlier radices), butterfly
// j in [0, N/R):
//   in  : x[j + r*N/R],                r = 0..R-1
//   tw  : x_r *= exp(DIR*2*pi*i * (j mod Ns) * r / (Ns*R))
//   out : y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns]
// Twiddles come from a table tw[k] = exp(-2*pi*i*k/(N*TWS)), k < N*TWS,
// computed in FP64 on the host and rounded once to FP32.
#pragma once

#include "shx_common.h"

namespace shx {

template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = cmul_i<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  // even / odd 4-point DFTs, then combine with w8^k = exp(DIR*i*pi*k/4)
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  const float c = 0.70710678118654752440f;
  // w8^1 o1
  float2 w1 = make_float2(c * (o1.x - DIR * o1.y), c * (o1.y + DIR * o1.x));
  // w8^2 o2 = (DIR*i) o2
  float2 w2 = cmul_i<DIR>(o2);
  // w8^3 o3 = c*(-1 + DIR*i) o3
  float2 w3 = make_float2(c * (-o3.x - DIR * o3.y), c * (-o3.y + DIR * o3.x));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1);
  v[5] = csub(e1, w1);
  v[2] = cadd(e2, w2);
  v[6] = csub(e2, w2);
  v[3] = cadd(e3, w3);
  v[7] = csub(e3, w3);
}

template <int DIR, int R>
__device__ __forceinline__ void dft_radix(float2* v) {
  if constexpr (R == 2) {
    dft2<DIR>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else {
    dft8<DIR>(v);
  }
}

template <int N, int Ns>
struct StageRadix {
  static constexpr int value = ((N / Ns) % 8 == 0) ? 8 : (((N / Ns) % 4 == 0) ? 4 : 2);
};

// LDS index of linear element e with one padding slot every 2^PADSH elements
// (spreads the strided Stockham writes over the banks).
template <int PADSH>
__device__ __forceinline__ int lds_ix(int e) {
  return PADSH > 0 ? e + (e >> PADSH) : e;
}

// One workgroup-wide Stockham stage, then the remaining ones (compile-time
// recursion).  All threads of the block must call it.
//   BFAST: consecutive threads take consecutive transforms (column tiles),
//          otherwise consecutive butterflies of one transform (rows).

template <int N, int TWS, int DIR, int B, int T, bool BFAST, int PADSH = 0, int Ns = 1>
__device__ __forceinline__ void fft_lds(float2* buf, int bs, int es, const float2* __restrict__ tw, int tid) {
  if constexpr (Ns < N) {
    constexpr int R = StageRadix<N, Ns>::value;
    constexpr int NR = N / R;
    constexpr int NB = B * NR;
    constexpr int PER = (NB + T - 1) / T;
    constexpr int TWSTEP = (N / (Ns * R)) * TWS;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int q = tid + p * T;
      if ((NB % T) == 0 || q < NB) {
        const int b = BFAST ? (q % B) : (q / NR);
        const int j = BFAST ? (q / B) : (q % NR);
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][r] = buf[lds_ix<PADSH>(b * bs + (j + r * NR) * es)];
        if constexpr (Ns > 1) {
          const int jm = j & (Ns - 1);
#pragma unroll
          for (int r = 1; r < R; ++r) {
            float2 w = tw[jm * r * TWSTEP];
            if constexpr (DIR > 0) w.y = -w.y;
            v[p][r] = cmul(v[p][r], w);
          }
        }
        dft_radix<DIR, R>(v[p]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int q = tid + p * T;
      if ((NB % T) == 0 || q < NB) {
        const int b = BFAST ? (q % B) : (q / NR);
        const int j = BFAST ? (q / B) : (q % NR);
        const int jm = j & (Ns - 1);
        const int base = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lds_ix<PADSH>(b * bs + (base + r * Ns) * es)] = v[p][r];
      }
    }
    __syncthreads();
    fft_lds<N, TWS, DIR, B, T, BFAST, PADSH, Ns * R>(buf, bs, es, tw, tid);
  }
}

}  // namespace shx
