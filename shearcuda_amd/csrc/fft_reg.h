// Register-resident FFT building blocks for gfx950.
//
// A transform of N = R1 * R2 points is done in two passes with one LDS
// transpose between them (the four-step split, with the inner twiddles):
//   pass A  thread n1 (0..R1-1) holds x[n1 + R1*n2], n2 = 0..R2-1;
//           DFT of length R2 over n2, then multiply by w_N^(n1*k2)
//           -> y[n1][k2]
//   pass B  thread k2 (0..R2-1) holds y[n1][k2], n1 = 0..R1-1;
//           DFT of length R1 over n1 -> X[k2 + R2*k1], k1 = 0..R1-1
// After pass B thread k2 holds X at k2 + R2*k1: exactly the pass-A layout of
// a transform split as (R2, R1).  A forward transform, a pointwise product and
// an inverse transform therefore chain in registers.
//
// The short DFTs (2, 3, 4, 5, 8, 9, 16, 32) are fully unrolled with compile-time
// twiddles (computed below by constexpr series in FP64, rounded once to FP32);
// the pass twiddles w_N^m come from a table computed in FP64 on the host.
#pragma once

#include "fft_lds.h"

namespace shx {

// ---- compile-time twiddles ----------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ce_sin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 24; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}

constexpr double ce_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 24; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}

// exp(-2 pi i k / R) for k < R, angles reduced to [-pi, pi]
template <int R>
struct CTw {
  float re[R];
  float im[R];
  constexpr CTw() : re(), im() {
    for (int k = 0; k < R; ++k) {
      double a = -2.0 * kPi * k / R;
      if (a < -kPi) a += 2.0 * kPi;
      re[k] = (float)ce_cos(a);
      im[k] = (float)ce_sin(a);
    }
  }
};

template <int R>
constexpr CTw<R> kCTw{};

// a * exp(DIR * 2 pi i m / R) for a compile-time (after unrolling) m
template <int R, int DIR>
__device__ __forceinline__ float2 ctw_mul(float2 a, int m) {
  m %= R;
  if (m == 0) return a;
  if ((4 * m) % R == 0) {  // quarter turns: exact
    const int q = (4 * m) / R;
    if (q == 2) return make_float2(-a.x, -a.y);
    // exp(DIR*i*pi/2 * q): q=1 -> DIR*i, q=3 -> -DIR*i
    return (q == 1) == (DIR > 0) ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  }
  const float wr = kCTw<R>.re[m];
  const float wi = DIR < 0 ? kCTw<R>.im[m] : -kCTw<R>.im[m];
  return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v);

// length R = A*B in registers: x[a + A*b] -> X[kb + B*ka]
template <int A, int B, int DIR>
__device__ __forceinline__ void dft_split(float2* v) {
  constexpr int R = A * B;
  float2 t[A][B];
#pragma unroll
  for (int a = 0; a < A; ++a) {
#pragma unroll
    for (int b = 0; b < B; ++b) t[a][b] = v[a + A * b];
    dft<B, DIR>(t[a]);
#pragma unroll
    for (int kb = 1; kb < B; ++kb) t[a][kb] = ctw_mul<R, DIR>(t[a][kb], a * kb);
  }
#pragma unroll
  for (int kb = 0; kb < B; ++kb) {
    float2 u[A];
#pragma unroll
    for (int a = 0; a < A; ++a) u[a] = t[a][kb];
    dft<A, DIR>(u);
#pragma unroll
    for (int ka = 0; ka < A; ++ka) v[kb + B * ka] = u[ka];
  }
}

// length 3: X1,2 = a - (b+c)/2 -/+ DIR-signed i (sqrt3/2)(b - c)
template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  const float h = 0.86602540378443864676f;
  const float2 a = v[0], s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  const float2 m = make_float2(a.x - 0.5f * s.x, a.y - 0.5f * s.y);
  // DIR * i * h * d
  const float2 r = make_float2(-DIR * h * d.y, DIR * h * d.x);
  v[0] = cadd(a, s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

// direct DFT of a short odd length with compile-time twiddles
template <int R, int DIR>
__device__ __forceinline__ void dft_direct(float2* v) {
  float2 x[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float2 a = v[0];
#pragma unroll
    for (int n = 1; n < R; ++n) a = cadd(a, ctw_mul<R, DIR>(v[n], n * k));
    x[k] = a;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = x[k];
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 3) {
    dft3<DIR>(v);
  } else if constexpr (R == 5) {
    dft_direct<5, DIR>(v);
  } else if constexpr (R == 9) {
    dft_split<3, 3, DIR>(v);
  } else if constexpr (R == 2) {
    dft2<DIR>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    dft8<DIR>(v);
  } else if constexpr (R == 16) {
    dft_split<4, 4, DIR>(v);
  } else if constexpr (R == 32) {
    dft_split<4, 8, DIR>(v);
  } else {
    static_assert(R == 64, "unsupported register DFT length");
    dft_split<8, 8, DIR>(v);
  }
}

// N = R1 * R2 split used by the two-pass transforms (R1 <= R2 <= 32).
template <int N>
struct Split;
template <> struct Split<8> { static constexpr int R1 = 2, R2 = 4; };
template <> struct Split<24> { static constexpr int R1 = 3, R2 = 8; };
template <> struct Split<48> { static constexpr int R1 = 3, R2 = 16; };
template <> struct Split<80> { static constexpr int R1 = 5, R2 = 16; };
template <> struct Split<144> { static constexpr int R1 = 9, R2 = 16; };
template <> struct Split<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct Split<32> { static constexpr int R1 = 4, R2 = 8; };
template <> struct Split<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct Split<128> { static constexpr int R1 = 8, R2 = 16; };
template <> struct Split<256> { static constexpr int R1 = 16, R2 = 16; };
template <> struct Split<512> { static constexpr int R1 = 16, R2 = 32; };
template <> struct Split<1024> { static constexpr int R1 = 32, R2 = 32; };

// Pass A on registers v[RB] (thread index a in [0, RA)): DFT of length RB,
// then the pass twiddles w_N^(a*kb), N = RA*RB, from the table tw
// (tw[m*TWS] = exp(-2 pi i m / N)).
// For RB > 8 only w^r (r < 8) and w^(8q) are loaded and w^(8q + r) is one
// product of the two, which keeps the twiddle registers few.
// ---- lane-pair split of a length-2R DFT ----------------------------------------
// Partner lane (lane ^ 1) value through a DPP quad permutation [1, 0, 3, 2].
__device__ __forceinline__ float swap_pair(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float2 swap_pair(float2 v) { return make_float2(swap_pair(v.x), swap_pair(v.y)); }

// Length-2R DFT over a lane pair (h = lane & 1, both lanes active): the
// thread holds z[j] = x[2j + h] (j < R) and gets X[k + R h] (k < R):
//   X[k] = E0[k] + W^k E1[k],  X[k + R] = E0[k] - W^k E1[k],  W = w_2R.
template <int R, int DIR>
__device__ __forceinline__ void dft_pair(float2 (&z)[R], int h) {
  dft<R, DIR>(z);
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const float2 o = swap_pair(z[k]);
    const float2 e0 = h ? o : z[k], e1 = h ? z[k] : o;
    const float2 t = ctw_mul<2 * R, DIR>(e1, k);
    z[k] = h ? csub(e0, t) : cadd(e0, t);
  }
}

// Pass twiddles of a paired pass A (index a, outputs k2 = k + R h): w^(a r)
// (r < LO) and w^(a (R h + LO q)) (q < NQ), from the table tw (stride TWS).
template <int R>
struct PairTw {
  static constexpr int LO = R < 8 ? R : 8;
  static constexpr int NQ = R / LO;
  float2 lo[LO];
  float2 hi[NQ];
};

template <int R, int TWS>
__device__ __forceinline__ void pair_tw(PairTw<R>& t, int a, int h, const float2* __restrict__ tw) {
  constexpr int LO = PairTw<R>::LO, NQ = PairTw<R>::NQ;
#pragma unroll
  for (int r = 1; r < LO; ++r) t.lo[r] = tw[a * r * TWS];
#pragma unroll
  for (int q = 0; q < NQ; ++q) t.hi[q] = tw[a * (R * h + LO * q) * TWS];
}

template <int R, int DIR>
__device__ __forceinline__ void pair_tw_apply(float2 (&v)[R], PairTw<R> t) {
  constexpr int LO = PairTw<R>::LO, NQ = PairTw<R>::NQ;
  if constexpr (DIR > 0) {
#pragma unroll
    for (int r = 1; r < LO; ++r) t.lo[r].y = -t.lo[r].y;
#pragma unroll
    for (int q = 0; q < NQ; ++q) t.hi[q].y = -t.hi[q].y;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    v[q * LO] = cmul(v[q * LO], t.hi[q]);
#pragma unroll
    for (int r = 1; r < LO; ++r) v[q * LO + r] = cmul(v[q * LO + r], cmul(t.hi[q], t.lo[r]));
  }
}

// Pass twiddles of one thread: w^r (r < LO) and w^(LO q) (q < NHI).
template <int RB>
struct PassTw {
  static constexpr int LO = RB < 8 ? RB : 8;
  static constexpr int NHI = RB / LO;
  float2 lo[LO];
  float2 hi[NHI];
};

template <int RB, int DIR, int TWS>
__device__ __forceinline__ void passA_tw(PassTw<RB>& t, int a, const float2* __restrict__ tw) {
  constexpr int LO = PassTw<RB>::LO, NHI = PassTw<RB>::NHI;
#pragma unroll
  for (int r = 1; r < LO; ++r) t.lo[r] = tw[a * r * TWS];
#pragma unroll
  for (int q = 1; q < NHI; ++q) t.hi[q] = tw[a * q * LO * TWS];
}

// Opaque copy of loaded twiddles: products of its values cannot be hoisted
// out of a loop (where they would stay live in registers across iterations).
template <int RB>
__device__ __forceinline__ PassTw<RB> pin(PassTw<RB> t) {
#pragma unroll
  for (int i = 1; i < PassTw<RB>::LO; ++i) asm volatile("" : "+v"(t.lo[i].x), "+v"(t.lo[i].y));
#pragma unroll
  for (int i = 1; i < PassTw<RB>::NHI; ++i) asm volatile("" : "+v"(t.hi[i].x), "+v"(t.hi[i].y));
  return t;
}

template <int R>
__device__ __forceinline__ PairTw<R> pin(PairTw<R> t) {
#pragma unroll
  for (int i = 1; i < PairTw<R>::LO; ++i) asm volatile("" : "+v"(t.lo[i].x), "+v"(t.lo[i].y));
#pragma unroll
  for (int i = 0; i < PairTw<R>::NQ; ++i) asm volatile("" : "+v"(t.hi[i].x), "+v"(t.hi[i].y));
  return t;
}

template <int RB, int DIR>
__device__ __forceinline__ void passA_apply(float2 (&v)[RB], PassTw<RB> t) {
  constexpr int LO = PassTw<RB>::LO, NHI = PassTw<RB>::NHI;
  if constexpr (DIR > 0) {
#pragma unroll
    for (int r = 1; r < LO; ++r) t.lo[r].y = -t.lo[r].y;
#pragma unroll
    for (int q = 1; q < NHI; ++q) t.hi[q].y = -t.hi[q].y;
  }
  dft<RB, DIR>(v);
#pragma unroll
  for (int r = 1; r < LO; ++r) v[r] = cmul(v[r], t.lo[r]);
#pragma unroll
  for (int q = 1; q < NHI; ++q) {
    v[q * LO] = cmul(v[q * LO], t.hi[q]);
#pragma unroll
    for (int r = 1; r < LO; ++r) v[q * LO + r] = cmul(v[q * LO + r], cmul(t.hi[q], t.lo[r]));
  }
}

// The twiddle loads are issued before the DFT, and a scheduling barrier keeps
// them (and any loads of v the caller issued just before) ahead of the
// arithmetic, so all of them are in flight together.
template <int RA, int RB, int DIR, int TWS>
__device__ __forceinline__ void passA(float2 (&v)[RB], int a, const float2* __restrict__ tw) {
  PassTw<RB> t;
  passA_tw<RB, DIR, TWS>(t, a, tw);
  __builtin_amdgcn_sched_barrier(0);
  passA_apply<RB, DIR>(v, t);
}

}  // namespace shx
