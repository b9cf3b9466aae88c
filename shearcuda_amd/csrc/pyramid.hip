// A trous maxflat pyramid, one level per launch, as LDS-staged halo tiles.
//
// Replaces the reference's symExt + atrousSubsample + atrousConvolution chain
// (shearcuda/shear_cuda.cu:330-361,462-614,828-914, driven by
// shearcuda/src_cuda/atrousdec.cpp:41-131 and atrousrec.cpp:42-123).
//
// One level with dilation m (1, 2, 4, 8) and a symmetric filter of odd length
// LEN = 2*c0+1 computes
//   y(r,c) = sum_{a,b} h(a,b) x(fold(r + m(a-c0)), fold(c + m(b-c0)))
// where fold is the half-sample symmetric reflection (symext.m:18-31 with the
// level shifts of atrousdec.m:33-47 reduces to exactly this centred form).
// Outputs with the same (r mod m, c mod m) read the same sub-lattice, so a
// block owns a TH x TW tile of one polyphase sub-lattice and stages its input
// (with a halo of H = 9 sub-lattice samples) in LDS; the symmetric extension
// is folded into that loader.
//
// The maxflat filters have exact low rank (h0: 2, h1: 4, g0: 3, g1: 3), so
// h = sum_k v_k (x) w_k is applied as RANK vertical 1D passes into LDS
// followed by RANK horizontal 1D passes, skipping the taps 1 and LEN-2 that
// are zero in every maxflat filter.  The factors are compile-time constants
// (maxflat_factors.h), so every tap is an FMA with a literal operand.
#include "maxflat_factors.h"  // generated at build time by gen_factors.cpp
#include "pyramid.h"

namespace shx {

constexpr int HALO = 9;

// n is a power of two, so s mod 2n is s & (2n-1) for any sign of s.
__device__ __forceinline__ int fold_index(int s, int n) {
  const int t = s & (2 * n - 1);
  return t < n ? t : 2 * n - 1 - t;
}

template <int LEN>
__device__ __forceinline__ constexpr bool zero_tap(int a) {
  return a == 1 || a == LEN - 2;
}

// Vertical pass of one filter component over a strip of VS rows.
template <int LEN, int VS>
__device__ __forceinline__ void vpass(const float (&col)[VS + 2 * HALO], const float (&v)[LEN], float (&t)[VS]) {
  constexpr int off = HALO - (LEN - 1) / 2;
#pragma unroll
  for (int i = 0; i < VS; ++i) {
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < LEN; ++a) {
      if (zero_tap<LEN>(a)) continue;
      s = fmaf(v[a], col[i + a + off], s);
    }
    t[i] = s;
  }
}

// Horizontal pass: acc[j] += sum_b w[b] * row[j + b + off]
template <int LEN, int HS>
__device__ __forceinline__ void hpass(const float* __restrict__ row, const float (&w)[LEN], float (&acc)[HS]) {
  constexpr int off = HALO - (LEN - 1) / 2;
  float r[HS + LEN - 1];
#pragma unroll
  for (int i = 0; i < HS + LEN - 1; ++i) r[i] = row[i + off];
#pragma unroll
  for (int j = 0; j < HS; ++j) {
#pragma unroll
    for (int b = 0; b < LEN; ++b) {
      if (zero_tap<LEN>(b)) continue;
      acc[j] = fmaf(w[b], r[j + b], acc[j]);
    }
  }
}

// Load a (TH+2H) x (TW+2H) sub-lattice tile with symmetric folding.
template <int TH, int TW, int T>
__device__ __forceinline__ void load_tile(float* __restrict__ tile, const float* __restrict__ x, int L, int m, int pr,
                                          int pc, int R0, int C0, float th, float scale) {
  constexpr int W = TW + 2 * HALO;
  constexpr int NT = (TH + 2 * HALO) * W;
  for (int idx = threadIdx.x; idx < NT; idx += T) {
    const int rr = idx / W, cc = idx - rr * W;
    const int gr = fold_index(pr + m * (R0 - HALO + rr), L);
    const int gc = fold_index(pc + m * (C0 - HALO + cc), L);
    float v = x[(size_t)gr * L + gc];
    if (th >= 0.f) v = hard_keep(v, th);
    tile[idx] = v * scale;
  }
}

// ---------------------------------------------------------------------------
// Decomposition level: one input, two outputs (h0 lowpass, h1 highpass).
// grid: (ceil(L/m/TW), ceil(L/m/TH), m*m)
template <int TH, int TW, int VS, int HS, int T>
__global__ void __launch_bounds__(T) k_atrous_dec(const float* __restrict__ x, float* __restrict__ ylow,
                                                   float* __restrict__ yhigh, ImgStrides st3, int L, int m) {
  constexpr int W = TW + 2 * HALO;
  constexpr int WP = W | 1;  // odd row stride for the intermediates
  constexpr int R0N = mf::H0_RANK, R1N = mf::H1_RANK;
  static_assert(mf::H0_LEN == 13 && mf::H1_LEN == 19, "maxflat h0/h1 sizes");
  __shared__ float tile[(TH + 2 * HALO) * W];
  __shared__ float tv[R0N + R1N][TH * WP];
  const int ph = blockIdx.z % (m * m), img = blockIdx.z / (m * m);
  const int pr = ph / m, pc = ph - pr * m;
  x += (size_t)img * st3.a;
  ylow += (size_t)img * st3.b;
  yhigh += (size_t)img * st3.c;
  const int nsub = L / m;
  const int R0 = blockIdx.y * TH, C0 = blockIdx.x * TW;
  load_tile<TH, TW, T>(tile, x, L, m, pr, pc, R0, C0, -1.0f, 1.0f);
  __syncthreads();
  // vertical passes: items = W columns x TH/VS strips
  for (int it = threadIdx.x; it < W * (TH / VS); it += T) {
    const int cc = it % W, st = it / W;
    float col[VS + 2 * HALO];
#pragma unroll
    for (int i = 0; i < VS + 2 * HALO; ++i) col[i] = tile[(st * VS + i) * W + cc];
    float t[VS];
#pragma unroll
    for (int k = 0; k < R0N; ++k) {
      vpass<13, VS>(col, mf::H0_V[k], t);
#pragma unroll
      for (int i = 0; i < VS; ++i) tv[k][(st * VS + i) * WP + cc] = t[i];
    }
#pragma unroll
    for (int k = 0; k < R1N; ++k) {
      vpass<19, VS>(col, mf::H1_V[k], t);
#pragma unroll
      for (int i = 0; i < VS; ++i) tv[R0N + k][(st * VS + i) * WP + cc] = t[i];
    }
  }
  __syncthreads();
  // horizontal passes: items = TH rows x TW/HS strips, lanes along rows
  for (int it = threadIdx.x; it < TH * (TW / HS); it += T) {
    const int r = it % TH, st = it / TH;
    float a0[HS], a1[HS];
#pragma unroll
    for (int j = 0; j < HS; ++j) a0[j] = a1[j] = 0.f;
#pragma unroll
    for (int k = 0; k < R0N; ++k) hpass<13, HS>(&tv[k][r * WP + st * HS], mf::H0_W[k], a0);
#pragma unroll
    for (int k = 0; k < R1N; ++k) hpass<19, HS>(&tv[R0N + k][r * WP + st * HS], mf::H1_W[k], a1);
    const int R = R0 + r;
    if (R < nsub) {
      const size_t rowoff = (size_t)(pr + m * R) * L;
#pragma unroll
      for (int j = 0; j < HS; ++j) {
        const int C = C0 + st * HS + j;
        if (C < nsub) {
          const size_t o = rowoff + pc + (size_t)m * C;
          ylow[o] = a0[j];
          yhigh[o] = a1[j];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Reconstruction level: out = g0 * (sx * keep(xlow, th)) + g1 * (sy * yhigh).
template <int TH, int TW, int VS, int HS, int T>
__global__ void __launch_bounds__(T) k_atrous_rec(const float* __restrict__ xlow, const float* __restrict__ yhigh,
                                                   float* __restrict__ out, ImgStrides st3, int L, int m, LowThresh lt,
                                                   float sx, float sy) {
  constexpr int W = TW + 2 * HALO;
  constexpr int WP = W | 1;
  constexpr int RGN = mf::G0_RANK, RHN = mf::G1_RANK;
  static_assert(mf::G0_LEN == 19 && mf::G1_LEN == 13, "maxflat g0/g1 sizes");
  __shared__ float tile[2][(TH + 2 * HALO) * W];
  __shared__ float tv[RGN + RHN][TH * WP];
  const int ph = blockIdx.z % (m * m), img = blockIdx.z / (m * m);
  const int pr = ph / m, pc = ph - pr * m;
  xlow += (size_t)img * st3.a;
  yhigh += (size_t)img * st3.b;
  out += (size_t)img * st3.c;
  const float th_x = lt.mode == 0 ? -1.0f : (lt.mode == 1 ? lt.v : lt.dev[(size_t)img * lt.stride]);
  const int nsub = L / m;
  const int R0 = blockIdx.y * TH, C0 = blockIdx.x * TW;
  load_tile<TH, TW, T>(tile[0], xlow, L, m, pr, pc, R0, C0, th_x, sx);
  load_tile<TH, TW, T>(tile[1], yhigh, L, m, pr, pc, R0, C0, -1.0f, sy);
  __syncthreads();
  for (int it = threadIdx.x; it < 2 * W * (TH / VS); it += T) {
    const int which = it / (W * (TH / VS));
    const int rem = it - which * W * (TH / VS);
    const int cc = rem % W, st = rem / W;
    float col[VS + 2 * HALO];
#pragma unroll
    for (int i = 0; i < VS + 2 * HALO; ++i) col[i] = tile[which][(st * VS + i) * W + cc];
    float t[VS];
    if (which == 0) {
#pragma unroll
      for (int k = 0; k < RGN; ++k) {
        vpass<19, VS>(col, mf::G0_V[k], t);
#pragma unroll
        for (int i = 0; i < VS; ++i) tv[k][(st * VS + i) * WP + cc] = t[i];
      }
    } else {
#pragma unroll
      for (int k = 0; k < RHN; ++k) {
        vpass<13, VS>(col, mf::G1_V[k], t);
#pragma unroll
        for (int i = 0; i < VS; ++i) tv[RGN + k][(st * VS + i) * WP + cc] = t[i];
      }
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < TH * (TW / HS); it += T) {
    const int r = it % TH, st = it / TH;
    float a[HS];
#pragma unroll
    for (int j = 0; j < HS; ++j) a[j] = 0.f;
#pragma unroll
    for (int k = 0; k < RGN; ++k) hpass<19, HS>(&tv[k][r * WP + st * HS], mf::G0_W[k], a);
#pragma unroll
    for (int k = 0; k < RHN; ++k) hpass<13, HS>(&tv[RGN + k][r * WP + st * HS], mf::G1_W[k], a);
    const int R = R0 + r;
    if (R < nsub) {
      const size_t rowoff = (size_t)(pr + m * R) * L;
#pragma unroll
      for (int j = 0; j < HS; ++j) {
        const int C = C0 + st * HS + j;
        if (C < nsub) out[rowoff + pc + (size_t)m * C] = a[j];
      }
    }
  }
}

// Tile variants (sub-lattice tile TH x TW, vertical/horizontal strips VS/HS)
// chosen per plan by the autotuner.
template <int V>
struct PyrCfg;
template <>
struct PyrCfg<0> {
  static constexpr int TH = 16, TW = 32, VS = 4, HS = 2, T = 256;
};
template <>
struct PyrCfg<1> {
  static constexpr int TH = 32, TW = 32, VS = 8, HS = 4, T = 256;
};
template <>
struct PyrCfg<2> {
  static constexpr int TH = 32, TW = 64, VS = 8, HS = 8, T = 256;
};

template <int V>
static hipError_t dec_v(const float* x, float* ylow, float* yhigh, const ImgStrides& st, int nimg, int L, int m,
                        hipStream_t s) {
  using C = PyrCfg<V>;
  const int nsub = L / m;
  dim3 grid((nsub + C::TW - 1) / C::TW, (nsub + C::TH - 1) / C::TH, m * m * nimg);
  hipLaunchKernelGGL((k_atrous_dec<C::TH, C::TW, C::VS, C::HS, C::T>), grid, dim3(C::T), 0, s, x, ylow, yhigh, st, L,
                     m);
  return hipGetLastError();
}

template <int V>
static hipError_t rec_v(const float* xlow, const float* yhigh, float* out, const ImgStrides& st, int nimg, int L,
                        int m, const LowThresh& lt, float sx, float sy, hipStream_t s) {
  using C = PyrCfg<V>;
  const int nsub = L / m;
  dim3 grid((nsub + C::TW - 1) / C::TW, (nsub + C::TH - 1) / C::TH, m * m * nimg);
  hipLaunchKernelGGL((k_atrous_rec<C::TH, C::TW, C::VS, C::HS, C::T>), grid, dim3(C::T), 0, s, xlow, yhigh, out, st,
                     L, m, lt, sx, sy);
  return hipGetLastError();
}

hipError_t launch_atrous_dec(const float* x, float* ylow, float* yhigh, const ImgStrides& st, int nimg, int L, int m,
                             hipStream_t s, int variant) {
  if (m < 1 || L % m != 0 || nimg < 1 || m * m * nimg > 65535) return hipErrorInvalidValue;
  switch (variant) {
    case 0: return dec_v<0>(x, ylow, yhigh, st, nimg, L, m, s);
    case 1: return dec_v<1>(x, ylow, yhigh, st, nimg, L, m, s);
    case 2: return dec_v<2>(x, ylow, yhigh, st, nimg, L, m, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_atrous_rec(const float* xlow, const float* yhigh, float* out, const ImgStrides& st, int nimg, int L,
                             int m, const LowThresh& lt, float sx, float sy, hipStream_t s, int variant) {
  if (m < 1 || L % m != 0 || nimg < 1 || m * m * nimg > 65535) return hipErrorInvalidValue;
  switch (variant) {
    case 0: return rec_v<0>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    case 1: return rec_v<1>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    case 2: return rec_v<2>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace shx
