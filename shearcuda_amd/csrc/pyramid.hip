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
// The transposes (for gradients) use the same tiles: the adjoint of a level
// with symmetric filters is the same correlation applied to the zero-padded
// output gradient, evaluated on the extended grid [-9m, L+9m)^2, followed by
// a fold-add of the extended grid back onto [0, L)^2 (the transpose of the
// symmetric extension).
//
// A level applies a filter pair, (h0, h1) or (g0, g1).  On a common 19 x 19
// support the pair has joint rank 3 once the centre impulse is taken out of
// the second filter (maxflat.cpp), with shared basis vectors u_i:
//   F1 = sum_i u_i (x) a_i,   F2 = delta + sum_i u_i (x) b_i
// (vertical (x) horizontal), and, the filters being symmetric, also
//   F1 = sum_i a_i (x) u_i,   F2 = delta + sum_i b_i (x) u_i.
// One input, two outputs (decomposition): three vertical passes with u_i into
// LDS, then each output takes its three horizontal passes from the same
// loaded rows.  Two inputs, one output (reconstruction): three vertical
// passes a_i * x1 + b_i * x2 into LDS, then three horizontal passes with u_i.
// The identity term is read from the staged input tile.  The factors are
// compile-time constants (maxflat_factors.h): every tap is an FMA with a
// literal operand, and taps that are exactly zero are skipped at compile time.
#include "maxflat_factors.h"  // generated at build time by gen_factors.cpp
#include "pyramid.h"

namespace shx {

constexpr int HALO = 9;

// n is a power of two, so s mod 2n is s & (2n-1) for any sign of s.
__device__ __forceinline__ int fold_index(int s, int n) {
  const int t = s & (2 * n - 1);
  return t < n ? t : 2 * n - 1 - t;
}

// Joint factors of the pairs: 0 = (h0, h1), 1 = (g0, g1).
constexpr int PLEN = mf::PAIR_LEN, PRANK = mf::PAIR_RANK;
static_assert(PLEN == 2 * HALO + 1, "the pair factors span the halo");
template <int P>
struct Pair;
template <>
struct Pair<0> {
  static constexpr const float (&U)[PRANK][PLEN] = mf::PH_U;
  static constexpr const float (&A)[PRANK][PLEN] = mf::PH_A;
  static constexpr const float (&B)[PRANK][PLEN] = mf::PH_B;
};
template <>
struct Pair<1> {
  static constexpr const float (&U)[PRANK][PLEN] = mf::PG_U;
  static constexpr const float (&A)[PRANK][PLEN] = mf::PG_A;
  static constexpr const float (&B)[PRANK][PLEN] = mf::PG_B;
};

// t[i] += sum_a v[a] * col[i + a], i < VS (taps that are zero skipped).
template <int VS>
__device__ __forceinline__ void vacc(const float (&col)[VS + 2 * HALO], const float (&v)[PLEN], float (&t)[VS]) {
#pragma unroll
  for (int i = 0; i < VS; ++i) {
#pragma unroll
    for (int a = 0; a < PLEN; ++a) {
      if (v[a] == 0.0f) continue;
      t[i] = fmaf(v[a], col[i + a], t[i]);
    }
  }
}

// acc[j] += sum_b w[b] * row[j + b], j < HS.
template <int HS>
__device__ __forceinline__ void hacc(const float (&row)[HS + 2 * HALO], const float (&w)[PLEN], float (&acc)[HS]) {
#pragma unroll
  for (int j = 0; j < HS; ++j) {
#pragma unroll
    for (int b = 0; b < PLEN; ++b) {
      if (w[b] == 0.0f) continue;
      acc[j] = fmaf(w[b], row[j + b], acc[j]);
    }
  }
}

// Load a (TH+2H) x (TW+2H) sub-lattice tile centred on input rows R0.. and
// columns C0..: symmetric fold (forward) or zero padding (adjoint).  Split in
// two so that every global load of a block (one or two tiles) is issued
// before the first is used: tile_fetch loads this thread's NIT elements into
// registers (addresses clamped into the tile, no branch around a load),
// tile_store writes them (threshold, scale) to LDS; the LDS tile is padded to
// NIT * T elements so the stores need no guard either.
template <int TH, int TW, int T>
struct TileCfg {
  static constexpr int W = TW + 2 * HALO;
  static constexpr int NT = (TH + 2 * HALO) * W;
  static constexpr int NIT = (NT + T - 1) / T;
  static constexpr int PADDED = NIT * T;  // LDS tile size: every thread stores NIT elements
};

template <int TH, int TW, int T, bool ZERO, int NIT>
__device__ __forceinline__ void tile_fetch(float (&v)[NIT], const float* __restrict__ x, int L, int m, int pr, int pc,
                                           int R0, int C0) {
  using TC = TileCfg<TH, TW, T>;
  static_assert(NIT == TC::NIT, "register count of the tile");
#pragma unroll
  for (int i = 0; i < TC::NIT; ++i) {
    const int idx = min((int)threadIdx.x + i * T, TC::NT - 1);
    const int rr = idx / TC::W, cc = idx - rr * TC::W;
    int gr = pr + m * (R0 - HALO + rr);
    int gc = pc + m * (C0 - HALO + cc);
    if constexpr (ZERO) {
      const bool in = gr >= 0 && gr < L && gc >= 0 && gc < L;
      v[i] = x[in ? (size_t)gr * L + gc : 0];
      if (!in) v[i] = 0.0f;
    } else {
      gr = fold_index(gr, L);
      gc = fold_index(gc, L);
      v[i] = x[(size_t)gr * L + gc];
    }
  }
}

template <int TH, int TW, int T, int NIT>
__device__ __forceinline__ void tile_store(float* __restrict__ tile, const float (&v)[NIT], float th, int rule,
                                           float scale) {
  using TC = TileCfg<TH, TW, T>;
  static_assert(NIT == TC::NIT, "register count of the tile");
#pragma unroll
  for (int i = 0; i < TC::NIT; ++i) tile[threadIdx.x + i * T] = shrink(v[i], th, rule) * scale;
}

// Output grid of a level: the image (forward) or the extended grid of the
// adjoint (EXT), E = L + 2*HALO*m per side, in sub-lattice rows of `m`.
__device__ __forceinline__ int out_side(int L, int m, bool ext) { return ext ? L + 2 * HALO * m : L; }

// XCD-aware block order of the level kernels.  Workgroups go to the 8 XCDs in
// turn; the flat id is remapped so that each XCD takes a contiguous range of
// (sub-lattice phase fastest, then tile column, tile row, image): the m*m
// phases of one tile read the same 128-byte lines (a line interleaves m
// phases of a row), and neighbouring tiles share their halo lines, so they
// hit one L2 instead of fetching those lines once per XCD.
__device__ __forceinline__ void pyr_block(int mm, int& bx, int& by, int& ph, int& img) {
  const int nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  const int total = nx * ny * nz;
  int lid = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  if (total % 8 == 0) lid = (lid % 8) * (total / 8) + lid / 8;
  ph = __builtin_amdgcn_readfirstlane(lid % mm);
  int r = lid / mm;
  bx = __builtin_amdgcn_readfirstlane(r % nx);
  r /= nx;
  by = __builtin_amdgcn_readfirstlane(r % ny);
  img = __builtin_amdgcn_readfirstlane(r / ny);
}

// ---------------------------------------------------------------------------
// One input, two outputs with pair P: ya = F1 (x) x, yb = F2 (x) x.
//   forward decomposition level: P = 0 (h0, h1)
//   adjoint of a reconstruction level (ADJ): P = 1 (g0, g1) on the
//   zero-padded gradient, outputs on the extended grid.
// grid: (ceil(nout/TW), ceil(nout/TH), m*m*nimg), nout = out_side/m.
template <int P, bool ADJ, int TH, int TW, int VS, int HS, int T>
__global__ void __launch_bounds__(T) k_pyr_one2two(const float* __restrict__ x, float* __restrict__ ya,
                                                    float* __restrict__ yb, ImgStrides st3, int L, int m) {
  using F = Pair<P>;
  constexpr int W = TW + 2 * HALO;
  constexpr int WP = W | 1;  // odd row stride for the intermediates
  __shared__ float tile[TileCfg<TH, TW, T>::PADDED];
  __shared__ float tv[PRANK][TH * WP];
  int bx, by, ph, img;
  pyr_block(m * m, bx, by, ph, img);
  const int pr = ph / m, pc = ph - pr * m;
  x += (size_t)img * st3.a;
  ya += (size_t)img * st3.b;
  yb += (size_t)img * st3.c;
  const int E = out_side(L, m, ADJ);
  const int nout = E / m;
  const int R0 = by * TH, C0 = bx * TW;
  const int shift = ADJ ? HALO : 0;  // extended row R' sits on input row R' - HALO
  float v[TileCfg<TH, TW, T>::NIT];
  tile_fetch<TH, TW, T, ADJ>(v, x, L, m, pr, pc, R0 - shift, C0 - shift);
  __builtin_amdgcn_sched_barrier(0);  // the tile's loads in flight together
  tile_store<TH, TW, T>(tile, v, 0.0f, SHRINK_NONE, 1.0f);
  __syncthreads();
  // shared vertical passes: tv[i] = u_i (x) tile, every tile column
  for (int it = threadIdx.x; it < W * (TH / VS); it += T) {
    const int cc = it % W, st = it / W;
    float col[VS + 2 * HALO];
#pragma unroll
    for (int i = 0; i < VS + 2 * HALO; ++i) col[i] = tile[(st * VS + i) * W + cc];
#pragma unroll
    for (int k = 0; k < PRANK; ++k) {
      float t[VS];
#pragma unroll
      for (int i = 0; i < VS; ++i) t[i] = 0.f;
      vacc<VS>(col, F::U[k], t);
#pragma unroll
      for (int i = 0; i < VS; ++i) tv[k][(st * VS + i) * WP + cc] = t[i];
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < TH * (TW / HS); it += T) {
    const int r = it % TH, st = it / TH;
    float a0[HS], a1[HS];
#pragma unroll
    for (int j = 0; j < HS; ++j) {
      a0[j] = 0.f;
      a1[j] = tile[(r + HALO) * W + st * HS + j + HALO];  // identity term of F2
    }
#pragma unroll
    for (int k = 0; k < PRANK; ++k) {
      float row[HS + 2 * HALO];
#pragma unroll
      for (int i = 0; i < HS + 2 * HALO; ++i) row[i] = tv[k][r * WP + st * HS + i];
      hacc<HS>(row, F::A[k], a0);
      hacc<HS>(row, F::B[k], a1);
    }
    const int R = R0 + r;
    if (R < nout) {
      const size_t rowoff = (size_t)(pr + m * R) * E;
#pragma unroll
      for (int j = 0; j < HS; ++j) {
        const int Cc = C0 + st * HS + j;
        if (Cc < nout) {
          const size_t o = rowoff + pc + (size_t)m * Cc;
          ya[o] = a0[j];
          yb[o] = a1[j];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two inputs, one output with pair P:
//   out = F1 (x) (sa * keep(xa, th)) + F2 (x) (sb * xb).
//   forward reconstruction level: P = 1 (g0, g1)
//   adjoint of a decomposition level (ADJ): P = 0 (h0, h1) on the
//   zero-padded gradients, output on the extended grid.
template <int P, bool ADJ, int TH, int TW, int VS, int HS, int T>
__global__ void __launch_bounds__(T) k_pyr_two2one(const float* __restrict__ xa, const float* __restrict__ xb,
                                                    float* __restrict__ out, ImgStrides st3, int L, int m,
                                                    LowThresh lt, float sa, float sb) {
  using F = Pair<P>;
  constexpr int W = TW + 2 * HALO;
  constexpr int WP = W | 1;
  __shared__ float tile[2][TileCfg<TH, TW, T>::PADDED];
  __shared__ float tv[PRANK][TH * WP];
  int bx, by, ph, img;
  pyr_block(m * m, bx, by, ph, img);
  const int pr = ph / m, pc = ph - pr * m;
  xa += (size_t)img * st3.a;
  xb += (size_t)img * st3.b;
  out += (size_t)img * st3.c;
  const float th = low_value(lt, img);
  const int E = out_side(L, m, ADJ);
  const int nout = E / m;
  const int R0 = by * TH, C0 = bx * TW;
  const int shift = ADJ ? HALO : 0;
  float va[TileCfg<TH, TW, T>::NIT], vb[TileCfg<TH, TW, T>::NIT];
  tile_fetch<TH, TW, T, ADJ>(va, xa, L, m, pr, pc, R0 - shift, C0 - shift);
  tile_fetch<TH, TW, T, ADJ>(vb, xb, L, m, pr, pc, R0 - shift, C0 - shift);
  __builtin_amdgcn_sched_barrier(0);  // both tiles' loads in flight together
  tile_store<TH, TW, T>(tile[0], va, th, low_rule(lt), sa);
  tile_store<TH, TW, T>(tile[1], vb, 0.0f, SHRINK_NONE, sb);
  __syncthreads();
  // vertical passes of both inputs into the three shared intermediates:
  // tv[i] = a_i (x) tile0 + b_i (x) tile1
  for (int it = threadIdx.x; it < W * (TH / VS); it += T) {
    const int cc = it % W, st = it / W;
    float c0[VS + 2 * HALO], c1[VS + 2 * HALO];
#pragma unroll
    for (int i = 0; i < VS + 2 * HALO; ++i) {
      c0[i] = tile[0][(st * VS + i) * W + cc];
      c1[i] = tile[1][(st * VS + i) * W + cc];
    }
#pragma unroll
    for (int k = 0; k < PRANK; ++k) {
      float t[VS];
#pragma unroll
      for (int i = 0; i < VS; ++i) t[i] = 0.f;
      vacc<VS>(c0, F::A[k], t);
      vacc<VS>(c1, F::B[k], t);
#pragma unroll
      for (int i = 0; i < VS; ++i) tv[k][(st * VS + i) * WP + cc] = t[i];
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < TH * (TW / HS); it += T) {
    const int r = it % TH, st = it / TH;
    float a[HS];
#pragma unroll
    for (int j = 0; j < HS; ++j) a[j] = tile[1][(r + HALO) * W + st * HS + j + HALO];  // identity term of F2
#pragma unroll
    for (int k = 0; k < PRANK; ++k) {
      float row[HS + 2 * HALO];
#pragma unroll
      for (int i = 0; i < HS + 2 * HALO; ++i) row[i] = tv[k][r * WP + st * HS + i];
      hacc<HS>(row, F::U[k], a);
    }
    const int R = R0 + r;
    if (R < nout) {
      const size_t rowoff = (size_t)(pr + m * R) * E;
#pragma unroll
      for (int j = 0; j < HS; ++j) {
        const int Cc = C0 + st * HS + j;
        if (Cc < nout) out[rowoff + pc + (size_t)m * Cc] = a[j];
      }
    }
  }
}

// Transpose of the symmetric extension: out[s] = sum of Z[t'] over the
// extended grid points t' = t + H (H = HALO*m) with fold(t) = s, per axis.
__global__ void k_fold_add(const float* __restrict__ Z, float* __restrict__ out, int L, int H, long long zs,
                           long long os) {
  const int img = blockIdx.z;
  const int sc = blockIdx.x * blockDim.x + threadIdx.x, sr = blockIdx.y;
  if (sc >= L) return;
  const int E = L + 2 * H;
  const float* z = Z + (size_t)img * zs;
  int pre_r[8], pre_c[8], nr = 0, nc = 0;
  // preimages of s in [-H, L+H): t = s + 2Lk or t = -s-1 + 2Lk
  for (int k = -(H / (2 * L)) - 1; k <= H / (2 * L) + 1; ++k) {
    const int t1 = sr + 2 * L * k, t2 = -sr - 1 + 2 * L * k;
    if (t1 >= -H && t1 < L + H && nr < 8) pre_r[nr++] = t1 + H;
    if (t2 >= -H && t2 < L + H && nr < 8) pre_r[nr++] = t2 + H;
    const int u1 = sc + 2 * L * k, u2 = -sc - 1 + 2 * L * k;
    if (u1 >= -H && u1 < L + H && nc < 8) pre_c[nc++] = u1 + H;
    if (u2 >= -H && u2 < L + H && nc < 8) pre_c[nc++] = u2 + H;
  }
  float acc = 0.f;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) acc += z[(size_t)pre_r[i] * E + pre_c[j]];
  out[(size_t)img * os + (size_t)sr * L + sc] = acc;
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

template <int P, bool ADJ, int V>
static hipError_t one2two_v(const float* x, float* ya, float* yb, const ImgStrides& st, int nimg, int L, int m,
                            hipStream_t s) {
  using C = PyrCfg<V>;
  const int nout = (ADJ ? L + 2 * HALO * m : L) / m;
  dim3 grid((nout + C::TW - 1) / C::TW, (nout + C::TH - 1) / C::TH, m * m * nimg);
  hipLaunchKernelGGL((k_pyr_one2two<P, ADJ, C::TH, C::TW, C::VS, C::HS, C::T>), grid, dim3(C::T), 0, s, x, ya,
                     yb, st, L, m);
  return hipGetLastError();
}

template <int P, bool ADJ, int V>
static hipError_t two2one_v(const float* xa, const float* xb, float* out, const ImgStrides& st, int nimg, int L,
                            int m, const LowThresh& lt, float sa, float sb, hipStream_t s) {
  using C = PyrCfg<V>;
  const int nout = (ADJ ? L + 2 * HALO * m : L) / m;
  dim3 grid((nout + C::TW - 1) / C::TW, (nout + C::TH - 1) / C::TH, m * m * nimg);
  hipLaunchKernelGGL((k_pyr_two2one<P, ADJ, C::TH, C::TW, C::VS, C::HS, C::T>), grid, dim3(C::T), 0, s, xa, xb,
                     out, st, L, m, lt, sa, sb);
  return hipGetLastError();
}

static bool level_ok(int L, int m, int nimg) { return m >= 1 && L % m == 0 && nimg >= 1 && m * m * nimg <= 65535; }

hipError_t launch_atrous_dec(const float* x, float* ylow, float* yhigh, const ImgStrides& st, int nimg, int L, int m,
                             hipStream_t s, int variant) {
  if (!level_ok(L, m, nimg)) return hipErrorInvalidValue;
  switch (variant) {
    case 0: return one2two_v<0, false, 0>(x, ylow, yhigh, st, nimg, L, m, s);
    case 1: return one2two_v<0, false, 1>(x, ylow, yhigh, st, nimg, L, m, s);
    case 2: return one2two_v<0, false, 2>(x, ylow, yhigh, st, nimg, L, m, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_atrous_rec(const float* xlow, const float* yhigh, float* out, const ImgStrides& st, int nimg, int L,
                             int m, const LowThresh& lt, float sx, float sy, hipStream_t s, int variant) {
  if (!level_ok(L, m, nimg)) return hipErrorInvalidValue;
  switch (variant) {
    case 0: return two2one_v<1, false, 0>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    case 1: return two2one_v<1, false, 1>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    case 2: return two2one_v<1, false, 2>(xlow, yhigh, out, st, nimg, L, m, lt, sx, sy, s);
    default: return hipErrorInvalidValue;
  }
}

int adjoint_ext_side(int L, int m) { return L + 2 * HALO * m; }

hipError_t launch_atrous_dec_adjoint(const float* glow, const float* ghigh, float* gx, float* zext,
                                     const ImgStrides& st, long long zstride, int nimg, int L, int m, hipStream_t s) {
  if (!level_ok(L, m, nimg)) return hipErrorInvalidValue;
  // Z = C_h0^T glow + C_h1^T ghigh on the extended grid, then fold-add
  ImgStrides sz{st.a, st.b, zstride};
  hipError_t e = two2one_v<0, true, 0>(glow, ghigh, zext, sz, nimg, L, m, LowThresh{0, 0.f, nullptr, 0, 0}, 1.0f,
                                           1.0f, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fold_add, dim3((L + 255) / 256, L, nimg), dim3(256), 0, s, zext, gx, L, HALO * m, zstride,
                     st.c);
  return hipGetLastError();
}

hipError_t launch_atrous_rec_adjoint(const float* gout, float* glow, float* ghigh, float* zext, const ImgStrides& st,
                                     long long zstride, int nimg, int L, int m, hipStream_t s) {
  if (!level_ok(L, m, nimg)) return hipErrorInvalidValue;
  // Za = C_g0^T gout, Zb = C_g1^T gout on the extended grid, then fold-add each
  const long long E = adjoint_ext_side(L, m);
  float* za = zext;
  float* zb = zext + E * E;
  ImgStrides sz{st.a, zstride, zstride};
  hipError_t e = one2two_v<1, true, 0>(gout, za, zb, sz, nimg, L, m, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fold_add, dim3((L + 255) / 256, L, nimg), dim3(256), 0, s, za, glow, L, HALO * m, zstride,
                     st.b);
  hipLaunchKernelGGL(k_fold_add, dim3((L + 255) / 256, L, nimg), dim3(256), 0, s, zb, ghigh, L, HALO * m, zstride,
                     st.c);
  return hipGetLastError();
}

}  // namespace shx
