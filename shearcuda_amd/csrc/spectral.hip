// FFT-domain directional filtering of the NSST (2D), gfx950.
//
// Replaces the reference's cuFFT R2C/C2R plans + modulateAndNormalizeMany /
// modulateConjAndNormalize + sumVectors
// (shearcuda/src_cuda/shear_trans_cuda.cpp:220-287,
//  shearcuda/src_cuda/inverse_shear_cuda.cpp:222-286,
//  shearcuda/shear_cuda.cu:39-137,255-290).
//
// Data layout (row-major, L a power of two, N = L/2):
//   real plane      : L x L floats
//   packed spectrum : L x N float2.  After a real FFT along rows, columns v=0
//                     and v=N are real; they share slot 0 as (re, im).  The
//                     column pass transforms that packed column as one
//                     complex sequence.
// Forward decomposition of one scale j with D_j directions:
//   rows_r2c  : Y1 = RFFT_rows(y_j)
//   cols_dec  : T_k = IFFT_cols(H_k . FFT_cols(Y1))      (one block per k, tile)
//   rows_c2r  : c_k = IRFFT_rows(T_k)
// Inverse of one scale:
//   rows_r2c  : U_k = RFFT_rows(thresh(c_k))
//   cols_rec  : V = IFFT_cols(sum_k conj(H_k) . FFT_cols(U_k))
//   rows_c2r  : y_j = IRFFT_rows(V)
// H_k holds the normalised filter spectrum times 1/L^2 (the two unnormalised
// inverse passes each scale by L).  For the packed column, H_k(u,0) and
// H_k(u,N) act through alpha = (H(u,0)+H(u,N))/2 (stored in slot 0) and
// beta = (H(u,0)-H(u,N))/2 (stored separately):
//   Q(u) = alpha(u) P(u) + beta(u) conj(P(-u)).
#include "fft_lds.h"
#include "spectral.h"

namespace shx {

void band_map_groups(BandMap& m, int gsize) {
  int g = 0;
  for (int j = 0; j < m.nscales; ++j) {
    m.group_first[j] = g;
    for (int b = m.first[j]; b < m.first[j + 1]; b += gsize) m.group_band0[g++] = (unsigned short)b;
  }
  m.group_first[m.nscales] = g;
  m.group_band0[g] = (unsigned short)m.nbands;
  m.ngroups = g;
}

__device__ __forceinline__ float* plane_ptr(const PlaneSet& s, int p, size_t plane_elems, int b) {
  int j = 0;
#pragma unroll 1
  while (j + 1 < s.ncells && p >= s.off[j + 1]) ++j;
  return s.cell[j] + (size_t)(p - s.off[j]) * plane_elems + (size_t)b * s.img_stride;
}

__device__ __forceinline__ float thresh_of(const ThreshSpec& t, int p, int b) {
  return t.mode == 1 ? t.arg[p] : t.dev[(size_t)b * t.dev_stride + t.dev_off + p];
}

// ---------------------------------------------------------------------------
// Row-pass building blocks on a block of ROWS rows held in LDS: element k of
// row r (N = L/2 complex per row) is linear element r*N + k, at slot
// lds_ix<RPAD>(r*N + k).

constexpr int RPAD = 5;  // row buffers: one pad slot per 32 elements
constexpr int CPAD = 4;  // column tiles: one pad slot per 16 elements

template <int E, int PADSH>
constexpr int lds_size() {
  return E + (E >> PADSH) + 1;
}

// Packed half spectrum of a real row from the FFT of its (even, odd) pairs;
// the row starts at linear element e0 of buf.
template <int N>
__device__ __forceinline__ float2 r2c_post(const float2* buf, int e0, int k, const float2* __restrict__ tw) {
  const float2 z = buf[lds_ix<RPAD>(e0 + k)];
  if (k == 0) return make_float2(z.x + z.y, z.x - z.y);
  const float2 zc = cconj(buf[lds_ix<RPAD>(e0 + N - k)]);
  const float2 a = cscale(cadd(z, zc), 0.5f);
  const float2 b = cscale(csub(z, zc), 0.5f);
  const float2 wb = cmul(tw[k], b);  // X = a - i w^k b
  return make_float2(a.x + wb.y, a.y - wb.x);
}

// FFT input (even + i odd, times L) of the inverse of a packed half spectrum.
template <int N>
__device__ __forceinline__ float2 c2r_pre(const float2* buf, int e0, int k, const float2* __restrict__ tw) {
  const float2 x = buf[lds_ix<RPAD>(e0 + k)];
  if (k == 0) return make_float2(x.x + x.y, x.x - x.y);
  const float2 xc = cconj(buf[lds_ix<RPAD>(e0 + N - k)]);
  const float2 fe = cadd(x, xc);
  const float2 fo = cmulc(csub(x, xc), tw[k]);  // * w^{-k}
  return make_float2(fe.x - fo.y, fe.y + fo.x);  // fe + i fo
}

// In-place map of every row element through f(e0, k) (all reads, barrier,
// all writes).
template <int ROWS, int N, int T, class F>
__device__ __forceinline__ void rows_map(float2* buf, F f) {
  constexpr int PER = (ROWS * N + T - 1) / T;
  float2 z[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * T;
    if ((ROWS * N) % T == 0 || idx < ROWS * N) z[e] = f(idx - idx % N, idx % N);
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * T;
    if ((ROWS * N) % T == 0 || idx < ROWS * N) buf[lds_ix<RPAD>(idx)] = z[e];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Row real-to-complex FFT (packed output), optional hard threshold on load.
// grid: (L / ROWS, nplanes, nimg).
template <int L, int ROWS, int T>
__global__ void __launch_bounds__(T) k_rows_r2c(PlaneSet in, PlaneSet out, ThreshSpec th,
                                                 const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  __shared__ float2 buf[lds_size<ROWS * N, RPAD>()];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe, b)) + (size_t)row0 * N;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe, b)) + (size_t)row0 * N;
  const bool use_th = th.mode != 0;
  const float t = use_th ? thresh_of(th, p, b) : 0.0f;
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) {
    float2 v = src[idx];
    if (use_th) {
      v.x = hard_keep(v.x, t);
      v.y = hard_keep(v.y, t);
    }
    buf[lds_ix<RPAD>(idx)] = v;
  }
  __syncthreads();
  fft_lds<N, 2, -1, ROWS, T, false, RPAD>(buf, N, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) dst[idx] = r2c_post<N>(buf, idx - idx % N, idx % N, tw);
}

// ---------------------------------------------------------------------------
// Row complex-to-real inverse FFT of a packed spectrum (unnormalised: x L),
// optionally of the sum of several input planes.
template <int L, int ROWS, int T, bool SUM>
__global__ void __launch_bounds__(T) k_rows_c2r(PlaneSet in, PlaneSet out, SumTable sum,
                                                 const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  __shared__ float2 buf[lds_size<ROWS * N, RPAD>()];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const size_t pe = (size_t)L * L;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe, b)) + (size_t)row0 * N;
  if constexpr (SUM) {
    const int q0 = sum.first[p], q1 = sum.first[p + 1];
#pragma unroll
    for (int idx = tid; idx < ROWS * N; idx += T) {
      float2 a = make_float2(0.f, 0.f);
      for (int q = q0; q < q1; ++q)
        a = cadd(a, reinterpret_cast<const float2*>(plane_ptr(in, q, pe, b))[(size_t)row0 * N + idx]);
      buf[lds_ix<RPAD>(idx)] = a;
    }
  } else {
    const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe, b)) + (size_t)row0 * N;
#pragma unroll
    for (int idx = tid; idx < ROWS * N; idx += T) buf[lds_ix<RPAD>(idx)] = src[idx];
  }
  __syncthreads();
  rows_map<ROWS, N, T>(buf, [&](int e0, int k) { return c2r_pre<N>(buf, e0, k, tw); });
  fft_lds<N, 2, +1, ROWS, T, false, RPAD>(buf, N, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) dst[idx] = buf[lds_ix<RPAD>(idx)];
}

// ---------------------------------------------------------------------------
// Denoise fusion of the decomposition's last row pass with the
// reconstruction's first: c = C2R(T) is stored (the coefficients the caller
// receives), then U = R2C(keep(c, th)) is formed from the same rows in LDS.
template <int L, int ROWS, int T>
__global__ void __launch_bounds__(T) k_rows_c2r_r2c(PlaneSet Tin, PlaneSet band, PlaneSet Uout, ThreshSpec th,
                                                     const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  __shared__ float2 buf[lds_size<ROWS * N, RPAD>()];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(Tin, p, pe, b)) + (size_t)row0 * N;
  float2* cdst = reinterpret_cast<float2*>(plane_ptr(band, p, pe, b)) + (size_t)row0 * N;
  float2* udst = reinterpret_cast<float2*>(plane_ptr(Uout, p, pe, b)) + (size_t)row0 * N;
  const float t = th.mode != 0 ? thresh_of(th, p, b) : -1.0f;
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) buf[lds_ix<RPAD>(idx)] = src[idx];
  __syncthreads();
  rows_map<ROWS, N, T>(buf, [&](int e0, int k) { return c2r_pre<N>(buf, e0, k, tw); });
  fft_lds<N, 2, +1, ROWS, T, false, RPAD>(buf, N, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) {
    float2& v = buf[lds_ix<RPAD>(idx)];
    const float2 c = v;
    cdst[idx] = c;
    if (th.mode != 0) v = make_float2(hard_keep(c.x, t), hard_keep(c.y, t));
  }
  __syncthreads();
  fft_lds<N, 2, -1, ROWS, T, false, RPAD>(buf, N, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) udst[idx] = r2c_post<N>(buf, idx - idx % N, idx % N, tw);
}

// ---------------------------------------------------------------------------
// Column-tile helpers: tile of L rows x CT complex columns in LDS, element
// (u, c) at slot lds_ix<CPAD>(u*CT + c).
template <int L, int CT, int T>
__device__ __forceinline__ void tile_load(float2* buf, const float2* __restrict__ src) {
  constexpr int N = L / 2;
#pragma unroll
  for (int idx = threadIdx.x; idx < L * CT; idx += T) buf[lds_ix<CPAD>(idx)] = src[(size_t)(idx / CT) * N + idx % CT];
}

template <int L, int CT, int T>
__device__ __forceinline__ void tile_store(float2* __restrict__ dst, const float2* buf) {
  constexpr int N = L / 2;
#pragma unroll
  for (int idx = threadIdx.x; idx < L * CT; idx += T) dst[(size_t)(idx / CT) * N + idx % CT] = buf[lds_ix<CPAD>(idx)];
}

// Filter a transformed column tile by band p's spectrum (CONJ: by its
// conjugate).  Slot 0 of tile c0 == 0 is the packed (DC, Nyquist) column.
template <int L, int CT, bool CONJ>
__device__ __forceinline__ float2 filt(const float2* buf, int idx, int c0, const float2* __restrict__ H,
                                       const float2* __restrict__ Hb) {
  constexpr int N = L / 2;
  const int u = idx / CT, c = idx % CT;
  const float2 P = buf[lds_ix<CPAD>(idx)];
  const float2 h = H[(size_t)u * N + c];
  if (c0 + c == 0) {
    const float2 Pm = cconj(buf[lds_ix<CPAD>(((L - u) & (L - 1)) * CT)]);
    return CONJ ? cadd(cmulc(P, h), cmulc(Pm, Hb[u])) : cadd(cmul(h, P), cmul(Hb[u], Pm));
  }
  return CONJ ? cmulc(P, h) : cmul(h, P);
}

// Column pass of the decomposition: one block per (column tile, band, image).
template <int L, int CT, int T>
__global__ void __launch_bounds__(T) k_cols_dec(const float2* __restrict__ Y1, long long y1_stride,
                                                 float2* __restrict__ Tout, long long t_stride,
                                                 const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                 BandMap map, const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int PER = (L * CT + T - 1) / T;
  __shared__ float2 buf[lds_size<L * CT, CPAD>()];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CT;
  const int p = blockIdx.y, b = blockIdx.z;
  const int j = map.scale_of[p];
  tile_load<L, CT, T>(buf, Y1 + (size_t)b * y1_stride + (size_t)j * L * N + c0);
  __syncthreads();
  fft_lds<L, 1, -1, CT, T, true, CPAD>(buf, 1, CT, tw, tid);
  const float2* H = Hmain + (size_t)p * L * N + c0;
  const float2* Hb = Hbeta + (size_t)p * L;
  float2 q[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) q[e] = filt<L, CT, false>(buf, idx, c0, H, Hb);
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) buf[lds_ix<CPAD>(idx)] = q[e];
  }
  __syncthreads();
  fft_lds<L, 1, +1, CT, T, true, CPAD>(buf, 1, CT, tw, tid);
  tile_store<L, CT, T>(Tout + (size_t)b * t_stride + (size_t)p * L * N + c0, buf);
}

// Column pass of the reconstruction: one block per (column tile, direction
// group, image) forms IFFT_cols(sum_{k in group} conj(H_k) . FFT_cols(U_k));
// the row pass that follows sums the groups of a scale.
template <int L, int CT, int T>
__global__ void __launch_bounds__(T) k_cols_rec(const float2* __restrict__ U, long long u_stride,
                                                 float2* __restrict__ Q, long long q_stride,
                                                 const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                 BandMap map, const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int PER = (L * CT + T - 1) / T;
  __shared__ float2 buf[lds_size<L * CT, CPAD>()];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CT;
  const int g = blockIdx.y, b = blockIdx.z;
  float2 acc[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) acc[e] = make_float2(0.f, 0.f);
  const int p0 = map.group_band0[g], p1 = map.group_band0[g + 1];
  for (int p = p0; p < p1; ++p) {
    tile_load<L, CT, T>(buf, U + (size_t)b * u_stride + (size_t)p * L * N + c0);
    __syncthreads();
    fft_lds<L, 1, -1, CT, T, true, CPAD>(buf, 1, CT, tw, tid);
    const float2* H = Hmain + (size_t)p * L * N + c0;
    const float2* Hb = Hbeta + (size_t)p * L;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * T;
      if ((L * CT) % T == 0 || idx < L * CT) acc[e] = cadd(acc[e], filt<L, CT, true>(buf, idx, c0, H, Hb));
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) buf[lds_ix<CPAD>(idx)] = acc[e];
  }
  __syncthreads();
  fft_lds<L, 1, +1, CT, T, true, CPAD>(buf, 1, CT, tw, tid);
  tile_store<L, CT, T>(Q + (size_t)b * q_stride + (size_t)g * L * N + c0, buf);
}

// ---------------------------------------------------------------------------
// Host-side launchers, one instantiation per supported L.
template <int L>
struct SpectralCfg {
  static constexpr int N = L / 2;
  static constexpr int ROWS = (L >= 512) ? 4 : (2048 / L < 16 ? 2048 / L : 16);
  static constexpr int RT = 256;
  static constexpr int CT = (N >= 8) ? 8 : N;
  static constexpr int CTT = (L >= 512) ? 512 : 256;
};

template <int L>
struct Launch {
  using C = SpectralCfg<L>;
  static hipError_t rows_r2c(const PlaneSet& in, const PlaneSet& out, int np, int nimg, const ThreshSpec& th,
                             const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_rows_r2c<L, C::ROWS, C::RT>), dim3(L / C::ROWS, np, nimg), dim3(C::RT), 0, s, in, out, th,
                       tw);
    return hipGetLastError();
  }
  static hipError_t rows_c2r(const PlaneSet& in, const PlaneSet& out, int np, int nimg, const int* sum_first,
                             const float2* tw, hipStream_t s) {
    SumTable st{};
    if (sum_first) {
      for (int i = 0; i <= np; ++i) st.first[i] = sum_first[i];
      hipLaunchKernelGGL((k_rows_c2r<L, C::ROWS, C::RT, true>), dim3(L / C::ROWS, np, nimg), dim3(C::RT), 0, s, in,
                         out, st, tw);
    } else {
      hipLaunchKernelGGL((k_rows_c2r<L, C::ROWS, C::RT, false>), dim3(L / C::ROWS, np, nimg), dim3(C::RT), 0, s, in,
                         out, st, tw);
    }
    return hipGetLastError();
  }
  static hipError_t rows_c2r_r2c(const PlaneSet& Tin, const PlaneSet& band, const PlaneSet& U, int np, int nimg,
                                 const ThreshSpec& th, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_rows_c2r_r2c<L, C::ROWS, C::RT>), dim3(L / C::ROWS, np, nimg), dim3(C::RT), 0, s, Tin,
                       band, U, th, tw);
    return hipGetLastError();
  }
  static hipError_t cols_dec(const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                             const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_cols_dec<L, C::CT, C::CTT>), dim3(C::N / C::CT, map.nbands, nimg), dim3(C::CTT), 0, s, Y1,
                       ys, To, ts, Hm, Hb, map, tw);
    return hipGetLastError();
  }
  static hipError_t cols_rec(const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                             const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_cols_rec<L, C::CT, C::CTT>), dim3(C::N / C::CT, map.ngroups, nimg), dim3(C::CTT), 0, s, U,
                       us, Q, qs, Hm, Hb, map, tw);
    return hipGetLastError();
  }
};

#define SHX_BY_L(L, EXPR_T)                       \
  switch (L) {                                    \
    case 32: { using LT = Launch<32>; return EXPR_T; }     \
    case 64: { using LT = Launch<64>; return EXPR_T; }     \
    case 128: { using LT = Launch<128>; return EXPR_T; }   \
    case 256: { using LT = Launch<256>; return EXPR_T; }   \
    case 512: { using LT = Launch<512>; return EXPR_T; }   \
    case 1024: { using LT = Launch<1024>; return EXPR_T; } \
    default: return hipErrorInvalidValue;         \
  }

hipError_t launch_rows_r2c(int L, const PlaneSet& in, const PlaneSet& out, int np, int nimg, const ThreshSpec& th,
                           const float2* tw, hipStream_t s) {
  SHX_BY_L(L, LT::rows_r2c(in, out, np, nimg, th, tw, s))
}

hipError_t launch_rows_c2r(int L, const PlaneSet& in, const PlaneSet& out, int np, int nimg, const int* sum_first,
                           const float2* tw, hipStream_t s) {
  if (sum_first && np > SHX_MAX_SCALES) return hipErrorInvalidValue;
  SHX_BY_L(L, LT::rows_c2r(in, out, np, nimg, sum_first, tw, s))
}

hipError_t launch_rows_c2r_r2c(int L, const PlaneSet& Tin, const PlaneSet& band, const PlaneSet& U, int np, int nimg,
                               const ThreshSpec& th, const float2* tw, hipStream_t s) {
  SHX_BY_L(L, LT::rows_c2r_r2c(Tin, band, U, np, nimg, th, tw, s))
}

hipError_t launch_cols_dec(int L, const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                           const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
  SHX_BY_L(L, LT::cols_dec(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s))
}

hipError_t launch_cols_rec(int L, const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                           const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
  SHX_BY_L(L, LT::cols_rec(U, us, Q, qs, Hm, Hb, map, nimg, tw, s))
}

}  // namespace shx
