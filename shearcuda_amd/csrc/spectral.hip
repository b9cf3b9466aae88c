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

__device__ __forceinline__ int plane_cell(const PlaneSet& s, int p) {
  int j = 0;
#pragma unroll 1
  while (j + 1 < s.ncells && p >= s.off[j + 1]) ++j;
  return j;
}

__device__ __forceinline__ float* plane_ptr(const PlaneSet& s, int p, size_t plane_elems) {
  const int j = plane_cell(s, p);
  return s.cell[j] + (size_t)(p - s.off[j]) * plane_elems;
}

// ---------------------------------------------------------------------------
// Row real-to-complex FFT (packed output), optional hard threshold on load.
// grid: (L / ROWS, nplanes).  LDS: ROWS x (N + 1) float2.
template <int L, int ROWS, int T>
__global__ void __launch_bounds__(T) k_rows_r2c(PlaneSet in, PlaneSet out, BandThresh th, int use_th,
                                                 const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int BS = N + 1;
  __shared__ float2 buf[ROWS * BS];
  const int tid = threadIdx.x;
  const int p = blockIdx.y;
  const int row0 = blockIdx.x * ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe)) + (size_t)row0 * N;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe)) + (size_t)row0 * N;
  const float t = use_th ? th.v[p] : 0.0f;
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) {
    const int r = idx / N, i = idx % N;
    float2 v = src[idx];
    if (use_th) {
      v.x = hard_keep(v.x, t);
      v.y = hard_keep(v.y, t);
    }
    buf[r * BS + i] = v;
  }
  __syncthreads();
  fft_lds<N, 2, -1, ROWS, T, false>(buf, BS, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) {
    const int r = idx / N, k = idx % N;
    const float2 z = buf[r * BS + k];
    float2 o;
    if (k == 0) {
      o = make_float2(z.x + z.y, z.x - z.y);
    } else {
      const float2 zc = cconj(buf[r * BS + (N - k)]);
      const float2 a = cscale(cadd(z, zc), 0.5f);
      const float2 b = cscale(csub(z, zc), 0.5f);
      // X = a - i * w^k * b
      const float2 wb = cmul(tw[k], b);
      o = make_float2(a.x + wb.y, a.y - wb.x);
    }
    dst[idx] = o;
  }
}

// ---------------------------------------------------------------------------
// Row complex-to-real inverse FFT of a packed spectrum (unnormalised: x L).
template <int L, int ROWS, int T>
__global__ void __launch_bounds__(T) k_rows_c2r(PlaneSet in, PlaneSet out, const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int BS = N + 1;
  constexpr int PER = (ROWS * N + T - 1) / T;
  __shared__ float2 buf[ROWS * BS];
  const int tid = threadIdx.x;
  const int p = blockIdx.y;
  const int row0 = blockIdx.x * ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe)) + (size_t)row0 * N;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe)) + (size_t)row0 * N;
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) buf[(idx / N) * BS + idx % N] = src[idx];
  __syncthreads();
  float2 z[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((ROWS * N) % T == 0 || idx < ROWS * N) {
      const int r = idx / N, k = idx % N;
      const float2 x = buf[r * BS + k];
      if (k == 0) {
        z[e] = make_float2(x.x + x.y, x.x - x.y);
      } else {
        const float2 xc = cconj(buf[r * BS + (N - k)]);
        const float2 fe = cadd(x, xc);
        const float2 fo = cmulc(csub(x, xc), tw[k]);  // * w^{-k}
        z[e] = make_float2(fe.x - fo.y, fe.y + fo.x);  // fe + i*fo
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((ROWS * N) % T == 0 || idx < ROWS * N) buf[(idx / N) * BS + idx % N] = z[e];
  }
  __syncthreads();
  fft_lds<N, 2, +1, ROWS, T, false>(buf, BS, 1, tw, tid);
#pragma unroll
  for (int idx = tid; idx < ROWS * N; idx += T) dst[idx] = buf[(idx / N) * BS + idx % N];
}

// ---------------------------------------------------------------------------
// Column pass of the decomposition: one block per (column tile, band).
// grid: (N / CT, nbands).  LDS: L x CT float2.
template <int L, int CT, int T>
__global__ void __launch_bounds__(T) k_cols_dec(const float2* __restrict__ Y1, float2* __restrict__ Tout,
                                                 const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                 BandMap map, const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int PER = (L * CT + T - 1) / T;
  __shared__ float2 buf[L * CT];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CT;
  const int p = blockIdx.y;
  const int j = map.scale_of[p];
  const float2* src = Y1 + (size_t)j * L * N + c0;
#pragma unroll
  for (int idx = tid; idx < L * CT; idx += T) buf[idx] = src[(size_t)(idx / CT) * N + idx % CT];
  __syncthreads();
  fft_lds<L, 1, -1, CT, T, true>(buf, 1, CT, tw, tid);
  const float2* H = Hmain + (size_t)p * L * N + c0;
  float2 q[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) {
      const int u = idx / CT, c = idx % CT;
      const float2 P = buf[idx];
      const float2 h = H[(size_t)u * N + c];
      if (c0 + c == 0) {
        const float2 Pm = cconj(buf[((L - u) & (L - 1)) * CT]);
        q[e] = cadd(cmul(h, P), cmul(Hbeta[(size_t)p * L + u], Pm));
      } else {
        q[e] = cmul(h, P);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) buf[idx] = q[e];
  }
  __syncthreads();
  fft_lds<L, 1, +1, CT, T, true>(buf, 1, CT, tw, tid);
  float2* dst = Tout + (size_t)p * L * N + c0;
#pragma unroll
  for (int idx = tid; idx < L * CT; idx += T) dst[(size_t)(idx / CT) * N + idx % CT] = buf[idx];
}

// ---------------------------------------------------------------------------
// Column pass of the reconstruction: one block per (column tile, scale),
// accumulating conj(H_k) . FFT_cols(U_k) over the scale's directions.
template <int L, int CT, int T>
__global__ void __launch_bounds__(T) k_cols_rec(const float2* __restrict__ U, float2* __restrict__ V,
                                                 const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                 BandMap map, const float2* __restrict__ tw) {
  constexpr int N = L / 2;
  constexpr int PER = (L * CT + T - 1) / T;
  __shared__ float2 buf[L * CT];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CT;
  const int j = blockIdx.y;
  float2 acc[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) acc[e] = make_float2(0.f, 0.f);
  for (int p = map.first[j]; p < map.first[j + 1]; ++p) {
    const float2* src = U + (size_t)p * L * N + c0;
#pragma unroll
    for (int idx = tid; idx < L * CT; idx += T) buf[idx] = src[(size_t)(idx / CT) * N + idx % CT];
    __syncthreads();
    fft_lds<L, 1, -1, CT, T, true>(buf, 1, CT, tw, tid);
    const float2* H = Hmain + (size_t)p * L * N + c0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + e * T;
      if ((L * CT) % T == 0 || idx < L * CT) {
        const int u = idx / CT, c = idx % CT;
        const float2 P = buf[idx];
        const float2 h = H[(size_t)u * N + c];
        if (c0 + c == 0) {
          const float2 Pm = cconj(buf[((L - u) & (L - 1)) * CT]);
          acc[e] = cadd(acc[e], cadd(cmulc(P, h), cmulc(Pm, Hbeta[(size_t)p * L + u])));
        } else {
          acc[e] = cadd(acc[e], cmulc(P, h));
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = tid + e * T;
    if ((L * CT) % T == 0 || idx < L * CT) buf[idx] = acc[e];
  }
  __syncthreads();
  fft_lds<L, 1, +1, CT, T, true>(buf, 1, CT, tw, tid);
  float2* dst = V + (size_t)j * L * N + c0;
#pragma unroll
  for (int idx = tid; idx < L * CT; idx += T) dst[(size_t)(idx / CT) * N + idx % CT] = buf[idx];
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
static hipError_t launch_rows_r2c_t(const PlaneSet& in, const PlaneSet& out, int nplanes, const BandThresh* th,
                                    const float2* tw, hipStream_t s) {
  using C = SpectralCfg<L>;
  static_assert(L % C::ROWS == 0, "rows per block must divide L");
  BandThresh dummy;
  hipLaunchKernelGGL((k_rows_r2c<L, C::ROWS, C::RT>), dim3(L / C::ROWS, nplanes), dim3(C::RT), 0, s, in, out,
                     th ? *th : dummy, th ? 1 : 0, tw);
  return hipGetLastError();
}

template <int L>
static hipError_t launch_rows_c2r_t(const PlaneSet& in, const PlaneSet& out, int nplanes, const float2* tw,
                                    hipStream_t s) {
  using C = SpectralCfg<L>;
  hipLaunchKernelGGL((k_rows_c2r<L, C::ROWS, C::RT>), dim3(L / C::ROWS, nplanes), dim3(C::RT), 0, s, in, out, tw);
  return hipGetLastError();
}

template <int L>
static hipError_t launch_cols_dec_t(const float2* Y1, float2* Tout, const float2* Hm, const float2* Hb,
                                    const BandMap& map, const float2* tw, hipStream_t s) {
  using C = SpectralCfg<L>;
  hipLaunchKernelGGL((k_cols_dec<L, C::CT, C::CTT>), dim3(C::N / C::CT, map.nbands), dim3(C::CTT), 0, s, Y1, Tout,
                     Hm, Hb, map, tw);
  return hipGetLastError();
}

template <int L>
static hipError_t launch_cols_rec_t(const float2* U, float2* V, const float2* Hm, const float2* Hb,
                                    const BandMap& map, const float2* tw, hipStream_t s) {
  using C = SpectralCfg<L>;
  hipLaunchKernelGGL((k_cols_rec<L, C::CT, C::CTT>), dim3(C::N / C::CT, map.nscales), dim3(C::CTT), 0, s, U, V, Hm,
                     Hb, map, tw);
  return hipGetLastError();
}

hipError_t launch_rows_r2c(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, const BandThresh* th,
                           const float2* tw, hipStream_t s) {
#define CALL(LL) launch_rows_r2c_t<LL>(in, out, nplanes, th, tw, s)
  switch (L) {
    case 32: return CALL(32);
    case 64: return CALL(64);
    case 128: return CALL(128);
    case 256: return CALL(256);
    case 512: return CALL(512);
    case 1024: return CALL(1024);
    default: return hipErrorInvalidValue;
  }
#undef CALL
}

hipError_t launch_rows_c2r(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, const float2* tw,
                           hipStream_t s) {
#define CALL(LL) launch_rows_c2r_t<LL>(in, out, nplanes, tw, s)
  switch (L) {
    case 32: return CALL(32);
    case 64: return CALL(64);
    case 128: return CALL(128);
    case 256: return CALL(256);
    case 512: return CALL(512);
    case 1024: return CALL(1024);
    default: return hipErrorInvalidValue;
  }
#undef CALL
}

hipError_t launch_cols_dec(int L, const float2* Y1, float2* Tout, const float2* Hm, const float2* Hb,
                           const BandMap& map, const float2* tw, hipStream_t s) {
#define CALL(LL) launch_cols_dec_t<LL>(Y1, Tout, Hm, Hb, map, tw, s)
  switch (L) {
    case 32: return CALL(32);
    case 64: return CALL(64);
    case 128: return CALL(128);
    case 256: return CALL(256);
    case 512: return CALL(512);
    case 1024: return CALL(1024);
    default: return hipErrorInvalidValue;
  }
#undef CALL
}

hipError_t launch_cols_rec(int L, const float2* U, float2* V, const float2* Hm, const float2* Hb,
                           const BandMap& map, const float2* tw, hipStream_t s) {
#define CALL(LL) launch_cols_rec_t<LL>(U, V, Hm, Hb, map, tw, s)
  switch (L) {
    case 32: return CALL(32);
    case 64: return CALL(64);
    case 128: return CALL(128);
    case 256: return CALL(256);
    case 512: return CALL(512);
    case 1024: return CALL(1024);
    default: return hipErrorInvalidValue;
  }
#undef CALL
}

}  // namespace shx
