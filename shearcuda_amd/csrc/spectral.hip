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
// Every transform is a two-pass register FFT (fft_reg.h): the first pass reads
// global memory, the second writes it, with one padded LDS transpose between
// them; forward/inverse pairs chain in registers.
// H_k holds the normalised filter spectrum times 1/L^2 (the two unnormalised
// inverse passes each scale by L).  For the packed column, H_k(u,0) and
// H_k(u,N) act through alpha = (H(u,0)+H(u,N))/2 (stored in slot 0) and
// beta = (H(u,0)-H(u,N))/2 (stored separately):
//   Q(u) = alpha(u) P(u) + beta(u) conj(P(-u)).
#include <algorithm>
#include <vector>

#include "fft3.h"
#include "fft_reg.h"
#include "pyramid.h"
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
  // bands: grouped by filter plane when every plane is read by the same number of bands
  int nfilt = 0;
  for (int p = 0; p < m.nbands; ++p) nfilt = std::max(nfilt, (int)m.filt_of[p] + 1);
  std::vector<int> cnt(nfilt, 0);
  for (int p = 0; p < m.nbands; ++p) ++cnt[m.filt_of[p]];
  bool uniform = nfilt > 0 && m.nbands % nfilt == 0;
  for (int f = 0; uniform && f < nfilt; ++f) uniform = cnt[f] == m.nbands / nfilt;
  m.share = uniform ? m.nbands / nfilt : 1;
  int k = 0;
  if (m.share > 1) {
    for (int f = 0; f < nfilt; ++f)
      for (int p = 0; p < m.nbands; ++p)
        if (m.filt_of[p] == f) m.order[k++] = (unsigned char)p;
  } else {
    for (int p = 0; p < m.nbands; ++p) m.order[p] = (unsigned char)p;
  }
  // groups: two groups share when they read the same sequence of filter planes
  auto same = [&](int a, int b) {
    const int na = m.group_band0[a + 1] - m.group_band0[a], nb = m.group_band0[b + 1] - m.group_band0[b];
    if (na != nb) return false;
    for (int i = 0; i < na; ++i)
      if (m.filt_of[m.group_band0[a] + i] != m.filt_of[m.group_band0[b] + i]) return false;
    return true;
  };
  std::vector<int> cls(g, -1);
  std::vector<int> rep;
  for (int a = 0; a < g; ++a) {
    for (int c = 0; c < (int)rep.size() && cls[a] < 0; ++c)
      if (same(rep[c], a)) cls[a] = c;
    if (cls[a] < 0) {
      cls[a] = (int)rep.size();
      rep.push_back(a);
    }
  }
  const int ncls = (int)rep.size();
  std::vector<int> gcnt(ncls, 0);
  for (int a = 0; a < g; ++a) ++gcnt[cls[a]];
  bool guni = ncls > 0 && g % ncls == 0;
  for (int c = 0; guni && c < ncls; ++c) guni = gcnt[c] == g / ncls;
  m.gshare = guni ? g / ncls : 1;
  k = 0;
  if (m.gshare > 1) {
    for (int c = 0; c < ncls; ++c)
      for (int a = 0; a < g; ++a)
        if (cls[a] == c) m.gorder[k++] = (unsigned short)a;
  } else {
    for (int a = 0; a < g; ++a) m.gorder[a] = (unsigned short)a;
  }
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

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src), __shfl(v.y, src));
}

template <int E, int PADSH>
constexpr int lds_size() {
  return E + (E >> PADSH) + 1;
}

// Packed half spectrum P[k] of a real row from X = FFT(even + i odd) at k and
// N-k (k > 0): X_real[k] = (X[k] + X*[N-k])/2 - i w_L^k (X[k] - X*[N-k])/2.
__device__ __forceinline__ float2 r2c_post(float2 z, float2 zm, float2 w) {
  const float2 zc = cconj(zm);
  const float2 a = cscale(cadd(z, zc), 0.5f);
  const float2 b = cscale(csub(z, zc), 0.5f);
  const float2 wb = cmul(w, b);
  return make_float2(a.x + wb.y, a.y - wb.x);
}

// Inverse: FFT input (even + i odd, times L) from P at k and N-k (k > 0).
__device__ __forceinline__ float2 c2r_pre(float2 x, float2 xm, float2 w) {
  const float2 xc = cconj(xm);
  const float2 fe = cadd(x, xc);
  const float2 fo = cmulc(csub(x, xc), w);  // * w^{-k}
  return make_float2(fe.x - fo.y, fe.y + fo.x);
}

// ---------------------------------------------------------------------------
// Row passes.  Rows are transforms of N = L/2 complex points (the real row
// as (even, odd) pairs), split N = R1 x R2.  A block holds ROWS rows and
// T = ROWS * R2 threads; pass A uses ROWS * R1 of them.
template <int L, int ROWS_ = 0>
struct RowCfg {
  static constexpr int N = L / 2;
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int ROWS = ROWS_ ? ROWS_ : ((256 / R2) < L ? (256 / R2) : L);
  static constexpr int T = ROWS * R2;
  static constexpr int PADSH = ilog2(R1);
  static constexpr int LDS = lds_size<ROWS * N, PADSH>();
};

// Twiddles of row_fwd_from_A, loaded by the caller ahead of time: the pass-A
// twiddles of index n1 and w_L^k2.  (A store issued before a load also counts
// against the load's wait on this architecture, so loading these late would
// wait for earlier stores too.)
template <int L, int R1, int R2>
struct RowFwdTw {
  PassTw<R2> pt;
  float2 w0;
};
template <int L, int R1, int R2>
__device__ __forceinline__ RowFwdTw<L, R1, R2> row_fwd_tw(int n1, int k2, const float2* __restrict__ tw) {
  RowFwdTw<L, R1, R2> t;
  passA_tw<R2, -1, 2>(t.pt, n1, tw);
  t.w0 = tw[k2];
  return t;
}

// Forward real-row FFT from pass-A registers (thread (r, n1) holds
// z[n1 + R1 n2]) to the packed half spectrum, stored to dst (row r).
// All threads of the block must call it.
template <int L, int R1, int R2, int PADSH>
__device__ __forceinline__ void row_fwd_from_A(float2 (&v)[R2], bool validA, int r, int n1, float2* buf,
                                               float2* __restrict__ dst_rows, bool validB, int rB, int k2,
                                               const RowFwdTw<L, R1, R2>& t) {
  constexpr int N = L / 2;
  if (validA) {
    // products of the preloaded twiddles formed here, after the data:
    // computed earlier they would make the first wait cover the twiddle loads
    passA_apply<R2, -1>(v, pin(t.pt));
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[lds_ix<PADSH>(r * N + n1 + R1 * k)] = v[k];
  }
  __syncthreads();
  float2 u[R1];
#pragma unroll
  for (int a = 0; a < R1; ++a) u[a] = make_float2(0.f, 0.f);
  if (validB) {
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[lds_ix<PADSH>(rB * N + a + R1 * k2)];
    dft<R1, -1>(u);
  }
  const int lane = threadIdx.x & 63;
  const int src = (lane & ~(R2 - 1)) | ((R2 - k2) & (R2 - 1));
  float2* drow = dst_rows + (size_t)rB * N;
  // w_L^k for k = k2 + R2 k1: the table value w_L^k2 times the compile-time
  // w_L^(R2 k1) (one load per thread instead of one per output).  The values
  // are computed on every lane and only the stores are predicated.
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const float2 sh = shfl2(u[R1 - 1 - k1], src);
    const int k = k2 + R2 * k1;
    const float2 zm = (k2 == 0) ? u[(R1 - k1) & (R1 - 1)] : sh;
    const float2 val = (k == 0) ? make_float2(u[0].x + u[0].y, u[0].x - u[0].y)
                                : r2c_post(u[k1], zm, ctw_mul<L / R2, -1>(t.w0, k1));
    if (validB) drow[k] = val;
  }
}

// As row_fwd_from_A, but the packed half spectrum stays in registers: thread
// (rB, k2) gets P[k1] = P(k2 + R2 k1).  All threads of the block must call it.
template <int L, int R1, int R2, int PADSH>
__device__ __forceinline__ void row_fwd_to_P(float2 (&v)[R2], bool validA, int r, int n1, float2* buf, int rB,
                                             int k2, const RowFwdTw<L, R1, R2>& t, float2 (&P)[R1]) {
  constexpr int N = L / 2;
  if (validA) {
    passA_apply<R2, -1>(v, pin(t.pt));
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[lds_ix<PADSH>(r * N + n1 + R1 * k)] = v[k];
  }
  __syncthreads();
  float2 u[R1];
#pragma unroll
  for (int a = 0; a < R1; ++a) u[a] = buf[lds_ix<PADSH>(rB * N + a + R1 * k2)];
  dft<R1, -1>(u);
  const int lane = threadIdx.x & 63;
  const int src = (lane & ~(R2 - 1)) | ((R2 - k2) & (R2 - 1));
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const float2 sh = shfl2(u[R1 - 1 - k1], src);
    const int k = k2 + R2 * k1;
    const float2 zm = (k2 == 0) ? u[(R1 - k1) & (R1 - 1)] : sh;
    P[k1] = (k == 0) ? make_float2(u[0].x + u[0].y, u[0].x - u[0].y)
                     : r2c_post(u[k1], zm, ctw_mul<L / R2, -1>(t.w0, k1));
  }
}

// P[k] and P[N-k] of one packed row for k = n1 + R1 n2.  Only P[k] is
// loaded: P[N-k] = P((R1 - n1) + R1 (R2 - 1 - n2)) is the partner lane's
// value (lanes of a row are the R1 consecutive lanes n1 = 0 .. R1-1, all
// active together), or for n1 = 0 this lane's own P(R1 ((R2 - n2) mod R2)).
template <int N, int R1, int R2>
__device__ __forceinline__ void gather_row(const float2* __restrict__ row, int n1, float2 (&x)[R2],
                                           float2 (&xm)[R2]) {
  static_assert(64 % R1 == 0, "a row's lanes lie in one wave");
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) x[n2] = row[n1 + R1 * n2];
  const int lane = threadIdx.x & 63;
  const int src = (lane & ~(R1 - 1)) | ((R1 - n1) & (R1 - 1));
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    const float2 sh = shfl2(x[R2 - 1 - n2], src);
    xm[n2] = n1 == 0 ? x[(R2 - n2) & (R2 - 1)] : sh;
  }
}

// Inverse of a packed row: thread (r, n1) gathers P[k], P[N-k] for
// k = n1 + R1 n2 (gather(r, n1, x, xm) fills both arrays) and returns, in
// pass-B layout, z[k2 + R2 k1] (= L * (x[2k], x[2k+1])).  All threads of the
// block must call it.
template <int L, int R1, int R2, int PADSH, class GATHER>
__device__ __forceinline__ void row_inv_to_B(float2 (&u)[R1], bool validA, int r, int n1, float2* buf, bool validB,
                                             int rB, int k2, const float2* __restrict__ tw, GATHER gather) {
  constexpr int N = L / 2;
  if (validA) {
    float2 v[R2], xs[R2], xms[R2];
    gather(r, n1, xs, xms);
    // w_L^k for k = n1 + R1 n2: table value w_L^n1 times compile-time w_L^(R1 n2)
    const float2 w0 = tw[n1];
    PassTw<R2> pt;
    passA_tw<R2, +1, 2>(pt, n1, tw);
    __builtin_amdgcn_sched_barrier(0);  // every load above in flight before the arithmetic
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int k = n1 + R1 * n2;
      const float2 x = xs[n2], xm = xms[n2];
      v[n2] = (k == 0) ? make_float2(x.x + x.y, x.x - x.y) : c2r_pre(x, xm, ctw_mul<L / R1, -1>(w0, n2));
    }
    passA_apply<R2, +1>(v, pt);
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[lds_ix<PADSH>(r * N + n1 + R1 * k)] = v[k];
  }
  __syncthreads();
  if (validB) {
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[lds_ix<PADSH>(rB * N + a + R1 * k2)];
    dft<R1, +1>(u);
  }
}

// Twiddles of row_inv_to_B for pass-A index n1, loaded by the caller ahead of
// time (when its gather reads LDS, so the table loads would otherwise start
// only after it): w_L^n1 and the pass-A twiddles.
template <int R2>
struct RowInvTw {
  PassTw<R2> pt;
  float2 w0;
};
template <int L, int R1, int R2>
__device__ __forceinline__ RowInvTw<R2> row_inv_tw(int n1, const float2* __restrict__ tw) {
  RowInvTw<R2> t;
  t.w0 = tw[n1];
  passA_tw<R2, +1, 2>(t.pt, n1, tw);
  return t;
}

// row_inv_to_B with the twiddles preloaded (row_inv_tw of this thread's n1).
template <int L, int R1, int R2, int PADSH, class GATHER>
__device__ __forceinline__ void row_inv_to_B_tw(float2 (&u)[R1], bool validA, int r, int n1, float2* buf,
                                                bool validB, int rB, int k2, const RowInvTw<R2>& t, GATHER gather) {
  constexpr int N = L / 2;
  if (validA) {
    float2 v[R2], xs[R2], xms[R2];
    gather(r, n1, xs, xms);
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int k = n1 + R1 * n2;
      const float2 x = xs[n2], xm = xms[n2];
      v[n2] = (k == 0) ? make_float2(x.x + x.y, x.x - x.y) : c2r_pre(x, xm, ctw_mul<L / R1, -1>(t.w0, n2));
    }
    passA_apply<R2, +1>(v, t.pt);
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[lds_ix<PADSH>(r * N + n1 + R1 * k)] = v[k];
  }
  __syncthreads();
  if (validB) {
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[lds_ix<PADSH>(rB * N + a + R1 * k2)];
    dft<R1, +1>(u);
  }
}

// row_inv_to_B for long rows (L >= 1024), where holding P[N-k] beside P[k]
// would double pass A's registers (and halve the waves per SIMD): the
// gathered row goes through LDS once and each P[N-k] is read back from there.
// gather(r, n1, x) fills x[n2] = P(n1 + R1 n2) only.
template <int L, int R1, int R2, int PADSH, class GATHER>
__device__ __forceinline__ void row_inv_to_B_lds(float2 (&u)[R1], bool validA, int r, int n1, float2* buf,
                                                 bool validB, int rB, int k2, const float2* __restrict__ tw,
                                                 GATHER gather) {
  constexpr int N = L / 2;
  float2 v[R2];
  float2 w0 = make_float2(1.f, 0.f);
  PassTw<R2> pt;
  if (validA) {
    gather(r, n1, v);
    w0 = tw[n1];
    passA_tw<R2, +1, 2>(pt, n1, tw);
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) buf[lds_ix<PADSH>(r * N + n1 + R1 * n2)] = v[n2];
  }
  __syncthreads();
  if (validA) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int k = n1 + R1 * n2;
      const float2 xm = buf[lds_ix<PADSH>(r * N + ((N - k) & (N - 1)))];
      const float2 x = v[n2];
      v[n2] = (k == 0) ? make_float2(x.x + x.y, x.x - x.y) : c2r_pre(x, xm, ctw_mul<L / R1, -1>(w0, n2));
    }
    passA_apply<R2, +1>(v, pt);
  }
  __syncthreads();  // every partner read is done before buf takes pass A's output
  if (validA) {
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[lds_ix<PADSH>(r * N + n1 + R1 * k)] = v[k];
  }
  __syncthreads();
  if (validB) {
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[lds_ix<PADSH>(rB * N + a + R1 * k2)];
    dft<R1, +1>(u);
  }
}

// Loads P(n1 + R1 n2) of one packed row (the gather of row_inv_to_B_lds).
template <int N, int R1, int R2>
__device__ __forceinline__ void load_row_x(const float2* __restrict__ row, int n1, float2 (&x)[R2]) {
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) x[n2] = row[n1 + R1 * n2];
}

// Row real-to-complex FFT (packed output), optional hard threshold on load.
// grid: (L / ROWS, nplanes, nimg).
template <int L>
__global__ void __launch_bounds__(RowCfg<L>::T) k_rows_r2c(PlaneSet in, PlaneSet out, ThreshSpec th,
                                                            const float2* __restrict__ tw) {
  using C = RowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * C::ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe, b)) + (size_t)row0 * N;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe, b)) + (size_t)row0 * N;
  const int rule = th.mode == 0 ? SHRINK_NONE : (th.soft ? SHRINK_SOFT : SHRINK_HARD);
  const float t = rule ? thresh_of(th, p, b) : 0.0f;
  const int r = tid / R1, n1 = tid % R1;
  const bool validA = r < C::ROWS;
  float2 v[R2];
  if (validA) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) v[n2] = src[(size_t)r * N + n1 + R1 * n2];
  }
  const auto ftw = row_fwd_tw<L, R1, R2>(n1, tid % R2, tw);
  __builtin_amdgcn_sched_barrier(0);  // data and twiddle loads in flight together
  if (validA && rule) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) v[n2] = shrink2(v[n2], t, rule);
  }
  row_fwd_from_A<L, R1, R2, C::PADSH>(v, validA, r, n1, buf, dst, true, tid / R2, tid % R2, ftw);
}

// Row complex-to-real inverse FFT of a packed spectrum (unnormalised: x L),
// optionally of the sum of several input planes.
// The summing form runs on a few planes (one per scale), so it takes fewer
// rows per block (SumRows) to spread them over more blocks.
template <int L>
struct SumRows {
  static constexpr int v = RowCfg<L>::ROWS >= 4 ? 4 : RowCfg<L>::ROWS;
};

template <int L, bool SUM, int ROWS = (SUM ? SumRows<L>::v : RowCfg<L>::ROWS)>
__global__ void __launch_bounds__(ROWS * Split<L / 2>::R2) k_rows_c2r(PlaneSet in, PlaneSet out, SumTable sum,
                                                                  const float2* __restrict__ tw) {
  using C = RowCfg<L, ROWS>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * C::ROWS;
  const size_t pe = (size_t)L * L;
  float2* dst = reinterpret_cast<float2*>(plane_ptr(out, p, pe, b)) + (size_t)row0 * N;
  const int r = tid / R1, n1 = tid % R1, rB = tid / R2, k2 = tid % R2;
  float2 u[R1];
  if constexpr (SUM && L >= 1024) {
    const int q0 = sum.first[p], q1 = sum.first[p + 1];
    auto gather = [&](int rr, int nn, float2 (&x)[R2]) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) x[n2] = make_float2(0.f, 0.f);
      for (int q = q0; q < q1; ++q) {
        float2 a[R2];
        load_row_x<N, R1, R2>(reinterpret_cast<const float2*>(plane_ptr(in, q, pe, b)) + (size_t)(row0 + rr) * N, nn,
                              a);
        __builtin_amdgcn_sched_barrier(0);  // the group's loads in flight together
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) x[n2] = cadd(x[n2], a[n2]);
      }
    };
    row_inv_to_B_lds<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  } else if constexpr (SUM) {
    const int q0 = sum.first[p], q1 = sum.first[p + 1];
    // groups outermost: each group's R2 pairs of loads are in flight together
    auto gather = [&](int rr, int nn, float2 (&x)[R2], float2 (&xm)[R2]) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) x[n2] = xm[n2] = make_float2(0.f, 0.f);
      for (int q = q0; q < q1; ++q) {
        float2 a[R2], am[R2];
        gather_row<N, R1, R2>(reinterpret_cast<const float2*>(plane_ptr(in, q, pe, b)) + (size_t)(row0 + rr) * N, nn,
                              a, am);
        __builtin_amdgcn_sched_barrier(0);  // the group's loads in flight together
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
          x[n2] = cadd(x[n2], a[n2]);
          xm[n2] = cadd(xm[n2], am[n2]);
        }
      }
    };
    row_inv_to_B<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  } else if constexpr (L >= 1024) {
    const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe, b)) + (size_t)row0 * N;
    auto gather = [&](int rr, int nn, float2 (&x)[R2]) { load_row_x<N, R1, R2>(src + (size_t)rr * N, nn, x); };
    row_inv_to_B_lds<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  } else {
    const float2* src = reinterpret_cast<const float2*>(plane_ptr(in, p, pe, b)) + (size_t)row0 * N;
    auto gather = [&](int rr, int nn, float2 (&x)[R2], float2 (&xm)[R2]) {
      gather_row<N, R1, R2>(src + (size_t)rr * N, nn, x, xm);
    };
    row_inv_to_B<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) dst[(size_t)rB * N + k2 + R2 * k1] = u[k1];
}

// Denoise fusion of the decomposition's last row pass with the
// reconstruction's first: c = C2R(T) is stored (the coefficients the caller
// receives), then U = R2C(keep(c, th)) continues from the same registers with
// the split swapped (R2, R1).
template <int L>
__global__ void __launch_bounds__(RowCfg<L>::T) k_rows_c2r_r2c(PlaneSet Tin, PlaneSet band, PlaneSet Uout,
                                                                ThreshSpec th, const float2* __restrict__ tw) {
  using C = RowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * C::ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(Tin, p, pe, b)) + (size_t)row0 * N;
  // an empty band set: the caller takes no coefficients
  const bool store_c = band.ncells > 0;
  float2* cdst = store_c ? reinterpret_cast<float2*>(plane_ptr(band, p, pe, b)) + (size_t)row0 * N : nullptr;
  float2* udst = reinterpret_cast<float2*>(plane_ptr(Uout, p, pe, b)) + (size_t)row0 * N;
  const int rule = th.mode == 0 ? SHRINK_NONE : (th.soft ? SHRINK_SOFT : SHRINK_HARD);
  const float t = rule ? thresh_of(th, p, b) : 0.0f;
  const int r = tid / R1, n1 = tid % R1, rB = tid / R2, k2 = tid % R2;
  // threshold (and the coefficients out), then the forward transform split
  // (R2, R1): thread k2 is pass-A index n1' = k2
  auto finish = [&](float2 (&u)[R1], const RowFwdTw<L, R2, R1>& ftw) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const float2 c = u[k1];
      if (store_c) cdst[(size_t)rB * N + k2 + R2 * k1] = c;
      u[k1] = shrink2(c, t, rule);
    }
    __syncthreads();  // pass-B reads of buf are complete
    row_fwd_from_A<L, R2, R1, C::PADSH>(u, true, rB, k2, buf, udst, tid < C::ROWS * R1, tid / R1, tid % R1, ftw);
  };
  // The twiddles of the forward half (pass-A index k2) are loaded before any
  // store of this block: at the top, or at L >= 1024 after the inverse, so
  // that they are not live across it (141 -> 121 VGPRs, 3 -> 4 waves per
  // SIMD, 117 -> 105 us per 1024^2 band stack; at 512 the early load measured
  // faster).
  if constexpr (L >= 1024) {
    float2 u[R1];
    auto gather = [&](int rr, int nn, float2 (&x)[R2]) { load_row_x<N, R1, R2>(src + (size_t)rr * N, nn, x); };
    row_inv_to_B_lds<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
    const float2* twf = tw;
    asm volatile("" : "+s"(twf));  // an opaque copy: the loads stay here
    const auto ftw = row_fwd_tw<L, R2, R1>(k2, tid % R1, twf);
    finish(u, ftw);
  } else {
    const auto ftw = row_fwd_tw<L, R2, R1>(k2, tid % R1, tw);
    float2 u[R1];
    auto gather = [&](int rr, int nn, float2 (&x)[R2], float2 (&xm)[R2]) {
      gather_row<N, R1, R2>(src + (size_t)rr * N, nn, x, xm);
    };
    row_inv_to_B<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
    finish(u, ftw);
  }
}

// Backward of a soft-thresholded band read, fused the same way: g = C2R(T)
// is the gradient of the thresholded coefficients; with c the coefficients
// of the forward pass (cref) and t the band's threshold,
//   U = R2C(g * 1[|c| > t]),   gth += sum over the block's rows of -sign(c) 1[|c| > t] g
// (one atomic per block into gth[b * dev_stride + dev_off + p]).  th: mode 2.
template <int L>
__global__ void __launch_bounds__(RowCfg<L>::T) k_rows_c2r_mask_r2c(PlaneSet Tin, PlaneSet cref, PlaneSet Uout,
                                                                     ThreshSpec th, float* __restrict__ gth,
                                                                     const float2* __restrict__ tw) {
  using C = RowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2, NW = C::T / 64;
  static_assert(C::T % 64 == 0, "whole waves");
  __shared__ float2 buf[C::LDS];
  __shared__ float red[NW];
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * C::ROWS;
  const size_t pe = (size_t)L * L;
  const float2* src = reinterpret_cast<const float2*>(plane_ptr(Tin, p, pe, b)) + (size_t)row0 * N;
  const float2* csrc = reinterpret_cast<const float2*>(plane_ptr(cref, p, pe, b)) + (size_t)row0 * N;
  float2* udst = reinterpret_cast<float2*>(plane_ptr(Uout, p, pe, b)) + (size_t)row0 * N;
  const size_t ti = (size_t)b * th.dev_stride + th.dev_off + p;
  const float t = th.dev[ti];
  const int r = tid / R1, n1 = tid % R1, rB = tid / R2, k2 = tid % R2;
  const auto ftw = row_fwd_tw<L, R2, R1>(k2, tid % R1, tw);
  // the forward coefficients at this thread's pass-B positions, in flight
  // during the inverse transform
  float2 cv[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) cv[k1] = csrc[(size_t)rB * N + k2 + R2 * k1];
  float2 u[R1];
  if constexpr (L >= 1024) {
    auto gather = [&](int rr, int nn, float2 (&x)[R2]) { load_row_x<N, R1, R2>(src + (size_t)rr * N, nn, x); };
    row_inv_to_B_lds<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  } else {
    auto gather = [&](int rr, int nn, float2 (&x)[R2], float2 (&xm)[R2]) {
      gather_row<N, R1, R2>(src + (size_t)rr * N, nn, x, xm);
    };
    row_inv_to_B<L, R1, R2, C::PADSH>(u, r < C::ROWS, r, n1, buf, true, rB, k2, tw, gather);
  }
  float acc = 0.f;
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const float2 g = u[k1], c = cv[k1];
    const bool kx = fabsf(c.x) > t, ky = fabsf(c.y) > t;
    u[k1] = make_float2(kx ? g.x : 0.f, ky ? g.y : 0.f);
    acc += kx ? (c.x > 0.f ? -g.x : g.x) : 0.f;
    acc += ky ? (c.y > 0.f ? -g.y : g.y) : 0.f;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();  // pass-B reads of buf are complete; red is written
  row_fwd_from_A<L, R2, R1, C::PADSH>(u, true, rB, k2, buf, udst, tid < C::ROWS * R1, tid / R1, tid % R1, ftw);
  if (tid == 0) {
    float sum = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += red[w];
    if (sum != 0.f) atomicAdd(&gth[ti], sum);
  }
}

// ---------------------------------------------------------------------------
// Column passes: a tile of CT columns of a packed spectrum (L rows, row
// stride L/2 complex); each column is a transform of N = L points split
// R1 x R2.  T = CT * R2 threads; element (u, c) of the LDS tile is linear
// element u*CT + c.
// Column tile widths: 8; 16 where the tile fits the LDS (L <= 512); 4 where
// a block still fills a wave (4 R2 >= 64).  Narrow tiles use a quarter of
// each 128-byte row line; the XCD-aware block order keeps the neighbouring
// tiles that share a line in one L2.
template <int L>
struct ColTile {
  static constexpr bool wide = L / 2 >= 16 && L <= 512;
  static constexpr bool narrow = 4 * Split<L>::R2 >= 64;
  static constexpr int dflt = (L >= 1024) ? 8 : ((L / 2) < 16 ? (L / 2) : 16);
};

// LDS layout of a column tile: row u of the column transforms, column c at
//   u * CT + c + CT * (u / R1),
// one pad of CT elements after every R1 rows.  Every access run of the column
// passes is then a per-lane base plus compile-time offsets (no carry into the
// pad term, so no per-element address registers), and the lanes of a pass-B
// access (consecutive columns, rows R1 apart) fall on distinct bank pairs.
template <int L, int CT>
struct ColLds {
  static constexpr int R1 = Split<L>::R1, R2 = Split<L>::R2;
  static constexpr int BLK = R1 * CT + CT;  // R1 padded rows
  static constexpr int SIZE = (L / R1) * BLK;
  __device__ static __forceinline__ int ix(int u, int c) { return (u % R1) * CT + c + (u / R1) * BLK; }
};

template <int L, int CT_ = ColTile<L>::dflt>
struct ColCfg {
  static constexpr int R1 = Split<L>::R1, R2 = Split<L>::R2;
  static constexpr int CT = CT_;
  static constexpr int T = CT * R2;
  using Lds = ColLds<L, CT_>;
  static constexpr int LDS = Lds::SIZE;
};

// Column tile loads and stores through a buffer view of one tile (tile0:
// column c0 of row 0 of a plane with row stride L/2): element (row, c),
// c < CT, at byte (row * L/2 + c) * 8.  The per-lane part of the offset stays
// in one register for a whole run of accesses; the view ends with the tile.
template <int L, int CT>
__device__ __forceinline__ BufView col_view(const float2* tile0) {
  return BufView(tile0, (unsigned)(((L - 1) * (L / 2) + CT) * sizeof(float2)));
}
template <int L>
__device__ __forceinline__ unsigned col_off(int row, int c) {
  return (unsigned)((row * (L / 2) + c) * sizeof(float2));
}

// Pass-A registers of thread (c, n1 < R1): v[n2] = src(n1 + R1 n2, c).
template <int L, int CT_>
__device__ __forceinline__ void col_load_A(float2 (&v)[Split<L>::R2], const float2* src) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const BufView bv = col_view<L, CT>(src);
  const unsigned lane = col_off<L>(threadIdx.x / CT, threadIdx.x % CT);
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) v[n2] = bv.ld2(lane, col_off<L>(R1 * n2, 0));
}

// Pass-B registers of thread (c, k2): x[k1] = src(k2 + R2 k1, c).
template <int L, int CT_>
__device__ __forceinline__ void col_load_B(float2 (&x)[Split<L>::R1], const float2* src) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const BufView bv = col_view<L, CT>(src);
  const unsigned lane = col_off<L>(threadIdx.x / CT, threadIdx.x % CT);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) x[k1] = bv.ld2(lane, col_off<L>(R2 * k1, 0));
}

// Forward column transform: from global (row stride L/2) to pass-B layout
// registers u[k1] = X[k2 + R2 k1] of thread (c, k2).  Pass A (global loads,
// first DFT, LDS stores) and pass B (LDS loads, second DFT) are separate so a
// caller can issue other loads between them; a barrier must separate them.
//
// When R2 = 2 R1 the length-R2 DFTs of pass A (forward) and of the inverse's
// second pass run on lane pairs (dft_pair): every thread of the block takes
// part, each holding R1 values instead of R2 for half of the threads.
template <int L>
struct ColPair {
  static constexpr bool on = Split<L>::R2 == 2 * Split<L>::R1;
};

template <int L, int CT_>
__device__ __forceinline__ void col_fwd_A(const float2* __restrict__ src, float2* buf,
                                          const float2* __restrict__ tw) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT, NH = L / 2;
  const int tid = threadIdx.x;
  if constexpr (ColPair<L>::on) {
    // thread (c, n1, h): z[j] = x[n1 + R1 (2j + h)], outputs k2 = k + R1 h
    const int h = tid & 1, rest = tid >> 1;
    const int c = rest % CT, n1 = rest / CT;
    float2 v[R1];
    const BufView bv = col_view<L, CT>(src);
    const unsigned lane = col_off<L>(n1 + R1 * h, c);
#pragma unroll
    for (int j = 0; j < R1; ++j) v[j] = bv.ld2(lane, col_off<L>(2 * R1 * j, 0));
    PairTw<R1> t;
    pair_tw<R1, 1>(t, n1, h, tw);
    __builtin_amdgcn_sched_barrier(0);  // data and twiddle loads in flight together
    dft_pair<R1, -1>(v, h);
    pair_tw_apply<R1, -1>(v, t);
    float2* row = buf + n1 * CT + c + R1 * h * C::Lds::BLK;  // row n1 + R1 (k + R1 h)
#pragma unroll
    for (int k = 0; k < R1; ++k) row[k * C::Lds::BLK] = v[k];
    return;
  }
  const int c = tid % CT, n1 = tid / CT;
  if (n1 < R1) {
    float2 v[R2];
    col_load_A<L, CT_>(v, src);
    passA<R1, R2, -1, 1>(v, n1, tw);
    float2* row = buf + n1 * CT + c;  // row n1 + R1 k
#pragma unroll
    for (int k = 0; k < R2; ++k) row[k * C::Lds::BLK] = v[k];
  }
}

template <int L, int CT_>
__device__ __forceinline__ void col_fwd_B(float2 (&u)[Split<L>::R1], const float2* buf) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, CT = C::CT;
  const int tid = threadIdx.x;
  const int c = tid % CT, k2 = tid / CT;
  const float2* row = buf + c + k2 * C::Lds::BLK;  // row a + R1 k2
#pragma unroll
  for (int a = 0; a < R1; ++a) u[a] = row[a * CT];
  dft<R1, -1>(u);
}

template <int L, int CT_>
__device__ __forceinline__ void col_fwd_to_B(float2 (&u)[Split<L>::R1], const float2* __restrict__ src,
                                             float2* buf, const float2* __restrict__ tw) {
  col_fwd_A<L, CT_>(src, buf, tw);
  __syncthreads();
  col_fwd_B<L, CT_>(u, buf);
}

// Pass A of a forward column FFT whose input element (row, c) of this
// thread's column comes from load(row, c) (a permuted or LDS source).  SYNC:
// a barrier between the loads and the LDS stores (the source is buf itself).
template <int L, int CT_, bool SYNC, class LOAD>
__device__ __forceinline__ void col_fwd_A_src(LOAD load, float2* buf, const float2* __restrict__ tw) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const int tid = threadIdx.x;
  if constexpr (ColPair<L>::on) {
    const int h = tid & 1, rest = tid >> 1;
    const int c = rest % CT, n1 = rest / CT;
    float2 v[R1];
#pragma unroll
    for (int j = 0; j < R1; ++j) v[j] = load(n1 + R1 * (2 * j + h), c);
    PairTw<R1> t;
    pair_tw<R1, 1>(t, n1, h, tw);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (SYNC) __syncthreads();
    dft_pair<R1, -1>(v, h);
    pair_tw_apply<R1, -1>(v, t);
    float2* row = buf + n1 * CT + c + R1 * h * C::Lds::BLK;
#pragma unroll
    for (int k = 0; k < R1; ++k) row[k * C::Lds::BLK] = v[k];
  } else {
    const int c = tid % CT, n1 = tid / CT;
    float2 v[R2];
    if (n1 < R1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = load(n1 + R1 * n2, c);
    }
    if constexpr (SYNC) __syncthreads();
    if (n1 < R1) {
      passA<R1, R2, -1, 1>(v, n1, tw);
      float2* row = buf + n1 * CT + c;
#pragma unroll
      for (int k = 0; k < R2; ++k) row[k * C::Lds::BLK] = v[k];
    }
  }
}

// First half of an inverse column transform: pass-A twiddles on the pass-B
// layout registers q (thread (c, k2) holding Q[k2 + R2 k1]) and the LDS
// transpose, followed by a barrier (and preceded by one when buf may still be
// in use: IN_USE).
template <int L, int CT_, bool IN_USE>
__device__ __forceinline__ void col_inv_A(float2 (&q)[Split<L>::R1], float2* buf, const PassTw<Split<L>::R1>& t) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const int c = threadIdx.x % CT, k2 = threadIdx.x / CT;
  passA_apply<R1, +1>(q, t);
  if constexpr (IN_USE) __syncthreads();
  float2* row = buf + C::Lds::ix(k2, c);
#pragma unroll
  for (int k = 0; k < R1; ++k) row[k * (R2 / R1) * C::Lds::BLK] = q[k];
  __syncthreads();
}

// Second half: the last DFT, then store(row, c, value) for every output
// element (row of the inverse transform, column c).  SYNC: a barrier between
// the LDS reads and the stores (store may write buf).
template <int L, int CT_, bool SYNC, class STORE>
__device__ __forceinline__ void col_inv_B_to(const float2* buf, STORE store) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const int tid = threadIdx.x;
  if constexpr (ColPair<L>::on) {
    const int h = tid & 1, rest = tid >> 1;
    const int cp = rest % CT, kb = rest / CT;
    float2 w[R1];
    const float2* row = buf + h * CT + cp + kb * (R2 / R1) * C::Lds::BLK;
#pragma unroll
    for (int j = 0; j < R1; ++j) w[j] = row[((2 * j) % R1) * CT + ((2 * j) / R1) * C::Lds::BLK];
    dft_pair<R1, +1>(w, h);
    if constexpr (SYNC) __syncthreads();
#pragma unroll
    for (int k = 0; k < R1; ++k) store(kb + R1 * (k + R1 * h), cp, w[k]);
  } else {
    const int c = tid % CT, kb = tid / CT;
    float2 w[R2];
    if (kb < R1) {
      const float2* row = buf + c + kb * (R2 / R1) * C::Lds::BLK;
#pragma unroll
      for (int a = 0; a < R2; ++a) w[a] = row[(a % R1) * CT + (a / R1) * C::Lds::BLK];
      dft<R2, +1>(w);
    }
    if constexpr (SYNC) __syncthreads();
    if (kb < R1) {
#pragma unroll
      for (int k1 = 0; k1 < R2; ++k1) store(kb + R1 * k1, c, w[k1]);
    }
  }
}

// Inverse column transform from pass-B layout registers (thread (c, k2)
// holding Q[k2 + R2 k1]) to global (row stride L/2).  Starts with a barrier
// so buf may still be in use by the caller.  t: the first pass's twiddles
// (col_inv_tw), loaded by the caller ahead of time.
template <int L>
__device__ __forceinline__ PassTw<Split<L>::R1> col_inv_tw(int k2, const float2* __restrict__ tw) {
  PassTw<Split<L>::R1> t;
  passA_tw<Split<L>::R1, +1, 1>(t, k2, tw);
  return t;
}

template <int L, int CT_>
__device__ __forceinline__ void col_inv_from_B(float2 (&q)[Split<L>::R1], float2* __restrict__ dst, float2* buf,
                                               const PassTw<Split<L>::R1>& t) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT, NH = L / 2;
  const int tid = threadIdx.x;
  const int c = tid % CT, k2 = tid / CT;
  passA_apply<R1, +1>(q, t);
  __syncthreads();
  {
    float2* row = buf + C::Lds::ix(k2, c);  // row k2 + R2 k
#pragma unroll
    for (int k = 0; k < R1; ++k) row[k * (R2 / R1) * C::Lds::BLK] = q[k];
  }
  __syncthreads();
  if constexpr (ColPair<L>::on) {
    // thread (c, kb, h): z[j] = Y[2j + h] of output column kb, outputs k + R1 h
    const int h = tid & 1, rest = tid >> 1;
    const int cp = rest % CT, kb = rest / CT;
    float2 w[R1];
    const float2* row = buf + h * CT + cp + kb * (R2 / R1) * C::Lds::BLK;  // row 2j + h + R2 kb
#pragma unroll
    for (int j = 0; j < R1; ++j) w[j] = row[((2 * j) % R1) * CT + ((2 * j) / R1) * C::Lds::BLK];
    dft_pair<R1, +1>(w, h);
    const BufView bv = col_view<L, CT>(dst);
    const unsigned lane = col_off<L>(kb + R1 * R1 * h, cp);
#pragma unroll
    for (int k = 0; k < R1; ++k) bv.st2(w[k], lane, col_off<L>(R1 * k, 0));
    return;
  }
  const int kb = tid / CT;
  if (kb < R1) {
    float2 w[R2];
    const float2* row = buf + c + kb * (R2 / R1) * C::Lds::BLK;  // row a + R2 kb
#pragma unroll
    for (int a = 0; a < R2; ++a) w[a] = row[(a % R1) * CT + (a / R1) * C::Lds::BLK];
    dft<R2, +1>(w);
    const BufView bv = col_view<L, CT>(dst);
    const unsigned lane = col_off<L>(kb, c);
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) bv.st2(w[k1], lane, col_off<L>(R1 * k1, 0));
  }
}

// Band spectrum H in pass-B layout: h[k1] = H(k2 + R2 k1, c).  Issued as one
// batch of independent loads, ahead of the work that hides their latency.
template <int L, int CT_>
__device__ __forceinline__ void col_load_h(float2 (&h)[Split<L>::R1], const float2* __restrict__ H) {
  col_load_B<L, CT_>(h, H);
}

// Filter pass-B registers by the band spectrum h (col_load_h; CONJ: its
// conjugate).  In the block holding column 0 the packed (DC, Nyquist) column
// needs X(-u): it is exchanged through LDS (block-uniform branch), and the
// beta term (row Hb) is added on lane column 0 by a select.
template <int L, int CT_, bool CONJ, bool ACC>
__device__ __forceinline__ void col_filter(float2 (&u)[Split<L>::R1], float2 (&out)[Split<L>::R1], int c0,
                                           const float2 (&h)[Split<L>::R1], const float2* __restrict__ Hb,
                                           float2* buf) {
  using C = ColCfg<L, CT_>;
  constexpr int R1 = C::R1, R2 = C::R2, CT = C::CT;
  const int tid = threadIdx.x;
  const int c = tid % CT, k2 = tid / CT;
  if (c0 == 0) {
    float2 hb[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) hb[k1] = Hb[k2 + R2 * k1];
    __syncthreads();
    {
      float2* row = buf + C::Lds::ix(k2, c);  // row k2 + R2 k1
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) row[k1 * (R2 / R1) * C::Lds::BLK] = u[k1];
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int uu = k2 + R2 * k1;
      const float2 pm = cconj(buf[C::Lds::ix((L - uu) & (L - 1), 0)]);
      float2 q = CONJ ? cmulc(u[k1], h[k1]) : cmul(h[k1], u[k1]);
      const float2 e = CONJ ? cmulc(pm, hb[k1]) : cmul(hb[k1], pm);
      if (c == 0) q = cadd(q, e);
      out[k1] = ACC ? cadd(out[k1], q) : q;
    }
  } else {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const float2 q = CONJ ? cmulc(u[k1], h[k1]) : cmul(h[k1], u[k1]);
      out[k1] = ACC ? cadd(out[k1], q) : q;
    }
  }
}

// Wave-uniform reads of the byte and halfword tables of a BandMap kernel
// argument as aligned dwords, so they become scalar loads (the scalar unit has
// no sub-dword loads).
__device__ __forceinline__ int map_u8(const unsigned char* a, int i) {
  const unsigned w = reinterpret_cast<const unsigned*>(a)[i >> 2];
  return (w >> (8 * (i & 3))) & 0xffu;
}
__device__ __forceinline__ int map_u16(const unsigned short* a, int i) {
  const unsigned w = reinterpret_cast<const unsigned*>(a)[i >> 1];
  return (i & 1) ? (w >> 16) : (w & 0xffffu);
}
static_assert(offsetof(BandMap, scale_of) % 4 == 0 && offsetof(BandMap, group_band0) % 4 == 0 &&
                  offsetof(BandMap, filt_of) % 4 == 0 && offsetof(BandMap, order) % 4 == 0 &&
                  offsetof(BandMap, gorder) % 4 == 0,
              "BandMap tables must be dword aligned");

// XCD-aware block order of the column passes: workgroups are dispatched to
// the 8 XCDs in turn, so consecutive grid ids land in different L2 caches.
// The flat id is remapped so that each XCD takes a contiguous range of
// (tile fastest, then image, then plane): neighbouring column tiles, which
// share the 128-byte lines of narrow tiles, and the images sharing a filter
// plane then run in the same L2.
// share > 1: the planes come in runs of `share` that read the same filter
// planes (BandMap order / gorder); the run member is then the fastest index,
// so the blocks reading one filter tile are dispatched together.
__device__ __forceinline__ void col_block_order(int& tile, int& plane, int& img, int share = 1) {
  const int nt = gridDim.x, np = gridDim.y, ni = gridDim.z;
  const int total = nt * np * ni;
  int lid = blockIdx.x + nt * (blockIdx.y + np * blockIdx.z);
  if (total % 8 == 0) lid = (lid % 8) * (total / 8) + lid / 8;
  // the block index arithmetic runs on the vector unit; the results are
  // wave-uniform, so move them to scalar registers (scalar table lookups)
  const int m = lid % share;
  lid /= share;
  tile = __builtin_amdgcn_readfirstlane(lid % nt);
  const int r = lid / nt;
  img = __builtin_amdgcn_readfirstlane(r % ni);
  plane = __builtin_amdgcn_readfirstlane((r / ni) * share + m);
}

// Block order of the decomposition column passes (and of the reconstruction
// passes up to 512^2), after the same XCD remap:
// the TG column tiles that share a 128-byte line (TG = 16 / CT for tiles of
// CT complex columns) fastest, then the image, then the plane (band slot),
// then the tile group.  An XCD then runs every band of a few neighbouring
// tiles back to back, so the scale spectrum tile (read by all bands of its
// scale), each filter tile (read by the bands sharing it and by every image
// of a batch) come from its L2 after the first read, and each line is fetched
// once for the TG tiles that share it.  Measured against the tile-fastest
// order of the other column passes: 512^2, 8 columns, 28.8 -> 24.8 us
// (counted fetch 66 -> 47 MB); 1024^2, 8 columns, 91 -> 86 us; with the image
// after the tile group instead of before the plane, the batched mix ran 1.5%
// slower.  The reconstruction (planes = direction groups) shares only filter
// tiles; it takes this order for one image up to 512^2 (three-pass 4-column
// pass 29.3 -> 26.9 us) and the tile-fastest order for batches and at
// 1024^2, where it measured faster.
template <int TG>
__device__ __forceinline__ void col_block_order_lines(int& tile, int& plane, int& img) {
  const int nt = gridDim.x, np = gridDim.y, ni = gridDim.z;
  const int total = nt * np * ni;
  int lid = blockIdx.x + nt * (blockIdx.y + np * blockIdx.z);
  if (total % 8 == 0) lid = (lid % 8) * (total / 8) + lid / 8;
  const int tg = nt % TG == 0 ? TG : 1;
  const int m = lid % tg;
  lid /= tg;
  img = __builtin_amdgcn_readfirstlane(lid % ni);
  lid /= ni;
  plane = __builtin_amdgcn_readfirstlane(lid % np);
  tile = __builtin_amdgcn_readfirstlane((lid / np) * tg + m);
}

// Forward column FFT of every scale spectrum, in place: the column pass of the
// decomposition then filters the transformed scale once per band instead of
// transforming it again for each.  One block per (column tile, scale, image);
// each block reads its whole tile before writing it back.
template <int L, int CT_>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_cols_fwd(float2* __restrict__ Y1, long long y1_stride,
                                                            const float2* __restrict__ tw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2, R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  int tile, j, b;
  col_block_order(tile, j, b);
  const int c0 = tile * C::CT;
  float2* base = Y1 + (size_t)b * y1_stride + (size_t)j * L * NH + c0;
  float2 u[R1];
  col_fwd_to_B<L, CT_>(u, base, buf, tw);
  const BufView bv = col_view<L, C::CT>(base);
  const unsigned lane = col_off<L>(threadIdx.x / C::CT, threadIdx.x % C::CT);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) bv.st2(u[k1], lane, col_off<L>(R2 * k1, 0));
}

// Column pass of the decomposition: one block per (column tile, band, image).
// PRE: the scale spectrum is already column-transformed (k_cols_fwd).
template <int L, int CT_, bool PRE>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_cols_dec(const float2* __restrict__ Y1, long long y1_stride,
                                                            float2* __restrict__ Tout, long long t_stride,
                                                            const float2* __restrict__ Hmain,
                                                            const float2* __restrict__ Hbeta, BandMap map,
                                                            const float2* __restrict__ tw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2;
  __shared__ float2 buf[C::LDS];
  int tile, slot, b;
  col_block_order_lines<(16 / C::CT > 1 ? 16 / C::CT : 1)>(tile, slot, b);
  const int p = map_u8(map.order, slot);
  const int c0 = tile * C::CT;
  const int j = map_u8(map.scale_of, p);
  const int f = map_u8(map.filt_of, p);
  float2 u[C::R1], q[C::R1];
  const float2* src = Y1 + (size_t)b * y1_stride + (size_t)j * L * NH + c0;
  const int k2 = threadIdx.x / C::CT;
  float2 h[C::R1];
  PassTw<C::R1> ivt;
  if constexpr (PRE) {
    col_load_B<L, CT_>(u, src);
    col_load_h<L, CT_>(h, Hmain + (size_t)f * L * NH + c0);
    ivt = col_inv_tw<L>(k2, tw);
    __builtin_amdgcn_sched_barrier(0);  // all loads in flight before the arithmetic
  } else {
    col_load_h<L, CT_>(h, Hmain + (size_t)f * L * NH + c0);
    ivt = col_inv_tw<L>(k2, tw);
    col_fwd_to_B<L, CT_>(u, src, buf, tw);
  }
  col_filter<L, CT_, false, false>(u, q, c0, h, Hbeta + (size_t)f * L, buf);
  col_inv_from_B<L, CT_>(q, Tout + (size_t)b * t_stride + (size_t)p * L * NH + c0, buf, ivt);
}

// Column pass of the reconstruction: one block per (column tile, direction
// group, image) forms IFFT_cols(sum_{k in group} conj(H_k) . FFT_cols(U_k));
// the row pass that follows sums the groups of a scale.  For L >= 512 the
// running sum lives in a second LDS tile (each thread owns its elements), which
// keeps the register footprint of the 32-point passes below the spill point.
template <int L, int CT_>
struct ColRecAcc {
  static constexpr bool LDS = L >= 512;
  static constexpr int SIZE = LDS ? ColCfg<L, CT_>::LDS : 1;
};

template <int L, int CT_>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_cols_rec(const float2* __restrict__ U, long long u_stride,
                                                            float2* __restrict__ Q, long long q_stride,
                                                            const float2* __restrict__ Hmain,
                                                            const float2* __restrict__ Hbeta, BandMap map,
                                                            const float2* __restrict__ tw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2, R1 = C::R1, R2 = C::R2, CT = C::CT;
  __shared__ float2 buf[C::LDS];
  __shared__ float2 accl[ColRecAcc<L, CT_>::SIZE];
  int tile, slot, b;
  // one image up to 512^2: the line-grouped order (measured 29.6 -> 28.4 us
  // at 512^2, 4 columns); batches and 1024^2: tile fastest, then image (the
  // images reading one filter tile together; 112 against 115 us at 1024^2)
  if (L <= 512 && gridDim.z == 1)
    col_block_order_lines<(16 / C::CT > 1 ? 16 / C::CT : 1)>(tile, slot, b);
  else
    col_block_order(tile, slot, b, map.gshare);
  const int g = map_u16(map.gorder, slot);
  const int c0 = tile * CT;
  const int c = threadIdx.x % CT, k2 = threadIdx.x / CT;
  const int p0 = map_u16(map.group_band0, g), p1 = map_u16(map.group_band0, g + 1);
  const float2* ub = U + (size_t)b * u_stride + c0;
  float2 acc[R1];
#pragma unroll
  for (int k = 0; k < R1; ++k) acc[k] = make_float2(0.f, 0.f);
  for (int p = p0; p < p1; ++p) {
    float2 u[R1], h[R1];
    __syncthreads();  // previous direction's reads of buf are complete
    const int f = map_u8(map.filt_of, p);
    col_fwd_A<L, CT_>(ub + (size_t)p * L * NH, buf, tw);
    col_load_h<L, CT_>(h, Hmain + (size_t)f * L * NH + c0);  // in flight across pass B
    __syncthreads();
    col_fwd_B<L, CT_>(u, buf);
    if constexpr (ColRecAcc<L, CT_>::LDS) {
      float2 q[R1];
      col_filter<L, CT_, true, false>(u, q, c0, h, Hbeta + (size_t)f * L, buf);
      float2* row = accl + C::Lds::ix(k2, c);  // row k2 + R2 k1
      if (p == p0) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) row[k1 * (R2 / R1) * C::Lds::BLK] = q[k1];
      } else {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
          float2* e = row + k1 * (R2 / R1) * C::Lds::BLK;
          *e = cadd(*e, q[k1]);
        }
      }
    } else {
      col_filter<L, CT_, true, true>(u, acc, c0, h, Hbeta + (size_t)f * L, buf);
    }
  }
  if constexpr (ColRecAcc<L, CT_>::LDS) {
    const float2* row = accl + C::Lds::ix(k2, c);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) acc[k1] = row[k1 * (R2 / R1) * C::Lds::BLK];
  }
  col_inv_from_B<L, CT_>(acc, Q + (size_t)b * q_stride + (size_t)g * L * NH + c0, buf, col_inv_tw<L>(k2, tw));
}

// ---------------------------------------------------------------------------
// Column passes on the three-pass 512-point engine (fft3.h): eight values per
// thread instead of sixteen, so the blocks hold far fewer registers and the
// reconstruction keeps its running sum in registers (no second LDS tile).
// Same grids, band maps and results (to FP32 rounding) as k_cols_dec<PRE> and
// k_cols_rec.  Thread (c, tau) of a block of CT * 64 threads owns rows
// tau + 64 s of tile column c.
template <int L, int CT>
__global__ void __launch_bounds__(CT * 64) k_cols_dec3(const float2* __restrict__ Y1, long long y1_stride,
                                                   float2* __restrict__ Tout, long long t_stride,
                                                   const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                   BandMap map, const float2* __restrict__ tw) {
  static_assert(L == 512, "three-pass engine: 512 points");
  using E = Col3<CT>;
  constexpr int NH = L / 2;
  __shared__ float2 buf[E::LDS];
  int tile, slot, b;
  col_block_order_lines<(16 / CT > 1 ? 16 / CT : 1)>(tile, slot, b);
  const int p = map_u8(map.order, slot);
  const int j = map_u8(map.scale_of, p), f = map_u8(map.filt_of, p);
  const int c0 = tile * CT;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  const float2* src = Y1 + (size_t)b * y1_stride + (size_t)j * L * NH + c0;
  const BufView xv = col_view<L, CT>(src);
  const BufView hv = col_view<L, CT>(Hmain + (size_t)f * L * NH + c0);
  const unsigned lane = col_off<L>(tau, c);
  float2 x[8], h[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    x[k] = xv.ld2(lane, col_off<L>(64 * k, 0));
    h[k] = hv.ld2(lane, col_off<L>(64 * k, 0));
  }
  const Fft3Tw t = fft3_tw(tau, tw);
  __builtin_amdgcn_sched_barrier(0);  // data, filter and twiddle loads in flight together
  float2 q[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) q[k] = cmul(h[k], x[k]);
  if (c0 == 0 && c == 0) {
    // packed (DC, Nyquist) column: + beta(u) conj(X(-u))
    const float2* hb = Hbeta + (size_t)f * L;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int u = tau + 64 * k;
      const float2 pm = cconj(src[(size_t)((L - u) & (L - 1)) * NH]);
      q[k] = cadd(q[k], cmul(hb[u], pm));
    }
  }
  fft512_3<CT, +1>(q, buf, t);
  const BufView tv = col_view<L, CT>(Tout + (size_t)b * t_stride + (size_t)p * L * NH + c0);
#pragma unroll
  for (int k = 0; k < 8; ++k) tv.st2(q[k], lane, col_off<L>(64 * k, 0));
}

template <int L, int CT>
__global__ void __launch_bounds__(CT * 64) k_cols_rec3(const float2* __restrict__ U, long long u_stride,
                                                   float2* __restrict__ Q, long long q_stride,
                                                   const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                   BandMap map, const float2* __restrict__ tw) {
  static_assert(L == 512, "three-pass engine: 512 points");
  using E = Col3<CT>;
  constexpr int NH = L / 2;
  __shared__ float2 buf[E::LDS];
  int tile, slot, b;
  if (gridDim.z == 1)  // one image: 29.3 -> 26.9 us at 4 columns
    col_block_order_lines<(16 / CT > 1 ? 16 / CT : 1)>(tile, slot, b);
  else
    col_block_order(tile, slot, b, map.gshare);
  const int g = map_u16(map.gorder, slot);
  const int c0 = tile * CT;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  const int p0 = map_u16(map.group_band0, g), p1 = map_u16(map.group_band0, g + 1);
  const float2* ub = U + (size_t)b * u_stride + c0;
  const unsigned lane = col_off<L>(tau, c);
  const Fft3Tw t = fft3_tw(tau, tw);
  float2 acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = make_float2(0.f, 0.f);
  // software pipeline: the band tile of the next direction is loaded while
  // this direction is transformed (the filter tile is already in flight
  // across the transform: it is used last)
  float2 vn[8];
  {
    const BufView uv = col_view<L, CT>(ub + (size_t)p0 * L * NH);
#pragma unroll
    for (int k = 0; k < 8; ++k) vn[k] = uv.ld2(lane, col_off<L>(64 * k, 0));
  }
  for (int p = p0; p < p1; ++p) {
    const int f = map_u8(map.filt_of, p);
    const BufView hv = col_view<L, CT>(Hmain + (size_t)f * L * NH + c0);
    float2 v[8], h[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[k] = vn[k];
      h[k] = hv.ld2(lane, col_off<L>(64 * k, 0));
    }
    if (p + 1 < p1) {
      const BufView uv = col_view<L, CT>(ub + (size_t)(p + 1) * L * NH);
#pragma unroll
      for (int k = 0; k < 8; ++k) vn[k] = uv.ld2(lane, col_off<L>(64 * k, 0));
    }
    __builtin_amdgcn_sched_barrier(0);
    if (p > p0) __syncthreads();  // the previous direction's reads of buf are complete
    fft512_3<CT, -1>(v, buf, t);
    if (c0 == 0) {
      // packed (DC, Nyquist) column: + conj(beta(u)) X(-u)* (the partner row
      // of column 0 sits in another thread: exchanged through LDS)
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 8; ++k) buf[E::ix(tau + 64 * k, c)] = v[k];
      __syncthreads();
      if (c == 0) {
        const float2* hb = Hbeta + (size_t)f * L;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int u = tau + 64 * k;
          const float2 pm = cconj(buf[E::ix((L - u) & (L - 1), 0)]);
          acc[k] = cadd(acc[k], cmulc(pm, hb[u]));
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = cadd(acc[k], cmulc(v[k], h[k]));
  }
  __syncthreads();
  fft512_3<CT, +1>(acc, buf, t);
  const BufView qv = col_view<L, CT>(Q + (size_t)b * q_stride + (size_t)g * L * NH + c0);
#pragma unroll
  for (int k = 0; k < 8; ++k) qv.st2(acc[k], lane, col_off<L>(64 * k, 0));
}

// The same two passes at L = 1024 on the 16 x 8 x 8 engine with 128 threads
// per column (Col3m): the filter product is formed in the engine's spectral
// layout (rows q + 128 s), between the forward transform (natural in,
// spectral out) and the inverse (spectral in, natural out).
template <int CT>
__global__ void __launch_bounds__(CT * 128) k_cols_dec3k(const float2* __restrict__ Y1, long long y1_stride,
                                                     float2* __restrict__ Tout, long long t_stride,
                                                     const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                     BandMap map, const float2* __restrict__ tw) {
  constexpr int L = 1024, NH = L / 2;
  using E = Col3m<CT>;
  __shared__ float2 buf[E::LDS];
  int tile, slot, b;
  col_block_order_lines<(16 / CT > 1 ? 16 / CT : 1)>(tile, slot, b);
  const int p = map_u8(map.order, slot);
  const int j = map_u8(map.scale_of, p), f = map_u8(map.filt_of, p);
  const int c0 = tile * CT;
  const int c = E::col(), q = E::q();
  const float2* src = Y1 + (size_t)b * y1_stride + (size_t)j * L * NH + c0;
  const BufView xv = col_view<L, CT>(src);
  const BufView hv = col_view<L, CT>(Hmain + (size_t)f * L * NH + c0);
  const unsigned slane = col_off<L>(q, c);
  float2 x[8], h[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    x[k] = xv.ld2(slane, col_off<L>(128 * k, 0));
    h[k] = hv.ld2(slane, col_off<L>(128 * k, 0));
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = cmul(h[k], x[k]);
  if (c0 == 0 && c == 0) {
    // packed (DC, Nyquist) column: + beta(u) conj(X(-u))
    const float2* hb = Hbeta + (size_t)f * L;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int u = q + 128 * k;
      const float2 pm = cconj(src[(size_t)((L - u) & (L - 1)) * NH]);
      x[k] = cadd(x[k], cmul(hb[u], pm));
    }
  }
  fft1024_inv<CT>(x, buf, tw);
  const BufView tv = col_view<L, CT>(Tout + (size_t)b * t_stride + (size_t)p * L * NH + c0);
  const unsigned nlane = col_off<L>(E::in_row(q), c);
#pragma unroll
  for (int k = 0; k < 8; ++k) tv.st2(x[k], nlane, col_off<L>(128 * k, 0));
}

template <int CT>
__global__ void __launch_bounds__(CT * 128) k_cols_rec3k(const float2* __restrict__ U, long long u_stride,
                                                     float2* __restrict__ Q, long long q_stride,
                                                     const float2* __restrict__ Hmain, const float2* __restrict__ Hbeta,
                                                     BandMap map, const float2* __restrict__ tw) {
  constexpr int L = 1024, NH = L / 2;
  using E = Col3m<CT>;
  __shared__ float2 buf[E::LDS];
  int tile, slot, b;
  col_block_order(tile, slot, b, map.gshare);
  const int g = map_u16(map.gorder, slot);
  const int c0 = tile * CT;
  const int c = E::col(), q = E::q();
  const int p0 = map_u16(map.group_band0, g), p1 = map_u16(map.group_band0, g + 1);
  const float2* ub = U + (size_t)b * u_stride + c0;
  const unsigned nlane = col_off<L>(E::in_row(q), c);
  const unsigned slane = col_off<L>(q, c);
  float2 acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = make_float2(0.f, 0.f);
  for (int p = p0; p < p1; ++p) {
    const int f = map_u8(map.filt_of, p);
    const BufView uv = col_view<L, CT>(ub + (size_t)p * L * NH);
    const BufView hv = col_view<L, CT>(Hmain + (size_t)f * L * NH + c0);
    float2 v[8], h[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = uv.ld2(nlane, col_off<L>(128 * k, 0));
    __builtin_amdgcn_sched_barrier(0);
    if (p > p0) __syncthreads();  // the previous direction's reads of buf are complete
    fft1024_fwd<CT>(v, buf, tw);
    // the filter tile is read after the forward transform (not held across it)
#pragma unroll
    for (int k = 0; k < 8; ++k) h[k] = hv.ld2(slane, col_off<L>(128 * k, 0));
    if (c0 == 0) {
      // packed (DC, Nyquist) column: + conj(beta(u)) X(-u)*, the partner row
      // of column 0 exchanged through LDS
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 8; ++k) buf[E::ix(q + 128 * k, c)] = v[k];
      __syncthreads();
      if (c == 0) {
        const float2* hb = Hbeta + (size_t)f * L;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int u = q + 128 * k;
          const float2 pm = cconj(buf[E::ix((L - u) & (L - 1), 0)]);
          acc[k] = cadd(acc[k], cmulc(pm, hb[u]));
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = cadd(acc[k], cmulc(v[k], h[k]));
  }
  __syncthreads();
  fft1024_inv<CT>(acc, buf, tw);
  const BufView qv = col_view<L, CT>(Q + (size_t)b * q_stride + (size_t)g * L * NH + c0);
#pragma unroll
  for (int k = 0; k < 8; ++k) qv.st2(acc[k], nlane, col_off<L>(128 * k, 0));
}

// ---------------------------------------------------------------------------
// Host-side launchers, one instantiation per supported L.
template <int L>
struct Launch {
  using R = RowCfg<L>;
  static hipError_t rows_r2c(const PlaneSet& in, const PlaneSet& out, int np, int nimg, const ThreshSpec& th,
                             const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_rows_r2c<L>), dim3(L / R::ROWS, np, nimg), dim3(R::T), 0, s, in, out, th, tw);
    return hipGetLastError();
  }
  static hipError_t rows_c2r(const PlaneSet& in, const PlaneSet& out, int np, int nimg, const int* sum_first,
                             const float2* tw, hipStream_t s) {
    SumTable st{};
    if (sum_first) {
      using RS = RowCfg<L, SumRows<L>::v>;
      for (int i = 0; i <= np; ++i) st.first[i] = sum_first[i];
      hipLaunchKernelGGL((k_rows_c2r<L, true>), dim3(L / RS::ROWS, np, nimg), dim3(RS::T), 0, s, in, out, st, tw);
    } else {
      hipLaunchKernelGGL((k_rows_c2r<L, false>), dim3(L / R::ROWS, np, nimg), dim3(R::T), 0, s, in, out, st, tw);
    }
    return hipGetLastError();
  }
  static hipError_t rows_c2r_r2c(const PlaneSet& Tin, const PlaneSet& band, const PlaneSet& U, int np, int nimg,
                                 const ThreshSpec& th, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_rows_c2r_r2c<L>), dim3(L / R::ROWS, np, nimg), dim3(R::T), 0, s, Tin, band, U, th, tw);
    return hipGetLastError();
  }
  static hipError_t rows_c2r_mask_r2c(const PlaneSet& Tin, const PlaneSet& cref, const PlaneSet& U, int np, int nimg,
                                      const ThreshSpec& th, float* gth, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_rows_c2r_mask_r2c<L>), dim3(L / R::ROWS, np, nimg), dim3(R::T), 0, s, Tin, cref, U, th, gth,
                       tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_dec_ct(const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                                const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s,
                                bool pre) {
    if (pre)
      hipLaunchKernelGGL((k_cols_dec<L, CT, true>), dim3(L / 2 / CT, map.nbands, nimg), dim3(ColCfg<L, CT>::T), 0, s,
                         Y1, ys, To, ts, Hm, Hb, map, tw);
    else
      hipLaunchKernelGGL((k_cols_dec<L, CT, false>), dim3(L / 2 / CT, map.nbands, nimg), dim3(ColCfg<L, CT>::T), 0,
                         s, Y1, ys, To, ts, Hm, Hb, map, tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_fwd_ct(float2* Y1, long long ys, int nplanes, int nimg, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_cols_fwd<L, CT>), dim3(L / 2 / CT, nplanes, nimg), dim3(ColCfg<L, CT>::T), 0, s, Y1, ys,
                       tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_rec_ct(const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                                const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_cols_rec<L, CT>), dim3(L / 2 / CT, map.ngroups, nimg), dim3(ColCfg<L, CT>::T), 0, s, U, us,
                       Q, qs, Hm, Hb, map, tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_dec3_ct(const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                                 const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
    if constexpr (L == 512 && CT >= 4)
      hipLaunchKernelGGL((k_cols_dec3<L, CT>), dim3(L / 2 / CT, map.nbands, nimg), dim3(CT * 64), 0, s, Y1, ys, To,
                         ts, Hm, Hb, map, tw);
    if constexpr (L == 1024)
      hipLaunchKernelGGL((k_cols_dec3k<CT>), dim3(L / 2 / CT, map.nbands, nimg), dim3(CT * 128), 0, s, Y1, ys, To, ts,
                         Hm, Hb, map, tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_rec3_ct(const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                                 const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s) {
    if constexpr (L == 512 && CT >= 4)
      hipLaunchKernelGGL((k_cols_rec3<L, CT>), dim3(L / 2 / CT, map.ngroups, nimg), dim3(CT * 64), 0, s, U, us, Q, qs,
                         Hm, Hb, map, tw);
    if constexpr (L == 1024)
      hipLaunchKernelGGL((k_cols_rec3k<CT>), dim3(L / 2 / CT, map.ngroups, nimg), dim3(CT * 128), 0, s, U, us, Q, qs,
                         Hm, Hb, map, tw);
    return hipGetLastError();
  }
  // ct: column tile width (8 or 16; 0 = default for L); -4 / -8: the
  // three-pass engine with 4 / 8 columns (L = 512 and 1024, pre only)
  static hipError_t cols_dec(const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                             const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s, int ct,
                             bool pre) {
    if (ct < 0) {
      if ((L != 512 && L != 1024) || !pre) return hipErrorInvalidValue;
      if (ct == -2 && L == 1024) return cols_dec3_ct<2>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s);
      if (ct == -4) return cols_dec3_ct<4>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s);
      if (ct == -8) return cols_dec3_ct<8>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s);
      return hipErrorInvalidValue;
    }
    if (ct == 0) ct = ColTile<L>::dflt;
    if (ct == 4 && ColTile<L>::narrow)
      return cols_dec_ct<(ColTile<L>::narrow ? 4 : 8)>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s, pre);
    if (ct == 8) return cols_dec_ct<8>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s, pre);
    if (ct == 16 && ColTile<L>::wide)
      return cols_dec_ct<(ColTile<L>::wide ? 16 : 8)>(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s, pre);
    return hipErrorInvalidValue;
  }
  static hipError_t cols_fwd(float2* Y1, long long ys, int nplanes, int nimg, const float2* tw, hipStream_t s, int ct) {
    if (ct <= 0) ct = ColTile<L>::dflt;
    if (ct == 4 && ColTile<L>::narrow) return cols_fwd_ct<(ColTile<L>::narrow ? 4 : 8)>(Y1, ys, nplanes, nimg, tw, s);
    if (ct == 8) return cols_fwd_ct<8>(Y1, ys, nplanes, nimg, tw, s);
    if (ct == 16 && ColTile<L>::wide) return cols_fwd_ct<(ColTile<L>::wide ? 16 : 8)>(Y1, ys, nplanes, nimg, tw, s);
    return hipErrorInvalidValue;
  }
  static hipError_t cols_rec(const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                             const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s, int ct) {
    if (ct < 0) {
      if (L != 512 && L != 1024) return hipErrorInvalidValue;
      if (ct == -2 && L == 1024) return cols_rec3_ct<2>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
      if (ct == -4) return cols_rec3_ct<4>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
      if (ct == -8) return cols_rec3_ct<8>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
      return hipErrorInvalidValue;
    }
    if (ct == 0) ct = ColTile<L>::dflt;
    if (ct == 4 && ColTile<L>::narrow) return cols_rec_ct<(ColTile<L>::narrow ? 4 : 8)>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
    if (ct == 8) return cols_rec_ct<8>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
    if (ct == 16 && ColTile<L>::wide) return cols_rec_ct<(ColTile<L>::wide ? 16 : 8)>(U, us, Q, qs, Hm, Hb, map, nimg, tw, s);
    return hipErrorInvalidValue;
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

hipError_t launch_rows_c2r_mask_r2c(int L, const PlaneSet& Tin, const PlaneSet& cref, const PlaneSet& U, int np,
                                    int nimg, const ThreshSpec& th, float* gth, const float2* tw, hipStream_t s) {
  if (th.mode != 2 || !th.dev || !gth) return hipErrorInvalidValue;
  SHX_BY_L(L, LT::rows_c2r_mask_r2c(Tin, cref, U, np, nimg, th, gth, tw, s))
}

hipError_t launch_cols_dec(int L, const float2* Y1, long long ys, float2* To, long long ts, const float2* Hm,
                           const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s, int ct,
                           bool pre) {
  SHX_BY_L(L, LT::cols_dec(Y1, ys, To, ts, Hm, Hb, map, nimg, tw, s, ct, pre))
}

hipError_t launch_cols_fwd(int L, float2* Y1, long long ys, int nplanes, int nimg, const float2* tw, hipStream_t s,
                           int ct) {
  SHX_BY_L(L, LT::cols_fwd(Y1, ys, nplanes, nimg, tw, s, ct))
}

// ---------------------------------------------------------------------------
// DCT-domain pyramid decomposition.
//
// With the half-sample symmetric extension of the a trous pyramid and its
// symmetric filters, every level is diagonal in the 2D DCT-II basis.  With
// X = D2(x), the raw DCT-II X[k] = sum_n x[n] cos(pi k (2n+1) / 2L) on both
// axes, the band-pass plane of scale j is y_j = D2^-1(M_j . X) and the
// lowpass D2^-1(M_low . X), where M_j = H1(m_j) prod_{i<j} H0(m_i) multiplies
// the filters' cosine responses at the dilations m = 1, 2, 4, ... (built on
// the host in FP64, divided by L^2 for the two unnormalised inverse passes).
// The pyramid's S dependent levels and the scale spectra's two FFT passes
// become three passes:
//   k_dct2_cols     columns of x: D2 (two real columns as one complex column)
//   k_dct_rows_dec  per (row block, plane): rows of X: D2, times the plane's
//                   mask, D2^-1, then the forward real row FFT (scale planes:
//                   the packed spectrum rows k_rows_r2c writes; lowpass: the
//                   real row)
//   k_dct_cols_dec  per (column tile, plane): D2^-1 along the columns, then
//                   the forward column FFT (scale planes, in place: the
//                   spectra k_cols_fwd produces); lowpass: D2^-1 only
// DCT-II by an L-point FFT (Makhoul): v[n] = x[perm(n)], perm(n) = 2n for
// n < L/2 and 2L-1-2n above, V = FFT(v) and
//   X[k] = (e^{-i t_k} V[k] + e^{i t_k} V[L-k]) / 2,   t_k = pi k / (2L)
// (Re(e^{-i t_k} V[k]) for real x).  Inverse: V[k] = e^{i t_k} (X[k] - i X[L-k])
// with X[L] = 0, v = IFFT(V), x[perm(n)] = v[n].  dtw[k] = e^{-i t_k}.
__device__ __forceinline__ int dct_perm(int n, int L) { return n < L / 2 ? 2 * n : 2 * L - 1 - 2 * n; }

// Columns of real planes (pairs of columns as one complex column): X = D2.
// grid: (L/2 / CT, 1, nimg), walked in the XCD-aware order (neighbouring
// narrow tiles share 128-byte lines).
template <int L, int CT_>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_dct2_cols(const float* __restrict__ x, long long xs,
                                                             float* __restrict__ out, long long os,
                                                             const float2* __restrict__ tw,
                                                             const float2* __restrict__ dtw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2, R1 = C::R1, R2 = C::R2, CT = C::CT;
  __shared__ float2 buf[C::LDS];
  int tile, plane, b;
  col_block_order(tile, plane, b);
  const int c0 = tile * CT;
  const float2* src = reinterpret_cast<const float2*>(x + (size_t)b * xs) + c0;
  float2* dst = reinterpret_cast<float2*>(out + (size_t)b * os) + c0;
  const int c = threadIdx.x % CT, k2 = threadIdx.x / CT;
  const float2 w0 = dtw[k2];  // with the data loads, not after the transpose
  col_fwd_A_src<L, CT_, false>([&](int row, int cc) { return src[(size_t)dct_perm(row, L) * NH + cc]; }, buf, tw);
  __syncthreads();
  float2 u[R1];
  col_fwd_B<L, CT_>(u, buf);
  __syncthreads();  // pass-B reads of buf are complete
  {
    float2* row = buf + C::Lds::ix(k2, c);  // V at row k = k2 + R2 k1
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) row[k1 * (R2 / R1) * C::Lds::BLK] = u[k1];
  }
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    const float2 vm = buf[C::Lds::ix((L - k) & (L - 1), c)];
    const float2 e = ctw_mul<4 * L / R2, -1>(w0, k1);
    dst[(size_t)k * NH + c] = cscale(cadd(cmul(e, u[k1]), cmulc(vm, e)), 0.5f);
  }
}

// Rows of X: D2, times mask plane p, D2^-1 and (scale planes) the forward
// real row FFT.  Plane p < nspec: packed spectrum rows of spec plane p;
// p == nspec: the real rows of the lowpass plane.  grid: (L / ROWS, nspec + 1,
// nimg).  The rows of one block go through the row transforms of RowCfg<L, 4>.
template <int L>
using DctRowCfg = RowCfg<L, 4>;

template <int L>
__global__ void __launch_bounds__(DctRowCfg<L>::T) k_dct_rows_dec(const float* __restrict__ X, long long xs,
                                                                   float2* __restrict__ spec, long long ss, int nspec,
                                                                   float* __restrict__ low, long long ls,
                                                                   const float* __restrict__ mask,
                                                                   const float2* __restrict__ tw,
                                                                   const float2* __restrict__ dtw) {
  using C = DctRowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2, ROWS = C::ROWS;
  __shared__ float2 buf[C::LDS];
  // the row v' of the inverse in natural order, in buf between the FFTs
  float (*vrow)[L] = reinterpret_cast<float (*)[L]>(buf);
  static_assert(ROWS * L <= 2 * C::LDS, "v' fits in the transpose tile");
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const float* xr0 = X + (size_t)b * xs + (size_t)row0 * L;
  // D2 of each row: the real FFT of v = x[perm(.)], whose (even, odd) pairs
  // are z[n] = (x[perm(2n)], x[perm(2n+1)])
  const int r = tid / R1, n1 = tid % R1;
  const bool validA = r < ROWS;
  const int rB = tid / R2, k2 = tid % R2;
  // (x[4n], x[4n+2]) for n < N/2 and (x[4m+3], x[4m+1]), m = N-1-n, above:
  // one 16-byte load per pair
  float2 v[R2];
  if (validA) {
    const float4* xr = reinterpret_cast<const float4*>(xr0 + (size_t)r * L);
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = n1 + R1 * n2;
      if (n2 < R2 / 2) {
        const float4 f = xr[n];
        v[n2] = make_float2(f.x, f.z);
      } else {
        const float4 f = xr[N - 1 - n];
        v[n2] = make_float2(f.w, f.y);
      }
    }
  }
  const auto ftw = row_fwd_tw<L, R1, R2>(n1, k2, tw);
  const float2 w0 = dtw[k2];
  const float* mrow = mask + (size_t)p * L * L + (size_t)(row0 + rB) * L;
  float mk[R1], mm[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    mk[k1] = mrow[k];
    mm[k1] = mrow[k == 0 ? N : L - k];
  }
  __builtin_amdgcn_sched_barrier(0);
  float2 P[R1];
  row_fwd_to_P<L, R1, R2, C::PADSH>(v, validA, r, n1, buf, rB, k2, ftw, P);
  // mask and D2^-1 pre-twiddle: P'[k] = e^{i t_k} (m[k] W.x + i m[L-k] W.y),
  // W = e^{-i t_k} P[k]; k = 0 packs (m[0] X[0], m[L/2] V[L/2])
  float2 Q[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    const float2 e = ctw_mul<4 * L / R2, -1>(w0, k1);
    const float2 W = cmul(e, P[k1]);
    Q[k1] = (k == 0) ? make_float2(mk[k1] * P[k1].x, mm[k1] * P[k1].y)
                     : cmulc(make_float2(mk[k1] * W.x, mm[k1] * W.y), e);
  }
  // inverse real row FFT from registers (split swapped): thread (rB, k2) is
  // the pass-A index k2 holding Q(k2 + R2 k1); partners through shuffles
  const int lane = threadIdx.x & 63;
  const int srcl = (lane & ~(R2 - 1)) | ((R2 - k2) & (R2 - 1));
  auto gather = [&](int, int, float2 (&xq)[R1], float2 (&xm)[R1]) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const float2 sh = shfl2(Q[R1 - 1 - k1], srcl);
      xq[k1] = Q[k1];
      xm[k1] = (k2 == 0) ? Q[(R1 - k1) & (R1 - 1)] : sh;
    }
  };
  __syncthreads();  // pass-B reads of buf are complete
  const int rC = tid / R1, kc = tid % R1;
  const bool validC = tid < ROWS * R1;
  float2 z[R2];
  row_inv_to_B<L, R2, R1, C::PADSH>(z, true, rB, k2, buf, validC, rC, kc, tw, gather);
  __syncthreads();  // every pass-B read of buf is complete before v' overwrites it
  // z[k1] = L (v'[2m], v'[2m+1]), m = kc + R1 k1: v' in natural order to LDS
  if (validC) {
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) {
      const int m = kc + R1 * k1;
      vrow[rC][2 * m] = z[k1].x;
      vrow[rC][2 * m + 1] = z[k1].y;
    }
  }
  __syncthreads();
  // y[perm(n)] = v'[n]: the real row y as pairs (y[2n], y[2n+1]) = (v'[n], v'[L-1-n])
  float2 yz[R2];
  if (validA) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = n1 + R1 * n2;
      yz[n2] = make_float2(vrow[r][n], vrow[r][L - 1 - n]);
    }
  }
  if (p == nspec) {
    if (validA) {
      float2* lr = reinterpret_cast<float2*>(low + (size_t)b * ls + (size_t)(row0 + r) * L);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) lr[n1 + R1 * n2] = yz[n2];
    }
    return;
  }
  __syncthreads();  // every read of v' is complete before the forward pass writes buf
  // the forward twiddles again, through an opaque copy of the table pointer so
  // that the first load's registers are not kept live across the inverse
  const float2* tw2 = tw;
  asm volatile("" : "+s"(tw2));
  const auto ftw2 = row_fwd_tw<L, R1, R2>(n1, k2, tw2);
  float2* dst = spec + (size_t)b * ss + (size_t)p * L * N + (size_t)row0 * N;
  row_fwd_from_A<L, R1, R2, C::PADSH>(yz, validA, r, n1, buf, dst, true, rB, k2, ftw2);
}

// ---------------------------------------------------------------------------
// Lane-pair row transforms, for N = L/2 = R x R (L = 128, 512).  The two-pass
// row FFT gives each length-R DFT to a lane pair (dft_pair: each lane half of
// it), so a row takes 2R lanes holding H = R/2 values each instead of R lanes
// holding R: half the arithmetic per lane on the short critical path of the
// DCT-domain pyramid's row passes (fewer rows in flight than SIMDs).
// Thread (r, p, h): row r of the block, pair p, lane h of the pair.
//   forward pass A: pair p = n1, lane h holds z[n1 + R (2j + h)];
//     after it, A[n1][k2] for k2 = k + H h, stored at n1 + R k2;
//   forward pass B: pair p = k2, lane h reads A[2j + h][k2] and ends with
//     Z[kk], kk = k2 + R (k + H h) (the "kk layout");
//   inverse stage 1: from the kk layout, a decimation-in-frequency pair
//     butterfly leaves lane h the odd/even half n1 = 2m + h, stored at n1 + R k2;
//   inverse stage 2: pair p = n1, lane h reads k2 = 2j + h and ends with
//     z[n], n = n1 + R (k + H h) (the "n layout").
template <int L>
struct PRow {
  static constexpr int N = L / 2;
  static constexpr int R = Split<N>::R1, H = R / 2;
  static constexpr bool ok = Split<N>::R1 == Split<N>::R2 && H >= 2;
  static constexpr int LANES = 2 * R;
  static constexpr int ROWS = ok ? 64 / LANES : 1;
  static constexpr int T = ROWS * LANES;
  static constexpr int PADSH = ilog2(R);
  static constexpr int LDS = lds_size<ROWS * N, PADSH>();
};

// Twiddles of a thread, loaded once: forward pass A (w_N^(n1 k2)), the r2c /
// c2r factor w_L^kk (w0 = w_L^p times -i for the upper half), inverse stage 1
// (w_N^(+2 p m) and the lane's extra w_N^(+p h)).
template <int L>
struct PRowTw {
  PairTw<PRow<L>::H> fa;
  PassTw<PRow<L>::H> ia;
  float2 fh;  // w_N^(+p h)
  float2 wh;  // w_L^(p + R H h)
};
template <int L>
__device__ __forceinline__ PRowTw<L> prow_tw(int p, int h, const float2* __restrict__ tw) {
  using C = PRow<L>;
  PRowTw<L> t;
  pair_tw<C::H, 2>(t.fa, p, h, tw);
  passA_tw<C::H, +1, 4>(t.ia, p, tw);
  const float2 w2p = tw[2 * p];  // w_N^p
  t.fh = h ? cconj(w2p) : make_float2(1.f, 0.f);
  const float2 w0 = tw[p];  // w_L^p; w_L^(R H) = w_L^(L/4) = -i
  t.wh = h ? make_float2(w0.y, -w0.x) : w0;
  return t;
}

// Forward pass A (registers z, n1 = p) then pass B: Z in the kk layout.
// Starts by writing buf (the caller makes sure earlier reads of it are done).
template <int L>
__device__ __forceinline__ void prow_fwd(float2 (&z)[PRow<L>::H], int r, int p, int h, float2* buf,
                                         const PRowTw<L>& t) {
  using C = PRow<L>;
  constexpr int N = C::N, R = C::R, H = C::H;
  dft_pair<H, -1>(z, h);
  pair_tw_apply<H, -1>(z, pin(t.fa));
#pragma unroll
  for (int k = 0; k < H; ++k) buf[lds_ix<C::PADSH>(r * N + p + R * (k + H * h))] = z[k];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < H; ++j) z[j] = buf[lds_ix<C::PADSH>(r * N + (2 * j + h) + R * p)];
  dft_pair<H, -1>(z, h);
}

// Z[(N - kk) mod N] for the thread's kk-layout values (all lanes take part).
template <int L>
__device__ __forceinline__ void prow_partner(const float2 (&Z)[PRow<L>::H], int p, int h,
                                             float2 (&Zm)[PRow<L>::H]) {
  using C = PRow<L>;
  constexpr int R = C::R, H = C::H;
  const int lane = threadIdx.x & 63;
  const int src = (lane & ~(C::LANES - 1)) + 2 * ((R - p) & (R - 1)) + (1 - h);
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float2 sh1 = shfl2(Z[H - 1 - k], src);
    const float2 sh0 = swap_pair(Z[(H - k) & (H - 1)]);
    Zm[k] = p != 0 ? sh1 : (k == 0 ? Z[0] : sh0);
  }
}

// Packed half spectrum of the real row from Z (kk layout), in place.
template <int L>
__device__ __forceinline__ void prow_r2c_post(float2 (&Z)[PRow<L>::H], int p, int h, const PRowTw<L>& t) {
  using C = PRow<L>;
  constexpr int R = C::R, H = C::H;
  float2 Zm[H];
  prow_partner<L>(Z, p, h, Zm);
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const bool k0 = k == 0 && p == 0 && h == 0;
    const float2 z = Z[k];
    Z[k] = k0 ? make_float2(z.x + z.y, z.x - z.y) : r2c_post(z, Zm[k], ctw_mul<L / R, -1>(t.wh, k));
  }
}

// Inverse of a packed half spectrum P (kk layout) to z (n layout), times L.
// Starts by writing buf (the caller makes sure earlier reads of it are done).
template <int L>
__device__ __forceinline__ void prow_inv(float2 (&P)[PRow<L>::H], int r, int p, int h, float2* buf,
                                         const PRowTw<L>& t) {
  using C = PRow<L>;
  constexpr int N = C::N, R = C::R, H = C::H;
  float2 Pm[H];
  prow_partner<L>(P, p, h, Pm);
  float2 a[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const bool k0 = k == 0 && p == 0 && h == 0;
    const float2 x = P[k];
    a[k] = k0 ? make_float2(x.x + x.y, x.x - x.y) : c2r_pre(x, Pm[k], ctw_mul<L / R, -1>(t.wh, k));
  }
  // decimation in frequency across the pair: lane 0 keeps x0 + x1, lane 1
  // (x0 - x1) w_R^{-k}; each then takes its half of the outputs
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float2 o = swap_pair(a[k]);
    a[k] = h == 0 ? cadd(a[k], o) : ctw_mul<R, +1>(csub(o, a[k]), k);
  }
  passA_apply<H, +1>(a, pin(t.ia));
#pragma unroll
  for (int m = 0; m < H; ++m) buf[lds_ix<C::PADSH>(r * N + (2 * m + h) + R * p)] = cmul(a[m], t.fh);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < H; ++j) P[j] = buf[lds_ix<C::PADSH>(r * N + p + R * (2 * j + h))];
  dft_pair<H, +1>(P, h);
}

// k_dct_rows_dec with lane-pair row transforms (same grid roles, rows per
// block PRow<L>::ROWS).
template <int L>
__global__ void __launch_bounds__(PRow<L>::T) k_dct_rows_dec_p(const float* __restrict__ X, long long xs,
                                                                float2* __restrict__ spec, long long ss, int nspec,
                                                                float* __restrict__ low, long long ls,
                                                                const float* __restrict__ mask,
                                                                const float2* __restrict__ tw,
                                                                const float2* __restrict__ dtw) {
  using C = PRow<L>;
  constexpr int N = C::N, R = C::R, H = C::H, ROWS = C::ROWS;
  __shared__ float2 buf[C::LDS];
  __shared__ float vrow[ROWS][L];
  const int tid = threadIdx.x;
  const int r = tid / C::LANES, l = tid % C::LANES, p = l >> 1, h = l & 1;
  const int pl = blockIdx.y, b = blockIdx.z;
  const int row = blockIdx.x * ROWS + r;
  const float* xr = X + (size_t)b * xs + (size_t)row * L;
  // z[n] = (x[perm(2n)], x[perm(2n+1)]) for n = p + R (2j + h)
  float2 z[H];
  {
    const float4* x4 = reinterpret_cast<const float4*>(xr);
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int n = p + R * (2 * j + h);
      if (j < H / 2) {
        const float4 f = x4[n];
        z[j] = make_float2(f.x, f.z);
      } else {
        const float4 f = x4[N - 1 - n];
        z[j] = make_float2(f.w, f.y);
      }
    }
  }
  const PRowTw<L> t = prow_tw<L>(p, h, tw);
  // e_kk = e^{-i pi kk / 2L}, kk = p + R (k + H h): dtw[p] times e^{-i pi h / 8}
  // (R H = L/4) times the compile-time e^{-i pi R k / 2L}
  float2 eh = dtw[p];
  if (h) eh = cmul(eh, make_float2(0.92387953251128674f, -0.38268343236508977f));
  const float* mrow = mask + (size_t)pl * L * L + (size_t)row * L;
  float mk[H], mm[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int kk = p + R * (k + H * h);
    mk[k] = mrow[kk];
    mm[k] = mrow[kk == 0 ? N : L - kk];
  }
  __builtin_amdgcn_sched_barrier(0);
  prow_fwd<L>(z, r, p, h, buf, t);
  prow_r2c_post<L>(z, p, h, t);
  // mask and D2^-1 pre-twiddle (as k_dct_rows_dec)
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const bool k0 = k == 0 && p == 0 && h == 0;
    const float2 e = ctw_mul<4 * L / R, -1>(eh, k);
    const float2 W = cmul(e, z[k]);
    z[k] = k0 ? make_float2(mk[k] * z[k].x, mm[k] * z[k].y) : cmulc(make_float2(mk[k] * W.x, mm[k] * W.y), e);
  }
  __syncthreads();  // pass-B reads of buf are complete
  prow_inv<L>(z, r, p, h, buf, t);
  // z[k] = L (v'[2n], v'[2n+1]), n = p + R (k + H h)
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int n = p + R * (k + H * h);
    vrow[r][2 * n] = z[k].x;
    vrow[r][2 * n + 1] = z[k].y;
  }
  __syncthreads();
  // y pairs (y[2n], y[2n+1]) = (v'[n], v'[L-1-n]) for n = p + R (2j + h)
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const int n = p + R * (2 * j + h);
    z[j] = make_float2(vrow[r][n], vrow[r][L - 1 - n]);
  }
  if (pl == nspec) {
    float2* lr = reinterpret_cast<float2*>(low + (size_t)b * ls + (size_t)row * L);
#pragma unroll
    for (int j = 0; j < H; ++j) lr[p + R * (2 * j + h)] = z[j];
    return;
  }
  prow_fwd<L>(z, r, p, h, buf, t);
  prow_r2c_post<L>(z, p, h, t);
  float2* dst = spec + (size_t)b * ss + (size_t)pl * L * N + (size_t)row * N;
#pragma unroll
  for (int k = 0; k < H; ++k) dst[p + R * (k + H * h)] = z[k];
}

// Columns of the planes k_dct_rows_dec wrote: D2^-1, then (scale planes) the
// forward column FFT, in place.  Plane p < nspec: spec plane p (complex
// columns); p == nspec: the lowpass plane (pairs of real columns).
// grid: (L/2 / CT, nspec + 1, nimg), walked in the XCD-aware order.
template <int L, int CT_>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_dct_cols_dec(float2* __restrict__ spec, long long ss,
                                                                int nspec, float* __restrict__ low, long long ls,
                                                                const float2* __restrict__ tw,
                                                                const float2* __restrict__ dtw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2, R1 = C::R1, R2 = C::R2, CT = C::CT;
  __shared__ float2 buf[C::LDS];
  int tile, p, b;
  col_block_order(tile, p, b);
  const int c0 = tile * CT;
  float2* base = (p == nspec ? reinterpret_cast<float2*>(low + (size_t)b * ls)
                             : spec + (size_t)b * ss + (size_t)p * L * NH) + c0;
  const int c = threadIdx.x % CT, k2 = threadIdx.x / CT;
  // pre-twiddle V[k] = e^{i t_k} (X[k] - i X[L-k]) in pass-B layout
  float2 xk[R1], xm[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    xk[k1] = base[(size_t)k * NH + c];
    xm[k1] = k == 0 ? make_float2(0.f, 0.f) : base[(size_t)(L - k) * NH + c];
  }
  const float2 w0 = dtw[k2];
  const PassTw<R1> ivt = col_inv_tw<L>(k2, tw);
  __builtin_amdgcn_sched_barrier(0);
  float2 q[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const float2 e = ctw_mul<4 * L / R2, -1>(w0, k1);
    q[k1] = cmulc(make_float2(xk[k1].x + xm[k1].y, xk[k1].y - xm[k1].x), e);
  }
  col_inv_A<L, CT_, false>(q, buf, ivt);
  if (p == nspec) {
    col_inv_B_to<L, CT_, false>(buf, [&](int row, int cc, float2 val) { base[(size_t)dct_perm(row, L) * NH + cc] = val; });
    return;
  }
  // y[perm(n)] = v[n] back into the tile in natural row order (after every
  // pass-B read), then the forward FFT from there (its pass A writes the tile
  // after all its reads): one LDS tile, which at L = 1024 doubles the blocks
  // per CU against a second tile for y
  col_inv_B_to<L, CT_, true>(buf, [&](int row, int cc, float2 val) { buf[C::Lds::ix(dct_perm(row, L), cc)] = val; });
  __syncthreads();
  col_fwd_A_src<L, CT_, true>([&](int row, int cc) { return buf[C::Lds::ix(row, cc)]; }, buf, tw);
  __syncthreads();
  float2 u[R1];
  col_fwd_B<L, CT_>(u, buf);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) base[(size_t)(k2 + R2 * k1) * NH + c] = u[k1];
}

// The front's column kernels on the engines with 64 threads per column
// (fft3.h, ColEng: 512 = 8 x 8 x 8, 1024 = 16 x 8 x 8): thread (c, tau) of
// CT * 64 threads holds rows tau + 64 s of its complex column.  Same results
// as k_dct2_cols / k_dct_cols_dec to FP32 rounding.
template <int L, int CT>
__global__ void __launch_bounds__(CT * 64) k_dct2_cols3(const float* __restrict__ x, long long xs,
                                                    float* __restrict__ out, long long os,
                                                    const float2* __restrict__ tw,
                                                    const float2* __restrict__ dtw) {
  constexpr int NH = L / 2;
  using G = ColEng<L, CT>;
  using E = typename G::E;
  constexpr int NV = G::NV;
  __shared__ float2 buf[E::LDS];
  int tile, plane, b;
  col_block_order(tile, plane, b);
  const int c0 = tile * CT;
  const float2* src = reinterpret_cast<const float2*>(x + (size_t)b * xs) + c0;
  float2* dst = reinterpret_cast<float2*>(out + (size_t)b * os) + c0;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  float2 v[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) v[s] = src[(size_t)dct_perm(tau + 64 * s, L) * NH + c];
  const typename G::Tw t = G::twiddles(tau, tw);
  const float2 w0 = dtw[tau];
  __builtin_amdgcn_sched_barrier(0);
  G::template fft<-1>(v, buf, t);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) buf[E::ix(tau + 64 * s, c)] = v[s];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const int k = tau + 64 * s;
    const float2 vm = buf[E::ix((L - k) & (L - 1), c)];
    const float2 e = ctw_mul<4 * L / 64, -1>(w0, s);
    dst[(size_t)k * NH + c] = cscale(cadd(cmul(e, v[s]), cmulc(vm, e)), 0.5f);
  }
}

template <int L, int CT>
__global__ void __launch_bounds__(CT * 64) k_dct_cols_dec3(float2* __restrict__ spec, long long ss, int nspec,
                                                       float* __restrict__ low, long long ls,
                                                       const float2* __restrict__ tw,
                                                       const float2* __restrict__ dtw) {
  constexpr int NH = L / 2;
  using G = ColEng<L, CT>;
  using E = typename G::E;
  constexpr int NV = G::NV;
  __shared__ float2 buf[E::LDS];
  int tile, p, b;
  col_block_order(tile, p, b);
  const int c0 = tile * CT;
  float2* base = (p == nspec ? reinterpret_cast<float2*>(low + (size_t)b * ls)
                             : spec + (size_t)b * ss + (size_t)p * L * NH) + c0;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  // pre-twiddle V[k] = e^{i t_k} (X[k] - i X[L-k]), X[L] = 0
  float2 xk[NV], xm[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const int k = tau + 64 * s;
    xk[s] = base[(size_t)k * NH + c];
    xm[s] = k == 0 ? make_float2(0.f, 0.f) : base[(size_t)(L - k) * NH + c];
  }
  const typename G::Tw t = G::twiddles(tau, tw);
  const float2 w0 = dtw[tau];
  __builtin_amdgcn_sched_barrier(0);
  float2 v[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const float2 e = ctw_mul<4 * L / 64, -1>(w0, s);
    v[s] = cmulc(make_float2(xk[s].x + xm[s].y, xk[s].y - xm[s].x), e);
  }
  G::template fft<+1>(v, buf, t);
  if (p == nspec) {
#pragma unroll
    for (int s = 0; s < NV; ++s) base[(size_t)dct_perm(tau + 64 * s, L) * NH + c] = v[s];
    return;
  }
  // y[perm(n)] = v[n] into the tile in natural row order, then the forward
  // FFT of y (every read of the tile done before each overwrite)
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) buf[E::ix(dct_perm(tau + 64 * s, L), c)] = v[s];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) v[s] = buf[E::ix(tau + 64 * s, c)];
  __syncthreads();
  G::template fft<-1>(v, buf, t);
#pragma unroll
  for (int s = 0; s < NV; ++s) base[(size_t)(tau + 64 * s) * NH + c] = v[s];
}

// Column tiles: the single-plane first pass takes the narrowest tile (most
// blocks); the per-plane last pass the narrow tile.  At 512 and 1024 both run
// on the engines with 64 threads per column (2 and 4 columns; measured 8.6 ->
// 7.8 and 12.5 -> 10.8 us at 512^2, 13.9 -> 12.4 and 23.0 -> 22.0 us at
// 1024^2, against the two-pass kernels).
template <int L>
struct LaunchDct {
  static hipError_t dec(const float* x, long long xs, float* xwork, long long ws, float2* spec, long long ss,
                        int nspec, float* low, long long ls, const float* mask, int nimg, const float2* tw,
                        const float2* dtw, hipStream_t s, int step) {
    constexpr int ct = ColTile<L>::narrow ? 4 : 8;
    constexpr int ct1 = 2;
    if (step == 0) {
      if constexpr (L == 512 || L == 1024)
        hipLaunchKernelGGL((k_dct2_cols3<L, 2>), dim3(L / 2 / 2, 1, nimg), dim3(2 * 64), 0, s, x, xs, xwork, ws, tw,
                           dtw);
      else
        hipLaunchKernelGGL((k_dct2_cols<L, ct1>), dim3(L / 2 / ct1, 1, nimg), dim3(ColCfg<L, ct1>::T), 0, s, x, xs,
                           xwork, ws, tw, dtw);
    } else if (step == 1) {
      if constexpr (PRow<L>::ok)
        hipLaunchKernelGGL((k_dct_rows_dec_p<L>), dim3(L / PRow<L>::ROWS, nspec + 1, nimg), dim3(PRow<L>::T), 0, s,
                           xwork, ws, spec, ss, nspec, low, ls, mask, tw, dtw);
      else
        hipLaunchKernelGGL((k_dct_rows_dec<L>), dim3(L / DctRowCfg<L>::ROWS, nspec + 1, nimg),
                           dim3(DctRowCfg<L>::T), 0, s, xwork, ws, spec, ss, nspec, low, ls, mask, tw, dtw);
    } else {
      if constexpr (L == 512 || L == 1024)
        hipLaunchKernelGGL((k_dct_cols_dec3<L, 4>), dim3(L / 2 / 4, nspec + 1, nimg), dim3(4 * 64), 0, s, spec, ss,
                           nspec, low, ls, tw, dtw);
      else
        hipLaunchKernelGGL((k_dct_cols_dec<L, ct>), dim3(L / 2 / ct, nspec + 1, nimg), dim3(ColCfg<L, ct>::T), 0, s,
                           spec, ss, nspec, low, ls, tw, dtw);
    }
    return hipGetLastError();
  }
};

hipError_t launch_dct_dec(int L, int step, const float* x, long long xs, float* xwork, long long ws, float2* spec,
                          long long ss, int nspec, float* low, long long ls, const float* mask, int nimg,
                          const float2* tw, const float2* dtw, hipStream_t s) {
  if (step < 0 || step > 2) return hipErrorInvalidValue;
  switch (L) {
    case 64: return LaunchDct<64>::dec(x, xs, xwork, ws, spec, ss, nspec, low, ls, mask, nimg, tw, dtw, s, step);
    case 128: return LaunchDct<128>::dec(x, xs, xwork, ws, spec, ss, nspec, low, ls, mask, nimg, tw, dtw, s, step);
    case 256: return LaunchDct<256>::dec(x, xs, xwork, ws, spec, ss, nspec, low, ls, mask, nimg, tw, dtw, s, step);
    case 512: return LaunchDct<512>::dec(x, xs, xwork, ws, spec, ss, nspec, low, ls, mask, nimg, tw, dtw, s, step);
    case 1024: return LaunchDct<1024>::dec(x, xs, xwork, ws, spec, ss, nspec, low, ls, mask, nimg, tw, dtw, s, step);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// DCT-domain pyramid reconstruction.  The a trous reconstruction of the scale
// planes y_j and the lowpass is x = D2^-1( sum_j G_j . D2(y_j) + G_low . D2(low) )
// with G_j = G1(m_j) prod_{i<j} G0(m_i) (synthesis filters' cosine responses,
// divided by L^2 on the host), so the S dependent levels and the scale sum's
// row pass become three passes:
//   k_dct_rows_rec  per (row block, plane): y_j = C2R(sum of the scale's
//                   direction-group partials) (lowpass: the thresholded real
//                   row), then the row DCT-II; real rows out
//   k_dct_cols_rec  per (column tile, plane): column DCT-II, times the plane's
//                   mask, inverse column DCT-II, in place
//   k_dct_rows_out  per row block: sum of the planes, inverse row DCT-II
// (The lowpass hard threshold of the reconstruction is fused into the first.)

// Real DCT-II row values from the packed half spectrum P of v = y[perm(.)]
// (thread (rB, k2) holding P(k2 + R2 k1)): X[k] = Re(e_k V[k]),
// X[L-k] = -Im(e_k V[k]), X[0] = V[0], X[L/2] = V[L/2] / sqrt(2).
template <int L, int R1, int R2>
__device__ __forceinline__ void dct_row_store(const float2 (&P)[R1], int k2, float2 w0, float* __restrict__ drow) {
  constexpr int N = L / 2;
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    if (k == 0) {
      drow[0] = P[k1].x;
      drow[N] = P[k1].y * 0.70710678118654752f;
    } else {
      const float2 W = cmul(ctw_mul<4 * L / R2, -1>(w0, k1), P[k1]);
      drow[k] = W.x;
      drow[L - k] = -W.y;
    }
  }
}

// grid: (L / ROWS, nspec + 1, nimg).  Plane p < nspec: scale p from the
// direction-group partials gf.first[p] .. gf.first[p+1]-1 of Qg (packed rows,
// plane stride L*L/2 complex); p == nspec: the lowpass.  D: nspec + 1 real
// planes (stride ds per image).
template <int L>
__global__ void __launch_bounds__(DctRowCfg<L>::T) k_dct_rows_rec(const float2* __restrict__ Qg, long long qs,
                                                                   SumTable gf, int nspec,
                                                                   const float* __restrict__ low, long long ls,
                                                                   LowThresh lt, float* __restrict__ D, long long ds,
                                                                   const float2* __restrict__ tw,
                                                                   const float2* __restrict__ dtw) {
  using C = DctRowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2, ROWS = C::ROWS;
  __shared__ float2 buf[C::LDS];
  // the real row y as pairs (y[2i], y[2i+1]), in buf between the two FFTs
  // (one LDS tile per block: at L = 1024 the block count per CU is then set
  // by registers, not LDS)
  float2 (*yrow)[N] = reinterpret_cast<float2 (*)[N]>(buf);
  static_assert(ROWS * N <= C::LDS, "y fits in the transpose tile");
  const int tid = threadIdx.x;
  const int p = blockIdx.y, b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const int r = tid / R1, n1 = tid % R1;
  const bool validA = r < ROWS;
  const int rB = tid / R2, k2 = tid % R2;
  const float2 w0 = dtw[k2];
  // z[n] = (y[perm(2n)], y[perm(2n+1)]): (y[4n], y[4n+2]) for n < N/2 and
  // (y[4m+3], y[4m+1]), m = N-1-n, above
  float2 v[R2];
  if (p < nspec) {
    const int q0 = gf.first[p], q1 = gf.first[p + 1];
    const float2* qb = Qg + (size_t)b * qs;
    float2 u[R1];
    if constexpr (L >= 1024) {
      // one group per round trip and P[N-k] through LDS: the two-group
      // gather with register partners holds four R2-value arrays (256 VGPRs
      // at L = 1024).  Same sums in the same order.
      auto gather = [&](int rr, int nn, float2 (&x)[R2]) {
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) x[n2] = make_float2(0.f, 0.f);
        for (int q = q0; q < q1; ++q) {
          float2 a[R2];
          load_row_x<N, R1, R2>(qb + (size_t)q * L * N + (size_t)(row0 + rr) * N, nn, a);
          __builtin_amdgcn_sched_barrier(0);  // the group's loads in flight together
#pragma unroll
          for (int n2 = 0; n2 < R2; ++n2) x[n2] = cadd(x[n2], a[n2]);
        }
      };
      row_inv_to_B_lds<L, R1, R2, C::PADSH>(u, validA, r, n1, buf, true, rB, k2, tw, gather);
    } else {
    // the scale's partials summed two groups per round trip (an odd last
    // group pairs with itself at weight 0); P[N-k] of the sum from the
    // partner lane (as gather_row)
    auto gather = [&](int rr, int nn, float2 (&x)[R2], float2 (&xm)[R2]) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) x[n2] = make_float2(0.f, 0.f);
      for (int q = q0; q < q1; q += 2) {
        const int q2 = q + 1 < q1 ? q + 1 : q;
        const float wgt = q + 1 < q1 ? 1.0f : 0.0f;
        const float2* ra = qb + (size_t)q * L * N + (size_t)(row0 + rr) * N;
        const float2* rc = qb + (size_t)q2 * L * N + (size_t)(row0 + rr) * N;
        float2 a[R2], c[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
          a[n2] = ra[nn + R1 * n2];
          c[n2] = rc[nn + R1 * n2];
        }
        __builtin_amdgcn_sched_barrier(0);  // both groups' loads in flight together
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) x[n2] = cadd(cadd(x[n2], a[n2]), cscale(c[n2], wgt));
      }
      const int lane = threadIdx.x & 63;
      const int src = (lane & ~(R1 - 1)) | ((R1 - nn) & (R1 - 1));
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const float2 sh = shfl2(x[R2 - 1 - n2], src);
        xm[n2] = nn == 0 ? x[(R2 - n2) & (R2 - 1)] : sh;
      }
    };
    row_inv_to_B<L, R1, R2, C::PADSH>(u, validA, r, n1, buf, true, rB, k2, tw, gather);
    }
    __syncthreads();  // every pass-B read of buf is complete before y overwrites it
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) yrow[rB][k2 + R2 * k1] = u[k1];
    __syncthreads();
    if (validA) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = n1 + R1 * n2;
        const int m = n2 < R2 / 2 ? n : N - 1 - n;
        const float2 a = yrow[r][2 * m], c = yrow[r][2 * m + 1];
        v[n2] = n2 < R2 / 2 ? make_float2(a.x, c.x) : make_float2(c.y, a.y);
      }
    }
  } else {
    const int rule = low_rule(lt);
    const float th = low_value(lt, b);
    if (validA) {
      const float4* xr = reinterpret_cast<const float4*>(low + (size_t)b * ls + (size_t)(row0 + r) * L);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int n = n1 + R1 * n2;
        if (n2 < R2 / 2) {
          const float4 f = xr[n];
          v[n2] = shrink2(make_float2(f.x, f.z), th, rule);
        } else {
          const float4 f = xr[N - 1 - n];
          v[n2] = shrink2(make_float2(f.w, f.y), th, rule);
        }
      }
    }
  }
  __syncthreads();  // every read of y is complete before the forward pass writes buf
  // the forward pass's twiddles only now: loaded at the top they stay live
  // across the inverse pass
  const auto ftw = row_fwd_tw<L, R1, R2>(n1, k2, tw);
  float2 P[R1];
  row_fwd_to_P<L, R1, R2, C::PADSH>(v, validA, r, n1, buf, rB, k2, ftw, P);
  dct_row_store<L, R1, R2>(P, k2, w0, D + (size_t)b * ds + (size_t)p * L * L + (size_t)(row0 + rB) * L);
}

// Column DCT-II of plane p of D (pairs of real columns), times the plane's
// mask, then the inverse column DCT-II, in place.  grid: (L/2 / CT, nplanes,
// nimg), XCD-aware order.
template <int L, int CT_>
__global__ void __launch_bounds__(CT_ * Split<L>::R2) k_dct_cols_rec(float* __restrict__ D, long long ds,
                                                                const float* __restrict__ mask,
                                                                const float2* __restrict__ tw,
                                                                const float2* __restrict__ dtw) {
  using C = ColCfg<L, CT_>;
  constexpr int NH = L / 2, R1 = C::R1, R2 = C::R2, CT = C::CT;
  __shared__ float2 buf[C::LDS];
  int tile, p, b;
  col_block_order(tile, p, b);
  const int c0 = tile * CT;
  float2* base = reinterpret_cast<float2*>(D + (size_t)b * ds + (size_t)p * L * L) + c0;
  const int c = threadIdx.x % CT, k2 = threadIdx.x / CT;
  col_fwd_A_src<L, CT_, false>([&](int row, int cc) { return base[(size_t)dct_perm(row, L) * NH + cc]; }, buf, tw);
  const float2 w0 = dtw[k2];
  const PassTw<R1> ivt = col_inv_tw<L>(k2, tw);
  // the mask of this thread's two real columns at rows k = k2 + R2 k1
  const float2* mcol = reinterpret_cast<const float2*>(mask + (size_t)p * L * L) + c0 + c;
  float2 mk[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) mk[k1] = mcol[(size_t)(k2 + R2 * k1) * NH];
  __syncthreads();
  float2 u[R1];
  col_fwd_B<L, CT_>(u, buf);
  __syncthreads();  // pass-B reads of buf are complete
  float2* row = buf + C::Lds::ix(k2, c);  // row k2 + R2 k1
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) row[k1 * (R2 / R1) * C::Lds::BLK] = u[k1];
  __syncthreads();
  // X = D2 (two real columns as one complex column), times the mask
  float2 xm[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    const float2 vm = buf[C::Lds::ix((L - k) & (L - 1), c)];
    const float2 e = ctw_mul<4 * L / R2, -1>(w0, k1);
    const float2 X = cscale(cadd(cmul(e, u[k1]), cmulc(vm, e)), 0.5f);
    xm[k1] = make_float2(mk[k1].x * X.x, mk[k1].y * X.y);
  }
  __syncthreads();  // reads of the exchanged spectrum are complete
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) row[k1 * (R2 / R1) * C::Lds::BLK] = xm[k1];
  __syncthreads();
  // D2^-1 pre-twiddle V[k] = e^{i t_k} (X[k] - i X[L-k]), X[L] = 0
  float2 q[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    const float2 xl = k == 0 ? make_float2(0.f, 0.f) : buf[C::Lds::ix(L - k, c)];
    const float2 e = ctw_mul<4 * L / R2, -1>(w0, k1);
    q[k1] = cmulc(make_float2(xm[k1].x + xl.y, xm[k1].y - xl.x), e);
  }
  col_inv_A<L, CT_, true>(q, buf, ivt);
  col_inv_B_to<L, CT_, false>(buf, [&](int rr, int cc, float2 val) { base[(size_t)dct_perm(rr, L) * NH + cc] = val; });
}


// k_dct_cols_rec on the engines with 64 threads per column (ColEng): thread (c, tau)
// of CT * 64 threads holds rows tau + 64 s of its complex column; the two
// spectrum partners (V[L-k], then X[L-k]) are exchanged through the tile.
// Twice the waves of the two-pass kernel, each with half the values:
// measured 13.2 -> 11.7 us at 512^2 and 31.4 -> 29.5 us at 1024^2 (4
// columns), same results to rounding.
template <int L, int CT>
__global__ void __launch_bounds__(CT * 64) k_dct_cols_rec3(float* __restrict__ D, long long ds,
                                                       const float* __restrict__ mask,
                                                       const float2* __restrict__ tw,
                                                       const float2* __restrict__ dtw) {
  constexpr int NH = L / 2;
  using G = ColEng<L, CT>;
  using E = typename G::E;
  constexpr int NV = G::NV;
  __shared__ float2 buf[E::LDS];
  int tile, p, b;
  col_block_order(tile, p, b);
  const int c0 = tile * CT;
  float2* base = reinterpret_cast<float2*>(D + (size_t)b * ds + (size_t)p * L * L) + c0;
  const int c = threadIdx.x % CT, tau = threadIdx.x / CT;
  const float2* mcol = reinterpret_cast<const float2*>(mask + (size_t)p * L * L) + c0 + c;
  float2 v[NV], mk[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    v[s] = base[(size_t)dct_perm(tau + 64 * s, L) * NH + c];
    mk[s] = mcol[(size_t)(tau + 64 * s) * NH];
  }
  const typename G::Tw t = G::twiddles(tau, tw);
  const float2 w0 = dtw[tau];
  __builtin_amdgcn_sched_barrier(0);
  G::template fft<-1>(v, buf, t);
  // X = D2 (two real columns as one complex column) from V[k] and V[L-k]
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) buf[E::ix(tau + 64 * s, c)] = v[s];
  __syncthreads();
  float2 x[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const int k = tau + 64 * s;
    const float2 vm = buf[E::ix((L - k) & (L - 1), c)];
    const float2 e = ctw_mul<4 * L / 64, -1>(w0, s);
    const float2 X = cscale(cadd(cmul(e, v[s]), cmulc(vm, e)), 0.5f);
    x[s] = make_float2(mk[s].x * X.x, mk[s].y * X.y);
  }
  __syncthreads();  // reads of the exchanged spectrum are complete
#pragma unroll
  for (int s = 0; s < NV; ++s) buf[E::ix(tau + 64 * s, c)] = x[s];
  __syncthreads();
  // D2^-1 pre-twiddle V[k] = e^{i t_k} (X[k] - i X[L-k]), X[L] = 0
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const int k = tau + 64 * s;
    const float2 xl = k == 0 ? make_float2(0.f, 0.f) : buf[E::ix(L - k, c)];
    const float2 e = ctw_mul<4 * L / 64, -1>(w0, s);
    v[s] = cmulc(make_float2(x[s].x + xl.y, x[s].y - xl.x), e);
  }
  __syncthreads();
  G::template fft<+1>(v, buf, t);
#pragma unroll
  for (int s = 0; s < NV; ++s) base[(size_t)dct_perm(tau + 64 * s, L) * NH + c] = v[s];
}

// Sum of the nplanes planes of D, inverse row DCT-II, into x.
// grid: (L / ROWS, 1, nimg).
template <int L>
__global__ void __launch_bounds__(DctRowCfg<L>::T) k_dct_rows_out(const float* __restrict__ D, long long ds,
                                                                   int nplanes, float* __restrict__ x, long long xs,
                                                                   const float2* __restrict__ tw,
                                                                   const float2* __restrict__ dtw) {
  using C = DctRowCfg<L>;
  constexpr int N = C::N, R1 = C::R1, R2 = C::R2, ROWS = C::ROWS;
  static_assert(L % (4 * R2) == 0, "row chunking");
  __shared__ float2 buf[C::LDS];
  __shared__ float xrow[ROWS][L];
  const int tid = threadIdx.x;
  const int b = blockIdx.z;
  const int row0 = blockIdx.x * ROWS;
  const int r = tid / R1, n1 = tid % R1;
  const bool validA = r < ROWS;
  const int rB = tid / R2, k2 = tid % R2;
  const float2 e1 = dtw[n1];
  const RowInvTw<R2> itw = row_inv_tw<L, R1, R2>(n1, tw);
  // the summed row X (thread (rB, k2): 16-byte chunks k2 + R2 i), three
  // planes per round trip (past the last plane, the last again at weight 0)
  {
    constexpr int NC = L / 4 / R2;
    const float4* src = reinterpret_cast<const float4*>(D + (size_t)b * ds + (size_t)(row0 + rB) * L);
    float4 acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int pl = 0; pl < nplanes; pl += 3) {
      float4 a[3][NC];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int q = pl + j < nplanes ? pl + j : nplanes - 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) a[j][i] = src[(size_t)q * L * L / 4 + k2 + R2 * i];
      }
      __builtin_amdgcn_sched_barrier(0);  // the three planes' loads in flight together
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float wgt = pl + j < nplanes ? 1.0f : 0.0f;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          acc[i].x += wgt * a[j][i].x;
          acc[i].y += wgt * a[j][i].y;
          acc[i].z += wgt * a[j][i].z;
          acc[i].w += wgt * a[j][i].w;
        }
      }
    }
    float4* xr = reinterpret_cast<float4*>(&xrow[rB][0]);
#pragma unroll
    for (int i = 0; i < NC; ++i) xr[k2 + R2 * i] = acc[i];
  }
  __syncthreads();
  // packed spectrum of v: P'[k] = e^{i t_k} (X[k] - i X[L-k]) (k >= 1),
  // P'[0] = (X[0], sqrt(2) X[L/2]); partners P'[N-k], e_{N-k} = e^{-i pi/4} conj(e_k)
  auto gather = [&](int rr, int nn, float2 (&xq)[R2], float2 (&xmq)[R2]) {
    const float* X = xrow[rr];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int k = nn + R1 * n2;
      const float2 e = ctw_mul<4 * L / R1, -1>(e1, n2);
      if (k == 0) {
        xq[n2] = xmq[n2] = make_float2(X[0], 1.41421356237309505f * X[N]);
      } else {
        xq[n2] = cmulc(make_float2(X[k], -X[L - k]), e);
        // e_{N-k}^* = e^{i pi/4} e_k
        const float2 em = cmul(make_float2(0.70710678118654752f, 0.70710678118654752f), e);
        xmq[n2] = cmul(make_float2(X[N - k], -X[N + k]), em);
      }
    }
  };
  float2 u[R1];
  row_inv_to_B_tw<L, R1, R2, C::PADSH>(u, validA, r, n1, buf, true, rB, k2, itw, gather);
  __syncthreads();  // every read of xrow is complete
  // u[k1] = L (v[2m], v[2m+1]), m = k2 + R2 k1; x[perm(n)] = v[n]
  {
    float* vr = xrow[rB];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int m = k2 + R2 * k1;
      vr[2 * m] = u[k1].x;
      vr[2 * m + 1] = u[k1].y;
    }
  }
  __syncthreads();
  // (x[2m], x[2m+1]) = (v[m], v[L-1-m])
  float2* xo = reinterpret_cast<float2*>(x + (size_t)b * xs + (size_t)(row0 + rB) * L);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int m = k2 + R2 * k1;
    xo[m] = make_float2(xrow[rB][m], xrow[rB][L - 1 - m]);
  }
}

// k_dct_rows_rec with lane-pair row transforms.
template <int L>
__global__ void __launch_bounds__(PRow<L>::T) k_dct_rows_rec_p(const float2* __restrict__ Qg, long long qs,
                                                                SumTable gf, int nspec,
                                                                const float* __restrict__ low, long long ls,
                                                                LowThresh lt, float* __restrict__ D, long long ds,
                                                                const float2* __restrict__ tw,
                                                                const float2* __restrict__ dtw) {
  using C = PRow<L>;
  constexpr int N = C::N, R = C::R, H = C::H, ROWS = C::ROWS;
  __shared__ float2 buf[C::LDS];
  __shared__ float2 yrow[ROWS][N];
  const int tid = threadIdx.x;
  const int r = tid / C::LANES, l = tid % C::LANES, p = l >> 1, h = l & 1;
  const int pl = blockIdx.y, b = blockIdx.z;
  const int row = blockIdx.x * ROWS + r;
  const PRowTw<L> t = prow_tw<L>(p, h, tw);
  float2 eh = dtw[p];
  if (h) eh = cmul(eh, make_float2(0.92387953251128674f, -0.38268343236508977f));
  float2 z[H];
  if (pl < nspec) {
    // the scale's direction-group partials at kk = p + R (k + H h), two groups
    // per round trip, then the inverse row transform: y pairs in the n layout
    const int q0 = gf.first[pl], q1 = gf.first[pl + 1];
    const float2* qrow = Qg + (size_t)b * qs + (size_t)row * N;
#pragma unroll
    for (int k = 0; k < H; ++k) z[k] = make_float2(0.f, 0.f);
    for (int q = q0; q < q1; q += 2) {
      const int q2 = q + 1 < q1 ? q + 1 : q;
      const float wgt = q + 1 < q1 ? 1.0f : 0.0f;
      float2 a[H], c[H];
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const int kk = p + R * (k + H * h);
        a[k] = qrow[(size_t)q * L * N + kk];
        c[k] = qrow[(size_t)q2 * L * N + kk];
      }
      __builtin_amdgcn_sched_barrier(0);  // both groups' loads in flight together
#pragma unroll
      for (int k = 0; k < H; ++k) z[k] = cadd(cadd(z[k], a[k]), cscale(c[k], wgt));
    }
    prow_inv<L>(z, r, p, h, buf, t);
#pragma unroll
    for (int k = 0; k < H; ++k) yrow[r][p + R * (k + H * h)] = z[k];
    __syncthreads();  // also: every read of buf is complete
    // z[j] = (y[perm(2n)], y[perm(2n+1)]), n = p + R (2j + h)
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int n = p + R * (2 * j + h);
      const int m = j < H / 2 ? n : N - 1 - n;
      const float2 a = yrow[r][2 * m], c = yrow[r][2 * m + 1];
      z[j] = j < H / 2 ? make_float2(a.x, c.x) : make_float2(c.y, a.y);
    }
  } else {
    const int rule = low_rule(lt);
    const float th = low_value(lt, b);
    const float4* x4 = reinterpret_cast<const float4*>(low + (size_t)b * ls + (size_t)row * L);
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int n = p + R * (2 * j + h);
      if (j < H / 2) {
        const float4 f = x4[n];
        z[j] = shrink2(make_float2(f.x, f.z), th, rule);
      } else {
        const float4 f = x4[N - 1 - n];
        z[j] = shrink2(make_float2(f.w, f.y), th, rule);
      }
    }
  }
  prow_fwd<L>(z, r, p, h, buf, t);
  prow_r2c_post<L>(z, p, h, t);
  // real DCT-II values: X[kk] = Re(e_kk V[kk]), X[L-kk] = -Im(e_kk V[kk]),
  // X[0] = V[0], X[L/2] = V[L/2] / sqrt(2)
  float* drow = D + (size_t)b * ds + (size_t)pl * L * L + (size_t)row * L;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int kk = p + R * (k + H * h);
    if (kk == 0) {
      drow[0] = z[k].x;
      drow[N] = z[k].y * 0.70710678118654752f;
    } else {
      const float2 W = cmul(ctw_mul<4 * L / R, -1>(eh, k), z[k]);
      drow[kk] = W.x;
      drow[L - kk] = -W.y;
    }
  }
}

// k_dct_rows_out with lane-pair row transforms.
template <int L>
__global__ void __launch_bounds__(PRow<L>::T) k_dct_rows_out_p(const float* __restrict__ D, long long ds,
                                                                int nplanes, float* __restrict__ x, long long xs,
                                                                const float2* __restrict__ tw,
                                                                const float2* __restrict__ dtw) {
  using C = PRow<L>;
  constexpr int N = C::N, R = C::R, H = C::H, ROWS = C::ROWS;
  constexpr int NC = L / 4 / C::LANES;  // 16-byte chunks of a row per lane
  static_assert(NC * 4 * C::LANES == L, "row chunking");
  __shared__ float2 buf[C::LDS];
  __shared__ float xrow[ROWS][L];
  const int tid = threadIdx.x;
  const int r = tid / C::LANES, l = tid % C::LANES, p = l >> 1, h = l & 1;
  const int b = blockIdx.z;
  const int row = blockIdx.x * ROWS + r;
  const PRowTw<L> t = prow_tw<L>(p, h, tw);
  float2 eh = dtw[p];
  if (h) eh = cmul(eh, make_float2(0.92387953251128674f, -0.38268343236508977f));
  // the summed row X, three planes per round trip (past the last plane, the
  // last again at weight 0)
  {
    const float4* src = reinterpret_cast<const float4*>(D + (size_t)b * ds + (size_t)row * L);
    float4 acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int q0 = 0; q0 < nplanes; q0 += 3) {
      float4 a[3][NC];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int q = q0 + j < nplanes ? q0 + j : nplanes - 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) a[j][i] = src[(size_t)q * L * L / 4 + l + C::LANES * i];
      }
      __builtin_amdgcn_sched_barrier(0);  // the three planes' loads in flight together
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float wgt = q0 + j < nplanes ? 1.0f : 0.0f;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          acc[i].x += wgt * a[j][i].x;
          acc[i].y += wgt * a[j][i].y;
          acc[i].z += wgt * a[j][i].z;
          acc[i].w += wgt * a[j][i].w;
        }
      }
    }
    float4* xr = reinterpret_cast<float4*>(&xrow[r][0]);
#pragma unroll
    for (int i = 0; i < NC; ++i) xr[l + C::LANES * i] = acc[i];
  }
  __syncthreads();
  // P'[kk] = e^{i t_kk} (X[kk] - i X[L-kk]) (kk >= 1), P'[0] = (X[0], sqrt(2) X[L/2])
  float2 z[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int kk = p + R * (k + H * h);
    const float* X = xrow[r];
    z[k] = kk == 0 ? make_float2(X[0], 1.41421356237309505f * X[N])
                   : cmulc(make_float2(X[kk], -X[L - kk]), ctw_mul<4 * L / R, -1>(eh, k));
  }
  prow_inv<L>(z, r, p, h, buf, t);
  __syncthreads();  // every read of xrow is complete
  // z[k] = L (v[2n], v[2n+1]), n = p + R (k + H h); x[perm(n)] = v[n]
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int n = p + R * (k + H * h);
    xrow[r][2 * n] = z[k].x;
    xrow[r][2 * n + 1] = z[k].y;
  }
  __syncthreads();
  // (x[2m], x[2m+1]) = (v[m], v[L-1-m])
  float2* xo = reinterpret_cast<float2*>(x + (size_t)b * xs + (size_t)row * L);
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int m = p + R * (k + H * h);
    xo[m] = make_float2(xrow[r][m], xrow[r][L - 1 - m]);
  }
}

template <int L>
struct LaunchDctRec {
  static hipError_t rec(const float2* Qg, long long qs, const int* gfirst, int nspec, const float* low, long long ls,
                        const LowThresh& lt, float* D, long long ds, const float* mask, float* x, long long xs,
                        int nimg, const float2* tw, const float2* dtw, hipStream_t s, int step) {
    constexpr int ct = ColTile<L>::narrow ? 4 : 8;
    if (step == 0) {
      SumTable gf{};
      for (int i = 0; i <= nspec; ++i) gf.first[i] = gfirst[i];
      if constexpr (PRow<L>::ok)
        hipLaunchKernelGGL((k_dct_rows_rec_p<L>), dim3(L / PRow<L>::ROWS, nspec + 1, nimg), dim3(PRow<L>::T), 0, s,
                           Qg, qs, gf, nspec, low, ls, lt, D, ds, tw, dtw);
      else
        hipLaunchKernelGGL((k_dct_rows_rec<L>), dim3(L / DctRowCfg<L>::ROWS, nspec + 1, nimg),
                           dim3(DctRowCfg<L>::T), 0, s, Qg, qs, gf, nspec, low, ls, lt, D, ds, tw, dtw);
    } else if (step == 1) {
      if constexpr (L == 512 || L == 1024)
        hipLaunchKernelGGL((k_dct_cols_rec3<L, 4>), dim3(L / 2 / 4, nspec + 1, nimg), dim3(4 * 64), 0, s, D, ds,
                           mask, tw, dtw);
      else
        hipLaunchKernelGGL((k_dct_cols_rec<L, ct>), dim3(L / 2 / ct, nspec + 1, nimg), dim3(ColCfg<L, ct>::T), 0, s, D,
                           ds, mask, tw, dtw);
    } else {
      if constexpr (PRow<L>::ok)
        hipLaunchKernelGGL((k_dct_rows_out_p<L>), dim3(L / PRow<L>::ROWS, 1, nimg), dim3(PRow<L>::T), 0, s, D, ds,
                           nspec + 1, x, xs, tw, dtw);
      else
        hipLaunchKernelGGL((k_dct_rows_out<L>), dim3(L / DctRowCfg<L>::ROWS, 1, nimg), dim3(DctRowCfg<L>::T), 0, s,
                           D, ds, nspec + 1, x, xs, tw, dtw);
    }
    return hipGetLastError();
  }
};

hipError_t launch_dct_rec(int L, int step, const float2* Qg, long long qs, const int* gfirst, int nspec,
                          const float* low, long long ls, const LowThresh& lt, float* D, long long ds,
                          const float* mask, float* x, long long xs, int nimg, const float2* tw, const float2* dtw,
                          hipStream_t s) {
  if (step < 0 || step > 2 || nspec < 1 || nspec > SHX_MAX_SCALES) return hipErrorInvalidValue;
  switch (L) {
    case 64: return LaunchDctRec<64>::rec(Qg, qs, gfirst, nspec, low, ls, lt, D, ds, mask, x, xs, nimg, tw, dtw, s, step);
    case 128: return LaunchDctRec<128>::rec(Qg, qs, gfirst, nspec, low, ls, lt, D, ds, mask, x, xs, nimg, tw, dtw, s, step);
    case 256: return LaunchDctRec<256>::rec(Qg, qs, gfirst, nspec, low, ls, lt, D, ds, mask, x, xs, nimg, tw, dtw, s, step);
    case 512: return LaunchDctRec<512>::rec(Qg, qs, gfirst, nspec, low, ls, lt, D, ds, mask, x, xs, nimg, tw, dtw, s, step);
    case 1024: return LaunchDctRec<1024>::rec(Qg, qs, gfirst, nspec, low, ls, lt, D, ds, mask, x, xs, nimg, tw, dtw, s, step);
    default: return hipErrorInvalidValue;
  }
}

bool col_tile_ok(int L, int ct) {
  // three-pass engines: 512 (8 x 8 x 8) with 4 or 8 columns, 1024
  // (16 x 8 x 8 on 128 threads per column) with 2 or 4
  if (ct < 0) return (L == 512 && (ct == -4 || ct == -8)) || (L == 1024 && (ct == -2 || ct == -4));
  switch (L) {
#define SHX_CT_OK(LL) \
  case LL: return ct == 8 || (ct == 16 && ColTile<LL>::wide) || (ct == 4 && ColTile<LL>::narrow);
    SHX_CT_OK(32) SHX_CT_OK(64) SHX_CT_OK(128) SHX_CT_OK(256) SHX_CT_OK(512) SHX_CT_OK(1024)
#undef SHX_CT_OK
    default: return false;
  }
}

hipError_t launch_cols_rec(int L, const float2* U, long long us, float2* Q, long long qs, const float2* Hm,
                           const float2* Hb, const BandMap& map, int nimg, const float2* tw, hipStream_t s, int ct) {
  SHX_BY_L(L, LT::cols_rec(U, us, Q, qs, Hm, Hb, map, nimg, tw, s, ct))
}

}  // namespace shx
