// 3D line FFT passes and band-pass mask kernels (gfx950).
//
// Replaces, for the 3D path, the reference's per-band cuFFT 3D plans at
// nextpow2(n + 23) with zeroPad / modulateAndNormalize3D / applyHardThreshold /
// addVector (shearcuda3d/3DShearTrans/src_cuda/ShDecThRec_cuda.cpp:228-325,
// shearcuda/shear_cuda.cu:143-195,2284-2386), and the CPU Fourier-domain
// band-pass of shearcuda3d/3DBP/PrySDdec_onestep.m:56-125 and
// PrySDrec_onestep.m:56-118.
//
// Each line transform of N = R1 * R2 points is the two-pass register FFT of
// fft_reg.h: pass A (global -> registers, DFT R2, twiddles), one LDS
// transpose, pass B (DFT R1 -> global).  Zero padding is applied on load and
// the output window on store, so padded forward transforms and cropped
// inverse transforms cost no extra passes.
#include "fft_reg.h"
#include "spectral3d.h"

namespace shx {

__device__ __forceinline__ bool in_window(int j, int n) { return (unsigned)j < (unsigned)n; }

// ---------------------------------------------------------------------------
// Rows (contiguous axis).  A block holds LPB lines and T = LPB * R2 threads;
// pass A uses LPB * R1 of them.
template <int N>
struct Rows3Cfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int LPB = 256 / R2;
  static constexpr int T = LPB * R2;
  static constexpr int LD = N + 1;
  static constexpr int LDS = LPB * LD;
};

template <int N, int DIR>
__global__ void __launch_bounds__(256) k3_rows(Rows3 io, const float2* __restrict__ tw) {
  using C = Rows3Cfg<N>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int line0 = blockIdx.x * C::LPB;
  const float2* in = io.in + (size_t)blockIdx.y * io.in_bs;
  float2* out = io.out + (size_t)blockIdx.y * io.out_bs;
  const int la = tid / R1, n1 = tid % R1;
  if (la < C::LPB && line0 + la < io.nlines) {
    const float2* src = in + (size_t)(line0 + la) * io.in_ld;
    float2 v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int j = n1 + R1 * n2 - io.o_in;
      v[n2] = in_window(j, io.n_in) ? src[j] : make_float2(0.f, 0.f);
    }
    passA<R1, R2, DIR, 1>(v, n1, tw);
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[la * C::LD + n1 + R1 * k] = v[k];
  }
  __syncthreads();
  const int lb = tid / R2, k2 = tid % R2;
  if (line0 + lb < io.nlines) {
    float2 u[R1];
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[lb * C::LD + a + R1 * k2];
    dft<R1, DIR>(u);
    float2* dst = out + (size_t)(line0 + lb) * io.out_ld;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int j = k2 + R2 * k1 - io.o_out;
      if (in_window(j, io.n_out)) dst[j] = u[k1];
    }
  }
}

// Inverse rows of F_g . S for a group of bands; S is loaded once.
template <int N>
__global__ void __launch_bounds__(256) k3_rows_mul(Rows3Mul io, const float2* __restrict__ tw) {
  using C = Rows3Cfg<N>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int line0 = blockIdx.x * C::LPB;
  const int la = tid / R1, n1 = tid % R1;
  const bool vA = la < C::LPB && line0 + la < io.nlines;
  const int lb = tid / R2, k2 = tid % R2;
  const bool vB = line0 + lb < io.nlines;
  const size_t lineA = (size_t)(line0 + la) * N;
  float2 sv[R2];
  if (vA) {
    const float2* S = io.S + (size_t)blockIdx.y * io.s_bs + lineA;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) sv[n2] = cscale(S[n1 + R1 * n2], io.scale);
  }
  float2* outb = io.out + (size_t)blockIdx.y * io.out_bs + (size_t)(line0 + lb) * io.out_ld;
  for (int g = 0; g < io.G; ++g) {
    if (vA) {
      const float2* F = io.F + (size_t)g * io.f_gs + lineA;
      float2 v[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = cmul(sv[n2], F[n1 + R1 * n2]);
      passA<R1, R2, +1, 1>(v, n1, tw);
#pragma unroll
      for (int k = 0; k < R2; ++k) buf[la * C::LD + n1 + R1 * k] = v[k];
    }
    __syncthreads();
    if (vB) {
      float2 u[R1];
#pragma unroll
      for (int a = 0; a < R1; ++a) u[a] = buf[lb * C::LD + a + R1 * k2];
      dft<R1, +1>(u);
      float2* dst = outb + (size_t)g * io.out_gs;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int j = k2 + R2 * k1 - io.o_out;
        if (in_window(j, io.n_out)) dst[j] = u[k1];
      }
    }
    __syncthreads();
  }
}

// Forward rows of G band inputs times conj(F_g), summed over the group.
template <int N>
__global__ void __launch_bounds__(256) k3_rows_cmac(Rows3Cmac io, const float2* __restrict__ tw) {
  using C = Rows3Cfg<N>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int line0 = blockIdx.x * C::LPB;
  const int la = tid / R1, n1 = tid % R1;
  const bool vA = la < C::LPB && line0 + la < io.nlines;
  const int lb = tid / R2, k2 = tid % R2;
  const bool vB = line0 + lb < io.nlines;
  float2 acc[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) acc[k1] = make_float2(0.f, 0.f);
  const size_t lineB = (size_t)(line0 + lb) * N;
  for (int g = 0; g < io.G; ++g) {
    if (vA) {
      const float2* src = io.in + (size_t)blockIdx.y * io.in_bs + (size_t)g * io.in_gs + (size_t)(line0 + la) * io.in_ld;
      float2 v[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int j = n1 + R1 * n2 - io.o_in;
        v[n2] = in_window(j, io.n_in) ? src[j] : make_float2(0.f, 0.f);
      }
      passA<R1, R2, -1, 1>(v, n1, tw);
#pragma unroll
      for (int k = 0; k < R2; ++k) buf[la * C::LD + n1 + R1 * k] = v[k];
    }
    // band spectrum loads in flight across the barrier and the second pass
    float2 f[R1];
    const float2* F = io.F + (size_t)g * io.f_gs + lineB;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) f[k1] = vB ? F[k2 + R2 * k1] : make_float2(0.f, 0.f);
    __syncthreads();
    if (vB) {
      float2 u[R1];
#pragma unroll
      for (int a = 0; a < R1; ++a) u[a] = buf[lb * C::LD + a + R1 * k2];
      dft<R1, -1>(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) acc[k1] = cadd(acc[k1], cmulc(u[k1], f[k1]));
    }
    __syncthreads();
  }
  if (vB) {
    float2* dst = io.out + (size_t)blockIdx.y * io.out_bs + lineB;
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) acc[k1] = cscale(acc[k1], io.scale);
    if (io.accumulate) {  // every old value loaded before the first store
      float2 old[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) old[k1] = dst[k2 + R2 * k1];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) acc[k1] = cadd(old[k1], acc[k1]);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) dst[k2 + R2 * k1] = acc[k1];
  }
}

// ---------------------------------------------------------------------------
// Fused plane inverse: rows then cols of one k0 plane for a chunk of CW
// output columns, through an LDS tile of N rows x CW columns.  Blocks are
// ordered volume-fastest, then band, chunk, plane, so the blocks that share an
// S plane or an F plane run together, and each XCD (the low 3 bits of the
// dispatch order) takes a contiguous range of that order, so the sharing
// stays in one L2.
template <int N>
struct Plane3Cfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int T = N >= 80 ? 512 : 256;
  static constexpr int RC = T / R2;  // rows per row pass
  static constexpr int CC = T / R2;  // columns per column pass
};

template <int N, int CW>
__global__ void __launch_bounds__(Plane3Cfg<N>::T) k3_plane_inv(Plane3Inv io, int nchunk, int nblk, int nvol,
                                                                const float2* __restrict__ tw) {
  using C = Plane3Cfg<N>;
  constexpr int R1 = C::R1, R2 = C::R2, RC = C::RC, CC = C::CC, T = C::T;
  constexpr int CWP = CW + 1;
  constexpr int SCR = (RC * (N + 1) > N * CC) ? RC * (N + 1) : N * CC;
  __shared__ float2 tile[N * CWP];
  __shared__ float2 scr[SCR];
  // XCD-aware block order
  const int pid = blockIdx.x;
  const int per = nblk / 8;
  const int lid = (nblk % 8 == 0) ? (pid % 8) * per + pid / 8 : pid;
  int t = lid;
  const int v = t % nvol;
  t /= nvol;
  const int g = t % io.G;
  t /= io.G;
  const int ch = t % nchunk;
  const int k0 = t / nchunk;
  const int tid = threadIdx.x;
  const int col0 = ch * CW;
  const float2* Sp = io.S + (size_t)v * io.s_bs + (size_t)k0 * N * N;
  const float2* Fp = io.F + (size_t)g * io.f_gs + (size_t)k0 * N * N;
  // rows: inverse along axis 2, keep columns [o_out + col0, o_out + col0 + CW)
  for (int r0 = 0; r0 < N; r0 += RC) {
    const int ra = tid / R1, n1 = tid % R1;
    if (ra < RC && r0 + ra < N) {
      const float2* srow = Sp + (size_t)(r0 + ra) * N;
      const float2* frow = Fp + (size_t)(r0 + ra) * N;
      float2 vv[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) vv[n2] = cscale(cmul(srow[n1 + R1 * n2], frow[n1 + R1 * n2]), io.scale);
      passA<R1, R2, +1, 1>(vv, n1, tw);
#pragma unroll
      for (int k = 0; k < R2; ++k) scr[ra * (N + 1) + n1 + R1 * k] = vv[k];
    }
    __syncthreads();
    const int rb = tid / R2, k2 = tid % R2;
    if (r0 + rb < N) {
      float2 u[R1];
#pragma unroll
      for (int a = 0; a < R1; ++a) u[a] = scr[rb * (N + 1) + a + R1 * k2];
      dft<R1, +1>(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int j = k2 + R2 * k1 - io.o_out - col0;
        if (j >= 0 && j < CW) tile[(r0 + rb) * CWP + j] = u[k1];
      }
    }
    __syncthreads();
  }
  // cols: inverse along axis 1, keep rows [o_out, o_out + n_out)
  const int n = io.n_out;
  float2* op = io.out + (size_t)v * io.out_bs + (size_t)g * io.out_gs + (size_t)k0 * n * n + col0;
  for (int c0 = 0; c0 < CW; c0 += CC) {
    const int c = tid % CC, n1 = tid / CC;
    if (n1 < R1 && c0 + c < CW) {
      float2 vv[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) vv[n2] = tile[(n1 + R1 * n2) * CWP + c0 + c];
      passA<R1, R2, +1, 1>(vv, n1, tw);
#pragma unroll
      for (int k = 0; k < R2; ++k) scr[(n1 + R1 * k) * CC + c] = vv[k];
    }
    __syncthreads();
    const int k2 = tid / CC;
    if (c0 + c < CW && col0 + c0 + c < n) {
      float2 u[R1];
#pragma unroll
      for (int a = 0; a < R1; ++a) u[a] = scr[(a + R1 * k2) * CC + c];
      dft<R1, +1>(u);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int i1 = k2 + R2 * k1 - io.o_out;
        if (in_window(i1, n)) op[(size_t)i1 * n + c0 + c] = u[k1];
      }
    }
    __syncthreads();
  }
  (void)T;
}

// ---------------------------------------------------------------------------
// Cols (axis 1): a tile of CT adjacent columns, T = CT * R2 threads; element
// (m, c) of the tile at LDS slot m * CT + c.
template <int N, int CT>
struct Cols3Cfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int T = CT * R2;
  static constexpr int LDS = N * CT;
};

template <int N, int CT, int DIR>
__global__ void __launch_bounds__(CT * Split<N>::R2) k3_cols(Cols3 io, const float2* __restrict__ tw) {
  using C = Cols3Cfg<N, CT>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int c = tid % CT, n1 = tid / CT;
  const int c0 = blockIdx.x * CT;
  const size_t W = io.W;
  if (n1 < R1) {
    const float2* src = io.in + (size_t)blockIdx.z * io.in_bs + (size_t)blockIdx.y * io.in_ps + c0 + c;
    float2 v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int j = n1 + R1 * n2 - io.o_in;
      v[n2] = in_window(j, io.n_in) ? src[(size_t)j * W] : make_float2(0.f, 0.f);
    }
    passA<R1, R2, DIR, 1>(v, n1, tw);
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[(n1 + R1 * k) * CT + c] = v[k];
  }
  __syncthreads();
  const int k2 = tid / CT;
  float2 u[R1];
#pragma unroll
  for (int a = 0; a < R1; ++a) u[a] = buf[(a + R1 * k2) * CT + c];
  dft<R1, DIR>(u);
  float2* dst = io.out + (size_t)blockIdx.z * io.out_bs + (size_t)blockIdx.y * io.out_ps + c0 + c;
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int j = k2 + R2 * k1 - io.o_out;
    if (in_window(j, io.n_out)) dst[(size_t)j * W] = u[k1];
  }
}

// ---------------------------------------------------------------------------
// Pencils (axis 0): two real columns (2e, 2e+1) form one complex sequence
// z = a + i b; a tile holds CT such pairs.
template <int N, int CT>
__global__ void __launch_bounds__(CT * Split<N>::R2) k3_pencil_fwd(Pencil3Fwd io, const float2* __restrict__ tw) {
  using C = Cols3Cfg<N, CT>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int c = tid % CT, n1 = tid / CT;
  const int e = blockIdx.x * CT + c;  // pair index
  const size_t E = io.E;
  if (n1 < R1) {
    const float* src = io.in + (size_t)blockIdx.z * io.in_bs + (size_t)blockIdx.y * io.in_gs + 2 * (size_t)e;
    float2 v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int j = n1 + R1 * n2 - io.o_in;
      v[n2] = in_window(j, io.n_in) ? *reinterpret_cast<const float2*>(src + (size_t)j * E) : make_float2(0.f, 0.f);
    }
    passA<R1, R2, -1, 1>(v, n1, tw);
#pragma unroll
    for (int k = 0; k < R2; ++k) buf[(n1 + R1 * k) * CT + c] = v[k];
  }
  __syncthreads();
  const int k2 = tid / CT;
  float2 u[R1];
#pragma unroll
  for (int a = 0; a < R1; ++a) u[a] = buf[(a + R1 * k2) * CT + c];
  dft<R1, -1>(u);
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) buf[(k2 + R2 * k1) * CT + c] = u[k1];
  __syncthreads();
  float2* dst = io.out + (size_t)blockIdx.z * io.out_bs + (size_t)blockIdx.y * io.out_gs + 2 * (size_t)e;
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    const int k = k2 + R2 * k1;
    if (k <= N / 2) {
      const float2 z = u[k1];
      const float2 zm = cconj(buf[((N - k) % N) * CT + c]);
      const float2 s = cadd(z, zm), d = csub(z, zm);
      // A = (Z[k] + conj Z[N-k]) / 2, B = (Z[k] - conj Z[N-k]) / (2i)
      *reinterpret_cast<float4*>(dst + (size_t)k * E) = make_float4(0.5f * s.x, 0.5f * s.y, 0.5f * d.y, -0.5f * d.x);
    }
  }
}

template <int N, int CT>
__global__ void __launch_bounds__(CT * Split<N>::R2) k3_pencil_inv(Pencil3Inv io, const float2* __restrict__ tw) {
  using C = Cols3Cfg<N, CT>;
  constexpr int R1 = C::R1, R2 = C::R2;
  __shared__ float2 buf[C::LDS];
  const int tid = threadIdx.x;
  const int c = tid % CT, n1 = tid / CT, k2 = tid / CT;
  const int e = blockIdx.x * CT + c;
  const size_t E = io.E;
  const int b = blockIdx.z;
  float2 acc[R1];
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) acc[k1] = make_float2(0.f, 0.f);
  for (int g = 0; g < io.G; ++g) {
    const float t = io.mode == 1 ? io.th[(size_t)b * io.th_bs + io.th0 + g] : 0.0f;
    if (n1 < R1) {
      const float2* src = io.in + (size_t)b * io.in_bs + (size_t)g * io.in_gs + 2 * (size_t)e;
      // data and twiddle loads all in flight before the arithmetic
      PassTw<R2> pt;
      float2 v[R2];
      float4 ab[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int m = n1 + R1 * n2;
        const int km = m <= N / 2 ? m : N - m;
        ab[n2] = *reinterpret_cast<const float4*>(src + (size_t)km * E);
      }
      passA_tw<R2, +1, 1>(pt, n1, tw);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        const int m = n1 + R1 * n2;
        // Z[m] = A[m] + i B[m]; Z[N-m] = conj(A[m]) + i conj(B[m]); the
        // imaginary parts at m = 0 and N/2 are dropped as in a real inverse.
        if (m == 0 || m == N / 2)
          v[n2] = make_float2(ab[n2].x, ab[n2].z);
        else if (m < N / 2)
          v[n2] = make_float2(ab[n2].x - ab[n2].w, ab[n2].y + ab[n2].z);
        else
          v[n2] = make_float2(ab[n2].x + ab[n2].w, ab[n2].z - ab[n2].y);
      }
      passA_apply<R2, +1>(v, pt);
#pragma unroll
      for (int k = 0; k < R2; ++k) buf[(n1 + R1 * k) * CT + c] = v[k];
    }
    __syncthreads();
    float2 u[R1];
#pragma unroll
    for (int a = 0; a < R1; ++a) u[a] = buf[(a + R1 * k2) * CT + c];
    dft<R1, +1>(u);
    if (io.mode == 0) {
      float* dst = io.out + (size_t)b * io.out_bs + (size_t)g * io.out_gs + 2 * (size_t)e;
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int j = k2 + R2 * k1 - io.o_out;
        if (in_window(j, io.n_out)) *reinterpret_cast<float2*>(dst + (size_t)j * E) = u[k1];
      }
    } else if (io.mode == 1) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1)
        acc[k1] = cadd(acc[k1], make_float2(hard_keep(u[k1].x, t), hard_keep(u[k1].y, t)));
    } else {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) acc[k1] = cadd(acc[k1], u[k1]);
    }
    __syncthreads();
  }
  if (io.mode != 0) {
    float* dst = io.out + (size_t)b * io.out_bs + 2 * (size_t)e;
    if (!io.first) {  // add to out: every old value loaded before the first store
      float2 old[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int j = k2 + R2 * k1 - io.o_out;
        old[k1] = in_window(j, io.n_out) ? *reinterpret_cast<const float2*>(dst + (size_t)j * E)
                                         : make_float2(0.f, 0.f);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) acc[k1] = cadd(old[k1], acc[k1]);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int j = k2 + R2 * k1 - io.o_out;
      if (in_window(j, io.n_out)) *reinterpret_cast<float2*>(dst + (size_t)j * E) = acc[k1];
    }
  }
}

// ---------------------------------------------------------------------------
// Band-pass masks (half axis = memory axis 0).
__device__ __forceinline__ double lf_at(const BpMask& m, int n, int k0, int k1, int k2) {
  const int f1 = k1 <= n - k1 ? k1 : n - k1;
  const int f2 = k2 <= n - k2 ? k2 : n - k2;
  if (k0 > m.P || f1 > m.P || f2 > m.P) return 0.0;
  return m.v[k0] * m.v[f1] * m.v[f2];
}

__global__ void k3_bp_split(const float2* __restrict__ X, long long xs, float2* __restrict__ Hs, long long hs,
                            float2* __restrict__ Lp, long long ls, int n, int M, BpMask m, double hscale,
                            double lscale) {
  const long long HX = (long long)(n / 2 + 1) * n * n;
  const long long HL = (long long)(M / 2 + 1) * M * M;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  X += (size_t)b * xs;
  if (e < HX) {
    const int k2 = (int)(e % n), k1 = (int)((e / n) % n), k0 = (int)(e / ((long long)n * n));
    const double lf = lf_at(m, n, k0, k1, k2);
    const double h = sqrt(1.0 - lf * lf) * hscale;
    const float2 x = X[e];
    Hs[(size_t)b * hs + e] = make_float2((float)(x.x * h), (float)(x.y * h));
  } else if (e < HX + HL) {
    const long long q = e - HX;
    const int q2 = (int)(q % M), q1 = (int)((q / M) % M), q0 = (int)(q / ((long long)M * M));
    const int f1 = q1 <= M / 2 ? q1 : q1 - M, f2 = q2 <= M / 2 ? q2 : q2 - M;
    const int a1 = f1 < 0 ? -f1 : f1, a2 = f2 < 0 ? -f2 : f2;
    float2 r = make_float2(0.f, 0.f);
    if (q0 <= m.P && a1 <= m.P && a2 <= m.P) {
      const double lf = m.v[q0] * m.v[a1] * m.v[a2] * lscale;
      const int s1 = f1 < 0 ? f1 + n : f1, s2 = f2 < 0 ? f2 + n : f2;
      const float2 x = X[((size_t)q0 * n + s1) * n + s2];
      r = make_float2((float)(x.x * lf), (float)(x.y * lf));
    }
    Lp[(size_t)b * ls + q] = r;
  }
}

__global__ void k3_bp_merge(const float2* __restrict__ H, long long hs, const float2* __restrict__ Lp, long long ls,
                            float2* __restrict__ out, long long os, int n, int M, BpMask m, double hscale,
                            double lscale) {
  const long long HX = (long long)(n / 2 + 1) * n * n;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= HX) return;
  const int k2 = (int)(e % n), k1 = (int)((e / n) % n), k0 = (int)(e / ((long long)n * n));
  const double lf = lf_at(m, n, k0, k1, k2);
  const float2 h = H[(size_t)b * hs + e];
  const double hf = sqrt(1.0 - lf * lf) * hscale;
  double re = h.x * hf, im = h.y * hf;
  if (lf != 0.0) {
    const int f1 = k1 <= n / 2 ? k1 : k1 - n, f2 = k2 <= n / 2 ? k2 : k2 - n;
    const int s1 = f1 < 0 ? f1 + M : f1, s2 = f2 < 0 ? f2 + M : f2;
    const float2 l = Lp[(size_t)b * ls + ((size_t)k0 * M + s1) * M + s2];
    const double g = lf * lscale;
    re += l.x * g;
    im += l.y * g;
  }
  out[(size_t)b * os + e] = make_float2((float)re, (float)im);
}

// ---------------------------------------------------------------------------
// Launchers.
template <int N>
struct Launch3 {
  using RC = Rows3Cfg<N>;
  static hipError_t rows(int dir, const Rows3& io, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid((io.nlines + RC::LPB - 1) / RC::LPB, nb);
    if (dir < 0)
      hipLaunchKernelGGL((k3_rows<N, -1>), grid, dim3(RC::T), 0, s, io, tw);
    else
      hipLaunchKernelGGL((k3_rows<N, +1>), grid, dim3(RC::T), 0, s, io, tw);
    return hipGetLastError();
  }
  static hipError_t rows_mul(const Rows3Mul& io, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid((io.nlines + RC::LPB - 1) / RC::LPB, nb);
    hipLaunchKernelGGL((k3_rows_mul<N>), grid, dim3(RC::T), 0, s, io, tw);
    return hipGetLastError();
  }
  template <int CW>
  static hipError_t plane_cw(const Plane3Inv& io, int nb, const float2* tw, hipStream_t s) {
    const int nchunk = (io.n_out + CW - 1) / CW;
    const int nblk = nchunk * io.H * io.G * nb;
    hipLaunchKernelGGL((k3_plane_inv<N, CW>), dim3(nblk), dim3(Plane3Cfg<N>::T), 0, s, io, nchunk, nblk, nb, tw);
    return hipGetLastError();
  }
  static hipError_t plane_inv(const Plane3Inv& io, int nb, int cw, const float2* tw, hipStream_t s) {
    if constexpr (N >= 24) {
      if (cw == 0) cw = io.n_out >= 64 ? 64 : 32;
      if (cw == 64) return plane_cw<64>(io, nb, tw, s);
      if (cw == 32) return plane_cw<32>(io, nb, tw, s);
    }
    return hipErrorInvalidValue;
  }
  static hipError_t rows_cmac(const Rows3Cmac& io, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid((io.nlines + RC::LPB - 1) / RC::LPB, nb);
    hipLaunchKernelGGL((k3_rows_cmac<N>), grid, dim3(RC::T), 0, s, io, tw);
    return hipGetLastError();
  }
  template <int CT>
  static hipError_t cols_ct(int dir, const Cols3& io, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid(io.W / CT, io.P, nb);
    if (dir < 0)
      hipLaunchKernelGGL((k3_cols<N, CT, -1>), grid, dim3(Cols3Cfg<N, CT>::T), 0, s, io, tw);
    else
      hipLaunchKernelGGL((k3_cols<N, CT, +1>), grid, dim3(Cols3Cfg<N, CT>::T), 0, s, io, tw);
    return hipGetLastError();
  }
  static hipError_t cols(int dir, const Cols3& io, int nb, const float2* tw, hipStream_t s) {
    if (io.W % 16 == 0) return cols_ct<16>(dir, io, nb, tw, s);
    if (io.W % 8 == 0) return cols_ct<8>(dir, io, nb, tw, s);
    return hipErrorInvalidValue;
  }
  template <int CT>
  static hipError_t pfwd_ct(const Pencil3Fwd& io, int G, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid(io.E / 2 / CT, G, nb);
    hipLaunchKernelGGL((k3_pencil_fwd<N, CT>), grid, dim3(Cols3Cfg<N, CT>::T), 0, s, io, tw);
    return hipGetLastError();
  }
  static hipError_t pencil_fwd(const Pencil3Fwd& io, int G, int nb, const float2* tw, hipStream_t s) {
    if ((io.E / 2) % 16 == 0) return pfwd_ct<16>(io, G, nb, tw, s);
    if ((io.E / 2) % 8 == 0) return pfwd_ct<8>(io, G, nb, tw, s);
    return hipErrorInvalidValue;
  }
  template <int CT>
  static hipError_t pinv_ct(const Pencil3Inv& io, int nb, const float2* tw, hipStream_t s) {
    const dim3 grid(io.E / 2 / CT, 1, nb);
    hipLaunchKernelGGL((k3_pencil_inv<N, CT>), grid, dim3(Cols3Cfg<N, CT>::T), 0, s, io, tw);
    return hipGetLastError();
  }
  static hipError_t pencil_inv(const Pencil3Inv& io, int nb, const float2* tw, hipStream_t s) {
    if ((io.E / 2) % 16 == 0) return pinv_ct<16>(io, nb, tw, s);
    if ((io.E / 2) % 8 == 0) return pinv_ct<8>(io, nb, tw, s);
    return hipErrorInvalidValue;
  }
};

#define SHX3_BY_N(N, EXPR)                                   \
  switch (N) {                                               \
    case 8: { using LT = Launch3<8>; return EXPR; }          \
    case 16: { using LT = Launch3<16>; return EXPR; }        \
    case 24: { using LT = Launch3<24>; return EXPR; }        \
    case 32: { using LT = Launch3<32>; return EXPR; }        \
    case 48: { using LT = Launch3<48>; return EXPR; }        \
    case 64: { using LT = Launch3<64>; return EXPR; }        \
    case 80: { using LT = Launch3<80>; return EXPR; }        \
    case 128: { using LT = Launch3<128>; return EXPR; }      \
    case 144: { using LT = Launch3<144>; return EXPR; }      \
    default: return hipErrorInvalidValue;                    \
  }

bool fft3_supported(int N) {
  switch (N) {
    case 8: case 16: case 24: case 32: case 48: case 64: case 80: case 128: case 144: return true;
    default: return false;
  }
}

hipError_t launch3_rows(int N, int dir, const Rows3& io, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::rows(dir, io, nb, tw, s))
}
hipError_t launch3_rows_mul(int N, const Rows3Mul& io, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::rows_mul(io, nb, tw, s))
}
hipError_t launch3_rows_cmac(int N, const Rows3Cmac& io, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::rows_cmac(io, nb, tw, s))
}
bool plane3_supported(int N, int n_out) { return N >= 24 && fft3_supported(N) && n_out >= 8; }

hipError_t launch3_plane_inv(int N, const Plane3Inv& io, int nb, int cw, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::plane_inv(io, nb, cw, tw, s))
}

hipError_t launch3_cols(int N, int dir, const Cols3& io, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::cols(dir, io, nb, tw, s))
}
hipError_t launch3_pencil_fwd(int N, const Pencil3Fwd& io, int G, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::pencil_fwd(io, G, nb, tw, s))
}
hipError_t launch3_pencil_inv(int N, const Pencil3Inv& io, int nb, const float2* tw, hipStream_t s) {
  SHX3_BY_N(N, LT::pencil_inv(io, nb, tw, s))
}

hipError_t launch3_bp_split(const float2* X, long long xs, float2* Hs, long long hs, float2* Lp, long long ls, int n,
                            int M, const BpMask& m, double hscale, double lscale, int nb, hipStream_t s) {
  const long long tot = (long long)(n / 2 + 1) * n * n + (long long)(M / 2 + 1) * M * M;
  const dim3 grid((unsigned)((tot + 255) / 256), nb);
  hipLaunchKernelGGL(k3_bp_split, grid, dim3(256), 0, s, X, xs, Hs, hs, Lp, ls, n, M, m, hscale, lscale);
  return hipGetLastError();
}

hipError_t launch3_bp_merge(const float2* H, long long hs, const float2* Lp, long long ls, float2* out,
                            long long os, int n, int M, const BpMask& m, double hscale, double lscale, int nb,
                            hipStream_t s) {
  const long long tot = (long long)(n / 2 + 1) * n * n;
  const dim3 grid((unsigned)((tot + 255) / 256), nb);
  hipLaunchKernelGGL(k3_bp_merge, grid, dim3(256), 0, s, H, hs, Lp, ls, out, os, n, M, m, hscale, lscale);
  return hipGetLastError();
}

}  // namespace shx
