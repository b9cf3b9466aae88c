// Median absolute deviation of every plane of a stack, on the device.
//
// The noise scalars of the denoise (shearlet_toolbox_1/nsst_scalars.m:28-45,
// demo_denoise/shearcuda/shearcuda.cpp:205-271) are MAD/0.6745 of each band of
// the transform of unit white noise.  The reference copies every band to the
// host and takes two medians there; here one workgroup per plane finds the
// order statistics by radix selection over monotone integer keys of the
// values, so the result is the exact median of the data, as a sort gives it.
//
// Two arithmetic conventions, each reproducing one caller bit for bit:
//   SHX_MAD_FP64  values widened to double: m = (a + b) / 2 of the two middle
//                 values (even count), deviations |v - m| in double, their
//                 median the same way, / 0.6745 (MATLAB's median on the double
//                 transform of nsst_scalars.m:36-44; numpy's median).
//   SHX_MAD_FP32  float arithmetic as in the demo: m = (a + b) / 2 in float,
//                 |v - m| in float, median in float, then / .6745 in double
//                 (shearcuda.cpp:165-181, 246-252).
// Inputs must be finite.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "api_err.h"

namespace shx {
namespace {

constexpr int MT = 1024;                // threads per plane
constexpr int MBITS = 11;               // key bits resolved per histogram pass
constexpr int MBINS = 1 << MBITS;       // == 2 * MT: two bins per thread in the scan
static_assert(MBINS == 2 * MT, "scan assumes two bins per thread");

// Order-preserving key of a float: negative values flipped, sign bit set on
// the others, so that unsigned comparison of keys orders the values.
__device__ __forceinline__ uint32_t f32_key(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float f32_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Keys of the three selections.  Deviations are non-negative, so their bit
// patterns already order them.
struct KeyValue {
  static constexpr int W = 32;
  __device__ uint64_t operator()(float v) const { return f32_key(v); }
};
struct KeyDev64 {
  static constexpr int W = 64;
  double m;
  __device__ uint64_t operator()(float v) const { return (uint64_t)__double_as_longlong(fabs((double)v - m)); }
};
struct KeyDev32 {
  static constexpr int W = 32;
  float m;
  __device__ uint64_t operator()(float v) const { return __float_as_uint(fabsf(v - m)); }
};

template <bool VEC, class K, class F>
__device__ __forceinline__ void each_key(const float* __restrict__ p, int n, const K& kf, F&& fn) {
  int i0 = 0;
  if constexpr (VEC) {
    const float4* p4 = reinterpret_cast<const float4*>(p);
    const int n4 = n >> 2;
    for (int i = threadIdx.x; i < n4; i += MT) {
      const float4 v = p4[i];
      fn(kf(v.x));
      fn(kf(v.y));
      fn(kf(v.z));
      fn(kf(v.w));
    }
    i0 = n4 << 2;
  }
  for (int i = i0 + threadIdx.x; i < n; i += MT) fn(kf(p[i]));
}

struct SelShared {
  uint32_t hist[MBINS];
  uint32_t wsum[MT / 64];
  uint32_t res[3];
  uint64_t wmin[MT / 64];
};

// Key of the element of rank k (0-based) of the plane under kf; *after_eq is
// the number of elements with the same key ranked after it.
template <bool VEC, class K>
__device__ uint64_t block_select(const float* __restrict__ p, int n, uint32_t k, const K& kf, SelShared& sh,
                                 uint32_t* after_eq) {
  const int t = threadIdx.x;
  uint64_t prefix = 0, pmask = 0;
  uint32_t krem = k, cnt = 0;
  for (int left = K::W; left > 0;) {
    const int d = left < MBITS ? left : MBITS;
    const int shift = left - d;
    const uint32_t dmask = (1u << d) - 1u;
    sh.hist[2 * t] = 0;
    sh.hist[2 * t + 1] = 0;
    __syncthreads();
    each_key<VEC>(p, n, kf, [&](uint64_t key) {
      if ((key & pmask) == prefix) atomicAdd(&sh.hist[(uint32_t)(key >> shift) & dmask], 1u);
    });
    __syncthreads();
    // exclusive scan over the bins: thread t owns bins 2t and 2t+1
    const uint32_t h0 = sh.hist[2 * t], h1 = sh.hist[2 * t + 1], s = h0 + h1;
    uint32_t inc = s;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if ((t & 63) >= o) inc += y;
    }
    if ((t & 63) == 63) sh.wsum[t >> 6] = inc;
    __syncthreads();
    uint32_t before = inc - s;
    for (int w = 0; w < (t >> 6); ++w) before += sh.wsum[w];
    if (krem >= before && krem < before + h0) {
      sh.res[0] = 2 * t;
      sh.res[1] = before;
      sh.res[2] = h0;
    } else if (krem >= before + h0 && krem < before + s) {
      sh.res[0] = 2 * t + 1;
      sh.res[1] = before + h0;
      sh.res[2] = h1;
    }
    __syncthreads();
    krem -= sh.res[1];
    cnt = sh.res[2];
    prefix |= (uint64_t)sh.res[0] << shift;
    pmask |= (uint64_t)dmask << shift;
    left = shift;
    __syncthreads();  // hist and res are rewritten by the next pass
  }
  *after_eq = cnt - krem - 1;
  return prefix;
}

// Smallest key above `above` (exists whenever the caller asks for it).
template <bool VEC, class K>
__device__ uint64_t block_min_above(const float* __restrict__ p, int n, uint64_t above, const K& kf, SelShared& sh) {
  const int t = threadIdx.x;
  uint64_t mn = ~0ull;
  each_key<VEC>(p, n, kf, [&](uint64_t key) {
    if (key > above && key < mn) mn = key;
  });
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)mn, o, 64), hi = __shfl_xor((uint32_t)(mn >> 32), o, 64);
    const uint64_t y = ((uint64_t)hi << 32) | lo;
    mn = y < mn ? y : mn;
  }
  if ((t & 63) == 0) sh.wmin[t >> 6] = mn;
  __syncthreads();
  uint64_t r = sh.wmin[0];
  for (int w = 1; w < MT / 64; ++w) r = sh.wmin[w] < r ? sh.wmin[w] : r;
  __syncthreads();
  return r;
}

// The two middle keys (equal for an odd count).
template <bool VEC, class K>
__device__ void block_middle(const float* __restrict__ p, int n, const K& kf, SelShared& sh, uint64_t* lo,
                             uint64_t* hi) {
  uint32_t after = 0;
  if (n & 1) {
    *lo = *hi = block_select<VEC>(p, n, (uint32_t)(n / 2), kf, sh, &after);
    return;
  }
  *lo = block_select<VEC>(p, n, (uint32_t)(n / 2 - 1), kf, sh, &after);
  *hi = after > 0 ? *lo : block_min_above<VEC>(p, n, *lo, kf, sh);
}

template <bool VEC, bool FP32>
__global__ void __launch_bounds__(MT) k_band_mad(const float* __restrict__ planes, long long stride, int n,
                                                 double* __restrict__ mad) {
  __shared__ SelShared sh;
  const float* p = planes + (size_t)blockIdx.x * (size_t)stride;
  uint64_t a, b;
  block_middle<VEC>(p, n, KeyValue{}, sh, &a, &b);
  const float fa = f32_unkey((uint32_t)a), fb = f32_unkey((uint32_t)b);
  double r;
  if constexpr (FP32) {
    const float m = (n & 1) ? fa : (fa + fb) / 2.0f;
    block_middle<VEC>(p, n, KeyDev32{m}, sh, &a, &b);
    const float da = __uint_as_float((uint32_t)a), db = __uint_as_float((uint32_t)b);
    const float md = (n & 1) ? da : (da + db) / 2.0f;
    r = (double)md / 0.6745;
  } else {
    const double m = (n & 1) ? (double)fa : ((double)fa + (double)fb) / 2.0;
    block_middle<VEC>(p, n, KeyDev64{m}, sh, &a, &b);
    const double da = __longlong_as_double((long long)a), db = __longlong_as_double((long long)b);
    const double md = (n & 1) ? da : (da + db) / 2.0;
    r = md / 0.6745;
  }
  if (threadIdx.x == 0) mad[blockIdx.x] = r;
}

}  // namespace
}  // namespace shx

extern "C" int shx_band_mad(const float* planes, long long plane_elems, long long plane_stride, int nplanes,
                            int arith, double* mad, void* stream) {
  using namespace shx;
  if (!planes || !mad) return fail(SHX_ERR_ARG, "null pointer");
  if (plane_elems < 1 || plane_elems > 0x7fffffffLL) return fail(SHX_ERR_ARG, "plane size %lld outside [1, 2^31)", plane_elems);
  if (plane_stride < plane_elems) return fail(SHX_ERR_ARG, "plane stride %lld below the plane size %lld", plane_stride, plane_elems);
  if (nplanes < 0) return fail(SHX_ERR_ARG, "negative plane count");
  if (arith != SHX_MAD_FP64 && arith != SHX_MAD_FP32) return fail(SHX_ERR_ARG, "unknown arithmetic %d", arith);
  if (nplanes == 0) return SHX_OK;
  const int n = (int)plane_elems;
  const bool vec = ((uintptr_t)planes % 16 == 0) && (plane_stride % 4 == 0);
  hipStream_t s = (hipStream_t)stream;
  const dim3 g(nplanes), b(MT);
  if (vec) {
    if (arith == SHX_MAD_FP32)
      hipLaunchKernelGGL((k_band_mad<true, true>), g, b, 0, s, planes, plane_stride, n, mad);
    else
      hipLaunchKernelGGL((k_band_mad<true, false>), g, b, 0, s, planes, plane_stride, n, mad);
  } else {
    if (arith == SHX_MAD_FP32)
      hipLaunchKernelGGL((k_band_mad<false, true>), g, b, 0, s, planes, plane_stride, n, mad);
    else
      hipLaunchKernelGGL((k_band_mad<false, false>), g, b, 0, s, planes, plane_stride, n, mad);
  }
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}
