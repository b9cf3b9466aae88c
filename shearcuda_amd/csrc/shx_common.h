// Shared device/host helpers for the shx HIP library (gfx950, FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#ifndef SHX_MAX_SCALES
#define SHX_MAX_SCALES 8          // directional scales per dictionary
#endif
#define SHX_MAX_DIRS 32           // directions per scale (2^5)
#define SHX_MAX_BANDS (SHX_MAX_SCALES * SHX_MAX_DIRS)

namespace shx {

// ---- complex helpers (float2 = re, im) -------------------------------------
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by +i (dir>0) or -i (dir<0)
template <int DIR>
__host__ __device__ __forceinline__ float2 cmul_i(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Hard threshold used everywhere on the path: keep iff |v| > th (strict).
__device__ __forceinline__ float hard_keep(float v, float th) { return fabsf(v) > th ? v : 0.0f; }

// Soft threshold sign(v) max(|v| - th, 0), in the operation order of the
// stand-alone band kernel (so fused and unfused results agree bit for bit).
__device__ __forceinline__ float soft_keep(float v, float th) {
  const float a = fabsf(v) - th;
  return a > 0.f ? copysignf(a, v) : 0.f;
}

// Threshold rule of a fused read: 0 none, 1 hard, 2 soft (wave-uniform).
enum ShrinkRule { SHRINK_NONE = 0, SHRINK_HARD = 1, SHRINK_SOFT = 2 };
__device__ __forceinline__ float shrink(float v, float th, int rule) {
  return rule == SHRINK_SOFT ? soft_keep(v, th) : (rule == SHRINK_HARD ? hard_keep(v, th) : v);
}
__device__ __forceinline__ float2 shrink2(float2 v, float th, int rule) {
  return make_float2(shrink(v.x, th, rule), shrink(v.y, th, rule));
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Raw buffer view of `bytes` bytes from a wave-uniform base address.  An
// access takes a per-lane byte offset (one vector register) and a uniform
// byte offset (a scalar register), so a run of accesses at different uniform
// offsets shares one address register; accesses at or past `bytes` read zero
// and drop writes.  These are vector memory instructions (the descriptor only
// lives in scalar registers).
struct BufView {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ BufView(const void* base, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ float2 ld2(unsigned lane_off, unsigned uni_off) const {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, uni_off, 0));
  }
  __device__ __forceinline__ void st2(float2 v, unsigned lane_off, unsigned uni_off) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                          lane_off, uni_off, 0);
  }
};

}  // namespace shx
