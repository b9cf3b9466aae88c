// C ABI of the shx library (include/shx.h): dictionary management, the NSST
// decompose / threshold / reconstruct entry points, and their helpers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/shx.h"
#include "api_err.h"
#include "dict_dev.h"
#include "dict_host.h"
#include "meyer_dev.h"
#include "pyramid.h"
#include "rdwt.h"
#include "spectral.h"

using shx::BandMap;
using shx::PlaneSet;

// ---------------------------------------------------------------------------
// errors
namespace {
thread_local std::string g_err;
}  // namespace

int shx::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {
using shx::fail;

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Device allocation owned by a scope: freed on every return path.
// Accumulator of the reductions whose result the host reads back before the
// call returns (norms, the maximum of shx_thresh): a file-scope device array,
// so those calls allocate nothing.  A call holds ReduceSlot's lock from the
// clear to the read-back, so host threads take turns.  Host code never passes
// the array's name to a kernel: each device's copy is found with
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) double reduce_slot[1];

struct ReduceSlot {
  std::unique_lock<std::mutex> lock;
  void* p = nullptr;
};

int reduce_slot_acquire(ReduceSlot& rs) {
  static std::mutex addr_mu, use_mu;
  static void* addr[64] = {};
  int dev = -1;
  SHX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(SHX_ERR_UNSUPPORTED, "device %d out of range", dev);
  {
    std::lock_guard<std::mutex> g(addr_mu);
    if (!addr[dev]) {
      void* q = nullptr;
      const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(reduce_slot));
      if (e != hipSuccess || !q)
        return fail(SHX_ERR_HIP, "hipGetSymbolAddress(reduce_slot) on device %d failed: %s", dev,
                    hipGetErrorString(e));
      addr[dev] = q;
    }
    rs.p = addr[dev];
  }
  rs.lock = std::unique_lock<std::mutex>(use_mu);
  return SHX_OK;
}

struct DevMem {
  void* p = nullptr;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};
}  // namespace

// Per-image intermediate buffers of the NSST path.
struct Work {
  float* planes = nullptr;  // per image: S pyramid bands (ypyr) + 2 temporaries + the lowpass of a denoise
                            // whose caller takes no coefficients
  float2* spec = nullptr;   // per image: S packed spectra (Y1 / rec input of the row pass)
  float2* band = nullptr;   // per image: nbands packed spectra (T / U)
  float2* q = nullptr;      // per image: ngroups packed partial reconstructions
  float* ext = nullptr;     // per image: two extended-grid planes (adjoint levels)
  long long s_planes = 0, s_spec = 0, s_band = 0, s_q = 0, s_ext = 0;  // floats per image
  size_t bytes = 0;
  // schedule picked by the autotuner (defaults until shx_autotune runs)
  int pyr_variant = 0;  // pyramid tile shape
  int group = 4;        // directions per reconstruction column block
  int ct_dec = 0;       // column tile width of the decomposition / reconstruction
  int ct_rec = 0;       // column passes (0 = default for L)
  bool dct_front = false;  // decomposition front end in the DCT domain (else pyramid levels + scale FFTs)
  bool dct_back = false;   // reconstruction back end in the DCT domain (else scale row pass + pyramid levels)
  BandMap map{};        // the dictionary's band map with this plan's groups
  // side streams (one per scale) and events of the concurrent denoise: each
  // scale's spectral chain overlaps the pyramid levels
  hipStream_t lanes[SHX_MAX_SCALES] = {};
  hipEvent_t evs[4 * SHX_MAX_SCALES + 4] = {};
  int nlanes = 0;
};

struct shx_dict {
  int L = 0, S = 0, nbands = 0, shift = 0, device = -1;
  int nfilt = 0;  // stored filter spectrum planes (scales with identical filters share theirs)
  int ndir[SHX_MAX_SCALES] = {};
  int dsize[SHX_MAX_SCALES] = {};
  BandMap map{};
  float2* Hmain = nullptr;
  float2* Hbeta = nullptr;
  float2* tw = nullptr;
  float* dmask = nullptr;  // DCT-domain pyramid masks, S + 1 planes (L >= 64)
  float* rmask = nullptr;  // the same for the reconstruction filters
  float2* dtw = nullptr;   // DCT twiddles exp(-i pi k / 2L), k < L
  Work w;  // one image
  size_t bytes = 0;
  size_t plane() const { return (size_t)L * L; }
};

struct shx_batch {
  shx_dict* d = nullptr;
  int nimg = 0;
  Work w;
};

struct shx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t own = nullptr;
};

// ---------------------------------------------------------------------------
// small elementwise / reduction kernels
namespace {

struct ThreshTable {
  int n;
  float* ptr[64];
  long long len[64];
  float th[64];
  float invE[64];
};

// Sum (or max) over a 256-thread block, returned to every thread: one wave
// shuffle tree, then the four wave results through LDS.  Reductions below
// combine one value per block with an atomic, not one per wave, so a call's
// atomics to its single accumulator stay few.
template <class T, class Op>
__device__ __forceinline__ T block_reduce256(T v, Op op) {
  __shared__ T part[4];
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  return op(op(part[0], part[1]), op(part[2], part[3]));
}

__global__ void __launch_bounds__(256) k_hard_thresh(ThreshTable t, int opt, unsigned int* maxbits) {
  const int e = blockIdx.y;
  float* p = t.ptr[e];
  const long long n = t.len[e];
  const float th = t.th[e];
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = shx::hard_keep(p[i], th);
    p[i] = v;
    m = fmaxf(m, fabsf(v));
  }
  if (opt) {
    m = block_reduce256(m * t.invE[e], [](float a, float b) { return fmaxf(a, b); });
    if (threadIdx.x == 0) atomicMax(maxbits, __float_as_uint(m));
  }
}

// Residual step of the separation loop: r = (x - a) - b, ra = a + r, rb = b + r.
__global__ void k_sep_residual(const float* __restrict__ x, const float* __restrict__ a, const float* __restrict__ b,
                               float* __restrict__ r, float* __restrict__ ra, float* __restrict__ rb, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = (x[i] - a[i]) - b[i];
    r[i] = v;
    ra[i] = a[i] + v;
    rb[i] = b[i] + v;
  }
}

// Sum of squares of each plane (one block per plane, FP64 accumulation):
// the band norms of com_norm.m in one launch, in a fixed order.
__global__ void __launch_bounds__(256) k_plane_sumsq(const float* __restrict__ planes, long long pl,
                                                     double* __restrict__ out) {
  const float* p = planes + (size_t)blockIdx.x * pl;
  double s = 0.0;
  for (long long i = threadIdx.x; i < pl; i += 256) s += (double)p[i] * (double)p[i];
  const double d = block_reduce256(s, [](double u, double v) { return u + v; });
  if (threadIdx.x == 0) out[blockIdx.x] = d;
}

// The finest diagonal band of a one-level undecimated transform (yh: L rows
// of [lh hl hh], 3L columns) at even rows and columns: the decimated band.
__global__ void k_even_hh(const float* __restrict__ yh, int L, float* __restrict__ out) {
  const int h = L / 2;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < h * h; i += gridDim.x * blockDim.x) {
    const int r = i / h, c = i - r * h;
    out[i] = yh[(size_t)(2 * r) * 3 * L + 2 * L + 2 * c];
  }
}

__global__ void k_softth(float* x, long long n, float t) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    const float a = fabsf(v) - t;
    x[i] = a > 0.f ? copysignf(a, v) : 0.f;
  }
}

__global__ void __launch_bounds__(256) k_norm_pow(const float* x, long long n, float p, double* out) {
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float a = fabsf(x[i]);
    s += (p == 2.f) ? a * a : (p == 1.f ? a : powf(a, p));
  }
  const double d = block_reduce256((double)s, [](double u, double v) { return u + v; });
  if (threadIdx.x == 0) atomicAdd(out, d);
}

__global__ void __launch_bounds__(256) k_norm_pow_diff(const float* a, const float* b, long long n, float p,
                                                       double* out) {
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = fabsf(a[i] - b[i]);
    s += (p == 2.f) ? v * v : (p == 1.f ? v : powf(v, p));
  }
  const double d = block_reduce256((double)s, [](double u, double v) { return u + v; });
  if (threadIdx.x == 0) atomicAdd(out, d);
}

// Counter-based Gaussian noise: element i draws two uniforms from a
// splitmix64 hash of (seed, i) and uses the Box-Muller cosine branch.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_add_noise(float* x, long long n, float sigma, unsigned long long seed) {
  const unsigned long long key = mix64(seed);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long h = mix64(key ^ (unsigned long long)i);
    const float u1 = ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);  // (0, 1]
    const float u2 = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);  // [0, 1)
    x[i] += sigma * sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
  }
}

// Image type conversion of the demo path: float -> uint8 saturates to
// [0, 255] and truncates toward zero; uint8 -> float is exact.
__global__ void k_f32_to_u8(const float* src, unsigned char* dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = src[i];
    dst[i] = v > 255.f ? (unsigned char)255 : (v < 0.f ? (unsigned char)0 : (unsigned char)v);
  }
}

__global__ void k_u8_to_f32(const unsigned char* src, float* dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// Per-band soft threshold over batched coefficient layouts: element e of
// image i has band band_off + ((e mod rowlen) / seglen) (planes: rowlen =
// image elements, seglen = plane; wavelet yh rows: rowlen = 3nL, seglen = n).
struct BandLayout {
  long long img_elems, rowlen, seglen;
  int band_off;
};

__device__ __forceinline__ int band_of(const BandLayout& b, long long e) {
  const long long r = e % b.img_elems;
  return b.band_off + (int)((r % b.rowlen) / b.seglen);
}

__global__ void k_soft_bands(const float* __restrict__ x, float* __restrict__ y, long long n, BandLayout bl,
                             const float* __restrict__ th) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const float v = x[e];
    const float a = fabsf(v) - th[band_of(bl, e)];
    y[e] = a > 0.f ? copysignf(a, v) : 0.f;
  }
}

// gx = gy * 1[|x| > t];  gth[band] += sum -sign(x) * 1[|x| > t] * gy.
// A block owns a contiguous chunk; the threshold gradient is summed per band in
// LDS (one LDS atomic per wave when the wave's lanes share a band) and flushed
// with one global atomic per band touched by the block.
__global__ void __launch_bounds__(256) k_soft_bands_bwd(const float* __restrict__ x, const float* __restrict__ gy,
                                                        float* __restrict__ gx, float* __restrict__ gth, long long n,
                                                        BandLayout bl, const float* __restrict__ th, long long chunk,
                                                        int nbands) {
  __shared__ float acc[SHX_MAX_BANDS];
  for (int b = threadIdx.x; b < nbands; b += blockDim.x) acc[b] = 0.f;
  __syncthreads();
  const long long start = (long long)blockIdx.x * chunk;
  const long long end = start + chunk < n ? start + chunk : n;
  for (long long e0 = start; e0 < end; e0 += blockDim.x) {
    const long long e = e0 + threadIdx.x;
    int band = -1;
    float c = 0.f;
    if (e < end) {
      band = band_of(bl, e);
      const float v = x[e], g = gy[e];
      const bool keep = fabsf(v) > th[band];
      if (gx) gx[e] = keep ? g : 0.f;
      c = keep ? (v > 0.f ? -g : g) : 0.f;
    }
    const int b0 = __shfl(band, 0);
    if (__all(band == b0 || band < 0)) {
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
      if ((threadIdx.x & 63) == 0 && b0 >= 0) atomicAdd(&acc[b0], c);
    } else if (band >= 0 && c != 0.f) {
      atomicAdd(&acc[band], c);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbands; b += blockDim.x)
    if (acc[b] != 0.f) atomicAdd(&gth[b], acc[b]);
}

// Lowpass part of the shrink backward, in place on g: g = g * 1[|c| > t] and
// gth[img * stride] += sum -sign(c) 1[|c| > t] g over image img's plane of n
// elements (c at img * cs, g at img * gs).  grid (blocks, nimg).
__global__ void __launch_bounds__(256) k_low_mask_bwd(const float* __restrict__ c, long long cs, float* __restrict__ g,
                                                      long long gs, long long n, const float* __restrict__ th,
                                                      int stride, float* __restrict__ gth) {
  const int img = blockIdx.y;
  c += (size_t)img * cs;
  g += (size_t)img * gs;
  const float t = th[(size_t)img * stride];
  float acc = 0.f;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const float v = c[e], gv = g[e];
    const bool keep = fabsf(v) > t;
    g[e] = keep ? gv : 0.f;
    acc += keep ? (v > 0.f ? -gv : gv) : 0.f;
  }
  acc = block_reduce256(acc, [](float a, float b) { return a + b; });
  if (threadIdx.x == 0 && acc != 0.f) atomicAdd(&gth[(size_t)img * stride], acc);
}

// Per-band hard threshold y = x * 1[|x| > t] and its input gradient
// gx = gy * 1[|x| > t] (the threshold has no gradient almost everywhere).
__global__ void k_hard_bands(const float* __restrict__ x, const float* __restrict__ gy, float* __restrict__ y,
                             long long n, BandLayout bl, const float* __restrict__ th) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const bool keep = fabsf(x[e]) > th[band_of(bl, e)];
    y[e] = keep ? (gy ? gy[e] : x[e]) : 0.f;
  }
}

hipError_t pyr_dec(const float* x, int L, int nlev, float* const* y, float* tA, float* tB, hipStream_t s) {
  const float* in = x;
  for (int lev = 0; lev < nlev; ++lev) {
    float* low = (lev == nlev - 1) ? y[0] : ((lev & 1) ? tB : tA);
    hipError_t e = shx::launch_atrous_dec(in, low, y[nlev - lev], shx::ImgStrides{0, 0, 0}, 1, L, 1 << lev, s);
    if (e != hipSuccess) return e;
    in = low;
  }
  return hipSuccess;
}

// w: optional per-band weights (nlev+1)
hipError_t pyr_rec(const float* const* y, int L, int nlev, float* out, float* tA, float* tB, const double* w,
                   hipStream_t s) {
  const float* xl = y[0];
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int j = nlev - lev;
    float* o = (lev == 0) ? out : ((lev & 1) ? tA : tB);
    const bool first = (lev == nlev - 1);
    const float sx = (first && w) ? (float)w[0] : 1.0f;
    const float sy = w ? (float)w[j] : 1.0f;
    hipError_t e = shx::launch_atrous_rec(xl, y[j], o, shx::ImgStrides{0, 0, 0}, 1, L, 1 << lev,
                                          shx::LowThresh{0, 0.f, nullptr, 0, 0}, sx, sy, s);
    if (e != hipSuccess) return e;
    xl = o;
  }
  return hipSuccess;
}

int check_dict(const shx_dict* d) {
  if (!d) return fail(SHX_ERR_ARG, "null dictionary");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != d->device)
    return fail(SHX_ERR_ARG, "dictionary belongs to device %d, current device is %d", d->device, dev);
  return SHX_OK;
}

// largest number of direction groups any schedule uses (groups of 2)
int max_groups(const shx_dict* d) {
  int g = 0;
  for (int j = 0; j < d->S; ++j) g += (d->ndir[j] + 1) / 2;
  return g;
}

void set_group(const shx_dict* d, Work& w, int g) {
  w.group = g;
  w.map = d->map;
  shx::band_map_groups(w.map, g);
}

int alloc_work(const shx_dict* d, int nimg, Work& w) {
  const long long pl = (long long)d->L * d->L;
  set_group(d, w, nimg == 1 ? 4 : 4);
  w.s_planes = pl * (d->S + 3);
  w.s_spec = pl * d->S;
  w.s_band = pl * d->nbands;
  w.s_q = pl * max_groups(d);
  {
    const long long E = shx::adjoint_ext_side(d->L, 1 << (d->S - 1));
    w.s_ext = 2 * E * E;
  }
  const size_t n = (size_t)nimg;
  hipError_t e;
  if ((e = hipMalloc(&w.planes, sizeof(float) * w.s_planes * n)) != hipSuccess ||
      (e = hipMalloc(&w.spec, sizeof(float) * w.s_spec * n)) != hipSuccess ||
      (e = hipMalloc(&w.band, sizeof(float) * w.s_band * n)) != hipSuccess ||
      (e = hipMalloc(&w.q, sizeof(float) * w.s_q * n)) != hipSuccess ||
      (e = hipMalloc(&w.ext, sizeof(float) * w.s_ext * n)) != hipSuccess)
    return fail(SHX_ERR_ALLOC, "workspace allocation (%d images of %d^2) failed: %s", nimg, d->L, hipGetErrorString(e));
  w.bytes = sizeof(float) * n * (w.s_planes + w.s_spec + w.s_band + w.s_q + w.s_ext);
  for (int j = 0; j < d->S; ++j)
    if ((e = hipStreamCreateWithFlags(&w.lanes[j], hipStreamNonBlocking)) != hipSuccess)
      return fail(SHX_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
  w.nlanes = d->S;
  for (hipEvent_t& ev : w.evs)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return fail(SHX_ERR_HIP, "event creation failed: %s", hipGetErrorString(e));
  return SHX_OK;
}

void free_work(Work& w) {
  for (int j = 0; j < w.nlanes; ++j)
    if (w.lanes[j]) hipStreamDestroy(w.lanes[j]);
  for (hipEvent_t& ev : w.evs)
    if (ev) hipEventDestroy(ev);
  for (hipEvent_t& ev : w.evs) ev = nullptr;
  w.nlanes = 0;
  hipFree(w.planes);
  hipFree(w.spec);
  hipFree(w.band);
  hipFree(w.q);
  hipFree(w.ext);
  w.planes = nullptr;
  w.ext = nullptr;
  w.spec = w.band = w.q = nullptr;
  w.bytes = 0;
}

// Coefficients of nimg images: lowpass plane (+ per-image stride) and bands.
struct Coeffs {
  float* low;
  long long low_stride;
  PlaneSet bands;
};

Coeffs cell_coeffs(const shx_dict* d, float* const* coeff) {
  Coeffs c{};
  c.low = coeff[0];
  c.low_stride = 0;
  c.bands.ncells = d->S;
  for (int j = 0; j <= d->S; ++j) c.bands.off[j] = d->map.first[j];
  for (int j = 0; j < d->S; ++j) c.bands.cell[j] = coeff[j + 1];
  c.bands.img_stride = 0;
  return c;
}

// contiguous stacks: image i at base + i*(1+nbands)*L*L, lowpass first
// A denoise whose caller takes no coefficients: the lowpass (which the
// reconstruction reads) lives in the workspace and the band coefficients are
// not written (empty band set).
Coeffs work_coeffs(const shx_dict* d, const Work& w) {
  Coeffs c{};
  c.low = w.planes + (size_t)(d->S + 2) * d->plane();
  c.low_stride = w.s_planes;
  c.bands.ncells = 0;
  return c;
}

Coeffs stack_coeffs(const shx_dict* d, float* base) {
  const long long pl = (long long)d->L * d->L, st = pl * (1 + d->nbands);
  Coeffs c{};
  c.low = base;
  c.low_stride = st;
  c.bands = shx::single_cell(base + pl, d->nbands, st);
  return c;
}

// One kernel launch of a path, with the bytes it must move at minimum
// (FP32 inputs + outputs it touches, each once), for the profiler.
struct Stage {
  std::string name;
  double bytes;
  std::function<hipError_t()> run;
  // concurrent plans (run_stage_dag): the stream run() launches on, earlier
  // stages to wait for, and whether later stages wait for this one
  hipStream_t stream = nullptr;
  std::vector<int> waits;
  bool record = false;
};

// Thresholds of a reconstruction: none, host values (one image, kernel
// arguments) or a device table [nimg][1 + nbands] (lowpass first).
// The device table's row stride is 1 + nbands unless `stride` says otherwise
// (0: one row for every image); soft selects sign(c) max(|c| - t, 0) over the
// hard keep rule.
struct Thr {
  int mode = 0;
  const float* host = nullptr;
  const float* dev = nullptr;
  int stride = -1;
  bool soft = false;
};

int thr_stride(const shx_dict* d, const Thr& t) { return t.stride >= 0 ? t.stride : 1 + d->nbands; }

void stages_pyr_dec(shx_dict* d, const Work& w, int nimg, const float* x, long long xs, const Coeffs& c,
                    hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  const float* in = x;
  long long ins = xs;
  for (int lev = 0; lev < S; ++lev) {
    const bool last = lev == S - 1;
    float* low = last ? c.low : w.planes + (size_t)(S + (lev & 1)) * pl;
    const long long lows = last ? c.low_stride : w.s_planes;
    float* high = w.planes + (size_t)(S - lev - 1) * pl;
    const int m = 1 << lev;
    shx::ImgStrides is{ins, lows, w.s_planes};
    st.push_back({"atrous_dec_m" + std::to_string(m), 3 * P,
                  [=] { return shx::launch_atrous_dec(in, low, high, is, nimg, L, m, s, w.pyr_variant); }});
    in = low;
    ins = lows;
  }
}

// Scale spectra (row FFT, column FFT of every pyramid band-pass plane).
void stages_scale_fft(shx_dict* d, const Work& w, int nimg, hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const double P = 4.0 * L * L * nimg;
  PlaneSet yset = shx::single_cell(w.planes, S, w.s_planes);
  PlaneSet specset = shx::single_cell(reinterpret_cast<float*>(w.spec), S, w.s_spec);
  shx::ThreshSpec none{};
  st.push_back({"rows_r2c_scales", 2 * S * P,
                [=] { return shx::launch_rows_r2c(L, yset, specset, S, nimg, none, d->tw, s); }});
  st.push_back({"cols_fwd_scales", 2 * S * P,
                [=] { return shx::launch_cols_fwd(L, w.spec, w.s_spec / 2, S, nimg, d->tw, s, w.ct_dec); }});
}

void stage_cols_dec(shx_dict* d, const Work& w, int nimg, hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  st.push_back({"cols_dec", (S + 2.0 * nb) * P, [=] {
                  return shx::launch_cols_dec(L, w.spec, w.s_spec / 2, w.band, w.s_band / 2, d->Hmain, d->Hbeta,
                                              w.map, nimg, d->tw, s, w.ct_dec, true);
                }});
}

void stages_spec_dec(shx_dict* d, const Work& w, int nimg, hipStream_t s, std::vector<Stage>& st) {
  stages_scale_fft(d, w, nimg, s, st);
  stage_cols_dec(d, w, nimg, s, st);
}

// DCT-domain front end: the lowpass (c.low) and the scale spectra (w.spec)
// of the whole pyramid in three passes; the image's column DCT goes through
// the first plane of w.planes.
// mask: the decomposition masks (NULL), or the reconstruction masks for the
// transpose of the DCT-domain reconstruction (S^T).
void stages_dct_front(shx_dict* d, const Work& w, int nimg, const float* x, long long xs, const Coeffs& c,
                      hipStream_t s, std::vector<Stage>& st, const float* mask = nullptr) {
  const int L = d->L, S = d->S;
  const double P = 4.0 * L * L * nimg;
  const long long ss = w.s_spec / 2;
  if (!mask) mask = d->dmask;
  st.push_back({"dct_cols", 2 * P, [=] {
                  return shx::launch_dct_dec(L, 0, x, xs, w.planes, w.s_planes, w.spec, ss, S, c.low, c.low_stride,
                                             mask, nimg, d->tw, d->dtw, s);
                }});
  // reads the column DCT and the S + 1 masks, writes S spectra and the lowpass
  st.push_back({"dct_rows_dec", (2.0 * S + 3) * P, [=] {
                  return shx::launch_dct_dec(L, 1, x, xs, w.planes, w.s_planes, w.spec, ss, S, c.low, c.low_stride,
                                             mask, nimg, d->tw, d->dtw, s);
                }});
  st.push_back({"dct_cols_dec", 2.0 * (S + 1) * P, [=] {
                  return shx::launch_dct_dec(L, 2, x, xs, w.planes, w.s_planes, w.spec, ss, S, c.low, c.low_stride,
                                             mask, nimg, d->tw, d->dtw, s);
                }});
}

// Decomposition up to the band spectra T (w.band): either front end, then the
// column filter pass.
void stages_front_dec(shx_dict* d, const Work& w, int nimg, const float* x, long long xs, const Coeffs& c,
                      hipStream_t s, std::vector<Stage>& st) {
  if (w.dct_front) {
    stages_dct_front(d, w, nimg, x, xs, c, s, st);
  } else {
    stages_pyr_dec(d, w, nimg, x, xs, c, s, st);
    stages_scale_fft(d, w, nimg, s, st);
  }
  stage_cols_dec(d, w, nimg, s, st);
}

shx::ThreshSpec band_thresh(const shx_dict* d, const Thr& t) {
  shx::ThreshSpec th{};
  th.mode = t.mode;
  th.soft = t.soft ? 1 : 0;
  if (t.mode == 1)
    for (int b = 0; b < d->nbands; ++b) th.arg[b] = t.host[1 + b];
  if (t.mode == 2) {
    th.dev = t.dev;
    th.dev_stride = thr_stride(d, t);
    th.dev_off = 1;
  }
  return th;
}

// Reconstruction column pass (direction-group partials in w.q) and, unless
// the DCT-domain back end takes the partials itself, the scale row pass
// (scale planes y_j in w.planes).
void stages_spec_rec_tail(shx_dict* d, const Work& w, int nimg, hipStream_t s, std::vector<Stage>& st,
                          bool scale_rows = true) {
  const int L = d->L, S = d->S, nb = d->nbands, ng = w.map.ngroups;
  const double P = 4.0 * L * L * nimg;
  st.push_back({"cols_rec", (2.0 * nb + ng) * P, [=] {
                  return shx::launch_cols_rec(L, w.band, w.s_band / 2, w.q, w.s_q / 2, d->Hmain, d->Hbeta, w.map,
                                              nimg, d->tw, s, w.ct_rec);
                }});
  if (!scale_rows) return;
  PlaneSet qset = shx::single_cell(reinterpret_cast<float*>(w.q), ng, w.s_q);
  PlaneSet yset = shx::single_cell(w.planes, S, w.s_planes);
  std::vector<int> sf(w.map.group_first, w.map.group_first + S + 1);
  st.push_back({"rows_c2r_scales", (ng + S) * P,
                [=] { return shx::launch_rows_c2r(L, qset, yset, S, nimg, sf.data(), d->tw, s); }});
}

// The lowpass threshold of a reconstruction (lowpass first in the tables).
shx::LowThresh low_thresh(const shx_dict* d, const Thr& t) {
  const int soft = t.soft ? 1 : 0;
  shx::LowThresh lt{0, 0.f, nullptr, 0, 0};
  if (t.mode == 1) lt = shx::LowThresh{1, t.host[0], nullptr, 0, soft};
  if (t.mode == 2) lt = shx::LowThresh{2, 0.f, t.dev, thr_stride(d, t), soft};
  return lt;
}

void stages_pyr_rec(shx_dict* d, const Work& w, int nimg, const Coeffs& c, const Thr& t, float* x, long long xs,
                    hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  const shx::LowThresh lt = low_thresh(d, t);
  const float* xl = c.low;
  long long xls = c.low_stride;
  for (int lev = S - 1; lev >= 0; --lev) {
    const int j = S - lev, m = 1 << lev;
    const bool last = lev == 0;
    float* o = last ? x : w.planes + (size_t)(S + (lev & 1)) * pl;
    const long long os = last ? xs : w.s_planes;
    const float* yh = w.planes + (size_t)(j - 1) * pl;
    const shx::LowThresh tl = (lev == S - 1) ? lt : shx::LowThresh{0, 0.f, nullptr, 0, 0};
    shx::ImgStrides is{xls, w.s_planes, os};
    st.push_back({"atrous_rec_m" + std::to_string(m), 3 * P, [=] {
                    return shx::launch_atrous_rec(xl, yh, o, is, nimg, L, m, tl, 1.0f, 1.0f, s, w.pyr_variant);
                  }});
    xl = o;
    xls = os;
  }
}

// DCT-domain back end: from the direction-group partials (w.q) and the
// lowpass to the image in three passes (the scale planes' row DCTs go through
// w.planes' first S + 1 planes).
// mask: the reconstruction masks (NULL), or the decomposition masks for the
// transpose of the DCT-domain decomposition (A^T).
void stages_dct_back(shx_dict* d, const Work& w, int nimg, const Coeffs& c, const Thr& t, float* x, long long xs,
                     hipStream_t s, std::vector<Stage>& st, const float* mask = nullptr) {
  const int L = d->L, S = d->S, ng = w.map.ngroups;
  if (!mask) mask = d->rmask;
  const double P = 4.0 * L * L * nimg;
  const shx::LowThresh lt = low_thresh(d, t);
  std::vector<int> gf(w.map.group_first, w.map.group_first + S + 1);
  const long long qs = w.s_q / 2;
  // reads the partials and the lowpass, writes S + 1 row-transformed planes
  st.push_back({"dct_rows_rec", (ng + 1.0 + (S + 1)) * P, [=] {
                  return shx::launch_dct_rec(L, 0, w.q, qs, gf.data(), S, c.low, c.low_stride, lt, w.planes,
                                             w.s_planes, mask, x, xs, nimg, d->tw, d->dtw, s);
                }});
  // reads and writes every plane, reads its mask
  st.push_back({"dct_cols_rec", 3.0 * (S + 1) * P, [=] {
                  return shx::launch_dct_rec(L, 1, w.q, qs, gf.data(), S, c.low, c.low_stride, lt, w.planes,
                                             w.s_planes, mask, x, xs, nimg, d->tw, d->dtw, s);
                }});
  st.push_back({"dct_rows_out", (S + 2.0) * P, [=] {
                  return shx::launch_dct_rec(L, 2, w.q, qs, gf.data(), S, c.low, c.low_stride, lt, w.planes,
                                             w.s_planes, mask, x, xs, nimg, d->tw, d->dtw, s);
                }});
}

// From the band spectra U (w.band) to the image: the reconstruction column
// pass, then either back end.
void stages_back_rec(shx_dict* d, const Work& w, int nimg, const Coeffs& c, const Thr& t, float* x, long long xs,
                     hipStream_t s, std::vector<Stage>& st) {
  if (w.dct_back) {
    stages_spec_rec_tail(d, w, nimg, s, st, false);
    stages_dct_back(d, w, nimg, c, t, x, xs, s, st);
  } else {
    stages_spec_rec_tail(d, w, nimg, s, st);
    stages_pyr_rec(d, w, nimg, c, t, x, xs, s, st);
  }
}

// Forward transform: pyramid, spectral pass, band rows written to c.bands.
void stages_dec(shx_dict* d, const Work& w, int nimg, const float* x, long long xs, const Coeffs& c, hipStream_t s,
                std::vector<Stage>& st) {
  stages_front_dec(d, w, nimg, x, xs, c, s, st);
  const int L = d->L, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  PlaneSet out = c.bands;
  st.push_back({"rows_c2r_bands", 2.0 * nb * P,
                [=] { return shx::launch_rows_c2r(L, bandset, out, nb, nimg, nullptr, d->tw, s); }});
}

// Inverse transform of (optionally thresholded) coefficients.
void stages_rec(shx_dict* d, const Work& w, int nimg, const Coeffs& c, const Thr& t, float* x, long long xs,
                hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  PlaneSet in = c.bands;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  shx::ThreshSpec th = band_thresh(d, t);
  st.push_back({"rows_r2c_bands", 2.0 * nb * P,
                [=] { return shx::launch_rows_r2c(L, in, bandset, nb, nimg, th, d->tw, s); }});
  stages_back_rec(d, w, nimg, c, t, x, xs, s, st);
}

// Denoise: the decomposition's band row pass and the reconstruction's
// thresholded band row pass run as one kernel over each band row.
void stages_denoise(shx_dict* d, const Work& w, int nimg, const float* x, long long xs, const Coeffs& c,
                    const Thr& t, float* out, long long os, hipStream_t s, std::vector<Stage>& st) {
  stages_front_dec(d, w, nimg, x, xs, c, s, st);
  const int L = d->L, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  PlaneSet cb = c.bands;
  shx::ThreshSpec th = band_thresh(d, t);
  // reads T, writes the coefficients (when the caller takes them) and U (in
  // place of T, row by row)
  st.push_back({"rows_c2r_r2c_bands", (cb.ncells ? 3.0 : 2.0) * nb * P,
                [=] { return shx::launch_rows_c2r_r2c(L, bandset, cb, bandset, nb, nimg, th, d->tw, s); }});
  stages_back_rec(d, w, nimg, c, t, out, os, s, st);
}

// Planes [first, first + count) of a set whose range lies in one cell.
PlaneSet sub_planes(const PlaneSet& s, int first, int count, long long plane_elems) {
  int j = 0;
  while (j + 1 < s.ncells && first >= s.off[j + 1]) ++j;
  return shx::single_cell(s.cell[j] + (first - s.off[j]) * plane_elems, count, s.img_stride);
}

// Concurrent denoise (batched form): the pyramid decomposition and reconstruction run on s0;
// scale j's spectral chain (row R2C, column filter, fused row C2R/threshold/
// R2C, column reconstruction, summing row C2R) runs on side stream j as soon
// as the decomposition has produced its band-pass plane, and the
// reconstruction level of scale j waits for that chain.  Same kernels and
// results as stages_denoise; the per-scale launches overlap.
void stages_denoise_dag(shx_dict* d, Work& w, int nimg, const float* x, long long xs, const Coeffs& c, const Thr& t,
                        float* out, long long os, hipStream_t s0, std::vector<Stage>& st) {
  const int L = d->L, S = d->S, NH = L / 2;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  std::vector<int> made(S), done(S);
  if (w.dct_front) {
    // every scale spectrum comes out of the front end's last pass
    std::vector<Stage> tmp;
    stages_dct_front(d, w, nimg, x, xs, c, s0, tmp);
    for (Stage& g : tmp) {
      g.stream = s0;
      st.push_back(g);
    }
    st.back().record = true;
    for (int j = 0; j < S; ++j) made[j] = (int)st.size() - 1;
  } else {
    std::vector<Stage> tmp;
    stages_pyr_dec(d, w, nimg, x, xs, c, s0, tmp);
    for (int lev = 0; lev < S; ++lev) {
      tmp[lev].stream = s0;
      tmp[lev].record = true;
      made[S - 1 - lev] = (int)st.size();
      st.push_back(tmp[lev]);
    }
  }
  const shx::ThreshSpec thall = band_thresh(d, t);
  for (int j = 0; j < S; ++j) {
    hipStream_t sj = w.lanes[j];
    const int f = d->map.first[j], nb = d->ndir[j];
    const int g0 = w.map.group_first[j], ng = w.map.group_first[j + 1] - g0;
    // sub-map of scale j: bands 0..nb-1, its groups with band ranges relative to f
    BandMap m{};
    m.nscales = 1;
    m.nbands = nb;
    m.first[0] = 0;
    m.first[1] = nb;
    m.ngroups = ng;
    m.group_first[0] = 0;
    m.group_first[1] = ng;
    for (int g = 0; g <= ng; ++g) m.group_band0[g] = (unsigned short)(w.map.group_band0[g0 + g] - f);
    for (int p = 0; p < nb; ++p) {
      m.filt_of[p] = d->map.filt_of[f + p];
      m.order[p] = (unsigned char)p;
    }
    for (int g = 0; g < ng; ++g) m.gorder[g] = (unsigned short)g;
    m.share = m.gshare = 1;
    shx::ThreshSpec th = thall;
    if (th.mode == 1)
      for (int p = 0; p < nb; ++p) th.arg[p] = thall.arg[f + p];
    if (th.mode == 2) th.dev_off += f;
    PlaneSet ypl = shx::single_cell(w.planes + (size_t)j * pl, 1, w.s_planes);
    PlaneSet spc = shx::single_cell(reinterpret_cast<float*>(w.spec) + (size_t)j * pl, 1, w.s_spec);
    PlaneSet bnd = shx::single_cell(reinterpret_cast<float*>(w.band) + (size_t)f * pl, nb, w.s_band);
    PlaneSet cb = c.bands.ncells ? sub_planes(c.bands, f, nb, pl) : c.bands;
    PlaneSet qset = shx::single_cell(reinterpret_cast<float*>(w.q) + (size_t)g0 * pl, ng, w.s_q);
    const float2* Y1 = w.spec + (size_t)j * L * NH;
    float2* T = w.band + (size_t)f * L * NH;
    float2* Q = w.q + (size_t)g0 * L * NH;
    const float2* Hm = d->Hmain;  // the sub-map's filt_of indexes the whole table
    const float2* Hb = d->Hbeta;
    shx::ThreshSpec none{};
    const std::string sfx = "_s" + std::to_string(j);
    if (!w.dct_front) {
      Stage a{"rows_r2c" + sfx, 2 * P, [=] { return shx::launch_rows_r2c(L, ypl, spc, 1, nimg, none, d->tw, sj); }};
      a.stream = sj;
      a.waits = {made[j]};
      st.push_back(a);
      float2* Y1w = w.spec + (size_t)j * L * NH;
      Stage cf{"cols_fwd" + sfx, 2 * P,
               [=] { return shx::launch_cols_fwd(L, Y1w, w.s_spec / 2, 1, nimg, d->tw, sj, w.ct_dec); }};
      cf.stream = sj;
      st.push_back(cf);
    }
    Stage b{"cols_dec" + sfx, (1 + 2.0 * nb) * P, [=] {
              return shx::launch_cols_dec(L, Y1, w.s_spec / 2, T, w.s_band / 2, Hm, Hb, m, nimg, d->tw, sj, w.ct_dec,
                                          true);
            }};
    b.stream = sj;
    if (w.dct_front) b.waits = {made[j]};
    st.push_back(b);
    Stage cc{"rows_c2r_r2c" + sfx, (cb.ncells ? 3.0 : 2.0) * nb * P,
             [=] { return shx::launch_rows_c2r_r2c(L, bnd, cb, bnd, nb, nimg, th, d->tw, sj); }};
    cc.stream = sj;
    st.push_back(cc);
    Stage e{"cols_rec" + sfx, (2.0 * nb + ng) * P, [=] {
              return shx::launch_cols_rec(L, T, w.s_band / 2, Q, w.s_q / 2, Hm, Hb, m, nimg, d->tw, sj, w.ct_rec);
            }};
    e.stream = sj;
    if (w.dct_back) {  // the back end takes the partials of every scale at once
      e.record = true;
      done[j] = (int)st.size();
      st.push_back(e);
      continue;
    }
    st.push_back(e);
    const int sf[2] = {0, ng};
    Stage h{"rows_c2r_sum" + sfx, (ng + 1.0) * P,
            [=] { return shx::launch_rows_c2r(L, qset, ypl, 1, nimg, sf, d->tw, sj); }};
    h.stream = sj;
    h.record = true;
    done[j] = (int)st.size();
    st.push_back(h);
  }
  std::vector<Stage> tmp;
  if (w.dct_back) {
    stages_dct_back(d, w, nimg, c, t, out, os, s0, tmp);
    tmp[0].waits = done;
    for (Stage& g : tmp) {
      g.stream = s0;
      st.push_back(g);
    }
  } else {
    stages_pyr_rec(d, w, nimg, c, t, out, os, s0, tmp);
    for (int k = 0; k < S; ++k) {
      tmp[k].stream = s0;
      tmp[k].waits = {done[k]};
      st.push_back(tmp[k]);
    }
  }
}

// Adjoint of the pyramid decomposition: gradients of the lowpass (c.low) and
// of the pyramid bands (workspace ypyr planes) -> gradient of the image gx.
void stages_pyr_dec_adjoint(shx_dict* d, const Work& w, int nimg, const Coeffs& c, float* gx, long long gxs,
                            hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  const float* glow = c.low;
  long long gls = c.low_stride;
  for (int lev = S - 1; lev >= 0; --lev) {
    const int m = 1 << lev;
    const float* gh = w.planes + (size_t)(S - lev - 1) * pl;
    float* o = (lev == 0) ? gx : w.planes + (size_t)(S + (lev & 1)) * pl;
    const long long os = (lev == 0) ? gxs : w.s_planes;
    shx::ImgStrides is{gls, w.s_planes, os};
    st.push_back({"atrous_dec_adj_m" + std::to_string(m), 3 * P, [=] {
                    return shx::launch_atrous_dec_adjoint(glow, gh, o, w.ext, is, w.s_ext, nimg, L, m, s);
                  }});
    glow = o;
    gls = os;
  }
}

// Adjoint of the pyramid reconstruction: gradient of the output image gout ->
// gradients of the lowpass (c.low) and of the pyramid bands (ypyr planes).
void stages_pyr_rec_adjoint(shx_dict* d, const Work& w, int nimg, const float* gout, long long gos, const Coeffs& c,
                            hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  const float* gi = gout;
  long long gis = gos;
  for (int lev = 0; lev < S; ++lev) {
    const int m = 1 << lev;
    const bool last = lev == S - 1;
    float* glow = last ? c.low : w.planes + (size_t)(S + (lev & 1)) * pl;
    const long long gls = last ? c.low_stride : w.s_planes;
    float* gh = w.planes + (size_t)(S - lev - 1) * pl;
    shx::ImgStrides is{gis, gls, w.s_planes};
    st.push_back({"atrous_rec_adj_m" + std::to_string(m), 3 * P, [=] {
                    return shx::launch_atrous_rec_adjoint(gi, glow, gh, w.ext, is, w.s_ext, nimg, L, m, s);
                  }});
    gi = glow;
    gis = gls;
  }
}

// The pyramid part of A^T from the band spectra U (w.band): the directional
// reconstruction's column pass, then either the transposed pyramid levels or,
// when the plan's decomposition runs in the DCT domain, its transpose there.
// A DCT-domain level is x -> D2^-1(M . D2 x) with D2 the raw 2D DCT-II.  With
// C the 1D DCT-II matrix, C C^T = diag(L, L/2, ..., L/2) = Lam, so C^T =
// C^-1 Lam and the diagonal Lam commutes with the diagonal mask: the
// transpose is D2^-1(M . D2 y), the same operator.  So A^T is the DCT-domain
// reconstruction run with the decomposition masks, and S^T the DCT-domain
// decomposition run with the reconstruction masks.
void stages_dec_adjoint_tail(shx_dict* d, const Work& w, int nimg, const Coeffs& c, float* gx, long long gxs,
                             hipStream_t s, std::vector<Stage>& st) {
  if (w.dct_front && d->dmask) {
    stages_spec_rec_tail(d, w, nimg, s, st, false);
    stages_dct_back(d, w, nimg, c, Thr{}, gx, gxs, s, st, d->dmask);
  } else {
    stages_spec_rec_tail(d, w, nimg, s, st);
    stages_pyr_dec_adjoint(d, w, nimg, c, gx, gxs, s, st);
  }
}

// The pyramid part of S^T, up to the band spectra T (w.band).
void stages_rec_adjoint_front(shx_dict* d, const Work& w, int nimg, const float* gx, long long gxs, const Coeffs& c,
                              hipStream_t s, std::vector<Stage>& st) {
  if (w.dct_back && d->rmask) {
    stages_dct_front(d, w, nimg, gx, gxs, c, s, st, d->rmask);
    stage_cols_dec(d, w, nimg, s, st);
  } else {
    stages_pyr_rec_adjoint(d, w, nimg, gx, gxs, c, s, st);
    stages_spec_dec(d, w, nimg, s, st);
  }
}

// A^T: coefficient gradients -> image gradient (directional stage: its
// adjoint is the directional reconstruction, the filters form a tight frame).
void stages_dec_adjoint(shx_dict* d, const Work& w, int nimg, const Coeffs& c, float* gx, long long gxs,
                        hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  PlaneSet in = c.bands;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  shx::ThreshSpec none{};
  st.push_back({"rows_r2c_bands", 2.0 * nb * P,
                [=] { return shx::launch_rows_r2c(L, in, bandset, nb, nimg, none, d->tw, s); }});
  stages_dec_adjoint_tail(d, w, nimg, c, gx, gxs, s, st);
}

// S^T: image gradient -> coefficient gradients.
void stages_rec_adjoint(shx_dict* d, const Work& w, int nimg, const float* gx, long long gxs, const Coeffs& c,
                        hipStream_t s, std::vector<Stage>& st) {
  stages_rec_adjoint_front(d, w, nimg, gx, gxs, c, s, st);
  const int L = d->L, nb = d->nbands;
  const double P = 4.0 * L * L * nimg;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  PlaneSet out = c.bands;
  st.push_back({"rows_c2r_bands", 2.0 * nb * P,
                [=] { return shx::launch_rows_c2r(L, bandset, out, nb, nimg, nullptr, d->tw, s); }});
}

// Backward of the soft-threshold shrink out = S(soft(A x, th)) given the
// forward coefficients c: S^T out_bar to the band spectra, the mask
// 1[|c| > t] and the threshold gradient between S^T's last row pass and A^T's
// first (one fused kernel per band row), the lowpass's mask in place, then
// A^T.  The lowpass gradient lives in workspace plane S + 2.
void stages_shrink_backward(shx_dict* d, const Work& w, int nimg, const Coeffs& c, const Thr& t, const float* gout,
                            long long gos, float* gth, float* gx, long long gxs, hipStream_t s,
                            std::vector<Stage>& st) {
  const int L = d->L, S = d->S, nb = d->nbands;
  const long long pl = (long long)L * L;
  const double P = 4.0 * pl * nimg;
  Coeffs g{};
  g.low = w.planes + (size_t)(S + 2) * pl;
  g.low_stride = w.s_planes;
  stages_rec_adjoint_front(d, w, nimg, gout, gos, g, s, st);
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  PlaneSet cb = c.bands;
  shx::ThreshSpec th = band_thresh(d, t);
  // reads T and the coefficients, writes U in place of T
  st.push_back({"rows_c2r_mask_r2c_bands", 3.0 * nb * P, [=] {
                  return shx::launch_rows_c2r_mask_r2c(L, bandset, cb, bandset, nb, nimg, th, gth, d->tw, s);
                }});
  const float* clow = c.low;
  const long long cls = c.low_stride;
  float* glow = g.low;
  const long long gls = g.low_stride;
  const float* thd = t.dev;
  const int ts = thr_stride(d, t);
  st.push_back({"low_mask", 3.0 * P, [=] {
                  // few blocks per image: each adds its sum to one threshold
                  // gradient with an atomic, and all images may share it
                  const int blocks = (int)std::min<long long>(32, (pl + 255) / 256);
                  hipLaunchKernelGGL(k_low_mask_bwd, dim3(blocks, nimg), dim3(256), 0, s, clow, cls, glow, gls, pl,
                                     thd, ts, gth);
                  return hipGetLastError();
                }});
  stages_dec_adjoint_tail(d, w, nimg, g, gx, gxs, s, st);
}

int run_stages(const std::vector<Stage>& st) {
  for (const Stage& g : st) {
    hipError_t e = g.run();
    if (e != hipSuccess) return fail(SHX_ERR_HIP, "%s launch failed: %s", g.name.c_str(), hipGetErrorString(e));
  }
  return SHX_OK;
}

// Runs stages that carry their own streams: every wait becomes
// hipStreamWaitEvent on an event recorded after the awaited stage (under
// graph capture these are the graph's edges).
int run_stage_dag(const std::vector<Stage>& st, const Work& w) {
  std::vector<int> ev_of(st.size(), -1);
  int nev = 0;
  const int maxev = (int)(sizeof(w.evs) / sizeof(w.evs[0]));
  for (size_t i = 0; i < st.size(); ++i) {
    const Stage& g = st[i];
    for (int j : g.waits) {
      if (j < 0 || j >= (int)i || ev_of[j] < 0) return fail(SHX_ERR_ARG, "stage %s waits for an unrecorded stage", g.name.c_str());
      SHX_HIP(hipStreamWaitEvent(g.stream, w.evs[ev_of[j]], 0));
    }
    hipError_t e = g.run();
    if (e != hipSuccess) return fail(SHX_ERR_HIP, "%s launch failed: %s", g.name.c_str(), hipGetErrorString(e));
    if (g.record) {
      if (nev >= maxev) return fail(SHX_ERR_ARG, "too many concurrent stages");
      ev_of[i] = nev;
      SHX_HIP(hipEventRecord(w.evs[nev++], g.stream));
    }
  }
  return SHX_OK;
}

std::vector<float> nsst_ht_thresholds(const shx_dict* d, double sigma, const double* T, const double* sc) {
  std::vector<float> th(1 + d->nbands);
  th[0] = (float)(T[0] * sigma * sc[0]);
  for (int j = 0; j < d->S; ++j) {
    const double Tj = (j < d->S - 1) ? T[1] : T[2];
    for (int k = 0; k < d->ndir[j]; ++k) {
      const int b = d->map.first[j] + k;
      th[1 + b] = (float)(Tj * sigma * sc[1 + b]);
    }
  }
  return th;
}

int check_cells(const shx_dict* d, const void* const* coeff) {
  if (!coeff) return fail(SHX_ERR_ARG, "null coefficient cells");
  for (int j = 0; j <= d->S; ++j)
    if (!coeff[j]) return fail(SHX_ERR_ARG, "coefficient cell %d is null", j);
  return SHX_OK;
}

int profile_stages(const std::vector<Stage>& st, hipStream_t s, int reps, int max_stages, double* stage_ms,
                   double* stage_bytes, char* names, int name_len, int* nstages) {
  const int n = (int)st.size();
  *nstages = n;
  if (n > max_stages) return fail(SHX_ERR_ARG, "%d stages, room for %d", n, max_stages);
  std::vector<hipEvent_t> ev(n + 1);
  for (auto& e : ev) SHX_HIP(hipEventCreate(&e));
  std::vector<double> acc(n, 0.0);
  int rc = SHX_OK;
  for (int r = 0; r < reps && rc == SHX_OK; ++r) {
    if (hipEventRecord(ev[0], s) != hipSuccess) {
      rc = fail(SHX_ERR_HIP, "event record failed");
      break;
    }
    for (int i = 0; i < n && rc == SHX_OK; ++i) {
      hipError_t e = st[i].run();
      if (e == hipSuccess) e = hipEventRecord(ev[i + 1], s);
      if (e != hipSuccess) rc = fail(SHX_ERR_HIP, "%s: %s", st[i].name.c_str(), hipGetErrorString(e));
    }
    if (rc != SHX_OK) break;
    if (hipEventSynchronize(ev[n]) != hipSuccess) {
      rc = fail(SHX_ERR_HIP, "event sync failed");
      break;
    }
    for (int i = 0; i < n; ++i) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      acc[i] += ms;
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  if (rc != SHX_OK) return rc;
  for (int i = 0; i < n; ++i) {
    stage_ms[i] = acc[i] / reps;
    if (stage_bytes) stage_bytes[i] = st[i].bytes;
    if (names && name_len > 0) {
      strncpy(names + (size_t)i * name_len, st[i].name.c_str(), name_len - 1);
      names[(size_t)i * name_len + name_len - 1] = 0;
    }
  }
  return SHX_OK;
}

// Mean duration (ms) of `reps` replays of a stage list captured as one graph
// on stream cs: the way plans run in practice (no host launch cost between
// kernels), so short kernels are compared by their own durations.
double time_stages_graph(const std::vector<Stage>& st, hipStream_t cs, int reps) {
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1.0;
  const int rc = run_stages(st);
  if (hipStreamEndCapture(cs, &g) != hipSuccess || rc != SHX_OK) {
    if (g) hipGraphDestroy(g);
    return -1.0;
  }
  double ms = -1.0;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess && hipEventCreate(&a) == hipSuccess &&
      hipEventCreate(&b) == hipSuccess && hipGraphLaunch(ex, cs) == hipSuccess &&
      hipEventRecord(a, cs) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; ++r) ok = hipGraphLaunch(ex, cs) == hipSuccess;
    float t = 0.f;
    if (ok && hipEventRecord(b, cs) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
        hipEventElapsedTime(&t, a, b) == hipSuccess)
      ms = t / reps;
  }
  hipStreamSynchronize(cs);
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (ex) hipGraphExecDestroy(ex);
  hipGraphDestroy(g);
  return ms;
}

// Autotune one plan: pyramid tile shape, then reconstruction direction
// groups, each by timing graph replays of the kernels it affects on the
// plan's own workspace and scratch images.  The timing runs on a private
// stream after the caller's stream has drained.
int autotune_work(shx_dict* d, Work& w, int nimg, int reps, hipStream_t caller, int* chosen) {
  const long long pl = (long long)d->L * d->L;
  SHX_HIP(hipStreamSynchronize(caller));
  hipStream_t s = nullptr;
  SHX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      hipStreamSynchronize(s);
      hipStreamDestroy(s);
    }
  } guard{s};
  float* img = nullptr;
  if (hipMalloc(&img, sizeof(float) * 2 * pl * nimg) != hipSuccess)
    return fail(SHX_ERR_ALLOC, "autotune scratch allocation failed");
  hipMemsetAsync(img, 0, sizeof(float) * 2 * pl * nimg, s);
  hipMemsetAsync(w.planes, 0, sizeof(float) * w.s_planes * nimg, s);
  hipMemsetAsync(w.band, 0, sizeof(float) * w.s_band * nimg, s);
  float* coeff = nullptr;
  if (hipMalloc(&coeff, sizeof(float) * pl * (1 + d->nbands) * nimg) != hipSuccess) {
    hipFree(img);
    return fail(SHX_ERR_ALLOC, "autotune scratch allocation failed");
  }
  hipMemsetAsync(coeff, 0, sizeof(float) * pl * (1 + d->nbands) * nimg, s);
  Coeffs c = stack_coeffs(d, coeff);
  // candidates are timed in three interleaved rounds; each keeps its best time
  double tp[shx::PYR_VARIANTS];
  for (double& t : tp) t = 1e30;
  for (int round = 0; round < 3; ++round)
    for (int v = 0; v < shx::PYR_VARIANTS; ++v) {
      w.pyr_variant = v;
      std::vector<Stage> st;
      stages_pyr_dec(d, w, nimg, img, pl, c, s, st);
      stages_pyr_rec(d, w, nimg, c, Thr{}, img + pl * nimg, pl, s, st);
      const double t = time_stages_graph(st, s, reps);
      if (t > 0 && t < tp[v]) tp[v] = t;
    }
  int bestv = 0;
  for (int v = 1; v < shx::PYR_VARIANTS; ++v)
    if (tp[v] < tp[bestv]) bestv = v;
  w.pyr_variant = bestv;
  // reconstruction: directions per block x column tile; decomposition: tile
  // column tiles: two-pass engine 4/8/16 columns, three-pass engines (-8/-4/-2)
  const int gcand[3] = {2, 4, 8};
  int ccand[6], nct = 0;
  for (int ct : {8, 16, 4, -8, -4, -2})
    if (shx::col_tile_ok(d->L, ct)) ccand[nct++] = ct;
  double tg[3][6];
  for (auto& r : tg)
    for (double& t : r) t = 1e30;
  for (int round = 0; round < 3; ++round)
    for (int i = 0; i < 3; ++i)
      for (int c = 0; c < nct; ++c) {
        set_group(d, w, gcand[i]);
        w.ct_rec = ccand[c];
        std::vector<Stage> st;
        stages_spec_rec_tail(d, w, nimg, s, st);
        const double t = time_stages_graph(st, s, reps);
        if (t > 0 && t < tg[i][c]) tg[i][c] = t;
      }
  int bi = 0, bc = 0;
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < nct; ++c)
      if (tg[i][c] < tg[bi][bc]) bi = i, bc = c;
  set_group(d, w, gcand[bi]);
  w.ct_rec = ccand[bc];
  double td[6] = {1e30, 1e30, 1e30, 1e30, 1e30, 1e30};
  for (int round = 0; round < 3; ++round)
    for (int c = 0; c < nct; ++c) {
      w.ct_dec = ccand[c];
      std::vector<Stage> st;
      stages_spec_dec(d, w, nimg, s, st);
      const double t = time_stages_graph(st, s, reps);
      if (t > 0 && t < td[c]) td[c] = t;
    }
  int bd = 0;
  for (int c = 1; c < nct; ++c)
    if (td[c] < td[bd]) bd = c;
  w.ct_dec = ccand[bd];
  // decomposition front end: pyramid levels + scale FFTs, or the DCT domain
  if (d->dmask) {
    double tf[2] = {1e30, 1e30};
    for (int round = 0; round < 3; ++round)
      for (int f = 0; f < 2; ++f) {
        w.dct_front = f == 1;
        std::vector<Stage> st;
        if (w.dct_front) {
          stages_dct_front(d, w, nimg, img, pl, c, s, st);
        } else {
          stages_pyr_dec(d, w, nimg, img, pl, c, s, st);
          stages_scale_fft(d, w, nimg, s, st);
        }
        const double t = time_stages_graph(st, s, reps);
        if (t > 0 && t < tf[f]) tf[f] = t;
      }
    w.dct_front = tf[1] < tf[0];
    // reconstruction back end: scale row pass + pyramid levels, or the DCT domain
    double tb[2] = {1e30, 1e30};
    for (int round = 0; round < 3; ++round)
      for (int f = 0; f < 2; ++f) {
        std::vector<Stage> st;
        if (f == 1) {
          stages_dct_back(d, w, nimg, c, Thr{}, img + pl * nimg, pl, s, st);
        } else {
          const int L = d->L, S = d->S, ng = w.map.ngroups;
          PlaneSet qset = shx::single_cell(reinterpret_cast<float*>(w.q), ng, w.s_q);
          PlaneSet yset = shx::single_cell(w.planes, S, w.s_planes);
          std::vector<int> sf(w.map.group_first, w.map.group_first + S + 1);
          st.push_back({"rows_c2r_scales", 0.0,
                        [=] { return shx::launch_rows_c2r(L, qset, yset, S, nimg, sf.data(), d->tw, s); }});
          stages_pyr_rec(d, w, nimg, c, Thr{}, img + pl * nimg, pl, s, st);
        }
        const double t = time_stages_graph(st, s, reps);
        if (t > 0 && t < tb[f]) tb[f] = t;
      }
    w.dct_back = tb[1] < tb[0];
  }
  hipStreamSynchronize(s);
  hipFree(img);
  hipFree(coeff);
  if (chosen) {
    chosen[0] = w.pyr_variant;
    chosen[1] = w.group;
    chosen[2] = w.ct_dec;
    chosen[3] = w.ct_rec;
  }
  return SHX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* shx_last_error(void) { return g_err.c_str(); }
int shx_version(void) { return 200; }

double shx_maxflat_factor_error(void) { return shx::maxflat_factor_error(); }

int shx_maxflat_filter(int which, double* out, int* len) {
  int n = 0;
  const double* h = shx::maxflat_filter(which, &n);
  if (!h) return fail(SHX_ERR_ARG, "filter index %d out of range", which);
  if (len) *len = n;
  if (out) memcpy(out, h, sizeof(double) * n * n);
  return SHX_OK;
}

int shx_prepare_shearing_filters(int L, int nscales, const int* ndir, const int* dsize, const double* const* w_s,
                                 int shift, shx_dict** out) {
  return shx_prepare_shearing_filters_ex(L, nscales, ndir, dsize, w_s, shift, 0, out);
}

}  // extern "C"

namespace {
// Dictionary from spatial filters in host memory, or in device memory
// (w_on_device: the spectra then come from them without a host round trip,
// except that the host-spectra flag copies them down first).
int prepare_dict(int L, int nscales, const int* ndir, const int* dsize, const double* const* w_s, bool w_on_device,
                 const int* same_as, int shift, int flags, shx_dict** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  if (flags & ~SHX_DICT_HOST_SPECTRA) return fail(SHX_ERR_ARG, "unknown dictionary flags 0x%x", flags);
  const bool host_spectra = (flags & SHX_DICT_HOST_SPECTRA) != 0;
  *out = nullptr;
  if (!is_pow2(L) || L < 32 || L > 1024)
    return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported (power of two in [32, 1024])", L);
  if (nscales < 1 || nscales > SHX_MAX_SCALES)
    return fail(SHX_ERR_ARG, "number of scales %d outside [1, %d]", nscales, SHX_MAX_SCALES);
  if (!ndir || !dsize || !w_s) return fail(SHX_ERR_ARG, "null filter description");
  if (shift != SHX_SHIFT_NONE && shift != SHX_SHIFT_CENTER) return fail(SHX_ERR_ARG, "unknown shift mode %d", shift);
  if (L % (1 << (nscales - 1)) != 0) return fail(SHX_ERR_ARG, "L must be divisible by the coarsest dilation");
  if (!shx::maxflat_ok()) return fail(SHX_ERR_HIP, "maxflat factorisation failed");
  auto d = new shx_dict();
  d->L = L;
  d->S = nscales;
  d->shift = shift;
  int nb = 0;
  for (int j = 0; j < nscales; ++j) {
    if (!is_pow2(ndir[j]) || ndir[j] > SHX_MAX_DIRS) {
      delete d;
      return fail(SHX_ERR_ARG, "scale %d: %d directions (power of two <= %d required)", j, ndir[j], SHX_MAX_DIRS);
    }
    if (dsize[j] < 1 || dsize[j] > L || !w_s[j]) {
      delete d;
      return fail(SHX_ERR_ARG, "scale %d: filter size %d invalid for L=%d", j, dsize[j], L);
    }
    d->ndir[j] = ndir[j];
    d->dsize[j] = dsize[j];
    d->map.first[j] = nb;
    for (int k = 0; k < ndir[j]; ++k) d->map.scale_of[nb + k] = (unsigned char)j;
    nb += ndir[j];
  }
  d->map.first[nscales] = nb;
  d->map.nscales = nscales;
  d->map.nbands = nb;
  // scales with identical filters (same direction count and size, equal
  // values) read one stored set of spectra
  std::vector<int> base(nscales, -1);
  for (int j = 0; j < nscales; ++j) {
    int src = -1;
    for (int i = 0; i < j && src < 0; ++i) {
      if (base[i] < 0 || ndir[i] != ndir[j] || dsize[i] != dsize[j]) continue;
      if (same_as) {
        if (same_as[j] == i) src = i;
      } else if (!w_on_device &&
                 memcmp(w_s[i], w_s[j], sizeof(double) * (size_t)ndir[j] * dsize[j] * dsize[j]) == 0) {
        src = i;
      }
    }
    base[j] = src >= 0 ? base[src] : d->nfilt;
    if (src < 0) d->nfilt += ndir[j];
    for (int k = 0; k < ndir[j]; ++k) d->map.filt_of[d->map.first[j] + k] = (unsigned char)(base[j] + k);
  }
  if (d->nfilt > 255) {
    delete d;
    return fail(SHX_ERR_ARG, "%d distinct filter planes (at most 255)", d->nfilt);
  }
  shx::band_map_groups(d->map, 4);
  // the first scale with a given set of planes computes their spectra
  std::vector<bool> owner(nscales, true);
  for (int j = 0; j < nscales; ++j)
    for (int i = 0; i < j; ++i)
      if (base[i] == base[j]) owner[j] = false;
  d->nbands = nb;
  if (hipGetDevice(&d->device) != hipSuccess) {
    delete d;
    return fail(SHX_ERR_HIP, "no HIP device");
  }
  const int N = L / 2, nf = d->nfilt;
  std::vector<float> hm(host_spectra ? (size_t)nf * L * N * 2 : 0), hb(host_spectra ? (size_t)nf * L * 2 : 0);
  std::vector<float> tw((size_t)L * 2);
  for (int j = 0; host_spectra && j < nscales; ++j) {
    if (!owner[j]) continue;
    std::vector<std::complex<double>> H;
    std::vector<double> wh;
    const double* wj = w_s[j];
    if (w_on_device) {
      wh.resize((size_t)ndir[j] * dsize[j] * dsize[j]);
      if (hipMemcpy(wh.data(), w_s[j], wh.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        delete d;
        return fail(SHX_ERR_HIP, "scale %d: filter download failed", j);
      }
      wj = wh.data();
    }
    shx::scale_spectra_fp64(L, ndir[j], dsize[j], wj, shift, H);
    for (size_t i = 0; i < H.size(); ++i)
      if (!std::isfinite(H[i].real()) || !std::isfinite(H[i].imag())) {
        delete d;
        return fail(SHX_ERR_ARG, "scale %d: filter spectra vanish at some frequency (division by zero)", j);
      }
    shx::pack_spectra_fp32(L, ndir[j], H, hm.data() + (size_t)base[j] * L * N * 2, hb.data() + (size_t)base[j] * L * 2);
  }
  shx::twiddles_fp32(L, tw.data());
  // DCT-domain front end (L >= 64): masks and twiddles
  std::vector<float> dmask, rmask, dtw;
  if (L >= 64) {
    dmask.resize((size_t)(nscales + 1) * L * L);
    rmask.resize(dmask.size());
    dtw.resize((size_t)2 * L);
    shx::dct_pyramid_masks(L, nscales, 0, dmask.data());
    shx::dct_pyramid_masks(L, nscales, 1, rmask.data());
    shx::dct_twiddles_fp32(L, dtw.data());
  }
  const size_t bytes_hm = sizeof(float2) * (size_t)nf * L * N, bytes_hb = sizeof(float2) * (size_t)nf * L;
  const size_t bytes_dm = dmask.size() * sizeof(float), bytes_dt = dtw.size() * sizeof(float);
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&d->Hmain, bytes_hm)) != hipSuccess || (e = hipMalloc(&d->Hbeta, bytes_hb)) != hipSuccess ||
      (e = hipMalloc(&d->tw, tw.size() * sizeof(float))) != hipSuccess ||
      (bytes_dm && ((e = hipMalloc(&d->dmask, bytes_dm)) != hipSuccess ||
                    (e = hipMalloc(&d->rmask, bytes_dm)) != hipSuccess ||
                    (e = hipMalloc(&d->dtw, bytes_dt)) != hipSuccess))) {
    shx_clear_shearing_filters(d);
    return fail(SHX_ERR_ALLOC, "device allocation failed: %s", hipGetErrorString(e));
  }
  if (int r = alloc_work(d, 1, d->w)) {
    shx_clear_shearing_filters(d);
    return r;
  }
  d->bytes = bytes_hm + bytes_hb + tw.size() * sizeof(float) + 2 * bytes_dm + bytes_dt + d->w.bytes;
  for (int j = 0; !host_spectra && j < nscales; ++j) {
    if (!owner[j]) continue;
    float2* hm_j = d->Hmain + (size_t)base[j] * L * N;
    float2* hb_j = d->Hbeta + (size_t)base[j] * L;
    const int r = w_on_device ? shx::spectra_device_from_dev(L, ndir[j], dsize[j], w_s[j], shift, hm_j, hb_j, j)
                              : shx::spectra_device(L, ndir[j], dsize[j], w_s[j], shift, hm_j, hb_j, j);
    if (r) {
      shx_clear_shearing_filters(d);
      return r;
    }
  }
  if ((host_spectra && ((e = hipMemcpy(d->Hmain, hm.data(), bytes_hm, hipMemcpyHostToDevice)) != hipSuccess ||
                        (e = hipMemcpy(d->Hbeta, hb.data(), bytes_hb, hipMemcpyHostToDevice)) != hipSuccess)) ||
      (e = hipMemcpy(d->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
      (bytes_dm && ((e = hipMemcpy(d->dmask, dmask.data(), bytes_dm, hipMemcpyHostToDevice)) != hipSuccess ||
                    (e = hipMemcpy(d->rmask, rmask.data(), bytes_dm, hipMemcpyHostToDevice)) != hipSuccess ||
                    (e = hipMemcpy(d->dtw, dtw.data(), bytes_dt, hipMemcpyHostToDevice)) != hipSuccess))) {
    shx_clear_shearing_filters(d);
    return fail(SHX_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
  }
  *out = d;
  return SHX_OK;
}
}  // namespace

extern "C" {

int shx_prepare_shearing_filters_ex(int L, int nscales, const int* ndir, const int* dsize, const double* const* w_s,
                                    int shift, int flags, shx_dict** out) {
  return prepare_dict(L, nscales, ndir, dsize, w_s, false, nullptr, shift, flags, out);
}

int shx_meyer_filters(int n1, int level, double* w_dev, void* stream) {
  return shx::meyer_filters_device(n1, level, w_dev, (hipStream_t)stream);
}

int shx_shearing_filters_myer(int L, int nscales, const int* dcomp, const int* dsize, int shift, int flags,
                              shx_dict** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!dcomp || !dsize) return fail(SHX_ERR_ARG, "null filter description");
  if (nscales < 1 || nscales > SHX_MAX_SCALES)
    return fail(SHX_ERR_ARG, "number of scales %d outside [1, %d]", nscales, SHX_MAX_SCALES);
  std::vector<DevMem> w(nscales);
  std::vector<const double*> wp(nscales);
  std::vector<int> ndir(nscales), same_as(nscales, -1);
  for (int j = 0; j < nscales; ++j) {
    // equal parameters give equal filters: compute them once
    for (int i = 0; i < j && same_as[j] < 0; ++i)
      if (dcomp[i] == dcomp[j] && dsize[i] == dsize[j]) same_as[j] = same_as[i] >= 0 ? same_as[i] : i;
    if (dcomp[j] < 0 || (1 << dcomp[j]) > SHX_MAX_DIRS)
      return fail(SHX_ERR_ARG, "scale %d: 2^%d directions (at most %d)", j, dcomp[j], SHX_MAX_DIRS);
    if (dsize[j] < 2 || dsize[j] > L) return fail(SHX_ERR_ARG, "scale %d: filter size %d invalid for L=%d", j, dsize[j], L);
    ndir[j] = 1 << dcomp[j];
    if (same_as[j] >= 0) {
      wp[j] = wp[same_as[j]];
      continue;
    }
    SHX_HIP(hipMalloc(&w[j].p, sizeof(double) * (size_t)ndir[j] * dsize[j] * dsize[j]));
    if (int r = shx::meyer_filters_device(dsize[j], dcomp[j], w[j].as<double>(), nullptr)) return r;
    wp[j] = w[j].as<double>();
  }
  return prepare_dict(L, nscales, ndir.data(), dsize, wp.data(), true, same_as.data(), shift, flags, out);
}

int shx_clear_shearing_filters(shx_dict* d) {
  if (!d) return SHX_OK;
  hipFree(d->Hmain);
  hipFree(d->Hbeta);
  hipFree(d->tw);
  hipFree(d->dmask);
  hipFree(d->rmask);
  hipFree(d->dtw);
  free_work(d->w);
  delete d;
  return SHX_OK;
}

int shx_dict_info(const shx_dict* d, int* L, int* nscales, int* ndir) {
  if (!d) return fail(SHX_ERR_ARG, "null dictionary");
  if (L) *L = d->L;
  if (nscales) *nscales = d->S;
  if (ndir)
    for (int j = 0; j < d->S; ++j) ndir[j] = d->ndir[j];
  return SHX_OK;
}

size_t shx_dict_bytes(const shx_dict* d) { return d ? d->bytes : 0; }

int shx_shear_trans(shx_dict* d, const float* x, float* const* coeff, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x) return fail(SHX_ERR_ARG, "null input");
  if (int r = check_cells(d, (const void* const*)coeff)) return r;
  std::vector<Stage> st;
  stages_dec(d, d->w, 1, x, 0, cell_coeffs(d, coeff), (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_nsst_scalars(shx_dict* d, const float* noise, int arith, double* scalars, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!noise || !scalars) return fail(SHX_ERR_ARG, "null noise image or output");
  if (arith != SHX_MAD_FP64 && arith != SHX_MAD_FP32) return fail(SHX_ERR_ARG, "unknown arithmetic %d", arith);
  const size_t pl = d->plane(), np = 1 + (size_t)d->nbands;
  hipStream_t s = (hipStream_t)stream;
  float* stack = nullptr;
  double* mad = nullptr;
  hipError_t e = hipMalloc(&stack, sizeof(float) * pl * np);
  if (e == hipSuccess) e = hipMalloc(&mad, sizeof(double) * np);
  if (e != hipSuccess) {
    if (stack) (void)hipFree(stack);
    return fail(SHX_ERR_ALLOC, "scalar workspace allocation failed: %s", hipGetErrorString(e));
  }
  std::vector<float*> cells(d->S + 1);
  cells[0] = stack;
  for (int j = 0; j < d->S; ++j) cells[j + 1] = stack + (1 + (size_t)d->map.first[j]) * pl;
  std::vector<Stage> st;
  stages_dec(d, d->w, 1, noise, 0, cell_coeffs(d, cells.data()), s, st);
  int r = run_stages(st);
  if (r == SHX_OK) r = shx_band_mad(stack, (long long)pl, (long long)pl, (int)np, arith, mad, stream);
  if (r == SHX_OK && ((e = hipMemcpyAsync(scalars, mad, sizeof(double) * np, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                      (e = hipStreamSynchronize(s)) != hipSuccess))
    r = fail(SHX_ERR_HIP, "scalar download failed: %s", hipGetErrorString(e));
  if (r != SHX_OK) (void)hipStreamSynchronize(s);
  (void)hipFree(stack);
  (void)hipFree(mad);
  return r;
}

int shx_inverse_shear(shx_dict* d, const float* const* coeff, float* x, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x) return fail(SHX_ERR_ARG, "null output");
  if (int r = check_cells(d, (const void* const*)coeff)) return r;
  std::vector<Stage> st;
  stages_rec(d, d->w, 1, cell_coeffs(d, const_cast<float* const*>(coeff)), Thr{}, x, 0, (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_inverse_shear_thresh(shx_dict* d, const float* const* coeff, const float* th, float* x, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !th) return fail(SHX_ERR_ARG, "null thresholds or output");
  if (int r = check_cells(d, (const void* const*)coeff)) return r;
  std::vector<Stage> st;
  Thr t;
  t.mode = 1;
  t.host = th;
  stages_rec(d, d->w, 1, cell_coeffs(d, const_cast<float* const*>(coeff)), t, x, 0, (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_denoise(shx_dict* d, const float* x, float* const* coeff, double sigma, const double* T, const double* sc,
                float* out, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !out || !T || !sc) return fail(SHX_ERR_ARG, "null input, output or threshold scalars");
  if (coeff)
    if (int r = check_cells(d, (const void* const*)coeff)) return r;
  std::vector<float> th = nsst_ht_thresholds(d, sigma, T, sc);
  Thr t;
  t.mode = 1;
  t.host = th.data();
  // one image: the kernels are short, so the single-stream chain beats the
  // concurrent form (whose cross-stream edges cost more than the overlap wins)
  std::vector<Stage> st;
  stages_denoise(d, d->w, 1, x, 0, coeff ? cell_coeffs(d, coeff) : work_coeffs(d, d->w), t, out, 0,
                 (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_denoise_profile(shx_dict* d, const float* x, float* const* coeff, double sigma, const double* T,
                        const double* sc, float* out, int reps, int max_stages, double* stage_ms, double* stage_bytes,
                        char* names, int name_len, int* nstages, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !out || !T || !sc || !stage_ms || !nstages || reps < 1)
    return fail(SHX_ERR_ARG, "invalid profiling arguments");
  if (coeff)
    if (int r = check_cells(d, (const void* const*)coeff)) return r;
  std::vector<float> th = nsst_ht_thresholds(d, sigma, T, sc);
  Thr t;
  t.mode = 1;
  t.host = th.data();
  std::vector<Stage> st;
  stages_denoise(d, d->w, 1, x, 0, coeff ? cell_coeffs(d, coeff) : work_coeffs(d, d->w), t, out, 0,
                 (hipStream_t)stream, st);
  return profile_stages(st, (hipStream_t)stream, reps, max_stages, stage_ms, stage_bytes, names, name_len, nstages);
}

// ---- batched form ------------------------------------------------------------
int shx_batch_create(shx_dict* d, int nimg, shx_batch** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (int r = check_dict(d)) return r;
  if (nimg < 1 || nimg > 65535) return fail(SHX_ERR_ARG, "batch of %d images", nimg);
  auto b = new shx_batch();
  b->d = d;
  b->nimg = nimg;
  if (int r = alloc_work(d, nimg, b->w)) {
    delete b;
    return r;
  }
  *out = b;
  return SHX_OK;
}

int shx_batch_destroy(shx_batch* b) {
  if (!b) return SHX_OK;
  free_work(b->w);
  delete b;
  return SHX_OK;
}

size_t shx_batch_bytes(const shx_batch* b) { return b ? b->w.bytes : 0; }

static int batch_args(shx_batch* b, const void* x, const void* c, const void* o) {
  if (!b) return fail(SHX_ERR_ARG, "null batch");
  if (int r = check_dict(b->d)) return r;
  if (!x || !c || !o) return fail(SHX_ERR_ARG, "null device buffer");
  return SHX_OK;
}

int shx_batch_shear_trans(shx_batch* b, const float* x, float* coeff, void* stream) {
  if (int r = batch_args(b, x, coeff, coeff)) return r;
  shx_dict* d = b->d;
  std::vector<Stage> st;
  stages_dec(d, b->w, b->nimg, x, (long long)d->L * d->L, stack_coeffs(d, coeff), (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_batch_inverse_shear(shx_batch* b, const float* coeff, const float* th_dev, float* out, void* stream) {
  if (int r = batch_args(b, coeff, coeff, out)) return r;
  shx_dict* d = b->d;
  Thr t;
  if (th_dev) {
    t.mode = 2;
    t.dev = th_dev;
  }
  std::vector<Stage> st;
  stages_rec(d, b->w, b->nimg, stack_coeffs(d, const_cast<float*>(coeff)), t, out, (long long)d->L * d->L,
             (hipStream_t)stream, st);
  return run_stages(st);
}

// The batched denoise: with the DCT-domain front end every scale's spectrum
// comes out of one pass, so the per-scale side-stream chains of
// stages_denoise_dag only add cross-stream edges (measured 2% slower at 32 x
// 512^2 and 16 x 1024^2) and the single-stream chain runs; with the level
// pyramid each scale's chain starts as soon as its level is done.
static int run_batch_denoise(shx_dict* d, shx_batch* b, const float* x, float* coeff, const Thr& t, float* out,
                             hipStream_t s) {
  std::vector<Stage> st;
  const long long pl = (long long)d->L * d->L;
  const Coeffs c = coeff ? stack_coeffs(d, coeff) : work_coeffs(d, b->w);
  if (b->w.dct_front) {
    stages_denoise(d, b->w, b->nimg, x, pl, c, t, out, pl, s, st);
    return run_stages(st);
  }
  stages_denoise_dag(d, b->w, b->nimg, x, pl, c, t, out, pl, s, st);
  return run_stage_dag(st, b->w);
}

int shx_batch_denoise(shx_batch* b, const float* x, float* coeff, const float* th_dev, float* out, void* stream) {
  if (int r = batch_args(b, x, out, out)) return r;
  if (!th_dev) return fail(SHX_ERR_ARG, "null threshold table");
  shx_dict* d = b->d;
  Thr t;
  t.mode = 2;
  t.dev = th_dev;
  return run_batch_denoise(d, b, x, coeff, t, out, (hipStream_t)stream);
}

static int shrink_thr(shx_batch* b, const float* th_dev, int th_stride, Thr& t) {
  if (!th_dev) return fail(SHX_ERR_ARG, "null threshold table");
  if (th_stride != 0 && th_stride < 1 + b->d->nbands)
    return fail(SHX_ERR_ARG, "threshold row stride %d (0, or at least 1 + nbands = %d)", th_stride,
                1 + b->d->nbands);
  t.mode = 2;
  t.dev = th_dev;
  t.stride = th_stride;
  t.soft = true;
  return SHX_OK;
}

int shx_batch_shrink(shx_batch* b, const float* x, float* coeff, const float* th_dev, int th_stride, float* out,
                     void* stream) {
  if (int r = batch_args(b, x, out, out)) return r;
  Thr t;
  if (int r = shrink_thr(b, th_dev, th_stride, t)) return r;
  return run_batch_denoise(b->d, b, x, coeff, t, out, (hipStream_t)stream);
}

int shx_batch_shrink_backward(shx_batch* b, const float* coeff, const float* th_dev, int th_stride,
                              const float* out_bar, float* x_bar, float* gth_dev, void* stream) {
  if (int r = batch_args(b, coeff, out_bar, x_bar)) return r;
  if (!gth_dev) return fail(SHX_ERR_ARG, "null threshold gradient");
  Thr t;
  if (int r = shrink_thr(b, th_dev, th_stride, t)) return r;
  shx_dict* d = b->d;
  std::vector<Stage> st;
  const long long pl = (long long)d->L * d->L;
  stages_shrink_backward(d, b->w, b->nimg, stack_coeffs(d, const_cast<float*>(coeff)), t, out_bar, pl, gth_dev,
                         x_bar, pl, (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_autotune(shx_dict* d, shx_batch* b, int reps, int* chosen, void* stream) {
  if (int r = check_dict(d)) return r;
  if (b && b->d != d) return fail(SHX_ERR_ARG, "batch belongs to another dictionary");
  if (reps < 1) reps = 3;
  Work& w = b ? b->w : d->w;
  return autotune_work(d, w, b ? b->nimg : 1, reps, (hipStream_t)stream, chosen);
}

int shx_set_plan_options(shx_dict* d, shx_batch* b, int dct_front, int dct_back) {
  if (!d) return fail(SHX_ERR_ARG, "null dictionary");
  if (b && b->d != d) return fail(SHX_ERR_ARG, "batch belongs to another dictionary");
  if ((dct_front != 0 && dct_front != 1) || (dct_back != 0 && dct_back != 1))
    return fail(SHX_ERR_ARG, "plan options %d, %d (0 pyramid levels, 1 DCT domain)", dct_front, dct_back);
  if ((dct_front || dct_back) && !d->dmask)
    return fail(SHX_ERR_UNSUPPORTED, "the DCT-domain pyramid needs L >= 64 (L = %d)", d->L);
  Work& w = b ? b->w : d->w;
  w.dct_front = dct_front == 1;
  w.dct_back = dct_back == 1;
  return SHX_OK;
}

int shx_get_plan_options(const shx_dict* d, const shx_batch* b, int* dct_front, int* dct_back) {
  if (!d) return fail(SHX_ERR_ARG, "null dictionary");
  if (b && b->d != d) return fail(SHX_ERR_ARG, "batch belongs to another dictionary");
  const Work& w = b ? b->w : d->w;
  if (dct_front) *dct_front = w.dct_front ? 1 : 0;
  if (dct_back) *dct_back = w.dct_back ? 1 : 0;
  return SHX_OK;
}

int shx_set_schedule(shx_dict* d, shx_batch* b, int pyr_variant, int group, int ct_dec, int ct_rec) {
  if (!d) return fail(SHX_ERR_ARG, "null dictionary");
  if (pyr_variant < 0 || pyr_variant >= shx::PYR_VARIANTS) return fail(SHX_ERR_ARG, "no pyramid variant %d", pyr_variant);
  if (group != 2 && group != 4 && group != 8) return fail(SHX_ERR_ARG, "direction group must be 2, 4 or 8");
  for (int ct : {ct_dec, ct_rec})
    if (ct != 0 && !shx::col_tile_ok(d->L, ct))
      return fail(SHX_ERR_ARG, "column tile %d not available for L = %d (0, 4, 8 or 16 by size)", ct, d->L);
  Work& w = b ? b->w : d->w;
  w.pyr_variant = pyr_variant;
  set_group(d, w, group);
  w.ct_dec = ct_dec;
  w.ct_rec = ct_rec;
  return SHX_OK;
}

// ---- adjoints ----------------------------------------------------------------------
int shx_shear_trans_adjoint(shx_dict* d, const float* const* coeff_bar, float* x_bar, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x_bar) return fail(SHX_ERR_ARG, "null output");
  if (int r = check_cells(d, (const void* const*)coeff_bar)) return r;
  std::vector<Stage> st;
  stages_dec_adjoint(d, d->w, 1, cell_coeffs(d, const_cast<float* const*>(coeff_bar)), x_bar, 0,
                     (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_inverse_shear_adjoint(shx_dict* d, const float* x_bar, float* const* coeff_bar, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x_bar) return fail(SHX_ERR_ARG, "null input");
  if (int r = check_cells(d, (const void* const*)coeff_bar)) return r;
  std::vector<Stage> st;
  stages_rec_adjoint(d, d->w, 1, x_bar, 0, cell_coeffs(d, coeff_bar), (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_batch_shear_trans_adjoint(shx_batch* b, const float* coeff_bar, float* x_bar, void* stream) {
  if (int r = batch_args(b, coeff_bar, coeff_bar, x_bar)) return r;
  shx_dict* d = b->d;
  std::vector<Stage> st;
  stages_dec_adjoint(d, b->w, b->nimg, stack_coeffs(d, const_cast<float*>(coeff_bar)), x_bar,
                     (long long)d->L * d->L, (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_batch_inverse_shear_adjoint(shx_batch* b, const float* x_bar, float* coeff_bar, void* stream) {
  if (int r = batch_args(b, x_bar, coeff_bar, coeff_bar)) return r;
  shx_dict* d = b->d;
  std::vector<Stage> st;
  stages_rec_adjoint(d, b->w, b->nimg, x_bar, (long long)d->L * d->L, stack_coeffs(d, coeff_bar),
                     (hipStream_t)stream, st);
  return run_stages(st);
}

// ---- graphs --------------------------------------------------------------------
int shx_capture_begin(void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail(SHX_ERR_ARG, "capture needs a non-default stream");
  SHX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  return SHX_OK;
}

int shx_capture_end(void* stream, shx_graph** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  hipGraph_t g = nullptr;
  SHX_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return fail(SHX_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  }
  auto gr = new shx_graph();
  gr->graph = g;
  gr->exec = ex;
  *out = gr;
  return SHX_OK;
}

int shx_graph_launch(shx_graph* g, void* stream) {
  if (!g) return fail(SHX_ERR_ARG, "null graph");
  SHX_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return SHX_OK;
}

int shx_graph_parallel(shx_graph* const* parts, int n, shx_graph** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!parts || n < 1) return fail(SHX_ERR_ARG, "no graphs to combine");
  for (int i = 0; i < n; ++i)
    if (!parts[i] || !parts[i]->graph) return fail(SHX_ERR_ARG, "graph %d is null", i);
  hipGraph_t g = nullptr;
  SHX_HIP(hipGraphCreate(&g, 0));
  for (int i = 0; i < n; ++i) {
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddChildGraphNode(&node, g, nullptr, 0, parts[i]->graph);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      return fail(SHX_ERR_HIP, "adding graph %d failed: %s", i, hipGetErrorString(e));
    }
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return fail(SHX_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  }
  auto gr = new shx_graph();
  gr->graph = g;
  gr->exec = ex;
  *out = gr;
  return SHX_OK;
}

int shx_graph_destroy(shx_graph* g) {
  if (!g) return SHX_OK;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return SHX_OK;
}

int shx_denoise_graph(shx_dict* d, const float* x, float* const* coeff, double sigma, const double* T,
                      const double* sc, float* out, shx_graph** g) {
  if (!g) return fail(SHX_ERR_ARG, "null output pointer");
  *g = nullptr;
  if (int r = check_dict(d)) return r;
  hipStream_t s = nullptr;
  SHX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  int rc = shx_capture_begin(s);
  if (rc == SHX_OK) {
    int rc2 = shx_denoise(d, x, coeff, sigma, T, sc, out, s);
    shx_graph* gr = nullptr;
    rc = shx_capture_end(s, &gr);
    if (rc2 != SHX_OK) {
      shx_graph_destroy(gr);
      rc = rc2;
    } else if (rc == SHX_OK) {
      *g = gr;
    }
  }
  hipStreamDestroy(s);
  return rc;
}

// ---- thresholds, pyramid, helpers -------------------------------------------------
int shx_thresh(float* const* cells, int ncells, const int* cell_planes, const int64_t* plane_elems, double lambda,
               int option, const double* E, int E_rows, int E_cols, const double* sc, int n_sc, int opt, double* maxi,
               void* stream) {
  if (!cells || !cell_planes || !plane_elems) return fail(SHX_ERR_ARG, "null cell description");
  if (option != 1 && option != 2) return fail(SHX_ERR_ARG, "option must be 1 (wavelets) or 2 (shearlets)");
  hipStream_t s = (hipStream_t)stream;
  std::vector<float*> ptr;
  std::vector<long long> len;
  std::vector<float> th, invE;
  if (option == 1) {
    if (ncells < 2) return fail(SHX_ERR_ARG, "Number of cell elements of 'Ct' should be at least 2");
    for (int w = 0; w < cell_planes[1]; ++w) {
      ptr.push_back(cells[1] + (size_t)w * plane_elems[1]);
      len.push_back(plane_elems[1]);
      th.push_back((float)lambda);
      invE.push_back(1.0f);
    }
  } else {
    if (!E || !sc) return fail(SHX_ERR_ARG, "null E or sc");
    if (n_sc != E_rows) return fail(SHX_ERR_ARG, "Number of elmenents of 'sc'' does not match number of rows of 'E'");
    if (ncells != E_rows)
      return fail(SHX_ERR_ARG, "Number of cell elements of 'Ct' does not match number of rows of 'E'");
    for (int c = 0; c < ncells; ++c) {
      if (cell_planes[c] > E_cols) return fail(SHX_ERR_ARG, "cell %d has more planes than E has columns", c);
      for (int w = 0; w < cell_planes[c]; ++w) {
        const double e = E[c + (size_t)w * E_rows];
        ptr.push_back(cells[c] + (size_t)w * plane_elems[c]);
        len.push_back(plane_elems[c]);
        th.push_back((float)(sc[c] * lambda * e));
        invE.push_back((float)(1.0 / e));
      }
    }
  }
  ReduceSlot slot;
  unsigned int* dmax = nullptr;
  if (opt == 1) {
    if (int r = reduce_slot_acquire(slot)) return r;
    dmax = static_cast<unsigned int*>(slot.p);
    SHX_HIP(hipMemsetAsync(dmax, 0, sizeof(unsigned int), s));
  }
  for (size_t b = 0; b < ptr.size(); b += 64) {
    ThreshTable t{};
    t.n = (int)std::min<size_t>(64, ptr.size() - b);
    long long mx = 0;
    for (int i = 0; i < t.n; ++i) {
      t.ptr[i] = ptr[b + i];
      t.len[i] = len[b + i];
      t.th[i] = th[b + i];
      t.invE[i] = invE[b + i];
      mx = std::max(mx, len[b + i]);
    }
    // about 16 elements per thread: with opt, one atomic per block
    const int blocks = (int)std::max<long long>(1, std::min<long long>(1024, (mx + 4095) / 4096));
    hipLaunchKernelGGL(k_hard_thresh, dim3(blocks, t.n), dim3(256), 0, s, t, opt == 1 ? 1 : 0, dmax);
    SHX_HIP(hipGetLastError());
  }
  if (maxi) *maxi = 0.0;
  if (opt == 1) {
    unsigned int bits = 0;
    SHX_HIP(hipMemcpyAsync(&bits, dmax, sizeof bits, hipMemcpyDeviceToHost, s));
    SHX_HIP(hipStreamSynchronize(s));
    float f;
    memcpy(&f, &bits, sizeof f);
    if (maxi) *maxi = f;
  }
  return SHX_OK;
}

int shx_atrousdec(const float* x, int L, int nlevels, float* const* y, float* work, void* stream) {
  if (!x || !y) return fail(SHX_ERR_ARG, "null input or output");
  if (!is_pow2(L) || L < 32) return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported", L);
  if (nlevels < 1 || nlevels > 9 || L % (1 << (nlevels - 1)) != 0)
    return fail(SHX_ERR_ARG, "invalid number of levels %d", nlevels);
  for (int j = 0; j <= nlevels; ++j)
    if (!y[j]) return fail(SHX_ERR_ARG, "output plane %d is null", j);
  float* w = work;
  if (!w) SHX_HIP(hipMalloc(&w, sizeof(float) * 2 * L * L));
  hipError_t e = pyr_dec(x, L, nlevels, y, w, w + (size_t)L * L, (hipStream_t)stream);
  if (!work) {
    hipStreamSynchronize((hipStream_t)stream);
    hipFree(w);
  }
  SHX_HIP(e);
  return SHX_OK;
}

}  // extern "C"

// ---- allocation-free pyramid decomposition ---------------------------------------
// Level temporaries of shx_atrousdec_noalloc: two planes of at most
// SHX_NOALLOC_MAX_L^2 floats in a file-scope device array.  Every device the
// library runs on holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in): the
// device address of each device's copy comes from hipGetSymbolAddress, taken
// once per device with that device current.
static __device__ __align__(256) float scratch[2 * (size_t)SHX_NOALLOC_MAX_L * SHX_NOALLOC_MAX_L];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`.
int scratch_address(float** out) {
  int dev = -1;
  SHX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SHX_ERR_UNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SHX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SHX_OK;
}
}  // namespace

extern "C" {

int shx_atrousdec_noalloc(const float* x, int L, int nlevels, float* const* y, void* stream) {
  if (!x || !y) return fail(SHX_ERR_ARG, "null input or output");
  if (!is_pow2(L) || L < 32) return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported", L);
  if (L > SHX_NOALLOC_MAX_L)
    return fail(SHX_ERR_ARG, "image size %d exceeds the %d x %d planes of the static scratch (use shx_atrousdec)", L,
                SHX_NOALLOC_MAX_L, SHX_NOALLOC_MAX_L);
  if (nlevels < 1 || nlevels > 9 || L % (1 << (nlevels - 1)) != 0)
    return fail(SHX_ERR_ARG, "invalid number of levels %d", nlevels);
  for (int j = 0; j <= nlevels; ++j)
    if (!y[j]) return fail(SHX_ERR_ARG, "output plane %d is null", j);
  float* w = nullptr;
  if (int r = scratch_address(&w)) return r;
  // every level writes its whole temporary before the next level reads it, so
  // the scratch needs no clearing
  SHX_HIP(pyr_dec(x, L, nlevels, y, w, w + (size_t)L * L, (hipStream_t)stream));
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_atrousrec(const float* const* y, int L, int nlevels, float* x, float* work, void* stream) {
  if (!x || !y) return fail(SHX_ERR_ARG, "null input or output");
  if (!is_pow2(L) || L < 32) return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported", L);
  if (nlevels < 1 || nlevels > 9 || L % (1 << (nlevels - 1)) != 0)
    return fail(SHX_ERR_ARG, "invalid number of levels %d", nlevels);
  for (int j = 0; j <= nlevels; ++j)
    if (!y[j]) return fail(SHX_ERR_ARG, "input plane %d is null", j);
  float* w = work;
  if (!w) SHX_HIP(hipMalloc(&w, sizeof(float) * 2 * L * L));
  hipError_t e = pyr_rec(y, L, nlevels, x, w, w + (size_t)L * L, nullptr, (hipStream_t)stream);
  if (!work) {
    hipStreamSynchronize((hipStream_t)stream);
    hipFree(w);
  }
  SHX_HIP(e);
  return SHX_OK;
}

int shx_re_weight_band(const float* x, int L, const double* coeff, int ncoeff, float* out, float* work,
                       void* stream) {
  if (!x || !out || !coeff) return fail(SHX_ERR_ARG, "null input, weights or output");
  if (ncoeff <= 1) return fail(SHX_ERR_ARG, "There should be at least one band");
  if (!is_pow2(L) || L < 32) return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported", L);
  const int nlev = ncoeff - 1;
  if (nlev > 9 || L % (1 << (nlev - 1)) != 0) return fail(SHX_ERR_ARG, "invalid number of levels %d", nlev);
  const size_t pl = (size_t)L * L;
  float* w = work;
  if (!w) SHX_HIP(hipMalloc(&w, sizeof(float) * (ncoeff + 2) * pl));
  float* y[10];
  for (int j = 0; j < ncoeff; ++j) y[j] = w + j * pl;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = pyr_dec(x, L, nlev, y, w + ncoeff * pl, w + (ncoeff + 1) * pl, s);
  if (e == hipSuccess) e = pyr_rec(y, L, nlev, out, w + ncoeff * pl, w + (ncoeff + 1) * pl, coeff, s);
  if (!work) {
    hipStreamSynchronize(s);
    hipFree(w);
  }
  SHX_HIP(e);
  return SHX_OK;
}

}  // extern "C"

int shx_hard_bands(const float* x, float* y, int nimg, int64_t img_elems, int64_t rowlen, int64_t seglen,
                   int band_off, const float* th_dev, void* stream) {
  if (!x || !y || !th_dev) return fail(SHX_ERR_ARG, "null input, output or thresholds");
  if (nimg < 1 || img_elems < 1 || rowlen < 1 || seglen < 1 || img_elems % rowlen || rowlen % seglen)
    return fail(SHX_ERR_ARG, "invalid band layout");
  const long long n = (long long)nimg * img_elems;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(8192, (n + 255) / 256));
  hipLaunchKernelGGL(k_hard_bands, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, nullptr, y, n,
                     BandLayout{img_elems, rowlen, seglen, band_off}, th_dev);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_hard_bands_backward(const float* x, const float* gy, float* gx, int nimg, int64_t img_elems, int64_t rowlen,
                            int64_t seglen, int band_off, const float* th_dev, void* stream) {
  if (!x || !gy || !gx || !th_dev) return fail(SHX_ERR_ARG, "null input, gradient or thresholds");
  if (nimg < 1 || img_elems < 1 || rowlen < 1 || seglen < 1 || img_elems % rowlen || rowlen % seglen)
    return fail(SHX_ERR_ARG, "invalid band layout");
  const long long n = (long long)nimg * img_elems;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(8192, (n + 255) / 256));
  hipLaunchKernelGGL(k_hard_bands, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, gy, gx, n,
                     BandLayout{img_elems, rowlen, seglen, band_off}, th_dev);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_soft_bands(const float* x, float* y, int nimg, int64_t img_elems, int64_t rowlen, int64_t seglen,
                   int band_off, const float* th_dev, void* stream) {
  if (!x || !y || !th_dev) return fail(SHX_ERR_ARG, "null input, output or thresholds");
  if (nimg < 1 || img_elems < 1 || rowlen < 1 || seglen < 1 || img_elems % rowlen || rowlen % seglen)
    return fail(SHX_ERR_ARG, "invalid band layout");
  const long long n = (long long)nimg * img_elems;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(8192, (n + 255) / 256));
  hipLaunchKernelGGL(k_soft_bands, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, y, n,
                     BandLayout{img_elems, rowlen, seglen, band_off}, th_dev);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_soft_bands_backward(const float* x, const float* gy, float* gx, float* gth_dev, int nimg, int64_t img_elems,
                            int64_t rowlen, int64_t seglen, int band_off, const float* th_dev, void* stream) {
  if (!x || !gy || !gth_dev || !th_dev) return fail(SHX_ERR_ARG, "null input, gradient or thresholds");
  if (nimg < 1 || img_elems < 1 || rowlen < 1 || seglen < 1 || img_elems % rowlen || rowlen % seglen)
    return fail(SHX_ERR_ARG, "invalid band layout");
  const long long nbands = band_off + rowlen / seglen;
  if (band_off < 0 || nbands > SHX_MAX_BANDS) return fail(SHX_ERR_ARG, "%lld bands (at most %d)", nbands, SHX_MAX_BANDS);
  const long long n = (long long)nimg * img_elems;
  // chunks of whole waves, about 2048 blocks
  long long chunk = std::max<long long>(4096, (n + 2047) / 2048);
  chunk = (chunk + 255) / 256 * 256;
  const int blocks = (int)((n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_soft_bands_bwd, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, gy, gx, gth_dev, n,
                     BandLayout{img_elems, rowlen, seglen, band_off}, th_dev, chunk, (int)nbands);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

// ---- undecimated wavelet transform ---------------------------------------------
namespace {
// taps of the analysis (a0, a1) or synthesis (g0, g1) passes from the scaling
// filter h (mrdwt_r.h:134-140, mirdwt_r.h:136-141)
shx::RdwtFilt rdwt_taps(const double* h, int lh, bool synthesis, double scale) {
  shx::RdwtFilt f{};
  f.lh = lh;
  for (int j = 0; j < lh; ++j) {
    if (!synthesis) {
      f.f0[j] = (float)(scale * h[j]);
      f.f1[j] = (float)(scale * h[lh - 1 - j] * (((lh - 1 - j) % 2 == 0) ? -1.0 : 1.0));
    } else {
      f.f0[j] = (float)(scale * h[j] / 2.0);
      f.f1[j] = (float)(scale * h[lh - 1 - j] / 2.0 * ((j % 2 == 1) ? -1.0 : 1.0));
    }
  }
  return f;
}

int rdwt_check(int m, int n, const double* h, int lh, int L, int nimg) {
  if (!h) return fail(SHX_ERR_ARG, "null filter");
  if (lh < 2 || lh > shx::RDWT_MAX_TAPS || lh % 2) return fail(SHX_ERR_ARG, "filter length %d (even, <= %d)", lh,
                                                               shx::RDWT_MAX_TAPS);
  if (L < 1 || L > 12) return fail(SHX_ERR_ARG, "invalid number of levels %d", L);
  if (m < 1 || n < 1 || m % (1 << L) || n % (1 << L))
    return fail(SHX_ERR_ARG, "The row and the column dimension must be divisible by 2^L");
  if (nimg < 1) return fail(SHX_ERR_ARG, "batch of %d images", nimg);
  return SHX_OK;
}

template <class F>
int with_tmp(float* work, size_t floats, hipStream_t s, F f) {
  float* w = work;
  if (!w) SHX_HIP(hipMalloc(&w, sizeof(float) * floats));
  hipError_t e = f(w);
  if (!work) {
    hipStreamSynchronize(s);
    hipFree(w);
  }
  SHX_HIP(e);
  return SHX_OK;
}
}  // namespace

extern "C" {

int shx_mrdwt(const float* x, int m, int n, const double* h, int lh, int L, float* yl, float* yh, float* work,
              int nimg, void* stream) {
  if (int r = rdwt_check(m, n, h, lh, L, nimg)) return r;
  if (!x || !yl || !yh) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  const shx::RdwtFilt f = rdwt_taps(h, lh, false, 1.0);
  return with_tmp(work, 2 * (size_t)m * n * nimg, s,
                  [&](float* t) { return shx::rdwt_analysis(x, yl, yh, t, m, n, L, f, nimg, s); });
}

int shx_mirdwt(const float* yl, const float* yh, int m, int n, const double* h, int lh, int L, float* x, float* work,
               int nimg, void* stream) {
  if (int r = rdwt_check(m, n, h, lh, L, nimg)) return r;
  if (!x || !yl || !yh) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  const shx::RdwtFilt f = rdwt_taps(h, lh, true, 1.0);
  return with_tmp(work, 2 * (size_t)m * n * nimg, s,
                  [&](float* t) { return shx::rdwt_synthesis(yl, yh, x, t, m, n, L, f, nimg, s); });
}

int shx_mrdwt_adjoint(const float* yl_bar, const float* yh_bar, int m, int n, const double* h, int lh, int L,
                      float* x_bar, float* work, int nimg, void* stream) {
  if (int r = rdwt_check(m, n, h, lh, L, nimg)) return r;
  if (!x_bar || !yl_bar || !yh_bar) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  const shx::RdwtFilt f = rdwt_taps(h, lh, true, 2.0);  // synthesis passes with the analysis taps
  return with_tmp(work, 2 * (size_t)m * n * nimg, s,
                  [&](float* t) { return shx::rdwt_synthesis(yl_bar, yh_bar, x_bar, t, m, n, L, f, nimg, s); });
}

int shx_mirdwt_adjoint(const float* x_bar, int m, int n, const double* h, int lh, int L, float* yl_bar,
                       float* yh_bar, float* work, int nimg, void* stream) {
  if (int r = rdwt_check(m, n, h, lh, L, nimg)) return r;
  if (!x_bar || !yl_bar || !yh_bar) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  const shx::RdwtFilt f = rdwt_taps(h, lh, false, 0.5);  // analysis passes with the synthesis taps
  return with_tmp(work, 2 * (size_t)m * n * nimg, s,
                  [&](float* t) { return shx::rdwt_analysis(x_bar, yl_bar, yh_bar, t, m, n, L, f, nimg, s); });
}

int shx_norm_pow(const float* x, int64_t n, double p, double* out, void* stream) {
  if (!x || !out) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  ReduceSlot slot;
  if (int r = reduce_slot_acquire(slot)) return r;
  double* acc = static_cast<double*>(slot.p);
  SHX_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
  const int blocks = (int)std::max<long long>(1, std::min<long long>(1024, (n + 4095) / 4096));
  if (n > 0) hipLaunchKernelGGL(k_norm_pow, dim3(blocks), dim3(256), 0, s, x, (long long)n, (float)p, acc);
  SHX_HIP(hipGetLastError());
  SHX_HIP(hipMemcpyAsync(out, acc, sizeof(double), hipMemcpyDeviceToHost, s));
  SHX_HIP(hipStreamSynchronize(s));
  return SHX_OK;
}

int shx_norm_pow_diff(const float* a, const float* b, int64_t n, double p, double* out, void* stream) {
  if (!a || !b || !out) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  ReduceSlot slot;
  if (int r = reduce_slot_acquire(slot)) return r;
  double* acc = static_cast<double*>(slot.p);
  SHX_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
  const int blocks = (int)std::max<long long>(1, std::min<long long>(1024, (n + 4095) / 4096));
  if (n > 0) hipLaunchKernelGGL(k_norm_pow_diff, dim3(blocks), dim3(256), 0, s, a, b, (long long)n, (float)p, acc);
  SHX_HIP(hipGetLastError());
  SHX_HIP(hipMemcpyAsync(out, acc, sizeof(double), hipMemcpyDeviceToHost, s));
  SHX_HIP(hipStreamSynchronize(s));
  return SHX_OK;
}

int shx_add_noise(float* x, int64_t n, double sigma, uint64_t seed, void* stream) {
  if (!x) return fail(SHX_ERR_ARG, "null input");
  if (n < 0) return fail(SHX_ERR_ARG, "negative length");
  const int blocks = (int)std::max<long long>(1, std::min<long long>(4096, (n + 255) / 256));
  if (n > 0)
    hipLaunchKernelGGL(k_add_noise, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, (long long)n, (float)sigma,
                       (unsigned long long)seed);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_convolution(shx_dict* d, const float* const* y, float* const* coeff, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!y) return fail(SHX_ERR_ARG, "null pyramid cells");
  for (int j = 0; j <= d->S; ++j)
    if (!y[j]) return fail(SHX_ERR_ARG, "pyramid cell %d is null", j);
  if (int r = check_cells(d, (const void* const*)coeff)) return r;
  hipStream_t s = (hipStream_t)stream;
  const size_t pl = sizeof(float) * d->plane();
  const Work& w = d->w;
  for (int j = 0; j < d->S; ++j)
    SHX_HIP(hipMemcpyAsync(w.planes + (size_t)j * d->plane(), y[j + 1], pl, hipMemcpyDeviceToDevice, s));
  if (coeff[0] != y[0]) SHX_HIP(hipMemcpyAsync(coeff[0], y[0], pl, hipMemcpyDeviceToDevice, s));
  std::vector<Stage> st;
  stages_spec_dec(d, w, 1, s, st);
  const int L = d->L, nb = d->nbands;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  PlaneSet out = cell_coeffs(d, coeff).bands;
  st.push_back({"rows_c2r_bands", 8.0 * nb * L * L,
                [=] { return shx::launch_rows_c2r(L, bandset, out, nb, 1, nullptr, d->tw, s); }});
  return run_stages(st);
}

int shx_applyfilters(shx_dict* d, const float* const* coeff, float* const* y, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!y) return fail(SHX_ERR_ARG, "null pyramid cells");
  for (int j = 1; j <= d->S; ++j)
    if (!y[j]) return fail(SHX_ERR_ARG, "pyramid cell %d is null", j);
  if (int r = check_cells(d, (const void* const*)coeff)) return r;
  hipStream_t s = (hipStream_t)stream;
  const Work& w = d->w;
  const int L = d->L, nb = d->nbands;
  PlaneSet in = cell_coeffs(d, const_cast<float* const*>(coeff)).bands;
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(w.band), nb, w.s_band);
  shx::ThreshSpec none{};
  std::vector<Stage> st;
  st.push_back({"rows_r2c_bands", 8.0 * nb * L * L,
                [=] { return shx::launch_rows_r2c(L, in, bandset, nb, 1, none, d->tw, s); }});
  stages_spec_rec_tail(d, w, 1, s, st);
  if (int r = run_stages(st)) return r;
  const size_t pl = sizeof(float) * d->plane();
  for (int j = 0; j < d->S; ++j)
    SHX_HIP(hipMemcpyAsync(y[j + 1], w.planes + (size_t)j * d->plane(), pl, hipMemcpyDeviceToDevice, s));
  return SHX_OK;
}

int shx_convert(void* dst, int dst_type, const void* src, int src_type, int64_t n, void* stream) {
  if (!dst || !src) return fail(SHX_ERR_ARG, "null buffer");
  if (n < 0) return fail(SHX_ERR_ARG, "negative length");
  const int blocks = (int)std::max<long long>(1, std::min<long long>(4096, (n + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  if (src_type == SHX_TYPE_FLOAT && dst_type == SHX_TYPE_UINT8) {
    if (n > 0)
      hipLaunchKernelGGL(k_f32_to_u8, dim3(blocks), dim3(256), 0, s, (const float*)src, (unsigned char*)dst,
                         (long long)n);
  } else if (src_type == SHX_TYPE_UINT8 && dst_type == SHX_TYPE_FLOAT) {
    if (n > 0)
      hipLaunchKernelGGL(k_u8_to_f32, dim3(blocks), dim3(256), 0, s, (const unsigned char*)src, (float*)dst,
                         (long long)n);
  } else if (src_type == dst_type && (src_type == SHX_TYPE_FLOAT || src_type == SHX_TYPE_UINT8)) {
    const size_t es = src_type == SHX_TYPE_FLOAT ? sizeof(float) : 1;
    if (n > 0 && dst != src) SHX_HIP(hipMemcpyAsync(dst, src, es * (size_t)n, hipMemcpyDeviceToDevice, s));
  } else {
    return fail(SHX_ERR_UNSUPPORTED, "conversion %d -> %d not supported", src_type, dst_type);
  }
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

int shx_softth(float* x, int64_t n, double t, void* stream) {
  if (!x) return fail(SHX_ERR_ARG, "null input");
  const int blocks = (int)std::max<long long>(1, std::min<long long>(4096, (n + 255) / 256));
  if (n > 0) hipLaunchKernelGGL(k_softth, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, (long long)n, (float)t);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The shipped separation loop, shearcuda/separate.m:41-257 (GPU branch) and
// crack_detection/separateHairlineCrack.m, run natively: the same sequence of
// library calls as the Python mirror, over one workspace allocated per call.
namespace {

std::vector<double> on_filter(const double* f, int n) {
  double ss = 0.0;
  for (int i = 0; i < n; ++i) ss += f[i] * f[i];
  const double r = std::sqrt(ss);
  std::vector<double> o(n);
  for (int i = 0; i < n; ++i) o[i] = f[i] / r;
  return o;
}

}  // namespace

extern "C" int shx_separate(shx_dict* d, const float* img, int level2, int itermax, double stop, double gamma,
                            const double* coeff, int ncoeff, const double* qmf, int lqmf, double delta_max,
                            const double* E, int E_rows, int E_cols, float* C, float* P, double* info,
                            void* stream) {
  if (int r = check_dict(d)) return r;
  if (!img || !coeff || !qmf || !C || !P) return fail(SHX_ERR_ARG, "null image, weights, filter or output");
  const int L = d->L, S = d->S, nb = d->nbands;
  if (itermax < 2) return fail(SHX_ERR_ARG, "itermax must be at least 2 (the schedule divides by 1 - itermax)");
  if (level2 < 1 || L % (1 << level2) != 0) return fail(SHX_ERR_ARG, "invalid number of wavelet levels %d", level2);
  if (lqmf < 2 || lqmf % 2) return fail(SHX_ERR_ARG, "wavelet filter length %d must be even", lqmf);
  if (ncoeff <= 1 || ncoeff > 10) return fail(SHX_ERR_ARG, "need 2..10 band weights");
  int maxd = 1;
  for (int j = 0; j < S; ++j) maxd = std::max(maxd, d->ndir[j]);
  if (E && E_rows != S + 1) return fail(SHX_ERR_ARG, "E has %d rows, the dictionary needs %d", E_rows, S + 1);
  if (E && E_cols < maxd) return fail(SHX_ERR_ARG, "E has %d columns, the dictionary needs %d", E_cols, maxd);
  hipStream_t s = (hipStream_t)stream;
  const size_t pl = (size_t)L * L, nyh = (size_t)3 * L * level2 * L;
  // workspace: x r ra rb t yl ll | stack (1+nb planes) | yh | wc (3 planes) | work | hh
  const size_t nwork = std::max<size_t>(2 * pl, (size_t)(ncoeff + 2) * pl);
  const size_t total = 7 * pl + (1 + (size_t)nb) * pl + nyh + 3 * pl + nwork + pl / 4;
  float* ws = nullptr;
  double* dmad = nullptr;
  SHX_HIP(hipMalloc(&ws, sizeof(float) * total));
  if (hipMalloc(&dmad, sizeof(double)) != hipSuccess) {
    (void)hipFree(ws);
    return fail(SHX_ERR_ALLOC, "separation workspace allocation failed");
  }
  float *x = ws, *r = ws + pl, *ra = ws + 2 * pl, *rb = ws + 3 * pl, *t = ws + 4 * pl, *yl = ws + 5 * pl,
        *ll = ws + 6 * pl, *stack = ws + 7 * pl;
  float* yh = stack + (1 + (size_t)nb) * pl;
  float* wc = yh + nyh;
  float* work = wc + 3 * pl;
  float* hh = work + nwork;
  std::vector<float*> cells(S + 1);
  std::vector<int> cplanes(S + 1);
  std::vector<int64_t> celems(S + 1, (int64_t)pl);
  cells[0] = stack;
  cplanes[0] = 1;
  for (int j = 0; j < S; ++j) {
    cells[j + 1] = stack + (1 + (size_t)d->map.first[j]) * pl;
    cplanes[j + 1] = d->ndir[j];
  }
  float* wcells[2] = {yl, yh};
  int wplanes[2] = {1, 1};
  int64_t welems[2] = {(int64_t)pl, (int64_t)nyh};
  const std::vector<double> ones(S + 1, 1.0);
  const double zero = 0.0;
  const double haar_raw[2] = {1.0, 1.0}, d4_raw[4] = {.482962913145, .836516303738, .224143868042, -.129409522551};
  const std::vector<double> haar = on_filter(haar_raw, 2), d4 = on_filter(d4_raw, 4);
  std::vector<double> Eh;
  int rc = SHX_OK;
  auto ok = [&](int v) { return (rc = v) == SHX_OK; };
  const int eb = (int)std::min<size_t>(4096, (pl + 255) / 256);
  double sigma = 0.0, dmax = 0.0, lam = 0.0;
  do {
    // band norms (com_norm.m:38-53): transform of an impulse of height L at the centre
    if (!E) {
      Eh.assign((size_t)(S + 1) * maxd, 0.0);
      std::vector<float> imp(pl, 0.f);
      imp[(size_t)(L / 2) * L + L / 2] = (float)L;
      if (hipMemcpyAsync(t, imp.data(), sizeof(float) * pl, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(SHX_ERR_HIP, "impulse upload failed");
        break;
      }
      if (!ok(shx_shear_trans(d, t, cells.data(), stream))) break;
      // the stack is contiguous: lowpass, then the bands scale by scale
      double* dsq = reinterpret_cast<double*>(wc);  // 3 planes of floats hold 1 + nb doubles
      hipLaunchKernelGGL(k_plane_sumsq, dim3(1 + nb), dim3(256), 0, s, stack, (long long)pl, dsq);
      std::vector<double> sq(1 + (size_t)nb);
      if (hipMemcpyAsync(sq.data(), dsq, sizeof(double) * sq.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(SHX_ERR_HIP, "band norm download failed");
        break;
      }
      for (int c = 0, k = 0; c <= S; ++c)
        for (int w = 0; w < cplanes[c]; ++w, ++k) Eh[c + (size_t)w * (S + 1)] = std::sqrt(sq[k]) / std::sqrt((double)pl);
      E = Eh.data();
      E_cols = maxd;
    }
    if (!ok(shx_re_weight_band(img, L, coeff, ncoeff, x, work, stream))) break;     // separate.m:74
    if (delta_max > 0) {                                                              // separateHairlineCrack.m:63-66
      sigma = 2.5;
      dmax = delta_max;
    } else {
      if (!ok(shx_mrdwt(x, L, L, d4.data(), 4, 1, t, yh, work, 1, stream))) break;   // separate.m:81-101
      hipLaunchKernelGGL(k_even_hh, dim3((unsigned)std::min<size_t>(1024, (pl / 4 + 255) / 256)), dim3(256), 0, s,
                         yh, L, hh);
      if (!ok(shx_band_mad(hh, (long long)(pl / 4), (long long)(pl / 4), 1, SHX_MAD_FP64, dmad, stream))) break;
      if (hipMemcpyAsync(&sigma, dmad, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = fail(SHX_ERR_HIP, "noise level download failed");
        break;
      }
      double b1 = 0.0, b2 = 0.0;
      if (!ok(shx_shear_trans(d, x, cells.data(), stream))) break;                   // :116-117
      if (!ok(shx_thresh(cells.data(), S + 1, cplanes.data(), celems.data(), 0.0, 2, E, S + 1, E_cols, ones.data(),
                         S + 1, 1, &b1, stream)))
        break;
      if (!ok(shx_mrdwt(x, L, L, qmf, lqmf, level2, yl, yh, work, 1, stream))) break; // :127-131
      if (!ok(shx_thresh(wcells, 2, wplanes, welems, 0.0, 1, E, S + 1, E_cols, &zero, 1, 1, &b2, stream))) break;
      dmax = std::min(b1, b2);                                                        // :133
    }
    const double stop_t = stop * sigma;
    lam = std::pow(dmax / stop_t, 1.0 / (1 - itermax));                              // :135
    double delta = dmax;
    if (hipMemsetAsync(C, 0, sizeof(float) * pl, s) != hipSuccess ||
        hipMemsetAsync(P, 0, sizeof(float) * pl, s) != hipSuccess) {
      rc = fail(SHX_ERR_HIP, "clearing the parts failed");
      break;
    }
    for (int it = 0; it < itermax && rc == SHX_OK; ++it) {
      hipLaunchKernelGGL(k_sep_residual, dim3(eb), dim3(256), 0, s, x, C, P, r, ra, rb, (long long)pl);  // :164-172
      if (!ok(shx_shear_trans(d, ra, cells.data(), stream))) break;                  // :175-177
      if (!ok(shx_thresh(cells.data(), S + 1, cplanes.data(), celems.data(), 1.4 * delta, 2, E, S + 1, E_cols,
                         ones.data(), S + 1, 0, nullptr, stream)))
        break;
      if (!ok(shx_inverse_shear(d, cells.data(), t, stream))) break;                 // :187
      if (!ok(shx_mrdwt(t, L, L, haar.data(), 2, 1, ll, wc, work, 1, stream))) break; // :189, TVCorrection_cuda
      if (!ok(shx_softth(wc, (int64_t)(3 * pl), gamma, stream))) break;
      if (!ok(shx_mirdwt(ll, wc, L, L, haar.data(), 2, 1, C, work, 1, stream))) break;
      if (!ok(shx_mrdwt(rb, L, L, qmf, lqmf, level2, yl, yh, work, 1, stream))) break; // :206-212
      if (!ok(shx_thresh(wcells, 2, wplanes, welems, 1.4 * delta, 1, E, S + 1, E_cols, &zero, 1, 0, nullptr, stream)))
        break;
      if (!ok(shx_mirdwt(yl, yh, L, L, qmf, lqmf, level2, P, work, 1, stream))) break;
      delta *= lam;                                                                    // :252
    }
  } while (false);
  if (rc == SHX_OK && hipGetLastError() != hipSuccess) rc = fail(SHX_ERR_HIP, "separation kernel launch failed");
  (void)hipStreamSynchronize(s);
  (void)hipFree(ws);
  (void)hipFree(dmad);
  if (rc == SHX_OK && info) {
    info[0] = sigma;
    info[1] = dmax;
    info[2] = lam;
  }
  return rc;
}
