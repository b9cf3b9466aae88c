// C ABI of the 3D shearlet path (shx3_* in include/shx.h): the Fourier-domain
// band-pass pyramid, the directional cone filtering with its fused hard
// threshold and sum, the materialised analysis and its transpose.
//
// Pipeline of one level l (band-pass subband of side n, FFT grid N >= n + fs/2,
// fs the filter side):
//   S       = FFT_N(zero-pad(sub_l))                     pencil / cols / rows
//   for each group of G bands:
//     T_g   = IFFT_rows(F_g . S / N^3), window [fs/2, fs/2 + n)
//     U_g   = IFFT_cols(T_g), same window
//     rec_l (+)= sum_g keep(IFFT_pencil(U_g), th_g), same window
// which is the reference's per-band pad / R2C / modulateAndNormalize3D / C2R /
// crop / applyHardThreshold / addVector loop
// (shearcuda3d/3DShearTrans/src_cuda/ShDecThRec_cuda.cpp:228-325) on a grid of
// n + 12 (rounded to a supported FFT size) instead of nextpow2(n + 23), with
// the filter spectra computed once per dictionary instead of once per band
// and call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "api_err.h"
#include "dict_host.h"
#include "spectral3d.h"

using shx::BpMask;
using shx::fail;

namespace {

constexpr int kSizes[] = {8, 16, 24, 32, 48, 64, 80, 128, 144};

int fft_grid(int need) {
  for (int s : kSizes)
    if (s >= need) return s;
  return 0;
}

long long spec_elems(int N) { return (long long)(N / 2 + 1) * N * N; }

// Band-pass parameters (w, tbw, D) of level l (shearcuda3d/3DBP/PyrNDDec_mm.m:52-70).
bool pyr_params(double mode, int L, int l, double* w, double* tbw, double* D) {
  if (mode == 1.0) {
    *w = l == 0 ? 0.5 : 0.25;
    *tbw = l == 0 ? 1.0 / 6.0 : 1.0 / 12.0;
    *D = l == 0 ? 1.0 : 2.0;
  } else if (mode == 1.5) {
    *w = l == 0 ? 0.5 : 3.0 / 8.0;
    *tbw = l == 0 ? 1.0 / 7.0 : 1.0 / 9.0;
    *D = l == 0 ? 1.5 : 2.0;
  } else if (mode == 2.0) {
    *w = 1.0 / 3.0;
    *tbw = 1.0 / 7.0;
    *D = 2.0;
  } else {
    return false;
  }
  (void)L;
  return true;
}

// 1D passband factor on a grid of side nr (3DBP/PrySDdec_onestep.m:73-79).
BpMask passband(int nr, double w, double tbw) {
  BpMask m{};
  const double hi = (double)nr / 2 * (w + tbw), lo = (double)nr / 2 * (w - tbw);
  m.P = (int)std::ceil(hi);
  const double den = std::ceil(hi) - std::floor(lo);
  for (int k = 0; k <= m.P && k <= shx::BP_MAX_P; ++k) {
    const double x = (m.P - k) / den;
    double t = 0.5 * (1.0 - std::cos(M_PI * x));
    if (x <= 0) t = 0.0;
    if (x >= 1) t = 1.0;
    m.v[k] = std::sqrt(t);
  }
  return m;
}

}  // namespace

struct Level3 {
  int n = 0;      // band-pass subband side
  int N = 0;      // FFT grid of the directional filtering
  int M = 0;      // lowpass side after this level
  int ndir = 0;   // directions per axis
  int nb = 0;     // bands (3 cones x ndir^2)
  int band0 = 0;  // global index of the first band
  double D = 1;
  BpMask mask{};
  float2* F = nullptr;  // nb filter spectra, each spec_elems(N)
  long long sub_off = 0, band_off = 0;
};

struct shx3_dict {
  int n = 0, L = 0, fs = 0, device = -1, nbands = 0, n_low = 0;
  double mode = 1;
  Level3 lv[SHX3_MAX_LEVELS];
  long long sub_low = 0, sub_total = 0, band_total = 0;
  std::vector<float2*> spectra;
  std::map<int, float2*> tw;
  size_t bytes = 0;
  const float2* twid(int N) const { return tw.at(N); }
};

struct shx3_plan {
  const shx3_dict* d = nullptr;
  int nvol = 0, G = 1, Gmax = 1;
  // fused plane kernel columns per block (0 auto, 32, 64) or -1: separate row
  // and column passes, the default (measured faster at 128^3: the fused form
  // recomputes row FFTs per column chunk and runs at 2 waves per SIMD)
  int plane_cols = -1;
  float2 *X = nullptr, *Hs = nullptr, *Lp0 = nullptr, *Lp1 = nullptr, *S = nullptr, *tA = nullptr, *tB = nullptr;
  long long sX = 0, sS = 0, sA = 0, sB = 0;
  float *subs = nullptr, *rec = nullptr;
  size_t bytes = 0;
};

namespace {

int check_dev(const shx3_dict* d) {
  if (!d) return fail(SHX_ERR_ARG, "null 3D dictionary");
  int dev = -1;
  SHX_HIP(hipGetDevice(&dev));
  if (dev != d->device) return fail(SHX_ERR_ARG, "3D dictionary belongs to device %d, current is %d", d->device, dev);
  return SHX_OK;
}

#define SHX_LAUNCH(expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(SHX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Real volume of side n (placed at FFT index o) -> half spectrum on grid N.
// Uses tA, tB.  G volumes-or-bands per batch entry: in_gs / out_gs strides.
int fwd3(const shx3_dict* d, shx3_plan* p, const float* x, long long xs, int n, int N, int o, float2* out,
         long long os, int nb, hipStream_t s) {
  const int H = N / 2 + 1;
  const float2* tw = d->twid(N);
  shx::Pencil3Fwd pf{x, 0, xs, n, o, p->tA, 0, p->sA, n * n};
  SHX_LAUNCH(shx::launch3_pencil_fwd(N, pf, 1, nb, tw, s));
  shx::Cols3 c{p->tA, (long long)n * n, p->sA, n, o, p->tB, (long long)N * n, p->sB, N, 0, n, H};
  SHX_LAUNCH(shx::launch3_cols(N, -1, c, nb, tw, s));
  shx::Rows3 r{p->tB, n, p->sB, n, o, out, N, os, N, 0, H * N};
  SHX_LAUNCH(shx::launch3_rows(N, -1, r, nb, tw, s));
  return SHX_OK;
}

// Half spectrum on grid N -> real window [o, o + n)^3 (unnormalised).
int inv3(const shx3_dict* d, shx3_plan* p, const float2* in, long long is, int N, int n, int o, float* out,
         long long os, int nb, hipStream_t s) {
  const int H = N / 2 + 1;
  const float2* tw = d->twid(N);
  shx::Rows3 r{in, N, is, N, 0, p->tA, n, p->sA, n, o, H * N};
  SHX_LAUNCH(shx::launch3_rows(N, +1, r, nb, tw, s));
  shx::Cols3 c{p->tA, (long long)N * n, p->sA, N, 0, p->tB, (long long)n * n, p->sB, n, o, n, H};
  SHX_LAUNCH(shx::launch3_cols(N, +1, c, nb, tw, s));
  shx::Pencil3Inv pi{p->tB, 0, p->sB, out, 0, os, n, o, n * n, 1, nullptr, 0, 0, 0, 1};
  SHX_LAUNCH(shx::launch3_pencil_inv(N, pi, nb, tw, s));
  return SHX_OK;
}

int pyr_dec(shx3_plan* p, const float* x, float* subs, hipStream_t s) {
  const shx3_dict* d = p->d;
  const int nv = p->nvol;
  int rc = fwd3(d, p, x, (long long)d->n * d->n * d->n, d->n, d->n, 0, p->X, p->sX, nv, s);
  if (rc) return rc;
  const float2* cur = p->X;
  for (int l = 0; l < d->L; ++l) {
    const Level3& v = d->lv[l];
    float2* lp = (cur == p->Lp0) ? p->Lp1 : p->Lp0;
    const bool last = l == d->L - 1;
    const double M3 = (double)v.M * v.M * v.M, n3 = (double)v.n * v.n * v.n;
    const double lscale = std::pow(v.D, -1.5) / (last ? M3 : 1.0);
    SHX_LAUNCH(shx::launch3_bp_split(cur, p->sX, p->Hs, p->sX, lp, p->sX, v.n, v.M, v.mask, 1.0 / n3, lscale, nv, s));
    if ((rc = inv3(d, p, p->Hs, p->sX, v.n, v.n, 0, subs + v.sub_off, d->sub_total, nv, s))) return rc;
    cur = lp;
  }
  return inv3(d, p, cur, p->sX, d->n_low, d->n_low, 0, subs + d->sub_low, d->sub_total, nv, s);
}

int pyr_rec(shx3_plan* p, const float* subs, float* x, hipStream_t s) {
  const shx3_dict* d = p->d;
  const int nv = p->nvol;
  int rc = fwd3(d, p, subs + d->sub_low, d->sub_total, d->n_low, d->n_low, 0, p->Lp0, p->sX, nv, s);
  if (rc) return rc;
  float2* cur = p->Lp0;
  for (int l = d->L - 1; l >= 0; --l) {
    const Level3& v = d->lv[l];
    if ((rc = fwd3(d, p, subs + v.sub_off, d->sub_total, v.n, v.n, 0, p->Hs, p->sX, nv, s))) return rc;
    float2* nxt = (cur == p->Lp0) ? p->Lp1 : p->Lp0;
    const double fin = l == 0 ? 1.0 / ((double)v.n * v.n * v.n) : 1.0;
    SHX_LAUNCH(shx::launch3_bp_merge(p->Hs, p->sX, cur, p->sX, nxt, p->sX, v.n, v.M, v.mask, fin,
                                     std::pow(v.D, 1.5) * fin, nv, s));
    cur = nxt;
  }
  return inv3(d, p, cur, p->sX, d->n, d->n, 0, x, (long long)d->n * d->n * d->n, nv, s);
}

// Directional filtering of every level.  mode 0: store each band to
// bands (band layout); mode 1: thresholded sum into rec (subband layout).
int directional(shx3_plan* p, const float* subs, const float* th, float* bands, float* rec, int mode,
                hipStream_t s) {
  const shx3_dict* d = p->d;
  const int nv = p->nvol, o = d->fs / 2;
  for (int l = 0; l < d->L; ++l) {
    const Level3& v = d->lv[l];
    const int n = v.n, N = v.N, H = N / 2 + 1;
    const float2* tw = d->twid(N);
    int rc = fwd3(d, p, subs + v.sub_off, d->sub_total, n, N, 0, p->S, p->sS, nv, s);
    if (rc) return rc;
    const long long fstr = spec_elems(N), n3 = (long long)n * n * n;
    for (int g0 = 0; g0 < v.nb; g0 += p->G) {
      const int G = std::min(p->G, v.nb - g0);
      const float scale = (float)(1.0 / ((double)N * N * N));
      if (p->plane_cols >= 0 && shx::plane3_supported(N, n)) {
        // rows and cols of each plane in one kernel, straight into tA
        shx::Plane3Inv pi{p->S, p->sS, v.F + (size_t)g0 * fstr, fstr, scale, p->tA, (long long)H * n * n, p->sA,
                          n, o, G, H};
        SHX_LAUNCH(shx::launch3_plane_inv(N, pi, nv, p->plane_cols, tw, s));
      } else {
        shx::Rows3Mul rm{p->S, p->sS, v.F + (size_t)g0 * fstr, fstr, scale, p->tB, n, (long long)H * N * n, p->sB, n,
                         o, H * N, G};
        SHX_LAUNCH(shx::launch3_rows_mul(N, rm, nv, tw, s));
        shx::Cols3 c{p->tB, (long long)N * n, p->sB, N, 0, p->tA, (long long)n * n, p->sA, n, o, n, G * H};
        SHX_LAUNCH(shx::launch3_cols(N, +1, c, nv, tw, s));
      }
      shx::Pencil3Inv pi{};
      pi.in = p->tA;
      pi.in_gs = (long long)H * n * n;
      pi.in_bs = p->sA;
      pi.n_out = n;
      pi.o_out = o;
      pi.E = n * n;
      pi.G = G;
      if (mode == 0) {
        pi.out = bands + v.band_off + (size_t)g0 * n3;
        pi.out_gs = n3;
        pi.out_bs = d->band_total;
        pi.mode = 0;
      } else {
        pi.out = rec + v.sub_off;
        pi.out_bs = d->sub_total;
        pi.th = th;
        pi.th_bs = d->nbands;
        pi.th0 = v.band0 + g0;
        pi.mode = 1;
        pi.first = g0 == 0;
      }
      SHX_LAUNCH(shx::launch3_pencil_inv(N, pi, nv, tw, s));
    }
  }
  return SHX_OK;
}

// Transpose of the directional filtering: sub_bar_l = sum_b corr(bands_b, F_b).
int directional_adjoint(shx3_plan* p, const float* bands, float* subs_bar, hipStream_t s) {
  const shx3_dict* d = p->d;
  const int nv = p->nvol, o = d->fs / 2;
  for (int l = 0; l < d->L; ++l) {
    const Level3& v = d->lv[l];
    const int n = v.n, N = v.N, H = N / 2 + 1;
    const float2* tw = d->twid(N);
    const long long fstr = spec_elems(N), n3 = (long long)n * n * n;
    for (int g0 = 0; g0 < v.nb; g0 += p->G) {
      const int G = std::min(p->G, v.nb - g0);
      shx::Pencil3Fwd pf{bands + v.band_off + (size_t)g0 * n3, n3, d->band_total, n, o, p->tA,
                         (long long)H * n * n, p->sA, n * n};
      SHX_LAUNCH(shx::launch3_pencil_fwd(N, pf, G, nv, tw, s));
      shx::Cols3 c{p->tA, (long long)n * n, p->sA, n, o, p->tB, (long long)N * n, p->sB, N, 0, n, G * H};
      SHX_LAUNCH(shx::launch3_cols(N, -1, c, nv, tw, s));
      shx::Rows3Cmac rc{p->tB, n, (long long)H * N * n, p->sB, n, o, v.F + (size_t)g0 * fstr, fstr,
                        (float)(1.0 / ((double)N * N * N)), p->S, p->sS, H * N, G, g0 > 0};
      SHX_LAUNCH(shx::launch3_rows_cmac(N, rc, nv, tw, s));
    }
    int rc = inv3(d, p, p->S, p->sS, N, n, 0, subs_bar + v.sub_off, d->sub_total, nv, s);
    if (rc) return rc;
  }
  return SHX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int shx3_prepare(int n, int nlevels, double pyr_mode, const int* ndir, int fsize, const float* const* filters,
                 shx3_dict** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (nlevels < 1 || nlevels > SHX3_MAX_LEVELS) return fail(SHX_ERR_ARG, "nlevels %d not in [1, %d]", nlevels, SHX3_MAX_LEVELS);
  if (!ndir || !filters) return fail(SHX_ERR_ARG, "null direction counts or filters");
  if (fsize < 2 || fsize > 64 || (fsize * fsize / 2) % 8 != 0 || fsize % 8 != 0)
    return fail(SHX_ERR_UNSUPPORTED, "filter side %d not supported (multiple of 8, at most 64)", fsize);
  auto d = new shx3_dict();
  d->n = n;
  d->L = nlevels;
  d->fs = fsize;
  d->mode = pyr_mode;
  hipGetDevice(&d->device);
  double side = n;
  long long off = 0, boff = 0;
  for (int l = 0; l < nlevels; ++l) {
    Level3& v = d->lv[l];
    double w, tbw, D;
    if (!pyr_params(pyr_mode, nlevels, l, &w, &tbw, &D)) {
      delete d;
      return fail(SHX_ERR_ARG, "unsupported Pyr_mode %g (1, 1.5 or 2)", pyr_mode);
    }
    const double M = side / D;
    if (side != std::floor(side) || M != std::floor(M)) {
      delete d;
      return fail(SHX_ERR_ARG,
                  "The downsampling factor must be able to divide the size of the FFT matrix (level %d: %g / %g)",
                  l + 1, side, D);
    }
    v.n = (int)side;
    v.M = (int)M;
    v.D = D;
    v.mask = passband(v.n, w, tbw);
    if (v.mask.P > shx::BP_MAX_P) {
      delete d;
      return fail(SHX_ERR_UNSUPPORTED, "volume side %d too large", n);
    }
    v.ndir = ndir[l];
    if (v.ndir < 1 || fsize / v.ndir < 1) {
      delete d;
      return fail(SHX_ERR_ARG, "level %d: direction count %d invalid for filter side %d", l + 1, v.ndir, fsize);
    }
    v.nb = 3 * v.ndir * v.ndir;
    v.band0 = d->nbands;
    d->nbands += v.nb;
    v.N = fft_grid(std::max(v.n + fsize / 2, v.n + fsize - 1 - fsize / 2));
    if (!shx::fft3_supported(v.n) || !v.N || v.n % 8 != 0) {
      delete d;
      return fail(SHX_ERR_UNSUPPORTED, "level %d side %d (FFT grid %d) not supported", l + 1, v.n, v.N);
    }
    if (!filters[l]) {
      delete d;
      return fail(SHX_ERR_ARG, "filters of level %d are null", l + 1);
    }
    v.sub_off = off;
    off += (long long)v.n * v.n * v.n;
    v.band_off = boff;
    boff += (long long)v.nb * v.n * v.n * v.n;
    side = M;
  }
  d->n_low = (int)side;
  if (!shx::fft3_supported(d->n_low) || d->n_low % 8 != 0) {
    int nl = d->n_low;
    delete d;
    return fail(SHX_ERR_UNSUPPORTED, "lowpass side %d not supported (at least 8)", nl);
  }
  d->sub_low = off;
  d->sub_total = off + (long long)d->n_low * d->n_low * d->n_low;
  d->band_total = boff;

  // twiddle tables for every FFT size in use
  std::vector<int> sizes;
  for (int l = 0; l < nlevels; ++l) {
    sizes.push_back(d->lv[l].n);
    sizes.push_back(d->lv[l].N);
  }
  sizes.push_back(d->n_low);
  for (int N : sizes) {
    if (d->tw.count(N)) continue;
    std::vector<float> t(2 * N);
    shx::twiddles_fp32(N, t.data());
    float2* p = nullptr;
    if (hipMalloc(&p, sizeof(float2) * N) != hipSuccess ||
        hipMemcpy(p, t.data(), sizeof(float2) * N, hipMemcpyHostToDevice) != hipSuccess) {
      if (p) hipFree(p);
      shx3_destroy(d);
      return fail(SHX_ERR_ALLOC, "twiddle table allocation failed");
    }
    d->tw[N] = p;
    d->bytes += sizeof(float2) * N;
  }

  // filter spectra: FFT_N of each spatial filter (placed at the origin),
  // shared by levels with the same direction count and grid
  std::map<std::pair<int, int>, std::pair<float2*, const float*>> made;
  const int CH = 16;
  for (int l = 0; l < nlevels; ++l) {
    Level3& v = d->lv[l];
    auto key = std::make_pair(v.ndir, v.N);
    if (made.count(key) && made[key].second == filters[l]) {
      v.F = made[key].first;
      continue;
    }
    const long long fstr = spec_elems(v.N), f3 = (long long)fsize * fsize * fsize;
    const int H = v.N / 2 + 1;
    if (hipMalloc(&v.F, sizeof(float2) * fstr * v.nb) != hipSuccess) {
      v.F = nullptr;
      shx3_destroy(d);
      return fail(SHX_ERR_ALLOC, "filter spectra allocation (%d bands of %d^3) failed", v.nb, v.N);
    }
    d->spectra.push_back(v.F);
    d->bytes += sizeof(float2) * fstr * v.nb;
    made[key] = std::make_pair(v.F, filters[l]);
    float* fdev = nullptr;
    shx3_plan tmp{};
    tmp.sA = (long long)H * fsize * fsize;
    tmp.sB = (long long)H * v.N * fsize;
    if (hipMalloc(&fdev, sizeof(float) * f3 * CH) != hipSuccess ||
        hipMalloc(&tmp.tA, sizeof(float2) * tmp.sA * CH) != hipSuccess ||
        hipMalloc(&tmp.tB, sizeof(float2) * tmp.sB * CH) != hipSuccess) {
      hipFree(fdev);
      hipFree(tmp.tA);
      hipFree(tmp.tB);
      shx3_destroy(d);
      return fail(SHX_ERR_ALLOC, "filter staging allocation failed");
    }
    int rc = SHX_OK;
    for (int b0 = 0; b0 < v.nb && rc == SHX_OK; b0 += CH) {
      const int cnt = std::min(CH, v.nb - b0);
      if (hipMemcpy(fdev, filters[l] + (size_t)b0 * f3, sizeof(float) * f3 * cnt, hipMemcpyHostToDevice) !=
          hipSuccess) {
        rc = fail(SHX_ERR_HIP, "filter upload failed");
        break;
      }
      rc = fwd3(d, &tmp, fdev, f3, fsize, v.N, 0, v.F + (size_t)b0 * fstr, fstr, cnt, nullptr);
    }
    if (rc == SHX_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(SHX_ERR_HIP, "filter spectra failed");
    hipFree(fdev);
    hipFree(tmp.tA);
    hipFree(tmp.tB);
    if (rc) {
      shx3_destroy(d);
      return rc;
    }
  }
  *out = d;
  return SHX_OK;
}

int shx3_destroy(shx3_dict* d) {
  if (!d) return SHX_OK;
  for (float2* p : d->spectra) hipFree(p);
  for (auto& kv : d->tw) hipFree(kv.second);
  delete d;
  return SHX_OK;
}

int shx3_info(const shx3_dict* d, int* sides, int* fft_sides, int* band_first) {
  if (!d) return fail(SHX_ERR_ARG, "null 3D dictionary");
  for (int l = 0; l < d->L; ++l) {
    if (sides) sides[l] = d->lv[l].n;
    if (fft_sides) fft_sides[l] = d->lv[l].N;
    if (band_first) band_first[l] = d->lv[l].band0;
  }
  if (sides) sides[d->L] = d->n_low;
  if (band_first) band_first[d->L] = d->nbands;
  return SHX_OK;
}

long long shx3_subs_elems(const shx3_dict* d) { return d ? d->sub_total : 0; }
long long shx3_bands_elems(const shx3_dict* d) { return d ? d->band_total : 0; }
size_t shx3_dict_bytes(const shx3_dict* d) { return d ? d->bytes : 0; }

int shx3_plan_create(const shx3_dict* d, int nvol, int group, shx3_plan** out) {
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
  *out = nullptr;
  int rc = check_dev(d);
  if (rc) return rc;
  if (nvol < 1 || nvol > 65535) return fail(SHX_ERR_ARG, "nvol %d out of range", nvol);
  if (group < 1 || group > 64) return fail(SHX_ERR_ARG, "band group %d not in [1, 64]", group);
  auto p = new shx3_plan();
  p->d = d;
  p->nvol = nvol;
  p->G = p->Gmax = group;
  long long sA = spec_elems(d->n), sB = spec_elems(d->n), sS = 0;
  for (int l = 0; l < d->L; ++l) {
    const Level3& v = d->lv[l];
    const long long H = v.N / 2 + 1;
    sA = std::max(sA, group * H * v.n * v.n);
    sB = std::max(sB, group * H * v.N * v.n);
    sS = std::max(sS, spec_elems(v.N));
  }
  p->sX = spec_elems(d->n);
  p->sA = sA;
  p->sB = sB;
  p->sS = sS;
  struct A {
    void** ptr;
    size_t bytes;
  } allocs[] = {
      {(void**)&p->X, sizeof(float2) * p->sX},   {(void**)&p->Hs, sizeof(float2) * p->sX},
      {(void**)&p->Lp0, sizeof(float2) * p->sX}, {(void**)&p->Lp1, sizeof(float2) * p->sX},
      {(void**)&p->S, sizeof(float2) * p->sS},   {(void**)&p->tA, sizeof(float2) * p->sA},
      {(void**)&p->tB, sizeof(float2) * p->sB},  {(void**)&p->subs, sizeof(float) * d->sub_total},
      {(void**)&p->rec, sizeof(float) * d->sub_total},
  };
  for (auto& a : allocs) {
    const size_t b = a.bytes * nvol;
    if (hipMalloc(a.ptr, b) != hipSuccess) {
      *a.ptr = nullptr;
      shx3_plan_destroy(p);
      return fail(SHX_ERR_ALLOC, "3D workspace allocation (%zu bytes) failed", b);
    }
    p->bytes += b;
  }
  *out = p;
  return SHX_OK;
}

int shx3_plan_destroy(shx3_plan* p) {
  if (!p) return SHX_OK;
  void* ptrs[] = {p->X, p->Hs, p->Lp0, p->Lp1, p->S, p->tA, p->tB, p->subs, p->rec};
  for (void* q : ptrs)
    if (q) hipFree(q);
  delete p;
  return SHX_OK;
}

size_t shx3_plan_bytes(const shx3_plan* p) { return p ? p->bytes : 0; }

int shx3_set_schedule(shx3_plan* p, int group, int plane_cols) {
  if (!p) return fail(SHX_ERR_ARG, "null 3D plan");
  if (group < 1 || group > p->Gmax)
    return fail(SHX_ERR_ARG, "band group %d not in [1, %d] (the plan's workspace size)", group, p->Gmax);
  if (plane_cols != -1 && plane_cols != 0 && plane_cols != 32 && plane_cols != 64)
    return fail(SHX_ERR_ARG, "plane columns %d not in {-1, 0, 32, 64}", plane_cols);
  p->G = group;
  p->plane_cols = plane_cols;
  return SHX_OK;
}

int shx3_pyr_dec(shx3_plan* p, const float* x, float* subs, void* stream) {
  if (!p || !x || !subs) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  return rc ? rc : pyr_dec(p, x, subs, (hipStream_t)stream);
}

int shx3_pyr_rec(shx3_plan* p, const float* subs, float* x, void* stream) {
  if (!p || !x || !subs) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  return rc ? rc : pyr_rec(p, subs, x, (hipStream_t)stream);
}

int shx3_dec_thresh_rec(shx3_plan* p, const float* subs, const float* th, float* rec, void* stream) {
  if (!p || !subs || !th || !rec) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  return rc ? rc : directional(p, subs, th, nullptr, rec, 1, (hipStream_t)stream);
}

int shx3_analysis(shx3_plan* p, const float* subs, float* bands, void* stream) {
  if (!p || !subs || !bands) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  return rc ? rc : directional(p, subs, nullptr, bands, nullptr, 0, (hipStream_t)stream);
}

int shx3_analysis_adjoint(shx3_plan* p, const float* bands, float* subs_bar, void* stream) {
  if (!p || !bands || !subs_bar) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  return rc ? rc : directional_adjoint(p, bands, subs_bar, (hipStream_t)stream);
}

int shx3_denoise(shx3_plan* p, const float* x, const float* th, float* out, void* stream) {
  if (!p || !x || !th || !out) return fail(SHX_ERR_ARG, "null argument");
  int rc = check_dev(p->d);
  if (rc) return rc;
  const shx3_dict* d = p->d;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = pyr_dec(p, x, p->subs, s))) return rc;
  if ((rc = directional(p, p->subs, th, nullptr, p->rec, 1, s))) return rc;
  const size_t low = sizeof(float) * d->n_low * d->n_low * d->n_low, pitch = sizeof(float) * d->sub_total;
  SHX_HIP(hipMemcpy2DAsync(p->rec + d->sub_low, pitch, p->subs + d->sub_low, pitch, low, p->nvol,
                           hipMemcpyDeviceToDevice, s));
  return pyr_rec(p, p->rec, out, s);
}

}  // extern "C"
