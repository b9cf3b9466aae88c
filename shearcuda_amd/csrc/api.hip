// C ABI of the shx library (include/shx.h): dictionary management, the NSST
// decompose / threshold / reconstruct entry points, and their helpers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/shx.h"
#include "dict_host.h"
#include "pyramid.h"
#include "spectral.h"

using shx::BandMap;
using shx::BandThresh;
using shx::PlaneSet;

// ---------------------------------------------------------------------------
// errors
namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define SHX_HIP(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(SHX_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
}  // namespace

struct shx_dict {
  int L = 0, S = 0, nbands = 0, shift = 0, device = -1;
  int ndir[SHX_MAX_SCALES] = {};
  int dsize[SHX_MAX_SCALES] = {};
  BandMap map{};
  float2* Hmain = nullptr;
  float2* Hbeta = nullptr;
  float2* tw = nullptr;
  float* planes = nullptr;  // S band planes (ypyr) + 2 pyramid temporaries
  float2* spec = nullptr;   // S packed spectra
  float2* band = nullptr;   // nbands packed spectra
  size_t bytes = 0;
  size_t plane() const { return (size_t)L * L; }
  float* ypyr(int j) const { return planes + (size_t)(j - 1) * plane(); }  // j = 1..S
  float* tmpA() const { return planes + (size_t)S * plane(); }
  float* tmpB() const { return planes + (size_t)(S + 1) * plane(); }
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

__global__ void k_hard_thresh(ThreshTable t, int opt, unsigned int* maxbits) {
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
    m *= t.invE[e];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(m));
  }
}

__global__ void k_softth(float* x, long long n, float t) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    const float a = fabsf(v) - t;
    x[i] = a > 0.f ? copysignf(a, v) : 0.f;
  }
}

__global__ void k_norm_pow(const float* x, long long n, float p, double* out) {
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float a = fabsf(x[i]);
    s += (p == 2.f) ? a * a : (p == 1.f ? a : powf(a, p));
  }
  double d = s;
  for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, d);
}

hipError_t pyr_dec(const float* x, int L, int nlev, float* const* y, float* tA, float* tB, hipStream_t s) {
  const float* in = x;
  for (int lev = 0; lev < nlev; ++lev) {
    float* low = (lev == nlev - 1) ? y[0] : ((lev & 1) ? tB : tA);
    hipError_t e = shx::launch_atrous_dec(in, low, y[nlev - lev], L, 1 << lev, shx::dec_filters(), s);
    if (e != hipSuccess) return e;
    in = low;
  }
  return hipSuccess;
}

// w: optional per-band weights (nlev+1), th_low: threshold of y[0] (< 0: none)
hipError_t pyr_rec(const float* const* y, int L, int nlev, float* out, float* tA, float* tB, float th_low,
                   const double* w, hipStream_t s) {
  const float* xl = y[0];
  for (int lev = nlev - 1; lev >= 0; --lev) {
    const int j = nlev - lev;
    float* o = (lev == 0) ? out : ((lev & 1) ? tA : tB);
    const bool first = (lev == nlev - 1);
    const float sx = (first && w) ? (float)w[0] : 1.0f;
    const float sy = w ? (float)w[j] : 1.0f;
    hipError_t e = shx::launch_atrous_rec(xl, y[j], o, L, 1 << lev, shx::rec_filters(), first ? th_low : -1.0f, sx,
                                          sy, s);
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

PlaneSet band_cells(const shx_dict* d, float* const* coeff) {
  PlaneSet p{};
  p.ncells = d->S;
  for (int j = 0; j <= d->S; ++j) p.off[j] = d->map.first[j];
  for (int j = 0; j < d->S; ++j) p.cell[j] = coeff[j + 1];
  return p;
}

// One kernel launch of a path, with the bytes it must move at minimum
// (FP32 inputs + outputs it touches, each once), for the profiler.
struct Stage {
  std::string name;
  double bytes;
  std::function<hipError_t()> run;
};

void stages_dec(shx_dict* d, const float* x, float* const* coeff, hipStream_t s, std::vector<Stage>& st) {
  const int L = d->L, S = d->S;
  const double P = (double)L * L * 4.0;  // bytes per real plane
  const float* in = x;
  for (int lev = 0; lev < S; ++lev) {
    float* low = (lev == S - 1) ? coeff[0] : ((lev & 1) ? d->tmpB() : d->tmpA());
    float* high = d->ypyr(S - lev);
    const int m = 1 << lev;
    st.push_back({"atrous_dec_m" + std::to_string(m), 3 * P,
                  [=] { return shx::launch_atrous_dec(in, low, high, L, m, shx::dec_filters(), s); }});
    in = low;
  }
  PlaneSet yset = shx::single_cell(d->planes, S);
  PlaneSet specset = shx::single_cell(reinterpret_cast<float*>(d->spec), S);
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(d->band), d->nbands);
  PlaneSet out = band_cells(d, coeff);
  const int nb = d->nbands;
  st.push_back({"rows_r2c_scales", 2 * S * P, [=] { return shx::launch_rows_r2c(L, yset, specset, S, nullptr, d->tw, s); }});
  st.push_back({"cols_dec", (S + 2.0 * nb) * P,
                [=] { return shx::launch_cols_dec(L, d->spec, d->band, d->Hmain, d->Hbeta, d->map, d->tw, s); }});
  st.push_back({"rows_c2r_bands", 2.0 * nb * P, [=] { return shx::launch_rows_c2r(L, bandset, out, nb, d->tw, s); }});
}

void stages_rec(shx_dict* d, const float* const* coeff, const float* th, float* x, hipStream_t s,
                std::vector<Stage>& st) {
  const int L = d->L, S = d->S, nb = d->nbands;
  const double P = (double)L * L * 4.0;
  PlaneSet in = band_cells(d, const_cast<float* const*>(coeff));
  PlaneSet bandset = shx::single_cell(reinterpret_cast<float*>(d->band), nb);
  PlaneSet specset = shx::single_cell(reinterpret_cast<float*>(d->spec), S);
  PlaneSet yset = shx::single_cell(d->planes, S);
  const bool use_th = th != nullptr;
  BandThresh bt{};
  if (use_th)
    for (int b = 0; b < nb; ++b) bt.v[b] = th[1 + b];
  st.push_back({"rows_r2c_bands", 2.0 * nb * P,
                [=] { return shx::launch_rows_r2c(L, in, bandset, nb, use_th ? &bt : nullptr, d->tw, s); }});
  st.push_back({"cols_rec", (2.0 * nb + S) * P,
                [=] { return shx::launch_cols_rec(L, d->band, d->spec, d->Hmain, d->Hbeta, d->map, d->tw, s); }});
  st.push_back({"rows_c2r_scales", 2 * S * P, [=] { return shx::launch_rows_c2r(L, specset, yset, S, d->tw, s); }});
  const float th_low = use_th ? th[0] : -1.0f;
  const float* xl = coeff[0];
  for (int lev = S - 1; lev >= 0; --lev) {
    const int j = S - lev, m = 1 << lev;
    float* o = (lev == 0) ? x : ((lev & 1) ? d->tmpA() : d->tmpB());
    const float* yh = d->ypyr(j);
    const float t = (lev == S - 1) ? th_low : -1.0f;
    st.push_back({"atrous_rec_m" + std::to_string(m), 3 * P, [=] {
                    return shx::launch_atrous_rec(xl, yh, o, L, m, shx::rec_filters(), t, 1.0f, 1.0f, s);
                  }});
    xl = o;
  }
}

int run_stages(const std::vector<Stage>& st) {
  for (const Stage& g : st) {
    hipError_t e = g.run();
    if (e != hipSuccess) return fail(SHX_ERR_HIP, "%s launch failed: %s", g.name.c_str(), hipGetErrorString(e));
  }
  return SHX_OK;
}

int dict_inverse(shx_dict* d, const float* const* coeff, const float* th, float* x, hipStream_t s) {
  std::vector<Stage> st;
  stages_rec(d, coeff, th, x, s, st);
  return run_stages(st);
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

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* shx_last_error(void) { return g_err.c_str(); }
int shx_version(void) { return 100; }

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
  if (!out) return fail(SHX_ERR_ARG, "null output pointer");
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
  d->nbands = nb;
  if (hipGetDevice(&d->device) != hipSuccess) {
    delete d;
    return fail(SHX_ERR_HIP, "no HIP device");
  }
  const int N = L / 2;
  std::vector<float> hm((size_t)nb * L * N * 2), hb((size_t)nb * L * 2), tw((size_t)L * 2);
  for (int j = 0; j < nscales; ++j) {
    std::vector<std::complex<double>> H;
    shx::scale_spectra_fp64(L, ndir[j], dsize[j], w_s[j], shift, H);
    for (size_t i = 0; i < H.size(); ++i)
      if (!std::isfinite(H[i].real()) || !std::isfinite(H[i].imag())) {
        delete d;
        return fail(SHX_ERR_ARG, "scale %d: filter spectra vanish at some frequency (division by zero)", j);
      }
    shx::pack_spectra_fp32(L, ndir[j], H, hm.data() + (size_t)d->map.first[j] * L * N * 2,
                           hb.data() + (size_t)d->map.first[j] * L * 2);
  }
  shx::twiddles_fp32(L, tw.data());
  const size_t pl = (size_t)L * L * sizeof(float);
  const size_t bytes_hm = hm.size() * sizeof(float), bytes_hb = hb.size() * sizeof(float);
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&d->Hmain, bytes_hm)) != hipSuccess || (e = hipMalloc(&d->Hbeta, bytes_hb)) != hipSuccess ||
      (e = hipMalloc(&d->tw, tw.size() * sizeof(float))) != hipSuccess ||
      (e = hipMalloc(&d->planes, pl * (nscales + 2))) != hipSuccess ||
      (e = hipMalloc(&d->spec, pl * nscales)) != hipSuccess || (e = hipMalloc(&d->band, pl * nb)) != hipSuccess) {
    shx_clear_shearing_filters(d);
    return fail(SHX_ERR_ALLOC, "device allocation failed: %s", hipGetErrorString(e));
  }
  d->bytes = bytes_hm + bytes_hb + tw.size() * sizeof(float) + pl * (nscales + 2) + pl * nscales + pl * nb;
  if ((e = hipMemcpy(d->Hmain, hm.data(), bytes_hm, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d->Hbeta, hb.data(), bytes_hb, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
    shx_clear_shearing_filters(d);
    return fail(SHX_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
  }
  *out = d;
  return SHX_OK;
}

int shx_clear_shearing_filters(shx_dict* d) {
  if (!d) return SHX_OK;
  hipFree(d->Hmain);
  hipFree(d->Hbeta);
  hipFree(d->tw);
  hipFree(d->planes);
  hipFree(d->spec);
  hipFree(d->band);
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
  if (!x || !coeff) return fail(SHX_ERR_ARG, "null input or output");
  for (int j = 0; j <= d->S; ++j)
    if (!coeff[j]) return fail(SHX_ERR_ARG, "coefficient cell %d is null", j);
  std::vector<Stage> st;
  stages_dec(d, x, coeff, (hipStream_t)stream, st);
  return run_stages(st);
}

int shx_inverse_shear(shx_dict* d, const float* const* coeff, float* x, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !coeff) return fail(SHX_ERR_ARG, "null input or output");
  for (int j = 0; j <= d->S; ++j)
    if (!coeff[j]) return fail(SHX_ERR_ARG, "coefficient cell %d is null", j);
  return dict_inverse(d, coeff, nullptr, x, (hipStream_t)stream);
}

int shx_inverse_shear_thresh(shx_dict* d, const float* const* coeff, const float* th, float* x, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !coeff || !th) return fail(SHX_ERR_ARG, "null input, thresholds or output");
  for (int j = 0; j <= d->S; ++j)
    if (!coeff[j]) return fail(SHX_ERR_ARG, "coefficient cell %d is null", j);
  return dict_inverse(d, coeff, th, x, (hipStream_t)stream);
}

int shx_denoise(shx_dict* d, const float* x, float* const* coeff, double sigma, const double* T, const double* sc,
                float* out, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!T || !sc) return fail(SHX_ERR_ARG, "null threshold scalars");
  std::vector<float> th = nsst_ht_thresholds(d, sigma, T, sc);
  if (int r = shx_shear_trans(d, x, coeff, stream)) return r;
  return shx_inverse_shear_thresh(d, coeff, th.data(), out, stream);
}

int shx_denoise_profile(shx_dict* d, const float* x, float* const* coeff, double sigma, const double* T,
                        const double* sc, float* out, int reps, int max_stages, double* stage_ms, double* stage_bytes,
                        char* names, int name_len, int* nstages, void* stream) {
  if (int r = check_dict(d)) return r;
  if (!x || !coeff || !out || !T || !sc || !stage_ms || !nstages || reps < 1)
    return fail(SHX_ERR_ARG, "invalid profiling arguments");
  hipStream_t s = (hipStream_t)stream;
  std::vector<float> th = nsst_ht_thresholds(d, sigma, T, sc);
  std::vector<Stage> st;
  stages_dec(d, x, coeff, s, st);
  stages_rec(d, coeff, th.data(), out, s, st);
  const int n = (int)st.size();
  *nstages = n;
  if (n > max_stages) return fail(SHX_ERR_ARG, "%d stages, room for %d", n, max_stages);
  std::vector<hipEvent_t> ev(n + 1);
  for (auto& e : ev) SHX_HIP(hipEventCreate(&e));
  std::vector<double> acc(n, 0.0);
  int rc = SHX_OK;
  for (int r = 0; r < reps && rc == SHX_OK; ++r) {
    if (hipEventRecord(ev[0], s) != hipSuccess) { rc = fail(SHX_ERR_HIP, "event record failed"); break; }
    for (int i = 0; i < n && rc == SHX_OK; ++i) {
      hipError_t e = st[i].run();
      if (e == hipSuccess) e = hipEventRecord(ev[i + 1], s);
      if (e != hipSuccess) rc = fail(SHX_ERR_HIP, "%s: %s", st[i].name.c_str(), hipGetErrorString(e));
    }
    if (rc != SHX_OK) break;
    if (hipEventSynchronize(ev[n]) != hipSuccess) { rc = fail(SHX_ERR_HIP, "event sync failed"); break; }
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
  unsigned int* dmax = nullptr;
  if (opt == 1) {
    SHX_HIP(hipMalloc(&dmax, sizeof(unsigned int)));
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
    const int blocks = (int)std::min<long long>(1024, (mx + 255) / 256);
    hipLaunchKernelGGL(k_hard_thresh, dim3(blocks, t.n), dim3(256), 0, s, t, opt == 1 ? 1 : 0, dmax);
    SHX_HIP(hipGetLastError());
  }
  if (maxi) *maxi = 0.0;
  if (opt == 1) {
    unsigned int bits = 0;
    SHX_HIP(hipMemcpyAsync(&bits, dmax, sizeof bits, hipMemcpyDeviceToHost, s));
    SHX_HIP(hipStreamSynchronize(s));
    hipFree(dmax);
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

int shx_atrousrec(const float* const* y, int L, int nlevels, float* x, float* work, void* stream) {
  if (!x || !y) return fail(SHX_ERR_ARG, "null input or output");
  if (!is_pow2(L) || L < 32) return fail(SHX_ERR_UNSUPPORTED, "image size %d not supported", L);
  if (nlevels < 1 || nlevels > 9 || L % (1 << (nlevels - 1)) != 0)
    return fail(SHX_ERR_ARG, "invalid number of levels %d", nlevels);
  float* w = work;
  if (!w) SHX_HIP(hipMalloc(&w, sizeof(float) * 2 * L * L));
  hipError_t e = pyr_rec(y, L, nlevels, x, w, w + (size_t)L * L, -1.0f, nullptr, (hipStream_t)stream);
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
  if (e == hipSuccess) e = pyr_rec(y, L, nlev, out, w + ncoeff * pl, w + (ncoeff + 1) * pl, -1.0f, coeff, s);
  if (!work) {
    hipStreamSynchronize(s);
    hipFree(w);
  }
  SHX_HIP(e);
  return SHX_OK;
}

int shx_norm_pow(const float* x, int64_t n, double p, double* out, void* stream) {
  if (!x || !out) return fail(SHX_ERR_ARG, "null input or output");
  hipStream_t s = (hipStream_t)stream;
  double* acc = nullptr;
  SHX_HIP(hipMalloc(&acc, sizeof(double)));
  SHX_HIP(hipMemsetAsync(acc, 0, sizeof(double), s));
  const int blocks = (int)std::min<long long>(2048, (n + 255) / 256);
  if (n > 0) hipLaunchKernelGGL(k_norm_pow, dim3(blocks), dim3(256), 0, s, x, (long long)n, (float)p, acc);
  SHX_HIP(hipGetLastError());
  SHX_HIP(hipMemcpyAsync(out, acc, sizeof(double), hipMemcpyDeviceToHost, s));
  SHX_HIP(hipStreamSynchronize(s));
  hipFree(acc);
  return SHX_OK;
}

int shx_softth(float* x, int64_t n, double t, void* stream) {
  if (!x) return fail(SHX_ERR_ARG, "null input");
  const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
  if (n > 0) hipLaunchKernelGGL(k_softth, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, (long long)n, (float)t);
  SHX_HIP(hipGetLastError());
  return SHX_OK;
}

}  // extern "C"
