// shearcuda_amd.hpp — C++ entry points of the demo_denoise library, on the
// shx C ABI (shx.h).
//
// Mirrors the classes of demo_denoise/shearcuda (shearcuda.h:13-48,
// gpuimage.h, gpucelldata.h, gpushearfilters.h, gpusheardictionary.h) and the
// SHFM0002 filter file of demoImageDenoise/shearconfig.cpp:25-70: the same
// class and method names, argument meaning and bool returns, so a caller of
// the reference's demo library switches by changing its includes and linking
// libshx.so.  Host code only (compile with any C++17 compiler plus the HIP
// runtime headers); all computation runs in libshx.so.
//
// Differences from the reference, by design:
//   * data type gpuFLOAT (FP32) only for transforms; gpuUINT8 for image I/O;
//   * shearlet coefficients are real FP32 planes (the reference's demo keeps
//     complex planes whose imaginary parts are rounding noise of a C2C FFT of
//     real data);
//   * images and filter planes are row-major (element (r, c) at r * width + c);
//     ShearConfig::readFromFile transposes the file's column-major planes;
//   * the a trous filters must be the maxflat filters (the only pyramid
//     filters of the reference's configurations); other filters are rejected;
//   * addNoise and computeShearNorm draw from the library's counter-based
//     Gaussian generator (seeds 1234 and 4321 as in the reference), not cuRAND.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "shx.h"

namespace shearcuda_amd {

enum gpuTYPE { gpuFLOAT = 0, gpuCFLOAT = 1, gpuDOUBLE = 2, gpuCDOUBLE = 3, gpuINT32 = 4, gpuUINT8 = 5, gpuNOTDEF = 20 };
typedef enum gpuTYPE gpuTYPE_t;

inline int elemSize(gpuTYPE_t t) {
  switch (t) {
    case gpuFLOAT: case gpuINT32: return 4;
    case gpuCFLOAT: case gpuDOUBLE: return 8;
    case gpuCDOUBLE: return 16;
    case gpuUINT8: return 1;
    default: return 0;
  }
}

inline bool shxOk(int rc, const char* what) {
  if (rc == SHX_OK) return true;
  std::cerr << what << " failed: " << shx_last_error() << std::endl;
  return false;
}

// A thin wrapper of a device image (gpuimage.h): dims[0] = width (fastest).
class GpuImage {
 public:
  GpuImage() : m_type(gpuNOTDEF), m_d_ptr(nullptr), m_auto_delete(false), m_dims{0, 0, 0} {}
  GpuImage(const GpuImage& src) : GpuImage() {
    if (src.m_d_ptr) {
      allocate(src.m_type, src.m_dims[0], src.m_dims[1], src.m_dims[2]);
      hipMemcpy(m_d_ptr, src.m_d_ptr, size(), hipMemcpyDeviceToDevice);
    }
  }
  GpuImage(gpuTYPE_t type, int ndims, const int* dims) : GpuImage() { allocate(type, ndims, dims); }
  GpuImage(gpuTYPE_t type, int dimX, int dimY, int dimZ = 1) : GpuImage() { allocate(type, dimX, dimY, dimZ); }
  GpuImage& operator=(const GpuImage& src) {
    if (this != &src) {
      freeData();
      if (src.m_d_ptr) {
        allocate(src.m_type, src.m_dims[0], src.m_dims[1], src.m_dims[2]);
        hipMemcpy(m_d_ptr, src.m_d_ptr, size(), hipMemcpyDeviceToDevice);
      }
    }
    return *this;
  }
  ~GpuImage() { freeData(); }

  int size() const { return numElements() * elemSize(m_type); }
  void* getPtr() const { return m_d_ptr; }
  int width() const { return m_dims[0]; }
  int height() const { return m_dims[1]; }
  int depth() const { return m_dims[2]; }
  const int* dims() const { return m_dims; }
  gpuTYPE_t type() const { return m_type; }
  int numDims() const { return m_dims[2] > 1 ? 3 : 2; }
  int numElements() const { return m_dims[0] * m_dims[1] * m_dims[2]; }

  void setData(gpuTYPE_t type, int dimX, int dimY, int dimZ, void* d_ptr) {
    freeData();
    m_type = type;
    m_dims[0] = dimX;
    m_dims[1] = dimY;
    m_dims[2] = dimZ;
    m_d_ptr = d_ptr;
    m_auto_delete = false;
  }
  void allocate(gpuTYPE_t type, int dimX, int dimY, int dimZ = 1) {
    freeData();
    m_type = type;
    m_dims[0] = dimX;
    m_dims[1] = dimY;
    m_dims[2] = dimZ;
    if (numElements() > 0 && hipMalloc(&m_d_ptr, size()) != hipSuccess) {
      std::cerr << "GpuImage::allocate() failed" << std::endl;
      m_d_ptr = nullptr;
    }
    m_auto_delete = m_d_ptr != nullptr;
  }
  void allocate(gpuTYPE_t type, int ndims, const int* dims) {
    allocate(type, dims[0], ndims > 1 ? dims[1] : 1, ndims > 2 ? dims[2] : 1);
  }
  void transferToDevice(const void* src) { hipMemcpy(m_d_ptr, src, size(), hipMemcpyHostToDevice); }
  void transferFromDevice(void* dst) const { hipMemcpy(dst, m_d_ptr, size(), hipMemcpyDeviceToHost); }
  void freeData() {
    if (m_auto_delete && m_d_ptr) hipFree(m_d_ptr);
    m_d_ptr = nullptr;
    m_auto_delete = false;
  }
  // Convert between gpuFLOAT and gpuUINT8 on the device (gpuimage.cpp:121-140).
  bool convert(gpuTYPE_t dstType) {
    void* d_dst = nullptr;
    if (hipMalloc(&d_dst, (size_t)numElements() * elemSize(dstType)) != hipSuccess) return false;
    if (!shxOk(shx_convert(d_dst, dstType, m_d_ptr, m_type, numElements(), nullptr), "Image conversion")) {
      hipFree(d_dst);
      return false;
    }
    hipDeviceSynchronize();
    freeData();
    m_d_ptr = d_dst;
    m_auto_delete = true;
    m_type = dstType;
    return true;
  }

 private:
  gpuTYPE_t m_type;
  void* m_d_ptr;
  bool m_auto_delete;
  int m_dims[3];
};

// A cell array of device images (gpucelldata.h).
class GpuCellData {
 public:
  int numElem() const { return (int)m_GpuData.size(); }
  void setNumElem(int n) { m_GpuData.resize(n); }
  GpuImage* elem(int idx) { return &m_GpuData[idx]; }
  const GpuImage* elem(int idx) const { return &m_GpuData[idx]; }

 private:
  std::vector<GpuImage> m_GpuData;
};

// Spatial shearing filters of every scale plus the a trous filters
// (gpushearfilters.h).  filters[j]: numDir[j] row-major planes of
// sizeShearing[j]^2 values of dataType (gpuFLOAT or gpuDOUBLE), host memory.
class GpuShearFilters {
 public:
  GpuShearFilters(int version, gpuTYPE_t dataType, int numScales, const int* sizeShearing, const int* numDir,
                  const void* const* filters, const void* const* atrousFilters, const int* atrousLen)
      : m_nVersion(version), m_type(dataType), m_valid(true) {
    for (int j = 0; j < numScales; ++j) {
      m_size.push_back(sizeShearing[j]);
      m_numDir.push_back(numDir[j]);
      m_w.push_back(toDouble(filters[j], (size_t)numDir[j] * sizeShearing[j] * sizeShearing[j]));
    }
    // the pyramid runs the built-in maxflat filters: accept only those
    for (int k = 0; k < 4 && m_valid; ++k) {
      int len = 0;
      shx_maxflat_filter(k, nullptr, &len);
      std::vector<double> ref((size_t)len * len);
      shx_maxflat_filter(k, ref.data(), &len);
      std::vector<double> got = toDouble(atrousFilters[k], (size_t)atrousLen[k] * atrousLen[k]);
      if (atrousLen[k] != len) {
        m_valid = false;
        break;
      }
      for (size_t i = 0; i < got.size(); ++i)
        if (std::fabs(got[i] - ref[i]) > 1e-6 * (1.0 + std::fabs(ref[i]))) m_valid = false;
    }
    if (!m_valid) std::cerr << "GpuShearFilters: only the maxflat a trous filters are supported" << std::endl;
  }

  gpuTYPE_t dataType() const { return m_type; }
  int numScales() const { return (int)m_size.size(); }
  int numDirections(int scale) const { return m_numDir[scale]; }
  int filterSize(int scale) const { return m_size[scale]; }
  bool hasData() const { return m_valid && !m_size.empty(); }
  const double* hostFilters(int scale) const { return m_w[scale].data(); }
  int version() const { return m_nVersion; }

 private:
  std::vector<double> toDouble(const void* src, size_t n) const {
    std::vector<double> out(n);
    if (m_type == gpuDOUBLE)
      std::memcpy(out.data(), src, n * sizeof(double));
    else
      for (size_t i = 0; i < n; ++i) out[i] = ((const float*)src)[i];
    return out;
  }
  int m_nVersion;
  gpuTYPE_t m_type;
  bool m_valid;
  std::vector<int> m_size, m_numDir;
  std::vector<std::vector<double>> m_w;
};

// Dictionary of one image size (gpusheardictionary.h): the prepared spectra
// live in an shx_dict.  norm() holds the per-band noise scalars,
// E[idxScale + idxDirection * (numScales + 1)] (lowpass row 0).
class GpuShearDictionary {
 public:
  GpuShearDictionary() : m_dict(nullptr), m_pBaseFilters(nullptr), m_nFilterLen(0) {}
  ~GpuShearDictionary() { shx_clear_shearing_filters(m_dict); }
  GpuShearDictionary(const GpuShearDictionary&) = delete;
  GpuShearDictionary& operator=(const GpuShearDictionary&) = delete;

  gpuTYPE_t dataType() const { return gpuFLOAT; }
  gpuTYPE_t realType() const { return gpuFLOAT; }
  int numScales() const { return m_pBaseFilters ? m_pBaseFilters->numScales() : 0; }
  int filterLen() const { return m_nFilterLen; }
  int numDirections(int scale) const { return m_pBaseFilters->numDirections(scale); }
  bool hasData() const { return m_dict != nullptr; }
  int maxDirections() const {
    int m = 1;
    for (int j = 0; j < numScales(); ++j) m = std::max(m, numDirections(j));
    return m;
  }
  double* norm() { return m_dNorm.data(); }
  const double* norm() const { return m_dNorm.data(); }
  shx_dict* handle() const { return m_dict; }

  // Prepare the filter spectra for len x len images (gpusheardictionary.cpp:54-132).
  bool prepare(const GpuShearFilters* filters, int len) {
    if (hasData()) {
      std::cerr << "GpuShearDictionary::prepare() should only be called once" << std::endl;
      return false;
    }
    if (!filters || !filters->hasData()) {
      std::cerr << "GpuShearDictionary::prepare() has been called without valid filters" << std::endl;
      return false;
    }
    const int S = filters->numScales();
    std::vector<int> nd(S), ds(S);
    std::vector<const double*> w(S);
    for (int j = 0; j < S; ++j) {
      nd[j] = filters->numDirections(j);
      ds[j] = filters->filterSize(j);
      w[j] = filters->hostFilters(j);
    }
    if (!shxOk(shx_prepare_shearing_filters(len, S, nd.data(), ds.data(), w.data(), SHX_SHIFT_NONE, &m_dict),
               "GpuShearDictionary::prepare()"))
      return false;
    m_pBaseFilters = filters;
    m_nFilterLen = len;
    m_dNorm.assign((size_t)(S + 1) * maxDirections(), 0.0);
    return true;
  }
  // Temporary buffers of the transforms (gpusheardictionary.cpp:134-164): the
  // library keeps its workspaces in the dictionary; temp receives the
  // pyramid planes used by atrousdec / atrousrec callers.
  bool prepareTempBuffers(GpuCellData& temp) const {
    if (!hasData()) return false;
    temp.setNumElem(4 + numScales() + 1);
    for (int i = 0; i < 3; ++i) temp.elem(i)->allocate(gpuFLOAT, m_nFilterLen, m_nFilterLen, 2 * maxDirections());
    for (int j = 0; j <= numScales(); ++j) temp.elem(3 + j)->allocate(gpuFLOAT, m_nFilterLen, m_nFilterLen);
    return true;
  }

 private:
  shx_dict* m_dict;
  const GpuShearFilters* m_pBaseFilters;
  int m_nFilterLen;
  std::vector<double> m_dNorm;
};

// The demo's transform object (shearcuda.h:13-48).
class Shearcuda {
 public:
  Shearcuda() : m_nDevice(-1), m_ms(0.f), m_start(nullptr), m_stop(nullptr) {}
  ~Shearcuda() { shutdown(); }

  bool init() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
      std::cerr << "No HIP device available" << std::endl;
      return false;
    }
    m_nDevice = 0;
    if (hipSetDevice(m_nDevice) != hipSuccess) return false;
    hipEventCreate(&m_start);
    hipEventCreate(&m_stop);
    return true;
  }
  void shutdown() {
    if (m_start) hipEventDestroy(m_start);
    if (m_stop) hipEventDestroy(m_stop);
    m_start = m_stop = nullptr;
  }

  bool addNoise(GpuImage& gpuImage, double sigma) const {
    if (gpuImage.type() != gpuFLOAT) return false;
    return shxOk(shx_add_noise((float*)gpuImage.getPtr(), gpuImage.numElements(), sigma, 1234ULL, nullptr),
                 "Shearcuda::addNoise()");
  }

  // coeff: numScales + 1 cells; cell 0 lowpass (L x L), cell j L x L x D_j.
  bool shearTrans(const GpuImage& gpuImage, const GpuShearDictionary& shearDict, GpuCellData& temp,
                  GpuCellData& coeff) const {
    (void)temp;
    if (gpuImage.type() != shearDict.realType()) {
      std::cerr << "Data types mismatch in Shearcuda::shearTrans()" << std::endl;
      return false;
    }
    if (gpuImage.width() != shearDict.filterLen() || gpuImage.height() != shearDict.filterLen()) {
      std::cerr << "Image size mismatch in Shearcuda::shearTrans()" << std::endl;
      return false;
    }
    std::vector<float*> cells;
    if (!allocCoeff(shearDict, coeff, cells)) return false;
    return shxOk(shx_shear_trans(shearDict.handle(), (const float*)gpuImage.getPtr(), cells.data(), nullptr),
                 "Shearcuda::shearTrans()");
  }

  // Hard threshold in place: band (s, k) keeps |c| > sc * lambda * E, with
  // sc[0] for the lowpass, sc[1] for scales 1..S-1 and sc[2] for the finest
  // (shearcuda.cpp:320-358).
  bool hardThreshold(GpuCellData& coeff, double lambda, const double* E, const double* sc) const {
    const int ns = coeff.numElem();
    int maxDir = 1;
    std::vector<float*> cells(ns);
    std::vector<int> planes(ns);
    std::vector<int64_t> elems(ns);
    std::vector<double> scv(ns);
    for (int s = 0; s < ns; ++s) {
      cells[s] = (float*)coeff.elem(s)->getPtr();
      planes[s] = coeff.elem(s)->depth();
      elems[s] = (int64_t)coeff.elem(s)->width() * coeff.elem(s)->height();
      scv[s] = s == 0 ? sc[0] : (s < ns - 1 ? sc[1] : sc[2]);
      maxDir = std::max(maxDir, planes[s]);
    }
    return shxOk(shx_thresh(cells.data(), ns, planes.data(), elems.data(), lambda, 2, E, ns, maxDir, scv.data(), ns,
                            0, nullptr, nullptr),
                 "Shearcuda::hardThreshold()");
  }

  bool inverseShear(const GpuCellData& coeff, const GpuShearDictionary& shearDict, GpuCellData& temp,
                    GpuImage& gpuImage) const {
    (void)temp;
    if (coeff.numElem() != shearDict.numScales() + 1) return false;
    if (gpuImage.numElements() == 0) gpuImage.allocate(shearDict.realType(), shearDict.filterLen(), shearDict.filterLen());
    std::vector<float*> cells(coeff.numElem());
    for (int j = 0; j < coeff.numElem(); ++j) cells[j] = (float*)coeff.elem(j)->getPtr();
    return shxOk(shx_inverse_shear(shearDict.handle(), cells.data(), (float*)gpuImage.getPtr(), nullptr),
                 "Shearcuda::inverseShear()");
  }

  // Noise scalars of every band: transform of a unit-variance noise image
  // (seed 4321), median absolute deviation / 0.6745 per band in the
  // reference's float arithmetic (shearcuda.cpp:205-271), computed on the
  // device by shx_band_mad (exact medians, the same values as the host loop).
  bool computeShearNorm(GpuShearDictionary& shearDict, GpuCellData& temp) const {
    const int L = shearDict.filterLen();
    GpuImage x(gpuFLOAT, L, L);
    hipMemset(x.getPtr(), 0, x.size());
    if (!shxOk(shx_add_noise((float*)x.getPtr(), x.numElements(), 1.0, 4321ULL, nullptr), "noise")) return false;
    GpuCellData Ct;
    if (!shearTrans(x, shearDict, temp, Ct)) return false;
    const int ns = Ct.numElem();
    int maxd = 1;
    for (int s = 0; s < ns; ++s) maxd = std::max(maxd, Ct.elem(s)->depth());
    GpuImage mad(gpuDOUBLE, maxd, 1);
    std::vector<double> host(maxd);
    double* E = shearDict.norm();
    for (int s = 0; s < ns; ++s) {
      const int depth = Ct.elem(s)->depth();
      if (!shxOk(shx_band_mad((const float*)Ct.elem(s)->getPtr(), (long long)L * L, (long long)L * L, depth,
                              SHX_MAD_FP32, (double*)mad.getPtr(), nullptr),
                 "Shearcuda::computeShearNorm()"))
        return false;
      if (hipMemcpy(host.data(), mad.getPtr(), sizeof(double) * depth, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
      for (int k = 0; k < depth; ++k) E[s + k * ns] = host[k];
    }
    return true;
  }

  // y: numLevels + 1 images (y[0] lowpass, coarse to fine); h0 / h1 must be
  // the maxflat filters (the library's pyramid).
  bool atrousdec(const GpuImage& x, const GpuImage& h0, const GpuImage& h1, int numLevels, GpuImage* y,
                 GpuImage* tempBuffer) const {
    (void)h0;
    (void)h1;
    (void)tempBuffer;
    std::vector<float*> yp(numLevels + 1);
    for (int j = 0; j <= numLevels; ++j) {
      if (y[j].numElements() != x.numElements()) y[j].allocate(gpuFLOAT, x.width(), x.height());
      yp[j] = (float*)y[j].getPtr();
    }
    return shxOk(shx_atrousdec((const float*)x.getPtr(), x.width(), numLevels, yp.data(), nullptr, nullptr),
                 "Shearcuda::atrousdec()");
  }
  bool atrousrec(const GpuImage* y, const GpuImage& g0, const GpuImage& g1, int numLevels, GpuImage& outputImage,
                 GpuImage* tempBuffer) const {
    (void)g0;
    (void)g1;
    (void)tempBuffer;
    std::vector<const float*> yp(numLevels + 1);
    for (int j = 0; j <= numLevels; ++j) yp[j] = (const float*)y[j].getPtr();
    if (outputImage.numElements() == 0) outputImage.allocate(gpuFLOAT, y[0].width(), y[0].height());
    return shxOk(shx_atrousrec(yp.data(), y[0].width(), numLevels, (float*)outputImage.getPtr(), nullptr, nullptr),
                 "Shearcuda::atrousrec()");
  }
  // Directional bands of pyramid levels y (shearcuda.cpp:549-624); d receives
  // numScales + 1 images (d[0] a copy of y[0]).
  bool convolution(const GpuImage* y, const GpuShearDictionary& shear, GpuImage* tempBuffer, GpuImage* d) const {
    (void)tempBuffer;
    const int S = shear.numScales(), L = shear.filterLen();
    std::vector<const float*> yp(S + 1);
    std::vector<float*> dp(S + 1);
    for (int j = 0; j <= S; ++j) {
      yp[j] = (const float*)y[j].getPtr();
      const int depth = j == 0 ? 1 : shear.numDirections(j - 1);
      if (d[j].numElements() != L * L * depth) d[j].allocate(gpuFLOAT, L, L, depth);
      dp[j] = (float*)d[j].getPtr();
    }
    return shxOk(shx_convolution(shear.handle(), yp.data(), dp.data(), nullptr), "Shearcuda::convolution()");
  }
  // Conjugate-filtered band sums of every scale (shearcuda.cpp:719-800).
  bool applyfilters(const GpuImage* d, const GpuShearDictionary& shear, GpuImage* tempBuffer, GpuImage* y) const {
    (void)tempBuffer;
    const int S = shear.numScales(), L = shear.filterLen();
    std::vector<const float*> dp(S + 1);
    std::vector<float*> yp(S + 1, nullptr);
    for (int j = 0; j <= S; ++j) {
      dp[j] = (const float*)d[j].getPtr();
      if (j > 0) {
        if (y[j].width() != L || y[j].height() != L) return false;
        yp[j] = (float*)y[j].getPtr();
      }
    }
    return shxOk(shx_applyfilters(shear.handle(), dp.data(), yp.data(), nullptr), "Shearcuda::applyfilters()");
  }

  // Lp norm (p-th root taken) of numElem FP32 values (shearcuda.cpp:626-670).
  double getNorm(const void* image_ptr, void* scratch_buf, double p, int numElem, gpuTYPE_t type_signal) const {
    (void)scratch_buf;
    if (type_signal != gpuFLOAT) return -1.0;
    double s = 0.0;
    if (!shxOk(shx_norm_pow((const float*)image_ptr, numElem, p, &s, nullptr), "Shearcuda::getNorm()")) return -1.0;
    return std::pow(s, 1.0 / p);
  }
  double getNormErr(const void* imageA_ptr, const void* imageB_ptr, void* scratch_buf, double p, int numElem,
                    gpuTYPE_t type_signal) const {
    (void)scratch_buf;
    if (type_signal != gpuFLOAT) return -1.0;
    double s = 0.0;
    if (!shxOk(shx_norm_pow_diff((const float*)imageA_ptr, (const float*)imageB_ptr, numElem, p, &s, nullptr),
               "Shearcuda::getNormErr()"))
      return -1.0;
    return std::pow(s, 1.0 / p);
  }

  // Device sanity check (shearcuda.cpp:121-128).
  void quickTest() const {
    GpuImage a(gpuFLOAT, 16, 16);
    std::vector<float> h(256, 1.0f);
    a.transferToDevice(h.data());
    std::cout << "quickTest: |ones(16)|_2 = " << getNorm(a.getPtr(), nullptr, 2.0, 256, gpuFLOAT) << std::endl;
  }
  // Wall time between startTimers and stopTimers (the reference's GpuTimes
  // keeps per-operation totals; here one elapsed figure, elapsedMs()).
  void startTimers() const { hipEventRecord(m_start, nullptr); }
  void stopTimers() const {
    hipEventRecord(m_stop, nullptr);
    hipEventSynchronize(m_stop);
    hipEventElapsedTime(&m_ms, m_start, m_stop);
  }
  float elapsedMs() const { return m_ms; }

 private:
  static bool allocCoeff(const GpuShearDictionary& d, GpuCellData& coeff, std::vector<float*>& cells) {
    const int S = d.numScales(), L = d.filterLen();
    if (coeff.numElem() != S + 1) coeff.setNumElem(S + 1);
    cells.resize(S + 1);
    for (int j = 0; j <= S; ++j) {
      const int depth = j == 0 ? 1 : d.numDirections(j - 1);
      GpuImage* e = coeff.elem(j);
      if (e->type() != gpuFLOAT || e->width() != L || e->height() != L || e->depth() != depth)
        e->allocate(gpuFLOAT, L, L, depth);
      if (!e->getPtr()) return false;
      cells[j] = (float*)e->getPtr();
    }
    return true;
  }
  int m_nDevice;
  mutable float m_ms;
  hipEvent_t m_start, m_stop;
};

// SHFM0002 filter file (demoImageDenoise/shearconfig.cpp:25-70): "SHFM0002",
// uint16 numScales, uint16 dataType (0 single, 2 double), per scale uint16
// size, uint16 numDir and the column-major planes, then the four a trous
// filters (uint16 length and column-major data).  Planes are transposed to
// row-major on reading.
class ShearConfig {
 public:
  bool readFromFile(const std::string& fileName) {
    FILE* f = std::fopen(fileName.c_str(), "rb");
    if (!f) return false;
    bool ok = read(f);
    std::fclose(f);
    return ok;
  }
  int version() const { return m_nVersion; }
  gpuTYPE_t dataType() const { return m_dataType; }
  int numScales() const { return (int)m_sizeShearing.size(); }
  const int* sizeShearing() const { return m_sizeShearing.data(); }
  const int* numDir() const { return m_numDir.data(); }
  const void* const* filters() const {
    m_fptr.clear();
    for (auto& v : m_filter) m_fptr.push_back(v.data());
    return m_fptr.data();
  }
  const void* const* atrousFilters() const {
    m_aptr.clear();
    for (auto& v : m_atrous) m_aptr.push_back(v.data());
    return m_aptr.data();
  }
  const int* atrousLen() const { return m_atrousLen; }

 private:
  template <class T>
  static bool readv(FILE* f, T* v, size_t n) {
    return std::fread(v, sizeof(T), n, f) == n;
  }
  bool readPlanes(FILE* f, int len, int count, std::vector<unsigned char>& out) {
    const int es = elemSize(m_dataType);
    std::vector<unsigned char> raw((size_t)count * len * len * es);
    if (!readv(f, raw.data(), raw.size())) return false;
    out.resize(raw.size());
    for (int k = 0; k < count; ++k)
      for (int r = 0; r < len; ++r)
        for (int c = 0; c < len; ++c)
          std::memcpy(&out[(((size_t)k * len + r) * len + c) * es], &raw[(((size_t)k * len + c) * len + r) * es], es);
    return true;
  }
  bool read(FILE* f) {
    char head[8];
    if (!readv(f, head, 8) || std::memcmp(head, "SHFM0002", 8) != 0) return false;
    m_nVersion = 1;
    uint16_t ns = 0, dt = 0;
    if (!readv(f, &ns, 1) || !readv(f, &dt, 1)) return false;
    if (dt != 0 && dt != 2) return false;
    m_dataType = (gpuTYPE_t)dt;
    m_sizeShearing.resize(ns);
    m_numDir.resize(ns);
    m_filter.resize(ns);
    for (int j = 0; j < ns; ++j) {
      uint16_t len = 0, nd = 0;
      if (!readv(f, &len, 1) || !readv(f, &nd, 1)) return false;
      m_sizeShearing[j] = len;
      m_numDir[j] = nd;
      if (!readPlanes(f, len, nd, m_filter[j])) return false;
    }
    m_atrous.resize(4);
    for (int j = 0; j < 4; ++j) {
      uint16_t len = 0;
      if (!readv(f, &len, 1)) return false;
      m_atrousLen[j] = len;
      if (!readPlanes(f, len, 1, m_atrous[j])) return false;
    }
    return true;
  }
  int m_nVersion = -1;
  gpuTYPE_t m_dataType = gpuFLOAT;
  std::vector<int> m_sizeShearing, m_numDir;
  std::vector<std::vector<unsigned char>> m_filter, m_atrous;
  int m_atrousLen[4] = {0, 0, 0, 0};
  mutable std::vector<const void*> m_fptr, m_aptr;
};

}  // namespace shearcuda_amd
