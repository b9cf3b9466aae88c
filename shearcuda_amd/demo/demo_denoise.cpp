// Command-line form of the reference's image denoising demo
// (demo_denoise/demoImageDenoise/mainwindow.cpp:133-261, without the Qt UI):
// reads an SHFM0002 filter file and an 8-bit L x L image, adds N(0, sigma^2)
// noise, computes the per-band noise norms, and denoises with
// shearTrans -> hardThreshold (Tscalars {0, 3, 4}) -> inverseShear through the
// C++ entry points of include/shearcuda_amd.hpp.
//
// usage: demo_denoise FILTERS.bin IMAGE.u8 L SIGMA OUT.f32 [NOISY.f32]
//   IMAGE.u8  L*L bytes, row-major; OUT.f32 / NOISY.f32 receive FP32 images.
// Prints the noisy and denoised PSNR (dB, peak 255) and the denoise time.
//        demo_denoise --check-config FILTERS.bin R C
// reads the filter file only (no GPU) and prints, per scale, its size,
// direction count and element (R, C) of direction 0 (row-major, after the
// reader's transpose), then the a trous filter lengths.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "shearcuda_amd.hpp"

using namespace shearcuda_amd;

static double calculatePsnr(const Shearcuda& sc, const GpuImage& a, const GpuImage& b, double& mse) {
  const int num = a.numElements();
  const double nd = sc.getNormErr(a.getPtr(), b.getPtr(), nullptr, 2.0, num, a.type());
  const double sq = nd * nd;
  mse = sq / num;
  return 10 * std::log10(255.0 * 255.0 * num / sq);
}

static bool writeFile(const char* path, const GpuImage& img) {
  std::vector<float> h(img.numElements());
  img.transferFromDevice(h.data());
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
  std::fclose(f);
  return ok;
}

static int checkConfig(const char* path, int r, int c) {
  ShearConfig config;
  if (!config.readFromFile(path)) {
    std::fprintf(stderr, "Failed to read shearlet configuration from file %s\n", path);
    return 1;
  }
  std::printf("scales %d type %d\n", config.numScales(), (int)config.dataType());
  for (int j = 0; j < config.numScales(); ++j) {
    const int n = config.sizeShearing()[j];
    const size_t i = (size_t)r * n + c;
    const double v = config.dataType() == gpuDOUBLE ? ((const double*)config.filters()[j])[i]
                                                    : ((const float*)config.filters()[j])[i];
    std::printf("scale %d size %d dirs %d elem %.9e\n", j, n, config.numDir()[j], v);
  }
  std::printf("atrous %d %d %d %d\n", config.atrousLen()[0], config.atrousLen()[1], config.atrousLen()[2],
              config.atrousLen()[3]);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 5 && std::string(argv[1]) == "--check-config") return checkConfig(argv[2], std::atoi(argv[3]), std::atoi(argv[4]));
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s FILTERS.bin IMAGE.u8 L SIGMA OUT.f32 [NOISY.f32]\n", argv[0]);
    return 2;
  }
  const int L = std::atoi(argv[3]);
  const double sigma = std::atof(argv[4]);
  Shearcuda sc;
  if (!sc.init()) return 1;
  ShearConfig config;
  if (!config.readFromFile(argv[1])) {
    std::fprintf(stderr, "Failed to read shearlet configuration from file %s\n", argv[1]);
    return 1;
  }
  std::vector<unsigned char> pix((size_t)L * L);
  FILE* f = std::fopen(argv[2], "rb");
  if (!f || std::fread(pix.data(), 1, pix.size(), f) != pix.size()) {
    std::fprintf(stderr, "Failed to read %d x %d image %s\n", L, L, argv[2]);
    return 1;
  }
  std::fclose(f);

  GpuShearFilters gpuFilters(config.version(), config.dataType(), config.numScales(), config.sizeShearing(),
                             config.numDir(), config.filters(), config.atrousFilters(), config.atrousLen());
  GpuShearDictionary gpuShear;
  if (!gpuShear.prepare(&gpuFilters, L)) return 1;
  GpuCellData gpuTemp;
  gpuShear.prepareTempBuffers(gpuTemp);
  if (!sc.computeShearNorm(gpuShear, gpuTemp)) return 1;

  GpuImage gpuResult(gpuFLOAT, L, L);
  GpuImage gpuImage(gpuUINT8, L, L);
  gpuImage.transferToDevice(pix.data());
  if (!gpuImage.convert(gpuFLOAT)) return 1;
  GpuImage gpuNoisyImage(gpuImage);
  if (!sc.addNoise(gpuNoisyImage, sigma)) return 1;
  double noisyMse = 0, resultMse = 0;
  const double noisyPsnr = calculatePsnr(sc, gpuNoisyImage, gpuImage, noisyMse);

  GpuCellData gpuCoeff;
  hipDeviceSynchronize();
  sc.startTimers();
  if (!sc.shearTrans(gpuNoisyImage, gpuShear, gpuTemp, gpuCoeff)) return 1;
  const double Tscalars[3] = {0, 3, 4};
  if (!sc.hardThreshold(gpuCoeff, sigma, gpuShear.norm(), Tscalars)) return 1;
  if (!sc.inverseShear(gpuCoeff, gpuShear, gpuTemp, gpuResult)) return 1;
  sc.stopTimers();
  const double resultPsnr = calculatePsnr(sc, gpuResult, gpuImage, resultMse);

  if (!writeFile(argv[5], gpuResult)) return 1;
  if (argc > 6 && !writeFile(argv[6], gpuNoisyImage)) return 1;
  std::printf("Noisy image PSNR = %.4f dB (MSE = %.4f)\n", noisyPsnr, noisyMse);
  std::printf("Denoised image PSNR = %.4f dB (MSE = %.4f)\n", resultPsnr, resultMse);
  std::printf("Denoise time elapsed: %.3f ms\n", sc.elapsedMs());
  sc.shutdown();
  return 0;
}
