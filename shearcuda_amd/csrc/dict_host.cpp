// FP64 host construction of the shearing-filter spectra.
//
// H_k = FFT_LxL(pad(w_k)) with pad = zero padding at the top-left corner
// (shearlet_toolbox_1/nsst_dec2.m:50, fft2(w_s(:,:,k), L, L)) or, for
// SHX_SHIFT_CENTER, the padded filter rolled by -n1/2 in both axes
// (shearcuda/src_cuda/prepare_shearing_filters_cuda.cpp:254 with the circular
// shift of shearcuda/shear_cuda.cu:2303-2305); then every frequency is divided
// by sqrt(sum_k |H_k|^2) (nsst_dec2.m:54-58, shear_cuda.cu:2399-2423).  The
// common 1/L factor of nsst_dec2.m:50 cancels in that division.
// Only the half spectrum v = 0..L/2 is formed: the filters are real.
#include <algorithm>
#include <cmath>
#include <complex>
#include <thread>
#include <vector>

#include "dict_host.h"

namespace shx {
namespace {

using cd = std::complex<double>;

struct Fft {
  int n = 0, lg = 0;
  std::vector<cd> w;     // exp(-2 pi i k / n), k < n/2
  std::vector<int> rev;  // bit reversal
  explicit Fft(int n_) : n(n_) {
    while ((1 << lg) < n) ++lg;
    w.resize(n / 2 > 0 ? n / 2 : 1);
    for (int k = 0; k < n / 2; ++k) w[k] = std::polar(1.0, -2.0 * M_PI * k / n);
    rev.resize(n);
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b)
        if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
  }
  // in-place forward transform of n points at stride s
  void run(cd* a, int s) const {
    for (int i = 0; i < n; ++i)
      if (i < rev[i]) std::swap(a[(size_t)i * s], a[(size_t)rev[i] * s]);
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len)
        for (int k = 0; k < half; ++k) {
          const cd t = w[(size_t)k * step] * a[(size_t)(i + k + half) * s];
          const cd u = a[(size_t)(i + k) * s];
          a[(size_t)(i + k) * s] = u + t;
          a[(size_t)(i + k + half) * s] = u - t;
        }
    }
  }
};

}  // namespace

void scale_spectra_fp64(int L, int D, int n1, const double* w, int shift, std::vector<std::complex<double>>& H) {
  const int N = L / 2, NC = N + 1;
  H.assign((size_t)D * L * NC, cd(0.0, 0.0));
  const int s = shift ? n1 / 2 : 0;
  const Fft fft(L);
  auto work = [&](int k0, int k1) {
    std::vector<cd> A((size_t)L * L);
    std::vector<cd> col(L);
    for (int k = k0; k < k1; ++k) {
      std::fill(A.begin(), A.end(), cd(0.0, 0.0));
      const double* wk = w + (size_t)k * n1 * n1;
      for (int r = 0; r < n1; ++r)
        for (int c = 0; c < n1; ++c) {
          const int rr = ((r - s) % L + L) % L, cc = ((c - s) % L + L) % L;
          A[(size_t)rr * L + cc] += wk[(size_t)r * n1 + c];
        }
      for (int r = 0; r < n1; ++r) {
        const int rr = ((r - s) % L + L) % L;
        fft.run(&A[(size_t)rr * L], 1);
      }
      cd* Hk = H.data() + (size_t)k * L * NC;
      for (int v = 0; v < NC; ++v) {
        for (int u = 0; u < L; ++u) col[u] = A[(size_t)u * L + v];
        fft.run(col.data(), 1);
        for (int u = 0; u < L; ++u) Hk[(size_t)u * NC + v] = col[u];
      }
    }
  };
  const int nt = std::max(1, std::min<int>(D, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int k0 = (int)((long)D * t / nt), k1 = (int)((long)D * (t + 1) / nt);
    th.emplace_back(work, k0, k1);
  }
  for (auto& t : th) t.join();
  // normalise every bin so that sum_k |H_k|^2 = 1
  for (size_t i = 0; i < (size_t)L * NC; ++i) {
    double d = 0.0;
    for (int k = 0; k < D; ++k) d += std::norm(H[(size_t)k * L * NC + i]);
    d = std::sqrt(d);
    for (int k = 0; k < D; ++k) H[(size_t)k * L * NC + i] /= d;
  }
}

void pack_spectra_fp32(int L, int D, const std::vector<std::complex<double>>& H, float* main_out, float* beta_out) {
  const int N = L / 2, NC = N + 1;
  const double sc = 1.0 / ((double)L * (double)L);
  for (int k = 0; k < D; ++k) {
    const cd* Hk = H.data() + (size_t)k * L * NC;
    float* mk = main_out + (size_t)k * L * N * 2;
    float* bk = beta_out + (size_t)k * L * 2;
    for (int u = 0; u < L; ++u) {
      const cd h0 = Hk[(size_t)u * NC], hn = Hk[(size_t)u * NC + N];
      const cd alpha = 0.5 * (h0 + hn) * sc, beta = 0.5 * (h0 - hn) * sc;
      mk[((size_t)u * N) * 2 + 0] = (float)alpha.real();
      mk[((size_t)u * N) * 2 + 1] = (float)alpha.imag();
      bk[(size_t)u * 2 + 0] = (float)beta.real();
      bk[(size_t)u * 2 + 1] = (float)beta.imag();
      for (int v = 1; v < N; ++v) {
        const cd h = Hk[(size_t)u * NC + v] * sc;
        mk[((size_t)u * N + v) * 2 + 0] = (float)h.real();
        mk[((size_t)u * N + v) * 2 + 1] = (float)h.imag();
      }
    }
  }
}

void twiddles_fp32(int L, float* out) {
  for (int k = 0; k < L; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)L;
    out[2 * k + 0] = (float)std::cos(a);
    out[2 * k + 1] = (float)std::sin(a);
  }
}

void dct_twiddles_fp32(int L, float* out) {
  for (int k = 0; k < L; ++k) {
    const double a = -M_PI * (double)k / (2.0 * L);
    out[2 * k + 0] = (float)std::cos(a);
    out[2 * k + 1] = (float)std::sin(a);
  }
}

}  // namespace shx
