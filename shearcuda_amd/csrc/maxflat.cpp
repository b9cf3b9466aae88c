// Maxflat pyramid filters and their exact low-rank separable factors (host).
//
// The quadrant tables are numeric data that must equal the reference's filter
// design (shearlet_toolbox_1/atrousfilters.m:109-164); they are mirrored to
// 13x13 (h0, g1) and 19x19 (h1, g0) as atrousfilters.m:167-175 does.  Each full
// filter is symmetric, so h = sum_k lambda_k u_k u_k^T; a cyclic Jacobi
// eigensolver in FP64 finds the nonzero terms (h0: 2, h1: 4, g0: 3, g1: 3)
// and the kernels use v_k = u_k, w_k = lambda_k u_k (written as FP32
// compile-time constants by gen_factors.cpp).
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include <algorithm>

#include "pyramid.h"

namespace shx {
namespace {

// h0: 7x7 quadrant (atrousfilters.m maxflat case)
static const double kH0Q[7][7] = {
    {-7.900496718847182e-07, 0.0, 1.4220894093924927e-05, 2.5281589500310983e-05, -4.9773129328737247e-05, -0.00022753430550279883, -0.00033182086219158167},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {1.4220894093924927e-05, 0.0, -0.0002559760936906487, -0.00045506861100559767, 0.0008959163279172705, 0.004095617499050379, 0.00597277551944847},
    {2.5281589500310983e-05, 0.0, -0.00045506861100559767, 0.0009765625, 0.0015927401385195919, -0.0087890625, -0.01795090623402861},
    {-4.9773129328737247e-05, 0.0, 0.0008959163279172705, 0.0015927401385195919, -0.0031357071477104465, -0.014334661246676327, -0.020904714318069645},
    {-0.00022753430550279883, 0.0, 0.004095617499050379, -0.0087890625, -0.014334661246676327, 0.0791015625, 0.16155815610625748},
    {-0.00033182086219158167, 0.0, 0.00597277551944847, -0.01795090623402861, -0.020904714318069645, 0.16155815610625748, 0.3177420190660832},
};
// h1: 10x10 quadrant (atrousfilters.m maxflat case)
static const double kH1Q[10][10] = {
    {6.391587676622346e-10, 0.0, -1.7257286726880333e-08, -3.067962084778726e-08, 1.3805829381504267e-07, 5.522331752601707e-07, 3.3747582932565985e-07, -1.9328161134105974e-06, -5.6949046198705095e-06, -7.649452131381623e-06},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-1.7257286726880333e-08, 0.0, 4.65946741625769e-07, 8.283497628902559e-07, -3.727573933006152e-06, -1.4910295732024608e-05, -9.111847391792816e-06, 5.2186035062086126e-05, 0.00015376242473650378, 0.00020653520754730382},
    {-3.067962084778726e-08, 0.0, 8.283497628902559e-07, -2.9917573832012203e-07, -6.6267981031220475e-06, 5.3851632897621965e-06, 4.049868144081346e-05, -1.884807151416769e-05, -0.00023692226948222173, -0.0003769812640795245},
    {1.3805829381504267e-07, 0.0, -3.727573933006152e-06, -6.6267981031220475e-06, 2.9820591464049215e-05, 0.00011928236585619686, 7.289477913434253e-05, -0.000417488280496689, -0.0012300993978920302, -0.0016522816603784306},
    {5.522331752601707e-07, 0.0, -1.4910295732024608e-05, 5.3851632897621965e-06, 0.00011928236585619686, -9.693293921571956e-05, -0.0007289762659346422, 0.00033926528725501844, 0.004264600850679991, 0.006785662753431441},
    {3.3747582932565985e-07, 0.0, -9.111847391792816e-06, 4.049868144081346e-05, 7.289477913434253e-05, -0.0007289762659346422, -0.001468581806076247, 0.002551416930771248, 0.01181401175635013, 0.017093222023136675},
    {-1.9328161134105974e-06, 0.0, 5.2186035062086126e-05, -1.884807151416769e-05, -0.000417488280496689, 0.00033926528725501844, 0.002551416930771248, -0.0011874285053925643, -0.01492610297737997, -0.023749819637010044},
    {-5.6949046198705095e-06, 0.0, 0.00015376242473650378, -0.00023692226948222173, -0.0012300993978920302, 0.004264600850679991, 0.01181401175635013, -0.01492610297737997, -0.0826466923977296, -0.12203257624594532},
    {-7.649452131381623e-06, 0.0, 0.00020653520754730382, -0.0003769812640795245, -0.0016522816603784306, 0.006785662753431441, 0.017093222023136675, -0.023749819637010044, -0.12203257624594532, 0.821896776039774},
};
// g0: 10x10 quadrant (atrousfilters.m maxflat case)
static const double kG0Q[10][10] = {
    {-6.391587676622346e-10, 0.0, 1.7257286726880333e-08, 3.067962084778726e-08, -1.3805829381504267e-07, -5.522331752601707e-07, -3.3747582932565985e-07, 1.9328161134105974e-06, 5.6949046198705095e-06, 7.649452131381623e-06},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {1.7257286726880333e-08, 0.0, -4.65946741625769e-07, -8.283497628902559e-07, 3.727573933006152e-06, 1.4910295732024608e-05, 9.111847391792816e-06, -5.2186035062086126e-05, -0.00015376242473650378, -0.00020653520754730382},
    {3.067962084778726e-08, 0.0, -8.283497628902559e-07, -1.2809236054493144e-06, 6.6267981031220475e-06, 2.305662489808766e-05, 1.0064497559808503e-05, -8.06981871433068e-05, -0.00021814634152337594, -0.00028666046030363884},
    {-1.3805829381504267e-07, 0.0, 3.727573933006152e-06, 6.6267981031220475e-06, -2.9820591464049215e-05, -0.00011928236585619686, -7.289477913434253e-05, 0.000417488280496689, 0.0012300993978920302, 0.0016522816603784306},
    {-5.522331752601707e-07, 0.0, 1.4910295732024608e-05, 2.305662489808766e-05, -0.00011928236585619686, -0.00041501924816557786, -0.00018116095607655303, 0.0014525673685795225, 0.0039266341474207675, 0.005159888285465499},
    {-3.3747582932565985e-07, 0.0, 9.111847391792816e-06, 1.0064497559808503e-05, -7.289477913434253e-05, -0.00018116095607655303, 0.001468581806076247, 0.0006340633462679356, -0.01181401175635013, -0.021745034491193898},
    {1.9328161134105974e-06, 0.0, -5.2186035062086126e-05, -8.06981871433068e-05, 0.000417488280496689, 0.0014525673685795225, 0.0006340633462679356, -0.005083985790028328, -0.013743219515972684, -0.018059608999129246},
    {5.6949046198705095e-06, 0.0, -0.00015376242473650378, -0.00021814634152337594, 0.0012300993978920302, 0.0039266341474207675, -0.01181401175635013, -0.013743219515972684, 0.0826466923977296, 0.1638988884584603},
    {7.649452131381623e-06, 0.0, -0.00020653520754730382, -0.00028666046030363884, 0.0016522816603784306, 0.005159888285465499, -0.021745034491193898, -0.018059608999129246, 0.1638988884584603, 0.31358726209239235},
};
// g1: 7x7 quadrant (atrousfilters.m maxflat case)
static const double kG1Q[7][7] = {
    {-7.900496718847182e-07, 0.0, 1.4220894093924927e-05, 2.5281589500310983e-05, -4.9773129328737247e-05, -0.00022753430550279883, -0.00033182086219158167},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {1.4220894093924927e-05, 0.0, -0.0002559760936906487, -0.00045506861100559767, 0.0008959163279172705, 0.004095617499050379, 0.00597277551944847},
    {2.5281589500310983e-05, 0.0, -0.00045506861100559767, -0.0009765625, 0.0015927401385195919, 0.0087890625, 0.01329909376597139},
    {-4.9773129328737247e-05, 0.0, 0.0008959163279172705, 0.0015927401385195919, -0.0031357071477104465, -0.014334661246676327, -0.020904714318069645},
    {-0.00022753430550279883, 0.0, 0.004095617499050379, 0.0087890625, -0.014334661246676327, -0.0791015625, -0.1196918438937425},
    {-0.00033182086219158167, 0.0, 0.00597277551944847, 0.01329909376597139, -0.020904714318069645, -0.1196918438937425, 0.8177420190660831},
};

struct Full {
  int len = 0;
  std::vector<double> h;  // len x len
};

Full mirror(const double* q, int n) {
  Full f;
  f.len = 2 * n - 1;
  f.h.assign((size_t)f.len * f.len, 0.0);
  for (int r = 0; r < f.len; ++r)
    for (int c = 0; c < f.len; ++c) {
      const int qr = r < n ? r : 2 * n - 2 - r;
      const int qc = c < n ? c : 2 * n - 2 - c;
      f.h[(size_t)r * f.len + c] = q[qr * n + qc];
    }
  return f;
}

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
void jacobi(int n, std::vector<double> a, std::vector<double>& eval, std::vector<double>& evec) {
  evec.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) evec[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        const double v = a[(size_t)p * n + q] * a[(size_t)p * n + q];
        tot += v;
        if (p != q) off += v;
      }
    if (off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = evec[(size_t)k * n + p], vkq = evec[(size_t)k * n + q];
          evec[(size_t)k * n + p] = c * vkp - s * vkq;
          evec[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  eval.resize(n);
  for (int i = 0; i < n; ++i) eval[i] = a[(size_t)i * n + i];
}

struct Factors {
  int len = 0, rank = 0;
  std::vector<double> v, w;  // rank x len
  double err = 0.0;
};

Factors factorize(const Full& f) {
  const int n = f.len;
  std::vector<double> ev, V;
  jacobi(n, f.h, ev, V);
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return std::fabs(ev[x]) > std::fabs(ev[y]); });
  const double top = std::fabs(ev[idx[0]]);
  Factors F;
  F.len = n;
  for (int i : idx) {
    if (std::fabs(ev[i]) <= 1e-12 * top) break;
    ++F.rank;
    for (int a = 0; a < n; ++a) {
      double u = V[(size_t)a * n + i];
      if (a == 1 || a == n - 2) u = 0.0;  // exact zero rows of every maxflat filter
      F.v.push_back(u);
    }
    for (int a = 0; a < n; ++a) F.w.push_back(ev[i] * F.v[(size_t)(F.rank - 1) * n + a]);
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      double s = 0.0;
      for (int k = 0; k < F.rank; ++k) s += F.v[(size_t)k * n + r] * F.w[(size_t)k * n + c];
      F.err = std::max(F.err, std::fabs(s - f.h[(size_t)r * n + c]));
    }
  return F;
}

struct Maxflat {
  Full full[4];
  Factors fac[4];
  bool ok = false;
};

const Maxflat& maxflat() {
  static Maxflat M;
  static std::once_flag once;
  std::call_once(once, [] {
    M.full[0] = mirror(&kH0Q[0][0], 7);
    M.full[1] = mirror(&kH1Q[0][0], 10);
    M.full[2] = mirror(&kG0Q[0][0], 10);
    M.full[3] = mirror(&kG1Q[0][0], 7);
    for (int i = 0; i < 4; ++i) M.fac[i] = factorize(M.full[i]);
    // ranks the pyramid kernels are specialised for
    const int want[4] = {2, 4, 3, 3}, len[4] = {13, 19, 19, 13};
    M.ok = true;
    for (int i = 0; i < 4; ++i) M.ok = M.ok && M.fac[i].rank == want[i] && M.fac[i].len == len[i];
  });
  return M;
}

}  // namespace

bool maxflat_ok() { return maxflat().ok; }

const double* maxflat_filter(int which, int* len) {
  if (which < 0 || which > 3) return nullptr;
  const Full& f = maxflat().full[which];
  if (len) *len = f.len;
  return f.h.data();
}

bool maxflat_factors_fp64(int which, int* len, int* rank, const double** v, const double** w) {
  if (which < 0 || which > 3) return false;
  const Factors& F = maxflat().fac[which];
  *len = F.len;
  *rank = F.rank;
  *v = F.v.data();
  *w = F.w.data();
  return true;
}

double maxflat_factor_error() {
  const Maxflat& M = maxflat();
  if (!M.ok) return INFINITY;
  double e = 0.0;
  for (int i = 0; i < 4; ++i) e = std::max(e, M.fac[i].err);
  return e;
}

}  // namespace shx
