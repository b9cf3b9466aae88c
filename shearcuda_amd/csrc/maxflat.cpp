// Maxflat pyramid filters and their exact low-rank separable factors (host).
//
// The quadrant tables are numeric data that must equal the reference's filter
// design (shearlet_toolbox_1/atrousfilters.m:109-164); they are mirrored to
// 13x13 (h0, g1) and 19x19 (h1, g0) as atrousfilters.m:167-175 does.
//
// A pyramid level applies a filter pair: (h0, h1) in the decomposition,
// (g0, g1) in the reconstruction.  Zero-padded to 19x19, the columns of
// [h0 | h1 - delta] span only a 3-dimensional space, and so do those of
// [g0 | g1 - delta] (delta: the unit impulse at the centre).  With an
// orthonormal basis u_1..u_3 of that space (pivoted modified Gram-Schmidt in
// FP64), each filter F of the pair (the second one minus delta) is
//   F = sum_i u_i (F u_i)^T,
// exactly, and since every maxflat filter is symmetric (F = F^T) also
//   F = sum_i (F u_i) u_i^T.
// A level therefore needs three separable passes shared by both filters, plus
// the identity term (written as FP32 compile-time constants by
// gen_factors.cpp).
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
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

Full pad_to(const Full& f, int n) {
  Full g;
  g.len = n;
  g.h.assign((size_t)n * n, 0.0);
  const int o = (n - f.len) / 2;
  for (int r = 0; r < f.len; ++r)
    for (int c = 0; c < f.len; ++c) g.h[(size_t)(r + o) * n + c + o] = f.h[(size_t)r * f.len + c];
  return g;
}

struct PairFactors {
  int len = 0, rank = 0;
  std::vector<double> u, a, b;  // rank x len: basis, A u_i, (B - delta) u_i
  double err = 0.0;
};

// Joint factors of (A, B - delta) on a common odd side n (see the header).
PairFactors pair_factorize(const Full& A0, const Full& B0, int n) {
  const Full A = pad_to(A0, n);
  const Full B = [&] {
    Full t = pad_to(B0, n);
    t.h[(size_t)(n / 2) * n + n / 2] -= 1.0;
    return t;
  }();
  // the 2n columns of [A | B]
  std::vector<std::vector<double>> cols;
  for (const Full* f : {&A, &B})
    for (int c = 0; c < n; ++c) {
      std::vector<double> v(n);
      for (int r = 0; r < n; ++r) v[r] = f->h[(size_t)r * n + c];
      cols.push_back(v);
    }
  auto norm = [](const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
  };
  double top = 0.0;
  for (const auto& v : cols) top = std::max(top, norm(v));
  PairFactors P;
  P.len = n;
  std::vector<std::vector<double>> basis;
  while ((int)basis.size() < n) {
    int best = -1;
    double bn = 0.0;
    for (size_t j = 0; j < cols.size(); ++j) {
      const double t = norm(cols[j]);
      if (t > bn) bn = t, best = (int)j;
    }
    if (best < 0 || bn <= 1e-13 * top) break;
    std::vector<double> q = cols[best];
    for (double& x : q) x /= bn;
    for (int pass = 0; pass < 2; ++pass)  // re-orthogonalise against the earlier vectors
      for (const auto& e : basis) {
        double d = 0.0;
        for (int r = 0; r < n; ++r) d += e[r] * q[r];
        for (int r = 0; r < n; ++r) q[r] -= d * e[r];
      }
    const double qn = norm(q);
    for (double& x : q) x /= qn;
    for (auto& v : cols) {
      double d = 0.0;
      for (int r = 0; r < n; ++r) d += q[r] * v[r];
      for (int r = 0; r < n; ++r) v[r] -= d * q[r];
    }
    basis.push_back(q);
  }
  P.rank = (int)basis.size();
  for (const auto& q : basis) P.u.insert(P.u.end(), q.begin(), q.end());
  for (const Full* f : {&A, &B}) {
    std::vector<double>& out = (f == &A) ? P.a : P.b;
    for (const auto& q : basis)
      for (int c = 0; c < n; ++c) {  // (F^T q)[c]; F is symmetric
        double s = 0.0;
        for (int r = 0; r < n; ++r) s += f->h[(size_t)r * n + c] * q[r];
        out.push_back(s);
      }
  }
  for (const Full* f : {&A, &B}) {
    const std::vector<double>& w = (f == &A) ? P.a : P.b;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        double s = 0.0;
        for (int k = 0; k < P.rank; ++k) s += P.u[(size_t)k * n + r] * w[(size_t)k * n + c];
        P.err = std::max(P.err, std::fabs(s - f->h[(size_t)r * n + c]));
      }
  }
  return P;
}

struct Maxflat {
  Full full[4];
  PairFactors pair[2];  // (h0, h1), (g0, g1)
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
    M.pair[0] = pair_factorize(M.full[0], M.full[1], 19);
    M.pair[1] = pair_factorize(M.full[2], M.full[3], 19);
    // the rank the pyramid kernels are specialised for
    M.ok = M.pair[0].rank == 3 && M.pair[1].rank == 3;
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

bool maxflat_pair_fp64(int pair, int* len, int* rank, const double** u, const double** a, const double** b) {
  if (pair < 0 || pair > 1) return false;
  const PairFactors& P = maxflat().pair[pair];
  *len = P.len;
  *rank = P.rank;
  *u = P.u.data();
  *a = P.a.data();
  *b = P.b.data();
  return true;
}

namespace {
// R[kr][kc] = sum_{a,b} h[a][b] cos(pi kr (a-c) m / L) cos(pi kc (b-c) m / L),
// c the centre tap: the eigenvalues of the symmetric filter h dilated by m
// under the half-sample symmetric extension (period 2L), on the DCT-II basis.
std::vector<double> dct_response(const Full& f, int m, int L) {
  const int n = f.len, c = n / 2;
  const double pi = 3.14159265358979323846;
  std::vector<double> C((size_t)L * n);
  for (int k = 0; k < L; ++k)
    for (int t = 0; t < n; ++t) {
      long long q = ((long long)k * (t - c) * m) % (2LL * L);  // angle pi q / L, reduced exactly
      if (q < 0) q += 2LL * L;
      C[(size_t)k * n + t] = std::cos(pi * (double)q / L);
    }
  std::vector<double> T((size_t)n * L);  // T = h C^T
  for (int a = 0; a < n; ++a)
    for (int k = 0; k < L; ++k) {
      double s = 0.0;
      for (int b = 0; b < n; ++b) s += f.h[(size_t)a * n + b] * C[(size_t)k * n + b];
      T[(size_t)a * L + k] = s;
    }
  std::vector<double> R((size_t)L * L, 0.0);
  for (int kr = 0; kr < L; ++kr) {
    double* r = &R[(size_t)kr * L];
    for (int a = 0; a < n; ++a) {
      const double ca = C[(size_t)kr * n + a];
      const double* t = &T[(size_t)a * L];
      for (int kc = 0; kc < L; ++kc) r[kc] += ca * t[kc];
    }
  }
  return R;
}
}  // namespace

void dct_pyramid_masks(int L, int nscales, int pair, float* out) {
  const Maxflat& M = maxflat();
  const Full& f0 = M.full[pair == 0 ? 0 : 2];  // h0 / g0
  const Full& f1 = M.full[pair == 0 ? 1 : 3];  // h1 / g1
  const size_t pe = (size_t)L * L;
  const double inv = 1.0 / ((double)L * L);
  // the 2S level responses are independent: one host thread each
  std::vector<std::vector<double>> H0(nscales), H1(nscales);
  {
    std::vector<std::thread> th;
    for (int lev = 0; lev < nscales; ++lev) {
      th.emplace_back([&, lev] { H1[lev] = dct_response(f1, 1 << lev, L); });
      th.emplace_back([&, lev] { H0[lev] = dct_response(f0, 1 << lev, L); });
    }
    for (auto& t : th) t.join();
  }
  std::vector<double> P(pe, 1.0);
  for (int lev = 0; lev < nscales; ++lev) {
    float* o = out + (size_t)(nscales - 1 - lev) * pe;
    for (size_t i = 0; i < pe; ++i) o[i] = (float)(H1[lev][i] * P[i] * inv);
    for (size_t i = 0; i < pe; ++i) P[i] *= H0[lev][i];
  }
  float* o = out + (size_t)nscales * pe;
  for (size_t i = 0; i < pe; ++i) o[i] = (float)(P[i] * inv);
}

double maxflat_factor_error() {
  const Maxflat& M = maxflat();
  if (!M.ok) return INFINITY;
  return std::max(M.pair[0].err, M.pair[1].err);
}

}  // namespace shx
