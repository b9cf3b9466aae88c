// Launch interface of the a trous pyramid kernels.
#pragma once

#include "shx_common.h"

namespace shx {

// Separable factors of a symmetric LEN x LEN filter of rank RANK:
//   h(a,b) = sum_k v[k][a] * w[k][b]
template <int LEN, int RANK>
struct SepFilter {
  float v[RANK][LEN];
  float w[RANK][LEN];
};

struct DecFilters {
  SepFilter<13, 2> h0;
  SepFilter<19, 4> h1;
};

struct RecFilters {
  SepFilter<19, 3> g0;
  SepFilter<13, 3> g1;
};

// Maxflat filter tables (FP64) and their separable factors, built once.
const DecFilters& dec_filters();
const RecFilters& rec_filters();
// Full FP64 filters, row-major LEN x LEN: which = 0 h0, 1 h1, 2 g0, 3 g1.
const double* maxflat_filter(int which, int* len);
// Largest |h - sum_k v_k w_k^T| over the four filters (FP64 factors).
double maxflat_factor_error();
// True when every filter has the rank the kernels are specialised for.
bool maxflat_ok();

// Per-image element strides of the three planes a level touches
// (dec: x, ylow, yhigh; rec: xlow, yhigh, out).  0 for a single image.
struct ImgStrides {
  long long a, b, c;
};

// Hard threshold applied to the rec level's xlow input: none, one value, or
// one per image in device memory (dev[img * stride]).
struct LowThresh {
  int mode;  // 0 none, 1 v, 2 dev
  float v;
  const float* dev;
  int stride;
};

hipError_t launch_atrous_dec(const float* x, float* ylow, float* yhigh, const ImgStrides& st, int nimg, int L, int m,
                             const DecFilters& f, hipStream_t s);
hipError_t launch_atrous_rec(const float* xlow, const float* yhigh, float* out, const ImgStrides& st, int nimg, int L,
                             int m, const RecFilters& f, const LowThresh& lt, float sx, float sy, hipStream_t s);

}  // namespace shx
