// Launch interface of the FFT-domain directional filtering kernels.
#pragma once

#include "shx_common.h"

namespace shx {

// A set of planes spread over up to SHX_MAX_SCALES base pointers ("cells"):
// plane p lives in cell j with off[j] <= p < off[j+1], at
// cell[j] + (p - off[j]) * plane_elems.
struct PlaneSet {
  int ncells;
  int off[SHX_MAX_SCALES + 2];
  float* cell[SHX_MAX_SCALES + 1];
};

// Per-band hard threshold (keep iff |v| > v[p]).
struct BandThresh {
  float v[SHX_MAX_BANDS];
};

// Band p -> scale, and the first band of each scale.
struct BandMap {
  int nscales;
  int nbands;
  int first[SHX_MAX_SCALES + 1];
  unsigned char scale_of[SHX_MAX_BANDS];
};

inline PlaneSet single_cell(float* base, int nplanes) {
  PlaneSet s{};
  s.ncells = 1;
  s.off[0] = 0;
  s.off[1] = nplanes;
  s.cell[0] = base;
  return s;
}

hipError_t launch_rows_r2c(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, const BandThresh* th,
                           const float2* tw, hipStream_t s);
hipError_t launch_rows_c2r(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, const float2* tw,
                           hipStream_t s);
hipError_t launch_cols_dec(int L, const float2* Y1, float2* Tout, const float2* Hm, const float2* Hb,
                           const BandMap& map, const float2* tw, hipStream_t s);
hipError_t launch_cols_rec(int L, const float2* U, float2* V, const float2* Hm, const float2* Hb,
                           const BandMap& map, const float2* tw, hipStream_t s);

}  // namespace shx
