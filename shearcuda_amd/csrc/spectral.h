// Launch interface of the FFT-domain directional filtering kernels.
#pragma once

#include "shx_common.h"

namespace shx {

// A set of planes spread over up to SHX_MAX_SCALES+1 base pointers ("cells"):
// plane p of image b lives in cell j with off[j] <= p < off[j+1], at
//   cell[j] + (p - off[j]) * plane_elems + b * img_stride.
struct PlaneSet {
  int ncells;
  int off[SHX_MAX_SCALES + 2];
  float* cell[SHX_MAX_SCALES + 1];
  long long img_stride;  // floats between consecutive images (0: one image)
};

// Thresholds for a row pass: none, per band in the kernel arguments (one
// image), or per image and band in device memory (dev_stride 0: one row of
// band thresholds for every image); hard (keep iff |c| > t) or soft
// (sign(c) max(|c| - t, 0)).
struct ThreshSpec {
  int mode;              // 0 none, 1 arg[p], 2 dev[b * dev_stride + dev_off + p]
  int dev_stride;
  int dev_off;
  int soft;              // 0 hard, 1 soft
  const float* dev;
  float arg[SHX_MAX_BANDS];
};

// Output plane p of a summing row pass reads input planes first[p]..first[p+1]-1.
struct SumTable {
  int first[SHX_MAX_SCALES + 1];
};

// Band p -> scale, first band of each scale, the stored filter plane of each
// band, and the direction groups the reconstruction column pass splits each
// scale into.  Scales whose spatial filters are identical share their
// spectra: filt_of[p] is the plane of the filter spectrum table band p reads.
// The column passes dispatch their planes in the orders below, so that the
// `share` bands (and the `gshare` groups) reading the same filter planes are
// neighbours in the grid and run on one XCD, whose L2 then serves the second
// read of each filter tile.
struct BandMap {
  int nscales;
  int nbands;
  int first[SHX_MAX_SCALES + 1];
  unsigned char scale_of[SHX_MAX_BANDS];
  unsigned char filt_of[SHX_MAX_BANDS];
  int ngroups;
  int group_first[SHX_MAX_SCALES + 1];        // first group of each scale
  unsigned short group_band0[SHX_MAX_BANDS + 1];  // band range of group g: [band0[g], band0[g+1])
  int share, gshare;                           // bands / groups per shared filter plane set (1: none)
  unsigned char order[SHX_MAX_BANDS];          // dispatch slot -> band
  unsigned short gorder[SHX_MAX_BANDS];        // dispatch slot -> group
};

// Fill the group fields of a BandMap: each scale's directions are split into
// groups of at most `gsize` consecutive directions; then the dispatch orders
// of bands and groups from filt_of.
void band_map_groups(BandMap& m, int gsize);

inline PlaneSet single_cell(float* base, int nplanes, long long img_stride = 0) {
  PlaneSet s{};
  s.ncells = 1;
  s.off[0] = 0;
  s.off[1] = nplanes;
  s.cell[0] = base;
  s.img_stride = img_stride;
  return s;
}

// All launches take nimg images; workspace strides are per image in floats.
hipError_t launch_rows_r2c(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, int nimg,
                           const ThreshSpec& th, const float2* tw, hipStream_t s);
// out plane p = C2R of the sum of input planes sum_first[p] .. sum_first[p+1]-1
// (sum_first == nullptr: input plane p).
hipError_t launch_rows_c2r(int L, const PlaneSet& in, const PlaneSet& out, int nplanes, int nimg,
                           const int* sum_first, const float2* tw, hipStream_t s);
// Denoise fusion: c = C2R(T) written to `band`, then U = R2C(keep(c, th)).
hipError_t launch_rows_c2r_r2c(int L, const PlaneSet& T, const PlaneSet& band, const PlaneSet& U, int nplanes,
                               int nimg, const ThreshSpec& th, const float2* tw, hipStream_t s);
// Backward of a soft-thresholded fused read (th.mode 2): U = R2C(C2R(T) *
// 1[|c| > t]) with c the forward coefficients (cref), and the threshold
// gradient sum -sign(c) 1[|c| > t] C2R(T) added into gth at the threshold's
// index.
hipError_t launch_rows_c2r_mask_r2c(int L, const PlaneSet& Tin, const PlaneSet& cref, const PlaneSet& U, int nplanes,
                                    int nimg, const ThreshSpec& th, float* gth, const float2* tw, hipStream_t s);
// ct: column tile width of the block (4, 8 or 16 where col_tile_ok; 0 = the
// default for L).  pre: Y1 is already column-transformed (launch_cols_fwd).
bool col_tile_ok(int L, int ct);
hipError_t launch_cols_dec(int L, const float2* Y1, long long y1_stride, float2* Tout, long long t_stride,
                           const float2* Hm, const float2* Hb, const BandMap& map, int nimg, const float2* tw,
                           hipStream_t s, int ct = 0, bool pre = false);
// In-place forward column FFT of nplanes packed spectra (stride y1_stride per image).
hipError_t launch_cols_fwd(int L, float2* Y1, long long y1_stride, int nplanes, int nimg, const float2* tw,
                           hipStream_t s, int ct = 0);
// DCT-domain pyramid decomposition (L >= 64), steps 0..2: 0: xwork = column
// DCT-II of x; 1: rows (DCT-II, mask plane p, inverse, row FFT) into the
// packed spec planes p < nspec and the real lowpass rows; 2: columns (inverse
// DCT-II, column FFT) in place.  mask: nspec + 1 planes of L*L, plane nspec
// the lowpass.  dtw[k] = e^{-i pi k / (2L)}, k < L.
hipError_t launch_dct_dec(int L, int step, const float* x, long long xs, float* xwork, long long ws, float2* spec,
                          long long ss, int nspec, float* low, long long ls, const float* mask, int nimg,
                          const float2* tw, const float2* dtw, hipStream_t s);
// DCT-domain pyramid reconstruction (L >= 64), steps 0..2: 0: D planes
// p < nspec = row DCT-II of C2R(sum of Qg planes gfirst[p]..gfirst[p+1]-1),
// plane nspec = row DCT-II of the lowpass (hard threshold lt); 1: column
// DCT-II, times mask plane p, inverse column DCT-II, in place; 2: x = inverse
// row DCT-II of the sum of the nspec + 1 planes.  mask: nspec + 1 planes of
// L*L.  lt is declared with the pyramid kernels.
struct LowThresh;
hipError_t launch_dct_rec(int L, int step, const float2* Qg, long long qs, const int* gfirst, int nspec,
                          const float* low, long long ls, const LowThresh& lt, float* D, long long ds,
                          const float* mask, float* x, long long xs, int nimg, const float2* tw, const float2* dtw,
                          hipStream_t s);
hipError_t launch_cols_rec(int L, const float2* U, long long u_stride, float2* Q, long long q_stride,
                           const float2* Hm, const float2* Hb, const BandMap& map, int nimg, const float2* tw,
                           hipStream_t s, int ct = 0);

}  // namespace shx
