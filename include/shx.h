/*
 * shx.h — C ABI of the MI355X (gfx950) nonsubsampled shearlet transform.
 *
 * Drop-in boundary for the 2D NSST hot path of ShearCuda.  Every entry point
 * takes plain device pointers (FP32, row-major L x L planes) and sizes, plus a
 * HIP stream (hipStream_t passed as void*; NULL = the default stream), and
 * returns SHX_OK or a negative status; shx_last_error() then holds a message.
 * No entry point exits the process.
 *
 * Coefficient layout ("cells", as in the MATLAB cell array the reference
 * returns, shearcuda/src_cuda/shear_trans_cuda.cpp:204-216): a host array of
 * S+1 device pointers.  Cell 0 is the lowpass (one L x L plane); cell j
 * (1..S, coarse to fine) holds D_j contiguous L x L planes, direction k at
 * offset k*L*L.  Array element [r][c] is MATLAB element (r+1, c+1).
 *
 * Thread-safety: a dictionary owns its workspaces; use one dictionary from one
 * stream at a time.  Distinct dictionaries are independent.
 */
#ifndef SHX_H
#define SHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHX_OK 0
#define SHX_ERR_ARG (-1)        /* invalid argument (size, count, pointer)   */
#define SHX_ERR_HIP (-2)        /* HIP runtime error                        */
#define SHX_ERR_ALLOC (-3)      /* device allocation failed                 */
#define SHX_ERR_UNSUPPORTED (-4)/* unsupported size or option               */

#define SHX_SHIFT_NONE 0   /* fft2(w, L, L): shearlet_toolbox_1/nsst_dec2.m:50 */
#define SHX_SHIFT_CENTER 1 /* padded filter rolled by -n1/2: the MEX GPU path,
                              prepare_shearing_filters_cuda.cpp:254            */

typedef struct shx_dict shx_dict;
typedef struct shx_batch shx_batch;
typedef struct shx_graph shx_graph;

/* Message of the last failing call on this thread ("" if none). */
const char* shx_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int shx_version(void);

/* Shearlet dictionary: normalised FP32 spectra of the directional filters of
 * every scale (computed in FP64 on the device from the spatial filters), FFT
 * twiddles, the DCT-domain pyramid masks and workspaces.
 * Replaces prepare_shearing_filters_cuda(w_s, L)
 * (shearcuda/src_cuda/prepare_shearing_filters_cuda.cpp:61-249).
 *   L        image side, power of two in [32, 1024]
 *   nscales  number of directional scales S (1..8)
 *   ndir[j]  directions of scale j (power of two, 1..32)
 *   dsize[j] spatial filter side n1 of scale j (<= L)
 *   w_s[j]   host FP64 array ndir[j] x dsize[j] x dsize[j] (row-major)
 *   shift    SHX_SHIFT_NONE or SHX_SHIFT_CENTER */
int shx_prepare_shearing_filters(int L, int nscales, const int* ndir, const int* dsize, const double* const* w_s,
                                 int shift, shx_dict** out);
/* The same with flags.  SHX_DICT_HOST_SPECTRA: spectra by an FP64 FFT on the
 * host and uploaded, instead of the FP64 direct transform on the device (both
 * round the same FP64 values once; they agree to FP32 rounding). */
#define SHX_DICT_HOST_SPECTRA 1
int shx_prepare_shearing_filters_ex(int L, int nscales, const int* ndir, const int* dsize, const double* const* w_s,
                                    int shift, int flags, shx_dict** out);
/* The whole dictionary from its parameters, everything on the device: the
 * 2^dcomp[j] spatial Meyer filters of side dsize[j] of every scale (FP64,
 * shx_meyer_filters), then their spectra.  Replaces
 * shearing_filters_Myer(dsize, dcomp, L, 'GPUsingle')
 * (shearcuda/shearing_filters_Myer.m:29-81, GPU branch, which builds the
 * filters on the host and calls prepare_shearing_filters_cuda).  flags as in
 * shx_prepare_shearing_filters_ex. */
int shx_shearing_filters_myer(int L, int nscales, const int* dcomp, const int* dsize, int shift, int flags,
                              shx_dict** out);
/* Spatial Meyer shearing filters of one scale on the device (FP64): the
 * pseudo-polar grid of an n1 x n1 block, 2^level Meyer windows over its 2n1
 * lines, each re-assembled on the block and brought to space by a centred
 * inverse DFT (shearlet_toolbox_1/shearing_filters_Myer.m:18-27 with
 * gen_x_y_cordinates.m:22-111, windowing.m:11-27, meyer_wind.m:11-19,
 * rec_from_pol.m:16-35).  w_dev: device array of 2^level * n1 * n1 doubles
 * (element [k][r][c] = MATLAB w_s(r+1, c+1, k+1)); 2*n1 must be a multiple
 * of 2^level.  Allocates scratch and synchronises the stream. */
int shx_meyer_filters(int n1, int level, double* w_dev, void* stream);
/* Frees a dictionary (clear_shearing_filters_cuda.cpp:51-96). NULL is a no-op. */
int shx_clear_shearing_filters(shx_dict* dict);
/* Dictionary shape: L, S and ndir[0..S-1] (ndir may be NULL). */
int shx_dict_info(const shx_dict* dict, int* L, int* nscales, int* ndir);
/* Device bytes held by the dictionary (spectra + workspaces). */
size_t shx_dict_bytes(const shx_dict* dict);

/* Forward NSST: atrousdec + per-scale FFT x filter bank -> bands.
 * Replaces shear_trans_cuda(x, 'maxflat', shear, Ct)
 * (shearcuda/src_cuda/shear_trans_cuda.cpp:56-217, convolution :220-287). */
int shx_shear_trans(shx_dict* dict, const float* x, float* const* coeff, void* stream);

/* Inverse NSST.  Replaces inverse_shear_cuda(Ct, 'maxflat', shear)
 * (shearcuda/src_cuda/inverse_shear_cuda.cpp:57-219, applyfilters :222-286). */
int shx_inverse_shear(shx_dict* dict, const float* const* coeff, float* x, void* stream);

/* Inverse NSST of hard-thresholded bands without modifying them: band b
 * (b = 0 lowpass, then scale-major, direction-minor) is read as
 * c * (|c| > th[b]).  th has 1 + sum(ndir) entries (host memory).  The
 * threshold is fused into the reconstruction's first read of each band. */
int shx_inverse_shear_thresh(shx_dict* dict, const float* const* coeff, const float* th, float* x, void* stream);

/* Plain denoise: shear_trans, nsst_HT thresholds T*sigma*s (T = Tscalars[0]
 * lowpass, [1] scales 1..S-1, [2] finest; shearlet_toolbox_1/nsst_HT.m:32-47)
 * fused into the inverse.  scalars: 1 + sum(ndir) per-band noise scalars
 * (lowpass first).  coeff: caller-owned band storage (receives the
 * unthresholded coefficients), or NULL when the caller needs only the image:
 * then the band coefficients are never written to memory (the reference's
 * demo_denoise/demoImageDenoise/mainwindow.cpp:223-233 drops them after the
 * reconstruction too). */
int shx_denoise(shx_dict* dict, const float* x, float* const* coeff, double sigma, const double* Tscalars,
                const double* scalars, float* out, void* stream);

/* Measurement hook: runs shx_denoise `reps` times with a HIP event after
 * every kernel launch on `stream` and returns each launch's mean duration
 * (stage_ms, milliseconds), the bytes it must move at minimum (stage_bytes:
 * each FP32 input and output it touches, once) and its name (names: max_stages
 * strings of name_len chars).  *nstages receives the number of launches.
 * No counterpart in the reference (its GpuTimes singleton,
 * shearcuda/src_cuda/GpuTimes.cpp:59-127, timed host calls with device
 * syncs). */
int shx_denoise_profile(shx_dict* dict, const float* x, float* const* coeff, double sigma, const double* Tscalars,
                        const double* scalars, float* out, int reps, int max_stages, double* stage_ms,
                        double* stage_bytes, char* names, int name_len, int* nstages, void* stream);

/* ---- batched form ----------------------------------------------------------
 * A batch plan runs nimg images of the dictionary's size L through every
 * kernel of the path at once (the image index is one more grid dimension), so
 * small images fill the GPU.  Buffers are contiguous device arrays:
 *   x, out : nimg x L x L
 *   coeff  : nimg x (1 + sum ndir) x L x L   (per image: lowpass, then the
 *            bands scale-major, direction-minor)
 *   th_dev : nimg x (1 + sum ndir) FP32 thresholds in device memory
 *            (per image: lowpass first), keep iff |c| > th.
 * No counterpart in the reference (its MEX entry points take one image). */
int shx_batch_create(shx_dict* dict, int nimg, shx_batch** out);
int shx_batch_destroy(shx_batch* batch);
size_t shx_batch_bytes(const shx_batch* batch);
int shx_batch_shear_trans(shx_batch* batch, const float* x, float* coeff, void* stream);
/* th_dev may be NULL (plain inverse). */
int shx_batch_inverse_shear(shx_batch* batch, const float* coeff, const float* th_dev, float* out, void* stream);
/* coeff may be NULL (no coefficients returned or written), as in shx_denoise. */
int shx_batch_denoise(shx_batch* batch, const float* x, float* coeff, const float* th_dev, float* out,
                      void* stream);

/* ---- adjoints (gradients) -------------------------------------------------------
 * Transposes of the analysis A = shear_trans and the synthesis
 * S = inverse_shear, for gradient computations (no counterpart in the
 * reference).  A^T maps coefficient gradients (cells) to an image gradient;
 * S^T maps an image gradient to coefficient gradients.  The directional stage
 * is a tight frame, so its transpose is the directional reconstruction; the
 * pyramid levels are transposed exactly, including the symmetric extension
 * (zero-padded correlation on the extended grid, then fold-add). */
int shx_shear_trans_adjoint(shx_dict* dict, const float* const* coeff_bar, float* x_bar, void* stream);
int shx_inverse_shear_adjoint(shx_dict* dict, const float* x_bar, float* const* coeff_bar, void* stream);
/* Batched forms (contiguous stacks as in shx_batch_shear_trans). */
int shx_batch_shear_trans_adjoint(shx_batch* batch, const float* coeff_bar, float* x_bar, void* stream);
int shx_batch_inverse_shear_adjoint(shx_batch* batch, const float* x_bar, float* coeff_bar, void* stream);

/* Soft-threshold shrinkage, one step of the unrolled separation
 * (shearcuda/separate.m:162-257 with the threshold of
 * shearcuda/SoftThresh.m:11-13, per band):
 *   out = S(soft(A x, th)),  soft(c, t) = sign(c) max(|c| - t, 0).
 * The soft threshold is applied as the reconstruction reads each band and the
 * lowpass (no thresholded copy of the coefficients is made); the result is
 * shx_soft_bands followed by shx_batch_inverse_shear up to FP32 rounding
 * (the fused row pass splits its FFT differently; thresholded values are
 * identical).
 * th_dev: device FP32, th_dev[i * th_stride + p] is the threshold of band p
 * (lowpass p = 0) of image i; th_stride 0 shares one row among the images,
 * otherwise th_stride >= 1 + nbands.  coeff (may be NULL) receives A x, the
 * unthresholded stack, which the backward takes.
 * Backward, given those coefficients c and the output gradient out_bar:
 *   x_bar = A^T(1[|c| > t] . S^T out_bar),
 *   gth_dev[i * th_stride + p] += sum over band p of image i of
 *                                 -sign(c) 1[|c| > t] (S^T out_bar)
 * (th_stride 0: every image adds into the one row; gth_dev is not cleared).
 * The mask and the threshold gradient are applied between S^T's last row
 * pass and A^T's first, in one kernel per band row.  No counterpart in the
 * reference (learned thresholds). */
int shx_batch_shrink(shx_batch* batch, const float* x, float* coeff, const float* th_dev, int th_stride, float* out,
                     void* stream);
int shx_batch_shrink_backward(shx_batch* batch, const float* coeff, const float* th_dev, int th_stride,
                              const float* out_bar, float* x_bar, float* gth_dev, void* stream);

/* Autotuner: times the pyramid tile shapes, the reconstruction's
 * directions-per-block grouping, the column tile width of the
 * decomposition and reconstruction column passes and the plan options
 * (shx_set_plan_options) on the plan's own workspace
 * (the dictionary's one-image plan when batch is NULL) and keeps the fastest.
 * chosen (may be NULL, else 4 ints) receives {pyramid variant, group,
 * decomposition column tile, reconstruction column tile}.  Each candidate is
 * timed as a captured graph replay on a private stream.  Allocates scratch
 * and synchronises: call it outside graph capture.  No counterpart in the
 * reference. */
int shx_autotune(shx_dict* dict, shx_batch* batch, int reps, int* chosen, void* stream);
/* Sets a schedule explicitly (pyramid variant 0..2, group 2/4/8, column tiles
 * 4 (L >= 128), 8, or 16 (L <= 512); 0 = default; negative: the three-pass
 * column FFT with |ct| columns, -4 or -8 at L = 512, -2 or -4 at L = 1024). */
int shx_set_schedule(shx_dict* dict, shx_batch* batch, int pyr_variant, int group, int ct_dec, int ct_rec);
/* Plan options (both picked by shx_autotune; defaults 0, 0; L >= 64 for 1).
 * With the pyramid's half-sample symmetric extension every a trous level is
 * a diagonal mask on the 2D DCT-II coefficients, so the dependent levels of
 * a pyramid collapse into three passes.  dct_front: the decomposition
 * (shear_trans, denoise) computes the lowpass and the scale spectra that way
 * instead of running the atrousdec.m levels and then the FFT of every
 * band-pass plane.  dct_back: the reconstruction (inverse_shear*, denoise)
 * goes from the direction-group partials to the image that way instead of
 * the scale row pass and the atrousrec.m levels.  Every option gives the same
 * results to FP32 rounding.  No counterpart in the reference. */
int shx_set_plan_options(shx_dict* dict, shx_batch* batch, int dct_front, int dct_back);
int shx_get_plan_options(const shx_dict* dict, const shx_batch* batch, int* dct_front, int* dct_back);

/* ---- HIP graphs --------------------------------------------------------------
 * Any sequence of shx calls that neither allocates nor synchronises
 * (shx_shear_trans, shx_inverse_shear*, shx_denoise, shx_batch_*) can be
 * recorded between shx_capture_begin/end on a non-default stream and replayed
 * with one shx_graph_launch; e.g. several batch plans of different sizes in
 * one launch.  Replays reuse the recorded pointers and thresholds. */
int shx_capture_begin(void* stream);
int shx_capture_end(void* stream, shx_graph** out);
int shx_graph_launch(shx_graph* graph, void* stream);
int shx_graph_destroy(shx_graph* graph);
/* One graph holding copies of n recorded graphs as child nodes with no order
 * between them, so one launch runs them all and the runtime may run them
 * concurrently (e.g. the batch plans of a mixed-size batch).  The parts stay
 * owned by the caller and may be destroyed afterwards.  No counterpart in the
 * reference. */
int shx_graph_parallel(shx_graph* const* parts, int n, shx_graph** out);
/* Records shx_denoise on an internal stream into a graph. */
int shx_denoise_graph(shx_dict* dict, const float* x, float* const* coeff, double sigma, const double* Tscalars,
                      const double* scalars, float* out, shx_graph** out_graph);

/* In-place hard threshold of coefficient cells, thresh_cuda(Ct, lambda,
 * option, E, sc, opt) (shearcuda/src_cuda/thresh_cuda.cpp:57-217,
 * shearcuda/thresh.m:28-55).
 *   cells[i] holds cell_planes[i] planes of plane_elems[i] floats.
 *   option 1 (wavelets): only cells[1], th = lambda.
 *   option 2 (shearlets): th(s,w) = sc[s] * lambda * E[s + w*E_rows]
 *       (E column-major E_rows x E_cols doubles, E_rows = ncells).
 *   opt 1: *maxi = max over thresholded bands of max|c|/E (else 0); the
 *       maximum accumulates in a file-scope device word (no allocation; host
 *       threads take turns on it) and the call synchronises the stream, so
 *       an opt 1 call cannot be recorded into a graph (opt 0 can). */
int shx_thresh(float* const* cells, int ncells, const int* cell_planes, const int64_t* plane_elems, double lambda,
               int option, const double* E, int E_rows, int E_cols, const double* sc, int n_sc, int opt,
               double* maxi, void* stream);

/* A trous maxflat pyramid on an L x L image (L power of two >= 32):
 * y[0] lowpass, y[1..nlevels] coarse -> fine; each an L x L device plane.
 * Replaces atrousdec / atrousrec (shearcuda/src_cuda/atrousdec.cpp:41-131,
 * atrousrec.cpp:42-123; demo_denoise/shearcuda/shearcuda.cpp:401-546).
 * work: 2 planes of scratch (L*L floats each) or NULL (allocated per call). */
int shx_atrousdec(const float* x, int L, int nlevels, float* const* y, float* work, void* stream);
/* shx_atrousdec without any allocation: the two level temporaries live in a
 * static device array of 2 * SHX_NOALLOC_MAX_L^2 floats (8 MiB) that every
 * device the library runs on holds until the process ends; L above
 * SHX_NOALLOC_MAX_L is rejected (SHX_ERR_ARG).  The results equal
 * shx_atrousdec's bit for bit.  It calls no hipMalloc/hipFree and does not
 * synchronise.  All calls on one device share that device's array, so calls
 * that overlap on one device (concurrent streams or host threads) race; serialise
 * them.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
#define SHX_NOALLOC_MAX_L 1024
int shx_atrousdec_noalloc(const float* x, int L, int nlevels, float* const* y, void* stream);
int shx_atrousrec(const float* const* y, int L, int nlevels, float* x, float* work, void* stream);

/* re_weight_band_cuda(x, coeff): atrousrec(coeff(j) .* atrousdec(x)) with
 * numel(coeff)-1 levels (shearcuda/src_cuda/re_weight_band_cuda.cpp:59-183).
 * work: ncoeff + 2 planes or NULL (allocated per call). */
int shx_re_weight_band(const float* x, int L, const double* coeff, int ncoeff, float* y, float* work, void* stream);

/* ---- undecimated periodic 2D wavelet transform (Rice Wavelet Toolbox) -------
 * [yl, yh] = mrdwt(x, h, L) and x = mirdwt(yl, yh, h, L)
 * (shearcuda/src/mrdwt_r.h:70-207, mirdwt_r.h:92-218; GPU entry points
 * shearcuda/src_cuda/mrdwt_cuda.cpp:58-253, mirdwt_cuda.cpp:58-255).
 * x, yl: nimg x m x n; yh: nimg x m x 3nL with level l in columns
 * [3n(l-1), 3nl) as [lh hl hh]; h: lh (even) scaling-filter taps (host FP64);
 * m and n divisible by 2^L.  work: 2*m*n*nimg floats or NULL (allocated).
 * The *_adjoint entry points apply the transposes (mrdwt^T, mirdwt^T); they
 * have no counterpart in the reference and serve gradient computations. */
int shx_mrdwt(const float* x, int m, int n, const double* h, int lh, int L, float* yl, float* yh, float* work,
              int nimg, void* stream);
int shx_mirdwt(const float* yl, const float* yh, int m, int n, const double* h, int lh, int L, float* x, float* work,
               int nimg, void* stream);
int shx_mrdwt_adjoint(const float* yl_bar, const float* yh_bar, int m, int n, const double* h, int lh, int L,
                      float* x_bar, float* work, int nimg, void* stream);
int shx_mirdwt_adjoint(const float* x_bar, int m, int n, const double* h, int lh, int L, float* yl_bar,
                       float* yh_bar, float* work, int nimg, void* stream);

/* Soft threshold with one threshold per band (th_dev, device FP32) over a
 * batched layout: element e of each image (img_elems elements) belongs to
 * band band_off + ((e mod rowlen) / seglen).  NSST stacks: rowlen =
 * img_elems, seglen = L*L; wavelet yh: rowlen = 3nL, seglen = n, band_off 1.
 * y = sign(x) max(|x| - t, 0) (shearcuda/SoftThresh.m:11-13, per band).
 * Backward: gx = gy * 1[|x| > t] (gx may be NULL) and
 * gth[band] += sum -sign(x) 1[|x| > t] gy (gth not cleared).  No
 * counterpart in the reference (learned thresholds). */
int shx_soft_bands(const float* x, float* y, int nimg, int64_t img_elems, int64_t rowlen, int64_t seglen,
                   int band_off, const float* th_dev, void* stream);
int shx_soft_bands_backward(const float* x, const float* gy, float* gx, float* gth_dev, int nimg, int64_t img_elems,
                            int64_t rowlen, int64_t seglen, int band_off, const float* th_dev, void* stream);

/* Per-band hard threshold y = x * 1[|x| > t] over the same batched layouts
 * (nsst_HT.m:39 / thresh_cuda's keep rule, one threshold per band), and its
 * input gradient gx = gy * 1[|x| > t].  The backward has no counterpart in the
 * reference (gradients through the hard threshold). */
int shx_hard_bands(const float* x, float* y, int nimg, int64_t img_elems, int64_t rowlen, int64_t seglen,
                   int band_off, const float* th_dev, void* stream);
int shx_hard_bands_backward(const float* x, const float* gy, float* gx, int nimg, int64_t img_elems, int64_t rowlen,
                            int64_t seglen, int band_off, const float* th_dev, void* stream);

/* Sum |x|^p over n elements, result (p-th root not taken) in *out (host).
 * norm_cuda's reduction (shearcuda/src_cuda/norm_cuda.cpp:172-215).
 * Accumulates in a file-scope device word (no allocation) and synchronises
 * the stream (not capturable). */
int shx_norm_pow(const float* x, int64_t n, double p, double* out, void* stream);

/* Median absolute deviation / 0.6745 of each of nplanes planes (plane i at
 * planes + i*plane_stride, plane_elems finite floats) into mad[i] (device
 * FP64), on the device: the per-band step of the noise scalars, which the
 * reference takes on the host (shearlet_toolbox_1/nsst_scalars.m:36-44;
 * demo_denoise/shearcuda/shearcuda.cpp:241-268 with its median at :165-201).
 * The medians are exact order statistics (mean of the two middle values for
 * an even count).  arith SHX_MAD_FP64: values widened to double (MATLAB /
 * numpy on the double transform); SHX_MAD_FP32: float arithmetic as the demo,
 * final division in double.  Each call is bit-identical to that host
 * computation. */
#define SHX_MAD_FP64 0
#define SHX_MAD_FP32 1
int shx_band_mad(const float* planes, long long plane_elems, long long plane_stride, int nplanes, int arith,
                 double* mad, void* stream);

/* Noise scalars of a dictionary: the transform of the given noise image
 * (L x L device FP32, normally unit white noise), then shx_band_mad of every
 * band, into scalars[0 .. nbands] on the host (lowpass first, then the bands
 * coarse -> fine); synchronises the stream.  Replaces nsst_scalars(L, shear,
 * 'maxflat') (shearlet_toolbox_1/nsst_scalars.m:28-45; arith SHX_MAD_FP64) and
 * the per-band loop of Shearcuda::computeShearNorm
 * (demo_denoise/shearcuda/shearcuda.cpp:237-268; SHX_MAD_FP32).  Allocates a
 * coefficient stack for the call. */
int shx_nsst_scalars(shx_dict* dict, const float* noise, int arith, double* scalars, void* stream);

/* The shipped image separation loop, [C, P] = separate(img, level2, itermax,
 * stop, gamma, coeff, 0, shear) (shearcuda/separate.m:41-257, GPU branch),
 * run natively over one workspace allocated per call:
 *   img       L x L device image (L = the dictionary's size)
 *   level2    levels of the undecimated wavelet of the point part
 *   coeff     ncoeff re_weight_band weights (ncoeff - 1 pyramid levels)
 *   qmf       lqmf orthonormal taps of that wavelet (host FP64; separate.m
 *             uses Symmlet 4)
 *   delta_max > 0: fixed starting threshold with sigma = 2.5
 *             (crack_detection/separateHairlineCrack.m:63-66); else the noise
 *             level and deltamax are estimated as separate.m:81-133 does
 *   E         band norms, column-major E_rows x E_cols host FP64 with
 *             E_rows = S+1 and E_cols >= max ndir, or NULL to compute them as
 *             com_norm.m:38-53 does (E_rows, E_cols ignored)
 *   C, P      L x L device outputs (curve part, point part)
 *   info      NULL or 3 doubles: sigma, deltamax, lambda.
 * The same library calls in the same order as the Python mirror
 * separation.separate; synchronises the stream before returning. */
int shx_separate(shx_dict* dict, const float* img, int level2, int itermax, double stop, double gamma,
                 const double* coeff, int ncoeff, const double* qmf, int lqmf, double delta_max, const double* E,
                 int E_rows, int E_cols, float* C, float* P, double* info, void* stream);

/* In-place soft threshold sign(v)*max(|v|-t, 0) (softth_cuda.cpp:49-112). */
int shx_softth(float* x, int64_t n, double t, void* stream);

/* Sum |a - b|^p over n elements into *out (host): the reduction of
 * Shearcuda::getNormErr (demo_denoise/shearcuda/shearcuda.cpp:672-717).
 * Accumulates in a file-scope device word (no allocation) and synchronises
 * the stream (not capturable). */
int shx_norm_pow_diff(const float* a, const float* b, int64_t n, double p, double* out, void* stream);

/* x += N(0, sigma^2) noise from a counter-based generator (splitmix64 of
 * (seed, index), Box-Muller): Shearcuda::addNoise
 * (demo_denoise/shearcuda/shearcuda.cpp:140-163), which draws from cuRAND
 * with seed 1234; the streams differ, the distribution is the same. */
int shx_add_noise(float* x, int64_t n, double sigma, uint64_t seed, void* stream);

/* The directional halves of the transform on given pyramid levels, as
 * Shearcuda::convolution / Shearcuda::applyfilters
 * (demo_denoise/shearcuda/shearcuda.cpp:549-624,719-800).  y: S+1 device
 * planes, y[0] the lowpass, y[j] (1..S, coarse to fine) the band-pass level
 * of scale j.  convolution: coeff cells = directional bands of y (coeff[0]
 * receives a copy of y[0]).  applyfilters: y[1..S] = the conjugate-filtered
 * band sums of coeff (the input of the pyramid reconstruction). */
int shx_convolution(shx_dict* d, const float* const* y, float* const* coeff, void* stream);
/* Device image conversion between FP32 and 8-bit (GpuImage::convert,
 * demo_denoise/shearcuda/gpuimage.cpp:121-140 with shear_cuda.cu:2039-2064):
 * float -> uint8 saturates to [0, 255] and truncates. */
#define SHX_TYPE_FLOAT 0
#define SHX_TYPE_UINT8 5
int shx_convert(void* dst, int dst_type, const void* src, int src_type, int64_t n, void* stream);
int shx_applyfilters(shx_dict* d, const float* const* coeff, float* const* y, void* stream);

/* Largest |h - sum_k v_k w_k^T| of the separable factors the pyramid kernels
 * use for the four maxflat filters (FP64; runs on the host, no GPU needed). */
double shx_maxflat_factor_error(void);
/* Copy maxflat filter which (0 h0, 1 h1, 2 g0, 3 g1) into out (row-major,
 * len*len doubles, len returned); out may be NULL to query len. */
int shx_maxflat_filter(int which, double* out, int* len);

/* ---- 3D shearlet volume transform -------------------------------------------
 * Volumes are n^3 FP32 arrays on the device; a volume and the filters applied
 * to it share one memory order (the transform is the same for any order).
 * Batches of nvol volumes are contiguous.
 *
 * Subband block ("subs", per volume shx3_subs_elems() floats): the band-pass
 * subbands of the Fourier-domain pyramid, finest first, each side_l^3, then
 * the lowpass side_L^3 (the cell L of DoPyrDec,
 * shearcuda3d/3DShearTrans/DoPyrDec.m:1-6 -> 3DBP/PyrNDDec_mm.m:48-109).
 * Band block ("bands", per volume shx3_bands_elems() floats): the shearlet
 * bands of every level, level-major; within a level cone-major (cones
 * 1, 2, 3 of GetMeyerBasedFilter's output), then the MATLAB linear
 * (column-major) index over the (l2, l1) direction cell; each side_l^3.
 * Thresholds th: device FP32 [nvol][nbands] in band order, a band kept where
 * |c| > th (shearcuda3d/DenoiseDemo/ThresholdShCoeff.m:7 zeroes |c| <= th).
 */
#define SHX3_MAX_LEVELS 6
typedef struct shx3_dict shx3_dict;
typedef struct shx3_plan shx3_plan;

/* Dictionary for volumes of side n: nlevels band-pass levels with
 * Pyr_mode 1, 1.5 or 2 (3DBP/PyrNDDec_mm.m:52-70), ndir[l] directions per
 * axis at level l (3 * ndir[l]^2 bands), and the spatial cone filters of each
 * level (host FP32, 3*ndir[l]^2 filters of fsize^3 in band order), the output
 * of shearcuda3d/3DShearTrans/GetMeyerBasedFilter.m:12-151.  Levels passing
 * the same filter pointer with the same direction count share spectra.
 * Supported: band-pass sides in {8..128} (powers of two), fsize a multiple of
 * 8 up to 64.  Replaces the per-call filter padding and R2C FFT of
 * ShDecThRec_cuda.cpp:282-296. */
int shx3_prepare(int n, int nlevels, double pyr_mode, const int* ndir, int fsize, const float* const* filters,
                 shx3_dict** out);
int shx3_destroy(shx3_dict* d);
/* sides[nlevels+1] (subband sides, lowpass last), fft_sides[nlevels] (FFT
 * grid of each level's directional filtering), band_first[nlevels+1]. */
int shx3_info(const shx3_dict* d, int* sides, int* fft_sides, int* band_first);
long long shx3_subs_elems(const shx3_dict* d);
long long shx3_bands_elems(const shx3_dict* d);
size_t shx3_dict_bytes(const shx3_dict* d);

/* Workspace for batches of nvol volumes; group = bands per directional launch. */
int shx3_plan_create(const shx3_dict* d, int nvol, int group, shx3_plan** out);
int shx3_plan_destroy(shx3_plan* p);
size_t shx3_plan_bytes(const shx3_plan* p);
/* Schedule of the directional loop: bands per launch (1 .. the group the plan
 * was created with) and the fused plane kernel's output columns per block
 * (0 automatic, 32 or 64; -1, the default, runs the row and column passes as
 * two kernels). */
int shx3_set_schedule(shx3_plan* p, int group, int plane_cols);

/* DoPyrDec / DoPyrRec on the GPU (3DBP/PyrNDDec_mm.m:48-109,
 * PyrNDRec_mm.m:46-95): x (nvol volumes) <-> subs. */
int shx3_pyr_dec(shx3_plan* p, const float* x, float* subs, void* stream);
int shx3_pyr_rec(shx3_plan* p, const float* subs, float* x, void* stream);
/* ShDecThRec_cuda(F, BP, level, ..., nsstScalars, sigma, T)
 * (shearcuda3d/3DShearTrans/src_cuda/ShDecThRec_cuda.cpp:74-174,178-335):
 * for every level, the sum over cones and directions of the hard-thresholded
 * 'same' convolutions of subs level l, written to the level slots of rec
 * (subband layout; the lowpass slot is not touched). */
int shx3_dec_thresh_rec(shx3_plan* p, const float* subs, const float* th, float* rec, void* stream);
/* ShDec for every cone (3DShearTrans/ShDec.m:18-34): all bands. */
int shx3_analysis(shx3_plan* p, const float* subs, float* bands, void* stream);
/* Transpose of shx3_analysis: level slots of subs_bar = sum over the level's
 * bands of the correlation with the band filter.  No counterpart in the
 * reference (gradients and adjoint checks). */
int shx3_analysis_adjoint(shx3_plan* p, const float* bands, float* subs_bar, void* stream);
/* DenoiseDemo/main.m:100-150: DoPyrDec, ShDecThRec, lowpass passed through,
 * DoPyrRec, on nvol volumes. */
int shx3_denoise(shx3_plan* p, const float* x, const float* th, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SHX_H */
