/*
 * shwfs.h -- C ABI of the MI355X Shack-Hartmann wavefront-sensor library (libshwfs.so).
 *
 * Plain C types only: every array argument is a device pointer on the plan's device,
 * every stream is a hipStream_t passed as void*.  All functions return 0 on success or a
 * negative error code; shwfs_last_error() returns the message of the last failure on the
 * calling thread.  No function allocates device memory except the *_create calls, the set-up
 * calls shwfs_plan_set_pupil (resampled plans), shwfs_plan_set_resampling and shwfs_plan_bind_slopec,
 * and the first VJP call on a non-Parseval or resampled plan (its scratch).
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * SPECULA checkout).  The reference path is Python; the ctypes binding a maintainer
 * adds on the reference side is shown in INTEGRATION.md.
 */
#ifndef SHWFS_H
#define SHWFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHWFS_OK              0
#define SHWFS_ERR_ARG        -1   /* invalid argument / unsupported geometry            */
#define SHWFS_ERR_HIP        -2   /* HIP runtime error                                  */
#define SHWFS_ERR_UNSUPPORTED -3  /* geometry outside the compiled kernel instantiations */

typedef struct shwfs_plan shwfs_plan;
typedef struct shslopec_plan shslopec_plan;

/* Geometry of one SH sensor, derived on the host by the restatement of
 * SH.set_in_ef / calc_trigger_geometry (processing_objects/sh.py:112-323). */
typedef struct shwfs_geom {
    int32_t n_lenses;      /* subapertures per side: Lenslet.dimx == dimy (sh.py:67)          */
    int32_t ef_size;       /* input ElectricField side [px]                                   */
    int32_t np_sub;        /* phase px per subaperture side: SH._ovs_np_sub (sh.py:218)       */
    int32_t fft_size;      /* reference zero-padded FFT size N: SH._fft_size (sh.py:219)      */
    int32_t spec_size;     /* M: power-spectrum grid, power of two >= 2*np_sub-1             */
    int32_t subap_npx;     /* CCD px per subaperture side: SH._subap_npx                     */
    int32_t lgs;           /* 1: per-subaperture ConvolutionKernel correlation (sh.py:382-394) */
    int32_t parseval;      /* 1: sum of the binned image = norm_const * sum(A^2) exactly      */
    int32_t max_frames;    /* largest n_frames accepted by shwfs_forward                      */
    int32_t variant;       /* kernel choice: 0 auto, 1 register GEMV binning, 2 MFMA binning  */
                           /* (FP32), 3 MFMA binning from split bf16 pieces (hi + lo)        */
    double  k_rad_per_nm;  /* 2*pi/lambda (electric_field.py:65)                              */
    double  tilt_rad;      /* pi/N: half-pixel tilt coefficient (sh.py:469-477)               */
    double  tilt_a0;       /* first tilt abscissa: -ceil(np_sub/2)                            */
    double  norm_const;    /* parseval mode: sum over all CCD px of the unscaled image / sum(A^2) */
    double  pixel_area;    /* ElectricField.pixel_pitch**2 (electric_field.py:87-89)           */
} shwfs_geom;

/* Library identification. */
const char* shwfs_version(void);
const char* shwfs_last_error(void);

/* Create a sensor plan on `device`.
 *   row_w, col_w : binning operators (host, FP64).
 *       NGS: real, subap_npx x spec_size, row-major:  B = row_w . S . col_w^T, where S is the
 *            spec_size^2 power spectrum of one subaperture.  Encodes fftshift (sh.py:396),
 *            FoV mask (sh.py:399), FoV cut (sh.py:405) and the CCD rebin (lib/toccd.py:33-63).
 *       LGS: complex interleaved (re,im), subap_npx x (2*np_sub-1):  B = Re(row_w . D . col_w^T)
 *            with D the kernel-weighted autocorrelation on the lag grid.
 * Replaces the geometry/state that SH.setup() builds (sh.py:439-467). */
int shwfs_plan_create(const shwfs_geom* geom, const double* row_w, const double* col_w,
                      int device, shwfs_plan** out);

/* As shwfs_plan_create for a FoV mask that is not separable (make_mask with square=False
 * clipping the cut window, sh.py:282): the mask is given as n_terms separable terms, row_w and
 * col_w holding n_terms operators of the shape above one after the other, and
 * B = sum_k row_k . S . col_k^T (NGS) or Re sum_k row_k . D . col_k^T (LGS).  Such a plan runs on
 * the any-size path (variant 0 or 4) and normalises the image by its own sum (parseval = 0);
 * n_terms = 1 is shwfs_plan_create.  Replaces the same set-up as shwfs_plan_create with
 * squaremask=False (sh.py:50, 282). */
int shwfs_plan_create_masked(const shwfs_geom* geom, int n_terms, const double* row_w, const double* col_w,
                             int device, shwfs_plan** out);
int shwfs_plan_destroy(shwfs_plan* plan);

/* LGS only: per-subaperture kernel spectra sampled on the lag grid, device complex64
 * interleaved, layout [n_lenses^2][2*np_sub-1][2*np_sub-1], subaperture s = row*n + col,
 * lag index (ly + np_sub-1, lx + np_sub-1).  Copied into the plan (device-to-device).  The
 * spectra must be those of real kernels (Hermitian: K(-l) = conj K(l)), as ConvolutionKernel's
 * are; the kernels use only the lx >= 0 half.
 * Replaces ConvolutionKernel.kernels as read at sh.py:383-388. */
int shwfs_set_kernels(shwfs_plan* plan, const float* kspec_lags, void* stream);

/* SH forward for n_frames ElectricFields: one CCD intensity per frame, normalised to
 * S0 * masked_area (sh.py:422-428).  A, phase_nm: ef_size x ef_size float32 per frame
 * (a_frame_stride, phase_frame_stride may be 0 to share one amplitude or phase); ccd:
 * (n*subap_npx)^2 float32 per frame.
 * Replaces SH.trigger_code + SH.post_trigger (processing_objects/sh.py:348-428). */
int shwfs_forward(shwfs_plan* plan, int n_frames,
                  const float* A, int64_t a_frame_stride,
                  const float* phase_nm, int64_t phase_frame_stride,
                  double S0, float* ccd, int64_t ccd_frame_stride, void* stream);

/* Register a pupil amplitude that does not change between calls: its statistics (sum(A),
 * sum(A^2) and the Parseval flux scale for S0) are computed once here, on `stream`, and later
 * shwfs_forward / shwfs_forward_slopes / shwfs_vjp calls with the same A pointer, frame stride and
 * S0 reuse them instead of recomputing them per call.  For a single-frame pupil (frame stride 0,
 * no resampling) it also builds the lit-first subaperture order used by the Parseval NGS and LGS
 * kernels (subapertures with a nonzero pupil sample first): a performance hint only, the images
 * are the same with or without it.  The caller must register A again after changing its
 * contents; A = NULL forgets the registration.  No reference counterpart: SH recomputes the masked
 * area from in_ef.A on every trigger (processing_objects/sh.py:422-428). */
int shwfs_plan_set_pupil(shwfs_plan* plan, const float* A, int64_t a_frame_stride, int n_frames,
                         double S0, void* stream);

/* 1 when (A, a_frame_stride, S0) is the pupil registered with shwfs_plan_set_pupil, 0 otherwise.
 * Calls of shwfs_forward with another amplitude or S0 keep the registration (its statistics live
 * in buffers of their own); shwfs_plan_set_resampling and shwfs_plan_set_pupil(NULL) drop it.
 * No reference counterpart (the registration is a performance hint). */
int shwfs_plan_pupil_registered(const shwfs_plan* plan, const float* A, int64_t a_frame_stride, double S0);

/* Field resampling for fov_ovs != 1, phase rotation or shift (SH.setup, sh.py:448-458, and the
 * interpolation block of SH.trigger_code, sh.py:350-360).  After this call shwfs_forward and
 * shwfs_forward_slopes take A and phase_nm on the input grid (ef_in x ef_in; the plan's
 * ef_size is the resampled side): the phase is extended beyond the pupil edge with the sources
 * ext_src1 / ext_src2 (host int32 maps, ef_in^2, -1 = none; pixel p takes phase[src1], or
 * 2 phase[src1] - phase[src2] when src2 >= 0, extrapolate_edge_pixel.py:4-29), then A and the
 * phase are resampled bilinearly at input coordinates y = a0 Y + a1 X + a2, x = a3 Y + a4 X + a5
 * (clamped; affine = host double[6]) of output pixel (Y, X) (Interp2D, interp2d.py:48-122).
 * The masked area stays that of the input amplitude.  NULL maps: no extrapolation. */
int shwfs_plan_set_resampling(shwfs_plan* plan, int ef_in, const int32_t* ext_src1, const int32_t* ext_src2,
                              const double* affine);

/* Optional timing hook: when both events are non-null (hipEvent_t), shwfs_forward records
 * `start` immediately before and `stop` immediately after the main SH kernel launch, on the
 * call's stream, so that the caller can time that kernel alone.  Pass nulls to disable.
 * No reference counterpart (measurement support). */
int shwfs_plan_set_timing_events(shwfs_plan* plan, void* start, void* stop);

/* Launch-shape knob of the NGS MFMA path: each block streams `fpb` consecutive frames of its
 * subaperture tile, prefetching the next frame's pupil samples (1 to 64; default 4).
 * Results do not depend on it.  No reference counterpart (tuning). */
int shwfs_plan_set_frames_per_block(shwfs_plan* plan, int fpb);

/* ShSlopec plan.  Either `block_origin` (n_sub linear pixel offsets of square np_sub^2
 * blocks, row stride frame_w) or `idx` (n_sub x np_sub^2 linear pixel indices, int32) must
 * be given (host pointers).  xw, yw, mw: np_sub^2 weights (host, FP64): the x/y CoG weights
 * and mask_weighted of ShSlopec.computeXYweights (processing_objects/sh_slopec.py:86-132).
 * Replaces the state built by ShSlopec.__init__/set_xy_weights (sh_slopec.py:26-84). */
int shslopec_plan_create(int n_sub, int np_sub, int frame_h, int frame_w,
                         const int32_t* block_origin, const int32_t* idx,
                         const double* xw, const double* yw, const double* mw,
                         int max_frames, int device, shslopec_plan** out);
int shslopec_plan_destroy(shslopec_plan* plan);

/* Thresholded centre of gravity for n_frames pixel frames.
 *   thr_mode: 0 = subtract thr then clamp at 0 (thr = 0 when no threshold is set),
 *             1 = pedestal: zero the pixels below thr (sh_slopec.py:374-386)
 *   slopes  : float32, 2*n_sub per frame, [sx..., sy...] (data_objects/slopes.py:27-32)
 *   flux    : FP64, n_sub per frame (pre-threshold subaperture sums, sh_slopec.py:365)
 *   totals  : FP64, 2 per frame: sum and mean of flux (sh_slopec.py:432-434)
 * Replaces ShSlopec.calc_slopes_nofor (processing_objects/sh_slopec.py:308-438). */
int shslopec_forward(shslopec_plan* plan, int n_frames,
                     const float* pixels, int64_t pixel_frame_stride,
                     float thr, int thr_mode,
                     float* slopes, int64_t slopes_frame_stride,
                     double* flux, double* totals, void* stream);

/* Fused SH forward + ShSlopec: the centre of gravity of each valid subaperture is taken from
 * the binned CCD tile while it is still on chip, so the CCD frames are written once and never
 * read back.  Results equal shwfs_forward followed by shslopec_forward on the CCD frames
 * (same outputs and conventions as those two calls; the moment sums are FP64), to FP32 rounding:
 * the flux of a subaperture is its exact value from the Parseval identity of this geometry
 * (frame scale x norm_const x the subaperture's sum(A^2)) rather than the sum of its FP32 pixels,
 * and without threshold the moments keep the FP32 image's rounding-level negative pixels, which the
 * reference's clamp at zero (sh_slopec.py:382-386) would remove from its exact, non-negative image.
 * shwfs_plan_bind_slopec checks that the pair can be fused: Parseval normalisation and an NGS
 * MFMA plan with np_sub 8 and 8 or 16 px per subaperture, or an LGS plan on a specialised LGS
 * kernel (np_sub 6, 8, 12 or 16, at most 16 px); ShSlopec plan made from block origins lying on
 * the lenslet grid of this CCD, with the default separable weights.  It returns
 * SHWFS_ERR_UNSUPPORTED otherwise (then call the two functions).  The bound ShSlopec plan must
 * outlive its use; bind NULL to unbind.  Replaces the SH -> ShSlopec hand-over of one loop step
 * (processing_objects/sh.py:348-428 followed by sh_slopec.py:308-438). */
int shwfs_plan_bind_slopec(shwfs_plan* plan, shslopec_plan* slopec);
int shwfs_forward_slopes(shwfs_plan* plan, int n_frames,
                         const float* A, int64_t a_frame_stride,
                         const float* phase_nm, int64_t phase_frame_stride,
                         double S0, float* ccd, int64_t ccd_frame_stride,
                         float thr, int thr_mode,
                         float* slopes, int64_t slopes_frame_stride,
                         double* flux, double* totals, void* stream);

/* ---- adjoint (vector-Jacobian products) ------------------------------------------------
 * No reference counterpart: the reference builds interaction matrices by push-pull
 * (processing_objects/im_rec_calibrator.py:42-82, lib/modal_pushpull_signal.py:35-65).
 *
 * shslopec_vjp: dL/dpixels from dL/dslopes for the CoG of shslopec_forward (moments are
 * recomputed; the dim-subaperture factor is held fixed).  grad_pixels is fully overwritten
 * (zero outside the subapertures), frame stride = frame_h*frame_w. */
int shslopec_vjp(shslopec_plan* plan, int n_frames, const float* pixels, int64_t pixel_frame_stride,
                 float thr, int thr_mode, const float* grad_slopes, int64_t grad_slopes_stride,
                 float* grad_pixels, void* stream);

/* shwfs_vjp: dL/dphase_nm from dL/d(CCD intensity) for shwfs_forward, for every plan: NGS and LGS
 * (kernel spectra set), specialised kernels and the any-size path (variant 4, round FoV masks).
 * NGS plans with (np_sub, spec_size) in {(4,8),(6,16),(8,16),(16,32)} use register-resident
 * kernels; the others (LGS, any np_sub, spec_size a power of two in [16, 1024]) run full M-point
 * FFTs per subaperture with per-block scratch (at most 2 GiB, allocated by the first call and kept
 * by the plan).  phase_frame_stride may be 0 to evaluate several cotangents at one phase.  Any normalisation: with the Parseval scale (no cut, all-ones mask) the scale does not
 * depend on the phase; otherwise the image I = c(phi) B is normalised by its own sum and the VJP
 * is taken exactly (the forward runs first into plan scratch to get c and <g, I> / sum(I)).  A
 * resampled plan (shwfs_plan_set_resampling) takes A, phase and returns the gradient on the input
 * grid: the sensor-grid gradient goes through the transpose of the resampling (float atomics:
 * summation order not fixed).  The first call on a non-Parseval or resampled plan allocates its
 * scratch (kept by the plan). */
int shwfs_vjp(shwfs_plan* plan, int n_frames, const float* A, int64_t a_frame_stride,
              const float* phase_nm, int64_t phase_frame_stride, double S0,
              const float* grad_ccd, int64_t grad_ccd_stride, float* grad_phase, int64_t grad_phase_stride,
              void* stream);

/* As shwfs_vjp, but the gradient is returned on the sensor grid (the plan's ef_size, i.e. with
 * respect to the resampled phase; same as shwfs_vjp for a plan without resampling).  With it the
 * slope Jacobian of a resampled sensor is block diagonal again, so shwfs_im_from_jacobian applies
 * to modes resampled with shwfs_resample.  No reference counterpart (the reference builds
 * interaction matrices by push-pull only, im_rec_calibrator.py:58-63). */
int shwfs_vjp_sensor_grid(shwfs_plan* plan, int n_frames, const float* A, int64_t a_frame_stride,
                          const float* phase_nm, int64_t phase_frame_stride, double S0,
                          const float* grad_ccd, int64_t grad_ccd_stride, float* grad_phase_sensor,
                          int64_t grad_phase_stride, void* stream);

/* The plan's field resampling (edge extrapolation + bilinear resampling, sh.py:350-360) applied to
 * n_frames phase-like frames on the input grid (in_frame_stride 0: one frame for all) into the
 * sensor grid.  Errors for a plan without resampling. */
int shwfs_resample(shwfs_plan* plan, int n_frames, const float* in, int64_t in_frame_stride, float* out,
                   int64_t out_frame_stride, void* stream);

/* Interaction matrix from the block-diagonal slope Jacobian (any np_sub and pupil width; with
 * field resampling, the sensor-grid Jacobian and resampled modes):
 *   im[m][slot]           = scale * sum_{p in subaperture} jx[p] * modes[m][p]
 *   im[m][n_valid + slot] = scale * sum_{p in subaperture} jy[p] * modes[m][p]
 * jx, jy: ef x ef maps (d slope / d phase_nm); modes: n_modes maps of ef x ef (stride
 * mode_stride); slot: n_lenses^2 int32, slope index of each subaperture or -1; scale = 2
 * reproduces the push-pull convention IM = (s(+a) - s(-a)) / a.  Precondition: the slots in
 * [0, n_valid) are each claimed by one subaperture.  Slot values outside [0, n_valid) are ignored
 * (never used as an index); a slot in [0, n_valid) that no subaperture claims gets zero rows with
 * `work`, and is left unwritten without it.
 * work: optional device workspace of n_modes * 2 * n_lenses^2 floats followed by n_valid int32
 * (16-byte aligned); with it the results are written in grid order and reordered into slot
 * order by a second pass (coalesced stores); NULL writes slot order directly. */
int shwfs_im_from_jacobian(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy,
                           int n_modes, const float* modes, int64_t mode_stride, const int32_t* slot,
                           int n_valid, float scale, float* im, void* work, void* stream);

/* Same interaction matrix without any allocation and without a workspace argument: no reference
 * counterpart.  For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The grid-order results go through a
 * library-owned 16 MiB device scratch array (one per GPU, held until the process ends), modes in
 * chunks of (16 MiB - n_valid * 4 bytes) / (8 n_lenses^2 bytes) (327 modes per chunk at 80 x 80);
 * results equal shwfs_im_from_jacobian with a workspace bit for bit.  Rejects with SHWFS_ERR_ARG a
 * geometry whose single mode does not fit (n_valid * 4 + 8 n_lenses^2 > 16 MiB).  Calls that
 * overlap on one GPU (other streams, other threads) share the scratch and race: serialise them. */
int shwfs_im_from_jacobian_noalloc(int n_lenses, int np_sub, int ef_size, const float* jx, const float* jy,
                                   int n_modes, const float* modes, int64_t mode_stride, const int32_t* slot,
                                   int n_valid, float scale, float* im, void* stream);

/* ---- multi-sensor batch -----------------------------------------------------------------
 * No reference counterpart (the reference triggers one SH object at a time, each optionally a
 * captured CUDA graph, base_processing_obj.py:114-155).  Persistent launches per body type (LGS
 * np 16 / N 128 tiles, then NGS MFMA np 8 tiles of up to 4 streamed frames; one launch per body
 * variant present, those with fewer work items first) run every (sensor, frame, subaperture tile)
 * of several sensors of uneven shapes; sensors with field
 * resampling have their phases resampled before it and sensors normalised by their own image sum
 * are rescaled after it.  Any other plan (other np, the any-size path, round FoV masks) runs its
 * own launch in the same sequence (n_frames <= its max_frames).  All buffers are fixed at creation
 * (graph friendly): per sensor A (input grid, static), phase (n_frames x input grid), ccd
 * (n_frames x ccd^2), S0.
 * grid_blocks: persistent blocks (0: one per work item); order: 0 heaviest tiles first,
 * 1 sensor by sensor, 2 round robin over sensors. */
typedef struct shwfs_batch shwfs_batch;
int shwfs_batch_create(int n_sensors, shwfs_plan* const* plans, const int32_t* n_frames,
                       const float* const* A, const float* const* phase, float* const* ccd,
                       const double* S0, int grid_blocks, int order, int device, shwfs_batch** out);
int shwfs_batch_forward(shwfs_batch* batch, void* stream);
/* Runs the ShSlopec of batch sensor `sensor` in the SH launch's epilogue, as shwfs_forward_slopes
 * does for one plan (ShSlopec.calc_slopes_nofor, sh_slopec.py:308-438): the sensor's plan must have
 * its ShSlopec plan bound (shwfs_plan_bind_slopec), Parseval normalisation, and run in one of the
 * batch's persistent bodies (NGS MFMA np 8 with 8 or 16 px per subaperture and 16-byte aligned CCD
 * rows, or LGS np 16 / M 32); every later shwfs_batch_forward then also writes its slopes (frames
 * `slopes_stride` apart), flux and totals.  thr_mode 0 subtracts and clamps the threshold, 1 is
 * the pedestal mode.  A set-up call (synchronous): call it before capturing the batch in a graph. */
int shwfs_batch_set_slopes(shwfs_batch* batch, int sensor, float thr, int thr_mode, float* slopes,
                           int64_t slopes_stride, double* flux, double* totals);
/* Per-sensor completion time [us] of the last shwfs_batch_forward, from its start (synchronises
 * the stream). */
int shwfs_batch_latencies(shwfs_batch* batch, double* us, void* stream);
int shwfs_batch_info(shwfs_batch* batch, int32_t* n_items, int32_t* grid);
int shwfs_batch_destroy(shwfs_batch* batch);

/* ---- detector post-step (CCD.trigger_code, processing_objects/ccd.py:198-280) -------------
 * Noise sources use a counter-based generator keyed by their seed and indexed by
 * (frame * n_pixels + pixel, counter, draw): results are reproducible for a given counter, and
 * statistically (not bit-) equivalent to numpy's default_rng streams of the reference. */
#define SHCCD_PHOTON         0x001  /* photon_noise: Poisson, then ADU quantisation            */
#define SHCCD_READOUT        0x002  /* readout_noise: + readout_level * N(0,1)                 */
#define SHCCD_EXCESS         0x004  /* excess_noise: EM gamma of shape delta*x, scale emccd    */
#define SHCCD_DARK           0x008  /* darkcurrent_noise                                       */
#define SHCCD_BACKGROUND     0x010  /* background_noise                                        */
#define SHCCD_CIC            0x020  /* cic_noise: + Bernoulli(cic_level)                       */
#define SHCCD_KEEP_BIAS      0x040  /* _keep_ADU_bias                                          */
#define SHCCD_BG_REMOVE_AVG  0x080  /* _bg_remove_average                                      */
#define SHCCD_KEEP_DARK      0x100  /* _do_not_remove_dark                                     */
#define SHCCD_OFFSET_APPLIED 0x200  /* the input already holds background + dark: not added
                                       again (their subtraction after the ADU step is kept)     */

typedef struct shccd_params {
    int32_t in_rows, in_cols;      /* integrated frame = size * binning (ccd.py:134-136)     */
    int32_t out_rows, out_cols;    /* pixel frame = size // binning                          */
    int32_t binning;
    int32_t flags;                 /* SHCCD_* bits                                           */
    double qe;                     /* quantum_eff (apply_qe, ccd.py:275-280)                 */
    double background_level, darkcurrent_level, readout_level, cic_level;
    double adu_gain, adu_bias, emccd_gain, excess_delta;
    uint64_t photon_seed, readout_seed, excess_seed, cic_seed;
} shccd_params;

/* Intensity.sum(in_i, factor) into the integration buffer (ccd.py:200-205):
 * integrated[f] += in_i[f] * factor over n_px pixels per frame. */
int shccd_integrate(int n_frames, const float* in_i, int64_t in_stride, double factor, float* integrated,
                    int64_t int_stride, int64_t n_px, void* stream);
/* Readout (apply_binning, apply_qe, apply_noise, ccd.py:207-273) of n_frames frames into
 * `pixels`.  Input: the integration buffer (read, then zeroed as ccd.py:217 does) and/or the
 * current intensity `in_i * factor` (the last integration step fused into the readout; either
 * pointer may be NULL, not both).  pixel_gains: out_rows x out_cols or NULL.  `counter`
 * selects the random numbers of this readout (the caller advances it). */
int shccd_readout(const shccd_params* params, int n_frames, float* integrated, int64_t int_stride,
                  const float* in_i, int64_t in_stride, double factor, const float* pixel_gains,
                  uint64_t counter, float* pixels, int64_t out_stride, void* stream);

/* Readout with charge-transfer-efficiency matrices (cte_noise, ccd.py:223-224):
 *   frame = cte_rows . (binned signal * QE + background + dark) . cte_cols
 * followed by CIC, photon noise, EM gain, readout noise, pixel gains and ADU conversion exactly as
 * shccd_readout.  cte_rows: out_rows x out_rows, cte_cols: out_cols x out_cols, row-major float32
 * (the reference's cte_mat[:, :, 0] and cte_mat[:, :, 1]).  work: 2 * n_frames * out_rows *
 * out_cols floats of device scratch owned by the caller.  Same random numbers as shccd_readout
 * for the same counter. */
int shccd_readout_cte(const shccd_params* params, int n_frames, float* integrated, int64_t int_stride,
                      const float* in_i, int64_t in_stride, double factor, const float* cte_rows,
                      const float* cte_cols, float* work, const float* pixel_gains, uint64_t counter,
                      float* pixels, int64_t out_stride, void* stream);

/* The detector noise generator: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11) of n
 * counter/key sets counters_keys[6 i .. 6 i + 5] = (c0, c1, c2, c3, k0, k1) into out[4 i .. 4 i + 3]
 * (device memory).  No counterpart in the reference (its streams are numpy's default_rng): exposes
 * the generator that every shccd_readout* noise draw comes from, for its known-answer test. */
int shccd_philox(int n, const uint32_t* counters_keys, uint32_t* out, void* stream);

/* ---- push-pull calibration ------------------------------------------------------------------
 * Fields of a modal push-pull around phi0 (n_px floats): out[k] = phi0 + amp * modes[k],
 * out[n_modes + k] = phi0 - amp * modes[k] (2 n_modes frames, out_stride apart).  Replaces the
 * command history of modal_pushpull_signal.py:35-65 projected through the DM (one push and one
 * pull per mode). */
int shwfs_pushpull_fields(int n_modes, const float* phi0, int64_t n_px, const float* modes, int64_t mode_stride,
                          float amp, float* out, int64_t out_stride, void* stream);
/* Interaction-matrix rows from the slopes of those fields, as ImRecCalibrator accumulates them
 * (im_rec_calibrator.py:58-63, im[mode] += slopes / command):
 * im[k] = slopes[k] / amp + slopes[n_modes + k] / (-amp). */
int shwfs_pushpull_rows(int n_modes, const float* slopes, int64_t slope_stride, int n_slopes, float amp, float* im,
                        int64_t im_stride, void* stream);

/* ---- FP32 matrix product on the matrix cores ----------------------------------------------
 *   C[i][j] = alpha * sum_k A[i][k] op(B)[k][j] + beta * C[i][j] + bias[j]     (i < m, j < n)
 * A row-major m x k (lda); op(B) = B row-major k x n (ldb) when trans_b == 0, B^T with B row-major
 * n x k when trans_b == 1; bias: n floats or NULL; C row-major (ldc).  Device pointers.
 * Replaces the command -> phase projection of the DM, phase[idx] = (sign * cmd) @ ifunc
 * (processing_objects/dm.py:72-86), and the reconstruction / modal filtering products of
 * Slopec.post_trigger (processing_objects/slopec.py:92-104). */
int shgemm_f32(int m, int n, int k, float alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
               int trans_b, float beta, const float* bias, float* C, int64_t ldc, void* stream);

/* ---- HIP graphs -----------------------------------------------------------------------------
 * Capture of library calls into a replayable HIP graph.  Replaces the CUDA-graph capture of a
 * processing object's trigger_code (base_processing_obj.py:114-131, launched at :145-155): begin a
 * capture on a non-default stream, issue any library calls on that stream, end it; each replay is
 * one launch on any stream of the device.  Thread-local capture mode.  Replays read and write the
 * buffers that the captured calls named (the caller keeps them alive and in place). */
typedef struct shwfs_graph shwfs_graph;
int shwfs_graph_begin(void* stream);
int shwfs_graph_end(void* stream, shwfs_graph** out);
int shwfs_graph_launch(shwfs_graph* graph, void* stream);
int shwfs_graph_destroy(shwfs_graph* graph);
/* Number of nodes (kernel launches) of a captured graph: a replay pays off only when it replaces
 * enough direct launches (the processing objects capture only above a node-count threshold). */
int shwfs_graph_node_count(const shwfs_graph* graph, int64_t* n_nodes);

#ifdef __cplusplus
}
#endif
#endif /* SHWFS_H */
