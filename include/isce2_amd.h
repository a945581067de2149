/*
 * isce2_amd.h — C ABI of the MI355X dense offset-tracking engine.
 *
 * The engine computes, for every window, what PyCuAmpcor's cuAmpcorChunk::run
 * computes (contrib/PyCuAmpcor/src/cuAmpcorChunk.cpp:12-217): the integer
 * correlation peak, SNR and covariance on the raw correlation surface, and
 * the sub-pixel offset from the oversampled surface.
 *
 * Conventions
 *   - x = row = down = azimuth, y = column = across = range; row-major images.
 *   - Every pointer argument of ampcor_run* is DEVICE memory on the context's
 *     device; sizes are element counts; pitches are in elements.
 *   - Functions return 0 (AMPCOR_OK) or a negative AMPCOR_E* code and never
 *     exit the process; ampcor_last_error() describes the last failure.
 *   - A context is not thread-safe; use one per host thread.
 */
#ifndef ISCE2_AMD_H
#define ISCE2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMPCOR_OK            0
#define AMPCOR_E_ARG        -1   /* invalid argument / size */
#define AMPCOR_E_RANGE      -2   /* a window reads outside its image */
#define AMPCOR_E_DEVICE     -3   /* HIP runtime error */
#define AMPCOR_E_NOMEM      -4   /* device allocation failed */
#define AMPCOR_E_UNSUPPORTED -5  /* configuration outside the engine's limits */

typedef struct ampcor_ctx ampcor_ctx;

/* Per-window parameters. Mirrors the fields of cuAmpcorParameter that the
 * per-window pipeline reads (contrib/PyCuAmpcor/src/cuAmpcorParameter.h:32-158);
 * the values are the *Raw user settings, before setupParameters. */
typedef struct ampcor_params {
    int32_t window_h;      /* windowSizeHeightRaw */
    int32_t window_w;      /* windowSizeWidthRaw */
    int32_t hs_down;       /* halfSearchRangeDownRaw */
    int32_t hs_across;     /* halfSearchRangeAcrossRaw */
    int32_t raw_ovs;       /* rawDataOversamplingFactor */
    int32_t zoom;          /* zoomWindowSize (corrSurfaceZoomInWindow), clamped as setupParameters */
    int32_t surf_ovs;      /* oversamplingFactor (corrSurfaceOverSamplingFactor) */
    int32_t surf_method;   /* oversamplingMethod: 0 FFT, 1 sinc */
    int32_t corr_stat;     /* corrStatWindowSize */
    int32_t deramp;        /* derampMethod: 0 amplitude chips, 1 deramp, other: complex chips */
    int32_t algorithm;     /* 0 frequency domain, 1 time domain (identical sums) */
    int32_t image_complex; /* 1: complex64 rasters, 0: float32 rasters */
} ampcor_params;

/* Derived sizes, as cuAmpcorParameter::setupParameters computes them
 * (contrib/PyCuAmpcor/src/cuAmpcorParameter.cpp:70-98). */
typedef struct ampcor_sizes {
    int32_t zoom, half_zoom;
    int32_t search_h, search_w;          /* raw search chip */
    int32_t zoom_in_h, zoom_in_w;        /* raw zoom-in search chip */
    int32_t os_win_h, os_win_w;          /* oversampled template */
    int32_t os_search_h, os_search_w;    /* oversampled zoom-in search chip */
    int32_t raw_corr_h, raw_corr_w;      /* raw correlation surface */
    int32_t os_corr_h, os_corr_w;        /* oversampled correlation surface (before trim) */
    int32_t surf_out;                    /* oversampled surface side */
    int32_t corr_raw_zoom_h, corr_raw_zoom_w; /* SNR patch */
    int32_t fused;                       /* 1 if the fused per-window kernels cover this config */
    int32_t fft_corr_raw_h, fft_corr_raw_w; /* staged correlation transform sizes (>= search chip, */
    int32_t fft_corr_os_h, fft_corr_os_w;   /* smallest 2^a 3^b 5^c: any length >= the chip gives the same lags) */
} ampcor_sizes;

/* Execution modes of ampcor_run_windows. */
#define AMPCOR_MODE_AUTO    0   /* fused when supported, else staged */
#define AMPCOR_MODE_STAGED  1   /* one batched kernel per pipeline stage (any size) */
#define AMPCOR_MODE_FUSED   2   /* per-window fused kernels (window <= 64 class) */
/* Flag OR-ed into the mode: no host synchronization inside the call. Chip
 * bounds are then checked on the device (a window whose chips leave their image
 * gets NaN offset/SNR/covariance instead of AMPCOR_E_RANGE) and the precision-1
 * FP64 re-run is skipped (pass margin outputs to ampcor_run_windows_debug and
 * re-run from them). Required on a stream under HIP graph capture; a captured
 * call must not allocate, so run the identical call once before capturing it. */
#define AMPCOR_FLAG_NOSYNC  0x100

/* Create / destroy an engine context on HIP device `device`.
 * Replaces cuAmpcorController's ctor/dtor and gpuDeviceInit
 * (contrib/PyCuAmpcor/src/cuAmpcorController.cpp:19-29,43; cudaUtil.cpp:8-27). */
int ampcor_create(int32_t device, ampcor_ctx **out);
int ampcor_destroy(ampcor_ctx *ctx);

/* Validate parameters and compute derived sizes; builds FFT plans.
 * Replaces cuAmpcorParameter::setupParameters (cuAmpcorParameter.cpp:70-109). */
int ampcor_setup(ampcor_ctx *ctx, const ampcor_params *params, ampcor_sizes *sizes_out);

/* Run the per-window pipeline for n windows.
 * Replaces cuAmpcorController::runAmpcor's chunk loop and cuAmpcorChunk::run
 * (cuAmpcorController.cpp:107-137; cuAmpcorChunk.cpp:12-217), with images
 * already resident in device memory (GDALImage::loadToDevice is the caller's).
 *   ref_img/sec_img : complex64 (interleaved float pairs) or float32 images,
 *                     [h][pitch] row-major.
 *   ref_d/ref_a/sec_d/sec_a : int32[n] start pixels of each window's reference
 *                     chip and secondary search chip (cuAmpcorParameter.cpp:155-205).
 *   out_offset : float[n][2] (down, across); out_snr : float[n];
 *   out_cov : float[n][3] (cov_down, cov_across, cov_cross).
 *   chunk : windows per launch (0 = engine default: 1024 for the staged
 *           pipeline, 32768 for the fused kernels); mode : AMPCOR_MODE_*.
 *   stream : hipStream_t (NULL = default stream). The call is asynchronous
 *   with respect to the host; synchronize the stream before reading outputs.
 * A window whose chips leave their image fails the call with AMPCOR_E_RANGE
 * (the lowest such window is named). Under precision 0 and 2 this is checked
 * before any kernel runs; under precision 1 it is reported with the FP64
 * re-run's window counts (the call's one host synchronization), after the
 * FP32 kernels have run on window (0, 0) in its place, and before the re-run. */
int ampcor_run_windows(ampcor_ctx *ctx,
                       const void *ref_img, int64_t ref_h, int64_t ref_w, int64_t ref_pitch,
                       const void *sec_img, int64_t sec_h, int64_t sec_w, int64_t sec_pitch,
                       const int32_t *ref_d, const int32_t *ref_a,
                       const int32_t *sec_d, const int32_t *sec_a, int32_t n,
                       float *out_offset, float *out_snr, float *out_cov,
                       int32_t chunk, int32_t mode, void *stream);

/* Debug/parity entry: same as ampcor_run_windows but also returns the integer
 * stage results per window (int32[n][2] each): raw-surface argmax (i_offsetInit
 * before adjustment), zoom start (i_offsetInitAdjusted), zoom shift
 * (i_maxLocShift) and oversampled-surface argmax (i_offsetZoomIn), plus the raw
 * peak value (r_maxval). Names follow CUAMPCOR_DEBUG (cuAmpcorChunk.cpp:82-198).
 * The fused path also reports each window's argmax margins (float[n]).
 * out_offset_fp64 (double[n][2], down/across) is an FP64 sub-pixel refinement
 * with no counterpart in the reference (whose offset is the bare argmax on the
 * 1/(surf_ovs*raw_ovs) grid, cuOffset.cu:137-187): the FFT-interpolated surface
 * is evaluated in FP64 at the peak and its axis neighbours, and a parabola per
 * axis gives the vertex; needs corrSurfaceOverSamplingMethod 0 and a zoom
 * window of at most 64 (else AMPCOR_E_UNSUPPORTED). out_offset stays the
 * reference's grid offset.
 * Any of the extra pointers may be NULL. */
int ampcor_run_windows_debug(ampcor_ctx *ctx,
                       const void *ref_img, int64_t ref_h, int64_t ref_w, int64_t ref_pitch,
                       const void *sec_img, int64_t sec_h, int64_t sec_w, int64_t sec_pitch,
                       const int32_t *ref_d, const int32_t *ref_a,
                       const int32_t *sec_d, const int32_t *sec_a, int32_t n,
                       float *out_offset, float *out_snr, float *out_cov,
                       int32_t *out_maxloc, int32_t *out_start, int32_t *out_shift,
                       int32_t *out_zoom_idx, float *out_maxval,
                       float *out_margin_coarse, float *out_margin_fine,
                       double *out_offset_fp64,
                       int32_t chunk, int32_t mode, void *stream);

/* Allocation-free variant of ampcor_run_windows. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree and no allocating library routine: the fused
 * kernels' intermediates live in one file-scope device array of 256 MiB per
 * GPU (each GPU the library runs on holds it until the process ends), taken in
 * chunks of up to ~1,800 windows at window 64 (per window: 145 KB of
 * oversampled chips plus a few stage values). Any n is accepted; windows run
 * chunk after chunk. The tables the kernels read are built by ampcor_setup.
 * Covers the configurations of the fused kernels only (ampcor_sizes.fused == 1;
 * otherwise AMPCOR_E_UNSUPPORTED). FP32 kernels only: the context's precision
 * setting is not applied (no FP64 re-run). No host synchronization: window
 * bounds are checked on the device, and a window whose chips leave their image
 * gets NaN offset/SNR/covariance. Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (any context, any stream) race:
 * serialize them. Outputs equal ampcor_run_windows' under precision 0. */
int ampcor_run_windows_noalloc(ampcor_ctx *ctx,
                       const void *ref_img, int64_t ref_h, int64_t ref_w, int64_t ref_pitch,
                       const void *sec_img, int64_t sec_h, int64_t sec_w, int64_t sec_pitch,
                       const int32_t *ref_d, const int32_t *ref_a,
                       const int32_t *sec_d, const int32_t *sec_a, int32_t n,
                       float *out_offset, float *out_snr, float *out_cov, void *stream);

/* Near-tie selection for FP64 re-runs driven by the caller (no counterpart in
 * the reference; for batch drivers that capture their re-runs in a HIP graph).
 * From the margins that ampcor_run_windows_debug writes, idx[0..cap) gets the
 * indices of the first cap windows in order whose margin_coarse < tau_coarse
 * or margin_fine < tau_fine (NaN counts as below), then, while slots remain,
 * the first windows that are not. Needs 0 <= cap <= n. Device pointers; one
 * workgroup on `stream`, no host synchronization, no allocation (capturable). */
int ampcor_select_near_ties(ampcor_ctx *ctx, const float *margin_coarse, const float *margin_fine, int32_t n,
                            float tau_coarse, float tau_fine, int32_t cap, int32_t *idx, void *stream);

/* Precision control. No counterpart in the reference (it runs FP32 only).
 *   mode 0: FP32 kernels only.
 *   mode 1 (default): FP32 kernels; windows whose raw-surface or oversampled-
 *     surface argmax margin (top-1 minus top-2 of the normalized surface) is
 *     below tau_coarse / tau_fine are re-run through an FP64 pipeline, so the
 *     integer peaks are those of an FP64 evaluation.
 *   mode 2: FP64 pipeline for every window.
 * Default thresholds: 2e-6 for both (surface values are of order 1 and FP32
 * rounding moves them by a few 1e-7). Negative taus keep the current
 * thresholds. ampcor_precision_stats returns
 * the windows checked and re-run since the last call. */
int ampcor_set_precision(ampcor_ctx *ctx, int32_t mode, float tau_coarse, float tau_fine);
int ampcor_precision_stats(ampcor_ctx *ctx, int64_t *checked, int64_t *rerun);

/* Radix schedule of the staged pipeline's length-n FFTs (radices from
 * {2,3,4,5,7,8,11,13,16,17}, product n, at most 14 passes), for autotuning.
 * No counterpart in the reference (cuFFT plans internally). */
int ampcor_set_fft_radices(ampcor_ctx *ctx, int32_t n, const int32_t *radices, int32_t count);

/* Launch-shape options, for autotuning. No counterpart in the reference.
 *   "fine_threads": 512 or 1024 threads per workgroup of the window-64 fused
 *                   fine kernel (default 1024).
 *   "fft_block_elems": 1024, 2048 (default), 4096 or 8192 complex elements per
 *                   workgroup of the staged pipeline's row/column FFT kernels.
 *   "fft_cols_width": 0, 8 (default) or 16 columns per workgroup of the staged
 *                   pipeline's FP32 column FFTs (0: as many as fft_block_elems allows). */
int ampcor_set_option(ampcor_ctx *ctx, const char *name, int64_t value);

/* Per-kernel timing with HIP events recorded on the launch stream around every
 * kernel of ampcor_run_windows. No counterpart in the reference (its timing is
 * host wall clock, cuDenseOffsets.py:465,484-485); for benchmarks and profiling.
 * ampcor_profile(ctx, 1) resets and enables, (ctx, 0) disables.
 * ampcor_profile_report synchronizes the recorded events and writes lines
 * "kernel_name launches total_ms windows\n" into buf (NUL-terminated). */
int ampcor_profile(ampcor_ctx *ctx, int32_t enable);
int ampcor_profile_report(ampcor_ctx *ctx, char *buf, int64_t len);
/* Restrict the event timing to one kernel name (NULL or "" = every kernel),
 * so that a timed region carries the fewest event records. */
int ampcor_profile_only(ampcor_ctx *ctx, const char *name);

/* Window start pixels of a regular grid, in HOST memory (no device is used).
 * Replaces cuAmpcorParameter::setStartPixels, the constant- and varying-gross
 * overloads (contrib/PyCuAmpcor/src/cuAmpcorParameter.cpp:167-205).
 * Window i = row * n_across + col:
 *   ref_d[i] = start_down0 + row * skip_down,  ref_a[i] = start_across0 + col * skip_across,
 *   sec_d[i] = ref_d[i] + gross_d - hs_down,   sec_a[i] = ref_a[i] + gross_a - hs_across,
 * with gross_d/gross_a = gross_down[i]/gross_across[i] when those int32[n] arrays
 * are given (setVaryingGrossOffset), else gross_down0/gross_across0
 * (setConstantGrossOffset). Outputs are int32[n_down * n_across]. The half
 * search ranges come from params (each axis its own, see ampcor_setup). */
int ampcor_start_pixels(const ampcor_params *params, int32_t n_down, int32_t n_across,
                        int32_t start_down0, int32_t start_across0, int32_t skip_down, int32_t skip_across,
                        int32_t gross_down0, int32_t gross_across0,
                        const int32_t *gross_down, const int32_t *gross_across,
                        int32_t *ref_d, int32_t *ref_a, int32_t *sec_d, int32_t *sec_a);

/* The reference's range check on HOST start arrays (no device is used).
 * Replaces cuAmpcorParameter::checkPixelInImageRange
 * (contrib/PyCuAmpcor/src/cuAmpcorParameter.cpp:275-331), with its strict end
 * test: a window fails when start < 0 or start + chip size >= image size (so a
 * chip ending on the image's last row or column fails too). Returns AMPCOR_OK,
 * or AMPCOR_E_RANGE with *bad_window = the first failing window and *which =
 * the first failing test: 0/1 reference start down/across, 2/3 reference end
 * down/across, 4/5 secondary start down/across, 6/7 secondary end down/across
 * (the reference's order); ampcor_last_error(NULL) names it. Where the reference
 * exits the process, this returns. */
int ampcor_check_range(const ampcor_params *params, const int32_t *ref_d, const int32_t *ref_a,
                       const int32_t *sec_d, const int32_t *sec_a, int32_t n,
                       int64_t ref_h, int64_t ref_w, int64_t sec_h, int64_t sec_w,
                       int32_t *bad_window, int32_t *which);

/* Text of the last error on this context (or, when ctx is NULL, of the last
 * failure of a call that takes no context: ampcor_create, ampcor_start_pixels,
 * ampcor_check_range). */
const char *ampcor_last_error(const ampcor_ctx *ctx);

/* Library version string. */
const char *ampcor_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ISCE2_AMD_H */
