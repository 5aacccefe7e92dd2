/*
 * sarlabtdc_amd.h -- C ABI of the MI355X time-domain-correlation (TDC) SAR
 * image former.
 *
 * The reference exposes a C++ class (tdc-lib/include/tdcprocessor.h:42-135)
 * and a pybind11 module (py-sarlabtdc/src/interface.cpp:9-23) and has no C
 * ABI. These entry points carry exactly the arguments of that class, as plain
 * pointers and sizes, so any FFI (ctypes, cgo, JNI, N-API) can bind them. The
 * C++ class in tdcprocessor.h and the Python module _sarlabtdc are thin
 * layers over this header.
 *
 * Conventions
 *   - Every function returns an int status (TDC_OK or a TDC_ERR_* code).
 *     No exception crosses this boundary. tdc_last_error() returns a
 *     thread-local message for the last failing call on this thread.
 *   - Complex data are interleaved float pairs (re, im), the layout of
 *     std::complex<float> and of HIP's float2.
 *   - Host arrays passed to tdc_set_* are borrowed, not copied, until the
 *     next tdc_start / tdc_prepare on the handle (the reference keeps the
 *     same non-owning pointers, tdcprocessor.cpp:169-217).
 *   - A handle is not thread safe; one handle per host thread.
 */
#ifndef SARLABTDC_AMD_H
#define SARLABTDC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The C++ layer maps INVALID_ARGUMENT to std::invalid_argument
 * and every other error to std::runtime_error, as the reference throws
 * (tdcprocessor.cpp:48-52, :189-195, :149-154). */
#define TDC_OK 0
#define TDC_ERR_INVALID_ARGUMENT 1
#define TDC_ERR_RUNTIME 2
#define TDC_ERR_DEVICE 3
#define TDC_ERR_STATE 4
#define TDC_ERR_UNSUPPORTED 5
#define TDC_ERR_OUT_OF_MEMORY 6

/* Image-formation forms (additive; the reference has only the plain one). */
#define TDC_FORM_PLAIN 0     /* direct per-term FP32 reference phase: the drop-in */
#define TDC_FORM_PRECISION 1 /* FP64 range/anchor + FP32 rotation recurrence */

/* Engines of the precision form (additive). All compute the same sums to
 * FP32-level accuracy. The range-profile engine takes each chirp's
 * oversampled spectrum once (an FFT in LDS) and interpolates it at every
 * pixel's range with an 8- to 12-tap Gaussian kernel (a type-2 non-uniform
 * FFT; up to 8192 samples per chirp; see tdc_set_profile_kernel). The matrix
 * engine forms the 16-sample group sums of every 128-sample block with FP16
 * hi/lo MFMA (32x32x16). The vector engine runs a one-sample rotation
 * recurrence on the VALU. */
#define TDC_ENGINE_AUTO 0    /* range profile when the chirp has at most 8192 samples, else
                                matrix; vector when pixels_per_lane > 1 is forced */
#define TDC_ENGINE_VECTOR 1
#define TDC_ENGINE_MATRIX 2
#define TDC_ENGINE_PROFILE 3

/* Raw sample storage on the device for the precision form (additive). */
#define TDC_RAW_C64 0  /* complex64 */
#define TDC_RAW_F16 1  /* two IEEE half values per sample */
#define TDC_RAW_BF16 2 /* two bfloat16 values per sample */

/* Range window families (additive; the reference hard-codes the periodic
 * Hamming window with A = 25/46, tdc-lib/include/tuning.h:31). */
#define TDC_WIN_HAMMING 0 /* A - (1-A) cos(2 pi m / Ns), p0 = A (25/46 default) */
#define TDC_WIN_HANN 1    /* A = 0.5 */
#define TDC_WIN_TAYLOR 2  /* p0 = nbar, p1 = peak sidelobe level in dB (positive) */

typedef struct tdc_processor tdc_processor;

/* tdc_stats.per_chirp bits. RANGE: the antenna position was the same for all
 * samples of every chirp, so |pos - pixel| was computed once per (pixel,
 * chirp); the image is bit-identical to the per-term range. WINDOW: velocity
 * and attitude were constant too, so the Doppler window was evaluated once
 * per (pixel, chirp) and scaled the chirp's sum. */
#define TDC_PER_CHIRP_RANGE 1
#define TDC_PER_CHIRP_WINDOW 2
/* Not per chirp: the antenna moved within chirps, the range was evaluated per
 * term, and its offset from the chirp's first position came from the
 * two-term expansion of the square root (the motion was small enough against
 * the shortest range for a remainder below 5e-8 cycles of phase). */
#define TDC_PER_TERM_SERIES 4

typedef struct tdc_stats {
    double kernel_ms;       /* device time of the last run divided by its passes, each pass the
                               staging, correlate-sum and split reduction kernels, back to back
                               (HIP events) */
    double main_kernel_ms;  /* device time of the correlate-sum kernel alone (after tdc_start:
                               of its slabs, with the waits for their parts of the grid) */
    double end_to_end_ms;   /* host wall time of the last tdc_start (H2D + compute + D2H) */
    double terms;           /* pixels * chirps * samples of the last run */
    int32_t form;           /* TDC_FORM_* actually run */
    int32_t raw_storage;    /* TDC_RAW_* actually run */
    int32_t doppler_active; /* 1 when the Doppler window / skip test ran */
    int32_t chirp_splits;   /* chirp ranges the pixels were split over */
    int32_t pixels_per_lane;
    int32_t workgroups;
    int64_t skipped_chunks;   /* (pixel, chunk) pairs skipped; only counted when doppler_active */
    int64_t processed_chunks; /* (pixel, chunk) pairs processed */
    int32_t engine;           /* TDC_ENGINE_* actually run (precision form), else 0 */
    int32_t per_chirp;        /* plain form, TDC_PER_CHIRP_* bits: what ran once per (pixel,
                                 chirp) because the inputs it depends on were constant
                                 within every chirp (stop-and-hop) */
    int32_t profile_taps;         /* range-profile engine: taps of its kernel, else 0 */
    int32_t profile_oversampling; /* range-profile engine: M / samples (2 or 4), else 0 */
    int32_t profile_incremental;  /* range-profile engine: 1 when the range advanced by FP32
                                     increments within blocks of 8 chirps (see
                                     tdc_set_profile_exact_range), else 0 */
    int32_t reserved;
} tdc_stats;

/* Message for the last failing call on this thread ("" when none). */
const char *tdc_last_error(void);

/* ABI version: major * 10000 + minor * 100 + patch. */
int tdc_abi_version(void);

/* Number of visible HIP devices (0 when none); never fails. */
int tdc_device_count(void);

/* Replaces TdcProcessor::TdcProcessor(int gpuNum), tdcprocessor.cpp:39-55:
 * selects the device, fails with TDC_ERR_DEVICE when it cannot. */
int tdc_create(int gpu_num, tdc_processor **out);

/* Replaces TdcProcessor::~TdcProcessor, tdcprocessor.cpp:57. NULL is a no-op. */
int tdc_destroy(tdc_processor *h);

/* Replaces TdcProcessor::setRawData, tdcprocessor.cpp:169-209 (declared
 * tdcprocessor.h:104-107). raw_iq [n_pri][n_samples][2], pri_times [n_pri]
 * (kept for parity; the reference uploads but never reads it),
 * sample_times [n_samples], position/velocity [n_pri][n_samples][3],
 * attitude [n_pri][n_samples][4] quaternions (scalar first, body to local),
 * body_boresight [3]. n_pri <= 0 or n_samples <= 0 gives
 * TDC_ERR_INVALID_ARGUMENT. Computes the centre wavelength
 * c / ((f0 + f_end) / 2) in double, as the reference does. */
int tdc_set_raw_data(tdc_processor *h, const float *raw_iq, const float *pri_times,
                     const float *sample_times, const float *position, const float *velocity,
                     const float *attitude, int n_pri, int n_samples, float mod_rate,
                     float start_freq, const float *body_boresight);

/* Replaces TdcProcessor::setFocusGrid, tdcprocessor.cpp:211-217:
 * grid [n_rows][n_cols][3] (x, y, z) in the local frame. */
int tdc_set_focus_grid(tdc_processor *h, const float *grid, int n_rows, int n_cols);

/* Replaces TdcProcessor::start, tdcprocessor.cpp:59-167 (tdcprocessor.h:67-68).
 * Forms the image: upload, one fused correlate-sum launch (plus a split
 * reduction), download into the processor-owned image buffer. */
int tdc_start(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
              int dop_centroid_win, int apply_range_win, int doppler_win_taper);

/* Replaces TdcProcessor::imageBuffer, tdcprocessor.cpp:240-243: *image_iq
 * receives the processor-owned [n_rows][n_cols][2] image, valid until the
 * next start or destroy; NULL before the first start (as the reference). */
int tdc_image_buffer(const tdc_processor *h, const float **image_iq);

/* Shape of the focus grid last set (no reference counterpart; the reference's
 * Python wrapper keeps it itself, tdcprocessor_wrapper.h:56-57). */
int tdc_image_shape(const tdc_processor *h, int *n_rows, int *n_cols);

/* ---- additive API: no counterpart in the reference ------------------------ */

/* Range window used when start(..., apply_range_win = 1). */
int tdc_set_range_window(tdc_processor *h, int kind, double p0, double p1);
/* The range window of (kind, p0, p1) as the library applies it: evaluated in
 * FP64, periodic (DFT-even), rounded once to window[n_samples]. Host only; no
 * device or handle needed. */
int tdc_range_window(int kind, double p0, double p1, int n_samples, float *window);

/* Form and device raw storage used by the next start / prepare. */
int tdc_set_form(tdc_processor *h, int form);
int tdc_set_raw_storage(tdc_processor *h, int storage);
/* Engine of the precision form used by the next start / prepare. */
int tdc_set_engine(tdc_processor *h, int engine);

/* FP64 inputs for the precision form: raw_iq complex64 [n_pri][n_samples],
 * sample_times [n_samples] (uniformly spaced), chirp_position [n_pri][3]
 * (the phase centre of each chirp: stop-and-hop), grid set by
 * tdc_set_focus_grid_f64. Selects TDC_FORM_PRECISION. */
int tdc_set_raw_data_f64(tdc_processor *h, const float *raw_iq, const double *pri_times,
                         const double *sample_times, const double *chirp_position, int n_pri,
                         int n_samples, double mod_rate, double start_freq);
int tdc_set_focus_grid_f64(tdc_processor *h, const double *grid, int n_rows, int n_cols);

/* Split of tdc_start for callers that keep inputs resident in HBM:
 * tdc_prepare uploads the inputs and sizes the device buffers; tdc_run forms
 * the image on the device `iterations` times, each one a full pass over the
 * resident inputs (the kernels that fold windows into the raw samples and
 * build per-chirp records, then the correlate-sum, then the split
 * reduction); tdc_fetch_image copies it into the image buffer. tdc_start
 * forms the same image as prepare + run(1) + fetch, bit for bit, but streams
 * the focus grid behind the chirp inputs in two pixel slabs and launches
 * each slab's correlate-sum as soon as its part of the grid has landed (once
 * the image has at least 8192 (tile, split) workgroups; below that, one
 * slab). Uploads go through page-locked double buffers on a copy stream of
 * the processor's own. */
int tdc_prepare(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
                int dop_centroid_win, int apply_range_win, int doppler_win_taper);
int tdc_run(tdc_processor *h, int iterations);
int tdc_fetch_image(tdc_processor *h);

/* Statistics of the last run (timings from HIP events on the processor's
 * stream). */
int tdc_get_stats(const tdc_processor *h, tdc_stats *out);

/* Interpolation kernel of the range-profile engine: (taps, oversampling) =
 * (8, 2), (10, 2), (12, 2), (6, 4) or (8, 4); (0, 0) = automatic, (6, 4) while
 * the fourfold transform fits (chirps of up to 4096 samples) and the image has
 * at least 32768 pixels (the larger table then pays for itself), else (10, 2).
 * The kernel's error against the direct sum is ~1e-5, ~1e-6, ~1e-7, ~4e-5 and
 * ~2e-6 of a chirp's profile (rel-L2); each two taps fewer save 8 of ~74
 * instructions per (pixel, chirp), fourfold oversampling doubles the per-chirp
 * table. */
int tdc_set_profile_kernel(tdc_processor *h, int taps, int oversampling);

/* Range of the range-profile engine. exact = 0 (the default): within each
 * block of 8 chirps the first chirp's range, bin position and anchor phase
 * are evaluated in FP64 and the next chirps' advanced by FP32 increments from
 * the chirps' position steps, when a bound on that error (from the largest
 * step, the positions' and pixels' extents and their smallest distance) stays
 * below 2e-6 cycles per step; otherwise, and with exact = 1, every (pixel,
 * chirp) is evaluated in FP64. tdc_stats.profile_incremental reports which
 * ran. */
int tdc_set_profile_exact_range(tdc_processor *h, int exact);

/* Force a launch configuration for tuning and tests: pixels per lane
 * (0 = automatic; 1, 2 or 4 for the plain form, up to 8 for the precision
 * form) and the chirp split count (0 = automatic). */
int tdc_set_launch(tdc_processor *h, int pixels_per_lane, int chirp_splits);

/* Device memory the processor may use for one image's buffers, in bytes
 * (0 = no limit, the default). No counterpart in the reference, whose
 * start() stages 128-chirp chunks (tdcprocessor.cpp:245-345). When a
 * precision-form image would exceed it, tdc_start streams the chirps through
 * the device in slabs of whole chirp splits, two buffer sets in turn (one
 * filled over PCIe while the other is formed), and adds each slab's splits in
 * split order into an FP64 image: the image is the one a resident pass with
 * the same split count forms, bit for bit (tdc_stats.chirp_splits reports
 * that count). tdc_prepare refuses such an image. The plain form keeps its
 * whole aperture resident and ignores the budget. */
int tdc_set_device_budget(tdc_processor *h, unsigned long long bytes);

/* ---- allocation-free entry point ------------------------------------------ */
/* No counterpart in the reference. For callers that cannot allocate device
 * memory at the point of the call (serving loops, harnesses that forbid
 * hipMalloc): forms the image tdc_start forms for the precision form with
 * the range-profile engine (FP64 inputs or tdc_set_form(TDC_FORM_PRECISION),
 * at most 8192 samples per chirp, any raw storage) and writes it into the
 * caller's host buffer image_iq [n_rows][n_cols][2] (image_floats floats).
 * It calls no device allocator: every device buffer lies in one static
 * device array of TDC_NOALLOC_SCRATCH_BYTES per GPU, held by each GPU the
 * library runs on until the process ends. The chirp split count is the one
 * tdc_start would choose, halved until the buffers fit (tdc_stats.chirp_splits
 * reports it); with that count forced on tdc_start (tdc_set_launch) the two
 * images are bit-identical. Inputs too large for the array, even with one
 * split, give TDC_ERR_INVALID_ARGUMENT and a message with the bytes needed
 * (1024 chirps x 512 samples x 512^2 pixels needs ~55 MB with 8 splits). Calls on one GPU share the array and are serialised inside the
 * process; the array keeps the last call's contents. Blocks until the image
 * is in image_iq. The processor's prepared state (tdc_prepare) must be
 * prepared again afterwards. */
#define TDC_NOALLOC_SCRATCH_BYTES (64ull << 20)
int tdc_start_noalloc(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
                      int dop_centroid_win, int apply_range_win, int doppler_win_taper,
                      float *image_iq, size_t image_floats);

/* ---- serving form (additive; no counterpart in the reference) ------------- */
/* A server owns one GPU stream and a scheduler thread. Jobs of any shape are
 * submitted at any time; the scheduler admits them in arrival order while
 * their device footprint fits the HBM budget, keeps their inputs, partial and
 * final images resident, and advances all admitted jobs together in short
 * steps (continuous batching): each step is a persistent correlate-sum
 * kernel per tile family it holds, draining a table of (job, pixel tile,
 * chirp range) tiles, replayed as a hipGraph per grid-size bucket and set of
 * families (`graphs` counts those captured). A job's tile decomposition and
 * reduction order depend only on its own shape, so its image is the same
 * bit for bit alone or in any mix. */

typedef struct tdc_server tdc_server;

#define TDC_JOB_QUEUED 0
#define TDC_JOB_ADMITTED 1
#define TDC_JOB_DONE 2
#define TDC_JOB_FAILED 3

typedef struct tdc_job_desc {
    const float *raw_iq;          /* [n_pri][n_samples][2] complex64, windows applied */
    const double *sample_times;   /* [n_samples], uniformly spaced */
    const double *chirp_position; /* [n_pri][3] phase centre per chirp */
    const double *grid;           /* [n_rows][n_cols][3] */
    int32_t n_pri, n_samples, n_rows, n_cols;
    double mod_rate, start_freq;
    int32_t form;                 /* TDC_FORM_PLAIN or TDC_FORM_PRECISION */
    int32_t raw_storage;          /* TDC_RAW_* (precision form) */
} tdc_job_desc;

typedef struct tdc_job_info {
    int64_t id;
    int32_t state;        /* TDC_JOB_* */
    int32_t status;       /* TDC_OK, or the error that failed the job */
    double t_submit_s;    /* server clock, seconds since tdc_server_create */
    double t_admit_s;
    double t_first_step_s;
    double t_done_s;      /* image copied to host memory */
    double terms;         /* pixels * chirps * samples */
    int64_t device_bytes; /* HBM footprint charged against the budget */
    int32_t tiles;
    int32_t steps;        /* steps the job's tiles ran in */
} tdc_job_info;

typedef struct tdc_server_stats {
    int64_t submitted, completed, failed, steps;
    double device_busy_ms; /* summed device time of the steps (HIP events) */
    double terms_done;
    int64_t bytes_in_use, bytes_peak;
    int32_t queued, active, graphs;
} tdc_server_stats;

/* hbm_budget_bytes: device memory the admitted jobs may hold together;
 * step_terms: work per step (pixels*chirps*samples; 0 = default 4e10). */
int tdc_server_create(int gpu_num, double hbm_budget_bytes, double step_terms, tdc_server **out);
/* Queues a job and returns at once. The arrays stay borrowed until the job
 * is DONE or FAILED. A job larger than the whole budget fails at submit. */
int tdc_server_submit(tdc_server *s, const tdc_job_desc *job, int64_t *job_id);
/* *done = 1 once the job is DONE or FAILED. */
int tdc_server_poll(tdc_server *s, int64_t job_id, int *done);
/* Blocks until the job ends, copies its [n_rows][n_cols][2] image into
 * image_iq (may be NULL) and its record into info (may be NULL), then
 * forgets the job. Returns the job's status. */
int tdc_server_wait(tdc_server *s, int64_t job_id, float *image_iq, tdc_job_info *info);
int tdc_server_get_stats(tdc_server *s, tdc_server_stats *out);
/* Seconds on the server clock (the clock of tdc_job_info). */
int tdc_server_clock(tdc_server *s, double *seconds);
/* Finishes the steps in flight, fails the jobs still queued, frees everything. */
int tdc_server_destroy(tdc_server *s);

#ifdef __cplusplus
}
#endif

#endif /* SARLABTDC_AMD_H */
