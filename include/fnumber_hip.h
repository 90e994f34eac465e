/*
 * fnumber_hip.h -- C ABI of libfnumber_hip.so, the MI355X (gfx950) Fourier-domain
 * delay-and-sum beamformer with frequency-dependent F-number.
 *
 * This is the drop-in boundary under the reference's MEX gateways:
 *   +cuda/gpu_bf_das_pw_rf.cu:110  mexFunction (steered plane waves)
 *   +cuda/gpu_bf_saft_rf.cu:107    mexFunction (full synthetic aperture)
 * A MEX shim (f_number_amd/csrc/mex/) parses the reference's positional
 * argument lists, evaluates MATLAB-side objects (windows, F-numbers,
 * normalizations) into the plain values below and calls fn_das_pw/fn_das_sa.
 * Every entry point returns an fn_status (0 = success) and never throws across
 * the ABI; fn_status_id() gives the reference's MEX error identifier suffix
 * ("InvalidRFData", ...) and fn_last_error() a message (thread-local).
 *
 * Conventions
 *   - Images are complex float, interleaved (re, im), laid out [n_x][n_z]
 *     (z fastest), i.e. MATLAB's column-major [N_z x N_x]
 *     (gpu_bf_das_pw_rf.cu:1077-1079).  Weights are float [n_x][n_z].
 *   - RF data are column-major [n_samples x n_elements x n_tx] (n_tx = steering
 *     angles for PW, transmitting elements for SA), float32 or float64.
 *   - Frequency bins are 0-based DFT indices n of order N_dft.
 *   - No function allocates memory the caller must free, except plans
 *     (fn_plan_destroy).
 */
#ifndef FNUMBER_HIP_H
#define FNUMBER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FN_ABI_VERSION 1

/* ---- status codes (MEX error identifiers: gpu_bf_das_pw_rf.cu:112-485,
 *      gpu_bf_saft_rf.cu:111-592, gpu_bf_error_handling.cu:9-70) ---------- */
typedef enum fn_status {
    FN_OK = 0,
    FN_ERR_NRHS = 1,                     /* "nrhs" */
    FN_ERR_NLHS = 2,                     /* "nlhs" */
    FN_ERR_NO_DEVICES = 3,               /* "NoCUDADevices" */
    FN_ERR_LATERAL_POSITIONS = 4,        /* "InvalidLateralPositions" */
    FN_ERR_AXIAL_POSITIONS = 5,          /* "InvalidAxialPositions" */
    FN_ERR_RF_DATA = 6,                  /* "InvalidRFData" */
    FN_ERR_SAMPLING_RATE = 7,            /* "InvalidSamplingRate" */
    FN_ERR_STEERING_ANGLE = 8,           /* "InvalidSteeringAngle" */
    FN_ERR_ELEMENT_WIDTH = 9,            /* "InvalidElementWidth" */
    FN_ERR_ELEMENT_PITCH = 10,           /* "InvalidElementPitch" */
    FN_ERR_SOUND_SPEED = 11,             /* "InvalidSoundSpeed" */
    FN_ERR_FREQUENCY_BOUNDS = 12,        /* "InvalidFrequencyBounds" */
    FN_ERR_FREQUENCY_BOUND = 13,         /* "InvalidFrequencyBound" */
    FN_ERR_TIME_SHIFT = 14,              /* "InvalidTimeShift" */
    FN_ERR_WINDOW = 15,                  /* "InvalidWindow" */
    FN_ERR_F_NUMBER = 16,                /* "InvalidFNumber" */
    FN_ERR_NORMALIZATION = 17,           /* "InvalidNormalization" */
    FN_ERR_FREQUENCY_WINDOW = 18,        /* "InvalidFrequencyWindow" */
    FN_ERR_DEVICE = 19,                  /* "InvalidCUDADevice" (id kept for drop-in) */
    FN_ERR_VERBOSITY = 20,               /* "InvalidVerbosityLevel" */
    FN_ERR_TX_F_NUMBER = 21,             /* "InvalidTXFNumber" */
    FN_ERR_RX_F_NUMBER = 22,             /* "InvalidRXFNumber" */
    FN_ERR_HIP = 23,                     /* runtime failure: "CUDAError" (id kept) */
    FN_ERR_OUT_OF_MEMORY = 24,           /* "OutOfMemory" */
    FN_ERR_ARGUMENT = 25,                /* "InvalidArgument": null pointer, bad enum, size */
    FN_ERR_UNSUPPORTED = 26              /* "Unsupported": a valid request this build cannot serve */
} fn_status;

/* window ids: the MEX numbering (gpu_bf_windows.cuh:14-20; the class map of
 * gpu_bf_helper_functions.cu:56-97 sends windows.triangular to 4) */
enum { FN_WINDOW_BOXCAR = 0, FN_WINDOW_HANN = 1, FN_WINDOW_TUKEY = 2, FN_WINDOW_COSINE = 3,
       FN_WINDOW_TRIANGULAR = 4 };
/* normalizations (+normalizations/on.m, off.m, windowed.m) */
enum { FN_NORMALIZATION_OFF = 0, FN_NORMALIZATION_ON = 1, FN_NORMALIZATION_WINDOWED = 2 };
/* semantics: CPU follows das_pw.m/das_sa.m (N_pad from positions_z(1), depth clamp
 * of the PW F-number, untruncated Hann/triangular, continuous boxcar);
 * MEX follows the +cuda kernels (positions_z[1], no clamp, every window
 * truncated at |t| >= 1, element-snapped boxcar). */
enum { FN_SEMANTICS_CPU = 0, FN_SEMANTICS_MEX = 1 };
/* arithmetic of the element x transmit contraction */
enum { FN_PRECISION_FP32 = 0,    /* FP32 MFMA operands, FP32 accumulation */
       FN_PRECISION_BF16X3 = 1,  /* BF16 hi+lo split operands, 3 MFMA passes, FP32 accumulation */
       FN_PRECISION_FP16X3 = 2,  /* FP16 hi+lo split operands, 3 passes */
       FN_PRECISION_FP16 = 3 };  /* FP16 operands, 1 pass */
/* output scale: MEX returns 2/N_dft * sum (gpu_bf_das_pw_rf.cu:1118-1119);
 * the MATLAB CPU path returns 2 * sum (das_pw.m:384,467) */
enum { FN_SCALE_MEX = 0, FN_SCALE_CPU = 1 };
enum { FN_DTYPE_F32 = 0, FN_DTYPE_F64 = 1 };

typedef struct fn_window {
    int32_t id;          /* FN_WINDOW_* */
    double parameter;    /* Tukey cosine fraction in (0,1); ignored otherwise */
} fn_window;

/* F-number objects of the +f_numbers package (C++ port, replaces the
 * mexCallMATLAB round trip of gpu_bf_helper_functions.cu:35-51). */
enum {
    FN_F_NUMBER_CONSTANT = 0,            /* constant(value): p0 = value */
    FN_F_NUMBER_GRATING_ANGLE_LB = 1,    /* grating.angle_lb(chi_deg, F_ub, delta_deg): p0,p1,p2 */
    FN_F_NUMBER_DISTANCE_CONSTANT = 2,   /* distance.constant(delta_deg, F_ub): p0,p1 (F_ub may be inf) */
    FN_F_NUMBER_SZABO = 3,               /* directivity.szabo(width_over_pitch): p0 */
    FN_F_NUMBER_DIRECTIVITY_SOFT = 4,    /* directivity.soft(width_over_pitch, att_dB): p0,p1 */
    FN_F_NUMBER_DIRECTIVITY_RIGID = 5,   /* directivity.rigid(width_over_pitch, att_dB): p0,p1 */
    FN_F_NUMBER_PERROT = 6,              /* directivity.perrot(width_over_pitch): p0 */
    FN_F_NUMBER_DISTANCE_NONE = 7        /* distance.none: no parameters */
};
typedef struct fn_f_number {
    int32_t kind;
    double p0, p1, p2;
} fn_f_number;

/* Steered plane-wave beamforming (das_pw.m:1, gpu_bf_das_pw_rf.cu:110).
 * steering_angles holds n_angles angles; the result is the compounded image
 * sum_a das_pw(angle a) (examples.m:70-78).  One angle with 2-D RF is the
 * reference call exactly. */
typedef struct fn_pw_params {
    uint32_t struct_size;            /* sizeof(fn_pw_params) */
    const double *positions_x;       /* [n_x] lateral voxel positions (m) */
    int32_t n_x;
    const double *positions_z;       /* [n_z] axial voxel positions (m), > 0 */
    int32_t n_z;
    int32_t n_samples;               /* N_t */
    int32_t n_elements;              /* N_el */
    int32_t n_angles;                /* N_a >= 1 */
    const double *steering_angles;   /* [n_angles] rad, in (-pi/2, pi/2) */
    double f_s, element_width, element_pitch, c_0;
    double f_lb, f_ub;               /* Hz; the MEX defaults are DBL_MIN and f_s/2 */
    double index_t0;                 /* >= 0 */
    fn_window window;
    const double *f_number_values;   /* [n_f] F per band bin n_lb..n_ub (fn_pw_frequency_grid) */
    int32_t n_f;
    int32_t normalization;           /* FN_NORMALIZATION_* */
    fn_window normalization_window;  /* window of normalizations.windowed */
    int32_t semantics;               /* FN_SEMANTICS_* */
    int32_t precision;               /* FN_PRECISION_* */
    int32_t output_scale;            /* FN_SCALE_* */
    int32_t index_gpu;
    int32_t verbosity;               /* 0..3 (gpu_bf_das_pw_rf.cu:469-485) */
    /* ensemble (no counterpart in the reference): n_frames acquisitions of the
     * same geometry beamformed in one pass, each frame compounding its
     * n_angles; RF [n_frames][n_angles][n_elements][n_samples], images
     * [n_frames][n_x][n_z].  0 or 1: one frame (the reference call). */
    int32_t n_frames;
    /* virtual-source transmits (no counterpart in the reference; its
     * gpu_bf_das_pw_rf.cu:63 lists diverging waves as to do): NULL for plane
     * waves, else [n_angles][2] = (x_vs, z_vs) in m per transmit, replacing the
     * steering angles (which are then not read).  z_vs < 0: diverging wave,
     * transmit time (|r - r_vs| - |r_vs|) / c (zero where the wavefront crosses
     * the array centre); z_vs > 0: focused wave, (|r_vs| +- |r - r_vs|) / c
     * (+ below the focus).  Both plus index_t0 / f_s. */
    const double *virtual_sources;
    /* line-by-line scan (README.md:31-33 names the application; the reference
     * has no code for it): 0 off; 1: grid column i is formed from transmit i
     * alone (n_x == n_angles, e.g. one focused virtual source per line), so an
     * image of n_x lines is one plan, one spectra launch and one beamforming
     * launch.  FP32, one frame, forward only: the tangent, adjoint, vjp and
     * focused-signal calls and windowed normalization (CPU semantics) refuse it. */
    int32_t line_scan;
} fn_pw_params;

/* Full synthetic aperture (das_sa.m:1, gpu_bf_saft_rf.cu:107).
 * RF is [n_samples x n_elements(rx) x n_elements(tx)]. */
typedef struct fn_sa_params {
    uint32_t struct_size;
    const double *positions_x;
    int32_t n_x;
    const double *positions_z;
    int32_t n_z;
    int32_t n_samples;
    int32_t n_elements;
    double f_s, element_width, element_pitch, c_0;
    double f_lb, f_ub;
    double index_t0;
    fn_window window_tx, window_rx;
    const double *f_number_tx_values;   /* [n_f] */
    const double *f_number_rx_values;   /* [n_f] */
    int32_t n_f;
    int32_t normalization_tx, normalization_rx;
    fn_window normalization_window_tx, normalization_window_rx;
    int32_t semantics, precision, output_scale, index_gpu, verbosity;
} fn_sa_params;

/* ---- library ---------------------------------------------------------- */
int32_t fn_abi_version(void);
const char *fn_status_id(int32_t status);
const char *fn_last_error(void);
/* run-time report of this thread's last fn_das_pw / fn_das_sa call made with
 * verbosity >= 1 (empty otherwise): device properties, launch parameters, memory
 * consumption and per-stage times, the sections gpu_bf_das_pw_rf.cu:664-689,
 * 875-903 and 1128-1144 print; the MEX gateways print it after their status lines */
const char *fn_last_report(void);
/* number of HIP devices (gpu_bf_das_pw_rf.cu:115-121) */
int32_t fn_device_count(int32_t *count);

/* ---- host-side configuration (no GPU involved) ------------------------- */
/* Frequency grid: N_pad, odd N_dft and the band [n_lb, n_ub]
 * (das_pw.m:343-358, gpu_bf_das_pw_rf.cu:525-552). n_f = n_ub - n_lb + 1. */
int32_t fn_pw_frequency_grid(const fn_pw_params *p, int32_t *N_dft, int32_t *n_lb, int32_t *n_ub);
int32_t fn_sa_frequency_grid(const fn_sa_params *p, int32_t *N_dft, int32_t *n_lb, int32_t *n_ub);
/* element pitch over wavelength p*f_n/c_0 for the band bins: the argument the
 * reference passes to compute_values (das_pw.m:368, gpu_bf_das_pw_rf.cu:562-575). */
int32_t fn_element_pitch_over_lambda(double element_pitch, double f_s, double c_0, int32_t N_dft,
                                     int32_t n_lb, int32_t n_ub, double *out);
/* compute_values of an F-number object (+f_numbers/f_number.m:50-114). */
int32_t fn_f_number_values(const fn_f_number *F, const double *element_pitch_over_lambda, int32_t n,
                           double *out);
/* window samples of t (|t| is taken; +windows/window.m:48-93 for CPU
 * semantics, gpu_bf_windows.cu:15-103 for MEX semantics). */
int32_t fn_window_samples(const fn_window *w, int32_t semantics, const double *t, int32_t n, double *out);
/* argument checks of the MEX gateways, in their order, without running anything */
int32_t fn_pw_validate(const fn_pw_params *p);
int32_t fn_sa_validate(const fn_sa_params *p);

/* ---- one-shot entry points (host buffers in and out) ------------------- */
/* Replaces the body of gpu_bf_das_pw_rf.cu:110 after argument parsing:
 * H2D, spectra, fused beamforming, D2H.  image: complex float [n_x][n_z];
 * weights (may be NULL): sum of receive apodization weights of the last bin
 * n_ub per voxel (the MEX's second output). */
int32_t fn_das_pw(const fn_pw_params *p, const void *data_rf, int32_t rf_dtype, float *image,
                  float *weights);
/* Replaces gpu_bf_saft_rf.cu:107.  weights_tx / weights_rx may be NULL. */
int32_t fn_das_sa(const fn_sa_params *p, const void *data_rf, int32_t rf_dtype, float *image,
                  float *weights_tx, float *weights_rx);

/* No counterpart in the reference: an entry point that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Plane waves, one
 * frame: fn_das_pw's computation on device buffers -- d_rf (rf_dtype) in the
 * layout of fn_plan_execute, d_image complex float [n_x][n_z], d_weights (may
 * be NULL) float [n_x][n_z] -- enqueued on `stream` (NULL: the default stream)
 * without host synchronisation.  The plan tables and the spectra workspace live
 * in a fixed 32 MiB device array of the library (one per GPU, held until the
 * process ends); a call that needs more returns FN_ERR_UNSUPPORTED before any
 * device work (the FP32 spectra take 16 KiB per band bin and group of 16 angles
 * at 128 elements, so about 2000 (bin, angle group) pairs fit: 11 angles over
 * 542 bins take 8.9 MB, 75 angles over 1455 bins would take 119 MB; the 16-bit
 * forms need twice that per pair).  Ensembles, line scans and windowed normalization (CPU
 * semantics) also return FN_ERR_UNSUPPORTED.  The host arrays of p are read
 * before the call returns.  Every call on one GPU uses that GPU's array, so
 * calls that overlap on one GPU (other host threads, or other streams still
 * running an earlier call) race. */
int32_t fn_das_pw_noalloc(const fn_pw_params *p, const void *d_rf, int32_t rf_dtype, float *d_image,
                          float *d_weights, void *stream);

/* ---- plans: device-resident execution ---------------------------------- */
/* A plan holds the validated configuration, the device copies of the grid and
 * F-number tables and the spectra workspace for one shape.  fn_plan_execute
 * takes device pointers and enqueues on `stream` (a hipStream_t; NULL is the
 * default stream, as in HIP) without host synchronisation.  Every fn_plan_*
 * call that takes a stream orders its work on that stream only. */
typedef struct fn_plan fn_plan;
typedef struct fn_plan_info {
    int32_t N_dft, n_lb, n_ub, n_f;
    int32_t n_tx;                 /* angles (PW) or elements (SA) */
    int64_t n_voxels;
    int64_t workspace_bytes;
    double flops_per_frame;       /* 8 * n_f * n_voxels * n_elements * n_tx */
    double useful_flops_per_frame;/* same over the elements whose weight can be nonzero (the aperture
                                     bounds; every element of an untruncated window) and, for the
                                     synthetic aperture, the active transmitters */
    int32_t tile_voxels, bin_split, m_tiles, k_steps;
} fn_plan_info;
int32_t fn_plan_pw_create(const fn_pw_params *p, fn_plan **plan);
int32_t fn_plan_sa_create(const fn_sa_params *p, fn_plan **plan);
int32_t fn_plan_info_get(const fn_plan *plan, fn_plan_info *info);
/* d_rf: device RF (rf_dtype); d_image: complex float [n_x][n_z]; weights may be NULL */
int32_t fn_plan_execute(fn_plan *plan, const void *d_rf, int32_t rf_dtype, float *d_image,
                        float *d_weights, float *d_weights_rx, void *stream);
/* Per-stage timing with HIP events recorded on the execution stream around
 * the spectra kernel and the beamforming kernel; totals since the last reset. */
int32_t fn_plan_set_profiling(fn_plan *plan, int32_t enable);
int32_t fn_plan_get_profile(fn_plan *plan, double *spectra_ms, double *beamform_ms, int32_t *n_runs);
int32_t fn_plan_destroy(fn_plan *plan);

/* ---- tuning and batches (no counterpart in the reference) ----------------- */
/* Tiling of a plan's beamforming launch: m_tiles transmit tiles of 16 rows per
 * workgroup (plane waves 1..5, ensembles 1..4; synthetic aperture fixed) and the
 * number of bin chunks split across workgroups (bin_split, 1..n_f). */
typedef struct fn_tune_result {
    double best_ms;               /* per execution, HIP events on the tuning stream */
    double default_ms;            /* the plan's initial tiling */
    int32_t m_tiles, bin_split;   /* chosen (and applied) */
    int32_t n_candidates;
} fn_tune_result;
/* Times the candidate tilings on d_rf / d_image (contents overwritten) and keeps the fastest. */
int32_t fn_plan_autotune(fn_plan *plan, const void *d_rf, int32_t rf_dtype, float *d_image, int32_t repetitions,
                         fn_tune_result *result, void *stream);
/* Applies a tiling (e.g. one stored from an earlier fn_plan_autotune). */
int32_t fn_plan_set_tiling(fn_plan *plan, int32_t m_tiles, int32_t bin_split);

/* A batch executes many plans (same device) as one hipGraph: the launch
 * sequences of all plans are captured once, spread over n_streams forked
 * streams, and replayed by one graph launch per fn_batch_execute.  The graph
 * is captured again when any buffer pointer differs from the captured ones.
 * Arrays have one entry per plan; rf_dtypes FN_DTYPE_*; weights may be NULL. */
typedef struct fn_batch fn_batch;
int32_t fn_batch_create(fn_plan *const *plans, int32_t n_plans, int32_t n_streams, fn_batch **batch);
int32_t fn_batch_execute(fn_batch *batch, const void *const *d_rf, const int32_t *rf_dtypes, float *const *d_images,
                         void *stream);
/* Plans that share a kernel class (mode, element count class, transmit tiles,
 * window classes, precision; not ensembles or windowed normalization) are
 * beamformed by persistent launches over all their work items -- uneven grids,
 * angle counts and bin chunkings in one kernel whose workgroups claim
 * (plan, voxel tile, bin chunk, transmit group) items from a device counter,
 * longest chunks first.  A class is cut into launches of at most 8 plans, each
 * waiting only for its own plans' spectra.  enable = 0 launches every plan on
 * its own (each still a graph node); the next execute captures again.
 * Default: disabled (one launch per plan measured faster on the mixed
 * ensemble + job-queue batch since the pipelined tile staging). */
int32_t fn_batch_set_persistent(fn_batch *batch, int32_t enable);
/* Row stacks: single-frame plane-wave plans (up to 128 elements, at most 48
 * transmit rows each) that differ only in their steering angles and RF --
 * equal voxel grid, sample count, band, F-numbers, window, normalization,
 * semantics, precision, speed of sound and delay offsets -- are beamformed by
 * one launch over their stacked transmit rows (16 per matrix-core tile), each
 * row added to its own plan's image.  The per-(voxel, bin, element) weights
 * are then generated once for up to 64 rows of several plans.  No counterpart
 * in the reference (one MEX call per acquisition).  Default: enabled; enable =
 * 0 beamforms every plan on its own; the next execute captures again.
 * fn_batch_stacks reports the stacks of the last capture and the plans in them. */
int32_t fn_batch_set_stacking(fn_batch *batch, int32_t enable);
int32_t fn_batch_stacks(const fn_batch *batch, int32_t *n_stacks, int32_t *n_stacked_plans);
/* graph nodes and the number of captures so far */
int32_t fn_batch_info(const fn_batch *batch, int64_t *n_nodes, int32_t *n_captures);
int32_t fn_batch_destroy(fn_batch *batch);

/* Focused RF signal of the last voxel the CPU loop processes (third output of
 * das_pw, das_pw.m:477-481; fourth of das_sa, das_sa.m:501-505; on the GPU the
 * reference returns neither): the focused spectrum of that voxel -- normalized
 * apodization x phase x analytic spectra, summed over the elements (SA: and the
 * transmitters; PW plans with several angles: and the angles) -- placed at its
 * bins and inverse-DFT'd to d_signal, complex float [N_dft], in the CPU path's
 * scale.  Uses the spectra of the plan's last execute (same stream). */
int32_t fn_plan_focused_signal(fn_plan *plan, float *d_signal, void *stream);

/* ---- differentiation form (plane-wave plans) ---------------------------- */
/* No counterpart in the reference (its MEX only beamforms); these extend the
 * plan of fn_plan_pw_create with the time-of-flight parameters
 *   theta = (c_tof, delay_k): tau_tx = geometric_tx / c_tof + i_t0 / f_s,
 *                              tau_rx(k) = geometric_rx(k) / c_tof + delay_k
 * (das_pw.m:336, 339-340, 427-428 with c_0 -> c_tof and the offsets added).
 * Frequency grid, F-numbers, apertures and apodization stay those of the
 * plan's c_0, so the image is smooth in theta.
 * c_tof <= 0 restores c_0; delay_offsets (host, [n_elements], seconds) NULL = 0.
 * Also applies to fn_plan_execute (SA plans: the offset of element a delays
 * both its transmit and its receive). */
int32_t fn_plan_set_tof(fn_plan *plan, double c_tof, const double *delay_offsets);
/* image and its tangent dI/dc_tof (same layout), FP32; d_weights may be NULL */
int32_t fn_plan_execute_tangent(fn_plan *plan, const void *d_rf, int32_t rf_dtype, float *d_image,
                                float *d_dimage_dc, float *d_weights, void *stream);
/* adjoint (backprojection) of the RF -> image map of fn_plan_execute:
 * d_y complex float [n_x][n_z] -> d_rf_adj float [n_tx][n_el][n_t] with
 * Re sum conj(A x) y = sum x (A^H y) (same output scale as the forward). */
int32_t fn_plan_adjoint(fn_plan *plan, const float *d_y, float *d_rf_adj, void *stream);
/* vector-Jacobian products for a cotangent d_g (complex float [n_x][n_z]):
 * d_vjp_delay[k] = sum_v Re(conj(g_v) dI_v / d delay_k)   (double [n_el], device)
 * and, when d_dimage_dc (from fn_plan_execute_tangent) and d_vjp_c are given,
 * d_vjp_c[0] = sum_v Re(conj(g_v) dI_v / dc_tof)           (double, device).
 * Recomputes the spectra of d_rf. */
int32_t fn_plan_vjp(fn_plan *plan, const void *d_rf, int32_t rf_dtype, const float *d_g,
                    const float *d_dimage_dc, double *d_vjp_delay, double *d_vjp_c, void *stream);

/* One-shot adjoint on host buffers (plans the shape, uploads, runs, copies back):
 * image_y complex float interleaved [n_x][n_z] -> rf_adj float [n_angles][n_elements][n_samples]
 * (MATLAB [N_t x N_el x N_a]), with the time-of-flight parameters of fn_plan_set_tof. */
int32_t fn_das_pw_adjoint(const fn_pw_params *p, double c_tof, const double *delay_offsets,
                          const float *image_y, float *rf_adj);

/* ---- synthetic acquisitions (seeded; the reference ships no data) ------ */
typedef struct fn_synth_params {
    uint32_t struct_size;
    uint64_t seed;                 /* scatterer seed; noise uses seed + 1000 */
    int32_t n_points, n_speckle;   /* point scatterers (amplitude point_amplitude) + Gaussian speckle */
    double point_amplitude;
    double x_min, x_max, z_min, z_max;
    int32_t n_samples, n_elements;
    double f_s, element_pitch, c_0;
    double f_c, fractional_bandwidth;  /* Gaussian-modulated pulse */
    double snr_db;                 /* white noise relative to signal RMS; +inf: none */
    int32_t mode;                  /* 0: plane waves (angles), 1: synthetic aperture, 2: virtual sources */
    int32_t n_angles;
    const double *steering_angles;
    int32_t on_device;             /* 1: rf is a device pointer, computed on the GPU */
    int32_t index_gpu;
    /* medium: a layer 0 <= z < layer_thickness at layer_c over c_0 (0: homogeneous);
     * straight rays, the layer's share of a path is min(1, layer_thickness / z_s) */
    double layer_thickness, layer_c;
    const double *delay_offsets;   /* per-element receive delay offsets (s), NULL = 0 */
    /* flow: scatterers within tube_radius of the axis through (tube_x, tube_z) at
     * tube_angle (rad from +x) move along it with v = tube_v_max (1 - (r/R)^2),
     * displaced by v * frame_time and wrapped over the lateral extent; R = 0: none */
    double tube_x, tube_z, tube_radius, tube_angle, tube_v_max;
    double frame_time;
    /* mode 2: virtual-source transmits, [n_angles][2] = (x_vs, z_vs) in m, with the
     * transmit times of fn_pw_params.virtual_sources (homogeneous medium) */
    const double *virtual_sources;
} fn_synth_params;
/* rf: float32 column-major [n_samples x n_elements x (n_angles | n_elements)] */
int32_t fn_synth_rf(const fn_synth_params *p, float *rf);
/* the seeded scatterer list itself: xs, zs, amps of length n_points + n_speckle */
int32_t fn_synth_scatterers(const fn_synth_params *p, double *xs, double *zs, double *amps);

#ifdef __cplusplus
}
#endif
#endif /* FNUMBER_HIP_H */
