/*
 * oracle.h -- FP64 CPU restatement of the Fourier-domain delay-and-sum
 * beamformer with frequency-dependent F-number (das_pw.m / das_sa.m CPU branch).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the f_number_amd product links, loads
 * or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the timed
 * CPU baseline, never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * function (closed forms, invariants, brute force, the numpy transliteration
 * in oracle/das_numpy.py) lives in tests/test_oracle_*.py.
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* window ids: the reference MEX numbering (gpu_bf_windows.cuh:14-20,
 * gpu_bf_helper_functions.cu:56-97) */
enum { OR_WIN_BOXCAR = 0, OR_WIN_HANN = 1, OR_WIN_TUKEY = 2, OR_WIN_COSINE = 3, OR_WIN_TRIANGULAR = 4 };
/* semantics: CPU = das_pw.m/das_sa.m branch; MEX = +cuda kernels' choices (quirks Q1,Q3,Q4,Q5) */
enum { OR_SEM_CPU = 0, OR_SEM_MEX = 1 };
enum { OR_NORM_OFF = 0, OR_NORM_ON = 1, OR_NORM_WINDOWED = 2 };
/* F-number kinds (+f_numbers package) */
enum {
    OR_FN_CONSTANT = 0,          /* +f_numbers/constant.m */
    OR_FN_GRATING_ANGLE_LB = 1,  /* +f_numbers/+grating/angle_lb.m */
    OR_FN_DISTANCE_CONSTANT = 2, /* +f_numbers/+distance/constant.m */
    OR_FN_SZABO = 3,             /* +f_numbers/+directivity/szabo.m */
    OR_FN_DIRECTIVITY_SOFT = 4,  /* +f_numbers/+directivity/soft.m (obliquity soft) */
    OR_FN_DIRECTIVITY_RIGID = 5, /* +f_numbers/+directivity/rigid.m */
    OR_FN_PERROT = 6,            /* +f_numbers/+directivity/perrot.m */
    OR_FN_DISTANCE_NONE = 7      /* +f_numbers/+distance/none.m */
};

typedef struct {
    int kind;
    double p0, p1, p2; /* parameters, meaning per kind (see oracle.c) */
} or_fnumber_t;

/* das_pw.m:343-358 / das_sa.m:324-338 (CPU) and gpu_bf_das_pw_rf.cu:525-552 (MEX). */
int or_grid(int n_t, int n_el, double pitch, double z_pad, double f_s, double c_0,
            double f_lb, double f_ub, int *N_dft, int *n_lb, int *n_ub);

/* compute_values of one F-number object for an array of pitch-to-wavelength ratios. */
int or_fnumber_values(const or_fnumber_t *F, const double *r, int n, double *out);

/* window samples of |t| under CPU (+windows/<name>.m) or MEX (gpu_bf_windows.cu) semantics. */
double or_window(int id, double param, double t, int semantics);

/* Band-limited spectra X[ch][n - n_lb] = sum_t x[ch][t] e^{-j 2 pi n t / N}
 * (das_pw.m:383-384 without the factor 2).  mode 0: direct DFT with exact
 * integer index reduction; mode 1: Bluestein FFT. rf is [n_ch][n_t]
 * (column-major [n_t x n_ch]), float32 when rf_f32 != 0, else double. */
int or_spectra(const void *rf, int rf_f32, int n_t, int n_ch, int N, int n_lb, int n_ub,
               int mode, int n_threads, double *out_re, double *out_im);

typedef struct {
    /* grid */
    int n_x, n_z;
    const double *x, *z;
    /* RF: column-major [n_t x n_el x n_tx] (n_tx = angles for PW, tx elements for SA) */
    int n_t, n_el, n_tx;
    const void *rf;
    int rf_f32;
    double f_s, width, pitch, c_0, f_lb, f_ub, index_t0;
    const double *angles; /* PW only: n_tx steering angles (rad) */
    /* receive config (PW: the only window/F-number) */
    int win_rx; double win_rx_param;
    or_fnumber_t f_rx;
    int norm_rx; int norm_rx_win; double norm_rx_win_param;
    /* transmit config (SA only) */
    int win_tx; double win_tx_param;
    or_fnumber_t f_tx;
    int norm_tx; int norm_tx_win; double norm_tx_win_param;
    int semantics;
    int spectra_mode;
    int n_threads;
    /* optional voxel subset: when n_sel > 0 only voxels (sel_ix[i], sel_iz[i]) are
       computed and written to image[i]; otherwise image is [n_x][n_z] (z fastest) */
    int n_sel;
    const int *sel_ix, *sel_iz;
    /* outputs (CPU scale: 2*sum, das_pw.m:384,467) */
    double *image_re, *image_im;
    /* grid info out */
    int N_dft, n_lb, n_ub;
    /* optional (NULL: not computed; full grid only): focused RF signal of the last
       voxel the CPU loop processes, [N_dft] complex (das_pw.m:477-481, das_sa.m:501-505),
       angles summed for PW */
    double *signal_re, *signal_im;
    /* optional (PW): [n_tx][2] virtual sources (x_vs, z_vs) replacing the steering
       angles: diverging (z_vs < 0) t_tx = (|r - r_vs| - |r_vs|) / c, focused
       (z_vs > 0) t_tx = (|r_vs| + sign(z - z_vs) |r - r_vs|) / c */
    const double *virtual_sources;
    /* threads of the spectra stage (0: n_threads); lets a timing run keep the voxel
       loop single-threaded without a single-threaded DFT of every channel */
    int n_threads_spectra;
} or_das_args_t;

/* das_pw.m:302-474, compounded over the n_tx angles (examples.m:70-78). */
int or_das_pw(or_das_args_t *a);
/* das_sa.m:289-498 */
int or_das_sa(or_das_args_t *a);

/* gradient helper: the image as a function of the speed of sound with N_dft,
   band and apertures frozen at c_frozen (used by finite-difference checks of
   the differentiation form).  Same as or_das_pw but TOF uses c_tof. */
int or_das_pw_c(or_das_args_t *a, double c_tof, const double *delay_offsets);

/* Derivatives of the PW image with respect to the time-of-flight parameters
 * (speed of sound c_tof of the geometric delays, per-element receive delay
 * offsets), apertures and apodization frozen at c_0; and the adjoint of the
 * RF -> image map.  Each output is optional (NULL pointer = not computed):
 *   dIdc            image-shaped dI/dc_tof
 *   g -> vjp_delay  vjp_delay[k] = sum_v Re(conj(g_v) dI_v / d delay_k)
 *   g + dIdc -> vjp_c = sum_v Re(conj(g_v) dI_v / dc_tof)
 *   y -> adj_out    A^H y, [n_tx][n_el][n_t], with <A x, y> = Re sum conj(A x) y = <x, A^H y>
 * Scale: the image of or_das_pw (2 * sum). */
typedef struct {
    double *dIdc_re, *dIdc_im;
    const double *g_re, *g_im;
    double *vjp_delay;
    double vjp_c;
    const double *y_re, *y_im;
    double *adj_out;
} or_pw_diff_t;
int or_das_pw_diff(or_das_args_t *a, double c_tof, const double *delay_offsets, or_pw_diff_t *d);

#ifdef __cplusplus
}
#endif
#endif
