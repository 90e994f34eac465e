/*
 * oracle.c -- FP64 CPU restatement of the reference's MATLAB CPU path
 * (das_pw.m:302-485, das_sa.m:289-509) and of its configuration classes
 * (+f_numbers, +windows, +normalizations).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built by oracle/Makefile with
 * -O2 -ffp-contract=off and no fast-math so that every floating-point
 * operation is evaluated in the order the MATLAB expressions give.
 *
 * Parity status per function (pins are in tests/test_oracle_*.py):
 *   or_grid              pinned: integer closed forms at the configs
 *   or_fnumber_values    pinned: geometric invariants (grating-lobe angle,
 *                        angular distance), branch edges, closed forms;
 *                        directivity kinds: pinned against an exact root
 *                        within fminbnd's TolX (1e-4), MATLAB's fminbnd
 *                        iterates themselves are parity unpinned
 *   or_window            pinned: closed-form samples, symmetry, truncation
 *   or_spectra           pinned: brute-force long-double DFT, FFT vs DFT,
 *                        pure-tone bins, Parseval
 *   or_das_pw/or_das_sa  pinned: numpy transliteration of the MATLAB
 *                        statements (oracle/das_numpy.py), single-element /
 *                        single-bin closed forms, linearity, time-shift
 *                        phase, point-target focus, SA transmit/receive
 *                        reciprocity
 *   aperture/apod        pinned: per-element weights extracted from images of
 *                        unit impulses, against an element-scan construction
 *                        of the aperture (membership |x_k - x| <= h, half-widths
 *                        to the outer faces of the outermost members, side from
 *                        the sign of x_k - x), CPU and MEX semantics, all four
 *                        windows, normalization on and off, mirror symmetry;
 *                        the OR_PLANT_* variants below are built by the tests
 *                        to show that these pins fail on such mistakes
 *   memory / UB          the CPU oracle tests also run against a build with
 *                        AddressSanitizer and UndefinedBehaviorSanitizer
 *                        (OR_PLANT_OOB shows that this build reports a write
 *                        past the end of an output)
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* frequency grid                                                            */
/* ------------------------------------------------------------------------- */
/* das_pw.m:343 (N_samples_t_pad), :353 (N_points_dft), :356-357 (band indices,
 * converted to 0-based bins); identical in das_sa.m:324-338.  The MEX
 * (gpu_bf_das_pw_rf.cu:525-552) uses the same formulas with z_pad taken from
 * positions_z[1] instead of positions_z(1) (quirk Q1): the caller passes the
 * axial position to use. */
int or_grid(int n_t, int n_el, double pitch, double z_pad, double f_s, double c_0,
            double f_lb, double f_ub, int *N_dft, int *n_lb, int *n_ub)
{
    const double M = (n_el - 1) / 2.0;
    const double x_first = (0 - M) * pitch;
    const double x_last = ((n_el - 1) - M) * pitch;
    const double span = x_last - x_first; /* diff( positions_ctr_x( [ 1, end ] ) ) */
    const double pad = ceil((sqrt(span * span + z_pad * z_pad) - z_pad) * f_s / c_0);
    const int n_pad = (int)pad;
    const int N = n_t + n_pad + 1 - ((n_t + n_pad) % 2);
    *N_dft = N;
    *n_lb = (int)ceil(f_lb * N / f_s);
    *n_ub = (int)floor(f_ub * N / f_s);
    return (*n_lb <= *n_ub) ? 0 : -1;
}

/* ------------------------------------------------------------------------- */
/* F-numbers                                                                 */
/* ------------------------------------------------------------------------- */
static double deg2rad(double d) { return (M_PI / 180.0) * d; }

/* +f_numbers/+distance/constant.m:90-121 (properties :72-74). */
static double fn_distance_constant(double delta_deg, double F_ub, double r)
{
    const double s = sin(deg2rad(delta_deg));
    const double o = 1 + cos(deg2rad(delta_deg));
    const int ind_lb = r >= 1 / o;
    const int ind_ub = r < 1 / s;
    const int ind_transition = !ind_lb && r >= 0.5;
    const int ind_possible = ind_lb && ind_ub;
    double v = 0.0;
    if (ind_transition) {
        const double q = 1 / r - 1;
        v = sqrt(1 / (q * q) - 1) / 2;
    }
    if (ind_possible) {
        const double temp = r * r;
        v = (sqrt(2 * o * temp - 1) + o * s * temp) / (2 - 2 * (s * s) * temp);
        v = fmin(v, F_ub);
    }
    if (!ind_ub) v = F_ub;
    return v;
}

/* +f_numbers/+grating/angle_lb.m:44-75 with the properties set in
 * +f_numbers/+grating/grating.m:72-75. */
static double fn_angle_lb(double chi_deg, double F_ub, double delta_deg, double r)
{
    const double thresh = sin(deg2rad(chi_deg));
    const double lower = 1 / (1 + thresh);
    const double q = 2 * F_ub;
    const double upper = 1 / (1 / sqrt(1 + q * q) + thresh);
    double v = 0.0;
    if (r > lower && r < upper) {
        const double d = 1 / r - thresh;
        v = sqrt(1 / (d * d) - 1) / 2;
    }
    if (r >= upper) v = F_ub;
    if (delta_deg >= 0) {
        const double aa = fn_distance_constant(delta_deg, INFINITY, r);
        v = fmax(v, aa);
    }
    return v;
}

/* MATLAB sinc: sin(pi x)/(pi x), 1 at 0 */
static double msinc(double x)
{
    if (x == 0.0) return 1.0;
    return sin(M_PI * x) / (M_PI * x);
}

typedef struct { double w; int soft; double thresh; } dir_ctx_t;

/* objective of +f_numbers/+directivity/directivity.m:103-112 */
static double dir_obj(const dir_ctx_t *c, double theta)
{
    const double obl = c->soft ? cos(theta) : 1.0; /* +obliquity/soft.m, rigid.m */
    return fabs(obl * msinc(c->w * sin(theta)) - c->thresh);
}

/* Brent's derivative-free minimiser on [a,b] (golden section + parabolic
 * interpolation; R. P. Brent, "Algorithms for Minimization without
 * Derivatives", 1973, ch. 5 -- the method MATLAB's fminbnd documents) with
 * fminbnd's default TolX = 1e-4. */
static double brent_min(const dir_ctx_t *ctx, double a, double b, double tol)
{
    const double cgold = 0.5 * (3.0 - sqrt(5.0));
    const double eps = sqrt(2.220446049250313e-16);
    double v = a + cgold * (b - a), w = v, x = v, e = 0.0, d = 0.0;
    double fx = dir_obj(ctx, x), fv = fx, fw = fx;
    for (int it = 0; it < 500; ++it) {
        const double xm = 0.5 * (a + b);
        const double tol1 = eps * fabs(x) + tol / 3.0;
        const double tol2 = 2.0 * tol1;
        if (fabs(x - xm) <= tol2 - 0.5 * (b - a)) break;
        int golden = 1;
        if (fabs(e) > tol1) {
            double r = (x - w) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = fabs(q);
            const double etemp = e;
            e = d;
            if (!(fabs(p) >= fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x))) {
                d = p / q;
                const double u = x + d;
                if ((u - a) < tol2 || (b - u) < tol2) d = (xm >= x) ? tol1 : -tol1;
                golden = 0;
            }
        }
        if (golden) {
            e = (x >= xm) ? a - x : b - x;
            d = cgold * e;
        }
        const double u = (fabs(d) >= tol1) ? x + d : x + ((d >= 0) ? tol1 : -tol1);
        const double fu = dir_obj(ctx, u);
        if (fu <= fx) {
            if (u >= x) a = x; else b = x;
            v = w; fv = fw; w = x; fw = fx; x = u; fx = fu;
        } else {
            if (u < x) a = u; else b = u;
            if (fu <= fw || w == x) { v = w; fv = fw; w = u; fw = fu; }
            else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
        }
    }
    return x;
}

/* +f_numbers/+directivity/directivity.m:91-120 */
static double fn_directivity(double width_over_pitch, int soft, double att_dB, double r)
{
    dir_ctx_t c;
    c.w = width_over_pitch * r;
    c.soft = soft;
    c.thresh = pow(10.0, -att_dB / 20);
    const double alpha = brent_min(&c, 0.0, M_PI / 2, 1e-4);
    return 1 / (2 * tan(alpha));
}

/* f_number.m:50-114 dispatch to compute_values_scalar of the concrete class.
 * Parameters: CONSTANT p0=value (constant.m:72); GRATING_ANGLE_LB p0=chi_deg,
 * p1=F_ub, p2=delta_deg (grating.m:36-79); DISTANCE_CONSTANT p0=delta_deg,
 * p1=F_ub (constant.m:38-78); SZABO p0=width_over_pitch (szabo.m:83-86);
 * DIRECTIVITY_SOFT/RIGID p0=width_over_pitch, p1=attenuation_dB;
 * PERROT p0=width_over_pitch (perrot.m: soft, 3 dB); DISTANCE_NONE
 * (none.m: distance.constant(0, inf)). */
int or_fnumber_values(const or_fnumber_t *F, const double *r, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        if (!(r[i] > 0)) return -1; /* f_number.m:100 mustBePositive */
        switch (F->kind) {
        case OR_FN_CONSTANT: out[i] = F->p0; break;
        case OR_FN_GRATING_ANGLE_LB: out[i] = fn_angle_lb(F->p0, F->p1, F->p2, r[i]); break;
        case OR_FN_DISTANCE_CONSTANT: out[i] = fn_distance_constant(F->p0, F->p1, r[i]); break;
        case OR_FN_SZABO: out[i] = F->p0 * r[i] / 1.2; break;
        case OR_FN_DIRECTIVITY_SOFT: out[i] = fn_directivity(F->p0, 1, F->p1, r[i]); break;
        case OR_FN_DIRECTIVITY_RIGID: out[i] = fn_directivity(F->p0, 0, F->p1, r[i]); break;
        case OR_FN_PERROT: out[i] = fn_directivity(F->p0, 1, 3.0, r[i]); break;
        case OR_FN_DISTANCE_NONE: out[i] = fn_distance_constant(0.0, INFINITY, r[i]); break;
        default: return -2;
        }
    }
#ifdef OR_PLANT_OOB
    /* planted mistake (the sanitizer test shows that its build reports it): one
       element written past the end of out */
    if (n > 0) out[n] = out[n - 1];
#endif
    return 0;
}

/* ------------------------------------------------------------------------- */
/* windows                                                                   */
/* ------------------------------------------------------------------------- */
/* CPU: samples of |t| by +windows/boxcar.m:62, hann.m:62, tukey.m:101-105,
 * triangular.m:62 (hann and triangular are NOT truncated, quirk Q4).
 * MEX: gpu_bf_windows.cu:15-103 (apply() truncates every window at |t|>=1). */
double or_window(int id, double param, double t, int semantics)
{
    const double ta = fabs(t);
    if (semantics == OR_SEM_MEX) {
        if (ta >= 1.0) return 0.0;
        switch (id) {
        case OR_WIN_BOXCAR: return 1.0;
        case OR_WIN_HANN: return (1.0 + cos(M_PI * t)) / 2.0;
        case OR_WIN_TUKEY: {
            const double thr = 1.0 - param;
            if (ta <= thr) return 1.0;
            return (1.0 + cos(M_PI * (ta - thr) / param)) / 2.0;
        }
        case OR_WIN_COSINE: return cos(M_PI * t / 2.0) / (4.0 / M_PI);
        case OR_WIN_TRIANGULAR: return 1.0 - ta;
        default: return NAN;
        }
    }
    switch (id) {
    case OR_WIN_BOXCAR: return (double)(ta < 1);
    case OR_WIN_HANN: return (1 + cos(M_PI * ta)) / 2;
    case OR_WIN_TUKEY: {
        const double frac_rect = 1 - param;
        double s = (double)(ta < 1);
        if (ta < 1 && ta > frac_rect) {
            const double u = (ta - frac_rect) / param;
            s = (1 + cos(M_PI * u)) / 2;
        }
        return s;
    }
    case OR_WIN_TRIANGULAR: return 1 - ta;
    default: return NAN; /* cosine has no MATLAB class */
    }
}

/* ------------------------------------------------------------------------- */
/* spectra                                                                   */
/* ------------------------------------------------------------------------- */
static inline double rf_at(const void *rf, int f32, long idx)
{
    return f32 ? (double)((const float *)rf)[idx] : ((const double *)rf)[idx];
}

/* in-place iterative radix-2 FFT of length m (power of two), sign -1 forward */
static void fft_pow2(double *re, double *im, int m, int sign, const double *wr, const double *wi)
{
    for (int i = 1, j = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        const int step = m / len;
        for (int i = 0; i < m; i += len) {
            for (int k = 0; k < len / 2; ++k) {
                const double cr = wr[k * step];
                const double ci = sign * wi[k * step];
                const double ur = re[i + k], ui = im[i + k];
                const double xr = re[i + k + len / 2], xi = im[i + k + len / 2];
                const double vr = xr * cr - xi * ci;
                const double vi = xr * ci + xi * cr;
                re[i + k] = ur + vr; im[i + k] = ui + vi;
                re[i + k + len / 2] = ur - vr; im[i + k + len / 2] = ui - vi;
            }
        }
    }
}

/* X[ch][n-n_lb] = sum_{t<n_t} x[ch][t] e^{-j2 pi n t/N}: the bins das_pw.m:383-384
 * selects from fft(data_RF, N_points_dft, 1) (zero padding to N). */
int or_spectra(const void *rf, int rf_f32, int n_t, int n_ch, int N, int n_lb, int n_ub,
               int mode, int n_threads, double *out_re, double *out_im)
{
    const int nf = n_ub - n_lb + 1;
    if (nf <= 0 || n_t > N) return -1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    if (mode == 0) {
        /* direct DFT, twiddle index (n t) mod N reduced exactly in integers */
        double *cw = (double *)malloc(sizeof(double) * N);
        double *sw = (double *)malloc(sizeof(double) * N);
        for (int m = 0; m < N; ++m) {
            const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)N;
            cw[m] = (double)cosl(ang);
            sw[m] = (double)sinl(ang);
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int ch = 0; ch < n_ch; ++ch) {
            for (int f = 0; f < nf; ++f) {
                const long n = n_lb + f;
                double ar = 0.0, ai = 0.0;
                long m = 0;
                for (int t = 0; t < n_t; ++t) {
                    const double x = rf_at(rf, rf_f32, (long)ch * n_t + t);
                    ar += x * cw[m];
                    ai -= x * sw[m];
                    m += n;
                    if (m >= N) m %= N;
                }
                out_re[(long)ch * nf + f] = ar;
                out_im[(long)ch * nf + f] = ai;
            }
        }
        free(cw);
        free(sw);
        return 0;
    }
    /* Bluestein chirp-z: X(n) = conj(b_n) sum_t (x_t conj(b_t)) b_{n-t},
       b_m = e^{j pi m^2 / N}, evaluated by a power-of-two circular convolution. */
    int m2 = 1;
    while (m2 < 2 * N - 1) m2 <<= 1;
    double *br = (double *)malloc(sizeof(double) * N), *bi = (double *)malloc(sizeof(double) * N);
    for (long m = 0; m < N; ++m) {
        const long k = (m * m) % (2L * N); /* e^{j pi m^2/N} is 2N-periodic in m^2 */
        const long double ang = 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
        br[m] = (double)cosl(ang);
        bi[m] = (double)sinl(ang);
    }
    double *twr = (double *)malloc(sizeof(double) * m2 / 2), *twi = (double *)malloc(sizeof(double) * m2 / 2);
    for (int k = 0; k < m2 / 2; ++k) {
        const long double ang = 2.0L * 3.141592653589793238462643383279502884L * k / m2;
        twr[k] = (double)cosl(ang);
        twi[k] = (double)sinl(ang);
    }
    /* FFT of the chirp filter h_m = b_m for |m| < N (circular) */
    double *hr = (double *)calloc(m2, sizeof(double)), *hi = (double *)calloc(m2, sizeof(double));
    for (int m = 0; m < N; ++m) {
        hr[m] = br[m]; hi[m] = bi[m];
        if (m > 0) { hr[m2 - m] = br[m]; hi[m2 - m] = bi[m]; }
    }
    fft_pow2(hr, hi, m2, -1, twr, twi);
#pragma omp parallel
    {
        double *ar = (double *)malloc(sizeof(double) * m2), *ai = (double *)malloc(sizeof(double) * m2);
#pragma omp for schedule(dynamic, 4)
        for (int ch = 0; ch < n_ch; ++ch) {
            memset(ar, 0, sizeof(double) * m2);
            memset(ai, 0, sizeof(double) * m2);
            for (int t = 0; t < n_t; ++t) {
                const double x = rf_at(rf, rf_f32, (long)ch * n_t + t);
                ar[t] = x * br[t];   /* x_t * conj(b_t) */
                ai[t] = -x * bi[t];
            }
            fft_pow2(ar, ai, m2, -1, twr, twi);
            for (int k = 0; k < m2; ++k) {
                const double pr = ar[k] * hr[k] - ai[k] * hi[k];
                const double pi_ = ar[k] * hi[k] + ai[k] * hr[k];
                ar[k] = pr; ai[k] = pi_;
            }
            fft_pow2(ar, ai, m2, +1, twr, twi);
            for (int f = 0; f < nf; ++f) {
                const int n = n_lb + f;
                const double cr = ar[n] / m2, ci = ai[n] / m2;
                /* times conj(b_n) */
                out_re[(long)ch * nf + f] = cr * br[n] + ci * bi[n];
                out_im[(long)ch * nf + f] = ci * br[n] - cr * bi[n];
            }
        }
        free(ar);
        free(ai);
    }
    free(br); free(bi); free(twr); free(twi); free(hr); free(hi);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* apertures                                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int valid;
    int lb, ub;        /* 0-based element indices (das_pw.m:399-400 minus one) */
    double WL, WR;     /* snapped half-widths (das_pw.m:407,411) */
    double h;          /* desired half-width z/(2F) (das_pw.m:387) */
} aperture_t;

/* das_pw.m:387, 399-411 for one (bin, z) at lateral position x. */
static aperture_t aperture(double x, double h, double M, double pitch, int n_el, double width)
{
    aperture_t ap;
    ap.h = h;
#ifdef OR_PLANT_CEIL_FLOOR
    /* planted mistake (tests/test_oracle_aperture.py builds a variant with it and
       checks that the aperture pins catch it): rounding directions exchanged */
    const double a = floor(M + (x - h) / pitch);
    const double b = ceil(M + (x + h) / pitch);
#else
    const double a = ceil(M + (x - h) / pitch);
    const double b = floor(M + (x + h) / pitch);
#endif
    const double lbd = fmax(a, 0.0);          /* max( ceil(...) + 1, 1 ) - 1 */
    const double ubd = fmin(b, n_el - 1.0);   /* min( floor(...) + 1, N ) - 1 */
    ap.valid = lbd <= ubd;
    ap.lb = ap.valid ? (int)lbd : 0;
    ap.ub = ap.valid ? (int)ubd : -1;
    ap.WL = ap.WR = 0.0;
    if (ap.valid) {
        const double w2 = width / 2;
        const double xlb = ((ap.lb - M) * pitch) - w2; /* positions_lbs_x */
        const double xub = ((ap.ub - M) * pitch) + w2; /* positions_ubs_x */
#ifdef OR_PLANT_SWAP_W
        /* planted mistake (see OR_PLANT_CEIL_FLOOR): half-widths exchanged */
        ap.WL = xub - x;
        ap.WR = x - xlb;
#else
        ap.WL = x - xlb;
        ap.WR = xub - x;
#endif
    }
    return ap;
}

/* receive (or transmit) apodization of element k for one aperture:
 * CPU: das_pw.m:436-454; MEX: das_kernel_phase_shifts.cu:90-125. */
static double apod(const aperture_t *ap, double d, int win, double param, int semantics)
{
    if (semantics == OR_SEM_CPU) {
        if (win == OR_WIN_BOXCAR) return or_window(win, param, d / ap->h, OR_SEM_CPU);
        if (d < 0) return or_window(win, param, d / ap->WL, OR_SEM_CPU);
        return or_window(win, param, d / ap->WR, OR_SEM_CPU);
    }
    /* MEX: the window is applied with the snapped half-width of the side */
    if (d < 0) return or_window(win, param, d / ap->WL, OR_SEM_MEX);
    return or_window(win, param, d / ap->WR, OR_SEM_MEX);
}

/* normalizations.windowed (windowed.m:81-86): after "on", each element column
 * is tapered over the valid-frequency rows by window((i - L/2)/(L/2)) with
 * L = number of rows with positive weight in that column. */
static void windowed_taper(double *w, int nfv, int n_el, int win, double param)
{
    for (int k = 0; k < n_el; ++k) {
        int cnt = 0;
        for (int i = 0; i < nfv; ++i) cnt += w[(long)i * n_el + k] > 0;
        const double L2 = cnt / 2.0;
        for (int i = 0; i < nfv; ++i) {
            const double t = ((double)i - L2) / L2;
            w[(long)i * n_el + k] *= or_window(win, param, t, OR_SEM_CPU);
        }
    }
}

/* F value for one bin at one depth: das_pw.m:368-377 (clamp only with CPU
 * semantics and only for PW, quirk Q3). */
static double f_at(double F_n, double f_n, double z, double c_0, int clamp)
{
    if (!clamp) return F_n;
    const double F_ub = sqrt(f_n * z / (2.88 * c_0));
    return fmin(F_n, F_ub);
}

/* focused RF signal (das_pw.m:477-481): s[t] = ifft of the focused spectrum
   placed at its bins, s[t] = (1/N) sum_i F[i] e^{+j 2 pi (n_lb + i) t / N} */
static void signal_idft(const double *fr, const double *fi, int nf, int nlb, int N, double *out_re, double *out_im)
{
    for (int t = 0; t < N; ++t) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < nf; ++i) {
            const long m = ((long)(nlb + i) * t) % N;
            const double ang = 2 * M_PI * (double)m / N, c = cos(ang), s = sin(ang);
            sr += fr[i] * c - fi[i] * s;
            si += fr[i] * s + fi[i] * c;
        }
        out_re[t] = sr / N;
        out_im[t] = si / N;
    }
}

/* ------------------------------------------------------------------------- */
/* das_pw                                                                    */
/* ------------------------------------------------------------------------- */
static int das_pw_impl(or_das_args_t *a, double c_tof, const double *delay_offsets, or_pw_diff_t *dx)
{
    const int n_el = a->n_el;
    const double M = (n_el - 1) / 2.0;
    const double z_pad = (a->semantics == OR_SEM_MEX && a->n_z > 1) ? a->z[1] : a->z[0];
    int N, nlb, nub;
    if (or_grid(a->n_t, n_el, a->pitch, z_pad, a->f_s, a->c_0, a->f_lb, a->f_ub, &N, &nlb, &nub)) return -1;
    a->N_dft = N; a->n_lb = nlb; a->n_ub = nub;
    const int nf = nub - nlb + 1;
    const double c0 = a->c_0;

    /* frequency axis and F-number values: das_pw.m:361-371 */
    double *fax = (double *)malloc(sizeof(double) * nf);
    double *om = (double *)malloc(sizeof(double) * nf);
    double *r = (double *)malloc(sizeof(double) * nf);
    double *Fv = (double *)malloc(sizeof(double) * nf);
    for (int i = 0; i < nf; ++i) {
        fax[i] = a->f_s * (double)(nlb + i) / N;
        om[i] = 2 * M_PI * fax[i];
        r[i] = a->pitch * fax[i] / c0;
    }
    if (or_fnumber_values(&a->f_rx, r, nf, Fv)) { free(fax); free(om); free(r); free(Fv); return -2; }

    /* spectra of every channel of every angle (das_pw.m:383-384, factor 2 applied below) */
    const long nch = (long)n_el * a->n_tx;
    double *Xr = (double *)malloc(sizeof(double) * nch * nf);
    double *Xi = (double *)malloc(sizeof(double) * nch * nf);
    or_spectra(a->rf, a->rf_f32, a->n_t, (int)nch, N, nlb, nub, a->spectra_mode,
               a->n_threads_spectra > 0 ? a->n_threads_spectra : a->n_threads, Xr, Xi);

    const int clamp = (a->semantics == OR_SEM_CPU);
    const int nvox = a->n_sel > 0 ? a->n_sel : a->n_x * a->n_z;
    for (long v = 0; v < nvox; ++v) { a->image_re[v] = 0.0; a->image_im[v] = 0.0; }
    /* the last voxel the CPU loop processes: last lateral position with a valid
       frequency, its last such axial position (das_pw.m:393-474) */
    long v_sig = -1;
    double *sig_fr = NULL, *sig_fi = NULL;
    if (a->signal_re && a->n_sel == 0) {
        for (int ix = a->n_x - 1; ix >= 0 && v_sig < 0; --ix)
            for (int iz = a->n_z - 1; iz >= 0 && v_sig < 0; --iz)
                for (int i = 0; i < nf; ++i) {
                    const double F = f_at(Fv[i], fax[i], a->z[iz], c0, clamp);
                    if (aperture(a->x[ix], a->z[iz] / (2 * F), M, a->pitch, n_el, a->width).valid) {
                        v_sig = (long)ix * a->n_z + iz;
                        break;
                    }
                }
        sig_fr = (double *)calloc(nf, sizeof(double));
        sig_fi = (double *)calloc(nf, sizeof(double));
    }
    const int want_dc = dx && dx->dIdc_re;
    const int want_vjp = dx && dx->g_re;
    const int want_adj = dx && dx->y_re;
    const long nz_adj = want_adj ? (long)a->n_tx * n_el * nf : 0;
    double *Zr_all = NULL, *Zi_all = NULL;
    double vjp_c = 0.0;
    if (want_dc)
        for (long v = 0; v < nvox; ++v) { dx->dIdc_re[v] = 0.0; dx->dIdc_im[v] = 0.0; }
    if (want_vjp)
        for (int k = 0; k < n_el; ++k) dx->vjp_delay[k] = 0.0;
    if (want_adj) {
        Zr_all = (double *)calloc(nz_adj, sizeof(double));
        Zi_all = (double *)calloc(nz_adj, sizeof(double));
    }

#ifdef _OPENMP
    if (a->n_threads > 0) omp_set_num_threads(a->n_threads);
#endif
#pragma omp parallel
    {
        double *tof = (double *)malloc(sizeof(double) * n_el);
        double *dtof = (double *)malloc(sizeof(double) * n_el);  /* d tof / d c_tof */
        double *vjp = want_vjp ? (double *)calloc(n_el, sizeof(double)) : NULL;
        double *Zr = want_adj ? (double *)calloc(nz_adj, sizeof(double)) : NULL;
        double *Zi = want_adj ? (double *)calloc(nz_adj, sizeof(double)) : NULL;
        double vjp_c_t = 0.0;
        double *w = (double *)malloc(sizeof(double) * (long)nf * n_el);
        int *vidx = (int *)malloc(sizeof(int) * nf);
        aperture_t *aps = (aperture_t *)malloc(sizeof(aperture_t) * nf);
#pragma omp for schedule(dynamic, 1)
        for (long v = 0; v < nvox; ++v) {
            const int ix = a->n_sel > 0 ? a->sel_ix[v] : (int)(v / a->n_z);
            const int iz = a->n_sel > 0 ? a->sel_iz[v] : (int)(v % a->n_z);
            const double x = a->x[ix], z = a->z[iz];
            /* apertures (das_pw.m:387, 399-411) */
            int nfv = 0;
            for (int i = 0; i < nf; ++i) {
                const double F = f_at(Fv[i], fax[i], z, c0, clamp);
                const double h = z / (2 * F);
                aps[i] = aperture(x, h, M, a->pitch, n_el, a->width);
                if (aps[i].valid) vidx[nfv++] = i;
            }
            if (nfv == 0) continue; /* das_pw.m:414 */
            /* apodization weights of the valid frequencies (das_pw.m:436-457) */
            for (int j = 0; j < nfv; ++j) {
                const aperture_t *ap = &aps[vidx[j]];
                double s = 0.0;
                for (int k = 0; k < n_el; ++k) {
                    const double d = ((k - M) * a->pitch) - x; /* dist_lateral */
                    const double wk = apod(ap, d, a->win_rx, a->win_rx_param, a->semantics);
                    w[(long)j * n_el + k] = wk;
                    s += wk;
                }
                if (a->norm_rx != OR_NORM_OFF)
                    for (int k = 0; k < n_el; ++k) w[(long)j * n_el + k] = w[(long)j * n_el + k] / s;
            }
            /* the MEX reads the taper window but never applies it (gpu_bf_das_pw_rf.cu:421-427) */
            if (a->norm_rx == OR_NORM_WINDOWED && a->semantics == OR_SEM_CPU)
                windowed_taper(w, nfv, n_el, a->norm_rx_win, a->norm_rx_win_param);

            double acc_r = 0.0, acc_i = 0.0, dc_r = 0.0, dc_i = 0.0;
            const double gr = want_vjp ? dx->g_re[v] : 0.0, gi = want_vjp ? dx->g_im[v] : 0.0;
            const double yr = want_adj ? dx->y_re[v] : 0.0, yi = want_adj ? dx->y_im[v] : 0.0;
            for (int ia = 0; ia < a->n_tx; ++ia) {
                double t_geo;
                if (a->virtual_sources) {
                    /* virtual source: spherical wavefront from (x_vs, z_vs) */
                    const double xv = a->virtual_sources[2 * ia], zv = a->virtual_sources[2 * ia + 1];
                    const double r = sqrt((x - xv) * (x - xv) + (z - zv) * (z - zv));
                    const double r0 = sqrt(xv * xv + zv * zv);
                    t_geo = (zv < 0 ? r - r0 : r0 + (z >= zv ? r : -r)) / c_tof;
                } else {
                    const double th = a->angles[ia];
                    const double ex = -sin(th), ez = cos(th);
                    const double sg = (ex > 0) - (ex < 0);
                    const double xref = sg * ((0 - M) * a->pitch);
                    /* das_pw.m:336, 339-340, 427-428 */
                    t_geo = (ex * (x - xref) + ez * z) / c_tof;
                }
                const double t_in = t_geo + a->index_t0 / a->f_s;
                const double zc = z / c_tof;
                for (int k = 0; k < n_el; ++k) {
                    const double lat = (((k - M) * a->pitch) - x) / c_tof;
                    const double t_rx = sqrt(lat * lat + zc * zc);
                    tof[k] = t_in + t_rx;
                    if (delay_offsets) tof[k] += delay_offsets[k];
                    /* geometric times scale as 1 / c: d/dc = -(t_geo + t_rx) / c */
                    dtof[k] = -(t_geo + t_rx) / c_tof;
                }
                double img_r = 0.0, img_i = 0.0;
                for (int j = 0; j < nfv; ++j) {
                    const int i = vidx[j];
                    double fr = 0.0, fi = 0.0;
                    for (int k = 0; k < n_el; ++k) {
                        const double arg = om[i] * tof[k];
                        const double wr = cos(arg), wi = sin(arg);
                        const long ch = (long)ia * n_el + k;
                        const double xr = 2 * Xr[ch * nf + i], xi = 2 * Xi[ch * nf + i];
                        const double ww = w[(long)j * n_el + k];
                        /* (window .* weights) .* data */
                        const double pr = ww * wr, pi_ = ww * wi;
                        const double tr = pr * xr - pi_ * xi, ti = pr * xi + pi_ * xr;
                        fr += tr;
                        fi += ti;
                        if (want_dc || want_vjp) {
                            /* d/dtau of the term: j om term */
                            const double dr = -om[i] * ti, di = om[i] * tr;
                            if (want_dc) { dc_r += dtof[k] * dr; dc_i += dtof[k] * di; }
                            if (want_vjp) vjp[k] += gr * dr + gi * di; /* Re(conj(g) dI/d delay_k) */
                        }
                        if (want_adj) {
                            /* adjoint: y conj(2 ww e^{j arg}) accumulated per (angle, element, bin) */
                            const long zi = ((long)ia * n_el + k) * nf + i;
                            Zr[zi] += 2 * ww * (yr * wr + yi * wi);
                            Zi[zi] += 2 * ww * (yi * wr - yr * wi);
                        }
                    }
                    img_r += fr;
                    img_i += fi;
                    if (v == v_sig) { sig_fr[i] += fr; sig_fi[i] += fi; }
                }
                acc_r += img_r;
                acc_i += img_i;
            }
            a->image_re[v] = acc_r;
            a->image_im[v] = acc_i;
            if (want_dc) { dx->dIdc_re[v] = dc_r; dx->dIdc_im[v] = dc_i; }
            if (want_vjp && want_dc) vjp_c_t += gr * dc_r + gi * dc_i;
        }
#pragma omp critical
        {
            if (want_vjp) for (int k = 0; k < n_el; ++k) dx->vjp_delay[k] += vjp[k];
            vjp_c += vjp_c_t;
            if (want_adj) for (long q = 0; q < nz_adj; ++q) { Zr_all[q] += Zr[q]; Zi_all[q] += Zi[q]; }
        }
        free(tof); free(dtof); free(w); free(vidx); free(aps); free(vjp); free(Zr); free(Zi);
    }
    if (want_vjp) dx->vjp_c = vjp_c;
    if (want_adj) {
        /* band-limited inverse DFT of Z back to the RF samples t < n_t:
           out[ia][k][t] = Re sum_n Z[ia][k][n] e^{+j 2 pi n t / N}, index n t reduced mod N exactly */
        const long nch2 = (long)a->n_tx * n_el;
#pragma omp parallel for schedule(static)
        for (long ch = 0; ch < nch2; ++ch)
            for (int t = 0; t < a->n_t; ++t) {
                double acc = 0.0;
                for (int i = 0; i < nf; ++i) {
                    const long m = ((long)(nlb + i) * t) % N;
                    const double ang = 2 * M_PI * (double)m / N;
                    acc += Zr_all[ch * nf + i] * cos(ang) - Zi_all[ch * nf + i] * sin(ang);
                }
                dx->adj_out[ch * a->n_t + t] = acc;
            }
        free(Zr_all);
        free(Zi_all);
    }
    if (sig_fr) {
        signal_idft(sig_fr, sig_fi, nf, nlb, N, a->signal_re, a->signal_im);
        free(sig_fr);
        free(sig_fi);
    }
    free(fax); free(om); free(r); free(Fv); free(Xr); free(Xi);
    return 0;
}

int or_das_pw(or_das_args_t *a) { return das_pw_impl(a, a->c_0, NULL, NULL); }

int or_das_pw_c(or_das_args_t *a, double c_tof, const double *delay_offsets)
{
    return das_pw_impl(a, c_tof, delay_offsets, NULL);
}

int or_das_pw_diff(or_das_args_t *a, double c_tof, const double *delay_offsets, or_pw_diff_t *d)
{
    return das_pw_impl(a, c_tof, delay_offsets, d);
}

/* ------------------------------------------------------------------------- */
/* das_sa                                                                    */
/* ------------------------------------------------------------------------- */
/* das_sa.m:297-498.  RF is [n_t x n_rx x n_tx] with n_rx == n_tx == n_el.
 * A frequency contributes at a voxel when both its transmit and its receive
 * apertures are non-empty (das_sa.m:414; the reference's own statements
 * :485-486 are only defined when both masks agree, which holds when
 * transmit and receive share F-number and window). */
int or_das_sa(or_das_args_t *a)
{
    const int n_el = a->n_el;
    const double M = (n_el - 1) / 2.0;
    const double z_pad = (a->semantics == OR_SEM_MEX && a->n_z > 1) ? a->z[1] : a->z[0];
    int N, nlb, nub;
    if (a->n_tx != n_el) return -3;
    if (or_grid(a->n_t, n_el, a->pitch, z_pad, a->f_s, a->c_0, a->f_lb, a->f_ub, &N, &nlb, &nub)) return -1;
    a->N_dft = N; a->n_lb = nlb; a->n_ub = nub;
    const int nf = nub - nlb + 1;
    const double c0 = a->c_0;
    double *fax = (double *)malloc(sizeof(double) * nf);
    double *om = (double *)malloc(sizeof(double) * nf);
    double *r = (double *)malloc(sizeof(double) * nf);
    double *Ftx = (double *)malloc(sizeof(double) * nf);
    double *Frx = (double *)malloc(sizeof(double) * nf);
    for (int i = 0; i < nf; ++i) {
        fax[i] = a->f_s * (double)(nlb + i) / N;
        om[i] = 2 * M_PI * fax[i];
        r[i] = a->pitch * fax[i] / c0;
    }
    if (or_fnumber_values(&a->f_tx, r, nf, Ftx) || or_fnumber_values(&a->f_rx, r, nf, Frx)) {
        free(fax); free(om); free(r); free(Ftx); free(Frx);
        return -2;
    }
    const long nch = (long)n_el * n_el;
    double *Xr = (double *)malloc(sizeof(double) * nch * nf);
    double *Xi = (double *)malloc(sizeof(double) * nch * nf);
    or_spectra(a->rf, a->rf_f32, a->n_t, (int)nch, N, nlb, nub, a->spectra_mode,
               a->n_threads_spectra > 0 ? a->n_threads_spectra : a->n_threads, Xr, Xi);

    const int nvox = a->n_sel > 0 ? a->n_sel : a->n_x * a->n_z;
    /* last voxel the CPU loop processes (das_sa.m:375-498): both apertures valid */
    long v_sig = -1;
    double *sig_fr = NULL, *sig_fi = NULL;
    if (a->signal_re && a->n_sel == 0) {
        for (int ix = a->n_x - 1; ix >= 0 && v_sig < 0; --ix)
            for (int iz = a->n_z - 1; iz >= 0 && v_sig < 0; --iz)
                for (int i = 0; i < nf; ++i) {
                    const double zz = a->z[iz], xx = a->x[ix];
                    if (aperture(xx, zz / (2 * Ftx[i]), M, a->pitch, n_el, a->width).valid &&
                        aperture(xx, zz / (2 * Frx[i]), M, a->pitch, n_el, a->width).valid) {
                        v_sig = (long)ix * a->n_z + iz;
                        break;
                    }
                }
        sig_fr = (double *)calloc(nf, sizeof(double));
        sig_fi = (double *)calloc(nf, sizeof(double));
    }
#ifdef _OPENMP
    if (a->n_threads > 0) omp_set_num_threads(a->n_threads);
#endif
#pragma omp parallel
    {
        double *tsc = (double *)malloc(sizeof(double) * n_el);
        double *wtx = (double *)malloc(sizeof(double) * n_el);
        double *wrx = (double *)malloc(sizeof(double) * n_el);
        double *ftr = (double *)malloc(sizeof(double) * n_el);
        double *fti = (double *)malloc(sizeof(double) * n_el);
        double *er = (double *)malloc(sizeof(double) * n_el);
        double *ei = (double *)malloc(sizeof(double) * n_el);
        double *wtx_all = (double *)malloc(sizeof(double) * (long)nf * n_el);
        double *wrx_all = (double *)malloc(sizeof(double) * (long)nf * n_el);
        int *vidx = (int *)malloc(sizeof(int) * nf);
#pragma omp for schedule(dynamic, 1)
        for (long v = 0; v < nvox; ++v) {
            const int ix = a->n_sel > 0 ? a->sel_ix[v] : (int)(v / a->n_z);
            const int iz = a->n_sel > 0 ? a->sel_iz[v] : (int)(v % a->n_z);
            const double x = a->x[ix], z = a->z[iz];
            int nfv = 0;
            for (int i = 0; i < nf; ++i) {
                /* das_sa.m:368-369, 384-414 (no depth clamp: das_sa.m:354-358) */
                const aperture_t aptx = aperture(x, z / (2 * Ftx[i]), M, a->pitch, n_el, a->width);
                const aperture_t aprx = aperture(x, z / (2 * Frx[i]), M, a->pitch, n_el, a->width);
                if (!(aptx.valid && aprx.valid)) continue;
                double st = 0.0, sr = 0.0;
                for (int k = 0; k < n_el; ++k) {
                    const double d = ((k - M) * a->pitch) - x;
                    const double at = apod(&aptx, d, a->win_tx, a->win_tx_param, a->semantics);
                    const double ar = apod(&aprx, d, a->win_rx, a->win_rx_param, a->semantics);
                    wtx_all[(long)nfv * n_el + k] = at;
                    wrx_all[(long)nfv * n_el + k] = ar;
                    st += at; sr += ar;
                }
                /* das_sa.m:479-480 */
                for (int k = 0; k < n_el; ++k) {
                    if (a->norm_tx != OR_NORM_OFF) wtx_all[(long)nfv * n_el + k] /= st;
                    if (a->norm_rx != OR_NORM_OFF) wrx_all[(long)nfv * n_el + k] /= sr;
                }
                vidx[nfv++] = i;
            }
            if (nfv == 0) { a->image_re[v] = 0.0; a->image_im[v] = 0.0; continue; }
            /* MEX semantics: the taper window is read but not applied (gpu_bf_saft_rf.cu:485-490) */
            if (a->semantics == OR_SEM_CPU) {
                if (a->norm_tx == OR_NORM_WINDOWED)
                    windowed_taper(wtx_all, nfv, n_el, a->norm_tx_win, a->norm_tx_win_param);
                if (a->norm_rx == OR_NORM_WINDOWED)
                    windowed_taper(wrx_all, nfv, n_el, a->norm_rx_win, a->norm_rx_win_param);
            }
            /* das_sa.m:431 */
            const double zc = z / c0;
            for (int k = 0; k < n_el; ++k) {
                const double lat = (((k - M) * a->pitch) - x) / c0;
                tsc[k] = sqrt(lat * lat + zc * zc) + 0.5 * a->index_t0 / a->f_s;
            }
            double acc_r = 0.0, acc_i = 0.0;
            for (int j = 0; j < nfv; ++j) {
                const int i = vidx[j];
                for (int k = 0; k < n_el; ++k) {
                    const double arg = om[i] * tsc[k];
                    er[k] = cos(arg); ei[k] = sin(arg);
                    wtx[k] = wtx_all[(long)j * n_el + k];
                    wrx[k] = wrx_all[(long)j * n_el + k];
                }
                /* das_sa.m:485: sum over tx (3rd dim) */
                for (int rx = 0; rx < n_el; ++rx) {
                    double sr_ = 0.0, si_ = 0.0;
                    for (int tx = 0; tx < n_el; ++tx) {
                        const long ch = (long)tx * n_el + rx;
                        const double xr = 2 * Xr[ch * nf + i], xi = 2 * Xi[ch * nf + i];
                        const double pr = wtx[tx] * er[tx], pi_ = wtx[tx] * ei[tx];
                        sr_ += pr * xr - pi_ * xi;
                        si_ += pr * xi + pi_ * xr;
                    }
                    ftr[rx] = sr_; fti[rx] = si_;
                }
                /* das_sa.m:486: sum over rx (2nd dim) */
                double fr = 0.0, fi = 0.0;
                for (int rx = 0; rx < n_el; ++rx) {
                    const double pr = wrx[rx] * er[rx], pi_ = wrx[rx] * ei[rx];
                    fr += pr * ftr[rx] - pi_ * fti[rx];
                    fi += pr * fti[rx] + pi_ * ftr[rx];
                }
                acc_r += fr;
                acc_i += fi;
                if (v == v_sig) { sig_fr[i] = fr; sig_fi[i] = fi; }
            }
            a->image_re[v] = acc_r;
            a->image_im[v] = acc_i;
        }
        free(tsc); free(wtx); free(wrx); free(ftr); free(fti); free(er); free(ei);
        free(wtx_all); free(wrx_all); free(vidx);
    }
    if (sig_fr) {
        signal_idft(sig_fr, sig_fi, nf, nlb, N, a->signal_re, a->signal_im);
        free(sig_fr);
        free(sig_fi);
    }
    free(fax); free(om); free(r); free(Ftx); free(Frx); free(Xr); free(Xi);
    return 0;
}
