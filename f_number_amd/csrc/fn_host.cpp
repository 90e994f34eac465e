// fn_host.cpp -- host-side configuration of the beamformer: status codes,
// argument checks, frequency grid, F-number family and window samples.
// No GPU code here; everything is usable without a device.
#include "fn_internal.h"

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace fnh {

static thread_local char g_last_error[512] = "";

int fail(int status, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error() { g_last_error[0] = '\0'; }

// ---------------------------------------------------------------------------
// frequency grid: das_pw.m:343 (N_samples_t_pad), :353 (odd N_points_dft),
// :356-357 (band indices, here 0-based); gpu_bf_das_pw_rf.cu:525-552.
// ---------------------------------------------------------------------------
int frequency_grid(int n_t, int n_el, double pitch, double z_pad, double f_s, double c_0, double f_lb,
                   double f_ub, Grid *g)
{
    // gpu_bf_das_pw_rf.cu:301-303 precede the index check of :548-551
    if (!(f_lb > 0)) return fail(FN_ERR_FREQUENCY_BOUND, "Lower frequency bound must be positive!");
    if (!(f_ub > f_lb)) return fail(FN_ERR_FREQUENCY_BOUND, "Upper frequency bound must exceed the lower frequency bound!");
    if (!(f_ub <= f_s / 2)) return fail(FN_ERR_FREQUENCY_BOUND, "Upper frequency bound must not exceed the Nyquist frequency!");
    if (n_t < 1 || n_el < 1) return fail(FN_ERR_RF_DATA, "RF data must not be empty");
    const double M = (n_el - 1) / 2.0;
    const double span = ((n_el - 1) - M) * pitch - (0 - M) * pitch;
    const double pad = std::ceil((std::sqrt(span * span + z_pad * z_pad) - z_pad) * f_s / c_0);
    if (!(pad >= 0 && pad < 1e8)) return fail(FN_ERR_AXIAL_POSITIONS, "invalid zero padding");
    const int n_pad = (int)pad;
    const long N = (long)n_t + n_pad + 1 - (((long)n_t + n_pad) % 2);
    if (N > (1L << 24)) return fail(FN_ERR_RF_DATA, "DFT order %ld too large", N);
    g->N = (int)N;
    const double lb = std::ceil(f_lb * N / f_s);
    const double ub = std::floor(f_ub * N / f_s);
    // gpu_bf_das_pw_rf.cu:548-551: 0 < n_lb <= n_ub < N/2+1
    if (!(lb > 0 && lb <= ub && ub < (double)(N / 2 + 1)))
        return fail(FN_ERR_FREQUENCY_BOUNDS, "Frequency bounds lead to invalid indices!");
    g->n_lb = (int)lb;
    g->n_ub = (int)ub;
    g->n_f = g->n_ub - g->n_lb + 1;
    return FN_OK;
}

// ---------------------------------------------------------------------------
// F-numbers
// ---------------------------------------------------------------------------
static double rad(double deg) { return (M_PI / 180.0) * deg; }

// +f_numbers/+distance/constant.m:90-121 (properties set at :72-74): the
// F-number that keeps the first grating lobe a fixed angle delta away from the
// main-lobe edge, with the transition branch below r = 1/(1+cos delta).
static double distance_constant(double delta_deg, double F_ub, double r)
{
    const double s = std::sin(rad(delta_deg));
    const double o = 1 + std::cos(rad(delta_deg));
    const bool ge_lb = r >= 1 / o;
    const bool lt_ub = r < 1 / s;
    double v = 0.0;
    if (!ge_lb && r >= 0.5) {
        const double u = 1 / r - 1;
        v = std::sqrt(1 / (u * u) - 1) / 2;
    }
    if (ge_lb && lt_ub) {
        const double r2 = r * r;
        v = (std::sqrt(2 * o * r2 - 1) + o * s * r2) / (2 - 2 * (s * s) * r2);
        v = std::fmin(v, F_ub);
    }
    if (!lt_ub) v = F_ub;
    return v;
}

// +f_numbers/+grating/angle_lb.m:44-75 with +grating/grating.m:72-75: the
// smallest F-number whose first grating lobe stays beyond chi, capped at F_ub
// once impossible, then max'ed with distance.constant(delta, inf).
static double grating_angle_lb(double chi_deg, double F_ub, double delta_deg, double r)
{
    const double thr = std::sin(rad(chi_deg));
    const double lower = 1 / (1 + thr);
    const double two_f = 2 * F_ub;
    const double upper = 1 / (1 / std::sqrt(1 + two_f * two_f) + thr);
    double v = 0.0;
    if (r > lower && r < upper) {
        const double u = 1 / r - thr;
        v = std::sqrt(1 / (u * u) - 1) / 2;
    }
    if (r >= upper) v = F_ub;
    if (delta_deg >= 0) v = std::fmax(v, distance_constant(delta_deg, INFINITY, r));
    return v;
}

static double sinc_pi(double x) { return x == 0.0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x); }

// +f_numbers/+directivity/directivity.m:91-120: alpha minimises
// |O(theta) sinc(w sin theta) - 10^(-att/20)| on [0, pi/2] (fminbnd, TolX 1e-4),
// F = 1 / (2 tan alpha).  The minimiser is Brent's golden-section/parabolic
// method (R. P. Brent, Algorithms for Minimization without Derivatives, 1973,
// ch. 5), the method fminbnd documents.
static double directivity(double width_over_pitch, bool soft, double att_db, double r)
{
    const double w = width_over_pitch * r;
    const double thresh = std::pow(10.0, -att_db / 20);
    auto f = [&](double th) {
        const double obl = soft ? std::cos(th) : 1.0;  // +obliquity/soft.m, rigid.m
        return std::fabs(obl * sinc_pi(w * std::sin(th)) - thresh);
    };
    double a = 0.0, b = M_PI / 2;
    const double tol = 1e-4;
    const double golden = 0.5 * (3.0 - std::sqrt(5.0));
    const double sqrt_eps = std::sqrt(DBL_EPSILON);
    double x = a + golden * (b - a), wv = x, v = x;
    double fx = f(x), fw = fx, fv = fx;
    double d = 0.0, e = 0.0;
    for (int it = 0; it < 500; ++it) {
        const double m = 0.5 * (a + b);
        const double tol1 = sqrt_eps * std::fabs(x) + tol / 3.0;
        const double tol2 = 2.0 * tol1;
        if (std::fabs(x - m) <= tol2 - 0.5 * (b - a)) break;
        bool take_golden = true;
        if (std::fabs(e) > tol1) {
            double rr = (x - wv) * (fx - fv);
            double q = (x - v) * (fx - fw);
            double p = (x - v) * q - (x - wv) * rr;
            q = 2.0 * (q - rr);
            if (q > 0.0) p = -p;
            q = std::fabs(q);
            const double e_old = e;
            e = d;
            const bool ok = std::fabs(p) < std::fabs(0.5 * q * e_old) && p > q * (a - x) && p < q * (b - x);
            if (ok) {
                d = p / q;
                const double u = x + d;
                if ((u - a) < tol2 || (b - u) < tol2) d = (m >= x) ? tol1 : -tol1;
                take_golden = false;
            }
        }
        if (take_golden) {
            e = (x >= m) ? a - x : b - x;
            d = golden * e;
        }
        const double u = (std::fabs(d) >= tol1) ? x + d : x + (d >= 0 ? tol1 : -tol1);
        const double fu = f(u);
        if (fu <= fx) {
            if (u >= x) a = x; else b = x;
            v = wv; fv = fw; wv = x; fw = fx; x = u; fx = fu;
        } else {
            if (u < x) a = u; else b = u;
            if (fu <= fw || wv == x) { v = wv; fv = fw; wv = u; fw = fu; }
            else if (fu <= fv || v == x || v == wv) { v = u; fv = fu; }
        }
    }
    return 1 / (2 * std::tan(x));
}

int f_number_values(const fn_f_number *F, const double *r, int n, double *out)
{
    if (!F || (!r && n > 0) || (!out && n > 0) || n < 0) return fail(FN_ERR_ARGUMENT, "null argument");
    for (int i = 0; i < n; ++i) {
        // f_number.m:100 mustBePositive(element_pitch_over_lambda)
        if (!(r[i] > 0)) return fail(FN_ERR_F_NUMBER, "element_pitch_over_lambda must be positive");
        double v;
        switch (F->kind) {
        case FN_F_NUMBER_CONSTANT: v = F->p0; break;                                   // constant.m:72
        case FN_F_NUMBER_GRATING_ANGLE_LB: v = grating_angle_lb(F->p0, F->p1, F->p2, r[i]); break;
        case FN_F_NUMBER_DISTANCE_CONSTANT: v = distance_constant(F->p0, F->p1, r[i]); break;
        case FN_F_NUMBER_SZABO: v = F->p0 * r[i] / 1.2; break;                         // szabo.m:83-86
        case FN_F_NUMBER_DIRECTIVITY_SOFT: v = directivity(F->p0, true, F->p1, r[i]); break;
        case FN_F_NUMBER_DIRECTIVITY_RIGID: v = directivity(F->p0, false, F->p1, r[i]); break;
        case FN_F_NUMBER_PERROT: v = directivity(F->p0, true, 3.0, r[i]); break;       // perrot.m: soft, 3 dB
        case FN_F_NUMBER_DISTANCE_NONE: v = distance_constant(0.0, INFINITY, r[i]); break;  // none.m
        default: return fail(FN_ERR_F_NUMBER, "unknown F-number kind %d", (int)F->kind);
        }
        out[i] = v;
    }
    return FN_OK;
}

// ---------------------------------------------------------------------------
// windows (samples of |t|)
// ---------------------------------------------------------------------------
double window_sample(int id, double param, double t, int semantics)
{
    const double a = std::fabs(t);
    if (semantics == FN_SEMANTICS_MEX) {
        // gpu_bf_windows.cu:15-24: every window is zero for |l| >= L/2
        if (a >= 1.0) return 0.0;
        switch (id) {
        case FN_WINDOW_BOXCAR: return 1.0;
        case FN_WINDOW_HANN: return 0.5 * (1.0 + std::cos(M_PI * a));
        case FN_WINDOW_TUKEY: return a <= 1.0 - param ? 1.0 : 0.5 * (1.0 + std::cos(M_PI * (a - (1.0 - param)) / param));
        case FN_WINDOW_COSINE: return std::cos(0.5 * M_PI * a) * (M_PI / 4.0);
        case FN_WINDOW_TRIANGULAR: return 1.0 - a;
        default: return NAN;
        }
    }
    switch (id) {
    case FN_WINDOW_BOXCAR: return a < 1.0 ? 1.0 : 0.0;                 // boxcar.m:62
    case FN_WINDOW_HANN: return 0.5 * (1.0 + std::cos(M_PI * a));     // hann.m:62 (no truncation)
    case FN_WINDOW_TUKEY:                                             // tukey.m:101-105
        if (!(a < 1.0)) return 0.0;
        if (a > 1.0 - param) return 0.5 * (1.0 + std::cos(M_PI * ((a - (1.0 - param)) / param)));
        return 1.0;
    case FN_WINDOW_TRIANGULAR: return 1.0 - a;                        // triangular.m:62 (no truncation)
    default: return NAN;
    }
}

int check_window(const fn_window &w, int status)
{
    // gpu_bf_das_pw_rf.cu:341-349: known id, parameter in [0, 1]
    if (!(w.id >= 0 && w.id <= FN_WINDOW_TRIANGULAR)) return fail(status, "Unknown window function!");
    if (!(w.parameter >= 0.0 && w.parameter <= 1.0)) return fail(status, "Invalid window parameter!");
    if (w.id == FN_WINDOW_TUKEY && !(w.parameter > 0.0 && w.parameter < 1.0))
        return fail(status, "Tukey cosine fraction must lie in (0, 1) (tukey.m:46-57)");
    return FN_OK;
}

// ---------------------------------------------------------------------------
// argument checks: gpu_bf_das_pw_rf.cu:127-485 (PW), gpu_bf_saft_rf.cu:124-592 (SA)
// ---------------------------------------------------------------------------
static int check_common(const double *px, int nx, const double *pz, int nz, int n_t, int n_el, double f_s,
                        double width, double pitch, double c_0, double f_lb, double f_ub, double t0,
                        int semantics, int precision, int scale, int index_gpu, int verbosity)
{
    if (!px || nx < 1) return fail(FN_ERR_LATERAL_POSITIONS, "Lateral positions must be a real-valued row vector!");
    for (int i = 0; i < nx; ++i)
        if (!std::isfinite(px[i])) return fail(FN_ERR_LATERAL_POSITIONS, "positions_x must contain finite x-coordinates!");
    if (!pz || nz < 1) return fail(FN_ERR_AXIAL_POSITIONS, "Axial positions must be a real-valued row vector!");
    for (int i = 0; i < nz; ++i)
        if (!(pz[i] > 0 && std::isfinite(pz[i])))
            return fail(FN_ERR_AXIAL_POSITIONS, "positions_z must contain positive and finite z-coordinates!");
    if (n_t < 2 || n_el < 2) return fail(FN_ERR_RF_DATA, "data_RF must contain at least two temporal and two spatial samples!");
    if (!(f_s > 0 && std::isfinite(f_s))) return fail(FN_ERR_SAMPLING_RATE, "Sampling rate must be positive!");
    if (!(width > 0 && std::isfinite(width))) return fail(FN_ERR_ELEMENT_WIDTH, "Element width must be positive and finite!");
    if (!(pitch > width && std::isfinite(pitch)))
        return fail(FN_ERR_ELEMENT_PITCH, "Element pitch must be finite and larger than element width!");
    if (!(c_0 > 0 && std::isfinite(c_0))) return fail(FN_ERR_SOUND_SPEED, "Average speed of sound must be positive and finite!");
    if (!(f_lb > 0)) return fail(FN_ERR_FREQUENCY_BOUND, "Lower frequency bound must be positive!");
    if (!(f_ub > f_lb)) return fail(FN_ERR_FREQUENCY_BOUND, "Upper frequency bound must exceed the lower frequency bound!");
    if (!(f_ub <= f_s / 2)) return fail(FN_ERR_FREQUENCY_BOUND, "Upper frequency bound must not exceed the Nyquist frequency!");
    if (!(t0 >= 0 && std::isfinite(t0))) return fail(FN_ERR_TIME_SHIFT, "Time index must be nonnegative!");
    if (semantics != FN_SEMANTICS_CPU && semantics != FN_SEMANTICS_MEX) return fail(FN_ERR_ARGUMENT, "unknown semantics %d", semantics);
    if (precision < FN_PRECISION_FP32 || precision > FN_PRECISION_FP16) return fail(FN_ERR_ARGUMENT, "unknown precision %d", precision);
    if (scale != FN_SCALE_MEX && scale != FN_SCALE_CPU) return fail(FN_ERR_ARGUMENT, "unknown output scale %d", scale);
    if (index_gpu < 0) return fail(FN_ERR_DEVICE, "Selected device does not exist!");
    if (verbosity < 0 || verbosity >= 4) return fail(FN_ERR_VERBOSITY, "Selected verbosity level does not exist!");
    return FN_OK;
}

static int check_f_values(const double *F, int n, int status)
{
    if (!F) return fail(status, "F-number values missing");
    // gpu_bf_das_pw_rf.cu:578-586: 0 <= F <= DBL_MAX
    for (int i = 0; i < n; ++i)
        if (!(F[i] >= 0 && F[i] <= DBL_MAX)) return fail(status, "Values of F-number must be nonnegative!");
    return FN_OK;
}

static int check_norm(int norm, const fn_window &w)
{
    if (norm < FN_NORMALIZATION_OFF || norm > FN_NORMALIZATION_WINDOWED)
        return fail(FN_ERR_NORMALIZATION, "Normalization must be off, on or windowed");
    if (norm == FN_NORMALIZATION_WINDOWED) return check_window(w, FN_ERR_FREQUENCY_WINDOW);
    return FN_OK;
}

double z_pad_of(const double *pz, int nz, int semantics)
{
    // quirk: the MEX pads with positions_z[1] (gpu_bf_das_pw_rf.cu:525), the
    // CPU path with positions_z(1) (das_pw.m:343); a single depth falls back
    // to the first entry instead of reading past the end.
    return (semantics == FN_SEMANTICS_MEX && nz > 1) ? pz[1] : pz[0];
}

int validate_pw(const fn_pw_params *p, Grid *g)
{
    clear_error();
    if (!p || p->struct_size != sizeof(fn_pw_params)) return fail(FN_ERR_ARGUMENT, "fn_pw_params size mismatch");
    int s = check_common(p->positions_x, p->n_x, p->positions_z, p->n_z, p->n_samples, p->n_elements, p->f_s,
                         p->element_width, p->element_pitch, p->c_0, p->f_lb, p->f_ub, p->index_t0, p->semantics,
                         p->precision, p->output_scale, p->index_gpu, p->verbosity);
    if (s) return s;
    if (p->n_angles < 1 || (!p->steering_angles && !p->virtual_sources))
        return fail(FN_ERR_STEERING_ANGLE, "Steering angle missing");
    if (p->n_frames < 0 || (long)std::max(1, p->n_frames) * p->n_angles > (1L << 24))
        return fail(FN_ERR_ARGUMENT, "n_frames must be in 0..2^24 / n_angles (got %d)", p->n_frames);
    if (p->virtual_sources) {
        for (int a = 0; a < p->n_angles; ++a) {
            const double xv = p->virtual_sources[2 * a], zv = p->virtual_sources[2 * a + 1];
            if (!(std::isfinite(xv) && std::isfinite(zv) && zv != 0.0))
                return fail(FN_ERR_STEERING_ANGLE, "virtual source %d must be finite with z_vs != 0", a);
        }
    } else {
        for (int a = 0; a < p->n_angles; ++a) {
            const double th = p->steering_angles[a];
            if (!(th > -M_PI / 2 && th < M_PI / 2))
                return fail(FN_ERR_STEERING_ANGLE, "Steering angle must be larger than negative pi over two and smaller than pi over two!");
        }
    }
    if ((s = check_window(p->window, FN_ERR_WINDOW))) return s;
    if (p->semantics == FN_SEMANTICS_CPU && p->window.id == FN_WINDOW_COSINE)
        return fail(FN_ERR_WINDOW, "the cosine window has no MATLAB class (CPU semantics)");
    if ((s = check_norm(p->normalization, p->normalization_window))) return s;
    if (p->line_scan) {
        if (p->line_scan != 1) return fail(FN_ERR_ARGUMENT, "line_scan must be 0 or 1 (got %d)", p->line_scan);
        if (p->n_angles != p->n_x)
            return fail(FN_ERR_ARGUMENT, "line scan: one transmit per grid column (n_angles %d, n_x %d)", p->n_angles, p->n_x);
        if (p->n_x > 65535) return fail(FN_ERR_UNSUPPORTED, "line scan: at most 65535 lines per plan");
        if (p->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "line scan: one frame per plan");
        if (p->precision != FN_PRECISION_FP32) return fail(FN_ERR_UNSUPPORTED, "line scan: FP32 only");
        if (p->normalization == FN_NORMALIZATION_WINDOWED && p->semantics == FN_SEMANTICS_CPU)
            return fail(FN_ERR_UNSUPPORTED, "line scan: normalizations.windowed (CPU semantics) is not supported");
    }
    Grid gg;
    if ((s = frequency_grid(p->n_samples, p->n_elements, p->element_pitch, z_pad_of(p->positions_z, p->n_z, p->semantics),
                            p->f_s, p->c_0, p->f_lb, p->f_ub, &gg)))
        return s;
    if (p->n_f != gg.n_f) return fail(FN_ERR_F_NUMBER, "expected %d F-number values (band %d..%d), got %d", gg.n_f, gg.n_lb, gg.n_ub, p->n_f);
    if ((s = check_f_values(p->f_number_values, p->n_f, FN_ERR_F_NUMBER))) return s;
    if (g) *g = gg;
    return FN_OK;
}

int validate_sa(const fn_sa_params *p, Grid *g)
{
    clear_error();
    if (!p || p->struct_size != sizeof(fn_sa_params)) return fail(FN_ERR_ARGUMENT, "fn_sa_params size mismatch");
    int s = check_common(p->positions_x, p->n_x, p->positions_z, p->n_z, p->n_samples, p->n_elements, p->f_s,
                         p->element_width, p->element_pitch, p->c_0, p->f_lb, p->f_ub, p->index_t0, p->semantics,
                         p->precision, p->output_scale, p->index_gpu, p->verbosity);
    if (s) return s;
    if ((s = check_window(p->window_tx, FN_ERR_WINDOW))) return s;
    if ((s = check_window(p->window_rx, FN_ERR_WINDOW))) return s;
    if (p->semantics == FN_SEMANTICS_CPU && (p->window_tx.id == FN_WINDOW_COSINE || p->window_rx.id == FN_WINDOW_COSINE))
        return fail(FN_ERR_WINDOW, "the cosine window has no MATLAB class (CPU semantics)");
    if ((s = check_norm(p->normalization_tx, p->normalization_window_tx))) return s;
    if ((s = check_norm(p->normalization_rx, p->normalization_window_rx))) return s;
    Grid gg;
    if ((s = frequency_grid(p->n_samples, p->n_elements, p->element_pitch, z_pad_of(p->positions_z, p->n_z, p->semantics),
                            p->f_s, p->c_0, p->f_lb, p->f_ub, &gg)))
        return s;
    if (p->n_f != gg.n_f) return fail(FN_ERR_F_NUMBER, "expected %d F-number values (band %d..%d), got %d", gg.n_f, gg.n_lb, gg.n_ub, p->n_f);
    if ((s = check_f_values(p->f_number_tx_values, p->n_f, FN_ERR_TX_F_NUMBER))) return s;
    if ((s = check_f_values(p->f_number_rx_values, p->n_f, FN_ERR_RX_F_NUMBER))) return s;
    if (g) *g = gg;
    return FN_OK;
}

}  // namespace fnh

// ---------------------------------------------------------------------------
// C ABI (host part)
// ---------------------------------------------------------------------------
using namespace fnh;

extern "C" {

int32_t fn_abi_version(void) { return FN_ABI_VERSION; }

const char *fn_status_id(int32_t s)
{
    static const char *ids[] = {"OK", "nrhs", "nlhs", "NoCUDADevices", "InvalidLateralPositions",
                                "InvalidAxialPositions", "InvalidRFData", "InvalidSamplingRate",
                                "InvalidSteeringAngle", "InvalidElementWidth", "InvalidElementPitch",
                                "InvalidSoundSpeed", "InvalidFrequencyBounds", "InvalidFrequencyBound",
                                "InvalidTimeShift", "InvalidWindow", "InvalidFNumber", "InvalidNormalization",
                                "InvalidFrequencyWindow", "InvalidCUDADevice", "InvalidVerbosityLevel",
                                "InvalidTXFNumber", "InvalidRXFNumber", "CUDAError", "OutOfMemory",
                                "InvalidArgument", "Unsupported"};
    if (s < 0 || s >= (int)(sizeof(ids) / sizeof(ids[0]))) return "Unknown";
    return ids[s];
}

const char *fn_last_error(void) { return g_last_error; }

int32_t fn_pw_frequency_grid(const fn_pw_params *p, int32_t *N_dft, int32_t *n_lb, int32_t *n_ub)
{
    if (!p || !N_dft || !n_lb || !n_ub) return fail(FN_ERR_ARGUMENT, "null argument");
    if (p->struct_size != sizeof(fn_pw_params)) return fail(FN_ERR_ARGUMENT, "fn_pw_params size mismatch");
    if (!p->positions_z || p->n_z < 1) return fail(FN_ERR_AXIAL_POSITIONS, "Axial positions missing");
    Grid g;
    int s = frequency_grid(p->n_samples, p->n_elements, p->element_pitch, z_pad_of(p->positions_z, p->n_z, p->semantics),
                           p->f_s, p->c_0, p->f_lb, p->f_ub, &g);
    if (s) return s;
    *N_dft = g.N; *n_lb = g.n_lb; *n_ub = g.n_ub;
    return FN_OK;
}

int32_t fn_sa_frequency_grid(const fn_sa_params *p, int32_t *N_dft, int32_t *n_lb, int32_t *n_ub)
{
    if (!p || !N_dft || !n_lb || !n_ub) return fail(FN_ERR_ARGUMENT, "null argument");
    if (p->struct_size != sizeof(fn_sa_params)) return fail(FN_ERR_ARGUMENT, "fn_sa_params size mismatch");
    if (!p->positions_z || p->n_z < 1) return fail(FN_ERR_AXIAL_POSITIONS, "Axial positions missing");
    Grid g;
    int s = frequency_grid(p->n_samples, p->n_elements, p->element_pitch, z_pad_of(p->positions_z, p->n_z, p->semantics),
                           p->f_s, p->c_0, p->f_lb, p->f_ub, &g);
    if (s) return s;
    *N_dft = g.N; *n_lb = g.n_lb; *n_ub = g.n_ub;
    return FN_OK;
}

int32_t fn_element_pitch_over_lambda(double pitch, double f_s, double c_0, int32_t N, int32_t n_lb, int32_t n_ub,
                                     double *out)
{
    if (!out || N < 1 || n_ub < n_lb) return fail(FN_ERR_ARGUMENT, "invalid band");
    // das_pw.m:361,368: p * (f_s * n / N) / c_0
    for (int n = n_lb; n <= n_ub; ++n) out[n - n_lb] = pitch * (f_s * (double)n / N) / c_0;
    return FN_OK;
}

int32_t fn_f_number_values(const fn_f_number *F, const double *r, int32_t n, double *out)
{
    clear_error();
    return f_number_values(F, r, n, out);
}

int32_t fn_window_samples(const fn_window *w, int32_t semantics, const double *t, int32_t n, double *out)
{
    clear_error();
    if (!w || (!t && n > 0) || (!out && n > 0) || n < 0) return fail(FN_ERR_ARGUMENT, "null argument");
    int s = check_window(*w, FN_ERR_WINDOW);
    if (s) return s;
    if (semantics != FN_SEMANTICS_CPU && semantics != FN_SEMANTICS_MEX) return fail(FN_ERR_ARGUMENT, "unknown semantics");
    if (semantics == FN_SEMANTICS_CPU && w->id == FN_WINDOW_COSINE) return fail(FN_ERR_WINDOW, "no MATLAB cosine window");
    for (int i = 0; i < n; ++i) out[i] = window_sample(w->id, w->parameter, t[i], semantics);
    return FN_OK;
}

int32_t fn_pw_validate(const fn_pw_params *p) { return validate_pw(p, nullptr); }
int32_t fn_sa_validate(const fn_sa_params *p) { return validate_sa(p, nullptr); }

}  // extern "C"
