// fn_synth.hip -- seeded synthetic RF acquisitions (the reference ships no
// data: data_RF.mat is missing, .MISSING_LARGE_BLOBS:1).
//
// Medium: n_points point scatterers of amplitude point_amplitude plus
// n_speckle diffuse scatterers with N(0,1) amplitudes, all uniform in
// [x_min, x_max] x [z_min, z_max].  Echo: Gaussian-modulated pulse
// exp(-t^2 / (2 sigma^2)) cos(2 pi f_c t) whose -6 dB fractional bandwidth is
// fractional_bandwidth, placed at the exact transmit + receive delay:
//   plane wave a:  tau = (e_x (x_s - x_ref) + e_z z_s) / c + |r_k - r_s| / c   (das_pw.m:336,427)
//   element i -> k: tau = |r_i - r_s| / c + |r_k - r_s| / c                     (das_sa.m:431)
//   virtual source (x_v, z_v): diverging (z_v < 0) (|r_s - r_v| - |r_v|) / c,
//                  focused (z_v > 0) (|r_v| + sign(z_s - z_v) |r_s - r_v|) / c  (+ receive)
// Layered medium (straight rays): a path of length L from the surface to depth
// z_s spends the share f = min(1, h / z_s) of its length in the layer, so its
// time is L / c_0 + f L (1 / c_layer - 1 / c_0); the plane-wave transmit path to
// depth z_s crosses the layer over h / e_z.  Receive delay offsets add per element.
// Flow: scatterers inside the tube move along its axis (parabolic profile).
// White Gaussian noise at snr_db relative to the RMS of the clean signal.
// Random numbers: splitmix64 (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), stateless per noise sample.
#include "fn_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double u01(uint64_t z) { return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// N(0,1) sample for noise index i (Box-Muller, cosine branch)
__host__ __device__ inline double gauss_at(uint64_t seed, uint64_t i)
{
    const uint64_t base = mix64(seed * 0xD1B54A32D192ED03ull) + 2 * i;
    const double u1 = u01(mix64(base)), u2 = u01(mix64(base + 1));
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

struct Geo {
    int n_t, n_el, n_tx, mode, n_s;
    double f_s, pitch, c0, M, fc, sigma, half;  // half: pulse half-support (s)
    double h, slow;                              // layer thickness, 1/c_layer - 1/c_0 (0: homogeneous)
    const double *delay;                         // per-element receive offsets (device or host), may be null
};

// time along a straight path of length L from the surface to depth zs
__host__ __device__ inline double path_time(const Geo &G, double L, double zs)
{
    double t = L / G.c0;
    if (G.h > 0.0) t += L * fmin(1.0, G.h / zs) * G.slow;
    return t;
}

__host__ __device__ inline double tx_delay(const Geo &G, const double *txp, int a, double xs, double zs)
{
    if (G.mode == 0) {
        const double ex = txp[3 * a], ez = txp[3 * a + 1], xr = txp[3 * a + 2];
        double t = (ex * (xs - xr) + ez * zs) / G.c0;
        if (G.h > 0.0) t += fmin(zs, G.h) / ez * G.slow;
        return t;
    }
    if (G.mode == 2) {  // virtual source (x_v, z_v, |r_v|), homogeneous medium
        const double xv = txp[3 * a], zv = txp[3 * a + 1], r0 = txp[3 * a + 2];
        const double r = sqrt((xs - xv) * (xs - xv) + (zs - zv) * (zs - zv));
        return (zv < 0.0 ? r - r0 : r0 + (zs >= zv ? r : -r)) / G.c0;
    }
    const double dx = ((double)a - G.M) * G.pitch - xs;
    return path_time(G, sqrt(dx * dx + zs * zs), zs);
}

__host__ __device__ inline double rx_delay(const Geo &G, int k, double xs, double zs)
{
    const double dx = ((double)k - G.M) * G.pitch - xs;
    return path_time(G, sqrt(dx * dx + zs * zs), zs) + (G.delay ? G.delay[k] : 0.0);
}

__host__ __device__ inline double pulse(const Geo &G, double t)
{
    return exp(-t * t / (2.0 * G.sigma * G.sigma)) * cos(2.0 * M_PI * G.fc * t);
}

// one block per channel; samples t in [t_lo, t_lo + 256 * SPT) split in contiguous
// runs of SPT per thread; scatterers streamed through LDS in their index order
constexpr int SPT = 16;

__global__ __launch_bounds__(256) void synth_signal_kernel(Geo G, const double *__restrict__ xs,
                                                           const double *__restrict__ zs,
                                                           const double *__restrict__ amp,
                                                           const double *__restrict__ txp, float *__restrict__ rf)
{
    __shared__ double tau_s[256];
    __shared__ double amp_s[256];
    const int ch = blockIdx.x;
    const int k = ch % G.n_el, a = ch / G.n_el;
    for (int t_lo = 0; t_lo < G.n_t; t_lo += 256 * SPT) {
        const int t0 = t_lo + threadIdx.x * SPT;
        double acc[SPT];
#pragma unroll
        for (int i = 0; i < SPT; ++i) acc[i] = 0.0;
        const double tmin = (double)t0 / G.f_s - G.half, tmax = (double)(t0 + SPT - 1) / G.f_s + G.half;
        for (int s0 = 0; s0 < G.n_s; s0 += 256) {
            __syncthreads();
            const int s = s0 + threadIdx.x;
            if (s < G.n_s) {
                tau_s[threadIdx.x] = tx_delay(G, txp, a, xs[s], zs[s]) + rx_delay(G, k, xs[s], zs[s]);
                amp_s[threadIdx.x] = amp[s];
            }
            __syncthreads();
            const int ns = min(256, G.n_s - s0);
            for (int j = 0; j < ns; ++j) {
                const double tau = tau_s[j];
                if (tau < tmin || tau > tmax) continue;
#pragma unroll
                for (int i = 0; i < SPT; ++i) {
                    const double dt = (double)(t0 + i) / G.f_s - tau;
                    if (fabs(dt) <= G.half) acc[i] += amp_s[j] * pulse(G, dt);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            if (t0 + i < G.n_t) rf[(size_t)ch * G.n_t + t0 + i] = (float)acc[i];
    }
}

__global__ void sumsq_kernel(const float *__restrict__ x, long n, double *__restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) s += (double)x[i] * x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void noise_kernel(float *__restrict__ x, long n, uint64_t seed, double sigma)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = (float)((double)x[i] + sigma * gauss_at(seed, (uint64_t)i));
}

int build_geo(const fn_synth_params *p, Geo *G, std::vector<double> *txp)
{
    using fnh::fail;
    if (!p || p->struct_size != sizeof(fn_synth_params)) return fail(FN_ERR_ARGUMENT, "fn_synth_params size mismatch");
    if (p->n_samples < 2 || p->n_elements < 2) return fail(FN_ERR_RF_DATA, "invalid RF size");
    if (!(p->f_s > 0 && p->c_0 > 0 && p->element_pitch > 0 && p->f_c > 0 && p->fractional_bandwidth > 0))
        return fail(FN_ERR_ARGUMENT, "invalid acquisition parameters");
    if (p->n_points < 0 || p->n_speckle < 0) return fail(FN_ERR_ARGUMENT, "negative scatterer count");
    G->n_t = p->n_samples;
    G->n_el = p->n_elements;
    G->mode = p->mode;
    G->n_s = p->n_points + p->n_speckle;
    G->f_s = p->f_s;
    G->pitch = p->element_pitch;
    G->c0 = p->c_0;
    G->M = (p->n_elements - 1) / 2.0;
    G->h = 0.0;
    G->slow = 0.0;
    G->delay = nullptr;
    if (p->layer_thickness > 0.0) {
        if (!(p->layer_c > 0.0)) return fail(FN_ERR_ARGUMENT, "layer_c must be positive");
        G->h = p->layer_thickness;
        G->slow = 1.0 / p->layer_c - 1.0 / p->c_0;
    }
    if (p->tube_radius < 0.0) return fail(FN_ERR_ARGUMENT, "tube_radius must be >= 0");
    G->fc = p->f_c;
    // -6 dB half-bandwidth bw f_c / 2 = sigma_f sqrt(2 ln 2), sigma = 1 / (2 pi sigma_f)
    const double sigma_f = p->fractional_bandwidth * p->f_c / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    G->sigma = 1.0 / (2.0 * M_PI * sigma_f);
    G->half = 6.0 * G->sigma;
    if (p->mode == 0) {
        if (p->n_angles < 1 || !p->steering_angles) return fail(FN_ERR_STEERING_ANGLE, "steering angles missing");
        G->n_tx = p->n_angles;
        txp->resize(3 * (size_t)p->n_angles);
        for (int a = 0; a < p->n_angles; ++a) {
            const double th = p->steering_angles[a];
            const double ex = -std::sin(th), ez = std::cos(th);
            const double sg = (ex > 0) - (ex < 0);
            (*txp)[3 * a] = ex;
            (*txp)[3 * a + 1] = ez;
            (*txp)[3 * a + 2] = sg * ((0 - G->M) * p->element_pitch);
        }
    } else if (p->mode == 2) {
        if (p->n_angles < 1 || !p->virtual_sources) return fail(FN_ERR_STEERING_ANGLE, "virtual sources missing");
        if (p->layer_thickness > 0.0) return fail(FN_ERR_ARGUMENT, "virtual sources: homogeneous media only");
        G->n_tx = p->n_angles;
        txp->resize(3 * (size_t)p->n_angles);
        for (int a = 0; a < p->n_angles; ++a) {
            const double xv = p->virtual_sources[2 * a], zv = p->virtual_sources[2 * a + 1];
            if (!(std::isfinite(xv) && std::isfinite(zv) && zv != 0.0))
                return fail(FN_ERR_STEERING_ANGLE, "virtual source %d must be finite with z_vs != 0", a);
            (*txp)[3 * a] = xv;
            (*txp)[3 * a + 1] = zv;
            (*txp)[3 * a + 2] = std::sqrt(xv * xv + zv * zv);
        }
    } else if (p->mode == 1) {
        G->n_tx = p->n_elements;
        txp->assign(3, 0.0);
    } else {
        return fail(FN_ERR_ARGUMENT, "unknown synth mode %d", p->mode);
    }
    return FN_OK;
}

void scatterers(const fn_synth_params *p, double *xs, double *zs, double *amp)
{
    uint64_t st = p->seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return u01(z ^ (z >> 31));
    };
    const int n = p->n_points + p->n_speckle;
    for (int s = 0; s < n; ++s) {
        xs[s] = p->x_min + next() * (p->x_max - p->x_min);
        zs[s] = p->z_min + next() * (p->z_max - p->z_min);
        if (s < p->n_points) {
            amp[s] = p->point_amplitude;
        } else {
            const double u1 = next(), u2 = next();
            amp[s] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
        }
    }
    // flow: axial coordinate u along the tube, radial w; v(w) = v_max (1 - (w/R)^2),
    // u wrapped over [-W/2, W/2) with W the lateral extent of the field
    if (p->tube_radius > 0.0 && p->frame_time != 0.0) {
        const double ca = std::cos(p->tube_angle), sa = std::sin(p->tube_angle);
        const double W = p->x_max - p->x_min, R = p->tube_radius;
        for (int s = 0; s < n; ++s) {
            const double dx = xs[s] - p->tube_x, dz = zs[s] - p->tube_z;
            const double u = dx * ca + dz * sa, w = -dx * sa + dz * ca;
            if (std::fabs(w) >= R) continue;
            double u2 = u + p->tube_v_max * (1.0 - (w / R) * (w / R)) * p->frame_time;
            u2 -= W * std::floor(u2 / W + 0.5);
            xs[s] = p->tube_x + u2 * ca - w * sa;
            zs[s] = p->tube_z + u2 * sa + w * ca;
        }
    }
}

}  // namespace

extern "C" {

int32_t fn_synth_scatterers(const fn_synth_params *p, double *xs, double *zs, double *amps)
{
    fnh::clear_error();
    if (!p || !xs || !zs || !amps) return fnh::fail(FN_ERR_ARGUMENT, "null argument");
    scatterers(p, xs, zs, amps);
    return FN_OK;
}

int32_t fn_synth_rf(const fn_synth_params *p, float *rf)
{
    using fnh::fail;
    fnh::clear_error();
    if (!rf) return fail(FN_ERR_ARGUMENT, "null rf");
    Geo G;
    std::vector<double> txp;
    int s = build_geo(p, &G, &txp);
    if (s) return s;
    std::vector<double> xs(G.n_s), zs(G.n_s), amp(G.n_s);
    scatterers(p, xs.data(), zs.data(), amp.data());
    std::vector<double> delays;
    if (p->delay_offsets) delays.assign(p->delay_offsets, p->delay_offsets + G.n_el);
    const long n_ch = (long)G.n_el * G.n_tx;
    const long n = n_ch * G.n_t;
    const bool noisy = std::isfinite(p->snr_db);
    if (!p->on_device) {
        // host twin of synth_signal_kernel: same delays, same per-sample summation order
        if (!delays.empty()) G.delay = delays.data();
        std::vector<double> sig((size_t)n, 0.0);
        std::atomic<long> next_ch{0};
        auto worker = [&]() {
            for (long ch = next_ch++; ch < n_ch; ch = next_ch++) {
                const int k = (int)(ch % G.n_el), a = (int)(ch / G.n_el);
                double *row = sig.data() + ch * G.n_t;
                for (int si = 0; si < G.n_s; ++si) {
                    const double tau = tx_delay(G, txp.data(), a, xs[si], zs[si]) + rx_delay(G, k, xs[si], zs[si]);
                    const long lo = std::max(0L, (long)std::ceil((tau - G.half) * G.f_s) - 1);
                    const long hi = std::min((long)G.n_t - 1, (long)std::floor((tau + G.half) * G.f_s) + 1);
                    for (long t = lo; t <= hi; ++t) {
                        const double dt = (double)t / G.f_s - tau;
                        if (std::fabs(dt) <= G.half) row[t] += amp[si] * pulse(G, dt);
                    }
                }
            }
        };
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < nth; ++i) pool.emplace_back(worker);
        worker();
        for (auto &th : pool) th.join();
        double ss = 0.0;
        for (long i = 0; i < n; ++i) {
            const double v = (double)(float)sig[i];
            ss += v * v;
        }
        const double sigma = noisy ? std::sqrt(ss / n) * std::pow(10.0, -p->snr_db / 20.0) : 0.0;
        for (long i = 0; i < n; ++i) {
            double v = (double)(float)sig[i];
            if (noisy) v += sigma * gauss_at(p->seed + 1000, (uint64_t)i);
            rf[i] = noisy ? (float)v : (float)sig[i];
        }
        return FN_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return fail(FN_ERR_NO_DEVICES, "Could not find any HIP devices!");
    if (p->index_gpu < 0 || p->index_gpu >= count) return fail(FN_ERR_DEVICE, "Selected device does not exist!");
    if (hipSetDevice(p->index_gpu) != hipSuccess) return fail(FN_ERR_HIP, "hipSetDevice failed");
    double *d_x = nullptr, *d_z = nullptr, *d_a = nullptr, *d_t = nullptr, *d_part = nullptr, *d_dl = nullptr;
    const int n_red = 1024;
    hipError_t e = hipMalloc(&d_x, sizeof(double) * std::max(1, G.n_s));
    if (e == hipSuccess && !delays.empty()) {
        e = hipMalloc(&d_dl, sizeof(double) * delays.size());
        if (e == hipSuccess) e = hipMemcpy(d_dl, delays.data(), sizeof(double) * delays.size(), hipMemcpyHostToDevice);
        G.delay = d_dl;
    }
    if (e == hipSuccess) e = hipMalloc(&d_z, sizeof(double) * std::max(1, G.n_s));
    if (e == hipSuccess) e = hipMalloc(&d_a, sizeof(double) * std::max(1, G.n_s));
    if (e == hipSuccess) e = hipMalloc(&d_t, sizeof(double) * txp.size());
    if (e == hipSuccess) e = hipMalloc(&d_part, sizeof(double) * n_red);
    if (e == hipSuccess && G.n_s) e = hipMemcpy(d_x, xs.data(), sizeof(double) * G.n_s, hipMemcpyHostToDevice);
    if (e == hipSuccess && G.n_s) e = hipMemcpy(d_z, zs.data(), sizeof(double) * G.n_s, hipMemcpyHostToDevice);
    if (e == hipSuccess && G.n_s) e = hipMemcpy(d_a, amp.data(), sizeof(double) * G.n_s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_t, txp.data(), sizeof(double) * txp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(synth_signal_kernel, dim3((unsigned)n_ch), dim3(256), 0, 0, G, d_x, d_z, d_a, d_t, rf);
        e = hipGetLastError();
    }
    if (e == hipSuccess && noisy) {
        hipLaunchKernelGGL(sumsq_kernel, dim3(n_red), dim3(256), 0, 0, rf, n, d_part);
        e = hipGetLastError();
        std::vector<double> part(n_red);
        if (e == hipSuccess) e = hipMemcpy(part.data(), d_part, sizeof(double) * n_red, hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            double ss = 0.0;
            for (double v : part) ss += v;
            const double sigma = std::sqrt(ss / n) * std::pow(10.0, -p->snr_db / 20.0);
            hipLaunchKernelGGL(noise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, rf, n, p->seed + 1000, sigma);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    hipFree(d_x); hipFree(d_z); hipFree(d_a); hipFree(d_t); hipFree(d_part); hipFree(d_dl);
    if (e != hipSuccess) return fail(FN_ERR_HIP, "synthetic RF: %s", hipGetErrorString(e));
    return FN_OK;
}

}  // extern "C"
