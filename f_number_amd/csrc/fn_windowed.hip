// fn_windowed.hip -- row counts of the windowed normalization.
//
// normalizations.windowed (windowed.m:81-86) first normalizes the apodization
// of every valid frequency (row) to unit sum ("on", on.m:60-61) and then tapers
// every element column k over the valid rows i = 0, 1, ... by
//   window((i - L_k / 2) / (L_k / 2)),   L_k = #rows whose normalized weight is > 0.
// L_k depends on the geometry and the F-number tables only, so it is computed
// once per plan, in FP64 with the CPU path's formulas (das_pw.m:393-457,
// das_sa.m:383-480): the beamforming kernels then read h = L_k / 2 per (voxel,
// element) and count the valid rows of their voxel as they walk the bins.
// Plane waves: the rows are the bins with a non-empty receive aperture; the
// synthetic aperture: the bins where both apertures are non-empty.
#include "fn_kernels.hip.h"

#include <hip/hip_runtime.h>

namespace fnk {

// windows.*.compute_samples_scalar in CPU semantics (windows/boxcar.m:62,
// hann.m:62, tukey.m:101-105, triangular.m:62), of |t|
__device__ double window_cpu64(int id, double param, double t)
{
    const double ta = fabs(t);
    switch (id) {
    case 0: return ta < 1.0 ? 1.0 : 0.0;                                  // boxcar
    case 1: return (1.0 + cos(3.141592653589793 * ta)) / 2.0;            // hann
    case 2: {                                                             // tukey
        const double rect = 1.0 - param;
        if (!(ta < 1.0)) return 0.0;
        if (ta > rect) return (1.0 + cos(3.141592653589793 * ((ta - rect) / param))) / 2.0;
        return 1.0;
    }
    case 4: return 1.0 - ta;                                              // triangular
    default: return 0.0;
    }
}

struct NwSide {
    int win;          // receive / transmit window id
    double param;     // its parameter
    const double *F;  // F-number per bin
    float *h_out;     // L / 2 per (voxel, element), or null when this side is not windowed
};

// apodization of element k (distance d = x_k - x) for the aperture of F at depth
// z, and whether that aperture is non-empty: the aperture bounds of the
// beamforming kernels (make_aperture, FP64) so that both count the same rows;
// boxcar on the open interval |x_k - x| / h < 1 (das_pw.m:441), other windows of
// d over the snapped half-width of its side (das_pw.m:445-452)
__device__ void apod64(const BeamArgs &A, double x, double z, double F, int k, double d, int win, double param,
                       bool *valid, double *w)
{
    const double xp = x / A.pitch, zp = z / (2.0 * A.pitch);
    const double hp = zp / F;
    const double ua = A.M + (xp - hp), ub = A.M + (xp + hp);
    const double lbd = fmax(ceil(ua), 0.0), ubd = fmin(floor(ub), (double)(A.n_el - 1));
    *valid = lbd <= ubd;
    if (!*valid) {
        *w = 0.0;
        return;
    }
    if (win == 0) {
        const double lo = fmax(floor(ua) + 1.0, 0.0), hi = fmin(ceil(ub) - 1.0, (double)(A.n_el - 1));
        *w = ((double)k >= lo && (double)k <= hi) ? 1.0 : 0.0;
        return;
    }
    const double wl = x - ((lbd - A.M) * A.pitch - 0.5 * A.width);
    const double wr = ((ubd - A.M) * A.pitch + 0.5 * A.width) - x;
    *w = window_cpu64(win, param, d < 0.0 ? d / wl : d / wr);
}

__device__ double block_sum128(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1];
}

// one block of 128 threads per voxel, thread k = element k
__global__ __launch_bounds__(128) void windowed_rows_kernel(BeamArgs A, int mode, NwSide rx, NwSide tx)
{
    __shared__ double red[2];
    const long v = blockIdx.x;
    const int k = threadIdx.x;
    const bool elem = k < A.n_el;
    const int ix = (int)(v / A.n_z), iz = (int)(v - (long)ix * A.n_z);
    const double x = A.px[ix], z = A.pz[iz];
    const double d = ((double)k - A.M) * A.pitch - x;
    const double clamp_z = sqrt(z / (2.88 * A.c0));
    int L_rx = 0, L_tx = 0;
    for (int b = 0; b < A.n_f; ++b) {
        double Fr = rx.F[b];
        if (A.clamp) Fr = fmin(Fr, A.sqrt_f[b] * clamp_z);  // das_pw.m:374-377
        bool vr, vt = true;
        double wr, wt = 0.0;
        apod64(A, x, z, Fr, k, d, rx.win, rx.param, &vr, &wr);
        if (mode == MODE_SA) apod64(A, x, z, tx.F[b], k, d, tx.win, tx.param, &vt, &wt);
        if (!elem) wr = wt = 0.0;
        const double sr = block_sum128(wr, red);
        double st = 0.0;
        if (mode == MODE_SA) st = block_sum128(wt, red);
        if (vr && vt) {  // a row of this voxel; positive normalized weights (NaN for a zero sum is not)
            L_rx += (wr / sr > 0.0) ? 1 : 0;
            L_tx += (wt / st > 0.0) ? 1 : 0;
        }
    }
    if (!elem) return;
    const size_t o = (size_t)v * A.n_el + k;
    if (rx.h_out) rx.h_out[o] = 0.5f * (float)L_rx;
    if (tx.h_out) tx.h_out[o] = 0.5f * (float)L_tx;
}

hipError_t launch_windowed_rows(const BeamArgs &A, int mode, int win_rx, double par_rx, float *h_rx, int win_tx,
                                double par_tx, float *h_tx, hipStream_t s)
{
    const long n_vox = (long)A.n_x * A.n_z;
    const NwSide rx{win_rx, par_rx, A.Frx, h_rx}, tx{win_tx, par_tx, mode == MODE_SA ? A.Ftx : A.Frx, h_tx};
    hipLaunchKernelGGL(windowed_rows_kernel, dim3((unsigned)n_vox), dim3(128), 0, s, A, mode, rx, tx);
    return hipGetLastError();
}

}  // namespace fnk
