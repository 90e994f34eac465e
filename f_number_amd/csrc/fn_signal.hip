// fn_signal.hip -- focused RF signal of the last voxel (das_pw.m:477-481,
// das_sa.m:501-505): the band-limited spectrum that the CPU path focuses for the
// last voxel it processes,
//   F(n) = 2 sum_angles sum_k w_nk e^{j w_n tau_k} X_n(k)        (plane waves)
//   F(n) = 2 sum_tx sum_rx w_tx w_rx e^{j w_n (tau_tx + tau_rx)} X_n(rx, tx)   (SA)
// with the normalized (and, for normalizations.windowed, tapered) apodization of
// the beamforming kernels, zero outside the voxel's valid bins; and its inverse
// DFT of length N: s[t] = (1/N) sum_n F(n) e^{+j 2 pi n t / N} (MATLAB ifft).
// Multi-angle plans sum the angles' focused spectra (the compounded signal).
#include "fn_kernels.hip.h"

#include <hip/hip_runtime.h>

namespace fnk {

// receive / transmit window of any class, with the support test
__device__ __forceinline__ float window_rt(int wc, float t, const WinCoef &w)
{
    if (wc == WC_RANGE) return 1.0f;
    if (wc == WC_LINEAR) return window_eval<WC_LINEAR>(t, w);
    return window_eval<WC_COS>(t, w);  // every cosine class has the general flat / cut form
}

__device__ float2 block_sum2_128(float2 v, float2 *red)
{
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o);
        v.y += __shfl_xor(v.y, o);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return make_float2(red[0].x + red[1].x, red[0].y + red[1].y);
}

struct SigArgs {
    int mode, wcr, wct, ix, iz, MT_all, KI, nw;
    float2 *F;  // focused spectrum [n_f]
};

// one block of 128 threads per bin, thread k = (receive) element k
__global__ __launch_bounds__(128) void focused_spectrum_kernel(BeamArgs A, SigArgs S)
{
    __shared__ float2 red[2];
    __shared__ float wtx_s[128];
    __shared__ uint32_t qtx_s[128];
    const int b = blockIdx.x, k = threadIdx.x;
    const bool elem = k < A.n_el;
    const double x = A.px[S.ix], z = A.pz[S.iz];
    const double xp = x / A.pitch, zp = z / (2.0 * A.pitch);
    const double clamp_z = sqrt(z / (2.88 * A.c0));
    const long v = (long)S.ix * A.n_z + S.iz;
    auto F_rx = [&](int bb) {
        double Fr = A.Frx[bb];
        if (A.clamp) Fr = fmin(Fr, A.sqrt_f[bb] * clamp_z);  // das_pw.m:374-377
        return Fr;
    };
    auto valid_at = [&](int bb) {
        bool ok = make_aperture(x, xp, zp, F_rx(bb), A, A.support_rx).valid;
        if (S.mode == MODE_SA) ok = ok && make_aperture(x, xp, zp, A.Ftx[bb], A, A.support_tx).valid;
        return ok;
    };
    const Aperture ap = make_aperture(x, xp, zp, F_rx(b), A, A.support_rx);
    Aperture at{};
    if (S.mode == MODE_SA) at = make_aperture(x, xp, zp, A.Ftx[b], A, A.support_tx);
    const bool valid = ap.valid && (S.mode != MODE_SA || at.valid);
    if (!valid) {  // block-uniform
        if (k == 0) S.F[b] = make_float2(0.f, 0.f);
        return;
    }
    int row = 0;  // windowed normalization: valid rows before this bin
    if (S.nw)
        for (int bb = 0; bb < b; ++bb) row += valid_at(bb) ? 1 : 0;

    const uint32_t n = (uint32_t)(A.n_lb + b);
    const float dk = (float)(((double)k - A.M) * A.pitch - x);
    float w = 0.f;
    if (elem) {
        w = window_rt(S.wcr, fabsf(dk) * (dk < 0.f ? ap.inv_wl : ap.inv_wr), A.wrx);
        w = (k >= ap.klo && k <= ap.khi) ? w : 0.f;
    }
    const float asum = block_sum2_128(make_float2(w, 0.f), red).x;
    float sc = 2.0f * (A.norm_rx ? 1.0f / asum : 1.0f);  // analytic spectra: 2 x DFT (das_pw.m:384)
    if (S.nw && A.nw_rx) w *= elem ? taper_at(row, A.nw_rx[v * A.n_el + k], A.wn_rx) : 0.f;

    // receive time of flight of element k (das_pw.m:339-340, 427; das_sa.m:431)
    const double zc = z / A.ctof;
    const double t_half = (S.mode == MODE_SA) ? 0.5 * A.t0 / A.f_s : 0.0;
    uint32_t qrx = 0;
    if (elem) {
        const double lat = (((double)k - A.M) * A.pitch - x) / A.ctof;
        qrx = frac32(A.f_s * (sqrt(lat * lat + zc * zc) + t_half + A.delay[k]) / A.N);
    }
    float ar = 0.f, ai = 0.f;
    if (S.mode == MODE_PW) {
        for (int a = 0; a < A.n_tx; ++a) {
            const uint32_t qtx = frac32(A.f_s * (pw_tx_time(A, a, x, z) + A.t0 / A.f_s) / A.N);
            if (!elem) continue;
            const float rv = rev_of(n * qrx + n * qtx);
            const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
            const float2 X = A.X[(((size_t)b * S.MT_all + (a >> 4)) * S.KI + (k >> 2)) * 64 + (a & 15) + 16 * (k & 3)];
            ar += cs * X.x - sn * X.y;
            ai += cs * X.y + sn * X.x;
        }
    } else {
        // transmit element k's apodization and phase word, shared through LDS
        float wt = 0.f;
        if (elem) {
            wt = window_rt(S.wct, fabsf(dk) * (dk < 0.f ? at.inv_wl : at.inv_wr), A.wtx);
            wt = (k >= at.klo && k <= at.khi) ? wt : 0.f;
        }
        const float tsum = block_sum2_128(make_float2(wt, 0.f), red).x;
        sc *= A.norm_tx ? 1.0f / tsum : 1.0f;
        if (S.nw && A.nw_tx) wt *= elem ? taper_at(row, A.nw_tx[v * A.n_el + k], A.wn_tx) : 0.f;
        wtx_s[k] = wt;
        qtx_s[k] = qrx;  // the same time of flight formula for the transmitting element
        __syncthreads();
        if (elem)
            for (int i = 0; i < A.n_el; ++i) {
                const float rv = rev_of(n * qrx + n * qtx_s[i]);
                const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                const float2 X =
                    A.X[(((size_t)b * S.MT_all + (i >> 4)) * S.KI + (k >> 2)) * 64 + (i & 15) + 16 * (k & 3)];
                ar += wtx_s[i] * (cs * X.x - sn * X.y);
                ai += wtx_s[i] * (cs * X.y + sn * X.x);
            }
    }
    const float2 tot = block_sum2_128(make_float2(w * ar, w * ai), red);
    if (k == 0) S.F[b] = make_float2(sc * tot.x, sc * tot.y);
}

// s[t] = (1/N) sum_b F[b] e^{+j 2 pi (n_lb + b) t / N}, t = 0 .. N-1
__global__ __launch_bounds__(256) void signal_idft_kernel(const float2 *__restrict__ F, const float2 *__restrict__ tw,
                                                          float2 *__restrict__ out, int N, int n_lb, int n_f)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= N) return;
    long m = ((long)n_lb * t) % N;
    float sr = 0.f, si = 0.f;
    for (int b = 0; b < n_f; ++b) {
        const float2 f = F[b], w = tw[m];
        sr += f.x * w.x - f.y * w.y;
        si += f.x * w.y + f.y * w.x;
        m += t;
        if (m >= N) m -= N;
    }
    const float inv = 1.0f / (float)N;
    out[t] = make_float2(sr * inv, si * inv);
}

hipError_t launch_focused_signal(const BeamArgs &A, int mode, int wcr, int wct, int MT_all, int KI, int nw, int ix,
                                 int iz, float2 *F, const float2 *tw, float2 *out, hipStream_t s)
{
    const SigArgs S{mode, wcr, wct, ix, iz, MT_all, KI, nw, F};
    hipLaunchKernelGGL(focused_spectrum_kernel, dim3((unsigned)A.n_f), dim3(128), 0, s, A, S);
    hipLaunchKernelGGL(signal_idft_kernel, dim3((unsigned)((A.N + 255) / 256)), dim3(256), 0, s, F, tw, out, A.N,
                       A.n_lb, A.n_f);
    return hipGetLastError();
}

}  // namespace fnk
