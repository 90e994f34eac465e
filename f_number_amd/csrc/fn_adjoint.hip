// fn_adjoint.hip -- adjoint (backprojection) of the plane-wave beamformer and the
// vector-Jacobian products of the differentiation form.
//
// The forward map (das_pw.m:383-467 with the time of flight of das_pw.m:336,
// 339-340, 427-428) is linear in the RF data x:
//   I(v) = s sum_n sum_a T_n(v,a) sum_k [a_nk(v) / w_n(v)] e^{j w_n tau_rx(v,k)} X_n(k,a),
//   X_n(k,a) = sum_t x[a][k][t] e^{-j 2 pi n t / N},
// with s the output scale (2 or 2/N), T_n = e^{j w_n tau_tx}, w_n(v) the
// normalising weight sum.  Its adjoint for the real inner product
// <u, y> = Re sum conj(u) y is
//   Z_n(k,a)      = s sum_v y(v) conj(T_n(v,a) a_nk(v) e^{j w_n tau_rx(v,k)}) / w_n(v)   (voxel contraction)
//   (A^H y)[a][k][t] = Re sum_n Z_n(k,a) e^{+j 2 pi n t / N}                          (band-limited inverse DFT)
//
// Voxel contraction on the matrix cores (v_mfma_f32_16x16x4f32, Gauss three
// multiplications): per bin, D[element][angle] += sum_v conj(B)[element][v] U[v][angle]
// with U = s y conj(T) / w.  The contraction index (voxels) is the MFMA K, so
// lane l holds the pairs (voxel 4s + l/16 of its 16-voxel group, element
// l%16 + 16 mt); apertures are computed per voxel in the forward layout and
// exchanged with lane shuffles.  A block owns a chunk of 8 bins and a voxel
// range; it sums the wave results of each bin in LDS and writes one partial
// per (range, bin), reduced by a second kernel.  Phase tiles are recomputed
// per bin, never stored.
//
// Vector-Jacobian products for a cotangent g (J real, dJ = sum Re(conj(g) dI)):
//   d J / d delay_k = sum_n w_n sum_a Im(conj(X_n(k,a)) Z_n(k,a))   with Z from y = g
//                   = sum_n w_n sum_a (Xr Zi - Xi Zr)
//   d J / d c_tof   = sum_v Re(conj(g(v)) dI(v)/dc)                 (dI/dc from the tangent kernel)
#include "fn_kernels.hip.h"

#include <hip/hip_runtime.h>

namespace fnk {

constexpr int ADJ_C = 8;       // bins per block
constexpr int ADJ_WAVES = 4;   // 256 threads, one wave per SIMD (the accumulators need the full register file)
constexpr int ADJ_ET = 8;      // element tiles of 16 (128 elements)
constexpr int ADJ_PAD = 17;    // LDS row stride (floats) of one 16 x 16 tile
constexpr int ADJ_TILE = 16 * ADJ_PAD;
constexpr size_t ADJ_LDS = (size_t)ADJ_C * 2 * ADJ_ET * ADJ_TILE * sizeof(float);

template <int WCR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void adjoint_kernel(AdjArgs S)
{
    extern __shared__ __align__(16) float zacc[];  // [ADJ_C][2][ADJ_ET][16][ADJ_PAD]
    const BeamArgs &A = S.B;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cidx = lane & 15, q = lane >> 4;
    const long n_vox = (long)A.n_x * A.n_z;
    const int b0 = blockIdx.y * ADJ_C;
    const int nb = min(ADJ_C, A.n_f - b0);
    const int at = blockIdx.z;
    const long vr0 = (long)blockIdx.x * S.vox_per_range;
    const long vr1 = min(vr0 + S.vox_per_range, n_vox);

    for (int i = threadIdx.x; i < ADJ_C * 2 * ADJ_ET * ADJ_TILE; i += 64 * ADJ_WAVES) zacc[i] = 0.f;
    __syncthreads();

    const float pf = (float)A.pitch;
    const int a_lane = 16 * at + cidx;  // angle of this lane's U column
    double ex = 0.0, ez = 0.0, xr = 0.0;
    if (a_lane < A.n_tx) {
        ex = A.txp[3 * a_lane];
        ez = A.txp[3 * a_lane + 1];
        xr = A.txp[3 * a_lane + 2];
    }

    for (long base = vr0 + 16 * wave; base < vr1; base += 16 * ADJ_WAVES) {
        // forward layout: voxel base + cidx computes the apertures
        const long vf = min(base + cidx, n_vox - 1);
        const int ixf = (int)(vf / A.n_z), izf = (int)(vf - (long)ixf * A.n_z);
        const double xf = A.px[ixf], zf = A.pz[izf];
        const double xpf = xf / A.pitch, zpf = zf / (2.0 * A.pitch);
        const double clamp_zf = sqrt(zf / (2.88 * A.c0));

        // adjoint layout: voxels base + 4s + q, elements cidx + 16 mt
        uint32_t Qrx[4][ADJ_ET], Qtx[4];
        float dg[4], yr[4], yi[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const long u = base + 4 * s + q;
            const bool ok = u < vr1;
            const long uu = ok ? u : n_vox - 1;
            const int ix = (int)(uu / A.n_z), iz = (int)(uu - (long)ix * A.n_z);
            const double x = A.px[ix], z = A.pz[iz];
            const float2 yv = S.y[uu];  // uu is clamped into the image
            const bool use = ok && a_lane < A.n_tx;
            yr[s] = use ? yv.x : 0.f;
            yi[s] = use ? yv.y : 0.f;
            const double zc = z / A.ctof;
#pragma unroll
            for (int mt = 0; mt < ADJ_ET; ++mt) {
                const int e = cidx + 16 * mt;
                const double lat = (((double)e - A.M) * A.pitch - x) / A.ctof;
                const double tau = sqrt(lat * lat + zc * zc) + (e < A.n_el ? A.delay[e] : 0.0);
                Qrx[s][mt] = frac32(A.f_s * tau / A.N);
            }
            const double ttx = (ex * (x - xr) + ez * z) / A.ctof + A.t0 / A.f_s;
            Qtx[s] = frac32(A.f_s * ttx / A.N);
            dg[s] = (float)(((double)cidx - A.M) * A.pitch - x);
        }

        for (int c = 0; c < nb; ++c) {
            const int b = b0 + c;
            const uint32_t n = (uint32_t)(A.n_lb + b);
            double Fr = A.Frx[b];
            if (A.clamp) Fr = fmin(Fr, A.sqrt_f[b] * clamp_zf);  // das_pw.m:374-377
            const Aperture ap = make_aperture(xf, xpf, zpf, Fr, A, A.support_rx);

            // Gauss: conj(B) U = (br - j bi)(ur + j ui):
            //   K1 = (br - bi) ur, K2 = br (ui - ur), K3 = -bi (ur + ui); Re = K1 - K3, Im = K1 + K2
            floatx4 K1[ADJ_ET], K2[ADJ_ET], K3[ADJ_ET];
#pragma unroll
            for (int mt = 0; mt < ADJ_ET; ++mt) {
                K1[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                K2[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                K3[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
            // one voxel step (4 voxels x 128 elements) at a time: weights, their
            // sum, U, then the MFMAs of the step
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int src = 4 * s + q;
                const int klo = __shfl(ap.klo, src), khi = __shfl(ap.khi, src);
                const float iwl = __shfl(ap.inv_wl, src), iwr = __shfl(ap.inv_wr, src);
                float wv[ADJ_ET];
                float asum = 0.f;
#pragma unroll
                for (int mt = 0; mt < ADJ_ET; ++mt) {
                    const int e = cidx + 16 * mt;
                    const float d = fmaf((float)(16 * mt), pf, dg[s]);
                    const float t = fabsf(d) * (d < 0.f ? iwl : iwr);
                    float w = window_eval<WCR>(t, A.wrx);
                    w = (e >= klo && e <= khi) ? w : 0.f;
                    wv[mt] = w;
                    asum += w;
                }
                // weight sum over the 128 elements of voxel 4s + q (16 lanes x 8 tiles)
                asum += __shfl_xor(asum, 1);
                asum += __shfl_xor(asum, 2);
                asum += __shfl_xor(asum, 4);
                asum += __shfl_xor(asum, 8);
                const bool valid = klo <= khi;
                const float sc = valid ? A.out_scale * (A.norm_rx ? 1.0f / asum : 1.0f) : 0.f;
                // U = s y conj(T) / w
                const float rt = rev_of(n * Qtx[s]);
                const float st = __builtin_amdgcn_sinf(rt), ct = __builtin_amdgcn_cosf(rt);
                const float ur = sc * (yr[s] * ct + yi[s] * st);
                const float ui = sc * (yi[s] * ct - yr[s] * st);
                const float u2 = ui - ur, u3 = ur + ui;
#pragma unroll
                for (int mt = 0; mt < ADJ_ET; ++mt) {
                    const float rv = rev_of(n * Qrx[s][mt]);
                    const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                    const float br = wv[mt] * cs, bi = wv[mt] * sn;
                    K1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(br - bi, ur, K1[mt], 0, 0, 0);
                    K2[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(br, u2, K2[mt], 0, 0, 0);
                    K3[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(-bi, u3, K3[mt], 0, 0, 0);
                }
            }
            // lane holds D[4q + r][cidx] = Z(element 16 mt + 4q + r, angle cidx) of this wave's voxels
            float *zc_re = zacc + (size_t)(c * 2 + 0) * ADJ_ET * ADJ_TILE;
            float *zc_im = zacc + (size_t)(c * 2 + 1) * ADJ_ET * ADJ_TILE;
#pragma unroll
            for (int mt = 0; mt < ADJ_ET; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = mt * ADJ_TILE + (4 * q + r) * ADJ_PAD + cidx;
                    atomicAdd(&zc_re[o], K1[mt][r] - K3[mt][r]);
                    atomicAdd(&zc_im[o], K1[mt][r] + K2[mt][r]);
                }
        }
    }
    __syncthreads();
    // partial of this (range, bin chunk, angle tile): [bin][plane][element 128][angle 16]
    const int per_bin = 2 * ADJ_ET * 16 * 16;
    for (int i = threadIdx.x; i < nb * per_bin; i += 64 * ADJ_WAVES) {
        const int c = i / per_bin, rem = i - c * per_bin;
        const int plane = rem / (ADJ_ET * 256), r2 = rem - plane * (ADJ_ET * 256);
        const int mt = r2 >> 8, row = (r2 >> 4) & 15, col = r2 & 15;
        const float val = zacc[(size_t)(c * 2 + plane) * ADJ_ET * ADJ_TILE + mt * ADJ_TILE + row * ADJ_PAD + col];
        const size_t dst = (((size_t)blockIdx.x * A.n_f + (b0 + c)) * S.n_at + at) * per_bin + rem;
        S.P[dst] = val;
    }
}

// Z[b][a][k] = sum over voxel ranges of the partials (one thread per output)
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(const float *__restrict__ P, float2 *__restrict__ Z,
                                                             int n_vr, int n_f, int n_at, int n_tx, int n_el)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long total = (long)n_f * n_tx * n_el;
    if (i >= total) return;
    const int k = (int)(i % n_el);
    const long ba = i / n_el;
    const int a = (int)(ba % n_tx), b = (int)(ba / n_tx);
    const int at = a >> 4, col = a & 15;
    const int per_bin = 2 * ADJ_ET * 256;
    const size_t off_re = (size_t)k * 16 + col, off_im = (size_t)ADJ_ET * 256 + off_re;
    float re = 0.f, im = 0.f;
    for (int r = 0; r < n_vr; ++r) {
        const size_t base = (((size_t)r * n_f + b) * n_at + at) * per_bin;
        re += P[base + off_re];
        im += P[base + off_im];
    }
    Z[i] = make_float2(re, im);
}

// (A^H y)[ch][t] = Re sum_b Z[b][ch] e^{+j 2 pi (n_lb + b) t / N}, ch = a n_el + k.
// blockIdx.y = channel, blockIdx.x * 256 + tid = sample; twiddles e^{+j 2 pi m / N} in LDS,
// index (n t) mod N advanced exactly in integers.
__global__ __launch_bounds__(256) void adjoint_idft_kernel(const float2 *__restrict__ Z, const float2 *__restrict__ tw,
                                                           float *__restrict__ out, int n_t, int n_ch, int N, int n_lb,
                                                           int n_f)
{
    extern __shared__ float2 tws[];
    for (int m = threadIdx.x; m < N; m += 256) tws[m] = tw[m];
    __syncthreads();
    const int ch = blockIdx.y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n_t) return;
    long m = ((long)n_lb * t) % N;
    float acc = 0.f;
    for (int b = 0; b < n_f; ++b) {
        const float2 z = Z[(size_t)b * n_ch + ch];
        const float2 w = tws[m];
        acc = fmaf(z.x, w.x, acc);
        acc = fmaf(-z.y, w.y, acc);
        m += t;
        if (m >= N) m -= N;
    }
    out[(size_t)ch * n_t + t] = acc;
}

// d J / d delay_k = sum_b w_b sum_a (Xr Zi - Xi Zr); one block per element, FP64 sums.
// X in the operand order of the spectra: (((b MT + a/16) KI + k/4) 64 + (a & 15) + 16 (k & 3)).
__global__ __launch_bounds__(256) void vjp_delay_kernel(const float2 *__restrict__ X, const float2 *__restrict__ Z,
                                                        double *__restrict__ out, int n_f, int n_tx, int n_el, int MT,
                                                        int KI, int n_lb, double w_unit)
{
    __shared__ double red[256];
    const int k = blockIdx.x;
    double acc = 0.0;
    for (long i = threadIdx.x; i < (long)n_f * n_tx; i += 256) {
        const int b = (int)(i / n_tx), a = (int)(i - (long)b * n_tx);
        const float2 x = X[(((size_t)b * MT + (a >> 4)) * KI + (k >> 2)) * 64 + (a & 15) + 16 * (k & 3)];
        const float2 z = Z[((size_t)b * n_tx + a) * n_el + k];
        acc += w_unit * (double)(n_lb + b) * ((double)x.x * z.y - (double)x.y * z.x);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

// sum_v Re(conj(g) d): stage 1 per-block partials, stage 2 one block (deterministic)
__global__ __launch_bounds__(256) void dot_re_partial_kernel(const float2 *__restrict__ g, const float2 *__restrict__ d,
                                                             double *__restrict__ part, long n)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float2 a = g[i], b = d[i];
        acc += (double)a.x * b.x + (double)a.y * b.y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_adjoint(int wcr, dim3 grid, hipStream_t s, const AdjArgs &S)
{
    const void *k = wcr == WC_RANGE ? reinterpret_cast<const void *>(adjoint_kernel<WC_RANGE>)
                  : wcr == WC_LINEAR ? reinterpret_cast<const void *>(adjoint_kernel<WC_LINEAR>)
                                     : reinterpret_cast<const void *>(adjoint_kernel<WC_COS>);
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ADJ_LDS);
    if (e != hipSuccess) return e;
    switch (wcr) {
    case WC_RANGE: hipLaunchKernelGGL(adjoint_kernel<WC_RANGE>, grid, dim3(64 * ADJ_WAVES), ADJ_LDS, s, S); break;
    case WC_LINEAR: hipLaunchKernelGGL(adjoint_kernel<WC_LINEAR>, grid, dim3(64 * ADJ_WAVES), ADJ_LDS, s, S); break;
    default: hipLaunchKernelGGL(adjoint_kernel<WC_COS>, grid, dim3(64 * ADJ_WAVES), ADJ_LDS, s, S); break;
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const float *P, float2 *Z, int n_vr, int n_f, int n_at, int n_tx, int n_el,
                                 hipStream_t s)
{
    const long total = (long)n_f * n_tx * n_el;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P, Z, n_vr, n_f,
                       n_at, n_tx, n_el);
    return hipGetLastError();
}

hipError_t launch_adjoint_idft(const float2 *Z, const float2 *tw, float *out, int n_t, int n_ch, int N, int n_lb,
                               int n_f, hipStream_t s)
{
    const size_t lds = (size_t)N * sizeof(float2);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(adjoint_idft_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(adjoint_idft_kernel, dim3((unsigned)((n_t + 255) / 256), (unsigned)n_ch), dim3(256), lds, s, Z,
                       tw, out, n_t, n_ch, N, n_lb, n_f);
    return hipGetLastError();
}

hipError_t launch_vjp_delay(const float2 *X, const float2 *Z, double *out, int n_f, int n_tx, int n_el, int MT, int KI,
                            int n_lb, double w_unit, hipStream_t s)
{
    hipLaunchKernelGGL(vjp_delay_kernel, dim3((unsigned)n_el), dim3(256), 0, s, X, Z, out, n_f, n_tx, n_el, MT, KI,
                       n_lb, w_unit);
    return hipGetLastError();
}

hipError_t launch_dot_re(const float2 *g, const float2 *d, long n, double *part, int n_part, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(dot_re_partial_kernel, dim3((unsigned)n_part), dim3(256), 0, s, g, d, part, n);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, part, n_part, out);
    return hipGetLastError();
}

}  // namespace fnk
