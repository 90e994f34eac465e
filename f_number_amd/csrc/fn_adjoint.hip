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
// lane l holds the pairs (voxel 4s + l/16 of a 16-voxel group, element
// l%16 + 16 mt).  A block of 4 waves owns 4 consecutive bins (one per wave)
// and a voxel range: per 16-voxel group the block computes the phase words
// (FP64 time of flight reduced to 2^-32 revolutions) once into LDS, every wave
// contracts the group for its own bin and keeps its 16 x 128 x 16 result in
// registers across the whole range -- no atomics; one partial per (range, bin)
// is written at the end and reduced by a second kernel.  Apertures: each wave
// computes those of its bin for 64 voxels at once (one voxel per lane, from
// per-voxel FP64 terms precomputed with the phase words) every fourth group and
// keeps them in LDS for the four groups.
// Phase tiles are recomputed per bin, never stored.
//
// Vector-Jacobian products for a cotangent g (J real, dJ = sum Re(conj(g) dI)):
//   d J / d delay_k = sum_n w_n sum_a Im(conj(X_n(k,a)) Z_n(k,a))   with Z from y = g
//                   = sum_n w_n sum_a (Xr Zi - Xi Zr)
//   d J / d c_tof   = sum_v Re(conj(g(v)) dI(v)/dc)                 (dI/dc from the tangent kernel)
#include "fn_kernels.hip.h"

#include <hip/hip_runtime.h>

namespace fnk {

constexpr int ADJ_WAVES = 4;   // bins per block, one per wave (two blocks per CU)
constexpr int ADJ_ET = 8;      // element tiles of 16 (128 elements)
constexpr int ADJ_QS = 144;    // LDS words per voxel of the receive phase table ([mt][cidx] + pad: lanes
                               // 0-31 of a read meet 32 distinct banks)
constexpr int ADJ_GV = 32;     // voxels per LDS group (one barrier each): two 16-voxel contraction groups

// FULL: no per-element support test on a 128-element array (no padded
// elements).  Either the support is every element (the test reduces to the
// aperture's validity, which the voxel scale already carries), or it is the
// snapped aperture [lb, ub] of a window cut at t = 1 with pitch > width / 2: an
// element outside a valid snapped aperture then has t > 1 (with W the snapped
// half-width on its side, t - 1 >= (pitch - width/2) / (h + width/2)), so the
// window's own cut zeroes it.
template <int WCR, bool FULL>
__global__ __launch_bounds__(64 * ADJ_WAVES) __attribute__((amdgpu_waves_per_eu(2))) void adjoint_kernel(AdjArgs S)
{
    // group tables, double-buffered: the next group is loaded while this one is contracted
    __shared__ __align__(16) uint32_t q_rx[2][ADJ_GV * ADJ_QS];  // [voxel i][mt][cidx]: element cidx + 16 mt
    __shared__ uint32_t q_tx[2][ADJ_GV * 16];      // [voxel i][angle column]
    // cotangent times s / w_n(v) of each wave's bin for the group's voxels (0 outside the range)
    __shared__ float2 ys[2][ADJ_WAVES][ADJ_GV];
    // per wave, for 64 voxels: the window argument of element e on each side as e r1 / 16 + r0
    // and e l1 / 16 + l0 (r1, l1 the slopes of 16 elements; for WC_COS_RF with the window's
    // quarter-period scale and offset folded in), and the support klo | khi << 16
    __shared__ __align__(16) uint4 apx[ADJ_WAVES][64];
    __shared__ uint32_t akk[ADJ_WAVES][64];
    const BeamArgs &A = S.B;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int cidx = lane & 15, q = lane >> 4;
    const long n_vox = (long)A.n_x * A.n_z;
    const int b = blockIdx.x * ADJ_WAVES + wave;
    const bool active = b < A.n_f;
    const int at = blockIdx.z;
    const long vr0 = (long)blockIdx.y * S.vox_per_range;
    const long vr1 = min(vr0 + S.vox_per_range, n_vox);
    const float c16 = (float)cidx * 0.0625f;  // element cidx in units of 16 elements
    const uint32_t n = (uint32_t)(A.n_lb + (active ? b : 0));
    double Fb = 0.0, sqf = 0.0;
    if (active) {
        Fb = A.Frx[b];
        sqf = A.sqrt_f[b];
    }

    floatx4 K1[ADJ_ET], K2[ADJ_ET], K3[ADJ_ET];
#pragma unroll
    for (int mt = 0; mt < ADJ_ET; ++mt) {
        K1[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        K2[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
        K3[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
    }

    const int qtx_stride = 16 * S.n_at;
    // the phase words and voxel data of the group at `base`: fetched into registers
    // at the top of the previous group's step and written to LDS buffer `buf` after
    // its contraction, so the global loads are in flight during that contraction
    static_assert(ADJ_WAVES == 4 && ADJ_GV % 16 == 0, "group loader assumes 256 threads");
    constexpr int RX4 = ADJ_GV * 32 / 256, TX1 = ADJ_GV * 16 / 256;  // uint4 / words per thread
    struct GroupRegs {
        uint4 rx[RX4];
        uint32_t tx[TX1];
        float2 ys;
    };
    static_assert(ADJ_WAVES * ADJ_GV <= 256, "one scale per thread");
    auto fetch_group = [&](long base) {
        GroupRegs r;
#pragma unroll
        for (int h = 0; h < RX4; ++h) {
            const int it = tid + 256 * h, i = it >> 5, j4 = it & 31;  // ADJ_GV voxels x 32 uint4
            const long u = base + i;
            r.rx[h] = make_uint4(0u, 0u, 0u, 0u);
            if (u < n_vox) r.rx[h] = reinterpret_cast<const uint4 *>(S.qrx + (size_t)u * 128)[j4];
        }
#pragma unroll
        for (int h = 0; h < TX1; ++h) {
            const int it = tid + 256 * h, iv = it >> 4, col = it & 15;  // ADJ_GV voxels x 16 angle columns
            const long uv = base + iv;
            r.tx[h] = uv < n_vox ? S.qtx[(size_t)uv * qtx_stride + 16 * at + col] : 0u;
        }
        r.ys = make_float2(0.f, 0.f);
        if (tid < ADJ_WAVES * ADJ_GV) {
            const int bs = blockIdx.x * ADJ_WAVES + tid / ADJ_GV;
            const long u3 = base + tid % ADJ_GV;
            if (bs < A.n_f && u3 < vr1) {
                const float sc = S.vsc[(size_t)bs * n_vox + u3];
                const float2 y = S.y[u3];
                r.ys = make_float2(sc * y.x, sc * y.y);
            }
        }
        return r;
    };
    auto store_group = [&](const GroupRegs &r, int buf) {
#pragma unroll
        for (int h = 0; h < RX4; ++h) {
            const int it = tid + 256 * h, i = it >> 5, j4 = it & 31;
            *reinterpret_cast<uint4 *>(&q_rx[buf][i * ADJ_QS + 4 * j4]) = r.rx[h];
        }
#pragma unroll
        for (int h = 0; h < TX1; ++h) q_tx[buf][tid + 256 * h] = r.tx[h];
        if (tid < ADJ_WAVES * ADJ_GV) ys[buf][tid / ADJ_GV][tid % ADJ_GV] = r.ys;
    };
    // per-voxel aperture terms of voxel `base + lane` (clamped into the grid)
    auto load_geo = [&](long base) {
        const double *g = S.vgeo + 4 * (size_t)min(base + lane, n_vox - 1);
        return make_double4(g[0], g[1], g[2], g[3]);
    };
    const bool a_ok = 16 * at + cidx < A.n_tx;
    const double x0 = -A.M * A.pitch;
    double4 geo = make_double4(0.0, 0.0, 0.0, 0.0);
    if (vr0 < vr1) {
        store_group(fetch_group(vr0), 0);
        if (active) geo = load_geo(vr0);
    }
    __syncthreads();
    int cur = 0;
    for (long gbase = vr0; gbase < vr1; gbase += ADJ_GV, cur ^= 1) {
        const bool more = gbase + ADJ_GV < vr1;
        GroupRegs next{};
        if (more) next = fetch_group(gbase + ADJ_GV);
#pragma unroll
        for (int hb = 0; hb < ADJ_GV / 16; ++hb) {
        const long base = gbase + 16 * hb;  // contraction group of 16 voxels
        const int o = 16 * hb;              // its offset in the LDS group
        const int gq = (int)((base - vr0) >> 4) & 3;           // group within the 64 voxels of apx
        if (active && base < vr1) {
            if (gq == 0) {
                // apertures of this wave's bin for voxels base .. base + 63, one per lane (forward formulas)
                double Fr = Fb;
                if (A.clamp) Fr = fmin(Fr, sqf * geo.w);  // das_pw.m:374-377
                const Aperture ap = make_aperture(geo.x, geo.y, geo.z, Fr, A, A.support_rx);
                // d = x_e - x = (x_0 - x) + e pitch; right side d / W_R, left side -d / W_L
                // (WC_COS_RF: times q_s, plus q_o)
                const double sl = WCR == WC_COS_RF ? (double)A.wrx.q_s : 1.0;
                const double qo = WCR == WC_COS_RF ? (double)A.wrx.q_o : 0.0;
                const double wr = (double)ap.inv_wr * sl, wl = (double)ap.inv_wl * sl, dx = x0 - geo.x;
                akk[wave][lane] = (uint32_t)ap.klo | ((uint32_t)ap.khi << 16);
                apx[wave][lane] = make_uint4(__float_as_uint((float)(16.0 * A.pitch * wr)),
                                             __float_as_uint((float)(-16.0 * A.pitch * wl)),
                                             __float_as_uint((float)(dx * wr + qo)), __float_as_uint((float)(qo - dx * wl)));
                if (base + 64 < vr1) geo = load_geo(base + 64);  // consumed four groups later
            }

            // element tiles [t_lo, t_hi] of 16 that meet the support of some voxel of the
            // group (wave-uniform): every other tile has zero window for all 16 voxels
            int t_lo, t_hi;
            {
                const uint32_t kk = akk[wave][16 * gq + cidx];
                const int kl = (int)(kk & 0xffffu), kh = (int)(kk >> 16);
                const int lo = row_min16(kl <= kh ? kl : 0x7fff), hi = row_max16(kl <= kh ? kh : -1);
                t_lo = __builtin_amdgcn_readfirstlane(lo) >> 4;
                t_hi = __builtin_amdgcn_readfirstlane(hi) >> 4;  // -1: no voxel of the group has elements
            }

            // Gauss: conj(B) U = (br - j bi)(ur + j ui):
            //   K1 = (br - bi) ur, K2 = br (ui - ur), K3 = bi (ur + ui); Re = K1 + K3, Im = K1 + K2
            // Per voxel step s (voxel 4 s + q of the group): U, and the window argument as
            // max(mt R1 + R0, mt L1 + L0) over the element tile mt (each side's term is the
            // larger exactly on its own side, so no per-element select)
            float Ur[4], U2[4], U3[4], R0[4], R1[4], L0[4], L1[4];
            int Klo[4], Khi[4];
            int qoff[4];
    #pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 4 * s + q;  // voxel of this lane in step s
                const uint4 apv = apx[wave][16 * gq + i];  // written by this wave (LDS operations of a wave are ordered)
                if (!FULL) {
                    const uint32_t kk = akk[wave][16 * gq + i];
                    Klo[s] = (int)(kk & 0xffffu);
                    Khi[s] = (int)(kk >> 16);
                }
                R1[s] = __uint_as_float(apv.x);
                L1[s] = __uint_as_float(apv.y);
                R0[s] = fmaf(c16, R1[s], __uint_as_float(apv.z));  // element cidx of tile 0
                L0[s] = fmaf(c16, L1[s], __uint_as_float(apv.w));
                // U = s y conj(T) / w, s / w_n(v) from the per-plan table (0 outside a valid aperture)
                const float2 yv = a_ok ? ys[cur][wave][o + i] : make_float2(0.f, 0.f);
                const float rt = rev_of(n * q_tx[cur][(o + i) * 16 + cidx]);
                const float st = __builtin_amdgcn_sinf(rt), ct = __builtin_amdgcn_cosf(rt);
                const float ur = fmaf(yv.x, ct, yv.y * st);
                const float ui = fmaf(yv.y, ct, -(yv.x * st));
                Ur[s] = ur;
                U2[s] = ui - ur;
                U3[s] = ur + ui;
                qoff[s] = (o + i) * ADJ_QS + cidx;
            }
            // element tile outer, voxel step inner: four independent chains per tile
    #pragma unroll
            for (int mt = 0; mt < ADJ_ET; ++mt) {
                if (mt < t_lo || mt > t_hi) continue;
                const int e = cidx + 16 * mt;
    #pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const uint32_t Qv = q_rx[cur][qoff[s] + 16 * mt];
                    const float ta = mt == 0 ? R0[s] : fmaf((float)mt, R1[s], R0[s]);
                    const float tb = mt == 0 ? L0[s] : fmaf((float)mt, L1[s], L0[s]);
                    // max(ta, tb) as a signed-integer max of the bit patterns: fmaxf would
                    // first canonicalize both operands (two more instructions per weight),
                    // since they come from another basic block.  One of the two is >= 0 (each
                    // side's term is non-negative on its own side), and on non-negative floats
                    // the integer order is the float order, so the result is max(ta, tb)
                    const float tm = __int_as_float(max(__float_as_int(ta), __float_as_int(tb)));
                    float w;
                    if (WCR == WC_COS_RF) {
                        const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fmed3f(tm, 0.f, 0.25f));
                        w = c * c;
                    } else {
                        w = window_eval<WCR>(tm, A.wrx);
                    }
                    if (!FULL) w = (e >= Klo[s] && e <= Khi[s]) ? w : 0.f;
                    const float rv = rev_of(n * Qv);
                    const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                    const float br = w * cs, bi = w * sn;
                    K1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(br - bi, Ur[s], K1[mt], 0, 0, 0);
                    K2[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(br, U2[s], K2[mt], 0, 0, 0);
                    K3[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi, U3[s], K3[mt], 0, 0, 0);
                }
            }
        }  // active
        }  // hb
        // buffer cur ^ 1 was last read two groups ago (before the previous barrier)
        if (more) store_group(next, cur ^ 1);
        __syncthreads();  // group `gbase` consumed, group `gbase + ADJ_GV` stored
    }
    if (!active) return;
    // lane holds D[4q + r][cidx] = Z(element 16 mt + 4q + r, angle column cidx) of the range;
    // partial layout [range][bin][angle tile][plane][element 128][angle 16]
    const int per_bin = 2 * ADJ_ET * 256;
    float *dst = S.P + (((size_t)blockIdx.y * A.n_f + b) * S.n_at + at) * per_bin;
#pragma unroll
    for (int mt = 0; mt < ADJ_ET; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = (16 * mt + 4 * q + r) * 16 + cidx;
            dst[o] = K1[mt][r] + K3[mt][r];
            dst[ADJ_ET * 256 + o] = K1[mt][r] + K2[mt][r];
        }
}

// Phase words of every voxel for the voxel contraction: FP64 time of flight
// (das_pw.m:336, 339-340, 427-428 with c_tof and the delay offsets) reduced to
// frac(f_s tau / N) in units of 2^-32.  qrx[v][e] holds element e (zero-padded
// to 128 elements), qtx[v][a] angle a (16 n_at columns).
__global__ __launch_bounds__(256) void adjoint_tables_kernel(BeamArgs A, uint32_t *__restrict__ qrx,
                                                             uint32_t *__restrict__ qtx, double *__restrict__ vgeo,
                                                             int n_at)
{
    const long n_vox = (long)A.n_x * A.n_z;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int n_cols = 128 + 16 * n_at + 1;
    if (i >= n_vox * n_cols) return;
    const long v = i / n_cols;
    const int col = (int)(i - v * n_cols);
    const int ix = (int)(v / A.n_z), iz = (int)(v - (long)ix * A.n_z);
    const double x = A.px[ix], z = A.pz[iz];
    if (col == 128 + 16 * n_at) {
        // aperture terms, the same expressions as the forward kernel's prologue
        vgeo[4 * v] = x;
        vgeo[4 * v + 1] = x / A.pitch;
        vgeo[4 * v + 2] = z / (2.0 * A.pitch);
        vgeo[4 * v + 3] = sqrt(z / (2.88 * A.c0));
    } else if (col < 128) {
        const int e = col;
        const double lat = (((double)e - A.M) * A.pitch - x) / A.ctof, zc = z / A.ctof;
        const double tau = sqrt(lat * lat + zc * zc) + (e < A.n_el ? A.delay[e] : 0.0);
        qrx[v * 128 + e] = frac32(A.f_s * tau / A.N);
    } else {
        const int a = col - 128;
        double tau = 0.0;
        if (a < A.n_tx) tau = pw_tx_time(A, a, x, z) + A.t0 / A.f_s;
        qtx[v * 16 * n_at + a] = frac32(A.f_s * tau / A.N);
    }
}

// Z[b][a][k] = sum over voxel ranges of the partials; one thread per (bin, angle
// tile, element, angle column), the column fastest so the partial reads coalesce
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(const float *__restrict__ P, float2 *__restrict__ Z,
                                                             int n_vr, int n_f, int n_at, int n_tx, int n_el)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long total = (long)n_f * n_at * 16 * ADJ_ET * 16;
    if (i >= total) return;
    const int col = (int)(i & 15), k = (int)((i >> 4) % (16 * ADJ_ET));
    const long bat = i / (16 * 16 * ADJ_ET);  // b n_at + at
    const int at = (int)(bat % n_at), b = (int)(bat / n_at);
    const int a = 16 * at + col;
    if (a >= n_tx || k >= n_el) return;
    const int per_bin = 2 * ADJ_ET * 256;
    const size_t off_re = (size_t)k * 16 + col, off_im = (size_t)ADJ_ET * 256 + off_re;
    float re = 0.f, im = 0.f;
    for (int r = 0; r < n_vr; ++r) {
        const size_t base = (((size_t)r * n_f + b) * n_at + at) * per_bin;
        re += P[base + off_re];
        im += P[base + off_im];
    }
    Z[((size_t)b * n_tx + a) * n_el + k] = make_float2(re, im);
}

// (A^H y)[ch][t] = Re sum_b Z[b][ch] e^{+j 2 pi (n_lb + b) t / N}, ch = a n_el + k, as a
// real product on the matrix cores: D[t][ch] += sum_k A[t][k] B[k][ch] over k = (bin, part),
//   A[t][(b, 0)] = Re w,  A[t][(b, 1)] = -Im w,  w = e^{+j 2 pi m / N}, m = (n_lb + b) t mod N,
//   B[(b, 0)][ch] = Re Z[b][ch],  B[(b, 1)][ch] = Im Z[b][ch].
// v_mfma_f32_16x16x4f32: lane l supplies A[t0 + l % 16][k0 + l / 16] and
// B[k0 + l / 16][c0 + l % 16], so one K step of 4 covers two bins.  A block of 4 waves
// owns 64 samples x 64 channels; a wave 16 samples (1 tile) x 64 channels (4 tiles), with
// the next step's spectra loaded while the current step's products run.
// Twiddles in LDS; each lane advances its index m by 2t mod N per step in integers.
constexpr int IDFT_WT = 1, IDFT_WC = 4;

__global__ __launch_bounds__(256) void adjoint_idft_kernel(const float2 *__restrict__ Z, const float2 *__restrict__ tw,
                                                           float *__restrict__ out, int n_t, int n_ch, int N, int n_lb,
                                                           int n_f)
{
    extern __shared__ float2 tws[];
    for (int m = threadIdx.x; m < N; m += 256) tws[m] = tw[m];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, q = lane >> 4;  // q = k within the step: bin q / 2, part q % 2
    const int part = q & 1, db = q >> 1;
    const int t0 = blockIdx.x * (64 * IDFT_WT) + wave * (16 * IDFT_WT);
    const int c0 = blockIdx.y * 64;
    int m[IDFT_WT], dm[IDFT_WT];
    bool tok[IDFT_WT];
#pragma unroll
    for (int i = 0; i < IDFT_WT; ++i) {
        const int t = t0 + 16 * i + r16;
        tok[i] = t < n_t;
        m[i] = (int)(((long)(n_lb + db) * t) % N);
        dm[i] = (int)((2L * t) % N);
    }
    int ch[IDFT_WC];
#pragma unroll
    for (int j = 0; j < IDFT_WC; ++j) ch[j] = c0 + 16 * j + r16;
    floatx4 acc[IDFT_WT][IDFT_WC];
#pragma unroll
    for (int i = 0; i < IDFT_WT; ++i)
#pragma unroll
        for (int j = 0; j < IDFT_WC; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto load_z = [&](int b, float *bv) {
#pragma unroll
        for (int j = 0; j < IDFT_WC; ++j) {
            float2 z = make_float2(0.f, 0.f);
            if (b < n_f && ch[j] < n_ch) z = Z[(size_t)b * n_ch + ch[j]];
            bv[j] = part ? z.y : z.x;
        }
    };
    float bn[IDFT_WC];
    load_z(db, bn);
    for (int b0 = 0; b0 < n_f; b0 += 2) {
        float a[IDFT_WT], bv[IDFT_WC];
#pragma unroll
        for (int j = 0; j < IDFT_WC; ++j) bv[j] = bn[j];
        load_z(b0 + 2 + db, bn);  // next step's spectra (zeros past the band)
#pragma unroll
        for (int i = 0; i < IDFT_WT; ++i) {
            const float2 w = tws[m[i]];
            a[i] = tok[i] ? (part ? -w.y : w.x) : 0.f;
            m[i] += dm[i];
            if (m[i] >= N) m[i] -= N;
        }
#pragma unroll
        for (int i = 0; i < IDFT_WT; ++i)
#pragma unroll
            for (int j = 0; j < IDFT_WC; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bv[j], acc[i][j], 0, 0, 0);
    }
    // lane holds D[t0 + 16 i + 4 q + r][ch[j]], r = 0..3: four consecutive samples of one channel
    const bool vec = (n_t & 3) == 0;
#pragma unroll
    for (int i = 0; i < IDFT_WT; ++i)
#pragma unroll
        for (int j = 0; j < IDFT_WC; ++j) {
            const int t = t0 + 16 * i + 4 * q;
            if (ch[j] >= n_ch || t >= n_t) continue;
            float *dst = out + (size_t)ch[j] * n_t + t;
            if (vec) {  // t + 3 < n_t and 16-byte aligned
                *reinterpret_cast<float4 *>(dst) = make_float4(acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (t + r < n_t) dst[r] = acc[i][j][r];
            }
        }
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

// Normalization of the voxel contraction per (bin, voxel): s / w_n(v) for a valid
// receive aperture, 0 otherwise, with w_n(v) the sum of the apodization weights over
// the array (das_pw.m:400-410, normalizations.on; s alone for off).  It depends on
// the geometry, the F-numbers and the window only -- not on the time of flight --
// and is computed once per plan, with the contraction kernel's aperture and element
// arithmetic (elements cidx + 16 mt, FULL as there).
template <int WCR, bool FULL>
__global__ __launch_bounds__(256) void adjoint_scale_kernel(BeamArgs A, float *__restrict__ vsc)
{
    const long n_vox = (long)A.n_x * A.n_z;
    const long v = (long)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= n_vox) return;
    const int ix = (int)(v / A.n_z), iz = (int)(v - (long)ix * A.n_z);
    const double x = A.px[ix], z = A.pz[iz];
    double Fr = A.Frx[b];
    if (A.clamp) Fr = fmin(Fr, A.sqrt_f[b] * sqrt(z / (2.88 * A.c0)));  // das_pw.m:374-377
    const Aperture ap = make_aperture(x, x / A.pitch, z / (2.0 * A.pitch), Fr, A, A.support_rx);
    float sc = 0.f;
    if (ap.klo <= ap.khi) {
        const float pf = (float)A.pitch;
        const float x0f = (float)(-A.M * A.pitch - x);
        const float sl = WCR == WC_COS_RF ? A.wrx.q_s : 1.0f;
        const float iwl = ap.inv_wl * sl, iwr = ap.inv_wr * sl;
        float asum = 0.f;
        for (int e = 0; e < 16 * ADJ_ET; ++e) {
            const int cidx = e & 15, mt = e >> 4;
            const float d = fmaf((float)(16 * mt), pf, fmaf((float)cidx, pf, x0f));
            float w;
            if (WCR == WC_COS_RF) {
                w = window_rf(fabsf(d), d < 0.f ? iwl : iwr, A.wrx.q_o);
            } else {
                w = window_eval<WCR>(fabsf(d) * (d < 0.f ? iwl : iwr), A.wrx);
            }
            if (!FULL) w = (e >= ap.klo && e <= ap.khi) ? w : 0.f;
            asum += w;
        }
        sc = A.out_scale * (A.norm_rx ? 1.0f / asum : 1.0f);
    }
    vsc[(size_t)b * n_vox + v] = sc;
}

template <int WCR, bool FULL>
static void launch_adjoint_class(bool scale, dim3 grid, hipStream_t s, const AdjArgs &S)
{
    if (scale) {
        const long n_vox = (long)S.B.n_x * S.B.n_z;
        hipLaunchKernelGGL((adjoint_scale_kernel<WCR, FULL>), dim3((unsigned)((n_vox + 255) / 256), (unsigned)S.B.n_f),
                           dim3(256), 0, s, S.B, const_cast<float *>(S.vsc));
    }
    hipLaunchKernelGGL((adjoint_kernel<WCR, FULL>), grid, dim3(64 * ADJ_WAVES), 0, s, S);
}

hipError_t launch_adjoint(int wcr, bool tables, bool scale, dim3 grid, hipStream_t s, const AdjArgs &S)
{
    const long n_vox = (long)S.B.n_x * S.B.n_z;
    const long n_tab = n_vox * (128 + 16 * S.n_at + 1);
    if (tables)
        hipLaunchKernelGGL(adjoint_tables_kernel, dim3((unsigned)((n_tab + 255) / 256)), dim3(256), 0, s, S.B,
                           const_cast<uint32_t *>(S.qrx), const_cast<uint32_t *>(S.qtx), const_cast<double *>(S.vgeo),
                           S.n_at);
    const bool full = range_free(S.B, 16 * ADJ_ET);
    // the forward's element-loop classes: WC_COS_RF holds when the support test is
    // free here too (128 elements); WC_COS_ALL evaluates as the general cosine
    if (wcr == WC_COS_RF && full) {
        launch_adjoint_class<WC_COS_RF, true>(scale, grid, s, S);
        return hipGetLastError();
    }
    if (wcr == WC_COS_RF || wcr == WC_COS_ALL) wcr = WC_COS;
    switch (wcr) {
    case WC_RANGE:
        if (full) launch_adjoint_class<WC_RANGE, true>(scale, grid, s, S);
        else launch_adjoint_class<WC_RANGE, false>(scale, grid, s, S);
        break;
    case WC_LINEAR:
        if (full) launch_adjoint_class<WC_LINEAR, true>(scale, grid, s, S);
        else launch_adjoint_class<WC_LINEAR, false>(scale, grid, s, S);
        break;
    default:
        if (full) launch_adjoint_class<WC_COS, true>(scale, grid, s, S);
        else launch_adjoint_class<WC_COS, false>(scale, grid, s, S);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const float *P, float2 *Z, int n_vr, int n_f, int n_at, int n_tx, int n_el,
                                 hipStream_t s)
{
    const long total = (long)n_f * n_at * 16 * ADJ_ET * 16;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, P, Z, n_vr, n_f,
                       n_at, n_tx, n_el);
    return hipGetLastError();
}

hipError_t launch_adjoint_idft(const float2 *Z, const float2 *tw, float *out, int n_t, int n_ch, int N, int n_lb,
                               int n_f, hipStream_t s)
{
    const size_t lds = (size_t)N * sizeof(float2);
    static size_t lds_done = 0;  // raised once per process and size (graph capture issues no attribute calls)
    if (lds > 65536 && lds > lds_done) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(adjoint_idft_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        lds_done = lds;
    }
    hipLaunchKernelGGL(adjoint_idft_kernel, dim3((unsigned)((n_t + 64 * IDFT_WT - 1) / (64 * IDFT_WT)), (unsigned)((n_ch + 63) / 64)),
                       dim3(256), lds, s, Z, tw, out, n_t, n_ch, N, n_lb, n_f);
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
