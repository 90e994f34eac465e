// fn_kernels.hip.h -- device code of libfnumber_hip.so for gfx950 (CDNA4).
//
// Two kernels form the hot path of das_pw / das_sa:
//
//  spectra_kernel   band-limited real-to-complex DFT of every RF channel
//                   (replaces cuFFT in gpu_bf_das_pw_rf.cu:917-925 and the
//                   bin selection of das_pw.m:383-384).  Twiddle index
//                   (n t) mod N is reduced exactly in integers; the output is
//                   written directly in the MFMA operand order the
//                   beamforming kernel stages into LDS.
//
//  beamform_kernel  for every band bin n: generates the receive
//                   apodization x phase tile B_n(v,k) = a_{n,k}(v) e^{j w_n tau_k(v)}
//                   in registers (replaces das_kernel_phase_shifts.cu:8-147 /
//                   saft_kernel_phase_shifts.cu:8-199, which wrote it to HBM),
//                   contracts it against the transmit spectra X_n(k,a) on
//                   v_mfma_f32_16x16x4f32 (replaces the cublasCgemm GEMV of
//                   gpu_bf_das_pw_rf.cu:1034-1041), then applies the transmit
//                   phase / transmit apodization, the normalization
//                   (das_kernel_normalization.cu:8-34) and accumulates the
//                   image over bins in registers.
//
// Phases: tau is formed in FP64 and reduced to q = frac(f_s tau / N) in
// units of 2^-32 (uint32).  The phase of bin n is then the exact wrap-around
// integer product n*Q mod 2^32, converted to revolutions and evaluated with
// the hardware sine/cosine (v_sin_f32/v_cos_f32 take revolutions).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include <type_traits>

namespace fnk {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// window classes for the element loop (the generic window below covers the
// rest with runtime coefficients)
// WC_COS_ALL: the raised cosine A + B cos(pi t) over every element with no flat
// part and no cut (CPU-semantics Hann, das_pw.m:449-452 with hann.m:62) on a
// full array (n_el = 4 KI): no support test, the 1/2 of the argument folded
// into the per-bin reciprocals.
// WC_COS_RF: the Hann / Tukey window cut at t = 1 where that cut already zeroes
// every element outside the support (range_free()), so no per-element support
// test; evaluated as cos^2 of a clamped quarter-period argument:
//   1/2 + 1/2 cos(2 pi u) = cos^2(pi u),  u = clamp((t - t_off) s_half, 0, 1/2),
// which is 1 on the flat part (u = 0) and 0 from the cut on (u = 1/2).
enum { WC_RANGE = 0, WC_LINEAR = 1, WC_COS = 2, WC_COS_ALL = 3, WC_COS_RF = 4 };
enum { MODE_PW = 0, MODE_SA = 1 };

// Generic |t| window: 1 for t <= t_flat; for t_flat < t < t_cut:
// A + B cos(pi (t - t_off) s) (cosine class) or A + C t (linear class);
// 0 for t >= t_cut.  Coefficients per window (CPU semantics from
// +windows/<name>.m, MEX semantics from gpu_bf_windows.cu):
//   tukey(alpha): t_flat = t_off = 1 - alpha, t_cut = 1, A = B = 1/2, s = 1/alpha
//   hann:         t_flat = -1, t_off = 0, t_cut = inf (CPU) | 1 (MEX), A = B = 1/2, s = 1
//   triangular:   t_flat = -1, t_cut = inf | 1, A = 1, B = 0, C = -1 (linear class)
//   boxcar:       constant 1 on its integer support (range class)
struct WinCoef {
    float t_flat, t_cut, t_off, A, B, C, s_half;  // cosine argument: (t - t_off) * s_half revolutions
    float q_s, q_o;  // WC_COS_RF: quarter-period argument t q_s + q_o = (t - t_off) s_half / 2
};

// element support of a window for one aperture
enum { SUPPORT_OPEN_H = 0,   // boxcar, CPU: |x_k - x| / h < 1 (das_pw.m:441)
       SUPPORT_SNAPPED = 1,  // [k_lb, k_ub] (MEX: apply() truncation, das_kernel_phase_shifts.cu:94-125)
       SUPPORT_ALL = 2 };    // CPU non-boxcar windows: every element (das_pw.m:449-452)

struct BeamArgs {
    const float2 *X;        // spectra in operand order [n_f][MT][KI][64]
    const void *Xlp;        // 16-bit spectra planes (mixed precision) [n_f][planes][MT][KI/8][64][8]
    const float *xscale;    // per-bin scale of the 16-bit planes
    const double *px, *pz;  // voxel positions
    const double *Frx;      // F-number per band bin (receive)
    const double *Ftx;      // F-number per band bin (transmit, SA)
    const double *txp;      // PW: per angle (e_x, e_z, x_ref)
    const double *sqrt_f;   // sqrt(f_n) per band bin (depth clamp)
    float2 *image;
    float *weights, *weights_tx;
    int n_x, n_z, n_el, n_tx, n_f, n_lb, N;
    int bin_chunk;          // bins per blockIdx.y
    int tx_group;           // M-tiles per blockIdx.z group (PW angle groups)
    int clamp;              // CPU-semantics depth clamp of the PW F-number (das_pw.m:374-377)
    int support_rx, support_tx;        // SUPPORT_* of the receive / transmit window
    int norm_rx, norm_tx;
    // partial images: non-null when several workgroups share a voxel (bin chunks,
    // transmit groups): workgroup (by, bz) writes slot by * slot_z + bz of
    // [n_slots][n_vox] (the tangent: dI/dc in slots n_slots .. 2 n_slots - 1), and
    // reduce_partials_kernel sums the slots in a fixed order (reproducible images)
    float2 *partial;
    int slot_z, n_slots;
    double M, pitch, width, c0, f_s, t0;
    // ensemble: F = ens_frames frames in ens_slots = F rounded up to a multiple of 4
    // frame slots; its transmit rows are angle-major, row = angle ens_slots + slot, so the
    // four rows a lane holds per M-tile share one angle and one transmit phase (rows of
    // slots >= F have zero spectra and no image); images [frame][n_vox]
    int ens_slots, ens_frames;
    const int *row_frame;   // row stack: frame of every transmit row (ens_slots 0)
    float2 *const *frame_image;  // row stack: image of every frame
    int tx_kind;            // TX_PLANE: txp = (e_x, e_z, x_ref); TX_VIRTUAL: txp = (x_vs, z_vs, |r_vs|)
    double ctof;            // speed of sound of the time of flight (c0 unless set)
    const double *delay;    // per-element delay offsets (s), [n_el]
    float *dimage;          // tangent: dI / dc_tof, interleaved complex [n_vox]
    WinCoef wrx, wtx;
    float out_scale;
    // windowed normalization (windowed.m:81-86): half the number of positive rows
    // L / 2 per (voxel, element), receive and (SA) transmit, and the taper windows
    const float *nw_rx, *nw_tx;  // [n_vox][n_el]; 0 where L = 0
    WinCoef wn_rx, wn_tx;
};

// adjoint voxel contraction (fn_adjoint.hip)
struct AdjArgs {
    BeamArgs B;
    const float2 *y;        // cotangent image [n_vox]
    float *P;               // partials [n_vr][n_f][n_at][2][128][16]
    const uint32_t *qrx;    // receive phase words [n_vox][element 128]
    const uint32_t *qtx;    // transmit phase words [n_vox][16 n_at]
    const double *vgeo;     // per voxel: x, x / pitch, z / (2 pitch), sqrt(z / (2.88 c0))
    const float *vsc;       // per bin and voxel [n_f][n_vox]: s / w_n(v), 0 outside a valid aperture
    long vox_per_range;     // multiple of 64
    int n_at;               // angle tiles of 16
};

struct SpecArgs {
    int n_t, n_ch, N, n_lb, n_f, n_el, n_tx, MT, KI;
    // output rows: row a of this RF goes to row row0 + a of an operand layout of MT_out
    // tiles (a row stack of several acquisitions); write_pad also writes the zero rows
    // past n_tx of the last tile (a plan of its own)
    int row0, MT_out, write_pad;
    // ensemble: output row R = angle S + slot reads RF row slot P + angle when slot < F
    // (F frames of P angles, S frame slots; a row of a slot >= F is zero); ens_slots 0:
    // output row = RF row
    int ens_slots, ens_frames, ens_angles;
};

__device__ __forceinline__ uint32_t frac32(double q)
{
    const double f = q - floor(q);
    return (uint32_t)(int64_t)rint(f * 4294967296.0);
}

// The phase word P (units of 2^-32 revolutions) as the float 1 + P 2^-32 with 23
// fraction bits, in one instruction: bits 0x3F800000 | (P >> 9).  v_sin_f32 /
// v_cos_f32 take revolutions and drop the integer part.  Truncating the low 9 bits
// errs by at most 2^-23 revolutions (7.5e-7 rad).
__device__ __forceinline__ float rev_of(uint32_t P)
{
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, P, 9u));
}

// Per (voxel, bin) aperture: das_pw.m:387, 399-411 (CPU) and
// das_kernel_phase_shifts.cu:90-125 (MEX).  All bounds in FP64, as the oracle.
struct Aperture {
    int klo, khi;           // element support (inclusive); empty when klo > khi
    float inv_wl, inv_wr;   // |1 / W_L|, |1 / W_R| (snapped half-widths)
    bool valid;             // non-empty aperture: the bin contributes
};

// xp = x / pitch, zp = z / (2 pitch) are per-voxel constants; h / pitch = zp / F.
__device__ __forceinline__ Aperture make_aperture(double x, double xp, double zp, double F, const BeamArgs &A,
                                                  int support)
{
    Aperture ap;
    const double hp = zp / F;  // h / pitch with h = z / (2F) (das_pw.m:387); F = 0 gives inf
    const double ua = A.M + (xp - hp);
    const double ub = A.M + (xp + hp);
    const double lbd = fmax(ceil(ua), 0.0);
    const double ubd = fmin(floor(ub), (double)(A.n_el - 1));
    ap.valid = lbd <= ubd;
    ap.klo = 1;
    ap.khi = 0;
    ap.inv_wl = ap.inv_wr = 0.0f;
    if (ap.valid) {
        const int lb = (int)lbd, ubi = (int)ubd;
        // snapped half-widths (das_pw.m:407,411): x - (x_lb - w/2), (x_ub + w/2) - x.
        // Correctly rounded reciprocals: with untruncated windows the normalising
        // sum can nearly cancel and amplifies any error in 1/W.
        const double wl = x - (((double)lb - A.M) * A.pitch - 0.5 * A.width);
        const double wr = (((double)ubi - A.M) * A.pitch + 0.5 * A.width) - x;
        ap.inv_wl = (float)fabs(1.0 / wl);
        ap.inv_wr = (float)fabs(1.0 / wr);
        if (support == SUPPORT_OPEN_H) {
            // {k : |x_k - x| / h < 1} = open interval (ua, ub) in element units
            const double lo = fmax(floor(ua) + 1.0, 0.0);
            const double hi = fmin(ceil(ub) - 1.0, (double)(A.n_el - 1));
            if (lo <= hi) { ap.klo = (int)lo; ap.khi = (int)hi; }
        } else if (support == SUPPORT_SNAPPED) {
            ap.klo = lb;
            ap.khi = ubi;
        } else {
            ap.klo = 0;
            ap.khi = A.n_el - 1;
        }
    }
    return ap;
}

// An aperture in three words, for moving it between lanes: klo | valid << 15 | khi << 16
// (0 <= klo, khi < 256), then the two reciprocal half-widths.
struct ApWord {
    uint32_t kk;
    float iwl, iwr;
};

__device__ __forceinline__ ApWord pack_aperture(const Aperture &a)
{
    return ApWord{(uint32_t)a.klo | ((uint32_t)a.valid << 15) | ((uint32_t)a.khi << 16), a.inv_wl, a.inv_wr};
}

__device__ __forceinline__ Aperture unpack_aperture(const ApWord &w)
{
    Aperture a;
    a.klo = (int)(w.kk & 0x7fffu);
    a.khi = (int)(w.kk >> 16);
    a.valid = ((w.kk >> 15) & 1u) != 0u;
    a.inv_wl = w.iwl;
    a.inv_wr = w.iwr;
    return a;
}

// row r (16 lanes, wave-uniform r) of x in every row: v_permlane32_swap with itself
// gives [x0 x1 x0 x1] and [x2 x3 x2 x3] (rows), v_permlane16_swap of the chosen one
// with itself [y0 y0 y2 y2] and [y1 y1 y3 y3]
__device__ __forceinline__ uint32_t row_broadcast(uint32_t x, int r)
{
    const auto s = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    const uint32_t y = (r & 2) ? s[1] : s[0];
    const auto p = __builtin_amdgcn_permlane16_swap(y, y, false, false);
    return (r & 1) ? p[1] : p[0];
}

// the aperture word of row r (lanes 16 r + c) in every row, for lane c of the row:
// on the vector ALU (two lane swaps per word) or through the LDS crossbar (ds_bpermute,
// no LDS storage)
template <bool SWAP>
__device__ __forceinline__ ApWord row_aperture(const ApWord &w, int r, int c)
{
    if constexpr (SWAP) {
        return ApWord{row_broadcast(w.kk, r), __uint_as_float(row_broadcast(__float_as_uint(w.iwl), r)),
                      __uint_as_float(row_broadcast(__float_as_uint(w.iwr), r))};
    } else {
        const int src = ((r << 4) | c) << 2;
        return ApWord{(uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)w.kk),
                      __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(w.iwl))),
                      __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(w.iwr)))};
    }
}

// sum over the 16 lanes of a DPP row, result in every lane of the row:
// quad xor 1, quad xor 2, half-row mirror (quad <-> quad), row mirror (half <-> half)
__device__ __forceinline__ float row_sum16(float v)
{
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, true));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, true));  // row_mirror
    return v;
}

// min / max over the 16 lanes of a DPP row, result in every lane of the row
__device__ __forceinline__ int row_min16(int v)
{
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));
    return v;
}

__device__ __forceinline__ int row_max16(int v)
{
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));
    return v;
}

// The receive window needs no per-element support test on an array of
// n_full elements (no padding): support on every element, or the snapped
// aperture of a window cut at t = 1 when pitch > width / 2 (elements outside a
// valid snapped aperture have t - 1 >= (pitch - width/2) / (h + width/2) > 0;
// the 1e-3 margin keeps that far above float rounding).
inline bool range_free_side(const BeamArgs &A, int support, const WinCoef &w, int n_full)
{
    if (A.n_el != n_full) return false;
    if (support == SUPPORT_ALL) return true;
    return support == SUPPORT_SNAPPED && w.t_cut <= 1.0f && 0.5 * A.width <= 0.999 * A.pitch;
}

inline bool range_free(const BeamArgs &A, int n_full) { return range_free_side(A, A.support_rx, A.wrx, n_full); }

// a cosine window cut at t = 1 whose support test the cut implies (WC_COS_RF):
// raised cosine (A = B = 1/2) with the flat part ending where the taper starts
inline bool cos_rf_window(const BeamArgs &A, int support, const WinCoef &w, int n_full)
{
    return range_free_side(A, support, w, n_full) && w.t_cut == 1.f && w.A == 0.5f && w.B == 0.5f &&
           (w.t_flat == w.t_off || (w.t_flat < 0.f && w.t_off == 0.f)) && w.q_s > 0.f && std::isfinite(w.q_s) &&
           std::isfinite(w.q_o);
}

// WC_COS_RF with the quarter-period slope folded into the inverse half-width:
// ad = |x_k - x|, iw_q = q_s / W on the element's side; the same value as
// window_eval<WC_COS_RF>(ad / W) with one multiplication less per element.
__device__ __forceinline__ float window_rf(float ad, float iw_q, float q_o)
{
    const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fmed3f(fmaf(ad, iw_q, q_o), 0.f, 0.25f));
    return c * c;
}

template <int WC>
__device__ __forceinline__ float window_eval(float t, const WinCoef &w)
{
    if (WC == WC_RANGE) return 1.0f;
    if (WC == WC_COS_RF) {
        const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fmed3f(fmaf(t, w.q_s, w.q_o), 0.f, 0.25f));
        return c * c;
    }
    float v;
    if (WC == WC_LINEAR) {
        v = w.A + w.C * t;
    } else {
        const float u = __builtin_amdgcn_fractf((t - w.t_off) * w.s_half);  // revolutions, reduced to [0, 1)
        v = w.A + w.B * __builtin_amdgcn_cosf(u);
    }
    v = (t <= w.t_flat) ? 1.0f : v;
    v = (t < w.t_cut) ? v : 0.0f;
    return v;
}

// every window form of WinCoef at once (|t| of the argument): flat part, raised
// cosine plus linear term, cut -- the taper of the windowed normalization,
// evaluated in CPU semantics (windows/*.m compute_samples_scalar)
__device__ __forceinline__ float window_any(float t, const WinCoef &w)
{
    const float ta = fabsf(t);
    float u = (ta - w.t_off) * w.s_half;
    u = u - rintf(u);
    float v = fmaf(w.C, ta, w.A) + w.B * __builtin_amdgcn_cosf(u);
    v = (ta <= w.t_flat) ? 1.0f : v;
    return (ta < w.t_cut) ? v : 0.0f;
}

// taper factor of a row: window((i - L/2) / (L/2)); columns without positive
// rows (h = 0) keep their zero weights
__device__ __forceinline__ float taper_at(int row, float h, const WinCoef &w)
{
    return h > 0.f ? window_any(((float)row - h) / h, w) : 0.f;
}

// transmit models of the plane-wave mode
enum { TX_PLANE = 0, TX_VIRTUAL = 1 };

// geometric transmit time of row a at voxel (x, z), without the i_t0 / f_s shift:
// plane wave (das_pw.m:336) (e_x (x - x_ref) + e_z z) / c; virtual source at
// r_vs = (x_vs, z_vs): diverging (z_vs < 0) (|r - r_vs| - |r_vs|) / c, zero where
// the wavefront crosses the array centre; focused (z_vs > 0)
// (|r_vs| +- |r - r_vs|) / c, through the focus at |r_vs| / c (+ below it).
__host__ __device__ __forceinline__ double pw_tx_time(const BeamArgs &A, int a, double x, double z)
{
    const double p0 = A.txp[3 * a], p1 = A.txp[3 * a + 1], p2 = A.txp[3 * a + 2];
    if (A.tx_kind == TX_PLANE) return (p0 * (x - p2) + p1 * z) / A.ctof;
    const double dx = x - p0, dz = z - p1;
    const double r = sqrt(dx * dx + dz * dz);
    return (p1 < 0.0 ? r - p2 : p2 + (z >= p1 ? r : -r)) / A.ctof;
}

// -------------------------------------------------------------------------
// band-limited DFT
// -------------------------------------------------------------------------
// Band-limited DFT on the matrix cores: for every bin n and channel ch,
//   X_n(ch) = sum_t x_ch[t] e^{-j 2 pi n t / N}
// as two real products D[bin][ch] += W[bin][t] R[t][ch] (W = cos, then -sin) on
// v_mfma_f32_16x16x4f32, K = samples.  A block (4 waves) owns 64 bins and one channel
// group of the operand layout -- 16 transmit rows a = 16 A .. 16 A + 15 times 4 receive
// elements k = 4 K4 .. 4 K4 + 3 -- so that each bin's output is one contiguous 64-value
// block ((b MT + A) KI + K4) 64 + (a & 15) + 16 (k & 3).  Wave w holds bins 16 w ..
// 16 w + 15 of the block (MFMA M) and the group's 64 channels as 4 N-tiles (N-tile j =
// element 4 K4 + j, column i = row 16 A + i).  Lane l supplies W[bin l % 16][t0 + l / 16]
// from the LDS twiddle table at the exactly advanced index (n t) mod N, and
// R[t0 + l / 16][channel] from the RF tile staged in LDS.  Rows and elements of the
// padded operand layout get zeros.  Block grid: (ceil(n_f / 64), MT KI).
template <typename T>
__global__ __launch_bounds__(256) void spectra_kernel(const T *__restrict__ rf, const float2 *__restrict__ twiddle,
                                                           float2 *__restrict__ X, SpecArgs S)
{
    extern __shared__ float smem[];
    float2 *tw = reinterpret_cast<float2 *>(smem);  // [N]
    float *xs = smem + 2 * S.N;                     // [64 channels][65]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    const int grp = blockIdx.y, A = grp / S.KI, K4 = grp - A * S.KI;
    const int b0 = blockIdx.x * 64;
    for (int m = tid; m < S.N; m += 256) tw[m] = twiddle[m];

    const int b_l = b0 + 16 * wave + r16;                  // this lane's bin (A operand row)
    const int n = (b_l < S.n_f) ? S.n_lb + b_l : 0;        // bins past the band are computed as n = 0 and dropped
    const int dm = (int)((4LL * n) % S.N);                 // index step per K step (4 samples)
    floatx4 accR[4], accI[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        accR[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        accI[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    // RF chunks of 64 samples x 64 channels: thread tid holds samples tid & 63 of channels
    // (tid >> 6) + 4 i, i = 0..15; the next chunk is loaded into registers while the
    // current one is contracted
    // (channel cl = (tid >> 6) + 4 i is row 16 A + (cl & 15), element 4 K4 + (cl >> 4), so the
    // per-i offsets from the thread's base are wave-uniform)
    const int tl_own = tid & 63, cl_own = tid >> 6;
    const int a_own = 16 * A + cl_own;
    // RF row of output rows a_own + 4 u (ensembles: angle-major output rows over frame
    // slots, the slots past the last frame read nothing)
    size_t roff[4];
    bool rok[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int R = a_own + 4 * u;
        int in = R;
        rok[u] = R < S.n_tx;
        if (S.ens_slots > 0) {
            const int slot = R % S.ens_slots;
            in = slot * S.ens_angles + R / S.ens_slots;
            rok[u] = rok[u] && slot < S.ens_frames;
        }
        roff[u] = rok[u] ? ((size_t)in * S.n_el + 4 * K4) * S.n_t + tl_own : 0;
    }
    float pre[16];
    auto load_chunk = [&](int t0) {
        const bool tok = t0 + tl_own < S.n_t;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int ke = i >> 2;
            const bool ok = tok && rok[i & 3] && 4 * K4 + ke < S.n_el;
            pre[i] = ok ? (float)rf[roff[i & 3] + (size_t)ke * S.n_t + t0] : 0.f;
        }
    };
    load_chunk(0);
    for (int t0 = 0; t0 < S.n_t; t0 += 64) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) xs[(cl_own + 4 * i) * 65 + tl_own] = pre[i];
        __syncthreads();
        if (t0 + 64 < S.n_t) load_chunk(t0 + 64);
        int m = (int)(((int64_t)n * (t0 + q)) % S.N);
        const int n_steps = (min(64, S.n_t - t0) + 3) >> 2;  // samples past n_t are staged as zeros
        for (int st = 0; st < n_steps; ++st) {
            const float2 w = tw[m];
            m += dm;
            if (m >= S.N) m -= S.N;
            const int tl = 4 * st + q;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float xv = xs[(16 * j + r16) * 65 + tl];
                accR[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(w.x, xv, accR[j], 0, 0, 0);
                accI[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-w.y, xv, accI[j], 0, 0, 0);
            }
        }
    }
    // lane holds D[bin 16 wave + 4 q + r][column r16] of N-tile j
    const int a_row = 16 * A + r16, g_row = S.row0 + a_row;  // row in this RF, row in the output
    const bool put = a_row < S.n_tx || S.write_pad;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int b = b0 + 16 * wave + 4 * q + r;
        if (b >= S.n_f || !put) continue;
        float2 *dst = X + (((size_t)b * S.MT_out + (g_row >> 4)) * S.KI + K4) * 64 + (g_row & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) dst[16 * j] = make_float2(accR[j][r], accI[j][r]);
    }
}

// -------------------------------------------------------------------------
// fused beamforming
// -------------------------------------------------------------------------
// block: 256 threads = 4 waves; wave = 16 voxels (MFMA N); lane group g = lane/16.
// M rows of the MFMA are transmit indices (angles for PW, transmitting elements
// for SA), 16 per M-tile; lane l holds D[4g + r][l & 15] = Y(tx = 16 mt + 4g + r,
// voxel l & 15).  K is the receive element:
//   FP32  (v_mfma_f32_16x16x4f32):      lane group g supplies element 4j + g at step j
//   16-bit (v_mfma_f32_16x16x32_{bf16,f16}): lane group g supplies elements
//          32c + 8g + s, s = 0..7, of chunk c (8 packed values per lane)
// Both give every lane KI elements j = 0..KI-1 (kelem()).
enum { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP16X3 = 2, PREC_FP16 = 3 };

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Prec {
    static constexpr int planes = (PREC == PREC_FP16) ? 2 : 4;  // Xr, Xi (+ lo parts)
    static constexpr bool split = (PREC == PREC_BF16X3 || PREC == PREC_FP16X3);
};

template <int PREC>
__device__ __forceinline__ int kelem(int j, int g)
{
    return PREC == PREC_FP32 ? 4 * j + g : 32 * (j >> 3) + 8 * g + (j & 7);
}

// bytes of one bin of staged spectra per transmit M-tile and element step
template <int PREC>
__host__ __device__ constexpr int stage_bytes_per_tile_step()
{
    return PREC == PREC_FP32 ? 64 * 8 : Prec<PREC>::planes * 64 * 2;
}

template <int PREC>
__device__ __forceinline__ void pack8(const float *v, u32x4 &hi, u32x4 &lo)
{
    if (PREC == PREC_BF16X3) {
        // high parts rounded in pairs (one packed conversion), read back as floats from
        // the packed word (a bf16 is the upper half of its float), low parts from the
        // residuals
        bf16x8 l;
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
            const bf16x2 hp = bf16x2{(__bf16)v[s], (__bf16)v[s + 1]};
            const uint32_t hw = __builtin_bit_cast(uint32_t, hp);
            hi[s / 2] = hw;
            l[s] = (__bf16)(v[s] - __uint_as_float(hw << 16));
            l[s + 1] = (__bf16)(v[s + 1] - __uint_as_float(hw & 0xffff0000u));
        }
        lo = __builtin_bit_cast(u32x4, l);
    } else {
        f16x8 h, l;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            h[s] = (_Float16)v[s];
            if (PREC == PREC_FP16X3) l[s] = (_Float16)(v[s] - (float)h[s]);
        }
        hi = __builtin_bit_cast(u32x4, h);
        if (PREC == PREC_FP16X3) lo = __builtin_bit_cast(u32x4, l);
    }
}

template <int PREC>
__device__ __forceinline__ floatx4 mfma16(const u32x4 &a, const u32x4 &b, const floatx4 &c)
{
    if (PREC == PREC_BF16X3)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ u32x4 neg16(const u32x4 &v)
{
    const unsigned m = 0x80008000u;
    return u32x4{v[0] ^ m, v[1] ^ m, v[2] ^ m, v[3] ^ m};
}

// Register budget from the LDS footprint: tiles of up to 80 KiB let two blocks
// share a CU, so registers are capped at two waves per SIMD; small tiles leave
// the compiler free, larger ones allow one block per CU only.
template <int MT, int KI>
struct WavesPerEU {
    static constexpr int bytes = MT * KI * 64 * 8;
    static constexpr int value = (MT <= 2) ? 1 : (bytes <= 81920 ? 2 : 1);
};

// One-launch-per-plan kernel: the smallest tiles (MT x KI <= 32) are held to 3
// waves per SIMD where they fit without spilling (FP32, fp16, and the split forms
// with the cosine windows WC_COS_ALL / WC_COS_RF; not the ensemble epilogue);
// otherwise WavesPerEU.
template <int MT, int KI, int WCR, int PREC, bool ENS>
struct BeamWavesPerEU {
    static constexpr bool fits3 =
        !ENS && (PREC == PREC_FP32 || PREC == PREC_FP16 || WCR == WC_COS_ALL || WCR == WC_COS_RF);
    static constexpr int value = (MT * KI <= 32 && fits3) ? 3 : WavesPerEU<MT, KI>::value;
};

// FP32 staging: a bin's spectra tile [MT][KI][64] float2 is held as STAGE_NC
// chunks of KI / STAGE_NC element steps, each its own LDS object ([MT][KI / STAGE_NC][64]),
// so that a copy into one chunk (LDS-DMA) and reads of another are disjoint for the
// compiler: its wait before a read covers only the copies into the chunk read.
constexpr int STAGE_NC = 4;

// s_waitcnt immediate (gfx9 encoding) that waits for at most n outstanding vector
// memory operations and leaves the export and LDS/scalar counters unconstrained
constexpr int waitcnt_vm(int n) { return (n & 0xf) | (((n >> 4) & 0x3) << 14) | (0x7 << 4) | (0xf << 8); }

template <int MT, int KI, int C>
__device__ __forceinline__ float2 *stage_chunk()
{
    __shared__ __align__(16) float2 buf[MT * (KI / STAGE_NC) * 64];
    return buf;
}

// 16-bit forms (KI = 32): chunk c holds [plane][M-tile][64] 16-byte groups of eight halves
template <int PREC, int MT, int C>
__device__ __forceinline__ u32x4 *stage_chunk_lp()
{
    __shared__ __align__(16) u32x4 buf[Prec<PREC>::planes * MT * 64];
    return buf;
}

template <int PREC, int MT>
__device__ __forceinline__ u32x4 *stage_chunk_lp_at(int c)
{
    switch (c) {
    case 0: return stage_chunk_lp<PREC, MT, 0>();
    case 1: return stage_chunk_lp<PREC, MT, 1>();
    case 2: return stage_chunk_lp<PREC, MT, 2>();
    default: return stage_chunk_lp<PREC, MT, 3>();
    }
}

template <int MT, int KI>
__device__ __forceinline__ float2 *stage_chunk_at(int c)
{
    switch (c) {
    case 0: return stage_chunk<MT, KI, 0>();
    case 1: return stage_chunk<MT, KI, 1>();
    case 2: return stage_chunk<MT, KI, 2>();
    default: return stage_chunk<MT, KI, 3>();
    }
}

// Occupancy target of the persistent batch kernel.  The claim loop around the
// tile costs registers; left alone the compiler settles small plane-wave tiles
// (MT x KI <= 64, where LDS admits 3+ workgroups per CU) at 2 waves per SIMD
// where the one-launch-per-plan kernel reaches 3.  A target of 3 is stated where
// the tile still fits without spilling: one transmit tile in the 16-bit forms (fp16,
// and the split forms with the cosine windows WC_COS_ALL / WC_COS_RF).
template <int MODE, int MT, int KI, int WCR, int PREC>
struct PersistWavesPerEU {
    static constexpr bool fits3 =
        MT == 1 && (PREC == PREC_FP16 || (PREC != PREC_FP32 && (WCR == WC_COS_ALL || WCR == WC_COS_RF)));
    static constexpr int value = (MODE == MODE_PW && MT <= 3 && MT * KI <= 64 && fits3) ? 3 : WavesPerEU<MT, KI>::value;
};

// TAN (plane waves, FP32): also accumulates the tangent dI/dc_tof.  Geometric
// delays scale as 1/c, so d/dc e^{j w tau} = j w (-(tau_tx,g + tau_rx,g) / c) e^{j w tau}:
// a second contraction with B' = B tau_rx,g and the row factor tau_tx,g.
// ENS (plane waves): the transmit rows are (angle, frame slot) pairs of an ensemble
// sharing one geometry, angle-major; every row keeps its own image accumulator over
// the bins and is added to its frame's image at the end (frame = row % ens_slots when
// below ens_frames, or row_frame[row] for a row stack of acquisitions with their own
// images: a stack's rows are distinct angles, each row with its own transmit phase).
// NW: windowed normalization (FP32, one bin chunk per block so that every lane
// counts its voxel's valid rows from the first bin).
// One workgroup's share of a plan: voxel tile bx (64 voxels), bin chunk by,
// transmit group bz.  beamform_kernel runs it for its block index; the
// persistent batch kernel for every work item it claims.
template <int MODE, int MT, int KI, int WCR, int WCT, int PREC, bool TAN = false, bool ENS = false, bool NW = false,
          bool LBL = false, bool STK = false, typename Args = BeamArgs>
__device__ __forceinline__ void beamform_tile(Args A, const int bx, const int by, const int bz, const int tid)
{
    static_assert(!STK || ENS, "row stacks run the ensemble form");
    static_assert(!TAN || (MODE == MODE_PW && PREC == PREC_FP32), "tangent: plane waves, FP32");
    static_assert(!ENS || (MODE == MODE_PW && !TAN), "ensemble: plane waves");
    static_assert(!NW || (PREC == PREC_FP32 && !TAN && !ENS), "windowed normalization: FP32 forward");
    static_assert(!LBL || (MODE == MODE_PW && MT == 1 && !TAN && !ENS && !NW), "line scan: plane-wave forward, one tile");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the wave index is uniform: held in a scalar register, so the addresses of the
    // wave's tile copies are formed on the scalar unit
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, c = lane & 15;
    const long n_vox = (long)A.n_x * A.n_z;
    // line scan: block bz covers depths of grid column bz, the line of transmit row bz
    const long v_loc = (long)bx * 64 + wave * 16 + c;
    const long v = LBL ? (long)bz * A.n_z + v_loc : v_loc;
    const bool vok = LBL ? v_loc < A.n_z : v < n_vox;
    const long vv = vok ? v : (LBL ? (long)bz * A.n_z + A.n_z - 1 : n_vox - 1);
    const int ix = (int)(vv / A.n_z), iz = (int)(vv - (long)ix * A.n_z);
    const double x = A.px[ix], z = A.pz[iz];
    const double xp = x / A.pitch, zp = z / (2.0 * A.pitch);
    const double clamp_z = sqrt(z / (2.88 * A.c0));  // sqrt(f z / (2.88 c)) = sqrt(f) clamp_z (das_pw.m:374)
    const int mt0 = LBL ? (bz >> 4) : bz * A.tx_group;  // first M-tile of this transmit group

    // receive phase increments: tau_rx = sqrt(((x_k-x)/c)^2 + (z/c)^2) (das_pw.m:339-340,427)
    // plus the element's delay offset
    const double zc = z / A.ctof;
    const double t_half = (MODE == MODE_SA) ? 0.5 * A.t0 / A.f_s : 0.0;  // das_sa.m:431
    uint32_t Qrx[KI];
    float Trx[TAN ? KI : 1];  // tangent: geometric receive time (s)
#pragma unroll
    for (int j = 0; j < KI; ++j) {
        const int k = kelem<PREC>(j, g);
        const double lat = (((double)k - A.M) * A.pitch - x) / A.ctof;
        const double tg = sqrt(lat * lat + zc * zc);
        const double tau = tg + t_half + (k < A.n_el ? A.delay[k] : 0.0);
        Qrx[j] = frac32(A.f_s * tau / A.N);
        if (TAN) Trx[TAN ? j : 0] = (float)tg;
    }
    // transmit phase increments for the rows this lane holds
    uint32_t Qtx[MT][4];
    float Ttx[TAN ? MT : 1][4];  // tangent: geometric transmit time (s)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int a = 16 * (mt0 + mt) + 4 * g + r;
            double tau = 0.0, tg = 0.0;
            if (a < A.n_tx) {
                if (MODE == MODE_PW) {
                    // das_pw.m:336 (or a virtual source) + i_t0 / f_s
                    tg = pw_tx_time(A, a, x, z);
                    tau = tg + A.t0 / A.f_s;
                } else {
                    // transmitting element a: its geometric delay and its delay offset
                    const double lat = (((double)a - A.M) * A.pitch - x) / A.ctof;
                    tau = sqrt(lat * lat + zc * zc) + t_half + A.delay[a];
                }
            }
            Qtx[mt][r] = frac32(A.f_s * tau / A.N);
            if (TAN) Ttx[TAN ? mt : 0][r] = (float)tg;
        }
    // line scan: the phase increment of the block's own transmit row
    const uint32_t Qln = LBL ? frac32(A.f_s * (pw_tx_time(A, bz, x, z) + A.t0 / A.f_s) / A.N) : 0u;
    const int k0 = kelem<PREC>(0, g);
    const float dg = (float)(((double)k0 - A.M) * A.pitch - x);  // x_k - x of element j = 0
    const float pf = (float)A.pitch;
    const float dg_tx = (float)((4.0 * g - A.M) * A.pitch - x);  // SA transmit rows, mt = r = 0

    float Ir = 0.f, Ii = 0.f, w_last = 0.f, wtx_last = 0.f;
    int nw_row = 0;  // windowed normalization: valid rows of this voxel before the current bin
    float Dr = 0.f, Di = 0.f;  // tangent accumulators
    float Er[ENS ? MT : 1][4], Ei[ENS ? MT : 1][4];  // ensemble: per-row image accumulators
    if (ENS) {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) Er[ENS ? mt : 0][r] = Ei[ENS ? mt : 0][r] = 0.f;
    }
    // w_n = n * 2 pi f_s / N; the tangent factor -w_n / c_tof is applied per bin
    const float w_unit = (float)(2.0 * 3.141592653589793 * A.f_s / A.N);
    const float inv_c = (float)(1.0 / A.ctof);
    const int b_begin = by * A.bin_chunk;
    const int b_end = min(b_begin + A.bin_chunk, A.n_f);
    const int MT_all = (A.n_tx + 15) >> 4;  // stride of the operand layout
    // staged bytes per bin: this group's tiles of every plane
    constexpr int tile_bytes = stage_bytes_per_tile_step<PREC>() * KI;  // one M-tile, all elements
    constexpr int n_f4 = MT * tile_bytes / 16;
    const int mt_here = min(MT, MT_all - mt0);

    // Staging pipeline: chunk c of bin b is copied by LDS-DMA (global_load_lds_dwordx4,
    // 1 KB per wave instruction: two element steps of one M-tile in FP32, one plane of
    // one M-tile over 32 elements in the 16-bit forms) STAGE_NC - 1 chunks ahead of its
    // contraction, into the chunk array the block finished last, so the next bin's
    // tile streams in during this bin's contraction.  A chunk is awaited on this wave's
    // own copies (vmcnt) and one barrier.  Blocks whose group holds fewer than MT tiles
    // (zero rows to fill), the windowed form, and tilings whose copies do not divide
    // evenly over the four waves copy synchronously; the 16-bit forms at KI < 32 keep
    // one dynamic LDS tile.
    static_assert(STAGE_NC == 4, "the awaited counts below assume two chunks in flight behind the current one");
    constexpr int SJC = KI / STAGE_NC;  // FP32 element steps per chunk
    constexpr int KC = KI / 8;          // 16-bit: chunks of 32 elements
    constexpr int PL = Prec<PREC>::planes;
    constexpr bool LP_CHUNKED = (PREC != PREC_FP32) && KC == STAGE_NC;
    constexpr bool CHUNKED = (PREC == PREC_FP32) || LP_CHUNKED;
    constexpr int UNITS = (PREC == PREC_FP32) ? MT * KI / 8 : PL * MT;  // 1 KB copies per chunk
    constexpr bool PIPE = CHUNKED && !NW && (PREC != PREC_FP32 || SJC % 2 == 0);
    // 1 KB copies per wave per chunk: PIPE_LMAX for the first UNITS % 4 waves, PIPE_L
    // for the others; the awaited counts use PIPE_L, which no wave exceeds in copies
    // issued after the awaited chunk (so every wait covers at least that chunk)
    constexpr int PIPE_L = UNITS / 4, PIPE_LMAX = (UNITS + 3) / 4;
    const bool pipe = PIPE && mt_here == MT;         // block-uniform
    const int n_chunks = STAGE_NC * (b_end - b_begin);
    auto issue_chunk = [&](auto cq_const, int bq) {
        constexpr int cq = decltype(cq_const)::value;
        if constexpr (PREC == PREC_FP32) {
            const float4 *src = reinterpret_cast<const float4 *>(A.X) + ((size_t)bq * MT_all + mt0) * (tile_bytes / 16);
            float4 *dst = reinterpret_cast<float4 *>(stage_chunk<MT, KI, cq>());
#pragma unroll
            for (int i = 0; i < PIPE_LMAX; ++i) {
                const int u = wave + 4 * i;  // unit: element steps cq SJC + 2 jp, + 1 of M-tile mt
                if (PIPE_LMAX != PIPE_L && u >= UNITS) break;  // wave-uniform
                const int mt = u / (SJC / 2), jp = u - mt * (SJC / 2);
                __builtin_amdgcn_global_load_lds(
                    (__attribute__((address_space(1))) const void *)(src + (mt * KI + cq * SJC + 2 * jp) * 32 + lane),
                    (__attribute__((address_space(3))) void *)(dst + (mt * SJC + 2 * jp) * 32), 16, 0, 0);
            }
        } else if constexpr (LP_CHUNKED) {
            // planes [p][mt][c][64] of 16-byte groups per bin; unit (p, mt) of chunk cq
            const u32x4 *src = reinterpret_cast<const u32x4 *>(A.Xlp) + (size_t)bq * PL * MT_all * KC * 64;
            u32x4 *dst = stage_chunk_lp<PREC, MT, cq>();
#pragma unroll
            for (int i = 0; i < PIPE_LMAX; ++i) {
                const int u = wave + 4 * i;
                if (PIPE_LMAX != PIPE_L && u >= UNITS) break;  // wave-uniform
                const int pp = u / MT, mt = u - pp * MT;
                __builtin_amdgcn_global_load_lds(
                    (__attribute__((address_space(1))) const void *)(src + ((pp * MT_all + mt0 + mt) * KC + cq) * 64 + lane),
                    (__attribute__((address_space(3))) void *)(dst + (pp * MT + mt) * 64), 16, 0, 0);
            }
        }
    };
    if (pipe) {
        if (n_chunks > 0) issue_chunk(std::integral_constant<int, 0>{}, b_begin);
        if (n_chunks > 1) issue_chunk(std::integral_constant<int, 1>{}, b_begin);
        if (n_chunks > 2) issue_chunk(std::integral_constant<int, 2>{}, b_begin);
    }
    // top of a chunk's contraction: await chunk q of the sequence, then hand the
    // positions of chunk q - 1 to chunk q + 3 (the last chunk of a bin hands them
    // over after its contraction, see below)
    auto chunk_begin = [&](auto cc_const, int b) {
        constexpr int cc = decltype(cc_const)::value;
        const int q = STAGE_NC * (b - b_begin) + cc;
        // copies issued after chunk q's: those of chunks q + 1, q + 2 (when they exist)
        const int later = min(2, n_chunks - 1 - q);
        if (later >= 2) __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * PIPE_L));
        else if (later == 1) __builtin_amdgcn_s_waitcnt(waitcnt_vm(PIPE_L));
        else __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
        // chunk q landed for every wave; every wave is done with chunk q - 1.  A bare
        // barrier: the wait above is the only one needed (no release fence, which
        // would also wait for the copies still in flight); the empty asm keeps the
        // compiler from moving LDS accesses across it
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // chunk q + 3 goes where chunk q - 1 was: chunk (cc + 3) % 4 of bin b + (cc + 3) / 4;
        // the last chunk's copy is issued after its contraction, so that the compiler's
        // own wait before reading the last chunk -- whose copy it tracks from the start
        // of the bin -- does not also wait for it
        if (cc != STAGE_NC - 1 && q + STAGE_NC - 1 < n_chunks)
            issue_chunk(std::integral_constant<int, (cc + STAGE_NC - 1) % STAGE_NC>{}, b + (cc + STAGE_NC - 1) / STAGE_NC);
    };
    auto chunk_end = [&](auto cc_const, int b) {
        constexpr int cc = decltype(cc_const)::value;
        if (cc == STAGE_NC - 1 && STAGE_NC * (b - b_begin) + cc + STAGE_NC - 1 < n_chunks)
            issue_chunk(std::integral_constant<int, (cc + STAGE_NC - 1) % STAGE_NC>{}, b + (cc + STAGE_NC - 1) / STAGE_NC);
    };

    // Receive apertures (FP64 bounds, a division and two reciprocals per voxel and
    // bin).  The four lane groups g hold the same 16 voxels, so for plane waves group g
    // computes the aperture of bin b0 + g once per four bins and each bin takes its own
    // from group (b - b0) & 3: bitwise the same apertures.  The FP32 instances of three
    // or more tiles move it by lane swaps, the others by a lane permute (each measured
    // the faster for them).  The synthetic aperture's FP32 instances do the same for
    // their transmit apertures; its 16-bit forms measured slower with it (and the split
    // ones and the windowed normalization have no registers to spare for it).
    constexpr bool QUAD = MODE == MODE_PW || (MODE == MODE_SA && !NW && PREC == PREC_FP32);
    constexpr bool SWAP = PREC == PREC_FP32 && MT >= 3;
    // The plane-wave FP32 instances also take the four bins' receive support unions over
    // the wave (for the element-step skip) from the group that computed each aperture,
    // once per four bins (0.6-2 % faster); the synthetic aperture and the 16-bit forms
    // measured no faster or slower with it, and the tangent instances have no registers
    // to spare for it.
    constexpr bool QUNION = QUAD && MODE == MODE_PW && PREC == PREC_FP32 && WCR != WC_COS_ALL && !TAN;
    const int b_last = b_end - 1;
    auto quad_rx = [&](int b0) {
        const int bq = min(b0 + g, b_last);
        double Fr = A.Frx[bq];
        if (A.clamp) Fr = fmin(Fr, A.sqrt_f[bq] * clamp_z);  // das_pw.m:374-377
        return pack_aperture(make_aperture(x, xp, zp, Fr, A, A.support_rx));
    };
    auto quad_tx = [&](int b0) {  // das_sa.m:384-400
        const int bq = min(b0 + g, b_last);
        return pack_aperture(make_aperture(x, xp, zp, A.Ftx[bq], A, A.support_tx));
    };
    ApWord q_rx{0u, 0.f, 0.f}, q_tx{0u, 0.f, 0.f};  // this lane's bin of the current four
    int q_ulo = 0, q_uhi = 4 * KI - 1;  // lane group g: the receive support union of bin b0 + g

    for (int b = b_begin; b < b_end; ++b) {
        const int n = A.n_lb + b;
        const int qb = (b - b_begin) & 3;
        if (QUAD && qb == 0) {
            q_rx = quad_rx(b);
            if (MODE == MODE_SA) q_tx = quad_tx(b);
            if (QUNION) {
                // support unions of the four bins over the wave's 16 voxels, one per lane group
                const Aperture aq = unpack_aperture(q_rx);
                const bool ne = aq.klo <= aq.khi;
                q_ulo = row_min16(ne ? aq.klo : 0x7fff);
                q_uhi = row_max16(ne ? aq.khi : -1);
            }
        }
        double Fr = QUAD ? 0.0 : A.Frx[b];
        if (!QUAD && A.clamp) Fr = fmin(Fr, A.sqrt_f[b] * clamp_z);  // das_pw.m:374-377
        const Aperture ap = QUAD ? unpack_aperture(row_aperture<SWAP>(q_rx, qb, c))
                                 : make_aperture(x, xp, zp, Fr, A, A.support_rx);
        if (CHUNKED && !pipe) {
            __syncthreads();
            if constexpr (PREC == PREC_FP32) {
                const float4 *src = reinterpret_cast<const float4 *>(A.X) + ((size_t)b * MT_all + mt0) * (tile_bytes / 16);
                const int n4 = mt_here * (tile_bytes / 16);
                for (int e = tid; e < n_f4; e += 256) {
                    const int mt = e / (KI * 32), rem = e - mt * (KI * 32), j = rem >> 5;
                    float4 *dst = reinterpret_cast<float4 *>(stage_chunk_at<MT, KI>(j / SJC));
                    dst[(mt * SJC + j % SJC) * 32 + (rem & 31)] = (e < n4) ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
                }
            } else if constexpr (LP_CHUNKED) {
                constexpr int tile4 = KC * 64;  // 16-byte groups per tile per plane
                const u32x4 *src = reinterpret_cast<const u32x4 *>(A.Xlp) + (size_t)b * PL * MT_all * tile4;
                for (int e = tid; e < PL * MT * tile4; e += 256) {
                    const int pp = e / (MT * tile4), rem = e - pp * (MT * tile4);
                    const int mt = rem / tile4, r2 = rem - mt * tile4, cq = r2 >> 6;
                    u32x4 *dst = stage_chunk_lp_at<PREC, MT>(cq);
                    dst[(pp * MT + mt) * 64 + (r2 & 63)] =
                        (mt < mt_here) ? src[((size_t)pp * MT_all + mt0 + mt) * tile4 + r2] : u32x4{0u, 0u, 0u, 0u};
                }
            }
            __syncthreads();
        }
        if (!CHUNKED) {
            // 16-bit planes [p][mt][c][64][8] per bin in one dynamic LDS tile
            __syncthreads();
            float4 *dst = reinterpret_cast<float4 *>(smem_raw);
            constexpr int tile4 = tile_bytes / PL / 16;  // float4 per tile per plane
            const float4 *src = reinterpret_cast<const float4 *>(A.Xlp) + (size_t)b * PL * MT_all * tile4;
            for (int e = tid; e < n_f4; e += 256) {
                const int pp = e / (MT * tile4), rem = e - pp * (MT * tile4);
                const int mt = rem / tile4;
                dst[e] = (mt < mt_here) ? src[((size_t)pp * MT_all + mt0 + mt) * tile4 + (rem - mt * tile4)]
                                        : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            __syncthreads();
        }

        // elements [u_lo, u_hi] meeting the support of some voxel of this wave
        // (wave-uniform): element steps outside carry zero window for all 16 voxels
        int u_lo = 0, u_hi = 4 * KI - 1;
        if (WCR != WC_COS_ALL) {
            if (QUNION) {
                u_lo = __builtin_amdgcn_readlane(q_ulo, 16 * qb);
                u_hi = __builtin_amdgcn_readlane(q_uhi, 16 * qb);  // -1: none
            } else {
                const bool ne = ap.klo <= ap.khi;
                u_lo = __builtin_amdgcn_readfirstlane(row_min16(ne ? ap.klo : 0x7fff));
                u_hi = __builtin_amdgcn_readfirstlane(row_max16(ne ? ap.khi : -1));  // -1: none
            }
        }
        constexpr bool RTEST = WCR != WC_COS_ALL && WCR != WC_COS_RF;  // per-element support test

        floatx4 accR[MT], accI[MT];
        floatx4 tanR[TAN ? MT : 1], tanI[TAN ? MT : 1];
        float asum = 0.f;
        // recompute x_k - x per bin (one FMA each) instead of holding KI registers
        float dgb = dg;
        asm volatile("" : "+v"(dgb));
        if constexpr (PREC == PREC_FP32) {
            // Gauss three-multiplication complex product on the matrix cores:
            //   K1 = sum_k Br (Xr + Xi), K2 = sum_k Xr (Bi - Br), K3 = sum_k Xi (Br + Bi)
            //   Re Y = K1 - K3, Im Y = K1 + K2
            floatx4 K1[MT], K2[MT], K3[MT];
            floatx4 T1[TAN ? MT : 1], T2[TAN ? MT : 1], T3[TAN ? MT : 1];
            // line scan: one transmit row contributes, so the product runs on the vector ALUs
            // (the lane's elements 4 j + g against row bz of the staged tile, read as a
            // broadcast) instead of a 16-row matrix-core product of which one row is used
            float Lr = 0.f, Li = 0.f;
            const int lrow = bz & 15;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                K1[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                K2[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                K3[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                if (TAN) {
                    T1[TAN ? mt : 0] = floatx4{0.f, 0.f, 0.f, 0.f};
                    T2[TAN ? mt : 0] = floatx4{0.f, 0.f, 0.f, 0.f};
                    T3[TAN ? mt : 0] = floatx4{0.f, 0.f, 0.f, 0.f};
                }
            }
            const int klo = ap.klo - g, khi = ap.khi - g;  // support in units of 4j
            const float dstep = 4.0f * pf;
            const float ul = ap.inv_wl * A.wrx.s_half, ur = ap.inv_wr * A.wrx.s_half;
            const float ql = ap.inv_wl * A.wrx.q_s, qr = ap.inv_wr * A.wrx.q_s;  // WC_COS_RF slopes
            // WC_COS_ALL: the cosine argument |d| / (2 W_side) of element step j without a
            // per-element side select: max(d ur, -d ul) with d = dgb + j dstep (each term is
            // positive exactly on its own side)
            const float ar0 = dgb * ur, ar1 = dstep * ur, al0 = -dgb * ul, al1 = -dstep * ul;
            // the other classes the same way: |d| / W_side = max(d / W_R, -d / W_L), and for
            // WC_COS_RF the cosine argument |d| q_s / W_side + q_o
            const float qo = WCR == WC_COS_RF ? A.wrx.q_o : 0.f;
            const float sr = WCR == WC_COS_RF ? qr : ap.inv_wr, sl = WCR == WC_COS_RF ? ql : ap.inv_wl;
            const float tr0 = fmaf(dgb, sr, qo), tr1 = dstep * sr, tl0 = fmaf(-dgb, sl, qo), tl1 = -dstep * sl;
            const int j_hi = min(u_hi >> 2, KI - 1);  // elements 4j + g, g = 0..3
            auto chunk_step = [&](auto cc_const) {
            constexpr int cc = decltype(cc_const)::value;
            if (pipe) chunk_begin(cc_const, b);
            const float2 *xs = stage_chunk<MT, KI, cc>();
            // one element step: elements 4 j + g of this lane, against the chunk's tiles
            auto step = [&](const int j) __attribute__((always_inline)) {
                const float jf = (float)j;
                float a;
                if (WCR == WC_COS_ALL) {
                    const float u = __builtin_amdgcn_fractf(fmaxf(fmaf(jf, ar1, ar0), fmaf(jf, al1, al0)));
                    a = fmaf(A.wrx.B, __builtin_amdgcn_cosf(u), A.wrx.A);
                } else {
                    // (wide arrays keep the per-element side select: fewer live registers)
                    const float d = fmaf(jf, dstep, dgb);
                    const float t = KI <= 32 ? fmaxf(fmaf(jf, tr1, tr0), fmaf(jf, tl1, tl0))
                                             : fmaf(fabsf(d), d < 0.f ? sl : sr, qo);
                    if (WCR == WC_COS_RF) {
                        const float c = __builtin_amdgcn_cosf(__builtin_amdgcn_fmed3f(t, 0.f, 0.25f));
                        a = c * c;
                    } else {
                        a = window_eval<WCR>(t, A.wrx);
                        if (RTEST) a = (4 * j >= klo && 4 * j <= khi) ? a : 0.f;
                    }
                }
                const float rv = rev_of((uint32_t)n * Qrx[j]);
                const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                asum += a;  // the normalizing sum is taken before the taper (on, then windowed)
                if (NW && A.nw_rx) {  // null when only the transmit side is windowed (SA)
                    const int k = 4 * j + g;
                    a *= k < A.n_el ? taper_at(nw_row, A.nw_rx[vv * A.n_el + k], A.wn_rx) : 0.f;
                }
                const float br = a * cs, bi = a * sn;
                const float bd = bi - br, bp = br + bi;
                if constexpr (LBL) {
                    const float2 xv = xs[(j - cc * SJC) * 64 + lrow + 16 * g];
                    Lr = fmaf(xv.x, br, Lr);
                    Lr = fmaf(-xv.y, bi, Lr);
                    Li = fmaf(xv.x, bi, Li);
                    Li = fmaf(xv.y, br, Li);
                    return;
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const float2 xv = xs[(mt * SJC + (j - cc * SJC)) * 64 + lane];
                    const float xsum = xv.x + xv.y;
                    K1[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xsum, br, K1[mt], 0, 0, 0);
                    K2[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.x, bd, K2[mt], 0, 0, 0);
                    K3[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, bp, K3[mt], 0, 0, 0);
                    if (TAN) {
                        const float tr = Trx[TAN ? j : 0];
                        T1[TAN ? mt : 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xsum, br * tr, T1[TAN ? mt : 0], 0, 0, 0);
                        T2[TAN ? mt : 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.x, bd * tr, T2[TAN ? mt : 0], 0, 0, 0);
                        T3[TAN ? mt : 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, bp * tr, T3[TAN ? mt : 0], 0, 0, 0);
                    }
                }
                        };
            const int j_a = max(u_lo >> 2, cc * SJC), j_b = min(j_hi, cc * SJC + SJC - 1);
            if constexpr (WCR == WC_COS_ALL && !LBL && !TAN && !NW) {
                // every element step (no support test): unrolled, so the phase words are
                // read from fixed registers and the first step starts the accumulators
#pragma unroll
                for (int jj = 0; jj < SJC; ++jj) step(cc * SJC + jj);
            } else if constexpr (KI > 32 && !LBL) {
                // wide arrays: every step unrolled under its own uniform guard, so that the
                // 48 / 64 phase words stay in registers (a runtime index would move them to
                // scratch memory)
#pragma unroll
                for (int jj = 0; jj < SJC; ++jj) {
                    const int j = cc * SJC + jj;
                    if (j >= j_a && j <= j_b) step(j);
                }
            } else if constexpr (LBL || MT > 2) {
                // (three or more tiles: a step's 9+ products keep the matrix cores busy)
#pragma clang loop unroll(disable)
                for (int j = j_a; j <= j_b; ++j) step(j);
            } else {
                // one or two tiles: pairs of element steps, unrolled -- phase words in fixed
                // registers and two independent chains for the vector ALUs.  A pair that
                // meets [j_a, j_b] runs whole; its other step lies outside the support of
                // every voxel of the wave, where the support test (or, for WC_COS_RF, the
                // window's cut) gives zero weight.
                static_assert(SJC % 2 == 0, "element steps per chunk come in pairs");
#pragma unroll
                for (int pp = 0; pp < SJC / 2; ++pp) {
                    const int j0 = cc * SJC + 2 * pp;
                    if (j0 + 1 >= j_a && j0 <= j_b) {
                        step(j0);
                        step(j0 + 1);
                    }
                }
            }
            if (pipe) chunk_end(cc_const, b);
            };  // chunk_step
            chunk_step(std::integral_constant<int, 0>{});
            chunk_step(std::integral_constant<int, 1>{});
            chunk_step(std::integral_constant<int, 2>{});
            chunk_step(std::integral_constant<int, 3>{});
            if constexpr (LBL) {
                accR[0] = floatx4{Lr, 0.f, 0.f, 0.f};
                accI[0] = floatx4{Li, 0.f, 0.f, 0.f};
            } else
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                accR[mt] = K1[mt] - K3[mt];
                accI[mt] = K1[mt] + K2[mt];
                if (TAN) {
                    // Y' = sum_k B tau_rx,g X, then Y' += tau_tx,g Y (row factor)
                    tanR[TAN ? mt : 0] = T1[TAN ? mt : 0] - T3[TAN ? mt : 0];
                    tanI[TAN ? mt : 0] = T1[TAN ? mt : 0] + T2[TAN ? mt : 0];
                }
            }
        } else {
            // 16-bit operands, FP32 accumulation: Re Y = Xr Br - Xi Bi, Im Y = Xr Bi + Xi Br;
            // split forms add hi*lo + lo*hi passes (x = hi + lo)
            constexpr int pl = Prec<PREC>::planes;
            // (p, mt) group of chunk cc: per-chunk arrays [p][mt][64] (KI = 32) or the dynamic
            // tile [p][mt][c][64]
            constexpr int XS = LP_CHUNKED ? 64 : KC * 64;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                accR[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
                accI[mt] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
            const float ul = ap.inv_wl * A.wrx.s_half, ur = ap.inv_wr * A.wrx.s_half;
            const float ql = ap.inv_wl * A.wrx.q_s, qr = ap.inv_wr * A.wrx.q_s;  // WC_COS_RF slopes
            // WC_COS_ALL: |d| / (2 W_side) = max(d ur, -d ul), d = dgb + (k - k0) pitch
            const float ar0 = dgb * ur, ar1 = pf * ur, al0 = -dgb * ul, al1 = -pf * ul;
            auto chunk_lp = [&](auto cc_const) {
                constexpr int cc = decltype(cc_const)::value;
                if (pipe) chunk_begin(cc_const, b);
                const u32x4 *xq;
                if constexpr (LP_CHUNKED) xq = stage_chunk_lp<PREC, MT, cc>();
                else xq = reinterpret_cast<const u32x4 *>(smem_raw) + cc * 64;
                if (cc >= (u_lo >> 5) && cc <= (u_hi >> 5)) {  // elements 32 cc .. 32 cc + 31
                float br[8], bi[8];
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const int j = 8 * cc + s;
                    const int k = kelem<PREC>(j, g);
                    const float d = fmaf((float)(k - k0), pf, dgb);
                    float a;
                    if (WCR == WC_COS_ALL) {
                        const float e = (float)(k - k0);
                        const float u = __builtin_amdgcn_fractf(fmaxf(fmaf(e, ar1, ar0), fmaf(e, al1, al0)));
                        a = fmaf(A.wrx.B, __builtin_amdgcn_cosf(u), A.wrx.A);
                    } else if (WCR == WC_COS_RF) {
                        a = window_rf(fabsf(d), d < 0.f ? ql : qr, A.wrx.q_o);
                    } else {
                        const float t = fabsf(d) * (d < 0.f ? ap.inv_wl : ap.inv_wr);
                        a = window_eval<WCR>(t, A.wrx);
                        if (RTEST) a = (k >= ap.klo && k <= ap.khi) ? a : 0.f;
                    }
                    const float rv = rev_of((uint32_t)n * Qrx[j]);
                    const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                    br[s] = a * cs;
                    bi[s] = a * sn;
                    asum += a;
                }
                u32x4 brh, brl, bih, bil;
                pack8<PREC>(br, brh, brl);
                pack8<PREC>(bi, bih, bil);
                const u32x4 nbih = neg16(bih);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const u32x4 xrh = xq[(0 * MT + mt) * XS + lane];
                    const u32x4 xih = xq[(1 * MT + mt) * XS + lane];
                    accR[mt] = mfma16<PREC>(xrh, brh, accR[mt]);
                    accR[mt] = mfma16<PREC>(xih, nbih, accR[mt]);
                    accI[mt] = mfma16<PREC>(xrh, bih, accI[mt]);
                    accI[mt] = mfma16<PREC>(xih, brh, accI[mt]);
                    if (Prec<PREC>::split) {
                        const u32x4 xrl = xq[(2 * MT + mt) * XS + lane];
                        const u32x4 xil = xq[(3 * MT + mt) * XS + lane];
                        const u32x4 nbil = neg16(bil);
                        accR[mt] = mfma16<PREC>(xrh, brl, accR[mt]);
                        accR[mt] = mfma16<PREC>(xrl, brh, accR[mt]);
                        accR[mt] = mfma16<PREC>(xih, nbil, accR[mt]);
                        accR[mt] = mfma16<PREC>(xil, nbih, accR[mt]);
                        accI[mt] = mfma16<PREC>(xrh, bil, accI[mt]);
                        accI[mt] = mfma16<PREC>(xrl, bih, accI[mt]);
                        accI[mt] = mfma16<PREC>(xih, brl, accI[mt]);
                        accI[mt] = mfma16<PREC>(xil, brh, accI[mt]);
                    }
                }
                }  // chunk in the support
                if (pipe) chunk_end(cc_const, b);
            };  // chunk_lp
            // every chunk of 32 elements: KC = 1, 2, 4, 6 or 8
            chunk_lp(std::integral_constant<int, 0>{});
            if constexpr (KC > 1) chunk_lp(std::integral_constant<int, 1>{});
            if constexpr (KC > 2) {
                chunk_lp(std::integral_constant<int, 2>{});
                chunk_lp(std::integral_constant<int, 3>{});
            }
            if constexpr (KC > 4) {
                chunk_lp(std::integral_constant<int, 4>{});
                chunk_lp(std::integral_constant<int, 5>{});
            }
            if constexpr (KC > 6) {
                chunk_lp(std::integral_constant<int, 6>{});
                chunk_lp(std::integral_constant<int, 7>{});
            }
            static_assert(KC == 1 || KC == 2 || KC == 4 || KC == 6 || KC == 8, "chunks of 32 elements");
            (void)pl;
        }
        asum += __shfl_xor(asum, 16);
        asum += __shfl_xor(asum, 32);
        const float xscale = (PREC == PREC_FP32) ? 1.0f : A.xscale[b];  // spectra were stored / xscale

        if (MODE == MODE_PW) {
            if (ap.valid) {
                // v_rcp_f32 (1 ulp) instead of the correctly rounded division's ten instructions
                const float sc = (A.norm_rx ? __builtin_amdgcn_rcpf(asum) : 1.0f) * xscale;
                // padded transmit rows have zero spectra, so every row is summed
                float pr = 0.f, pi = 0.f;  // tangent: sum_rows T (tau_tx,g Y + Y')
                if constexpr (LBL) {
                    // the lane's partial product over its elements, at the line's own transmit phase
                    const float rv = rev_of((uint32_t)n * Qln);
                    const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                    const float yr = accR[0][0] * sc, yi = accI[0][0] * sc;
                    Ir += cs * yr - sn * yi;
                    Ii += cs * yi + sn * yr;
                } else {
                float Sr = 0.f, Si = 0.f;  // sum over the rows of T Y, scaled once below
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    // ensemble: the four rows of this lane in the tile share their angle
                    // (angle-major rows over frame slots), so one transmit phase, scaled by
                    // sc; a row stack's rows (STK) each have their own
                    float sn0 = 0.f, cs0 = 0.f;
                    constexpr bool shared = ENS && !STK;
                    if (shared) {
                        const float rv = rev_of((uint32_t)n * Qtx[mt][0]);
                        sn0 = __builtin_amdgcn_sinf(rv) * sc;
                        cs0 = __builtin_amdgcn_cosf(rv) * sc;
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (ENS) {
                            if (!shared) {
                                const float rv = rev_of((uint32_t)n * Qtx[mt][r]);
                                sn0 = __builtin_amdgcn_sinf(rv) * sc;
                                cs0 = __builtin_amdgcn_cosf(rv) * sc;
                            }
                            Er[ENS ? mt : 0][r] = fmaf(cs0, accR[mt][r], fmaf(-sn0, accI[mt][r], Er[ENS ? mt : 0][r]));
                            Ei[ENS ? mt : 0][r] = fmaf(cs0, accI[mt][r], fmaf(sn0, accR[mt][r], Ei[ENS ? mt : 0][r]));
                            continue;
                        }
                        const float rv = rev_of((uint32_t)n * Qtx[mt][r]);
                        const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                        {
                            Sr = fmaf(cs, accR[mt][r], fmaf(-sn, accI[mt][r], Sr));
                            Si = fmaf(cs, accI[mt][r], fmaf(sn, accR[mt][r], Si));
                        }
                        if (TAN) {
                            const float tt = Ttx[TAN ? mt : 0][r];
                            const float ur = fmaf(tt, accR[mt][r], tanR[TAN ? mt : 0][r]);
                            const float ui = fmaf(tt, accI[mt][r], tanI[TAN ? mt : 0][r]);
                            pr += cs * ur - sn * ui;
                            pi += cs * ui + sn * ur;
                        }
                    }
                }
                if (!ENS) {
                    Ir = fmaf(sc, Sr, Ir);
                    Ii = fmaf(sc, Si, Ii);
                }
                }
                if (TAN) {
                    // dI += sc (j w_n) (-1 / c) P
                    const float f = -sc * (float)n * w_unit * inv_c;
                    Dr = fmaf(-f, pi, Dr);
                    Di = fmaf(f, pr, Di);
                }
            }
            if (NW) nw_row += ap.valid ? 1 : 0;
        } else {
            // transmit aperture and apodization (das_sa.m:384-400, 438-455)
            const Aperture at = QUAD ? unpack_aperture(row_aperture<SWAP>(q_tx, qb, c))
                                     : make_aperture(x, xp, zp, A.Ftx[b], A, A.support_tx);
            float wtx[MT][4];
            float tsum = 0.f;
            // WC_COS_RF: cos^2 of the quarter-period argument max(k r1 + r0, k l1 + l0) for
            // transmitter offset k = 16 t + r from this lane's first row (each side's term
            // is the larger exactly on its own side); the window's cut is the support test
            const float sr = at.inv_wr * A.wtx.q_s, sl = at.inv_wl * A.wtx.q_s;
            const float tr1 = pf * sr, tl1 = -pf * sl;
            const float tr0 = fmaf(dg_tx, sr, A.wtx.q_o), tl0 = fmaf(-dg_tx, sl, A.wtx.q_o);
            auto tx_weight = [&](int t, int r) __attribute__((always_inline)) {
                const int i = 16 * t + 4 * g + r;
                const float kf = (float)(16 * t + r);
                if (WCT == WC_COS_RF) {
                    const float c = __builtin_amdgcn_cosf(
                        __builtin_amdgcn_fmed3f(fmaxf(fmaf(kf, tr1, tr0), fmaf(kf, tl1, tl0)), 0.f, 0.25f));
                    return c * c;
                }
                const float d = fmaf(kf, pf, dg_tx);
                const float tt = fabsf(d) * (d < 0.f ? at.inv_wl : at.inv_wr);
                const float w = window_eval<WCT>(tt, A.wtx);
                return (i >= at.klo && i <= at.khi) ? w : 0.f;
            };
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float w = tx_weight(mt0 + mt, r);
                    wtx[mt][r] = w;
                    tsum += w;
                }
            if (MT_all > MT) {
                // transmit rows split over groups (arrays beyond 128 elements): the
                // normalization sums the window over every transmitter, not only this
                // group's (das_sa.m:479)
                tsum = 0.f;
                for (int t = 0; t < MT_all; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) tsum += tx_weight(t, r);
            }
            tsum += __shfl_xor(tsum, 16);
            tsum += __shfl_xor(tsum, 32);
            if (NW && A.nw_tx) {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = 16 * (mt0 + mt) + 4 * g + r;
                        wtx[mt][r] *= i < A.n_el ? taper_at(nw_row, A.nw_tx[vv * A.n_el + i], A.wn_tx) : 0.f;
                    }
            }
            if (ap.valid && at.valid) {
                const float sc = (A.norm_rx ? __builtin_amdgcn_rcpf(asum) : 1.0f) *
                                 (A.norm_tx ? __builtin_amdgcn_rcpf(tsum) : 1.0f) * xscale;
                float pr = 0.f, pi = 0.f;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float rv = rev_of((uint32_t)n * Qtx[mt][r]);
                        const float sn = __builtin_amdgcn_sinf(rv), cs = __builtin_amdgcn_cosf(rv);
                        const float yr = accR[mt][r] * wtx[mt][r], yi = accI[mt][r] * wtx[mt][r];
                        pr += cs * yr - sn * yi;
                        pi += cs * yi + sn * yr;
                    }
                Ir = fmaf(sc, pr, Ir);
                Ii = fmaf(sc, pi, Ii);
            }
            if (NW) nw_row += (ap.valid && at.valid) ? 1 : 0;
            if (b == A.n_f - 1) wtx_last = tsum;
        }
        if (b == A.n_f - 1) w_last = asum;
    }
    if (ENS) {
        // rows of this lane -> their frames; consecutive rows of one frame (row stacks) are
        // summed first
        if (vok) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                float sr = 0.f, si = 0.f;
                int f_run = -1;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * (mt0 + mt) + 4 * g + r;
                    int f = -1;
                    if (row < A.n_tx) {
                        if (STK) {
                            f = A.row_frame[row];
                        } else {
                            const int slot = row % A.ens_slots;
                            f = slot < A.ens_frames ? slot : -1;
                        }
                    }
                    if (f != f_run) {
                        if (f_run >= 0) {
                            float2 *img = A.frame_image ? A.frame_image[f_run] : A.image + (size_t)f_run * n_vox;
                            atomicAdd(&img[v].x, A.out_scale * sr);
                            atomicAdd(&img[v].y, A.out_scale * si);
                        }
                        f_run = f;
                        sr = si = 0.f;
                    }
                    sr += Er[ENS ? mt : 0][r];
                    si += Ei[ENS ? mt : 0][r];
                }
                if (f_run >= 0) {
                    float2 *img = A.frame_image ? A.frame_image[f_run] : A.image + (size_t)f_run * n_vox;
                    atomicAdd(&img[v].x, A.out_scale * sr);
                    atomicAdd(&img[v].y, A.out_scale * si);
                }
            }
            if (g == 0 && b_end == A.n_f && bz == 0 && A.weights) A.weights[v] = w_last;
        }
        return;
    }
    Ir += __shfl_xor(Ir, 16);
    Ir += __shfl_xor(Ir, 32);
    Ii += __shfl_xor(Ii, 16);
    Ii += __shfl_xor(Ii, 32);
    if (TAN) {
        Dr += __shfl_xor(Dr, 16);
        Dr += __shfl_xor(Dr, 32);
        Di += __shfl_xor(Di, 16);
        Di += __shfl_xor(Di, 32);
    }
    if (g == 0 && vok) {
        const float2 val = make_float2(A.out_scale * Ir, A.out_scale * Ii);
        float2 *dimg = reinterpret_cast<float2 *>(A.dimage);
        if (A.partial) {
            const long slot = (long)by * A.slot_z + (LBL ? 0 : bz);
            A.partial[slot * n_vox + v] = val;
            if (TAN) A.partial[((long)A.n_slots + slot) * n_vox + v] = make_float2(A.out_scale * Dr, A.out_scale * Di);
        } else {
            A.image[v] = val;
            if (TAN) dimg[v] = make_float2(A.out_scale * Dr, A.out_scale * Di);
        }
        if (b_end == A.n_f && (LBL || bz == 0)) {
            if (A.weights) A.weights[v] = w_last;
            if (MODE == MODE_SA && A.weights_tx) A.weights_tx[v] = wtx_last;
        }
    }
}

template <int MODE, int MT, int KI, int WCR, int WCT, int PREC, bool TAN = false, bool ENS = false, bool NW = false,
          bool LBL = false, bool STK = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BeamWavesPerEU<MT, KI, WCR, PREC, ENS>::value)))
void beamform_kernel(BeamArgs A)
{
    // small tiles take the arguments by value (their register count decides the
    // occupancy); LDS holds large FP32 ones at <= 2 waves per SIMD, where the
    // reference form leaves the compiler the registers it schedules best with
    using Args =
        typename std::conditional<(MT * KI <= 64 || PREC != PREC_FP32), BeamArgs, const BeamArgs &>::type;
    beamform_tile<MODE, MT, KI, WCR, WCT, PREC, TAN, ENS, NW, LBL, STK, Args>(A, blockIdx.x, blockIdx.y, blockIdx.z,
                                                                           threadIdx.x);
}

// Persistent batch form: the work items (job, voxel tile, bin chunk, transmit
// group) of every plan of one kernel class -- uneven grids, angle counts and band
// chunkings -- in one launch.  Workgroups claim items from a counter (reset to 0
// before the launch) until the list is exhausted, so every workgroup exits.
template <int MODE, int MT, int KI, int WCR, int WCT, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PersistWavesPerEU<MODE, MT, KI, WCR, PREC>::value)))
void beamform_persistent_kernel(const BeamArgs *__restrict__ jobs, const int4 *__restrict__ items, int n_items,
                                int *__restrict__ counter)
{
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        // wave-uniform item: its table row and the job's arguments are read with scalar loads
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();  // read by every thread before the next claim overwrites it
        if (it >= n_items) return;
        const int4 w = items[it];
        const BeamArgs &A = jobs[__builtin_amdgcn_readfirstlane(w.x)];
        // the thread index, opaque to the compiler per item: values derived from it are
        // recomputed per tile rather than hoisted out of the loop and held across tiles
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        beamform_tile<MODE, MT, KI, WCR, WCT, PREC, false, false, false, false, false, const BeamArgs &>(
            A, __builtin_amdgcn_readfirstlane(w.y),
                                                    __builtin_amdgcn_readfirstlane(w.z),
                                                    __builtin_amdgcn_readfirstlane(w.w), tid);
    }
}

static __global__ void reset_counter_kernel(int *c) { *c = 0; }

// sum of the partial images of every (bin chunk, transmit group) slot in slot
// order: out[v] = sum_s P[s][v]; dout (tangent) from slots n_slots .. 2 n_slots - 1
static __global__ __launch_bounds__(256) void reduce_partials_kernel(const float2 *__restrict__ P, int n_slots,
                                                                     long n_vox, float2 *__restrict__ out,
                                                                     float2 *__restrict__ dout)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vox) return;
    float2 a = P[v];
    for (int k = 1; k < n_slots; ++k) {
        const float2 q = P[(long)k * n_vox + v];
        a.x += q.x;
        a.y += q.y;
    }
    out[v] = a;
    if (dout) {
        float2 d = P[(long)n_slots * n_vox + v];
        for (int k = 1; k < n_slots; ++k) {
            const float2 q = P[(long)(n_slots + k) * n_vox + v];
            d.x += q.x;
            d.y += q.y;
        }
        dout[v] = d;
    }
}


// -------------------------------------------------------------------------
// 16-bit spectra planes for the mixed-precision contraction
// -------------------------------------------------------------------------
// One block per bin: reads the FP32 operand-ordered spectra of the bin, picks a
// power-of-two scale (16-bit float variants: largest component in [2^14, 2^15)),
// and writes planes [p][mt][c][64][8] (p: Xr hi, Xi hi, Xr lo, Xi lo) with
// lane (a & 15) + 16 ((k & 31) >> 3), slot k & 7, chunk k >> 5.
template <int PREC>
__global__ __launch_bounds__(256) void lowp_planes_kernel(const float2 *__restrict__ X, void *__restrict__ Xlp,
                                                          float *__restrict__ xscale, int MT_all, int KI)
{
    __shared__ float red[256];
    const int b = blockIdx.x;
    const int n_pairs = MT_all * KI * 64;
    const float2 *Xb = X + (size_t)b * n_pairs;
    float m = 0.f;
    for (int e = threadIdx.x; e < n_pairs; e += 256) {
        const float2 v = Xb[e];
        m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const float mx = red[0];
    float scale = 1.0f;
    if (PREC != PREC_BF16X3 && mx > 0.f) {
        int e2;
        frexpf(mx, &e2);              // mx = f 2^e2, f in [0.5, 1)
        scale = ldexpf(1.0f, e2 - 15);  // mx / scale in [2^14, 2^15)
    }
    if (threadIdx.x == 0) xscale[b] = scale;
    const float inv = 1.0f / scale;
    constexpr int pl = Prec<PREC>::planes;
    const int KC = KI / 8;
    const int plane_elems = MT_all * KC * 64 * 8;  // 16-bit values per plane
    unsigned short *out = reinterpret_cast<unsigned short *>(Xlp) + (size_t)b * pl * plane_elems;
    for (int o = threadIdx.x; o < plane_elems; o += 256) {
        const int s = o & 7, ln = (o >> 3) & 63, q = o >> 9;  // q = mt * KC + c
        const int mt = q / KC, cc = q - mt * KC;
        const int row = ln & 15, gg = ln >> 4;
        const int k = 32 * cc + 8 * gg + s;
        const float2 v = Xb[(mt * KI + (k >> 2)) * 64 + row + 16 * (k & 3)];
        const float vr = v.x * inv, vi = v.y * inv;
        if (PREC == PREC_BF16X3) {
            const __bf16 rh = (__bf16)vr, ih = (__bf16)vi;
            const __bf16 rl = (__bf16)(vr - (float)rh), il = (__bf16)(vi - (float)ih);
            out[o] = __builtin_bit_cast(unsigned short, rh);
            out[plane_elems + o] = __builtin_bit_cast(unsigned short, ih);
            out[2 * plane_elems + o] = __builtin_bit_cast(unsigned short, rl);
            out[3 * plane_elems + o] = __builtin_bit_cast(unsigned short, il);
        } else {
            const _Float16 rh = (_Float16)vr, ih = (_Float16)vi;
            out[o] = __builtin_bit_cast(unsigned short, rh);
            out[plane_elems + o] = __builtin_bit_cast(unsigned short, ih);
            if (PREC == PREC_FP16X3) {
                const _Float16 rl = (_Float16)(vr - (float)rh), il = (_Float16)(vi - (float)ih);
                out[2 * plane_elems + o] = __builtin_bit_cast(unsigned short, rl);
                out[3 * plane_elems + o] = __builtin_bit_cast(unsigned short, il);
            }
        }
    }
}

}  // namespace fnk
