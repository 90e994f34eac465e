// fn_core.hip -- plans, kernel dispatch and the GPU entry points of the C ABI
// (fn_das_pw / fn_das_sa / fn_plan_*).  Replaces the orchestration of
// gpu_bf_das_pw_rf.cu:700-1122 and gpu_bf_saft_rf.cu:800-1292: device
// buffers, spectra, per-frequency beamforming and the output scale.
#include "fn_internal.h"
#include "fn_kernels.hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace fnk {
#define FN_DECL_LAUNCHERS(m, k)                                                                      \
    hipError_t launch_beamform_m##m##_k##k##_p0(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &); \
    hipError_t launch_beamform_m##m##_k##k##_p1(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &); \
    hipError_t launch_beamform_m##m##_k##k##_p2(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &); \
    hipError_t launch_beamform_m##m##_k##k##_p3(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
FN_DECL_LAUNCHERS(0, 8)
FN_DECL_LAUNCHERS(0, 16)
FN_DECL_LAUNCHERS(0, 32)
FN_DECL_LAUNCHERS(1, 8)
FN_DECL_LAUNCHERS(1, 16)
FN_DECL_LAUNCHERS(1, 32)
FN_DECL_LAUNCHERS(0, 48)
FN_DECL_LAUNCHERS(0, 64)
FN_DECL_LAUNCHERS(1, 48)
FN_DECL_LAUNCHERS(1, 64)
#undef FN_DECL_LAUNCHERS
#define FN_DECL_ENS(k)                                                                              \
    hipError_t launch_beamform_ens_k##k##_p0(int, int, dim3, size_t, hipStream_t, const BeamArgs &, int); \
    hipError_t launch_beamform_ens_k##k##_p1(int, int, dim3, size_t, hipStream_t, const BeamArgs &, int); \
    hipError_t launch_beamform_ens_k##k##_p2(int, int, dim3, size_t, hipStream_t, const BeamArgs &, int); \
    hipError_t launch_beamform_ens_k##k##_p3(int, int, dim3, size_t, hipStream_t, const BeamArgs &, int);
FN_DECL_ENS(8)
FN_DECL_ENS(16)
FN_DECL_ENS(32)
#undef FN_DECL_ENS
// the last argument: 1 for a row stack (per-row transmit phases), 0 for an ensemble
typedef hipError_t (*EnsLauncher)(int, int, dim3, size_t, hipStream_t, const BeamArgs &, int);
// [KI = 8, 16, 32][precision]
static const EnsLauncher kEnsLaunchers[3][4] = {
    {launch_beamform_ens_k8_p0, launch_beamform_ens_k8_p1, launch_beamform_ens_k8_p2, launch_beamform_ens_k8_p3},
    {launch_beamform_ens_k16_p0, launch_beamform_ens_k16_p1, launch_beamform_ens_k16_p2, launch_beamform_ens_k16_p3},
    {launch_beamform_ens_k32_p0, launch_beamform_ens_k32_p1, launch_beamform_ens_k32_p2, launch_beamform_ens_k32_p3}};
hipError_t launch_beamform_nw_m0_k8(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_nw_m0_k16(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_nw_m0_k32(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_nw_m1_k8(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_nw_m1_k16(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_nw_m1_k32(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
#define FN_DECL_PERSIST(m, k)                                                                                \
    hipError_t launch_beamform_persistent_m##m##_k##k##_p0(int, int, int, int, size_t, hipStream_t,               \
                                                           const BeamArgs *, const int4 *, int, int *, int *);    \
    hipError_t launch_beamform_persistent_m##m##_k##k##_p1(int, int, int, int, size_t, hipStream_t,               \
                                                           const BeamArgs *, const int4 *, int, int *, int *);    \
    hipError_t launch_beamform_persistent_m##m##_k##k##_p2(int, int, int, int, size_t, hipStream_t,               \
                                                           const BeamArgs *, const int4 *, int, int *, int *);    \
    hipError_t launch_beamform_persistent_m##m##_k##k##_p3(int, int, int, int, size_t, hipStream_t,               \
                                                           const BeamArgs *, const int4 *, int, int *, int *);
FN_DECL_PERSIST(0, 8) FN_DECL_PERSIST(0, 16) FN_DECL_PERSIST(0, 32)
FN_DECL_PERSIST(1, 8) FN_DECL_PERSIST(1, 16) FN_DECL_PERSIST(1, 32)
typedef hipError_t (*PersistLauncher)(int, int, int, int, size_t, hipStream_t, const BeamArgs *, const int4 *, int,
                                      int *, int *);
#define FN_PERSIST_ROW(m, k)                                                                                 \
    {launch_beamform_persistent_m##m##_k##k##_p0, launch_beamform_persistent_m##m##_k##k##_p1,                 \
     launch_beamform_persistent_m##m##_k##k##_p2, launch_beamform_persistent_m##m##_k##k##_p3}
static const PersistLauncher kPersistLaunchers[2][3][4] = {
    {FN_PERSIST_ROW(0, 8), FN_PERSIST_ROW(0, 16), FN_PERSIST_ROW(0, 32)},
    {FN_PERSIST_ROW(1, 8), FN_PERSIST_ROW(1, 16), FN_PERSIST_ROW(1, 32)}};
hipError_t launch_windowed_rows(const BeamArgs &A, int mode, int win_rx, double par_rx, float *h_rx, int win_tx,
                                double par_tx, float *h_tx, hipStream_t s);
hipError_t launch_focused_signal(const BeamArgs &A, int mode, int wcr, int wct, int MT_all, int KI, int nw, int ix,
                                 int iz, float2 *F, const float2 *tw, float2 *out, hipStream_t s);
hipError_t launch_beamform_tan_k8(int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_tan_k16(int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_tan_k32(int, int, dim3, size_t, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_lbl_k8(int, dim3, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_lbl_k16(int, dim3, hipStream_t, const BeamArgs &);
hipError_t launch_beamform_lbl_k32(int, dim3, hipStream_t, const BeamArgs &);
typedef hipError_t (*LblLauncher)(int, dim3, hipStream_t, const BeamArgs &);
static const LblLauncher kLblLaunchers[3] = {launch_beamform_lbl_k8, launch_beamform_lbl_k16, launch_beamform_lbl_k32};
hipError_t launch_adjoint(int wcr, bool tables, bool scale, dim3 grid, hipStream_t s, const AdjArgs &S);
hipError_t launch_adjoint_reduce(const float *P, float2 *Z, int n_vr, int n_f, int n_at, int n_tx, int n_el,
                                 hipStream_t s);
hipError_t launch_adjoint_idft(const float2 *Z, const float2 *tw, float *out, int n_t, int n_ch, int N, int n_lb,
                               int n_f, hipStream_t s);
hipError_t launch_vjp_delay(const float2 *X, const float2 *Z, double *out, int n_f, int n_tx, int n_el, int MT, int KI,
                            int n_lb, double w_unit, hipStream_t s);
hipError_t launch_dot_re(const float2 *g, const float2 *d, long n, double *part, int n_part, double *out, hipStream_t s);
typedef hipError_t (*BeamLauncher)(int, int, int, dim3, size_t, hipStream_t, const BeamArgs &);
// [mode][KI = 8, 16, 32] (windowed normalization; [mode][KI = 8 .. 64][precision] below)
static const BeamLauncher kNwLaunchers[2][3] = {
    {launch_beamform_nw_m0_k8, launch_beamform_nw_m0_k16, launch_beamform_nw_m0_k32},
    {launch_beamform_nw_m1_k8, launch_beamform_nw_m1_k16, launch_beamform_nw_m1_k32}};
#define FN_BEAM_ROW(m, k) \
    {launch_beamform_m##m##_k##k##_p0, launch_beamform_m##m##_k##k##_p1, launch_beamform_m##m##_k##k##_p2, \
     launch_beamform_m##m##_k##k##_p3}
static const BeamLauncher kBeamLaunchers[2][5][4] = {
    {FN_BEAM_ROW(0, 8), FN_BEAM_ROW(0, 16), FN_BEAM_ROW(0, 32), FN_BEAM_ROW(0, 48), FN_BEAM_ROW(0, 64)},
    {FN_BEAM_ROW(1, 8), FN_BEAM_ROW(1, 16), FN_BEAM_ROW(1, 32), FN_BEAM_ROW(1, 48), FN_BEAM_ROW(1, 64)}};
#undef FN_BEAM_ROW
// zero fill used instead of hipMemsetAsync so the launch sequence stays
// capturable into a graph
__global__ void fill_zero_kernel(float *p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}
}  // namespace fnk

using fnh::fail;

namespace {

constexpr size_t kMaxLds = 160 * 1024;  // gfx950 LDS per workgroup

struct Run {
    hipEvent_t e0, e1, e2;
};

}  // namespace

struct fn_plan {
    int mode = 0;  // 0: plane waves, 1: synthetic aperture
    int device = 0;
    fnh::Grid g;
    int n_x = 0, n_z = 0, n_t = 0, n_el = 0, n_tx = 0;
    long n_vox = 0;
    int KI = 0, MT_all = 0, MT_tmpl = 0, n_groups = 1, bin_split = 1;
    int n_frames = 1;  // ensemble frames (transmit rows = frames x angles)
    int ens_angles = 0;  // ensemble: angles per frame
    int wcr = 0, wct = 0;
    int nw = 0;                 // windowed normalization (CPU semantics): taper kernels, one bin chunk
    int lbl = 0;                // line scan: grid column i from transmit row i (groups = columns)
    float *d_nw = nullptr;      // L / 2 per (voxel, element): receive, then (SA) transmit
    int executed = 0;           // spectra of an execute are in d_X (focused signal)
    float2 *d_sigF = nullptr;   // focused spectrum of the signal voxel [n_f]
    std::vector<double> h_px, h_pz, h_Frx, h_Ftx, h_sqf;  // host copies (signal voxel search)
    std::vector<double> h_delay;                          // delay offsets of the last set_tof (empty: zeros)
    int precision = FN_PRECISION_FP32;
    int verbosity = 0;
    fnk::BeamArgs args{};
    fnk::SpecArgs sargs{};
    size_t lds_beam = 0, lds_spec = 0;
    double *d_px = nullptr, *d_pz = nullptr, *d_Frx = nullptr, *d_Ftx = nullptr, *d_txp = nullptr, *d_sqf = nullptr;
    float2 *d_tw = nullptr, *d_X = nullptr;
    void *d_Xlp = nullptr;    // 16-bit spectra planes (mixed precision)
    float *d_xscale = nullptr;
    double *d_delay = nullptr;  // per-element delay offsets (zeros unless set)
    // adjoint / vector-Jacobian workspace (allocated on first use)
    float *d_P = nullptr;       // voxel-range partials of the voxel contraction
    float2 *d_Z = nullptr;      // reduced contraction [n_f][n_tx][n_el]
    uint32_t *d_qtab = nullptr; // per-voxel phase words of the contraction
    double *d_part = nullptr;   // reduction partials
    int adj_vr = 0, adj_at = 0;
    long tof_serial = 1;                // bumped by every set_tof
    long adj_tab_serial = 0;            // tof_serial the phase tables were built for
    hipStream_t adj_tab_stream = nullptr;  // stream that built them (reuse is ordered only there)
    float *d_vsc = nullptr;                // adjoint normalization s / w_n(v) [n_f][n_vox]
    bool adj_sc_built = false;
    hipStream_t adj_sc_stream = nullptr;
    long adj_vpr = 0;
    size_t x_bytes = 0, xlp_bytes = 0;
    float2 *d_pimg = nullptr;   // partial images [slot][n_vox] (several workgroups per voxel)
    size_t pimg_cap = 0;        // float2 capacity of d_pimg
    float2 *d_ptan = nullptr;   // tangent partials: image slots, then dI/dc slots
    size_t ptan_cap = 0;
    hipStream_t stream = nullptr;
    int profiling = 0;
    std::vector<Run> pending, spare;
    double spec_ms = 0.0, beam_ms = 0.0;
    int runs = 0;
    double flops = 0.0, useful_flops = 0.0;
};

namespace {

// voxel tiles of 64 per beamforming group: the whole grid, or (line scan) one column
long beam_tiles(const fn_plan *P) { return P->lbl ? (P->n_z + 63) / 64 : (P->n_vox + 63) / 64; }

// workgroups that share a voxel: bin chunks times transmit groups (a line scan's
// groups are its columns, one per voxel); ensembles add their frames' rows atomically
int beam_slots(const fn_plan *P) { return P->n_frames > 1 ? 1 : P->bin_split * (P->lbl ? 1 : P->n_groups); }

int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory) return fail(FN_ERR_OUT_OF_MEMORY, "%s: out of device memory", what);
    return fail(FN_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_CHECK(call, what)                          \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// partial-image buffer of the plan's tiling (grown, never shrunk; freed with the plan)
int ensure_partials(fn_plan *P)
{
    const int slots = beam_slots(P);
    const size_t need = slots > 1 ? (size_t)slots * P->n_vox : 0;
    if (need <= P->pimg_cap) return FN_OK;
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    if (P->d_pimg) HIP_CHECK(hipFree(P->d_pimg), "hipFree");  // synchronizes: no launch still reads it
    P->d_pimg = nullptr;
    P->pimg_cap = 0;
    HIP_CHECK(hipMalloc(&P->d_pimg, need * sizeof(float2)), "partial images");
    P->pimg_cap = need;
    return FN_OK;
}

int select_device(int index)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) return fail(FN_ERR_NO_DEVICES, "Could not find any HIP devices!");
    // gpu_bf_das_pw_rf.cu:460 tests (index >= 0 || index < count), which always
    // passes; the intended check is applied here.
    if (index < 0 || index >= count) return fail(FN_ERR_DEVICE, "Selected device does not exist!");
    HIP_CHECK(hipSetDevice(index), "hipSetDevice");
    return FN_OK;
}

// window -> (class, coefficients, support); see fn_kernels.hip.h for the forms
int setup_window(const fn_window &w, int semantics, fnk::WinCoef *c, int *wc, int *support)
{
    const bool cpu = semantics == FN_SEMANTICS_CPU;
    *c = fnk::WinCoef{-1.f, INFINITY, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    switch (w.id) {
    case FN_WINDOW_BOXCAR:
        *wc = fnk::WC_RANGE;
        *support = cpu ? fnk::SUPPORT_OPEN_H : fnk::SUPPORT_SNAPPED;
        return FN_OK;
    case FN_WINDOW_TUKEY:
        *wc = fnk::WC_COS;
        c->t_flat = c->t_off = (float)(1.0 - w.parameter);
        c->t_cut = 1.f;
        c->A = c->B = 0.5f;
        c->s_half = (float)(0.5 / w.parameter);
        c->q_s = (float)(0.25 / w.parameter);
        c->q_o = (float)(-(1.0 - w.parameter) * 0.25 / w.parameter);
        *support = fnk::SUPPORT_SNAPPED;  // the window is zero beyond the snapped aperture
        return FN_OK;
    case FN_WINDOW_HANN:
        *wc = fnk::WC_COS;
        c->t_cut = cpu ? INFINITY : 1.f;
        c->A = c->B = 0.5f;
        c->s_half = 0.5f;
        c->q_s = 0.25f;
        *support = cpu ? fnk::SUPPORT_ALL : fnk::SUPPORT_SNAPPED;
        return FN_OK;
    case FN_WINDOW_TRIANGULAR:
        *wc = fnk::WC_LINEAR;
        c->t_cut = cpu ? INFINITY : 1.f;
        c->A = 1.f;
        c->C = -1.f;
        *support = cpu ? fnk::SUPPORT_ALL : fnk::SUPPORT_SNAPPED;
        return FN_OK;
    default:
        // id 3 (cosine) has no windows.* class (gpu_bf_helper_functions.cu:56-97 never yields it)
        return fail(FN_ERR_WINDOW, "window id %d is not reachable from any windows.* class", (int)w.id);
    }
}

// taper of the windowed normalization in CPU semantics as window_any()
// coefficients (|t|: flat part, A + C t + B cos(2 pi (t - t_off) s_half), cut)
int taper_coef(const fn_window &w, fnk::WinCoef *c)
{
    *c = fnk::WinCoef{-1.f, INFINITY, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    switch (w.id) {
    case FN_WINDOW_BOXCAR: c->t_cut = 1.f; return FN_OK;                       // |t| < 1
    case FN_WINDOW_HANN: c->A = c->B = 0.5f; c->s_half = 0.5f; return FN_OK;   // untruncated
    case FN_WINDOW_TUKEY:
        c->t_flat = c->t_off = (float)(1.0 - w.parameter);
        c->t_cut = 1.f;
        c->A = c->B = 0.5f;
        c->s_half = (float)(0.5 / w.parameter);
        return FN_OK;
    case FN_WINDOW_TRIANGULAR: c->A = 1.f; c->C = -1.f; return FN_OK;          // 1 - |t|, untruncated
    default: return fail(FN_ERR_WINDOW, "window id %d is not reachable from any windows.* class", (int)w.id);
    }
}

// windowed normalization: L / 2 per (voxel, element) from the FP64 row pass,
// once per plan (geometry and F-number tables only)
int setup_windowed(fn_plan *P, bool rx, const fn_window &wrx, const fn_window &nwin_rx, bool tx, const fn_window &wtx,
                   const fn_window &nwin_tx)
{
    if (P->precision != FN_PRECISION_FP32)
        return fail(FN_ERR_UNSUPPORTED, "normalizations.windowed runs in FP32 only");
    if (P->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "normalizations.windowed: single-frame plans only");
    const size_t n = (size_t)P->n_vox * P->n_el;
    HIP_CHECK(hipMalloc(&P->d_nw, n * (rx + tx) * sizeof(float)), "windowed normalization rows");
    float *h_rx = rx ? P->d_nw : nullptr, *h_tx = tx ? P->d_nw + (rx ? n : 0) : nullptr;
    int s;
    if (rx && (s = taper_coef(nwin_rx, &P->args.wn_rx))) return s;
    if (tx && (s = taper_coef(nwin_tx, &P->args.wn_tx))) return s;
    hipError_t e = fnk::launch_windowed_rows(P->args, P->mode, wrx.id, wrx.parameter, h_rx, wtx.id, wtx.parameter, h_tx,
                                             P->stream);
    if (e != hipSuccess) return hip_fail(e, "windowed_rows_kernel");
    HIP_CHECK(hipStreamSynchronize(P->stream), "windowed_rows_kernel");
    P->args.nw_rx = h_rx;
    P->args.nw_tx = h_tx;
    return FN_OK;
}

// forward element loops without per-element support tests: the raised cosine
// over every element of a full array (CPU Hann) takes WC_COS_ALL, a cosine
// window whose cut implies the support takes WC_COS_RF; tangent and adjoint
// launch the general classes
void use_cos_all(fn_plan *P)
{
    if (P->nw) return;
    const fnk::WinCoef &w = P->args.wrx;
    if (P->wcr == fnk::WC_COS) {
        if (P->args.support_rx == fnk::SUPPORT_ALL && P->n_el == 4 * P->KI && w.t_off == 0.f && w.t_flat < 0.f &&
            std::isinf(w.t_cut))
            P->wcr = fnk::WC_COS_ALL;
        else if (fnk::cos_rf_window(P->args, P->args.support_rx, w, 4 * P->KI))
            P->wcr = fnk::WC_COS_RF;
    }
    // the synthetic aperture's transmit side: the same cut-window class for its
    // transmitters (the SA kernels launch it; every other SA path evaluates the window
    // with its support test, which WC_COS_RF maps to)
    if (P->mode == 1 && P->wct == fnk::WC_COS &&
        fnk::cos_rf_window(P->args, P->args.support_tx, P->args.wtx, 4 * P->KI))
        P->wct = fnk::WC_COS_RF;
}

int wcr_general(int wcr) { return (wcr == fnk::WC_COS_ALL || wcr == fnk::WC_COS_RF) ? fnk::WC_COS : wcr; }

// default transmit tiles per plane-wave workgroup that fit the register file:
// ensembles hold per-row image accumulators, so 4; FP32 and fp16 at 128 elements 5
// (FP32 with one 8-byte reload per bin from scratch in the full-array cosine class)
int tile_cap(int n_frames, int precision, int KI)
{
    if (n_frames <= 1) return 5;
    if ((precision == FN_PRECISION_FP16 || precision == FN_PRECISION_FP32) && KI == 32) return 5;
    return 4;
}

// transmit tiles per group of a row stack (its own kernel instances, per-row phases)
int stack_tile_cap(int precision, int KI) { return (precision == FN_PRECISION_FP16 && KI == 32) ? 5 : 4; }

// element steps per lane: 4 KI >= n_el elements in four lane groups
int ki_of(int n_el)
{
    if (n_el <= 32) return 8;
    if (n_el <= 64) return 16;
    if (n_el <= 128) return 32;
    if (n_el <= 192) return 48;
    if (n_el <= 256) return 64;
    return 0;
}

// index of a KI in the launcher tables
int ki_index(int KI) { return KI == 8 ? 0 : KI == 16 ? 1 : KI == 32 ? 2 : KI == 48 ? 3 : 4; }

// transmit tiles per synthetic-aperture group: every transmitter up to 128
// elements, groups of 4 tiles beyond (the staged tile within 128 KiB)
int sa_mt(int KI) { return KI <= 32 ? KI / 4 : 4; }

// transmit tiles per plane-wave group beyond 128 elements: the staged FP32 tile
// (MT x KI x 512 bytes) within 80 KiB
int pw_mt_wide(int KI) { return KI <= 32 ? 5 : 81920 / (KI * 512); }

struct Common {
    const double *px, *pz;
    int n_x, n_z, n_t, n_el, n_tx;
    double f_s, width, pitch, c0, t0;
    int semantics, precision, scale, index_gpu, verbosity;
};

void destroy(fn_plan *p)
{
    if (!p) return;
    hipFree(p->d_px); hipFree(p->d_pz); hipFree(p->d_Frx); hipFree(p->d_Ftx); hipFree(p->d_txp); hipFree(p->d_sqf);
    hipFree(p->d_tw); hipFree(p->d_X); hipFree(p->d_Xlp); hipFree(p->d_xscale);
    hipFree(p->d_delay); hipFree(p->d_P); hipFree(p->d_Z); hipFree(p->d_part); hipFree(p->d_qtab); hipFree(p->d_vsc); hipFree(p->d_nw);
    hipFree(p->d_sigF); hipFree(p->d_pimg); hipFree(p->d_ptan);
    for (auto &r : p->pending) { hipEventDestroy(r.e0); hipEventDestroy(r.e1); hipEventDestroy(r.e2); }
    for (auto &r : p->spare) { hipEventDestroy(r.e0); hipEventDestroy(r.e1); hipEventDestroy(r.e2); }
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
}

template <typename T>
int upload(T **dst, const T *src, size_t n, const char *what)
{
    HIP_CHECK(hipMalloc(dst, sizeof(T) * n), what);
    HIP_CHECK(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice), what);
    return FN_OK;
}

#ifndef FN_FEW_ROWS_BLOCKS
#define FN_FEW_ROWS_BLOCKS 8192L
#endif
static bool few_rows_pw(const fn_plan *P)
{
    return P->mode == 0 && P->n_frames == 1 && !P->lbl && P->MT_tmpl <= 2;
}

// shared configuration of plan creation: tiling, launch shape and the scalar
// arguments of the kernels; no device work and no allocation (the
// no-allocation entry point runs it on a plan held on the stack)
int configure_common(fn_plan *P, const Common &c, const fnh::Grid &g)
{
    if (c.precision < FN_PRECISION_FP32 || c.precision > FN_PRECISION_FP16)
        return fail(FN_ERR_ARGUMENT, "unknown precision %d", c.precision);
    P->device = c.index_gpu;
    P->g = g;
    P->n_x = c.n_x; P->n_z = c.n_z; P->n_t = c.n_t; P->n_el = c.n_el; P->n_tx = c.n_tx;
    P->n_vox = (long)c.n_x * c.n_z;
    P->verbosity = c.verbosity;
    P->precision = c.precision;
    P->KI = ki_of(c.n_el);
    if (!P->KI) return fail(FN_ERR_UNSUPPORTED, "at most 256 array elements are supported (got %d)", c.n_el);
    if (P->KI > 32) {
        // beyond 128 elements: the one-launch forward kernel of single-frame plans
        if (P->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "ensembles: at most 128 elements (got %d)", c.n_el);
        if (P->lbl) return fail(FN_ERR_UNSUPPORTED, "line scans: at most 128 elements (got %d)", c.n_el);
        if (P->nw)
            return fail(FN_ERR_UNSUPPORTED, "normalizations.windowed (CPU semantics): at most 128 elements (got %d)",
                        c.n_el);
    }
    // operand rows: an ensemble's angles x frame slots (frames rounded up to a multiple
    // of 4, angle-major), otherwise the transmit rows
    const int op_rows = P->n_frames > 1 ? P->ens_angles * P->args.ens_slots : c.n_tx;
    P->MT_all = (op_rows + 15) / 16;
    if (P->lbl) {
        // line scan: one group per grid column, staging the M-tile of its transmit row
        P->MT_tmpl = 1;
        P->n_groups = c.n_x;
    } else if (P->mode == 0) {
        // at most 5 transmit tiles per group keep the LDS tile within 80 KiB
        // (ensembles: 4-5, for the per-row image accumulators)
        const int max_tiles = P->nw ? 2 : std::min(tile_cap(P->n_frames, c.precision, P->KI), pw_mt_wide(P->KI));
        const int groups = (P->MT_all + max_tiles - 1) / max_tiles;
        P->MT_tmpl = (P->MT_all + groups - 1) / groups;
        P->n_groups = (P->MT_all + P->MT_tmpl - 1) / P->MT_tmpl;
    } else {
        P->MT_tmpl = sa_mt(P->KI);
        P->n_groups = (P->MT_all + P->MT_tmpl - 1) / P->MT_tmpl;
    }
    // staged spectra per bin: FP32 interleaved complex (8 bytes per value) or
    // 16-bit planes (Xr, Xi and, for the split forms, their residuals)
    const int planes = c.precision == FN_PRECISION_FP16 ? 2 : 4;
    const size_t step_bytes = c.precision == FN_PRECISION_FP32 ? 64 * sizeof(float2) : (size_t)planes * 64 * 2;
    P->lds_beam = (size_t)P->MT_tmpl * P->KI * step_bytes;
    P->lds_spec = (size_t)g.N * sizeof(float2) + 64 * 65 * sizeof(float);
    if (P->lds_beam > kMaxLds) return fail(FN_ERR_UNSUPPORTED, "transmit tile exceeds LDS");
    if (P->lds_spec > kMaxLds) return fail(FN_ERR_UNSUPPORTED, "DFT order %d too large for the LDS twiddle table", g.N);

    // bin split: enough workgroups to fill 256 CUs several times over
    const long tiles = beam_tiles(P);
    const long blocks = tiles * P->n_groups;
    int split = 1;  // windowed normalization: every block walks all bins (valid-row counts)
    if (P->lbl) {
        // line scans (vector-ALU bound, one column per group): about 8192 blocks below
        // 4096, measured best over bin splits 1..64 at 64-256 lines x 256-1024 depths
        if (blocks < 4096) split = (int)std::min<long>({(8192 + blocks - 1) / blocks, 16L, (long)std::max(1, g.n_f / 8)});
    } else if (!P->nw && blocks < (few_rows_pw(P) ? FN_FEW_ROWS_BLOCKS : 2048L)) {
        // single-frame plane waves of one or two transmit tiles per group (vector-ALU
        // bound element loops): about FN_FEW_ROWS_BLOCKS blocks, otherwise about 2048
        const long target = few_rows_pw(P) ? FN_FEW_ROWS_BLOCKS : 2048L;
        split = (int)std::min<long>((target + blocks - 1) / blocks, std::max(1, g.n_f / 8));
    } else if (P->mode == 0 && !P->nw && P->n_frames == 1) {
        // a group's band spectra beyond 32 MiB (many transmit tiles, a long band): the bins
        // are split so that the workgroups running together walk about 4 MiB of them (an
        // XCD's L2); at 75 angles x 1455 bins this cut the fetches beyond L2 from 152-179 GB
        // to 74-93 GB per launch at the same duration
        const double set = (double)g.n_f * (double)P->lds_beam;
        if (set > 32.0 * 1048576.0)
            split = (int)std::min<double>({16.0, std::ceil(set / (4.0 * 1048576.0)), (double)std::max(1, g.n_f / 8)});
    }
    const int chunk = (g.n_f + split - 1) / split;
    P->bin_split = (g.n_f + chunk - 1) / chunk;
    P->x_bytes = (size_t)g.n_f * P->MT_all * P->KI * 64 * sizeof(float2);
    if (c.precision != FN_PRECISION_FP32) P->xlp_bytes = (size_t)g.n_f * planes * P->MT_all * P->KI * 64 * 2;

    fnk::BeamArgs &A = P->args;
    A.ctof = c.c0;
    A.dimage = nullptr;
    A.n_x = c.n_x; A.n_z = c.n_z; A.n_el = c.n_el; A.n_tx = op_rows;
    A.n_f = g.n_f; A.n_lb = g.n_lb; A.N = g.N;
    A.bin_chunk = chunk;
    A.tx_group = P->MT_tmpl;
    A.M = (c.n_el - 1) / 2.0;
    A.pitch = c.pitch; A.width = c.width; A.c0 = c.c0; A.f_s = c.f_s; A.t0 = c.t0;
    // das_pw.m:384,467 (CPU: 2 * sum) vs gpu_bf_das_pw_rf.cu:1118-1119 (MEX: 2 / N * sum)
    A.out_scale = (c.scale == FN_SCALE_CPU) ? 2.0f : (float)(2.0 / g.N);

    fnk::SpecArgs &S = P->sargs;
    S.n_t = c.n_t; S.n_ch = c.n_el * c.n_tx; S.N = g.N; S.n_lb = g.n_lb; S.n_f = g.n_f;
    S.n_el = c.n_el; S.n_tx = op_rows; S.MT = P->MT_all; S.KI = P->KI;
    S.row0 = 0; S.MT_out = P->MT_all; S.write_pad = 1;
    S.ens_slots = P->args.ens_slots;
    S.ens_frames = P->args.ens_frames;
    S.ens_angles = P->n_frames > 1 ? P->ens_angles : 0;

    P->flops = 8.0 * g.n_f * (double)P->n_vox * c.n_el * (P->lbl ? 1 : c.n_tx);
    return FN_OK;
}

// shared part of plan creation: configuration, then the device tables and workspace
int create_common(fn_plan *P, const Common &c, const fnh::Grid &g, const double *Frx, const double *Ftx)
{
    int s = configure_common(P, c, g);
    if (s) return s;
    if ((s = select_device(c.index_gpu))) return s;

    HIP_CHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate");
    if ((s = upload(&P->d_px, c.px, c.n_x, "positions_x"))) return s;
    if ((s = upload(&P->d_pz, c.pz, c.n_z, "positions_z"))) return s;
    if ((s = upload(&P->d_Frx, Frx, g.n_f, "F-number"))) return s;
    if (Ftx && (s = upload(&P->d_Ftx, Ftx, g.n_f, "F-number tx"))) return s;
    // sqrt(f_n) per bin for the depth clamp of the F-number
    std::vector<double> sqf(g.n_f);
    for (int b = 0; b < g.n_f; ++b) sqf[b] = std::sqrt(c.f_s * (double)(g.n_lb + b) / g.N);
    if ((s = upload(&P->d_sqf, sqf.data(), sqf.size(), "sqrt(f)"))) return s;
    P->h_px.assign(c.px, c.px + c.n_x);
    P->h_pz.assign(c.pz, c.pz + c.n_z);
    P->h_Frx.assign(Frx, Frx + g.n_f);
    if (Ftx) P->h_Ftx.assign(Ftx, Ftx + g.n_f);
    P->h_sqf = sqf;
    // twiddles e^{+j 2 pi m / N} rounded once from long double (conjugated in the kernel)
    std::vector<float2> tw(g.N);
    for (int m = 0; m < g.N; ++m) {
        const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)g.N;
        tw[m] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    if ((s = upload(&P->d_tw, tw.data(), tw.size(), "twiddles"))) return s;
    HIP_CHECK(hipMalloc(&P->d_X, P->x_bytes), "spectra workspace");
    HIP_CHECK(hipMemset(P->d_X, 0, P->x_bytes), "spectra workspace");  // padding rows stay zero
    if (c.precision != FN_PRECISION_FP32) {
        HIP_CHECK(hipMalloc(&P->d_Xlp, P->xlp_bytes), "16-bit spectra planes");
        HIP_CHECK(hipMalloc(&P->d_xscale, (size_t)g.n_f * sizeof(float)), "spectra scales");
    }

    HIP_CHECK(hipMalloc(&P->d_delay, (size_t)c.n_el * sizeof(double)), "delay offsets");
    HIP_CHECK(hipMemset(P->d_delay, 0, (size_t)c.n_el * sizeof(double)), "delay offsets");

    fnk::BeamArgs &A = P->args;
    A.X = P->d_X;
    A.delay = P->d_delay;
    A.Xlp = P->d_Xlp;
    A.xscale = P->d_xscale;
    A.px = P->d_px; A.pz = P->d_pz; A.Frx = P->d_Frx; A.Ftx = P->d_Ftx; A.txp = P->d_txp; A.sqrt_f = P->d_sqf;
    return ensure_partials(P);
}

// Elements whose weight can be nonzero for one voxel and bin: the aperture of
// das_pw.m:387-411 (das_sa.m:384-400) -- the open boxcar support (CPU), the snapped
// [lb, ub] of a window that vanishes beyond it, or every element of an untruncated
// window (CPU Hann and triangular) on a non-empty aperture; 0 where the aperture is
// empty (the bin is skipped).
int active_elements(double x, double z, double F, int support, double pitch, int n_el)
{
    const double M = (n_el - 1) / 2.0;
    const double hp = z / (2.0 * pitch * F);  // h / pitch (F = 0: the whole array)
    const double ua = M + (x / pitch - hp), ub = M + (x / pitch + hp);
    const double lb = std::fmax(std::ceil(ua), 0.0), ubd = std::fmin(std::floor(ub), n_el - 1.0);
    if (lb > ubd) return 0;
    if (support == fnk::SUPPORT_OPEN_H) {
        const double lo = std::fmax(std::floor(ua) + 1.0, 0.0), hi = std::fmin(std::ceil(ub) - 1.0, n_el - 1.0);
        return hi >= lo ? (int)(hi - lo) + 1 : 0;
    }
    if (support == fnk::SUPPORT_ALL) return n_el;
    return (int)(ubd - lb) + 1;
}

// Useful complex multiply-adds per frame for the roofline (8 flop each): the sum over
// voxels and bins of the active receive elements times the transmit rows (plane
// waves), or times the active transmitters (synthetic aperture), from a sample of
// about 64 x 64 voxels and every bin, scaled to the grid.
double useful_macs(const fn_plan *P, const double *px, int nx, const double *pz, int nz, const double *Frx,
                   const double *Ftx, double f_s, double c0, double pitch)
{
    const fnh::Grid &g = P->g;
    const int sx = std::max(1, nx / 64), sz = std::max(1, nz / 64);
    double sum = 0.0;
    long n = 0;
    for (int ix = 0; ix < nx; ix += sx)
        for (int iz = 0; iz < nz; iz += sz, ++n)
            for (int b = 0; b < g.n_f; ++b) {
                double Fr = Frx[b];
                if (P->args.clamp) {
                    const double fn = f_s * (double)(g.n_lb + b) / g.N;
                    Fr = std::fmin(Fr, std::sqrt(fn * pz[iz] / (2.88 * c0)));  // das_pw.m:374-377
                }
                const int k = active_elements(px[ix], pz[iz], Fr, P->args.support_rx, pitch, P->n_el);
                if (!k) continue;
                if (P->mode == 1) {
                    const int t = active_elements(px[ix], pz[iz], Ftx[b], P->args.support_tx, pitch, P->n_el);
                    sum += (double)k * t;
                } else {
                    sum += (double)k * (P->lbl ? 1 : P->n_tx);  // a line scan's voxel meets one row
                }
            }
    return n ? sum / n * (double)P->n_vox : 0.0;
}

hipError_t launch_beam(const fn_plan *P, dim3 grid, const fnk::BeamArgs &A, hipStream_t s)
{
    const int kidx = ki_index(P->KI);  // ensembles, line scans and windowed plans stop at KI = 32
    if (P->lbl) return fnk::kLblLaunchers[kidx](P->wcr, grid, s, A);
    if (P->nw) return fnk::kNwLaunchers[P->mode][kidx](P->MT_tmpl, P->wcr, P->wct, grid, P->lds_beam, s, A);
    if (P->n_frames > 1) return fnk::kEnsLaunchers[kidx][P->precision](P->MT_tmpl, P->wcr, grid, P->lds_beam, s, A, 0);
    return fnk::kBeamLaunchers[P->mode][kidx][P->precision](P->MT_tmpl, P->wcr, P->wct, grid, P->lds_beam, s, A);
}

// raise a kernel's dynamic-LDS limit once per process and size (no attribute
// calls inside captured launch sequences)
hipError_t lds_limit(const void *kern, size_t lds, size_t &done)
{
    if (lds <= 65536 || lds <= done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) done = lds;
    return e;
}

// the band spectra of RF d_rf into the operand layout X described by S
int launch_spectra_into(const fn_plan *P, const void *d_rf, int dtype, hipStream_t s, float2 *X,
                        const fnk::SpecArgs &S)
{
    // 64 bins x one operand group (16 rows x 4 elements) per block
    dim3 sg((unsigned)((S.n_f + 63) / 64), (unsigned)(S.MT * S.KI));
    static size_t done_f = 0, done_d = 0;
    HIP_CHECK(dtype == FN_DTYPE_F32
                  ? lds_limit(reinterpret_cast<const void *>(fnk::spectra_kernel<float>), P->lds_spec, done_f)
                  : lds_limit(reinterpret_cast<const void *>(fnk::spectra_kernel<double>), P->lds_spec, done_d),
              "spectra LDS limit");
    if (dtype == FN_DTYPE_F32) {
        hipLaunchKernelGGL(fnk::spectra_kernel<float>, sg, dim3(256), P->lds_spec, s,
                           reinterpret_cast<const float *>(d_rf), P->d_tw, X, S);
    } else {
        hipLaunchKernelGGL(fnk::spectra_kernel<double>, sg, dim3(256), P->lds_spec, s,
                           reinterpret_cast<const double *>(d_rf), P->d_tw, X, S);
    }
    HIP_CHECK(hipGetLastError(), "spectra_kernel");
    return FN_OK;
}

int launch_spectra(fn_plan *P, const void *d_rf, int dtype, hipStream_t s)
{
    P->executed = 1;
    return launch_spectra_into(P, d_rf, dtype, s, P->d_X, P->sargs);
}

// the beamforming arguments of one execute (image and weight buffers)
fnk::BeamArgs beam_args(const fn_plan *P, float *d_image, float *d_w, float *d_w_rx)
{
    fnk::BeamArgs A = P->args;
    A.image = reinterpret_cast<float2 *>(d_image);
    A.n_slots = beam_slots(P);
    A.slot_z = P->lbl ? 1 : P->n_groups;
    A.partial = A.n_slots > 1 ? P->d_pimg : nullptr;
    if (P->mode == 0) {
        A.weights = d_w;
        A.weights_tx = nullptr;
    } else {
        A.weights = d_w_rx;
        A.weights_tx = d_w;
    }
    return A;
}

// the 16-bit operand planes (and per-bin scales) of FP32 spectra X of MT tiles;
// nothing for FP32 plans
int launch_lowp_planes(int precision, const float2 *X, void *Xlp, float *xscale, int n_f, int MT, int KI, hipStream_t s)
{
    if (precision == FN_PRECISION_FP32) return FN_OK;
    const dim3 cg((unsigned)n_f), cb(256);
    switch (precision) {
    case FN_PRECISION_BF16X3:
        hipLaunchKernelGGL(fnk::lowp_planes_kernel<fnk::PREC_BF16X3>, cg, cb, 0, s, X, Xlp, xscale, MT, KI);
        break;
    case FN_PRECISION_FP16X3:
        hipLaunchKernelGGL(fnk::lowp_planes_kernel<fnk::PREC_FP16X3>, cg, cb, 0, s, X, Xlp, xscale, MT, KI);
        break;
    default:
        hipLaunchKernelGGL(fnk::lowp_planes_kernel<fnk::PREC_FP16>, cg, cb, 0, s, X, Xlp, xscale, MT, KI);
        break;
    }
    HIP_CHECK(hipGetLastError(), "lowp_planes_kernel");
    return FN_OK;
}

// after the beamforming launch: the partial images of a voxel's workgroups summed
// in slot order into the image (nothing when one workgroup owns each voxel)
int finish_beam(const fn_plan *P, const fnk::BeamArgs &A, hipStream_t s, float2 *dimage = nullptr)
{
    if (!A.partial) return FN_OK;
    hipLaunchKernelGGL(fnk::reduce_partials_kernel, dim3((unsigned)((P->n_vox + 255) / 256)), dim3(256), 0, s,
                       A.partial, A.n_slots, P->n_vox, A.image, dimage);
    HIP_CHECK(hipGetLastError(), "reduce_partials_kernel");
    return FN_OK;
}

// everything of an execute before the beamforming launch: the zeroed images of an
// ensemble (its frames' rows are added atomically), the band spectra and (16-bit
// forms) their planes
int prepare(fn_plan *P, const void *d_rf, int dtype, float *d_image, hipStream_t s)
{
    if (P->n_frames > 1) {
        const long n = 2 * P->n_vox * P->n_frames;
        hipLaunchKernelGGL(fnk::fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_image, n);
        HIP_CHECK(hipGetLastError(), "fill_zero_kernel");
    }
    const fnk::SpecArgs &S = P->sargs;
    if (int st = launch_spectra(P, d_rf, dtype, s)) return st;
    return launch_lowp_planes(P->precision, P->d_X, P->d_Xlp, P->d_xscale, S.n_f, S.MT, S.KI, s);
}

int execute(fn_plan *P, const void *d_rf, int dtype, float *d_image, float *d_w, float *d_w_rx, hipStream_t user)
{
    if (!P || !d_rf || !d_image) return fail(FN_ERR_ARGUMENT, "null argument");
    if (dtype != FN_DTYPE_F32 && dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    hipStream_t s = user;  // NULL: the default stream
    const fnk::BeamArgs A = beam_args(P, d_image, d_w, d_w_rx);
    Run run{};
    if (P->profiling) {
        if (!P->spare.empty()) {
            run = P->spare.back();
            P->spare.pop_back();
        } else {
            HIP_CHECK(hipEventCreate(&run.e0), "hipEventCreate");
            HIP_CHECK(hipEventCreate(&run.e1), "hipEventCreate");
            HIP_CHECK(hipEventCreate(&run.e2), "hipEventCreate");
        }
    }
    if (P->profiling) HIP_CHECK(hipEventRecord(run.e0, s), "hipEventRecord");
    if (int st = prepare(P, d_rf, dtype, d_image, s)) return st;
    if (P->profiling) HIP_CHECK(hipEventRecord(run.e1, s), "hipEventRecord");
    const long tiles = beam_tiles(P);
    hipError_t e = launch_beam(P, dim3((unsigned)tiles, (unsigned)P->bin_split, (unsigned)P->n_groups), A, s);
    if (e != hipSuccess) return hip_fail(e, "beamform_kernel");
    if (int st = finish_beam(P, A, s)) return st;
    if (P->profiling) {
        HIP_CHECK(hipEventRecord(run.e2, s), "hipEventRecord");
        P->pending.push_back(run);
    }
    return FN_OK;
}

int collect_profile(fn_plan *P)
{
    for (auto &r : P->pending) {
        HIP_CHECK(hipEventSynchronize(r.e2), "hipEventSynchronize");
        float a = 0.f, b = 0.f;
        HIP_CHECK(hipEventElapsedTime(&a, r.e0, r.e1), "hipEventElapsedTime");
        HIP_CHECK(hipEventElapsedTime(&b, r.e1, r.e2), "hipEventElapsedTime");
        P->spec_ms += a;
        P->beam_ms += b;
        P->runs += 1;
        P->spare.push_back(r);
    }
    P->pending.clear();
    return FN_OK;
}

// plane-wave configuration shared by plan creation and the no-allocation entry
// point: receive window, normalization, transmit rows (no device work)
int configure_pw(fn_plan *P, const fn_pw_params *p, Common *c)
{
    P->mode = 0;
    P->n_frames = std::max(1, p->n_frames);
    P->lbl = p->line_scan != 0;
    // windowed normalization: the CPU path tapers (windowed.m:81-86); the
    // reference MEX reads the window and normalizes as "on" (gpu_bf_das_pw_rf.cu:421-427)
    P->nw = p->normalization == FN_NORMALIZATION_WINDOWED && p->semantics == FN_SEMANTICS_CPU;
    int wsup = 0;
    if (int s = setup_window(p->window, p->semantics, &P->args.wrx, &P->wcr, &wsup)) return s;
    P->args.support_rx = wsup;
    P->args.norm_rx = p->normalization != FN_NORMALIZATION_OFF;
    P->args.clamp = p->semantics == FN_SEMANTICS_CPU;
    // transmit rows: every angle of every frame.  An ensemble's operand rows are
    // angle-major over frame slots (row = angle * slots + slot, slots = frames rounded up
    // to a multiple of 4), so that the four rows a lane holds per M-tile share one angle
    const int n_rows = p->n_angles * P->n_frames;
    P->args.ens_frames = P->n_frames > 1 ? P->n_frames : 0;
    P->args.ens_slots = P->n_frames > 1 ? (P->n_frames + 3) / 4 * 4 : 0;
    P->ens_angles = p->n_angles;
    P->args.tx_kind = p->virtual_sources ? fnk::TX_VIRTUAL : fnk::TX_PLANE;
    *c = Common{p->positions_x, p->positions_z, p->n_x, p->n_z, p->n_samples, p->n_elements, n_rows,
                p->f_s, p->element_width, p->element_pitch, p->c_0, p->index_t0,
                p->semantics, p->precision, p->output_scale, p->index_gpu, p->verbosity};
    return FN_OK;
}

int pw_plan(const fn_pw_params *p, fn_plan **out)
{
    fnh::Grid g;
    int s = fnh::validate_pw(p, &g);
    if (s) return s;
    fn_plan *P = new (std::nothrow) fn_plan();
    if (!P) return fail(FN_ERR_OUT_OF_MEMORY, "host allocation");
    Common c{};
    if ((s = configure_pw(P, p, &c))) { destroy(P); return s; }
    if ((s = create_common(P, c, g, p->f_number_values, nullptr))) { destroy(P); return s; }
    const int n_rows = P->args.n_tx;  // operand rows (ensembles: frame slots)
    // steering geometry in FP64: das_pw.m:323-327 (x_ref = sign(e_x) x_0), or the
    // virtual sources (x_vs, z_vs, |r_vs|)
    std::vector<double> txp(3 * (size_t)n_rows);
    const double M = (p->n_elements - 1) / 2.0;
    for (int row = 0; row < n_rows; ++row) {
        const int a = P->n_frames > 1 ? row / P->args.ens_slots : row;  // angle-major ensemble rows
        if (p->virtual_sources) {
            const double xv = p->virtual_sources[2 * a], zv = p->virtual_sources[2 * a + 1];
            txp[3 * row] = xv;
            txp[3 * row + 1] = zv;
            txp[3 * row + 2] = std::sqrt(xv * xv + zv * zv);
            continue;
        }
        const double th = p->steering_angles[a];
        const double ex = -std::sin(th), ez = std::cos(th);
        const double sg = (ex > 0) - (ex < 0);
        txp[3 * row] = ex;
        txp[3 * row + 1] = ez;
        txp[3 * row + 2] = sg * ((0 - M) * p->element_pitch);
    }
    if ((s = upload(&P->d_txp, txp.data(), txp.size(), "steering"))) { destroy(P); return s; }
    P->args.txp = P->d_txp;
    if (P->nw && (s = setup_windowed(P, true, p->window, p->normalization_window, false, p->window,
                                     p->normalization_window))) { destroy(P); return s; }
    use_cos_all(P);
    P->useful_flops = 8.0 * useful_macs(P, p->positions_x, p->n_x, p->positions_z, p->n_z, p->f_number_values,
                                        nullptr, p->f_s, p->c_0, p->element_pitch);
    *out = P;
    return FN_OK;
}

int sa_plan(const fn_sa_params *p, fn_plan **out)
{
    fnh::Grid g;
    int s = fnh::validate_sa(p, &g);
    if (s) return s;
    fn_plan *P = new (std::nothrow) fn_plan();
    if (!P) return fail(FN_ERR_OUT_OF_MEMORY, "host allocation");
    P->mode = 1;
    const bool cpu = p->semantics == FN_SEMANTICS_CPU;
    const bool nw_rx = cpu && p->normalization_rx == FN_NORMALIZATION_WINDOWED;
    const bool nw_tx = cpu && p->normalization_tx == FN_NORMALIZATION_WINDOWED;
    P->nw = nw_rx || nw_tx;
    int sup = 0;
    if ((s = setup_window(p->window_rx, p->semantics, &P->args.wrx, &P->wcr, &sup))) { destroy(P); return s; }
    P->args.support_rx = sup;
    if ((s = setup_window(p->window_tx, p->semantics, &P->args.wtx, &P->wct, &sup))) { destroy(P); return s; }
    P->args.support_tx = sup;
    P->args.norm_rx = p->normalization_rx != FN_NORMALIZATION_OFF;
    P->args.norm_tx = p->normalization_tx != FN_NORMALIZATION_OFF;
    P->args.clamp = 0;  // das_sa.m:354-358: no depth clamp
    Common c{p->positions_x, p->positions_z, p->n_x, p->n_z, p->n_samples, p->n_elements, p->n_elements,
             p->f_s, p->element_width, p->element_pitch, p->c_0, p->index_t0,
             p->semantics, p->precision, p->output_scale, p->index_gpu, p->verbosity};
    if ((s = create_common(P, c, g, p->f_number_rx_values, p->f_number_tx_values))) { destroy(P); return s; }
    P->useful_flops = 8.0 * useful_macs(P, p->positions_x, p->n_x, p->positions_z, p->n_z, p->f_number_rx_values,
                                        p->f_number_tx_values, p->f_s, p->c_0, p->element_pitch);
    if (P->nw && (s = setup_windowed(P, nw_rx, p->window_rx, p->normalization_window_rx, nw_tx, p->window_tx,
                                     p->normalization_window_tx))) { destroy(P); return s; }
    use_cos_all(P);
    *out = P;
    return FN_OK;
}

size_t dtype_size(int dtype) { return dtype == FN_DTYPE_F64 ? 8 : 4; }

// ---- tiling and autotuning ---------------------------------------------

// transmit tiles per group the autotuner may try: the default cap, and for fp16
// ensembles at 128 elements up to 8 (one wave per SIMD; each weight serves 128 rows)
int max_tiles(const fn_plan *P)
{
    if (P->mode == 1) return sa_mt(P->KI);
    if (P->lbl) return 1;
    if (P->nw) return std::min(P->MT_all, 2);
    if (P->n_frames > 1 && P->precision == FN_PRECISION_FP16 && P->KI == 32) return std::min(P->MT_all, 8);
    return std::min(P->MT_all, std::min(tile_cap(P->n_frames, P->precision, P->KI), pw_mt_wide(P->KI)));
}

int set_tiling(fn_plan *P, int mt, int split)
{
    if (!P) return fail(FN_ERR_ARGUMENT, "null plan");
    if (P->mode == 1) {
        if (mt != sa_mt(P->KI)) return fail(FN_ERR_ARGUMENT, "synthetic-aperture plans use %d transmit tiles", sa_mt(P->KI));
    } else if (mt < 1 || mt > max_tiles(P)) {
        return fail(FN_ERR_ARGUMENT, "m_tiles must be in 1..%d", max_tiles(P));
    }
    const int n_f = P->g.n_f;
    if (split < 1 || split > n_f) return fail(FN_ERR_ARGUMENT, "bin_split must be in 1..%d", n_f);
    if (P->nw && split != 1) return fail(FN_ERR_ARGUMENT, "windowed normalization runs with bin_split 1");
    const int planes = P->precision == FN_PRECISION_FP16 ? 2 : 4;
    const size_t step_bytes = P->precision == FN_PRECISION_FP32 ? 64 * sizeof(float2) : (size_t)planes * 64 * 2;
    const size_t lds = (size_t)mt * P->KI * step_bytes;
    if (lds > kMaxLds) return fail(FN_ERR_ARGUMENT, "transmit tile exceeds LDS");
    const int chunk = (n_f + split - 1) / split;
    P->MT_tmpl = mt;
    P->n_groups = P->lbl ? P->n_x : (P->MT_all + mt - 1) / mt;
    P->bin_split = (n_f + chunk - 1) / chunk;
    P->lds_beam = lds;
    P->args.tx_group = mt;
    P->args.bin_chunk = chunk;
    return ensure_partials(P);
}

int autotune(fn_plan *P, const void *d_rf, int dtype, float *d_image, int reps, fn_tune_result *res, hipStream_t s)
{
    if (!P || !d_rf || !d_image || !res) return fail(FN_ERR_ARGUMENT, "null argument");
    reps = std::max(1, reps);
    const int mt0 = P->MT_tmpl, split0 = P->bin_split;
    std::vector<std::pair<int, int>> cand;
    const int n_f = P->g.n_f;
    const int mt_lo = P->mode == 1 ? sa_mt(P->KI) : 1, mt_hi = max_tiles(P);
    for (int mt = mt_lo; mt <= mt_hi; ++mt) {
        // distinct group counts only: the tile count of a balanced split
        if (P->mode == 0) {
            const int groups = (P->MT_all + mt - 1) / mt;
            if ((P->MT_all + groups - 1) / groups != mt) continue;
        }
        for (int split : {1, 2, 4, 8, 16, split0})
            if (split <= n_f && (!P->nw || split == 1) &&
                std::find(cand.begin(), cand.end(), std::make_pair(mt, split)) == cand.end())
                cand.emplace_back(mt, split);
    }
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0), "hipEventCreate");
    HIP_CHECK(hipEventCreate(&e1), "hipEventCreate");
    const int prof = P->profiling;
    P->profiling = 0;
    double best = 1e300, dflt = -1.0;
    int best_mt = mt0, best_split = split0, st = FN_OK;
    for (auto &c : cand) {
        if ((st = set_tiling(P, c.first, c.second))) break;
        if ((st = execute(P, d_rf, dtype, d_image, nullptr, nullptr, s))) break;  // warm-up
        hipEventRecord(e0, s);
        for (int r = 0; r < reps && !st; ++r) st = execute(P, d_rf, dtype, d_image, nullptr, nullptr, s);
        hipEventRecord(e1, s);
        if (st) break;
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) { st = hip_fail(e, "autotune timing"); break; }
        const double per = ms / reps;
        if (c.first == mt0 && c.second == split0) dflt = per;
        if (per < best) { best = per; best_mt = c.first; best_split = c.second; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    P->profiling = prof;
    if (st) { set_tiling(P, mt0, split0); return st; }
    if ((st = set_tiling(P, best_mt, best_split))) return st;
    res->best_ms = best;
    res->default_ms = dflt;
    res->m_tiles = P->MT_tmpl;
    res->bin_split = P->bin_split;
    res->n_candidates = (int32_t)cand.size();
    return FN_OK;
}

// ---- differentiation form ----------------------------------------------

int set_tof(fn_plan *P, double c_tof, const double *delays)
{
    if (!P) return fail(FN_ERR_ARGUMENT, "null plan");
    if (!(c_tof == c_tof) || c_tof > DBL_MAX) return fail(FN_ERR_ARGUMENT, "c_tof must be finite");
    std::vector<double> d(P->n_el, 0.0);
    if (delays)
        for (int k = 0; k < P->n_el; ++k) {
            if (!(std::fabs(delays[k]) <= 1.0)) return fail(FN_ERR_ARGUMENT, "delay offset %d must be finite and |d| <= 1 s", k);
            d[k] = delays[k];
        }
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    // kernels of earlier calls on any stream may still read the delay offsets
    HIP_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    HIP_CHECK(hipMemcpy(P->d_delay, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice), "delay offsets");
    P->h_delay = d;
    P->args.ctof = c_tof > 0 ? c_tof : P->args.c0;
    ++P->tof_serial;
    return FN_OK;
}

int check_pw_fp32(const fn_plan *P, const char *what)
{
    if (!P) return fail(FN_ERR_ARGUMENT, "null plan");
    if (P->mode != 0) return fail(FN_ERR_UNSUPPORTED, "%s: plane-wave plans only", what);
    if (P->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "%s: single-frame plans only", what);
    if (P->n_el > 128) return fail(FN_ERR_UNSUPPORTED, "%s: at most 128 elements", what);
    if (P->nw) return fail(FN_ERR_UNSUPPORTED, "%s: normalizations.windowed is not differentiated", what);
    if (P->lbl) return fail(FN_ERR_UNSUPPORTED, "%s: line-scan plans are forward only", what);
    return FN_OK;
}

int execute_tangent(fn_plan *P, const void *d_rf, int dtype, float *d_image, float *d_dimage, float *d_w,
                    hipStream_t user)
{
    if (int st = check_pw_fp32(P, "tangent")) return st;
    if (!d_rf || !d_image || !d_dimage) return fail(FN_ERR_ARGUMENT, "null argument");
    if (dtype != FN_DTYPE_F32 && dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    hipStream_t s = user;  // NULL: the default stream
    if (int st = launch_spectra(P, d_rf, dtype, s)) return st;
    // groups of at most two transmit tiles keep the tangent accumulators in registers
    const int mt = std::min(P->MT_all, 2);
    const int groups = (P->MT_all + mt - 1) / mt;
    fnk::BeamArgs A = P->args;
    A.image = reinterpret_cast<float2 *>(d_image);
    A.dimage = d_dimage;
    A.weights = d_w;
    A.weights_tx = nullptr;
    A.tx_group = mt;
    // bin chunks: at least the plan's, and about 8192 workgroups (at most 8 chunks) when
    // the grid is smaller -- 11 angles, 384 x 384, 542 bins: 17.5 ms in one chunk,
    // 15.7-15.8 ms in 4 or 8
    const long tiles = (P->n_vox + 63) / 64;
    int split = P->bin_split;
    if (tiles * groups * split < 8192)
        split = std::max(split, (int)std::min<long>({(8192 + tiles * groups - 1) / (tiles * groups), 8L,
                                                    (long)std::max(1, P->g.n_f / 8)}));
    const int chunk = (P->g.n_f + split - 1) / split;
    split = (P->g.n_f + chunk - 1) / chunk;
    A.bin_chunk = chunk;
    // partial images and tangents of the workgroups that share a voxel, summed in slot order
    A.n_slots = split * groups;
    A.slot_z = groups;
    A.partial = nullptr;
    if (A.n_slots > 1) {
        const size_t need = 2 * (size_t)A.n_slots * P->n_vox;
        if (need > P->ptan_cap) {
            if (P->d_ptan) HIP_CHECK(hipFree(P->d_ptan), "hipFree");
            P->d_ptan = nullptr;
            P->ptan_cap = 0;
            HIP_CHECK(hipMalloc(&P->d_ptan, need * sizeof(float2)), "tangent partial images");
            P->ptan_cap = need;
        }
        A.partial = P->d_ptan;
    }
    const size_t lds = (size_t)mt * P->KI * 64 * sizeof(float2);
    const dim3 grid((unsigned)tiles, (unsigned)split, (unsigned)groups);
    const int wcr = wcr_general(P->wcr);
    hipError_t e = P->KI == 8 ? fnk::launch_beamform_tan_k8(mt, wcr, grid, lds, s, A)
                 : P->KI == 16 ? fnk::launch_beamform_tan_k16(mt, wcr, grid, lds, s, A)
                               : fnk::launch_beamform_tan_k32(mt, wcr, grid, lds, s, A);
    if (e != hipSuccess) return hip_fail(e, "beamform tangent kernel");
    return finish_beam(P, A, s, reinterpret_cast<float2 *>(d_dimage));
}

// voxel contraction Z_n(k,a) = s sum_v y conj(T B) / w into P->d_Z
#ifndef FN_ADJ_BLOCKS
// target workgroups of the adjoint voxel contraction: about sixteen rounds of two blocks
// per CU (11 angles, 384 x 384, 542 bins: 10.2 ms at 2048, 9.8 at 4096, 9.6 at 8192,
// 10.8 at 1024; more ranges enlarge the partials the reduction reads, 0.5 GB there)
#define FN_ADJ_BLOCKS 8192L
#endif

int contract_adjoint(fn_plan *P, const float *d_y, hipStream_t s)
{
    const int n_f = P->g.n_f;
    const int n_chunks = (n_f + 3) / 4;  // bins per block of the contraction kernel (one per wave)
    if (!P->d_Z) {
        P->adj_at = (P->n_tx + 15) / 16;
        const long groups64 = (P->n_vox + 63) / 64;  // ranges are whole 64-voxel aperture blocks
        // FN_ADJ_BLOCKS workgroups (rounded down so the last round is not a short
        // tail), bounded by the voxel blocks
        long vr = FN_ADJ_BLOCKS / ((long)n_chunks * P->adj_at);
        vr = std::max(1L, std::min(vr, groups64));
        const long blocks_per_range = (groups64 + vr - 1) / vr;
        P->adj_vpr = blocks_per_range * 64;
        P->adj_vr = (int)((groups64 + blocks_per_range - 1) / blocks_per_range);
        const size_t p_floats = (size_t)P->adj_vr * n_f * P->adj_at * 2 * 128 * 16;
        HIP_CHECK(hipMalloc(&P->d_P, p_floats * sizeof(float)), "adjoint partials");
        HIP_CHECK(hipMalloc(&P->d_Z, (size_t)n_f * P->n_tx * P->n_el * sizeof(float2)), "adjoint contraction");
        HIP_CHECK(hipMalloc(&P->d_part, 1024 * sizeof(double)), "reduction partials");
        // phase words (128 + 16 n_at per voxel) followed by 4 doubles of aperture terms per voxel
        HIP_CHECK(hipMalloc(&P->d_qtab, (size_t)P->n_vox * (128 + 16 * P->adj_at + 8) * sizeof(uint32_t)),
                  "phase tables");
        HIP_CHECK(hipMalloc(&P->d_vsc, (size_t)n_f * P->n_vox * sizeof(float)), "adjoint normalization");
    }
    fnk::AdjArgs S{};
    S.B = P->args;
    S.y = reinterpret_cast<const float2 *>(d_y);
    S.P = P->d_P;
    S.qrx = P->d_qtab;
    S.qtx = P->d_qtab + (size_t)P->n_vox * 128;
    S.vgeo = reinterpret_cast<const double *>(P->d_qtab + (size_t)P->n_vox * (128 + 16 * P->adj_at));
    S.vsc = P->d_vsc;
    S.vox_per_range = P->adj_vpr;
    S.n_at = P->adj_at;
    // bin quads fastest: the workgroups dispatched together share a voxel range, whose
    // phase-word tables then stay in L2
    const dim3 grid((unsigned)n_chunks, (unsigned)P->adj_vr, (unsigned)P->adj_at);
    // the phase tables depend on the geometry and the time of flight only: rebuilt
    // after set_tof, or when this call's stream is not the one that built them; the
    // normalization table on the geometry only: built once per stream
    const bool tables = P->adj_tab_serial != P->tof_serial || P->adj_tab_stream != s;
    const bool scale = !P->adj_sc_built || P->adj_sc_stream != s;
    hipError_t e = fnk::launch_adjoint(P->wcr, tables, scale, grid, s, S);
    if (e == hipSuccess && tables) {
        P->adj_tab_serial = P->tof_serial;
        P->adj_tab_stream = s;
    }
    if (e == hipSuccess && scale) {
        P->adj_sc_built = true;
        P->adj_sc_stream = s;
    }
    if (e != hipSuccess) return hip_fail(e, "adjoint_kernel");
    e = fnk::launch_adjoint_reduce(P->d_P, P->d_Z, P->adj_vr, n_f, P->adj_at, P->n_tx, P->n_el, s);
    if (e != hipSuccess) return hip_fail(e, "adjoint_reduce_kernel");
    return FN_OK;
}

int adjoint(fn_plan *P, const float *d_y, float *d_rf_adj, hipStream_t user)
{
    if (int st = check_pw_fp32(P, "adjoint")) return st;
    if (!d_y || !d_rf_adj) return fail(FN_ERR_ARGUMENT, "null argument");
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    hipStream_t s = user;  // NULL: the default stream
    if (int st = contract_adjoint(P, d_y, s)) return st;
    hipError_t e = fnk::launch_adjoint_idft(P->d_Z, P->d_tw, d_rf_adj, P->n_t, P->n_tx * P->n_el, P->g.N, P->g.n_lb,
                                            P->g.n_f, s);
    if (e != hipSuccess) return hip_fail(e, "adjoint_idft_kernel");
    return FN_OK;
}

int vjp(fn_plan *P, const void *d_rf, int dtype, const float *d_g, const float *d_dimage, double *d_vjp_delay,
        double *d_vjp_c, hipStream_t user)
{
    if (int st = check_pw_fp32(P, "vjp")) return st;
    if (!d_rf || !d_g || !d_vjp_delay) return fail(FN_ERR_ARGUMENT, "null argument");
    if (dtype != FN_DTYPE_F32 && dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    hipStream_t s = user;  // NULL: the default stream
    if (int st = launch_spectra(P, d_rf, dtype, s)) return st;
    if (int st = contract_adjoint(P, d_g, s)) return st;
    const double w_unit = 2.0 * 3.14159265358979323846 * P->args.f_s / P->g.N;
    hipError_t e = fnk::launch_vjp_delay(P->d_X, P->d_Z, d_vjp_delay, P->g.n_f, P->n_tx, P->n_el, P->MT_all, P->KI,
                                         P->g.n_lb, w_unit, s);
    if (e != hipSuccess) return hip_fail(e, "vjp_delay_kernel");
    if (d_dimage && d_vjp_c) {
        const int n_part = (int)std::min<long>(1024, (P->n_vox + 255) / 256);
        e = fnk::launch_dot_re(reinterpret_cast<const float2 *>(d_g), reinterpret_cast<const float2 *>(d_dimage),
                               P->n_vox, P->d_part, n_part, d_vjp_c, s);
        if (e != hipSuccess) return hip_fail(e, "dot_re_kernel");
    }
    return FN_OK;
}

// a bin of voxel (ix, iz) with non-empty aperture(s): the FP64 bounds of
// make_aperture (das_pw.m:399-414, das_sa.m:384-414)
bool bin_valid(const fn_plan *P, int ix, int iz, int b)
{
    const fnk::BeamArgs &A = P->args;
    const double x = P->h_px[ix], z = P->h_pz[iz];
    const double xp = x / A.pitch, zp = z / (2.0 * A.pitch);
    auto ok = [&](double F) {
        const double hp = zp / F;
        const double ua = A.M + (xp - hp), ub = A.M + (xp + hp);
        return std::fmax(std::ceil(ua), 0.0) <= std::fmin(std::floor(ub), (double)(A.n_el - 1));
    };
    double Fr = P->h_Frx[b];
    if (A.clamp) Fr = std::fmin(Fr, P->h_sqf[b] * std::sqrt(z / (2.88 * A.c0)));  // das_pw.m:374-377
    bool v = ok(Fr);
    if (P->mode == 1) v = v && ok(P->h_Ftx[b]);
    return v;
}

// focused RF signal of the last voxel the CPU loop processes (das_pw.m:477-481,
// das_sa.m:501-505) from the spectra of the plan's last execute: complex [N_dft]
int focused_signal(fn_plan *P, float *d_signal, hipStream_t s)
{
    if (!P || !d_signal) return fail(FN_ERR_ARGUMENT, "null argument");
    if (P->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "focused signal: single-frame plans only");
    if (P->lbl) return fail(FN_ERR_UNSUPPORTED, "focused signal: not available for line-scan plans");
    if (P->n_el > 128) return fail(FN_ERR_UNSUPPORTED, "focused signal: at most 128 elements");
    if (!P->executed) return fail(FN_ERR_ARGUMENT, "focused signal: execute the plan first");
    int vx = -1, vz = -1;  // last lateral position with a valid bin, its last such depth
    for (int ix = P->n_x - 1; ix >= 0 && vx < 0; --ix)
        for (int iz = P->n_z - 1; iz >= 0 && vx < 0; --iz)
            for (int b = 0; b < P->g.n_f; ++b)
                if (bin_valid(P, ix, iz, b)) {
                    vx = ix;
                    vz = iz;
                    break;
                }
    if (vx < 0) return fail(FN_ERR_ARGUMENT, "focused signal: no voxel has a non-empty aperture");
    HIP_CHECK(hipSetDevice(P->device), "hipSetDevice");
    if (!P->d_sigF) HIP_CHECK(hipMalloc(&P->d_sigF, (size_t)P->g.n_f * sizeof(float2)), "signal spectrum");
    hipError_t e = fnk::launch_focused_signal(P->args, P->mode, P->wcr, P->wct, P->MT_all, P->KI, P->nw, vx, vz,
                                              P->d_sigF, P->d_tw, reinterpret_cast<float2 *>(d_signal), s);
    if (e != hipSuccess) return hip_fail(e, "focused_spectrum_kernel");
    return FN_OK;
}

// run-time report of the last one-shot call of this thread with verbosity >= 1:
// the device, the launch parameters, the memory in use and per-stage times (the
// sections gpu_bf_das_pw_rf.cu:664-689, 875-903, 1128-1144 print); the gateways
// print it after their own status section
thread_local std::string t_report;

void report_line(std::string &r, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    r += buf;
}

const char *kRule = " --------------------------------------------------------------------------------\n";

void report_bytes(std::string &r, const char *what, double bytes, double total)
{
    report_line(r, "     %-31s: %12.2f KiByte (%6.2f %%)\n", what, bytes / 1024.0, total > 0 ? 100.0 * bytes / total : 0.0);
}

void report_time(std::string &r, const char *what, double ms, double total)
{
    report_line(r, "     %-34s: %10.3f ms (%6.2f %%)\n", what, ms, total > 0 ? 100.0 * ms / total : 0.0);
}

void compose_report(const fn_plan *P, size_t rf_bytes, size_t img_bytes, size_t w_bytes, int n_w, const float *ms)
{
    std::string &r = t_report;
    r.clear();
    hipDeviceProp_t prop{};
    int count = 0, drv = 0, rt = 0;
    size_t mem_free = 0, mem_total = 0;
    hipGetDeviceProperties(&prop, P->device);
    hipGetDeviceCount(&count);
    hipDriverGetVersion(&drv);
    hipRuntimeGetVersion(&rt);
    hipMemGetInfo(&mem_free, &mem_total);
    static const char *kPrec[] = {"fp32", "bf16x3", "fp16x3", "fp16"};
    static const char *kWin[] = {"range", "linear", "cosine", "cosine, full array", "cosine, cut at the aperture"};
    r += kRule;
    report_line(r, " Selected device %d of %d: %s (%s)\n", P->device, count, prop.name, prop.gcnArchName);
    r += kRule;
    report_line(r, "  %-20s: %-18d %-22s: %d\n", "compute units", prop.multiProcessorCount, "wavefront size", prop.warpSize);
    report_line(r, "  %-20s: %-11.2f MiByte %-22s: %.2f KiByte\n", "total global memory",
                (double)prop.totalGlobalMem / 1048576.0, "LDS per compute unit",
                (double)prop.maxSharedMemoryPerMultiProcessor / 1024.0);
    report_line(r, "  %-20s: %-18d %-22s: %d.%d / %d.%d\n", "registers per block", prop.regsPerBlock,
                "driver / runtime", drv / 10000000, (drv / 100000) % 100, rt / 10000000, (rt / 100000) % 100);
    report_line(r, "  %-20s: %-18s %-22s: %s\n", "contraction", kPrec[P->precision], "mode",
                P->mode == 0 ? (P->n_frames > 1 ? "plane waves, ensemble" : "plane waves") : "synthetic aperture");
    r += kRule;
    r += " launch parameters\n";
    r += kRule;
    const long tiles = beam_tiles(P);
    report_line(r, "  %-20s: %-6ld x 64 voxels %-11s %-22s: %d x %d of 16 rows\n", "voxel tiles", tiles, "",
                "transmit groups x tiles", P->n_groups, P->MT_tmpl);
    report_line(r, "  %-20s: %-3d of %-5d bins %-8s %-22s: %d (4 elements each)\n", "bin chunks", P->bin_split,
                P->args.bin_chunk, "", "element steps / lane", P->KI);
    report_line(r, "  %-20s: %-18ld %-22s: %s\n", "workgroups", tiles * P->bin_split * P->n_groups, "receive window class",
                kWin[P->wcr]);
    const double b_pos = 8.0 * (P->n_x + P->n_z);
    const double b_tab = 8.0 * P->g.n_f * (P->d_Ftx ? 3 : 2) + (P->d_txp ? 24.0 * P->n_tx : 0.0) + 8.0 * P->g.N +
                         8.0 * P->n_el;
    const double b_spec = (double)P->x_bytes + (double)P->xlp_bytes + (P->xlp_bytes ? 4.0 * P->g.n_f : 0.0);
    const double b_part = 8.0 * (double)P->pimg_cap;
    const double b_tot = b_pos + b_tab + rf_bytes + b_spec + b_part + img_bytes + (double)w_bytes * n_w;
    r += kRule;
    report_line(r, " memory consumption (available global memory: %.2f MiByte (%.2f %%))\n", mem_free / 1048576.0,
                mem_total ? 100.0 * mem_free / mem_total : 0.0);
    r += kRule;
    report_bytes(r, "voxel positions", b_pos, b_tot);
    report_bytes(r, "F-numbers, steering, twiddles", b_tab, b_tot);
    report_bytes(r, "RF signals", (double)rf_bytes, b_tot);
    report_bytes(r, "band spectra (workspace)", b_spec, b_tot);
    if (b_part > 0) report_bytes(r, "partial images", b_part, b_tot);
    report_bytes(r, "DAS image", (double)img_bytes, b_tot);
    if (n_w) report_bytes(r, "weights", (double)w_bytes * n_w, b_tot);
    report_line(r, "     %-31s  %s\n", "-------------------------------", "----------------------------");
    report_bytes(r, "total", b_tot, b_tot);
    const double t_tot = ms[0] + ms[1] + ms[2] + ms[3];
    r += kRule;
    r += " performance statistics:\n";
    r += kRule;
    report_time(r, "memory transfer: host to device", ms[0], t_tot);
    report_time(r, "band-limited DFT (spectra)", ms[1], t_tot);
    report_time(r, "beamforming kernel", ms[2], t_tot);
    report_time(r, "memory transfer: device to host", ms[3], t_tot);
    report_line(r, "     %-34s  %s\n", "----------------------------------", "------------------------");
    report_time(r, "total", t_tot, t_tot);
}

// one-shot: H2D, execute, D2H (gpu_bf_das_pw_rf.cu:730-1122); with verbosity >= 1
// the stages are timed with HIP events and reported
int one_shot(fn_plan *P, const void *rf, int dtype, float *image, float *w0, float *w1)
{
    const size_t rf_bytes = (size_t)P->n_t * P->n_el * P->n_tx * dtype_size(dtype);
    const size_t img_bytes = (size_t)P->n_vox * 2 * sizeof(float) * P->n_frames;
    const size_t w_bytes = (size_t)P->n_vox * sizeof(float);
    const bool rep = P->verbosity > 0;
    void *d_rf = nullptr;
    float *d_img = nullptr, *d_w0 = nullptr, *d_w1 = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int s = FN_OK;
    hipError_t e = hipMalloc(&d_rf, rf_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_img, img_bytes);
    if (e == hipSuccess && w0) e = hipMalloc(&d_w0, w_bytes);
    if (e == hipSuccess && w1) e = hipMalloc(&d_w1, w_bytes);
    for (int i = 0; i < 4 && rep && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    if (rep) P->profiling = 1;
    if (e == hipSuccess && rep) e = hipEventRecord(ev[0], P->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_rf, rf, rf_bytes, hipMemcpyHostToDevice, P->stream);
    if (e == hipSuccess && rep) e = hipEventRecord(ev[1], P->stream);
    if (e != hipSuccess) s = hip_fail(e, "one-shot buffers");
    if (!s) s = execute(P, d_rf, dtype, d_img, d_w0, d_w1, P->stream);
    if (!s) {
        e = rep ? hipEventRecord(ev[2], P->stream) : hipSuccess;
        if (e == hipSuccess) e = hipMemcpyAsync(image, d_img, img_bytes, hipMemcpyDeviceToHost, P->stream);
        if (e == hipSuccess && w0) e = hipMemcpyAsync(w0, d_w0, w_bytes, hipMemcpyDeviceToHost, P->stream);
        if (e == hipSuccess && w1) e = hipMemcpyAsync(w1, d_w1, w_bytes, hipMemcpyDeviceToHost, P->stream);
        if (e == hipSuccess && rep) e = hipEventRecord(ev[3], P->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
        if (e != hipSuccess) s = hip_fail(e, "one-shot results");
    }
    if (!s && rep) {
        float ms[4] = {0.f, 0.f, 0.f, 0.f};
        hipEventElapsedTime(&ms[0], ev[0], ev[1]);
        hipEventElapsedTime(&ms[3], ev[2], ev[3]);
        collect_profile(P);
        ms[1] = (float)P->spec_ms;
        ms[2] = (float)P->beam_ms;
        compose_report(P, rf_bytes, img_bytes, w_bytes, (w0 ? 1 : 0) + (w1 ? 1 : 0), ms);
    }
    for (auto &x : ev)
        if (x) hipEventDestroy(x);
    hipFree(d_rf); hipFree(d_img); hipFree(d_w0); hipFree(d_w1);
    return s;
}

}  // namespace

// ---- the no-allocation entry point (fn_das_pw_noalloc) ---------------------
// Every device table and the spectra workspace of one call live in `scratch`, a
// file-scope array of this code object: each device holds its own copy at its
// own address, which host code obtains once per device with hipGetSymbolAddress
// and then uses exclusively (the array's name in host code is only a stand-in).
// Calls on one device share the array and race if they overlap.
namespace {
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace fnk {
// host arrays reach the scratch as kernel arguments (captured at launch, so the
// caller's arrays are free to change once the call returns): 448 doubles a launch
struct UploadChunk {
    double v[448];
};

__global__ void upload_chunk_kernel(double *dst, UploadChunk c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = c.v[i];
}

// the tables pw_plan computes on the host, computed here from the raw inputs:
// twiddles e^{+j 2 pi m / N}, sqrt(f_n) per bin, the steering rows
// (das_pw.m:323-327: e_x = -sin, e_z = cos, x_ref = sign(e_x) x_0; or the
// virtual sources (x_vs, z_vs, |r_vs|)) and zero delay offsets
__global__ void noalloc_tables_kernel(float2 *tw, int N, double *sqf, int n_f, int n_lb, double f_s, double *txp,
                                      const double *raw_tx, int n_tx, int virt, double M, double pitch, double *delay,
                                      int n_el)
{
    const int n = max(max(N, n_f), max(n_tx, n_el));
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < N) {
            double sn, cs;
            sincospi(2.0 * (double)i / (double)N, &sn, &cs);
            tw[i] = make_float2((float)cs, (float)sn);
        }
        if (i < n_f) sqf[i] = sqrt(f_s * (double)(n_lb + i) / N);
        if (i < n_tx) {
            if (virt) {
                const double xv = raw_tx[2 * i], zv = raw_tx[2 * i + 1];
                txp[3 * i] = xv;
                txp[3 * i + 1] = zv;
                txp[3 * i + 2] = sqrt(xv * xv + zv * zv);
            } else {
                const double th = raw_tx[i];
                const double ex = -sin(th), ez = cos(th);
                const double sg = (ex > 0) - (ex < 0);
                txp[3 * i] = ex;
                txp[3 * i + 1] = ez;
                txp[3 * i + 2] = sg * ((0 - M) * pitch);
            }
        }
        if (i < n_el) delay[i] = 0.0;
    }
}
}  // namespace fnk

namespace {

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device`, looked up on the first call on that
// device (which is current: select_device ran) and kept
int scratch_on_device(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(FN_ERR_DEVICE, "device index %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(FN_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[device]);
    return FN_OK;
}

int upload_chunked(double *dst, const double *src, long n, hipStream_t s)
{
    fnk::UploadChunk c;
    for (long i0 = 0; i0 < n; i0 += 448) {
        const int m = (int)std::min<long>(448, n - i0);
        std::memcpy(c.v, src + i0, (size_t)m * sizeof(double));
        hipLaunchKernelGGL(fnk::upload_chunk_kernel, dim3(1), dim3(256), 0, s, dst + i0, c, m);
        HIP_CHECK(hipGetLastError(), "upload_chunk_kernel");
    }
    return FN_OK;
}

// sub-allocation of the scratch: 256-byte aligned pieces
struct Carve {
    size_t used = 0;
    size_t take(size_t bytes)
    {
        const size_t at = used;
        used += (bytes + 255) & ~(size_t)255;
        return at;
    }
};

int das_pw_noalloc(const fn_pw_params *p, const void *d_rf, int dtype, float *d_image, float *d_w, hipStream_t s)
{
    fnh::Grid g;
    int st = fnh::validate_pw(p, &g);
    if (st) return st;
    if (!d_rf) return fail(FN_ERR_RF_DATA, "RF data missing");
    if (!d_image) return fail(FN_ERR_ARGUMENT, "image buffer missing");
    if (dtype != FN_DTYPE_F32 && dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    if (p->n_frames > 1) return fail(FN_ERR_UNSUPPORTED, "fn_das_pw_noalloc: one frame per call (use a plan for ensembles)");
    if (p->line_scan) return fail(FN_ERR_UNSUPPORTED, "fn_das_pw_noalloc: line scans need a plan");
    if (p->normalization == FN_NORMALIZATION_WINDOWED && p->semantics == FN_SEMANTICS_CPU)
        return fail(FN_ERR_UNSUPPORTED, "fn_das_pw_noalloc: normalizations.windowed (CPU semantics) needs a plan");
    // the configuration of fn_plan_pw_create on a plan held on the stack: the same
    // tiling and kernels, with every device pointer in the scratch
    fn_plan P;
    Common c{};
    if ((st = configure_pw(&P, p, &c))) return st;
    if ((st = configure_common(&P, c, g))) return st;
    use_cos_all(&P);
    const int n_f = g.n_f, n_tx = c.n_tx, raw_per_tx = p->virtual_sources ? 2 : 1;
    Carve cv;
    const size_t o_X = cv.take(P.x_bytes);
    const size_t o_Xlp = cv.take(P.xlp_bytes);
    const size_t o_xs = cv.take(P.precision != FN_PRECISION_FP32 ? (size_t)n_f * sizeof(float) : 0);
    const size_t o_px = cv.take((size_t)c.n_x * sizeof(double));
    const size_t o_pz = cv.take((size_t)c.n_z * sizeof(double));
    const size_t o_F = cv.take((size_t)n_f * sizeof(double));
    const size_t o_sqf = cv.take((size_t)n_f * sizeof(double));
    const size_t o_raw = cv.take((size_t)n_tx * raw_per_tx * sizeof(double));
    const size_t o_txp = cv.take((size_t)n_tx * 3 * sizeof(double));
    const size_t o_tw = cv.take((size_t)g.N * sizeof(float2));
    const size_t o_dl = cv.take((size_t)c.n_el * sizeof(double));
    const int slots = beam_slots(&P);
    const size_t o_pi = cv.take(slots > 1 ? (size_t)slots * P.n_vox * sizeof(float2) : 0);
    if (cv.used > kScratchBytes)
        return fail(FN_ERR_UNSUPPORTED,
                    "fn_das_pw_noalloc: this call needs %zu bytes of device scratch, the entry point holds %zu "
                    "(spectra workspace %zu bytes); use a plan",
                    cv.used, kScratchBytes, P.x_bytes + P.xlp_bytes);
    if ((st = select_device(p->index_gpu))) return st;
    unsigned char *base = nullptr;
    if ((st = scratch_on_device(p->index_gpu, &base))) return st;
    P.d_X = reinterpret_cast<float2 *>(base + o_X);
    P.d_Xlp = P.xlp_bytes ? base + o_Xlp : nullptr;
    P.d_xscale = P.xlp_bytes ? reinterpret_cast<float *>(base + o_xs) : nullptr;
    P.d_px = reinterpret_cast<double *>(base + o_px);
    P.d_pz = reinterpret_cast<double *>(base + o_pz);
    P.d_Frx = reinterpret_cast<double *>(base + o_F);
    P.d_sqf = reinterpret_cast<double *>(base + o_sqf);
    P.d_txp = reinterpret_cast<double *>(base + o_txp);
    P.d_tw = reinterpret_cast<float2 *>(base + o_tw);
    P.d_delay = reinterpret_cast<double *>(base + o_dl);
    P.d_pimg = slots > 1 ? reinterpret_cast<float2 *>(base + o_pi) : nullptr;
    double *raw = reinterpret_cast<double *>(base + o_raw);
    if ((st = upload_chunked(P.d_px, p->positions_x, c.n_x, s))) return st;
    if ((st = upload_chunked(P.d_pz, p->positions_z, c.n_z, s))) return st;
    if ((st = upload_chunked(P.d_Frx, p->f_number_values, n_f, s))) return st;
    if ((st = upload_chunked(raw, p->virtual_sources ? p->virtual_sources : p->steering_angles, (long)n_tx * raw_per_tx,
                             s)))
        return st;
    const int n_tab = std::max(std::max(g.N, n_f), std::max(n_tx, c.n_el));
    hipLaunchKernelGGL(fnk::noalloc_tables_kernel, dim3((unsigned)((n_tab + 255) / 256)), dim3(256), 0, s, P.d_tw,
                       g.N, P.d_sqf, n_f, g.n_lb, c.f_s, P.d_txp, raw, n_tx, p->virtual_sources ? 1 : 0,
                       (p->n_elements - 1) / 2.0, p->element_pitch, P.d_delay, c.n_el);
    HIP_CHECK(hipGetLastError(), "noalloc_tables_kernel");
    fnk::BeamArgs &A = P.args;
    A.X = P.d_X;
    A.Xlp = P.d_Xlp;
    A.xscale = P.d_xscale;
    A.px = P.d_px; A.pz = P.d_pz; A.Frx = P.d_Frx; A.Ftx = nullptr; A.txp = P.d_txp; A.sqrt_f = P.d_sqf;
    A.delay = P.d_delay;
    // spectra (every operand slot of the workspace is written, padding as zeros) and
    // the 16-bit planes, the beamformer, and the sum of the partial images
    if ((st = prepare(&P, d_rf, dtype, d_image, s))) return st;
    const fnk::BeamArgs B = beam_args(&P, d_image, d_w, nullptr);
    hipError_t e = launch_beam(&P, dim3((unsigned)beam_tiles(&P), (unsigned)P.bin_split, (unsigned)P.n_groups), B, s);
    if (e != hipSuccess) return hip_fail(e, "beamform_kernel");
    return finish_beam(&P, B, s);
}

}  // namespace

extern "C" {

int32_t fn_das_pw_noalloc(const fn_pw_params *p, const void *d_rf, int32_t rf_dtype, float *d_image, float *d_weights,
                          void *stream)
{
    fnh::clear_error();
    return das_pw_noalloc(p, d_rf, rf_dtype, d_image, d_weights, static_cast<hipStream_t>(stream));
}

int32_t fn_device_count(int32_t *count)
{
    if (!count) return fail(FN_ERR_ARGUMENT, "null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return FN_OK;
}

int32_t fn_plan_pw_create(const fn_pw_params *p, fn_plan **plan)
{
    if (!plan) return fail(FN_ERR_ARGUMENT, "null plan pointer");
    *plan = nullptr;
    return pw_plan(p, plan);
}

int32_t fn_plan_sa_create(const fn_sa_params *p, fn_plan **plan)
{
    if (!plan) return fail(FN_ERR_ARGUMENT, "null plan pointer");
    *plan = nullptr;
    return sa_plan(p, plan);
}

int32_t fn_plan_info_get(const fn_plan *P, fn_plan_info *info)
{
    if (!P || !info) return fail(FN_ERR_ARGUMENT, "null argument");
    info->N_dft = P->g.N;
    info->n_lb = P->g.n_lb;
    info->n_ub = P->g.n_ub;
    info->n_f = P->g.n_f;
    info->n_tx = P->n_tx;
    info->n_voxels = P->n_vox;
    info->workspace_bytes = (int64_t)P->x_bytes;
    info->flops_per_frame = P->flops;
    info->useful_flops_per_frame = P->useful_flops;
    info->tile_voxels = 64;
    info->bin_split = P->bin_split;
    info->m_tiles = P->MT_tmpl;
    info->k_steps = P->KI;
    return FN_OK;
}

int32_t fn_plan_execute(fn_plan *P, const void *d_rf, int32_t dtype, float *d_image, float *d_w, float *d_w_rx,
                        void *stream)
{
    fnh::clear_error();
    return execute(P, d_rf, dtype, d_image, d_w, d_w_rx, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_autotune(fn_plan *P, const void *d_rf, int32_t dtype, float *d_image, int32_t repetitions,
                         fn_tune_result *result, void *stream)
{
    fnh::clear_error();
    return autotune(P, d_rf, dtype, d_image, repetitions, result, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_set_tiling(fn_plan *P, int32_t m_tiles, int32_t bin_split)
{
    fnh::clear_error();
    return set_tiling(P, m_tiles, bin_split);
}

int32_t fn_plan_set_tof(fn_plan *P, double c_tof, const double *delay_offsets)
{
    fnh::clear_error();
    return set_tof(P, c_tof, delay_offsets);
}

int32_t fn_plan_execute_tangent(fn_plan *P, const void *d_rf, int32_t dtype, float *d_image, float *d_dimage_dc,
                                float *d_w, void *stream)
{
    fnh::clear_error();
    return execute_tangent(P, d_rf, dtype, d_image, d_dimage_dc, d_w, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_focused_signal(fn_plan *P, float *d_signal, void *stream)
{
    fnh::clear_error();
    return focused_signal(P, d_signal, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_adjoint(fn_plan *P, const float *d_y, float *d_rf_adj, void *stream)
{
    fnh::clear_error();
    return adjoint(P, d_y, d_rf_adj, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_vjp(fn_plan *P, const void *d_rf, int32_t dtype, const float *d_g, const float *d_dimage_dc,
                    double *d_vjp_delay, double *d_vjp_c, void *stream)
{
    fnh::clear_error();
    return vjp(P, d_rf, dtype, d_g, d_dimage_dc, d_vjp_delay, d_vjp_c, reinterpret_cast<hipStream_t>(stream));
}

int32_t fn_plan_set_profiling(fn_plan *P, int32_t enable)
{
    if (!P) return fail(FN_ERR_ARGUMENT, "null plan");
    P->profiling = enable ? 1 : 0;
    return FN_OK;
}

int32_t fn_plan_get_profile(fn_plan *P, double *spectra_ms, double *beamform_ms, int32_t *n_runs)
{
    if (!P) return fail(FN_ERR_ARGUMENT, "null plan");
    int s = collect_profile(P);
    if (s) return s;
    if (spectra_ms) *spectra_ms = P->spec_ms;
    if (beamform_ms) *beamform_ms = P->beam_ms;
    if (n_runs) *n_runs = P->runs;
    P->spec_ms = P->beam_ms = 0.0;
    P->runs = 0;
    return FN_OK;
}

int32_t fn_plan_destroy(fn_plan *P)
{
    if (P) {
        hipSetDevice(P->device);
        hipStreamSynchronize(P->stream);
    }
    destroy(P);
    return FN_OK;
}

const char *fn_last_report(void) { return t_report.c_str(); }

int32_t fn_das_pw(const fn_pw_params *p, const void *data_rf, int32_t rf_dtype, float *image, float *weights)
{
    fnh::clear_error();
    t_report.clear();
    if (!data_rf) return fail(FN_ERR_RF_DATA, "RF data missing");
    if (!image) return fail(FN_ERR_ARGUMENT, "image buffer missing");
    if (rf_dtype != FN_DTYPE_F32 && rf_dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    fn_plan *P = nullptr;
    int s = pw_plan(p, &P);
    if (s) return s;
    s = one_shot(P, data_rf, rf_dtype, image, weights, nullptr);
    destroy(P);
    return s;
}

int32_t fn_das_pw_adjoint(const fn_pw_params *p, double c_tof, const double *delay_offsets, const float *image_y,
                          float *rf_adj)
{
    fnh::clear_error();
    if (!image_y || !rf_adj) return fail(FN_ERR_ARGUMENT, "image and output buffers are required");
    fn_plan *P = nullptr;
    int s = pw_plan(p, &P);
    if (s) return s;
    s = set_tof(P, c_tof, delay_offsets);
    const size_t img_bytes = (size_t)P->n_vox * 2 * sizeof(float);
    const size_t rf_bytes = (size_t)P->n_t * P->n_el * P->n_tx * sizeof(float);
    float *d_y = nullptr, *d_rf = nullptr;
    if (!s) {
        hipError_t e = hipMalloc(&d_y, img_bytes);
        if (e == hipSuccess) e = hipMalloc(&d_rf, rf_bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(d_y, image_y, img_bytes, hipMemcpyHostToDevice, P->stream);
        if (e != hipSuccess) s = hip_fail(e, "adjoint buffers");
    }
    if (!s) s = adjoint(P, d_y, d_rf, P->stream);
    if (!s) {
        hipError_t e = hipMemcpyAsync(rf_adj, d_rf, rf_bytes, hipMemcpyDeviceToHost, P->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
        if (e != hipSuccess) s = hip_fail(e, "adjoint result");
    }
    hipFree(d_y);
    hipFree(d_rf);
    destroy(P);
    return s;
}

int32_t fn_das_sa(const fn_sa_params *p, const void *data_rf, int32_t rf_dtype, float *image, float *weights_tx,
                  float *weights_rx)
{
    fnh::clear_error();
    t_report.clear();
    if (!data_rf) return fail(FN_ERR_RF_DATA, "RF data missing");
    if (!image) return fail(FN_ERR_ARGUMENT, "image buffer missing");
    if (rf_dtype != FN_DTYPE_F32 && rf_dtype != FN_DTYPE_F64) return fail(FN_ERR_RF_DATA, "RF data must be float32 or float64");
    fn_plan *P = nullptr;
    int s = sa_plan(p, &P);
    if (s) return s;
    s = one_shot(P, data_rf, rf_dtype, image, weights_tx, weights_rx);
    destroy(P);
    return s;
}

// ---- batches: many plans captured into one graph ----------------------

// plans of one kernel class (mode, element steps, transmit tiles, window classes,
// precision) beamformed by one persistent launch over their work items
struct BatchClass {
    std::vector<int> plans;
    fnk::BeamArgs *d_jobs = nullptr;  // one BeamArgs per plan (its buffers)
    int4 *d_items = nullptr;          // (job, voxel tile, bin chunk, transmit group), longest chunks first
    int *d_counter = nullptr;
    int n_items = 0, grid = 0;
};

// Row stack: plane-wave plans of a batch that differ only in their transmit rows
// (steering angles) and RF share every receive-side quantity -- voxel grid,
// frequency grid, F-numbers, window, normalization, element geometry and delays.
// Their rows are stacked into one operand (16 rows per matrix-core tile, a plan's
// rows need not start a tile) and beamformed by one ensemble launch whose epilogue
// adds every row to its own plan's image.  The apodization x phase values, one per
// (voxel, bin, element) generated on the vector ALUs, then serve all stacked rows
// of a transmit group (up to 64) instead of the few rows of one plan.
struct RowStack {
    std::vector<int> plans;  // members (batch indices), stacked in this order
    std::vector<int> row0;   // first stacked row of each member
    int rows = 0, MT_all = 0, MT_tmpl = 0, n_groups = 0, bin_split = 1, chunk = 0;
    size_t lds_beam = 0;
    float2 *d_X = nullptr;
    void *d_Xlp = nullptr;
    float *d_xscale = nullptr;
    double *d_txp = nullptr;
    int *d_row_frame = nullptr;
    float2 **d_frame_image = nullptr;
};

struct fn_batch {
    std::vector<fn_plan *> plans;
    int device = 0;
    int persistent = 0;                // plans sharing a kernel class run as one persistent launch
    int stacking = 1;                  // row stacks of plans that share their receive side
    std::vector<RowStack> stacks;
    std::vector<BatchClass> classes;
    std::vector<int> singles;          // plans launched on their own
    hipStream_t pstream = nullptr;     // stream of the persistent launches inside the capture
    std::vector<hipStream_t> streams;  // forked worker streams (default priority)
    hipStream_t capture = nullptr;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void *> rf;
    std::vector<int32_t> dtypes;
    std::vector<float *> images;
    int captures = 0;
    int64_t n_nodes = 0;
};

static void batch_release_graph(fn_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (auto &c : b->classes) {
        hipFree(c.d_jobs);
        hipFree(c.d_items);
        hipFree(c.d_counter);
    }
    for (auto &r : b->stacks) {
        hipFree(r.d_X);
        hipFree(r.d_Xlp);
        hipFree(r.d_xscale);
        hipFree(r.d_txp);
        hipFree(r.d_row_frame);
        hipFree(r.d_frame_image);
    }
    b->stacks.clear();
    b->classes.clear();
    b->singles.clear();
}

// plans whose rows can share one beamforming launch: everything but the transmit
// rows and the RF is equal (single-frame plane-wave plans up to 128 elements)
static bool row_stackable(const fn_plan *P)
{
    // beyond three tiles of rows a plan fills its own transmit groups (the ensemble
    // kernel holds at most four tiles per group against five of the plain kernel)
    return P->mode == 0 && P->n_frames == 1 && !P->lbl && !P->nw && P->KI <= 32 && P->n_tx <= 48;
}

static bool same_receive_side(const fn_plan *a, const fn_plan *b)
{
    const fnk::BeamArgs &x = a->args, &y = b->args;
    auto zeros = [](const std::vector<double> &v) { return std::all_of(v.begin(), v.end(), [](double d) { return d == 0.0; }); };
    const bool delays = (a->h_delay.empty() || zeros(a->h_delay)) && (b->h_delay.empty() || zeros(b->h_delay))
                            ? true
                            : a->h_delay == b->h_delay;
    return a->device == b->device && a->precision == b->precision && a->n_x == b->n_x && a->n_z == b->n_z &&
           a->n_t == b->n_t && a->n_el == b->n_el && a->KI == b->KI && a->g.N == b->g.N && a->g.n_lb == b->g.n_lb &&
           a->g.n_f == b->g.n_f && a->wcr == b->wcr && a->h_px == b->h_px && a->h_pz == b->h_pz &&
           a->h_Frx == b->h_Frx && delays && std::memcmp(&x.wrx, &y.wrx, sizeof(x.wrx)) == 0 &&
           x.support_rx == y.support_rx && x.norm_rx == y.norm_rx && x.clamp == y.clamp &&
           x.out_scale == y.out_scale && x.tx_kind == y.tx_kind && x.pitch == y.pitch && x.width == y.width &&
           x.c0 == y.c0 && x.ctof == y.ctof && x.f_s == y.f_s && x.t0 == y.t0 && x.M == y.M;
}

// group the stackable plans (two or more per stack), size each stack's tiling and
// build its tables: stacked steering rows, row -> plan, plan -> image; outside any capture
static int batch_plan_stacks(fn_batch *b, float *const *d_images, std::vector<char> &taken)
{
    const int n = (int)b->plans.size();
    for (int i = 0; i < n; ++i) {
        if (taken[i] || !row_stackable(b->plans[i])) continue;
        RowStack r;
        for (int j = i; j < n; ++j)
            if (!taken[j] && row_stackable(b->plans[j]) && same_receive_side(b->plans[i], b->plans[j]))
                r.plans.push_back(j);
        if (r.plans.size() < 2) continue;
        for (int j : r.plans) {
            r.row0.push_back(r.rows);
            r.rows += b->plans[j]->n_tx;
            taken[j] = 1;
        }
        const fn_plan *P0 = b->plans[i];
        r.MT_all = (r.rows + 15) / 16;
        const int max_tiles = stack_tile_cap(P0->precision, P0->KI);  // the stack kernel's register budget
        const int groups = (r.MT_all + max_tiles - 1) / max_tiles;
        r.MT_tmpl = (r.MT_all + groups - 1) / groups;
        r.n_groups = (r.MT_all + r.MT_tmpl - 1) / r.MT_tmpl;
        const int planes = P0->precision == FN_PRECISION_FP16 ? 2 : 4;
        const size_t step_bytes = P0->precision == FN_PRECISION_FP32 ? 64 * sizeof(float2) : (size_t)planes * 64 * 2;
        r.lds_beam = (size_t)r.MT_tmpl * P0->KI * step_bytes;
        // bin chunks: enough workgroups to fill the GPU several times over
        const long blocks = beam_tiles(P0) * r.n_groups;
        int split = 1;
        if (blocks < 2048) split = (int)std::min<long>((2048 + blocks - 1) / blocks, std::max(1, P0->g.n_f / 8));
        r.chunk = (P0->g.n_f + split - 1) / split;
        r.bin_split = (P0->g.n_f + r.chunk - 1) / r.chunk;
        b->stacks.push_back(r);  // freed with the graph, also on failure
        RowStack &R = b->stacks.back();
        const size_t n_f = (size_t)P0->g.n_f;
        const size_t x_bytes = n_f * R.MT_all * P0->KI * 64 * sizeof(float2);
        std::vector<int> rf(16 * (size_t)R.MT_all, -1);
        std::vector<float2 *> imgs;
        for (size_t m = 0; m < R.plans.size(); ++m) {
            for (int a = 0; a < b->plans[R.plans[m]]->n_tx; ++a) rf[R.row0[m] + a] = (int)m;
            imgs.push_back(reinterpret_cast<float2 *>(d_images[R.plans[m]]));
        }
        hipError_t e = hipMalloc(&R.d_X, x_bytes);
        if (e == hipSuccess) e = hipMemset(R.d_X, 0, x_bytes);  // rows past the last plan stay zero
        if (e == hipSuccess && P0->precision != FN_PRECISION_FP32) {
            e = hipMalloc(&R.d_Xlp, n_f * planes * R.MT_all * P0->KI * 64 * 2);
            if (e == hipSuccess) e = hipMalloc(&R.d_xscale, n_f * sizeof(float));
        }
        if (e == hipSuccess) e = hipMalloc(&R.d_txp, 3 * sizeof(double) * R.rows);
        for (size_t m = 0; m < R.plans.size() && e == hipSuccess; ++m) {
            const fn_plan *P = b->plans[R.plans[m]];
            e = hipMemcpy(R.d_txp + 3 * (size_t)R.row0[m], P->d_txp, 3 * sizeof(double) * P->n_tx, hipMemcpyDeviceToDevice);
        }
        if (e == hipSuccess) e = hipMalloc(&R.d_row_frame, rf.size() * sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(R.d_row_frame, rf.data(), rf.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&R.d_frame_image, imgs.size() * sizeof(float2 *));
        if (e == hipSuccess)
            e = hipMemcpy(R.d_frame_image, imgs.data(), imgs.size() * sizeof(float2 *), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "row stack tables");
    }
    return FN_OK;
}

// one stack on stream s: the members' images zeroed, their spectra into the stacked
// operand (each at its first row), the 16-bit planes, one ensemble launch
static int batch_run_stack(fn_batch *b, const RowStack &R, const void *const *d_rf, const int32_t *dtypes,
                           float *const *d_images, hipStream_t s)
{
    const fn_plan *P0 = b->plans[R.plans[0]];
    for (size_t m = 0; m < R.plans.size(); ++m) {
        fn_plan *P = b->plans[R.plans[m]];
        const long nz = 2 * P->n_vox;
        hipLaunchKernelGGL(fnk::fill_zero_kernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s,
                           d_images[R.plans[m]], nz);
        HIP_CHECK(hipGetLastError(), "fill_zero_kernel");
        fnk::SpecArgs S = P->sargs;
        S.row0 = R.row0[m];
        S.MT_out = R.MT_all;
        S.write_pad = 0;
        if (int st = launch_spectra_into(P, d_rf[R.plans[m]], dtypes[R.plans[m]], s, R.d_X, S)) return st;
    }
    if (int st = launch_lowp_planes(P0->precision, R.d_X, R.d_Xlp, R.d_xscale, P0->g.n_f, R.MT_all, P0->KI, s))
        return st;
    fnk::BeamArgs A = P0->args;
    A.X = R.d_X;
    A.Xlp = R.d_Xlp;
    A.xscale = R.d_xscale;
    A.txp = R.d_txp;
    A.n_tx = R.rows;
    A.tx_group = R.MT_tmpl;
    A.bin_chunk = R.chunk;
    A.ens_frames = 0;
    A.ens_slots = 0;
    A.row_frame = R.d_row_frame;
    A.frame_image = R.d_frame_image;
    A.image = nullptr;
    A.weights = nullptr;
    A.weights_tx = nullptr;
    A.partial = nullptr;
    A.n_slots = 1;
    A.slot_z = R.n_groups;
    const dim3 grid((unsigned)beam_tiles(P0), (unsigned)R.bin_split, (unsigned)R.n_groups);
    hipError_t e = fnk::kEnsLaunchers[ki_index(P0->KI)][P0->precision](R.MT_tmpl, P0->wcr, grid, R.lds_beam, s, A, 1);
    if (e != hipSuccess) return hip_fail(e, "beamform_kernel (row stack)");
    return FN_OK;
}

// kernel class of a plan for the persistent launches; -1: launched on its own
// (ensembles, windowed normalization)
static long batch_class_key(const fn_plan *P)
{
    if (P->n_frames > 1 || P->nw || P->lbl || P->KI > 32) return -1;  // persistent kernels up to 128 elements
    return ((((long)P->mode * 64 + P->KI) * 16 + P->MT_tmpl) * 8 + P->wcr) * 8 * 4 + P->wct * 4 + P->precision;
}

// group the plans into classes, upload each class's job and work-item tables and
// size its grid (resident workgroups x CUs); outside any capture
static int batch_plan_classes(fn_batch *b, float *const *d_images)
{
    std::vector<char> taken(b->plans.size(), 0);
    if (b->stacking)
        if (int st = batch_plan_stacks(b, d_images, taken)) return st;
    std::vector<std::pair<long, std::vector<int>>> groups;
    for (int i = 0; i < (int)b->plans.size(); ++i) {
        if (taken[i]) continue;
        const long key = b->persistent ? batch_class_key(b->plans[i]) : -1;
        if (key < 0) { b->singles.push_back(i); continue; }
        auto it = std::find_if(groups.begin(), groups.end(), [&](const auto &g) { return g.first == key; });
        if (it == groups.end()) groups.push_back({key, {i}});
        else it->second.push_back(i);
    }
    // a class launch waits for the spectra of all its plans: classes are cut into
    // launches of at most kPersistPlans plans, so that one launch beamforms while
    // the worker streams prepare the next
    constexpr size_t kPersistPlans = 8;
    std::vector<std::vector<int>> launches;
    for (auto &g : groups) {
        if (g.second.size() < 2) { b->singles.push_back(g.second[0]); continue; }
        const size_t n_l = (g.second.size() + kPersistPlans - 1) / kPersistPlans;
        for (size_t l = 0; l < n_l; ++l) {
            const size_t lo = g.second.size() * l / n_l, hi = g.second.size() * (l + 1) / n_l;
            launches.emplace_back(g.second.begin() + lo, g.second.begin() + hi);
        }
    }
    int n_cu = 256;
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, b->device);
    for (auto &plans : launches) {
        BatchClass c;
        c.plans = plans;
        std::vector<fnk::BeamArgs> jobs;
        std::vector<std::pair<int, int4>> items;  // (bins of the chunk, item)
        for (int j = 0; j < (int)c.plans.size(); ++j) {
            const fn_plan *P = b->plans[c.plans[j]];
            jobs.push_back(beam_args(P, d_images[c.plans[j]], nullptr, nullptr));
            const int tiles = (int)((P->n_vox + 63) / 64), chunk = P->args.bin_chunk;
            for (int bz = 0; bz < P->n_groups; ++bz)
                for (int by = 0; by < P->bin_split; ++by) {
                    const int bins = std::min(chunk, P->g.n_f - by * chunk);
                    for (int bx = 0; bx < tiles; ++bx) items.push_back({bins, make_int4(j, bx, by, bz)});
                }
        }
        std::stable_sort(items.begin(), items.end(), [](const auto &a, const auto &b2) { return a.first > b2.first; });
        std::vector<int4> it4(items.size());
        for (size_t q = 0; q < items.size(); ++q) it4[q] = items[q].second;
        c.n_items = (int)it4.size();
        const fn_plan *P0 = b->plans[c.plans[0]];
        const int kidx = P0->KI == 8 ? 0 : (P0->KI == 16 ? 1 : 2);
        int occ = 0;
        hipError_t e = fnk::kPersistLaunchers[P0->mode][kidx][P0->precision](P0->MT_tmpl, P0->wcr, P0->wct, 0,
                                                                            P0->lds_beam, nullptr, nullptr, nullptr, 0,
                                                                            nullptr, &occ);
        if (e != hipSuccess) return hip_fail(e, "persistent kernel occupancy");
        c.grid = std::max(1, std::min(c.n_items, std::max(1, occ) * n_cu));
        e = hipMalloc(&c.d_jobs, jobs.size() * sizeof(fnk::BeamArgs));
        if (e == hipSuccess) e = hipMalloc(&c.d_items, it4.size() * sizeof(int4));
        if (e == hipSuccess) e = hipMalloc(&c.d_counter, sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(c.d_jobs, jobs.data(), jobs.size() * sizeof(fnk::BeamArgs), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c.d_items, it4.data(), it4.size() * sizeof(int4), hipMemcpyHostToDevice);
        b->classes.push_back(c);  // freed with the graph, also on failure
        if (e != hipSuccess) return hip_fail(e, "persistent batch tables");
    }
    return FN_OK;
}

int32_t fn_batch_create(fn_plan *const *plans, int32_t n_plans, int32_t n_streams, fn_batch **out)
{
    fnh::clear_error();
    if (!plans || !out || n_plans < 1) return fail(FN_ERR_ARGUMENT, "plans and batch pointer required");
    *out = nullptr;
    n_streams = std::max(1, std::min(n_streams, 16));
    auto *b = new (std::nothrow) fn_batch();
    if (!b) return fail(FN_ERR_OUT_OF_MEMORY, "host allocation");
    b->device = plans[0] ? plans[0]->device : 0;
    for (int i = 0; i < n_plans; ++i) {
        if (!plans[i]) { delete b; return fail(FN_ERR_ARGUMENT, "plan %d is null", i); }
        if (plans[i]->device != b->device) { delete b; return fail(FN_ERR_ARGUMENT, "plans must share one device"); }
        b->plans.push_back(plans[i]);
    }
    int st = FN_OK;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->pstream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int i = 0; i < n_streams && e == hipSuccess; ++i) {
        hipStream_t w = nullptr;
        hipEvent_t j = nullptr;
        e = hipStreamCreateWithFlags(&w, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (w) b->streams.push_back(w);
        if (j) b->joins.push_back(j);
    }
    if (e != hipSuccess) st = hip_fail(e, "batch streams");
    if (st) { fn_batch_destroy(b); return st; }
    *out = b;
    return FN_OK;
}

int32_t fn_batch_execute(fn_batch *b, const void *const *d_rf, const int32_t *dtypes, float *const *d_images,
                         void *stream)
{
    fnh::clear_error();
    if (!b || !d_rf || !dtypes || !d_images) return fail(FN_ERR_ARGUMENT, "null argument");
    const size_t n = b->plans.size();
    for (size_t i = 0; i < n; ++i)
        if (!d_rf[i] || !d_images[i]) return fail(FN_ERR_ARGUMENT, "buffers of plan %d missing", (int)i);
    HIP_CHECK(hipSetDevice(b->device), "hipSetDevice");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bool same = b->exec && b->rf.size() == n &&
                      std::equal(b->rf.begin(), b->rf.end(), d_rf) &&
                      std::equal(b->dtypes.begin(), b->dtypes.end(), dtypes) &&
                      std::equal(b->images.begin(), b->images.end(), d_images);
    if (!same) {
        batch_release_graph(b);
        // eager pass: validates every job and raises kernel LDS limits outside the capture
        for (size_t i = 0; i < n; ++i) {
            const int prof = b->plans[i]->profiling;
            b->plans[i]->profiling = 0;
            int st = execute(b->plans[i], d_rf[i], dtypes[i], d_images[i], nullptr, nullptr, s);
            b->plans[i]->profiling = prof;
            if (st) return st;
        }
        HIP_CHECK(hipStreamSynchronize(s), "batch warm-up");
        if (int st = batch_plan_classes(b, d_images)) { batch_release_graph(b); return st; }
        // the row stacks once outside the capture (raises their kernels' LDS limits)
        for (const RowStack &R : b->stacks)
            if (int st = batch_run_stack(b, R, d_rf, dtypes, d_images, s)) { batch_release_graph(b); return st; }
        if (!b->stacks.empty()) HIP_CHECK(hipStreamSynchronize(s), "row stack warm-up");
        HIP_CHECK(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        int st = FN_OK;
        const size_t W = b->streams.size();
        hipError_t e = hipEventRecord(b->fork, b->capture);
        for (size_t w = 0; w < W && e == hipSuccess; ++w) e = hipStreamWaitEvent(b->streams[w], b->fork, 0);
        // row stacks and plans on their own (ensembles and the like, usually the longest)
        // first, one worker each; the persistent classes' spectra on the other workers
        const size_t n_single = b->stacks.size() + b->singles.size();
        size_t slot = 0;
        for (const RowStack &R : b->stacks) {
            if (e != hipSuccess || st) break;
            st = batch_run_stack(b, R, d_rf, dtypes, d_images, b->streams[slot++ % W]);
        }
        for (int i : b->singles) {
            if (e != hipSuccess || st) break;
            fn_plan *P = b->plans[i];
            hipStream_t ws = b->streams[slot++ % W];
            st = prepare(P, d_rf[i], dtypes[i], d_images[i], ws);
            const long tiles = beam_tiles(P);
            const fnk::BeamArgs A = beam_args(P, d_images[i], nullptr, nullptr);
            if (!st) e = launch_beam(P, dim3((unsigned)tiles, (unsigned)P->bin_split, (unsigned)P->n_groups), A, ws);
            if (!st && e == hipSuccess) st = finish_beam(P, A, ws);
        }
        const size_t w0 = n_single < W ? n_single : 0, n_w = n_single < W ? W - n_single : W;
        slot = 0;
        for (auto &c : b->classes) {
            if (e != hipSuccess || st) break;
            std::vector<char> used(W, 0);
            for (int i : c.plans) {
                if (st) break;
                const size_t w = w0 + slot++ % n_w;
                st = prepare(b->plans[i], d_rf[i], dtypes[i], d_images[i], b->streams[w]);
                used[w] = 1;
            }
            // the class's launch waits for exactly its own plans' spectra
            for (size_t w = 0; w < W && e == hipSuccess && !st; ++w) {
                if (!used[w]) continue;
                e = hipEventRecord(b->joins[w], b->streams[w]);
                if (e == hipSuccess) e = hipStreamWaitEvent(b->pstream, b->joins[w], 0);
            }
            if (e != hipSuccess || st) break;
            const fn_plan *P0 = b->plans[c.plans[0]];
            const int kidx = P0->KI == 8 ? 0 : (P0->KI == 16 ? 1 : 2);
            e = fnk::kPersistLaunchers[P0->mode][kidx][P0->precision](P0->MT_tmpl, P0->wcr, P0->wct, c.grid,
                                                                     P0->lds_beam, b->pstream, c.d_jobs, c.d_items,
                                                                     c.n_items, c.d_counter, nullptr);
            for (int i : c.plans) {
                if (e != hipSuccess || st) break;
                st = finish_beam(b->plans[i], beam_args(b->plans[i], d_images[i], nullptr, nullptr), b->pstream);
            }
        }
        for (size_t w = 0; w < W && e == hipSuccess && !st; ++w) {
            e = hipEventRecord(b->joins[w], b->streams[w]);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->capture, b->joins[w], 0);
        }
        if (!b->classes.empty() && e == hipSuccess && !st) {
            e = hipEventRecord(b->fork, b->pstream);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->capture, b->fork, 0);
        }
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->capture, &g);
        if (e == hipSuccess) e = e2;
        if (st) { if (g) hipGraphDestroy(g); return st; }
        if (e != hipSuccess) { if (g) hipGraphDestroy(g); return hip_fail(e, "batch capture"); }
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { batch_release_graph(b); return hip_fail(e, "hipGraphInstantiate"); }
        size_t nodes = 0;
        if (hipGraphGetNodes(g, nullptr, &nodes) == hipSuccess) b->n_nodes = (int64_t)nodes;
        b->rf.assign(d_rf, d_rf + n);
        b->dtypes.assign(dtypes, dtypes + n);
        b->images.assign(d_images, d_images + n);
        b->captures += 1;
    }
    HIP_CHECK(hipGraphLaunch(b->exec, s), "hipGraphLaunch");
    return FN_OK;
}

int32_t fn_batch_info(const fn_batch *b, int64_t *n_nodes, int32_t *n_captures)
{
    if (!b) return fail(FN_ERR_ARGUMENT, "null batch");
    if (n_nodes) *n_nodes = b->n_nodes;
    if (n_captures) *n_captures = b->captures;
    return FN_OK;
}

int32_t fn_batch_destroy(fn_batch *b)
{
    if (!b) return FN_OK;
    hipSetDevice(b->device);
    for (hipStream_t w : b->streams) hipStreamSynchronize(w);
    if (b->capture) hipStreamSynchronize(b->capture);
    if (b->pstream) hipStreamSynchronize(b->pstream);
    batch_release_graph(b);
    for (hipStream_t w : b->streams) hipStreamDestroy(w);
    for (hipEvent_t j : b->joins) hipEventDestroy(j);
    if (b->fork) hipEventDestroy(b->fork);
    if (b->capture) hipStreamDestroy(b->capture);
    if (b->pstream) hipStreamDestroy(b->pstream);
    delete b;
    return FN_OK;
}

int32_t fn_batch_set_stacking(fn_batch *b, int32_t enable)
{
    fnh::clear_error();
    if (!b) return fail(FN_ERR_ARGUMENT, "null batch");
    if (b->stacking != (enable != 0)) {
        hipSetDevice(b->device);
        batch_release_graph(b);  // captured again at the next execute
    }
    b->stacking = enable != 0;
    return FN_OK;
}

int32_t fn_batch_stacks(const fn_batch *b, int32_t *n_stacks, int32_t *n_stacked_plans)
{
    if (!b) return fail(FN_ERR_ARGUMENT, "null batch");
    int32_t k = 0;
    for (const RowStack &r : b->stacks) k += (int32_t)r.plans.size();
    if (n_stacks) *n_stacks = (int32_t)b->stacks.size();
    if (n_stacked_plans) *n_stacked_plans = k;
    return FN_OK;
}

int32_t fn_batch_set_persistent(fn_batch *b, int32_t enable)
{
    fnh::clear_error();
    if (!b) return fail(FN_ERR_ARGUMENT, "null batch");
    if (b->persistent != (enable != 0)) {
        hipSetDevice(b->device);
        batch_release_graph(b);  // captured again at the next execute
    }
    b->persistent = enable != 0;
    return FN_OK;
}

}  // extern "C"
