// fn_beamform_inst.hip -- explicit instantiations of beamform_kernel for one
// (mode, KI, precision) triple, compiled once per triple (-DINST_MODE=0|1
// -DINST_KI=8|16|32|48|64 -DINST_PREC=0..3) so the instantiation set builds in parallel.
// KI = 48 and 64 (arrays of up to 192 and 256 elements) provide the one-launch
// forward kernel only: plane waves in groups of at most 3 / 2 transmit tiles (the
// staged tile within 80 KiB), synthetic aperture in groups of 4 transmit tiles.
#include "fn_kernels.hip.h"

#ifndef INST_MODE
#error "INST_MODE must be defined"
#endif
#ifndef INST_KI
#error "INST_KI must be defined"
#endif
#ifndef INST_PREC
#define INST_PREC 0
#endif

#define FN_WIDE (INST_KI > 32)
// transmit tiles per plane-wave group: up to 5 (tile_cap); wide arrays while the
// staged FP32 tile (MT x KI x 512 bytes) stays within 80 KiB
#define FN_PW_MT_MAX (INST_KI <= 32 ? 5 : 81920 / (INST_KI * 512))

#define FN_CAT6(a, b, c, d, e, f) a##b##c##d##e##f
#define FN_LAUNCHER(m, k, p) FN_CAT6(launch_beamform_m, m, _k, k, _p, p)

namespace fnk {

// Raise the dynamic-LDS limit of a kernel once per process and size (`done` is a
// static of the calling launcher, one per kernel instantiation), so that a launch
// sequence captured into a graph issues no attribute calls.
template <typename K>
static hipError_t set_lds_once(K kern, size_t lds, size_t &done)
{
    if (lds <= 65536 || lds <= done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e == hipSuccess) done = lds;
    return e;
}

// The FP32 kernels and the 16-bit ones at KI = 32 stage their spectra tile in static
// LDS (per-chunk arrays); the 16-bit forms at KI < 32 take it as dynamic LDS of the
// plan's size.
static inline size_t dyn_lds(size_t lds) { return (INST_PREC == 0 || INST_KI == 32) ? 0 : lds; }

template <int MT, int WCR, int WCT>
static hipError_t launch_one(dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    auto kern = beamform_kernel<INST_MODE, MT, INST_KI, WCR, WCT, INST_PREC>;
    static size_t lds_done = 0;
    lds = dyn_lds(lds);
    if (hipError_t e = set_lds_once(kern, lds, lds_done)) return e;
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, A);
    return hipGetLastError();
}

#if INST_MODE == 0
// plane waves: any number of transmit M-tiles per group (1..8), receive window class
template <int MT>
static hipError_t launch_mt(int wcr, dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    switch (wcr) {
    case WC_RANGE: return launch_one<MT, WC_RANGE, WC_RANGE>(grid, lds, s, A);
    case WC_LINEAR: return launch_one<MT, WC_LINEAR, WC_RANGE>(grid, lds, s, A);
    case WC_COS_ALL: return launch_one<MT, WC_COS_ALL, WC_RANGE>(grid, lds, s, A);
    case WC_COS_RF: return launch_one<MT, WC_COS_RF, WC_RANGE>(grid, lds, s, A);
    default: return launch_one<MT, WC_COS, WC_RANGE>(grid, lds, s, A);
    }
}

#if INST_PREC == 0 && !FN_WIDE
// tangent (dI / dc_tof) variant: groups of at most two transmit tiles
#define FN_TAN_LAUNCHER(k) FN_CAT4(launch_beamform_tan, _k, k, )
#define FN_CAT4(a, b, c, d) a##b##c##d
template <int MT, int WCR>
static hipError_t launch_tan_one(dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    auto kern = beamform_kernel<0, MT, INST_KI, WCR, WC_RANGE, 0, true>;
    (void)lds;  // FP32: static LDS
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t FN_TAN_LAUNCHER(INST_KI)(int mt, int wcr, dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    if (mt == 1) {
        switch (wcr) {
        case WC_RANGE: return launch_tan_one<1, WC_RANGE>(grid, lds, s, A);
        case WC_LINEAR: return launch_tan_one<1, WC_LINEAR>(grid, lds, s, A);
        default: return launch_tan_one<1, WC_COS>(grid, lds, s, A);
        }
    }
    if (mt == 2) {
        switch (wcr) {
        case WC_RANGE: return launch_tan_one<2, WC_RANGE>(grid, lds, s, A);
        case WC_LINEAR: return launch_tan_one<2, WC_LINEAR>(grid, lds, s, A);
        default: return launch_tan_one<2, WC_COS>(grid, lds, s, A);
        }
    }
    return hipErrorInvalidValue;
}

// line scan: block z = grid column = its transmit row, one transmit tile
#define FN_LBL_LAUNCHER(k) FN_CAT4(launch_beamform_lbl, _k, k, )
template <int WCR>
static hipError_t launch_lbl_one(dim3 grid, hipStream_t s, const BeamArgs &A)
{
    auto kern = beamform_kernel<0, 1, INST_KI, WCR, WC_RANGE, 0, false, false, false, true>;
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, A);  // FP32: static LDS
    return hipGetLastError();
}

hipError_t FN_LBL_LAUNCHER(INST_KI)(int wcr, dim3 grid, hipStream_t s, const BeamArgs &A)
{
    switch (wcr) {
    case WC_RANGE: return launch_lbl_one<WC_RANGE>(grid, s, A);
    case WC_LINEAR: return launch_lbl_one<WC_LINEAR>(grid, s, A);
    case WC_COS_ALL: return launch_lbl_one<WC_COS_ALL>(grid, s, A);
    case WC_COS_RF: return launch_lbl_one<WC_COS_RF>(grid, s, A);
    default: return launch_lbl_one<WC_COS>(grid, s, A);
    }
}
#endif

#if INST_MODE == 0 && !FN_WIDE
// ensemble variant: frames x angles as transmit rows, up to four tiles per group (fp16
// at 128 elements up to eight)
#define FN_ENS_LAUNCHER(k, p) FN_CAT6(launch_beamform_ens, _k, k, _p, p, )
template <int MT, int WCR, bool STK>
static hipError_t launch_ens_one(dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    auto kern = beamform_kernel<0, MT, INST_KI, WCR, WC_RANGE, INST_PREC, false, true, false, false, STK>;
    static size_t lds_done = 0;
    lds = dyn_lds(lds);
    if (hipError_t e = set_lds_once(kern, lds, lds_done)) return e;
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, s, A);
    return hipGetLastError();
}

template <int MT, bool STK>
static hipError_t launch_ens_mt(int wcr, dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    switch (wcr) {
    case WC_RANGE: return launch_ens_one<MT, WC_RANGE, STK>(grid, lds, s, A);
    case WC_LINEAR: return launch_ens_one<MT, WC_LINEAR, STK>(grid, lds, s, A);
    case WC_COS_ALL: return launch_ens_one<MT, WC_COS_ALL, STK>(grid, lds, s, A);
    case WC_COS_RF: return launch_ens_one<MT, WC_COS_RF, STK>(grid, lds, s, A);
    default: return launch_ens_one<MT, WC_COS, STK>(grid, lds, s, A);
    }
}

// stack != 0: a row stack (rows of several acquisitions, each its own transmit phase)
hipError_t FN_ENS_LAUNCHER(INST_KI, INST_PREC)(int mt, int wcr, dim3 grid, size_t lds, hipStream_t s,
                                                const BeamArgs &A, int stack)
{
    if (stack) {
        // groups of at most the default tile count (the row stacks' tiling)
        switch (mt) {
        case 1: return launch_ens_mt<1, true>(wcr, grid, lds, s, A);
        case 2: return launch_ens_mt<2, true>(wcr, grid, lds, s, A);
        case 3: return launch_ens_mt<3, true>(wcr, grid, lds, s, A);
        case 4: return launch_ens_mt<4, true>(wcr, grid, lds, s, A);
#if INST_PREC == 3 && INST_KI == 32
        case 5: return launch_ens_mt<5, true>(wcr, grid, lds, s, A);
#endif
        default: return hipErrorInvalidValue;
        }
    }
    switch (mt) {
    case 1: return launch_ens_mt<1, false>(wcr, grid, lds, s, A);
    case 2: return launch_ens_mt<2, false>(wcr, grid, lds, s, A);
    case 3: return launch_ens_mt<3, false>(wcr, grid, lds, s, A);
    case 4: return launch_ens_mt<4, false>(wcr, grid, lds, s, A);
#if INST_PREC == 0 && INST_KI == 32
    // FP32 at 128 elements: 5 tiles (the default)
    case 5: return launch_ens_mt<5, false>(wcr, grid, lds, s, A);
#endif
#if INST_PREC == 3 && INST_KI == 32
    // fp16 at 128 elements: 5 tiles (the default), and 6-8 at one wave per SIMD
    // (offered to the autotuner; at 64 elements 7 and 8 tiles would spill)
    case 5: return launch_ens_mt<5, false>(wcr, grid, lds, s, A);
    case 6: return launch_ens_mt<6, false>(wcr, grid, lds, s, A);
    case 7: return launch_ens_mt<7, false>(wcr, grid, lds, s, A);
    case 8: return launch_ens_mt<8, false>(wcr, grid, lds, s, A);
#endif
    default: return hipErrorInvalidValue;
    }
}
#endif

hipError_t FN_LAUNCHER(INST_MODE, INST_KI, INST_PREC)(int mt, int wcr, int wct, dim3 grid, size_t lds, hipStream_t s,
                                            const BeamArgs &A)
{
    (void)wct;
    if (mt < 1 || mt > FN_PW_MT_MAX) return hipErrorInvalidValue;
    switch (mt) {
    case 1: return launch_mt<1>(wcr, grid, lds, s, A);
    case 2: return launch_mt<2>(wcr, grid, lds, s, A);
#if FN_PW_MT_MAX >= 3
    case 3: return launch_mt<3>(wcr, grid, lds, s, A);
#endif
#if FN_PW_MT_MAX >= 5
    case 4: return launch_mt<4>(wcr, grid, lds, s, A);
    case 5: return launch_mt<5>(wcr, grid, lds, s, A);  // at most 5 tiles per group (tile_cap)
#endif
    default: return hipErrorInvalidValue;
    }
}
#else
// synthetic aperture: transmit rows = elements, MT = 4 KI / 16 = KI / 4 tiles (all
// transmitters in one group); wide arrays: groups of 4 tiles
constexpr int SA_MT = INST_KI <= 32 ? INST_KI / 4 : 4;
template <int WCR>
static hipError_t launch_wct(int wct, dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    switch (wct) {
    case WC_RANGE: return launch_one<SA_MT, WCR, WC_RANGE>(grid, lds, s, A);
    case WC_LINEAR: return launch_one<SA_MT, WCR, WC_LINEAR>(grid, lds, s, A);
    case WC_COS_RF: return launch_one<SA_MT, WCR, WC_COS_RF>(grid, lds, s, A);
    default: return launch_one<SA_MT, WCR, WC_COS>(grid, lds, s, A);
    }
}

hipError_t FN_LAUNCHER(INST_MODE, INST_KI, INST_PREC)(int mt, int wcr, int wct, dim3 grid, size_t lds, hipStream_t s,
                                            const BeamArgs &A)
{
    if (mt != SA_MT) return hipErrorInvalidValue;
    switch (wcr) {
    case WC_RANGE: return launch_wct<WC_RANGE>(wct, grid, lds, s, A);
    case WC_LINEAR: return launch_wct<WC_LINEAR>(wct, grid, lds, s, A);
    case WC_COS_ALL: return launch_wct<WC_COS_ALL>(wct, grid, lds, s, A);
    case WC_COS_RF: return launch_wct<WC_COS_RF>(wct, grid, lds, s, A);
    default: return launch_wct<WC_COS>(wct, grid, lds, s, A);
    }
}
#endif

#if !FN_WIDE
// persistent batch launches: one kernel per class over the work items of many
// plans (grid <= 0: only report the resident workgroups per CU in *occupancy)
#define FN_PERSIST_LAUNCHER(m, k, p) FN_CAT6(launch_beamform_persistent_m, m, _k, k, _p, p)
template <int MT, int WCR, int WCT>
static hipError_t launch_persist_one(int grid, size_t lds, hipStream_t s, const BeamArgs *jobs, const int4 *items,
                                     int n_items, int *counter, int *occupancy)
{
    auto kern = beamform_persistent_kernel<INST_MODE, MT, INST_KI, WCR, WCT, INST_PREC>;
    static size_t lds_done = 0;
    lds = dyn_lds(lds);
    if (hipError_t e = set_lds_once(kern, lds, lds_done)) return e;
    if (grid <= 0) {
        int n = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 256, lds);
        *occupancy = n;
        return e;
    }
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(256), lds, s, jobs, items, n_items, counter);
    return hipGetLastError();
}

template <int MT, int WCT>
static hipError_t launch_persist_wcr(int wcr, int grid, size_t lds, hipStream_t s, const BeamArgs *jobs,
                                     const int4 *items, int n_items, int *counter, int *occ)
{
    switch (wcr) {
    case WC_RANGE: return launch_persist_one<MT, WC_RANGE, WCT>(grid, lds, s, jobs, items, n_items, counter, occ);
    case WC_LINEAR: return launch_persist_one<MT, WC_LINEAR, WCT>(grid, lds, s, jobs, items, n_items, counter, occ);
    case WC_COS_ALL: return launch_persist_one<MT, WC_COS_ALL, WCT>(grid, lds, s, jobs, items, n_items, counter, occ);
    case WC_COS_RF: return launch_persist_one<MT, WC_COS_RF, WCT>(grid, lds, s, jobs, items, n_items, counter, occ);
    default: return launch_persist_one<MT, WC_COS, WCT>(grid, lds, s, jobs, items, n_items, counter, occ);
    }
}

hipError_t FN_PERSIST_LAUNCHER(INST_MODE, INST_KI, INST_PREC)(int mt, int wcr, int wct, int grid, size_t lds,
                                                            hipStream_t s, const BeamArgs *jobs, const int4 *items,
                                                            int n_items, int *counter, int *occupancy)
{
#if INST_MODE == 0
    (void)wct;
    switch (mt) {
    case 1: return launch_persist_wcr<1, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    case 2: return launch_persist_wcr<2, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    case 3: return launch_persist_wcr<3, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    case 4: return launch_persist_wcr<4, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    case 5: return launch_persist_wcr<5, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    default: return hipErrorInvalidValue;
    }
#else
    if (mt != INST_KI / 4) return hipErrorInvalidValue;
    switch (wct) {
    case WC_RANGE: return launch_persist_wcr<INST_KI / 4, WC_RANGE>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    case WC_LINEAR: return launch_persist_wcr<INST_KI / 4, WC_LINEAR>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    default: return launch_persist_wcr<INST_KI / 4, WC_COS>(wcr, grid, lds, s, jobs, items, n_items, counter, occupancy);
    }
#endif
}

#endif  // !FN_WIDE

#if INST_PREC == 0 && !FN_WIDE
// windowed-normalization variant (FP32, general window classes): plane-wave
// groups of at most two transmit tiles, synthetic aperture at its fixed tiling
#define FN_NW_LAUNCHER(m, k) FN_CAT6(launch_beamform_nw_m, m, _k, k, , )
template <int MT, int WCR, int WCT>
static hipError_t launch_nw_one(dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    auto kern = beamform_kernel<INST_MODE, MT, INST_KI, WCR, WCT, 0, false, false, true>;
    (void)lds;  // FP32: static LDS
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, A);
    return hipGetLastError();
}

template <int MT, int WCT>
static hipError_t launch_nw_wcr(int wcr, dim3 grid, size_t lds, hipStream_t s, const BeamArgs &A)
{
    switch (wcr) {
    case WC_RANGE: return launch_nw_one<MT, WC_RANGE, WCT>(grid, lds, s, A);
    case WC_LINEAR: return launch_nw_one<MT, WC_LINEAR, WCT>(grid, lds, s, A);
    case WC_COS: return launch_nw_one<MT, WC_COS, WCT>(grid, lds, s, A);
    default: return hipErrorInvalidValue;
    }
}

hipError_t FN_NW_LAUNCHER(INST_MODE, INST_KI)(int mt, int wcr, int wct, dim3 grid, size_t lds, hipStream_t s,
                                             const BeamArgs &A)
{
#if INST_MODE == 0
    (void)wct;
    if (mt == 1) return launch_nw_wcr<1, WC_RANGE>(wcr, grid, lds, s, A);
    if (mt == 2) return launch_nw_wcr<2, WC_RANGE>(wcr, grid, lds, s, A);
    return hipErrorInvalidValue;
#else
    if (mt != INST_KI / 4) return hipErrorInvalidValue;
    switch (wct) {
    case WC_RANGE: return launch_nw_wcr<INST_KI / 4, WC_RANGE>(wcr, grid, lds, s, A);
    case WC_LINEAR: return launch_nw_wcr<INST_KI / 4, WC_LINEAR>(wcr, grid, lds, s, A);
    case WC_COS: return launch_nw_wcr<INST_KI / 4, WC_COS>(wcr, grid, lds, s, A);
    default: return hipErrorInvalidValue;
    }
#endif
}
#endif

}  // namespace fnk
