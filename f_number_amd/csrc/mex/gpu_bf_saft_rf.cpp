// gpu_bf_saft_rf.cpp -- MEX gateway cuda.gpu_bf_saft_rf over libfnumber_hip.
//
// Drop-in for +cuda/gpu_bf_saft_rf.cu:107 (called from das_sa.m:285):
//   [image, weights_tx, weights_rx] = gpu_bf_saft_rf(positions_x, positions_z,
//       data_RF, f_s, element_width, element_pitch, c_0, f_bounds, index_t0,
//       window_tx, window_rx, F_number_tx, F_number_rx, normalization_tx,
//       normalization_rx, index_gpu, verbosity)
// data_RF is [N_t x N_rx x N_tx] with N_rx == N_tx.  Arguments 0-6 are required.
// Argument 17 (optional) selects the semantics: 'cpu' (default, das_sa.m:289-509)
// or 'mex'.
// Build (MATLAB): mex -R2017b gpu_bf_saft_rf.cpp -I<repo>/include
//                 -L<repo>/f_number_amd -lfnumber_hip
#include "fn_mex_common.h"

#include <vector>

static const char *const GW = "gpu_bf_saft_rf";

static void window_arg(int nrhs, const mxArray *prhs[], int i, const char *what, fn_window *w)
{
    using namespace fnmex;
    *w = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (!present(nrhs, prhs, i)) return;
    if (!(isa(prhs[i], "windows.window") && mxGetNumberOfElements(prhs[i]) == 1))
        fail(GW, "InvalidWindow", what);
    if (!window_of(prhs[i], w)) fail(GW, "InvalidWindow", "Could not read window parameters!");
}

static mxArray *fnumber_arg(int nrhs, const mxArray *prhs[], int i, const char *ident, bool *own)
{
    using namespace fnmex;
    *own = false;
    if (present(nrhs, prhs, i)) {
        if (!(isa(prhs[i], "f_numbers.f_number") && mxGetNumberOfElements(prhs[i]) == 1))
            fail(GW, ident, "F-number must be a single object of the class f_numbers.f_number!");
        return const_cast<mxArray *>(prhs[i]);
    }
    mxArray *F = nullptr;
    mxArray *rhs[3] = {mxCreateDoubleScalar(70.0), mxCreateDoubleScalar(1.5), mxCreateDoubleScalar(10.0)};
    mexCallMATLAB(1, &F, 3, rhs, "f_numbers.grating.angle_lb");
    for (mxArray *r : rhs) mxDestroyArray(r);
    *own = true;
    return F;
}

static void norm_arg(int nrhs, const mxArray *prhs[], int i, int32_t *code, fn_window *w)
{
    using namespace fnmex;
    *code = FN_NORMALIZATION_ON;
    *w = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (!present(nrhs, prhs, i)) return;
    if (!(isa(prhs[i], "normalizations.normalization") && mxGetNumberOfElements(prhs[i]) == 1))
        fail(GW, "InvalidNormalization", "Normalization must be a single object of the class normalizations.normalization!");
    int c;
    if (!normalization_of(prhs[i], &c, w)) fail(GW, "InvalidFrequencyWindow", "Could not read window parameters!");
    *code = c;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    using namespace fnmex;
    int32_t n_dev = 0;
    fn_device_count(&n_dev);
    if (n_dev < 1) fail(GW, "NoCUDADevices", "Could not find any HIP devices!");
    if (nrhs < 7) fail(GW, "nrhs", "At least 7 inputs are required.");
    if (nrhs > 18) fail(GW, "nrhs", "At most 18 inputs are allowed.");
    if (nlhs > 3) fail(GW, "nlhs", "At most 3 outputs are allowed.");
    if (!(is_real_double(prhs[0]) && mxGetNumberOfDimensions(prhs[0]) == 2 && mxGetM(prhs[0]) == 1))
        fail(GW, "InvalidLateralPositions", "Lateral positions must be a real-valued row vector of type double!");
    if (!(is_real_double(prhs[1]) && mxGetNumberOfDimensions(prhs[1]) == 2 && mxGetM(prhs[1]) == 1))
        fail(GW, "InvalidAxialPositions", "Axial positions must be a real-valued row vector of type double!");
    if (!(is_real_double(prhs[2]) && mxGetNumberOfDimensions(prhs[2]) == 3))
        fail(GW, "InvalidRFData", "RF data must be a real-valued array of type double with 3 dimensions!");
    const mwSize *dims = mxGetDimensions(prhs[2]);
    if (dims[1] != dims[2]) fail(GW, "InvalidRFData", "Size of RF data must be equal along 2nd and 3rd dimensions!");
    if (!is_real_double_scalar(prhs[3])) fail(GW, "InvalidSamplingRate", "Sampling rate must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[4])) fail(GW, "InvalidElementWidth", "Element width must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[5])) fail(GW, "InvalidElementPitch", "Element pitch must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[6])) fail(GW, "InvalidSoundSpeed", "Speed of sound must be a real-valued scalar of type double!");

    fn_sa_params p{};
    p.struct_size = sizeof(fn_sa_params);
    p.positions_x = mxGetPr(prhs[0]);
    p.n_x = (int32_t)mxGetN(prhs[0]);
    p.positions_z = mxGetPr(prhs[1]);
    p.n_z = (int32_t)mxGetN(prhs[1]);
    p.n_samples = (int32_t)dims[0];
    p.n_elements = (int32_t)dims[1];
    p.f_s = *mxGetPr(prhs[3]);
    p.element_width = *mxGetPr(prhs[4]);
    p.element_pitch = *mxGetPr(prhs[5]);
    p.c_0 = *mxGetPr(prhs[6]);
    p.f_lb = 2.2250738585072014e-308;
    p.f_ub = p.f_s / 2;
    if (present(nrhs, prhs, 7)) {
        if (!(is_real_double(prhs[7]) && mxGetNumberOfElements(prhs[7]) == 2))
            fail(GW, "InvalidFrequencyBounds", "Frequency bounds must be a real-valued vector with two entries of type double!");
        p.f_lb = mxGetPr(prhs[7])[0];
        p.f_ub = mxGetPr(prhs[7])[1];
    }
    if (present(nrhs, prhs, 8)) {
        if (!is_real_double_scalar(prhs[8])) fail(GW, "InvalidTimeShift", "Time index must be a real-valued scalar of type double!");
        p.index_t0 = *mxGetPr(prhs[8]);
    }
    window_arg(nrhs, prhs, 9, "Transmit window must be a single object of the class windows.window!", &p.window_tx);
    window_arg(nrhs, prhs, 10, "Receive window must be a single object of the class windows.window!", &p.window_rx);
    bool own_tx, own_rx;
    mxArray *Ftx = fnumber_arg(nrhs, prhs, 11, "InvalidFNumber", &own_tx);
    mxArray *Frx = fnumber_arg(nrhs, prhs, 12, "InvalidFNumber", &own_rx);
    norm_arg(nrhs, prhs, 13, &p.normalization_tx, &p.normalization_window_tx);
    norm_arg(nrhs, prhs, 14, &p.normalization_rx, &p.normalization_window_rx);
    p.index_gpu = 0;
    if (present(nrhs, prhs, 15)) {
        if (!is_real_double_scalar(prhs[15])) fail(GW, "InvalidCUDADevice", "Device index must be a real-valued scalar of type double!");
        p.index_gpu = (int32_t)*mxGetPr(prhs[15]);
    }
    if (p.index_gpu < 0 || p.index_gpu >= n_dev) fail(GW, "InvalidCUDADevice", "Selected device does not exist!");
    p.verbosity = 1;
    if (present(nrhs, prhs, 16)) {
        if (!is_real_double_scalar(prhs[16])) fail(GW, "InvalidVerbosityLevel", "Verbosity level must be a real-valued scalar of type double!");
        p.verbosity = (int32_t)*mxGetPr(prhs[16]);
    }
    p.semantics = semantics_arg(nrhs, prhs, 17, GW);
    p.precision = FN_PRECISION_FP32;
    p.output_scale = FN_SCALE_MEX;

    int32_t N = 0, lb = 0, ub = 0;
    check(fn_sa_frequency_grid(&p, &N, &lb, &ub), GW);
    std::vector<double> r(ub - lb + 1);
    check(fn_element_pitch_over_lambda(p.element_pitch, p.f_s, p.c_0, N, lb, ub, r.data()), GW);
    std::vector<double> Ft = compute_values(Ftx, r), Fr = compute_values(Frx, r);
    if (p.verbosity > 0) {
        // status section (gpu_bf_saft_rf.cu, as gpu_bf_das_pw_rf.cu:609-632), before the computation
        status_common("gpu_bf_saft_rf: DAS, synthetic aperture", p.n_x, p.n_z, p.n_samples, p.n_elements, p.f_s, p.c_0,
                      p.f_lb, p.f_ub, lb, ub, p.index_t0);
        mexPrintf("  %-20s: %-6d %12s %-22s: %d\n", "number tx channels", p.n_elements, "", "DFT order", N);
        mexPrintf("  %-20s: %s\n", "window (tx)", present(nrhs, prhs, 9) ? name_of(prhs[9]).c_str() : "boxcar");
        mexPrintf("  %-20s: %s\n", "window (rx)", present(nrhs, prhs, 10) ? name_of(prhs[10]).c_str() : "boxcar");
        mexPrintf("  %-20s: %s\n", "F-number (tx)", name_of(Ftx).c_str());
        mexPrintf("  %-20s: %s\n", "F-number (rx)", name_of(Frx).c_str());
        mexPrintf("  %-20s: %s\n", "normalization (tx)", present(nrhs, prhs, 13) ? name_of(prhs[13]).c_str() : "on");
        mexPrintf("  %-20s: %s\n", "normalization (rx)", present(nrhs, prhs, 14) ? name_of(prhs[14]).c_str() : "on");
        mexPrintf("  %-20s: %s\n", "semantics", p.semantics == FN_SEMANTICS_MEX ? "mex" : "cpu");
    }
    if (own_tx) mxDestroyArray(Ftx);
    if (own_rx) mxDestroyArray(Frx);
    p.f_number_tx_values = Ft.data();
    p.f_number_rx_values = Fr.data();
    p.n_f = (int32_t)Ft.size();

    const size_t nv = (size_t)p.n_x * p.n_z;
    std::vector<float> img(2 * nv), wt(nv), wr(nv);
    check(fn_das_sa(&p, mxGetPr(prhs[2]), FN_DTYPE_F64, img.data(), wt.data(), wr.data()), GW);
    // device, launch, memory and timing sections (gpu_bf_saft_rf.cu, as gpu_bf_das_pw_rf.cu:664-689, 875-903,
    // 1128-1144)
    if (p.verbosity > 0) mexPrintf("%s", fn_last_report());
    plhs[0] = image_out(img, p.n_z, p.n_x);
    if (nlhs > 1) plhs[1] = weights_out(wt, p.n_z, p.n_x);
    if (nlhs > 2) plhs[2] = weights_out(wr, p.n_z, p.n_x);
}
