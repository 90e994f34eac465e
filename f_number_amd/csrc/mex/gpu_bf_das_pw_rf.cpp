// gpu_bf_das_pw_rf.cpp -- MEX gateway cuda.gpu_bf_das_pw_rf over libfnumber_hip.
//
// Drop-in for +cuda/gpu_bf_das_pw_rf.cu:110 (called from das_pw.m:298):
//   [image, weights] = gpu_bf_das_pw_rf(positions_x, positions_z, data_RF, f_s,
//       steering_angle, element_width, element_pitch, c_0, f_bounds, index_t0,
//       window, F_number, normalization, index_gpu, verbosity)
// Arguments 0-7 are required; [] selects a default (boxcar window,
// grating.angle_lb(70, 1.5, 10), normalization on, GPU 0, verbosity 1, as the
// reference gateway).  Extension: steering_angle may hold N_a angles when
// data_RF is [N_t x N_el x N_a]; the result is then the compounded image.
// Argument 15 (optional) selects the semantics: 'cpu' (default: the MATLAB CPU
// path, das_pw.m:302-485) or 'mex' (the reference CUDA kernels' choices).
// Build (MATLAB): mex -R2017b gpu_bf_das_pw_rf.cpp -I<repo>/include
//                 -L<repo>/f_number_amd -lfnumber_hip
#include "fn_mex_common.h"

#include <cmath>
#include <vector>

static const char *const GW = "gpu_bf_das_pw_rf";

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    using namespace fnmex;
    int32_t n_dev = 0;
    fn_device_count(&n_dev);
    if (n_dev < 1) fail(GW, "NoCUDADevices", "Could not find any HIP devices!");
    if (nrhs < 8) fail(GW, "nrhs", "At least 8 inputs are required.");
    if (nrhs > 16) fail(GW, "nrhs", "At most 16 inputs are allowed.");
    if (nlhs > 2) fail(GW, "nlhs", "At most 2 outputs are allowed.");

    if (!(is_real_double(prhs[0]) && mxGetNumberOfDimensions(prhs[0]) == 2 && mxGetM(prhs[0]) == 1))
        fail(GW, "InvalidLateralPositions", "Lateral positions must be a real-valued row vector of type double!");
    if (!(is_real_double(prhs[1]) && mxGetNumberOfDimensions(prhs[1]) == 2 && mxGetM(prhs[1]) == 1))
        fail(GW, "InvalidAxialPositions", "Axial positions must be a real-valued row vector of type double!");
    const mwSize nd = mxGetNumberOfDimensions(prhs[2]);
    if (!(is_real_double(prhs[2]) && (nd == 2 || nd == 3)))
        fail(GW, "InvalidRFData", "RF data must be a real-valued matrix of type double!");
    const mwSize *dims = mxGetDimensions(prhs[2]);
    const int n_t = (int)dims[0], n_el = (int)dims[1], n_a = nd == 3 ? (int)dims[2] : 1;
    if (!is_real_double_scalar(prhs[3])) fail(GW, "InvalidSamplingRate", "Sampling rate must be a real-valued scalar of type double!");
    if (!(is_real_double(prhs[4]) && (int)mxGetNumberOfElements(prhs[4]) == n_a))
        fail(GW, "InvalidSteeringAngle", "Steering angle must be a real-valued scalar of type double (one per RF data set)!");
    if (!is_real_double_scalar(prhs[5])) fail(GW, "InvalidElementWidth", "Element width must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[6])) fail(GW, "InvalidElementPitch", "Element pitch must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[7])) fail(GW, "InvalidSoundSpeed", "Speed of sound must be a real-valued scalar of type double!");

    fn_pw_params p{};
    p.struct_size = sizeof(fn_pw_params);
    p.positions_x = mxGetPr(prhs[0]);
    p.n_x = (int32_t)mxGetN(prhs[0]);
    p.positions_z = mxGetPr(prhs[1]);
    p.n_z = (int32_t)mxGetN(prhs[1]);
    p.n_samples = n_t;
    p.n_elements = n_el;
    p.n_angles = n_a;
    p.steering_angles = mxGetPr(prhs[4]);
    p.f_s = *mxGetPr(prhs[3]);
    p.element_width = *mxGetPr(prhs[5]);
    p.element_pitch = *mxGetPr(prhs[6]);
    p.c_0 = *mxGetPr(prhs[7]);
    p.f_lb = 2.2250738585072014e-308;  // DBL_MIN (gpu_bf_das_pw_rf.cu:282)
    p.f_ub = p.f_s / 2;
    if (present(nrhs, prhs, 8)) {
        if (!(is_real_double(prhs[8]) && mxGetNumberOfElements(prhs[8]) == 2))
            fail(GW, "InvalidFrequencyBounds", "Frequency bounds must be a real-valued vector with two entries of type double!");
        p.f_lb = mxGetPr(prhs[8])[0];
        p.f_ub = mxGetPr(prhs[8])[1];
    }
    p.index_t0 = 0.0;
    if (present(nrhs, prhs, 9)) {
        if (!is_real_double_scalar(prhs[9])) fail(GW, "InvalidTimeShift", "Time index must be a real-valued scalar of type double!");
        p.index_t0 = *mxGetPr(prhs[9]);
    }
    p.window = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (present(nrhs, prhs, 10)) {
        if (!(isa(prhs[10], "windows.window") && mxGetNumberOfElements(prhs[10]) == 1))
            fail(GW, "InvalidWindow", "Receive window must be a single object of the class windows.window!");
        if (!window_of(prhs[10], &p.window)) fail(GW, "InvalidWindow", "Could not read window parameters!");
    }
    // default F-number of the gateway (gpu_bf_das_pw_rf.cu:360-365)
    mxArray *F = nullptr;
    bool own_F = false;
    if (present(nrhs, prhs, 11)) {
        if (!(isa(prhs[11], "f_numbers.f_number") && mxGetNumberOfElements(prhs[11]) == 1))
            fail(GW, "InvalidFNumber", "F-number must be a single object of the class f_numbers.f_number!");
        F = const_cast<mxArray *>(prhs[11]);
    } else {
        mxArray *rhs[3] = {mxCreateDoubleScalar(70.0), mxCreateDoubleScalar(1.5), mxCreateDoubleScalar(10.0)};
        mexCallMATLAB(1, &F, 3, rhs, "f_numbers.grating.angle_lb");
        for (mxArray *r : rhs) mxDestroyArray(r);
        own_F = true;
    }
    p.normalization = FN_NORMALIZATION_ON;
    p.normalization_window = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (present(nrhs, prhs, 12)) {
        if (!(isa(prhs[12], "normalizations.normalization") && mxGetNumberOfElements(prhs[12]) == 1))
            fail(GW, "InvalidNormalization", "Normalization must be a single object of the class normalizations.normalization!");
        int code;
        if (!normalization_of(prhs[12], &code, &p.normalization_window))
            fail(GW, "InvalidFrequencyWindow", "Could not read window parameters!");
        p.normalization = code;
    }
    p.index_gpu = 0;
    if (present(nrhs, prhs, 13)) {
        if (!is_real_double_scalar(prhs[13])) fail(GW, "InvalidCUDADevice", "Device index must be a real-valued scalar of type double!");
        p.index_gpu = (int32_t)*mxGetPr(prhs[13]);
    }
    if (p.index_gpu < 0 || p.index_gpu >= n_dev) fail(GW, "InvalidCUDADevice", "Selected device does not exist!");
    p.verbosity = 1;
    if (present(nrhs, prhs, 14)) {
        if (!is_real_double_scalar(prhs[14])) fail(GW, "InvalidVerbosityLevel", "Verbosity level must be a real-valued scalar of type double!");
        p.verbosity = (int32_t)*mxGetPr(prhs[14]);
    }
    p.semantics = semantics_arg(nrhs, prhs, 15, GW);
    p.precision = FN_PRECISION_FP32;
    p.output_scale = FN_SCALE_MEX;

    // band, then the F-number values through MATLAB (gpu_bf_das_pw_rf.cu:525-587)
    int32_t N = 0, lb = 0, ub = 0;
    check(fn_pw_frequency_grid(&p, &N, &lb, &ub), GW);
    std::vector<double> r(ub - lb + 1);
    check(fn_element_pitch_over_lambda(p.element_pitch, p.f_s, p.c_0, N, lb, ub, r.data()), GW);
    std::vector<double> Fv = compute_values(F, r);
    if (p.verbosity > 0) {
        // status section (gpu_bf_das_pw_rf.cu:609-632), before the computation
        status_common("gpu_bf_das_pw_rf: DAS, steered plane waves", p.n_x, p.n_z, n_t, n_el, p.f_s, p.c_0, p.f_lb,
                      p.f_ub, lb, ub, p.index_t0);
        double a0 = p.steering_angles[0], a1 = a0;
        for (int a = 1; a < n_a; ++a) {
            a0 = std::fmin(a0, p.steering_angles[a]);
            a1 = std::fmax(a1, p.steering_angles[a]);
        }
        const double deg = 180.0 / 3.14159265358979323846;
        mexPrintf("  %-20s: %d, %.2f..%.2f deg %-4s %-22s: %d\n", "angles of incidence", n_a, a0 * deg, a1 * deg, "",
                  "DFT order", N);
        // the transmit reference is the array edge the wave reaches first (gpu_bf_das_pw_rf.cu:519)
        const double x_ref = 0.5 * (n_el - 1) * p.element_pitch * 1e3;
        if (n_a == 1)
            mexPrintf("  %-20s: %7.2f mm\n", "reference position", -std::sin(a0) >= 0 ? -x_ref : x_ref);
        else
            mexPrintf("  %-20s: +-%.2f mm (by the sign of each angle)\n", "reference position", x_ref);
        mexPrintf("  %-20s: %s\n", "window", present(nrhs, prhs, 10) ? name_of(prhs[10]).c_str() : "boxcar");
        mexPrintf("  %-20s: %s\n", "F-number", name_of(F).c_str());
        mexPrintf("  %-20s: %s\n", "normalization", present(nrhs, prhs, 12) ? name_of(prhs[12]).c_str() : "on");
        mexPrintf("  %-20s: %s\n", "semantics", p.semantics == FN_SEMANTICS_MEX ? "mex" : "cpu");
    }
    if (own_F) mxDestroyArray(F);
    p.f_number_values = Fv.data();
    p.n_f = (int32_t)Fv.size();

    std::vector<float> img(2 * (size_t)p.n_x * p.n_z), w((size_t)p.n_x * p.n_z);
    check(fn_das_pw(&p, mxGetPr(prhs[2]), FN_DTYPE_F64, img.data(), w.data()), GW);
    // device, launch, memory and timing sections (gpu_bf_das_pw_rf.cu:664-689, 875-903, 1128-1144)
    if (p.verbosity > 0) mexPrintf("%s", fn_last_report());
    plhs[0] = image_out(img, p.n_z, p.n_x);
    if (nlhs > 1) plhs[1] = weights_out(w, p.n_z, p.n_x);
}
