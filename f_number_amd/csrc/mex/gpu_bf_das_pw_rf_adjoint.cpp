// gpu_bf_das_pw_rf_adjoint.cpp -- MEX gateway of the plane-wave adjoint
// (backprojection) over libfnumber_hip.  No counterpart in the reference, whose
// +cuda package only beamforms (gpu_bf_das_pw_rf.cu:110); the argument list
// follows that gateway so MATLAB code can call both alike:
//   data_RF_adj = gpu_bf_das_pw_rf_adjoint(positions_x, positions_z, image_y,
//       rf_size, f_s, steering_angle, element_width, element_pitch, c_0, f_bounds,
//       index_t0, window, F_number, normalization, index_gpu, verbosity,
//       c_tof, delay_offsets, semantics)
// image_y: complex single [N_z x N_x] (the forward gateway's output type);
// rf_size: [N_t, N_el]; steering_angle: N_a angles.  Returns double
// [N_t x N_el x N_a] with Re sum conj(A x) y = sum x (A^H y) for the forward
// map A of gpu_bf_das_pw_rf (output scale 2/N_dft), time of flight with c_tof
// ([] = c_0) and per-element receive delay offsets ([] = 0).  Arguments 0-8
// are required; [] selects the defaults of gpu_bf_das_pw_rf.
// Build (MATLAB): mex -R2017b gpu_bf_das_pw_rf_adjoint.cpp -I<repo>/include
//                 -L<repo>/f_number_amd -lfnumber_hip
#include "fn_mex_common.h"

#include <vector>

static const char *const GW = "gpu_bf_das_pw_rf_adjoint";

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    using namespace fnmex;
    int32_t n_dev = 0;
    fn_device_count(&n_dev);
    if (n_dev < 1) fail(GW, "NoCUDADevices", "Could not find any HIP devices!");
    if (nrhs < 9) fail(GW, "nrhs", "At least 9 inputs are required.");
    if (nrhs > 19) fail(GW, "nrhs", "At most 19 inputs are allowed.");
    if (nlhs > 1) fail(GW, "nlhs", "At most 1 output is allowed.");
    if (!(is_real_double(prhs[0]) && mxGetNumberOfDimensions(prhs[0]) == 2 && mxGetM(prhs[0]) == 1))
        fail(GW, "InvalidLateralPositions", "Lateral positions must be a real-valued row vector of type double!");
    if (!(is_real_double(prhs[1]) && mxGetNumberOfDimensions(prhs[1]) == 2 && mxGetM(prhs[1]) == 1))
        fail(GW, "InvalidAxialPositions", "Axial positions must be a real-valued row vector of type double!");
    const int n_x = (int)mxGetN(prhs[0]), n_z = (int)mxGetN(prhs[1]);
    if (!(mxIsClass(prhs[2], "single") && mxIsComplex(prhs[2]) && (int)mxGetM(prhs[2]) == n_z &&
          (int)mxGetN(prhs[2]) == n_x))
        fail(GW, "InvalidImage", "Image must be a complex single matrix of size N_z x N_x!");
    if (!(is_real_double(prhs[3]) && mxGetNumberOfElements(prhs[3]) == 2))
        fail(GW, "InvalidRFData", "RF size must be the real-valued vector [N_t, N_el]!");
    const int n_t = (int)mxGetPr(prhs[3])[0], n_el = (int)mxGetPr(prhs[3])[1];
    if (!(n_t >= 2 && n_el >= 2)) fail(GW, "InvalidRFData", "RF size must be at least [2, 2]!");
    if (!is_real_double_scalar(prhs[4])) fail(GW, "InvalidSamplingRate", "Sampling rate must be a real-valued scalar of type double!");
    if (!(is_real_double(prhs[5]) && mxGetNumberOfElements(prhs[5]) >= 1))
        fail(GW, "InvalidSteeringAngle", "Steering angles must be real-valued of type double!");
    if (!is_real_double_scalar(prhs[6])) fail(GW, "InvalidElementWidth", "Element width must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[7])) fail(GW, "InvalidElementPitch", "Element pitch must be a real-valued scalar of type double!");
    if (!is_real_double_scalar(prhs[8])) fail(GW, "InvalidSoundSpeed", "Speed of sound must be a real-valued scalar of type double!");
    const int n_a = (int)mxGetNumberOfElements(prhs[5]);

    fn_pw_params p{};
    p.struct_size = sizeof(fn_pw_params);
    p.positions_x = mxGetPr(prhs[0]);
    p.n_x = n_x;
    p.positions_z = mxGetPr(prhs[1]);
    p.n_z = n_z;
    p.n_samples = n_t;
    p.n_elements = n_el;
    p.n_angles = n_a;
    p.steering_angles = mxGetPr(prhs[5]);
    p.f_s = *mxGetPr(prhs[4]);
    p.element_width = *mxGetPr(prhs[6]);
    p.element_pitch = *mxGetPr(prhs[7]);
    p.c_0 = *mxGetPr(prhs[8]);
    p.f_lb = 2.2250738585072014e-308;
    p.f_ub = p.f_s / 2;
    if (present(nrhs, prhs, 9)) {
        if (!(is_real_double(prhs[9]) && mxGetNumberOfElements(prhs[9]) == 2))
            fail(GW, "InvalidFrequencyBounds", "Frequency bounds must be a real-valued vector with two entries of type double!");
        p.f_lb = mxGetPr(prhs[9])[0];
        p.f_ub = mxGetPr(prhs[9])[1];
    }
    if (present(nrhs, prhs, 10)) {
        if (!is_real_double_scalar(prhs[10])) fail(GW, "InvalidTimeShift", "Time index must be a real-valued scalar of type double!");
        p.index_t0 = *mxGetPr(prhs[10]);
    }
    p.window = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (present(nrhs, prhs, 11)) {
        if (!(isa(prhs[11], "windows.window") && mxGetNumberOfElements(prhs[11]) == 1))
            fail(GW, "InvalidWindow", "Receive window must be a single object of the class windows.window!");
        if (!window_of(prhs[11], &p.window)) fail(GW, "InvalidWindow", "Could not read window parameters!");
    }
    mxArray *F = nullptr;
    bool own_F = false;
    if (present(nrhs, prhs, 12)) {
        if (!(isa(prhs[12], "f_numbers.f_number") && mxGetNumberOfElements(prhs[12]) == 1))
            fail(GW, "InvalidFNumber", "F-number must be a single object of the class f_numbers.f_number!");
        F = const_cast<mxArray *>(prhs[12]);
    } else {
        mxArray *rhs[3] = {mxCreateDoubleScalar(70.0), mxCreateDoubleScalar(1.5), mxCreateDoubleScalar(10.0)};
        mexCallMATLAB(1, &F, 3, rhs, "f_numbers.grating.angle_lb");
        for (mxArray *r : rhs) mxDestroyArray(r);
        own_F = true;
    }
    p.normalization = FN_NORMALIZATION_ON;
    p.normalization_window = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (present(nrhs, prhs, 13)) {
        if (!(isa(prhs[13], "normalizations.normalization") && mxGetNumberOfElements(prhs[13]) == 1))
            fail(GW, "InvalidNormalization", "Normalization must be a single object of the class normalizations.normalization!");
        int code;
        if (!normalization_of(prhs[13], &code, &p.normalization_window))
            fail(GW, "InvalidFrequencyWindow", "Could not read window parameters!");
        p.normalization = code;
    }
    p.index_gpu = 0;
    if (present(nrhs, prhs, 14)) {
        if (!is_real_double_scalar(prhs[14])) fail(GW, "InvalidCUDADevice", "Device index must be a real-valued scalar of type double!");
        p.index_gpu = (int32_t)*mxGetPr(prhs[14]);
    }
    if (p.index_gpu < 0 || p.index_gpu >= n_dev) fail(GW, "InvalidCUDADevice", "Selected device does not exist!");
    p.verbosity = 1;
    if (present(nrhs, prhs, 15)) {
        if (!is_real_double_scalar(prhs[15])) fail(GW, "InvalidVerbosityLevel", "Verbosity level must be a real-valued scalar of type double!");
        p.verbosity = (int32_t)*mxGetPr(prhs[15]);
    }
    double c_tof = 0.0;
    if (present(nrhs, prhs, 16)) {
        if (!is_real_double_scalar(prhs[16])) fail(GW, "InvalidSoundSpeed", "c_tof must be a real-valued scalar of type double!");
        c_tof = *mxGetPr(prhs[16]);
    }
    const double *delays = nullptr;
    if (present(nrhs, prhs, 17)) {
        if (!(is_real_double(prhs[17]) && (int)mxGetNumberOfElements(prhs[17]) == n_el))
            fail(GW, "InvalidDelayOffsets", "Delay offsets must be real-valued of type double, one per element!");
        delays = mxGetPr(prhs[17]);
    }
    p.semantics = semantics_arg(nrhs, prhs, 18, GW);
    p.precision = FN_PRECISION_FP32;
    p.output_scale = FN_SCALE_MEX;

    int32_t N = 0, lb = 0, ub = 0;
    check(fn_pw_frequency_grid(&p, &N, &lb, &ub), GW);
    std::vector<double> r(ub - lb + 1);
    check(fn_element_pitch_over_lambda(p.element_pitch, p.f_s, p.c_0, N, lb, ub, r.data()), GW);
    std::vector<double> Fv = compute_values(F, r);
    if (own_F) mxDestroyArray(F);
    p.f_number_values = Fv.data();
    p.n_f = (int32_t)Fv.size();

    // complex single [N_z x N_x] (separate parts) -> interleaved [n_x][n_z]
    const size_t nv = (size_t)n_x * n_z;
    std::vector<float> y(2 * nv);
    const float *re = static_cast<const float *>(mxGetData(prhs[2]));
    const float *im = static_cast<const float *>(mxGetImagData(prhs[2]));
    for (size_t i = 0; i < nv; ++i) {
        y[2 * i] = re[i];
        y[2 * i + 1] = im[i];
    }
    std::vector<float> adj((size_t)n_t * n_el * n_a);
    check(fn_das_pw_adjoint(&p, c_tof, delays, y.data(), adj.data()), GW);
    const mwSize dims[3] = {(mwSize)n_t, (mwSize)n_el, (mwSize)n_a};
    mxArray *out = mxCreateNumericArray(n_a > 1 ? 3 : 2, dims, mxDOUBLE_CLASS, mxREAL);
    double *o = mxGetPr(out);
    for (size_t i = 0; i < adj.size(); ++i) o[i] = adj[i];
    plhs[0] = out;
}
