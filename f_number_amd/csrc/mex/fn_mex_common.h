// fn_mex_common.h -- helpers of the MEX gateways that put libfnumber_hip.so
// under the reference's das_pw.m / das_sa.m (cuda.gpu_bf_das_pw_rf,
// cuda.gpu_bf_saft_rf).  MATLAB-side objects (windows, F-numbers,
// normalizations) are evaluated through MATLAB exactly where the reference's
// gateways evaluate them (gpu_bf_helper_functions.cu:10-138); the core
// receives plain values only.
#pragma once

#include "mex.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/fnumber_hip.h"

namespace fnmex {

// MATLAB's isa(), which also recognises subclasses (gpu_bf_helper_functions.cu:10-30)
inline bool isa(const mxArray *obj, const char *cls)
{
    mxArray *lhs[1] = {nullptr};
    mxArray *rhs[2] = {mxDuplicateArray(obj), mxCreateString(cls)};
    mexCallMATLAB(1, lhs, 2, rhs, "isa");
    const bool r = mxIsLogicalScalarTrue(lhs[0]);
    mxDestroyArray(lhs[0]);
    mxDestroyArray(rhs[0]);
    mxDestroyArray(rhs[1]);
    return r;
}

// compute_values(F, element_pitch_over_lambda) evaluated in MATLAB
// (gpu_bf_helper_functions.cu:35-51); copies the result and frees the temporaries.
inline std::vector<double> compute_values(const mxArray *F, const std::vector<double> &r)
{
    mxArray *lhs[1] = {nullptr};
    mxArray *rarr = mxCreateDoubleMatrix(r.size(), 1, mxREAL);
    std::memcpy(mxGetPr(rarr), r.data(), r.size() * sizeof(double));
    mxArray *rhs[2] = {mxDuplicateArray(F), rarr};
    mexCallMATLAB(1, lhs, 2, rhs, "compute_values");
    std::vector<double> out(r.size());
    if (mxIsDouble(lhs[0]) && mxGetNumberOfElements(lhs[0]) == r.size())
        std::memcpy(out.data(), mxGetPr(lhs[0]), r.size() * sizeof(double));
    else
        out.assign(r.size(), -1.0);  // rejected by the core's F-number check
    mxDestroyArray(lhs[0]);
    mxDestroyArray(rhs[0]);
    mxDestroyArray(rhs[1]);
    return out;
}

// window object -> fn_window (gpu_bf_helper_functions.cu:56-97); false if unknown
inline bool window_of(const mxArray *w, fn_window *out)
{
    if (mxIsClass(w, "windows.boxcar")) { *out = fn_window{FN_WINDOW_BOXCAR, 0.0}; return true; }
    if (mxIsClass(w, "windows.hann")) { *out = fn_window{FN_WINDOW_HANN, 0.0}; return true; }
    if (mxIsClass(w, "windows.triangular")) { *out = fn_window{FN_WINDOW_TRIANGULAR, 0.0}; return true; }
    if (mxIsClass(w, "windows.tukey")) {
        const mxArray *fc = mxGetProperty(w, 0, "fraction_cosine");
        if (!fc || !mxIsDouble(fc) || mxGetNumberOfElements(fc) != 1) return false;
        *out = fn_window{FN_WINDOW_TUKEY, *mxGetPr(fc)};
        return true;
    }
    return false;
}

inline bool is_real_double(const mxArray *a) { return mxIsDouble(a) && !mxIsComplex(a); }
inline bool is_real_double_scalar(const mxArray *a) { return is_real_double(a) && mxGetNumberOfElements(a) == 1; }
inline bool present(int nrhs, const mxArray *prhs[], int i) { return nrhs > i && !mxIsEmpty(prhs[i]); }

// core status -> "<gateway>:<Id>" (mexErrMsgIdAndTxt does not return)
inline void check(int status, const char *gateway)
{
    if (status == FN_OK) return;
    char id[160];
    std::snprintf(id, sizeof(id), "%s:%s", gateway, fn_status_id(status));
    mexErrMsgIdAndTxt(id, "%s", fn_last_error());
}

inline void fail(const char *gateway, const char *ident, const char *msg)
{
    char id[160];
    std::snprintf(id, sizeof(id), "%s:%s", gateway, ident);
    mexErrMsgIdAndTxt(id, "%s", msg);
}

// normalization object -> (FN_NORMALIZATION_*, window of windowed)
inline bool normalization_of(const mxArray *n, int *code, fn_window *w)
{
    *w = fn_window{FN_WINDOW_BOXCAR, 0.0};
    if (mxIsClass(n, "normalizations.off")) { *code = FN_NORMALIZATION_OFF; return true; }
    if (mxIsClass(n, "normalizations.windowed")) {
        *code = FN_NORMALIZATION_WINDOWED;
        const mxArray *win = mxGetProperty(n, 0, "window");
        return win && window_of(win, w);
    }
    *code = FN_NORMALIZATION_ON;
    return true;
}

// optional semantics argument: 'cpu' (default) or 'mex'
inline int32_t semantics_arg(int nrhs, const mxArray *prhs[], int i, const char *gateway)
{
    if (!present(nrhs, prhs, i)) return FN_SEMANTICS_CPU;
    char buf[8] = {0};
    if (mxGetString(prhs[i], buf, sizeof(buf)) != 0) fail(gateway, "InvalidArgument", "semantics must be 'cpu' or 'mex'");
    if (std::strcmp(buf, "cpu") == 0) return FN_SEMANTICS_CPU;
    if (std::strcmp(buf, "mex") == 0) return FN_SEMANTICS_MEX;
    fail(gateway, "InvalidArgument", "semantics must be 'cpu' or 'mex'");
    return FN_SEMANTICS_CPU;
}

// the name MATLAB gives an object: string(object), then char() of that string,
// as the reference's gateways name windows, F-numbers and normalizations in
// their status section (gpu_bf_helper_functions.cu:113-138)
inline std::string name_of(const mxArray *obj)
{
    mxArray *rhs = mxDuplicateArray(obj), *str = nullptr, *chr = nullptr;
    mexCallMATLAB(1, &str, 1, &rhs, "string");
    mexCallMATLAB(1, &chr, 1, &str, "char");
    char buf[160] = {0};
    mxGetString(chr, buf, sizeof(buf));
    mxDestroyArray(rhs);
    mxDestroyArray(str);
    mxDestroyArray(chr);
    return buf;
}

inline void rule(char c)
{
    char line[82];
    std::memset(line, c, 80);
    line[80] = 0;
    mexPrintf(" %s\n", line);
}

// the status lines shared by both gateways (gpu_bf_das_pw_rf.cu:609-632): geometry,
// sampling, medium and band; each gateway adds its transmit lines and names
inline void status_common(const char *title, int n_x, int n_z, int n_t, int n_el, double f_s, double c_0,
                          double f_lb, double f_ub, int32_t lb, int32_t ub, double index_t0)
{
    rule('=');
    mexPrintf(" %s (libfnumber_hip, ABI %d), status information\n", title, fn_abi_version());
    rule('=');
    mexPrintf("  %-20s: %5d x %-5d %5s %-22s: %d\n", "lattice geometry", n_x, n_z, "", "number of samples", n_t);
    mexPrintf("  %-20s: %-11d %7s %-22s: %5.1f MHz\n", "number rx channels", n_el, "", "sampling rate", f_s / 1e6);
    mexPrintf("  %-20s: %7.2f m/s %7s %-22s: %5.1f MHz (%d)\n", "average sound speed", c_0, "",
              "lower frequency bound", f_lb / 1e6, lb);
    mexPrintf("  %-20s: %-6.2f %12s %-22s: %5.1f MHz (%d)\n", "additional shift", index_t0, "",
              "upper frequency bound", f_ub / 1e6, ub);
    mexPrintf("  %-20s: %-6d\n", "number of frequencies", ub - lb + 1);
}

// the image of the core (interleaved complex float [n_x][n_z]) as MATLAB's
// complex single [N_z x N_x] with separate parts (gpu_bf_das_pw_rf.cu:1077-1122)
inline mxArray *image_out(const std::vector<float> &img, int n_z, int n_x)
{
    mxArray *a = mxCreateNumericMatrix(n_z, n_x, mxSINGLE_CLASS, mxCOMPLEX);
    float *re = static_cast<float *>(mxGetData(a));
    float *im = static_cast<float *>(mxGetImagData(a));
    const size_t n = (size_t)n_z * n_x;
    for (size_t i = 0; i < n; ++i) {
        re[i] = img[2 * i];
        im[i] = img[2 * i + 1];
    }
    return a;
}

inline mxArray *weights_out(const std::vector<float> &w, int n_z, int n_x)
{
    mxArray *a = mxCreateNumericMatrix(n_z, n_x, mxSINGLE_CLASS, mxREAL);
    std::memcpy(mxGetData(a), w.data(), w.size() * sizeof(float));
    return a;
}

}  // namespace fnmex
