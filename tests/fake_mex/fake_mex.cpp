// fake_mex.cpp -- in-process stand-in for the MATLAB side of a MEX call
// (TEST INFRASTRUCTURE).  Arrays are plain C++ objects; MATLAB objects are an
// class name plus properties; mexCallMATLAB serves the functions the gateways
// call back: isa, compute_values (through libfnumber_hip's F-number port, as
// the classes in +f_numbers would), the default F-number constructor, and
// string / char for the names in the status section.  Errors
// raised with mexErrMsgIdAndTxt unwind to fm_call, which reports the id.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fnumber_hip.h"

struct mxArray_tag {
    std::string cls;  // "double", "single", "char", "logical" or a class name
    std::vector<std::string> supers;
    std::vector<mwSize> dims{0, 0};
    std::vector<double> pr;
    std::vector<float> fr, fi;
    bool complex_ = false;
    std::string str;
    std::map<std::string, std::shared_ptr<mxArray_tag>> props;
};

struct MexError {
    std::string id, msg;
};

static size_t numel(const mxArray *a)
{
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}

extern "C" {
bool mxIsDouble(const mxArray *a) { return a->cls == "double"; }
bool mxIsComplex(const mxArray *a) { return a->complex_; }
bool mxIsEmpty(const mxArray *a) { return numel(a) == 0; }
bool mxIsClass(const mxArray *a, const char *name) { return a->cls == name; }
bool mxIsLogicalScalarTrue(const mxArray *a) { return a->cls == "logical" && !a->pr.empty() && a->pr[0] != 0; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
size_t mxGetM(const mxArray *a) { return a->dims[0]; }
size_t mxGetN(const mxArray *a)
{
    size_t n = 1;
    for (size_t i = 1; i < a->dims.size(); ++i) n *= a->dims[i];
    return n;
}
size_t mxGetNumberOfElements(const mxArray *a) { return numel(a); }
double *mxGetPr(const mxArray *a) { return const_cast<double *>(a->pr.data()); }
void *mxGetData(const mxArray *a)
{
    if (a->cls == "single") return const_cast<float *>(a->fr.data());
    return const_cast<double *>(a->pr.data());
}
void *mxGetImagData(const mxArray *a) { return const_cast<float *>(a->fi.data()); }
mxArray *mxGetProperty(const mxArray *a, mwIndex, const char *name)
{
    auto it = a->props.find(name);
    if (it == a->props.end()) return nullptr;
    return new mxArray_tag(*it->second);
}
int mxGetString(const mxArray *a, char *buf, mwSize len)
{
    if (a->cls != "char" || a->str.size() + 1 > len) return 1;
    std::memcpy(buf, a->str.c_str(), a->str.size() + 1);
    return 0;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    auto *a = new mxArray_tag;
    a->cls = "double";
    a->dims = {m, n};
    a->pr.assign(m * n, 0.0);
    a->complex_ = c == mxCOMPLEX;
    return a;
}
mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    a->pr[0] = v;
    return a;
}
mxArray *mxCreateString(const char *s)
{
    auto *a = new mxArray_tag;
    a->cls = "char";
    a->str = s;
    a->dims = {1, std::strlen(s)};
    return a;
}
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c)
{
    auto *a = new mxArray_tag;
    a->dims = {m, n};
    if (cls == mxSINGLE_CLASS) {
        a->cls = "single";
        a->fr.assign(m * n, 0.f);
        if (c == mxCOMPLEX) a->fi.assign(m * n, 0.f);
    } else {
        a->cls = "double";
        a->pr.assign(m * n, 0.0);
    }
    a->complex_ = c == mxCOMPLEX;
    return a;
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c)
{
    mxArray *a = mxCreateNumericMatrix(0, 0, cls, c);
    a->dims.assign(dims, dims + ndim);
    const size_t n = numel(a);
    if (cls == mxSINGLE_CLASS) {
        a->fr.assign(n, 0.f);
        if (c == mxCOMPLEX) a->fi.assign(n, 0.f);
    } else {
        a->pr.assign(n, 0.0);
    }
    return a;
}
mxArray *mxDuplicateArray(const mxArray *a) { return new mxArray_tag(*a); }
void mxDestroyArray(mxArray *a) { delete a; }

static bool isa_impl(const mxArray *o, const std::string &cls)
{
    if (o->cls == cls) return true;
    for (auto &s : o->supers)
        if (s == cls) return true;
    return false;
}

static double prop(const mxArray *o, const char *name, double dflt)
{
    auto it = o->props.find(name);
    if (it == o->props.end() || it->second->pr.empty()) return dflt;
    return it->second->pr[0];
}

int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *name)
{
    const std::string fn = name;
    if (fn == "isa" && nrhs == 2 && nlhs == 1) {
        auto *r = mxCreateDoubleScalar(isa_impl(prhs[0], prhs[1]->str) ? 1.0 : 0.0);
        r->cls = "logical";
        plhs[0] = r;
        return 0;
    }
    if (fn == "f_numbers.grating.angle_lb" && nlhs == 1) {
        auto *o = new mxArray_tag;
        o->cls = fn;
        o->supers = {"f_numbers.grating.grating", "f_numbers.f_number"};
        o->dims = {1, 1};
        const char *names[3] = {"angle_lb_deg", "F_number_ub", "distance_deg"};
        const double dflt[3] = {45.0, 3.0, 10.0};
        for (int i = 0; i < 3; ++i)
            o->props[names[i]] = std::shared_ptr<mxArray_tag>(mxCreateDoubleScalar(i < nrhs ? prhs[i]->pr[0] : dflt[i]));
        plhs[0] = o;
        return 0;
    }
    if (fn == "compute_values" && nrhs == 2 && nlhs == 1) {
        const mxArray *o = prhs[0];
        fn_f_number F{};
        if (o->cls == "f_numbers.grating.angle_lb") {
            F = fn_f_number{FN_F_NUMBER_GRATING_ANGLE_LB, prop(o, "angle_lb_deg", 45), prop(o, "F_number_ub", 3),
                            prop(o, "distance_deg", 10)};
        } else if (o->cls == "f_numbers.constant") {
            F = fn_f_number{FN_F_NUMBER_CONSTANT, prop(o, "value", 1), 0, 0};
        } else if (o->cls == "f_numbers.distance.constant") {
            F = fn_f_number{FN_F_NUMBER_DISTANCE_CONSTANT, prop(o, "distance_deg", 5), prop(o, "F_number_ub", 3), 0};
        } else {
            mexErrMsgIdAndTxt("fake_mex:UnknownClass", "compute_values: unknown class %s", o->cls.c_str());
        }
        mxArray *r = mxCreateDoubleMatrix(numel(prhs[1]), 1, mxREAL);
        fn_f_number_values(&F, prhs[1]->pr.data(), (int32_t)numel(prhs[1]), r->pr.data());
        plhs[0] = r;
        return 0;
    }
    if (fn == "string" && nrhs == 1 && nlhs == 1) {
        // the class name and its numeric properties, standing in for the classes' string methods
        const mxArray *o = prhs[0];
        std::string t = o->cls;
        std::string sep = " (";
        for (const auto &kv : o->props) {
            if (kv.second->cls != "double" || kv.second->pr.size() != 1) continue;
            char b[64];
            std::snprintf(b, sizeof(b), "%s%s = %.2f", sep.c_str(), kv.first.c_str(), kv.second->pr[0]);
            t += b;
            sep = ", ";
        }
        if (sep == ", ") t += ")";
        auto *r = mxCreateString(t.c_str());
        r->cls = "string";
        plhs[0] = r;
        return 0;
    }
    if (fn == "char" && nrhs == 1 && nlhs == 1 && prhs[0]->cls == "string") {
        plhs[0] = mxCreateString(prhs[0]->str.c_str());
        return 0;
    }
    mexErrMsgIdAndTxt("fake_mex:UnknownFunction", "mexCallMATLAB: %s", name);
    return 1;
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw MexError{id, buf};
}

int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int n = std::vprintf(fmt, ap);
    va_end(ap);
    std::fflush(stdout);
    return n;
}

// ---- harness for the tests (ctypes) ---------------------------------------
mxArray *fm_double(const double *data, size_t ndim, const size_t *dims)
{
    auto *a = new mxArray_tag;
    a->cls = "double";
    a->dims.assign(dims, dims + ndim);
    a->pr.assign(data, data + numel(a));
    return a;
}
mxArray *fm_empty() { return mxCreateDoubleMatrix(0, 0, mxREAL); }
mxArray *fm_string(const char *s) { return mxCreateString(s); }
mxArray *fm_object(const char *cls, const char *supers)
{
    auto *a = new mxArray_tag;
    a->cls = cls;
    a->dims = {1, 1};
    std::string s = supers ? supers : "";
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find(',', p);
        if (q == std::string::npos) q = s.size();
        if (q > p) a->supers.push_back(s.substr(p, q - p));
        p = q + 1;
    }
    return a;
}
void fm_set_prop(mxArray *o, const char *name, mxArray *v) { o->props[name] = std::shared_ptr<mxArray_tag>(v); }
int fm_call(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs, char *err_id, size_t n_err)
{
    try {
        mexFunction(nlhs, plhs, nrhs, const_cast<const mxArray **>(prhs));
    } catch (const MexError &e) {
        std::snprintf(err_id, n_err, "%s", e.id.c_str());
        return 1;
    }
    return 0;
}
size_t fm_numel(const mxArray *a) { return numel(a); }
size_t fm_dim(const mxArray *a, size_t i) { return i < a->dims.size() ? a->dims[i] : 1; }
int fm_is_single_complex(const mxArray *a) { return a->cls == "single" && a->complex_; }
void fm_get_single(const mxArray *a, float *re, float *im)
{
    std::memcpy(re, a->fr.data(), a->fr.size() * sizeof(float));
    if (im && !a->fi.empty()) std::memcpy(im, a->fi.data(), a->fi.size() * sizeof(float));
}
int fm_is_double(const mxArray *a) { return a->cls == "double"; }
void fm_get_double(const mxArray *a, double *out) { std::memcpy(out, a->pr.data(), a->pr.size() * sizeof(double)); }
// complex single matrix [m x n] from separate parts (column-major)
mxArray *fm_single_complex(const float *re, const float *im, size_t m, size_t n)
{
    mxArray *a = mxCreateNumericMatrix(m, n, mxSINGLE_CLASS, mxCOMPLEX);
    std::memcpy(a->fr.data(), re, m * n * sizeof(float));
    std::memcpy(a->fi.data(), im, m * n * sizeof(float));
    return a;
}
void fm_free(mxArray *a) { delete a; }
}
