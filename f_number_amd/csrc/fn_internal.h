// fn_internal.h -- declarations shared by the host and device translation units
// of libfnumber_hip.so.  Not part of the public ABI (include/fnumber_hip.h).
#pragma once

#include "../../include/fnumber_hip.h"

#include <cmath>
#include <cstdint>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace fnh {

struct Grid {
    int N = 0;      // odd DFT order
    int n_lb = 0;   // first band bin (0-based)
    int n_ub = -1;  // last band bin
    int n_f = 0;
};

int fail(int status, const char *fmt, ...);
void clear_error();
int frequency_grid(int n_t, int n_el, double pitch, double z_pad, double f_s, double c_0, double f_lb,
                   double f_ub, Grid *g);
int f_number_values(const fn_f_number *F, const double *r, int n, double *out);
double window_sample(int id, double param, double t, int semantics);
int check_window(const fn_window &w, int status);
double z_pad_of(const double *pz, int nz, int semantics);
int validate_pw(const fn_pw_params *p, Grid *g);
int validate_sa(const fn_sa_params *p, Grid *g);

}  // namespace fnh
