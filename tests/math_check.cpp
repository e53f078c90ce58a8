// math_check.cpp -- TEST INFRASTRUCTURE ONLY: the host build of the event
// loop's f64 log and sin/cos (3photons-rust_amd/csrc/tp_math.h, the same
// source the kernels compile) beside the host libm the oracle uses and an
// x87 long-double reference, for tests/test_math.py.
//
//   g++ -O2 -std=c++20 -ffp-contract=off -shared -fPIC -I<csrc> math_check.cpp
#include "tp_math.h"

#include <cmath>
#include <cstdint>

namespace {
// error of y against the long-double value ref, in ulps of the double nearest ref
double ulp_err(double y, long double ref)
{
    const double r = (double)ref;
    if (r == 0.0)
        return y == 0.0 ? 0.0 : INFINITY;
    int e;
    std::frexp(r, &e);
    const long double ulp = std::ldexp(1.0L, e - 53);
    return (double)(((long double)y - ref) / ulp);
}
} // namespace

extern "C" {

void tpm_eval_log(const double *x, int64_t n, double *ours, double *libm, double *err)
{
    for (int64_t i = 0; i < n; ++i) {
        ours[i] = tpm::log_pos(x[i]);
        libm[i] = std::log(x[i]);
        err[i] = ulp_err(ours[i], logl((long double)x[i]));
    }
}

void tpm_eval_sincos(const double *x, int64_t n, double *s, double *c, double *s_libm, double *c_libm,
                     double *s_err, double *c_err)
{
    for (int64_t i = 0; i < n; ++i) {
        tpm::sincos_f64(x[i], s[i], c[i]);
        s_libm[i] = std::sin(x[i]);
        c_libm[i] = std::cos(x[i]);
        s_err[i] = ulp_err(s[i], sinl((long double)x[i]));
        c_err[i] = ulp_err(c[i], cosl((long double)x[i]));
    }
}
}
