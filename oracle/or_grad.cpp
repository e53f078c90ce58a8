// or_grad.cpp -- f64 CPU gradients of the event integrand for checking the
// hand-written GPU adjoint.  TEST INFRASTRUCTURE ONLY.
//
// F_c(u) = w * |M|^2_c(p(u)) * cut(p(u)) with the oracle's per-event code
// (or_event.h, the restatement of evgen.rs / evcut.rs / spinor.rs /
// matelems.rs) instantiated with forward-mode dual numbers: a completely
// different differentiation method (tangents pushed through every operation
// in evaluation order) from the reverse sweep on the GPU.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {
const int N = 12;
struct Dual {
    double v;
    double d[N];
    Dual(double x = 0.) : v(x) { std::memset(d, 0, sizeof d); }
};
inline Dual operator+(const Dual &a, const Dual &b)
{
    Dual r(a.v + b.v);
    for (int i = 0; i < N; ++i)
        r.d[i] = a.d[i] + b.d[i];
    return r;
}
inline Dual operator-(const Dual &a, const Dual &b)
{
    Dual r(a.v - b.v);
    for (int i = 0; i < N; ++i)
        r.d[i] = a.d[i] - b.d[i];
    return r;
}
inline Dual operator*(const Dual &a, const Dual &b)
{
    Dual r(a.v * b.v);
    for (int i = 0; i < N; ++i)
        r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
inline Dual operator/(const Dual &a, const Dual &b)
{
    Dual r(a.v / b.v);
    for (int i = 0; i < N; ++i)
        r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
inline Dual operator-(const Dual &a)
{
    Dual r(-a.v);
    for (int i = 0; i < N; ++i)
        r.d[i] = -a.d[i];
    return r;
}
inline Dual &operator*=(Dual &a, const Dual &b) { return a = a * b; }
inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
inline Dual chain(const Dual &a, double v, double dv)
{
    Dual r(v);
    for (int i = 0; i < N; ++i)
        r.d[i] = dv * a.d[i];
    return r;
}
inline Dual dsqrt(const Dual &a)
{
    double s = std::sqrt(a.v);
    return chain(a, s, s > 0 ? 0.5 / s : 0.);
}
inline Dual dsin(const Dual &a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline Dual dcos(const Dual &a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual dlog(const Dual &a) { return chain(a, std::log(a.v), 1. / a.v); }
inline Dual dfabs(const Dual &a) { return chain(a, std::fabs(a.v), a.v < 0 ? -1. : 1.); }
} // namespace

#define OR_REAL Dual
#define OR_SQRT dsqrt
#define OR_SIN dsin
#define OR_COS dcos
#define OR_LOG dlog
#define OR_FABS dfabs
#include "or_event.h"

namespace {
void phys_from(const double *k, int sort, or_phys &ph)
{
    ph.e_total = k[0];
    ph.cut_beam = k[1];
    ph.cut_pp = k[2];
    ph.e_min = k[3];
    ph.cut_plane = k[4];
    ph.g_a = k[5];
    ph.g_bp = k[6];
    ph.g_bm = k[7];
    for (int c = 0; c < 5; ++c)
        ph.sc[c] = 0.;
    ph.sort = sort;
}
} // namespace

extern "C" {

// k = [e_total, cut_beam, cut_pp, e_min, cut_plane, g_a, g_bp, g_bm, w].
// For n points u[n][12]: F[n][5], dF_du[n][5][12], dF_dp[n][5][12]
// (sorted photons, photon-major X,Y,Z,E); zeros for cut points.
void or_grad(const double *k, int sort, uint64_t n, const double *u, double *F, double *dF_du, double *dF_dp)
{
    or_phys ph;
    phys_from(k, sort, ph);
    const double w = k[8];
    for (uint64_t e = 0; e < n; ++e) {
        Dual uu[12];
        for (int j = 0; j < 12; ++j) {
            uu[j] = Dual(u[12 * e + j]);
            uu[j].d[j] = 1.;
        }
        Dual p[5][4];
        or_generate(&ph, uu, p);
        bool keep = or_keep(&ph, p);
        Dual m2[5][8], m[5];
        or_m2(&ph, p, m2);
        or_m2_sums(m2, m);
        // momentum gradient: fresh tangents on the 12 photon components
        Dual pp[5][4];
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 4; ++j) {
                pp[i][j] = Dual(p[i][j].v);
                if (i >= 2)
                    pp[i][j].d[4 * (i - 2) + j] = 1.;
            }
        Dual mp2[5][8], mp[5];
        or_m2(&ph, pp, mp2);
        or_m2_sums(mp2, mp);
        for (int c = 0; c < 5; ++c) {
            F[5 * e + c] = keep ? w * m[c].v : 0.;
            for (int j = 0; j < 12; ++j) {
                dF_du[60 * e + 12 * c + j] = keep ? w * m[c].d[j] : 0.;
                dF_dp[60 * e + 12 * c + j] = keep ? w * mp[c].d[j] : 0.;
            }
        }
    }
}

// w * |M|^2_c at given momenta p[n][5][4] (no cut), for finite differences
// in the momentum components.
void or_m2_at(const double *k, uint64_t n, const double *p, double *F)
{
    or_phys ph;
    phys_from(k, 1, ph);
    for (uint64_t e = 0; e < n; ++e) {
        Dual pp[5][4];
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 4; ++j)
                pp[i][j] = Dual(p[20 * e + 4 * i + j]);
        Dual m2[5][8], m[5];
        or_m2(&ph, pp, m2);
        or_m2_sums(m2, m);
        for (int c = 0; c < 5; ++c)
            F[5 * e + c] = k[8] * m[c].v;
    }
}

}
