// tp_physics.h -- per-event path of the 3photons event loop on gfx950:
// RAMBO (evgen), cuts (evcut), spinor products (spinor), helicity amplitudes
// and matrix elements (matelems), per-event weight (resacc integrate).
//
// Everything lives in registers; the only per-event memory reads are the
// L1-resident log and sin/cos tables of tp_math.h, whose routines return the
// correctly rounded value but within ~0.002 ulp of a rounding boundary (the
// reference calls the host libm; the two agree on > 99.8 % of arguments).  The
// f64 path keeps the reference's operation order exactly (the translation
// unit is compiled with -ffp-contract=off); it differs from the reference only
// where the reference computes known zeros or exact negations:
//   * s(j,i) is taken as -s(i,j) (exactly equal but for the sign of a zero);
//   * the helicity sums skip the zero amplitudes (adding +-0 to a non-zero sum
//     is exact, so the sums are unchanged);
//   * gemv terms multiplied by the zero Y/Z beam components are dropped (they
//     add +-0);
//   * the beam/photon-plane cut is skipped when cut_plane <= 0 (it can never
//     reject then).
// Citations are to HadrienG2/3photons-rust src/ files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tp_math.h"

#include <type_traits>

namespace tpd {

// Per-run constants, computed on the host with the same IEEE operations.
struct Params {
    double e_total;
    double cut_beam, cut_pp, e_min, cut_plane;
    double g_a, g_bp, g_bm;
    double sc[5];
    double el_x, el_e;            // incoming electron X and E (evgen.rs:66-69)
    double xx0, f0re, f0im;       // electron spinor: sqrt(E+Z), (X, Y)/sqrt(E+Z)
    double xx1, f1re, f1im;       // positron spinor
    double ns01re, ns01im;        // -RAC8 * s(e-, e+)
    double nt01re, nt01im;        // -RAC8 * t(e-, e+)
    // ns01re and nt01re times g_a, g_bp and g_bm (the f64 and mixed
    // amplitudes' real scale factors; the imaginary parts are +-0)
    double ka_s, ka_t, kp_s, kp_t, km_s, km_t;
    int sort;
    int pad;
};

template <class T> struct C {
    T re, im;
};
template <class T> __device__ __forceinline__ C<T> mk(T re, T im) { return C<T>{re, im}; }
template <class T> __device__ __forceinline__ C<T> cmul(C<T> a, C<T> b)
{
    return C<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// IEEE division and square root: in f64 through tpm::div_n / tpm::sqrt_n
// (the same correctly rounded results in the event loop's range, fewer
// instructions), in f32 the plain operations.
template <class T> __device__ __forceinline__ T dv(T a, T b)
{
    if constexpr (std::is_same<T, double>::value)
        return tpm::div_n(a, b);
    else
        return a / b;
}
template <class T> __device__ __forceinline__ T sq(T x)
{
    if constexpr (std::is_same<T, double>::value)
        return tpm::sqrt_n(x);
    else
        return sqrt(x);
}
// num-complex naive division (spinor.rs divides with it).
template <class T> __device__ __forceinline__ C<T> cdiv(C<T> a, C<T> b)
{
    T n = b.re * b.re + b.im * b.im;
    T re = a.re * b.re + a.im * b.im;
    T im = a.im * b.re - a.re * b.im;
    return C<T>{dv(re, n), dv(im, n)};
}
template <class T> __device__ __forceinline__ C<T> cadd(C<T> a, C<T> b) { return C<T>{a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ C<T> csq(C<T> z) { return cmul(z, z); }
template <class T> __device__ __forceinline__ C<T> neg(C<T> z) { return C<T>{-z.re, -z.im}; }
// t = -conj(s) (spinor.rs:63-66)
template <class T> __device__ __forceinline__ C<T> tof(C<T> s) { return C<T>{-s.re, s.im}; }

// ---- RAMBO: evgen.rs:174-207 (generate_raw) and 89-118 (generate) ---------

// Massless momentum of one photon from (cos theta, phi, e1, e2) uniforms.
// SCALED: the uniforms come as u 2^53, the integers xoshiro256+'s x >> 11
// holds (xo_uniform53), and the 2^-53 factors fold into the operations
// exactly: 2u - 1 = fma(u', 2^-52, -1) (2u is exact, so the reference's two
// roundings are one), 2 pi u = (2 pi 2^-53) u' and u1 u2 = (u1' u2') 2^-106,
// each rounded once as before.
template <bool SCALED>
__device__ __forceinline__ void raw_photon_f64(double uc, double uphi, double ue1, double ue2, double q[4])
{
    constexpr double S = SCALED ? 0x1.0p-53 : 1.;
    const double two_pi = 2. * 3.14159265358979323846;
    double cos_t = __builtin_fma(uc, 2. * S, -1.);
    double phi = (two_pi * S) * uphi;
    double em = ue1 * ue2;
    if constexpr (SCALED)
        em = em * (S * S);
    double sphi, cphi;
    tpm::sincos_f64(phi, sphi, cphi);
    double st = sq(1. - cos_t * cos_t);
    double en = -tpm::log_pos(em + 2.2250738585072014e-308);
    q[0] = en * (st * sphi);
    q[1] = en * (st * cphi);
    q[2] = en * cos_t;
    q[3] = en;
}

// Conformal map of three massless q onto the sqrt(s) rest frame, then the
// energy sort.  p[i] = photon i (X, Y, Z, E).
__device__ __forceinline__ void conformal_f64(const Params &P, const double q[3][4], double p[3][4])
{
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r[k] = (q[0][k] + q[1][k]) + q[2][k];
    double rn2 = r[3] * r[3] - ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    double alpha = dv(P.e_total, rn2);
    double rn = sq(rn2);
    double beta = dv(1., rn + r[3]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double rq = q[i][0] * r[0];
        rq = q[i][1] * r[1] + rq;
        rq = q[i][2] * r[2] + rq;
        p[i][3] = alpha * (r[3] * q[i][3] - rq);
        double b = beta * rq - q[i][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            p[i][k] = alpha * (rn * q[i][k] + b * r[k]);
    }
}

template <class T> __device__ __forceinline__ void swap4(T a[4], T b[4])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T t = a[k];
        a[k] = b[k];
        b[k] = t;
    }
}

// Sort by decreasing energy with strict '>' swaps over (0,1),(0,2),(1,2)
// (evgen.rs:109-118).
template <class T> __device__ __forceinline__ void sort3(T p[3][4])
{
    if (p[1][3] > p[0][3])
        swap4(p[0], p[1]);
    if (p[2][3] > p[0][3])
        swap4(p[0], p[2]);
    if (p[2][3] > p[1][3])
        swap4(p[1], p[2]);
}

// Slot of each photon after the reference's energy sort (evgen.rs:109-118:
// strict '>' swaps over the slot pairs (0,1), (0,2), (1,2)), from pairwise
// comparisons of the unsorted energies.  It reproduces the swap sequence
// exactly, ties and NaN included: step 2 compares photon 2 with the winner a
// of step 1, step 3 compares the loser b with whatever step 2 left in slot 2.
// The kernels keep the photons where RAMBO made them and apply the slots where
// they store the event (no register shuffles).
__host__ __device__ __forceinline__ void sort_slots(const double p[3][4], int sort, int d[3])
{
    if (!sort) {
        d[0] = 0;
        d[1] = 1;
        d[2] = 2;
        return;
    }
    const double e0 = p[0][3], e1 = p[1][3], e2 = p[2][3];
    const bool c01 = e1 > e0, c02 = e2 > e0, c12 = e2 > e1, g01 = e0 > e1;
    const bool s02 = c01 ? c12 : c02;                        // step 2 swaps
    const bool s12 = s02 ? (c01 || g01) : (c01 ? c02 : c12); // step 3 swaps
    const int da = s02 ? (s12 ? 1 : 2) : 0;                  // winner of step 1
    const int db = s12 ? 2 : 1;                              // loser of step 1
    d[0] = c01 ? db : da;
    d[1] = c01 ? da : db;
    d[2] = s02 ? 0 : (s12 ? 1 : 2);
}

// p sorted into its slots (for outputs that show the sorted event).
template <class T> __device__ __forceinline__ void to_slots(const T p[3][4], const int d[3], T ps[3][4])
{
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const T x0 = p[0][j], x1 = p[1][j], x2 = p[2][j];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            ps[k][j] = d[0] == k ? x0 : (d[1] == k ? x1 : x2);
    }
}

// ---- Cuts: evcut.rs:42-96 -------------------------------------------------
// On the unsorted photons with their slots d: every test but the energy and
// plane cuts treats the photons symmetrically (each dot product is the same
// IEEE expression for (i,j) and (j,i)), the energy cut takes the slot-2
// energy as the reference does after sorting, and the plane cut's normal
// n = p_slot0 x p_slot1 only changes sign (exactly) when its two photons are
// taken in the other order, which |num| and |n| do not see.  All tests are
// evaluated (no early exit), which changes nothing: they have no side effects.
__device__ __forceinline__ bool keep_f64(const Params &P, const double p[3][4], const int d[3])
{
    // (selects between loaded values, never between array elements: the
    // latter can turn into an address select that moves p to scratch)
    const double E0 = p[0][3], E1 = p[1][3], E2 = p[2][3];
    double emin;
    if (P.sort) {
        emin = d[0] == 2 ? E0 : (d[1] == 2 ? E1 : E2);
    } else {
        emin = E0;
        emin = (emin < E1) ? emin : E1;
        emin = (emin < E2) ? emin : E2;
    }
    bool ok = !(emin < P.e_min);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double num = p[i][0] * P.el_x;
        double den = p[i][3] * P.el_e;
        ok = ok & !(fabs(num) > P.cut_beam * den);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) {
            double num = (p[i][0] * p[j][0] + p[i][1] * p[j][1]) + p[i][2] * p[j][2];
            double den = p[i][3] * p[j][3];
            ok = ok & !(num > P.cut_pp * den);
        }
    if (P.cut_plane > 0.) {
        // the two photons outside slot 2
        double a[3], b[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double x0 = p[0][k], x1 = p[1][k], x2 = p[2][k];
            a[k] = d[0] == 2 ? x1 : x0;
            b[k] = d[2] == 2 ? x1 : x2;
        }
        double n0 = a[1] * b[2] - a[2] * b[1];
        double n1 = a[2] * b[0] - a[0] * b[2];
        double n2 = a[0] * b[1] - a[1] * b[0];
        double num = P.el_x * n0;
        double nn = sq((n0 * n0 + n1 * n1) + n2 * n2);
        ok = ok & !(fabs(num) < P.cut_plane * (P.el_e * nn));
    }
    return ok;
}

// ---- Spinor products and amplitudes: spinor.rs:31-162 ----------------------
//
// Particle indices follow the reference: 0 = e-, 1 = e+, 2..4 = photons.
template <class T> struct Spin {
    C<T> s[5][5]; // only off-diagonal entries are set
};

// a b - c d: with FAST one fma (the f64 event loop), otherwise the
// reference's two products and a difference.
template <bool FAST, class T> __device__ __forceinline__ T mul_sub(T a, T b, T c, T d)
{
    if constexpr (FAST)
        return fma(a, b, -(c * d));
    else
        return a * b - c * d;
}
template <bool FAST, class T> __device__ __forceinline__ T mul_add(T a, T b, T c, T d)
{
    if constexpr (FAST)
        return fma(a, b, c * d);
    else
        return a * b + c * d;
}

template <class T, bool FAST = false>
__device__ __forceinline__ void spinor_products(T xx0, T f0re, T f0im, T xx1, T f1re, T f1im, const T p[3][4],
                                                Spin<T> &S)
{
    T xx[3];
    C<T> fx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xx[k] = sq(p[k][3] + p[k][2]);
        // f32::MIN_POSITIVE / f64::MIN_POSITIVE
        const T minpos = sizeof(T) == 8 ? (T)2.2250738585072014e-308 : (T)1.17549435e-38f;
        if (xx[k] > minpos) {
            if constexpr (FAST) {
                const T y = tpm::rcp_n(xx[k]);
                fx[k] = mk<T>(p[k][0] * y, p[k][1] * y);
            } else {
                fx[k] = mk<T>(dv(p[k][0], xx[k]), dv(p[k][1], xx[k]));
            }
        }
        else
            fx[k] = mk<T>(sq((T)2. * p[k][3]), (T)0.);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        C<T> s0k = mk<T>(mul_sub<FAST>(f0re, xx[k], fx[k].re, xx0), mul_sub<FAST>(f0im, xx[k], fx[k].im, xx0));
        C<T> s1k = mk<T>(mul_sub<FAST>(f1re, xx[k], fx[k].re, xx1), mul_sub<FAST>(f1im, xx[k], fx[k].im, xx1));
        S.s[0][2 + k] = s0k;
        S.s[2 + k][0] = neg(s0k);
        S.s[1][2 + k] = s1k;
        S.s[2 + k][1] = neg(s1k);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = k + 1; l < 3; ++l) {
            C<T> skl = mk<T>(mul_sub<FAST>(fx[k].re, xx[l], fx[l].re, xx[k]),
                             mul_sub<FAST>(fx[k].im, xx[l], fx[l].im, xx[k]));
            S.s[2 + k][2 + l] = skl;
            S.s[2 + l][2 + k] = neg(skl);
        }
}

// Complex products of the amplitude algebra.  With FAST they contract
// multiply-adds into fma and divide by multiplying with one reciprocal of
// |b|^2 (in f64 the refined one, in f32 the hardware's), so the result moves
// by an ulp or so (the f64 event loops, the mixed-precision path, the
// adjoints).  Without FAST they are the reference's plain operations (the
// all-f32 parity path).
__device__ __forceinline__ C<double> cmul_fast(C<double> a, C<double> b)
{
    return C<double>{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ C<double> cdiv_fast(C<double> a, C<double> b)
{
    const double y = tpm::rcp_n(fma(b.re, b.re, b.im * b.im));
    return C<double>{fma(a.re, b.re, a.im * b.im) * y, fma(a.im, b.re, -(a.re * b.im)) * y};
}
__device__ __forceinline__ C<float> cmul_fast(C<float> a, C<float> b)
{
    return C<float>{fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ C<float> cdiv_fast(C<float> a, C<float> b)
{
    const float inv = __builtin_amdgcn_rcpf(fmaf(b.re, b.re, b.im * b.im));
    return C<float>{fmaf(a.re, b.re, a.im * b.im) * inv, fmaf(a.im, b.re, -(a.re * b.im)) * inv};
}
template <bool FAST, class T> __device__ __forceinline__ C<T> xmul(C<T> a, C<T> b)
{
    if constexpr (FAST)
        return cmul_fast(a, b);
    else
        return cmul(a, b);
}
template <bool FAST, class T> __device__ __forceinline__ C<T> xdiv(C<T> a, C<T> b)
{
    if constexpr (FAST)
        return cdiv_fast(a, b);
    else
        return cdiv(a, b);
}
template <bool FAST, class T> __device__ __forceinline__ T xnorm(C<T> z) { return mul_add<FAST>(z.re, z.re, z.im, z.im); }
template <bool FAST, class T> __device__ __forceinline__ C<T> xsq(C<T> z) { return xmul<FAST>(z, z); }

// Amplitude constants.  Plain (the all-f32 parity path): -RAC8*x*y/z is
// ((-RAC8*x)*y)/z with the per-run factors (-RAC8) s(e-,e+) and (-RAC8)
// t(e-,e+) (ns01, nt01), and the couplings multiply each amplitude after it.
// FAST: the beams run along x (evgen.rs:66-69), so s(e-,e+) and t(e-,e+) are
// real (imaginary parts +-0) for every configuration; each amplitude is its
// spinor-product expression scaled by one real constant, (-RAC8) s or t times
// its coupling, formed on the host (k_*; a few ulp from the reference's order
// of operations).
template <class T> struct AmpK {
    C<T> ns01, nt01;
    T g_a, g_bp, g_bm;
    T ka_s, ka_t, kp_s, kp_t, km_s, km_t; // FAST: ns01 g_a, nt01 g_a, ns01 g_bp, nt01 g_bp, ns01 g_bm, nt01 g_bm
};
template <class T> __device__ __forceinline__ C<T> rscale(C<T> z, T k) { return C<T>{z.re * k, z.im * k}; }
// The factor c (plain: complex, FAST: real k) times z.
template <bool FAST, class T> __device__ __forceinline__ C<T> kmul(C<T> c, T k, C<T> z)
{
    if constexpr (FAST)
        return rscale(z, k);
    else
        return cmul(c, z);
}

template <class T, bool FAST>
__device__ __forceinline__ C<T> a_ppm(const Spin<T> &S, const AmpK<T> &K, int k1, int k2, int k3)
{
    C<T> num = kmul<FAST>(K.ns01, K.ka_s, xsq<FAST>(S.s[0][k3]));
    C<T> den = xmul<FAST>(xmul<FAST>(xmul<FAST>(S.s[0][k1], S.s[0][k2]), S.s[1][k1]), S.s[1][k2]);
    return xdiv<FAST>(num, den);
}
template <class T, bool FAST>
__device__ __forceinline__ C<T> a_pmm(const Spin<T> &S, const AmpK<T> &K, int k1, int k2, int k3)
{
    C<T> num = kmul<FAST>(K.nt01, K.ka_t, xsq<FAST>(tof(S.s[1][k1])));
    C<T> den = xmul<FAST>(xmul<FAST>(xmul<FAST>(tof(S.s[1][k2]), tof(S.s[1][k3])), tof(S.s[0][k2])), tof(S.s[0][k3]));
    return xdiv<FAST>(num, den);
}
template <class T, bool FAST>
__device__ __forceinline__ C<T> bp_ppm(const Spin<T> &S, const AmpK<T> &K, int k1, int k2, int k3)
{
    return kmul<FAST>(K.nt01, K.kp_t, xsq<FAST>(xmul<FAST>(tof(S.s[k1][k2]), S.s[k3][0])));
}
template <class T, bool FAST>
__device__ __forceinline__ C<T> bp_pmm(const Spin<T> &S, const AmpK<T> &K, int k1, int k2, int k3)
{
    return kmul<FAST>(K.ns01, K.kp_s, xsq<FAST>(xmul<FAST>(tof(S.s[k1][1]), S.s[k2][k3])));
}
template <class T, bool FAST> __device__ __forceinline__ C<T> bm_ppp(const Spin<T> &S, const AmpK<T> &K)
{
    const int k1 = 2, k2 = 3, k3 = 4;
    C<T> s = cadd(cadd(xsq<FAST>(xmul<FAST>(tof(S.s[k1][k2]), tof(S.s[k3][1]))), xsq<FAST>(xmul<FAST>(tof(S.s[k1][k3]), tof(S.s[k2][1])))),
                  xsq<FAST>(xmul<FAST>(tof(S.s[k2][k3]), tof(S.s[k1][1]))));
    return kmul<FAST>(K.ns01, K.km_s, s);
}
template <class T, bool FAST> __device__ __forceinline__ C<T> bm_mmm(const Spin<T> &S, const AmpK<T> &K)
{
    const int k1 = 2, k2 = 3, k3 = 4;
    C<T> s = cadd(cadd(xsq<FAST>(xmul<FAST>(S.s[k1][0], S.s[k2][k3])), xsq<FAST>(xmul<FAST>(S.s[k2][0], S.s[k1][k3]))),
                  xsq<FAST>(xmul<FAST>(S.s[k3][0], S.s[k1][k2])));
    return kmul<FAST>(K.nt01, K.km_t, s);
}

// FAST: the A amplitudes' denominators.  a_ppm(k1, k2, k3) divides by
// D{k1,k2} = s(0,k1) s(0,k2) s(1,k1) s(1,k2) and a_pmm(k1, k2, k3) by
// t(1,k2) t(1,k3) t(0,k2) t(0,k3) = conj D{k2,k3} (t = -conj s, the four signs
// cancel), so the six amplitudes share three products D and their 1/|D|^2:
// num / D = num conj(D) / |D|^2 and num / conj(D) = num D / |D|^2.
template <class T> struct ADen {
    C<T> d;
    T inv; // 1 / |d|^2
};
__device__ __forceinline__ double rcp_amp(double x) { return tpm::rcp_n(x); }
__device__ __forceinline__ float rcp_amp(float x) { return __builtin_amdgcn_rcpf(x); }
template <class T> __device__ __forceinline__ ADen<T> a_den(const Spin<T> &S, int a, int b)
{
    const C<T> d = xmul<true>(xmul<true>(xmul<true>(S.s[0][a], S.s[0][b]), S.s[1][a]), S.s[1][b]);
    return ADen<T>{d, rcp_amp(xnorm<true>(d))};
}
template <class T> __device__ __forceinline__ C<T> a_ppm_d(const Spin<T> &S, const AmpK<T> &K, const ADen<T> &D, int k3)
{
    return rscale(xmul<true>(xsq<true>(S.s[0][k3]), mk<T>(D.d.re, -D.d.im)), K.ka_s * D.inv);
}
template <class T> __device__ __forceinline__ C<T> a_pmm_d(const Spin<T> &S, const AmpK<T> &K, const ADen<T> &D, int k1)
{
    return rscale(xmul<true>(xsq<true>(tof(S.s[1][k1])), D.d), K.ka_t * D.inv);
}

// Helicity sums of the five matrix-element contributions (matelems.rs:53-86):
// m[0] = sum|A g_a|^2, m[1] = sum|B+ g_b+|^2, m[2] = sum|B- g_b-|^2,
// m[3] + i m[4] = sum (2 A g_a) conj(B+ g_b+).  Helicities MMP..PPM carry A and
// B+ (spinor.rs:72-98), MMM and PPP carry B- only (spinor.rs:100-113).  The
// interference terms are summed without their factor 2, which multiplies the
// sums at the end: doubling is exact in binary, so every partial sum is the
// reference's halved and the results are the same bits.
template <class T, bool FAST>
__device__ __forceinline__ void helicity_sums(const Spin<T> &S, const AmpK<T> &K, T m[5])
{
    // (A, B+) for h = MMP, MPM, MPP, PMM, PMP, PPM, each folded into the sums
    // as soon as it is formed (same order of additions as a loop over h).
    T sA = 0, sP = 0, sR = 0, sI = 0;
    auto fold = [&](int h, C<T> a, C<T> b) {
        C<T> A = a, Pp = b;
        if constexpr (!FAST) {
            A = rscale(a, K.g_a);
            Pp = rscale(b, K.g_bp);
        }
        C<T> X = xmul<FAST>(A, mk<T>(Pp.re, -Pp.im));
        T nA = xnorm<FAST>(A);
        T nP = xnorm<FAST>(Pp);
        if (h == 0) {
            sA = nA;
            sP = nP;
            sR = X.re;
            sI = X.im;
        } else {
            sA = sA + nA;
            sP = sP + nP;
            sR = sR + X.re;
            sI = sI + X.im;
        }
    };
    if constexpr (FAST) {
        const ADen<T> d23 = a_den(S, 2, 3), d24 = a_den(S, 2, 4), d34 = a_den(S, 3, 4);
        fold(0, a_pmm_d(S, K, d23, 4), bp_pmm<T, FAST>(S, K, 4, 2, 3));
        fold(1, a_pmm_d(S, K, d24, 3), bp_pmm<T, FAST>(S, K, 3, 2, 4));
        fold(2, a_ppm_d(S, K, d34, 2), bp_ppm<T, FAST>(S, K, 3, 4, 2));
        fold(3, a_pmm_d(S, K, d34, 2), bp_pmm<T, FAST>(S, K, 2, 3, 4));
        fold(4, a_ppm_d(S, K, d24, 3), bp_ppm<T, FAST>(S, K, 2, 4, 3));
        fold(5, a_ppm_d(S, K, d23, 4), bp_ppm<T, FAST>(S, K, 2, 3, 4));
    } else {
        fold(0, a_pmm<T, FAST>(S, K, 4, 2, 3), bp_pmm<T, FAST>(S, K, 4, 2, 3));
        fold(1, a_pmm<T, FAST>(S, K, 3, 2, 4), bp_pmm<T, FAST>(S, K, 3, 2, 4));
        fold(2, a_ppm<T, FAST>(S, K, 3, 4, 2), bp_ppm<T, FAST>(S, K, 3, 4, 2));
        fold(3, a_pmm<T, FAST>(S, K, 2, 3, 4), bp_pmm<T, FAST>(S, K, 2, 3, 4));
        fold(4, a_ppm<T, FAST>(S, K, 2, 4, 3), bp_ppm<T, FAST>(S, K, 2, 4, 3));
        fold(5, a_ppm<T, FAST>(S, K, 2, 3, 4), bp_ppm<T, FAST>(S, K, 2, 3, 4));
    }
    C<T> Bm = bm_mmm<T, FAST>(S, K);
    C<T> Bp = bm_ppp<T, FAST>(S, K);
    if constexpr (!FAST) {
        Bm = rscale(Bm, K.g_bm);
        Bp = rscale(Bp, K.g_bm);
    }
    m[0] = sA;
    m[1] = sP;
    m[2] = xnorm<FAST>(Bm) + xnorm<FAST>(Bp);
    m[3] = (T)2. * sR;
    m[4] = (T)2. * sI;
}

// ---- mixed precision: f32 spinor/amplitude algebra ------------------------
//
// The mixed path runs the amplitudes with FAST complex products (fma and a
// hardware reciprocal, 1 ulp): it trades the reference's rounding for speed
// and is measured by its error against the f64 path on the same stream.
// After the cuts every spinor product of the shipped configuration has
// |s|^2 = 2 p_i.p_j >= 2.5 GeV^2, and every amplitude denominator stays far
// inside the f32 range.

// ez[k] = E_k + Z_k is formed in f64 by the caller: for a photon close to -Z
// the f32 difference of two rounded values could go negative.
__device__ __forceinline__ void spinor_products_f32(const Params &P, const float p[3][4], const float ez[3],
                                                    Spin<float> &S)
{
    const float xx0 = (float)P.xx0, f0re = (float)P.f0re, f0im = (float)P.f0im;
    const float xx1 = (float)P.xx1, f1re = (float)P.f1re, f1im = (float)P.f1im;
    float xx[3];
    C<float> fx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xx[k] = __builtin_amdgcn_sqrtf(ez[k]);
        if (xx[k] > 1.17549435e-38f) {
            const float inv = __builtin_amdgcn_rcpf(xx[k]);
            fx[k] = mk<float>(p[k][0] * inv, p[k][1] * inv);
        } else {
            fx[k] = mk<float>(__builtin_amdgcn_sqrtf(2.f * p[k][3]), 0.f);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        C<float> s0k = mk<float>(fmaf(f0re, xx[k], -(fx[k].re * xx0)), fmaf(f0im, xx[k], -(fx[k].im * xx0)));
        C<float> s1k = mk<float>(fmaf(f1re, xx[k], -(fx[k].re * xx1)), fmaf(f1im, xx[k], -(fx[k].im * xx1)));
        S.s[0][2 + k] = s0k;
        S.s[2 + k][0] = neg(s0k);
        S.s[1][2 + k] = s1k;
        S.s[2 + k][1] = neg(s1k);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = k + 1; l < 3; ++l) {
            C<float> skl = mk<float>(fmaf(fx[k].re, xx[l], -(fx[l].re * xx[k])),
                                     fmaf(fx[k].im, xx[l], -(fx[l].im * xx[k])));
            S.s[2 + k][2 + l] = skl;
            S.s[2 + l][2 + k] = neg(skl);
        }
}

// Per-event weight: dot(m, sigma_contribs) as a left fold (resacc.rs:126).
__device__ __forceinline__ double weight_f64(const Params &P, const double m[5])
{
    double w = m[0] * P.sc[0];
    w = w + m[1] * P.sc[1];
    w = w + m[2] * P.sc[2];
    w = w + m[3] * P.sc[3];
    w = w + m[4] * P.sc[4];
    return w;
}

// RAMBO event from 12 uniforms (photons sorted) and its cut decision.
// RAMBO event from 12 uniforms: photons as RAMBO makes them (unsorted), their
// sort slots d, and the cut decision.
template <bool SCALED = false>
__device__ __forceinline__ bool generate_f64(const Params &P, const double u[12], double p[3][4], int d[3])
{
    double q[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        raw_photon_f64<SCALED>(u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3], q[i]);
    conformal_f64(P, q, p);
    sort_slots(p, P.sort, d);
    return keep_f64(P, p, d);
}

// ---- faster-evgen re-rolls from accept bits -------------------------------------
// The stream's accept bits (bit i of acc1: the point (2u - 1) of positions
// bbase + i and bbase + i + 1 passes the unit-disc test), valid below blimit.
// acc1 null: none at hand.
struct RerollBits {
    const uint32_t *acc1;
    uint64_t bbase, blimit;
};

// Pair starts open to a RANF take(2) sequence that starts L = 55 - a
// positions before its round ends (bit k: a pair may start k positions on):
// even offsets whose pair fits before the round's end, then from each later
// round's start (offset L + 55 m) the even offsets of that round (its 55th
// position is left over, ranf.rs:89-92).
__device__ __forceinline__ uint64_t ranf_pair_starts(uint32_t L)
{
    const uint64_t E = 0x5555555555555555ull;
    auto low = [](uint32_t n) -> uint64_t { return n >= 64 ? ~0ull : (1ull << n) - 1ull; };
    uint64_t c = (E & low(L - 1)) | ((E << L) & low(L + 53));
    if (L + 55 < 64)
        c |= (E << (L + 55)) & low(L + 108);
    return c;
}

// Offsets from the re-roll start q of the first accepted pairs among the
// starts `cand` (k[0], k[1], k[2]: the first, second and third); false when
// the bits near q are not at hand or the 64 positions hold fewer than F
// accepted pairs.  Reads acc1 words up to position q + 95.
__device__ __forceinline__ bool accepted_pairs(const RerollBits &b, uint64_t q, uint64_t cand, uint32_t F,
                                               uint32_t k[3])
{
    if (!b.acc1 || q + 96 > b.blimit)
        return false;
    const uint64_t r = q - b.bbase;
    const uint32_t *w = b.acc1 + (r >> 5);
    const uint32_t s = (uint32_t)(r & 31);
    const uint64_t lo = ((uint64_t)w[1] << 32) | w[0];
    uint64_t y = (s ? (lo >> s) | ((uint64_t)w[2] << (64 - s)) : lo) & cand;
    if ((uint32_t)__popcll(y) < F)
        return false;
    k[0] = (uint32_t)__builtin_ctzll(y);
    y &= y - 1;
    k[1] = y ? (uint32_t)__builtin_ctzll(y) : 0u;
    y &= y - 1;
    k[2] = y ? (uint32_t)__builtin_ctzll(y) : 0u;
    return true;
}

// The pair of the j-th failed point (j = failed points before it).
__device__ __forceinline__ uint32_t pair_of(const uint32_t k[3], uint32_t j) { return j == 0 ? k[0] : j == 1 ? k[1] : k[2]; }

// ---- faster-evgen (evgen.rs:143-173, 221-249) --------------------------------
// Reads the RANF stream expanded in HBM.  Stream position q holds the value
// at index 55 - q % 55 of round q / 55; random_array::<N> first skips the
// round's leftovers when fewer than N remain, then returns its N values in
// ascending array index, i.e. descending stream position (ranf.rs:78-102).
struct RanfWalk {
    const int32_t *buf; // stream values from position qbase on
    uint64_t qbase;
    uint64_t q;         // next position
    RerollBits rb;
    __device__ __forceinline__ void take(int n, double *out)
    {
        const uint32_t rem = 55u - (uint32_t)(q % 55u);
        if (rem < (uint32_t)n)
            q += rem;
        const int32_t *v = buf + (q - qbase);
        for (int k = 0; k < n; ++k)
            out[k] = (double)v[n - 1 - k] * 1e-9;
        q += (uint64_t)n;
    }
    // The re-rolls of the failed points (bit i of fail: point i), each the
    // next accepted take(2) pair, as the re-roll loop takes them; false
    // (nothing taken) when the accept bits cannot tell.
    __device__ __forceinline__ bool rerolls(uint32_t fail, double x[3], double y[3], double r2[3])
    {
        uint32_t k[3];
        const uint32_t F = (uint32_t)__popc(fail);
        if (!accepted_pairs(rb, q, ranf_pair_starts(55u - (uint32_t)(q % 55u)), F, k))
            return false;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if ((fail >> i) & 1u) {
                const int32_t *v = buf + (q - qbase) + pair_of(k, (uint32_t)__popc(fail & ((1u << i) - 1u)));
                x[i] = 2. * ((double)v[1] * 1e-9) - 1.;
                y[i] = 2. * ((double)v[0] * 1e-9) - 1.;
                r2[i] = 0. + (x[i] * x[i] + y[i] * y[i]);
            }
        q += (uint64_t)pair_of(k, F - 1) + 2;
        return true;
    }
};

// The xoshiro256+ stream (standard.rs) expanded into HBM as its f64
// uniforms: random_array::<N> is N consecutive draws in array order
// (standard.rs:33-35, no round to discard).
struct XoWalk {
    const double *buf; // uniforms from position qbase on
    uint64_t qbase;
    uint64_t q;        // next position
    RerollBits rb;
    __device__ __forceinline__ void take(int n, double *out)
    {
        const double *v = buf + (q - qbase);
        for (int k = 0; k < n; ++k)
            out[k] = v[k];
        q += (uint64_t)n;
    }
    // As RanfWalk::rerolls; every even offset is a pair start.
    __device__ __forceinline__ bool rerolls(uint32_t fail, double x[3], double y[3], double r2[3])
    {
        uint32_t k[3];
        const uint32_t F = (uint32_t)__popc(fail);
        if (!accepted_pairs(rb, q, 0x5555555555555555ull, F, k))
            return false;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if ((fail >> i) & 1u) {
                const double *v = buf + (q - qbase) + pair_of(k, (uint32_t)__popc(fail & ((1u << i) - 1u)));
                x[i] = 2. * v[0] - 1.;
                y[i] = 2. * v[1] - 1.;
                r2[i] = 0. + (x[i] * x[i] + y[i] * y[i]);
            }
        q += (uint64_t)pair_of(k, F - 1) + 2;
        return true;
    }
};

// Massless photons from 9 + 6 draws and unit-disc rejection re-rolls, in the
// reference's operation order (matrix fills column-major, nalgebra sums).
template <class Walk> __device__ __forceinline__ void raw_photons_faster_f64(Walk &w, double q[3][4])
{
    double r9[9], r6[6], x[3], y[3], r2[3];
    w.take(9, r9);
    w.take(6, r6);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        x[i] = 2. * r6[i] - 1.;
        y[i] = 2. * r6[3 + i] - 1.;
        r2[i] = (0. + x[i] * x[i]) + y[i] * y[i];
    }
    const double min2 = 2.2250738585072014e-308 * 2.2250738585072014e-308;
    uint32_t fail = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        fail |= (r2[i] > 1. || r2[i] < min2) ? 1u << i : 0u;
    if (fail && !w.rerolls(fail, x, y, r2)) {
        for (int i = 0; i < 3; ++i)
            while (r2[i] > 1. || r2[i] < min2) {
                double r[2];
                w.take(2, r);
                x[i] = 2. * r[0] - 1.;
                y[i] = 2. * r[1] - 1.;
                r2[i] = 0. + (x[i] * x[i] + y[i] * y[i]);
            }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double nrm = dv(1., sq(r2[i]));
        const double cx = x[i] * nrm, cy = y[i] * nrm;
        const double cos_t = 2. * r9[i] - 1.;
        const double em = r9[3 + i] * r9[6 + i];
        const double st = sq(1. - cos_t * cos_t);
        const double en = -tpm::log_pos(em + 2.2250738585072014e-308);
        q[i][0] = en * (st * cx);
        q[i][1] = en * (st * cy);
        q[i][2] = en * cos_t;
        q[i][3] = en * 1.;
    }
}

template <class Walk> __device__ __forceinline__ bool generate_faster_f64(const Params &P, Walk &w, double p[3][4], int d[3])
{
    double q[3][4];
    raw_photons_faster_f64(w, q);
    conformal_f64(P, q, p);
    sort_slots(p, P.sort, d);
    return keep_f64(P, p, d);
}

// Helicity sums m[5] of an event that passed the cuts.
__device__ __forceinline__ void matrix_elements_f64(const Params &P, const double p[3][4], double m[5])
{
    Spin<double> S;
    spinor_products<double, true>(P.xx0, P.f0re, P.f0im, P.xx1, P.f1re, P.f1im, p, S);
    const AmpK<double> K{mk(P.ns01re, P.ns01im), mk(P.nt01re, P.nt01im), P.g_a, P.g_bp, P.g_bm,
                         P.ka_s, P.ka_t, P.kp_s, P.kp_t, P.km_s, P.km_t};
    helicity_sums<double, true>(S, K, m);
}

// Mixed precision keeps RAMBO (massless q, conformal map), the sort and the
// cuts in f64: the conformal map divides by M^2 = R_E^2 - |R|^2, which
// cancels for nearly collinear q and would turn f32 rounding of q into
// percent-level momentum errors on rare events.  So the mixed path sees the
// same momenta and cut decisions as the f64 path, and only the spinor and
// amplitude algebra runs in f32.
__device__ __forceinline__ void matrix_elements_mixed(const Params &P, const double p[3][4], double m[5])
{
    float pf[3][4], ez[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            pf[i][j] = (float)p[i][j];
        ez[i] = (float)fmax(p[i][3] + p[i][2], 0.);
    }
    Spin<float> S;
    spinor_products_f32(P, pf, ez, S);
    float mf[5];
    const AmpK<float> K{mk((float)P.ns01re, (float)P.ns01im), mk((float)P.nt01re, (float)P.nt01im),
                        (float)P.g_a, (float)P.g_bp, (float)P.g_bm, (float)P.ka_s, (float)P.ka_t,
                        (float)P.kp_s, (float)P.kp_t, (float)P.km_s, (float)P.km_t};
    helicity_sums<float, true>(S, K, mf);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        m[c] = (double)mf[c];
}

template <int PREC> __device__ __forceinline__ void matrix_elements(const Params &P, const double p[3][4], double m[5])
{
    if (PREC == 0)
        matrix_elements_f64(P, p, m);
    else
        matrix_elements_mixed(P, p, m);
}

// ---- all-f32 mode (numeric.rs:6-14, feature f32) ----------------------------
// Every quantity in f32 with the reference's operation order and plain IEEE
// operations (no contraction: the translation unit is built with
// -ffp-contract=off; FAST = false in the amplitude algebra).  Run constants
// are computed on the host in f32 arithmetic.
struct FParams {
    float e_total, cut_beam, cut_pp, e_min, cut_plane;
    float g_a, g_bp, g_bm;
    float sc[5];
    float el_x, el_e;
    float xx0, f0re, f0im, xx1, f1re, f1im;
    float ns01re, ns01im, nt01re, nt01im;
    int sort;
};

// sin, cos and ln of an f32 argument are evaluated in f64 and rounded once:
// the correctly rounded f32 result in all but astronomically rare cases,
// which is what the host libm's f32 functions return (within their < 1 ulp
// error bound) far more often than a faster f32 device routine would.
__device__ __forceinline__ void raw_photon_f32(float uc, float uphi, float ue1, float ue2, float q[4])
{
    const float two_pi = 2.f * 3.14159265358979323846f;
    float cos_t = 2.f * uc - 1.f;
    float phi = two_pi * uphi;
    float em = ue1 * ue2;
    double sd, cd;
    tpm::sincos_f64((double)phi, sd, cd);
    const float sphi = (float)sd, cphi = (float)cd;
    float st = sqrtf(1.f - cos_t * cos_t);
    float en = -(float)tpm::log_pos((double)(em + 1.17549435e-38f));
    q[0] = en * (st * sphi);
    q[1] = en * (st * cphi);
    q[2] = en * cos_t;
    q[3] = en;
}

__device__ __forceinline__ void conformal_f32(const FParams &P, const float q[3][4], float p[3][4])
{
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        r[k] = (q[0][k] + q[1][k]) + q[2][k];
    float rn2 = r[3] * r[3] - ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    float alpha = P.e_total / rn2;
    float rn = sqrtf(rn2);
    float beta = 1.f / (rn + r[3]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float rq = q[i][0] * r[0];
        rq = q[i][1] * r[1] + rq;
        rq = q[i][2] * r[2] + rq;
        p[i][3] = alpha * (r[3] * q[i][3] - rq);
        float b = beta * rq - q[i][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            p[i][k] = alpha * (rn * q[i][k] + b * r[k]);
    }
}

__device__ __forceinline__ bool keep_f32(const FParams &P, const float p[3][4])
{
    float emin;
    if (P.sort) {
        emin = p[2][3];
    } else {
        emin = p[0][3];
        emin = (emin < p[1][3]) ? emin : p[1][3];
        emin = (emin < p[2][3]) ? emin : p[2][3];
    }
    bool ok = !(emin < P.e_min);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float num = p[i][0] * P.el_x;
        float den = p[i][3] * P.el_e;
        ok = ok && !(fabsf(num) > P.cut_beam * den);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) {
            float num = (p[i][0] * p[j][0] + p[i][1] * p[j][1]) + p[i][2] * p[j][2];
            float den = p[i][3] * p[j][3];
            ok = ok && !(num > P.cut_pp * den);
        }
    if (P.cut_plane > 0.f) {
        float n0 = p[0][1] * p[1][2] - p[0][2] * p[1][1];
        float n1 = p[0][2] * p[1][0] - p[0][0] * p[1][2];
        float n2 = p[0][0] * p[1][1] - p[0][1] * p[1][0];
        float num = P.el_x * n0;
        float nn = sqrtf((n0 * n0 + n1 * n1) + n2 * n2);
        ok = ok && !(fabsf(num) < P.cut_plane * (P.el_e * nn));
    }
    return ok;
}

// One f32 event from 12 uniforms: photon momenta (sorted when P.sort) and
// the cut decision.
__device__ __forceinline__ bool generate_f32(const FParams &P, const float u[12], float p[3][4])
{
    float q[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        raw_photon_f32(u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3], q[i]);
    conformal_f32(P, q, p);
    if (P.sort)
        sort3(p);
    return keep_f32(P, p);
}

// Helicity sums m[5] of an f32 event that passed the cuts.
__device__ __forceinline__ void matrix_elements_f32(const FParams &P, const float p[3][4], float m[5])
{
    Spin<float> S;
    spinor_products<float>(P.xx0, P.f0re, P.f0im, P.xx1, P.f1re, P.f1im, p, S);
    const AmpK<float> K{mk(P.ns01re, P.ns01im), mk(P.nt01re, P.nt01im), P.g_a, P.g_bp, P.g_bm, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    helicity_sums<float, false>(S, K, m);
}

// ---- faster-evgen in f32 (evgen.rs:143-173, 221-249 with Float = f32) -------
// The RANF values as (n as f32) * 1e-9 (ranf.rs:99), random_array's round
// discards as in RanfWalk.
struct RanfWalkF32 {
    const int32_t *buf;
    uint64_t qbase;
    uint64_t q;
    RerollBits rb;
    __device__ __forceinline__ void take(int n, float *out)
    {
        const uint32_t rem = 55u - (uint32_t)(q % 55u);
        if (rem < (uint32_t)n)
            q += rem;
        const int32_t *v = buf + (q - qbase);
        for (int k = 0; k < n; ++k)
            out[k] = (float)v[n - 1 - k] * 1e-9f;
        q += (uint64_t)n;
    }
    // As RanfWalk::rerolls, in f32.
    __device__ __forceinline__ bool rerolls(uint32_t fail, float x[3], float y[3], float r2[3])
    {
        uint32_t k[3];
        const uint32_t F = (uint32_t)__popc(fail);
        if (!accepted_pairs(rb, q, ranf_pair_starts(55u - (uint32_t)(q % 55u)), F, k))
            return false;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if ((fail >> i) & 1u) {
                const int32_t *v = buf + (q - qbase) + pair_of(k, (uint32_t)__popc(fail & ((1u << i) - 1u)));
                x[i] = 2.f * ((float)v[1] * 1e-9f) - 1.f;
                y[i] = 2.f * ((float)v[0] * 1e-9f) - 1.f;
                r2[i] = 0.f + (x[i] * x[i] + y[i] * y[i]);
            }
        q += (uint64_t)pair_of(k, F - 1) + 2;
        return true;
    }
};

// The xoshiro128+ stream (standard.rs with the f32 feature) expanded into HBM
// as its f32 uniforms: random_array::<N> is N consecutive draws in array
// order (standard.rs:33-35, no round to discard).
struct Xo128WalkF32 {
    const float *buf; // uniforms from position qbase on
    uint64_t qbase;
    uint64_t q;       // next position
    RerollBits rb;
    __device__ __forceinline__ void take(int n, float *out)
    {
        const float *v = buf + (q - qbase);
        for (int k = 0; k < n; ++k)
            out[k] = v[k];
        q += (uint64_t)n;
    }
    // As XoWalk::rerolls, in f32.
    __device__ __forceinline__ bool rerolls(uint32_t fail, float x[3], float y[3], float r2[3])
    {
        uint32_t k[3];
        const uint32_t F = (uint32_t)__popc(fail);
        if (!accepted_pairs(rb, q, 0x5555555555555555ull, F, k))
            return false;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if ((fail >> i) & 1u) {
                const float *v = buf + (q - qbase) + pair_of(k, (uint32_t)__popc(fail & ((1u << i) - 1u)));
                x[i] = 2.f * v[0] - 1.f;
                y[i] = 2.f * v[1] - 1.f;
                r2[i] = 0.f + (x[i] * x[i] + y[i] * y[i]);
            }
        q += (uint64_t)pair_of(k, F - 1) + 2;
        return true;
    }
};

// f32::MIN_POSITIVE squared underflows to 0 in f32, so the re-roll test
// r2 < MIN_POSITIVE^2 never fires, as in the reference's f32 build.
template <class Walk> __device__ __forceinline__ void raw_photons_faster_f32(Walk &w, float q[3][4])
{
    float r9[9], r6[6], x[3], y[3], r2[3];
    w.take(9, r9);
    w.take(6, r6);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        x[i] = 2.f * r6[i] - 1.f;
        y[i] = 2.f * r6[3 + i] - 1.f;
        r2[i] = (0.f + x[i] * x[i]) + y[i] * y[i];
    }
    const float minpos = 1.17549435e-38f;
    const float min2 = minpos * minpos;
    uint32_t fail = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        fail |= (r2[i] > 1.f || r2[i] < min2) ? 1u << i : 0u;
    if (fail && !w.rerolls(fail, x, y, r2)) {
        for (int i = 0; i < 3; ++i)
            while (r2[i] > 1.f || r2[i] < min2) {
                float r[2];
                w.take(2, r);
                x[i] = 2.f * r[0] - 1.f;
                y[i] = 2.f * r[1] - 1.f;
                r2[i] = 0.f + (x[i] * x[i] + y[i] * y[i]);
            }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float nrm = 1.f / sqrtf(r2[i]);
        const float cx = x[i] * nrm, cy = y[i] * nrm;
        const float cos_t = 2.f * r9[i] - 1.f;
        const float em = r9[3 + i] * r9[6 + i];
        const float st = sqrtf(1.f - cos_t * cos_t);
        const float en = -(float)tpm::log_pos((double)(em + minpos)); // correctly rounded f32 ln
        q[i][0] = en * (st * cx);
        q[i][1] = en * (st * cy);
        q[i][2] = en * cos_t;
        q[i][3] = en * 1.f;
    }
}

template <class Walk> __device__ __forceinline__ bool generate_faster_f32(const FParams &P, Walk &w, float p[3][4])
{
    float q[3][4];
    raw_photons_faster_f32(w, q);
    conformal_f32(P, q, p);
    if (P.sort)
        sort3(p);
    return keep_f32(P, p);
}

// One f32 event from 12 uniforms: cut decision and helicity sums m[5].
__device__ __forceinline__ bool event_f32(const FParams &P, const float u[12], float m[5])
{
    float p[3][4];
    if (!generate_f32(P, u, p))
        return false;
    matrix_elements_f32(P, p, m);
    return true;
}

__device__ __forceinline__ float weight_f32(const FParams &P, const float m[5])
{
    float w = m[0] * P.sc[0];
    w = w + m[1] * P.sc[1];
    w = w + m[2] * P.sc[2];
    w = w + m[3] * P.sc[3];
    w = w + m[4] * P.sc[4];
    return w;
}

} // namespace tpd
