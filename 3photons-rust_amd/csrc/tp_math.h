// tp_math.h -- f64 log and sin/cos for the RAMBO inputs of the event loop
// (evgen.rs:174-207: -ln(u1 u2 + MIN_POSITIVE), sin and cos of 2 pi u).
//
// Both functions are table-driven and finish with an exact (double-double)
// sum, so the result is the correctly rounded value except within ~0.002 ulp
// of a rounding boundary: they agree bit for bit with the host libm the
// oracle calls in all but a few cases per thousand, at a third of the
// instructions of a general-purpose double-double implementation.
//
//   log_pos(x), x a positive normal double: P. T. P. Tang's table-driven
//   method (ACM TOMS 16(4), 1990).  x = m 2^k, m in [0.5, 1); entry j (the
//   top 8 fraction bits) holds INV, 1/c rounded to 9 bits, so that
//   r = m INV - 1 is exact with |r| < 2^-8, and -log(INV) as a head on the
//   2^-42 grid plus a tail.  log x = (k LN2_HI + head) + r + log1p(r) - r +
//   (k LN2_LO + tail); the first bracket is exact, it meets r in an exact
//   Fast2Sum (the head is 0 or at least 2^-9, r below 2^-8), and
//   log1p(r) - r is its Taylor series to r^8.
//
//   sincos_f64(x, s, c), |x| < 12.5 (so |j| <= 255): Cody-Waite reduction
//   x = j pi/64 + r with pi/64 in three parts (r as rh + rl), then
//   sin(a + r) = S + C r + [S (cos r - 1) + C (sin r - r)] and
//   cos(a + r) = C - S r + [C (cos r - 1) - S (sin r - r)] with S, C =
//   sin, cos of a = j pi/64 as double-double pairs; C r is formed exactly
//   and meets S in a Fast2Sum (|S| >= sin(pi/64) > |r| unless S = 0), the
//   bracket is small and carries the Taylor series of sin r - r (to r^7)
//   and cos r - 1 (to r^8), |r| <= pi/128.
//
// The tables (tp_mathtab.h) come from gen_tp_mathtab.py.  The same source
// builds for the host (g++), where the tests measure it against the libm.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TPM_FN __device__ __forceinline__
#define TPM_TABLE static __constant__ const
#else
#include <cmath>
#define TPM_FN inline
#define TPM_TABLE static const
#endif

#include "tp_mathtab.h"

namespace tpm {

TPM_FN uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
TPM_FN double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
TPM_FN double fma_(double a, double b, double c)
{
#if defined(__HIPCC__)
    return __builtin_fma(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}
TPM_FN double rint_(double x)
{
#if defined(__HIPCC__)
    return __builtin_rint(x);
#else
    return std::rint(x);
#endif
}

// a / b and sqrt(x) where operands, quotient and root lie well inside the
// normal range (exponents within about +-700; sqrt also takes an exact 0),
// which is the case everywhere in the event loop.  On the device they are the
// hardware's IEEE-correct Newton sequences (v_rcp_f64 / v_rsq_f64 seed, two
// refinements, residual correction) without the operand scaling and
// special-value fix-ups that only act outside that range, so they return the
// same correctly rounded values as '/' and sqrt() in 8 and 13 instructions
// instead of 11 and 17.
#if defined(__HIPCC__)
TPM_FN double div_n(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}
// 1 / b for the same b: the refined reciprocal of div_n, within 1 ulp
// (not always the nearest double).  a * rcp_n(b) is within 1.5 ulp of a / b:
// for quotients that may move by an ulp or so (the amplitude algebra).
TPM_FN double rcp_n(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.);
    return __builtin_fma(y, e, y);
}
TPM_FN double sqrt_n(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0. ? x : g;
}
#else
TPM_FN double div_n(double a, double b) { return a / b; }
TPM_FN double rcp_n(double b) { return 1. / b; }
TPM_FN double sqrt_n(double x) { return std::sqrt(x); }
#endif

// Natural logarithm of a positive normal double.
TPM_FN double log_pos(double x)
{
    const uint64_t b = bits_of(x);
    const int k = (int)(b >> 52) - 1022;                                   // x = m 2^k
    const double m = from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull); // [0.5, 1)
    const double *T = LOG_TAB[(b >> 44) & 255];
    const double r = fma_(m, T[0], -1.);                                   // exact
    const double kd = (double)k;
    const double hi = fma_(kd, LN2_HI, T[1]);                              // exact
    const double lo = fma_(kd, LN2_LO, T[2]);
    // log1p(r) - r = r^2 (-1/2 + r/3 - r^2/4 + r^3/5 - r^4/6 + r^5/7 - r^6/8)
    double q = -0.125;
    q = fma_(q, r, 1. / 7.);
    q = fma_(q, r, -1. / 6.);
    q = fma_(q, r, 0.2);
    q = fma_(q, r, -0.25);
    q = fma_(q, r, 1. / 3.);
    q = fma_(q, r, -0.5);
    const double p = (r * r) * q;
    const double s = hi + r;
    const double e = (hi - s) + r; // Fast2Sum
    return s + ((p + lo) + e);
}

// sin and cos of |x| < 12.5.
TPM_FN void sincos_f64(double x, double &sn, double &cs)
{
    const double jd = rint_(x * INV_PI64);
    const double r1 = fma_(-jd, PI64_1, x); // exact
    const double t = jd * PI64_2;           // exact
    const double rh = r1 - t;
    const double rl = fma_(-jd, PI64_3, (r1 - rh) - t);
    const double *T = SINCOS_TAB[(int)jd & 127];
    const double s_hi = T[0], s_lo = T[1], c_hi = T[2], c_lo = T[3];
    const double z = rh * rh;
    // sin r - rh = rh^3 (-1/6 + z/120 - z^2/5040) + rl
    double ps = -1. / 5040.;
    ps = fma_(ps, z, 1. / 120.);
    ps = fma_(ps, z, -1. / 6.);
    const double sr = fma_(rh * z, ps, rl);
    // cos r - 1 = z (-1/2 + z/24 - z^2/720 + z^3/40320) - rh rl
    double pc = 1. / 40320.;
    pc = fma_(pc, z, -1. / 720.);
    pc = fma_(pc, z, 1. / 24.);
    pc = fma_(pc, z, -0.5);
    const double cm = fma_(z, pc, -(rh * rl));
    // sin: S + C rh (exact product, Fast2Sum) + small terms
    const double bs = c_hi * rh;
    const double bs_lo = fma_(c_hi, rh, -bs);
    const double ss = s_hi + bs;
    const double se = (s_hi - ss) + bs;
    const double s_small = fma_(s_hi, cm, fma_(c_hi, sr, fma_(c_lo, rh, s_lo)));
    sn = ss + ((s_small + bs_lo) + se);
    // cos: C - S rh (exact product, Fast2Sum) + small terms
    const double bc = -(s_hi * rh);
    const double bc_lo = fma_(-s_hi, rh, -bc);
    const double cc = c_hi + bc;
    const double ce = (c_hi - cc) + bc;
    const double c_small = fma_(c_hi, cm, fma_(-s_hi, sr, fma_(-s_lo, rh, c_lo)));
    cs = cc + ((c_small + bc_lo) + ce);
}

} // namespace tpm
