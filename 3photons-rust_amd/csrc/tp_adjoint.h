// tp_adjoint.h -- hand-written reverse mode of the per-event integrand
//
//   F_c(u) = w * |M|^2_c(p(u)) * cut(p(u)),   c = A, B+, B-, R_MX, I_MX
//
// through RAMBO (evgen.rs:89-207), the energy sort, the spinor products and
// the helicity amplitudes (spinor.rs:31-162, matelems.rs:53-86).  w is the
// constant massless-RAMBO weight, so dw/du = 0; the cut is piecewise constant,
// so it only gates.  One call differentiates L = sum_c lam_c |M|^2_c.
//
// Complex adjoints follow z_bar = (dL/dRe z, dL/dIm z); for a holomorphic
// w = f(z) this gives z_bar += w_bar * conj(f'(z)).  Every amplitude is a
// monomial or a ratio of monomials in spinor products v, so
// dA/dv = e_v A / v with e_v its exponent, and
// v_bar += e_v * (A_bar conj(A)) * conj(1/v).
// The t = -conj(s) and s(j,i) = -s(i,j) variables are folded back onto the
// nine independent products s(0,k), s(1,k), s(k,l) at the end.
#pragma once

#include "tp_physics.h"

namespace tpd {

template <class T> struct AdjParams {
    T e_total, cut_beam, cut_pp, e_min, cut_plane;
    T g_a, g_bp, g_bm, w;
    T el_x, el_e;
    T xx0, f0re, f0im, xx1, f1re, f1im;
    C<T> ns01, nt01;
    int sort;
};

template <class T> __host__ __device__ inline AdjParams<T> adj_params(const Params &P, double w)
{
    AdjParams<T> A;
    A.e_total = (T)P.e_total;
    A.cut_beam = (T)P.cut_beam;
    A.cut_pp = (T)P.cut_pp;
    A.e_min = (T)P.e_min;
    A.cut_plane = (T)P.cut_plane;
    A.g_a = (T)P.g_a;
    A.g_bp = (T)P.g_bp;
    A.g_bm = (T)P.g_bm;
    A.w = (T)w;
    A.el_x = (T)P.el_x;
    A.el_e = (T)P.el_e;
    A.xx0 = (T)P.xx0;
    A.f0re = (T)P.f0re;
    A.f0im = (T)P.f0im;
    A.xx1 = (T)P.xx1;
    A.f1re = (T)P.f1re;
    A.f1im = (T)P.f1im;
    A.ns01 = C<T>{(T)P.ns01re, (T)P.ns01im};
    A.nt01 = C<T>{(T)P.nt01re, (T)P.nt01im};
    A.sort = P.sort;
    return A;
}

// f64: the event loop's correctly rounded routines (tp_math.h); f32: the device library.
__device__ __forceinline__ void tsincos(double x, double *s, double *c) { tpm::sincos_f64(x, *s, *c); }
__device__ __forceinline__ void tsincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ double tlog(double x) { return tpm::log_pos(x); }
__device__ __forceinline__ float tlog(float x) { return logf(x); }
template <class T> __device__ __forceinline__ T minpos() { return sizeof(T) == 8 ? (T)2.2250738585072014e-308 : (T)1.17549435e-38f; }

template <class T> __device__ __forceinline__ C<T> cconj(C<T> z) { return C<T>{z.re, -z.im}; }
template <class T> __device__ __forceinline__ C<T> cscale(T r, C<T> z) { return C<T>{r * z.re, r * z.im}; }
template <class T> __device__ __forceinline__ C<T> caddv(C<T> a, C<T> b) { return C<T>{a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ C<T> cinv(C<T> z)
{
    const T n = z.re * z.re + z.im * z.im;
    return C<T>{dv(z.re, n), dv(-z.im, n)};
}
// Re(a * conj(b))
template <class T> __device__ __forceinline__ T cdot(C<T> a, C<T> b) { return a.re * b.re + a.im * b.im; }

// Adjoint bookkeeping over directed pairs (i, j) of particles: sb[i][j] holds
// dL/ds(i,j) and tb[i][j] dL/dt(i,j) as if they were independent variables.
template <class T> struct SpinAdj {
    C<T> sb[5][5], tb[5][5];
};

template <class T> struct SpinVal {
    C<T> s[5][5];   // values (off-diagonal)
    C<T> is[5][5];  // 1 / s(i,j)
};

template <class T> __device__ __forceinline__ C<T> tval(const SpinVal<T> &V, int i, int j)
{
    return C<T>{-V.s[i][j].re, V.s[i][j].im};
}
// 1/t = 1/(-conj(s)) = -conj(1/s)
template <class T> __device__ __forceinline__ C<T> tinv(const SpinVal<T> &V, int i, int j)
{
    return C<T>{-V.is[i][j].re, V.is[i][j].im};
}

// v_bar += e * gamma * conj(1/v)
template <class T> __device__ __forceinline__ void add_s(SpinAdj<T> &D, const SpinVal<T> &V, int i, int j, T e, C<T> g)
{
    D.sb[i][j] = caddv(D.sb[i][j], cscale(e, xmul<true>(g, cconj(V.is[i][j]))));
}
template <class T> __device__ __forceinline__ void add_t(SpinAdj<T> &D, const SpinVal<T> &V, int i, int j, T e, C<T> g)
{
    D.tb[i][j] = caddv(D.tb[i][j], cscale(e, xmul<true>(g, cconj(tinv(V, i, j)))));
}

template <class T> __device__ __forceinline__ T log_sigmoid(T x) { return -(fmax(-x, (T)0) + log1p(exp(-fabs(x)))); }
template <class T> __device__ __forceinline__ T sigmoid(T x) { return (T)1 / ((T)1 + exp(-x)); }

// Forward value and reverse sweep of one event.  Returns false (and zeros)
// for a cut event.  u[12] uniforms, lam[5] channel weights; outputs m[5]
// (|M|^2 per channel), du[12] = dL/du * w, dp[12] = dL/dp_sorted * w.
//
// SOFT = 1 and 2 replace the energy, beam and photon-pair cuts
// (evcut.rs:42-96) by the smooth factor S = prod_k sigmoid(margin_k / tau)
// over the dimensionless margins (E_i - E_min) / E_min, cut_beam - |X_i| / E_i
// and cut_pp - cos(theta_ij), and m[c] becomes |M|^2_c * S.  The plane cut
// stays a hard gate.  These are training integrands of the learned proposal,
// whose gradient must see the cut boundaries:
//   SOFT 1: L = S sum_c lam_c |M|^2_c; events with a margin below
//           max(-30 tau, -1/2) are dropped (S < 1e-13 for tau <= 1/60);
//   SOFT 2: L = log(min(f, fcap)) + log S with f = sum_c lam_c |M|^2_c,
//           whose gradient keeps pointing back into the cuts however far
//           outside an event lies; the cap keeps the soft and collinear
//           singularities just outside the cuts from dominating (events with
//           a photon of zero energy or f <= 0 are dropped).
//
// event_adjoint_sweeps runs the forward pass once and then NLAM reverse
// sweeps over the stored forward state: NLAM 1 with the given lam, NLAM 5
// (SOFT 0 only) with lam = e_0 .. e_4, the five coupling components.  After
// each sweep it calls sink(c, m, du, dp); a cut event calls it NLAM times
// with zeros.  With CONVERGE a cut event does not leave early: it runs the
// same passes on its (possibly non-finite) values and hands the sink zeros,
// so that every lane of a wavefront reaches each sink call at the same
// point of the program (a sink that synchronises the workgroup needs that;
// the lanes of a wavefront execute both paths anyway).
template <class T, int SOFT, int NLAM, bool CONVERGE = false, class Sink>
__device__ __forceinline__ bool event_adjoint_sweeps(const AdjParams<T> &P, const T u[12], const T *lam_in, T m[5],
                                                     T du[12], T dp[12], T tau, T fcap, Sink &&sink)
{
    static_assert(NLAM == 1 || (NLAM == 5 && SOFT == 0), "five sweeps need the plain integrand");
    const T one = (T)1, two = (T)2, two_pi = (T)(2. * 3.14159265358979323846);
    // ---- forward: massless photons from the uniforms (evgen.rs:174-207)
    T ct[3], sp[3], cp[3], st[3], en[3], em[3], q[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ct[i] = two * u[4 * i] - one;
        const T ph = two_pi * u[4 * i + 1];
        em[i] = u[4 * i + 2] * u[4 * i + 3];
        tsincos(ph, &sp[i], &cp[i]);
        st[i] = sq(one - ct[i] * ct[i]);
        en[i] = -tlog(em[i] + minpos<T>());
        q[i][0] = en[i] * (st[i] * sp[i]);
        q[i][1] = en[i] * (st[i] * cp[i]);
        q[i][2] = en[i] * ct[i];
        q[i][3] = en[i];
    }
    // ---- forward: conformal map (evgen.rs:93-106)
    T R[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        R[k] = (q[0][k] + q[1][k]) + q[2][k];
    const T rn2 = R[3] * R[3] - ((R[0] * R[0] + R[1] * R[1]) + R[2] * R[2]);
    const T alpha = dv(P.e_total, rn2);
    const T rn = sq(rn2);
    const T beta = dv(one, rn + R[3]);
    T rq[3], b[3], p[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        rq[i] = (q[i][0] * R[0] + q[i][1] * R[1]) + q[i][2] * R[2];
        p[i][3] = alpha * (R[3] * q[i][3] - rq[i]);
        b[i] = beta * rq[i] - q[i][3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            p[i][k] = alpha * (rn * q[i][k] + b[i] * R[k]);
    }
    // ---- sort (evgen.rs:109-118) as a permutation: ps[j] = p[perm[j]]
    int perm[3] = {0, 1, 2};
    if (P.sort) {
        T e0 = p[0][3], e1 = p[1][3], e2 = p[2][3];
        if (e1 > e0) { T t = e0; e0 = e1; e1 = t; int k = perm[0]; perm[0] = perm[1]; perm[1] = k; }
        if (e2 > e0) { T t = e0; e0 = e2; e2 = t; int k = perm[0]; perm[0] = perm[2]; perm[2] = k; }
        if (e2 > e1) { T t = e1; e1 = e2; e2 = t; int k = perm[1]; perm[1] = perm[2]; perm[2] = k; }
    }
    T ps[3][4];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            ps[j][k] = perm[j] == 0 ? p[0][k] : (perm[j] == 1 ? p[1][k] : p[2][k]);
    // ---- cuts (evcut.rs:42-96)
    bool keep = true;
    T lsig = (T)0;                                     // log S (SOFT)
    T gE[3] = {(T)0, (T)0, (T)0}, gB[3] = {(T)0, (T)0, (T)0}, gP[3] = {(T)0, (T)0, (T)0}; // dlogS/dmargin
    if (SOFT) {
        // SOFT 1 keeps E_i >= E_min / 2 > 0; SOFT 2 only needs E_i > 0
        const T far = SOFT == 1 ? fmax((T)-30 * tau, (T)-0.5) : (T)-1 + (T)1e-12;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            keep = keep && (ps[i][3] - P.e_min) / P.e_min > far;
        if (keep) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const T mE = (ps[i][3] - P.e_min) / P.e_min;
                const T mB = P.cut_beam - fabs(ps[i][0]) / ps[i][3];
                lsig = lsig + log_sigmoid(mE / tau) + log_sigmoid(mB / tau);
                gE[i] = sigmoid(-mE / tau) / tau;
                gB[i] = sigmoid(-mB / tau) / tau;
                keep = keep && (SOFT == 2 || mB > far);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int i = k == 2 ? 1 : 0, j = k == 0 ? 1 : 2;
                const T cs = ((ps[i][0] * ps[j][0] + ps[i][1] * ps[j][1]) + ps[i][2] * ps[j][2]) / (ps[i][3] * ps[j][3]);
                const T mP = P.cut_pp - cs;
                lsig = lsig + log_sigmoid(mP / tau);
                gP[k] = sigmoid(-mP / tau) / tau;
                keep = keep && (SOFT == 2 || mP > far);
            }
        }
        if (P.cut_plane > (T)0) {
            const T n0 = ps[0][1] * ps[1][2] - ps[0][2] * ps[1][1];
            const T n1 = ps[0][2] * ps[1][0] - ps[0][0] * ps[1][2];
            const T n2 = ps[0][0] * ps[1][1] - ps[0][1] * ps[1][0];
            keep = keep && !(fabs(P.el_x * n0) < P.cut_plane * (P.el_e * sqrt((n0 * n0 + n1 * n1) + n2 * n2)));
        }
    } else {
        T emin = ps[2][3];
        if (!P.sort)
            emin = fmin(fmin(ps[0][3], ps[1][3]), ps[2][3]);
        keep = !(emin < P.e_min);
#pragma unroll
        for (int i = 0; i < 3; ++i)
            keep = keep && !(fabs(ps[i][0] * P.el_x) > P.cut_beam * (ps[i][3] * P.el_e));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = i + 1; j < 3; ++j)
                keep = keep && !(((ps[i][0] * ps[j][0] + ps[i][1] * ps[j][1]) + ps[i][2] * ps[j][2]) >
                                 P.cut_pp * (ps[i][3] * ps[j][3]));
        if (P.cut_plane > (T)0) {
            const T n0 = ps[0][1] * ps[1][2] - ps[0][2] * ps[1][1];
            const T n1 = ps[0][2] * ps[1][0] - ps[0][0] * ps[1][2];
            const T n2 = ps[0][0] * ps[1][1] - ps[0][1] * ps[1][0];
            keep = keep && !(fabs(P.el_x * n0) < P.cut_plane * (P.el_e * sqrt((n0 * n0 + n1 * n1) + n2 * n2)));
        }
    }
    if (!CONVERGE && !keep) {
#pragma unroll
        for (int c = 0; c < 5; ++c)
            m[c] = (T)0;
#pragma unroll
        for (int j = 0; j < 12; ++j)
            du[j] = dp[j] = (T)0;
        for (int c = 0; c < NLAM; ++c)
            sink(c, m, du, dp);
        return false;
    }
    // ---- forward: spinor products (spinor.rs:31-54) on the sorted photons
    T xx[3];
    C<T> fx[3];
    bool small[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xx[k] = sq(ps[k][3] + ps[k][2]);
        small[k] = !(xx[k] > minpos<T>());
        fx[k] = small[k] ? C<T>{sq(two * ps[k][3]), (T)0} : C<T>{dv(ps[k][0], xx[k]), dv(ps[k][1], xx[k])};
    }
    SpinVal<T> V;
    const C<T> f0 = C<T>{P.f0re, P.f0im}, f1 = C<T>{P.f1re, P.f1im};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const C<T> s0k = C<T>{f0.re * xx[k] - fx[k].re * P.xx0, f0.im * xx[k] - fx[k].im * P.xx0};
        const C<T> s1k = C<T>{f1.re * xx[k] - fx[k].re * P.xx1, f1.im * xx[k] - fx[k].im * P.xx1};
        V.s[0][2 + k] = s0k;
        V.s[2 + k][0] = neg(s0k);
        V.s[1][2 + k] = s1k;
        V.s[2 + k][1] = neg(s1k);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = k + 1; l < 3; ++l) {
            const C<T> skl = C<T>{fx[k].re * xx[l] - fx[l].re * xx[k], fx[k].im * xx[l] - fx[l].im * xx[k]};
            V.s[2 + k][2 + l] = skl;
            V.s[2 + l][2 + k] = neg(skl);
        }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j)
            if (i != j && !(i < 2 && j < 2))
                V.is[i][j] = cinv(V.s[i][j]);
    // ---- forward: amplitudes (spinor.rs:72-162), without couplings
    Spin<T> S;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j)
            S.s[i][j] = V.s[i][j];
    // h = MMP, MPM, MPP, PMM, PMP, PPM
    const int A_pmm = 0, A_ppm = 1;
    const int aty[6] = {A_pmm, A_pmm, A_ppm, A_pmm, A_ppm, A_ppm};
    const int ak[6][3] = {{4, 2, 3}, {3, 2, 4}, {3, 4, 2}, {2, 3, 4}, {2, 4, 3}, {2, 3, 4}};
    // ns01 and nt01 are real (the beams run along x): the amplitudes without
    // couplings scale by their real parts
    const AmpK<T> K{P.ns01, P.nt01, P.g_a, P.g_bp, P.g_bm, P.ns01.re, P.nt01.re, P.ns01.re, P.nt01.re, P.ns01.re, P.nt01.re};
    C<T> a[6], bp[6];
#pragma unroll
    for (int h = 0; h < 6; ++h) {
        const int k1 = ak[h][0], k2 = ak[h][1], k3 = ak[h][2];
        if (aty[h] == A_pmm) {
            a[h] = a_pmm<T, true>(S, K, k1, k2, k3);
            bp[h] = bp_pmm<T, true>(S, K, k1, k2, k3);
        } else {
            a[h] = a_ppm<T, true>(S, K, k1, k2, k3);
            bp[h] = bp_ppm<T, true>(S, K, k1, k2, k3);
        }
    }
    const C<T> bmm = bm_mmm<T, true>(S, K), bpp = bm_ppp<T, true>(S, K);
    // helicity sums (matelems.rs:65-86)
    {
        T sA = 0, sP = 0, sR = 0, sI = 0;
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const C<T> Ag = cscale(P.g_a, a[h]), Pg = cscale(P.g_bp, bp[h]);
            const C<T> X = xmul<true>(cscale(two, Ag), cconj(Pg));
            sA = sA + (Ag.re * Ag.re + Ag.im * Ag.im);
            sP = sP + (Pg.re * Pg.re + Pg.im * Pg.im);
            sR = sR + X.re;
            sI = sI + X.im;
        }
        const C<T> Bm = cscale(P.g_bm, bmm), Bp = cscale(P.g_bm, bpp);
        m[0] = sA;
        m[1] = sP;
        m[2] = (Bm.re * Bm.re + Bm.im * Bm.im) + (Bp.re * Bp.re + Bp.im * Bp.im);
        m[3] = sR;
        m[4] = sI;
    }
#pragma unroll 1
    for (int sw = 0; sw < NLAM; ++sw) {
        T lam[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
            lam[k] = NLAM == 1 ? lam_in[k] : (k == sw ? one : (T)0);
        // SOFT 1: L = S f, f = sum_c lam_c m_c: the amplitude sweep below runs with
        // lam * S and f S dlogS/dp joins the momentum adjoint.  SOFT 2: L = log f +
        // log S: the sweep runs with lam / f and dlogS/dp joins.
        T lm[5];
        T fS = (T)0;
        {
            const T S = SOFT ? exp(lsig) : (T)1;
            T f = (T)0;
#pragma unroll
            for (int c = 0; c < 5; ++c)
                f = f + lam[c] * m[c];
            if (SOFT == 2 && !(f > (T)0 && f < (T)INFINITY)) {
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    m[c] = (T)0;
#pragma unroll
                for (int j = 0; j < 12; ++j)
                    du[j] = dp[j] = (T)0;
                return false;
            }
            const bool capped = SOFT == 2 && f > fcap;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                lm[c] = SOFT == 2 ? (capped ? (T)0 : lam[c] / f) : lam[c] * S;
                if (SOFT)
                    m[c] = capped ? m[c] * (fcap / f) * S : m[c] * S;
            }
            fS = SOFT == 2 ? (T)1 : f * S;
        }
        // ---- reverse: amplitudes -> spinor products
        SpinAdj<T> D;
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = 0; j < 5; ++j)
                D.sb[i][j] = D.tb[i][j] = C<T>{(T)0, (T)0};
        const T kappa = two * P.g_a * P.g_bp;
        const C<T> Xbar = C<T>{lm[3], lm[4]};
        const T cA = two * lm[0] * P.g_a * P.g_a, cP = two * lm[1] * P.g_bp * P.g_bp;
        const T cM = two * lm[2] * P.g_bm * P.g_bm;
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const int k1 = ak[h][0], k2 = ak[h][1], k3 = ak[h][2];
            const C<T> abar = caddv(cscale(cA, a[h]), cscale(kappa, xmul<true>(Xbar, bp[h])));
            const C<T> pbar = caddv(cscale(cP, bp[h]), cscale(kappa, xmul<true>(cconj(Xbar), a[h])));
            const C<T> ga = xmul<true>(abar, cconj(a[h])), gp = xmul<true>(pbar, cconj(bp[h]));
            if (aty[h] == A_ppm) {
                // a_ppm = ns01 s(0,k3)^2 / (s(0,k1) s(0,k2) s(1,k1) s(1,k2))
                add_s(D, V, 0, k3, two, ga);
                add_s(D, V, 0, k1, -one, ga);
                add_s(D, V, 0, k2, -one, ga);
                add_s(D, V, 1, k1, -one, ga);
                add_s(D, V, 1, k2, -one, ga);
                // bp_ppm = nt01 (t(k1,k2) s(k3,0))^2
                add_t(D, V, k1, k2, two, gp);
                add_s(D, V, k3, 0, two, gp);
            } else {
                // a_pmm = nt01 t(1,k1)^2 / (t(1,k2) t(1,k3) t(0,k2) t(0,k3))
                add_t(D, V, 1, k1, two, ga);
                add_t(D, V, 1, k2, -one, ga);
                add_t(D, V, 1, k3, -one, ga);
                add_t(D, V, 0, k2, -one, ga);
                add_t(D, V, 0, k3, -one, ga);
                // bp_pmm = ns01 (t(k1,1) s(k2,k3))^2
                add_t(D, V, k1, 1, two, gp);
                add_s(D, V, k2, k3, two, gp);
            }
        }
        {
            // bm_mmm = nt01 sum_j Y_j^2 with Y = s(k1,0)s(k2,k3), s(k2,0)s(k1,k3), s(k3,0)s(k1,k2)
            const C<T> bbar = cscale(cM, bmm);
            const C<T> Y1 = xmul<true>(V.s[2][0], V.s[3][4]), Y2 = xmul<true>(V.s[3][0], V.s[2][4]), Y3 = xmul<true>(V.s[4][0], V.s[2][3]);
            const C<T> d1 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.nt01, Y1))));
            const C<T> d2 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.nt01, Y2))));
            const C<T> d3 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.nt01, Y3))));
            D.sb[2][0] = caddv(D.sb[2][0], xmul<true>(d1, cconj(V.s[3][4])));
            D.sb[3][4] = caddv(D.sb[3][4], xmul<true>(d1, cconj(V.s[2][0])));
            D.sb[3][0] = caddv(D.sb[3][0], xmul<true>(d2, cconj(V.s[2][4])));
            D.sb[2][4] = caddv(D.sb[2][4], xmul<true>(d2, cconj(V.s[3][0])));
            D.sb[4][0] = caddv(D.sb[4][0], xmul<true>(d3, cconj(V.s[2][3])));
            D.sb[2][3] = caddv(D.sb[2][3], xmul<true>(d3, cconj(V.s[4][0])));
        }
        {
            // bm_ppp = ns01 sum_j Y_j^2 with Y = t(k1,k2)t(k3,1), t(k1,k3)t(k2,1), t(k2,k3)t(k1,1)
            const C<T> bbar = cscale(cM, bpp);
            const C<T> t23 = tval(V, 2, 3), t41 = tval(V, 4, 1), t24 = tval(V, 2, 4), t31 = tval(V, 3, 1);
            const C<T> t34 = tval(V, 3, 4), t21 = tval(V, 2, 1);
            const C<T> Y1 = xmul<true>(t23, t41), Y2 = xmul<true>(t24, t31), Y3 = xmul<true>(t34, t21);
            const C<T> d1 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.ns01, Y1))));
            const C<T> d2 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.ns01, Y2))));
            const C<T> d3 = xmul<true>(bbar, cconj(cscale(two, xmul<true>(P.ns01, Y3))));
            D.tb[2][3] = caddv(D.tb[2][3], xmul<true>(d1, cconj(t41)));
            D.tb[4][1] = caddv(D.tb[4][1], xmul<true>(d1, cconj(t23)));
            D.tb[2][4] = caddv(D.tb[2][4], xmul<true>(d2, cconj(t31)));
            D.tb[3][1] = caddv(D.tb[3][1], xmul<true>(d2, cconj(t24)));
            D.tb[3][4] = caddv(D.tb[3][4], xmul<true>(d3, cconj(t21)));
            D.tb[2][1] = caddv(D.tb[2][1], xmul<true>(d3, cconj(t34)));
        }
        // fold onto the independent products: s(j,i) = -s(i,j), t(i,j) = -conj(s(i,j)),
        // t(j,i) = conj(s(i,j))
        C<T> g0[3], g1[3], gkl[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int K = 2 + k;
            g0[k] = caddv(caddv(D.sb[0][K], neg(D.sb[K][0])), caddv(neg(cconj(D.tb[0][K])), cconj(D.tb[K][0])));
            g1[k] = caddv(caddv(D.sb[1][K], neg(D.sb[K][1])), caddv(neg(cconj(D.tb[1][K])), cconj(D.tb[K][1])));
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = k + 1; l < 3; ++l) {
                const int K = 2 + k, L = 2 + l;
                gkl[k][l] =
                    caddv(caddv(D.sb[K][L], neg(D.sb[L][K])), caddv(neg(cconj(D.tb[K][L])), cconj(D.tb[L][K])));
            }
        // ---- reverse: spinor products -> sorted photon momenta
        T xxb[3] = {(T)0, (T)0, (T)0};
        C<T> fxb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // s0k = f0 xx_k - fx_k xx0 ; s1k = f1 xx_k - fx_k xx1
            xxb[k] = xxb[k] + cdot(g0[k], f0) + cdot(g1[k], f1);
            fxb[k] = caddv(cscale(-P.xx0, g0[k]), cscale(-P.xx1, g1[k]));
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = k + 1; l < 3; ++l) {
                // skl = fx_k xx_l - fx_l xx_k
                fxb[k] = caddv(fxb[k], cscale(xx[l], gkl[k][l]));
                fxb[l] = caddv(fxb[l], cscale(-xx[k], gkl[k][l]));
                xxb[l] = xxb[l] + cdot(gkl[k][l], fx[k]);
                xxb[k] = xxb[k] - cdot(gkl[k][l], fx[l]);
            }
        T pb[3][4]; // adjoint of the sorted momenta
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pb[k][0] = pb[k][1] = pb[k][2] = pb[k][3] = (T)0;
            if (small[k]) {
                pb[k][3] = pb[k][3] + fxb[k].re / fx[k].re;
            } else {
                pb[k][0] = fxb[k].re / xx[k];
                pb[k][1] = fxb[k].im / xx[k];
                xxb[k] = xxb[k] - (fxb[k].re * fx[k].re + fxb[k].im * fx[k].im) / xx[k];
            }
            // sqrt(E + Z) has no derivative at a photon exactly along -Z
            const T h = small[k] ? (T)0 : xxb[k] / (two * xx[k]);
            pb[k][3] = pb[k][3] + h;
            pb[k][2] = pb[k][2] + h;
        }
        if (SOFT) { // fS * dlogS/dp of the sorted momenta
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const T E = ps[i][3];
                pb[i][3] = pb[i][3] + fS * gE[i] / P.e_min;
                pb[i][0] = pb[i][0] - fS * gB[i] * ((ps[i][0] >= (T)0 ? one : -one) / E);
                pb[i][3] = pb[i][3] + fS * gB[i] * (fabs(ps[i][0]) / (E * E));
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int i = k == 2 ? 1 : 0, j = k == 0 ? 1 : 2;
                const T EE = ps[i][3] * ps[j][3];
                const T cs = ((ps[i][0] * ps[j][0] + ps[i][1] * ps[j][1]) + ps[i][2] * ps[j][2]) / EE;
                const T g = fS * gP[k];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T pic = ps[i][c], pjc = ps[j][c];
                    pb[i][c] = pb[i][c] - g * pjc / EE;
                    pb[j][c] = pb[j][c] - g * pic / EE;
                }
                pb[i][3] = pb[i][3] + g * cs / ps[i][3];
                pb[j][3] = pb[j][3] + g * cs / ps[j][3];
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                dp[4 * k + j] = P.w * pb[k][j];
        // ---- reverse: sort
        T pbu[3][4];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                pbu[i][j] = perm[0] == i ? pb[0][j] : (perm[1] == i ? pb[1][j] : pb[2][j]);
        // ---- reverse: conformal map
        T qb[3][4], Rb[4] = {(T)0, (T)0, (T)0, (T)0};
        T alb = (T)0, rnb = (T)0, beb = (T)0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T bb = (T)0, rqb = (T)0;
            qb[i][0] = qb[i][1] = qb[i][2] = qb[i][3] = (T)0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                alb = alb + pbu[i][k] * (rn * q[i][k] + b[i] * R[k]);
                const T t = alpha * pbu[i][k];
                rnb = rnb + t * q[i][k];
                qb[i][k] = qb[i][k] + t * rn;
                bb = bb + t * R[k];
                Rb[k] = Rb[k] + t * b[i];
            }
            alb = alb + pbu[i][3] * (R[3] * q[i][3] - rq[i]);
            const T t = alpha * pbu[i][3];
            Rb[3] = Rb[3] + t * q[i][3];
            qb[i][3] = qb[i][3] + t * R[3];
            rqb = rqb - t;
            beb = beb + bb * rq[i];
            rqb = rqb + beta * bb;
            qb[i][3] = qb[i][3] - bb;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                qb[i][k] = qb[i][k] + rqb * R[k];
                Rb[k] = Rb[k] + rqb * q[i][k];
            }
        }
        const T db = -beb * beta * beta;
        rnb = rnb + db;
        Rb[3] = Rb[3] + db;
        T rn2b = rnb / (two * rn);
        rn2b = rn2b - alb * alpha / rn2;
        Rb[3] = Rb[3] + two * R[3] * rn2b;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Rb[k] = Rb[k] - two * R[k] * rn2b;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k)
                qb[i][k] = qb[i][k] + Rb[k];
        // ---- reverse: generation
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T enb = qb[i][0] * (st[i] * sp[i]) + qb[i][1] * (st[i] * cp[i]) + qb[i][2] * ct[i] + qb[i][3];
            const T stb = en[i] * (qb[i][0] * sp[i] + qb[i][1] * cp[i]);
            const T spb = en[i] * st[i] * qb[i][0];
            const T cpb = en[i] * st[i] * qb[i][1];
            T ctb = en[i] * qb[i][2];
            if (st[i] > (T)0)
                ctb = ctb - stb * ct[i] / st[i];
            const T phb = spb * cp[i] - cpb * sp[i];
            const T emb = -enb / (em[i] + minpos<T>());
            du[4 * i] = P.w * (two * ctb);
            du[4 * i + 1] = P.w * (two_pi * phb);
            du[4 * i + 2] = P.w * (emb * u[4 * i + 3]);
            du[4 * i + 3] = P.w * (emb * u[4 * i + 2]);
        }
        if constexpr (CONVERGE) {
            // du and dp are rewritten by every sweep, so they are zeroed in
            // place; m feeds the later sweeps and goes out as a copy
            T mz[5];
#pragma unroll
            for (int c = 0; c < 5; ++c)
                mz[c] = keep ? m[c] : (T)0;
#pragma unroll
            for (int j = 0; j < 12; ++j) {
                du[j] = keep ? du[j] : (T)0;
                dp[j] = keep ? dp[j] : (T)0;
            }
            sink(sw, mz, du, dp);
        } else {
            sink(sw, m, du, dp);
        }
    }
    return keep;
}

template <class T, int SOFT = 0>
__device__ __forceinline__ bool event_adjoint(const AdjParams<T> &P, const T u[12], const T lam[5], T m[5], T du[12],
                                              T dp[12], T tau = (T)0, T fcap = (T)INFINITY)
{
    return event_adjoint_sweeps<T, SOFT, 1>(P, u, lam, m, du, dp, tau, fcap, [](int, const T *, const T *, const T *) {});
}

} // namespace tpd
