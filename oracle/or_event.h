/*
 * or_event.h -- per-event physics of the 3photons event loop, restated on the
 * CPU as the ORACLE of this repository.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C ABI, the CLI) includes, links or calls this file.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it, as the checker.
 *
 * The file is plain C99 and also compiles as C++ so that oracle/opcount.cpp can
 * instantiate it with an operation-counting scalar type.  The scalar type is
 * OR_REAL (double by default, float for the all-single-precision mode); the
 * transcendental functions are the host libm ones (glibc), exactly as the
 * reference binary calls them through Rust's std.
 *
 * Every operation follows the reference's evaluation order: no FMA contraction
 * (build with -ffp-contract=off), no reassociation, left-to-right sums starting
 * from zero where nalgebra folds from zero.  Citations are to the reference
 * sources (HadrienG2/3photons-rust, src/...).
 */
#ifndef OR_EVENT_H
#define OR_EVENT_H

#include <float.h>
#include <math.h>

#ifndef OR_REAL
#define OR_REAL double
#endif

#ifndef OR_SQRT
#if defined(OR_F32)
#define OR_SQRT sqrtf
#define OR_SIN sinf
#define OR_COS cosf
#define OR_LOG logf
#define OR_FABS fabsf
#else
#define OR_SQRT sqrt
#define OR_SIN sin
#define OR_COS cos
#define OR_LOG log
#define OR_FABS fabs
#endif
#endif

#if defined(OR_F32)
/* f32::MIN_POSITIVE, f32 PI and SQRT_2 (numeric.rs:6-14 selects f32). */
#define OR_MIN_POSITIVE FLT_MIN
#define OR_PI 3.14159265358979323846f
#define OR_SQRT_2 1.41421356237309504880f
#else
#define OR_MIN_POSITIVE DBL_MIN
#define OR_PI 3.14159265358979323846
#define OR_SQRT_2 1.41421356237309504880
#endif

typedef OR_REAL or_real;

/* Particle rows (event.rs:10-51) and 4-momentum columns (momentum.rs:12-22). */
enum { OR_EM = 0, OR_EP = 1, OR_X = 0, OR_Y = 1, OR_Z = 2, OR_E = 3 };

/* Everything the per-event code reads besides the event itself.  Filled by
 * tp_oracle.c from the configuration exactly as coupling.rs:23-33 and
 * resacc.rs:59-100 compute it. */
typedef struct {
    or_real e_total;
    or_real cut_beam, cut_pp, e_min, cut_plane; /* evcut.rs:11-23 */
    or_real g_a, g_bp, g_bm;                   /* coupling.rs:10-20 */
    or_real sc[5];                             /* resacc.rs:94-100 sigma_contribs */
    int sort;                                  /* 0 = no-photon-sorting feature */
} or_phys;

/* ---- num-complex 0.4 arithmetic (naive division, powi(2) == z*z) -------- */
typedef struct {
    or_real re, im;
} or_cplx;

static inline or_cplx or_c(or_real re, or_real im)
{
    or_cplx z;
    z.re = re;
    z.im = im;
    return z;
}
static inline or_cplx or_cmul(or_cplx a, or_cplx b)
{
    return or_c(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline or_cplx or_cdiv(or_cplx a, or_cplx b)
{
    or_real n = b.re * b.re + b.im * b.im;
    or_real re = a.re * b.re + a.im * b.im;
    or_real im = a.im * b.re - a.re * b.im;
    return or_c(re / n, im / n);
}
static inline or_cplx or_cadd(or_cplx a, or_cplx b) { return or_c(a.re + b.re, a.im + b.im); }
static inline or_cplx or_csub(or_cplx a, or_cplx b) { return or_c(a.re - b.re, a.im - b.im); }
static inline or_cplx or_crmul(or_real r, or_cplx z) { return or_c(r * z.re, r * z.im); }
static inline or_cplx or_csq(or_cplx z) { return or_cmul(z, z); }
static inline or_cplx or_cconj(or_cplx z) { return or_c(z.re, -z.im); }
static inline or_real or_cnorm2(or_cplx z) { return z.re * z.re + z.im * z.im; }

/* ---- RAMBO, default branch: evgen.rs:174-207 then evgen.rs:89-131 -------
 * u[12] are the twelve uniforms in stream order; photon i consumes
 * (cos theta, phi, e1, e2) = u[4i..4i+3] because nalgebra's from_fn walks the
 * 3x3 parameter matrix column-major and `rng.random() * rng.random()`
 * evaluates its left operand first (evgen.rs:182-187).
 * p receives the 5x4 event matrix (e-, e+, photons sorted by decreasing E). */
static inline void or_generate(const or_phys *ph, const or_real u[12], or_real p[5][4])
{
    or_real q[3][4];
    int i, k;
    for (i = 0; i < 3; ++i) {
        or_real cos_t = (or_real)2. * u[4 * i] - (or_real)1.;
        or_real phi = ((or_real)2. * OR_PI) * u[4 * i + 1];
        or_real em = u[4 * i + 2] * u[4 * i + 3];
        or_real cphi = OR_COS(phi);
        or_real sphi = OR_SIN(phi);
        or_real st = OR_SQRT((or_real)1. - cos_t * cos_t);
        or_real en = -OR_LOG(em + OR_MIN_POSITIVE);
        q[i][OR_X] = en * (st * sphi);
        q[i][OR_Y] = en * (st * cphi);
        q[i][OR_Z] = en * cos_t;
        q[i][OR_E] = en * (or_real)1.;
    }
    /* column_sum folds from a zero vector (evgen.rs:94). */
    or_real r[4];
    for (k = 0; k < 4; ++k)
        r[k] = (((or_real)0. + q[0][k]) + q[1][k]) + q[2][k];
    /* norm_squared = 0 + ((x*x + y*y) + z*z) (evgen.rs:95). */
    or_real rn2 = r[OR_E] * r[OR_E] -
                  ((or_real)0. + ((r[OR_X] * r[OR_X] + r[OR_Y] * r[OR_Y]) + r[OR_Z] * r[OR_Z]));
    or_real alpha = ph->e_total / rn2;
    or_real rn = OR_SQRT(rn2);
    or_real beta = (or_real)1. / (rn + r[OR_E]);
    or_real pe[3], pxyz[3][3];
    for (i = 0; i < 3; ++i) {
        /* gemv as a sequence of axpy over columns (evgen.rs:103). */
        or_real rq = q[i][OR_X] * r[OR_X];
        rq = q[i][OR_Y] * r[OR_Y] + rq;
        rq = q[i][OR_Z] * r[OR_Z] + rq;
        pe[i] = alpha * (r[OR_E] * q[i][OR_E] - rq);
        or_real b = beta * rq - q[i][OR_E];
        for (k = 0; k < 3; ++k)
            pxyz[i][k] = alpha * (rn * q[i][k] + b * r[k]);
    }
    /* Sort by decreasing energy with strict '>' swaps (evgen.rs:109-118). */
    if (ph->sort) {
        int a, c;
        for (a = 0; a < 2; ++a)
            for (c = a + 1; c < 3; ++c)
                if (pe[c] > pe[a]) {
                    or_real t = pe[a];
                    pe[a] = pe[c];
                    pe[c] = t;
                    for (k = 0; k < 3; ++k) {
                        t = pxyz[a][k];
                        pxyz[a][k] = pxyz[c][k];
                        pxyz[c][k] = t;
                    }
                }
    }
    /* Incoming momenta, beam along X (evgen.rs:66-69). */
    or_real half = ph->e_total / (or_real)2.;
    p[OR_EM][OR_X] = -ph->e_total / (or_real)2.;
    p[OR_EM][OR_Y] = 0.;
    p[OR_EM][OR_Z] = 0.;
    p[OR_EM][OR_E] = half;
    p[OR_EP][OR_X] = half;
    p[OR_EP][OR_Y] = 0.;
    p[OR_EP][OR_Z] = 0.;
    p[OR_EP][OR_E] = half;
    for (i = 0; i < 3; ++i) {
        for (k = 0; k < 3; ++k)
            p[2 + i][k] = pxyz[i][k];
        p[2 + i][OR_E] = pe[i];
    }
}

/* ---- Cuts: evcut.rs:42-96 ----------------------------------------------- */
static inline int or_keep(const or_phys *ph, const or_real p[5][4])
{
    int i, j;
    /* min_photon_energy (event.rs:95-110). */
    or_real emin;
    if (ph->sort) {
        emin = p[4][OR_E];
    } else {
        emin = p[2][OR_E];
        for (i = 3; i < 5; ++i)
            emin = (emin < p[i][OR_E]) ? emin : p[i][OR_E];
    }
    if (emin < ph->e_min)
        return 0;
    const or_real *el = p[OR_EM];
    /* (beam, photon) angles: gemv then |num| > cut * (E_i * E_el). */
    for (i = 0; i < 3; ++i) {
        const or_real *q = p[2 + i];
        or_real num = q[OR_X] * el[OR_X];
        num = q[OR_Y] * el[OR_Y] + num;
        num = q[OR_Z] * el[OR_Z] + num;
        or_real den = q[OR_E] * el[OR_E];
        if (OR_FABS(num) > ph->cut_beam * den)
            return 0;
    }
    /* (photon, photon) angles: 3-vector dot = (a + b) + c. */
    for (i = 0; i < 2; ++i)
        for (j = i + 1; j < 3; ++j) {
            const or_real *a = p[2 + i], *b = p[2 + j];
            or_real num = (a[OR_X] * b[OR_X] + a[OR_Y] * b[OR_Y]) + a[OR_Z] * b[OR_Z];
            or_real den = a[OR_E] * b[OR_E];
            if (num > ph->cut_pp * den)
                return 0;
        }
    /* (beam, normal to the photon plane) angle (evcut.rs:80-92). */
    const or_real *a = p[2], *b = p[3];
    or_real n[3];
    n[0] = a[OR_Y] * b[OR_Z] - a[OR_Z] * b[OR_Y];
    n[1] = a[OR_Z] * b[OR_X] - a[OR_X] * b[OR_Z];
    n[2] = a[OR_X] * b[OR_Y] - a[OR_Y] * b[OR_X];
    or_real num = (el[OR_X] * n[0] + el[OR_Y] * n[1]) + el[OR_Z] * n[2];
    or_real nn = OR_SQRT((or_real)0. + ((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]));
    or_real den = el[OR_E] * nn;
    if (OR_FABS(num) < ph->cut_plane * den)
        return 0;
    return 1;
}

/* ---- Spinor products: spinor.rs:31-54 (full 5x5 Gram matrix) ------------ */
typedef struct {
    or_cplx s[5][5];
} or_spinor;

static inline void or_spinor_new(or_spinor *sp, const or_real p[5][4])
{
    or_real xx[5];
    or_cplx fx[5];
    int i, j;
    for (i = 0; i < 5; ++i) {
        xx[i] = OR_SQRT(p[i][OR_E] + p[i][OR_Z]);
        if (xx[i] > OR_MIN_POSITIVE)
            fx[i] = or_c(p[i][OR_X] / xx[i], p[i][OR_Y] / xx[i]);
        else
            fx[i] = or_c(OR_SQRT((or_real)2. * p[i][OR_E]), (or_real)0.);
    }
    for (i = 0; i < 5; ++i)
        for (j = 0; j < 5; ++j) {
            or_cplx a = or_c(fx[i].re * xx[j], fx[i].im * xx[j]);
            or_cplx b = or_c(fx[j].re * xx[i], fx[j].im * xx[i]);
            sp->s[i][j] = or_csub(a, b);
        }
}
#define OR_S(sp, i, j) ((sp)->s[i][j])
/* t(i, j) = -conj(s(i, j)) (spinor.rs:63-66) */
static inline or_cplx or_t(const or_spinor *sp, int i, int j)
{
    or_cplx c = or_cconj(sp->s[i][j]);
    return or_c(-c.re, -c.im);
}

/* Amplitudes, spinor.rs:120-161.  -RAC8 * x * y / z parses as ((-RAC8*x)*y)/z. */
#define OR_NRAC8 (-((or_real)2. * OR_SQRT_2))
static inline or_cplx or_a_ppm(const or_spinor *sp, int k1, int k2, int k3)
{
    or_cplx num = or_cmul(or_crmul(OR_NRAC8, OR_S(sp, OR_EM, OR_EP)), or_csq(OR_S(sp, OR_EM, k3)));
    or_cplx den = or_cmul(or_cmul(or_cmul(OR_S(sp, OR_EM, k1), OR_S(sp, OR_EM, k2)), OR_S(sp, OR_EP, k1)),
                          OR_S(sp, OR_EP, k2));
    return or_cdiv(num, den);
}
static inline or_cplx or_a_pmm(const or_spinor *sp, int k1, int k2, int k3)
{
    or_cplx num = or_cmul(or_crmul(OR_NRAC8, or_t(sp, OR_EM, OR_EP)), or_csq(or_t(sp, OR_EP, k1)));
    or_cplx den = or_cmul(or_cmul(or_cmul(or_t(sp, OR_EP, k2), or_t(sp, OR_EP, k3)), or_t(sp, OR_EM, k2)),
                          or_t(sp, OR_EM, k3));
    return or_cdiv(num, den);
}
static inline or_cplx or_bp_ppm(const or_spinor *sp, int k1, int k2, int k3)
{
    return or_cmul(or_crmul(OR_NRAC8, or_t(sp, OR_EM, OR_EP)), or_csq(or_cmul(or_t(sp, k1, k2), OR_S(sp, k3, OR_EM))));
}
static inline or_cplx or_bp_pmm(const or_spinor *sp, int k1, int k2, int k3)
{
    return or_cmul(or_crmul(OR_NRAC8, OR_S(sp, OR_EM, OR_EP)), or_csq(or_cmul(or_t(sp, k1, OR_EP), OR_S(sp, k2, k3))));
}
static inline or_cplx or_bm_ppp(const or_spinor *sp, int k1, int k2, int k3)
{
    or_cplx s = or_cadd(or_cadd(or_csq(or_cmul(or_t(sp, k1, k2), or_t(sp, k3, OR_EP))),
                                or_csq(or_cmul(or_t(sp, k1, k3), or_t(sp, k2, OR_EP)))),
                        or_csq(or_cmul(or_t(sp, k2, k3), or_t(sp, k1, OR_EP))));
    return or_cmul(or_crmul(OR_NRAC8, OR_S(sp, OR_EM, OR_EP)), s);
}
static inline or_cplx or_bm_mmm(const or_spinor *sp, int k1, int k2, int k3)
{
    or_cplx s = or_cadd(or_cadd(or_csq(or_cmul(OR_S(sp, k1, OR_EM), OR_S(sp, k2, k3))),
                                or_csq(or_cmul(OR_S(sp, k2, OR_EM), OR_S(sp, k1, k3)))),
                        or_csq(or_cmul(OR_S(sp, k3, OR_EM), OR_S(sp, k1, k2))));
    return or_cmul(or_crmul(OR_NRAC8, or_t(sp, OR_EM, OR_EP)), s);
}

/* Helicity amplitudes h = MMM, MMP, MPM, MPP, PMM, PMP, PPM, PPP
 * (spinor.rs:168-180; dispatch tables at spinor.rs:72-116). */
static inline void or_amplitudes(const or_spinor *sp, or_cplx a[8], or_cplx bp[8], or_cplx bm[8])
{
    or_cplx z = or_c(0., 0.);
    int h;
    for (h = 0; h < 8; ++h)
        a[h] = bp[h] = bm[h] = z;
    a[1] = or_a_pmm(sp, 4, 2, 3);
    a[2] = or_a_pmm(sp, 3, 2, 4);
    a[3] = or_a_ppm(sp, 3, 4, 2);
    a[4] = or_a_pmm(sp, 2, 3, 4);
    a[5] = or_a_ppm(sp, 2, 4, 3);
    a[6] = or_a_ppm(sp, 2, 3, 4);
    bp[1] = or_bp_pmm(sp, 4, 2, 3);
    bp[2] = or_bp_pmm(sp, 3, 2, 4);
    bp[3] = or_bp_ppm(sp, 3, 4, 2);
    bp[4] = or_bp_pmm(sp, 2, 3, 4);
    bp[5] = or_bp_ppm(sp, 2, 4, 3);
    bp[6] = or_bp_ppm(sp, 2, 3, 4);
    bm[0] = or_bm_mmm(sp, 2, 3, 4);
    bm[7] = or_bm_ppp(sp, 2, 3, 4);
}

/* Matrix elements per helicity (matelems.rs:53-81): rows A, B_P, B_M, R_MX,
 * I_MX; zero amplitudes are multiplied by the couplings like the reference. */
static inline void or_m2(const or_phys *ph, const or_real p[5][4], or_real m2[5][8])
{
    or_spinor sp;
    or_cplx a[8], bp[8], bm[8];
    int h;
    or_spinor_new(&sp, p);
    or_amplitudes(&sp, a, bp, bm);
    for (h = 0; h < 8; ++h) {
        or_cplx aa = or_c(a[h].re * ph->g_a, a[h].im * ph->g_a);
        or_cplx pp = or_c(bp[h].re * ph->g_bp, bp[h].im * ph->g_bp);
        or_cplx mm = or_c(bm[h].re * ph->g_bm, bm[h].im * ph->g_bm);
        or_cplx mx = or_cmul(or_crmul((or_real)2., aa), or_cconj(pp));
        m2[0][h] = or_cnorm2(aa);
        m2[1][h] = or_cnorm2(pp);
        m2[2][h] = or_cnorm2(mm);
        m2[3][h] = mx.re;
        m2[4][h] = mx.im;
    }
}

/* m2_sums: column_sum over the 8 helicities from zero (matelems.rs:84-86). */
static inline void or_m2_sums(const or_real m2[5][8], or_real m[5])
{
    int c, h;
    for (c = 0; c < 5; ++c) {
        or_real s = 0.;
        for (h = 0; h < 8; ++h)
            s = s + m2[c][h];
        m[c] = s;
    }
}

/* Per-event weight: 5-vector dot through nalgebra's generic path, i.e. a
 * left fold from zero (resacc.rs:126). */
static inline or_real or_weight(const or_phys *ph, const or_real m[5])
{
    or_real w = 0.;
    int c;
    for (c = 0; c < 5; ++c)
        w = w + m[c] * ph->sc[c];
    return w;
}

#endif /* OR_EVENT_H */
