/*
 * tp_oracle.c -- CPU oracle for the 3photons event loop (random streams,
 * batch driver, accumulation, finalisation).  The per-event physics is in
 * or_event.h.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/, the smoke check and the CPU
 * baseline of bench.py through ctypes; never part of the product path.
 *
 * Build: oracle/Makefile (-O2 -ffp-contract=off, no fast-math, glibc libm).
 * The same file builds the all-f32 oracle with -DOR_F32.
 */
#define _GNU_SOURCE
#include "tp_oracle.h"

#if defined(OR_F32)
#define OR_REAL float
#endif
#include "or_event.h"

#include <ctype.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* Random streams                                                            */
/* ======================================================================== */

#define RANF_MODULO 1000000000
#define RANF_DEFAULT_SEED 234612947 /* ranf.rs:28-32 */
#define XOSHIRO_DEFAULT_SEED 12345u /* standard.rs:21-24 */

/* IRN55 round (ranf.rs:106-119). */
static void ranf_reset(or_rng_t *g)
{
    int i;
    for (i = 1; i < 25; ++i) {
        g->numbers[i] -= g->numbers[i + 31];
        if (g->numbers[i] < 0)
            g->numbers[i] += RANF_MODULO;
    }
    for (i = 25; i < 56; ++i) {
        g->numbers[i] -= g->numbers[i - 24];
        if (g->numbers[i] < 0)
            g->numbers[i] += RANF_MODULO;
    }
}

/* IN55 initialisation + 10 warm-up rounds (ranf.rs:36-66).  Rust release
 * builds wrap on i32 overflow; the arithmetic is done in uint32 to mirror that
 * for the (unused by the default configuration) negative-seed cases. */
static void ranf_seeded(or_rng_t *g, int32_t seed)
{
    int i;
    memset(g->numbers, 0, sizeof g->numbers);
    g->index = 55;
    g->ranf_seed = seed;
    g->numbers[55] = seed;
    int32_t j = seed, k = 1;
    for (i = 1; i < 55; ++i) {
        int ii = (21 * i) % 55;
        g->numbers[ii] = k;
        k = (int32_t)((uint32_t)j - (uint32_t)k);
        if (k < 0)
            k = (int32_t)((uint32_t)k + (uint32_t)RANF_MODULO);
        j = g->numbers[ii];
    }
    for (i = 0; i < 10; ++i)
        ranf_reset(g);
}

/* random_array::<N> (ranf.rs:78-102): N values taken in ascending index order
 * from the top of the unconsumed part of the round; a round with fewer than N
 * values left is discarded. */
static void ranf_array(or_rng_t *g, int n, or_scalar *out)
{
    int i;
    if ((int)g->index < n) {
        ranf_reset(g);
        g->index = 55;
    }
    g->index -= (uint32_t)n;
    for (i = 0; i < n; ++i)
        out[i] = (or_scalar)g->numbers[g->index + 1 + i] * (or_scalar)1e-9;
}

static int32_t ranf_next_int(or_rng_t *g)
{
    if (g->index < 1) {
        ranf_reset(g);
        g->index = 55;
    }
    g->index -= 1;
    return g->numbers[g->index + 1];
}

/* SplitMix64 (rand_xoshiro 0.6.0 seed_from_u64 expands the seed with it). */
static uint64_t splitmix64_next(uint64_t *x)
{
    *x += 0x9e3779b97f4a7c15ull;
    uint64_t z = *x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* xoshiro256+ next_u64: output s0 + s3 before the state update. */
static uint64_t xoshiro256p_next(uint64_t *s)
{
    uint64_t result = s[0] + s[3];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* xoshiro128+ next_u32. */
static uint32_t xoshiro128p_next(uint32_t *s)
{
    uint32_t result = s[0] + s[3];
    uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    return result;
}

/* Published 2^128 / 2^64 jump polynomials of xoshiro256+ / xoshiro128+. */
static void xoshiro256p_jump(uint64_t *s)
{
    static const uint64_t J[4] = {0x180ec6d33cfd0abaull, 0xd5a61266f0c9392cull, 0xa9582618e03fc9aaull,
                                  0x39abdc4529b1661cull};
    uint64_t a[4] = {0, 0, 0, 0};
    int i, b;
    for (i = 0; i < 4; ++i)
        for (b = 0; b < 64; ++b) {
            if (J[i] & (1ull << b)) {
                a[0] ^= s[0];
                a[1] ^= s[1];
                a[2] ^= s[2];
                a[3] ^= s[3];
            }
            xoshiro256p_next(s);
        }
    memcpy(s, a, sizeof a);
}

static void xoshiro128p_jump(uint32_t *s)
{
    static const uint32_t J[4] = {0x8764000bu, 0xf542d2d3u, 0x6fa035c3u, 0x77f2db5bu};
    uint32_t a[4] = {0, 0, 0, 0};
    int i, b;
    for (i = 0; i < 4; ++i)
        for (b = 0; b < 32; ++b) {
            if (J[i] & (1u << b)) {
                a[0] ^= s[0];
                a[1] ^= s[1];
                a[2] ^= s[2];
                a[3] ^= s[3];
            }
            xoshiro128p_next(s);
        }
    memcpy(s, a, sizeof a);
}

void OR_FN(or_rng_seeded)(int32_t kind, uint64_t seed, or_rng_t *g)
{
    memset(g, 0, sizeof *g);
    g->kind = kind;
    if (kind == OR_RNG_RANF) {
        ranf_seeded(g, (int32_t)seed);
    } else {
        /* seed_from_u64: SplitMix64 output bytes, little-endian, fill the
         * seed array; read back as u64 (256+) or u32 (128+) words. */
        uint64_t x = seed;
        int i;
        for (i = 0; i < 4; ++i)
            g->xs[i] = splitmix64_next(&x);
        uint64_t y = seed;
        uint64_t a = splitmix64_next(&y), b = splitmix64_next(&y);
        g->xs32[0] = (uint32_t)a;
        g->xs32[1] = (uint32_t)(a >> 32);
        g->xs32[2] = (uint32_t)b;
        g->xs32[3] = (uint32_t)(b >> 32);
    }
}

void OR_FN(or_rng_new)(int32_t kind, or_rng_t *g)
{
    OR_FN(or_rng_seeded)(kind, kind == OR_RNG_RANF ? (uint64_t)RANF_DEFAULT_SEED : XOSHIRO_DEFAULT_SEED, g);
}

uint64_t OR_FN(or_rng_next_raw)(or_rng_t *g)
{
    if (g->kind == OR_RNG_RANF)
        return (uint64_t)(uint32_t)ranf_next_int(g);
#if defined(OR_F32)
    return xoshiro128p_next(g->xs32);
#else
    return xoshiro256p_next(g->xs);
#endif
}

/* One uniform: RANF int * 1e-9 (ranf.rs:99), or rand's Standard
 * distribution: (u64 >> 11) * 2^-53 for f64, (u32 >> 8) * 2^-24 for f32. */
or_scalar OR_FN(or_rng_random)(or_rng_t *g)
{
    if (g->kind == OR_RNG_RANF)
        return (or_scalar)ranf_next_int(g) * (or_scalar)1e-9;
#if defined(OR_F32)
    return (float)(xoshiro128p_next(g->xs32) >> 8) * 0x1.0p-24f;
#else
    return (double)(xoshiro256p_next(g->xs) >> 11) * 0x1.0p-53;
#endif
}

static void rng_array(or_rng_t *g, int n, or_scalar *out)
{
    int i;
    if (g->kind == OR_RNG_RANF) {
        ranf_array(g, n, out);
        return;
    }
    for (i = 0; i < n; ++i)
        out[i] = OR_FN(or_rng_random)(g);
}

void OR_FN(or_rng_fill)(or_rng_t *g, uint64_t n, or_scalar *out)
{
    uint64_t i;
    for (i = 0; i < n; ++i)
        out[i] = OR_FN(or_rng_random)(g);
}

void OR_FN(or_rng_skip)(or_rng_t *g, uint64_t n)
{
    uint64_t i;
    for (i = 0; i < n; ++i)
        (void)OR_FN(or_rng_next_raw)(g);
}

/* jump(): RANF reseeds with seed + 123456 (ranf.rs:135-140, i32 wrapping);
 * xoshiro applies its published jump polynomial (standard.rs:52-55). */
void OR_FN(or_rng_jump)(or_rng_t *g)
{
    if (g->kind == OR_RNG_RANF) {
        int32_t s = (int32_t)((uint32_t)g->ranf_seed + 123456u);
        ranf_seeded(g, s);
        return;
    }
#if defined(OR_F32)
    xoshiro128p_jump(g->xs32);
#else
    xoshiro256p_jump(g->xs);
#endif
}

/* ======================================================================== */
/* Configuration and constants                                               */
/* ======================================================================== */

static int set_err(char *err, int len, int code, const char *msg, const char *what)
{
    if (err && len > 0)
        snprintf(err, (size_t)len, "%s%s", msg, what ? what : "");
    return code;
}

/* First whitespace-delimited token of each non-blank line (config.rs:69-71). */
int OR_FN(or_config_load)(const char *path, or_config_t *cfg, char *err, int err_len)
{
    static const char *names[18] = {"num_events", "e_total", "beam_photons_cut", "photon_photon_cut",
                                    "e_min", "beam_photon_plane_cut", "alpha", "alpha_z",
                                    "gev2_to_picobarn", "m_z0", "g_z0", "sin2_weinberg",
                                    "branching_ep_em", "beta_plus", "beta_moins", "num_bins",
                                    "impr", "plot"};
    char tok[18][128];
    char line[4096];
    int n = 0;
    FILE *f = fopen(path, "r");
    if (!f)
        return set_err(err, err_len, -1, "cannot open ", path);
    while (n < 18 && fgets(line, sizeof line, f)) {
        char *s = line;
        while (*s && isspace((unsigned char)*s))
            ++s;
        if (!*s)
            continue;
        int l = 0;
        while (s[l] && !isspace((unsigned char)s[l]) && l < 127)
            ++l;
        memcpy(tok[n], s, (size_t)l);
        tok[n][l] = 0;
        ++n;
    }
    fclose(f);
    if (n < 18)
        return set_err(err, err_len, -2, "missing configuration of ", names[n]);
    double *dst[14] = {&cfg->e_total, &cfg->cut_beam, &cfg->cut_pp, &cfg->e_min, &cfg->cut_plane,
                       &cfg->alpha, &cfg->alpha_z, &cfg->conv, &cfg->m_z0, &cfg->g_z0,
                       &cfg->sin2w, &cfg->br_ep_em, &cfg->beta_plus, &cfg->beta_minus};
    char *end;
    unsigned long long ne = strtoull(tok[0], &end, 10);
    if (*end || tok[0][0] == '-')
        return set_err(err, err_len, -3, "could not parse configuration of ", names[0]);
    cfg->num_events = ne;
    int i;
    for (i = 0; i < 14; ++i) {
        *dst[i] = strtod(tok[1 + i], &end);
        if (*end)
            return set_err(err, err_len, -3, "could not parse configuration of ", names[1 + i]);
    }
    cfg->num_bins = (int32_t)strtol(tok[15], &end, 10);
    if (*end)
        return set_err(err, err_len, -3, "could not parse configuration of ", names[15]);
    for (i = 16; i < 18; ++i) {
        char low[128];
        int c;
        for (c = 0; tok[i][c]; ++c)
            low[c] = (char)tolower((unsigned char)tok[i][c]);
        low[c] = 0;
        int v;
        if (!strcmp(low, ".true.") || !strcmp(tok[i], "true"))
            v = 1;
        else if (!strcmp(low, ".false.") || !strcmp(tok[i], "false"))
            v = 0;
        else
            return set_err(err, err_len, -3, "could not parse configuration of ", names[i]);
        if (i == 16)
            cfg->impr = v;
        else
            cfg->plot = v;
    }
    if (cfg->num_events == 0)
        return set_err(err, err_len, -4, "Please simulate at least one event", NULL);
    if (cfg->plot)
        return set_err(err, err_len, -5, "Plotting is not supported by this version", NULL);
    if (cfg->impr)
        return set_err(err, err_len, -6, "Individual result printing is not supported.", NULL);
    return 0;
}

#if defined(OR_F32)
#define R(x) ((float)(x))
/* LLVM folds powi on a float constant as (float)std::pow((double)x, n). */
#define RPOW(x, y) ((float)pow((double)(x), (double)(y)))
#else
#define R(x) ((double)(x))
#define RPOW(x, y) pow(x, y)
#endif

void OR_FN(or_consts)(const or_config_t *cfg, or_consts_t *k)
{
    or_scalar e_total = R(cfg->e_total), alpha = R(cfg->alpha), alpha_z = R(cfg->alpha_z);
    or_scalar sin2w = R(cfg->sin2w), m_z0 = R(cfg->m_z0), g_z0 = R(cfg->g_z0);
    or_scalar conv = R(cfg->conv), bp = R(cfg->beta_plus), bm = R(cfg->beta_minus);
    /* EventGenerator::new (evgen.rs:49-62) for 3 outgoing particles. */
    or_scalar z_n = R(2.) * OR_LOG(R(OR_PI) / R(2.));
    z_n -= R(2.) * OR_LOG(R(1.));
    z_n = z_n - OR_LOG(R(2.));
    or_scalar ln_weight = (R(2.) * R(3.) - R(4.)) * OR_LOG(e_total) + z_n;
    k->ev_weight = (or_scalar)exp(ln_weight);
#if defined(OR_F32)
    k->ev_weight = expf(ln_weight);
#endif
    /* Couplings (coupling.rs:23-33); powi(k) is binary powering. */
    or_scalar e2 = R(4.) * R(OR_PI) * alpha;
    or_scalar e2_z = R(4.) * R(OR_PI) * alpha_z;
    or_scalar cos2w = R(1.) - sin2w;
    or_scalar m2 = m_z0 * m_z0, m4 = m2 * m2;
    or_scalar g_beta = -OR_SQRT(e2_z / (R(4.) * cos2w * sin2w)) / m4;
    or_scalar x = OR_SQRT(e2);
    k->g_a = -(x * (x * x));
    k->g_bp = g_beta;
    k->g_bm = g_beta;
    /* ResultsAccumulator::new (resacc.rs:59-100). */
    or_scalar fact_com = R(1.) / R(6.) * conv;
    or_scalar relat_width = g_z0 / m_z0;
    or_scalar p_aa = R(2.);
    or_scalar p_ab = R(1.) - R(4.) * sin2w;
    or_scalar p_bb = p_ab + R(8.) * (sin2w * sin2w);
    or_scalar c_aa = fact_com * p_aa;
    or_scalar c_ab = fact_com * p_ab / m2;
    or_scalar c_bb = fact_com * p_bb / m4;
    or_scalar rr = e_total / m_z0;
    or_scalar dzeta = rr * rr;
    or_scalar delta = (dzeta - R(1.)) / relat_width;
    or_scalar propagator = R(1.) / (R(1.) + delta * delta);
    or_scalar n_ev = (or_scalar)cfg->num_events;
    /* (2 pi)^-5 is folded at compile time by LLVM with the host pow(). */
    or_scalar norm = (or_scalar)RPOW(R(2.) * R(OR_PI), R(-5.)) / n_ev;
    or_scalar norm_weight = k->ev_weight * norm;
    or_scalar com = norm_weight / R(4.);
    or_scalar aa = com * c_aa;
    or_scalar bb = com * c_bb * propagator / (relat_width * relat_width);
    or_scalar ab = com * c_ab * R(2.) * bp * propagator / relat_width;
    k->fact_com = fact_com;
    k->norm_weight = norm_weight;
    k->propagator = propagator;
    k->delta_z0 = delta;
    k->sc[0] = aa;
    k->sc[1] = bb * (bp * bp);
    k->sc[2] = bb * (bm * bm);
    k->sc[3] = ab * delta;
    k->sc[4] = -ab;
}

static void make_phys(const or_config_t *cfg, const or_opts_t *opt, or_phys *ph)
{
    or_consts_t k;
    OR_FN(or_consts)(cfg, &k);
    ph->e_total = R(cfg->e_total);
    ph->cut_beam = R(cfg->cut_beam);
    ph->cut_pp = R(cfg->cut_pp);
    ph->e_min = R(cfg->e_min);
    ph->cut_plane = R(cfg->cut_plane);
    ph->g_a = k.g_a;
    ph->g_bp = k.g_bp;
    ph->g_bm = k.g_bm;
    memcpy(ph->sc, k.sc, sizeof ph->sc);
    ph->sort = opt ? opt->sort : 1;
}

/* ======================================================================== */
/* faster-evgen event generation (evgen.rs:143-173, 221-249)                 */
/* ======================================================================== */

static void generate_raw_faster(or_rng_t *g, or_real q[3][4])
{
    or_real r9[9], r6[6], pts[3][2], r2[3];
    int i;
    rng_array(g, 9, r9);
    /* random_unit_2d_outgoing: 3x2 matrix filled column-major. */
    rng_array(g, 6, r6);
    for (i = 0; i < 3; ++i) {
        pts[i][0] = (or_real)2. * r6[i] - (or_real)1.;
        pts[i][1] = (or_real)2. * r6[3 + i] - (or_real)1.;
        r2[i] = ((or_real)0. + pts[i][0] * pts[i][0]) + pts[i][1] * pts[i][1];
    }
    const or_real min2 = OR_MIN_POSITIVE * OR_MIN_POSITIVE;
    for (i = 0; i < 3; ++i) {
        while (r2[i] > (or_real)1. || r2[i] < min2) {
            or_real r[2];
            rng_array(g, 2, r);
            pts[i][0] = (or_real)2. * r[0] - (or_real)1.;
            pts[i][1] = (or_real)2. * r[1] - (or_real)1.;
            r2[i] = (or_real)0. + (pts[i][0] * pts[i][0] + pts[i][1] * pts[i][1]);
        }
    }
    for (i = 0; i < 3; ++i) {
        or_real nrm = (or_real)1. / OR_SQRT(r2[i]);
        pts[i][0] *= nrm;
        pts[i][1] *= nrm;
    }
    for (i = 0; i < 3; ++i) {
        or_real cos_t = (or_real)2. * r9[i] - (or_real)1.;
        or_real em = r9[3 + i] * r9[6 + i];
        or_real st = OR_SQRT((or_real)1. - cos_t * cos_t);
        or_real en = -OR_LOG(em + OR_MIN_POSITIVE);
        q[i][OR_X] = en * (st * pts[i][0]);
        q[i][OR_Y] = en * (st * pts[i][1]);
        q[i][OR_Z] = en * cos_t;
        q[i][OR_E] = en * (or_real)1.;
    }
}

/* Conformal map + sort + event assembly from given massless q (the part of
 * or_generate after generate_raw), used by faster-evgen. */
static void generate_from_q(const or_phys *ph, or_real q[3][4], or_real p[5][4])
{
    /* Invert the default generate_raw: feeding q through or_generate would
     * need uniforms, so the conformal part is restated here verbatim. */
    or_real r[4];
    int i, k;
    for (k = 0; k < 4; ++k)
        r[k] = (((or_real)0. + q[0][k]) + q[1][k]) + q[2][k];
    or_real rn2 = r[OR_E] * r[OR_E] -
                  ((or_real)0. + ((r[OR_X] * r[OR_X] + r[OR_Y] * r[OR_Y]) + r[OR_Z] * r[OR_Z]));
    or_real alpha = ph->e_total / rn2;
    or_real rn = OR_SQRT(rn2);
    or_real beta = (or_real)1. / (rn + r[OR_E]);
    or_real pe[3], pxyz[3][3];
    for (i = 0; i < 3; ++i) {
        or_real rq = q[i][OR_X] * r[OR_X];
        rq = q[i][OR_Y] * r[OR_Y] + rq;
        rq = q[i][OR_Z] * r[OR_Z] + rq;
        pe[i] = alpha * (r[OR_E] * q[i][OR_E] - rq);
        or_real b = beta * rq - q[i][OR_E];
        for (k = 0; k < 3; ++k)
            pxyz[i][k] = alpha * (rn * q[i][k] + b * r[k]);
    }
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

/* ======================================================================== */
/* Batches, runs, merge, finalize                                            */
/* ======================================================================== */

static void acc_zero(or_acc_t *a) { memset(a, 0, sizeof *a); }

/* ResultsAccumulator::integrate (resacc.rs:121-129). */
static void acc_integrate(or_acc_t *a, const or_phys *ph, const or_real m[5])
{
    int c;
    a->selected += 1;
    for (c = 0; c < 5; ++c)
        a->spm2[c] += m[c];
    for (c = 0; c < 5; ++c)
        a->vars[c] += m[c] * m[c];
    or_real w = or_weight(ph, m);
    a->sigma += w;
    a->variance += w * w;
}

void OR_FN(or_merge)(or_acc_t *into, const or_acc_t *o)
{
    int c;
    into->selected += o->selected;
    for (c = 0; c < 5; ++c)
        into->spm2[c] += o->spm2[c];
    for (c = 0; c < 5; ++c)
        into->vars[c] += o->vars[c];
    into->sigma += o->sigma;
    into->variance += o->variance;
}

static void one_event(const or_phys *ph, int faster_evgen, or_rng_t *g, or_acc_t *acc)
{
    or_real p[5][4], m2[5][8], m[5];
    if (faster_evgen) {
        or_real q[3][4];
        generate_raw_faster(g, q);
        generate_from_q(ph, q, p);
    } else {
        or_real u[12];
        int i;
        for (i = 0; i < 12; ++i)
            u[i] = OR_FN(or_rng_random)(g);
        or_generate(ph, u, p);
    }
    if (!or_keep(ph, p))
        return;
    or_m2(ph, p, m2);
    or_m2_sums(m2, m);
    acc_integrate(acc, ph, m);
}

static void simulate_batch_phys(const or_phys *ph, int faster_evgen, or_rng_t *g, uint64_t n, or_acc_t *out)
{
    uint64_t e;
    acc_zero(out);
    for (e = 0; e < n; ++e)
        one_event(ph, faster_evgen, g, out);
}

void OR_FN(or_simulate_batch)(const or_config_t *cfg, const or_opts_t *opt, or_rng_t *g, uint64_t n,
                              or_acc_t *out)
{
    or_phys ph;
    make_phys(cfg, opt, &ph);
    simulate_batch_phys(&ph, opt ? opt->faster_evgen : 0, g, n, out);
}

void OR_FN(or_events_from_uniforms)(const or_config_t *cfg, const or_opts_t *opt, uint64_t n,
                                    const or_scalar *u, or_scalar *p, int32_t *keep, or_scalar *m2,
                                    or_scalar *m, or_scalar *w)
{
    or_phys ph;
    uint64_t e;
    make_phys(cfg, opt, &ph);
    for (e = 0; e < n; ++e) {
        or_real pp[5][4], mm2[5][8], mm[5];
        or_generate(&ph, u + 12 * e, pp);
        int k = or_keep(&ph, pp);
        or_m2(&ph, pp, mm2);
        or_m2_sums(mm2, mm);
        if (p)
            memcpy(p + 20 * e, pp, sizeof pp);
        if (keep)
            keep[e] = k;
        if (m2)
            memcpy(m2 + 40 * e, mm2, sizeof mm2);
        if (m)
            memcpy(m + 5 * e, mm, sizeof mm);
        if (w)
            w[e] = or_weight(&ph, mm);
    }
}

/* Spinor products s_ij of given momenta (spinor.rs:31-54): the full 5x5
 * matrix of each event, s[n][5][5][2] as (re, im). */
void OR_FN(or_spinors_at)(uint64_t n, const or_scalar *p, or_scalar *s)
{
    uint64_t e;
    for (e = 0; e < n; ++e) {
        or_real pp[5][4];
        or_spinor sp;
        int i, j;
        memcpy(pp, p + 20 * e, sizeof pp);
        or_spinor_new(&sp, pp);
        for (i = 0; i < 5; ++i)
            for (j = 0; j < 5; ++j) {
                s[50 * e + 10 * i + 2 * j] = sp.s[i][j].re;
                s[50 * e + 10 * i + 2 * j + 1] = sp.s[i][j].im;
            }
    }
}

/* Multi-threaded reproducible driver (multi_threading.rs:16-127): the
 * scheduling thread hands each batch a clone of the master stream and then
 * advances the master by the batch's draws (or jumps it in faster-threading
 * mode); workers simulate batches concurrently; results are folded in batch
 * order. */
typedef struct {
    const or_phys *ph;
    const or_opts_t *opt;
    or_rng_t *states;       /* per-batch start states */
    atomic_int *ready;      /* per-batch "state published" flags */
    or_acc_t *results;
    uint64_t num_events, nb;
    atomic_ullong next;
} mt_ctx;

static uint64_t batch_len(uint64_t N, uint64_t b)
{
    uint64_t start = b * OR_BATCH;
    uint64_t left = N - start;
    return left < OR_BATCH ? left : OR_BATCH;
}

static void *mt_worker(void *arg)
{
    mt_ctx *c = (mt_ctx *)arg;
    for (;;) {
        uint64_t b = atomic_fetch_add(&c->next, 1);
        if (b >= c->nb)
            break;
        while (!atomic_load_explicit(&c->ready[b], memory_order_acquire))
            ;
        or_rng_t g = c->states[b];
        simulate_batch_phys(c->ph, c->opt->faster_evgen, &g, batch_len(c->num_events, b), &c->results[b]);
    }
    return NULL;
}

/* Advance the master stream past one batch (evgen.rs:257-267). */
static void advance_master(const or_phys *ph, const or_opts_t *opt, or_rng_t *g, uint64_t n)
{
    if (opt->faster_threading) {
        OR_FN(or_rng_jump)(g);
    } else if (opt->faster_evgen) {
        or_acc_t scratch;
        simulate_batch_phys(ph, 1, g, n, &scratch); /* same draws as the batch */
    } else {
        OR_FN(or_rng_skip)(g, n * 12);
    }
}

int OR_FN(or_run)(const or_config_t *cfg, const or_opts_t *opt, or_acc_t *merged, or_acc_t *per_batch)
{
    or_phys ph;
    make_phys(cfg, opt, &ph);
    uint64_t N = cfg->num_events;
    if (N == 0)
        return -1;
    uint64_t nb = (N + OR_BATCH - 1) / OR_BATCH;
    or_acc_t *res = per_batch ? per_batch : (or_acc_t *)malloc(nb * sizeof(or_acc_t));
    if (!res)
        return -2;
    or_rng_t master;
    OR_FN(or_rng_new)(opt->rng_kind, &master);
    int threads = opt->threads < 1 ? 1 : opt->threads;
    if (threads == 1) {
        /* sequential.rs:15-40 semantics; batch partition identical. */
        uint64_t b;
        for (b = 0; b < nb; ++b) {
            if (opt->faster_threading) {
                or_rng_t g = master;
                simulate_batch_phys(&ph, opt->faster_evgen, &g, batch_len(N, b), &res[b]);
                OR_FN(or_rng_jump)(&master);
            } else {
                simulate_batch_phys(&ph, opt->faster_evgen, &master, batch_len(N, b), &res[b]);
            }
        }
    } else {
        mt_ctx c;
        c.ph = &ph;
        c.opt = opt;
        c.states = (or_rng_t *)malloc(nb * sizeof(or_rng_t));
        c.ready = (atomic_int *)calloc(nb, sizeof(atomic_int));
        c.results = res;
        c.num_events = N;
        c.nb = nb;
        atomic_init(&c.next, 0);
        pthread_t *tid = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
        int t;
        for (t = 0; t < threads; ++t)
            pthread_create(&tid[t], NULL, mt_worker, &c);
        uint64_t b;
        for (b = 0; b < nb; ++b) {
            c.states[b] = master;
            atomic_store_explicit(&c.ready[b], 1, memory_order_release);
            if (b + 1 < nb)
                advance_master(&ph, opt, &master, batch_len(N, b));
        }
        for (t = 0; t < threads; ++t)
            pthread_join(tid[t], NULL);
        free(tid);
        free(c.states);
        free((void *)c.ready);
    }
    /* ReproducibleAccumulator::get_merged_result: left fold from batch 0. */
    or_acc_t acc = res[0];
    uint64_t b;
    for (b = 1; b < nb; ++b)
        OR_FN(or_merge)(&acc, &res[b]);
    if (merged)
        *merged = acc;
    if (!per_batch)
        free(res);
    return 0;
}

/* ResultsAccumulator::finalize (resacc.rs:142-223). */
void OR_FN(or_finalize)(const or_config_t *cfg, const or_acc_t *acc_in, or_final_t *out)
{
    or_consts_t k;
    OR_FN(or_consts)(cfg, &k);
    or_acc_t a = *acc_in;
    or_scalar n_ev = (or_scalar)cfg->num_events;
    int c, s;
    for (c = 0; c < 5; ++c) {
        or_scalar v = (a.vars[c] - a.spm2[c] * a.spm2[c] / n_ev) / (n_ev - R(1.));
        a.vars[c] = OR_SQRT(v / n_ev) / OR_FABS(a.spm2[c] / n_ev);
    }
    or_scalar spm2[2][5], vars[2][5];
    for (s = 0; s < 2; ++s)
        for (c = 0; c < 5; ++c) {
            spm2[s][c] = a.spm2[c];
            vars[s][c] = a.vars[c];
        }
    or_scalar sin2w = R(cfg->sin2w), e_total = R(cfg->e_total), g_z0 = R(cfg->g_z0), m_z0 = R(cfg->m_z0);
    or_scalar polar_p = R(-2.) * sin2w;
    or_scalar polar_m = R(1.) + polar_p;
    or_scalar polars[2] = {polar_m, polar_p};
    for (c = 1; c < 5; ++c)
        for (s = 0; s < 2; ++s)
            spm2[s][c] *= polars[s];
    for (c = 1; c < 3; ++c)
        for (s = 0; s < 2; ++s)
            spm2[s][c] *= polars[s];
    or_scalar incident_flux = R(1.) / (R(2.) * (e_total * e_total));
    or_scalar scale = k.fact_com * incident_flux * k.norm_weight;
    for (s = 0; s < 2; ++s)
        for (c = 0; c < 5; ++c)
            spm2[s][c] *= scale;
    or_scalar gm_z0 = g_z0 * m_z0;
    or_scalar pz = k.propagator / gm_z0;
    for (c = 1; c < 5; ++c)
        for (s = 0; s < 2; ++s)
            spm2[s][c] *= pz;
    for (c = 1; c < 3; ++c)
        for (s = 0; s < 2; ++s)
            spm2[s][c] /= gm_z0;
    for (s = 0; s < 2; ++s)
        spm2[s][3] *= k.delta_z0;
#define COLSUM(col) ((R(0.) + spm2[0][col]) + spm2[1][col])
#define INC_NUM(col)                                                                                   \
    OR_SQRT(R(0.) + ((spm2[0][col] * vars[0][col]) * (spm2[0][col] * vars[0][col]) +                  \
                     (spm2[1][col] * vars[1][col]) * (spm2[1][col] * vars[1][col])))
    or_scalar beta_min = OR_SQRT(COLSUM(0) / COLSUM(1));
    or_scalar ss_denom = COLSUM(0);
    or_scalar ss_norm = R(1.) / (R(2.) * OR_SQRT(ss_denom));
    or_scalar ss_p = COLSUM(1) * ss_norm;
    or_scalar ss_m = COLSUM(2) * ss_norm;
    or_scalar inc_common = INC_NUM(0) / (R(2.) * OR_FABS(ss_denom));
    or_scalar inc_ss_p = INC_NUM(1) / OR_FABS(COLSUM(1)) + inc_common;
    or_scalar inc_ss_m = INC_NUM(2) / OR_FABS(COLSUM(2)) + inc_common;
    or_scalar variance = (a.variance - a.sigma * a.sigma / n_ev) / (n_ev - R(1.));
    or_scalar prec = OR_SQRT(variance / n_ev) / OR_FABS(a.sigma / n_ev);
    or_scalar sigma = a.sigma * incident_flux;
#undef COLSUM
#undef INC_NUM
    out->selected = a.selected;
    memcpy(out->spm2, spm2, sizeof spm2);
    memcpy(out->vars, vars, sizeof vars);
    out->sigma = sigma;
    out->prec = prec;
    out->variance = variance;
    out->beta_min = beta_min;
    out->ss_p = ss_p;
    out->inc_ss_p = inc_ss_p;
    out->ss_m = ss_m;
    out->inc_ss_m = inc_ss_m;
}
