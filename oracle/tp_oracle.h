/*
 * tp_oracle.h -- C API of the CPU oracle for the 3photons event loop.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, the smoke check and the CPU-baseline
 * leg of bench.py load liboracle.so (via ctypes), and only as the checker or
 * as the timed CPU baseline -- never as the product path.
 *
 * The oracle restates the reference (HadrienG2/3photons-rust) on the CPU:
 *   random streams   src/random/ranf.rs, src/random/standard.rs (+ rand 0.8.5,
 *                    rand_xoshiro 0.6.0 published algorithms)
 *   per-event path   src/evgen.rs, src/evcut.rs, src/spinor.rs, src/matelems.rs
 *                    (or_event.h)
 *   accumulation     src/resacc.rs:121-139, scheduling/{sequential,multi_threading}.rs
 *   finalisation     src/resacc.rs:142-223
 *   configuration    src/config.rs:57-128, src/coupling.rs, src/resacc.rs:59-117
 *
 * The double-precision build exports the names below; the all-f32 build
 * (-DOR_F32, liboracle_f32.so) exports the same names with an _f32 suffix and
 * float in place of or_real.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(OR_F32)
typedef float or_scalar;
#define OR_FN(name) name##_f32
#else
typedef double or_scalar;
#define OR_FN(name) name
#endif

/* Random stream kinds. */
enum { OR_RNG_RANF = 0, OR_RNG_XOSHIRO = 1 };

/* One generator: RANF (ranf.rs:12-22) or xoshiro (xoshiro256+ in f64 builds,
 * xoshiro128+ in f32 builds, standard.rs:8-11). */
typedef struct {
    int32_t kind;
    int32_t ranf_seed;      /* current seed, used by jump() (ranf.rs:14-15,135-140) */
    int32_t numbers[56];    /* ranf.rs:18 */
    uint32_t index;         /* ranf.rs:21 */
    uint32_t pad;
    uint64_t xs[4];         /* xoshiro256+ state */
    uint32_t xs32[4];       /* xoshiro128+ state */
} or_rng_t;

/* Run configuration: the 18 positional items of `valeurs` (config.rs:8-53). */
typedef struct {
    uint64_t num_events;
    double e_total, cut_beam, cut_pp, e_min, cut_plane;
    double alpha, alpha_z, conv, m_z0, g_z0, sin2w, br_ep_em, beta_plus, beta_minus;
    int32_t num_bins, impr, plot, pad;
} or_config_t;

/* Per-batch / merged accumulator (resacc.rs:16-34). */
typedef struct {
    uint64_t selected;
    or_scalar spm2[5], vars[5], sigma, variance;
} or_acc_t;

/* Final results (resfin.rs:26-62); index [spin][contribution]. */
typedef struct {
    uint64_t selected;
    or_scalar spm2[2][5], vars[2][5];
    or_scalar sigma, prec, variance, beta_min, ss_p, inc_ss_p, ss_m, inc_ss_m;
} or_final_t;

/* Host-side constants derived from the configuration. */
typedef struct {
    or_scalar ev_weight;                 /* evgen.rs:59-62 */
    or_scalar g_a, g_bp, g_bm;           /* coupling.rs:23-33 */
    or_scalar fact_com, norm_weight, propagator, delta_z0; /* resacc.rs:64-86 */
    or_scalar sc[5];                     /* sigma_contribs, resacc.rs:88-100 */
} or_consts_t;

/* Run options: the reference's compile-time features, as run-time flags. */
typedef struct {
    int32_t rng_kind;        /* OR_RNG_RANF (default) or OR_RNG_XOSHIRO (standard-random) */
    int32_t faster_threading;/* per-batch jump() instead of skip-ahead (multi_threading.rs:64-69) */
    int32_t sort;            /* 1 = sort photons (default), 0 = no-photon-sorting */
    int32_t threads;         /* worker threads; 1 = sequential.rs */
    int32_t faster_evgen;    /* 1 = faster-evgen event generation (evgen.rs:143-173) */
    int32_t pad;
} or_opts_t;

/* Parse `valeurs` (config.rs:57-128).  Returns 0 or a negative error code and
 * writes a message into err (err_len bytes). */
int OR_FN(or_config_load)(const char *path, or_config_t *cfg, char *err, int err_len);

/* Couplings, event weight and accumulator constants. */
void OR_FN(or_consts)(const or_config_t *cfg, or_consts_t *k);

/* Random streams. */
void OR_FN(or_rng_new)(int32_t kind, or_rng_t *g);                      /* RandomGenerator::new() */
void OR_FN(or_rng_seeded)(int32_t kind, uint64_t seed, or_rng_t *g);    /* seeded_new / seed_from_u64 */
or_scalar OR_FN(or_rng_random)(or_rng_t *g);
void OR_FN(or_rng_fill)(or_rng_t *g, uint64_t n, or_scalar *out);
void OR_FN(or_rng_skip)(or_rng_t *g, uint64_t n);
void OR_FN(or_rng_jump)(or_rng_t *g);
uint64_t OR_FN(or_rng_next_raw)(or_rng_t *g); /* raw integer output (RANF int, xoshiro u64/u32) */

/* Per-event path on given uniforms (12 per event).  Any output pointer may be
 * NULL.  p: [n][5][4]; keep: [n]; m2: [n][5][8]; m: [n][5]; w: [n]. */
void OR_FN(or_events_from_uniforms)(const or_config_t *cfg, const or_opts_t *opt, uint64_t n,
                                    const or_scalar *u, or_scalar *p, int32_t *keep,
                                    or_scalar *m2, or_scalar *m, or_scalar *w);

/* Spinor products of given momenta p[n][5][4] (spinor.rs:31-54):
 * s[n][5][5][2], (re, im) of s_ij. */
void OR_FN(or_spinors_at)(uint64_t n, const or_scalar *p, or_scalar *s);

/* One batch: the closure of main.rs:103-128 (advances *g like the reference). */
void OR_FN(or_simulate_batch)(const or_config_t *cfg, const or_opts_t *opt, or_rng_t *g,
                              uint64_t n, or_acc_t *out);

/* Whole run: scheduling::run_simulation minus finalize.  per_batch (optional,
 * ceil(N/10000) entries) receives every batch accumulator in batch order. */
int OR_FN(or_run)(const or_config_t *cfg, const or_opts_t *opt, or_acc_t *merged, or_acc_t *per_batch);

/* Left fold of batch accumulators in batch order (resacc.rs:133-139). */
void OR_FN(or_merge)(or_acc_t *into, const or_acc_t *other);

/* ResultsAccumulator::finalize (resacc.rs:142-223). */
void OR_FN(or_finalize)(const or_config_t *cfg, const or_acc_t *acc, or_final_t *out);

/* Batch size of the reference scheduler (scheduling/mod.rs:21). */
#define OR_BATCH 10000

#ifdef __cplusplus
}
#endif
#endif
