/*
 * tp.h -- C ABI of the MI355X (gfx950) engine for the 3photons event loop.
 *
 * The reference (HadrienG2/3photons-rust) runs its hot path through one Rust
 * closure,
 *
 *     simulate_events: impl Fn(usize, &mut RandomGenerator) -> ResultsAccumulator
 *                                                    (src/main.rs:103-128)
 *
 * which the scheduler calls once per 10 000-event batch (src/scheduling/mod.rs:21,
 * 31-59; sequential.rs:24-36; multi_threading.rs:44-71) before left-folding the
 * batch accumulators in batch order.  This header is what a Rust crate would
 * bind with `extern "C"` to move that work onto one MI355X:
 *
 *   tp_simulate_batch  -- drop-in for one call of the closure (same contract:
 *                         n events from *rng_state, rng_state advanced by the
 *                         draws those events consume, one accumulator back);
 *   tp_run             -- drop-in for scheduling::run_simulation_impl: every
 *                         batch of a run, per-batch accumulators in batch order
 *                         (the caller folds them, resacc.rs:133-139);
 *   tp_run_device      -- same run with the accumulators left in device memory.
 *
 * No C++ types, exceptions or allocations cross this boundary.  Errors are
 * negative tp_status codes with a message from tp_last_error().  A tp_ctx is
 * bound to one GPU and one HIP stream and must be used by one host thread at a
 * time; distinct contexts are independent.
 */
#ifndef TP_H
#define TP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
typedef enum {
    TP_OK = 0,
    TP_ERR_ARG = -1,         /* invalid argument (null pointer, bad size, bad enum) */
    TP_ERR_HIP = -2,         /* HIP runtime error (message has the HIP error string) */
    TP_ERR_NOMEM = -3,       /* device or host allocation failed */
    TP_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
    TP_ERR_CONFIG = -5,      /* configuration file missing/garbled/rejected */
    TP_ERR_IO = -6           /* output file could not be written */
} tp_status;

/* Random stream (src/random/mod.rs:13-16): RANF is the default, xoshiro the
 * `standard-random` feature (xoshiro256+ in f64 runs). */
typedef enum { TP_RNG_RANF = 0, TP_RNG_XOSHIRO = 1 } tp_rng_kind;

/* How consecutive batches get their streams (multi_threading.rs:58-70):
 * skip-ahead = reproducible default (batch b starts 120 000*b draws in),
 * jump = `faster-threading` (RANF reseeded with seed+123456*b, xoshiro jumped
 * 2^128 per batch). */
typedef enum { TP_SEED_SKIPAHEAD = 0, TP_SEED_JUMP = 1 } tp_seeding;

/* Arithmetic of the per-event path.
 *   TP_PREC_F64   -- IEEE double, no contraction, reference operation order.
 *   TP_PREC_MIXED -- f64 RAMBO, cuts and weight; f32 spinor products and
 *                    helicity amplitudes; Kahan-compensated f64 batch sums.
 *   TP_PREC_F32   -- everything in float (tp_adjoint only). */
typedef enum { TP_PREC_F64 = 0, TP_PREC_MIXED = 1, TP_PREC_F32 = 2 } tp_precision;

/* Everything the per-event path reads, all computed on the host exactly as
 * the reference does (coupling.rs:23-33, resacc.rs:59-100, evgen.rs:40-77). */
typedef struct {
    double e_total;                       /* sqrt(s) (GeV) */
    double cut_beam, cut_pp, e_min, cut_plane; /* evcut.rs:11-23 */
    double g_a, g_beta_p, g_beta_m;       /* coupling.rs:10-20 */
    double sigma_contribs[5];             /* resacc.rs:94-100 */
    double ev_weight;                     /* RAMBO event weight (evgen.rs:59-62) */
    int32_t sort_photons;                 /* 1 = sort by energy (default), 0 = no-photon-sorting */
    int32_t precision;                    /* tp_precision */
} tp_params;

/* One batch accumulator, field for field ResultsAccumulator (resacc.rs:16-34). */
typedef struct {
    uint64_t selected_events;
    double spm2[5];
    double vars[5];
    double sigma;
    double variance;
} tp_acc;

/* Generator states, field for field the reference structs.
 * RANF: ranf.rs:27-38 (numbers[0] unused, index = values left in the round,
 * seed = current seed for jump()).  xoshiro256+: rand_xoshiro's [u64; 4]. */
typedef struct {
    int32_t numbers[56];
    uint32_t index;
    int32_t seed;
} tp_ranf_state;

typedef struct {
    uint64_t s[4];
} tp_xoshiro_state;

/* Optional timing of a call, measured with HIP events on the context stream. */
typedef struct {
    double device_ms;   /* first kernel launch to last kernel end */
    double kernel_ms;   /* the fused event-loop kernel(s) only */
    uint64_t events;
    uint64_t batches;
} tp_stats;

/* Context: binds device `device` (HIP ordinal), creates its stream. */
int tp_create(int device, void **ctx_out);
int tp_destroy(void *ctx);
const char *tp_last_error(const void *ctx);

/* Initial stream of a run: RandomGenerator::new() (scheduling/mod.rs:40 ->
 * ranf.rs:44-48 seed 234612947, standard.rs:21-24 seed_from_u64(12345)).
 * state points to tp_ranf_state or tp_xoshiro_state. */
int tp_rng_init(int rng_kind, void *state);
/* Same with an explicit seed (RANF seeded_new / xoshiro seed_from_u64). */
int tp_rng_seed(int rng_kind, uint64_t seed, void *state);

/* RandomGenerator::skip (ranf.rs:137-143, standard.rs:194-200): advance the
 * state as n calls of random() would, in O(log n) (polynomial / GF(2)
 * skip-ahead instead of stepping). */
int tp_rng_skip(int rng_kind, void *state, uint64_t n);
/* RandomGenerator::jump (ranf.rs:151-156 reseed with seed+123456,
 * standard.rs:208-212 the 2^128 jump). */
int tp_rng_jump(int rng_kind, void *state);

/* Drop-in for one call of the per-batch closure (main.rs:103-128):
 * simulates n events (n <= 10 000 * 64) starting from *rng_state, writes one
 * accumulator to *out (host memory), and advances *rng_state by 12*n draws, as
 * `&mut rng` is advanced in sequential mode (sequential.rs:26-36). */
int tp_simulate_batch(void *ctx, const tp_params *params, int rng_kind, void *rng_state,
                      uint32_t n, tp_acc *out);

/* Drop-in for scheduling::run_simulation_impl (sequential.rs:15-40 and
 * multi_threading.rs:16-75): simulates num_events events from rng_state0 in
 * batches of 10 000 (last batch partial), writes the ceil(num_events/10 000)
 * batch accumulators in batch order to per_batch_out (host memory) and, if
 * merged_out is non-null, their left fold in batch order (resacc.rs:133-139;
 * multi_threading.rs:107-126).  stats may be null. */
int tp_run(void *ctx, const tp_params *params, int rng_kind, int seeding, const void *rng_state0,
           uint64_t num_events, tp_acc *per_batch_out, tp_acc *merged_out, tp_stats *stats);

/* tp_run with per_batch_dev (device memory, ceil(N/10 000) tp_acc) written on
 * the context stream and merged_dev (device memory, one tp_acc, may be null)
 * holding the in-order fold computed on the device.  Returns when the work is
 * enqueued; tp_sync() waits for it. */
int tp_run_device(void *ctx, const tp_params *params, int rng_kind, int seeding,
                  const void *rng_state0, uint64_t num_events, tp_acc *per_batch_dev,
                  tp_acc *merged_dev, tp_stats *stats);

/* Per-event diagnostics (the reference's debug-only Display of Event and
 * MEsContributions, event.rs:113-126, matelems.rs:89-105): runs the same
 * device code as the event loop on n events drawn from *rng_state (not
 * advanced) and writes, per event, the 12 uniforms u[n][12], the sorted
 * photon momenta p[n][3][4], the cut decision keep[n] and the helicity sums
 * m[n][5] (zeros for rejected events).  Host buffers; any may be null. */
int tp_events(void *ctx, const tp_params *params, int rng_kind, const void *rng_state, uint32_t n,
              double *u, double *p, int32_t *keep, double *m);

/* Adjoint of the event integrand.  No reference counterpart: the reference
 * only evaluates F_c(u) = |M|^2_c(p(u)) * cut(p(u)) * w for its five
 * contributions c (matelems.rs:53-86, evcut.rs:42-96, evgen.rs:89-207 with w
 * the constant RAMBO weight); this entry point also returns, for each of n
 * points, the hand-written reverse-mode derivatives
 *   dF_c/du_j  (j < 12, the uniforms in the evgen draw order) and
 *   dF_c/dp_k  (k < 12, the sorted photon momenta, photon-major X,Y,Z,E).
 * Points take their uniforms from u_in (n x 12) when it is non-null, else
 * from *rng_state (12 draws per point, not advanced).  precision is
 * TP_PREC_F64 (double outputs) or TP_PREC_F32 (float outputs, all-f32
 * arithmetic).  Outputs F[n][5], dF_du[n][5][12], dF_dp[n][5][12] and u_in
 * are device pointers when on_device is 1, host pointers when 0; any output
 * may be null.  Cut points have F = 0 and zero derivatives. */
int tp_adjoint(void *ctx, const tp_params *params, int rng_kind, const void *rng_state, const void *u_in,
               uint64_t n, int precision, int on_device, void *F, void *dF_du, void *dF_dp);

/* Wait for all work enqueued on the context stream; fills stats timings of the
 * last tp_run_device when stats is non-null. */
int tp_sync(void *ctx, tp_stats *stats);

/* Events per batch (scheduling/mod.rs:21). */
#define TP_BATCH_SIZE 10000

#ifdef __cplusplus
}
#endif
#endif /* TP_H */
