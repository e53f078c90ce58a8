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
    TP_ERR_IO = -6,          /* output file could not be written */
    TP_ERR_INTERNAL = -7     /* unexpected failure inside the library (message in tp_last_error) */
} tp_status;

/* Random stream (src/random/mod.rs:13-16): RANF is the default, xoshiro the
 * `standard-random` feature: xoshiro256+ in f64 runs, xoshiro128+ in all-f32
 * runs (standard.rs:8-11 picks the generator from the float type). */
typedef enum { TP_RNG_RANF = 0, TP_RNG_XOSHIRO = 1, TP_RNG_XOSHIRO128 = 2 } tp_rng_kind;

/* How consecutive batches get their streams (multi_threading.rs:58-70):
 * skip-ahead = reproducible default (batch b starts 120 000*b draws in),
 * jump = `faster-threading` (RANF reseeded with seed+123456*b, xoshiro jumped
 * 2^128 per batch). */
typedef enum { TP_SEED_SKIPAHEAD = 0, TP_SEED_JUMP = 1 } tp_seeding;

/* Arithmetic of the per-event path.
 *   TP_PREC_F64   -- IEEE double, no contraction, reference operation order.
 *   TP_PREC_MIXED -- f64 RAMBO, cuts and weight; f32 spinor products and
 *                    helicity amplitudes; Kahan-compensated f64 batch sums.
 *   TP_PREC_F32   -- the reference's `f32` feature (numeric.rs:6-14):
 *                    everything in float in the reference's order, batch
 *                    sums in event order (RANF or xoshiro128+ streams, run
 *                    constants from tp_params_from_config_f32); the tp_acc
 *                    fields then hold f32 values, to be folded with
 *                    tp_acc_merge_f32 and finalised with tp_finalize_f32.
 *                    tp_adjoint accepts it too (f32 arithmetic). */
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
    int32_t faster_evgen;                 /* 1 = `faster-evgen` generation (evgen.rs:143-173, 221-249),
                                             0 = default.  Available for every stream (RANF,
                                             xoshiro256+, xoshiro128+ in f32) in f64 and f32, with
                                             either seeding; not in mixed precision */
    int32_t reserved;                     /* 0 */
} tp_params;

/* One batch accumulator, field for field ResultsAccumulator (resacc.rs:16-34). */
typedef struct {
    uint64_t selected_events;
    double spm2[5];
    double vars[5];
    double sigma;
    double variance;
} tp_acc;

/* Generator states: the values of the reference's generators, not their
 * memory layout (a caller copies in and out).
 * RANF: ranf.rs:12-22 (numbers[0] unused, index = values left in the round
 * (usize there), seed = current seed for jump(), a field the reference has
 * only with faster-threading).  xoshiro256+: rand_xoshiro's four words. */
typedef struct {
    int32_t numbers[56];
    uint32_t index;
    int32_t seed;
} tp_ranf_state;

typedef struct {
    uint64_t s[4];
} tp_xoshiro_state;

/* xoshiro128+: rand_xoshiro's [u32; 4] (the all-f32 mode's generator). */
typedef struct {
    uint32_t s[4];
} tp_xoshiro128_state;

/* Optional timing of a call, measured with HIP events on the context stream. */
typedef struct {
    double device_ms;   /* first kernel launch to last kernel end */
    double kernel_ms;   /* the fused event-loop kernel(s) only */
    uint64_t events;
    uint64_t batches;
    uint64_t host_walk_chunks; /* faster-evgen with skip-ahead seeding: chunks of batches whose event
                                  starts came from the host walk (runs below 2^20 events, the
                                  closure, or device walkers that did not all meet) */
    uint64_t chunks;           /* faster-evgen: chunks of batches the run was split into (each one's
                                  stream expanded, walked and simulated in turn); 0 otherwise */
} tp_stats;

/* Context: binds device `device` (HIP ordinal), creates its stream. */
int tp_create(int device, void **ctx_out);
int tp_destroy(void *ctx);
const char *tp_last_error(const void *ctx);

/* Initial stream of a run: RandomGenerator::new() (scheduling/mod.rs:40 ->
 * ranf.rs:28-32 seed 234612947, standard.rs:21-24 seed_from_u64(12345)).
 * state points to tp_ranf_state or tp_xoshiro_state. */
int tp_rng_init(int rng_kind, void *state);
/* Same with an explicit seed (RANF seeded_new / xoshiro seed_from_u64). */
int tp_rng_seed(int rng_kind, uint64_t seed, void *state);

/* RandomGenerator::skip (ranf.rs:122-127, standard.rs:37-43): advance the
 * state as n calls of random() would, in O(log n) (polynomial / GF(2)
 * skip-ahead instead of stepping). */
int tp_rng_skip(int rng_kind, void *state, uint64_t n);
/* RandomGenerator::jump (ranf.rs:135-140 reseed with seed+123456,
 * standard.rs:51-55 the 2^128 jump). */
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
 * enqueued; tp_sync() waits for it.  A faster-evgen run returns when its
 * batches are done: its device stream walk reports through a flag read back
 * before the call returns. */
int tp_run_device(void *ctx, const tp_params *params, int rng_kind, int seeding,
                  const void *rng_state0, uint64_t num_events, tp_acc *per_batch_dev,
                  tp_acc *merged_dev, tp_stats *stats);

/* tp_run_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * runs the xoshiro256+ stream only (rng_kind TP_RNG_XOSHIRO, precision
 * TP_PREC_F64 or TP_PREC_MIXED, both seedings, no faster-evgen; anything else
 * returns TP_ERR_UNSUPPORTED) and at most 2^32 - 1 batches (TP_ERR_ARG
 * beyond).  Its skip-ahead tables sit in a static device array of 800 KiB
 * per GPU, filled once by tp_create and only read afterwards, so calls that
 * overlap on one GPU share it without racing; per_batch_dev and merged_dev
 * are the caller's, as for tp_run_device.  Returns when enqueued; tp_sync
 * waits. */
int tp_run_device_noalloc(void *ctx, const tp_params *params, int rng_kind, int seeding,
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

/* ---- Learned proposal ---------------------------------------------------
 * No reference counterpart: the reference samples RAMBO's 12 uniforms flat
 * (evgen.rs:174-207).  These entry points add a RealNVP normalising flow on
 * the 12-dimensional unit cube (6 affine coupling layers, each with a
 * 12 -> 64 -> 64 -> 24 hard-swish MLP on bf16 MFMA with f32 accumulation), trained
 * on the GPU by reverse KL through the adjoint of the event integrand, and a
 * weighted event loop that samples the mixture alpha q + (1 - alpha) U and
 * feeds the same per-batch accumulators (tp_acc) with every |M|^2 multiplied
 * by 1 / q_mix, so tp_finalize yields the cross-section and its error as for
 * a plain run.
 *
 * Parameters (f32, TP_FLOW_NPARAM values): per layer l, at l * 7328,
 *   W1[64][16] (inputs 12..15 padding), b1[64], W2[64][64], b2[64],
 *   W3[32][64], b3[32] (rows 0..11 = s, rows 16..27 = t, others padding).
 * Layer l keeps the features whose bit is set in
 *   {0x03F, 0xFC0, 0x555, 0xAAA, 0x1C7, 0xE38}[l]
 * and maps the others as z <- z exp(1.5 tanh(s_raw)) + t.
 * The flow's randomness is a counter-based SplitMix64 stream: draw k of key K
 * is mix64(K + (k + 1) 0x9E3779B97F4A7C15); sample i uses draws 12 i .. 12 i + 11. */
#define TP_FLOW_NPARAM 43968

typedef struct {
    double alpha;       /* flow share of the sampling mixture (0.7) */
    uint64_t batch;     /* samples per training step (65536; multiple of 16) */
    int32_t steps;      /* training steps (2000) */
    float lr;           /* Adam step size (3e-3) */
    float beta1, beta2, adam_eps; /* Adam moments (0.9, 0.999, 1e-8) */
    float grad_clip;    /* per-sample, per-feature clip of dloss/dz (100) */
    double eps_rel;     /* eps of log(f + eps), relative to the mean of f over the cube (1e-3) */
    double cut_width;   /* training integrand: cuts smoothed over this width tau of their
                           dimensionless margins (0.05) */
    int32_t target;     /* training integrand g, loss log q - log g:
                           0: f cut + eps (hard cuts, whose boundaries the gradient cannot see),
                           1: f S + eps, 2: min(f, fcap) S with log g = log min(f, fcap) + log S (2) */
    double cap_rel;     /* target 2: fcap relative to the mean of f cut over the cube (10) */
    uint64_t seed;      /* key of the training stream (1) */
} tp_flow_opts;

int tp_flow_default_opts(tp_flow_opts *opts);
/* Flow bound to a context; weights initialised from init_seed (hidden layers
 * uniform in +-1/sqrt(fan_in), output layer zero: the identity map). */
int tp_flow_create(void *ctx, uint64_t init_seed, void **flow_out);
int tp_flow_destroy(void *flow);
int tp_flow_get_params(void *flow, float *w);       /* host, TP_FLOW_NPARAM */
int tp_flow_set_params(void *flow, const float *w); /* host, TP_FLOW_NPARAM; resets Adam */
/* Reverse-KL training: per step, batch samples u ~ q, loss mean(log q(u) -
 * log g(u)) with g built from f = sum_c sigma_contribs[c] |M|^2_c as opts
 * target says, gradient through the flow and the hand-written adjoint of f,
 * one Adam update.
 * loss (host, steps values) may be null; stats gets the device time. */
int tp_flow_train(void *flow, const tp_params *params, const tp_flow_opts *opts, float *loss, tp_stats *stats);
/* n samples of the flow: u[n][12] (f64), z[n][12] = logit(u) (f32) and
 * log q(u) (f32), from the stream (key, samples sample0 ..) or, when v_in is
 * non-null, from the given base points v_in[n][16] (f32, columns 12..15
 * ignored).  Pointers are device pointers when on_device is 1; any output
 * may be null. */
int tp_flow_sample(void *flow, uint64_t key, uint64_t sample0, uint64_t n, const float *v_in, int on_device,
                   double *u, float *z, float *logq);
/* log q(u) of n given points u[n][12]. */
int tp_flow_density(void *flow, uint64_t n, const double *u, int on_device, float *logq);
/* Gradient of (1/n) sum_i [gz_i . z_i - sum_l sum_j s_lij] (z_i the flow
 * output of base point v_i; the second term is log q's Jacobian part) with
 * respect to the parameters (grad, TP_FLOW_NPARAM) and, per sample, to
 * z0_i = logit(v_i) (gv[n][16], without the 1/n).  n must be a multiple of 16. */
int tp_flow_grad(void *flow, uint64_t n, const float *v_in, const float *gz, int on_device, float *grad, float *gv);
/* The training integrand at n given points: f cut (target 0) or g (1, 2),
 * gz = d/dz[-sum log u(1-u) - log g] clipped to +-clip (n x 16, columns
 * 12..15 zero) and log g, with g as tp_flow_opts.target defines (RAMBO
 * weight 1, smoothed cuts of width tau, cap fcap). */
int tp_flow_integrand(void *flow, const tp_params *params, uint64_t n, const double *u, double eps, double tau,
                      double fcap, int target, float clip, int on_device, double *f, float *gz, float *logf);
/* Weighted event loop: num_events events in batches of 10 000, the first
 * round(alpha * n) events of every chunk of batches drawn from the flow (key,
 * draws continuing across chunks), the rest uniform (key ^ 0x5555...5),
 * every event weighted by 1 / (alpha q(u) + 1 - alpha); per-batch
 * accumulators and (if merged_dev is non-null) their in-order fold in device
 * memory, as tp_run_device.  Returns when enqueued; tp_sync waits. */
int tp_flow_run_device(void *flow, const tp_params *params, double alpha, uint64_t key, uint64_t num_events,
                       tp_acc *per_batch_dev, tp_acc *merged_dev, tp_stats *stats);
/* tp_flow_run_device with host outputs (per_batch_out: ceil(n / 10 000)
 * accumulators; merged_out may be null), like tp_run; returns when done. */
int tp_flow_run(void *flow, const tp_params *params, double alpha, uint64_t key, uint64_t num_events,
                tp_acc *per_batch_out, tp_acc *merged_out, tp_stats *stats);

/* Events per batch (scheduling/mod.rs:21). */
#define TP_BATCH_SIZE 10000

#ifdef __cplusplus
}
#endif
#endif /* TP_H */
