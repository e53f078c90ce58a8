// tp_rng_host.h -- host-side stream algebra for the engine: generator
// initialisation, skip-ahead and jump tables.
//
// RANF (src/random/ranf.rs) is Knuth's subtractive lagged-Fibonacci
// generator, X_g = X_{g-55} - X_{g-24} mod 1e9 in generation order; each
// IRN55 round computes 55 new values in index order and the stream emits them
// from index 55 down to 1.  The recurrence is linear over Z/1e9 with
// characteristic polynomial x^55 + x^31 - 1, so jumping n values ahead is
// multiplying the 55-value window by x^n mod that polynomial.
//
// xoshiro256+ (src/random/standard.rs via rand_xoshiro 0.6.0) has a state
// update that is linear over GF(2); skip-ahead by n steps is the 256x256 bit
// matrix M^n, built here by squaring and kept as 256 columns of 4 words.
#pragma once

#include <cstdint>
#include <vector>

namespace tp {

// ---- RANF ------------------------------------------------------------------
constexpr int32_t RANF_MODULO = 1000000000;
constexpr int32_t RANF_SEED = 234612947;

struct RanfState {
    int32_t numbers[56];
    uint32_t index;
    int32_t seed;
};

void ranf_seeded(RanfState &g, int32_t seed);     // IN55 + 10 warm-up rounds (ranf.rs:36-66)
void ranf_round(int32_t *numbers /*[56]*/);       // IRN55 (ranf.rs:106-119)

// Polynomial modulo x^55 + x^31 - 1 with coefficients mod 1e9.
using RanfPoly = std::vector<uint32_t>; // 55 coefficients, c[k] of x^k
RanfPoly ranf_poly_xpow(uint64_t n);       // x^n mod P
RanfPoly ranf_poly_mul(const RanfPoly &a, const RanfPoly &b);
// Apply a jump polynomial (x^(55 m)) to the round X[1..55]: returns X_{r+m}.
void ranf_apply_round_jump(const RanfPoly &c, const int32_t *x_in /*[56]*/, int32_t *x_out /*[56]*/);
// Advance a generator state by n single draws (same effect as n random()).
void ranf_advance(RanfState &g, uint64_t n);

// `faster-evgen` consumes a variable number of draws per event (9 + 6, plus
// 2 per rejected unit-disc point, each random_array::<N> call first
// discarding a round's leftovers when fewer than N remain: ranf.rs:78-102,
// evgen.rs:143-173,221-249), so event starts are found by walking the stream
// as EventGenerator::simulate_event_batch does (evgen.rs:257-267).  Stream
// positions q count values from the start of the state's current round
// (value q sits at index 55 - q % 55 of round q / 55), so the state g itself
// is at q0 = 55 - g.index.  starts[e] receives the position of event e's
// first random_array::<9> call (before any discard); on return g is the
// state after the last event and *q_end its position.
// faster-evgen stream walk: the start position of each event; g ends at the
// position after the last event.  Walks of 4e6 events or more use up to
// max_threads threads (1: sequential); the result does not depend on it.
// With f32 the unit-disc test runs in f32, as in the reference's f32 build.
void ranf_walk_faster_evgen(RanfState &g, uint64_t num_events, uint64_t *starts, uint64_t *q_end,
                            int max_threads = 16, bool f32 = false);
// The same walk from stream position q of g0's frame (q >= 55 - g0.index, the
// start of an event of the walk from g0): starts[] and *q_end in g0's frame.
void ranf_walk_faster_evgen_from(const RanfState &g0, uint64_t q, uint64_t num_events, uint64_t *starts,
                                 uint64_t *q_end, int max_threads = 16, bool f32 = false);

// ---- xoshiro256+ -------------------------------------------------------------
void xoshiro_seed(uint64_t seed, uint64_t s[4]); // seed_from_u64 (SplitMix64 expansion)
uint64_t xoshiro_next(uint64_t s[4]);

struct Gf2Mat {
    uint64_t col[256][4]; // column i = image of unit vector e_i
};
void gf2_identity(Gf2Mat &m);
void gf2_step(Gf2Mat &m);                    // one xoshiro256+ state update
void gf2_jump128(Gf2Mat &m);                 // the published 2^128 jump()
void gf2_mul(const Gf2Mat &a, const Gf2Mat &b, Gf2Mat &out); // out = a * b
void gf2_apply(const Gf2Mat &m, const uint64_t in[4], uint64_t out[4]);
// Tables T[k] = base^(2^k), k = 0..count-1.
std::vector<Gf2Mat> gf2_pow2_table(const Gf2Mat &base, int count);
// base^n from a pow2 table of base.
void gf2_pow_from_table(const std::vector<Gf2Mat> &table, uint64_t n, Gf2Mat &out);
// M^(2^k) for the single-step matrix M, k = 0..63 (built once, cached).
const std::vector<Gf2Mat> &xoshiro_step_pow2();
// Advance a state by n steps with the cached tables.
void xoshiro_advance(uint64_t s[4], uint64_t n);
// faster-evgen walk of the xoshiro256+ stream (standard.rs:33-35: each
// random_array::<N> is N consecutive draws, no discards): num_events event
// starts from position q of s0's frame (q = draws after s0, the start of an
// event), and *q_end the position after the last event.  Sequential.
void xoshiro_walk_faster_evgen(const uint64_t s0[4], uint64_t q, uint64_t num_events, uint64_t *starts,
                               uint64_t *q_end);

// ---- xoshiro128+ (the generator of the all-f32 mode, standard.rs with the
// f32 feature; rand_xoshiro 0.6.0): same structure on 4 x 32-bit words ------
void xoshiro128_seed(uint64_t seed, uint32_t s[4]); // seed_from_u64: SplitMix64 bytes, little-endian
uint32_t xoshiro128_next(uint32_t s[4]);
void xoshiro128_jump(uint32_t s[4]);                // the published 2^64 jump()

struct Gf2Mat128 {
    uint32_t col[128][4];
};
void gf2_128_identity(Gf2Mat128 &m);
void gf2_128_step(Gf2Mat128 &m);
void gf2_128_jump64(Gf2Mat128 &m);
void gf2_128_mul(const Gf2Mat128 &a, const Gf2Mat128 &b, Gf2Mat128 &out);
void gf2_128_apply(const Gf2Mat128 &m, const uint32_t in[4], uint32_t out[4]);
std::vector<Gf2Mat128> gf2_128_pow2_table(const Gf2Mat128 &base, int count);
void gf2_128_pow_from_table(const std::vector<Gf2Mat128> &table, uint64_t n, Gf2Mat128 &out);
const std::vector<Gf2Mat128> &xoshiro128_step_pow2(); // M^(2^k), k = 0..63 (cached)
void xoshiro128_advance(uint32_t s[4], uint64_t n);
// faster-evgen walk of the xoshiro128+ stream of the all-f32 mode, as
// xoshiro_walk_faster_evgen (the unit-disc test in f32).
void xoshiro128_walk_faster_evgen(const uint32_t s0[4], uint64_t q, uint64_t num_events, uint64_t *starts,
                                  uint64_t *q_end);

} // namespace tp
