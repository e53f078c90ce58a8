"""Stream algebra of the product host (tp_rng_*) against the oracle's
step-by-step streams, and the oracle's RANF against Knuth's recurrence."""
import ctypes as C

import numpy as np
import pytest

import trois_photons as tp


def ranf_equal(st: tp.RanfState, g) -> bool:
    # The observable state is the unconsumed part of the round plus the
    # rounds that follow; compare the next 200 draws via the oracle.
    return list(st.numbers)[1:] == list(g.numbers)[1:] and st.index == g.index


def test_init_matches_oracle(oracle):
    g = oracle.rng("ranf")
    st = tp.rng_state(tp.RNG_RANF)
    assert ranf_equal(st, g) and st.seed == g.ranf_seed == 234612947
    g = oracle.rng("xoshiro")
    st = tp.rng_state(tp.RNG_XOSHIRO)
    assert list(st.s) == list(g.xs)


def test_ranf_is_knuth_recurrence(oracle):
    """Generation order: X_g = X_{g-55} - X_{g-24} mod 1e9; each round of 55 is
    emitted from index 55 down to 1 (ranf.rs:92-135)."""
    g = oracle.rng("ranf")
    draws = oracle.raw(g, 55 * 40).astype(np.int64)
    rounds = draws.reshape(40, 55)[:, ::-1].reshape(-1)  # generation order
    for i in range(55, len(rounds)):
        assert rounds[i] == (rounds[i - 55] - rounds[i - 24]) % 1_000_000_000


@pytest.mark.parametrize("n", [0, 1, 7, 54, 55, 56, 109, 110, 12 * 9999 + 5, 120000, 1_234_567])
def test_ranf_skip(oracle, n):
    for pre in (0, 13):
        g = oracle.rng("ranf")
        oracle.skip(g, pre)
        st = tp.rng_state(tp.RNG_RANF)
        tp.rng_skip(tp.RNG_RANF, st, pre)
        oracle.skip(g, n)
        tp.rng_skip(tp.RNG_RANF, st, n)
        assert st.index == g.index
        # next draws agree
        h = oracle.rng("ranf")
        for i in range(56):
            h.numbers[i] = st.numbers[i]
        h.index = st.index
        assert np.array_equal(oracle.raw(h, 200), oracle.raw(g, 200))


@pytest.mark.parametrize("n", [0, 1, 12, 255, 256, 120000, 12 * 10_000 * 37 + 11, 10_000_019])
def test_xoshiro_skip(oracle, n):
    g = oracle.rng("xoshiro")
    st = tp.rng_state(tp.RNG_XOSHIRO)
    oracle.skip(g, n)
    tp.rng_skip(tp.RNG_XOSHIRO, st, n)
    assert list(st.s) == list(g.xs)


def test_xoshiro_skip_composes():
    a = tp.rng_state(tp.RNG_XOSHIRO)
    b = tp.rng_state(tp.RNG_XOSHIRO)
    tp.rng_skip(tp.RNG_XOSHIRO, a, 3 * 2**40 + 5)
    for _ in range(3):
        tp.rng_skip(tp.RNG_XOSHIRO, b, 2**40)
    tp.rng_skip(tp.RNG_XOSHIRO, b, 5)
    assert list(a.s) == list(b.s)


def test_jumps_match_oracle(oracle):
    g = oracle.rng("xoshiro")
    st = tp.rng_state(tp.RNG_XOSHIRO)
    for _ in range(3):
        oracle.jump(g)
        tp.rng_jump(tp.RNG_XOSHIRO, st)
        assert list(st.s) == list(g.xs)
    g = oracle.rng("ranf")
    st = tp.rng_state(tp.RNG_RANF)
    for _ in range(3):
        oracle.jump(g)
        tp.rng_jump(tp.RNG_RANF, st)
        assert ranf_equal(st, g) and st.seed == g.ranf_seed


def test_seeded(oracle):
    for seed in (0, 42, 2**63 + 5):
        g = oracle.rng("xoshiro", seed)
        st = tp.rng_state(tp.RNG_XOSHIRO, seed)
        assert list(st.s) == list(g.xs)
    g = oracle.rng("ranf", 1234567)
    st = tp.rng_state(tp.RNG_RANF, 1234567)
    assert ranf_equal(st, g)


# ---- xoshiro128+ (the all-f32 mode's stream) ----------------------------------

def test_xoshiro128_seed_skip_jump_match_oracle(oracle32):
    """seed_from_u64 (SplitMix64 bytes), GF(2) skip-ahead = stepping, and the
    published 2^64 jump, against the f32 oracle's generator."""
    for seed in (None, 0, 42):
        g = oracle32.rng("xoshiro", seed)
        st = tp.rng_state(tp.RNG_XOSHIRO128, seed)
        assert list(st.s) == list(g.xs32)
    for n in (0, 1, 12, 127, 128, 120000, 12 * 10_000 * 37 + 11):
        g = oracle32.rng("xoshiro")
        st = tp.rng_state(tp.RNG_XOSHIRO128)
        oracle32.skip(g, n)
        tp.rng_skip(tp.RNG_XOSHIRO128, st, n)
        assert list(st.s) == list(g.xs32), n
    g = oracle32.rng("xoshiro")
    st = tp.rng_state(tp.RNG_XOSHIRO128)
    for _ in range(3):
        oracle32.jump(g)
        tp.rng_jump(tp.RNG_XOSHIRO128, st)
        assert list(st.s) == list(g.xs32)
