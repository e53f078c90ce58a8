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
    emitted from index 55 down to 1 (ranf.rs:95-96,106-119)."""
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


# ---- jumps pinned by the step function alone --------------------------------
# The published jump() of xoshiro256+ advances the state as 2^128 calls of
# next(), and xoshiro128+'s as 2^64 (Blackman & Vigna, "Scrambled linear
# pseudorandom number generators", 2018).  Both generators are linear over
# GF(2), so the jump is the step matrix raised to that power: built here
# from the state update alone (no jump constants) by repeated squaring.

def _gf2_step_matrix(words: int, bits: int, a: int, rot: int) -> np.ndarray:
    """Column j = one state update of the unit state with bit j set; the
    update is t = s1 << a; s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t;
    s3 = rotl(s3, rot)."""
    mask = (1 << bits) - 1
    n = words * bits
    m = np.zeros((n, n), np.uint8)
    for j in range(n):
        s = [0] * 4
        s[j // bits] = 1 << (j % bits)
        t = (s[1] << a) & mask
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << rot) | (s[3] >> (bits - rot))) & mask
        for w in range(4):
            for b in range(bits):
                m[w * bits + b, j] = (s[w] >> b) & 1
    return m


def _gf2_pow2(m: np.ndarray, k: int) -> np.ndarray:
    f = m.astype(np.float64)
    for _ in range(k):
        f = np.mod(f @ f, 2.0)  # entries are sums of at most 256 ones: exact
    return f.astype(np.uint8)


def _apply(m: np.ndarray, state, bits: int):
    v = np.array([(state[w] >> b) & 1 for w in range(4) for b in range(bits)], np.uint8)
    r = (m.astype(np.int64) @ v.astype(np.int64)) % 2
    return [sum(int(r[w * bits + b]) << b for b in range(bits)) for w in range(4)]


def test_step_matrix_is_the_generator(oracle, oracle32):
    """The matrices above step the oracle's generators (so the jump checks
    below rest on the published state update, not on a table)."""
    for bits, a, rot, orc, kind, field in ((64, 17, 45, oracle, tp.RNG_XOSHIRO, "xs"),
                                            (32, 9, 11, oracle32, tp.RNG_XOSHIRO128, "xs32")):
        m = _gf2_step_matrix(4, bits, a, rot)
        g = orc.rng("xoshiro")
        s0 = list(getattr(g, field))
        orc.skip(g, 1)
        assert _apply(m, s0, bits) == list(getattr(g, field))


def test_xoshiro256_jump_is_two_to_the_128_steps(oracle):
    m = _gf2_pow2(_gf2_step_matrix(4, 64, 17, 45), 128)
    st = tp.rng_state(tp.RNG_XOSHIRO)
    g = oracle.rng("xoshiro")
    expect = _apply(m, list(st.s), 64)
    tp.rng_jump(tp.RNG_XOSHIRO, st)
    oracle.jump(g)
    assert list(st.s) == expect and list(g.xs) == expect


def test_xoshiro128_jump_is_two_to_the_64_steps(oracle32):
    """Also two GF(2) skips of 2^63 through tp_rng_skip."""
    m = _gf2_pow2(_gf2_step_matrix(4, 32, 9, 11), 64)
    st = tp.rng_state(tp.RNG_XOSHIRO128)
    expect = _apply(m, list(st.s), 32)
    sk = tp.rng_state(tp.RNG_XOSHIRO128)
    tp.rng_skip(tp.RNG_XOSHIRO128, sk, 2 ** 63)
    tp.rng_skip(tp.RNG_XOSHIRO128, sk, 2 ** 63)
    g = oracle32.rng("xoshiro")
    oracle32.jump(g)
    tp.rng_jump(tp.RNG_XOSHIRO128, st)
    assert list(st.s) == expect and list(sk.s) == expect and list(g.xs32) == expect
