"""faster-evgen: event start positions in the RANF stream depend on every
earlier draw (evgen.rs:143-173, 221-249: unit-disc re-rolls, and
random_array::<N> discarding a round's leftovers, ranf.rs:78-102).  Long
walks split the stream over host threads that start speculatively at round
boundaries and are joined where the true walk meets them; the result must
be the sequential walk's, bit for bit."""
import pytest


@pytest.fixture(scope="module")
def lib(walklib):
    return walklib


@pytest.mark.parametrize("n,seed,index,f32", [(4_000_000, 1234, 55, 0), (4_000_003, 98765, 17, 0), (5_500_000, 7, 0, 0),
                                               (4_000_001, 4321, 30, 1)])
def test_threaded_walk_equals_sequential(lib, n, seed, index, f32):
    # 8 threads, whatever the host has: the join logic is what is tested
    assert lib.tp_walk_mismatch(n, seed, index, 8, f32) == 0


@pytest.mark.parametrize("f32", [0, 1])
def test_ranf_walk_end_is_the_oracle_closure_advance(lib, oracle, oracle32, valeurs, f32):
    """The walk's end position for n events equals the stream advance of the
    oracle's closure over those n events (the f32 walk decides re-rolls on
    f32 values, as the reference's f32 build)."""
    import ctypes as C
    o = oracle32 if f32 else oracle
    ocfg = o.config(valeurs)
    opts = o.opts("ranf", faster_evgen=True)
    for n in (1, 7, 10000, 23456):
        g = o.rng("ranf")
        o.simulate_batch(ocfg, opts, g, n)
        h = o.rng("ranf")
        numbers = (C.c_int32 * 56)(*list(h.numbers))
        q = lib.tp_walk_end(numbers, h.index, n, 1, f32)
        o.skip(h, q - (55 - h.index))
        assert h.index == g.index and list(h.numbers)[1:] == list(g.numbers)[1:], (f32, n)


def test_xoshiro_walk_end_is_the_oracle_closure_advance(lib, oracle, valeurs):
    """faster-evgen on the xoshiro256+ stream (no discards): the host walk's
    end after n events equals the draws the oracle's closure consumed."""
    import ctypes as C
    ocfg = oracle.config(valeurs)
    opts = oracle.opts("xoshiro", faster_evgen=True)
    for n in (1, 7, 10000, 23456):
        g = oracle.rng("xoshiro")
        oracle.simulate_batch(ocfg, opts, g, n)
        h = oracle.rng("xoshiro")
        q = lib.tp_xo_walk_end((C.c_uint64 * 4)(*list(h.xs)), n)
        assert 15 * n <= q <= 40 * n
        oracle.skip(h, q)
        assert list(h.xs) == list(g.xs), n


@pytest.mark.parametrize("e0,n,seed,index", [(1, 1000, 1234, 55), (777, 50_000, 98765, 17), (123_457, 4_100_000, 7, 0)])
def test_walk_from_an_event_start(lib, e0, n, seed, index):
    """A long faster-evgen run is walked chunk by chunk, each chunk from the
    last start of the previous one (mid-round in general): the same starts
    and end position as the walk from the state itself."""
    assert lib.tp_walk_from_mismatch(e0, n, seed, index, 8) == 0


def test_xoshiro128_walk_end_is_the_f32_oracle_closure_advance(lib, oracle32, valeurs):
    """The same on the xoshiro128+ stream of the all-f32 mode (the unit-disc
    test in f32)."""
    import ctypes as C
    ocfg = oracle32.config(valeurs)
    opts = oracle32.opts("xoshiro", faster_evgen=True)
    for n in (1, 7, 10000, 23456):
        g = oracle32.rng("xoshiro")
        oracle32.simulate_batch(ocfg, opts, g, n)
        h = oracle32.rng("xoshiro")
        q = lib.tp_xo128_walk_end((C.c_uint32 * 4)(*list(h.xs32)), n)
        assert 15 * n <= q <= 40 * n
        oracle32.skip(h, q)
        assert list(h.xs32) == list(g.xs32), n
