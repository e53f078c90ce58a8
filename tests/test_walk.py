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


@pytest.mark.parametrize("n,seed,index", [(4_000_000, 1234, 55), (4_000_003, 98765, 17), (5_500_000, 7, 0)])
def test_threaded_walk_equals_sequential(lib, n, seed, index):
    # 8 threads, whatever the host has: the join logic is what is tested
    assert lib.tp_walk_mismatch(n, seed, index, 8) == 0


@pytest.mark.parametrize("e0,n,seed,index", [(1, 1000, 1234, 55), (777, 50_000, 98765, 17), (123_457, 4_100_000, 7, 0)])
def test_walk_from_an_event_start(lib, e0, n, seed, index):
    """A long faster-evgen run is walked chunk by chunk, each chunk from the
    last start of the previous one (mid-round in general): the same starts
    and end position as the walk from the state itself."""
    assert lib.tp_walk_from_mismatch(e0, n, seed, index, 8) == 0
