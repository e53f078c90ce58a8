"""faster-evgen: event start positions in the RANF stream depend on every
earlier draw (evgen.rs:143-173, 221-249: unit-disc re-rolls, and
random_array::<N> discarding a round's leftovers, ranf.rs:77-104).  Long
walks split the stream over host threads that start speculatively at round
boundaries and are joined where the true walk meets them; the result must
be the sequential walk's, bit for bit."""
import ctypes as C
import os
import subprocess

import pytest

from conftest import PKG

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("walk") / "libwalk.so"
    csrc = os.path.join(PKG, "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-shared", "-fPIC", "-I" + csrc,
                    os.path.join(HERE, "walk_check.cpp"), os.path.join(csrc, "tp_rng_host.cpp"), "-o", str(out)],
                   check=True)
    lib = C.CDLL(str(out))
    lib.tp_walk_mismatch.argtypes = [C.c_uint64, C.c_int, C.c_uint32, C.c_int]
    return lib


@pytest.mark.parametrize("n,seed,index", [(4_000_000, 1234, 55), (4_000_003, 98765, 17), (5_500_000, 7, 0)])
def test_threaded_walk_equals_sequential(lib, n, seed, index):
    # 8 threads, whatever the host has: the join logic is what is tested
    assert lib.tp_walk_mismatch(n, seed, index, 8) == 0
