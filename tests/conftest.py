"""Shared test setup: markers, import paths, build-on-demand, cached runs."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "3photons-rust_amd")

# faster-evgen runs go through the engine in chunks of 12 288 batches of
# 10 000 events (tp_stats.chunks reports how many): 1.24e8 events are two
# chunks, 12 288 + 112 batches.  Their first two batches, the two at the
# chunk edge and the last.
FE_TWO_CHUNKS = 124_000_000
FE_EDGE_BATCHES = (0, 1, 12287, 12288, 12399)
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    """Build the oracle and the engine if nothing is built yet."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "build", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    if not os.path.exists(os.path.join(PKG, "build", "libtp.so")):
        subprocess.run(["make", "-C", PKG, "-j8"], check=True, capture_output=True)


_ensure_built()


def golden_path(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def walklib(tmp_path_factory):
    """tests/walk_check.cpp with the host faster-evgen walk
    (3photons-rust_amd/csrc/tp_rng_host.cpp), built for the host."""
    import ctypes as C
    out = tmp_path_factory.mktemp("walk") / "libwalk.so"
    csrc = os.path.join(PKG, "csrc")
    here = os.path.dirname(os.path.abspath(__file__))
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-shared", "-fPIC", "-I" + csrc,
                    os.path.join(here, "walk_check.cpp"), os.path.join(csrc, "tp_rng_host.cpp"), "-o", str(out)],
                   check=True)
    lib = C.CDLL(str(out))
    lib.tp_walk_mismatch.argtypes = [C.c_uint64, C.c_int, C.c_uint32, C.c_int, C.c_int]
    lib.tp_walk_end.argtypes = [C.POINTER(C.c_int32), C.c_uint32, C.c_uint64, C.c_int, C.c_int]
    lib.tp_xo_walk_end.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
    lib.tp_xo_walk_end.restype = C.c_uint64
    lib.tp_xo128_walk_end.argtypes = [C.POINTER(C.c_uint32), C.c_uint64]
    lib.tp_xo128_walk_end.restype = C.c_uint64
    lib.tp_walk_end.restype = C.c_uint64
    lib.tp_walk_from_mismatch.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_uint32, C.c_int]
    return lib


@pytest.fixture(scope="session")
def oracle():
    from oracle_ffi import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle32():
    from oracle_ffi import Oracle
    return Oracle(f32=True)


@pytest.fixture(scope="session")
def valeurs():
    return golden_path("valeurs")


_RUN_CACHE: dict = {}


@pytest.fixture(scope="session")
def oracle_run(oracle, valeurs):
    """Cached full oracle runs at the shipped configuration (N = 1e7)."""
    def run(rng="ranf", faster_threading=False, sort=True, faster_evgen=False, n=None, per_batch=False):
        key = (rng, faster_threading, sort, faster_evgen, n, per_batch)
        if key not in _RUN_CACHE:
            cfg = oracle.config(valeurs, n)
            opts = oracle.opts(rng, faster_threading, sort, threads=os.cpu_count() or 8, faster_evgen=faster_evgen)
            merged, per = oracle.run(cfg, opts, per_batch)
            _RUN_CACHE[key] = (cfg, merged, per, oracle.finalize(cfg, merged))
        return _RUN_CACHE[key]
    return run
