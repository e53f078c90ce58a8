"""The event loop's f64 log and sin/cos (csrc/tp_math.h, built here for the
host from the same source the kernels compile) against an x87 long-double
reference and the host libm that the oracle, like the reference, calls
(evgen.rs:174-207).  Inputs follow the event loop: u1 u2 + MIN_POSITIVE and
fl(2 pi u) for u on the 2^-53 grid, plus the edges of the tables and of the
reductions."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

MIN_POSITIVE = 2.2250738585072014e-308
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("mathcheck") / "libmathcheck.so"
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-shared", "-fPIC",
                    "-I" + os.path.join(PKG, "csrc"), os.path.join(HERE, "math_check.cpp"), "-o", str(out)],
                   check=True)
    return C.CDLL(str(out))


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def uniforms(rng, n):
    return rng.integers(0, 2 ** 53, n, dtype=np.int64).astype(np.float64) * 2.0 ** -53


def eval_log(lib, x):
    x = np.ascontiguousarray(x, np.float64)
    ours, libm, err = (np.zeros_like(x) for _ in range(3))
    lib.tpm_eval_log(ptr(x), C.c_int64(len(x)), ptr(ours), ptr(libm), ptr(err))
    return ours, libm, err


def eval_sincos(lib, x):
    x = np.ascontiguousarray(x, np.float64)
    out = [np.zeros_like(x) for _ in range(6)]
    lib.tpm_eval_sincos(ptr(x), C.c_int64(len(x)), *map(ptr, out))
    return out


def log_edges():
    e = [MIN_POSITIVE, 2 * MIN_POSITIVE, 0.5, 0.25, 1.0, 1.5, 2.0, 1e300, 2.0 ** -1022 * 1.75,
         np.nextafter(1.0, 0.0), np.nextafter(0.5, 1.0), np.nextafter(1.0, 2.0), np.nextafter(2.0, 0.0)]
    for j in range(257):                     # every table interval, both ends
        m = 0.5 + j / 512
        for v in (m, np.nextafter(m, 0.0), np.nextafter(m, 1.0)):
            for k in (0, -1, -7, -60, -1000):
                e.append(np.ldexp(v, k))
    return np.array([v for v in e if MIN_POSITIVE <= v < np.inf])


def sincos_edges():
    two_pi = np.float64(2 * np.pi)
    e = [0.0, 5e-324, 1e-300, 1e-8, two_pi, np.nextafter(two_pi, 0.0), -1.0, -12.0, 12.0]
    for j in range(-255, 256):               # doubles next to every j pi/64, and the midpoints
        a = j * np.pi / 64
        for v in (a, np.nextafter(a, -np.inf), np.nextafter(a, np.inf), a + np.pi / 128, a - np.pi / 128):
            if abs(v) < 12.5:
                e.append(v)
    return np.array(e)


def test_log_is_correctly_rounded(lib):
    rng = np.random.default_rng(11)
    x = np.concatenate([uniforms(rng, 1_000_000) * uniforms(rng, 1_000_000) + MIN_POSITIVE, log_edges()])
    ours, libm, err = eval_log(lib, x)
    agree = float((ours == libm).mean())
    print(f"log: max |err| {np.abs(err).max():.4f} ulp; same bits as libm {agree:.6f}")
    assert np.abs(err).max() <= 0.502
    assert agree >= 0.999
    assert eval_log(lib, [1.0])[0][0] == 0.0
    assert eval_log(lib, [0.5])[0][0] == -np.log(2.0)


def test_sincos_is_correctly_rounded(lib):
    rng = np.random.default_rng(12)
    x = np.concatenate([np.float64(2 * np.pi) * uniforms(rng, 1_000_000), sincos_edges()])
    s, c, sl, cl, se, ce = eval_sincos(lib, x)
    print(f"sin: max |err| {np.abs(se).max():.4f} ulp, same bits as libm {(s == sl).mean():.6f}; "
          f"cos: max |err| {np.abs(ce).max():.4f} ulp, same bits as libm {(c == cl).mean():.6f}")
    assert np.abs(se).max() <= 0.51 and np.abs(ce).max() <= 0.51
    assert (s == sl).mean() >= 0.998 and (c == cl).mean() >= 0.998
    # the exact zeros and ones of the table
    s0, c0, *_ = eval_sincos(lib, [0.0, 5e-324])
    assert list(s0) == [0.0, 5e-324] and list(c0) == [1.0, 1.0]


def test_sincos_near_zeros(lib):
    """Next to pi/2, pi, 3 pi/2, 2 pi the result is the small reduced
    argument: the three-part pi/64 keeps its relative accuracy."""
    xs = []
    for a in (np.pi / 2, np.pi, 1.5 * np.pi, 2 * np.pi):
        v = a
        for _ in range(4):
            xs.append(v)
            v = np.nextafter(v, np.inf)
        v = a
        for _ in range(4):
            v = np.nextafter(v, -np.inf)
            xs.append(v)
    s, c, sl, cl, se, ce = eval_sincos(lib, np.array(xs))
    assert np.abs(se).max() <= 0.51 and np.abs(ce).max() <= 0.51
