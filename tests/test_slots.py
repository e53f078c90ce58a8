"""The event loop keeps the photons where RAMBO makes them and applies the
reference's energy sort (evgen.rs:109-118) as per-photon slots computed from
pairwise comparisons (tpd::sort_slots).  This builds that function for the
host from the kernels' own header and compares it with the literal swap
network on every triple of energies from {0, -0, 1, 2, -1, 1e300, NaN}, where
ties and NaN make the swap sequence differ from a plain descending sort.  The
GPU events tests cover the same function on real events."""
import ctypes as C
import os
import subprocess

from conftest import PKG, ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sort_slots_reproduce_the_swap_network(tmp_path):
    out = tmp_path / "libslots.so"
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-shared",
                    "-fPIC", "-I" + os.path.join(PKG, "csrc"), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(HERE, "slots_check.cpp"), "-o", str(out)], check=True, capture_output=True)
    lib = C.CDLL(str(out))
    n = C.c_int()
    bad = lib.tp_slots_mismatches(C.byref(n))
    assert n.value == 7 ** 3
    assert bad == 0
