"""The f32 adjoint kernel stages each workgroup's output rows in LDS and stores
them with consecutive lanes on consecutive addresses.  This builds the
store-index functions the kernel uses (3photons-rust_amd/csrc/tp_adjstore.h)
for the host and checks, for every point count n up to five workgroups and
seven points, that every lane's stage slot stays inside the stage and every
stored element lands inside the rows of the workgroup's own points and is
written exactly once -- lanes past n included."""
import os
import subprocess

from conftest import PKG

HERE = os.path.dirname(os.path.abspath(__file__))


def test_staged_store_indices_stay_in_their_rows(tmp_path):
    exe = tmp_path / "adjstore_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
                    "-I" + os.path.join(PKG, "csrc"), os.path.join(HERE, "adjstore_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
