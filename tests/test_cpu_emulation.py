"""The library's kernels compiled for the CPU (tests/cpu_emulation) under AddressSanitizer:
the real orchestration runs on small inputs against the oracle, so an out-of-bounds index
or a wrong buffer size fails here, on the CPU, before any GPU run.

The default runs every case but the two large ones (a few minutes); TD3_EMULATION=full
adds them (a 70k-particle tree and the row construction on many trees), quick runs two.
"""

import os
import subprocess
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_emulation")
QUICK = ["direct", "moments"]
DEFAULT = ["direct", "direct_noalloc", "strip", "smooth", "tree", "forest", "forest_resident", "moments", "walk", "fmm", "m2l_f32", "rows_overflow", "fmm_at",
           "coulomb", "push", "adjoint"]
FULL = DEFAULT + ["big_tree", "rows"]


@pytest.fixture(scope="module")
def emulated_library():
    r = subprocess.run(["bash", os.path.join(HERE, "build.sh")], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(f"CPU build of the kernels failed:\n{r.stderr[-4000:]}")
    return os.path.join(HERE, "libtd3_emu.so")


def test_kernels_under_address_sanitizer(emulated_library):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    mode = os.environ.get("TD3_EMULATION")
    cases = QUICK if mode == "quick" else FULL if mode == "full" else DEFAULT
    r = subprocess.run([sys.executable, os.path.join(HERE, "run_emulated.py"), *cases],
                       capture_output=True, text=True, env=env, timeout=3000)
    assert r.returncode == 0 and "EMULATION OK" in r.stdout, (r.stdout[-3000:] + r.stderr[-6000:])
