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
DEFAULT = ["forest_resident", "m2l_f32", "tree", "adjoint", "smooth", "fmm", "push", "forest",
           "strip", "rows_overflow", "walk", "coulomb", "fmm_at", "direct_noalloc", "direct",
           "moments"]
FULL = DEFAULT + ["big_tree", "rows"]


@pytest.fixture(scope="module")
def emulated_library():
    r = subprocess.run(["bash", os.path.join(HERE, "build.sh")], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(f"CPU build of the kernels failed:\n{r.stderr[-4000:]}")
    return os.path.join(HERE, "libtd3_emu.so")


# cases by measured cost, longest first, dealt to the worker processes in turn
WORKERS = 4


def _batches(cases):
    batches = [[] for _ in range(min(WORKERS, len(cases)))]
    for i, case in enumerate(cases):
        batches[i % len(batches)].append(case)
    return batches


def test_kernels_under_address_sanitizer(emulated_library):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    mode = os.environ.get("TD3_EMULATION")
    cases = QUICK if mode == "quick" else FULL if mode == "full" else DEFAULT
    # independent cases in a few processes at once (each process runs its cases in order)
    procs = [subprocess.Popen([sys.executable, os.path.join(HERE, "run_emulated.py"), *batch],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                              env=env)
             for batch in _batches(cases)]
    failures = []
    for batch, proc in zip(_batches(cases), procs):
        out, err = proc.communicate(timeout=3000)
        if proc.returncode != 0 or "EMULATION OK" not in out:
            failures.append(f"cases {batch}:\n{out[-3000:]}{err[-6000:]}")
    assert not failures, "\n\n".join(failures)
