"""GPU parity tests run on the CPU, against the SIMT build of the same HIP
sources (tests/simt: every GPU thread a fiber, wave intrinsics and block
barriers as the meeting points), under AddressSanitizer.

Not a GPU result: a host-side check that the kernels stay inside their
allocations and that workgroup-shared state is read before one thread
rewrites it.  The emulator runs each wave to its next meeting point before
the next wave starts, first to last and then last to first, which are the
two extreme schedules of the waves of one workgroup; a kernel that reads
shared data in one wave after another wave has already changed it fails
here deterministically (the GCMC deletion race was found this way).  The
parity bars are those of test_gpu_parity.py.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SUBSET = ("gcmc_trials or chunks or brownian_steps_with_overlap or "
          "ermak_steps_with_round_overlap or forces_fp64")


@pytest.fixture(scope="module")
def simt():
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    sys.path.insert(0, os.path.join(HERE, "simt"))
    import sim
    sim.build(asan=True)
    return sim


@pytest.mark.parametrize("order", ["forward", "reverse"])
def test_parity_subset_on_simt_build(simt, order):
    env = simt.env(asan=True)
    env.update(DML_SIMT="1", SIMT_ORDER=order)
    r = subprocess.run([sys.executable, "-u", "-m", "pytest", os.path.join(HERE, "test_gpu_parity.py"),
                        "-m", "gpu", "-q", "-x", "-s", "-p", "no:cacheprovider", "-k", SUBSET],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stdout + r.stderr, tail
