"""The reference-style driver built on the drop-in headers, run on the GPU: device
energies against host energies, run_batch against anneal seed for seed (and across
launch splits and state placement), and the 16-variable ground state reached through
branch_fractions (tools/gpu_anneal.cu --check, src/check_ising.cu A)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "check_dropin")


@pytest.mark.gpu
def test_dropin_driver_on_gpu():
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    out = subprocess.run([DRIVER], capture_output=True, text=True, timeout=240)
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("PASS") >= 12 and "FAIL" not in out.stdout
