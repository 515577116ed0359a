"""GPU parity of the ray engine on the gravimetry QUBO (stations as rows, every bit in
~200 of them, the builder's layer-lattice colouring): spins, p_void, tri-state maps
bit-exact against the oracle with the same classes and keys; FP64 energies within 1e-9
relative. The full 200,000-bit problem runs a short anneal, a small scene a long one."""
import numpy as np

import qublar_amd as qa
from test_gpu_parity import check_branches, gpu, say  # noqa: F401  (gpu is a fixture)

import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(nx=24, ny=20, nz=6, h=4.0, radius=14.0, body_lo=(30.0, 24.0, -16.0), body_hi=(60.0, 52.0, -4.0),
             sigma=3.0)


def test_gravity_small_fp64(gpu):
    P = qa.gravity_scene(**SMALL)
    _, out = check_branches(P, qa.schedule(50.0, 1.0, 40, 10), 16)
    assert 0 < out["branches"].sum() < out["branches"].size
    say(f"kernel_ms {out['report']['kernel_ms']:.2f}")


def test_gravity_small_fp32_recompute(gpu):
    P = qa.gravity_scene(**SMALL)
    _, out = check_branches(P, qa.schedule(50.0, 1.0, 40, 10), 12, seed_offset=3, precision=qa.FP32,
                            recompute_every=5)
    assert out["report"]["fp32_residual_err"] < 1e-3


def test_gravity_full_scene(gpu):
    """The precision study's gravity QUBO: 200,000 bits, 20,000 stations, 39.1M entries,
    2,586 classes; 16 branches for 8 + 2 sweeps, FP32 with recompute every 4."""
    P = qa.gravity_scene()
    _, out = check_branches(P, qa.schedule(50.0, 1.0, 8, 2), 16, precision=qa.FP32, recompute_every=4)
    say(f"kernel_ms {out['report']['kernel_ms']:.2f}")
    assert np.isfinite(out["energy"]).all()
