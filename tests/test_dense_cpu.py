"""CPU checks of the dense field engine's oracle: qlog against the C library's log, and the
trajectory pinned by what the mathematics fixes (brute-force ground states, zero
temperature ends where no single flip lowers the energy, the FP32 carry equals FP64 where
every float operation is exact, and a recompute returns the fields from scratch)."""
import itertools
import math

import numpy as np

import qublar_amd as qa


def test_qlog_within_3_ulp_of_libm(orc):
    rng = np.random.default_rng(3)
    u = (rng.integers(1, 2**53, 200000, dtype=np.int64).astype(np.float64)) * 2.0**-53
    q = orc.lib().orc_qlog
    worst = 0
    for v in u[:20000]:
        a, b = q(float(v)), math.log(float(v))
        worst = max(worst, abs(np.float64(a).view(np.int64) - np.float64(b).view(np.int64)))
    assert worst <= 3
    assert q(1.0) == 0.0 and q(0.5) == math.log(0.5) and q(0.0) == -math.inf


def brute_min(Q, h):
    n = len(h)
    best = math.inf
    for bits in itertools.product([0, 1], repeat=n):
        x = np.array(bits, np.float64)
        best = min(best, float(h @ x + 0.5 * x @ Q.astype(np.float64) @ x))
    return best


def test_dense_oracle_energy_and_ground_state(orc):
    Q, h = qa.dense_qubo(12, 5)
    best = brute_min(Q, h)
    sched = orc.schedule(1.0, 0.005, 300, 50)
    got = min(orc.dense_energy(Q, h, orc.dense_anneal(Q, h, sched, s)) for s in range(1, 17))
    assert abs(got - best) < 1e-9


def test_dense_oracle_zero_temperature_local_minimum(orc):
    Q, h = qa.dense_qubo(200, 8)
    x = orc.dense_anneal(Q, h, orc.schedule(1e-12, 1e-12, 30, 0), 3)
    e0 = orc.dense_energy(Q, h, x)
    for i in range(0, 200, 3):
        y = x.copy(); y[i] ^= 1
        assert orc.dense_energy(Q, h, y) > e0 - 1e-12


def test_dense_fp32_carry_equals_fp64_when_exact(orc):
    """Dyadic couplings with few bits: every FP32 field update is exact, so the FP32 carry
    gives the FP64 trajectory; a dropped, doubled or mis-signed update would not."""
    rng = np.random.default_rng(2)
    n = 150
    A = (rng.integers(-64, 65, (n, n)) / 256.0).astype(np.float32)
    Q = np.triu(A, 1); Q = (Q + Q.T).astype(np.float32)
    h = (rng.integers(-64, 65, n) / 256.0).astype(np.float32)
    sched = orc.schedule(0.8, 0.02, 40, 5)
    for seed in (1, 2, 3):
        a, fa = orc.dense_anneal(Q, h, sched, seed, 0, 0, fields=True)
        b, fb = orc.dense_anneal(Q, h, sched, seed, 1, 0, fields=True)
        assert np.array_equal(a, b) and np.array_equal(fa, fb) and 0 < a.sum() < n


def test_dense_recompute_returns_scratch_fields(orc):
    Q, h = qa.dense_qubo(256, 4)
    x, f = orc.dense_anneal(Q, h, orc.schedule(1.0, 0.05, 6, 0), 9, 1, 3, fields=True)
    scratch = h.astype(np.float64) + Q.astype(np.float64) @ x.astype(np.float64)
    assert np.allclose(f, scratch.astype(np.float32), rtol=0, atol=1e-6)
    # without a recompute the FP32 carry has drifted from the scratch fields by rounding only
    x2, f2 = orc.dense_anneal(Q, h, orc.schedule(1.0, 0.05, 6, 0), 9, 1, 0, fields=True)
    s2 = h.astype(np.float64) + Q.astype(np.float64) @ x2.astype(np.float64)
    assert 0 < np.max(np.abs(f2 - s2)) < 1e-4


def test_dense_create_refuses_n_beyond_the_lds_limit():
    # One replica's FP64 fields and spins and two staged 64 x 64 window blocks must fit 150 KB of
    # LDS: 32768 + 9 n <= 153600, so n = 13,425 is the largest accepted. The size is refused
    # before Q is read, so a small buffer stands in for the matrix.
    import ctypes as C
    Q = np.zeros(16, np.float32)
    h = np.zeros(16, np.float32)
    out = C.c_void_p()
    rc = qa.lib().qb_dense_create(13426, Q.ctypes.data, h.ctypes.data, C.byref(out))
    assert rc == qa.QB_E_LIMIT and not out.value
    assert "150 KB" in qa.lib().qb_last_error().decode()
