"""The exact region-of-interest posterior (tools/branches_blaze.py:45-53, 80-81 over
roi_qubo's QUBO, src/ising_recon.hpp:245-300) in the oracle, pinned by closed forms:
one bit, independent bits (J = 0: a product of logistic marginals, log Z a sum of
softplus terms), two coupled bits written out, beta = 0 (uniform), a large beta (the
posterior collapses onto the brute-force minimum); and the C ABI's refusals."""
import ctypes as C

import numpy as np
import pytest

import qublar_amd as qa


def softplus(x):
    return np.logaddexp(0.0, x)


def test_one_bit(orc):
    for h, beta in ((0.7, 1.0), (-2.5, 0.3), (0.0, 5.0)):
        r = orc.roi_enumerate(np.array([h]), np.zeros((1, 1)), beta)
        assert abs(r["marginal"][0] - 1.0 / (1.0 + np.exp(beta * h))) < 1e-15
        assert abs(r["log_z"] - softplus(-beta * h)) < 1e-14
        assert r["e_min"] == min(0.0, h) and r["map_index"] == (1 if h < 0 else 0)


def test_independent_bits(orc):
    rng = np.random.default_rng(3)
    h = rng.normal(0, 2, 12)
    for beta in (0.5, 1.0, 2.0):
        r = orc.roi_enumerate(h, np.zeros((12, 12)), beta)
        assert np.allclose(r["marginal"], 1.0 / (1.0 + np.exp(beta * h)), rtol=0, atol=1e-14)
        assert abs(r["log_z"] - softplus(-beta * h).sum()) < 1e-12
        assert abs(r["e_min"] - np.minimum(h, 0).sum()) < 1e-12
        assert r["map_index"] == sum(1 << i for i in range(12) if h[i] < 0)


def test_two_coupled_bits(orc):
    h0, h1, j, beta = 0.4, -0.9, 1.7, 1.3
    J = np.array([[0.0, j], [99.0, 0.0]])     # the lower triangle is not read
    r = orc.roi_enumerate(np.array([h0, h1]), J, beta)
    w = np.exp(-beta * np.array([0.0, h0, h1, h0 + h1 + j]))     # states 00, 10, 01, 11 (bit 0 first)
    Z = w.sum()
    assert abs(r["marginal"][0] - (w[1] + w[3]) / Z) < 1e-15
    assert abs(r["marginal"][1] - (w[2] + w[3]) / Z) < 1e-15
    assert abs(r["log_z"] - np.log(Z)) < 1e-14
    assert r["e_min"] == h1 and r["map_index"] == 2


def test_uniform_and_frozen_limits(orc):
    rng = np.random.default_rng(8)
    n = 11
    h = rng.normal(0, 1, n)
    J = np.triu(rng.normal(0, 1, (n, n)), 1)
    r0 = orc.roi_enumerate(h, J, 0.0)
    assert np.allclose(r0["marginal"], 0.5, atol=1e-15) and abs(r0["log_z"] - n * np.log(2)) < 1e-12
    # brute force of the minimum, independent of the oracle's loops
    X = (np.arange(1 << n)[:, None] >> np.arange(n)) & 1
    E = X @ h + np.einsum("ki,ij,kj->k", X, J, X)
    best = int(np.argmin(E))
    r = orc.roi_enumerate(h, J, 1e4)
    assert abs(r["e_min"] - E[best]) < 1e-12 and r["map_index"] == best
    assert np.allclose(r["marginal"], X[best], atol=1e-12)


def test_capi_refusals():
    L = qa.lib()
    res = qa.RoiResult()
    h = np.zeros(40)
    J = np.zeros(40 * 40)
    for n, beta in ((0, 1.0), (37, 1.0), (3, -1.0), (3, float("nan")), (3, float("inf"))):
        rc = L.qb_roi_enumerate(n, qa._p(h), qa._p(J), beta, None, None, C.byref(res))
        assert rc == qa.QB_E_ARG, (n, beta)
    h[1] = float("nan")
    assert L.qb_roi_enumerate(3, qa._p(h), qa._p(J), 1.0, None, None, C.byref(res)) == qa.QB_E_ARG
    assert "not finite" in L.qb_last_error().decode()
