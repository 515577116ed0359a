"""The allocation-free energy entry point (qb_energy_noalloc) against qb_energy, through the
C ABI on the GPU: two calls in a row, the second with fewer states, each bit-identical to
qb_energy on the same states; a call that needs more than the static scratch is refused
with QB_E_LIMIT and leaves the next call correct."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def states(n, K, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((K, n)) < rng.uniform(0.05, 0.6, size=(K, 1))).astype(np.uint8)


def test_noalloc_energy_equals_qb_energy_twice_then_smaller():
    P = qa.muon_scene(grid=40, sky_az=120, sky_th=24, log2_muons=22)
    prob = qa.Problem(P)
    op = orc_mod.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
    for K, seed in ((37, 1), (5, 2)):
        X = states(P.n_vars, K, seed)
        say(f"qb_energy_noalloc: {P.n_vars} bits, K {K}")
        got = prob.energy_noalloc(X)
        want = prob.energy(X)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        ref = np.array([orc_mod.binary_energy(op, x) for x in X])
        assert np.allclose(got[0], ref, rtol=1e-9, atol=0.0)
    # the full muon QUBO: a call within the scratch, then one beyond it
    F = qa.muon_scene()
    big = qa.Problem(F)
    X = states(F.n_vars, 12, 3)
    say(f"qb_energy_noalloc: {F.n_vars} bits, K 12")
    assert all(np.array_equal(g, w) for g, w in zip(big.energy_noalloc(X), big.energy(X)))
    Xbig = np.zeros((120, F.n_vars), np.uint8)
    tot = np.zeros(120)
    rc = qa.lib().qb_energy_noalloc(big.h, Xbig.ctypes.data, 120, tot.ctypes.data, None, None)
    assert rc == qa.QB_E_LIMIT
    assert b"scratch" in qa.lib().qb_last_error()
    X = states(P.n_vars, 3, 4)
    assert all(np.array_equal(g, w) for g, w in zip(prob.energy_noalloc(X), prob.energy(X)))
