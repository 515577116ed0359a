"""Closed-form pins of oracle functions that the reference's fixtures do not reach.

* energy_parts (the data / prior split of binary_energy, ising_recon.hpp:65-79):
  - at x = 0 the prior is 0 and the data term is sum_b w_b d_b^2 / 2;
  - for one set bit i the prior is lin(i) + lambda * |nbr(i)| (every bond of i unlike,
    each pair counted once);
  - with every bit set the prior is sum_i lin(i) (no unlike bond) and the data term is
    sum_b w_b (sum of a on ray b - d_b)^2 / 2;
  - data + prior = binary_energy, which the ree_cascade fixtures pin.
  A dropped 1/2, a flipped lambda sign or pairs counted twice each fail one of these.
* classify (ising_recon.hpp:302-309): p >= 0.9 NotThere, p <= 0.1 Exists, otherwise
  Undecided, compared in float.
* branch_fractions (ising_recon.hpp:148-166): branch b is anneal_branch with seed
  b + 1 + seed_offset whatever the thread count, and p_v is the float mean of the kept
  branches.
* tts99 (ising_weighted.hpp:271-275): infinite at p = 0, t at p >= 0.99, and
  t ln 0.01 / ln(1 - p) between -- at p = 1 - 10^-k that is 2 t / k.
"""
import math

import numpy as np
import pytest

from test_oracle_colour import random_ray_problem


@pytest.fixture(scope="module")
def prob(orc):
    return random_ray_problem(orc, n=49, n_rays=20, seed=5)


def _lin(prob):
    return np.full(prob.n_vars, prob.kappa)


def test_energy_parts_at_zero(orc, prob):
    data, prior = orc.energy_parts(prob, np.zeros(prob.n_vars, np.uint8))
    assert prior == 0.0
    assert data == pytest.approx(0.5 * float(np.sum(prob.w * prob.d ** 2)), rel=1e-14)


def test_energy_parts_single_bit(orc, prob):
    lam = prob.lam
    for i in (0, 7, 24, 48):
        x = np.zeros(prob.n_vars, np.uint8)
        x[i] = 1
        data, prior = orc.energy_parts(prob, x)
        deg = prob.nbr_ptr[i + 1] - prob.nbr_ptr[i]
        assert prior == pytest.approx(_lin(prob)[i] + lam * deg, rel=1e-14)
        # the data term of one set bit: only its rays move, by their a
        r = -prob.d.copy()
        for k in range(prob.row_ptr[i], prob.row_ptr[i + 1]):
            r[prob.ray[k]] += float(prob.a[k])
        assert data == pytest.approx(0.5 * float(np.sum(prob.w * r * r)), rel=1e-12)


def test_energy_parts_all_set(orc, prob):
    x = np.ones(prob.n_vars, np.uint8)
    data, prior = orc.energy_parts(prob, x)
    assert prior == pytest.approx(float(np.sum(_lin(prob))), rel=1e-14)
    ray_sum = np.bincount(prob.ray, weights=prob.a.astype(np.float64), minlength=prob.n_rays)
    assert data == pytest.approx(0.5 * float(np.sum(prob.w * (ray_sum - prob.d) ** 2)), rel=1e-12)


def test_energy_parts_sum_to_binary_energy(orc, prob):
    rng = np.random.default_rng(1)
    for _ in range(8):
        x = (rng.random(prob.n_vars) < 0.4).astype(np.uint8)
        data, prior = orc.energy_parts(prob, x)
        assert data + prior == pytest.approx(orc.binary_energy(prob, x), rel=1e-14)


def test_classify_thresholds(orc):
    L = orc.lib()
    f32 = np.float32
    cases = [(0.0, 0), (0.1, 0), (float(np.nextafter(f32(0.1), f32(1))), 2), (0.5, 2),
             (float(np.nextafter(f32(0.9), f32(0))), 2), (0.9, 1), (1.0, 1)]
    for p, want in cases:
        assert L.orc_classify(p, 0.1, 0.9) == want, p


def test_branch_fractions_is_mean_of_seeded_branches(orc, prob):
    sched = orc.schedule(3.0, 0.5, 6, 2)
    R, off = 5, 11
    p1, keep1 = orc.branch_fractions(prob, sched, R, 1, off)
    p3, keep3 = orc.branch_fractions(prob, sched, R, 3, off)
    assert np.array_equal(keep1, keep3) and np.array_equal(p1, p3)
    for b in range(R):
        assert np.array_equal(keep1[b], orc.anneal_branch(prob, sched, b + 1 + off))
    mean = (keep1.sum(axis=0).astype(np.float32) / np.float32(R)).astype(np.float32)
    assert np.array_equal(p1, mean)


def test_tts99_closed_form(orc):
    L = orc.lib()
    assert math.isinf(L.orc_tts99(3.0, 0.0))
    assert L.orc_tts99(3.0, 0.99) == 3.0 and L.orc_tts99(3.0, 1.0) == 3.0
    for k in (1, 2):
        p = 1.0 - 10.0 ** -k
        if p < 0.99:
            assert L.orc_tts99(3.0, p) == pytest.approx(2.0 * 3.0 / k, rel=1e-12)
    assert L.orc_tts99(1.0, 0.5) == pytest.approx(math.log(0.01) / math.log(0.5), rel=1e-15)
