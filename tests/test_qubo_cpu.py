"""CPU checks of the batched sparse-QUBO path: the host colouring and refusals of the
C ABI, the workload generator, and the oracle's trajectory for one instance pinned by
what the mathematics fixes (energy against a dense restatement, zero-temperature
descent ends where no single flip lowers the energy, brute-force ground states, order
freedom inside a class, and the replica keying)."""
import ctypes as C
import itertools

import numpy as np
import pytest

import qublar_amd as qa


def tiny_qubo(n, density, seed, replicas=8):
    """A dense-ish random QUBO small enough for brute force."""
    rng = np.random.default_rng(seed)
    rows, cols, ws = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < density:
                w = np.float32(rng.normal())
                rows += [i, j]; cols += [j, i]; ws += [w, w]
    order = np.lexsort((cols, rows))
    rows, cols, ws = np.array(rows)[order], np.array(cols)[order], np.array(ws, np.float32)[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    h = rng.normal(0, 0.5, n).astype(np.float32)
    return qa.QuboInstance(np.cumsum(row_ptr), cols, ws, h, replicas, seed)


def dense_energy(q, x):
    """E(x) = h.x + x^T U x with U the strict upper triangle of the coupling matrix."""
    Q = np.zeros((q.n, q.n))
    for i in range(q.n):
        for k in range(q.row_ptr[i], q.row_ptr[i + 1]):
            Q[i, q.col[k]] = q.J[k]
    U = np.triu(Q, 1)
    xf = np.asarray(x, np.float64)
    return float(q.h.astype(np.float64) @ xf + xf @ U @ xf)


@pytest.mark.parametrize("degree", [3, 8, 32, 64])
def test_generator_degree_and_symmetry(degree):
    q = qa.random_regular_qubo(3000, degree, 11)
    deg = np.diff(q.row_ptr)
    assert deg.max() <= degree and np.mean(deg == degree) > 0.95
    pairs = {}
    for i in range(q.n):
        for k in range(q.row_ptr[i], q.row_ptr[i + 1]):
            pairs[(i, int(q.col[k]))] = q.J[k]
    assert all(pairs[(j, i)] == w for (i, j), w in list(pairs.items())[:5000])


@pytest.mark.parametrize("degree", [3, 8, 32, 64])
def test_colouring_is_proper(degree):
    q = qa.random_regular_qubo(4000, degree, 5)
    order, cp = qa.colour_qubo(q)
    assert np.array_equal(np.sort(order), np.arange(q.n))
    col = np.empty(q.n, np.int64)
    for c in range(len(cp) - 1):
        col[order[cp[c]:cp[c + 1]]] = c
    i = np.repeat(np.arange(q.n), np.diff(q.row_ptr))
    assert np.all(col[i] != col[q.col])
    assert len(cp) - 1 <= degree + 1


def test_refusals():
    L = qa.lib()

    def colour(row_ptr, col, J, n):
        rp = np.asarray(row_ptr, np.int32); cl = np.asarray(col, np.int32); Jv = np.asarray(J, np.float32)
        o = np.zeros(n, np.int32); cp = np.zeros(n + 1, np.int32); nc = C.c_int()
        rc = L.qb_colour_qubo(n, qa._p(rp), qa._p(cl), qa._p(Jv), qa._p(o), qa._p(cp), C.byref(nc))
        return rc, L.qb_last_error().decode()

    assert colour([0, 1, 2], [1, 0], [0.5, 0.5], 2)[0] == qa.QB_OK
    rc, msg = colour([0, 1, 2], [1, 0], [0.5, 0.25], 2)
    assert rc == qa.QB_E_ARG and "symmetric" in msg
    rc, msg = colour([0, 1, 1], [0], [0.5], 2)
    assert rc == qa.QB_E_ARG and "itself" in msg
    rc, msg = colour([0, 1, 2], [5, 0], [0.5, 0.5], 2)
    assert rc == qa.QB_E_ARG and "column" in msg
    rc, msg = colour([0, 2, 4], [1, 1, 0, 0], [0.5, 0.5, 0.5, 0.5], 2)
    assert rc == qa.QB_E_ARG and "twice" in msg


def test_oracle_energy_matches_dense_form(orc):
    q = tiny_qubo(14, 0.4, 3)
    rng = np.random.default_rng(0)
    for _ in range(50):
        x = rng.integers(0, 2, q.n).astype(np.uint8)
        assert abs(orc.qubo_energy(q.row_ptr, q.col, q.J, q.h, x) - dense_energy(q, x)) < 1e-9


def test_oracle_zero_temperature_ends_in_a_local_minimum(orc):
    """At T -> 0 a flip is taken iff dE <= 0; when a sweep changes nothing, every single
    flip must raise the energy as the energy function computes it."""
    q = qa.random_regular_qubo(600, 8, 9, 4)
    order, cp = qa.colour_qubo(q)
    sched = orc.schedule(1e-12, 1e-12, 60, 0)
    X = orc.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, sched, 7, 4)
    for x in X:
        e0 = orc.qubo_energy(q.row_ptr, q.col, q.J, q.h, x)
        for i in range(0, q.n, 7):
            y = x.copy(); y[i] ^= 1
            assert orc.qubo_energy(q.row_ptr, q.col, q.J, q.h, y) > e0 - 1e-12


def test_oracle_finds_brute_force_ground_state(orc):
    q = tiny_qubo(12, 0.5, 21)
    best = min(dense_energy(q, x) for x in itertools.product([0, 1], repeat=q.n))
    order, cp = qa.colour_qubo(q)
    X = orc.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, orc.schedule(4.0, 0.01, 400, 50), 5, 16)
    got = min(orc.qubo_energy(q.row_ptr, q.col, q.J, q.h, x) for x in X)
    assert abs(got - best) < 1e-9


def test_oracle_order_free_inside_classes_and_replicas_independent(orc):
    q = qa.random_regular_qubo(800, 32, 4, 3)
    order, cp = qa.colour_qubo(q)
    rng = np.random.default_rng(1)
    perm = order.copy()
    for c in range(len(cp) - 1):
        perm[cp[c]:cp[c + 1]] = rng.permutation(perm[cp[c]:cp[c + 1]])
    sched = orc.schedule(3.0, 0.1, 20, 2)
    a = orc.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, sched, 40, 3)
    b = orc.qubo_anneal(q.row_ptr, q.col, q.J, q.h, perm, sched, 40, 3)
    assert np.array_equal(a, b) and 0 < a.sum() < a.size
    # replicas are independent: the first ones do not depend on how many follow, and the
    # two replicas of a draw pair (one Philox call) follow different trajectories
    c = orc.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, sched, 40, 5)
    assert np.array_equal(a, c[:3])
    assert not np.array_equal(c[0], c[1]) and not np.array_equal(c[2], c[3])
