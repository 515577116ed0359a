"""Pins the oracle's colour-class / Philox trajectory (the one the MI355X engine
reproduces), which has no counterpart in the reference, by invariants:

* permuting bits inside a colour class changes nothing (a class is independent);
* the integer and float paths make the same decisions on a +1 graph with
  lambda = -1 (gset_bench.cu:224-248, restated for the colour trajectory);
* the FP32 carry equals the FP64 carry where every FP32 operation is exact;
* a from-scratch recompute after the last sweep returns float(scratch residuals);
* local_field_dE = E(flip) - E(x) (check_local_field.cu:362-378);
* the ROI QUBO expansion is exact (check_ising.cu:536-559);
* frozen bits satisfy the freeze bound in every sampled context;
* where every ray is private to one bit the bound is the exact minimum of dE(0 -> 1)
  over contexts (brute force over all of them), and the frozen set equals the one built
  round by round from those minima, including a case that needs more than one round.
"""
import numpy as np
import pytest

from test_oracle_pins import er_graph


def greedy_colouring(n, conflicts):
    """Smallest-free-colour greedy in index order; conflicts[i] = set of bits."""
    colour = np.full(n, -1)
    for i in range(n):
        used = {colour[j] for j in conflicts[i] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    k = colour.max() + 1
    order = np.concatenate([np.flatnonzero(colour == c) for c in range(k)]).astype(np.int32)
    class_ptr = np.concatenate([[0], np.cumsum([np.sum(colour == c) for c in range(k)])]).astype(np.int32)
    return order, class_ptr


def ray_conflicts(prob):
    n = prob.n_vars
    by_ray = [[] for _ in range(prob.n_rays)]
    for i in range(n):
        for k in range(prob.row_ptr[i], prob.row_ptr[i + 1]):
            by_ray[prob.ray[k]].append(i)
    conf = [set() for _ in range(n)]
    for members in by_ray:
        for i in members:
            conf[i].update(members)
    for i in range(n):
        conf[i].update(prob.nbr[prob.nbr_ptr[i]:prob.nbr_ptr[i + 1]].tolist())
        conf[i].discard(i)
    return conf


def random_ray_problem(orc, n=60, n_rays=25, seed=0, exact=False, lam=0.8, kappa=0.3, grid=True):
    rng = np.random.default_rng(seed)
    row_ptr, ray, a = [0], [], []
    for _ in range(n):
        rays = np.sort(rng.choice(n_rays, size=rng.integers(1, 6), replace=False))
        for b in rays:
            ray.append(b)
            a.append(rng.integers(1, 64) / 256.0 if exact else rng.uniform(0.01, 0.3))
        row_ptr.append(len(ray))
    d = (rng.integers(0, 64, n_rays) / 64.0) if exact else rng.uniform(0.0, 0.8, n_rays)
    w = (rng.integers(1, 8, n_rays) * 1.0) if exact else rng.uniform(5.0, 40.0, n_rays)
    nbr_ptr, nbr = [0], []
    side = int(np.ceil(np.sqrt(n)))
    for i in range(n):
        if grid:
            x, y = i % side, i // side
            for j in (i - 1 if x > 0 else -1, i + 1 if x + 1 < side else -1, i - side, i + side):
                if 0 <= j < n and j != i:
                    nbr.append(j)
        nbr_ptr.append(len(nbr))
    return orc.FloatProblem(row_ptr, ray, a, d, w, nbr_ptr, nbr, lam, kappa)


def permute_within_classes(order, class_ptr, seed):
    rng = np.random.default_rng(seed)
    out = order.copy()
    for c in range(len(class_ptr) - 1):
        seg = out[class_ptr[c]:class_ptr[c + 1]]
        out[class_ptr[c]:class_ptr[c + 1]] = rng.permutation(seg)
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_float_class_permutation_invariance(orc, precision):
    prob = random_ray_problem(orc, seed=1)
    order, cp = greedy_colouring(prob.n_vars, ray_conflicts(prob))
    sched = orc.schedule(6.0, 0.2, 60, 10)
    m0 = orc.make_mode(orc.RNG_PHILOX, order, cp, precision, 7)
    m1 = orc.make_mode(orc.RNG_PHILOX, permute_within_classes(order, cp, 3), cp, precision, 7)
    for seed in (1, 2, 3):
        x0, r0 = orc.anneal_branch(prob, sched, seed, m0, residuals=True)
        x1, r1 = orc.anneal_branch(prob, sched, seed, m1, residuals=True)
        assert np.array_equal(x0, x1) and np.array_equal(r0, r1)
        assert 0 < x0.sum() < prob.n_vars   # the run is not trivial


def int_conflicts(g):
    conf = [set() for _ in range(g.n)]
    for u, v in zip(g.eu, g.ev):
        conf[u].add(v)
        conf[v].add(u)
    return conf


def test_int_class_permutation_invariance(orc):
    g = er_graph(orc, 120, 0.1, 5)
    order, cp = greedy_colouring(g.n, int_conflicts(g))
    sched = orc.schedule(3.0, 0.05, 200, 0)
    for seed in (1, 9):
        a = orc.anneal_int(g, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp))
        b = orc.anneal_int(g, sched, seed, orc.make_mode(orc.RNG_PHILOX, permute_within_classes(order, cp, 4), cp))
        assert np.array_equal(a[0], b[0]) and a[1] == b[1]
        assert a[1] == -orc.lib().orc_cut_of(g.c, a[0])   # tracked energy = -cut


def plus_one_pair(orc, n=150, m=900, seed=42):
    """A +1 graph as WeightedGraph and as the lambda = -1 BinaryProblem of
    gset_bench.cu:48-70 (neighbours in edge-insertion order)."""
    rng = np.random.default_rng(seed)
    seen, eu, ev = set(), [], []
    while len(eu) < m:
        u, v = rng.integers(0, n, 2)
        if u == v:
            continue
        a, b = min(u, v), max(u, v)
        if (a, b) in seen:
            continue
        seen.add((a, b)); eu.append(a); ev.append(b)
    g = orc.IntGraph(n, eu, ev, np.ones(m, np.int32))
    adj = [[] for _ in range(n)]
    for a, b in zip(eu, ev):
        adj[a].append(b); adj[b].append(a)
    nbr_ptr, nbr = [0], []
    for i in range(n):
        nbr.extend(adj[i]); nbr_ptr.append(len(nbr))
    prob = orc.FloatProblem(np.zeros(n + 1, np.int32), [], [], [], [], nbr_ptr, nbr, -1.0, 0.0)
    return g, prob


def test_int_equals_float_on_plus_one_graph(orc):
    """Both trajectories: the integer engine and anneal_branch agree decision for decision."""
    g, prob = plus_one_pair(orc)
    order, cp = greedy_colouring(g.n, int_conflicts(g))
    sched = orc.schedule(5.0, 1e-3, 150, 20)
    for mode in (orc.make_mode(), orc.make_mode(orc.RNG_PHILOX, order, cp)):
        for seed in (1, 2, 3, 4):
            xi, ei = orc.anneal_int(g, sched, seed, mode)
            xf = orc.anneal_branch(prob, sched, seed, mode)
            assert np.array_equal(xi, xf)
            assert float(ei) == orc.binary_energy(prob, xf)


def test_fp32_carry_equals_fp64_when_exact(orc):
    """With a, d on a 2^-8 grid and few flips, every FP32 residual update is exact,
    so the FP32 carry must reproduce the FP64 trajectory exactly."""
    prob = random_ray_problem(orc, seed=7, exact=True)
    order, cp = greedy_colouring(prob.n_vars, ray_conflicts(prob))
    sched = orc.schedule(4.0, 0.3, 50, 5)
    for seed in (1, 2, 3, 4, 5):
        x64, r64 = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, 0), residuals=True)
        x32, r32 = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, 1), residuals=True)
        assert np.array_equal(x64, x32) and np.array_equal(r64, r32)


def scratch_residuals(prob, x):
    r = -prob.d.copy()
    for i in range(prob.n_vars):
        if x[i]:
            for k in range(prob.row_ptr[i], prob.row_ptr[i + 1]):
                r[prob.ray[k]] += float(prob.a[k])
    return r


def test_recompute_returns_scratch_residuals(orc):
    prob = random_ray_problem(orc, seed=3)
    order, cp = greedy_colouring(prob.n_vars, ray_conflicts(prob))
    sched = orc.schedule(6.0, 0.5, 30, 0)
    x, r = orc.anneal_branch(prob, sched, 5, orc.make_mode(orc.RNG_PHILOX, order, cp, 1, 1), residuals=True)
    assert np.array_equal(r, scratch_residuals(prob, x).astype(np.float32).astype(np.float64))
    x, r = orc.anneal_branch(prob, sched, 5, orc.make_mode(orc.RNG_PHILOX, order, cp, 0, 0), residuals=True)
    assert np.allclose(r, scratch_residuals(prob, x), rtol=0, atol=1e-12)


def test_local_field_dE_is_energy_difference(orc):
    prob = random_ray_problem(orc, seed=9)
    rng = np.random.default_rng(2)
    for _ in range(20):
        x = rng.integers(0, 2, prob.n_vars).astype(np.uint8)
        r = scratch_residuals(prob, x)
        e0 = orc.binary_energy(prob, x)
        for i in rng.choice(prob.n_vars, 5, replace=False):
            y = x.copy(); y[i] ^= 1
            de = orc.lib().orc_local_field_dE(prob.c, x, r, int(i))
            assert abs(de - (orc.binary_energy(prob, y) - e0)) <= 1e-9 * max(1.0, abs(e0))


def test_roi_qubo_expansion_exact(orc):
    prob = random_ray_problem(orc, seed=4, n=40)
    rng = np.random.default_rng(8)
    x_bar = rng.integers(0, 2, prob.n_vars).astype(np.uint8)
    roi = np.sort(rng.choice(prob.n_vars, 10, replace=False)).astype(np.int32)
    n = len(roi)
    h = np.zeros(n); J = np.zeros(n * n)
    orc.lib().orc_roi_qubo(prob.c, roi, n, x_bar, h, J)
    J = J.reshape(n, n)
    base = x_bar.copy(); base[roi] = 0
    e0 = orc.binary_energy(prob, base)
    for _ in range(16):
        xr = rng.integers(0, 2, n)
        full = base.copy(); full[roi] = xr
        pred = h @ xr + xr @ np.triu(J, 1) @ xr
        assert abs(orc.binary_energy(prob, full) - e0 - pred) <= 1e-9 * max(1.0, abs(e0))


def test_freeze_bound_holds(orc):
    prob = random_ray_problem(orc, seed=12, n=80, kappa=2.5)
    frozen = np.zeros(prob.n_vars, np.uint8)
    rounds = orc.lib().orc_freeze_mask(prob.c, 1.0, frozen)
    assert rounds >= 1 and frozen.sum() > 0
    rng = np.random.default_rng(1)
    for _ in range(40):
        x = rng.integers(0, 2, prob.n_vars).astype(np.uint8)
        x[frozen == 1] = 0
        r = scratch_residuals(prob, x)
        for i in np.flatnonzero(frozen):
            assert orc.lib().orc_local_field_dE(prob.c, x, r, int(i)) >= 1.0 - 1e-9


def private_ray_problem(orc, n, seed, lam=1.0, edges=None):
    """n bits, each on 1-3 rays no other bit crosses, a per-bit field, and a random
    neighbour graph (or `edges`). With private rays a bit's data term does not depend on
    the other bits, so freeze_provable_rock's bound (src/ising_recon.hpp:189-202) is the
    exact minimum of dE(0 -> 1) over every context that holds the frozen bits at 0."""
    rng = np.random.default_rng(seed)
    row_ptr, ray, a, d, w = [0], [], [], [], []
    for i in range(n):
        for _ in range(int(rng.integers(1, 4))):
            ray.append(len(d))
            a.append(rng.uniform(0.05, 1.5))
            d.append(rng.uniform(0.0, 1.2))
            w.append(rng.uniform(0.5, 4.0))
        row_ptr.append(len(ray))
    if edges is None:
        edges = {(int(u), int(v)) for u, v in rng.integers(0, n, (2 * n, 2)) if u != v}
        edges = {(min(u, v), max(u, v)) for u, v in edges}
    adj = [[] for _ in range(n)]
    for u, v in sorted(edges):
        adj[u].append(v)
        adj[v].append(u)
    nbr_ptr = np.concatenate([[0], np.cumsum([len(x) for x in adj])])
    nbr = [j for x in adj for j in x]
    field = rng.uniform(-1.0, 6.0, n)
    return orc.FloatProblem(row_ptr, ray, a, d, w, nbr_ptr, nbr, lam, 0.0, field)


def brute_min_dE(orc, prob, frozen):
    """min over contexts x (frozen bits 0, the bit itself 0, every other bit free) of
    E(x + e_i) - E(x), from binary_energy alone."""
    n = prob.n_vars
    free = np.flatnonzero(frozen == 0)
    best = np.full(n, np.inf)
    for mask in range(1 << len(free)):
        x = np.zeros(n, np.uint8)
        x[free] = (mask >> np.arange(len(free))) & 1
        e0 = None
        for i in range(n):
            if x[i]:
                continue
            if e0 is None:
                e0 = orc.binary_energy(prob, x)
            y = x.copy()
            y[i] = 1
            best[i] = min(best[i], orc.binary_energy(prob, y) - e0)
    return best


def fixpoint_from_minima(orc, prob, margin):
    """The frozen set built round by round from brute-force minima: a bit freezes once its
    minimum dE over contexts consistent with the frozen set reaches the margin."""
    frozen = np.zeros(prob.n_vars, np.uint8)
    sets = []
    while True:
        m = brute_min_dE(orc, prob, frozen)
        new = (frozen == 0) & (m >= margin)
        if not new.any():
            return frozen, sets
        frozen = frozen | new.astype(np.uint8)
        sets.append(frozen.copy())


@pytest.mark.parametrize("seed,n,margin", [(1, 9, 0.5), (2, 11, 1.0), (3, 12, 2.0), (4, 10, 0.0)])
def test_freeze_bound_is_exact_with_private_rays(orc, seed, n, margin):
    """Both sides of the freeze: with private rays the oracle's bound equals the brute-force
    minimum of dE(0 -> 1) (a dropped lin term, a lost data term or lambda (rock - free)
    written as -lambda free all break the equality), and its frozen set equals the one built
    from those minima round by round (so it neither over- nor under-freezes)."""
    prob = private_ray_problem(orc, n, seed)
    got = np.zeros(n, np.uint8)
    rounds = orc.lib().orc_freeze_mask(prob.c, margin, got)
    want, sets = fixpoint_from_minima(orc, prob, margin)
    assert np.array_equal(got, want), (got, want)
    assert rounds >= 1
    for frozen in [np.zeros(n, np.uint8)] + sets:
        bound = np.zeros(n, np.float64)
        orc.lib().orc_freeze_bounds(prob.c, frozen, bound)
        m = brute_min_dE(orc, prob, frozen)
        live = frozen == 0
        assert np.allclose(bound[live], m[live], rtol=1e-12, atol=1e-12), (bound[live], m[live])
    assert 0 < want.sum() < n or margin == 0.0


def test_freeze_needs_more_than_one_round(orc):
    """A bit whose bound reaches the margin only once a later-indexed neighbour is frozen:
    one in-order round freezes the neighbour alone, the fixpoint freezes both, and the
    oracle runs the extra rounds (the last freezing nothing)."""
    prob = private_ray_problem(orc, 4, 7, lam=1.5, edges={(0, 1), (2, 3)})
    fld = prob.field.copy()
    bound0 = np.zeros(4, np.float64)
    orc.lib().orc_freeze_bounds(prob.c, np.zeros(4, np.uint8), bound0)
    data = bound0 - fld + 1.5     # each bit has one free neighbour at the start
    margin = 2.0
    fld[1] = margin + 1.5 + 1.0 - data[1]      # bit 1 clears the margin with its neighbour free
    fld[0] = margin - data[0]                  # bit 0 only with bit 1 frozen: lb - 1.5 < margin <= lb + 1.5
    fld[2] = fld[3] = -5.0
    prob = orc.FloatProblem(prob.row_ptr, prob.ray, prob.a, prob.d, prob.w, prob.nbr_ptr, prob.nbr, 1.5, 0.0, fld)
    got = np.zeros(4, np.uint8)
    rounds = orc.lib().orc_freeze_mask(prob.c, margin, got)
    want, sets = fixpoint_from_minima(orc, prob, margin)
    assert sets[0].tolist() == [0, 1, 0, 0]
    assert want.tolist() == [1, 1, 0, 0]
    assert np.array_equal(got, want)
    assert rounds == 3


def long_column_problem(orc, exact, seed=21, n=40, n_rays=120, lengths=(20, 70)):
    """Bits with 20-70 rays each, so dE sums run over several chains of 16 (with
    lengths=(150, 400), over 10-25 chains: several runs of 8 chain sums)."""
    rng = np.random.default_rng(seed)
    row_ptr, ray, a = [0], [], []
    for _ in range(n):
        rays = np.sort(rng.choice(n_rays, size=rng.integers(*lengths), replace=False))
        for b in rays:
            ray.append(b)
            a.append(rng.integers(1, 16) / 1024.0 if exact else rng.uniform(0.001, 0.05))
        row_ptr.append(len(ray))
    d = rng.integers(0, 256, n_rays) / 1024.0 if exact else rng.uniform(0.0, 0.3, n_rays)
    w = rng.integers(1, 4, n_rays) * 1.0 if exact else rng.uniform(1.0, 30.0, n_rays)
    return orc.FloatProblem(row_ptr, ray, a, d, w, np.zeros(n + 1, np.int32), [], 0.0, 0.05)


LONG = [dict(), dict(n_rays=600, lengths=(150, 400))]


@pytest.mark.parametrize("shape", LONG)
def test_segmented_sum_equals_sequential_when_exact(orc, shape):
    """With dyadic a, d, w every dE is exact in any order, so chains of 16 (their sums
    added in runs of 8) and one chain give the same trajectory; a dropped or repeated
    term, chain or run would not."""
    prob = long_column_problem(orc, exact=True, **shape)
    order, cp = greedy_colouring(prob.n_vars, ray_conflicts(prob))
    sched = orc.schedule(2.0, 0.05, 40, 5)
    for seed in (1, 2, 3):
        for prec in (0, 1):
            a = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, prec, 0, 16))
            b = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, prec, 0, 0))
            assert np.array_equal(a, b) and 0 < a.sum() < prob.n_vars


@pytest.mark.parametrize("shape", LONG)
def test_segmented_sum_is_a_rounding_change_only(orc, shape):
    """With generic floats the two orders differ only by rounding: over many states the
    segmented dE stays within 1e-12 relative of the sequential one (checked through
    one-sweep decisions at a temperature where every decision is far from the threshold)."""
    prob = long_column_problem(orc, exact=False, **shape)
    order, cp = greedy_colouring(prob.n_vars, ray_conflicts(prob))
    sched = orc.schedule(1e-6, 1e-6, 1, 0)   # T -> 0: accept iff dE <= 0
    for seed in (1, 2):
        a = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, 0, 0, 16))
        b = orc.anneal_branch(prob, sched, seed, orc.make_mode(orc.RNG_PHILOX, order, cp, 0, 0, 0))
        assert np.array_equal(a, b)
