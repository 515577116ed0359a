"""GPU parity: every result of the HIP engine, called through the C ABI, against the
CPU oracle run with the engine's own colouring and the same Philox keys.

Bar: integer problems bit-exact (spins, tracked energies, device energies); float
problems identical spins, p_void and tri-state maps, final FP64 energies within
1e-9 relative (the reduction order differs from the oracle's sequential sum).
"""
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    info = qa.device_info()
    say(f"device {info}")
    assert info["arch"].startswith("gfx950")
    return info


def oracle_graph(g):
    return orc_mod.IntGraph(g.n, g.eu, g.ev, g.ew)


def oracle_problem(P):
    return orc_mod.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa, P.field)


def osched(s):
    return orc_mod.schedule(s.t_hot, s.t_cold, s.anneal_sweeps, s.hold_sweeps)


# ---- integer path -------------------------------------------------------------------

def check_int_batch(g, sched, R, seed_base=1, force_global=False, updates_per_launch=1e9):
    G = qa.Graph(g, force_global=force_global)
    order = np.zeros(g.n, np.int32)
    plan = G.plan()
    cp = np.zeros(plan["n_classes"] + 1, np.int32)
    qa._check(qa.lib().qb_graph_colouring(G.h, order.ctypes.data, cp.ctypes.data))
    say(f"run_batch n={g.n} m={g.m} R={R} classes={plan['n_classes']} global={force_global}")
    s, e, rep = G.run_batch(sched, seed_base, R, updates_per_launch)
    og = oracle_graph(g)
    mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, order, cp)
    for r in range(R):
        x, en = orc_mod.anneal_int(og, osched(sched), seed_base + r, mode)
        xg = (s[r, :g.n] > 0).astype(np.uint8)
        assert np.array_equal(xg, x), f"replica {r}: {int((xg != x).sum())} spins differ"
        assert e[r] == en, f"replica {r}: energy {e[r]} vs {en}"
        assert np.all(np.abs(s[r, :g.n]) == 1)
    X = (s[:, :g.n] > 0).astype(np.uint8)
    assert np.array_equal(G.device_energies(X), e)
    return s, e, rep


def test_int_maxcut_800_sixteen_replicas(gpu):
    """The max-cut workload: ER n=800 p=0.06 +-1, 16 replicas, T 3.0 -> 0.05 over 1000 sweeps."""
    g = qa.er_graph(800, 0.06, 1)
    s, e, rep = check_int_batch(g, qa.schedule(3.0, 0.05, 1000, 0), 16)
    say(f"kernel_ms {rep['kernel_ms']:.2f} best cut {-e.min()}")


def test_int_chunked_launches_and_global_state(gpu):
    """The launch split and the state placement never change results (ising_gpu.cuh:19-21)."""
    g = qa.er_graph(300, 0.1, 4)
    sched = qa.schedule(3.0, 0.05, 120, 10)
    s0, e0, _ = check_int_batch(g, sched, 8, 5)
    s1, e1, rep = check_int_batch(g, sched, 8, 5, updates_per_launch=8 * 300 * 7)
    assert rep["launches"] > 10
    s2, e2, _ = check_int_batch(g, sched, 8, 5, force_global=True)
    assert np.array_equal(s0, s1) and np.array_equal(e0, e1)
    assert np.array_equal(s0, s2) and np.array_equal(e0, e2)


def test_int_k16_and_signed_weights(gpu):
    """K16 with weights in [-1000, 1000] (gpu_anneal.cu:166-169) and a signed sparse graph."""
    rng = np.random.default_rng(3)
    iu = np.triu_indices(16, 1)
    g = qa.GraphArrays(16, iu[0], iu[1], rng.integers(-1000, 1001, len(iu[0])))
    check_int_batch(g, qa.schedule(1000.0, 0.1, 400, 20), 8)
    eu, ev = [], []
    seen = set()
    while len(eu) < 3000:
        u, v = rng.integers(0, 300, 2)
        if u != v and (min(u, v), max(u, v)) not in seen:
            seen.add((min(u, v), max(u, v))); eu.append(min(u, v)); ev.append(max(u, v))
    w = rng.integers(1, 1001, 3000) * rng.choice([-1, 1], 3000)
    check_int_batch(qa.GraphArrays(300, eu, ev, w), qa.schedule(500.0, 0.1, 200, 20), 8)


# ---- float path ---------------------------------------------------------------------

def check_branches(P, sched, R, seed_offset=0, precision=qa.FP64, recompute_every=0, **kw):
    prob = qa.Problem(P)
    order, cp = prob.colouring()
    info = prob.info()
    say(f"anneal n={P.n_vars} rays={P.n_rays} classes={info['n_classes']} max_entries={info['max_entries_per_bit']}"
        f" R={R} prec={precision} k={recompute_every} {kw}")
    out = prob.anneal_branches(sched, R, seed_offset, precision, recompute_every, **kw)
    op = oracle_problem(P)
    mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, order, cp, precision, recompute_every)
    pv, br = orc_mod.branch_fractions(op, osched(sched), R, 16, seed_offset, mode)
    diff = (out["branches"] != br).sum(axis=1)
    assert not diff.any(), f"branches differ at {np.flatnonzero(diff)[:8]}: counts {diff[diff > 0][:8]}"
    assert np.array_equal(out["p_void"], pv)
    tri = np.where(pv >= np.float32(0.9), 1, np.where(pv <= np.float32(0.1), 0, 2))
    assert np.array_equal(out["tri"], tri.astype(np.uint8))
    assert np.array_equal(out["certainty"], np.maximum(pv, np.float32(1.0) - pv))
    e_ref = np.array([orc_mod.binary_energy(op, br[b]) for b in range(R)])
    assert np.allclose(out["energy"], e_ref, rtol=ENERGY_RTOL, atol=1e-12 * max(1.0, np.abs(e_ref).max()))
    return prob, out


def small_muon(**kw):
    base = dict(grid=32, sky_az=90, sky_th=20, log2_muons=22)
    base.update(kw)
    return qa.muon_scene(**base)


def test_ray_muon_small_fp64(gpu):
    P = small_muon()
    prob, out = check_branches(P, qa.schedule(50.0, 1.0, 30, 10), 16)
    assert 0 < out["branches"].sum() < out["branches"].size
    say(f"kernel_ms {out['report']['kernel_ms']:.2f}")


def test_ray_muon_small_fp32_recompute(gpu):
    P = small_muon()
    _, out = check_branches(P, qa.schedule(50.0, 1.0, 30, 10), 16, precision=qa.FP32, recompute_every=7)
    assert out["report"]["fp32_residual_err"] < 1e-3


def test_ray_full_muon_qubo(gpu):
    """The benchmark problem itself (306,328 bits, 115 classes, multi-pass bits of up to
    4,353 entries, two classes of ~90k ray-free bits handed out in claimed chunks) for
    10 sweeps of 16 branches, FP32 with recompute every 4."""
    P = qa.muon_scene()
    _, out = check_branches(P, qa.schedule(50.0, 1.0, 8, 2), 16, precision=qa.FP32, recompute_every=4)
    say(f"kernel_ms {out['report']['kernel_ms']:.2f}")


def test_ray_uneven_branch_groups_and_offset(gpu):
    """20 branches (a full group and a partial one) with seed_offset 5."""
    P = small_muon(grid=24, sky_az=60, sky_th=15)
    check_branches(P, qa.schedule(20.0, 1.0, 20, 5), 20, seed_offset=5, precision=qa.FP32)


def test_ray_launch_split_team_size_and_group_width_do_not_change_results(gpu):
    P = small_muon(grid=24, sky_az=60, sky_th=15)
    sched = qa.schedule(20.0, 1.0, 20, 5)
    _, a = check_branches(P, sched, 16, precision=qa.FP32, recompute_every=4)
    _, b = check_branches(P, sched, 16, precision=qa.FP32, recompute_every=4, sweeps_per_launch=3, workgroups=40)
    assert b["report"]["launches"] == 9
    assert np.array_equal(a["branches"], b["branches"])
    for bw in (4, 8, 16):
        _, c = check_branches(P, sched, 16, precision=qa.FP32, recompute_every=4, group_width=bw, workgroups=64)
        assert np.array_equal(a["branches"], c["branches"])


def test_ray_many_branches(gpu):
    """40 branches: 20 groups of 2 over the 8 XCD teams (uneven load per team)."""
    P = small_muon(grid=20, sky_az=45, sky_th=12)
    check_branches(P, qa.schedule(20.0, 1.0, 12, 4), 40, group_width=2, precision=qa.FP64, recompute_every=5)


def synthetic_problem(n=300, n_rays=400, seed=0, dup=False, field=True):
    """Mixed light and heavy columns, a linear field, 6-neighbour-like adjacency, and
    (dup=True) repeated rays inside some heavy columns."""
    rng = np.random.default_rng(seed)
    row_ptr, ray, a = [0], [], []
    for i in range(n):
        k = int(rng.integers(1, 12)) if rng.random() < 0.8 else int(rng.integers(17, 90))
        rs = list(rng.choice(n_rays, size=k, replace=False))
        if dup and k > 1 and rng.random() < 0.3:
            rs[-1] = rs[0]     # a repeated ray, in light and heavy columns
        for b in rs:
            ray.append(b)
            a.append(rng.uniform(0.005, 0.08))
        row_ptr.append(len(ray))
    side = 17
    nbr_ptr, nbr = [0], []
    for i in range(n):
        for j in (i - 1, i + 1, i - side, i + side):
            if 0 <= j < n and j != i and abs(j % side - i % side) <= 1:
                nbr.append(j)
        nbr_ptr.append(len(nbr))
    d = rng.uniform(0.0, 0.4, n_rays)
    w = rng.uniform(5.0, 60.0, n_rays)
    fld = rng.normal(0.5, 0.8, n) if field else None
    return qa.ProblemArrays(row_ptr, ray, a, d, w, nbr_ptr, nbr, 0.7, 1.3, fld)


@pytest.mark.parametrize("precision", [qa.FP64, qa.FP32])
def test_ray_synthetic_heavy_and_duplicate_columns(gpu, precision):
    P = synthetic_problem(dup=True)
    check_branches(P, qa.schedule(8.0, 0.3, 40, 10), 16, precision=precision, recompute_every=6)


def very_heavy_problem(seed=3):
    """Light bits plus a few columns longer than a workgroup's slots hold at group
    width 2 (> 2048 entries): three staged in LDS (2100, 2600, 4400 entries, one with
    a repeated ray) and one beyond it, evaluated pass by pass (6500 entries)."""
    rng = np.random.default_rng(seed)
    n, n_rays = 120, 8000
    heavy = {5: 2100, 40: 2600, 77: 4400, 101: 6500}
    row_ptr, ray, a = [0], [], []
    for i in range(n):
        k = heavy.get(i, int(rng.integers(1, 40)))
        rs = list(rng.choice(n_rays, size=k, replace=False))
        if i == 40:
            rs[-1] = rs[3]
        for b in rs:
            ray.append(b)
            a.append(rng.uniform(0.001, 0.02))
        row_ptr.append(len(ray))
    side = 11
    nbr_ptr, nbr = [0], []
    for i in range(n):
        for j in (i - 1, i + 1, i - side, i + side):
            if 0 <= j < n and j != i and abs(j % side - i % side) <= 1:
                nbr.append(j)
        nbr_ptr.append(len(nbr))
    d = rng.uniform(0.0, 0.3, n_rays)
    w = rng.uniform(2.0, 30.0, n_rays)
    return qa.ProblemArrays(row_ptr, ray, a, d, w, nbr_ptr, nbr, 0.9, 0.4)


@pytest.mark.parametrize("precision", [qa.FP64, qa.FP32])
def test_ray_very_heavy_columns(gpu, precision):
    P = very_heavy_problem()
    check_branches(P, qa.schedule(6.0, 0.2, 30, 6), 16, precision=precision, recompute_every=7)


def test_energy_of_arbitrary_states(gpu):
    P = synthetic_problem(seed=4)
    prob = qa.Problem(P)
    rng = np.random.default_rng(9)
    X = rng.integers(0, 2, (37, P.n_vars)).astype(np.uint8)
    say("qb_energy K=37")
    tot, dat, pri = prob.energy(X)
    op = oracle_problem(P)
    for k in range(37):
        d, q = orc_mod.energy_parts(op, X[k])
        assert abs(dat[k] - d) <= ENERGY_RTOL * max(1.0, abs(d))
        assert abs(pri[k] - q) <= ENERGY_RTOL * max(1.0, abs(q))
        assert abs(tot[k] - (d + q)) <= ENERGY_RTOL * max(1.0, abs(d + q))


def test_errors_are_reported_not_fatal(gpu):
    P = synthetic_problem(seed=2)
    prob = qa.Problem(P)
    with pytest.raises(qa.QublarError) as e:
        prob.anneal_branches(qa.schedule(-1.0, 1.0, 10, 0), 16)
    assert e.value.code == qa.QB_E_ARG
    with pytest.raises(qa.QublarError):
        prob.anneal_branches(qa.schedule(5.0, 1.0, 10, 0), 0)


# ---- the reference's own integer trajectory ---------------------------------------------

def check_int_reference(g, sched, R, seed_base=1, updates_per_launch=1e9):
    """qb_run_batch_reference against the oracle's reference mode (wising::anneal: index order,
    splitmix64 drawn on uphill proposals), replica by replica."""
    G = qa.Graph(g)
    say(f"run_batch_reference n={g.n} m={g.m} R={R} upl={updates_per_launch:g}")
    s, e, rep = G.run_batch_reference(sched, seed_base, R, updates_per_launch)
    og = oracle_graph(g)
    for r in range(R):
        x, en = orc_mod.anneal_int(og, osched(sched), seed_base + r)
        xg = (s[r, :g.n] > 0).astype(np.uint8)
        assert np.array_equal(xg, x), f"replica {r}: {int((xg != x).sum())} spins differ"
        assert e[r] == en, f"replica {r}: energy {e[r]} vs {en}"
    X = (s[:, :g.n] > 0).astype(np.uint8)
    assert np.array_equal(G.device_energies(X), e)
    return s, e, rep


def test_int_reference_trajectory_maxcut_800(gpu):
    """The max-cut workload in the reference's trajectory: each replica equals
    wising::anneal with its seed (src/ising_weighted.hpp:250-267)."""
    g = qa.er_graph(800, 0.06, 1)
    s, e, rep = check_int_reference(g, qa.schedule(3.0, 0.05, 1000, 0), 16)
    say(f"kernel_ms {rep['kernel_ms']:.2f} best cut {-e.min()}")


def test_int_reference_trajectory_split_launches_and_weights(gpu):
    """Signed weights, dense K16 and a launch split every few sweeps (the splitmix64 state
    carries across launches); results do not depend on the split."""
    rng = np.random.default_rng(3)
    iu = np.triu_indices(16, 1)
    k16 = qa.GraphArrays(16, iu[0], iu[1], rng.integers(-1000, 1001, len(iu[0])))
    check_int_reference(k16, qa.schedule(1000.0, 0.1, 300, 20), 8, 5, updates_per_launch=16 * 8 * 7)
    g = qa.er_graph(300, 0.1, 4)
    sched = qa.schedule(3.0, 0.05, 120, 10)
    s0, e0, _ = check_int_reference(g, sched, 12, 9)
    s1, e1, rep = check_int_reference(g, sched, 12, 9, updates_per_launch=12 * 300 * 11)
    assert np.array_equal(s0, s1) and np.array_equal(e0, e1) and rep["launches"] > 1


def test_int_reference_trajectory_one_replica_per_workgroup(gpu):
    """A graph whose replica state takes more than a quarter of the LDS limit runs one
    replica per workgroup."""
    g = qa.er_graph(9000, 0.0005, 6)
    check_int_reference(g, qa.schedule(2.0, 0.05, 6, 2), 3, 11)


def test_int_reference_trajectory_past_the_lds_limit(gpu):
    """A graph whose replica state (5 bytes a vertex) exceeds 160 KB of LDS runs with its
    spins and fields in global memory, replica by replica equal to wising::anneal, also
    across a launch split (the reference's run_batch has no size limit either,
    src/ising_gpu.cuh:396-458)."""
    g = qa.er_graph(33000, 0.0002, 8)
    sched = qa.schedule(2.5, 0.05, 14, 3)
    s0, e0, _ = check_int_reference(g, sched, 5, 21)
    s1, e1, rep = check_int_reference(g, sched, 5, 21, updates_per_launch=5 * 33000 * 4)
    assert rep["launches"] == 5
    assert np.array_equal(s0, s1) and np.array_equal(e0, e1)


def test_int_reference_trajectory_reproduces_published_k16_cuts(gpu):
    """gset_bench --check B on the device: the best of 8 replicas (seeds 1..8) of the three
    random K16 of gset_bench.cu:252-264 at Budget(5, 1e-3, 4000, 200) cut 10622, 12434 and
    5050 -- the numbers the reference recorded (experiments/p1_bench/check_log.txt:211)."""
    import ctypes as C

    for seed, want in zip((1, 2, 3), (10622, 12434, 5050)):
        s = C.c_uint64(seed)
        eu, ev, ew = [], [], []
        for i in range(16):
            for j in range(i + 1, 16):
                eu.append(i); ev.append(j)
                ew.append(int(orc_mod.lib().orc_splitmix(C.byref(s)) % 2001) - 1000)
        G = qa.Graph(qa.GraphArrays(16, eu, ev, ew))
        _, e, _ = G.run_batch_reference(qa.schedule(5.0, 1e-3, 4000, 200), 1, 8)
        say(f"K16 seed {seed}: best cut {-e.min()} (recorded {want})")
        assert -e.min() == want


# ---- the reference's own float trajectory -------------------------------------------------

def check_branches_reference(P, sched, R, seed_offset=0, sweeps_per_launch=0):
    """qb_anneal_branches_reference against the oracle's reference mode (anneal_branch:
    index order, splitmix64 on uphill proposals, local_field_dE's order), branch by branch,
    with p_void, tri-state and FP64 energies."""
    prob = qa.Problem(P)
    say(f"anneal_branches_reference n={P.n_vars} rays={P.n_rays} R={R} offset={seed_offset} "
        f"sweeps/launch={sweeps_per_launch}")
    out = prob.anneal_branches_reference(sched, R, seed_offset, sweeps_per_launch=sweeps_per_launch)
    op = oracle_problem(P)
    pv, br = orc_mod.branch_fractions(op, osched(sched), R, 16, seed_offset)
    diff = (out["branches"] != br).sum(axis=1)
    assert not diff.any(), f"branches differ at {np.flatnonzero(diff)[:8]}: counts {diff[diff > 0][:8]}"
    assert np.array_equal(out["p_void"], pv)
    tri = np.where(pv >= np.float32(0.9), 1, np.where(pv <= np.float32(0.1), 0, 2))
    assert np.array_equal(out["tri"], tri.astype(np.uint8))
    e_ref = np.array([orc_mod.binary_energy(op, br[b]) for b in range(R)])
    assert np.allclose(out["energy"], e_ref, rtol=ENERGY_RTOL, atol=1e-12 * max(1.0, np.abs(e_ref).max()))
    return out


def test_float_reference_trajectory_muon(gpu):
    """A small muon QUBO, 16 branches: each branch is anneal_branch with seed b + 1."""
    out = check_branches_reference(small_muon(), qa.schedule(50.0, 1.0, 30, 10), 16)
    assert 0 < out["branches"].sum() < out["branches"].size
    say(f"kernel_ms {out['report']['kernel_ms']:.2f}")


def test_float_reference_trajectory_split_launches(gpu):
    """Launches of 3 and of 7 sweeps continue from the spins, residuals and splitmix64 state
    the last launch left: every branch is still anneal_branch's, and the same as one launch."""
    P = small_muon(grid=24, sky_az=60, sky_th=15)
    sched = qa.schedule(30.0, 1.0, 16, 4)
    one = check_branches_reference(P, sched, 9, seed_offset=2, sweeps_per_launch=20)
    assert one["report"]["launches"] == 1
    for per in (3, 7):
        out = check_branches_reference(P, sched, 9, seed_offset=2, sweeps_per_launch=per)
        assert out["report"]["launches"] == (20 + per - 1) // per
        assert np.array_equal(out["branches"], one["branches"])


def test_float_reference_trajectory_fields_duplicates_and_offset(gpu):
    """Per-bit fields, repeated rays in a column (residual additions in CSR order), 13
    branches with seed_offset 4, and the very heavy columns."""
    check_branches_reference(synthetic_problem(dup=True), qa.schedule(8.0, 0.3, 40, 10), 13, seed_offset=4)
    check_branches_reference(very_heavy_problem(), qa.schedule(6.0, 0.2, 12, 3), 5, seed_offset=1)


# ---- freeze_provable_rock's fixpoint on the device ------------------------------------------

def test_freeze_equals_reference_fixpoint(gpu):
    """qb_freeze's frozen set equals the oracle's in-order freeze_provable_rock
    (src/ising_recon.hpp:187-205) for lambda >= 0, on a muon scene at several margins and on
    the full muon QUBO; lambda < 0 is refused."""
    for P, margins in ((small_muon(), (0.0, 2.0, 10.0)), (qa.muon_scene(), (10.0,))):
        prob = qa.Problem(P)
        op = oracle_problem(P)
        for margin in margins:
            got, rounds = prob.freeze(margin)
            want = np.zeros(P.n_vars, np.uint8)
            orc_rounds = orc_mod.lib().orc_freeze_mask(op.c, margin, want)
            say(f"freeze n={P.n_vars} margin {margin}: {int(got.sum())} frozen in {rounds} rounds "
                f"(in order: {orc_rounds})")
            assert np.array_equal(got, want)
    neg = small_muon(**{"lambda": -0.5})
    with pytest.raises(qa.QublarError):
        qa.Problem(neg).freeze(1.0)


def test_frozen_reduced_problem_anneals_as_the_oracle():
    """The frozen set on the device, the reduced problem on the host (qb_reduce_frozen,
    src/ising_recon.hpp:206-233) coloured by the full problem's colouring restricted to the
    kept bits, then annealed on the device: branches and p_void equal the oracle's on the
    reduced problem with that colouring, energies within 1e-9."""
    P = qa.muon_scene(grid=32, sky_az=72, sky_th=16, log2_muons=22)
    prob = qa.Problem(P)
    frozen, rounds = prob.freeze(10.0)
    order, cp = prob.colouring()
    colour = np.zeros(P.n_vars, np.int32)
    for c in range(len(cp) - 1):
        colour[order[cp[c]:cp[c + 1]]] = c
    R, kept = P.reduce_frozen(frozen)
    R.colour = colour[kept]
    say(f"frozen {int(frozen.sum())} of {P.n_vars} bits in {rounds} rounds; {R.n_vars} kept")
    assert 0 < R.n_vars < P.n_vars
    rp = qa.Problem(R)
    sched = qa.schedule(50.0, 1.0, 12, 4)
    out = rp.anneal_branches(sched, 8, 0, qa.FP32, 4)
    ro, rcp = rp.colouring()
    op = orc_mod.FloatProblem(R.row_ptr, R.ray, R.a, R.d, R.w, R.nbr_ptr, R.nbr, R.lam, R.kappa, R.field)
    mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, ro, rcp, orc_mod.FP32, 4)
    pv, br = orc_mod.branch_fractions(op, orc_mod.schedule(50.0, 1.0, 12, 4), 8, 8, 0, mode)
    assert np.array_equal(out["branches"], br) and np.array_equal(out["p_void"], pv)
    e = np.array([orc_mod.binary_energy(op, b) for b in br])
    assert np.allclose(out["energy"], e, rtol=1e-9, atol=0.0)


def test_team_barrier_shows_every_members_stores():
    """What the sweep's team barrier relies on, checked on this device: after every team
    barrier, each member's drained plain stores are seen by the others' L2-hitting loads --
    20,000 steps, every thread of every workgroup, no stale read."""
    stale, us = qa.team_barrier_check(20000)
    say(f"team barrier: {stale} stale reads, {us:.2f} us per step")
    assert stale == 0
