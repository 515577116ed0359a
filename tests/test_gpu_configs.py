"""GPU parity at the workload configurations' own shapes, through the C ABI, against the
CPU oracle with the engine's colouring and the same Philox keys.

- The muon QUBO at full size (306,328 bits, 75,600 rays, 3,829,376 entries) with the
  benchmark's exact settings: 16 branches, T 50 -> 1 over 200 sweeps + 50 at T = 1, FP32
  residuals with an FP64 recompute every 50 sweeps; and the same in FP64. This is the
  solve of src/check_ising.cu:233 through branch_fractions (src/ising_recon.hpp:148-166).
- The 256-instance batched workload (n log-uniform in [1k, 64k], degree 3/8/32/64, 8-32
  replicas) in the tuned launch shape and at replica blocks 16 and 32, which send the
  instances whose spin words exceed the workgroup's LDS to global spin words.
- The dense N = 4096 QUBO with 64 replicas, FP32 fields with an FP64 recompute.
- The serving stream: requests of 2k-235k voxels arriving at 50 per second, served by the
  continuous server, each against the oracle on that request alone.

Bar: spins bit-exact; p_void and the tri-state map identical; FP64 energies within
1e-9 relative (the device reduces in another order than the oracle's sequential sum).
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-9
THREADS = max(1, min(16, os.cpu_count() or 1))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def muon():
    P = qa.muon_scene()
    order, cp = P.colouring()
    op = orc_mod.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
    return P, order, cp, op


@pytest.mark.parametrize("precision,k", [(qa.FP32, 50), (qa.FP64, 0)])
def test_muon_full_benchmark_solve(muon, precision, k):
    """The benchmark solve itself: branches, p_void, tri-state and certainty identical to the
    oracle's colour mode, energies within 1e-9."""
    P, order, cp, op = muon
    sched = qa.schedule(50.0, 1.0, 200, 50)
    prob = qa.Problem(P)
    g_order, g_cp = prob.colouring()
    assert np.array_equal(g_order, order) and np.array_equal(g_cp, cp)
    say(f"muon {P.n_vars} bits, {P.n_rays} rays, 16 branches, 200+50 sweeps, precision {precision}, k {k}")
    out = prob.anneal_branches(sched, 16, 0, precision, k)
    say(f"device solve: kernel {out['report']['kernel_ms']:.1f} ms; oracle on {THREADS} threads")
    mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, order, cp, precision, k)
    pv, br = orc_mod.branch_fractions(op, orc_mod.schedule(50.0, 1.0, 200, 50), 16, THREADS, 0, mode)
    diff = (out["branches"] != br).sum(axis=1)
    assert not diff.any(), f"branches differ: {diff}"
    assert np.array_equal(out["p_void"], pv)
    tri = np.where(pv >= np.float32(0.9), 1, np.where(pv <= np.float32(0.1), 0, 2)).astype(np.uint8)
    assert np.array_equal(out["tri"], tri)
    assert np.array_equal(out["certainty"], np.maximum(pv, np.float32(1.0) - pv))
    e_ref = np.array([orc_mod.binary_energy(op, b) for b in br])
    assert np.allclose(out["energy"], e_ref, rtol=ENERGY_RTOL, atol=0.0)
    counts = np.bincount(tri, minlength=3)
    say(f"tri-state identical: exists {counts[0]}, not-there {counts[1]}, undecided {counts[2]}; "
        f"energies {e_ref.min():.3f} .. {e_ref.max():.3f}")
    assert counts[1] > 0 and 0 < br.sum() < br.size
    if precision == qa.FP32:
        assert out["report"]["fp32_residual_err"] < 1e-4


@pytest.fixture(scope="module")
def batch256():
    inst = qa.qubo_batch_workload(256, seed=2026, threads=THREADS)
    say(f"batched workload: {len(inst)} instances, {sum(q.n for q in inst)} bits, "
        f"max n {max(q.n for q in inst)}")
    B = qa.QuboBatch(inst)
    sched = qa.schedule(5.0, 0.05, 2, 1)
    o = orc_mod.schedule(5.0, 0.05, 2, 1)

    def task(k):
        q = inst[k]
        order, _ = B.colouring(k)
        return orc_mod.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, o, q.seed, q.replicas)

    with ThreadPoolExecutor(THREADS) as ex:
        ref = list(ex.map(task, range(len(inst))))
    return inst, B, sched, ref


@pytest.mark.parametrize("rb,block", [(0, 0), (16, 256), (32, 256)])
def test_batched_256_instances(batch256, rb, block):
    """Every replica of every instance of the 256-instance workload equals the oracle's, in
    the tuned shape and with replica blocks whose spin words leave LDS for the largest
    instances (more than 32,768 bits at 16 replicas a job, 16,384 at 32)."""
    inst, B, sched, ref = batch256
    say(f"batched 256: replica block {rb}, block {block}")
    out = B.anneal(sched, mode=0, replica_block=rb, block=block)
    rep = out["report"]
    say(f"ran with block {rep['block']}, replica block {rep['replica_block']}: kernel {rep['kernel_ms']:.1f} ms")
    if rb:
        assert rep["replica_block"] == rb
        word_bytes = {4: 1, 8: 1, 16: 2, 32: 4}[rb]   # 64 KB of LDS spin words per job
        words = [q.n for q in inst if q.n * word_bytes > 65536]
        say(f"{len(words)} instances run with global spin words")
        assert words
    for k, q in enumerate(inst):
        assert np.array_equal(out["x"][k], ref[k]), f"instance {k} (n {q.n}): {(out['x'][k] != ref[k]).sum()} differ"
    for k in (0, 100, 255):
        q = inst[k]
        e_ref = np.array([orc_mod.qubo_energy(q.row_ptr, q.col, q.J, q.h, x) for x in ref[k]])
        assert np.allclose(out["energy"][k], e_ref, rtol=ENERGY_RTOL, atol=1e-9)


def test_dense_4096_sixty_four_replicas():
    """The dense N = 4096 QUBO with 64 replicas: FP32 fields with an FP64 recompute every 2
    sweeps, spins per replica and the final field error equal to the oracle's."""
    n, R = 4096, 64
    Q, h = qa.dense_qubo(n, 7)
    D = qa.DenseQubo(Q, h)
    sched = qa.schedule(1.0, 0.02, 4, 1)
    say(f"dense n={n} R={R}")
    out = D.anneal(sched, R, seed_offset=0, precision=qa.FP32, recompute_every=2)
    say(f"kernel {out['report']['kernel_ms']:.1f} ms")
    o = orc_mod.schedule(1.0, 0.02, 4, 1)
    Qd = Q.astype(np.float64)

    def task(b):
        return orc_mod.dense_anneal(Q, h, o, b + 1, orc_mod.FP32, 2, fields=True)

    with ThreadPoolExecutor(THREADS) as ex:
        ref = list(ex.map(task, range(R)))
    for b, (x, f) in enumerate(ref):
        assert np.array_equal(out["x"][b], x), f"replica {b}: {(out['x'][b] != x).sum()} differ"
        e = orc_mod.dense_energy(Q, h, x)
        assert abs(out["energy"][b] - e) <= ENERGY_RTOL * max(1.0, abs(e))
        scratch = h.astype(np.float64) + Qd.T @ x.astype(np.float64)
        assert abs(out["field_err"][b] - np.max(np.abs(f - scratch))) < 1e-9
    assert 0 < sum(int(x.sum()) for x, _ in ref) < n * R


def test_serving_stream_against_the_oracle():
    """The serving stream at its request sizes: Poisson arrivals at 50 per second of ray x
    voxel requests from a pool whose voxels span the configuration's log-uniform [2k,
    300k] (one of about 250k), with 1k-75k rays and 8-16 branches, served by the
    continuous server (requests join running unions at stage boundaries). Every request's
    branches, p_void and tri-state map equal the oracle's colour mode run on that request
    alone, with its own colouring and its seed as key. The schedule is short so the oracle
    can run every request."""
    from qublar_amd.serving import ContinuousServer, Request

    shapes = [(24, 48, 8), (36, 90, 14), (52, 180, 30), (70, 250, 45), (90, 300, 60), (118, 360, 70)]
    pool = [qa.muon_scene(grid=g, sky_az=az, sky_th=th, log2_muons=24, seed=500 + i)
            for i, (g, az, th) in enumerate(shapes)]
    sizes = sorted(P.n_vars for P in pool)
    say(f"pool voxels {sizes}, rays {sorted(P.n_rays for P in pool)}")
    assert sizes[0] < 5000 and sizes[-1] >= 200_000
    rng = np.random.default_rng(5)
    t, reqs = 0.0, []
    while t < 0.3:
        P = pool[len(reqs) % len(pool)] if len(reqs) < len(pool) else pool[int(rng.integers(len(pool)))]
        reqs.append(Request(P, int(rng.integers(8, 17)), 300 + len(reqs), t, keep_branches=True))
        t += rng.exponential(1.0 / 50.0)
    sched = qa.schedule(50.0, 1.0, 10, 5)
    srv = ContinuousServer(sched, qa.FP32, 5, stage_sweeps=5)
    reports = srv.run(reqs)
    say(f"{len(reqs)} requests, {len(reports)} stages, unions joined {[q.batch for q in reqs]}")
    o = orc_mod.schedule(50.0, 1.0, 10, 5)
    colourings = {id(P): P.colouring() for P in pool}
    ops = {id(P): orc_mod.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
           for P in pool}
    for i, q in enumerate(reqs):
        order, cp = colourings[id(q.problem)]
        mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, order, cp, orc_mod.FP32, 5)
        pv, br = orc_mod.branch_fractions(ops[id(q.problem)], o, q.branches, THREADS, q.seed, mode)
        assert np.array_equal(q.branches_out, br), f"request {i} ({q.problem.n_vars} bits): branches differ"
        assert np.array_equal(q.p_void, pv)
        tri = np.where(pv >= np.float32(0.9), 1, np.where(pv <= np.float32(0.1), 0, 2)).astype(np.uint8)
        assert np.array_equal(q.tri, tri)
