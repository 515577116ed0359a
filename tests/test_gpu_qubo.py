"""GPU parity of the batched sparse-QUBO engine, called through the C ABI, against the
oracle's trajectory for every instance and replica of a batch: spins bit-exact, FP64
energies within 1e-9 relative (the device reduces in another order), in every launch
shape (one persistent launch or one launch per instance, every replica block)."""
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def instances():
    # (bits, degree, replicas): uneven sizes, degrees and replica counts, partial blocks
    specs = [(300, 3, 1), (1000, 8, 5), (2048, 32, 8), (1500, 64, 17), (700, 8, 32), (4000, 3, 12), (129, 64, 9)]
    return [qa.random_regular_qubo(n, d, 100 + i, r) for i, (n, d, r) in enumerate(specs)]


def oracle_x(q, order, sched):
    s = orc_mod.schedule(sched.t_hot, sched.t_cold, sched.anneal_sweeps, sched.hold_sweeps)
    return orc_mod.qubo_anneal(q.row_ptr, q.col, q.J, q.h, order, s, q.seed, q.replicas)


def test_qubo_batch_equals_oracle():
    inst = instances()
    B = qa.QuboBatch(inst)
    sched = qa.schedule(5.0, 0.05, 40, 5)
    say(f"batch of {len(inst)} instances, persistent, replica block 8")
    out = B.anneal(sched, mode=0, replica_block=8)
    for k, q in enumerate(inst):
        order, cp = B.colouring(k)
        o2, cp2 = qa.colour_qubo(q)
        assert np.array_equal(order, o2) and np.array_equal(cp, cp2)
        ref = oracle_x(q, order, sched)
        assert np.array_equal(out["x"][k], ref), f"instance {k}: {(out['x'][k] != ref).sum()} spins differ"
        e_ref = np.array([orc_mod.qubo_energy(q.row_ptr, q.col, q.J, q.h, x) for x in ref])
        assert np.allclose(out["energy"][k], e_ref, rtol=1e-9, atol=1e-9)
        assert 0 < ref.sum() < ref.size
    assert out["report"]["jobs"] == sum((q.replicas + 7) // 8 for q in inst)


def test_qubo_launch_shapes_do_not_change_results():
    inst = instances()
    B = qa.QuboBatch(inst)
    sched = qa.schedule(3.0, 0.1, 12, 2)
    base = B.anneal(sched, mode=0, replica_block=8)["x"]
    for mode in (0, 1):
        for rb, blk in ((4, 256), (8, 1024), (16, 256), (32, 256), (4, 1024), (0, 0)):
            say(f"mode {mode} replica block {rb} block {blk}")
            got = B.anneal(sched, mode=mode, replica_block=rb, block=blk, workgroups=37 if mode == 0 else 0)
            for k in range(len(inst)):
                assert np.array_equal(got["x"][k], base[k]), (mode, rb, k)
            if mode == 1:
                assert got["report"]["launches"] == len(inst)


@pytest.mark.parametrize("mode", [0, 1])
def test_recorded_graph_replays_equal_eager(mode):
    """The launch schedule recorded into a hipGraph and replayed: every replay, with eager
    anneals (other schedules, other shapes) and synchronisations between the replays,
    gives the eager result, spins and energies bit for bit; a new schedule length records
    a new graph."""
    inst = instances()
    B = qa.QuboBatch(inst)
    sched = qa.schedule(4.0, 0.1, 14, 2)
    other = qa.schedule(2.0, 0.2, 9, 1)
    # a recording holds the workspace it was made on, and a call that grows the workspace
    # records anew: grow it to the largest shape first
    B.anneal(other, mode=mode, replica_block=4, block=1024)
    eager = B.anneal(sched, mode=mode, replica_block=8, block=256)
    for rep in range(4):
        say(f"mode {mode}: graph replay {rep}")
        g = B.anneal(sched, mode=mode, replica_block=8, block=256, graph=True)
        assert g["report"]["graph_replayed"] == (1 if rep > 0 else 0)
        for k in range(len(inst)):
            assert np.array_equal(g["x"][k], eager["x"][k]), (rep, k)
            assert np.array_equal(g["energy"][k], eager["energy"][k]), (rep, k)
        B.anneal(other, mode=mode, replica_block=4, block=1024)
        B.anneal(other, mode=mode, replica_block=8, block=256)
    g2 = B.anneal(other, mode=mode, replica_block=8, block=256, graph=True)
    assert g2["report"]["graph_replayed"] == 0
    e2 = B.anneal(other, mode=mode, replica_block=8, block=256)
    for k in range(len(inst)):
        assert np.array_equal(g2["x"][k], e2["x"][k])
