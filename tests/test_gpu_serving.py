"""Requests annealed together give exactly what each gives alone.

Independent ray x voxel QUBOs are served in batches by annealing their block-diagonal
union: every request keeps its local bit indices as draw counters and its own seed as
draw key, and colour class c of the union is class c of every request, so request q's
branch b in the union is, spin for spin, its branch b annealed by itself with its own
seed offset. Checked here for requests of different sizes, seeds and branch counts, in
FP64 and in FP32 with recompute.
"""
import numpy as np
import pytest

import qublar_amd as qa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def requests():
    shapes = [dict(grid=16, sky_az=30, sky_th=10, log2_muons=20, seed=3),
              dict(grid=22, sky_az=45, sky_th=12, log2_muons=21, seed=8),
              dict(grid=18, sky_az=36, sky_th=9, log2_muons=20, seed=11)]
    return [qa.muon_scene(**s) for s in shapes], [5, 17, 1234567]


@pytest.mark.parametrize("precision,k", [(qa.FP64, 0), (qa.FP32, 4)])
def test_union_equals_alone(precision, k):
    probs, keys = requests()
    branches = [16, 8, 12]
    sched = qa.schedule(20.0, 1.0, 14, 4)
    alone = []
    for P, key, R in zip(probs, keys, branches):
        say(f"alone: {P.n_vars} bits, key {key}, {R} branches")
        alone.append(qa.Problem(P).anneal_branches(sched, R, key, precision, k)["branches"])
    U, offs = qa.union(probs, keys)
    say(f"union: {U.n_vars} bits, {U.n_rays} rays")
    mix = qa.Problem(U).anneal_branches(sched, max(branches), 0, precision, k)["branches"]
    for q, (P, R) in enumerate(zip(probs, branches)):
        b0 = offs[q][0]
        assert np.array_equal(mix[:R, b0:b0 + P.n_vars], alone[q]), f"request {q} differs in the mix"


def test_union_is_order_free():
    """The same requests in another order give the same per-request results."""
    probs, keys = requests()
    sched = qa.schedule(20.0, 1.0, 10, 2)
    U1, o1 = qa.union(probs, keys)
    U2, o2 = qa.union(probs[::-1], keys[::-1])
    say("two unions")
    m1 = qa.Problem(U1).anneal_branches(sched, 8, 0, qa.FP32, 3)["branches"]
    m2 = qa.Problem(U2).anneal_branches(sched, 8, 0, qa.FP32, 3)["branches"]
    n = len(probs)
    for q, P in enumerate(probs):
        a = o1[q][0]
        b = o2[n - 1 - q][0]
        assert np.array_equal(m1[:, a:a + P.n_vars], m2[:, b:b + P.n_vars])


def test_server_results_equal_alone():
    """The server's per-request p_void and tri-state maps, from one union of requests
    with 16, 8 and 12 branches, equal each request's own solve: a request averages its
    own branches, not the union's 16."""
    from qublar_amd.serving import Request, Server

    probs, keys = requests()
    branches = [16, 8, 12]
    sched = qa.schedule(20.0, 1.0, 14, 4)
    reqs = [Request(P, R, key, 0.0) for P, R, key in zip(probs, branches, keys)]
    say("server: one union of 3 requests")
    reports = Server(sched, qa.FP32, 4).run(reqs, batched=True)
    assert len(reports) == 1 and all(q.batch == 0 for q in reqs)
    for q in reqs:
        alone = qa.Problem(q.problem).anneal_branches(sched, q.branches, q.seed, qa.FP32, 4)
        assert np.array_equal(q.p_void, alone["p_void"]), f"request with {q.branches} branches"
        assert np.array_equal(q.tri, alone["tri"])


@pytest.mark.parametrize("precision,k", [(qa.FP64, 0), (qa.FP32, 4)])
def test_requests_joining_running_unions_equal_alone(precision, k):
    """Staged solves: request A runs alone for 8 sweeps, B joins it (A's state imported),
    then C joins both; every request's branches, p_void and tri-state equal its solve
    alone, although A and B changed unions mid-schedule and each union anneals 16
    branches while B asks for 8 and C for 12."""
    probs, keys = requests()
    branches = [16, 8, 12]
    sched = qa.schedule(20.0, 1.0, 18, 6)
    total = 24
    alone = [qa.Problem(P).anneal_branches(sched, R, key, precision, k) for P, key, R in zip(probs, keys, branches)]
    held = []

    def stage_of(members):
        U, offs = qa.union([probs[q] for q in members], [keys[q] for q in members])
        prob = qa.Problem(U)
        st = qa.Stage(prob, sched, max(branches[q] for q in members), precision, k, group_width=2)
        held.append((prob, st))
        return st, dict(zip(members, offs))

    s1, o1 = stage_of([0])
    s1.set_span(o1[0][0], probs[0].n_vars, 0, branches[0])
    s1.run(4)
    s1.run(4)
    say("A: 8 sweeps alone")
    s2, o2 = stage_of([0, 1])
    s2.import_span(s1, o1[0][0], o2[0][0], probs[0].n_vars, o1[0][1], o2[0][1], probs[0].n_rays)
    s2.set_span(o2[1][0], probs[1].n_vars, 0, branches[1])
    s2.run(8)
    say("A and B: 8 sweeps together")
    s3, o3 = stage_of([1, 2, 0])
    for q, src in ((0, (s2, o2)), (1, (s2, o2))):
        st, off = src
        s3.import_span(st, off[q][0], o3[q][0], probs[q].n_vars, off[q][1], o3[q][1], probs[q].n_rays)
    s3.set_span(o3[2][0], probs[2].n_vars, 0, branches[2])
    for _ in range(total // 4):
        s3.run(4)
    say("A, B and C: to the end of C's schedule")
    for q, P in enumerate(probs):
        out = s3.outputs(o3[q][0], P.n_vars, branches[q])
        assert np.array_equal(out["branches"], alone[q]["branches"]), f"request {q}"
        assert np.array_equal(out["p_void"], alone[q]["p_void"])
        assert np.array_equal(out["tri"], alone[q]["tri"])
        assert np.array_equal(out["certainty"], alone[q]["certainty"])
    # a span moves only at a multiple of the recompute period
    if k:
        s4, o4 = stage_of([0])
        s5, o5 = stage_of([0])
        s4.run(2)
        with pytest.raises(qa.QublarError):
            s5.import_span(s4, 0, 0, probs[0].n_vars, 0, 0, probs[0].n_rays)


def test_continuous_server_results_equal_alone():
    """Requests arriving over time are served by the continuous server: each joins a
    running union at a stage boundary (the unions and join points depend on timing), and
    each request's branches, p_void and tri-state equal its solve alone."""
    from qublar_amd.serving import ContinuousServer, Request

    probs, keys = requests()
    sched = qa.schedule(20.0, 1.0, 30, 10)
    specs = [(0, 16, 0.0), (1, 8, 0.02), (2, 12, 0.04), (0, 10, 0.06), (1, 16, 0.3), (2, 9, 0.31)]
    reqs = [Request(probs[p], R, 1000 + 7 * i, t, keep_branches=True) for i, (p, R, t) in enumerate(specs)]
    srv = ContinuousServer(sched, qa.FP32, 4, stage_sweeps=4)
    reports = srv.run(reqs)
    unions = sorted({r["union"] for r in reports})
    say(f"continuous: {len(reports)} stages over unions {unions}; joined unions {[q.batch for q in reqs]}")
    assert all(q.done >= q.arrival for q in reqs)
    for i, q in enumerate(reqs):
        alone = qa.Problem(q.problem).anneal_branches(sched, q.branches, q.seed, qa.FP32, 4)
        assert np.array_equal(q.branches_out, alone["branches"]), f"request {i}"
        assert np.array_equal(q.p_void, alone["p_void"])
        assert np.array_equal(q.tri, alone["tri"])
