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
