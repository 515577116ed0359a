"""CPU checks of libqublar_hip.so that need no GPU: the library loads and exports every
function include/qublar.h declares; the host problem builders and the colourings."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import qublar_amd as qa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qublar.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qb_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = qa.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_refuses_bad_arguments_without_a_device():
    import ctypes as C
    L = qa.lib()
    assert L.qb_problem_create(None, C.byref(C.c_void_p())) == qa.QB_E_ARG
    assert b"null" in L.qb_last_error()
    bad = qa.ProblemArrays([0, 1], [5], [0.1], [0.0], [1.0], [0, 0], [], 0.0, 0.0)   # ray 5 of 1
    d = bad.desc()
    assert L.qb_problem_create(C.byref(d), C.byref(C.c_void_p())) == qa.QB_E_ARG
    assert b"ray" in L.qb_last_error()


def test_muon_scene_matches_the_reference_problem_size():
    """The builder's 128^3 scene has the reference's published size: 306,328 variables,
    75,600 rays, 3,829,376 ray-voxel entries (experiments/p1_bench/check_log.txt:162)."""
    P = qa.muon_scene()
    assert (P.n_vars, P.n_rays, int(P.row_ptr[-1])) == (306328, 75600, 3829376)
    assert int(P.truth.sum()) > 0
    assert np.all(P.w >= 1.0) and np.all(np.isfinite(P.d))
    # 6-neighbour lists are symmetric
    i = np.repeat(np.arange(P.n_vars), np.diff(P.nbr_ptr))
    fwd = set(zip(i.tolist(), P.nbr.tolist()))
    assert all((b, a) in fwd for a, b in list(fwd)[:20000])


def colours_of(order, cp, n):
    col = np.empty(n, np.int64)
    for c in range(len(cp) - 1):
        col[order[cp[c]:cp[c + 1]]] = c
    return col


def assert_distance2_colouring(P, order, cp):
    n = P.n_vars
    assert np.array_equal(np.sort(order), np.arange(n))
    col = colours_of(order, cp, n)
    bit = np.repeat(np.arange(n), np.diff(P.row_ptr))
    key = P.ray.astype(np.int64) * (len(cp) + 1) + col[bit]
    assert len(np.unique(key)) == len(key), "two bits of one class share a ray"
    i = np.repeat(np.arange(n), np.diff(P.nbr_ptr))
    assert np.all(col[i] != col[P.nbr]), "neighbours share a class"


def test_muon_colouring_is_distance_two():
    P = qa.muon_scene()
    order, cp = P.colouring()
    assert_distance2_colouring(P, order, cp)
    lower = np.bincount(P.ray).max()
    assert lower <= len(cp) - 1 <= 2 * lower
    # heavy bits (more than 16 entries) come first inside every class
    deg = np.diff(P.row_ptr)[order]
    for c in range(len(cp) - 1):
        h = deg[cp[c]:cp[c + 1]] > 16
        assert not np.any(h[1:] & ~h[:-1])


def neighbour_problem(n, edges, rays=()):
    """A problem whose conflicts are the given neighbour pairs and rays (lists of bits)."""
    nb = [[] for _ in range(n)]
    for a, b in edges:
        nb[a].append(b)
        nb[b].append(a)
    ent = [[] for _ in range(n)]
    for r, bits in enumerate(rays):
        for i in bits:
            ent[i].append(r)
    row_ptr = np.cumsum([0] + [len(e) for e in ent])
    ray = [r for e in ent for r in e]
    nbr_ptr = np.cumsum([0] + [len(x) for x in nb])
    nbr = [j for x in nb for j in x]
    nr = max(1, len(rays))
    return qa.ProblemArrays(row_ptr, ray, np.full(len(ray), 0.5), np.zeros(nr), np.ones(nr), nbr_ptr, nbr, 0.3, 0.1)


def test_saturation_colouring_closed_cases():
    """DSatur is exact on bipartite graphs (Brelaz 1979): a 2-D grid and an even cycle
    take 2 classes, an odd cycle 3, and bits of one ray as many classes as the ray has
    bits. A greedy pass in index order needs 4 classes on the crown graph below (each
    a_i conflicts with every b_j, j != i); saturation order needs 2."""
    g = 9
    grid = [(r * g + c, r * g + c + 1) for r in range(g) for c in range(g - 1)] + \
           [(r * g + c, (r + 1) * g + c) for r in range(g - 1) for c in range(g)]
    for n, edges, k in [(g * g, grid, 2), (10, [(i, (i + 1) % 10) for i in range(10)], 2),
                        (9, [(i, (i + 1) % 9) for i in range(9)], 3)]:
        P = neighbour_problem(n, edges)
        order, cp = P.colouring()
        assert_distance2_colouring(P, order, cp)
        assert len(cp) - 1 == k
    crown = [(2 * i, 2 * j + 1) for i in range(4) for j in range(4) if i != j]
    P = neighbour_problem(8, crown)
    assert len(P.colouring()[1]) - 1 == 2
    P = neighbour_problem(12, [], rays=[list(range(7)), [6, 7, 8, 9, 10, 11]])
    order, cp = P.colouring()
    assert_distance2_colouring(P, order, cp)
    assert len(cp) - 1 == 7


def test_small_scene_colouring_is_distance_two():
    P = qa.muon_scene(grid=32, sky_az=90, sky_th=20)
    assert_distance2_colouring(P, *P.colouring())


def test_er_graph_matches_the_test_generator(orc):
    from test_oracle_pins import er_graph
    g = qa.er_graph(300, 0.05, 7)
    ref = er_graph(orc, 300, 0.05, 7)
    assert np.array_equal(g.eu, ref.eu) and np.array_equal(g.ev, ref.ev) and np.array_equal(g.ew, ref.ew)


def test_er_graph_workload_size():
    g = qa.er_graph(800, 0.06, 1)
    assert abs(g.m - 19176) < 4 * np.sqrt(19176)
    assert set(np.unique(g.ew).tolist()) == {-1, 1}


def test_graph_colouring_is_proper():
    g = qa.er_graph(800, 0.06, 1)
    order, cp = g.colouring()
    col = colours_of(order, cp, g.n)
    assert np.all(col[g.eu] != col[g.ev])
    assert len(cp) - 1 < 40


@pytest.mark.parametrize("eu,ev", [([0, 1, 2], [1, 2, 1]), ([0, 1, 0], [1, 2, 1]), ([0, 3, 2, 3], [3, 0, 1, 1])])
def test_graph_refuses_repeated_edges(eu, ev):
    """An edge listed twice, in either orientation, is refused (src/ising_weighted.hpp:124-133
    refuses it in the loader): the kernels would update the repeated neighbour twice."""
    g = qa.GraphArrays(4, eu, ev, np.ones(len(eu), np.int32))
    order = np.zeros(4, np.int32)
    cp = np.zeros(5, np.int32)
    rc = qa.lib().qb_colour_graph(C.byref(g.desc()), order.ctypes.data, cp.ctypes.data, C.byref(C.c_int()))
    assert rc == qa.QB_E_ARG
    assert "repeats an earlier edge" in qa.lib().qb_last_error().decode()
    ok = qa.GraphArrays(4, [0, 1, 2], [1, 2, 3], np.ones(3, np.int32))
    assert qa.lib().qb_colour_graph(C.byref(ok.desc()), order.ctypes.data, cp.ctypes.data, C.byref(C.c_int())) == 0


def test_reduce_frozen_energy_identity(orc):
    """freeze_provable_rock's reduction (src/ising_recon.hpp:206-233) through the C ABI, on the
    oracle's frozen set: the kept bits are the unfrozen ones in order, and for any state of the
    kept bits the reduced energy equals the full energy with the frozen bits at 0 -- the
    frozen bonds folded into the field exactly (rtol 1e-12: the prior sums in another order)."""
    P = qa.muon_scene(grid=28, sky_az=60, sky_th=12, log2_muons=22)
    op = orc.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
    frozen = np.zeros(P.n_vars, np.uint8)
    orc.lib().orc_freeze_mask(op.c, 10.0, frozen)
    assert 0 < frozen.sum() < P.n_vars
    R, kept = P.reduce_frozen(frozen)
    assert np.array_equal(kept, np.flatnonzero(frozen == 0))
    assert R.n_rays == P.n_rays and R.field is not None
    # a bit with k frozen neighbours carries kappa + lambda * k
    nb_frozen = np.add.reduceat(frozen[P.nbr].astype(np.int64), P.nbr_ptr[:-1])
    nb_frozen[np.diff(P.nbr_ptr) == 0] = 0
    assert np.allclose(R.field, P.kappa + P.lam * nb_frozen[kept], rtol=0, atol=1e-12)
    orr = orc.FloatProblem(R.row_ptr, R.ray, R.a, R.d, R.w, R.nbr_ptr, R.nbr, R.lam, R.kappa, R.field)
    rng = np.random.default_rng(2)
    for _ in range(6):
        x = (rng.random(R.n_vars) < rng.uniform(0.05, 0.5)).astype(np.uint8)
        full = np.zeros(P.n_vars, np.uint8)
        full[kept] = x
        e_full, e_red = orc.binary_energy(op, full), orc.binary_energy(orr, x)
        assert abs(e_full - e_red) <= 1e-12 * max(1.0, abs(e_full)), (e_full, e_red)
