"""Host side of batched serving: the union of independent requests colours each request
exactly as it is coloured alone (the precondition for identical trajectories), and the
union's arrays are the requests' arrays with offsets."""
import numpy as np

import qublar_amd as qa


def colours(P):
    order, cp = P.colouring()
    col = np.empty(P.n_vars, np.int64)
    for c in range(len(cp) - 1):
        col[order[cp[c]:cp[c + 1]]] = c
    return col


def test_union_colouring_equals_each_request():
    shapes = [dict(grid=16, sky_az=30, sky_th=10, log2_muons=20, seed=3),
              dict(grid=22, sky_az=45, sky_th=12, log2_muons=21, seed=8),
              dict(grid=18, sky_az=36, sky_th=9, log2_muons=20, seed=11)]
    probs = [qa.muon_scene(**s) for s in shapes]
    U, offs = qa.union(probs, [5, 17, 99])
    U.colour = None            # the engine's own colouring of the whole union
    cu = colours(U)
    for P, (b0, r0) in zip(probs, offs):
        assert np.array_equal(cu[b0:b0 + P.n_vars], colours(P))


def test_given_colouring_is_used_and_checked():
    P = qa.muon_scene(grid=16, sky_az=30, sky_th=10, log2_muons=20, seed=3)
    own = colours(P)
    # a valid colouring with the classes renumbered in reverse is taken as given
    K = int(own.max()) + 1
    P.colour = (K - 1 - own).astype(np.int32)
    assert np.array_equal(colours(P), K - 1 - own)
    # two bits that share a ray may not share a colour
    bad = own.copy()
    i = int(np.argmax(np.diff(P.row_ptr) > 0))
    ray = P.ray[P.row_ptr[i]]
    j = next(b for b in range(P.n_vars) if b != i and ray in P.ray[P.row_ptr[b]:P.row_ptr[b + 1]])
    bad[j] = bad[i]
    P.colour = bad
    try:
        P.colouring()
        raise AssertionError("a conflicting colouring was accepted")
    except qa.QublarError as e:
        assert "share ray" in str(e)


def test_union_arrays():
    probs = [qa.muon_scene(grid=16, sky_az=30, sky_th=10, log2_muons=20, seed=s) for s in (1, 2)]
    U, offs = qa.union(probs, [7, 9])
    assert U.n_vars == sum(P.n_vars for P in probs) and U.n_rays == sum(P.n_rays for P in probs)
    b1, r1 = offs[1]
    P = probs[1]
    k0 = int(U.row_ptr[b1])
    assert np.array_equal(U.ray[k0:k0 + int(P.row_ptr[-1])], P.ray + r1)
    assert np.array_equal(U.draw_bit[b1:], np.arange(P.n_vars))
    assert np.all(U.draw_key[:b1] == 7) and np.all(U.draw_key[b1:] == 9)
    assert np.allclose(U.field, P.kappa)


def test_branch_agreement_counts_the_requests_own_branches():
    from qublar_amd.serving import branch_agreement

    rng = np.random.default_rng(4)
    X = (rng.random((12, 500)) < rng.random(500)).astype(np.uint8)
    X[:, 0] = 1
    X[:, 1] = 0
    X[:11, 2] = 1        # 11 of 12: 0.9166... >= 0.9
    pv, tri = branch_agreement(X)
    assert pv.dtype == np.float32
    assert np.array_equal(pv, (X.sum(0) / 12.0).astype(np.float32))
    assert (pv[0], tri[0], pv[1], tri[1], tri[2]) == (1.0, 1, 0.0, 0, 1)
    assert np.array_equal(tri, np.where(pv >= 0.9, 1, np.where(pv <= 0.1, 0, 2)))
    # a request of 8 branches in a union of 12 sees only its 8
    pv8, _ = branch_agreement(X[:8])
    assert np.array_equal(pv8, (X[:8].sum(0) / 8.0).astype(np.float32))
