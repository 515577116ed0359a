"""CPU checks of the gravimetry scene builder (qb_scene_gravity): its entries against
the point-mass law of src/op_gravity.hpp:49-57 evaluated here by brute force over every
(station, cell) pair, the law's constants pinned by the published Bouguer plate
constant (a truncated slab of cells), the row truncation, the data (the true body's
anomaly plus noise of the declared sigma), and the builder's colouring (no two bits of
a class share a station or are neighbours, and each layer's lattice period is the
smallest that is)."""
import numpy as np
import pytest

import qublar_amd as qa

G = 6.67430e-11          # m^3 kg^-1 s^-2
MICROGAL = 1e-8          # m / s^2


def cell_centres(nx, ny, nz, h):
    b = np.arange(nx * ny * nz)
    i, j, k = b % nx, (b // nx) % ny, b // (nx * ny)
    return np.stack([(i + 0.5) * h, (j + 0.5) * h, -(k + 0.5) * h], 1)


def station_positions(nx, ny, h):
    s = np.arange(nx * ny)
    return np.stack([(s % nx + 0.5) * h, (s // nx + 0.5) * h, np.zeros(nx * ny)], 1)


def dense_operator(P, nx, ny, nz):
    A = np.zeros((P.n_rays, P.n_vars))
    for b in range(P.n_vars):
        A[P.ray[P.row_ptr[b]:P.row_ptr[b + 1]], b] = P.a[P.row_ptr[b]:P.row_ptr[b + 1]]
    return A


SMALL = dict(nx=14, ny=11, nz=5, h=4.0, radius=13.0, drho=800.0, sigma=3.0, body_lo=(16.0, 12.0, -14.0),
             body_hi=(32.0, 28.0, -4.0), seed=5)


def test_entries_follow_the_point_mass_law_within_the_radius():
    P = qa.gravity_scene(**SMALL)
    nx, ny, nz, h, R = SMALL["nx"], SMALL["ny"], SMALL["nz"], SMALL["h"], SMALL["radius"]
    c = cell_centres(nx, ny, nz, h)
    st = station_positions(nx, ny, h)
    diff = st[:, None, :] - c[None, :, :]
    r = np.sqrt((diff ** 2).sum(-1))
    expect = np.where(r <= R, G * SMALL["drho"] * h ** 3 * diff[..., 2] / r ** 3 / MICROGAL, 0.0)
    A = dense_operator(P, nx, ny, nz)
    assert np.array_equal(A != 0, r <= R)
    assert np.allclose(A, expect, rtol=1e-6, atol=0)
    assert P.row_ptr[-1] == int((r <= R).sum())
    # rows are listed by station index, ascending, inside every bit
    for b in range(0, P.n_vars, 37):
        rows = P.ray[P.row_ptr[b]:P.row_ptr[b + 1]]
        assert np.all(np.diff(rows) > 0)
    assert np.all(P.a > 0)   # every cell lies below every station


def test_truncated_bouguer_slab_fixes_the_constants():
    """Under a station in the middle of a wide grid, the cells of layers k0.. form a slab
    from depth z1 to z2; cut at radius R it pulls with 2 pi G rho ((z2 - z1) - (z2^2 -
    z1^2) / (2 R)) (a thin disc of radius sqrt(R^2 - z^2) gives 2 pi G sigma (1 - z / R)).
    2 pi G rho is the Bouguer plate constant, published as 0.04193 mGal per metre of
    1 g/cm^3 rock, which pins G, the microGal unit, the cell volume and the cut-off at
    once. Layers below k0 keep the point-mass approximation of each cell accurate."""
    bouguer = 41.93      # microGal per metre at 1000 kg/m^3
    for n, h, nz, R, k0 in ((41, 1.0, 12, 20.0, 4), (81, 0.5, 24, 20.0, 8)):
        P = qa.gravity_scene(nx=n, ny=n, nz=nz, h=h, radius=R, drho=1000.0, sigma=1.0)
        st = (n // 2) * n + n // 2
        bit = np.repeat(np.arange(P.n_vars), np.diff(P.row_ptr))
        sel = (P.ray == st) & (bit // (n * n) >= k0)
        z1, z2 = k0 * h, nz * h
        expect = bouguer * ((z2 - z1) - (z2 * z2 - z1 * z1) / (2 * R))
        assert abs(P.a[sel].astype(np.float64).sum() / expect - 1) < 1e-3


def test_data_is_the_body_anomaly_plus_declared_noise():
    s = dict(SMALL, nx=60, ny=50, nz=3, radius=10.0, sigma=2.5)
    P = qa.gravity_scene(**s)
    c = cell_centres(60, 50, 3, SMALL["h"])
    inside = np.all((c >= np.array(s["body_lo"])) & (c <= np.array(s["body_hi"])), 1)
    assert np.array_equal(P.truth.astype(bool), inside) and inside.sum() > 0
    A = dense_operator(P, 60, 50, 3)
    noise = P.d - A.astype(np.float64) @ P.truth
    assert abs(noise.mean()) < 4 * 2.5 / np.sqrt(len(noise))
    assert abs(noise.std() / 2.5 - 1.0) < 0.05
    assert np.all(P.w == 1.0 / 2.5 ** 2)
    assert P.lam == 2.0 and abs(P.kappa - np.log(999.0)) < 1e-12


def test_colouring_is_proper_and_each_layer_period_is_minimal():
    P = qa.gravity_scene(**SMALL)
    nx, ny, nz = SMALL["nx"], SMALL["ny"], SMALL["nz"]
    col = P.colour
    bit = np.repeat(np.arange(P.n_vars), np.diff(P.row_ptr))
    # no two bits of a class share a station
    key = P.ray.astype(np.int64) * (col.max() + 1) + col[bit]
    assert len(np.unique(key)) == len(key)
    # no neighbours share a colour
    nb_src = np.repeat(np.arange(P.n_vars), np.diff(P.nbr_ptr))
    assert np.all(col[nb_src] != col[P.nbr])
    # the library's own check accepts it and keeps the classes
    order, cp = P.colouring()
    assert len(cp) - 1 == len(np.unique(col))
    # layer k: colour = base_k + i % P_k + P_k (j % P_k); one period less makes two cells of
    # a class share a station somewhere in a big enough grid
    big = dict(SMALL, nx=60, ny=60)
    Pb = qa.gravity_scene(**big)
    colb = Pb.colour.reshape(nz, 60, 60)
    bitb = np.repeat(np.arange(Pb.n_vars), np.diff(Pb.row_ptr))
    for k in range(nz):
        lay = colb[k] - colb[k].min()
        per = int(np.flatnonzero(lay[0] == 0)[1])           # first repeat along x
        assert np.array_equal(lay, (np.arange(60)[None, :] % per) + per * (np.arange(60)[:, None] % per))
        if per > 2:
            smaller = per - 1
            cls = (np.arange(60)[None, :] % smaller) + smaller * (np.arange(60)[:, None] % smaller)
            cells = np.arange(k * 3600, (k + 1) * 3600)
            sel = np.isin(bitb, cells)
            c2 = cls.reshape(-1)[bitb[sel] - k * 3600]
            key = Pb.ray[sel].astype(np.int64) * 3600 + c2
            assert len(np.unique(key)) < len(key), f"layer {k}: period {smaller} would also be proper"


def test_scene_refusals():
    for bad in (dict(nx=0), dict(h=0.0), dict(sigma=0.0), dict(radius=-1.0), dict(nx=5000, ny=5000)):
        with pytest.raises(qa.QublarError):
            qa.gravity_scene(**bad)
    # the muon builder gives no colouring of its own
    assert qa.muon_scene(grid=16, sky_az=20, sky_th=8).colour is None


def test_default_scene_shape():
    """The precision study's gravity QUBO: 100 x 200 x 10 cells of 5 m, 20,000 stations,
    rows to 50 m."""
    P = qa.gravity_scene()
    assert (P.n_vars, P.n_rays) == (200_000, 20_000)
    deg = np.diff(P.row_ptr)
    assert deg.max() <= 317 and 150 < deg.mean() < 250
    assert P.colour.max() + 1 < 3000
    assert P.truth.sum() == 6 * 8 * 4
