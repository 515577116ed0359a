"""CPU pins of the oracle's problem-builder restatements, the checker of the device
builders: the float march of src/muon.cuh:66-134 against a slab test on every cell and
closed forms; the candidate exposure of src/muon.cuh:195-224 against the cos^2 sky law,
Beer-Lambert through a slab and its empty-medium identity; the rows of
src/op_muon.hpp:22-50 against the invariant d = 0 for counts that match the model; and
the stable assembly against numpy and, per bit, against the reference's own assemble
(src/engine.hpp:120-149, built from the reference's sources in oracle/_ref)."""
import ctypes as C

import numpy as np
import pytest

from test_dropin_cpu import march_segments

C_MIN, TH_MAX = np.float32(0.34202), np.float32(1.22173)


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


@pytest.mark.parametrize("seed", range(6))
def test_march_matches_slab_segments(orc, seed):
    rng = np.random.default_rng(seed)
    shape, voxel, lo = (9, 7, 11), 2.0, (-9.0, -7.0, -3.0)
    med = orc.medium(lo, voxel, shape, None)
    inside = seed % 2 == 0
    p = np.array([rng.uniform(-8, 8), rng.uniform(-6, 6), rng.uniform(-2, 18)] if inside else
                 [rng.uniform(-30, 30), rng.uniform(-30, 30), -10.0], np.float32)
    d = unit(rng.normal(size=3) + np.array([0.0, 0.0, 2.0]))
    tau, cells, segs = orc.march(med, p, d)
    assert tau == 0.0   # no medium: cells only
    want = march_segments(p.astype(np.float64), d.astype(np.float64), lo, voxel, shape)
    got = np.zeros_like(want)
    got[cells] = segs
    assert len(set(cells.tolist())) == len(cells)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * voxel)


def test_march_closed_forms(orc):
    shape, voxel, lo = (4, 5, 6), 1.5, (0.0, 0.0, 0.0)
    mu = np.full(np.prod(shape), 0.25, np.float32)
    med = orc.medium(lo, voxel, shape, mu)
    # straight up a column from the floor: one cell per layer, each a full voxel
    tau, cells, segs = orc.march(med, [2.2, 3.1, 0.0], [0.0, 0.0, 1.0])
    assert np.array_equal(cells, 1 + 2 * 4 + 4 * 5 * np.arange(6))
    assert np.allclose(segs, voxel) and np.isclose(tau, 0.25 * 6 * voxel, rtol=1e-6)
    # along x from a point inside: the march measures from the origin to the far face
    tau, cells, segs = orc.march(med, [0.7, 1.0, 1.0], [1.0, 0.0, 0.0])
    assert np.isclose(segs.sum(), 4 * voxel - 0.7, rtol=1e-6) and np.isclose(tau, 0.25 * segs.sum(), rtol=1e-6)
    # a ray that misses the box
    tau, cells, segs = orc.march(med, [-1.0, -1.0, -1.0], [-1.0, 0.0, 0.0])
    assert tau == 0.0 and len(cells) == 0


def test_hash_known_values(orc):
    """The Wang hash (src/lidar.cuh:90-97) on values worked by hand from its steps."""
    def wang(x):
        x = ((x ^ 61) ^ (x >> 16)) & 0xFFFFFFFF
        x = (x * 9) & 0xFFFFFFFF
        x = x ^ (x >> 4)
        x = (x * 0x27D4EB2D) & 0xFFFFFFFF
        return x ^ (x >> 15)
    L = orc.lib()
    for x in (0, 1, 61, 0x9E3779B9, 0xFFFFFFFF, 123456789):
        assert L.orc_hash_u32(x) == wang(x)
        assert L.orc_hash_unit(x) == np.float32((wang(x) >> 8) / 16777216.0)


def test_exposure_empty_medium_and_sky_law(orc):
    """No attenuation: every candidate is binned and detected (exp(0) = 1 > u). The zenith
    marginal follows p(cos theta) ~ cos^2 on [c_min, 1]: bin ith holds a fraction
    (cos^3 lo - cos^3 hi) / (1 - c_min^3) of the candidates; azimuth is uniform."""
    shape = (4, 4, 4)
    med = orc.medium((-4, -4, -4), 2.0, shape, np.zeros(64, np.float32))
    n_az, n_th, n = 12, 10, 1 << 17
    o, d = orc.muon_expose(med, [0, 0, 0], C_MIN, TH_MAX, n_az, n_th, n, 7)
    assert int(o.sum()) == n and np.array_equal(o, d)
    zen = o.reshape(n_th, n_az).sum(axis=1).astype(np.float64)
    edges = np.cos(np.arange(n_th + 1) / n_th * np.float64(TH_MAX))
    p = (edges[:-1] ** 3 - edges[1:] ** 3) / (1 - np.float64(C_MIN) ** 3)
    chi2 = ((zen - n * p) ** 2 / (n * p)).sum()
    assert chi2 < 60.0, chi2                     # 9 dof; the 1e-9 tail is far below 60
    az = o.reshape(n_th, n_az).sum(axis=0).astype(np.float64)
    assert ((az - n / n_az) ** 2 / (n / n_az)).sum() < 60.0


def test_exposure_beer_lambert_through_a_slab(orc):
    """A horizontal slab of thickness t and attenuation mu above the chamber: a candidate at
    zenith angle theta crosses mu t / cos(theta), so each zenith bin's detected fraction is
    the mean of exp(-mu t / cos theta) over its candidates."""
    nx = ny = 64
    nz = 8
    mu = np.zeros((nz, ny, nx), np.float32)
    mu[5] = 0.2                                    # one layer of 1 m cells: z in [5, 6)
    med = orc.medium((-32, -32, 0), 1.0, (nx, ny, nz), mu.reshape(-1))
    n_az, n_th, n = 8, 6, 1 << 16
    o, d = orc.muon_expose(med, [0.0, 0.0, 0.5], np.float32(np.cos(0.5)), np.float32(0.5), n_az, n_th, n, 3)
    O = o.reshape(n_th, n_az).sum(axis=1).astype(np.float64)
    D = d.reshape(n_th, n_az).sum(axis=1).astype(np.float64)
    th = (np.arange(n_th) + 0.5) / n_th * 0.5
    expect = np.exp(-0.2 / np.cos(th))
    frac = D / O
    sd = np.sqrt(expect * (1 - expect) / O)
    assert np.all(np.abs(frac - expect) < 6 * sd + 0.01), (frac, expect)


def test_rows_of_a_model_that_matches_its_counts(orc):
    """Rows (src/op_muon.hpp:22-50) where the counts follow the model, det = round(open
    exp(-tau0)): d = tau0 - (-log(det/open)) vanishes to within the rounding of det; rows
    of open < min_open or crossing no bit are dropped; with a_scale = mu and every crossed
    cell a bit, a row's entries sum to tau0."""
    shape, voxel, lo = (10, 10, 10), 2.0, (-10.0, -10.0, -2.0)
    mu = np.full(1000, 0.05, np.float32)
    med = orc.medium(lo, voxel, shape, mu)
    site_bit = np.arange(1000, dtype=np.int32)
    site_bit[::7] = -1
    site_bit[site_bit >= 0] = np.arange(int((site_bit >= 0).sum()), dtype=np.int32)
    n_az, n_th = 16, 8
    dirs = []
    for b in range(n_az * n_th):
        ith, iaz = divmod(b, n_az)
        th = np.float32((ith + 0.5) / n_th) * np.float32(1.0)
        az = np.float32((iaz + 0.5) * 6.28318530718 / n_az)
        dirs.append([np.sin(th) * np.cos(az), np.sin(th) * np.sin(az), np.cos(th)])
    dirs = np.array(dirs, np.float32)
    chamber = np.array([0.3, -0.2, 0.1], np.float32)
    open_ = np.full(n_az * n_th, 5000, np.uint64)
    open_[3] = 10                                  # below min_open
    tau0 = np.array([orc.march(med, chamber, dv)[0] for dv in dirs], np.float64)
    det = np.maximum(np.round(open_ * np.exp(-tau0)), 1).astype(np.uint64)
    d, w, eb, er, ea = orc.muon_rows(med, chamber, dirs, open_, det, site_bit, 30.0, np.float32(0.05))
    assert len(d) == n_az * n_th - 1
    assert np.all(np.abs(d) < 1.0 / w + 1e-12)
    assert np.array_equal(w, np.maximum(det[np.arange(len(det)) != 3], 1).astype(np.float64))
    assert np.all(np.diff(er) >= 0)
    full = site_bit.copy()
    full[:] = np.arange(1000)
    d2, w2, eb2, er2, ea2 = orc.muon_rows(med, chamber, dirs, open_, det, full, 30.0, np.float32(0.05))
    sums = np.bincount(er2, weights=ea2.astype(np.float64))
    keep = np.arange(len(tau0)) != 3
    assert np.allclose(sums, tau0[keep], rtol=1e-5)


def test_assemble_stable_against_numpy_and_the_reference(orc):
    rng = np.random.default_rng(4)
    n_bits, n = 50, 700
    eb = rng.integers(0, n_bits, n).astype(np.int32)
    er = np.sort(rng.integers(0, 90, n)).astype(np.int32)
    ea = rng.random(n).astype(np.float32)
    row_ptr, ray, a = orc.assemble_stable(n_bits, eb, er, ea)
    order = np.argsort(eb, kind="stable")
    assert np.array_equal(ray, er[order]) and np.array_equal(a, ea[order])
    assert np.array_equal(row_ptr, np.concatenate([[0], np.cumsum(np.bincount(eb, minlength=n_bits))]))
    R = orc.ref()
    if R is None:
        pytest.skip("reference shim not built")
    # the reference's assemble over one field of n_bits cells: the same entries per bit,
    # whatever order its std::sort leaves inside a bit
    nv, nnz, nn = C.c_int(), C.c_int(), C.c_int()
    dom = np.ones(n_bits, np.int8)
    rows = int(er.max()) + 1
    args = (n_bits, 1, 1, dom.tobytes(), rows, np.zeros(rows), np.ones(rows), n, eb, er, ea)
    R.ref_assemble(*args, C.byref(nv), C.byref(nnz), C.byref(nn), None, None, None, None, None, None)
    rp = np.zeros(nv.value + 1, np.int32)
    rr = np.zeros(nnz.value, np.int32)
    ra = np.zeros(nnz.value, np.float32)
    R.ref_assemble(*args, C.byref(nv), C.byref(nnz), C.byref(nn), rp.ctypes.data, rr.ctypes.data, ra.ctypes.data,
                   None, None, None)
    assert np.array_equal(rp, row_ptr)
    for b in range(n_bits):
        mine = sorted(zip(ray[row_ptr[b]:row_ptr[b + 1]].tolist(), a[row_ptr[b]:row_ptr[b + 1]].tolist()))
        theirs = sorted(zip(rr[rp[b]:rp[b + 1]].tolist(), ra[rp[b]:rp[b + 1]].tolist()))
        assert mine == theirs
