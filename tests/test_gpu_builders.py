"""The device problem builders against the CPU oracle and the host builders, through the
C ABI on the GPU.

- The candidate exposure (src/muon.cuh:195-224): every candidate is binned, and the
  open and detected counts equal the oracle's up to the candidates whose direction or
  detection draw falls within a float ulp of a bin edge or of exp(-tau) -- the device's
  cbrt, sin, cos, acos and exp against the host's libm, as in the reference (its kernel
  runs these on the device). The bar: at most 2e-4 of the candidates move.
- The rows, assembly and neighbour lists (src/op_muon.hpp:22-50, src/engine.hpp:61-149)
  from the same counts: identical to the oracle's rows and stable assembly, entry for
  entry (row_ptr, ray, a, w bit-exact; d within 1e-14 relative: the device's log).
- The muon scene built on the device: its bits, neighbour lists and truth identical to
  the host builder's, its rows those of its own exposure, and an anneal of it on the
  device equal to the oracle's.
- The gravimetry rows built on the device: the host builder's problem, array for array.
"""
import os
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, os.cpu_count() or 1))
C_MIN, TH_MAX = np.float32(0.34202), np.float32(1.22173)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pyramid(N):
    """The muon scene's media (float, as the device builder classifies its cells): true
    (with the void) and model attenuation, the domain's site_bit, voxel and corner."""
    voxel = np.float32(256.0 / N)
    lo = np.array([-128, -128, -10], np.float32)
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x = lo[0] + (i.astype(np.float32) + np.float32(0.5)) * voxel
    y = lo[1] + (j.astype(np.float32) + np.float32(0.5)) * voxel
    z = lo[2] + (k.astype(np.float32) + np.float32(0.5)) * voxel
    wz = np.float32(115.0) * (np.float32(1.0) - z / np.float32(139.0))
    rock = (z >= 0) & (z <= 139) & (np.abs(x) <= wz) & (np.abs(y) <= wz)
    hv = np.float32(15.0)
    vd = rock & (np.abs(x) <= hv) & (np.abs(y) <= hv) & (np.abs(z - np.float32(77.0)) <= hv)
    room = np.zeros_like(rock)
    for c in ((0, 0, 40), (-40, 0, 25), (40, 0, 25)):
        dx, dy, dz = x - np.float32(c[0]), y - np.float32(c[1]), z - np.float32(c[2])
        room |= dx * dx + dy * dy + dz * dz < np.float32(9.0)
    mu_rock = np.float32(0.046)
    mu_air = mu_rock * np.float32(1.2 / 2700.0)
    model = np.where(rock, mu_rock, mu_air).astype(np.float32).reshape(-1)
    true = np.where(rock & ~vd, mu_rock, mu_air).astype(np.float32).reshape(-1)
    dom = (rock & ~room).reshape(-1)
    site_bit = np.full(N ** 3, -1, np.int32)
    site_bit[dom] = np.arange(int(dom.sum()), dtype=np.int32)
    return true, model, site_bit, float(voxel), lo


def bin_dirs(n_az, n_th):
    out = []
    for b in range(n_az * n_th):
        ith, iaz = divmod(b, n_az)
        th = np.float32(np.float32(ith + 0.5) / np.float32(n_th) * TH_MAX)
        az = np.float32(np.float32(iaz + 0.5) * np.float32(6.28318530718) / np.float32(n_az))
        s = np.float32(np.sin(th))
        out.append([s * np.float32(np.cos(az)), s * np.float32(np.sin(az)), np.float32(np.cos(th))])
    return np.array(out, np.float32)


CHAMBERS = np.array([[0, 0, 40], [-40, 0, 25], [40, 0, 25]], np.float32)


def test_exposure_and_rows_against_the_oracle():
    N, n_az, n_th, n_cand = 40, 48, 20, 1 << 20
    true, model, site_bit, voxel, lo = pyramid(N)
    opens, dets = [], []
    for v, c in enumerate(CHAMBERS):
        say(f"exposure: view {v}, {n_cand} candidates, grid {N}")
        o, d = qa.muon_expose(qa.voxel_medium(lo, voxel, (N, N, N), true), c, n_az, n_th, n_cand, 11)
        ro, rd = orc_mod.muon_expose(orc_mod.medium(lo, voxel, (N, N, N), true), c, C_MIN, TH_MAX, n_az, n_th,
                                     n_cand, 11)
        assert int(o.sum()) == n_cand
        moved_o = int(np.abs(o.astype(np.int64) - ro.astype(np.int64)).sum())
        moved_d = int(np.abs(d.astype(np.int64) - rd.astype(np.int64)).sum())
        say(f"  candidates in another bin: {moved_o}, detections differing: {moved_d}")
        assert moved_o <= 2e-4 * n_cand and moved_d <= 2e-4 * n_cand
        opens.append(o)
        dets.append(d)
    dirs = np.concatenate([bin_dirs(n_az, n_th)] * 3)
    say("rows, assembly and neighbours on the device")
    P = qa.muon_rows_device(qa.voxel_medium(lo, voxel, (N, N, N), model), CHAMBERS, dirs, np.concatenate(opens),
                            np.concatenate(dets), site_bit, 30.0, np.float32(0.046))
    med = orc_mod.medium(lo, voxel, (N, N, N), model)
    d, w, eb, er, ea = [], [], [], [], []
    row0 = 0
    for v in range(3):
        rd = orc_mod.muon_rows(med, CHAMBERS[v], bin_dirs(n_az, n_th), opens[v], dets[v], site_bit, 30.0,
                               np.float32(0.046), row0)
        for acc, x in zip((d, w, eb, er, ea), rd):
            acc.append(x)
        row0 += len(rd[0])
    d, w, eb, er, ea = (np.concatenate(x) for x in (d, w, eb, er, ea))
    n_bits = int((site_bit >= 0).sum())
    row_ptr, ray, a = orc_mod.assemble_stable(n_bits, eb, er, ea)
    assert P.n_vars == n_bits and P.n_rays == len(d)
    assert np.array_equal(P.row_ptr, row_ptr) and np.array_equal(P.ray, ray) and np.array_equal(P.a, a)
    assert np.array_equal(P.w, w)
    assert np.allclose(P.d, d, rtol=1e-14, atol=1e-14)
    H = qa.muon_scene(grid=N, sky_az=n_az, sky_th=n_th)
    assert np.array_equal(P.nbr_ptr, H.nbr_ptr) and np.array_equal(P.nbr, H.nbr)
    say(f"identical: {n_bits} bits, {len(d)} rows, {len(ray)} entries")


def test_muon_scene_on_the_device():
    """The device pipeline at a small grid: the host builder's bits, neighbours and truth;
    its own rows (every bin of each chamber that crosses a bit); then branches annealed on
    the device equal the oracle's."""
    P, ms = qa.muon_scene_device(grid=32, sky_az=60, sky_th=15, log2_muons=22)
    H = qa.muon_scene(grid=32, sky_az=60, sky_th=15, log2_muons=22)
    say(f"device scene: {P.n_vars} bits, {P.n_rays} rays, {int(P.row_ptr[-1])} entries in {ms:.1f} ms")
    assert P.n_vars == H.n_vars and np.array_equal(P.nbr_ptr, H.nbr_ptr) and np.array_equal(P.nbr, H.nbr)
    assert np.array_equal(P.truth, H.truth)
    assert 0 < P.n_rays <= 3 * 60 * 15 and np.all(P.w >= 1.0) and np.all(np.isfinite(P.d))
    prob = qa.Problem(P)
    sched = qa.schedule(50.0, 1.0, 12, 4)
    out = prob.anneal_branches(sched, 8, 0, qa.FP32, 4)
    order, cp = prob.colouring()
    op = orc_mod.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
    mode = orc_mod.make_mode(orc_mod.RNG_PHILOX, order, cp, orc_mod.FP32, 4)
    pv, br = orc_mod.branch_fractions(op, orc_mod.schedule(50.0, 1.0, 12, 4), 8, THREADS, 0, mode)
    assert np.array_equal(out["branches"], br) and np.array_equal(out["p_void"], pv)


def test_muon_scene_on_the_device_at_full_size():
    """The 128^3 scene with 2^27 candidates per chamber: the host builder's 306,328 bits
    and neighbour lists; rows from every sky bin that crosses a bit."""
    P, ms = qa.muon_scene_device()
    H = qa.muon_scene()
    say(f"device scene: {P.n_vars} bits, {P.n_rays} rays, {int(P.row_ptr[-1])} entries in {ms:.1f} ms")
    assert P.n_vars == 306328 and np.array_equal(P.nbr, H.nbr) and np.array_equal(P.truth, H.truth)
    assert 70000 <= P.n_rays <= 75600 and np.all(np.diff(P.row_ptr) >= 0)


@pytest.mark.parametrize("shape", [(30, 40, 6), (100, 200, 10)])
def test_gravity_rows_on_the_device(shape):
    nx, ny, nz = shape
    G, ms = qa.gravity_scene_device(nx=nx, ny=ny, nz=nz)
    H = qa.gravity_scene(nx=nx, ny=ny, nz=nz)
    say(f"gravity {shape}: {int(G.row_ptr[-1])} entries on the device in {ms:.1f} ms")
    for k in ("row_ptr", "ray", "a", "d", "w", "nbr_ptr", "nbr", "truth", "colour"):
        assert np.array_equal(getattr(G, k), getattr(H, k)), k
