"""The drop-in C++ headers (include/argos/*) on the CPU: a reference-style driver
builds against them and passes its host checks; the model layer assembles problems
array-identical to the reference's own src/engine.hpp (compiled into oracle/_ref)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "check_dropin")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return DRIVER


def test_driver_host_checks(driver):
    out = subprocess.run([driver, "--host"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "FAIL" not in out.stdout and out.stdout.count("PASS") >= 6


RUDY_CORPUS = [
    # accepted: plain, weights, signs, blank lines, CR LF, tabs, an edge on the header line,
    # leading zeros, the 2^40 magnitude cap on a vertex token, trailing blank lines
    "3 2\n1 2\n2 3 -7\n",
    "4 3\n1 2 5\n3 4 +2\n1 4 -0\n",
    "\n\n  5 2\n\n\t1 5\t9\r\n2 4 -2147483648\r\n\n",
    "3 1 1 2\n",
    "3 2 1 2 4\n2 3\n",
    "3 1\n0001 003 2147483647\n",
    "2 1\n2 1 1\n\n\n",
    "6 0\n",
    "1 0\n",
    "3 2\n1 2\n3 2",
    # refused: non-integer / out-of-range weights, self-loops, duplicates in either
    # orientation, indices out of range, counts that differ, malformed headers and lines
    "3 1\n1 2 0.5\n", "3 1\n1 1 1\n", "3 2\n1 2 1\n2 1 1\n", "3 2\n1 2 1\n", "3 1\n1 4 1\n",
    "3 2\n1 2 1\n1 2 3\n", "3 1\n1 2 2147483648\n", "3 1\n1 2 1 4\n", "3 1\n1\n2 3\n",
    "3\n1\n1 2\n", "0 0\n", "-3 0\n", "3 -1\n", "x 1\n1 2\n", "3 1\n1 x\n", "3 1\n1 2 w\n",
    "3 1\n0 2\n", "3 1\n1 1099511627777\n", "3 1\n1 +\n", "3 1\n1 +-2\n", "", "   \n\n",
    "3 1 1\n2 3\n", "1073741825 0\n", "3 1\n1 2\f\n", "3 3\n1 2\n2 3\n",
]


def test_parse_rudy_equals_reference_loader(orc, driver, tmp_path):
    """The drop-in parse_rudy against the reference's own loader (src/ising_weighted.hpp:83-149,
    compiled into oracle/_ref) on a corpus of accepted and refused files: the same files are
    refused, and every accepted file parses to the same graph (edge lists, CSR, weight
    counts; one FNV-1a digest of all of them)."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference shim not built")
    R.ref_parse_rudy_digest.restype = C.c_int
    R.ref_parse_rudy_digest.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]
    paths, want = [], []
    for k, text in enumerate(RUDY_CORPUS):
        b = text.encode()
        path = tmp_path / f"g{k}.rudy"
        path.write_bytes(b)
        paths.append(str(path))
        dg = C.c_uint64()
        rc = R.ref_parse_rudy_digest(b, len(b), C.byref(dg))
        want.append("refused" if rc else f"ok {dg.value}")
    out = subprocess.run([driver, "--rudy-digest"] + paths, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    got = out.stdout.split("\n")[:len(paths)]
    for text, g, w in zip(RUDY_CORPUS, got, want):
        assert g == w, (text, g, w)
    assert sum(w.startswith("ok") for w in want) == 10


def test_host_anneal_equals_reference_code(orc, driver, tmp_path):
    """The drop-in wising::anneal is the host annealer, as the reference's: seed for seed it
    equals the reference's own src/ising_weighted.hpp:250-267 compiled into oracle/_ref
    (assignment and tracked energy), on a signed K16 and a sparse +-1 graph."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference shim not built")
    rng = np.random.default_rng(5)
    iu = np.triu_indices(16, 1)
    graphs = [(16, iu[0], iu[1], rng.integers(-1000, 1001, len(iu[0])), (5.0, 1e-3, 300, 20))]
    pairs = sorted({(int(min(u, v)), int(max(u, v))) for u, v in rng.integers(0, 200, (1500, 2)) if u != v})
    eu, ev = np.array(pairs).T
    graphs.append((200, eu, ev, rng.choice([-1, 1], len(eu)), (3.0, 0.05, 200, 10)))
    for k, (n, eu, ev, ew, b) in enumerate(graphs):
        path = tmp_path / f"g{k}.rudy"
        path.write_text(f"{n} {len(eu)}\n" + "".join(f"{u + 1} {v + 1} {w}\n" for u, v, w in zip(eu, ev, ew)))
        for seed in (1, 2, 7):
            out = subprocess.run([driver, "--anneal-int", str(path), *map(str, b), str(seed)], capture_output=True,
                                 text=True, timeout=60)
            assert out.returncode == 0, out.stderr
            _, e, _, xs = out.stdout.split()
            x = np.zeros(n, np.uint8)
            e_ref = C.c_int64()
            R.ref_anneal_int(n, len(eu), np.ascontiguousarray(eu, np.int32), np.ascontiguousarray(ev, np.int32),
                             np.ascontiguousarray(ew, np.int32), b[0], b[1], b[2], b[3], seed, x, C.byref(e_ref))
            assert int(e) == e_ref.value
            assert xs == "".join("1" if v else "0" for v in x)


def parse_dump(text):
    out = {}
    for line in text.strip().splitlines():
        key, *vals = line.split()
        if key == "a":
            out[key] = np.array([float.fromhex(v) for v in vals], np.float32)
        elif key == "adjoint":
            out[key] = float.fromhex(vals[0])
        else:
            out[key] = np.array([int(v) for v in vals], np.int32)
    return out


@pytest.mark.parametrize("shape,seed,rows", [((4, 3, 3), 9, 5), ((9, 7, 6), 3, 60), ((12, 10, 8), 77, 300)])
def test_assemble_matches_reference(orc, driver, shape, seed, rows):
    R = orc.ref()
    if R is None:
        pytest.skip("reference shim not built")
    nx, ny, nz = shape
    got = parse_dump(subprocess.run([driver, "--assemble", str(nx), str(ny), str(nz), str(seed), str(rows)],
                                    capture_output=True, text=True, check=True).stdout)
    # the same inputs, drawn in the same order, for the reference's assemble
    L = orc.lib()
    s = C.c_uint64(seed)
    sm = lambda: L.orc_splitmix(C.byref(s))
    u = lambda: (sm() >> 11) * 2.0 ** -53
    dom = np.array([1 if sm() % 4 != 0 else 0 for _ in range(nx * ny * nz)], np.int8)
    n_bits = int(dom.sum())
    d, w, eb, er, ea = [], [], [], [], []
    for r in range(rows):
        d.append(u()); w.append(1.0 + u())
        for _ in range(1 + sm() % 8):
            eb.append(sm() % n_bits); er.append(r); ea.append(np.float32(u()))
    eb, er, ea = np.array(eb, np.int32), np.array(er, np.int32), np.array(ea, np.float32)
    nv, nnz, nn = C.c_int(), C.c_int(), C.c_int()
    args = (nx, ny, nz, dom.tobytes(), rows, np.array(d), np.array(w), len(eb), eb, er, ea)
    R.ref_assemble(*args, C.byref(nv), C.byref(nnz), C.byref(nn), None, None, None, None, None, None)
    ref = dict(row_ptr=np.zeros(nv.value + 1, np.int32), ray=np.zeros(nnz.value, np.int32),
               a=np.zeros(nnz.value, np.float32), nbr_ptr=np.zeros(nv.value + 1, np.int32),
               nbr=np.zeros(nn.value, np.int32), var_voxel=np.zeros(nv.value, np.int32))
    R.ref_assemble(*args, C.byref(nv), C.byref(nnz), C.byref(nn), *[ref[k].ctypes.data for k in
                   ("row_ptr", "ray", "a", "nbr_ptr", "nbr", "var_voxel")])
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    mis = R.ref_adjoint_mismatch(n_bits, rows, len(eb), eb, er, ea, 7)
    assert got["adjoint"] == mis


def march_segments(p, d, lo, voxel, shape):
    """Per cell of the grid (x-fastest), the segment the reference's march gives it,
    from the slab test on every cell in double precision. The march measures cell exits
    from the origin p but ends at exit - entry of the grid, and its first segment starts
    at 0 (src/muon.cuh:90, 115-128): for p inside the grid these are the cells' lengths
    along the ray; for p outside, the first cell's segment includes the gap from p and the
    march stops early."""
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.T.reshape(-1), j.T.reshape(-1), k.T.reshape(-1)      # x fastest
    cell_lo = np.stack([lo[0] + i * voxel, lo[1] + j * voxel, lo[2] + k * voxel], 1)
    t0 = np.zeros(len(i))
    t1 = np.full(len(i), np.inf)
    for a in range(3):
        if d[a] == 0.0:
            inside = (p[a] >= cell_lo[:, a]) & (p[a] <= cell_lo[:, a] + voxel)
            t1 = np.where(inside, t1, -np.inf)
            continue
        ta = (cell_lo[:, a] - p[a]) / d[a]
        tb = (cell_lo[:, a] + voxel - p[a]) / d[a]
        t0 = np.maximum(t0, np.minimum(ta, tb))
        t1 = np.minimum(t1, np.maximum(ta, tb))
    seg = np.zeros(len(i))
    crossed = np.flatnonzero(t1 > t0)
    if len(crossed) == 0:
        return seg
    crossed = crossed[np.argsort(t1[crossed])]
    end = t1[crossed].max() - max(t0[crossed].min(), 0.0)
    done = 0.0
    for c in crossed:
        t = min(t1[c], end)
        seg[c] = max(t - done, 0.0)
        if t >= end:
            break
        done = t
    return seg


def parse_muon_dump(text):
    lines = [ln.split() for ln in text.strip().splitlines()]
    out = {"views": []}
    hexf = lambda v: np.array([float.fromhex(x) for x in v])
    for key, *vals in lines:
        if key == "grid":
            out["shape"] = tuple(int(v) for v in vals[:3])
            out["lo"] = hexf(vals[3:6])
            out["voxel"], out["a_scale"] = float.fromhex(vals[6]), float.fromhex(vals[7])
        elif key == "view":
            out["views"].append({"chamber": hexf(vals)})
        elif key == "dirs":
            out["views"][-1]["dirs"] = hexf(vals).reshape(-1, 3)
        elif key in ("open", "det"):
            out["views"][-1][key] = np.array([int(v) for v in vals])
        elif key in ("mu", "a", "d", "w"):
            out[key] = hexf(vals)
        else:
            out[key] = np.array([int(v) for v in vals])
    return out


@pytest.mark.parametrize("shape,seed,bins", [((6, 5, 4), 3, 12), ((14, 11, 9), 8, 60), ((20, 9, 13), 5, 90)])
def test_build_binary_problem_muon_rows(driver, shape, seed, bins):
    """build_binary_problem (src/ising_recon.hpp:52-62) over the muon rows
    (src/op_muon.hpp:22-50): every populated bin whose ray crosses a bit is one row with
    d = tau0 - t, w = max(det, 1); every bit the ray crosses carries a = a_per_metre * its
    segment. Segments, tau0 and the set of crossed cells are pinned by the slab test in
    double precision on every cell (the float march agrees to ~1e-5 m), for a chamber
    inside the grid and for one outside it."""
    P = parse_muon_dump(subprocess.run([driver, "--muon", *map(str, shape), str(seed), str(bins)],
                                       capture_output=True, text=True, check=True).stdout)
    dom = P["domain"].astype(bool)
    assert np.array_equal(P["var_voxel"], np.flatnonzero(dom))
    site_bit = np.full(dom.size, -1)
    site_bit[dom] = np.arange(dom.sum())
    # the problem's entries regrouped by row
    bit_of = np.repeat(np.arange(len(P["row_ptr"]) - 1), np.diff(P["row_ptr"]))
    rows = {}
    for b, r, a in zip(bit_of, P["ray"], P["a"]):
        rows.setdefault(int(r), {})[int(b)] = a
    row = 0
    crossing_rays = 0
    for v in P["views"]:
        for b in range(len(v["open"])):
            if v["open"][b] < 30:
                continue
            L = march_segments(v["chamber"], v["dirs"][b], P["lo"], P["voxel"], P["shape"])
            bits_crossed = (L > 1e-4) & dom
            ambiguous = (L > 0) & (L <= 1e-4) & dom
            if not bits_crossed.any():
                assert not ambiguous.any(), "a ray grazing bits only: pick another seed"
                continue
            crossing_rays += 1
            got = rows.get(row, {})
            want = {int(site_bit[c]): P["a_scale"] * L[c] for c in np.flatnonzero(bits_crossed)}
            assert set(want) <= set(got) and set(got) - set(want) <= {int(site_bit[c]) for c in np.flatnonzero(ambiguous)}
            for bit, a in want.items():
                assert abs(got[bit] - a) <= 1e-4 * abs(P["a_scale"]), (row, bit, got[bit], a)
            tau0 = float((P["mu"] * L).sum())
            det = max(float(v["det"][b]), 1.0)
            t = -np.log(det / float(v["open"][b]))
            assert abs(P["d"][row] - (tau0 - t)) < 1e-4, (row, P["d"][row], tau0 - t)
            assert P["w"][row] == det
            row += 1
    assert row == len(P["d"]) == len(P["w"]) and crossing_rays > 5
    # the same rows from the oracle's own float march (oracle/qb_oracle.c, a separate
    # restatement of src/muon.cuh:66-134 and src/op_muon.hpp:22-50): every entry's a, every
    # d and w, bit for bit; only the order inside a bit is std::sort's in the drop-in
    import oracle

    med = oracle.medium(P["lo"], P["voxel"], P["shape"], P["mu"].astype(np.float32))
    row0 = 0
    want_rows = {}
    for v in P["views"]:
        d, w, eb, er, ea = oracle.muon_rows(med, v["chamber"].astype(np.float32), v["dirs"].astype(np.float32),
                                            v["open"].astype(np.uint64), v["det"].astype(np.uint64),
                                            site_bit.astype(np.int32), 30.0, np.float32(P["a_scale"]), row0)
        for b, r, a in zip(eb, er, ea):
            want_rows.setdefault(int(r), {})[int(b)] = float(a)
        assert np.array_equal(P["d"][row0:row0 + len(d)], d) and np.array_equal(P["w"][row0:row0 + len(w)], w)
        row0 += len(d)
    assert row0 == len(P["d"])
    assert {r: {b: float(np.float32(a)) for b, a in bits.items()} for r, bits in rows.items()} == want_rows


def test_reference_drivers_link_and_fail_loudly_without_a_device():
    """The reference's own tools/gpu_anneal.cu and tools/gset_bench.cu, compiled unchanged
    against include/argos (make -C tests/cpp drivers), load the engine library. gset_bench's
    host-only checks (the rudy loader's refusals, tools/gset_bench.cu:313-333) pass; where
    there is no device its first anneal then stops with the engine's error, as there is no
    CPU fallback. The device half runs in tests/test_gpu_ref_drivers.py."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "gset_bench")):
        pytest.skip("the reference's drivers are built only where the reference tree exists")
    r = subprocess.run([os.path.join(ref, "gpu_anneal")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage: gpu_anneal" in r.stderr
    # line-buffered, so the lines printed before the abort are not lost with its stdio buffer
    r = subprocess.run(["stdbuf", "-oL", os.path.join(ref, "gset_bench"), "--check"], capture_output=True, text=True,
                       timeout=60)
    lines = [ln for ln in r.stdout.splitlines() if " PASS " in ln]
    assert any("refuses 5 malformed inputs" in ln for ln in lines)
    assert any("missing weight is +1" in ln for ln in lines)
    if not has_device():
        assert r.returncode != 0 and "EngineError" in r.stderr


def has_device():
    """True where the engine finds a GPU (then the drivers' device checks run instead of
    failing; tests/test_gpu_ref_drivers.py covers them there)."""
    import qublar_amd as qa

    try:
        qa.device_info()
        return True
    except qa.QublarError:
        return False
