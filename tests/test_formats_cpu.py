"""The export formats of qublar_amd.formats against the reference's writers and readers
(src/check_ising.cu:507-606, src/ledger_out.hpp:78-86, tools/export_machine_map.py,
tools/branches_blaze.py:34-43): byte layout, text formatting, thresholds and rounding."""
import json
import os

import numpy as np

from qublar_amd import formats as fm


def reference_load_meta(path):
    # tools/export_machine_map.py: dict(line.split(maxsplit=1) for non-empty lines)
    with open(path) as f:
        return dict(line.split(maxsplit=1) for line in f.read().splitlines() if line.strip())


def small_export(tmp_path, shape=(6, 6, 6)):
    rng = np.random.default_rng(2)
    n_cells = shape[0] * shape[1] * shape[2]
    var_voxel = np.sort(rng.choice(n_cells, 100, replace=False))
    br = (rng.random((12, 100)) < 0.3).astype(np.uint8)
    pv = (br.sum(0) / np.float32(12)).astype(np.float32)
    truth = (rng.random(100) < 0.1).astype(np.uint8)
    fm.write_ising_out(str(tmp_path), shape, 2.0, (-128.0, -128.0, -10.0), var_voxel, pv, br, 2.0, np.log(999.0),
                       p_void_noprior=pv[::-1].copy(), truth=truth)
    return var_voxel, br, pv, truth


def test_ising_out_layout(tmp_path):
    var_voxel, br, pv, truth = small_export(tmp_path)
    raw = np.fromfile(tmp_path / "ising_out_pvoid.bin", np.float32)
    assert raw.size == 216 and os.path.getsize(tmp_path / "ising_out_pvoid.bin") == 4 * 216
    inside = np.zeros(216, bool)
    inside[var_voxel] = True
    assert np.array_equal(raw[var_voxel], pv) and np.all(raw[~inside] == -1.0)
    assert np.array_equal(np.fromfile(tmp_path / "ising_out_branches.bin", np.uint8), br.reshape(-1))
    t = np.fromfile(tmp_path / "ising_out_truth.bin", np.uint8)
    assert np.array_equal(t[var_voxel], truth) and not t[~inside].any()
    # the meta as a C++ stream prints it, read as the reference's tools read it
    text = (tmp_path / "ising_out.meta").read_text()
    assert text == "nx 6\nvoxel 2\nlo -128 -128 -10\nbranches 12\nvars 100\nlambda 2\nkappa 6.90675\n"
    meta = reference_load_meta(tmp_path / "ising_out.meta")
    assert int(meta["nx"]) == 6 and meta["lo"] == "-128 -128 -10"
    back = fm.read_ising_out(str(tmp_path))
    assert back["p_void"].shape == (6, 6, 6) and np.array_equal(back["p_void"].reshape(-1), raw)
    assert np.array_equal(back["branches"], br)
    # x fastest: cell (i, j, k) at k * ny * nx + j * nx + i, as (z, y, x) in C order
    c = int(var_voxel[7])
    assert back["p_void"][c // 36, (c // 6) % 6, c % 6] == pv[7]


def test_non_cubic_grids_name_every_axis(tmp_path):
    fm.write_ising_out(str(tmp_path), (4, 3, 2), 5.0, (0.0, 0.0, -50.0), [0, 5, 23], [0.5, 1.0, 0.0],
                       np.zeros((2, 3), np.uint8), 2.0, 1.5)
    back = fm.read_ising_out(str(tmp_path))
    assert back["shape"] == (4, 3, 2) and back["p_void"].shape == (2, 3, 4)
    assert back["p_void"][1, 2, 3] == 0.0 and back["p_void"][0, 1, 1] == 1.0


def test_machine_map(tmp_path):
    shape = (3, 2, 2)
    below_hi = np.nextafter(np.float32(0.9), np.float32(0.0))
    above_lo = np.nextafter(np.float32(0.1), np.float32(1.0))
    p = np.array([0.0, 0.1, above_lo, 0.5, 0.9, below_hi, 1.0, 0.95], np.float32)
    fm.write_ising_out(str(tmp_path), shape, 2.0, (0.0, 0.0, 0.0), np.arange(8), p, np.zeros((1, 8), np.uint8),
                       2.0, 6.9)
    tri, trav = fm.write_machine_map(str(tmp_path))
    flat = tri.reshape(-1)
    assert list(flat[:8]) == [0, 0, 2, 2, 1, 2, 1, 1] and list(flat[8:]) == [255] * 4
    assert np.array_equal(trav.reshape(-1), (flat == 0).astype(np.uint8))
    assert np.array_equal(np.fromfile(tmp_path / "ising_out_tristate.bin", np.uint8), flat)
    meta = (tmp_path / "ising_out_machine.meta").read_text().splitlines()
    assert meta[0] == "nx 3" and meta[3] == "order zyx_C" and meta[7] == "classify_lo 0.1"
    assert meta[9] == "policy Exists=traversable; Undecided|NotThere|outside=block"


def test_roi_text_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n = 7
    h = rng.normal(0, 3, n)
    J = np.triu(rng.normal(0, 1, (n, n)), 1)
    p = rng.random(n).astype(np.float32)
    fm.write_roi(str(tmp_path / "roi.txt"), np.arange(100, 107), p, [0, 1, 0, 0, 1, 0, 0], h, J)
    lines = (tmp_path / "roi.txt").read_text().splitlines()
    assert lines[0] == "7" and lines[1].split()[0] == "100" and lines[8].endswith(" ")
    assert lines[1].split()[1] == "%.17g" % float(p[0])
    vox, pv, truth, h2, J2 = fm.read_roi(str(tmp_path / "roi.txt"))
    assert np.array_equal(h2, h) and np.array_equal(J2, J) and np.array_equal(pv, p.astype(np.float64))
    assert list(truth) == [0, 1, 0, 0, 1, 0, 0] and list(vox) == list(range(100, 107))


def test_cloud_u8_rounding_and_json(tmp_path):
    p = np.array([0.0, 1.0, 0.5, 0.5 / 254, 1.5 / 254, 0.49 / 254, 1.3, -0.2], np.float64)
    g = fm.cloud_u8(10, np.arange(8), p)
    # std::lround(254 * clamp(p, 0, 1)): halves away from zero
    assert list(g) == [0, 254, 127, 1, 2, 0, 254, 0, 255, 255]
    fm.write_cloud(str(tmp_path / "ledger" / "cloud.u8"), (2, 5, 1), 2.0, (0.0, 1.0, 2.0), np.arange(8), p, [3, 4])
    assert np.array_equal(np.fromfile(tmp_path / "ledger" / "cloud.u8", np.uint8), g)
    meta = json.loads((tmp_path / "ledger" / "cloud.json").read_text())
    assert meta["encoding"] == "u8: 255 outside the domain, round(254 p_void) inside"
    assert meta["classify"] == {"exists_max": 0.1, "not_there_min": 0.9} and meta["truth_voxels"] == [3, 4]


def test_cloud_json_bytes_for_the_reference_scene(tmp_path):
    """The reference scene's companion (kN 128, voxel 2.0f, lo (-128, -128, -10); check_ising.cu:
    597-604) byte for byte: ostream prints the floats 2 and -128 without a decimal point."""
    fm.write_cloud(str(tmp_path / "cloud_pyramid.u8"), (128, 128, 128), 2.0, (-128.0, -128.0, -10.0), np.arange(3),
                   np.zeros(3), [7, 4099])
    want = ('{"nx":128,"voxel_m":2,"lo_m":[-128,-128,-10],"encoding":"u8: 255 outside the domain, '
            'round(254 p_void) inside","classify":{"exists_max":0.1,"not_there_min":0.9},"truth_voxels":[7,4099]}\n')
    assert (tmp_path / "cloud_pyramid.json").read_text() == want
