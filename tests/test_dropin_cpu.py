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
