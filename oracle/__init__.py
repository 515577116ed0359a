"""ctypes binding of the CPU oracle (oracle/libqb_oracle.so).

TEST INFRASTRUCTURE ONLY. Importable from tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; the product never imports it.

The optional reference shim (oracle/_ref/libqb_ref.so, built from the reference's own
src/ising_weighted.hpp and src/engine.hpp) is exposed as `ref()`; it is None where the
reference tree was absent at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

RNG_REFERENCE = 0
RNG_PHILOX = 1
FP64 = 0
FP32 = 1

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")


class Problem(C.Structure):
    _fields_ = [("n_vars", C.c_int), ("n_rays", C.c_int),
                ("row_ptr", C.c_void_p), ("ray", C.c_void_p), ("a", C.c_void_p),
                ("d", C.c_void_p), ("w", C.c_void_p),
                ("nbr_ptr", C.c_void_p), ("nbr", C.c_void_p),
                ("lam", C.c_double), ("kappa", C.c_double), ("field", C.c_void_p)]


class Schedule(C.Structure):
    _fields_ = [("t_hot", C.c_double), ("t_cold", C.c_double),
                ("anneal_sweeps", C.c_int), ("hold_sweeps", C.c_int)]


class Mode(C.Structure):
    _fields_ = [("rng", C.c_int), ("order", C.c_void_p), ("class_ptr", C.c_void_p),
                ("n_classes", C.c_int), ("precision", C.c_int), ("recompute_every", C.c_int),
                ("segment", C.c_int)]


class Medium(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("voxel", C.c_float), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("mu", C.c_void_p)]


def medium(lo, voxel, shape, mu):
    """A voxel medium over mu (float32, x fastest; kept alive by the returned pair)."""
    mu = np.ascontiguousarray(mu, np.float32) if mu is not None else None
    m = Medium()
    m.lo[:] = [float(v) for v in lo]
    m.voxel = float(voxel)
    m.nx, m.ny, m.nz = (int(v) for v in shape)
    m.mu = mu.ctypes.data if mu is not None else None
    return m, mu


def march(med, p, d):
    """(tau, cells, segs) of src/muon.cuh's float march from p along d."""
    m, _ = med
    cap = 3 * (m.nx + m.ny + m.nz) + 16
    cells = np.zeros(cap, np.int32)
    segs = np.zeros(cap, np.float32)
    n = C.c_int()
    tau = lib().orc_march(C.byref(m), np.asarray(p, np.float32), np.asarray(d, np.float32), cells, segs, cap,
                          C.byref(n))
    return tau, cells[:n.value].copy(), segs[:n.value].copy()


def muon_expose(med, chamber, c_min, th_max, n_az, n_th, n_cand, seed):
    m, _ = med
    o = np.zeros(n_az * n_th, np.uint64)
    d = np.zeros(n_az * n_th, np.uint64)
    lib().orc_muon_expose(C.byref(m), np.asarray(chamber, np.float32), c_min, th_max, n_az, n_th, n_cand, seed, o, d)
    return o, d


def muon_rows(med, chamber, dirs, open_, det, site_bit, min_open, a_scale, row0=0):
    """(d, w, ent_bit, ent_row, ent_a) of one view's rows (src/op_muon.hpp:22-50)."""
    m, _ = med
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1)
    n_bins = dirs.size // 3
    cap_r, cap_e = n_bins, n_bins * (3 * (m.nx + m.ny + m.nz) + 16)
    d, w = np.zeros(cap_r), np.zeros(cap_r)
    eb, er, ea = np.zeros(cap_e, np.int32), np.zeros(cap_e, np.int32), np.zeros(cap_e, np.float32)
    nr, ne = C.c_int(), C.c_int()
    lib().orc_muon_rows(C.byref(m), np.asarray(chamber, np.float32), dirs, n_bins, np.ascontiguousarray(open_, np.uint64),
                        np.ascontiguousarray(det, np.uint64), np.ascontiguousarray(site_bit, np.int32), float(min_open),
                        float(a_scale), int(row0), d, w, cap_r, eb, er, ea, cap_e, C.byref(nr), C.byref(ne))
    return d[:nr.value], w[:nr.value], eb[:ne.value], er[:ne.value], ea[:ne.value]


def assemble_stable(n_bits, ent_bit, ent_row, ent_a):
    eb = np.ascontiguousarray(ent_bit, np.int32)
    row_ptr = np.zeros(n_bits + 1, np.int32)
    ray = np.zeros(len(eb), np.int32)
    a = np.zeros(len(eb), np.float32)
    lib().orc_assemble_stable(n_bits, len(eb), eb, np.ascontiguousarray(ent_row, np.int32),
                              np.ascontiguousarray(ent_a, np.float32), row_ptr, ray, a)
    return row_ptr, ray, a


class Graph(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int64), ("eu", C.c_void_p), ("ev", C.c_void_p),
                ("ew", C.c_void_p), ("row", C.c_void_p), ("col", C.c_void_p), ("w", C.c_void_p)]


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference shim when the reference tree exists)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference/src"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libqb_oracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        L.orc_splitmix.restype = C.c_uint64
        L.orc_splitmix.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_philox_uniform.restype = C.c_double
        L.orc_philox_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_qexp.restype = C.c_double
        L.orc_qexp.argtypes = [C.c_double]
        L.orc_temperature.restype = C.c_double
        L.orc_temperature.argtypes = [C.POINTER(Schedule), C.c_int]
        L.orc_binary_energy.restype = C.c_double
        L.orc_binary_energy.argtypes = [C.POINTER(Problem), _u8p]
        L.orc_energy_parts.argtypes = [C.POINTER(Problem), _u8p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_local_field_dE.restype = C.c_double
        L.orc_local_field_dE.argtypes = [C.POINTER(Problem), _u8p, _f64p, C.c_int]
        L.orc_anneal_branch.argtypes = [C.POINTER(Problem), C.POINTER(Schedule), C.c_uint64, C.POINTER(Mode),
                                        _u8p, C.c_void_p, C.c_void_p]
        L.orc_branch_fractions.argtypes = [C.POINTER(Problem), C.POINTER(Schedule), C.c_int, C.c_int,
                                           C.c_uint64, C.POINTER(Mode), C.c_void_p, C.c_void_p]
        L.orc_classify.restype = C.c_int
        L.orc_classify.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_freeze_mask.restype = C.c_int
        L.orc_freeze_mask.argtypes = [C.POINTER(Problem), C.c_double, _u8p]
        L.orc_freeze_bounds.argtypes = [C.POINTER(Problem), _u8p, _f64p]
        L.orc_roi_qubo.argtypes = [C.POINTER(Problem), _i32p, C.c_int, _u8p, _f64p, _f64p]
        L.orc_graph_csr.argtypes = [C.c_int, C.c_int64, _i32p, _i32p, _i32p, _i64p, _i32p, _i32p]
        L.orc_parse_rudy.restype = C.c_int
        L.orc_parse_rudy.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_size_t]
        L.orc_cut_of.restype = C.c_int64
        L.orc_cut_of.argtypes = [C.POINTER(Graph), _u8p]
        L.orc_energy_scratch.restype = C.c_int64
        L.orc_energy_scratch.argtypes = [C.POINTER(Graph), _u8p]
        L.orc_anneal_int.argtypes = [C.POINTER(Graph), C.POINTER(Schedule), C.c_uint64, C.POINTER(Mode),
                                     _u8p, C.POINTER(C.c_int64)]
        L.orc_tts99.restype = C.c_double
        L.orc_tts99.argtypes = [C.c_double, C.c_double]
        L.orc_tts99_bootstrap.argtypes = [_i32p, _f64p, C.c_size_t, C.c_int, C.c_uint64,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_sha256_hex.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        L.orc_qlog.restype = C.c_double
        L.orc_qlog.argtypes = [C.c_double]
        L.orc_dense_energy.restype = C.c_double
        L.orc_dense_energy.argtypes = [C.c_int, _f32p, _f32p, _u8p]
        L.orc_dense_anneal.argtypes = [C.c_int, _f32p, _f32p, C.POINTER(Schedule), C.c_uint64, C.c_int, C.c_int,
                                       _u8p, C.c_void_p]
        L.orc_qubo_energy.restype = C.c_double
        L.orc_qubo_energy.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _u8p]
        L.orc_qubo_anneal.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _i32p, C.POINTER(Schedule),
                                      C.c_uint64, C.c_int, _u8p]
        L.orc_hash_u32.restype = C.c_uint32
        L.orc_hash_u32.argtypes = [C.c_uint32]
        L.orc_hash_unit.restype = C.c_float
        L.orc_hash_unit.argtypes = [C.c_uint32]
        L.orc_march.restype = C.c_float
        L.orc_march.argtypes = [C.POINTER(Medium), _f32p, _f32p, _i32p, _f32p, C.c_int, C.POINTER(C.c_int)]
        L.orc_muon_expose.argtypes = [C.POINTER(Medium), _f32p, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int64,
                                      C.c_uint32, _u64p, _u64p]
        L.orc_muon_rows.argtypes = [C.POINTER(Medium), _f32p, _f32p, C.c_int, _u64p, _u64p, _i32p, C.c_double,
                                    C.c_float, C.c_int, _f64p, _f64p, C.c_int, _i32p, _i32p, _f32p, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_assemble_stable.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f32p, _i32p, _i32p, _f32p]
        L.orc_roi_enumerate.restype = C.c_double
        L.orc_roi_enumerate.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                        C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        _LIB = L
    return _LIB


def ref():
    """The reference shim, or None when it was not built (no reference tree)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libqb_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_anneal_int.argtypes = [C.c_int, C.c_int64, _i32p, _i32p, _i32p, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_uint64, _u8p, C.POINTER(C.c_int64)]
        R.ref_parse_rudy.restype = C.c_int
        R.ref_parse_rudy.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]
        R.ref_cut_of.restype = C.c_int64
        R.ref_cut_of.argtypes = [C.c_int, C.c_int64, _i32p, _i32p, _i32p, _u8p]
        R.ref_tts99_bootstrap.argtypes = [_i32p, _f64p, C.c_size_t, C.c_int, C.c_uint64,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_sha256_hex.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        R.ref_assemble.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int, _f64p, _f64p, C.c_int,
                                   _i32p, _i32p, _f32p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]
        R.ref_adjoint_mismatch.restype = C.c_double
        R.ref_adjoint_mismatch.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f32p, C.c_uint64]
        _REF = R
    return _REF


def _ptr(a):
    return None if a is None else a.ctypes.data


class FloatProblem:
    """Holds contiguous arrays of a BinaryProblem and its C view."""

    def __init__(self, row_ptr, ray, a, d, w, nbr_ptr, nbr, lam, kappa, field=None):
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.ray = np.ascontiguousarray(ray, np.int32)
        self.a = np.ascontiguousarray(a, np.float32)
        self.d = np.ascontiguousarray(d, np.float64)
        self.w = np.ascontiguousarray(w, np.float64)
        self.nbr_ptr = np.ascontiguousarray(nbr_ptr, np.int32)
        self.nbr = np.ascontiguousarray(nbr, np.int32)
        self.lam = float(lam)
        self.kappa = float(kappa)
        self.field = None if field is None else np.ascontiguousarray(field, np.float64)
        self.n_vars = len(self.row_ptr) - 1
        self.n_rays = len(self.d)
        self.c = Problem(self.n_vars, self.n_rays, _ptr(self.row_ptr), _ptr(self.ray), _ptr(self.a),
                         _ptr(self.d), _ptr(self.w), _ptr(self.nbr_ptr), _ptr(self.nbr), self.lam,
                         self.kappa, _ptr(self.field))


class IntGraph:
    """Edge list plus the symmetric CSR parse_rudy builds."""

    def __init__(self, n, eu, ev, ew):
        self.n = int(n)
        self.eu = np.ascontiguousarray(eu, np.int32)
        self.ev = np.ascontiguousarray(ev, np.int32)
        self.ew = np.ascontiguousarray(ew, np.int32)
        self.m = len(self.eu)
        self.row = np.zeros(self.n + 1, np.int64)
        self.col = np.zeros(2 * self.m, np.int32)
        self.w = np.zeros(2 * self.m, np.int32)
        lib().orc_graph_csr(self.n, self.m, self.eu, self.ev, self.ew, self.row, self.col, self.w)
        self.c = Graph(self.n, self.m, _ptr(self.eu), _ptr(self.ev), _ptr(self.ew), _ptr(self.row),
                       _ptr(self.col), _ptr(self.w))


ENGINE_SEGMENT = 16   # the MI355X engine sums dE in chains of 16 entries


def make_mode(rng=RNG_REFERENCE, order=None, class_ptr=None, precision=FP64, recompute_every=0, segment=None):
    """segment=None: 16 (the engine's order) for the Philox trajectory, 0 for the reference's."""
    if segment is None:
        segment = ENGINE_SEGMENT if rng == RNG_PHILOX else 0
    keep = []
    if order is not None:
        order = np.ascontiguousarray(order, np.int32)
        class_ptr = np.ascontiguousarray(class_ptr, np.int32)
        keep = [order, class_ptr]
    m = Mode(rng, _ptr(order), _ptr(class_ptr), 0 if class_ptr is None else len(class_ptr) - 1,
             precision, recompute_every, segment)
    m._keep = keep
    return m


def schedule(t_hot, t_cold, anneal, hold):
    return Schedule(float(t_hot), float(t_cold), int(anneal), int(hold))


def anneal_branch(prob: FloatProblem, sched, seed, mode=None, trace=False, residuals=False):
    mode = mode or make_mode()
    x = np.zeros(prob.n_vars, np.uint8)
    tr = np.zeros(sched.anneal_sweeps + sched.hold_sweeps, np.float64) if trace else None
    rf = np.zeros(prob.n_rays, np.float64) if residuals else None
    lib().orc_anneal_branch(C.byref(prob.c), C.byref(sched), C.c_uint64(seed), C.byref(mode), x,
                            _ptr(tr), _ptr(rf))
    out = [x]
    if trace:
        out.append(tr)
    if residuals:
        out.append(rf)
    return out[0] if len(out) == 1 else tuple(out)


def branch_fractions(prob: FloatProblem, sched, n_branches, n_threads=8, seed_offset=0, mode=None):
    """Returns (p_v float32[n], branches uint8[R, n])."""
    mode = mode or make_mode()
    keep = np.zeros((n_branches, prob.n_vars), np.uint8)
    p = np.zeros(prob.n_vars, np.float32)
    lib().orc_branch_fractions(C.byref(prob.c), C.byref(sched), n_branches, n_threads,
                               C.c_uint64(seed_offset), C.byref(mode), _ptr(keep), _ptr(p))
    return p, keep


def binary_energy(prob: FloatProblem, x):
    return lib().orc_binary_energy(C.byref(prob.c), np.ascontiguousarray(x, np.uint8))


def energy_parts(prob: FloatProblem, x):
    d, q = C.c_double(), C.c_double()
    lib().orc_energy_parts(C.byref(prob.c), np.ascontiguousarray(x, np.uint8), C.byref(d), C.byref(q))
    return d.value, q.value


def anneal_int(g: IntGraph, sched, seed, mode=None):
    mode = mode or make_mode()
    x = np.zeros(g.n, np.uint8)
    e = C.c_int64()
    lib().orc_anneal_int(C.byref(g.c), C.byref(sched), C.c_uint64(seed), C.byref(mode), x, C.byref(e))
    return x, e.value


def _qubo_arrays(row_ptr, col, J, h):
    return (np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32),
            np.ascontiguousarray(J, np.float32), np.ascontiguousarray(h, np.float32))


def qubo_anneal(row_ptr, col, J, h, order, sched, seed, R):
    """The batch engine's trajectory for one QUBO instance: R replicas, uint8[R, n]."""
    rp, cl, Jv, hv = _qubo_arrays(row_ptr, col, J, h)
    n = len(hv)
    x = np.zeros((R, n), np.uint8)
    lib().orc_qubo_anneal(n, rp, cl, Jv, hv, np.ascontiguousarray(order, np.int32), C.byref(sched),
                          C.c_uint64(seed), R, x)
    return x


def qubo_energy(row_ptr, col, J, h, x):
    rp, cl, Jv, hv = _qubo_arrays(row_ptr, col, J, h)
    return lib().orc_qubo_energy(len(hv), rp, cl, Jv, hv, np.ascontiguousarray(x, np.uint8))


def dense_anneal(Q, h, sched, seed, precision=FP64, recompute_every=0, fields=False):
    """One replica of the dense field engine's trajectory: x uint8[n] (and the carried fields)."""
    Qv = np.ascontiguousarray(Q, np.float32)
    hv = np.ascontiguousarray(h, np.float32)
    n = len(hv)
    x = np.zeros(n, np.uint8)
    f = np.zeros(n, np.float64) if fields else None
    lib().orc_dense_anneal(n, Qv.reshape(-1), hv, C.byref(sched), C.c_uint64(seed), int(precision),
                           int(recompute_every), x, _ptr(f))
    return (x, f) if fields else x


def dense_energy(Q, h, x):
    Qv = np.ascontiguousarray(Q, np.float32)
    hv = np.ascontiguousarray(h, np.float32)
    return lib().orc_dense_energy(len(hv), Qv.reshape(-1), hv, np.ascontiguousarray(x, np.uint8))


def roi_enumerate(h, J, beta=1.0):
    """Exact posterior of a small QUBO by brute force: dict(e_min, log_z, map_index,
    marginal)."""
    h = np.ascontiguousarray(h, np.float64)
    J = np.ascontiguousarray(J, np.float64)
    n = len(h)
    m = np.zeros(n, np.float64)
    lz = C.c_double()
    idx = C.c_uint64()
    e = lib().orc_roi_enumerate(n, h.ctypes.data, J.ctypes.data, float(beta), m.ctypes.data, C.byref(lz),
                                C.byref(idx))
    return dict(e_min=e, log_z=lz.value, map_index=idx.value, marginal=m)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def sha256_hex(data: bytes) -> str:
    out = C.create_string_buffer(65)
    lib().orc_sha256_hex(data, len(data), out)
    return out.value.decode()
