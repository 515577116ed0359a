"""Python binding of libqublar_hip.so, the MI355X simulated-annealing QUBO engine.

The engine is the C ABI of include/qublar.h; this module is a thin ctypes layer for
tests, benchmarks and scripting. Every call runs the HIP library: there is no CPU
fallback, and importing fails loudly when the library has not been built.

Mirrors the reference's entry points:
  Problem.anneal_branches   branch_fractions / anneal_branch (src/ising_recon.hpp:121-166)
  Problem.energy            binary_energy (src/ising_recon.hpp:65-79)
  Graph.run_batch           gising::run_batch (src/ising_gpu.cuh:396-458)
  Graph.device_energies     gising::device_energies (src/ising_gpu.cuh:461-475)
  Graph.plan                gising::make_plan (src/ising_gpu.cuh:74-101)
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqublar_hip.so")

QB_OK, QB_E_ARG, QB_E_HIP, QB_E_REFUSED, QB_E_NOMEM, QB_E_LIMIT, QB_E_TIMEOUT = range(7)
FP64, FP32 = 0, 1


class QublarError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qublar error {code}: {msg}")
        self.code = code


def build(jobs: int = 8) -> str:
    """Compile libqublar_hip.so for gfx950 in-tree (always recompiles what changed)."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"], check=True)
    return LIB_PATH


class ProblemDesc(C.Structure):
    _fields_ = [("n_vars", C.c_int), ("n_rays", C.c_int),
                ("row_ptr", C.c_void_p), ("ray", C.c_void_p), ("a", C.c_void_p),
                ("d", C.c_void_p), ("w", C.c_void_p),
                ("nbr_ptr", C.c_void_p), ("nbr", C.c_void_p),
                ("lam", C.c_double), ("kappa", C.c_double), ("field", C.c_void_p),
                ("draw_bit", C.c_void_p), ("draw_key", C.c_void_p), ("colour", C.c_void_p)]


class ProblemInfo(C.Structure):
    _fields_ = [("n_vars", C.c_int), ("n_rays", C.c_int), ("nnz", C.c_int64), ("n_nbr", C.c_int64),
                ("n_classes", C.c_int), ("max_class", C.c_int), ("min_class", C.c_int),
                ("max_entries_per_bit", C.c_int), ("max_bits_per_ray", C.c_int), ("device_bytes", C.c_int64)]


class Schedule(C.Structure):
    _fields_ = [("t_hot", C.c_double), ("t_cold", C.c_double),
                ("anneal_sweeps", C.c_int), ("hold_sweeps", C.c_int)]


class AnnealOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("recompute_every", C.c_int), ("workgroups", C.c_int),
                ("sweeps_per_launch", C.c_int), ("group_width", C.c_int)]


class Report(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("kernel_ms", C.c_double), ("launches", C.c_int),
                ("spin_updates", C.c_double), ("fp32_residual_err", C.c_double)]


class GraphDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int64), ("eu", C.c_void_p), ("ev", C.c_void_p), ("ew", C.c_void_p)]


class Plan(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int64), ("n_pad", C.c_int), ("B", C.c_int64), ("max_w", C.c_int64),
                ("field_bytes", C.c_int), ("weight_bytes", C.c_int), ("dense", C.c_int),
                ("shared_state", C.c_int), ("sigma", C.c_double), ("w_min", C.c_int64),
                ("refused", C.c_char * 64), ("n_classes", C.c_int)]


class QuboBatchDesc(C.Structure):
    _fields_ = [("n_instances", C.c_int), ("n_bits", C.c_void_p), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("J", C.c_void_p), ("h", C.c_void_p), ("replicas", C.c_void_p), ("seed", C.c_void_p)]


class QuboOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("replica_block", C.c_int), ("workgroups", C.c_int), ("block", C.c_int),
                ("graph", C.c_int)]


class QuboReport(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("kernel_ms", C.c_double), ("launches", C.c_int),
                ("replica_block", C.c_int), ("jobs", C.c_int), ("block", C.c_int), ("spin_updates", C.c_double),
                ("graph_replayed", C.c_int)]


class DenseReport(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("kernel_ms", C.c_double), ("spin_updates", C.c_double)]


class BatchReport(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("kernel_ms", C.c_double), ("launches", C.c_int),
                ("sweeps_per_launch", C.c_int)]


class RoiResult(C.Structure):
    _fields_ = [("e_min", C.c_double), ("log_z", C.c_double), ("p_map", C.c_double), ("map_index", C.c_uint64),
                ("states", C.c_double), ("kernel_ms", C.c_double)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int),
                ("xcds", C.c_int), ("hbm_bytes", C.c_int64), ("lds_bytes_per_cu", C.c_int)]


class MuonScene(C.Structure):
    _fields_ = [("log2_muons", C.c_int), ("rho", C.c_double), ("mu_per_rho", C.c_double),
                ("void_side", C.c_double), ("void_z", C.c_double), ("lam", C.c_double), ("kappa", C.c_double),
                ("min_open", C.c_double), ("seed", C.c_uint64), ("grid", C.c_int), ("sky_az", C.c_int),
                ("sky_th", C.c_int)]


class GravityScene(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h", C.c_double), ("radius", C.c_double),
                ("drho", C.c_double), ("sigma", C.c_double), ("body_lo", C.c_double * 3),
                ("body_hi", C.c_double * 3), ("lam", C.c_double), ("kappa", C.c_double), ("seed", C.c_uint64)]


class VoxelMedium(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("voxel", C.c_float), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("mu", C.c_void_p)]


class MuonSky(C.Structure):
    _fields_ = [("c_min", C.c_float), ("th_max", C.c_float), ("n_az", C.c_int), ("n_th", C.c_int)]


_LIB = None
# Every live device handle, closed at interpreter exit before the HIP runtime's own teardown
# (a handle's __del__ could otherwise call into HIP after the runtime has gone).
_HANDLES = weakref.WeakSet()


def _close_handles():
    for h in sorted(list(_HANDLES), key=lambda o: 0 if isinstance(o, Stage) else 1):   # stages before problems
        try:
            h.close()
        except Exception:
            pass


atexit.register(_close_handles)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise QublarError(QB_E_ARG, f"{LIB_PATH} is not built; run qublar_amd.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
        sig = {
            "qb_last_error": (C.c_char_p, []),
            "qb_version": (i32, []),
            "qb_get_device_info": (i32, [C.POINTER(DeviceInfo)]),
            "qb_copy_probe": (i32, [C.c_int64, i32, C.POINTER(C.c_double)]),
            "qb_team_barrier_check": (i32, [i32, i32, C.POINTER(u64), C.POINTER(dbl)]),
            "qb_problem_create": (i32, [C.POINTER(ProblemDesc), C.POINTER(vp)]),
            "qb_problem_destroy": (i32, [vp]),
            "qb_problem_get_info": (i32, [vp, C.POINTER(ProblemInfo)]),
            "qb_problem_colouring": (i32, [vp, vp, vp]),
            "qb_colour_problem": (i32, [C.POINTER(ProblemDesc), vp, vp, C.POINTER(i32)]),
            "qb_anneal_branches": (i32, [vp, C.POINTER(Schedule), i32, u64, C.POINTER(AnnealOpts), vp, vp, vp, vp,
                                         vp, C.POINTER(Report)]),
            "qb_energy": (i32, [vp, vp, i32, vp, vp, vp]),
            "qb_energy_noalloc": (i32, [vp, vp, i32, vp, vp, vp]),
            "qb_stage_create": (i32, [vp, C.POINTER(Schedule), i32, C.POINTER(AnnealOpts), C.POINTER(vp)]),
            "qb_stage_destroy": (i32, [vp]),
            "qb_stage_set_span": (i32, [vp, i32, i32, i32, i32]),
            "qb_stage_import": (i32, [vp, vp, i32, i32, i32, i32, i32, i32]),
            "qb_stage_run": (i32, [vp, i32, C.POINTER(Report)]),
            "qb_stage_outputs": (i32, [vp, i32, i32, i32, vp, vp, vp, vp]),
            "qb_stage_sweeps": (i32, [vp, C.POINTER(i32)]),
            "qb_freeze": (i32, [vp, dbl, vp, C.POINTER(i32)]),
            "qb_anneal_branches_reference": (i32, [vp, C.POINTER(Schedule), i32, u64, vp, vp, vp, vp, vp,
                                                   C.POINTER(Report)]),
            "qb_anneal_branches_reference_split": (i32, [vp, C.POINTER(Schedule), i32, u64, i32, vp, vp, vp, vp, vp,
                                                         C.POINTER(Report)]),
            "qb_graph_create": (i32, [C.POINTER(GraphDesc), i32, C.POINTER(vp)]),
            "qb_graph_destroy": (i32, [vp]),
            "qb_graph_get_plan": (i32, [vp, C.POINTER(Plan)]),
            "qb_graph_colouring": (i32, [vp, vp, vp]),
            "qb_colour_graph": (i32, [C.POINTER(GraphDesc), vp, vp, C.POINTER(i32)]),
            "qb_run_batch": (i32, [vp, C.POINTER(Schedule), u64, i32, dbl, vp, vp, C.POINTER(BatchReport)]),
            "qb_device_energies": (i32, [vp, vp, i32, vp]),
            "qb_run_batch_reference": (i32, [vp, C.POINTER(Schedule), u64, i32, dbl, vp, vp, C.POINTER(BatchReport)]),
            "qb_muon_scene_defaults": (i32, [C.POINTER(MuonScene)]),
            "qb_scene_muon": (i32, [C.POINTER(MuonScene), C.POINTER(vp)]),
            "qb_host_problem_view": (i32, [vp, C.POINTER(ProblemDesc)]),
            "qb_host_problem_truth": (i32, [vp, vp]),
            "qb_host_problem_destroy": (i32, [vp]),
            "qb_reduce_frozen": (i32, [C.POINTER(ProblemDesc), vp, C.POINTER(vp), vp, C.POINTER(i32)]),
            "qb_gravity_scene_defaults": (i32, [C.POINTER(GravityScene)]),
            "qb_scene_gravity": (i32, [C.POINTER(GravityScene), C.POINTER(vp)]),
            "qb_host_problem_colour": (i32, [vp, vp]),
            "qb_scene_er_graph": (i32, [i32, dbl, u64, vp, vp, vp, i64, C.POINTER(i64)]),
            "qb_muon_expose": (i32, [C.POINTER(VoxelMedium), vp, C.POINTER(MuonSky), i64, C.c_uint32, vp, vp]),
            "qb_muon_rows_device": (i32, [C.POINTER(VoxelMedium), i32, vp, vp, i32, vp, vp, vp, dbl, C.c_float, dbl,
                                          dbl, C.POINTER(vp)]),
            "qb_scene_muon_device": (i32, [C.POINTER(MuonScene), C.POINTER(vp), C.POINTER(dbl)]),
            "qb_scene_gravity_device": (i32, [C.POINTER(GravityScene), C.POINTER(vp), C.POINTER(dbl)]),
            "qb_qubo_batch_create": (i32, [C.POINTER(QuboBatchDesc), C.POINTER(vp)]),
            "qb_qubo_batch_destroy": (i32, [vp]),
            "qb_qubo_batch_anneal": (i32, [vp, C.POINTER(Schedule), C.POINTER(QuboOpts), vp, vp,
                                           C.POINTER(QuboReport)]),
            "qb_qubo_batch_colouring": (i32, [vp, i32, vp, vp, C.POINTER(i32)]),
            "qb_colour_qubo": (i32, [i32, vp, vp, vp, vp, vp, C.POINTER(i32)]),
            "qb_dense_create": (i32, [i32, vp, vp, C.POINTER(vp)]),
            "qb_dense_destroy": (i32, [vp]),
            "qb_roi_enumerate": (i32, [i32, vp, vp, dbl, vp, vp, C.POINTER(RoiResult)]),
            "qb_dense_anneal": (i32, [vp, C.POINTER(Schedule), i32, u64, i32, i32, vp, vp, vp, C.POINTER(DenseReport)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != QB_OK:
        raise QublarError(rc, lib().qb_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else a.ctypes.data


def device_info():
    d = DeviceInfo()
    _check(lib().qb_get_device_info(C.byref(d)))
    return dict(name=d.name.decode(), arch=d.arch.decode(), compute_units=d.compute_units, xcds=d.xcds,
                hbm_bytes=d.hbm_bytes, lds_bytes_per_cu=d.lds_bytes_per_cu)


def union(problems, keys, colours=None):
    """The block-diagonal union of independent problems (one lambda for all): bits,
    rays and neighbours of problem q follow those of problems 0 .. q-1, each bit keeps
    its local index as its draw counter and problem q's key as its draw key, and every
    bit's linear term is its problem's field (or kappa). Annealed with seed_offset 0,
    problem q's branch b in the union is, spin for spin, its branch b annealed alone
    with seed_offset keys[q]. The union is coloured by the problems' own colourings side
    by side (`colours`, or each problem's colours()), so its classes are theirs.
    Returns the union and each problem's (bit, ray) offsets."""
    lam = problems[0].lam
    if any(P.lam != lam for P in problems):
        raise ValueError("union: every problem must have the same lambda")
    row_ptr, ray, a, d, w, nbr_ptr, nbr, field, dbit, dkey, colour, offs = [np.zeros(1, np.int32)], [], [], [], \
        [], [np.zeros(1, np.int32)], [], [], [], [], [], []
    bit0 = ray0 = nnz0 = nn0 = 0
    if colours is None:
        colours = [P.colours() for P in problems]
    for P, key, col in zip(problems, keys, colours):
        offs.append((bit0, ray0))
        row_ptr.append(P.row_ptr[1:] + nnz0)
        ray.append(P.ray + ray0)
        a.append(P.a)
        d.append(P.d)
        w.append(P.w)
        nbr_ptr.append(P.nbr_ptr[1:] + nn0)
        nbr.append(P.nbr + bit0)
        field.append(P.field if P.field is not None else np.full(P.n_vars, P.kappa))
        dbit.append(P.draw_bit if P.draw_bit is not None else np.arange(P.n_vars, dtype=np.int32))
        dkey.append((P.draw_key if P.draw_key is not None else np.zeros(P.n_vars, np.uint64)) + np.uint64(key))
        colour.append(col)
        bit0 += P.n_vars
        ray0 += P.n_rays
        nnz0 += int(P.row_ptr[-1])
        nn0 += int(P.nbr_ptr[-1])
    cat = np.concatenate
    U = ProblemArrays(cat(row_ptr), cat(ray), cat(a), cat(d), cat(w), cat(nbr_ptr), cat(nbr), lam, 0.0, cat(field),
                      draw_bit=cat(dbit), draw_key=cat(dkey), colour=cat(colour))
    return U, offs


def team_barrier_check(steps=20000, workgroups=0):
    """The team barrier's visibility on this device: (stale reads, microseconds per step)."""
    stale = C.c_uint64()
    us = C.c_double()
    _check(lib().qb_team_barrier_check(int(steps), int(workgroups), C.byref(stale), C.byref(us)))
    return stale.value, us.value


def copy_probe(nbytes, reps=10):
    """Mean ms of a plain device copy of `nbytes` (reads and writes nbytes each)."""
    ms = C.c_double()
    _check(lib().qb_copy_probe(C.c_int64(int(nbytes)), int(reps), C.byref(ms)))
    return ms.value


def schedule(t_hot, t_cold, anneal, hold):
    return Schedule(float(t_hot), float(t_cold), int(anneal), int(hold))


class ProblemArrays:
    """A BinaryProblem held as contiguous numpy arrays (src/engine.hpp:30-49)."""

    def __init__(self, row_ptr, ray, a, d, w, nbr_ptr, nbr, lam, kappa, field=None, truth=None, draw_bit=None,
                 draw_key=None, colour=None):
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
        self.truth = truth
        self.draw_bit = None if draw_bit is None else np.ascontiguousarray(draw_bit, np.int32)
        self.draw_key = None if draw_key is None else np.ascontiguousarray(draw_key, np.uint64)
        self.colour = None if colour is None else np.ascontiguousarray(colour, np.int32)
        self.n_vars = len(self.row_ptr) - 1
        self.n_rays = len(self.d)

    def desc(self):
        return ProblemDesc(self.n_vars, self.n_rays, _p(self.row_ptr), _p(self.ray), _p(self.a), _p(self.d),
                           _p(self.w), _p(self.nbr_ptr), _p(self.nbr), self.lam, self.kappa, _p(self.field),
                           _p(self.draw_bit), _p(self.draw_key), _p(self.colour))

    def colours(self):
        """The engine's colour of every bit (the given colouring when there is one)."""
        if self.colour is not None:
            return self.colour
        order, cp = self.colouring()
        col = np.empty(self.n_vars, np.int32)
        for c in range(len(cp) - 1):
            col[order[cp[c]:cp[c + 1]]] = c
        return col

    def kwargs(self):
        return dict(row_ptr=self.row_ptr, ray=self.ray, a=self.a, d=self.d, w=self.w, nbr_ptr=self.nbr_ptr,
                    nbr=self.nbr, lam=self.lam, kappa=self.kappa, field=self.field)

    def reduce_frozen(self, frozen):
        """freeze_provable_rock's reduced problem (src/ising_recon.hpp:206-233) for a frozen
        set: (ProblemArrays of the unfrozen bits with the frozen bonds folded into their
        field, kept: reduced bit -> original bit)."""
        frozen = np.ascontiguousarray(frozen, np.uint8)
        kept = np.zeros(self.n_vars, np.int32)
        m = C.c_int()
        h = C.c_void_p()
        d = self.desc()
        _check(lib().qb_reduce_frozen(C.byref(d), _p(frozen), C.byref(h), _p(kept), C.byref(m)))
        R = _host_problem(h, False)
        R.truth = self.truth[kept[:m.value]] if self.truth is not None else None
        return R, kept[:m.value].copy()

    def colouring(self):
        """(order, class_ptr) of the engine's colour classes, computed on the host."""
        order = np.zeros(self.n_vars, np.int32)
        cp = np.zeros(self.n_vars + 1, np.int32)
        k = C.c_int()
        d = self.desc()
        _check(lib().qb_colour_problem(C.byref(d), _p(order), _p(cp), C.byref(k)))
        return order, cp[:k.value + 1].copy()


class Problem:
    """A float ray x voxel QUBO resident on the device."""

    def __init__(self, arrays: ProblemArrays):
        self.arrays = arrays
        self.h = C.c_void_p()
        d = arrays.desc()
        _check(lib().qb_problem_create(C.byref(d), C.byref(self.h)))
        _HANDLES.add(self)

    def close(self):
        if self.h:
            lib().qb_problem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = ProblemInfo()
        _check(lib().qb_problem_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in ProblemInfo._fields_}

    def colouring(self):
        n = self.arrays.n_vars
        info = self.info()
        order = np.zeros(n, np.int32)
        cp = np.zeros(info["n_classes"] + 1, np.int32)
        _check(lib().qb_problem_colouring(self.h, _p(order), _p(cp)))
        return order, cp

    def anneal_branches(self, sched, n_branches, seed_offset=0, precision=FP64, recompute_every=0, workgroups=0,
                        sweeps_per_launch=0, group_width=0, branches=True, outputs=True, energy=True):
        n = self.arrays.n_vars
        br = np.zeros((n_branches, n), np.uint8) if branches else None
        pv = np.zeros(n, np.float32) if outputs else None
        tri = np.zeros(n, np.uint8) if outputs else None
        cert = np.zeros(n, np.float32) if outputs else None
        en = np.zeros(n_branches, np.float64) if energy else None
        opts = AnnealOpts(precision, recompute_every, workgroups, sweeps_per_launch, group_width)
        rep = Report()
        _check(lib().qb_anneal_branches(self.h, C.byref(sched), n_branches, C.c_uint64(seed_offset), C.byref(opts),
                                        _p(br), _p(pv), _p(tri), _p(cert), _p(en), C.byref(rep)))
        return dict(branches=br, p_void=pv, tri=tri, certainty=cert, energy=en,
                    report={f: getattr(rep, f) for f, _ in Report._fields_})

    def freeze(self, margin):
        """freeze_provable_rock's frozen set on the device: (frozen uint8[n], rounds)."""
        fr = np.zeros(self.arrays.n_vars, np.uint8)
        rounds = C.c_int()
        _check(lib().qb_freeze(self.h, float(margin), _p(fr), C.byref(rounds)))
        return fr, rounds.value

    def anneal_branches_reference(self, sched, n_branches, seed_offset=0, branches=True, sweeps_per_launch=0):
        """branch_fractions in the reference's own trajectory (index order, splitmix64);
        sweeps_per_launch 0 lets the library split the launches."""
        n = self.arrays.n_vars
        br = np.zeros((n_branches, n), np.uint8) if branches else None
        pv = np.zeros(n, np.float32)
        tri = np.zeros(n, np.uint8)
        cert = np.zeros(n, np.float32)
        en = np.zeros(n_branches, np.float64)
        rep = Report()
        _check(lib().qb_anneal_branches_reference_split(self.h, C.byref(sched), n_branches, C.c_uint64(seed_offset),
                                                        int(sweeps_per_launch), _p(br), _p(pv), _p(tri), _p(cert),
                                                        _p(en), C.byref(rep)))
        return dict(branches=br, p_void=pv, tri=tri, certainty=cert, energy=en,
                    report={f: getattr(rep, f) for f, _ in Report._fields_})

    def energy(self, X):
        X = np.ascontiguousarray(X, np.uint8)
        K = X.shape[0]
        tot, dat, pri = np.zeros(K), np.zeros(K), np.zeros(K)
        _check(lib().qb_energy(self.h, _p(X), K, _p(tot), _p(dat), _p(pri)))
        return tot, dat, pri

    def energy_noalloc(self, X):
        """energy() through qb_energy_noalloc: the library's static device scratch, no
        device allocation in the call."""
        X = np.ascontiguousarray(X, np.uint8)
        K = X.shape[0]
        tot, dat, pri = np.zeros(K), np.zeros(K), np.zeros(K)
        _check(lib().qb_energy_noalloc(self.h, _p(X), K, _p(tot), _p(dat), _p(pri)))
        return tot, dat, pri


class Stage:
    """The branch state of a resident problem (usually a union of requests) kept on the
    device between calls (qb_stage_*): a few sweeps at a time, spans of bits with their own
    clocks, state imported from another stage, per-span outputs."""

    def __init__(self, problem, sched, n_branches, precision=FP64, recompute_every=0, workgroups=0, group_width=0):
        self.problem = problem          # keeps the problem alive while the stage exists
        self.h = C.c_void_p()
        self.n_branches = n_branches
        opts = AnnealOpts(precision, recompute_every, workgroups, 0, group_width)
        _check(lib().qb_stage_create(problem.h, C.byref(sched), n_branches, C.byref(opts), C.byref(self.h)))
        _HANDLES.add(self)

    def close(self):
        if self.h:
            lib().qb_stage_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_span(self, bit0, n_bits, sweeps_done=0, n_branches=0):
        _check(lib().qb_stage_set_span(self.h, int(bit0), int(n_bits), int(sweeps_done), int(n_branches)))

    def import_span(self, src, src_bit0, dst_bit0, n_bits, src_ray0, dst_ray0, n_rays):
        _check(lib().qb_stage_import(self.h, src.h, int(src_bit0), int(dst_bit0), int(n_bits), int(src_ray0),
                                     int(dst_ray0), int(n_rays)))

    def run(self, n_sweeps):
        rep = Report()
        _check(lib().qb_stage_run(self.h, int(n_sweeps), C.byref(rep)))
        return {f: getattr(rep, f) for f, _ in Report._fields_}

    @property
    def sweeps(self):
        v = C.c_int()
        _check(lib().qb_stage_sweeps(self.h, C.byref(v)))
        return v.value

    def outputs(self, bit0, n_bits, n_branches, branches=True):
        br = np.zeros((n_branches, n_bits), np.uint8) if branches else None
        pv = np.zeros(n_bits, np.float32)
        tri = np.zeros(n_bits, np.uint8)
        cert = np.zeros(n_bits, np.float32)
        _check(lib().qb_stage_outputs(self.h, int(bit0), int(n_bits), int(n_branches), _p(br), _p(pv), _p(tri),
                                      _p(cert)))
        return dict(branches=br, p_void=pv, tri=tri, certainty=cert)


class GraphArrays:
    def __init__(self, n, eu, ev, ew):
        self.n = int(n)
        self.eu = np.ascontiguousarray(eu, np.int32)
        self.ev = np.ascontiguousarray(ev, np.int32)
        self.ew = np.ascontiguousarray(ew, np.int32)
        self.m = len(self.eu)

    def desc(self):
        return GraphDesc(self.n, self.m, _p(self.eu), _p(self.ev), _p(self.ew))

    def colouring(self):
        order = np.zeros(self.n, np.int32)
        cp = np.zeros(self.n + 1, np.int32)
        k = C.c_int()
        d = self.desc()
        _check(lib().qb_colour_graph(C.byref(d), _p(order), _p(cp), C.byref(k)))
        return order, cp[:k.value + 1].copy()


class Graph:
    """A weighted integer graph resident on the device (DeviceGraph)."""

    def __init__(self, arrays: GraphArrays, force_global=False):
        self.arrays = arrays
        self.h = C.c_void_p()
        d = arrays.desc()
        _check(lib().qb_graph_create(C.byref(d), int(force_global), C.byref(self.h)))
        _HANDLES.add(self)

    def close(self):
        if self.h:
            lib().qb_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self):
        p = Plan()
        _check(lib().qb_graph_get_plan(self.h, C.byref(p)))
        out = {f: getattr(p, f) for f, _ in Plan._fields_}
        out["refused"] = p.refused.decode()
        return out

    def run_batch(self, sched, seed_base, R, updates_per_launch=1e9):
        n_pad = self.plan()["n_pad"]
        s = np.zeros((R, n_pad), np.int8)
        e = np.zeros(R, np.int64)
        rep = BatchReport()
        _check(lib().qb_run_batch(self.h, C.byref(sched), C.c_uint64(seed_base), R, float(updates_per_launch),
                                  _p(s), _p(e), C.byref(rep)))
        return s, e, {f: getattr(rep, f) for f, _ in BatchReport._fields_}

    def run_batch_reference(self, sched, seed_base, R, updates_per_launch=1e9):
        """run_batch in the reference's own trajectory: replica r is wising::anneal with seed
        seed_base + r (index order, splitmix64 on uphill proposals)."""
        n_pad = self.plan()["n_pad"]
        s = np.zeros((R, n_pad), np.int8)
        e = np.zeros(R, np.int64)
        rep = BatchReport()
        _check(lib().qb_run_batch_reference(self.h, C.byref(sched), C.c_uint64(seed_base), R, float(updates_per_launch),
                                            _p(s), _p(e), C.byref(rep)))
        return s, e, {f: getattr(rep, f) for f, _ in BatchReport._fields_}

    def device_energies(self, X):
        X = np.ascontiguousarray(X, np.uint8)
        out = np.zeros(X.shape[0], np.int64)
        _check(lib().qb_device_energies(self.h, _p(X), X.shape[0], _p(out)))
        return out


def _host_problem(h, coloured):
    """ProblemArrays (with .truth, and the builder's colouring when `coloured`) copied out
    of a host problem, which is destroyed."""
    try:
        v = ProblemDesc()
        _check(lib().qb_host_problem_view(h, C.byref(v)))
        n, nr = v.n_vars, v.n_rays

        def arr(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            ct = np.ctypeslib.as_ctypes_type(dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()

        row_ptr = arr(v.row_ptr, n + 1, np.int32)
        nbr_ptr = arr(v.nbr_ptr, n + 1, np.int32)
        truth = np.zeros(n, np.uint8)
        _check(lib().qb_host_problem_truth(h, _p(truth)))
        colour = None
        if coloured:
            colour = np.zeros(n, np.int32)
            _check(lib().qb_host_problem_colour(h, _p(colour)))
        field = arr(v.field, n, np.float64) if v.field else None
        return ProblemArrays(row_ptr, arr(v.ray, int(row_ptr[-1]), np.int32), arr(v.a, int(row_ptr[-1]), np.float32),
                             arr(v.d, nr, np.float64), arr(v.w, nr, np.float64), nbr_ptr,
                             arr(v.nbr, int(nbr_ptr[-1]), np.int32), v.lam, v.kappa, field, truth, colour=colour)
    finally:
        lib().qb_host_problem_destroy(h)


def muon_scene(**overrides):
    """The muon-tomography QUBO (host builder); returns ProblemArrays with .truth."""
    s = MuonScene()
    _check(lib().qb_muon_scene_defaults(C.byref(s)))
    for k, v in overrides.items():
        setattr(s, "lam" if k == "lambda" else k, v)
    h = C.c_void_p()
    _check(lib().qb_scene_muon(C.byref(s), C.byref(h)))
    return _host_problem(h, False)


def gravity_scene(**overrides):
    """The gravimetry QUBO (host builder, stations as rows, rows truncated at `radius`);
    returns ProblemArrays with .truth and the builder's layer-lattice colouring."""
    s = GravityScene()
    _check(lib().qb_gravity_scene_defaults(C.byref(s)))
    for k, v in overrides.items():
        if k in ("body_lo", "body_hi"):
            getattr(s, k)[:] = [float(t) for t in v]
        else:
            setattr(s, "lam" if k == "lambda" else k, v)
    h = C.c_void_p()
    _check(lib().qb_scene_gravity(C.byref(s), C.byref(h)))
    return _host_problem(h, True)


def voxel_medium(lo, voxel, shape, mu):
    """A qb_voxel_medium over mu (float32, x fastest); returns (struct, array) -- keep both."""
    mu = np.ascontiguousarray(mu, np.float32)
    m = VoxelMedium()
    m.lo[:] = [float(v) for v in lo]
    m.voxel = float(voxel)
    m.nx, m.ny, m.nz = (int(v) for v in shape)
    m.mu = mu.ctypes.data
    return m, mu


def muon_expose(medium, chamber, n_az, n_th, n_candidates, seed, c_min=0.34202, th_max=1.22173):
    """muon_expose_binned on the device: (open, det) uint64 counts per sky bin."""
    m, _keep = medium
    sky = MuonSky(c_min, th_max, n_az, n_th)
    o = np.zeros(n_az * n_th, np.uint64)
    d = np.zeros(n_az * n_th, np.uint64)
    _check(lib().qb_muon_expose(C.byref(m), _p(np.asarray(chamber, np.float32)), C.byref(sky), int(n_candidates),
                                int(seed) & 0xFFFFFFFF, _p(o), _p(d)))
    return o, d


def muon_rows_device(model, chambers, dirs, open_, det, site_bit, min_open, a_scale, lam=2.0, kappa=float(np.log(999.0))):
    """muon_rows + assemble + BitField::grid on the device (views x bins rays); ProblemArrays."""
    m, _keep = model
    ch = np.ascontiguousarray(chambers, np.float32).reshape(-1)
    dr = np.ascontiguousarray(dirs, np.float32).reshape(-1)
    n_views = ch.size // 3
    n_bins = dr.size // 3 // n_views
    o = np.ascontiguousarray(open_, np.uint64).reshape(-1)
    d = np.ascontiguousarray(det, np.uint64).reshape(-1)
    sb = np.ascontiguousarray(site_bit, np.int32)
    h = C.c_void_p()
    _check(lib().qb_muon_rows_device(C.byref(m), n_views, _p(ch), _p(dr), n_bins, _p(o), _p(d), _p(sb),
                                     float(min_open), float(a_scale), float(lam), float(kappa), C.byref(h)))
    return _host_problem(h, False)


def muon_scene_device(**overrides):
    """The reference's muon pipeline (candidate exposure, rows, assembly) on the device for
    the muon scene; returns (ProblemArrays with .truth, build milliseconds)."""
    s = MuonScene()
    _check(lib().qb_muon_scene_defaults(C.byref(s)))
    for k, v in overrides.items():
        setattr(s, "lam" if k == "lambda" else k, v)
    h = C.c_void_p()
    ms = C.c_double()
    _check(lib().qb_scene_muon_device(C.byref(s), C.byref(h), C.byref(ms)))
    return _host_problem(h, False), ms.value


def gravity_scene_device(**overrides):
    """gravity_scene with its rows built on the device; returns (ProblemArrays, build ms)."""
    s = GravityScene()
    _check(lib().qb_gravity_scene_defaults(C.byref(s)))
    for k, v in overrides.items():
        if k in ("body_lo", "body_hi"):
            getattr(s, k)[:] = [float(t) for t in v]
        else:
            setattr(s, "lam" if k == "lambda" else k, v)
    h = C.c_void_p()
    ms = C.c_double()
    _check(lib().qb_scene_gravity_device(C.byref(s), C.byref(h), C.byref(ms)))
    return _host_problem(h, True), ms.value


def er_graph(n=800, p=0.06, seed=1):
    """The Erdos-Renyi +-1 max-cut workload (host builder)."""
    cap = int(n * (n - 1) // 2 * min(1.0, p * 1.5 + 0.01)) + 1024
    eu = np.zeros(cap, np.int32); ev = np.zeros(cap, np.int32); ew = np.zeros(cap, np.int32)
    m = C.c_int64()
    _check(lib().qb_scene_er_graph(n, float(p), C.c_uint64(seed), _p(eu), _p(ev), _p(ew), cap, C.byref(m)))
    return GraphArrays(n, eu[:m.value], ev[:m.value], ew[:m.value])


# ---- batched sparse QUBOs ------------------------------------------------------------------

class QuboInstance:
    """One sparse QUBO E(x) = sum_i h_i x_i + sum_{i<j} J_ij x_i x_j, as a symmetric CSR
    (J_ij in row i and row j), with its replica count and seed."""

    def __init__(self, row_ptr, col, J, h, replicas, seed):
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.J = np.ascontiguousarray(J, np.float32)
        self.h = np.ascontiguousarray(h, np.float32)
        self.replicas = int(replicas)
        self.seed = int(seed)
        self.n = len(self.h)


def colour_qubo(q):
    """The engine's colouring of one instance, computed on the host: (order, class_ptr)."""
    order = np.zeros(q.n, np.int32)
    cp = np.zeros(q.n + 1, np.int32)
    nc = C.c_int()
    _check(lib().qb_colour_qubo(q.n, _p(q.row_ptr), _p(q.col), _p(q.J), _p(order), _p(cp), C.byref(nc)))
    return order, cp[:nc.value + 1].copy()


def random_regular_qubo(n, degree, seed, replicas=8):
    """A sparse QUBO on a random graph of the given degree (the union of degree / 2 random
    2-factors, plus a random perfect matching when the degree is odd; loops and repeated
    pairs are dropped and the bits left short are paired again at random, a few rounds,
    so all but a handful of bits have exactly the degree), J ~ N(0, 1) and h ~ N(0, 0.5)
    as float32, from numpy's PCG64 seeded with `seed`."""
    rng = np.random.default_rng(seed)
    us, vs = [], []
    for _ in range(degree // 2):
        perm = rng.permutation(n)
        us.append(np.arange(n)); vs.append(perm)
    if degree % 2:
        perm = rng.permutation(n)
        half = n // 2
        us.append(perm[:half]); vs.append(perm[half:2 * half])
    u = np.concatenate(us).astype(np.int64); v = np.concatenate(vs).astype(np.int64)
    keep = u != v
    a, b = np.minimum(u[keep], v[keep]), np.maximum(u[keep], v[keep])
    key = np.unique(a * n + b)
    for _ in range(30):
        deg = np.bincount(key // n, minlength=n) + np.bincount(key % n, minlength=n)
        stubs = np.repeat(np.arange(n), np.maximum(degree - deg, 0))
        if len(stubs) < 2:
            break
        stubs = rng.permutation(stubs)
        half = len(stubs) // 2
        u, v = stubs[:half], stubs[half:2 * half]
        keep = u != v
        a, b = np.minimum(u[keep], v[keep]), np.maximum(u[keep], v[keep])
        new = np.unique(a * n + b)
        new = new[~np.isin(new, key)]
        key = np.union1d(key, new)
    a, b = key // n, key % n
    w = rng.normal(0.0, 1.0, len(key)).astype(np.float32)
    rows = np.concatenate([a, b]); cols = np.concatenate([b, a]); ws = np.concatenate([w, w])
    order = np.lexsort((cols, rows))
    rows, cols, ws = rows[order], cols[order], ws[order]
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    h = rng.normal(0.0, 0.5, n).astype(np.float32)
    return QuboInstance(row_ptr, cols, ws, h, replicas, seed)


def qubo_batch_workload(n_instances=256, seed=2026, n_range=(1000, 65536), degrees=(3, 8, 32, 64),
                        replicas=(8, 32), threads=8):
    """The batched-QUBO workload: n log-uniform in n_range, degree drawn from `degrees`,
    replicas uniform in `replicas`, every instance from its own seed (derived from `seed`).
    The specs are drawn in order from one generator, then the instances are generated on
    `threads` host threads; each depends only on its spec, so the batch is the same for
    any thread count."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_instances):
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        d = int(rng.choice(degrees))
        r = int(rng.integers(replicas[0], replicas[1] + 1))
        specs.append((n, d, int(rng.integers(1, 2**62)), r))
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(lambda s: random_regular_qubo(*s), specs))


class QuboBatch:
    """Many independent sparse QUBOs resident on the device, annealed together."""

    def __init__(self, instances):
        self.instances = list(instances)
        k = len(self.instances)
        self._n = np.array([q.n for q in self.instances], np.int32)
        self._row = np.concatenate([q.row_ptr for q in self.instances]).astype(np.int32)
        self._col = np.concatenate([q.col for q in self.instances]).astype(np.int32)
        self._J = np.concatenate([q.J for q in self.instances]).astype(np.float32)
        self._h = np.concatenate([q.h for q in self.instances]).astype(np.float32)
        self._R = np.array([q.replicas for q in self.instances], np.int32)
        self._seed = np.array([q.seed for q in self.instances], np.uint64)
        d = QuboBatchDesc(k, _p(self._n), _p(self._row), _p(self._col), _p(self._J), _p(self._h), _p(self._R),
                          _p(self._seed))
        self.h = C.c_void_p()
        _check(lib().qb_qubo_batch_create(C.byref(d), C.byref(self.h)))
        _HANDLES.add(self)

    def close(self):
        if self.h:
            lib().qb_qubo_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def colouring(self, k):
        n = self.instances[k].n
        order = np.zeros(n, np.int32)
        cp = np.zeros(n + 1, np.int32)
        nc = C.c_int()
        _check(lib().qb_qubo_batch_colouring(self.h, int(k), _p(order), _p(cp), C.byref(nc)))
        return order, cp[:nc.value + 1].copy()

    def anneal(self, sched, mode=0, replica_block=0, workgroups=0, x=True, energy=True, block=0, graph=False):
        """mode 0: one persistent launch; 1: one launch per instance; graph: record the
        launch schedule into a hipGraph once and replay it. Returns x (per instance
        uint8[R, n]), energy (per instance float64[R]) and the report."""
        total_x = int(sum(q.replicas * q.n for q in self.instances))
        total_r = int(self._R.sum())
        xb = np.zeros(total_x, np.uint8) if x else None
        eb = np.zeros(total_r, np.float64) if energy else None
        rep = QuboReport()
        if (block == 0) != (replica_block == 0):
            block = block or 256
            replica_block = replica_block or 8
        o = QuboOpts(int(mode), int(replica_block), int(workgroups), int(block), int(bool(graph)))
        _check(lib().qb_qubo_batch_anneal(self.h, C.byref(sched), C.byref(o), _p(xb), _p(eb), C.byref(rep)))
        out = {"report": {f: getattr(rep, f) for f, _ in QuboReport._fields_}}
        if x:
            xs, off = [], 0
            for q in self.instances:
                xs.append(xb[off:off + q.replicas * q.n].reshape(q.replicas, q.n))
                off += q.replicas * q.n
            out["x"] = xs
        if energy:
            es, off = [], 0
            for q in self.instances:
                es.append(eb[off:off + q.replicas])
                off += q.replicas
            out["energy"] = es
        return out


# ---- dense QUBO field engine -----------------------------------------------------------------

def dense_qubo(n, seed):
    """A dense symmetric QUBO: Q_ij ~ N(0, 1/sqrt(n)) (standard deviation) for i < j, mirrored,
    and the diagonal as the linear term h; float32, from numpy's PCG64 seeded with `seed`."""
    rng = np.random.default_rng(seed)
    A = rng.normal(0.0, 1.0 / np.sqrt(n), (n, n)).astype(np.float32)
    Q = np.triu(A, 1)
    Q = Q + Q.T
    h = np.ascontiguousarray(np.diag(A))
    return np.ascontiguousarray(Q, np.float32), h.astype(np.float32)


class DenseQubo:
    """A dense QUBO resident on the device, annealed by the field engine."""

    def __init__(self, Q, h):
        self.Q = np.ascontiguousarray(Q, np.float32)
        self.hv = np.ascontiguousarray(h, np.float32)
        self.n = len(self.hv)
        self.h = C.c_void_p()
        _check(lib().qb_dense_create(self.n, _p(self.Q), _p(self.hv), C.byref(self.h)))
        _HANDLES.add(self)

    def close(self):
        if self.h:
            lib().qb_dense_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def anneal(self, sched, R, seed_offset=0, precision=FP64, recompute_every=0):
        x = np.zeros((R, self.n), np.uint8)
        e = np.zeros(R, np.float64)
        err = np.zeros(R, np.float64)
        rep = DenseReport()
        _check(lib().qb_dense_anneal(self.h, C.byref(sched), int(R), C.c_uint64(seed_offset), int(precision),
                                     int(recompute_every), _p(x), _p(e), _p(err), C.byref(rep)))
        return {"x": x, "energy": e, "field_err": err,
                "report": {f: getattr(rep, f) for f, _ in DenseReport._fields_}}


def roi_enumerate(h, J, beta=1.0):
    """The exact posterior of a region's QUBO (roi_qubo's h and J, J's upper triangle
    used) over all 2^n states on the device: dict(marginal, map_state, e_min, log_z,
    p_map, map_index, kernel_ms)."""
    h = np.ascontiguousarray(h, np.float64)
    n = len(h)
    J = np.ascontiguousarray(J, np.float64).reshape(n, n)
    m = np.zeros(n, np.float64)
    ms = np.zeros(n, np.uint8)
    r = RoiResult()
    _check(lib().qb_roi_enumerate(n, _p(h), _p(J), float(beta), _p(m), _p(ms), C.byref(r)))
    out = {f: getattr(r, f) for f, _ in RoiResult._fields_}
    out.update(marginal=m, map_state=ms)
    return out
