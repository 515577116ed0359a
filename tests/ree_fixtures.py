"""Readers for the reference's experiments/ree_cascade fixtures.

The reference (src/check_ree_qubo.cu:276-279) writes each QUBO as text: a header
"nv nterms offset P opt", then nv lines "i Qd_i cost_i/P", then nterms lines
"i j Qo_ij", every double in %.17g, the pairs in std::map order. Its annealer runs
the QUBO through a "pair-ray" BinaryProblem (check_ree_qubo.cu:172-192): one ray
(d = 0, w = 1) per off-diagonal term with a = float(sqrt|q|) on i and
float(+-sqrt|q|) on j, and a per-bit field Qd_i - 1/2 a^2 per ray touching i.
These files are read in place at test time; nothing of them is copied.
"""
from __future__ import annotations

import csv
import os

import numpy as np

REF = "/root/reference"
REE = os.path.join(REF, "experiments", "ree_cascade")


def available() -> bool:
    return os.path.isdir(REE)


def instance_ids():
    with open(os.path.join(REE, "instances.csv")) as f:
        return [row["id"] for row in csv.DictReader(f)]


def read_qubo(path):
    with open(path) as f:
        head = f.readline().split()
        nv, nt = int(head[0]), int(head[1])
        offset = float(head[2])
        qd = np.zeros(nv)
        for _ in range(nv):
            i, q, _c = f.readline().split()
            qd[int(i)] = float(q)
        pairs = []
        for _ in range(nt):
            i, j, q = f.readline().split()
            pairs.append((int(i), int(j), float(q)))
    return nv, offset, qd, pairs


def pair_ray_problem(nv, qd, pairs):
    """BinaryProblem arrays of the pair-ray encoding, in the reference's order."""
    rows = [[] for _ in range(nv)]
    field = [float(v) for v in qd]
    for r, (i, j, q) in enumerate(pairs):
        sq = float(np.sqrt(abs(q)))
        fa = float(np.float32(sq))
        fb = float(np.float32(sq if q >= 0 else -sq))
        rows[i].append((r, fa))
        rows[j].append((r, fb))
        field[i] -= 0.5 * fa * fa
        field[j] -= 0.5 * fa * fa
    row_ptr = [0]
    ray, a = [], []
    for i in range(nv):
        for r, v in rows[i]:
            ray.append(r)
            a.append(v)
        row_ptr.append(len(ray))
    n_rays = len(pairs)
    return dict(row_ptr=np.array(row_ptr, np.int32), ray=np.array(ray, np.int32),
                a=np.array(a, np.float32), d=np.zeros(n_rays), w=np.ones(n_rays),
                nbr_ptr=np.zeros(nv + 1, np.int32), nbr=np.zeros(0, np.int32), lam=0.0, kappa=0.0,
                field=np.array(field))


def qubo_energy(offset, qd, pairs, x):
    """qubo_energy (check_ree_qubo.cu:158-163), in its summation order."""
    e = offset
    for i in range(len(qd)):
        if x[i]:
            e += qd[i]
    for i, j, q in pairs:
        if x[i] and x[j]:
            e += q
    return e


def engine_runs():
    """{id: [energy string for seed 1..64]} from engine_runs.csv (%.12g)."""
    out = {}
    with open(os.path.join(REE, "engine_runs.csv")) as f:
        for row in csv.DictReader(f):
            out.setdefault(row["id"], {})[int(row["seed"])] = row["energy"]
    return out


def energy_trace(iid):
    """{seed: [energy string per sweep]} from energy_trace_<id>.csv (%.10g)."""
    out = {}
    with open(os.path.join(REE, f"energy_trace_{iid}.csv")) as f:
        for row in csv.DictReader(f):
            out.setdefault(int(row["seed"]), []).append(row["energy"])
    return out
