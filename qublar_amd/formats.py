"""On-disk formats of QuBLAR's posterior exports, written from this engine's results so
the reference's own tooling reads them unchanged.

- ``build/ising_out_*`` (src/check_ising.cu:507-526): float32 p_void grids in C order
  (z, y, x), -1 outside the unknown domain; the truth grid (one byte per voxel); every
  branch's spins in bit order, branch after branch; ``ising_out.meta``, text lines
  ``nx``, ``voxel``, ``lo``, ``branches``, ``vars``, ``lambda``, ``kappa`` printed as a C++
  stream prints them (6 significant digits).
- the region-of-interest QUBO text (src/check_ising.cu:527-592): n, then n lines
  ``voxel p_void truth h_i`` and n lines of J's row, every value with 17 significant
  digits and followed by a space.
- the machine map (docs/MACHINE-EXPORT.md:7-15, tools/export_machine_map.py): the
  tri-state grid (0 Exists, 1 NotThere, 2 Undecided, 255 outside), the traversable grid
  (1 only where Exists) and ``ising_out_machine.meta``.
- the ledger's cloud (src/ledger_out.hpp:78-86, src/check_ising.cu:596-606): one byte per
  voxel, 255 outside the domain, round(254 p) inside (halves away from zero), and its JSON
  companion.

Grids are x-fastest over nx * ny * nz cells; the reference's grids are cubes, and a
non-cubic grid adds ``ny`` and ``nz`` lines to the meta.
"""
import json
import os

import numpy as np

EXISTS, NOT_THERE, UNDECIDED, OUTSIDE = 0, 1, 2, 255
CLASSIFY_LO, CLASSIFY_HI = 0.1, 0.9


def _g6(v):
    """A number as a C++ ostream prints it by default (%g, 6 significant digits)."""
    return "%g" % v


def _g17(v):
    """A number as a C++ ostream prints it at precision 17 (%.17g)."""
    return "%.17g" % v


def voxel_grid(shape, var_voxel, values, outside, dtype):
    """A grid (C order z, y, x) holding values[i] at var_voxel[i] and `outside` elsewhere."""
    nx, ny, nz = shape
    g = np.full(nx * ny * nz, outside, dtype)
    g[np.asarray(var_voxel, np.int64)] = np.asarray(values).astype(dtype)
    return g.reshape(nz, ny, nx)


def write_ising_out(directory, shape, voxel, lo, var_voxel, p_void, branches, lam, kappa, p_void_noprior=None,
                    truth=None, prefix="ising_out"):
    """The posterior export of src/check_ising.cu:507-526 into `directory`: p_void (and
    p_void without the prior) per bit on the grid, the truth per bit as a byte grid,
    branches R x n_vars spins (bit order), and the meta."""
    os.makedirs(directory, exist_ok=True)
    path = lambda name: os.path.join(directory, f"{prefix}_{name}")
    voxel_grid(shape, var_voxel, np.asarray(p_void, np.float32), -1.0, np.float32).tofile(path("pvoid.bin"))
    if p_void_noprior is not None:
        voxel_grid(shape, var_voxel, np.asarray(p_void_noprior, np.float32), -1.0, np.float32).tofile(
            path("pvoid_noprior.bin"))
    if truth is not None:   # per bit; the grid holds 0 outside the domain
        voxel_grid(shape, var_voxel, np.asarray(truth, np.uint8), 0, np.uint8).tofile(path("truth.bin"))
    branches = np.ascontiguousarray(branches, np.uint8)
    branches.tofile(path("branches.bin"))
    nx, ny, nz = shape
    lines = [f"nx {nx}"]
    if not (nx == ny == nz):
        lines += [f"ny {ny}", f"nz {nz}"]
    lines += [f"voxel {_g6(voxel)}", "lo " + " ".join(_g6(v) for v in lo), f"branches {branches.shape[0]}",
              f"vars {len(var_voxel)}", f"lambda {_g6(lam)}", f"kappa {_g6(kappa)}"]
    with open(os.path.join(directory, f"{prefix}.meta"), "w") as f:
        f.write("\n".join(lines) + "\n")


def read_meta(path):
    """A meta file as the reference's tools read it: key -> rest of the line."""
    with open(path) as f:
        return dict(line.split(maxsplit=1) for line in f.read().splitlines() if line.strip())


def read_ising_out(directory, prefix="ising_out"):
    """The export back: dict(meta, shape, p_void grid, and the other grids / branches present)."""
    meta = read_meta(os.path.join(directory, f"{prefix}.meta"))
    nx = int(meta["nx"])
    shape = (nx, int(meta.get("ny", nx)), int(meta.get("nz", nx)))
    grid = (shape[2], shape[1], shape[0])
    out = {"meta": meta, "shape": shape,
           "p_void": np.fromfile(os.path.join(directory, f"{prefix}_pvoid.bin"), np.float32).reshape(grid)}
    extra = {"pvoid_noprior.bin": ("p_void_noprior", np.float32), "truth.bin": ("truth", np.uint8)}
    for name, (key, dt) in extra.items():
        p = os.path.join(directory, f"{prefix}_{name}")
        if os.path.exists(p):
            out[key] = np.fromfile(p, dt).reshape(grid)
    p = os.path.join(directory, f"{prefix}_branches.bin")
    if os.path.exists(p):
        out["branches"] = np.fromfile(p, np.uint8).reshape(int(meta["branches"]), int(meta["vars"]))
    return out


def classify_grid(p):
    """(tri-state, traversable) of a p_void grid (-1 outside), classify's thresholds
    (src/ising_recon.hpp:302-309) and the ahead policy: traversable only where Exists."""
    p = np.asarray(p, np.float32)
    inside = p >= 0.0
    tri = np.full(p.shape, OUTSIDE, np.uint8)
    tri[inside & (p >= np.float32(CLASSIFY_HI))] = NOT_THERE
    tri[inside & (p <= np.float32(CLASSIFY_LO))] = EXISTS
    tri[inside & (p > np.float32(CLASSIFY_LO)) & (p < np.float32(CLASSIFY_HI))] = UNDECIDED
    return tri, (tri == EXISTS).astype(np.uint8)


def write_machine_map(directory, prefix="ising_out"):
    """The machine map beside an export (docs/MACHINE-EXPORT.md:7-15): tri-state and
    traversable grids and their meta, derived from the p_void grid."""
    ex = read_ising_out(directory, prefix)
    tri, trav = classify_grid(ex["p_void"])
    tri.tofile(os.path.join(directory, f"{prefix}_tristate.bin"))
    trav.tofile(os.path.join(directory, f"{prefix}_traversable.bin"))
    meta = ex["meta"]
    text = (f"nx {meta['nx']}\n"
            f"voxel {meta['voxel']}\n"
            f"lo {meta['lo']}\n"
            f"order zyx_C\n"
            f"pvoid {prefix}_pvoid.bin float32\n"
            f"tristate {prefix}_tristate.bin uint8 0=Exists 1=NotThere 2=Undecided 255=outside\n"
            f"traversable {prefix}_traversable.bin uint8 1=Exists_only\n"
            f"classify_lo {CLASSIFY_LO}\n"
            f"classify_hi {CLASSIFY_HI}\n"
            f"policy Exists=traversable; Undecided|NotThere|outside=block\n"
            f"note does_not_include_evidence_budget; see check_ising / RESULTS\n")
    with open(os.path.join(directory, f"{prefix}_machine.meta"), "w") as f:
        f.write(text)
    return tri, trav


def write_roi(path, voxels, p_void, truth, h, J):
    """A region's conditional QUBO as text (src/check_ising.cu:575-592)."""
    n = len(h)
    J = np.asarray(J, np.float64).reshape(n, n)
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            f.write(f"{int(voxels[i])} {_g17(float(np.float32(p_void[i])))} {int(truth[i])} {_g17(float(h[i]))}\n")
        for i in range(n):
            f.write("".join(_g17(float(J[i, j])) + " " for j in range(n)) + "\n")


def read_roi(path):
    """(voxels, p_void, truth, h, J) of a region file."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0])
    rows = [ln.split() for ln in lines[1:1 + n]]
    J = np.array([[float(v) for v in ln.split()] for ln in lines[1 + n:1 + 2 * n]]).reshape(n, n)
    return (np.array([int(r[0]) for r in rows]), np.array([float(r[1]) for r in rows]),
            np.array([int(r[2]) for r in rows]), np.array([float(r[3]) for r in rows]), J)


def cloud_u8(n_voxels, var_voxel, p):
    """One byte per voxel: 255 outside the domain, round(254 p) inside, p clamped to
    [0, 1] and halves rounded away from zero."""
    g = np.full(n_voxels, 255, np.uint8)
    q = np.clip(np.asarray(p, np.float64), 0.0, 1.0) * 254.0
    g[np.asarray(var_voxel, np.int64)] = np.floor(q + 0.5).astype(np.uint8)
    return g


def write_cloud(path, shape, voxel, lo, var_voxel, p, truth_voxels=()):
    """The ledger's cloud and its JSON companion (src/check_ising.cu:596-606). The JSON is
    the reference's stream output byte for byte: keys in its order, numbers as a default
    ostream prints them (voxel 2 -> "2", lo -128 -> "-128"), no spaces. A non-cubic grid
    adds "ny" and "nz" after "nx" (the reference's grid is cubic)."""
    nx, ny, nz = shape
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    cloud_u8(nx * ny * nz, var_voxel, p).tofile(path)
    dims = f'"nx":{nx}' + ("" if nx == ny == nz else f',"ny":{ny},"nz":{nz}')
    text = ("{" + dims + f',"voxel_m":{_g6(voxel)},"lo_m":[' + ",".join(_g6(v) for v in lo) + "],"
            '"encoding":"u8: 255 outside the domain, round(254 p_void) inside",'
            f'"classify":{{"exists_max":{_g6(CLASSIFY_LO)},"not_there_min":{_g6(CLASSIFY_HI)}}},'
            '"truth_voxels":[' + ",".join(str(int(v)) for v in truth_voxels) + "]}\n")
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        f.write(text)
