#!/usr/bin/env python3
"""Benchmark: spin updates/s of the 306,328-bit muon-tomography QUBO, 16 branches.

A step is one whole job of the hot path: branch_fractions of the muon QUBO
(src/ising_recon.hpp:148-166) -- 16 branches annealed from all-zero under the
posterior schedule (T 50 -> 1 over 200 sweeps, then 50 sweeps at T = 1,
src/ising_recon.hpp:81-86), then the branch agreement (p_void, tri-state, certainty)
and each branch's FP64 energy, copied back to the host. The problem structure is
resident in HBM before the timed region; value = K * 16 * 250 * n_vars / wall.

Prints one JSON line. Also reports the sweep kernel against the MI355X HBM roofline
(algorithmic bytes per launch / launch duration from HIP events on the engine's
stream) and the CPU oracle -- the reference's own trajectory (index order, splitmix,
FP64), run on this host -- as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin updates/s, 306k-bit muon QUBO ×16 branches; p50/p99 batched solve ms"
HBM_PEAK_GBS = 8000.0
# The reference's published rate for this workload: branch_fractions, 16 branches x 250
# sweeps of the 306,328-bit muon QUBO in 7.0 s on 16 host threads
# (experiments/p1_bench/check_log.txt:162-163): 16 * 250 * 306,328 / 7.0.
REFERENCE_RATE = 16 * 250 * 306328 / 7.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_sweep(P, R, res_bytes):
    """Compulsory unique bytes of one sweep of all branches: the structure read once
    (CSR offsets, ray index + a per entry, neighbour CSR, d + w per ray) plus each
    branch's spins and residuals read and written once."""
    n, nnz, nn, nr = P.n_vars, int(P.row_ptr[-1]), int(P.nbr_ptr[-1]), P.n_rays
    structure = 4 * (n + 1) + 8 * nnz + 4 * (n + 1) + 4 * nn + 16 * nr
    state = R * 2 * (n / 8.0 + res_bytes * nr)
    return structure + state


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ray_team_kernel_muon16_traffic.md")


def profiled_traffic():
    """HBM-side bytes per sweep-kernel launch of this workload at the default settings,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the same
    counters on a plain 1 GiB copy); None when the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline(P, seconds, threads):
    """The oracle's reference trajectory (index order, splitmix64 drawn on uphill moves,
    FP64 residuals) -- anneal_branch restated -- on `threads` host threads, 16 branches,
    sweeps scaled so the sample runs about `seconds`."""
    import oracle

    op = oracle.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
    probe = oracle.schedule(50.0, 1.0, 2, 0)
    t0 = time.perf_counter()
    oracle.branch_fractions(op, probe, 16, threads)
    per_sweep = (time.perf_counter() - t0) / 2
    sweeps = int(max(2, min(250, seconds / max(per_sweep, 1e-6))))
    anneal = max(1, int(sweeps * 0.8))
    sched = oracle.schedule(50.0, 1.0, anneal, sweeps - anneal)
    t0 = time.perf_counter()
    oracle.branch_fractions(op, sched, 16, threads)
    wall = time.perf_counter() - t0
    rate = 16 * sweeps * P.n_vars / wall
    # one branch on one thread, a few sweeps: the single-thread rate
    s1 = max(1, min(10, int(seconds / 8 / max(per_sweep * 16 / threads, 1e-6))))
    t0 = time.perf_counter()
    oracle.anneal_branch(op, oracle.schedule(50.0, 1.0, s1, 0), 1)
    single = s1 * P.n_vars / (time.perf_counter() - t0)
    return dict(value=rate, unit="spin updates/s", cores=threads, kind="port",
                sample=f"16 branches x {sweeps} sweeps ({anneal} + {sweeps - anneal}) of the same QUBO, "
                       f"reference trajectory (index order, splitmix64, FP64), {wall:.1f} s",
                single_thread_value=single, host_cpus=os.cpu_count())


def serving_pool(n, seed):
    """n distinct ray x voxel requests of the serving stream: voxels log-uniform in
    [2k, 300k] (pyramid grid 24..128), rays log-uniform in [1k, 75k] (sky bins), the
    same kappa / lambda prior, each generated from its own seed."""
    import qublar_amd as qa

    rng = np.random.default_rng(seed)
    pool = []
    for i in range(n):
        voxels = float(np.exp(rng.uniform(np.log(2e3), np.log(3e5))))
        rays = float(np.exp(rng.uniform(np.log(1e3), np.log(7.5e4))))
        grid = int(np.clip(round((voxels / 0.146) ** (1.0 / 3.0)), 24, 128))
        th = int(np.clip(round(np.sqrt(rays / 15.0)), 4, 70))
        pool.append(qa.muon_scene(grid=grid, sky_az=min(360, 5 * th), sky_th=th, log2_muons=24, seed=1000 + i))
    return pool


def serving_stream(pool, rate, seconds, seed):
    """Poisson arrivals at `rate` per second over `seconds`: each request a pool problem
    drawn at random, 8-16 branches, its own seed."""
    from qublar_amd.serving import Request

    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        P = pool[int(rng.integers(len(pool)))]
        reqs.append(Request(P, int(rng.integers(8, 17)), 100 + len(reqs), t, keep_branches=True))
    return reqs


def serve(mode, pool, sched, rate, seconds, seed=11, stage_sweeps=10, k=10, max_union_bits=2_000_000):
    """Serves one stream in `mode` (continuous, unions or sequential); returns the summary
    and the requests (each with its results)."""
    import qublar_amd as qa
    from qublar_amd.serving import ContinuousServer, Server, latency_summary

    import threading

    reqs = serving_stream(pool, rate, seconds, seed)
    total = sched.anneal_sweeps + sched.hold_sweeps
    t0 = time.perf_counter()
    running = threading.Event()

    def progress():   # a line every 20 s while the stream is served
        while not running.wait(20.0):
            log(f"serving {mode}: {sum(q.done >= 0 for q in reqs)} of {len(reqs)} requests done, "
                f"{time.perf_counter() - t0:.0f} s")

    threading.Thread(target=progress, daemon=True).start()
    srv = None
    if mode == "continuous":
        srv = ContinuousServer(sched, qa.FP32, k, stage_sweeps=stage_sweeps, max_union_bits=max_union_bits)
        reports = srv.run(reqs)
    else:
        reports = Server(sched, qa.FP32, k).run(reqs, batched=mode == "unions")
    running.set()
    wall = time.perf_counter() - t0
    updates = sum(q.branches * total * q.problem.n_vars for q in reqs)
    kms = np.array([r["kernel_ms"] for r in reports])
    out = {"mode": {"continuous": "continuous batching (stage unions, state kept on the device)",
                    "unions": "a union per launch, formed when the GPU comes free",
                    "sequential": "one solve per launch"}[mode],
           "offered_req_per_s": rate, "arrival_window_s": seconds, "requests": len(reqs),
           "launches": len(reports), "sweep_kernel_ms": {"mean": float(kms.mean()), "max": float(kms.max())},
           "gpu_busy_fraction": float(kms.sum() / 1e3 / wall), "wall_s": wall,
           "spin_updates_per_s": updates / wall, "latency_ms": latency_summary(reqs)}
    if mode == "continuous":
        out["stage_sweeps"] = stage_sweeps
        out["max_union_bits"] = max_union_bits
        out["unions"] = len({r["union"] for r in reports})
        pl = srv.prep_log
        out["union_prep_s"] = {k: float(np.mean([p[k] for p in pl])) for k in ("assemble_s", "upload_s", "stage_s")}
        out["union_bits_mean"] = float(np.mean([p["bits"] for p in pl]))
        out["union_requests_mean"] = float(np.mean([p["requests"] for p in pl]))
    return out, reqs


def final_energies(pool, reqs, limit=64):
    """FP64 energies of the first `limit` requests' branches, each on its own problem."""
    import qublar_amd as qa

    probs = {}
    e = []
    for q in reqs[:limit]:
        if id(q.problem) not in probs:
            probs[id(q.problem)] = qa.Problem(q.problem)
        e.append(probs[id(q.problem)].energy(q.branches_out)[0])
    for p in probs.values():
        p.close()
    return e


def run_serving(args):
    """Poisson arrivals of ray x voxel requests (8-16 branches, annealed to T = 1) for
    --serve-seconds, served by the continuous server, then by unions formed at launch
    boundaries, then one solve per launch; prints one JSON line per mode with p50 / p99
    latency and the served spin-update rate, and checks that the modes end at the same
    final energies (every request's branches are its solve alone in each mode)."""
    import qublar_amd as qa

    pool = serving_pool(args.pool, 7)
    log("pool sizes:", sorted(P.n_vars for P in pool))
    sched = qa.schedule(50.0, 1.0, 200, 50)
    # the staged sweep kernel against the plain one on one union of 40 pool requests: 25
    # sweeps each, every bit at its own sweep 0 in the staged form
    rng = np.random.default_rng(3)
    probs = [pool[int(i)] for i in rng.integers(0, len(pool), 40)]
    U, offs = qa.union(probs, list(range(100, 140)))
    prob = qa.Problem(U)
    plain, staged = [], []
    for _ in range(3):
        plain.append(prob.anneal_branches(qa.schedule(50.0, 1.0, 20, 5), 16, 0, qa.FP32, 25, branches=False,
                                          outputs=False, energy=False)["report"]["kernel_ms"])
        st = qa.Stage(prob, sched, 16, qa.FP32, 25, group_width=2)
        for (b0, _), P in zip(offs, probs):
            st.set_span(b0, P.n_vars, 0, 12)
        staged.append(st.run(25)["kernel_ms"])
        st.close()
    prob.close()
    print(json.dumps({"workload": "serving", "staged_kernel_overhead": {
        "union_bits": U.n_vars, "sweeps": 25, "plain_kernel_ms": float(np.median(plain)),
        "staged_kernel_ms": float(np.median(staged))}}), flush=True)
    energies = {}
    for mode in ("continuous", "unions", "sequential"):
        out, reqs = serve(mode, pool, sched, args.rate, args.serve_seconds, stage_sweeps=args.stage_sweeps,
                          k=args.stage_sweeps)
        energies[mode] = final_energies(pool, reqs)
        same = all(np.array_equal(a, b) for a, b in zip(energies[mode], energies["continuous"]))
        out.update({"workload": "serving", "same_final_energies_as_continuous": same,
                    "data": "synthetic muon-tomography QUBOs; each request coloured on arrival on a host thread"})
        print(json.dumps(out), flush=True)


def precision_rows(qa, prob, R, sched):
    """FP64, then FP32 with an FP64 recompute every k in {1, 10, 100} sweeps and with
    none, the same keys and order: per run the final carry error, the energies against
    FP64's, the spins and tri-state cells that agree, and the time per solve."""
    runs = [("fp64", qa.FP64, 0), ("fp32 k=1", qa.FP32, 1), ("fp32 k=10", qa.FP32, 10),
            ("fp32 k=100", qa.FP32, 100), ("fp32 no recompute", qa.FP32, 0)]
    base = None
    rows = []
    for name, prec, k in runs:
        prob.anneal_branches(qa.schedule(50.0, 1.0, 2, 0), R, 0, prec, k, branches=False)   # warm
        log(f"{name}: {R} branches x {sched.anneal_sweeps + sched.hold_sweeps} sweeps ...")
        t = time.perf_counter()
        out = prob.anneal_branches(sched, R, 0, prec, k)
        wall = time.perf_counter() - t
        log(f"{name}: {wall:.2f} s")
        row = {"run": name, "solve_s": wall, "kernel_ms": out["report"]["kernel_ms"],
               "fp32_residual_err": out["report"]["fp32_residual_err"] if prec == qa.FP32 else 0.0,
               "energy_mean": float(np.mean(out["energy"]))}
        if base is None:
            base = out
        else:
            de = (out["energy"] - base["energy"]) / np.abs(base["energy"])
            row.update({"energy_rel_diff_vs_fp64": {"mean": float(de.mean()), "min": float(de.min()),
                                                    "max": float(de.max())},
                        "spins_equal_fraction": float(np.mean(out["branches"] == base["branches"])),
                        "branches_identical": int(np.sum(np.all(out["branches"] == base["branches"], axis=1))),
                        "tri_state_equal_fraction": float(np.mean(out["tri"] == base["tri"]))})
        rows.append((row, out))
    return rows


def run_precision(args):
    """The FP32 residual carry against full FP64 on the muon QUBO: 64 branches, T 50 -> 1
    over 400 sweeps + 100 at T = 1, the same keys and order. For FP32 with an FP64
    recompute every k sweeps (and with none) it reports the final carry error
    max |r_fp32 - r_fp64(scratch)|, the branches' final energies against the FP64 run's,
    how many spins and tri-state cells agree, and the time per solve."""
    import qublar_amd as qa

    prob = qa.Problem(qa.muon_scene())
    rows = precision_rows(qa, prob, 64, qa.schedule(50.0, 1.0, 400, 100))
    print(json.dumps({"workload": "precision", "problem": "muon QUBO 306,328 bits, 64 branches, 500 sweeps",
                      "rows": [r for r, _ in rows]}), flush=True)


def run_gravity_precision(args):
    """The same study on the gravimetry QUBO: 200,000 cells of 5 m under 20,000 surface
    stations, rows truncated at 50 m (39.1M entries, ~196 stations per bit), 16 branches,
    T 50 -> 1 over 200 sweeps + 50 at T = 1. Adds the spin-update rate of each run, the
    builder's and the upload's times, the recovered body, and the host oracle's rate."""
    import qublar_amd as qa

    t0 = time.perf_counter()
    P = qa.gravity_scene()
    t_scene = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob = qa.Problem(P)
    t_upload = time.perf_counter() - t0
    info = prob.info()
    log(f"gravity QUBO {P.n_vars} bits, {P.n_rays} stations, {int(P.row_ptr[-1])} entries, {info['n_classes']} "
        f"classes; scene {t_scene:.2f} s, upload {t_upload:.2f} s")
    R = args.gravity_branches
    A = args.gravity_sweeps * 4 // 5
    H = args.gravity_sweeps - A
    rows = precision_rows(qa, prob, R, qa.schedule(50.0, 1.0, A, H))
    updates = R * (A + H) * P.n_vars
    # gravity fixes mass well and depth poorly (src/op_gravity.hpp:13-14): count the cells
    # the branches agree on, those of the body, and those in the body's columns
    cols = P.truth.reshape(-1, P.n_rays).any(axis=0)
    for row, out in rows:
        row["spin_updates_per_s"] = updates / row["solve_s"]
        on = (out["tri"] == 1).reshape(-1, P.n_rays)
        row["cells_exists"] = int(on.sum())
        row["exists_in_body"] = int((on & (P.truth.reshape(-1, P.n_rays) == 1)).sum())
        row["exists_in_body_columns"] = int((on & cols[None, :]).sum())
    # the sweep is bound by its 2,586 dependent class steps, not by their width: more
    # branches per solve cost little more time
    wide = {}
    for R2 in ((32, 64) if R == 16 and A + H == 250 else ()):
        prob.anneal_branches(qa.schedule(50.0, 1.0, 2, 0), R2, 0, qa.FP32, 10, branches=False)   # warm
        t = time.perf_counter()
        out = prob.anneal_branches(qa.schedule(50.0, 1.0, A, H), R2, 0, qa.FP32, 10, branches=False)
        wall = time.perf_counter() - t
        log(f"fp32 k=10, {R2} branches: {wall:.2f} s")
        wide[str(R2)] = {"solve_s": wall, "spin_updates_per_s": R2 * (A + H) * P.n_vars / wall}
    result = {"workload": "gravity-precision",
              "problem": "gravimetry QUBO: 100 x 200 x 10 cells of 5 m, 20,000 stations, rows truncated at 50 m; "
                         f"{R} branches, T 50 -> 1 over {A} sweeps + {H} at T = 1",
              "n_vars": P.n_vars, "n_rays": P.n_rays, "entries": int(P.row_ptr[-1]),
              "colour_classes": info["n_classes"], "body_cells": int(P.truth.sum()),
              "scene_s": t_scene, "upload_s": t_upload, "rows": [r for r, _ in rows],
              "fp32_k10_more_branches": wide,
              "data": "synthetic: the body's truncated anomaly plus N(0, 10 microGal) noise"}
    if not args.no_cpu:
        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        log(f"cpu baseline on {threads} threads ...")
        result["cpu_baseline"] = cpu_baseline(P, args.cpu_seconds, threads)
    print(json.dumps(result), flush=True)


def run_maxcut(args):
    """The reference's GPU integer-annealer benchmark shape (tools/gpu_anneal.cu --bench,
    experiments/p2_gpu): a G1-like graph -- Erdos-Renyi n = 800, p = 0.06, all weights +1
    (m ~ 19,176) -- 16,384 replicas with seeds 1.., 4000 + 200 sweeps under scaled_budget
    (T_hot = sigma = sqrt(2 sum w^2 / n), T_cold = 0.1 w_min; src/ising_gpu.cuh:104-124),
    one run_batch per step; spin updates/s = R x sweeps x n / wall (state init to results
    on the host), against the reference's published 1.114e10 on an RTX 5060 Ti."""
    import qublar_amd as qa
    import oracle

    g = qa.er_graph(800, 0.06, 1)
    g = qa.GraphArrays(g.n, g.eu, g.ev, np.abs(g.ew))
    sigma = float(np.sqrt(2.0 * float(np.sum(g.ew.astype(np.float64) ** 2)) / g.n))
    A, H, R = 4000, 200, args.maxcut_replicas
    sched = qa.schedule(sigma, 0.1, A, H)
    G = qa.Graph(g)
    plan = G.plan()
    log(f"G1-like graph n={g.n} m={g.m}; {plan['n_classes']} colour classes; sigma {sigma:.3f}")
    G.run_batch(qa.schedule(sigma, 0.1, 2, 0), 1, R)   # warm
    walls, kms = [], []
    for st in range(args.steps):
        t = time.perf_counter()
        s, e, rep = G.run_batch(sched, 1, R)
        walls.append(time.perf_counter() - t)
        kms.append(rep["kernel_ms"])
        log(f"step {st}: {walls[-1]:.3f} s (kernels {rep['kernel_ms']:.1f} ms, {rep['launches']} launches); best cut {-e.min()}")
    updates = R * (A + H) * g.n
    value = updates / float(np.mean(walls))
    # the reference's own trajectory (index order, splitmix64 on uphill proposals), once
    G.run_batch_reference(qa.schedule(sigma, 0.1, 2, 0), 1, R)   # warm
    t = time.perf_counter()
    _, e_ref, rep_ref = G.run_batch_reference(sched, 1, R)
    wall_ref = time.perf_counter() - t
    log(f"reference trajectory: {wall_ref:.3f} s (kernels {rep_ref['kernel_ms']:.1f} ms); best cut {-e_ref.min()}")
    result = {"workload": "maxcut-g1", "metric": "spin updates/s, G1-like max-cut, 16,384 replicas x 4200 sweeps",
              "value": value, "unit": "spin updates/s", "steps": args.steps, "ms_per_step": float(np.mean(walls)) * 1e3,
              "kernel_ms": float(np.mean(kms)), "vs_reference_gpu": value / 1.114e10,
              "reference": "1.114e10 spin updates/s, RTX 5060 Ti (experiments/p2_gpu/summary.csv:2)",
              "config": {"n": g.n, "m": g.m, "replicas": R, "sweeps": A + H, "t_hot": sigma, "t_cold": 0.1,
                         "colour_classes": plan["n_classes"]},
              "best_cut": int(-e.min()), "mean_cut": float(-e.mean()), "dtype": "int32 fields, f64 accept",
              "reference_trajectory": {"wall_s": wall_ref, "kernel_ms": rep_ref["kernel_ms"],
                                       "spin_updates_per_s": updates / wall_ref,
                                       "vs_reference_gpu": updates / wall_ref / 1.114e10,
                                       "best_cut": int(-e_ref.min()), "mean_cut": float(-e_ref.mean()),
                                       "trajectory": "wising::anneal per replica (index order, splitmix64)"},
              "data": "synthetic Erdos-Renyi graph (the G-set files are not available)"}
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor

        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        og = oracle.IntGraph(g.n, g.eu, g.ev, g.ew)
        osched = oracle.schedule(sigma, 0.1, A, H)
        t = time.perf_counter()
        oracle.anneal_int(og, osched, 1)
        one = time.perf_counter() - t
        runs = max(threads, int(args.cpu_seconds / one) * threads // threads)
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda k: oracle.anneal_int(og, osched, k + 1), range(runs)))
        wall = time.perf_counter() - t
        result["cpu_baseline"] = {"value": runs * (A + H) * g.n / wall, "unit": "spin updates/s", "cores": threads,
                                  "kind": "port", "sample": f"{runs} replicas of the same schedule on {threads} "
                                  f"threads, the oracle's index-order splitmix trajectory, {wall:.1f} s"}
    print(json.dumps(result), flush=True)


def run_reference_float(args):
    """The float path in the reference's own trajectory (anneal_branch: index order,
    splitmix64) on a small muon QUBO (32^3 grid, 90 x 20 sky bins: ~5k bits) with
    --ref-branches branches, T 50 -> 1 over 200 sweeps + 50 at T = 1, against the host
    oracle running the same trajectory (branch_fractions restated) on min(16, nproc - 2)
    threads: like for like, bit for bit the same branches."""
    import qublar_amd as qa
    import oracle

    P = qa.muon_scene(grid=32, sky_az=90, sky_th=20, log2_muons=22)
    prob = qa.Problem(P)
    R, A, H = args.ref_branches, 200, 50
    sched = qa.schedule(50.0, 1.0, A, H)
    prob.anneal_branches_reference(qa.schedule(50.0, 1.0, 2, 0), R, 0, branches=False)   # warm
    t = time.perf_counter()
    out = prob.anneal_branches_reference(sched, R, 0, branches=False)
    wall = time.perf_counter() - t
    updates = R * (A + H) * P.n_vars
    log(f"reference trajectory: {R} branches x {P.n_vars} bits: {wall:.2f} s (kernel {out['report']['kernel_ms']:.0f} ms)")
    result = {"workload": "reference-float", "n_vars": P.n_vars, "branches": R, "sweeps": A + H,
              "solve_s": wall, "spin_updates_per_s": updates / wall,
              "trajectory": "anneal_branch per branch (index order, splitmix64 on uphill proposals)"}
    if not args.no_cpu:
        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        op = oracle.FloatProblem(P.row_ptr, P.ray, P.a, P.d, P.w, P.nbr_ptr, P.nbr, P.lam, P.kappa)
        rc = 4 * threads
        t = time.perf_counter()
        oracle.branch_fractions(op, oracle.schedule(50.0, 1.0, A, H), rc, threads)
        cw = time.perf_counter() - t
        result["cpu_baseline"] = {"value": rc * (A + H) * P.n_vars / cw, "unit": "spin updates/s", "cores": threads,
                                  "kind": "port", "sample": f"{rc} branches of the same trajectory, {cw:.1f} s"}
        result["vs_cpu"] = result["spin_updates_per_s"] / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


def run_roi(args):
    """The exact posterior of a region's QUBO over all 2^n states (the enumeration of
    tools/branches_blaze.py:45-53 over roi_qubo's h and J): n = --roi-n bits with h ~ N(0, 2),
    J_ij ~ N(0, 1); both device passes (minimum, then Z and the marginals) per step. The
    host baseline is the oracle's brute force (energy from scratch per state) at n = 21,
    as states per second."""
    import qublar_amd as qa
    import oracle

    n = args.roi_n
    rng = np.random.default_rng(31)
    h = rng.normal(0, 2, n)
    J = np.triu(rng.normal(0, 1, (n, n)), 1)
    qa.roi_enumerate(h[:16], J[:16, :16])    # warm
    walls, kms = [], []
    for s in range(args.steps):
        t = time.perf_counter()
        out = qa.roi_enumerate(h, J, 1.0)
        walls.append(time.perf_counter() - t)
        kms.append(out["kernel_ms"])
        log(f"step {s}: {walls[-1] * 1e3:.1f} ms (kernels {kms[-1]:.1f} ms)")
    states = 2.0 ** n
    result = {"workload": "roi-enumeration", "n": n, "states": states, "steps": args.steps,
              "ms_per_step": float(np.mean(walls)) * 1e3, "kernel_ms": float(np.mean(kms)),
              "states_per_s": states / float(np.mean(walls)), "kernel_states_per_s": states / (float(np.mean(kms)) / 1e3),
              "e_min": out["e_min"], "log_z": out["log_z"], "p_map": out["p_map"],
              "data": "synthetic h ~ N(0, 2), J ~ N(0, 1)"}
    if not args.no_cpu:
        m = 21
        t = time.perf_counter()
        oracle.roi_enumerate(h[:m], J[:m, :m], 1.0)
        wall = time.perf_counter() - t
        result["cpu_baseline"] = {"value": 2.0 ** m / wall, "unit": "states/s", "cores": 1, "kind": "port",
                                  "sample": f"the oracle's brute force at n = {m} ({wall:.1f} s)"}
    print(json.dumps(result), flush=True)


def qubo_cpu_baseline(inst, sched_args, seconds, threads):
    """The oracle's batch trajectory on host threads, one (instance, replica) per task, on
    the first instances of the batch with the sweep count cut so the sample runs about
    `seconds`: spin updates/s."""
    import oracle
    import qublar_amd as qa
    from concurrent.futures import ThreadPoolExecutor

    t_hot, t_cold = sched_args
    sample = inst[:32]
    orders = [qa.colour_qubo(q)[0] for q in sample]
    probe = oracle.schedule(t_hot, t_cold, 1, 0)
    t0 = time.perf_counter()
    oracle.qubo_anneal(sample[0].row_ptr, sample[0].col, sample[0].J, sample[0].h, orders[0], probe, 1, 1)
    per_upd = (time.perf_counter() - t0) / sample[0].n
    total_upd = sum(q.n * q.replicas for q in sample)
    sweeps = int(max(1, min(2000, seconds * threads / max(per_upd * total_upd, 1e-9))))
    sched = oracle.schedule(t_hot, t_cold, sweeps, 0)

    def task(k):
        q = sample[k]
        oracle.qubo_anneal(q.row_ptr, q.col, q.J, q.h, orders[k], sched, q.seed, q.replicas)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(task, range(len(sample))))
    wall = time.perf_counter() - t0
    return dict(value=total_upd * sweeps / wall, unit="spin updates/s", cores=threads, kind="port",
                sample=f"the batch's first {len(sample)} instances ({total_upd} replica-bits), {sweeps} sweeps of "
                       f"the same schedule's start, the engine's trajectory (oracle), {wall:.1f} s")


def run_batched(args):
    """256 sparse QUBOs (n log-uniform in [1k, 64k], degree 3/8/32/64, J ~ N(0,1),
    h ~ N(0,0.5), 8-32 replicas), 2000 sweeps T 5 -> 0.05: one persistent launch (replica
    block picked by timing each on a short run) against one launch per instance."""
    import qublar_amd as qa

    t0 = time.perf_counter()
    inst = qa.qubo_batch_workload(args.instances, seed=2026)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    B = qa.QuboBatch(inst)
    t_up = time.perf_counter() - t0
    n = np.array([q.n for q in inst]); nnz = np.array([len(q.col) for q in inst]); R = np.array([q.replicas for q in inst])
    log(f"batch: {len(inst)} instances, {n.sum()} bits, {nnz.sum()} entries, {int((n * R).sum())} replica-bits; "
        f"generate {t_gen:.1f} s, check+colour+upload {t_up:.1f} s")
    sweeps = args.batch_sweeps
    sched = qa.schedule(5.0, 0.05, sweeps, 0)
    upd = float((n * R).sum()) * sweeps
    first = B.anneal(sched, mode=0, replica_block=0, x=False, energy=False)
    rb, blk = first["report"]["replica_block"], first["report"]["block"]
    log(f"tuned shape: {blk} threads, replica block {rb}: {first['report']['kernel_ms']:.1f} ms")
    runs = []
    for s in range(args.steps):
        out = B.anneal(sched, mode=0, replica_block=rb, block=blk, x=False, energy=(s == args.steps - 1))
        runs.append(out["report"])
        log(f"persistent run {s}: kernel {out['report']['kernel_ms']:.1f} ms, wall {out['report']['wall_s'] * 1e3:.1f} ms")
    # the same comparison at a tenth of the sweeps: one persistent launch against one
    # launch per instance (each launch as wide as the instance's jobs)
    short = qa.schedule(5.0, 0.05, max(1, sweeps // 10), 0)
    pers_short = B.anneal(short, mode=0, replica_block=rb, block=blk, x=False, energy=False)["report"]
    log(f"persistent, {short.anneal_sweeps} sweeps: kernel {pers_short['kernel_ms']:.1f} ms")
    per = B.anneal(short, mode=1, replica_block=rb, block=blk, x=False, energy=False)["report"]
    log(f"one launch per instance, {short.anneal_sweeps} sweeps: kernel {per['kernel_ms']:.1f} ms, "
        f"{per['launches']} launches")
    # the short schedule's launches recorded into a hipGraph and replayed (the first call
    # records): host wall per call, eager against replayed, in both modes
    graphs = {}
    for mode in (0, 1):
        eager, replay = [], []
        B.anneal(short, mode=mode, replica_block=rb, block=blk, x=False, energy=False, graph=True)   # records
        for _ in range(3):
            eager.append(B.anneal(short, mode=mode, replica_block=rb, block=blk, x=False, energy=False)["report"])
            replay.append(B.anneal(short, mode=mode, replica_block=rb, block=blk, x=False, energy=False,
                                   graph=True)["report"])
        assert all(r["graph_replayed"] for r in replay)
        graphs["persistent" if mode == 0 else "per_instance"] = {
            "eager_wall_ms": float(np.mean([r["wall_s"] for r in eager])) * 1e3,
            "replay_wall_ms": float(np.mean([r["wall_s"] for r in replay])) * 1e3,
            "eager_kernel_ms": float(np.mean([r["kernel_ms"] for r in eager])),
            "replay_kernel_ms": float(np.mean([r["kernel_ms"] for r in replay]))}
        log(f"mode {mode}: eager {graphs['persistent' if mode == 0 else 'per_instance']}")
    upd_short = float((n * R).sum()) * short.anneal_sweeps
    kern = float(np.mean([r["kernel_ms"] for r in runs]))
    wall = float(np.mean([r["wall_s"] for r in runs]))
    jobs_words = sum(int(q.n) * ((q.replicas + rb - 1) // rb) for q in inst)
    per_sweep = float(np.sum(4 * (n + 1) + 8 * nnz + 8 * n)) + 8.0 * jobs_words
    alg = per_sweep * sweeps
    result = {
        "workload": "batched",
        "metric": "spin updates/s, batched sparse QUBOs",
        "value": upd / wall,
        "unit": "spin updates/s",
        "higher_is_better": True,
        "steps": args.steps,
        "kernel_ms": kern,
        "wall_ms": wall * 1e3,
        "kernel_updates_per_s": upd / (kern * 1e-3),
        "per_instance_launches": {"sweeps": short.anneal_sweeps, "kernel_ms": per["kernel_ms"],
                                  "updates_per_s": upd_short / (per["kernel_ms"] * 1e-3), "launches": per["launches"],
                                  "persistent_kernel_ms": pers_short["kernel_ms"]},
        "persistent_speedup_vs_per_instance": per["kernel_ms"] / pers_short["kernel_ms"],
        "recorded_graph": dict(graphs, sweeps=short.anneal_sweeps),
        "replica_block": rb,
        "block": blk,
        "jobs": runs[-1]["jobs"],
        "config": {"workload": f"{len(inst)} sparse QUBOs, n log-uniform in [1k, 64k], degree in {{3, 8, 32, 64}}, "
                               f"J ~ N(0, 1), h ~ N(0, 0.5), 8-32 replicas, {sweeps} sweeps T 5 -> 0.05",
                   "bits": int(n.sum()), "entries": int(nnz.sum()), "replica_bits": int((n * R).sum())},
        "roofline": {"bound": "hbm", "achieved": alg / (kern * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "qubo_batch_kernel", "algorithmic_bytes_per_launch": alg},
        "dtype": "f64 dE, f32 couplings",
        "data": "synthetic",
    }
    if not args.no_cpu:
        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        log(f"cpu baseline on {threads} threads")
        result["cpu_baseline"] = qubo_cpu_baseline(inst, (5.0, 0.05), args.cpu_seconds, threads)
    print(json.dumps(result), flush=True)


def run_dense_precision(args):
    """The dense half of the precision study: one dense symmetric QUBO, N = 4096, Q_ij ~
    N(0, 1/sqrt(N)) float32, 64 replicas, 5000 sweeps T 1 -> 0.01. Full-FP64 fields first,
    then FP32 fields carried incrementally with an FP64 recompute every k in {1, 10, 100}
    sweeps (and never): final field error against FP64 from scratch, energies against the
    FP64 run's, identical replicas, time."""
    import qublar_amd as qa

    n, R, sweeps = args.dense_n, 64, args.dense_sweeps
    Q, h = qa.dense_qubo(n, 4096)
    D = qa.DenseQubo(Q, h)
    sched = qa.schedule(1.0, 0.01, sweeps, 0)
    runs = [("fp64", qa.FP64, 0), ("fp32 k=1", qa.FP32, 1), ("fp32 k=10", qa.FP32, 10),
            ("fp32 k=100", qa.FP32, 100), ("fp32 no recompute", qa.FP32, 0)]
    base = None
    rows = []
    for name, prec, k in runs:
        out = D.anneal(sched, R, 0, prec, k)
        rep = out["report"]
        log(f"{name}: kernel {rep['kernel_ms']:.1f} ms, max field error {out['field_err'].max():.3g}")
        row = {"run": name, "kernel_ms": rep["kernel_ms"], "spin_updates_per_s": rep["spin_updates"] / (rep["kernel_ms"] * 1e-3),
               "field_err_max": float(out["field_err"].max()), "field_err_mean": float(out["field_err"].mean()),
               "energy_mean": float(out["energy"].mean())}
        if base is None:
            base = out
        else:
            de = (out["energy"] - base["energy"]) / np.abs(base["energy"])
            row.update({"energy_rel_diff_vs_fp64": {"mean": float(de.mean()), "min": float(de.min()), "max": float(de.max())},
                        "replicas_identical": int(np.sum(np.all(out["x"] == base["x"], axis=1))),
                        "spins_equal_fraction": float(np.mean(out["x"] == base["x"]))})
        rows.append(row)
    result = {"workload": "dense precision", "problem": f"dense QUBO N={n}, Q ~ N(0, 1/sqrt(N)) float32, {R} replicas, "
              f"{sweeps} sweeps T 1 -> 0.01", "rows": rows}
    if not args.no_cpu:
        import oracle
        from concurrent.futures import ThreadPoolExecutor
        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        few = max(1, min(sweeps, 400))
        osch = oracle.schedule(1.0, 0.01, few, 0)
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda b: oracle.dense_anneal(Q, h, osch, b + 1, 1, 10), range(threads)))
        wall = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": threads * few * n / wall, "unit": "spin updates/s", "cores": threads,
                                  "kind": "port", "sample": f"{threads} replicas x {few} sweeps, FP32 fields, k=10, "
                                  f"{wall:.1f} s"}
    print(json.dumps(result), flush=True)


def run_muon_frozen(args):
    """freeze_provable_rock then the anneal of the reduced problem (src/ising_recon.hpp:181-234;
    the reference's check_ising path, experiments/p1_bench/check_log.txt:196-198: 3.59 s + 0.16 s
    against 7.0 s for the full problem on 16 host threads). The frozen set is found on the
    device (qb_freeze, margin 10 nats), the reduced problem is built on the host and coloured by
    the full problem's colouring restricted to the kept bits, then annealed like the headline
    (16 branches, 200 + 50 sweeps, FP32 with an FP64 recompute every 50)."""
    import qublar_amd as qa

    P = qa.muon_scene()
    prob = qa.Problem(P)
    order, cp = prob.colouring()
    colour = np.zeros(P.n_vars, np.int32)
    for c in range(len(cp) - 1):
        colour[order[cp[c]:cp[c + 1]]] = c
    sched = qa.schedule(50.0, 1.0, 200, 50)
    prob.anneal_branches(sched, 16, 0, qa.FP32, 50, branches=False)   # warm
    full = []
    for _ in range(3):
        t = time.perf_counter()
        out_full = prob.anneal_branches(sched, 16, 0, qa.FP32, 50, branches=False)
        full.append(time.perf_counter() - t)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        frozen, rounds = prob.freeze(10.0)
        t1 = time.perf_counter()
        R, kept = P.reduce_frozen(frozen)
        R.colour = colour[kept]
        t2 = time.perf_counter()
        rp = qa.Problem(R)
        t3 = time.perf_counter()
        out = rp.anneal_branches(sched, 16, 0, qa.FP32, 50, branches=False)
        t4 = time.perf_counter()
        rp.close()
        runs.append({"freeze_s": t1 - t0, "reduce_s": t2 - t1, "create_s": t3 - t2, "anneal_s": t4 - t3,
                     "total_s": t4 - t0, "kernel_ms": out["report"]["kernel_ms"]})
        log(f"frozen {int(frozen.sum())} bits in {rounds} rounds; {runs[-1]}")
    best = min(runs, key=lambda r: r["total_s"])
    tri = np.zeros(P.n_vars, np.uint8)          # frozen bits are rock: Exists
    tri[kept] = out["tri"]
    agree = float(np.mean(tri == out_full["tri"]))
    n_red = int(R.n_vars)
    result = {"workload": "muon-frozen",
              "n_vars": P.n_vars, "frozen": int(frozen.sum()), "kept": n_red, "rounds": rounds,
              "full_solve_s": float(np.min(full)), "full_kernel_ms": out_full["report"]["kernel_ms"],
              "frozen_path": best,
              "speedup_vs_full": float(np.min(full)) / best["total_s"],
              "spin_updates_per_s_reduced": 16 * 250 * n_red / best["anneal_s"],
              "tri_state_agreement_with_full": agree,
              "not_there": {"full": int((out_full["tri"] == 1).sum()), "frozen_path": int((tri == 1).sum())},
              "reference": "3.59 s anneal + 0.16 s freeze against 7.0 s full, 16 host threads "
                           "(experiments/p1_bench/check_log.txt:196-198)"}
    print(json.dumps(result), flush=True)


def run_builders(args):
    """The problem builders on the device against the host builders: the muon QUBO (the
    reference's candidate exposure, 2^27 muons per chamber, rows, assembly, neighbour
    lists) and the gravimetry QUBO (39M-entry rows), build wall time each, best of 3."""
    import qublar_amd as qa

    out = {"workload": "builders"}
    for name, dev, host in (("muon 128^3, 3 x 2^27 candidates", lambda: qa.muon_scene_device(), qa.muon_scene),
                            ("gravity 100 x 200 x 10", lambda: qa.gravity_scene_device(), qa.gravity_scene)):
        dev()   # warm: code objects, allocator
        best, best_wall = None, None
        for _ in range(3):
            t = time.perf_counter()
            P, ms = dev()
            wall = time.perf_counter() - t
            if best is None or ms < best:
                best, best_wall = ms, wall
        t = time.perf_counter()
        H = host()
        host_s = time.perf_counter() - t
        out[name] = {"device_build_ms": best, "device_wall_incl_python_copy_s": best_wall, "host_build_s": host_s,
                     "n_vars": P.n_vars, "n_rays": P.n_rays, "entries": int(P.row_ptr[-1]),
                     "host_entries": int(H.row_ptr[-1])}
        log(name, out[name])
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--recompute", type=int, default=50, help="FP64 residual recompute every k sweeps")
    ap.add_argument("--workgroups", type=int, default=0, help="persistent workgroups (multiple of 8); 0 = one per CU")
    ap.add_argument("--group-width", type=int, default=0, help="branches per group (2, 4, 8, 16); 0 = auto")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dense-n", type=int, default=4096, help="dense precision: bits")
    ap.add_argument("--dense-sweeps", type=int, default=5000, help="dense precision: sweeps")
    ap.add_argument("--workload", choices=["muon", "serving", "precision", "batched", "dense-precision",
                                           "gravity-precision", "roi", "maxcut", "reference-float", "builders",
                                           "muon-frozen"],
                    default="muon",
                    help="muon: the headline line (default); serving: the batched serving stream; "
                         "batched: many sparse QUBOs in one persistent launch")
    ap.add_argument("--instances", type=int, default=256, help="batched: instances")
    ap.add_argument("--gravity-branches", type=int, default=16, help="gravity-precision: branches (replicas)")
    ap.add_argument("--gravity-sweeps", type=int, default=250,
                    help="gravity-precision: sweeps, 4/5 annealing T 50 -> 1 and 1/5 held at T = 1")
    ap.add_argument("--roi-n", type=int, default=32, help="roi: bits enumerated (2^n states)")
    ap.add_argument("--maxcut-replicas", type=int, default=16384, help="maxcut: replicas per batch")
    ap.add_argument("--ref-branches", type=int, default=4096, help="reference-float: branches")
    ap.add_argument("--batch-sweeps", type=int, default=2000, help="batched: sweeps")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: Poisson arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serving: arrival window")
    ap.add_argument("--pool", type=int, default=24, help="serving: distinct request problems")
    ap.add_argument("--stage-sweeps", type=int, default=10,
                    help="serving: sweeps per continuous-batching stage (also the FP64 recompute period there)")
    ap.add_argument("--headline-serve-seconds", type=float, default=8.0,
                    help="muon: arrival window of the batched-serving latency beside the headline (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain 1 GiB device copy (HBM counter calibration) and print its rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this engine has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import qublar_amd as qa

    info = qa.device_info()
    log("device", info)
    if args.workload == "serving":
        run_serving(args)
        return
    if args.workload == "precision":
        run_precision(args)
        return
    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "dense-precision":
        run_dense_precision(args)
        return
    if args.workload == "gravity-precision":
        run_gravity_precision(args)
        return
    if args.workload == "roi":
        run_roi(args)
        return
    if args.workload == "maxcut":
        run_maxcut(args)
        return
    if args.workload == "reference-float":
        run_reference_float(args)
        return
    if args.workload == "builders":
        run_builders(args)
        return
    if args.workload == "muon-frozen":
        run_muon_frozen(args)
        return
    if args.copy_probe:
        nbytes, reps = 1 << 30, 5
        ms = qa.copy_probe(nbytes, reps)
        print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms": ms, "read_plus_write_GBs": 2 * nbytes / ms / 1e6}),
              flush=True)
        return
    t0 = time.perf_counter()
    P = qa.muon_scene()
    t_scene = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob = qa.Problem(P)
    t_upload = time.perf_counter() - t0
    pinfo = prob.info()
    log(f"problem {P.n_vars} bits, {P.n_rays} rays, {int(P.row_ptr[-1])} entries; {pinfo['n_classes']} colour "
        f"classes; scene {t_scene:.2f} s, colour+upload {t_upload:.2f} s")
    R, A, H = 16, 200, 50
    sched = qa.schedule(50.0, 1.0, A, H)
    prec = qa.FP32 if args.precision == "fp32" else qa.FP64
    k = args.recompute if prec == qa.FP32 else 0

    def step():
        return prob.anneal_branches(sched, R, 0, prec, k, args.workgroups, group_width=args.group_width,
                                    branches=False)

    for w in range(args.warmup):
        out = step()
        log(f"warmup {w}: {out['report']['wall_s'] * 1e3:.1f} ms")
    walls, kms = [], []
    t_all = time.perf_counter()
    for s in range(args.steps):
        t1 = time.perf_counter()
        out = step()
        walls.append(time.perf_counter() - t1)
        kms.append(out["report"]["kernel_ms"])
        log(f"step {s}: {walls[-1] * 1e3:.1f} ms (sweep kernel {kms[-1]:.1f} ms)")
    total = time.perf_counter() - t_all
    updates = R * (A + H) * P.n_vars
    value = args.steps * updates / total
    n_found = int((out["tri"] == 1).sum())
    truth_hit = int(((out["tri"] == 1) & (P.truth == 1)).sum())
    res_bytes = 4 if prec == qa.FP32 else 8
    bytes_launch = (A + H) * algorithmic_bytes_per_sweep(P, R, res_bytes)
    launch_s = float(np.mean(kms)) / 1e3
    achieved = bytes_launch / launch_s / 1e9
    default_shape = prec == qa.FP32 and k == 50 and args.workgroups == 0 and args.group_width == 0
    traffic = profiled_traffic() if default_shape else None
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "spin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_RATE,
        "dtype": "f32" if prec == qa.FP32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": "muon-tomography QUBO, 306,328 bits x 75,600 rays x 3,829,376 entries, 16 branches, "
                        "T 50 -> 1 over 200 sweeps + 50 at T = 1",
            "n_vars": P.n_vars, "n_rays": P.n_rays, "entries": int(P.row_ptr[-1]),
            "colour_classes": pinfo["n_classes"], "branches": R, "sweeps": A + H,
            "arith": ("FP32 residual carry, FP64 recompute every %d sweeps, FP64 dE and energy" % k)
            if prec == qa.FP32 else "FP64 residuals, dE and energy",
        },
        "latency_ms": {"p50": float(np.percentile(walls, 50) * 1e3), "p99": float(np.percentile(walls, 99) * 1e3),
                       "kind": "the headline solve, one per call, sequential"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "ray_team_kernel", "launch_ms": launch_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_launch},
        "result_check": {"not_there": n_found, "not_there_in_void": truth_hit,
                         "fp32_residual_err": out["report"]["fp32_residual_err"],
                         "energy_mean": float(np.mean(out["energy"]))},
    }
    if args.headline_serve_seconds > 0:
        # the metric's batched solve latency: the serving stream (50 requests/s, 2k-300k
        # voxels, 8-16 branches, the same schedule) under continuous batching, beside one
        # solve per launch of the same requests; both end at the same final energies
        pool = serving_pool(args.pool, 7)
        lat = {}
        for mode in ("unions", "continuous", "sequential"):
            o, reqs = serve(mode, pool, sched, args.rate, args.headline_serve_seconds, stage_sweeps=args.stage_sweeps,
                            k=args.stage_sweeps)
            lat[mode] = (o, final_energies(pool, reqs, 16))
            log(f"serving {mode}: p50 {o['latency_ms']['p50_ms']:.0f} ms, p99 {o['latency_ms']['p99_ms']:.0f} ms")
        same = all(all(np.array_equal(a, b) for a, b in zip(lat[m][1], lat["sequential"][1]))
                   for m in ("unions", "continuous"))

        def summary(m):
            return {"p50": lat[m][0]["latency_ms"]["p50_ms"], "p99": lat[m][0]["latency_ms"]["p99_ms"],
                    "served_spin_updates_per_s": lat[m][0]["spin_updates_per_s"]}

        result["batched_latency_ms"] = {
            "p50": lat["unions"][0]["latency_ms"]["p50_ms"], "p99": lat["unions"][0]["latency_ms"]["p99_ms"],
            "kind": "serving stream (50 requests/s, 2k-300k voxels, 8-16 branches, this schedule): requests "
                    "batched into block-diagonal unions, one launch per union",
            "continuous_batching": summary("continuous"), "one_solve_per_launch": summary("sequential"),
            "requests": lat["unions"][0]["requests"], "arrival_window_s": args.headline_serve_seconds,
            "offered_req_per_s": args.rate, "same_final_energies_in_every_mode": same,
        }
    if not args.no_cpu:
        threads = max(1, min(16, (os.cpu_count() or 2) - 2))
        log(f"cpu baseline on {threads} threads ...")
        result["cpu_baseline"] = cpu_baseline(P, args.cpu_seconds, threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
