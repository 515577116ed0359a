"""Serving loops for ray x voxel QUBO solves on one GPU.

Requests (a problem, its branch count, its seed) arrive over time.

ContinuousServer batches continuously. The GPU anneals the block-diagonal union of the
requests in flight a stage (a few sweeps) at a time, with their branch state kept on the
device. Between stages, requests that have finished their schedule leave: their outputs
are reduced on the device over their own branches. When a host thread has prepared the
next union (the requests still running plus those that arrived and fit the HBM budget),
it replaces the current one at a stage boundary: the running requests' spins, residuals
and clocks are imported on the device, so nobody restarts, and newcomers start at their
own sweep 0. Every bit anneals at its own sweep (temperature, Philox counter) and keys
its draws by its request's seed and local index, so a request's result is, spin for spin,
its result annealed alone, whenever it joins and whatever it shares the GPU with.

Server is the earlier form: a union is formed only when the GPU comes free and annealed
to the end in one launch; or, in sequential mode, one request per launch, for comparison.

This is new around the reference: QuBLAR solves one problem per branch_fractions call
(src/ising_recon.hpp:148-166); batching many requests into one device solve has no
counterpart there.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import FP32, Problem, Stage, schedule, union


@dataclass
class Request:
    problem: object        # ProblemArrays
    branches: int
    seed: int
    arrival: float         # seconds after the start of the run
    done: float = -1.0
    batch: int = -1
    p_void: np.ndarray = field(default=None, repr=False)
    tri: np.ndarray = field(default=None, repr=False)
    colour: object = field(default=None, repr=False)   # a future of the request's colouring
    keep_branches: bool = False                          # ContinuousServer: also return the branches
    branches_out: np.ndarray = field(default=None, repr=False)


def device_bytes_estimate(P, branches):
    """Device bytes a request occupies in a union: structure (entries, colour-major copy,
    chain descriptors, ray lists, per-bit and per-ray arrays) and the branch state."""
    nnz = int(P.row_ptr[-1])
    groups = (branches + 1) // 2
    return 64 * nnz + 96 * P.n_vars + 32 * P.n_rays + groups * (2 * P.n_vars + 8 * P.n_rays)


class Server:
    def __init__(self, sched, precision=FP32, recompute_every=50, hbm_budget=64 << 30, max_batch_bits=2_000_000,
                 colour_threads=8):
        self.sched = sched
        self.precision = precision
        self.k = recompute_every if precision == FP32 else 0
        self.budget = hbm_budget
        self.max_bits = max_batch_bits
        self.pool = ThreadPoolExecutor(colour_threads)

    def _take(self, queue, now):
        """The batch of waiting requests (arrived by `now`) that fits the budget, in order."""
        batch, used, bits = [], 0, 0
        while queue and queue[0].arrival <= now:
            q = queue[0]
            b = device_bytes_estimate(q.problem, q.branches)
            if batch and (used + b > self.budget or bits + q.problem.n_vars > self.max_bits):
                break
            batch.append(queue.pop(0))
            used += b
            bits += q.problem.n_vars
        return batch

    def _prepare(self, batch):
        U, offs = union([q.problem for q in batch], [q.seed for q in batch], [q.colour.result() for q in batch])
        return Problem(U), offs, 0, max(q.branches for q in batch)

    def _admit(self, requests, t0):
        """Each request is coloured on a host thread as it arrives (in time order)."""
        for q in requests:
            wait = q.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            q.colour = self.pool.submit(q.problem.colours)

    def _finish(self, batch, prep, t0, index):
        prob, offs, seed, R = prep
        out = prob.anneal_branches(self.sched, R, seed, self.precision, self.k, branches=True, outputs=False,
                                   energy=False)
        for q, (b0, _) in zip(batch, offs):
            mine = out["branches"][:q.branches, b0:b0 + q.problem.n_vars]
            q.p_void, q.tri = branch_agreement(mine)
            if q.keep_branches:
                q.branches_out = mine.copy()
        now = time.perf_counter() - t0
        for q in batch:
            q.done = now
            q.batch = index
        return out["report"]

    def run(self, requests, batched=True):
        """Serves `requests` (sorted by arrival) in real time; returns per-batch reports."""
        queue = sorted(requests, key=lambda q: q.arrival)
        t0 = time.perf_counter()
        reports = []
        admit = threading.Thread(target=self._admit, args=(list(queue), t0), daemon=True)
        admit.start()
        if not batched:
            for i, q in enumerate(queue):
                wait = q.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                self._wait_admitted([q])
                reports.append(self._finish([q], self._prepare([q]), t0, i))
            return reports
        pending = None             # (batch, prepared) ready to launch
        index = 0
        while queue or pending:
            if pending is None:
                now = time.perf_counter() - t0
                if queue and queue[0].arrival > now:
                    time.sleep(queue[0].arrival - now)
                batch = self._take(queue, time.perf_counter() - t0)
                self._wait_admitted(batch)
                pending = (batch, self._prepare(batch))
            batch, prep = pending
            pending = None
            # while this batch anneals, prepare the next one from what has arrived by then
            nxt = {}

            def prepare_next():
                time.sleep(0.0)
                b = self._take(queue, time.perf_counter() - t0)
                if b:
                    self._wait_admitted(b)
                    nxt["v"] = (b, self._prepare(b))

            th = threading.Thread(target=prepare_next)
            th.start()
            reports.append(self._finish(batch, prep, t0, index))
            th.join()
            index += 1
            pending = nxt.get("v")
        return reports

    @staticmethod
    def _wait_admitted(batch):
        while any(q.colour is None for q in batch):
            time.sleep(0.0005)


@dataclass
class _Member:
    req: Request
    bit0: int          # the request's first bit and ray in the union
    ray0: int
    start: int         # own sweeps done when the union's stage started


class ContinuousServer:
    """Continuous batching over stages of `stage_sweeps` sweeps (see the module notes).

    With FP32 residuals and a recompute period k, stage_sweeps must be a multiple of k:
    requests join and move only at stage boundaries, so every request's FP64 recompute
    falls on its own k-th sweeps as when it is annealed alone."""

    def __init__(self, sched, precision=FP32, recompute_every=10, stage_sweeps=10, hbm_budget=64 << 30,
                 max_union_bits=2_000_000, colour_threads=8, group_width=2):
        self.sched = sched
        self.total = sched.anneal_sweeps + sched.hold_sweeps
        self.precision = precision
        self.k = recompute_every if precision == FP32 else 0
        if self.k and stage_sweeps % self.k:
            raise ValueError("stage_sweeps must be a multiple of the recompute period")
        self.stage_sweeps = stage_sweeps
        self.budget = hbm_budget
        self.max_bits = max_union_bits
        self.group_width = group_width
        self.pool = ThreadPoolExecutor(colour_threads)
        self.prep_log = []

    def _union(self, running, newcomers):
        """(problem, stage, members) of a new union: the running members (with their own
        sweeps done) followed by the newcomers."""
        t0 = time.perf_counter()
        reqs = [m.req for m in running] + newcomers
        U, offs = union([q.problem for q in reqs], [q.seed for q in reqs], [q.colour.result() for q in reqs])
        t1 = time.perf_counter()
        prob = Problem(U)
        t2 = time.perf_counter()
        R = max(q.branches for q in reqs)
        stage = Stage(prob, self.sched, R, self.precision, self.k, group_width=self.group_width)
        t3 = time.perf_counter()
        self.prep_log.append({"bits": U.n_vars, "requests": len(reqs), "assemble_s": t1 - t0, "upload_s": t2 - t1,
                              "stage_s": t3 - t2})
        members = [_Member(q, b0, r0, 0) for q, (b0, r0) in zip(reqs, offs)]
        return prob, stage, members, len(running)

    def _take(self, queue, now, used, bits, ready_only=True):
        """The arrived requests that fit the budget, in arrival order; with ready_only,
        only those whose colouring is done (the rest wait for the next union, so the
        stage loop never waits on a host colouring)."""
        batch = []
        k = 0
        while k < len(queue) and queue[k].arrival <= now:
            q = queue[k]
            if ready_only and (q.colour is None or not q.colour.done()):
                k += 1
                continue
            b = device_bytes_estimate(q.problem, q.branches)
            if (batch or used) and (used + b > self.budget or bits + q.problem.n_vars > self.max_bits):
                break
            batch.append(queue.pop(k))
            used += b
            bits += q.problem.n_vars
        return batch

    def run(self, requests):
        """Serves `requests` in real time; returns per-stage reports."""
        queue = sorted(requests, key=lambda q: q.arrival)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        threading.Thread(target=Server._admit, args=(self, list(queue), t0), daemon=True).start()
        reports = []
        cur = None                  # (problem, stage, members)
        prep = None                 # a future of the next union
        leaving = []                # running members the union in preparation leaves out
        union_index = 0
        n_left = len(queue)
        prep_est, stage_est = 0.5, 0.05   # running estimates of a preparation and a stage, s
        while n_left:
            if cur is None:
                if prep is None:
                    if queue and queue[0].arrival > clock():
                        time.sleep(queue[0].arrival - clock())
                    fresh = self._take(queue, clock(), 0, 0, ready_only=False)
                    Server._wait_admitted(fresh)
                    prep = self.pool.submit(self._union, [], fresh)
                prob, stage, members, _ = prep.result()
                prep = None
                for m in members:
                    stage.set_span(m.bit0, m.req.problem.n_vars, 0, m.req.branches)
                cur = (prob, stage, members)
                union_index += 1
            prob, stage, members = cur
            running = [m for m in members if m.req.done < 0]
            # Prepare the next union while this stage runs, when requests are waiting or
            # finished ones hold a large share of the union. It leaves out the members that
            # will finish before it is ready (their remaining sweeps fit the stages a
            # preparation takes), and the switch waits for them, so finished requests do
            # not idle in a union.
            if prep is None:
                horizon = self.stage_sweeps * (int(np.ceil(prep_est / max(stage_est, 1e-3))) + 1)
                keep = [m for m in running if self.total - (m.start + stage.sweeps) > horizon]
                used = sum(device_bytes_estimate(m.req.problem, m.req.branches) for m in keep)
                bits = sum(m.req.problem.n_vars for m in keep)
                fresh = self._take(queue, clock(), used, bits)
                idle_bits = sum(m.req.problem.n_vars for m in members if m.req.done >= 0)
                if (fresh or (running and idle_bits > bits)) and (keep or fresh):
                    Server._wait_admitted(fresh)
                    keep_ids = {id(m) for m in keep}
                    leaving = [m for m in running if id(m) not in keep_ids]
                    prep = self.pool.submit(self._union, keep, fresh)
            rep = stage.run(self.stage_sweeps)
            stage_est = 0.8 * stage_est + 0.2 * rep["wall_s"]
            rep["union"] = union_index
            rep["requests"] = len(running)
            reports.append(rep)
            done_sweeps = stage.sweeps
            now = clock()
            for m in members:
                if m.req.done < 0 and m.start + done_sweeps >= self.total:
                    out = stage.outputs(m.bit0, m.req.problem.n_vars, m.req.branches, branches=m.req.keep_branches)
                    m.req.p_void, m.req.tri = out["p_void"], out["tri"]
                    m.req.branches_out = out["branches"]
                    m.req.done = clock()
                    m.req.batch = union_index
                    n_left -= 1
            if prep is not None and prep.done() and all(m.req.done >= 0 for m in leaving):
                nprob, nstage, nmembers, n_running = prep.result()
                prep = None
                leaving = []
                last = self.prep_log[-1]
                prep_est = 0.7 * prep_est + 0.3 * (last["assemble_s"] + last["upload_s"] + last["stage_s"])
                old = {id(m.req): m for m in members}
                for nm in nmembers[:n_running]:
                    om = old[id(nm.req)]
                    nm.start = om.start + done_sweeps
                    nstage.import_span(stage, om.bit0, nm.bit0, nm.req.problem.n_vars, om.ray0, nm.ray0,
                                       nm.req.problem.n_rays)
                for nm in nmembers[n_running:]:
                    nstage.set_span(nm.bit0, nm.req.problem.n_vars, 0, nm.req.branches)
                stage.close()
                prob.close()
                cur = (nprob, nstage, nmembers)
                union_index += 1
            elif not any(m.req.done < 0 for m in members) and prep is None:
                stage.close()
                prob.close()
                cur = None
        return reports


def branch_agreement(branches):
    """p_void and the tri-state map of a request's own branches (R x n spins), as the
    engine computes them (src/ising_recon.hpp:148-166, 302-309): p = float(count) /
    float(R); 1 (NotThere) where p >= 0.9, 0 (Exists) where p <= 0.1, else 2. A request
    in a union averages its own R branches, not the union's."""
    R = branches.shape[0]
    pv = np.count_nonzero(branches, axis=0).astype(np.float32) / np.float32(R)
    tri = np.where(pv >= np.float32(0.9), 1, np.where(pv <= np.float32(0.1), 0, 2)).astype(np.uint8)
    return pv, tri


def latency_summary(requests):
    lat = np.array([q.done - q.arrival for q in requests])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
