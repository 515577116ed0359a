"""A serving loop for ray x voxel QUBO solves on one GPU.

Requests (a problem, its branch count, its seed) arrive over time. In batched mode every
launch anneals the block-diagonal union of the requests waiting when the GPU comes free,
admitted in arrival order while the union's device footprint stays within an HBM budget;
a host thread assembles and uploads the next union while the GPU anneals the current one,
so host preparation and device work overlap. A request's branches in a union are, spin
for spin, its branches annealed alone with its own seed (per-bit draw keys), and its
p_void averages its own branches, so batching changes when a result arrives, never what
it is. Sequential mode solves one request per
launch in arrival order, for comparison.

This is new around the reference: QuBLAR solves one problem per branch_fractions call
(src/ising_recon.hpp:148-166); batching many requests into one device solve has no
counterpart there.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import FP32, Problem, schedule, union


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
            q.p_void, q.tri = branch_agreement(out["branches"][:q.branches, b0:b0 + q.problem.n_vars])
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
