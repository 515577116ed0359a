"""The speculative-window bookkeeping of the device's reference-trajectory integer engine,
restated on the CPU and checked against the oracle's sequential loop (wising::anneal,
src/ising_weighted.hpp:250-267).

The device runs one wave per replica and decides 64 consecutive bits at once: the k-th
uphill proposal of a window draws splitmix64 at state + k * gamma, the first accepted lane
flips, and the later lanes keep their decisions up to the first lane whose field that flip
changed. This file replays exactly that bookkeeping (masks, prefix counts, the dirty bound)
with W lanes and requires every replica to end in the oracle's state and energy: a wrong
draw offset, a decision reused past a changed field or a window advanced too far shows up
as a different trajectory.
"""
import ctypes as C
import math

import numpy as np
import pytest

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def popc(m):
    return bin(m).count("1")


def ctz(m):
    return (m & -m).bit_length() - 1


def window_anneal(n, eu, ev, ew, temps, seed, W=64):
    """One replica by the device's window algorithm; returns (x, energy)."""
    nb = [[] for _ in range(n)]
    for u, v, w in zip(eu, ev, ew):
        nb[int(u)].append((int(v), int(w)))
        nb[int(v)].append((int(u), int(w)))
    full = (1 << W) - 1
    s = [-1] * n
    f = [-sum(w for _, w in nb[i]) for i in range(n)]
    st = (seed * 0x2545F4914F6CDD1D + 1) & M64
    e = 0
    for T in temps:
        i0 = 0
        while i0 < n:
            lanes = [i0 + j for j in range(W)]
            fp = [f[i] if i < n else 0 for i in lanes]
            dE = [-s[i] * fp[j] if i < n else 0 for j, i in enumerate(lanes)]
            m_up = sum(1 << j for j, i in enumerate(lanes) if i < n and dE[j] > 0)
            m_acc = 0
            for j, i in enumerate(lanes):
                if i >= n:
                    continue
                if dE[j] <= 0:
                    m_acc |= 1 << j
                else:
                    k = popc(m_up & ((1 << j) - 1))
                    u = (mix64((st + (k + 1) * GAMMA) & M64) >> 11) * 2.0 ** -53
                    if u < math.exp(-dE[j] / T):
                        m_acc |= 1 << j
            todo, ok = full, full
            while True:
                cand = m_acc & todo & ok
                if cand == 0:
                    st = (st + popc(m_up & todo & ok) * GAMMA) & M64
                    i0 += W if ok == full else ctz(~ok & full)
                    break
                js = ctz(cand)
                upto = (1 << (js + 1)) - 1
                st = (st + popc(m_up & todo & upto) * GAMMA) & M64
                i = lanes[js]
                e += dE[js]
                si = s[i]
                for v, w in nb[i]:
                    f[v] -= 2 * w * si
                s[i] = -si
                todo = full & ~upto
                dirty = sum(1 << j for j, q in enumerate(lanes) if q < n and f[q] != fp[j]) & todo
                if dirty:
                    ok &= (1 << ctz(dirty)) - 1
    return np.array([1 if v > 0 else 0 for v in s], np.uint8), e


def random_graph(n, p, wmax, seed):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(len(iu[0])) < p
    eu, ev = iu[0][keep], iu[1][keep]
    ew = rng.integers(-wmax, wmax + 1, len(eu)) if wmax > 1 else np.ones(len(eu), np.int64)
    return eu, ev, ew


@pytest.mark.parametrize("n,p,wmax,W,hot", [
    (70, 0.15, 1, 64, 3.0),     # sparse +-1 max-cut: long runs of reused decisions
    (70, 0.15, 1, 8, 3.0),      # many windows per sweep, the bound at a window's end
    (16, 1.0, 1000, 64, 1000.0),  # signed K16: every later lane is a neighbour
    (130, 0.05, 5, 64, 8.0),    # a partial last window, zero-sum fields
])
def test_window_bookkeeping_equals_sequential_loop(orc, n, p, wmax, W, hot):
    eu, ev, ew = random_graph(n, p, wmax, n + W)
    og = orc.IntGraph(n, eu, ev, ew)
    sched = orc.schedule(hot, 0.05, 30, 5)
    temp = orc.lib().orc_temperature
    temps = [temp(C.byref(sched), k) for k in range(35)]
    for seed in (1, 2, 9):
        x, e = window_anneal(n, eu, ev, ew, temps, seed, W)
        xo, eo = orc.anneal_int(og, sched, seed)
        assert np.array_equal(x, xo), f"seed {seed}: {int((x != xo).sum())} spins differ"
        assert e == eo


def test_window_bookkeeping_catches_a_stale_decision(orc):
    """The same replay with the dirty bound removed (a decision kept after its field
    changed) leaves the oracle's trajectory on a dense graph: the check has teeth."""
    n = 16
    eu, ev, ew = random_graph(n, 1.0, 1000, 5)
    og = orc.IntGraph(n, eu, ev, ew)
    sched = orc.schedule(1000.0, 0.05, 30, 5)
    temp = orc.lib().orc_temperature
    temps = [temp(C.byref(sched), k) for k in range(35)]

    def stale(seed):
        nbf = [[] for _ in range(n)]
        for u, v, w in zip(eu, ev, ew):
            nbf[int(u)].append((int(v), int(w)))
            nbf[int(v)].append((int(u), int(w)))
        s = [-1] * n
        f = [-sum(w for _, w in nbf[i]) for i in range(n)]
        st = (seed * 0x2545F4914F6CDD1D + 1) & M64
        for T in temps:   # one window covers all 16 bits; every decision kept
            dE = [-s[i] * f[i] for i in range(n)]
            k = 0
            for i in range(n):
                if dE[i] <= 0:
                    acc = True
                else:
                    k += 1
                    acc = (mix64((st + k * GAMMA) & M64) >> 11) * 2.0 ** -53 < math.exp(-dE[i] / T)
                if acc:
                    for v, w in nbf[i]:
                        f[v] -= 2 * w * s[i]
                    s[i] = -s[i]
            st = (st + k * GAMMA) & M64
        return np.array([1 if v > 0 else 0 for v in s], np.uint8)

    differs = 0
    for seed in (1, 2, 3, 4):
        xo, _ = orc.anneal_int(og, sched, seed)
        differs += not np.array_equal(stale(seed), xo)
    assert differs > 0
