"""CPU replays of three device schedules, checked against what they must preserve:

- the sweep kernel's light-bit chunks, claimed from a per-team counter: every task runs once,
  every lane's operands stay inside the class, and a class advances the counter by exactly
  n_chunks + n_waves, so every wave knows the next class's base (qb_kernels.hip, light_sweep);
- the dense engine's window decision from flip to flip: the same flips and fields, bit for bit,
  as deciding the 64 bits one step at a time (qb_dense.hip, decide_window);
- the dense team kernel's staging: each decision reads the buffer staged for its own window
  and sweep, and no staging outlives the last sweep (qb_dense.hip, dense_team_kernel)."""
import random

import numpy as np

K_TEAMS, K_LIGHT_U = 8, 2
K_CHUNK = 64 * K_LIGHT_U


def replay_light(n_lights, n_groups, team_size, team, seed):
    rng = random.Random(seed)
    my_groups = (n_groups - team + K_TEAMS - 1) // K_TEAMS if n_groups > team else 0
    if my_groups == 0:
        return
    n_waves = team_size * 4
    ctr = base = 0
    lb = 1000
    for n_light in n_lights:
        le = lb + n_light
        if n_light <= 0:
            continue
        tasks = n_light * my_groups
        n_chunks = (tasks + K_CHUNK - 1) // K_CHUNK
        seen = np.zeros(tasks, np.int32)
        got = {}
        for w in range(n_waves):   # every wave's first claim
            got[w], ctr = ctr, ctr + 1
        done = set()
        while len(done) < n_waves:
            w = rng.choice([v for v in range(n_waves) if v not in done])
            if not got[w] - base < n_chunks:
                done.add(w)
                continue
            c0 = (got[w] - base) * K_CHUNK
            for lane in range(64):
                for u in range(K_LIGHT_U):
                    t = c0 + lane + 64 * u
                    tt = t if t < tasks else c0
                    gi = tt // n_light
                    pos, grp = lb + tt - gi * n_light, team + gi * K_TEAMS
                    assert lb <= pos < le and gi < my_groups and grp < n_groups
                    if t < tasks:
                        seen[t] += 1
            got[w], ctr = ctr, ctr + 1   # the next claim
        assert (seen == 1).all()
        assert ctr - base == n_chunks + n_waves
        base += n_chunks + n_waves
        lb = le


def test_light_claims_cover_every_task_once_and_advance_the_counter_exactly():
    lights = [90154, 89706, 15155, 2540, 5, 0, 1, 127, 128, 129, 255, 256, 257]
    for n_groups in (1, 3, 8, 9, 17):
        for team in (0, 5, 7):
            for team_size in (1, 2, 32):
                replay_light(lights if team_size > 2 else lights[3:], n_groups, team_size, team,
                             100 * n_groups + team)


def decide_steps(wl, b, L, Q, f, x, T):
    f, x, flips = f.copy(), x.copy(), []
    for t in range(wl):
        dE = -np.float64(f[t]) if x[t] else np.float64(f[t])
        if dE <= 0.0 or dE * b < L[t]:
            dl = T(-1.0) if x[t] else T(1.0)
            f = (f + Q[t].astype(T) * dl).astype(T)
            x[t] ^= 1
            flips.append((t, float(dl)))
    return f, x, flips


def decide_flip_to_flip(wl, b, L, Q, f, x, T):
    f, x, flips, start = f.copy(), x.copy(), [], 0
    lanes = np.arange(64)
    while True:
        dE = np.where(x == 1, -f.astype(np.float64), f.astype(np.float64))
        accept = (lanes >= start) & (lanes < wl) & ((dE <= 0.0) | (dE * b < L))
        if not accept.any():
            return f, x, flips
        t = int(np.argmax(accept))
        dl = T(-1.0) if x[t] else T(1.0)
        f = (f + Q[t].astype(T) * dl).astype(T)
        x[t] ^= 1
        flips.append((t, float(dl)))
        start = t + 1


def test_dense_window_flip_to_flip_equals_stepping():
    rng = np.random.default_rng(11)
    for T in (np.float32, np.float64):
        for trial in range(600):
            wl = 64 if trial % 3 else int(rng.integers(1, 65))
            Q = rng.normal(0, 0.2, (64, 64)).astype(np.float32)
            Q = ((Q + Q.T) / 2).astype(np.float32)
            np.fill_diagonal(Q, 0)
            f = rng.normal(0, 1, 64).astype(T)
            x = rng.integers(0, 2, 64).astype(np.int64)
            b = float(10 ** rng.uniform(-2, 2))
            L = -np.log(rng.uniform(1e-12, 1, 64))
            a, c = decide_steps(wl, b, L, Q, f, x, T), decide_flip_to_flip(wl, b, L, Q, f, x, T)
            assert a[2] == c[2]
            assert np.array_equal(a[0].view(np.uint8), c[0].view(np.uint8)) and np.array_equal(a[1], c[1])


def test_dense_team_staging_feeds_each_decision_its_own_window():
    parts = 4
    for n_win in (4, 5, 7, 16, 64, 65):
        for n_sweeps in (1, 2, 3, 7):
            for part in range(parts):
                wb, we = part * n_win // parts, (part + 1) * n_win // parts
                buf = {0: (0, 0) if wb == 0 else None, 1: None}
                nbuf = 0
                for sweep in range(n_sweeps):
                    for w in range(n_win):
                        own = wb <= w < we
                        wn, swn = (w + 1, sweep) if w + 1 < n_win else (0, sweep + 1)
                        staging = wb <= wn < we and swn < n_sweeps
                        q = (nbuf ^ 1) if own else nbuf
                        if own:
                            assert buf[nbuf] == (w, sweep)
                            assert not staging or q != nbuf
                        if staging:   # lands before the window's closing barrier
                            buf[q] = (wn, swn)
                        if own:
                            nbuf ^= 1
                assert swn <= n_sweeps
