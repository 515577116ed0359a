"""Pins the CPU oracle against known answers: the reference's published fixtures,
closed forms, independent implementations and brute force.

Each test names what it pins and the reference line the arithmetic follows.
"""
import csv
import math
import os
import struct

import numpy as np
import pytest

from conftest import need_reference
import ree_fixtures as rf


# ---- primitives -------------------------------------------------------------------

def test_philox_random123_kat(orc):
    """Philox4x32-10 known-answer vectors of Random123 (Salmon et al., SC'11)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert tuple(orc.philox(ctr, key)) == want


def test_philox_matches_rocrand(orc):
    """96 vectors from rocRAND's independent header implementation (tests/golden)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "philox_rocrand.txt")
    n = 0
    for line in open(path):
        if line.startswith("#"):
            continue
        v = [int(t, 16) for t in line.split()]
        assert tuple(orc.philox(v[0:4], v[4:6])) == tuple(v[6:10])
        n += 1
    assert n == 96


def test_philox_uniform_layout(orc):
    """The draw is (o0 << 32 | o1) >> 11 scaled by 2^-53, counter (bit, sweep, 0, 0)."""
    seed, sweep, bit = 0x0123456789ABCDEF, 17, 4242
    o = orc.philox((bit, sweep, 0, 0), (seed & 0xFFFFFFFF, seed >> 32))
    z = (o[0] << 32) | o[1]
    assert orc.lib().orc_philox_uniform(seed, sweep, bit) == (z >> 11) * 2.0 ** -53


def test_splitmix64_known_outputs(orc):
    """SplitMix64 (Steele, Lea, Flood 2014) from state 0: published first outputs."""
    import ctypes as C
    s = C.c_uint64(0)
    got = [orc.lib().orc_splitmix(C.byref(s)) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def _ulps(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def test_qexp_accuracy_against_libm(orc):
    """qexp stays within 2 ulp of the C library exp over the acceptance range."""
    q = orc.lib().orc_qexp
    rng = np.random.default_rng(5)
    xs = np.concatenate([-np.logspace(-12, np.log10(700.0), 4000), rng.uniform(-700, 0, 4000),
                         rng.uniform(-1, 0, 4000), [0.0, -0.5 * math.log(2), -math.log(2), -1.0]])
    worst = max(_ulps(q(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2, worst
    assert q(0.0) == 1.0
    assert q(float("-inf")) == 0.0 and q(float("nan")) == 0.0 and q(-800.0) == 0.0


def test_schedule_closed_form(orc):
    """T(s) = t_hot (t_cold/t_hot)^(s/(A-1)) for s < A, then t_cold (ising_recon.hpp:129-131)."""
    s = orc.schedule(50.0, 1.0, 200, 50)
    f = orc.lib().orc_temperature
    import ctypes as C
    assert f(C.byref(s), 0) == 50.0
    assert f(C.byref(s), 199) == 50.0 * math.pow(1.0 / 50.0, 1.0)
    assert f(C.byref(s), 200) == 1.0 and f(C.byref(s), 249) == 1.0
    assert f(C.byref(s), 57) == 50.0 * math.pow(1.0 / 50.0, 57 / 199)


def test_sha256_fips_vectors(orc):
    """FIPS 180-4 example digests."""
    assert orc.sha256_hex(b"abc") == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    assert orc.sha256_hex(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"
    assert orc.sha256_hex(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == \
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"
    assert orc.sha256_hex(b"a" * 1000000) == \
        "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0"


# ---- float path: reference fixtures -------------------------------------------------

REE_SCHEDULE = (5.0, 1e-3, 400, 20)   # check_ree_qubo.cu:248


@pytest.mark.reference
def test_anneal_branch_reproduces_ree_cascade_runs(orc):
    """anneal_branch (reference trajectory) reproduces all 92 x 64 final energies of
    experiments/ree_cascade/engine_runs.csv (written by check_ree_qubo.cu:332-338)."""
    need_reference()
    runs = rf.engine_runs()
    sched = orc.schedule(*REE_SCHEDULE)
    bad, total = [], 0
    for iid in rf.instance_ids():
        nv, off, qd, pairs = rf.read_qubo(os.path.join(rf.REE, f"qubo_{iid}.txt"))
        prob = orc.FloatProblem(**rf.pair_ray_problem(nv, qd, pairs))
        _, br = orc.branch_fractions(prob, sched, 64, 8)
        for s in range(64):
            total += 1
            got = "%.12g" % rf.qubo_energy(off, qd, pairs, br[s])
            if got != runs[iid][s + 1]:
                bad.append((iid, s + 1, got, runs[iid][s + 1]))
    assert total == 5888
    assert not bad, bad[:5]


@pytest.mark.reference
@pytest.mark.parametrize("iid", ["MP_P507-A", "MP_P507-B"])
def test_anneal_branch_reproduces_energy_traces(orc, iid):
    """Per-sweep binary_energy + offset of 8 seeds x 420 sweeps equals
    energy_trace_<id>.csv (check_ree_qubo.cu:341-348)."""
    need_reference()
    nv, off, qd, pairs = rf.read_qubo(os.path.join(rf.REE, f"qubo_{iid}.txt"))
    prob = orc.FloatProblem(**rf.pair_ray_problem(nv, qd, pairs))
    sched = orc.schedule(*REE_SCHEDULE)
    want = rf.energy_trace(iid)
    for seed in range(1, 9):
        _, tr = orc.anneal_branch(prob, sched, seed, trace=True)
        got = ["%.10g" % (e + off) for e in tr]
        assert got == want[seed]


@pytest.mark.reference
def test_binary_energy_equals_direct_qubo(orc):
    """Pair-ray binary_energy equals the direct QUBO energy minus offset on random
    states, within the float rounding of a = float(sqrt|q|): the bound is the
    reference's own max_abs_diff over 1000 states (energy_check.csv, written by
    check_ree_qubo.cu:283-290), with a 2x allowance for our different states."""
    need_reference()
    with open(os.path.join(rf.REE, "energy_check.csv")) as f:
        bound = {r["id"]: float(r["max_abs_diff"]) for r in csv.DictReader(f)}
    rng = np.random.default_rng(11)
    for iid in ["MP_P507-A", "R8_3", "R12_9"]:
        nv, off, qd, pairs = rf.read_qubo(os.path.join(rf.REE, f"qubo_{iid}.txt"))
        prob = orc.FloatProblem(**rf.pair_ray_problem(nv, qd, pairs))
        worst = 0.0
        for _ in range(200):
            x = rng.integers(0, 2, nv).astype(np.uint8)
            worst = max(worst, abs(orc.binary_energy(prob, x) - (rf.qubo_energy(off, qd, pairs, x) - off)))
        assert worst <= 2.0 * bound[iid], (iid, worst, bound[iid])
        assert worst > 0.0   # the encoding is inexact; an exactly-zero diff would mean a dropped term


def sixteen_var_problem(orc):
    """The 16-variable oracle problem of check_ising.cu:119-147 (splitmix seed 7)."""
    import ctypes as C
    s = C.c_uint64(7)
    L = orc.lib()
    L.orc_uniform.restype = C.c_double
    L.orc_uniform.argtypes = [C.POINTER(C.c_uint64)]
    u = lambda: L.orc_uniform(C.byref(s))
    n, n_rays = 16, 12
    d, w = [], []
    for _ in range(n_rays):
        d.append(u() * 0.3)
        w.append(20.0 + 80.0 * u())
    row_ptr, ray, a = [0], [], []
    for _ in range(n):
        for b in range(n_rays):
            if u() < 0.35:
                ray.append(b)
                a.append(float(np.float32(0.02 + 0.1 * u())))
        row_ptr.append(len(ray))
    nbr_ptr, nbr = [0], []
    for i in range(n):
        x, y = i % 4, i // 4
        if x > 0: nbr.append(i - 1)
        if x < 3: nbr.append(i + 1)
        if y > 0: nbr.append(i - 4)
        if y < 3: nbr.append(i + 4)
        nbr_ptr.append(len(nbr))
    return orc.FloatProblem(row_ptr, ray, a, d, w, nbr_ptr, nbr, 0.7, 0.4)


def test_sixteen_variable_ground_state(orc):
    """check_ising.cu A: exhaustive 2^16 minimum is 7.404 (check_log.txt:157) and the
    MAP schedule reaches it within 8 seeds."""
    prob = sixteen_var_problem(orc)
    states = ((np.arange(1 << 16)[:, None] >> np.arange(16)) & 1).astype(np.uint8)
    exact = min(orc.binary_energy(prob, x) for x in states)
    assert "%.4g" % exact == "7.404"
    sched = orc.schedule(5.0, 1e-3, 400, 20)
    best = min(orc.binary_energy(prob, orc.anneal_branch(prob, sched, s)) for s in range(1, 9))
    assert abs(best - exact) < 1e-9


# ---- integer path -------------------------------------------------------------------

def random_complete(orc, n, wmax, seed):
    """random_complete of gset_bench.cu:252-264 / gpu_anneal.cu:51-63."""
    import ctypes as C
    s = C.c_uint64(seed)
    eu, ev, ew = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            w = int(orc.lib().orc_splitmix(C.byref(s)) % (2 * wmax + 1)) - wmax
            eu.append(i); ev.append(j); ew.append(w)
    return orc.IntGraph(n, eu, ev, ew)


def brute_max_cut(g):
    n = g.n
    states = ((np.arange(1 << n)[:, None] >> np.arange(n)) & 1).astype(np.int8)
    diff = states[:, g.eu] != states[:, g.ev]
    return int((diff * g.ew[None, :].astype(np.int64)).sum(1).max())


def test_k16_best_of_eight(orc):
    """gset_bench --check B: best-of-8 cuts 10622/12434/5050 against exact optima
    13777/12434/6376 (check_log.txt:211), Budget(5, 1e-3, 4000, 200)."""
    want_eng, want_opt = [10622, 12434, 5050], [13777, 12434, 6376]
    sched = orc.schedule(5.0, 1e-3, 4000, 200)
    for seed in (1, 2, 3):
        g = random_complete(orc, 16, 1000, seed)
        best = max(orc.lib().orc_cut_of(g.c, orc.anneal_int(g, sched, s)[0]) for s in range(1, 9))
        assert best == want_eng[seed - 1]
        assert brute_max_cut(g) == want_opt[seed - 1]


def test_energies_consistent_every_state_k16(orc):
    """cut_of (edge list) and energy_scratch (CSR) agree with a vectorised cut on all
    2^16 states of an integer-weighted K16 (the gray-code invariant of gset_bench.cu:288-309)."""
    g = random_complete(orc, 16, 1000, 2)
    states = ((np.arange(1 << 16)[:, None] >> np.arange(16)) & 1).astype(np.uint8)
    cut = ((states[:, g.eu] != states[:, g.ev]) * g.ew[None, :].astype(np.int64)).sum(1)
    for k in range(0, 1 << 16, 97):
        assert orc.lib().orc_cut_of(g.c, states[k]) == cut[k]
        assert orc.lib().orc_energy_scratch(g.c, states[k]) == -cut[k]


def parse(orc, text):
    import ctypes as C
    b = text.encode()
    n, m = C.c_int(), C.c_int64()
    cap = 64
    eu = np.zeros(cap, np.int32); ev = np.zeros(cap, np.int32); ew = np.zeros(cap, np.int32)
    err = C.create_string_buffer(256)
    rc = orc.lib().orc_parse_rudy(b, len(b), C.byref(n), C.byref(m), eu.ctypes.data, ev.ctypes.data,
                                  ew.ctypes.data, cap, err, 256)
    return rc, n.value, m.value, ew[:max(m.value, 0)]


def test_loader_refusals(orc):
    """gset_bench --check A: 5 malformed inputs refused; a missing weight is +1 and the
    sum of weights of '1 2 / 2 3 -7' is -6 (gset_bench.cu:313-334)."""
    bad = ["3 1\n1 2 0.5\n", "3 1\n1 1 1\n", "3 2\n1 2 1\n2 1 1\n", "3 2\n1 2 1\n", "3 1\n1 4 1\n"]
    assert all(parse(orc, t)[0] != 0 for t in bad)
    rc, n, m, ew = parse(orc, "3 2\n1 2\n2 3 -7\n")
    assert rc == 0 and m == 2 and int(ew.sum()) == -6
    R = orc.ref()
    if R is not None:
        import ctypes as C
        for t in bad:
            b = t.encode()
            assert R.ref_parse_rudy(b, len(b), C.byref(C.c_int()), C.byref(C.c_int64()), C.byref(C.c_int64())) != 0


def er_graph(orc, n=800, p=0.06, seed=1):
    """The max-cut workload generator: splitmix64 from `seed` over pairs i < j, keep
    the pair when u < p, weight sign from the next draw (+1 if its top bit is 0)."""
    import ctypes as C
    s = C.c_uint64(seed)
    L = orc.lib()
    eu, ev, ew = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            u = (L.orc_splitmix(C.byref(s)) >> 11) * 2.0 ** -53
            if u < p:
                eu.append(i); ev.append(j)
                ew.append(1 if (L.orc_splitmix(C.byref(s)) >> 63) == 0 else -1)
    return orc.IntGraph(n, eu, ev, ew)


def test_integer_anneal_equals_reference_code(orc):
    """The oracle's wising::anneal restatement equals the reference's own
    ising_weighted.hpp (compiled into oracle/_ref) seed for seed, bit for bit."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference shim not built")
    import ctypes as C
    cases = [(random_complete(orc, 16, 1000, 1), (5.0, 1e-3, 400, 20)),
             (er_graph(orc, 200, 0.1, 3), (3.0, 0.05, 300, 10))]
    for g, b in cases:
        sched = orc.schedule(*b)
        for seed in (1, 2, 7):
            x, e = orc.anneal_int(g, sched, seed)
            xr = np.zeros(g.n, np.uint8)
            er = C.c_int64()
            R.ref_anneal_int(g.n, g.m, g.eu, g.ev, g.ew, *b, seed, xr, C.byref(er))
            assert np.array_equal(x, xr) and e == er.value


def test_tts99_bootstrap_reproduces_summary(orc):
    """tts99_bootstrap over runs_gpu.csv / runs_cpu.csv success columns with
    t = wall / runs reproduces summary.csv's CIs (ising_weighted.hpp:279-303)."""
    need_reference()
    base = "/root/reference/experiments/p2_gpu"
    succ = {}
    for dev in ("gpu", "cpu"):
        with open(os.path.join(base, f"runs_{dev}.csv")) as f:
            for row in csv.DictReader(f):
                succ.setdefault((row["instance"], dev), []).append(int(row["success"]))
    import ctypes as C
    with open(os.path.join(base, "summary.csv")) as f:
        rows = list(csv.DictReader(f))
    for row in rows:
        s = np.array(succ[(row["instance"], row["device"])], np.int32)
        assert len(s) == int(row["runs"])
        t = np.full(len(s), float(row["wall_s"]) / len(s))
        lo, hi = C.c_double(), C.c_double()
        orc.lib().orc_tts99_bootstrap(s, t, len(s), 2000, 20260928, C.byref(lo), C.byref(hi))
        fmt = lambda v: "inf" if math.isinf(v) else "%.6f" % v
        assert fmt(lo.value) == row["tts99_ci_lo_s"] and fmt(hi.value) == row["tts99_ci_hi_s"]
