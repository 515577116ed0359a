"""The reference's own host drivers, compiled unchanged against the drop-in headers
(include/argos) and linked to this engine, run on the MI355X.

- tools/gpu_anneal.cu --check (:150-197): device energies against host cuts on three K16
  x 1000 states; run_batch against wising::anneal seed by seed on K16, a sparse +1 graph in
  shared and global plans, and a signed sparse graph.
- tools/gset_bench.cu --check (:311-378): the rudy loader's refusals; exact integer
  energies over all 2^16 states of three K16; wising::anneal equal to anneal_branch seed by
  seed on a +1 graph (anneal_branch in the reference's trajectory).
- tools/gpu_anneal.cu --bench (:215-385) on a G1-like graph: the six CSVs of its suite
  (runs_gpu, runs_cpu, summary, throughput, agreement, load), in its own schema, with
  every row consistent and the two engines identical on shared seeds.

The G-set files are not available, so the G1 lines of both self-checks report SKIP, as the
reference's do without data. The binaries are built from the reference tree by
`make -C tests/cpp drivers`; without them this file skips.
"""
import csv
import hashlib
import os
import subprocess

import numpy as np
import pytest

import qublar_amd as qa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.gpu


def driver(name):
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (make -C tests/cpp drivers needs the reference tree)")
    return path


def run(args, cwd=None, timeout=240):
    print("run", " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)), flush=True)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    return r


def check_lines(out):
    passed = [ln for ln in out.splitlines() if ln.rstrip().endswith("PASS") or " PASS " in ln]
    failed = [ln for ln in out.splitlines() if " FAIL " in ln or ln.rstrip().endswith("FAIL")]
    return passed, failed


def test_gpu_anneal_self_check():
    r = run([driver("gpu_anneal"), "--check", "no_gset_here"])
    passed, failed = check_lines(r.stdout)
    assert r.returncode == 0 and not failed, failed
    assert len(passed) == 5 and "all checks passed" in r.stdout
    assert "SKIP" in r.stdout   # G1 needs the G-set file


def test_gset_bench_self_check():
    r = run([driver("gset_bench"), "--check", "no_gset_here"])
    passed, failed = check_lines(r.stdout)
    assert r.returncode == 0 and not failed, failed
    assert len(passed) == 4 and "all checks passed" in r.stdout
    assert "integer engine = anneal_branch" in r.stdout


def write_rudy(path, g):
    with open(path, "w") as f:
        f.write(f"{g.n} {g.m}\n")
        for u, v, w in zip(g.eu, g.ev, g.ew):
            f.write(f"{int(u) + 1} {int(v) + 1} {int(w)}\n")


def read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def test_gpu_anneal_bench_suite_csvs(tmp_path):
    """The ADR-016 suite of the reference's harness on a G1-like graph (n 800, p 0.06, +1),
    16,384 replicas x (4000 + 200) sweeps: its CSVs, read back."""
    exe = driver("gpu_anneal")
    g = qa.er_graph(800, 0.06, 1)
    g = qa.GraphArrays(g.n, g.eu, g.ev, np.abs(g.ew))   # all weights +1, as G1's
    graph = tmp_path / "G1like.rudy"
    write_rudy(graph, g)
    sha = hashlib.sha256(graph.read_bytes()).hexdigest()
    target = 11400
    (tmp_path / "list.txt").write_text(f"G1like {graph} {target} {sha} {g.n} {g.m} {g.m} 0\n")
    # the harness joins outdir and file names with a backslash; "." keeps them in cwd
    r = run([exe, "--bench", "list.txt", ".", "16384", "8"], cwd=tmp_path)
    assert r.returncode == 0 and "done" in r.stdout
    f = lambda name: read_csv(tmp_path / (".\\" + name))
    load, summ, thr, agr = f("load.csv"), f("summary.csv"), f("throughput.csv"), f("agreement.csv")
    runs_gpu, runs_cpu = f("runs_gpu.csv"), f("runs_cpu.csv")
    assert len(load) == 1 and load[0]["sha_ok"] == "1" and load[0]["counts_ok"] == "1"
    assert int(load[0]["n"]) == 800 and int(load[0]["m"]) == g.m
    assert [row["device"] for row in summ] == ["gpu", "cpu"]
    gpu = summ[0]
    assert int(gpu["runs"]) == 16384 and gpu["all_consistent"] == "1" and summ[1]["all_consistent"] == "1"
    assert int(gpu["anneal_sweeps"]) == 4000 and int(gpu["hold_sweeps"]) == 200
    rate = float(gpu["rate_per_s"])
    print(f"harness: GPU rate {rate:.3e} spin updates/s, best cut {gpu['best_cut']}, "
          f"p_success {gpu['p_success']}, wall {gpu['wall_s']} s; runs on the other side {summ[1]['runs']}", flush=True)
    assert rate > 1e10
    assert int(gpu["best_cut"]) > target and float(gpu["p_success"]) > 0.5
    assert len(runs_gpu) == 16384 and all(row["consistent"] == "1" for row in runs_gpu)
    assert len(runs_cpu) == int(summ[1]["runs"]) and all(row["consistent"] == "1" for row in runs_cpu)
    shared = int(agr[0]["shared_seeds"])
    assert shared > 0 and int(agr[0]["identical"]) == shared and agr[0]["oracle_mismatches"] == "0"
    assert all(row["identical_to_cpu"] == "1" for row in runs_gpu[:shared])
    assert all(row["identical_to_cpu"] == "-1" for row in runs_gpu[shared:])
    # the per-seed hashes of the shared seeds are the same on both sides
    assert [row["assignment_sha256"] for row in runs_gpu[:shared]] == \
        [row["assignment_sha256"] for row in runs_cpu[:shared]]
    cuts = np.array([int(row["cut"]) for row in runs_gpu])
    assert int(gpu["best_cut"]) == cuts.max() and abs(float(gpu["mean_cut"]) - cuts.mean()) < 1e-3
    # throughput.csv's plan column holds commas of its own, unquoted: read its row from the end
    thr_rows = (tmp_path / ".\\throughput.csv").read_text().splitlines()
    assert len(thr) == 1 and len(thr_rows) == 2
    tail = thr_rows[1].split(",")[-10:]
    assert int(tail[-2]) == 16384 and int(tail[-1]) == int(summ[1]["runs"])
    assert abs(float(tail[1]) - rate) <= 1e-6 * rate   # rate_gpu_per_s, as in summary.csv


def test_gset_bench_suite_csvs(tmp_path):
    """The P1 suite of tools/gset_bench.cu --suite (:417-541) on the G1-like graph: cells B1,
    B10 and B100 (64, 64 and 16 runs of wising::anneal) and its runs, summary and load CSVs,
    read back. Through the drop-in headers every run is a one-replica device call."""
    exe = driver("gset_bench")
    g = qa.er_graph(800, 0.06, 1)
    g = qa.GraphArrays(g.n, g.eu, g.ev, np.abs(g.ew))
    graph = tmp_path / "G1like.rudy"
    write_rudy(graph, g)
    sha = hashlib.sha256(graph.read_bytes()).hexdigest()
    (tmp_path / "list.txt").write_text(f"G1like {graph} 11400 {sha} {g.n} {g.m} {g.m} 0\n")
    r = run(["stdbuf", "-oL", exe, "--suite", "list.txt", ".", "16", "60"], cwd=tmp_path)
    assert r.returncode == 0
    f = lambda name: read_csv(tmp_path / (".\\" + name))
    load, summ, runs = f("load.csv"), f("summary.csv"), f("runs.csv")
    assert len(load) == 1 and load[0]["sha_ok"] == "1" and load[0]["expect_ok"] == "1"
    assert [row["budget"] for row in summ] == ["B1", "B10", "B100"]
    planned = {"B1": 64, "B10": 64, "B100": 16}
    for row in summ:
        assert int(row["runs_done"]) == int(row["runs_planned"]) == planned[row["budget"]]
        assert row["all_consistent"] == "1"
    assert len(runs) == 144 and all(row["consistent"] == "1" for row in runs)
    # Budget(5, 1e-3, 4000, 200) on a graph whose best cuts are near 11,495
    b10 = [int(row["cut"]) for row in runs if row["budget"] == "B10"]
    assert max(b10) > 11000
    print("suite:", {row["budget"]: (row["best_cut"], row["p_success"], row["t_mean_s"]) for row in summ}, flush=True)
