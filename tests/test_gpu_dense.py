"""GPU parity of the dense field engine, called through the C ABI, against the oracle's
sequential trajectory: spins bit-exact per replica in FP64 fields and in the FP32 carry
(with and without the FP64 recompute), FP64 energies within 1e-9 relative, and the
reported final field error equal to the oracle's."""
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# Up to 64 replicas (and n >= 256) each replica is annealed by four workgroups sharing its
# fields; 66 replicas take one workgroup each. Both give the oracle's spins, also when the
# last window is short and odd (301, 259).
@pytest.mark.parametrize("n,R,prec,k", [(300, 6, qa.FP64, 0), (300, 6, qa.FP32, 3), (300, 6, qa.FP32, 0),
                                        (1024, 4, qa.FP32, 10), (300, 66, qa.FP32, 3), (200, 3, qa.FP64, 0),
                                        (301, 5, qa.FP32, 2), (259, 66, qa.FP64, 0), (520, 64, qa.FP32, 5),
                                        (1000, 7, qa.FP64, 0)])
def test_dense_equals_oracle(n, R, prec, k):
    Q, h = qa.dense_qubo(n, 7)
    D = qa.DenseQubo(Q, h)
    sched = qa.schedule(1.0, 0.02, 20, 4)
    say(f"dense n={n} R={R} precision={prec} k={k}")
    out = D.anneal(sched, R, seed_offset=2, precision=prec, recompute_every=k)
    osched = orc_mod.schedule(1.0, 0.02, 20, 4)
    for b in range(R):
        ref, f = orc_mod.dense_anneal(Q, h, osched, b + 1 + 2, prec, k, fields=True)
        assert np.array_equal(out["x"][b], ref), f"replica {b}: {(out['x'][b] != ref).sum()} spins differ"
        e = orc_mod.dense_energy(Q, h, ref)
        assert abs(out["energy"][b] - e) <= 1e-9 * max(1.0, abs(e))
        scratch = h.astype(np.float64) + Q.astype(np.float64).T @ ref.astype(np.float64)
        assert abs(out["field_err"][b] - np.max(np.abs(f - scratch))) < 1e-9
        assert 0 < ref.sum() < n
