"""GPU parity of the exact region-of-interest posterior (qb_roi_enumerate) against the
oracle's brute force, and at sizes beyond it against closed forms. The device carries
energies along Gray-code chunks of 4,096 states (~1e-12 from a scratch sum) and sums
chunk by chunk, then over blocks; the oracle sums 2^n terms with Kahan compensation. So
e_min agrees to 1e-11 relative, log Z to 1e-10 relative and marginals to 1e-10 absolute
(with a plain running sum over 2^25 terms the oracle alone could be 1e-9 off); the
minimum-energy state is the oracle's."""
import sys

import numpy as np
import pytest

import oracle as orc_mod
import qublar_amd as qa

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_roi(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    h = rng.normal(0, scale, n)
    J = np.triu(rng.normal(0, scale, (n, n)), 1) + np.tril(rng.normal(0, 5, (n, n)))   # lower part ignored
    return h, J


@pytest.mark.parametrize("n,beta", [(1, 1.0), (5, 0.7), (13, 1.0), (14, 2.0), (15, 1.0), (16, 1.0), (17, 0.5), (25, 1.0),
                                    (20, 1.0), (22, 3.0)])
def test_roi_equals_oracle(n, beta):
    h, J = random_roi(n, 100 + n)
    say(f"roi n={n} beta={beta}")
    g = qa.roi_enumerate(h, J, beta)
    o = orc_mod.roi_enumerate(h, J, beta)
    assert abs(g["e_min"] - o["e_min"]) <= 1e-11 * max(1.0, abs(o["e_min"]))
    assert abs(g["log_z"] - o["log_z"]) <= 1e-10 * max(1.0, abs(o["log_z"]))
    assert np.allclose(g["marginal"], o["marginal"], rtol=0, atol=1e-10)
    assert g["map_index"] == o["map_index"]
    assert np.array_equal(g["map_state"], (o["map_index"] >> np.arange(n)) & 1)
    assert abs(g["p_map"] - np.exp(-beta * o["e_min"] - o["log_z"])) < 1e-10
    assert g["states"] == 2.0 ** n


def test_roi_large_closed_forms():
    """n = 30 and 34 independent bits (J = 0): logistic marginals and a softplus log Z;
    beta = 0 gives 1/2 and n log 2."""
    rng = np.random.default_rng(5)
    for n in (30, 34):
        h = rng.normal(0, 2, n)
        say(f"roi independent n={n}")
        g = qa.roi_enumerate(h, np.zeros((n, n)), 1.0)
        assert np.allclose(g["marginal"], 1.0 / (1.0 + np.exp(h)), rtol=0, atol=1e-10)
        assert abs(g["log_z"] - np.logaddexp(0.0, -h).sum()) < 1e-9
        assert abs(g["e_min"] - np.minimum(h, 0).sum()) < 1e-10
        assert np.array_equal(g["map_state"], (h < 0).astype(np.uint8))
        say(f"  {g['kernel_ms']:.1f} ms for {g['states']:.3g} states")
    h, J = random_roi(28, 7)
    g = qa.roi_enumerate(h, J, 0.0)
    assert np.allclose(g["marginal"], 0.5, atol=1e-12) and abs(g["log_z"] - 28 * np.log(2)) < 1e-9
