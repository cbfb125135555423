"""Gate Gd1 on the HIP path: each DNMR coupling case against the 0+1D referee
(test_oracle_dnmr.py), and each HIP run at CFLtau 0.005 equal to the oracle's
within 1e-9 in e, pi^eta_eta and Pi."""
import pytest

import test_oracle_dnmr as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_dnmr_couplings_on_gpu(oracle, hip, name):
    h = G.gate(hip, name)["r2"][2]
    o = G.run2d(oracle, G.CASES[name][0], 0.005)
    for k in ("e", "phi", "Pi"):
        assert abs(h[k] - o[k]) <= 1e-9 * max(abs(o[k]), 1e-30), (k, h[k], o[k])
