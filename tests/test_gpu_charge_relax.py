"""Gate Gc7 on the HIP path: the uniform nu^x of a Bjorken state decays by the
derived-sign first-moment ODE on the run's own T history (test_oracle_charge_relax.py),
and the HIP nu(tau) history equals the oracle's within 1e-9 of nu_0."""
import numpy as np
import pytest

import test_oracle_charge_relax as R

pytestmark = pytest.mark.gpu


def test_first_moment_sign_on_gpu(oracle, hip):
    th = R.gc7_run(hip, oracle)
    nh = R.gc7_check(*th, oracle)
    to = R.gc7_run(oracle, oracle)
    assert np.array_equal(th[0], to[0])
    assert np.max(np.abs(nh - to[1])) <= 1e-9 * abs(to[1][0])
