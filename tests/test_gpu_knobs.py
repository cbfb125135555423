"""The knob matrix of test_oracle_knobs.py on the HIP path: for each knob set away
from its default, the same steps and counters as the oracle, final T and u within
1e-9 of the oracle's and every state row within 1e-7 of its scale."""
import numpy as np
import pytest

from helpers import rel_diff
from test_oracle_knobs import KNOBS, knob_run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(KNOBS))
def test_knob_parity(oracle, hip, name):
    rh, Uh, fh, L = knob_run(hip, name)
    ro, Uo, fo, _ = knob_run(oracle, name)
    for k in ("nsteps", "nhalvings", "nprimfail", "nvacuum", "tau"):
        assert rh[k] == ro[k], (name, k)
    dT = np.max(np.abs(fh["T"] - fo["T"])) / np.max(fo["T"])
    du = max(np.max(np.abs(fh[k] - fo[k])) for k in ("ux", "uy")) / max(np.max(np.abs(fo["ux"])), 1e-300)
    r = rel_diff(Uh, Uo)
    print(f"{name}: dT {dT:.1e} du {du:.1e} worst row {L.names[int(np.argmax(r))]} {np.max(r):.1e}", flush=True)
    assert dT <= 1e-9 and du <= 1e-9 and np.max(r) <= 1e-7, name
