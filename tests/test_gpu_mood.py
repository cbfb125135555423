"""MOOD on the HIP path: the sharp lumpy event of test_oracle_mood.py retries stages
with first-order faces on the dilated mask; the HIP run must retry the same
stages and end where the oracle ends (final T and u within 1e-9)."""
import numpy as np
import pytest

from test_oracle_mood import mood_run

pytestmark = pytest.mark.gpu


def test_mood_retries_match_the_oracle(oracle, hip):
    rh, fh = mood_run(hip)
    ro, fo = mood_run(oracle)
    print(f"HIP: steps {rh['nsteps']} MOOD {rh['nmood']} halvings {rh['nhalvings']} failures {rh['nprimfail']}; "
          f"oracle: {ro['nsteps']} {ro['nmood']} {ro['nhalvings']} {ro['nprimfail']}", flush=True)
    assert rh["ok"] and rh["nmood"] >= 1
    for k in ("nsteps", "nmood", "nhalvings", "nprimfail", "tau"):
        assert rh[k] == ro[k], k
    for k in ("T", "ux", "uy"):
        d = np.max(np.abs(fh[k] - fo[k])) / np.max(np.abs(fo[k]))
        print(f"  {k}: {d:.2e}", flush=True)
        assert d <= 1e-9, (k, d)
