"""Term switches on the HIP path: the Gt6/Gt7/Gt8 solves of test_oracle_terms.py
run by the library, each against the oracle's run (final fields within 1e-9 of
their scale), the homogeneous-preset identity bit for bit, and fivo_create's
refusal of a switch whose sector is off."""
import ctypes as C

import numpy as np
import pytest

import fivo2d as F
from fivo2d import _abi
import test_oracle_terms as G

pytestmark = pytest.mark.gpu


def _close(Uh, Uo, L, rows):
    for k in rows:
        a, b = Uh[:, L.idx(k)], Uo[:, L.idx(k)]
        d = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        assert d <= 1e-9, (k, d)


@pytest.mark.parametrize("terms", [None, {"fm_inertial": False}, "homogeneous", F.without("expansion")])
def test_charge_solves_match_oracle(oracle, hip, terms):
    kw = dict(consistent_fm=True) if terms is None else dict(consistent_fm=True, terms=terms)
    Uh, L = G._charge_run(hip, **kw)
    Uo, _ = G._charge_run(oracle, **kw)
    _close(Uh, Uo, L, ("Dtau", "E", "nux", "nuy"))


def test_homogeneous_preset_is_the_shipped_row_on_gpu(hip):
    Ush, _ = G._charge_run(hip)
    Uho, _ = G._charge_run(hip, consistent_fm=True, terms="homogeneous")
    assert np.array_equal(Ush, Uho)


def test_shear_vorticity_on_gpu(oracle, hip):
    for vort in (False, True):
        ph, po = G._swirl_run(hip, vort, 0.3), G._swirl_run(oracle, vort, 0.3)
        assert np.max(np.abs(ph - po)) <= 1e-9 * np.max(np.abs(po))


def test_create_refuses_switch_in_disabled_sector(hip):
    g = F.make_grid2d(16, 16)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_diff=True, kappa_coeff=0.1163)
    md = m.desc()
    md.terms = _abi.TERMS_DEFAULT & ~(1 << _abi.TERM_NAMES.index("fm_inertial"))
    gd = g.desc()
    out = C.c_void_p()
    rc = hip.create(C.byref(gd), C.byref(md), C.byref(out))
    assert rc < 0 and not out.value
    msg = hip.last_error(None)
    assert "fm_inertial" in msg and "consistent_fm" in msg, msg
