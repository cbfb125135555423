"""Gate Gt (test/test_terms2d.jl) on the host mirror and the oracle: the per-term
switches of the charm sectors and the medium's vorticity coupling.

Gt1  the register lists every switch once, in the bit order of include/fivo2d.h,
     and the defaults are all on except the two vorticity couplings.
Gt5  a switch that cannot act (its sector off), an unknown name and a typo are
     refused; a dict spec is accepted.
Gt6  on a solve: the default terms given explicitly are bit-identical to no terms,
     and every first-moment switch moves the evolved current.
Gt7  presets, without()/with by ingredient and the sector rule resolve as the
     reference documents (src/terms.jl); on a solve, consistent_fm with the
     'homogeneous' preset reproduces consistent_fm = false bit for bit — the
     shipped first-moment row is the homogeneous-medium reduction.
Gt8  the medium vorticity coupling (shear_vorticity) is inert on a round fireball
     up to the stencil's O(h^2) vorticity, and acts (> 100x that) on a swirling one.
"""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import _abi


def test_register_gt1():
    assert tuple(t[0] for t in F.TERM_REGISTER) == _abi.TERM_NAMES
    names = _abi.TERM_NAMES
    d = F.resolve_terms()
    for k, n in enumerate(names):
        assert bool((d >> k) & 1) == (n not in ("m2_vorticity", "shear_vorticity")), n


def _on(mask, name):
    return bool((mask >> _abi.TERM_NAMES.index(name)) & 1)


def test_refusals_gt5():
    base = dict(eos=F.LatticeHRGEOS(), enable_diff=True, kappa_coeff=0.1163)
    with pytest.raises(ValueError, match="consistent_fm"):
        F.build_model_2d(**base, terms={"fm_inertial": False})
    with pytest.raises(ValueError, match="consistent_m2"):
        F.build_model_2d(**base, terms={"m2_vorticity": True})
    with pytest.raises(ValueError, match="unknown term"):
        F.build_model_2d(**base, terms={"fm_inertia": False})
    with pytest.raises(ValueError, match="enable_diff"):
        F.build_model_2d(eos=F.LatticeHRGEOS(), terms={"nu_gradalpha": False})
    m = F.build_model_2d(**base, consistent_fm=True, terms={"fm_inertial": False})
    assert not _on(m.terms, "fm_inertial") and _on(m.terms, "fm_gradT")


def test_presets_and_ingredients_gt7():
    h = F.resolve_terms("homogeneous")
    assert _on(h, "nu_gradalpha") and _on(h, "m2_nu_gradient") and _on(h, "m2_bg_Dalpha")
    for n in ("fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion", "fm_dlnh", "m2_projector", "m2_accel_nu",
              "m2_bg_gradu", "shear_vorticity"):
        assert not _on(h, n), n
    assert F.resolve_terms("full") == (1 << 18) - 1 and F.resolve_terms("none") == 0
    w = F.resolve_terms(F.without("acceleration"))
    assert not _on(w, "fm_inertial") and not _on(w, "m2_projector") and not _on(w, "m2_accel_nu")
    assert _on(w, "fm_gradT") and _on(w, "m2_bg_gradu")
    w2 = F.resolve_terms(F.without("inertia", from_="full"))
    assert not _on(w2, "fm_inertial") and _on(w2, "m2_vorticity") and _on(w2, "shear_vorticity")
    v = F.resolve_terms({"with": ("vorticity",)})
    assert _on(v, "m2_vorticity") and _on(v, "shear_vorticity") and _on(v, "fm_inertial")
    g = F.resolve_terms(F.without("velocity_gradient"))
    assert not _on(g, "fm_nu_gradu") and not _on(g, "fm_expansion") and not _on(g, "m2_pi_sigma")
    assert _on(g, "fm_inertial") and _on(g, "fm_gradT")
    with pytest.raises(ValueError, match="ingredient"):
        F.resolve_terms(F.without("vortcity"))
    with pytest.raises(ValueError, match="preset"):
        F.resolve_terms("homogenous")
    # sector rule: implicit changes in an off sector are reset, explicit ones refused
    base = dict(eos=F.LatticeHRGEOS(), enable_diff=True, kappa_coeff=0.1163)
    assert F.build_model_2d(**base, terms="homogeneous").terms == _abi.TERMS_DEFAULT
    with pytest.raises(ValueError, match="consistent_fm"):
        F.build_model_2d(**base, terms={"preset": "homogeneous", "fm_dlnh": False})


def _charge_run(lib, **kw):
    """The Gt6/Gt7 solve: 24^2 on [-6, 6], a T lump and an offset fugacity lump,
    tau 0.4 -> 0.9 with run_sim_2d!'s defaults (test/test_terms2d.jl:322-334)."""
    g = F.make_grid2d(24, 24, xmax=6.0, ymax=6.0)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_diff=True, kappa_coeff=0.1163, **kw)
    X, Y = np.meshgrid(g.xC, g.yC, indexing="ij")
    T = 0.15 + 0.30 * np.exp(-(X ** 2 / 1.3 + Y ** 2 / 0.8) / 6)
    a = -4.0 + 0.8 * np.exp(-((X + 1) ** 2 + Y ** 2) / 4)
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(0.4, T=T.ravel(), alpha=a.ravel())
    r = s.run_sim_2d(tau0=0.4, taufinal=0.9)
    assert r["ok"]
    return s.get_state(), m.layout


def test_default_terms_bit_identical_and_fm_switches_act_gt6(oracle):
    U0, L = _charge_run(oracle, consistent_fm=True)
    U1, _ = _charge_run(oracle, consistent_fm=True, terms=dict(zip(_abi.TERM_NAMES, [
        bool((_abi.TERMS_DEFAULT >> k) & 1) for k in range(18)])))
    assert np.array_equal(U0, U1)
    nu = lambda U: np.concatenate([U[:, L.idx("nux")], U[:, L.idx("nuy")]])
    ref = nu(U0)
    sc = np.max(np.abs(ref))
    for k in ("nu_gradalpha", "fm_gradT", "fm_inertial", "fm_nu_gradu", "fm_expansion", "fm_dlnh"):
        Uk, _ = _charge_run(oracle, consistent_fm=True, terms={k: False})
        d = np.max(np.abs(nu(Uk) - ref)) / sc
        print(f"{k:13s} off moves nu by {d:.3e} of max|nu|")
        assert d > 1e-6, k


def test_homogeneous_preset_is_the_shipped_row_gt7(oracle):
    Ush, _ = _charge_run(oracle)
    Uho, _ = _charge_run(oracle, consistent_fm=True, terms="homogeneous")
    assert np.array_equal(Ush, Uho)


def _swirl_run(lib, vort, swirl):
    g = F.make_grid2d(32, 32, xmax=6.0, ymax=6.0)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.2,
                         terms={"shear_vorticity": vort})
    X, Y = np.meshgrid(g.xC, g.yC, indexing="ij")
    r2 = X ** 2 + Y ** 2
    T = 0.15 + 0.30 * np.exp(-r2 / 8)
    ux = -swirl * Y * np.exp(-r2 / 8)
    uy = swirl * X * np.exp(-r2 / 8)
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(0.6, T=T.ravel(), alpha=np.full(g.Ntot, -4.0), ux=ux.ravel(), uy=uy.ravel())
    r = s.run_sim_2d(tau0=0.6, taufinal=1.6)
    assert r["ok"]
    U = s.get_state()
    L = m.layout
    return np.concatenate([U[:, L.idx(k)] for k in ("pixx", "pixy", "piyy")])


def test_shear_vorticity_inert_without_swirl_acts_with_it_gt8(oracle):
    p0, p1 = _swirl_run(oracle, False, 0.0), _swirl_run(oracle, True, 0.0)
    da = np.max(np.abs(p1 - p0)) / np.max(np.abs(p0))
    s0, s1 = _swirl_run(oracle, False, 0.3), _swirl_run(oracle, True, 0.3)
    db = np.max(np.abs(s1 - s0)) / np.max(np.abs(s0))
    print(f"no swirl: {da:.2e}; swirl: {db:.2e} ({db / da:.0f}x)")
    assert da < 1e-4 and db > 100 * da
