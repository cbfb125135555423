"""Gate Gd1 (test/test_dnmr2d.jl:73-105): the medium's second-order (DNMR) couplings
on 2-D Bjorken flow against the 0+1D DNMR ODEs (test/analytic_referees.jl:115-129):

    de/dtau   = -(e + P + Pi - phi)/tau,                    phi = -pi^eta_eta
    tau_pi dphi/dtau = -phi + (4/3) eta/tau - (delta_pipi + tau_pipi/3) tau_pi phi/tau
                       + (2/3) lambda_piPi tau_pi Pi/tau
    tau_Pi dPi/dtau  = -Pi - zeta/tau - delta_PiPi tau_Pi Pi/tau + lambda_Pipi tau_Pi phi/tau

A 12^2 transversely uniform LatticeHRG state (T0 0.45, alpha -20), tau 0.6 -> 3, at
CFLtau 0.02, 0.01, 0.005 (first order: the operator split); the Richardson
extrapolation 2 r(0.005) - r(0.01) must sit within 1e-3 of the referee in e, phi and
Pi, the order above 0.7, and a referee with the coupling's sign flipped must miss by
ten times more. One case per coupling: tau_pipi, delta_PiPi, lambda_piPi, lambda_Pipi.
"""
import functools
import math

import numpy as np
import pytest

import fivo2d as F
from oracle_unit import OracleUnit

T0, A0, TAU0, TAU1 = 0.45, -20.0, 0.6, 3.0
HBARC = 0.1973269804
SH = dict(enable_shear=True, eta_over_s=0.2, tauShear_coeff=0.2)
BU = dict(enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0)
CASES = {
    "tau_pipi": (dict(SH, deltaShear_factor=4 / 3, taupi_pi_factor=10 / 7), dict(dpp=4 / 3, tpp=10 / 7),
                 dict(dpp=4 / 3, tpp=-10 / 7), "phi"),
    "delta_PiPi": (dict(BU, deltaPi_factor=2 / 3), dict(dPP=2 / 3), dict(dPP=-2 / 3), "Pi"),
    "lambda_piPi": (dict(SH, **BU, deltaShear_factor=4 / 3, lambda_pi_Pi_factor=6 / 5), dict(dpp=4 / 3, lpP=6 / 5),
                    dict(dpp=4 / 3, lpP=-6 / 5), "phi"),
    "lambda_Pipi": (dict(SH, **BU, deltaShear_factor=4 / 3, lambda_Pi_pi_factor=1.2), dict(dpp=4 / 3, lPp=1.2),
                    dict(dpp=4 / 3, lPp=-1.2), "Pi"),
}


@functools.lru_cache(maxsize=None)
def _ou():
    from conftest import ORACLE_LIB
    return OracleUnit(F.Lib(ORACLE_LIB, "orc_"))


def run2d(lib, km, cfltau):
    g = F.make_grid2d(12, 12, xmax=6.0, ymax=6.0)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), **km)
    s = F.Solver(g, m, lib=lib)
    assert s.init_from_prims(TAU0, T=np.full(g.Ntot, T0), alpha=np.full(g.Ntot, A0)) == 0
    r = s.run_sim_2d(tau0=TAU0, taufinal=TAU1, CFL=10.0, CFLtau=cfltau)
    assert r["ok"]
    f = s.fields_2d()
    return dict(e=f["e"][5, 5], phi=-f["pieta"][5, 5] if "pieta" in f else 0.0,
                Pi=f["Pi"][5, 5] if "Pi" in f else 0.0)


@functools.lru_cache(maxsize=None)
def referee(shear, bulk, dpp=0.0, tpp=0.0, lpP=0.0, dPP=0.0, lPp=0.0, n=2000):
    ou, eos = _ou(), F.LatticeHRGEOS()

    def eP(T):
        P, _, e = ou.eos_Pne(eos, T, A0 * T)
        return e, P

    def T_of_e(et):
        lo, hi = 0.01, 2.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if eP(mid)[0] < et:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    def coeffs(T):
        P, nn, e = ou.eos_Pne(eos, T, A0 * T)
        s = (e + P - A0 * T * nn) / T                                   # src/eos.jl:306-310
        eta = SH["eta_over_s"] * max(s, 0.0) * HBARC if shear else 0.0  # src/primitives.jl:55-64
        tp = eta / (T * s * SH["tauShear_coeff"]) if shear else 0.0
        zeta = BU["zeta_over_s"] / (1 + ((T - 0.175) / 0.024) ** 2) * HBARC * max(s, 0.0) if bulk else 0.0
        if bulk:
            delta = max(abs(1 / 3 - ou.cs2(eos, T)), 1e-8)
            tb = zeta / (T * s * BU["tauPi_coeff"]) / delta ** 2 + 0.1  # src/primitives.jl:98-108
        else:
            tb = 1.0
        return P, eta, tp, zeta, tb

    def f(tau, y):
        e, phi, Pi = y
        P, eta, tp, zeta, tb = coeffs(T_of_e(e))
        de = -(e + P + Pi - phi) / tau
        dphi = (-phi + (4 / 3) * eta / tau - (dpp + tpp / 3) * tp * phi / tau + (2 / 3) * lpP * tp * Pi / tau) / tp \
            if tp > 0 else 0.0
        dPi = (-Pi - zeta / tau - dPP * tb * Pi / tau + lPp * tb * phi / tau) / tb if (tb > 0 and zeta > 0) else 0.0
        return np.array([de, dphi, dPi])

    y = np.array([eP(T0)[0], 0.0, 0.0])
    h = (TAU1 - TAU0) / n
    tau = TAU0
    for _ in range(n):
        k1 = f(tau, y)
        k2 = f(tau + h / 2, y + h / 2 * k1)
        k3 = f(tau + h / 2, y + h / 2 * k2)
        k4 = f(tau + h, y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        tau += h
    return dict(e=y[0], phi=y[1], Pi=y[2])


def gate(lib, name):
    km, kr, kbad, key = CASES[name]
    shear, bulk = km.get("enable_shear", False), km.get("enable_bulk", False)
    ref = referee(shear, bulk, **kr)
    bad = referee(shear, bulk, **kbad)
    r2 = [run2d(lib, km, c) for c in (0.02, 0.01, 0.005)]
    on = {"e": True, "phi": shear, "Pi": bulk}
    ext = {k: 2 * r2[2][k] - r2[1][k] for k in on}
    rel = lambda x, y, o: x / y - 1 if o else 0.0
    eext = [rel(ext[k], ref[k], on[k]) for k in ("e", "phi", "Pi")]
    ebad = ext[key] / bad[key] - 1
    p = math.log2(max(abs(rel(r2[1][k], ref[k], on[k])) for k in on) /
                  max(abs(rel(r2[2][k], ref[k], on[k])) for k in on))
    print(f"{name:12s} extrapolated (e, phi, Pi) {eext[0]:+.1e} {eext[1]:+.1e} {eext[2]:+.1e} | order {p:.2f} | "
          f"wrong-sign referee {ebad:+.1e}", flush=True)
    assert max(abs(x) for x in eext) < 1e-3
    assert p > 0.7
    assert abs(ebad) > 10 * abs(eext[1] if key == "phi" else eext[2])
    return dict(r2=r2, err=max(abs(x) for x in eext), order=p, bad=abs(ebad))


def test_referee_converged():
    a = referee(True, True, dpp=4 / 3, lpP=6 / 5, n=1000)
    b = referee(True, True, dpp=4 / 3, lpP=6 / 5, n=2000)
    assert all(abs(a[k] - b[k]) <= 1e-7 * max(abs(b[k]), 1e-30) for k in ("e", "phi", "Pi"))


RESULTS = {}


@pytest.mark.parametrize("name", sorted(CASES))
def test_dnmr_couplings_gd1(oracle, name):
    RESULTS[name] = gate(oracle, name)


def test_dnmr_published_ranges():
    """README2D.md:283: extrapolated to <= 3e-4 at order 0.97-0.99; wrong-sign
    versions miss by 5e-3 to 1.7."""
    if len(RESULTS) < len(CASES):
        pytest.skip("needs every coupling case of this module")
    assert f"{max(r['err'] for r in RESULTS.values()):.0e}" == "3e-04"
    assert all(0.965 <= r["order"] < 0.995 for r in RESULTS.values())
    bad = sorted(r["bad"] for r in RESULTS.values())
    assert 5e-3 <= bad[0] < 6e-3 and f"{bad[-1]:.1f}" == "1.7"
