"""The oracle pinned against mathematics and the reference's own unit gates
(CPU only): Bessel K against 30-digit mpmath golden vectors, the EOS against
its defining identities, and the primitive recovery round trip of
test/test_primrec2d.jl:79-191."""
import os

import numpy as np
import pytest

import fivo2d as F
from oracle_unit import OracleUnit

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "besselkx.npz")


@pytest.fixture(scope="module")
def ou(oracle):
    return OracleUnit(oracle)


def test_besselkx_against_mpmath(ou):
    g = np.load(GOLD)
    xs, kx = g["x"], g["kx"]
    for nu in range(6):
        got = np.array([ou.besselkx(nu, x) for x in xs])
        rel = np.abs(got / kx[nu] - 1)
        # K0, K1 are minimax rationals accurate to a few ulp; the upward recurrence adds
        # a few ulp per order; the x >= 1e4 branch is the reference's 3-term series
        assert np.max(rel) < 2e-14 * (1 + nu), (nu, np.max(rel), xs[np.argmax(rel)])


def test_fast_k2_against_mpmath(ou):
    """src2d/bessel2d.jl's bound: fast K2 within 5e-15 of the exact value on [0.1, 1e4)."""
    g = np.load(GOLD)
    xs, kx = g["x"], g["kx"]
    sel = (xs >= 0.1) & (xs < 1e4)
    got = np.array([ou.besselk2x_fast(x) for x in xs[sel]])
    assert np.max(np.abs(got / kx[2][sel] - 1)) < 5e-15
    big = xs[xs >= 1e4]
    assert all(ou.besselk2x_fast(x) == ou.besselkx(2, x) for x in big)   # shared series


def test_besselkx_guards(ou):
    assert ou.besselkx(2, 0.0) == 0.0 and ou.besselkx(2, -1.0) == 0.0
    assert ou.besselkx(2, float("inf")) == 0.0 and ou.besselkx(2, float("nan")) == 0.0


def test_lattice_eos_identities(ou):
    eos = F.LatticeHRGEOS()
    for T in np.linspace(0.08, 0.8, 37):
        P, dP, d2P = ou.light(eos, T)
        h = 1e-5 * T
        Pp, dPp, _ = ou.light(eos, T + h)
        Pm, dPm, _ = ou.light(eos, T - h)
        assert abs((Pp - Pm) / (2 * h) / dP - 1) < 1e-8          # P' is the derivative of P
        assert abs((dPp - dPm) / (2 * h) / d2P - 1) < 1e-7       # P'' of P'
        Pt, n, e = ou.eos_Pne(eos, T, 0.0)
        assert Pt == P
        assert abs(e - (T * dP - P)) <= 1e-14 * abs(e)           # e = T P' - P
        assert abs(ou.cs2(eos, T) - dP / (T * d2P)) <= 1e-15
    # the fast-K2 charm density agrees with the AMOS-path one to round-off
    for T in (0.12, 0.2, 0.35, 0.6):
        for mu in (-3.0, 0.0, 1.0):
            a = ou.eos_Pne(eos, T, mu)
            b = ou.eos_Pne(eos, T, mu, two_d=True)
            assert a[0] == b[0] and a[2] == b[2]
            assert abs(a[1] - b[1]) <= 1e-14 * abs(a[1])


def test_conformal_eos(ou):
    eos0 = F.ConformalHQEOS(g_hq=0.0)
    for T in (0.1, 0.3, 0.6):
        P, n, e = ou.eos_Pne(eos0, T, 1.5)
        assert e == 3 * P and n == 0.0
        assert abs(P - (np.pi ** 2 / 90) * 40 * T ** 4 / 0.1973269804 ** 3) <= 1e-14 * P
    # the G1 EOS with its charm tracer: conformal to < 1e-6 at alpha = -20
    eos = F.ConformalHQEOS()
    worst = 0.0
    for T in np.linspace(0.05, 0.6, 23):
        P, n, e = ou.eos_Pne(eos, T, -20 * T)
        worst = max(worst, abs(e / (3 * P) - 1))
    assert worst < 1e-6


def _random_state(rng, eos, umax, diss):
    T = 0.12 + (0.55 - 0.12) * rng.random()
    phi = 4.0 * (2 * rng.random() - 1)
    mu = eos.m_hq + T * phi
    ux, uy = umax * (2 * rng.random() - 1), umax * (2 * rng.random() - 1)
    return T, mu, ux, uy, diss


@pytest.mark.parametrize("eosname", ["lattice", "conformal"])
@pytest.mark.parametrize("diss", [0.0, 0.1, 0.3])
def test_primrec_round_trip(ou, eosname, diss):
    """test/test_primrec2d.jl:79-118: prim -> cons -> prim, 0 failures, T/u/mu < 1e-9."""
    eos = F.LatticeHRGEOS() if eosname == "lattice" else F.ConformalHQEOS()
    rng = np.random.default_rng(20260901 + int(diss * 10))
    tau = 0.7
    worst = 0.0
    for _ in range(400):
        T, mu, ux, uy, _d = _random_state(rng, eos, 1.5, diss)
        P, n, _ = ou.eos_Pne(eos, T, mu, two_d=True)
        Pi, pixx, pixy, piyy = (diss * P * (2 * rng.random() - 1) for _ in range(4))
        ut = np.sqrt(1 + ux * ux + uy * uy)
        tx = (ux * pixx + uy * pixy) / ut
        ty = (ux * pixy + uy * piyy) / ut
        pieta = (ux * tx + uy * ty) / ut - pixx - piyy
        nux, nuy = (diss * n * (2 * rng.random() - 1) for _ in range(2))
        d7 = (nux, nuy, Pi, pixx, pixy, piyy, pieta)
        ok, cons = ou.prim_to_cons(eos, T, mu, ux, uy, d7, tau)
        assert ok
        r = ou.cons_to_prim(eos, cons[0] / tau, cons[1], cons[2], cons[3], d7, tau)
        assert r["ok"] == 1.0, r
        worst = max(worst, abs(r["T"] / T - 1), abs(r["ux"] - ux), abs(r["uy"] - uy),
                    abs(r["mu"] - mu) / max(abs(mu), 1.0))
    assert worst < 1e-9


def test_primrec_uy_zero_and_vacuum(ou):
    eos = F.LatticeHRGEOS()
    ok, cons = ou.prim_to_cons(eos, 0.3, 1.5 - 0.3 * 2, 0.8, 0.0, (0,) * 7, 1.0)
    r = ou.cons_to_prim(eos, cons[0], cons[1], cons[2], cons[3], (0,) * 7, 1.0)
    assert r["ok"] and r["uy"] == 0.0 and abs(r["ux"] - 0.8) < 1e-12
    v = ou.cons_to_prim(eos, 0.0, 0.0, 0.0, 1e-30, (0,) * 7, 1.0)
    assert v["ok"] == 0.0 and int(v["reason"]) == 1          # PRR_VACUUM
    # zero-flow cell
    ok, cons = ou.prim_to_cons(eos, 0.25, 1.5 - 0.25 * 3, 0.0, 0.0, (0,) * 7, 1.0)
    r = ou.cons_to_prim(eos, cons[0], cons[1], cons[2], cons[3], (0,) * 7, 1.0)
    assert r["ux"] == 0.0 and r["uy"] == 0.0 and abs(r["T"] / 0.25 - 1) < 1e-12
