"""Gate Gk (test/test_charge_dispersion2d.jl): the charge dispersion relation.

A uniform LatticeHRG background at rest (T0 0.35, alpha0 -2, tau0 3200 so the
Bjorken drift is frozen), periodic 12 fm x 1 fm box of 192 x 16 cells, charm
diffusion only. The charm is a tracer (e, P independent of alpha; d ln n / d alpha
= 1), so one Fourier mode of phi = dn/n and nu^x is closed:
    tau_n w^2 + i w - D_s k^2 = 0,   D_s = kappa/n = (D_sT/T) hbar c,
with the collision of the two roots at k* = 1 / (2 sqrt(D_s tau_n)). The exact
eigenmode of the slower root is seeded (alpha = alpha0 + A cos kx, nu^x =
Re[-i D_s n k A e^{ikx} / (1 - i w tau_n)]), and the complex Fourier amplitude of the
conserved charge row along one y-row gives Im w from its envelope and Re w from its
phase. tau_n comes from the solver's own coefficient (the oracle's probe
orc_diff_tauN), tau_N_coeff derived to put k* at a chosen multiple of the box's
fundamental: 2.5 k_1 (both tested harmonics overdamped) or 1.4 k_1 (= 0.733 1/fm,
the published collision, README2D.md:282).

Here on the oracle: Gk-0 (the setup) and Gk-a (the overdamped branch: rate within 2 %
of the slow root and of Fick times the IS excess 2(1 - sqrt(1 - x))/x, x = (k/k*)^2,
no oscillation, half-window split < 5 %, T drift < 5e-3, y spread < 1e-12). The GPU
test runs Gk-a to Gk-e and equality with the oracle.
"""
import cmath
import ctypes as C
import functools
import math

import numpy as np
import pytest

import fivo2d as F
from fivo2d import _abi

T0, A0, TAU0, AMP, LBOX, NX, NY = 0.35, -2.0, 3200.0, 1e-3, 12.0, 192, 16
HBARC = 0.1973269804
KFUND = 2 * math.pi / LBOX
C_NUM = 0.20       # the numerical-diffusion band of Gk-b (test/test_charge_dispersion2d.jl:304)


@functools.lru_cache(maxsize=None)
def _probe():
    from conftest import ORACLE_LIB
    d = F.Lib(ORACLE_LIB, "orc_").dll
    d.orc_diff_tauN.restype = C.c_double
    d.orc_diff_tauN.argtypes = [C.POINTER(_abi.EosDesc), C.c_double, C.c_double, C.c_double, C.c_double]
    return d


def tau_n(DsT, coeff, T=T0, alpha=A0):
    desc = F.LatticeHRGEOS().desc()
    return _probe().orc_diff_tauN(C.byref(desc), T, alpha, DsT, coeff)


def D_s(DsT, T=T0):
    return (DsT / T) * HBARC


def modes(k, D, tn):
    """The two roots, the slower (less damped; on the propagating branch Re w >= 0)
    first (test/test_charge_dispersion2d.jl:131-140)."""
    if tn <= 1e-14:
        return -1j * D * k * k, complex(math.inf, -math.inf)
    disc = cmath.sqrt(complex(4 * tn * D * k * k - 1.0))
    w1, w2 = (-1j + disc) / (2 * tn), (-1j - disc) / (2 * tn)
    if abs(w1.imag - w2.imag) < 1e-12 * max(abs(w1.imag), 1.0):
        return (w1, w2) if w1.real >= w2.real else (w2, w1)
    return (w1, w2) if w1.imag > w2.imag else (w2, w1)


def kstar(D, tn):
    return 1 / (2 * math.sqrt(D * tn))


def tauD_for_kstar(k_target, DsT=0.75):
    return 1 / (4 * D_s(DsT) * k_target ** 2) / tau_n(DsT, 1.0)


def fick_excess(x):
    return math.nan if x >= 1 else 2 * (1 - math.sqrt(1 - x)) / x


def _fit(x, y):
    x, y = np.asarray(x), np.asarray(y)
    mx, my = x.mean(), y.mean()
    return float(((x - mx) * (y - my)).sum() / ((x - mx) ** 2).sum())


def run_mode(lib, harmonic=1, DsT=0.75, tauD=1.0, trun=20.0, dts=0.25, CFL=0.15, N=NX, vacuum=True):
    """test/test_charge_dispersion2d.jl:195-285."""
    dx = LBOX / N
    g = F.make_grid2d(N, NY, xmax=LBOX / 2, ymax=NY * dx / 2)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_diff=True, kappa_coeff=DsT, tauN_coeff=tauD,
                         deltaN_factor=0.0, vacuum_n_lo=1e-6 if vacuum else 0.0,
                         vacuum_n_hi=2e-3 if vacuum else 0.0)
    from oracle_unit import OracleUnit
    from conftest import ORACLE_LIB
    n0 = OracleUnit(F.Lib(ORACLE_LIB, "orc_")).eos_Pne(F.LatticeHRGEOS(), T0, A0 * T0)[1]
    D, tn = D_s(DsT), tau_n(DsT, tauD)
    k = 2 * math.pi * harmonic / LBOX
    w = modes(k, D, tn)[0]
    nuc = -1j * D * n0 * k * AMP / (1 - 1j * w * tn)
    X, _ = np.meshgrid(g.xC, g.yC, indexing="ij")
    ph = np.exp(1j * k * X)
    s = F.Solver(g, m, lib=lib)
    nb = s.init_from_prims(TAU0, T=np.full(g.Ntot, T0), alpha=(A0 + AMP * ph.real).ravel(),
                           nux=(nuc * ph).real.ravel(), nuy=np.zeros(g.Ntot), bc=F.FIVO_BC_PERIODIC)
    assert nb == 0
    L = m.layout
    ng = g.nghost
    iy0 = ng + NY // 2
    xs = g.xC[ng:ng + g.Nx]
    phase = np.exp(-1j * k * xs)

    def amp():
        d = s.get_state()[:, L.idx("Dtau")].reshape(g.Nxtot, g.Nytot)[ng:ng + g.Nx, iy0]
        b = d.sum()
        return 2 * (d * phase).sum() / (b / g.Nx) / g.Nx, b / g.Nx

    def Tcell():
        return float(np.exp(s.get_work(("yT",))["yT"][g.lin(ng, iy0)]))

    tau = TAU0
    s.update_primitives_2d(tau, bc=F.FIVO_BC_PERIODIC)
    c0, b0 = amp()
    ts, la, phs, Ts = [0.0], [math.log(abs(c0))], [cmath.phase(c0)], [Tcell()]
    while tau < TAU0 + trun - 1e-9:
        r = s.run_sim_2d(tau0=tau, taufinal=min(tau + dts, TAU0 + trun), CFL=CFL, CFLtau=0.05,
                         bc=F.FIVO_BC_PERIODIC, reset_history=False)
        assert r["ok"], r
        tau = r["tau"]
        s.update_primitives_2d(tau, bc=F.FIVO_BC_PERIODIC)
        c, _ = amp()
        ts.append(tau - TAU0); la.append(math.log(abs(c))); phs.append(cmath.phase(c)); Ts.append(Tcell())
    _, b1 = amp()
    Dt = s.get_state()[:, L.idx("Dtau")].reshape(g.Nxtot, g.Nytot)[ng:ng + g.Nx, ng:ng + g.Ny]
    ysprd = float(np.max(np.abs(Dt - Dt[:, :1]) / np.maximum(np.abs(Dt[:, :1]), 1e-300)))
    s.close()
    for j in range(1, len(phs)):
        while phs[j] - phs[j - 1] > math.pi:
            phs[j] -= 2 * math.pi
        while phs[j] - phs[j - 1] < -math.pi:
            phs[j] += 2 * math.pi
    nn = len(ts)
    h = nn // 2
    wI = _fit(ts, la)
    wR = -_fit(ts, phs)
    wI1, wI2 = _fit(ts[:h], la[:h]), _fit(ts[h - 1:], la[h - 1:])
    return dict(k=k, D=D, tn=tn, n0=n0, dx=dx, w_meas=complex(wR, wI), w_ref=w, w_fast=modes(k, D, tn)[1],
                kstar=kstar(D, tn), split=abs(wI1 - wI2) / max(abs(wI), 1e-30), Tdrift=(Ts[-1] - Ts[0]) / Ts[0],
                ndrift=(b1 - b0) / b0, nsamp=nn, ysprd=ysprd, la=la, phs=phs)


def gk_a(lib):
    """test/test_charge_dispersion2d.jl:375-404."""
    tauD = tauD_for_kstar(2.5 * KFUND)
    out = []
    for harm in (1, 2):
        r = run_mode(lib, harmonic=harm, DsT=0.75, tauD=tauD, trun=18.0)
        g, gr = -r["w_meas"].imag, -r["w_ref"].imag
        xk = (r["k"] / r["kstar"]) ** 2
        fick = g / (r["D"] * r["k"] ** 2) / fick_excess(xk)
        print(f"    m={harm} k={r['k']:.4f} (k/k*={r['k'] / r['kstar']:.3f}) | gamma meas {g:.6f} ref {gr:.6f} "
              f"ratio {g / gr:.5f} | vs Fick x IS {fick:.5f} | split {r['split']:.4f} | Re w {r['w_meas'].real:.2e}",
              flush=True)
        assert abs(r["w_meas"].real) < 0.02 * g and r["split"] < 0.05 and r["Tdrift"] < 5e-3 and r["ysprd"] < 1e-12
        assert 0.98 < g / gr < 1.02 and 0.98 < fick < 1.02
        out.append(r)
    return out


def test_gk0_setup():
    from oracle_unit import OracleUnit
    from conftest import ORACLE_LIB
    ou, eos = OracleUnit(F.Lib(ORACLE_LIB, "orc_")), F.LatticeHRGEOS()
    P1, n1, e1 = ou.eos_Pne(eos, T0, -6.0 * T0)
    P2, n2, e2 = ou.eos_Pne(eos, T0, +1.0 * T0)
    assert abs(e2 - e1) / e1 < 1e-14 and abs(P2 - P1) / P1 < 1e-14      # the tracer does not back-react
    assert abs(math.log(n2 / n1) / 7.0 - 1) < 1e-12                      # d ln n / d alpha = 1
    n0 = ou.eos_Pne(eos, T0, A0 * T0)[1]
    assert n0 > 10 * 2e-3                                                # the vacuum ramp is inert
    # the collision of the propagating-branch runs is the published k* = 0.733 1/fm
    ks = kstar(D_s(0.75), tau_n(0.75, tauD_for_kstar(1.4 * KFUND)))
    assert f"{ks:.3f}" == "0.733"


def test_modes_solve_the_dispersion_relation():
    for k, D, tn in ((0.5, 0.2, 0.3), (3.0, 0.2, 0.3), (1.0, 0.1, 1e-3)):
        for w in modes(k, D, tn):
            assert abs(tn * w * w + 1j * w - D * k * k) < 1e-12 * max(1.0, abs(D * k * k))


def test_gk_a_overdamped_branch_on_the_oracle(oracle):
    gk_a(oracle)
