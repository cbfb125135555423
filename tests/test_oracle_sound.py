"""Gate Gs (test/test_sound2d.jl): sound speed and attenuation against linear theory.

A uniform LatticeHRG background at rest (alpha = -6, tau0 = 320 so the Bjorken
expansion is slow), periodic box of 6 fm, a 1e-3 travelling sound wave seeded as
the exact viscous eigenmode: delta u from omega / (w k) and delta pi from the
Israel-Stewart response, omega the sound root of
    (omega^2 - c_s^2 k^2)(1 - i omega tau_pi) + i omega D k^2 = 0,
D = ((4/3) eta + zeta) / (e + P) in natural units. The complex Fourier amplitude of
T along one row gives c_s from its phase and Gamma from its envelope; theory:
c_s = sqrt(eos_cs2), Gamma = k^2 ((4/3) eta + zeta) / (2 (e + P)) with zeta's
Lorentzian averaged over the run. Published: c_s to 0.05 % (README2D.md:285).

On the oracle: the ideal wave (c_s and the numerical damping floor) and one shear
viscosity; the GPU test runs the reference's whole set (three eta/s, bulk at two
temperatures across the Lorentzian).
"""
import cmath
import functools
import math

import numpy as np

import fivo2d as F
from oracle_unit import OracleUnit

INVFMGEV = 1 / 0.1973269804
ALPHA, TAU0, AMP, LBOX = -6.0, 320.0, 1e-3, 6.0


def lorentz(T):
    return 1 / (1 + ((T - 0.175) / 0.024) ** 2)


@functools.lru_cache(maxsize=None)
def _ou():
    from conftest import ORACLE_LIB
    return OracleUnit(F.Lib(ORACLE_LIB, "orc_"))


def sound_root(k, cs2, D, tauR):
    """The damped forward sound root (test/test_sound2d.jl:82-97)."""
    if tauR <= 1e-12:
        disc = cs2 * k ** 2 - (D * k ** 2 / 2) ** 2
        return math.sqrt(max(disc, 0.0)) - 1j * (D * k ** 2 / 2)
    c = np.array([-1j * cs2 * k ** 2, -(tauR * cs2 * k ** 2 + D * k ** 2), 1j]) / tauR
    Cm = np.array([[0, 0, -c[0]], [1, 0, -c[1]], [0, 1, -c[2]]], dtype=complex)
    best = None
    for w in np.linalg.eigvals(Cm):
        if w.real > 1e-9 and (best is None or w.imag > best.imag):
            best = w
    return best


def _fit(x, y):
    x, y = np.asarray(x), np.asarray(y)
    mx, my = x.mean(), y.mean()
    return float(((x - mx) * (y - my)).sum() / ((x - mx) ** 2).sum())


def sound(lib, T0, es, zs, N=128, trun=12.0, dtau=0.5):
    g = F.make_grid2d(N, N, xmax=LBOX / 2, ymax=LBOX / 2)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=es > 0, eta_over_s=es, tauShear_coeff=0.2,
                         deltaShear_factor=4 / 3, enable_bulk=zs > 0, zeta_over_s=zs, tauPi_coeff=15.0,
                         enable_diff=False)
    ou, eos = _ou(), F.LatticeHRGEOS()
    k = 2 * math.pi / LBOX
    P0, n0, e0 = ou.eos_Pne(eos, T0, ALPHA * T0)
    cs2 = ou.cs2(eos, T0)
    hT = 1e-5 * T0
    ep = ou.eos_Pne(eos, T0 + hT, ALPHA * (T0 + hT))[2]
    em = ou.eos_Pne(eos, T0 - hT, ALPHA * (T0 - hT))[2]
    dedT = (ep - em) / (2 * hT)
    s0 = (e0 + P0 - ALPHA * T0 * n0) / T0
    w_nat = (e0 + P0) * INVFMGEV
    etav, zetav = es * s0, zs * lorentz(T0) * s0
    D = ((4 / 3) * etav + zetav) / w_nat
    tauR = 5 * es / (T0 * INVFMGEV) if es > 0 else 0.0
    om = sound_root(k, cs2, D, tauR)
    de = dedT * AMP * T0 * INVFMGEV
    du = om * de / (w_nat * k)
    dpi = -(4 / 3) * etav * (1j * k * du) / (1 - 1j * om * tauR) if etav > 0 else 0j
    X, _ = np.meshgrid(g.xC, g.yC, indexing="ij")
    ph = np.exp(1j * k * X)
    pxx = (dpi * ph).real / INVFMGEV
    prims = dict(T=(T0 + AMP * T0 * ph.real).ravel(), alpha=np.full(g.Ntot, ALPHA), ux=(du * ph).real.ravel(),
                 uy=np.zeros(g.Ntot))
    if es > 0:
        prims.update(pixx=pxx.ravel(), piyy=(-pxx / 2).ravel(), pixy=np.zeros(g.Ntot))
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(TAU0, **prims)
    ng = g.nghost
    iy0 = ng + g.Ny // 2
    xs = g.xC[ng:ng + g.Nx]
    phase = np.exp(-1j * k * xs)

    def amp():
        T = np.exp(s.get_work(("yT",))["yT"]).reshape(g.Nxtot, g.Nytot)[ng:ng + g.Nx, iy0]
        return 2 * (T * phase).sum() / g.Nx, T.mean()

    tau = TAU0
    s.rhs_2d(tau, want_dU=False)
    c0, Tb = amp()
    ts, la, phs, lor = [0.0], [math.log(abs(c0) / Tb)], [cmath.phase(c0)], [lorentz(Tb)]
    while tau < TAU0 + trun - 1e-9:
        r = s.run_sim_2d(tau0=tau, taufinal=min(tau + dtau, TAU0 + trun), CFL=0.15, CFLtau=0.05,
                         bc=F.FIVO_BC_PERIODIC, reset_history=False)
        assert r["ok"]
        tau = r["tau"]
        s.update_primitives_2d(tau, bc=F.FIVO_BC_PERIODIC)
        c, Tb = amp()
        ts.append(tau - TAU0); la.append(math.log(abs(c) / Tb)); phs.append(cmath.phase(c)); lor.append(lorentz(Tb))
    for j in range(1, len(phs)):
        while phs[j] - phs[j - 1] > math.pi:
            phs[j] -= 2 * math.pi
        while phs[j] - phs[j - 1] < -math.pi:
            phs[j] += 2 * math.pi
    nn = len(ts)
    return dict(G=-_fit(ts, la), G1=-_fit(ts[:nn // 2], la[:nn // 2]), G2=-_fit(ts[nn // 2 - 1:], la[nn // 2 - 1:]),
                cs=-_fit(ts, phs) / k, cs_exact=math.sqrt(cs2), k=k, eP_nat=w_nat, s0=s0,
                lorbar=sum(lor) / len(lor))


def shear_ratio(r, base, es):
    Gp = 0.5 * r["k"] ** 2 * (4 / 3) * (es * r["s0"]) / r["eP_nat"]
    g = r["G"] - base["G"]
    split = abs((r["G1"] - base["G1"]) - (r["G2"] - base["G2"])) / g
    return g / Gp, split


def test_sound_speed_and_shear_attenuation_on_the_oracle(oracle):
    base = sound(oracle, 0.35, 0.0, 0.0)
    dcs = abs(base["cs"] - base["cs_exact"]) / base["cs_exact"]
    print(f"ideal: c_s {base['cs']:.5f} vs exact {base['cs_exact']:.5f} ({100 * dcs:.3f} %), floor {base['G']:.2e}")
    assert dcs < 0.02 and abs(base["G"]) < 5e-3
    assert round(100 * dcs, 2) <= 0.05                        # the published 0.05 %
    r = sound(oracle, 0.35, 0.02, 0.0)
    ratio, split = shear_ratio(r, base, 0.02)
    print(f"eta/s 0.02: Gamma / NS {ratio:.5f}, half-window split {split:.4f}")
    assert 0.98 < ratio < 1.02 and split < 0.01
