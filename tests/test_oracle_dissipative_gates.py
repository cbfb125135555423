"""Gates of the reference's dissipative sectors, run on the CPU oracle.

G0b (test/test_bjorken_bulk2d.jl): nonlinear bulk relaxation on Bjorken flow.
Transversally uniform, u = 0, theta = 1/tau: the solver's bulk sector reduces to
    de/dtau  = -(e + P + Pi)/tau
    dPi/dtau = (-Pi - zeta/tau)/tau_Pi
with zeta, tau_Pi the model's transport coefficients (src/primitives.jl:98-110)
at the local T. That ODE pair, integrated by RK4 far below the solver's error, is
the referee; the reference's own thresholds apply (dT < 1e-3, dPi < 1e-2 at
CFL 0.15; |Pi|/P > 5 %).

G3 / Gm-style symmetry and conservation (test/test_charge2d.jl,
test/test_consistent_m22d.jl): on a periodic box the total charge is conserved
to round-off, an event symmetric under x <-> y stays symmetric in every sector
(state rows map Sx <-> Sy, nu^x <-> nu^y, pi^xx <-> pi^yy, piQ^xx <-> piQ^yy),
and a transversally uniform state keeps pi_Q^xx = pi_Q^yy and pi_Q^xy = 0.
These pin sign and index mistakes in the flux, bulk, shear, charge and charm
second-moment code that a restatement alone would not.
"""
import functools
import math

import numpy as np
import pytest

import fivo2d as F
from fivo2d import ic
from oracle_unit import OracleUnit

HBARC = 0.1973269804
T0, ALPHA, TAU0 = 0.30, -4.2, 0.4


def bulk_model(zs):
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=False, enable_bulk=True, zeta_over_s=zs,
                            tauPi_coeff=15.0, enable_diff=False)


def run_bulk(lib, zs, tauf, N=24, CFL=0.15):
    g = F.make_grid2d(N, N, xmax=4.0, ymax=4.0)
    m = bulk_model(zs)
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(TAU0, **ic.uniform(g, T0, ALPHA))
    r = s.run_sim_2d(tau0=TAU0, taufinal=tauf, CFL=CFL, CFLtau=0.05, integrator="ssprk2")
    assert r["ok"]
    s.update_primitives_2d(r["tau"])
    w = s.get_work(("yT", "e", "P"))
    U = s.get_state()
    sx, sy = g.interior()
    T = np.exp(w["yT"].reshape(g.Nxtot, g.Nytot)[sx, sy])
    i0 = (g.nghost + N // 2) * g.Nytot + g.nghost + N // 2
    return dict(T=float(np.exp(w["yT"][i0])), e=float(w["e"][i0]), P=float(w["P"][i0]),
                Pi=float(U[i0, m.layout.idx("Pi")]), Tspread=float(np.max(np.abs(T - T.flat[0]) / T.flat[0])))


@functools.lru_cache(maxsize=None)
def _oracle_unit():
    from conftest import ORACLE_LIB
    return OracleUnit(F.Lib(ORACLE_LIB, "orc_"))


@functools.lru_cache(maxsize=None)
def _referee(zs, Cz, tauf, n):
    return bulk_referee(_oracle_unit(), F.LatticeHRGEOS(), zs, Cz, tauf, n, cached=False)


def bulk_referee(ou, eos, zs, Cz, tauf, n=4000, cached=True):
    """RK4 of the Bjorken + bulk pair; T from e by bisection (dP/dmu = 0 for this EOS)."""
    if cached:
        return _referee(zs, Cz, tauf, n)
    def T_of_e(et):
        lo, hi = 1e-4, 5.0
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            if ou.eos_Pne(eos, mid, 0.0)[2] < et:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    def rhs(tau, e, Pi):
        T = T_of_e(e)
        P, nn, ee = ou.eos_Pne(eos, T, ALPHA * T)
        s = max((ee + P - ALPHA * T * nn) / T, 0.0)                         # src/eos.jl:306-310
        zeta = zs / (1.0 + ((T - 0.175) / 0.024) ** 2) * HBARC * s          # src/primitives.jl:98-101
        cs2 = ou.cs2(eos, T)
        delta = max(abs(1 / 3 - cs2), 1e-8)
        tauPi = zeta / (T * s * Cz) / delta ** 2 + 0.1                      # src/primitives.jl:103-108
        return -(e + P + Pi) / tau, (-Pi - zeta / tau) / max(tauPi, 1e-12)

    e = ou.eos_Pne(eos, T0, ALPHA * T0)[2]
    Pi, tau = 0.0, TAU0
    h = (tauf - TAU0) / n
    for _ in range(n):
        k1 = rhs(tau, e, Pi)
        k2 = rhs(tau + h / 2, e + h / 2 * k1[0], Pi + h / 2 * k1[1])
        k3 = rhs(tau + h / 2, e + h / 2 * k2[0], Pi + h / 2 * k2[1])
        k4 = rhs(tau + h, e + h * k3[0], Pi + h * k3[1])
        e += h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        Pi += h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        tau += h
    return e, Pi, T_of_e(e)


def test_referee_rk4_is_converged(oracle):
    """The referee itself: halving its step changes T and Pi far below the gate's tolerances."""
    ou = OracleUnit(oracle)
    eos = F.LatticeHRGEOS()
    a = bulk_referee(ou, eos, 0.15, 15.0, 6.0, n=2000)
    b = bulk_referee(ou, eos, 0.15, 15.0, 6.0, n=4000)
    assert abs(a[2] - b[2]) / b[2] < 1e-8 and abs(a[1] - b[1]) / abs(b[1]) < 1e-6


# dPi the reference measured at CFL 0.15 (test/test_bjorken_bulk2d.jl:131), in percent
G0B_PUBLISHED_DPI = {0.05: 0.52, 0.15: 0.75, 0.30: 0.16}


@pytest.mark.parametrize("zs", [0.05, 0.15, 0.30])
def test_bulk_bjorken_gate_g0b(oracle, zs):
    ou = OracleUnit(oracle)
    eos = F.LatticeHRGEOS()
    P1, n1, e1 = ou.eos_Pne(eos, 0.30, 0.0)
    P2, n2, e2 = ou.eos_Pne(eos, 0.30, -2.0)
    assert abs(P1 - P2) / P1 < 1e-12 and abs(e1 - e2) / e1 < 1e-12   # charm is a pure tracer
    tauf = 6.0
    r = run_bulk(oracle, zs, tauf)
    ee, Pie, Te = bulk_referee(ou, eos, zs, 15.0, tauf)
    dT = abs(r["T"] - Te) / Te
    dPi = abs(r["Pi"] - Pie) / max(abs(Pie), 1e-30)
    print(f"zeta/s {zs}: T {r['T']:.8f} vs {Te:.8f} ({dT:.2e}) Pi {r['Pi']:+.6e} vs {Pie:+.6e} ({dPi:.2e}) "
          f"|Pi|/P {abs(r['Pi']) / r['P']:.3f}")
    assert r["Tspread"] < 1e-12
    assert dT < 1e-3 and dPi < 1e-2
    assert abs(r["Pi"]) / r["P"] > 0.05
    assert round(100 * dPi, 2) == G0B_PUBLISHED_DPI[zs]


def test_bulk_bjorken_converges_in_cfl(oracle):
    """The residual is the solver's time step: it falls with CFL and reaches the
    reference's bounds (dT < 1e-4, dPi < 1e-3) at CFL 0.025."""
    ou = OracleUnit(oracle)
    eos = F.LatticeHRGEOS()
    tauf = 6.0
    ee, Pie, Te = bulk_referee(ou, eos, 0.15, 15.0, tauf)
    eT, eP = [], []
    for cfl in (0.2, 0.1, 0.05, 0.025):
        r = run_bulk(oracle, 0.15, tauf, CFL=cfl)
        eT.append(abs(r["T"] - Te) / Te)
        eP.append(abs(r["Pi"] - Pie) / abs(Pie))
    print("CFL 0.2, 0.1, 0.05, 0.025: dT/T", eT, "dPi/Pi", eP)
    assert eP[3] < eP[2] < eP[1] < eP[0]
    assert eT[3] < 1e-4 and eP[3] < 1e-3


# ---- G2c-e: viscous (shear) Bjorken, test/test_dissipation2d.jl:108-241 -----------
def visc_referee(ou, eos, eta_over_s, Cs, T0v, alpha0, tau0, tauf, nsub=3000):
    """RK4 of de/dtau = -(e + P + pieta)/tau, tau_pi dpi^xx/dtau = -pi^xx (1 + delta_pi theta)
    + pi_NS^xx with pieta = -2 pi^xx, pi_NS^xx = 2 eta/(3 tau), eta = eta/s * s * hbar c,
    tau_pi = eta/(T s Cs), delta_pi = 4/3 tau_pi (src/primitives.jl:55-64,
    src2d/dissipation2d.jl:112-119)."""
    def T_of_e(et):
        lo, hi = 1e-4, 5.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if ou.eos_Pne(eos, mid, alpha0 * mid)[2] > et:
                hi = mid
            else:
                lo = mid
        return 0.5 * (lo + hi)

    def rhs(tau, y):
        e, pxx = y
        T = T_of_e(e)
        P, n, _ = ou.eos_Pne(eos, T, alpha0 * T)
        s = (e + P - alpha0 * T * n) / T
        eta = max(eta_over_s, 0.0) * max(s, 0.0) * HBARC
        taupi = eta / (T * s * Cs)
        th = 1.0 / tau
        de = -(e + P - 2 * pxx) / tau
        dp = (-pxx * (1 + 4.0 / 3.0 * taupi * th) + 2 * eta * th / 3) / taupi if taupi > 0 else 0.0
        return np.array([de, dp])

    y = np.array([ou.eos_Pne(eos, T0v, alpha0 * T0v)[2], 0.0])
    h = (tauf - tau0) / nsub
    tau = tau0
    for _ in range(nsub):
        k1 = rhs(tau, y)
        k2 = rhs(tau + h / 2, y + h / 2 * k1)
        k3 = rhs(tau + h / 2, y + h / 2 * k2)
        k4 = rhs(tau + h, y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        tau += h
    return y[0], y[1]


def run_visc_bjorken(lib, CFLtau=0.004, tau0=0.6, tauf=3.0, T0v=0.45, alpha0=-4.2, eta_over_s=0.15,
                     referee=True):
    g = F.make_grid2d(20, 20, xmax=5.0, ymax=5.0)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=eta_over_s, tauShear_coeff=0.2)
    s = F.Solver(g, m, lib=lib)
    assert s.init_from_prims(tau0, **ic.uniform(g, T0v, alpha0)) == 0
    r = s.run_sim_2d(tau0=tau0, taufinal=tauf, CFL=0.2, CFLtau=CFLtau)
    assert r["ok"]
    U = s.get_state()
    w = s.get_work(("ux", "uy"))
    L = m.layout
    sx, sy = g.interior()
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    Ein = U[inside.ravel(), L.idx("E")]
    i0 = g.lin(g.nghost + g.Nx // 2 - 1, g.nghost + g.Ny // 2 - 1)
    out = dict(tau=r["tau"], nsteps=r["nsteps"], primfail=r["nprimfail"], shear_res=r["max_shear_res"],
               maxU=max(np.max(np.abs(w["ux"][inside.ravel()])), np.max(np.abs(w["uy"][inside.ravel()]))),
               spread=(Ein.max() - Ein.min()) / abs(Ein.max()), e=U[i0, L.idx("E")],
               pixx=U[i0, L.idx("pixx")], pieta=U[i0, L.idx("pieta")])
    if referee:
        e_ref, pxx_ref = visc_referee(_oracle_unit(), F.LatticeHRGEOS(), eta_over_s, 0.2, T0v, alpha0, tau0, r["tau"])
        out.update(e_ref=e_ref, pixx_ref=pxx_ref, eerr=abs(out["e"] - e_ref) / e_ref,
                   perr=abs(out["pixx"] - pxx_ref) / max(abs(pxx_ref), 1e-12))
    return out


def test_viscous_bjorken_gate_g2c(oracle):
    r = run_visc_bjorken(oracle)
    print(f"steps {r['nsteps']} e {r['e']:.6e} vs {r['e_ref']:.6e} ({r['eerr']:.2e}) "
          f"pi^xx {r['pixx']:.6e} vs {r['pixx_ref']:.6e} ({r['perr']:.2e}) pieta {r['pieta']:.4e}")
    assert r["primfail"] == 0
    assert r["maxU"] == 0.0
    assert r["spread"] < 1e-12
    assert r["shear_res"] < 1e-12
    assert r["eerr"] < 2e-3 and r["perr"] < 2e-2


def test_shear_slows_bjorken_cooling_g2d(oracle):
    ideal = run_visc_bjorken(oracle, eta_over_s=0.0, referee=False)
    visc = run_visc_bjorken(oracle, eta_over_s=0.15, referee=False)
    print(f"e ideal {ideal['e']:.6e} viscous {visc['e']:.6e}; pieta {visc['pieta']:.4e}")
    assert visc["pieta"] < 0.0
    assert visc["e"] > ideal["e"]


def test_operator_split_converges_g2e(oracle):
    errs = [run_visc_bjorken(oracle, CFLtau=c)["eerr"] for c in (0.016, 0.008, 0.004)]
    p = math.log2(errs[1] / errs[2])
    print("errors", errs, "order", p)
    assert p > 0.7
    assert f"{p:.2f}" == "1.05"         # the published order (README2D.md:277)


# ---- symmetry and conservation with every sector on ---------------------------
def full_model():
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=True)


def symmetric_event(g, seed=4, nspots=40, sigma=0.6):
    """A lumpy event symmetric under x <-> y: the hot-spot sum plus its mirror."""
    p = ic.hotspot_event(g, seed=seed, nspots=nspots, sigma=sigma, Tpeak=0.45)
    T = p["T"].reshape(g.Nxtot, g.Nytot)
    a = p["alpha"].reshape(g.Nxtot, g.Nytot)
    T = 0.5 * (T + T.T)
    a = 0.5 * (a + a.T)
    return dict(T=T.ravel(), alpha=a.ravel(), ux=p["ux"], uy=p["uy"])


def transpose_rows(U, g, layout):
    """The state of the mirrored event: cells transposed, x and y components swapped."""
    nv = U.shape[1]
    V = U.reshape(g.Nxtot, g.Nytot, nv).transpose(1, 0, 2).copy()
    idx = {n: k for k, n in enumerate(layout.names)}
    for a, b in (("Sx", "Sy"), ("nux", "nuy"), ("pixx", "piyy"), ("pQxx", "pQyy")):
        if a in idx and b in idx:
            V[..., [idx[a], idx[b]]] = V[..., [idx[b], idx[a]]]
    return V.reshape(-1, nv)


def run_bump(lib, bc=F.FIVO_BC_OUTFLOW, enable_diff=True, N=48, box=8.0, tau0=0.6, tauf=2.0):
    """G3's Gaussian fugacity bump on a uniform T = 0.40 background
    (test/test_charge2d.jl:104-134): alpha = -4.5 + 1.5 exp(-(r/2)^2)."""
    g = F.make_grid2d(N, N, xmax=box, ymax=box)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_diff=enable_diff, kappa_coeff=0.1163, tauN_coeff=1.0)
    X, Y = np.meshgrid(g.xC, g.yC, indexing="ij")
    alpha = -4.5 + 1.5 * np.exp(-(np.hypot(X, Y) / 2.0) ** 2)
    s = F.Solver(g, m, lib=lib)
    assert s.init_from_prims(tau0, T=np.full(g.Ntot, 0.40), alpha=alpha.ravel(), bc=bc) == 0
    sx, sy = g.interior()
    Q0 = s.get_state()[:, 0].reshape(g.Nxtot, g.Nytot)[sx, sy].sum()
    r = s.run_sim_2d(tau0=tau0, taufinal=tauf, CFL=0.2, CFLtau=0.01, bc=bc)
    assert r["ok"]
    U = s.get_state()
    Q1 = U[:, 0].reshape(g.Nxtot, g.Nytot)[sx, sy].sum()
    return dict(g=g, m=m, s=s, U=U, r=r, Q0=Q0, Q1=Q1)


def test_charge_conserved_on_closed_domain_g3c(oracle):
    """The Dtau source is zero, so on a periodic box the interior sum of tau J^tau
    changes only by round-off; with outflow the charm background leaves."""
    per = run_bump(oracle, bc=F.FIVO_BC_PERIODIC)
    out = run_bump(oracle)
    dper = abs(per["Q1"] - per["Q0"]) / abs(per["Q0"])
    dout = abs(out["Q1"] - out["Q0"]) / abs(out["Q0"])
    print(f"charge drift: periodic {dper:.3e}, outflow {dout:.3e}; steps {per['r']['nsteps']}")
    assert per["r"]["nprimfail"] == 0 and out["r"]["nprimfail"] == 0
    assert dper < 1e-12
    assert dout > 1e3 * max(dper, 1e-16)      # the outflow run does lose charge: the box is open


def upgradient_cells(run):
    """G3d's count (test/test_charge2d.jl:171-183): interior cells off the edge ring
    with a resolvable central difference of alpha in x and a resolvable nu^x, and of
    those the ones whose nu^x points up the gradient."""
    g, L = run["g"], run["m"].layout
    alpha = run["s"].get_work(("alpha",))["alpha"]
    nux = run["U"][:, L.idx("nux")]
    ng, nyt = g.nghost, g.Nytot
    nsame = ntot = 0
    for ix in range(ng + 1, ng + g.Nx - 1):
        for iy in range(ng + 1, ng + g.Ny - 1):
            i = ix * nyt + iy
            dxa = alpha[i + nyt] - alpha[i - nyt]
            if abs(dxa) < 1e-10 or abs(nux[i]) < 1e-14:
                continue
            ntot += 1
            nsame += nux[i] * dxa > 0
    return nsame, ntot


def test_diffusion_runs_down_the_gradient_g3d(oracle):
    on = run_bump(oracle)
    off = run_bump(oracle, enable_diff=False)
    g, L = on["g"], on["m"].layout
    ic0 = g.lin(g.nghost + g.Nx // 2 - 1, g.nghost + g.Ny // 2 - 1)     # lin(g, ng+Nx/2, ng+Ny/2), 1-based
    peak_on, peak_off = on["U"][ic0, 0], off["U"][ic0, 0]
    print(f"central tau J^tau: diffusion off {peak_off:.6e} on {peak_on:.6e}")
    assert peak_on < peak_off
    nsame, ntot = upgradient_cells(on)
    print(f"cells with nu^x up the gradient: {nsame} / {ntot}")
    assert ntot > 100 and nsame == 0
    assert ntot == 2104                     # the published count (README2D.md:281: 0 of 2104)


def test_radial_bump_stays_mirror_symmetric_g3e(oracle):
    r = run_bump(oracle)
    g, m, U = r["g"], r["m"], r["U"]
    sx, sy = g.interior()
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    Ui = U[inside.ravel()]
    Vi = transpose_rows(U, g, m.layout)[inside.ravel()]
    worst = np.max(np.abs(Ui - Vi) / np.maximum(np.maximum(np.abs(Ui), np.abs(Vi)), 1e-30))
    print(f"max x<->y asymmetry {worst:.3e}")
    assert worst < 1e-12


def test_full_sectors_stay_mirror_symmetric(oracle):
    """Every sector on (shear, bulk, diffusion, charm second moment), a lumpy
    event symmetric under x <-> y: the state stays symmetric row by row, to the
    bound the reference's all-sector gate sets (x<->y < 1e-9,
    test/test_production_allsectors2d.jl:123-125); round-off of the ordered x, y
    sweeps is amplified by the Newton solves, which is why it is not 1e-12."""
    g = F.make_grid2d(32, 32, xmax=8.0, ymax=8.0)
    m = full_model()
    s = F.Solver(g, m, lib=oracle)
    s.init_from_prims(0.6, **symmetric_event(g))
    r = s.run_sim_2d(tau0=0.6, taufinal=1.2, integrator="ssprk3")
    assert r["ok"] and r["nsteps"] > 5
    U = s.get_state()
    sx, sy = g.interior()
    V = transpose_rows(U, g, m.layout)
    interior = np.zeros((g.Nxtot, g.Nytot), bool)
    interior[sx, sy] = True
    Ui, Vi = U[interior.ravel()], V[interior.ravel()]
    scale = np.maximum(np.max(np.abs(Ui), axis=0), 1e-300)
    asym = np.max(np.abs(Ui - Vi), axis=0) / scale
    print("x<->y asymmetry per row", dict(zip(m.layout.names, asym)))
    assert np.all(np.max(np.abs(Ui), axis=0)[[m.layout.idx(k) for k in ("pixy", "pQxx", "nux", "Pi")]] > 0)
    assert np.max(asym) < 1e-9, dict(zip(m.layout.names, asym))


def test_uniform_state_keeps_second_moment_isotropic(oracle):
    """Bjorken limit of the charm second moment: pi_Q^xx == pi_Q^yy and pi_Q^xy == 0."""
    g = F.make_grid2d(12, 12, xmax=4.0, ymax=4.0)
    m = full_model()
    s = F.Solver(g, m, lib=oracle)
    s.init_from_prims(0.6, **ic.uniform(g, 0.35, -3.0))
    r = s.run_sim_2d(tau0=0.6, taufinal=1.5, integrator="ssprk3")
    assert r["ok"]
    U = s.get_state()
    L = m.layout
    sx, sy = g.interior()
    inside = np.zeros((g.Nxtot, g.Nytot), bool)
    inside[sx, sy] = True
    q = U[inside.ravel()]
    xx, yy, xy = q[:, L.idx("pQxx")], q[:, L.idx("pQyy")], q[:, L.idx("pQxy")]
    assert np.max(np.abs(xx)) > 0                      # the moment is actually sourced
    assert np.array_equal(xx, yy) and np.all(xy == 0.0)
    assert np.array_equal(q[:, L.idx("pixx")], q[:, L.idx("piyy")]) and np.all(q[:, L.idx("pixy")] == 0.0)
