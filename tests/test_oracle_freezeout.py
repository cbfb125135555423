"""The oracle's freeze-out tracker and Cooper-Frye integral (freezeout2d.jl:68-192)
against what fixes them independently of the code:

  - prescribed temperature histories: the linearly interpolated crossing of a
    linear cooling law is the exact crossing; a re-heated cell loses its tau_fo,
    is counted once per re-heating and freezes again at its second crossing; a
    cell that is never hot never freezes; the flow is that of the sample that
    closes the crossing (freezeout2d.jl:97-119);
  - closed forms of the integral: a uniform surface at rest gives
    2 pi p_T 2 m_T K1(m_T/T) per p_T (Boltzmann) and the series over n of
    K1(n m_T/T) (Bose); a tilted surface tau_fo = tau0 + c.x with uniform flow
    gives I0, I1 of |u| p_T/T and K0, K1 of u^tau m_T/T -- a wrong sign of
    u.p or of the gradient term, a dropped (hbar c)^3, a wrong cell area or
    yield factor each move the result by orders beyond the tolerance;
  - the rim: one-sided gradients on a frozen disk (exact for a plane), and a
    frozen outer ring alone contributes nothing (freezeout2d.jl:151-169);
  - the negative-dSigma share: zero on a time-like surface, positive on a
    steep one, whose total Boltzmann yield stays the untilted one (the tilt's
    odd phi term sums to zero on the symmetric phi grid).
"""
import numpy as np
import pytest

import fivo2d as F
from fivo2d import freezeout as FO

import freezeout_cases as FC


def test_crossing_histories(oracle):
    sf, ex = FC.crossing_histories(oracle)
    t = sf["tau_fo"]
    assert np.array_equal(np.isnan(t), np.isnan(ex["tau_fo"]))
    ok = ~np.isnan(t)
    assert np.max(np.abs(t[ok] - ex["tau_fo"][ok])) < 1e-12
    assert np.max(np.abs(sf["uxf"] - ex["uxf"])) < 1e-15
    assert np.max(np.abs(sf["uyf"] + ex["uxf"])) < 1e-15
    assert (sf["nhot"], sf["nfroze"], sf["refroze"]) == (ex["nhot"], ex["nfroze"], ex["refroze"])


def _fo(lib, n=24, box=6.0):
    return FC.tracker(lib, n=n, box=box)


@pytest.mark.parametrize("bose", [False, True])
def test_static_surface_is_the_bessel_closed_form(oracle, bose):
    s, fo = _fo(oracle)
    n = fo.nx
    fo.set_surface(np.full((n, n), 2.5), np.zeros((n, n)), np.zeros((n, n)))
    NB, Nb, negf = fo.pion_yield(**FC.FINE)
    exp = FC.static_yield(2.5, (n - 2) ** 2, fo.dx, fo.dy, bose=bose)
    got = Nb if bose else NB
    assert abs(got / exp - 1) < 1e-12, (got, exp)
    assert negf == 0.0
    s.close()


@pytest.mark.parametrize("bose", [False, True])
@pytest.mark.parametrize("g,u", [((0.2, 0.0), (0.15, 0.0)), ((-0.1, 0.25), (0.12, -0.1)),
                                 ((0.3, -0.2), (-0.05, 0.14))])
def test_tilted_flowing_surface_is_the_bessel_closed_form(oracle, g, u, bose):
    s, fo = _fo(oracle)
    n = fo.nx
    t = FC.plane_surface(fo, 4.0, *g)
    fo.set_surface(t, np.full((n, n), u[0]), np.full((n, n), u[1]))
    NB, Nb, negf = fo.pion_yield(**FC.FINE)
    exp = FC.tilted_flow_yield(t[1:-1, 1:-1].ravel(), g[0], g[1], u[0], u[1], fo.dx, fo.dy, bose=bose)
    got = Nb if bose else NB
    assert abs(got / exp - 1) < 1e-11, (got, exp)
    # the same surface with the flow reversed moves the yield: the u.p coupling is live
    fo.set_surface(t, np.full((n, n), -u[0]), np.full((n, n), -u[1]))
    NB2, Nb2, _ = fo.pion_yield(**FC.FINE)
    exp2 = FC.tilted_flow_yield(t[1:-1, 1:-1].ravel(), g[0], g[1], -u[0], -u[1], fo.dx, fo.dy, bose=bose)
    assert abs((Nb2 if bose else NB2) / exp2 - 1) < 1e-11
    assert abs(exp2 / exp - 1) > 1e-3
    s.close()


def test_one_sided_gradients_at_the_rim(oracle):
    s, fo = _fo(oracle, n=30, box=7.5)
    X, Y = np.meshgrid(fo.xs, fo.ys, indexing="ij")
    disk = np.hypot(X - 0.4, Y + 0.3) < 5.2
    gx, gy, u = 0.25, -0.15, (0.1, 0.08)
    t = FC.plane_surface(fo, 3.0, gx, gy, mask=disk)
    n = fo.nx
    fo.set_surface(t, np.full((n, n), u[0]), np.full((n, n), u[1]))
    NB, Nb, _ = fo.pion_yield(**FC.FINE)
    take = FC.counted(disk)
    # the rim carries one-sided differences: a sizeable share of the cells
    inner = np.zeros_like(disk)
    inner[1:-1, 1:-1] = disk[1:-1, 1:-1] & disk[:-2, 1:-1] & disk[2:, 1:-1] & disk[1:-1, :-2] & disk[1:-1, 2:]
    assert np.sum(take & ~inner) > 20
    for bose, got in ((False, NB), (True, Nb)):
        exp = FC.tilted_flow_yield(t[take], gx, gy, u[0], u[1], fo.dx, fo.dy, bose=bose)
        assert abs(got / exp - 1) < 1e-11, (bose, got, exp)
    s.close()


def test_outer_ring_alone_contributes_nothing(oracle):
    s, fo = _fo(oracle, n=16)
    n = fo.nx
    t = np.full((n, n), np.nan)
    t[0, :] = t[-1, :] = t[:, 0] = t[:, -1] = 2.0
    fo.set_surface(t, np.zeros((n, n)), np.zeros((n, n)))
    assert fo.pion_yield(**FC.FINE) == (0.0, 0.0, 0.0)
    # an isolated frozen cell inside has no gradient and is skipped too
    t[7, 7] = 2.0
    fo.set_surface(t, np.zeros((n, n)), np.zeros((n, n)))
    assert fo.pion_yield(**FC.FINE)[:2] == (0.0, 0.0)
    # with one neighbour in x and one in y it counts, with one-sided gradients;
    # its two neighbours lack a neighbour in the other direction and do not
    t[8, 7] = t[7, 8] = 2.0
    fo.set_surface(t, np.zeros((n, n)), np.zeros((n, n)))
    NB = fo.pion_yield(**FC.FINE)[0]
    assert abs(NB / FC.static_yield(2.0, 1, fo.dx, fo.dy) - 1) < 1e-12
    s.close()


def test_negative_share(oracle):
    s, fo = _fo(oracle)
    n = fo.nx
    z = np.zeros((n, n))
    fo.set_surface(FC.plane_surface(fo, 4.0, 0.3, 0.0), z, z)
    assert fo.pion_yield(**FC.FINE)[2] == 0.0
    steep = FC.plane_surface(fo, 40.0, 3.0, 0.0)      # space-like where 3 p_x > m_T cosh(eta)
    fo.set_surface(steep, z, z)
    NB, _, negf = fo.pion_yield(**FC.FINE)
    assert 0.01 < negf < 0.5
    exp = FC.tilted_flow_yield(steep[1:-1, 1:-1].ravel(), 3.0, 0.0, 0.0, 0.0, fo.dx, fo.dy)
    assert abs(NB / exp - 1) < 1e-11
    s.close()


def test_reference_grid_defaults_and_refusals(oracle):
    """The reference's grid (40 x 40 x 24 on [-6, 6] x [0.1, 5] GeV) on a static
    surface: its eta spacing under-resolves the Boltzmann peak only at large p_T
    (width sqrt(T/m_T) ~ 0.18 < deta = 0.31 at 5 GeV), where the yield is
    exp(-30) smaller, so the sum still equals the closed form; the Bose yield
    lies above the Boltzmann one; out-of-range grids are refused."""
    s, fo = _fo(oracle)
    n = fo.nx
    fo.set_surface(np.full((n, n), 2.5), np.zeros((n, n)), np.zeros((n, n)))
    NB, Nb, _ = F.freezeout.pion_yield(fo)
    grid = dict(n_eta=40, n_pt=40, n_phi=24, eta_max=6.0, pt_lo=0.1, pt_hi=5.0)
    exp = FC.static_yield(2.5, (n - 2) ** 2, fo.dx, fo.dy, **grid)
    assert abs(NB / exp - 1) < 1e-10
    assert 1.05 < Nb / NB < 1.2
    for bad in (dict(n_eta=1), dict(n_pt=300), dict(pt_lo=5.0, pt_hi=1.0), dict(eta_max=0.0)):
        with pytest.raises(F.FivoError):
            fo.pion_yield(**bad)
    s.close()


def test_tracker_needs_a_positive_temperature(oracle):
    g = F.make_grid2d(12, 12, xmax=6.0, ymax=6.0)
    s = F.Solver(g, FC.medium(), lib=oracle)
    with pytest.raises(F.FivoError):
        FO.FreezeOut(s, T_fo=0.0)
    s.close()


def test_axisymmetric_surface_on_a_coarse_grid(oracle):
    """freeze_surface("axisym") on the oracle at N = 40, dtau = 0.1: every hot cell
    freezes by tau = 16, the surface keeps the profile's x <-> y and mirror
    symmetries to round-off, and the Bose yield sits above the Boltzmann one by
    the thermal ratio of a pion gas near T_fo."""
    import os
    csv = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "initial_profiles_physical.csv")
    fo = FO.freeze_surface("axisym", N=40, dtau=0.1, lib=oracle, profile_csv=csv, verbose=False)
    sf = fo.surface()
    NB, Nb, negf = FO.pion_yield(fo)
    t = sf["tau_fo"]
    assert sf["nhot"] == sf["nfroze"] > 200
    assert np.nanmax(np.abs(t - t.T)) < 1e-11 and np.nanmax(np.abs(t - t[::-1, :])) < 1e-11
    assert 1.09 < Nb / NB < 1.12 and 0.0 <= negf < 1e-3
    assert 900 < Nb < 1300
    fo.solver.close()
