"""The HIP freeze-out tracker and Cooper-Frye sweep (fivo_freezeout_*) against the
oracle and the closed forms of test_oracle_freezeout.py, and the reference's
axisymmetric control (freezeout2d.jl:79-86; published yields in
TWOD_PROGRAM.md:1847 and :2396-2431).

Integer outcomes (which cells are hot, frozen, re-heated) are compared exactly;
the integral, a sum of ~1e7..1e9 FP64 terms in another order, to 1e-12
relative on surfaces set directly and to 1e-9 on surfaces the two libraries
evolved themselves (their temperatures differ by round-off, so crossings move
by round-off)."""
import os

import numpy as np
import pytest

from fivo2d import freezeout as FO

import freezeout_cases as FC

pytestmark = pytest.mark.gpu

CSV = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "initial_profiles_physical.csv")


def test_crossing_histories_on_the_device(hip):
    sf, ex = FC.crossing_histories(hip)
    t = sf["tau_fo"]
    assert np.array_equal(np.isnan(t), np.isnan(ex["tau_fo"]))
    ok = ~np.isnan(t)
    assert np.max(np.abs(t[ok] - ex["tau_fo"][ok])) < 1e-12
    assert np.max(np.abs(sf["uxf"] - ex["uxf"])) < 1e-15
    assert (sf["nhot"], sf["nfroze"], sf["refroze"]) == (ex["nhot"], ex["nfroze"], ex["refroze"])


@pytest.mark.parametrize("g,u", [((0.0, 0.0), (0.0, 0.0)), ((-0.1, 0.25), (0.12, -0.1))])
def test_closed_forms_on_the_device(hip, g, u):
    s, fo = FC.tracker(hip, n=24)
    n = fo.nx
    t = FC.plane_surface(fo, 4.0, *g)
    fo.set_surface(t, np.full((n, n), u[0]), np.full((n, n), u[1]))
    NB, Nb, negf = fo.pion_yield(**FC.FINE)
    for bose, got in ((False, NB), (True, Nb)):
        exp = FC.tilted_flow_yield(t[1:-1, 1:-1].ravel(), g[0], g[1], u[0], u[1], fo.dx, fo.dy, bose=bose)
        assert abs(got / exp - 1) < 1e-11, (bose, got, exp)
    assert negf == 0.0
    s.close()


def _lumpy_surface(fo, seed):
    """A smooth bumpy tau_fo on a ragged frozen region with holes, steep enough
    in places for dSigma.p < 0, and a swirling flow."""
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(fo.xs, fo.ys, indexing="ij")
    t = 6.0 + 1.5 * np.cos(0.7 * X + 0.3) * np.sin(0.5 * Y) + 0.8 * X / fo.xs[-1]
    r = np.hypot(X, Y)
    region = (r < 0.8 * fo.xs[-1] + 0.6 * np.sin(5 * np.arctan2(Y, X))) & (rng.random(X.shape) > 0.03)
    t = np.where(region, t, np.nan)
    ux = 0.6 * np.tanh(X / 3) - 0.2 * Y / fo.xs[-1]
    uy = 0.6 * np.tanh(Y / 3) + 0.2 * X / fo.xs[-1]
    return t, ux, uy


@pytest.mark.parametrize("n,grid", [(40, {}), (97, dict(n_eta=33, n_pt=17, n_phi=40, eta_max=5.0)),
                                    (200, {})])
def test_device_sum_equals_the_oracle(hip, oracle, n, grid):
    out = []
    for lib in (hip, oracle):
        s, fo = FC.tracker(lib, n=n, box=10.0)
        fo.set_surface(*_lumpy_surface(fo, seed=n))
        out.append(fo.pion_yield(**grid))
        s.close()
    (hB, hb, hn), (oB, ob, on) = out
    print(f"  n={n}: HIP {hB:.12e} {hb:.12e} {hn:.6e} | oracle {oB:.12e} {ob:.12e} {on:.6e}", flush=True)
    assert abs(hB / oB - 1) < 1e-12 and abs(hb / ob - 1) < 1e-12
    assert abs(hn - on) < 1e-12 * max(on, 1e-30) + 1e-15
    assert hn > 0.0


def test_device_sum_is_reproducible(hip):
    s, fo = FC.tracker(hip, n=120, box=10.0)
    fo.set_surface(*_lumpy_surface(fo, seed=3))
    a = fo.pion_yield()
    b = fo.pion_yield()
    assert a == b
    s.close()


def test_evolved_surface_equals_the_oracle(hip, oracle):
    """The axisymmetric control on a coarse grid, evolved and recorded by each
    library: same hot, frozen and re-heated cells, the same surface to round-off."""
    res = {}
    for name, lib in (("hip", hip), ("oracle", oracle)):
        fo = FO.freeze_surface("axisym", N=48, dtau=0.1, taumax=14.0, lib=lib, profile_csv=CSV, verbose=False)
        res[name] = (fo.surface(), fo.pion_yield())
        fo.solver.close()
    (sh, yh), (so, yo) = res["hip"], res["oracle"]
    print(f"  N=48: hot {sh['nhot']} froze {sh['nfroze']} re-heated {sh['refroze']} | yields HIP {yh} oracle {yo}",
          flush=True)
    assert (sh["nhot"], sh["nfroze"], sh["refroze"]) == (so["nhot"], so["nfroze"], so["refroze"])
    th, to = sh["tau_fo"], so["tau_fo"]
    assert np.array_equal(np.isnan(th), np.isnan(to))
    ok = ~np.isnan(th)
    assert np.max(np.abs(th[ok] - to[ok])) < 1e-9
    assert np.max(np.abs(sh["uxf"] - so["uxf"])) < 1e-9
    for a, b in zip(yh[:2], yo[:2]):
        assert abs(a / b - 1) < 1e-9


def _control(hip, N, dtau, restart):
    fo = FO.freeze_surface("axisym", N=N, dtau=dtau, lib=hip, profile_csv=CSV, restart_history=restart,
                           verbose=False)
    sf = fo.surface()
    NB, Nb, negf = fo.pion_yield()
    t = sf["tau_fo"]
    sym = max(np.nanmax(np.abs(t - t.T)), np.nanmax(np.abs(t - t[::-1, :])))
    print(f"  N={N} dtau={dtau} history {'cleared per chunk' if restart else 'kept'}: hot {sf['nhot']} "
          f"froze {sf['nfroze']} re-heated {sf['refroze']} | dN/dy Boltzmann {NB:.2f} Bose {Nb:.2f} "
          f"(Bose/Boltzmann {Nb / NB:.4f}) neg {negf:.2e} | symmetry {sym:.1e}", flush=True)
    fo.solver.close()
    assert sf["nhot"] == sf["nfroze"] > 1000
    assert sym < 1e-9
    assert negf < 1e-3
    return Nb


@pytest.mark.parametrize("dtau,published", [(0.02, 1172.02), (0.05, 1180.96)])
def test_axisymmetric_control_published(hip, dtau, published):
    """freeze_surface("axisym") at the reference's settings (N = 300, box 18 fm,
    tau to 16 fm, eta/s = zeta/s = 0.1, interpolated crossing) with the
    relaxation history cleared at every chunk, as the reference ran when it
    published these Bose yields (TWOD_PROGRAM.md:2396-2431): 1172.02 at dtau =
    0.02 and 1180.96 at dtau = 0.05."""
    Nb = _control(hip, 300, dtau, True)
    assert abs(Nb / published - 1) < 1e-3


def test_axisymmetric_control_kept_history(hip):
    """The reference as it stands, history kept across chunks (main2D.jl:376): the
    dtau dependence the published runs carried (-0.76%, TWOD_PROGRAM.md:2401) is
    gone, as the restart-neutral fix says it should be (TWOD_PROGRAM.md:743)."""
    a = _control(hip, 300, 0.02, False)
    b = _control(hip, 300, 0.05, False)
    assert abs(a / b - 1) < 5e-4
    c = _control(hip, 200, 0.02, False)
    assert 0.97 < c / a < 1.0
