"""The background readout (tools/export_background2d.jl) on the oracle and the host
mirror: one slice's conversion and table layout, the maxima and their NaN rule, a
uniform-flow known answer, the exporter's output keys, time grid and file round
trip, and its |v| < 1 refusal."""
import math

import numpy as np
import pytest

import fivo2d as F
from fivo2d import background as B
from fivo2d import ic


def _event(g, seed=3):
    return ic.hotspot_event(g, seed=seed, nspots=40, sigma=0.6, Tpeak=0.4)


def _evolved(lib, nx=24, ny=20, steps=4):
    g = F.make_grid2d(nx, ny, xmax=8.0, ymax=8.0 * ny / nx)
    s = F.Solver(g, B.background_model(), lib=lib)
    s.init_from_prims(0.4, **_event(g))
    r = s.run_sim_2d(tau0=0.4, taufinal=10.0, CFL=0.2, CFLtau=0.02, max_steps=steps)
    assert r["ok"]
    s.update_primitives_2d(r["tau"])
    return g, s


def _tables(g, nt):
    return [np.full((g.Nx, g.Ny, nt), np.nan, order="F") for _ in range(4)]


def test_slice_conversion_layout_and_maxima(oracle):
    g, s = _evolved(oracle)
    T, vx, vy, n = _tables(g, 3)
    mu, mv = s.background_slice(1, 3, T, vx, vy, n)
    w = s.get_work(("yT", "ux", "uy", "n"))
    sx, sy = g.interior()
    grab = lambda a: a.reshape(g.Nxtot, g.Nytot)[sx, sy]
    ux, uy = grab(w["ux"]), grab(w["uy"])
    ut = np.sqrt(1 + ux * ux + uy * uy)
    assert np.array_equal(vx[:, :, 1], ux / ut) and np.array_equal(vy[:, :, 1], uy / ut)
    assert np.array_equal(n[:, :, 1], grab(w["n"]))
    assert np.max(np.abs(T[:, :, 1] / np.exp(grab(w["yT"])) - 1)) <= 2.3e-16
    # only slice k is written
    for k in (0, 2):
        assert np.all(np.isnan(T[:, :, k])) and np.all(np.isnan(vx[:, :, k]))
    # the velocity is the four-velocity divided by u^tau: u = v / sqrt(1 - v^2)
    v2 = vx[:, :, 1] ** 2 + vy[:, :, 1] ** 2
    assert np.max(np.abs(vx[:, :, 1] / np.sqrt(1 - v2) - ux)) <= 1e-14 * max(1.0, np.max(np.abs(ux)))
    assert mu == np.max(np.hypot(ux, uy)) and mv == np.max(np.hypot(vx[:, :, 1], vy[:, :, 1]))
    assert 0.0 < mv < 1.0 and mu > mv
    s.close()


def test_slice_maxima_propagate_nan(oracle):
    g, s = _evolved(oracle, steps=1)
    w = s.get_work(("ux",))
    w["ux"][g.lin(g.nghost + 5, g.nghost + 7)] = np.nan
    s.set_work(w)
    mu, mv = s.background_slice(0, 1, *_tables(g, 1))
    assert math.isnan(mu) and math.isnan(mv)
    s.close()


def test_slice_refuses_bad_tables(oracle):
    g, s = _evolved(oracle, steps=1)
    T, vx, vy, n = _tables(g, 2)
    with pytest.raises(ValueError, match="Fortran"):
        s.background_slice(0, 2, np.zeros((g.Nx, g.Ny, 2)), vx, vy, n)
    with pytest.raises(F.FivoError):
        s.background_slice(2, 2, T, vx, vy, n)
    s.close()


def test_uniform_flow_known_answer(oracle):
    """A uniform medium in uniform transverse flow: every slice cell must hold
    T0, v = u / sqrt(1 + u.u) and the EOS's n(T0, alpha0) at tau0."""
    from oracle_unit import OracleUnit
    U0X, U0Y, T0, A0 = 0.6, -0.35, 0.3, -5.0
    prims = lambda g: dict(T=np.full(g.Ntot, T0), alpha=np.full(g.Ntot, A0), ux=np.full(g.Ntot, U0X),
                           uy=np.full(g.Ntot, U0Y))
    res = F.export_background2d(prims, N=16, tau0=0.4, tauf=0.4, nt=1, xmax=4.0, lib=oracle, verbose=False)
    ut = math.sqrt(1 + U0X * U0X + U0Y * U0Y)
    assert np.max(np.abs(res["T_grid"] / T0 - 1)) < 1e-12
    assert np.max(np.abs(res["vx_grid"] - U0X / ut)) < 1e-12 and np.max(np.abs(res["vy_grid"] - U0Y / ut)) < 1e-12
    n0 = OracleUnit(oracle).eos_Pne(F.LatticeHRGEOS(), T0, A0 * T0)[1]
    assert np.max(np.abs(res["n_grid"] / n0 - 1)) < 1e-10
    assert res["_run"]["max_v"] == pytest.approx(math.hypot(U0X, U0Y) / ut, rel=1e-12)


def test_export_on_the_oracle(oracle, tmp_path):
    out = tmp_path / "bg.npz"
    res = F.export_background2d(_event, out=str(out), N=24, tau0=0.4, tauf=1.0, nt=4, xmax=8.0, lib=oracle,
                                verbose=False)
    d = F.load_background(str(out))
    keys = {"label", "xgrid", "ygrid", "tgrid", "T_grid", "vx_grid", "vy_grid", "n_grid", "ic", "N", "xmax",
            "tau0", "tauf", "DsT", "eta_over_s", "zeta_over_s", "velocity_convention", "note"}
    assert set(d) == keys and d["label"] == "fivo_2d_background"
    for k in ("T_grid", "vx_grid", "vy_grid", "n_grid"):
        assert d[k].shape == (24, 24, 4) and np.array_equal(d[k], res[k])
    assert np.array_equal(d["tgrid"], np.linspace(0.4, 1.0, 4))
    g = F.make_grid2d(24, 24, xmax=8.0, ymax=8.0)
    assert np.array_equal(d["xgrid"], g.xC[3:27]) and np.array_equal(d["ygrid"], g.yC[3:27])
    assert np.all(d["T_grid"] > 0) and np.max(np.hypot(d["vx_grid"], d["vy_grid"])) < 1
    # the flow develops: the rim accelerates between the first and last slices
    vmax = [np.max(np.hypot(d["vx_grid"][:, :, k], d["vy_grid"][:, :, k])) for k in range(4)]
    assert vmax[-1] > vmax[0] + 0.05
    # slice 0 is the initial state: T0 of the event on the interior
    T0 = _event(g)["T"].reshape(g.Nxtot, g.Nytot)[3:27, 3:27]
    hot = T0 > 0.05
    assert np.max(np.abs(d["T_grid"][:, :, 0][hot] / T0[hot] - 1)) < 1e-9


def test_export_refuses_superluminal_table(oracle, monkeypatch):
    """The exporter's |v| < 1 check (tools/export_background2d.jl:97): a slice that
    reports max |v| >= 1 (a planted fault in the host mirror) must stop the export."""
    monkeypatch.setattr(F.Solver, "background_slice", lambda self, *a, **k: (1.5, 1.2))
    with pytest.raises(F.FivoError, match="four-velocity"):
        F.export_background2d(_event, N=16, tau0=0.4, tauf=0.5, nt=2, xmax=8.0, lib=oracle, verbose=False)
