"""GPU tests of the background readout (fivo_background_slice, the device side of
tools/export_background2d.jl): the slice kernel against the oracle's slice on the
same primitives (non-square grids that do not fill the kernel's 16 x 16 tiles,
host and device tables), and the whole exporter against the oracle's run (final
T and v within 1e-9), and at the reference's default size."""
import math

import numpy as np
import pytest

import fivo2d as F
from fivo2d import background as B
from fivo2d import ic

pytestmark = pytest.mark.gpu


def _event(g):
    return ic.hotspot_event(g, seed=11, nspots=60, sigma=0.5, Tpeak=0.42)


def _ulps(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) / np.finfo(np.float64).eps


@pytest.mark.parametrize("shape", [(100, 70), (33, 65), (512, 512)])
def test_slice_kernel_matches_oracle_on_the_same_primitives(oracle, hip, shape):
    import torch
    nx, ny = shape
    g = F.make_grid2d(nx, ny, xmax=9.0, ymax=9.0 * ny / nx)
    m = B.background_model()
    so = F.Solver(g, m, lib=oracle)
    so.init_from_prims(0.4, **_event(g))
    r = so.run_sim_2d(tau0=0.4, taufinal=10.0, CFL=0.2, CFLtau=0.02, max_steps=2 if nx < 512 else 1)
    assert r["ok"]
    so.update_primitives_2d(r["tau"])
    w = so.get_work(("yT", "ux", "uy", "n"))
    sh = F.Solver(g, m, lib=hip)
    sh.set_work(w)
    nt, k = 3, 2
    ref = [np.full((nx, ny, nt), np.nan, order="F") for _ in range(4)]
    mo = so.background_slice(k, nt, *ref)
    host = [np.full((nx, ny, nt), np.nan, order="F") for _ in range(4)]
    mh = sh.background_slice(k, nt, *host)
    dev = [torch.full((nt * ny * nx,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(4)]
    md = sh.background_slice(k, nt, *dev)
    dev = [x.cpu().numpy().reshape(nt, ny, nx).transpose(2, 1, 0) for x in dev]
    for name, a, b, c in zip(("T", "vx", "vy", "n"), host, dev, ref):
        assert np.array_equal(a, b, equal_nan=True), name           # host and device tables: same bits
        assert np.all(np.isnan(a[:, :, :k])) and not np.any(np.isnan(a[:, :, k])), name
        u = _ulps(a[:, :, k], c[:, :, k])
        print(f"{shape} {name}: max {u:.1f} ulp from the oracle, {np.mean(a[:, :, k] == c[:, :, k]):.4f} equal",
              flush=True)
        assert u <= (1.0 if name == "T" else 0.0), name             # exp: 1 ulp; the rest exact
    assert md == mh
    assert abs(mh[0] / mo[0] - 1) < 4.5e-16 and abs(mh[1] / mo[1] - 1) < 4.5e-16
    so.close()
    sh.close()


def _smooth(g):
    """An elliptic fireball on a 0.15 GeV medium with an offset fugacity lump: no
    vacuum cells, whose recovery-failure flips make lumpy events ill-conditioned
    at the 1e-12 level (see the conditioning tests)."""
    X, Y = np.meshgrid(g.xC, g.yC, indexing="ij")
    T = 0.15 + 0.30 * np.exp(-(X ** 2 / 1.3 + Y ** 2 / 0.8) / 6)
    a = -4.0 + 0.8 * np.exp(-((X + 1) ** 2 + Y ** 2) / 4)
    return dict(T=T.ravel(), alpha=a.ravel())


def test_export_matches_oracle(oracle, hip):
    kw = dict(N=48, tau0=0.4, tauf=1.2, nt=5, xmax=8.0, verbose=False)
    a = F.export_background2d(_smooth, lib=hip, **kw)
    b = F.export_background2d(_smooth, lib=oracle, **kw)
    assert a["_run"]["steps"] == b["_run"]["steps"]
    dT = np.max(np.abs(a["T_grid"] / b["T_grid"] - 1))
    dv = max(np.max(np.abs(a["vx_grid"] - b["vx_grid"])), np.max(np.abs(a["vy_grid"] - b["vy_grid"])))
    nn = b["n_grid"] > 1e-12 * np.max(b["n_grid"])
    dn = np.max(np.abs(a["n_grid"][nn] / b["n_grid"][nn] - 1))
    print(f"export 48^2, 5 slices, {a['_run']['steps']} steps: max dT/T {dT:.2e}, max dv {dv:.2e}, "
          f"max dn/n {dn:.2e}", flush=True)
    assert dT <= 1e-9 and dv <= 1e-9 and dn <= 1e-9
    assert a["_run"]["max_v"] == pytest.approx(b["_run"]["max_v"], rel=1e-9)


def test_export_at_the_reference_defaults(hip, tmp_path):
    """200^2, +-16 fm, tau 0.4 -> 8 in 39 slices, SSPRK2, CFL 0.2, CFLtau 0.02, the
    exporter's medium; a seeded hot-spot event stands in for the reference's event
    file, which it does not ship."""
    out = tmp_path / "bg.npz"
    res = F.export_background2d(lambda g: ic.hotspot_event(g, seed=20, nspots=200, sigma=0.5, Tpeak=0.45),
                                out=str(out), lib=hip, verbose=True)
    d = F.load_background(str(out))
    run = res["_run"]
    print(f"defaults: {run['steps']} steps in {run['wall_s']:.2f} s, max |u| {run['max_u']:.3f}, "
          f"max |v| {run['max_v']:.4f}", flush=True)
    assert d["T_grid"].shape == (200, 200, 39) and np.all(np.isfinite(d["T_grid"]))
    assert 0.5 < run["max_v"] < 1.0 and run["max_u"] > run["max_v"]
    assert d["tgrid"][0] == 0.4 and d["tgrid"][-1] == 8.0
    # the medium cools: the hottest cell falls from slice to slice
    tmax = d["T_grid"].reshape(-1, 39).max(axis=0)
    assert np.all(np.diff(tmax) < 0)


def test_ensemble_export_equals_one_event_exports(hip):
    """export_backgrounds2d (every event advanced by one ensemble run per slice
    interval, slices written per event in HBM) against export_background2d run on
    each event alone, both with the reference exporter's SSPRK2: the same tables, bit
    for bit (the
    ensemble autotunes its face tile first; tuning restores the warm starts its
    primitive pass moves, so it changes nothing)."""
    ics = [lambda g, s=s: ic.hotspot_event(g, seed=s, nspots=40, sigma=0.6, Tpeak=0.42) for s in (31, 32, 33)]
    Ns = [40, 48, 32]
    kw = dict(tau0=0.4, tauf=1.0, nt=4, xmax=8.0)
    many = F.export_backgrounds2d(ics, N=Ns, lib=hip, **kw)
    for e, (icf, N) in enumerate(zip(ics, Ns)):
        one = F.export_background2d(icf, N=N, lib=hip, verbose=False, **kw)
        assert many[e]["_run"]["steps"] == one["_run"]["steps"], e
        for k in ("T_grid", "vx_grid", "vy_grid", "n_grid", "xgrid", "ygrid", "tgrid"):
            assert np.array_equal(many[e][k], one[k]), (e, k)
        print(f"event {e}: {one['_run']['steps']} steps, max |v| {one['_run']['max_v']:.4f}: identical", flush=True)
