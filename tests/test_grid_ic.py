"""initialize_from_grid_csv! (main2D.jl:238-301) in the host mirror, checked by
what the interpolation and the taper fix, not by retyping them:
- bilinear interpolation is exact for fields linear in x and y, and returns the
  nodal values where cell centres sit on file nodes;
- the normalised, edge-clamped blur leaves a constant field unchanged and keeps
  a field's range (it is a convex combination);
- beyond the file's extent the state is vacuum (T = T_MIN, alpha = 0), inside
  r_t0 = min(x_max, y_max) - taper_width it is untouched;
- the build goes through the library's set_cell! + finalize_ic! (the oracle here,
  the same entry point as the HIP library), with the reference's split of failed
  cells into vacuum and cells with matter.
"""
import numpy as np

import fivo2d as F
from fivo2d import ic


def write(tmp_path, xs, ys, T, A, name="ic.csv"):
    p = tmp_path / name
    ic.write_grid_csv(p, xs, ys, T, A)
    return str(p)


def test_bilinear_exact_for_linear_fields(tmp_path):
    xs = np.linspace(-9.0, 9.0, 37)
    ys = np.linspace(-8.0, 8.0, 33)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    T = 0.3 + 0.01 * X - 0.007 * Y
    A = -3.0 + 0.05 * X + 0.02 * Y
    csv = write(tmp_path, xs, ys, T, A)
    g = F.make_grid2d(40, 30, xmax=10.0, ymax=10.0)
    p = ic.grid_csv(g, csv, taper_width=0.0)
    Xc, Yc = np.meshgrid(g.xC, g.yC, indexing="ij")
    inside = (Xc > xs[0]) & (Xc < xs[-1]) & (Yc > ys[0]) & (Yc < ys[-1])
    Tp = p["T"].reshape(g.Nxtot, g.Nytot)
    Ap = p["alpha"].reshape(g.Nxtot, g.Nytot)
    assert np.max(np.abs(Tp - (0.3 + 0.01 * Xc - 0.007 * Yc))[inside]) < 1e-14
    assert np.max(np.abs(Ap - (-3.0 + 0.05 * Xc + 0.02 * Yc))[inside]) < 1e-13
    assert np.all(Tp[~inside] == F.T_MIN) and np.all(Ap[~inside] == 0.0)


def test_nodes_on_cell_centres_and_taper(tmp_path):
    g = F.make_grid2d(64, 64, xmax=8.0, ymax=8.0)
    xs, ys = g.xC[g.nghost + 4: g.nghost + 60], g.yC[g.nghost + 4: g.nghost + 60]
    rng = np.random.default_rng(3)
    T = 0.2 + 0.2 * rng.random((len(xs), len(ys)))
    A = -5.0 + rng.random((len(xs), len(ys)))
    csv = write(tmp_path, xs, ys, T, A)
    tw = 1.0
    p = ic.grid_csv(g, csv, taper_width=tw)
    Xc, Yc = np.meshgrid(g.xC, g.yC, indexing="ij")
    Tp = p["T"].reshape(g.Nxtot, g.Nytot)
    rt0 = min(xs[-1], ys[-1]) - tw
    sub = np.s_[g.nghost + 5: g.nghost + 59, g.nghost + 5: g.nghost + 59]
    untouched = (np.hypot(Xc, Yc) <= rt0)[sub]
    assert untouched.sum() > 1000
    assert np.allclose(Tp[sub][untouched], T[1:-1, 1:-1][untouched], rtol=0, atol=1e-15)
    # the taper only lowers T, toward T_MIN at the file's half extent
    assert np.all(Tp[sub] <= T[1:-1, 1:-1] + 1e-15)


def test_blur_keeps_constants_and_range(tmp_path):
    xs = np.linspace(-6.0, 6.0, 49)
    ys = np.linspace(-6.0, 6.0, 49)
    g = F.make_grid2d(32, 32, xmax=7.0, ymax=7.0)
    c = write(tmp_path, xs, ys, np.full((49, 49), 0.31), np.full((49, 49), -4.0), "c.csv")
    p0 = ic.grid_csv(g, c, taper_width=0.0, smooth_fm=0.5)
    Xc, Yc = np.meshgrid(g.xC, g.yC, indexing="ij")
    inside = ((Xc > xs[0]) & (Xc < xs[-1]) & (Yc > ys[0]) & (Yc < ys[-1])).ravel()
    assert np.max(np.abs(p0["T"][inside] - 0.31)) < 1e-15 and np.max(np.abs(p0["alpha"][inside] + 4.0)) < 1e-14
    rng = np.random.default_rng(5)
    T = 0.2 + 0.3 * rng.random((49, 49))
    r = write(tmp_path, xs, ys, T, -4.0 + rng.random((49, 49)), "r.csv")
    sharp = ic.grid_csv(g, r, taper_width=0.0)["T"][inside]
    soft = ic.grid_csv(g, r, taper_width=0.0, smooth_fm=0.6)["T"][inside]
    assert soft.min() >= T.min() and soft.max() <= T.max()
    assert np.std(soft) < 0.5 * np.std(sharp)


def test_build_on_the_library_counts_vacuum_apart(oracle, tmp_path):
    g = F.make_grid2d(48, 48, xmax=12.0, ymax=12.0)
    xs = np.linspace(-7.0, 7.0, 57)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    T = 0.05 + 0.4 * np.exp(-(X ** 2 + 0.5 * Y ** 2) / 8.0)
    A = -6.0 + 2.0 * np.exp(-(X ** 2 + Y ** 2) / 8.0)
    csv = write(tmp_path, xs, xs, T, A)
    m = F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, enable_bulk=True, enable_diff=True,
                         kappa_coeff=0.1163)
    s = F.Solver(g, m, lib=oracle)
    out = s.initialize_from_grid_csv(0.6, csv)
    print(out)
    ref = F.Solver(g, m, lib=oracle)
    nb = ref.init_from_prims(0.6, **ic.grid_csv(g, csv))
    assert out["nbad"] == 0 and out["nbad"] + out["nvacuum"] == nb
    assert np.array_equal(s.get_state(), ref.get_state())
    r = s.run_sim_2d(tau0=0.6, taufinal=0.8)
    assert r["ok"]
