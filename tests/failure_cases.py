"""Inputs that force the solver's failure and fallback branches, built the same way
for any library with the C ABI (the oracle on the CPU, the HIP library on the GPU).
Test infrastructure shared by test_failure_branches.py and
test_gpu_failure_branches.py.

  pathological_states   2000 cells of NaN, +-Inf, huge negatives, negative charge,
                        superluminal momentum, oversized currents and stresses,
                        through sanitize_stage_2d! and state_ok_2d
                        (src2d/floors2d.jl:218-227, :233-245 and the predicate after it)
  stale_warm_starts     warm starts at log(T_MIN) on cells with matter: the warm
                        Newton cannot recover them and cons_to_prim_2d! retries from
                        the fresh seed (src2d/primrec2d.jl:444-486)
  charge_row_unsolvable diffusion currents with nu^tau > J^tau: no n >= 0 closes the
                        charge row, the hydro rows converge, and the staged phi
                        fallback closes the solve (src2d/primrec2d.jl:723-738)
  superluminal_cell     one cell with |S| = 1.3 E: every stage is inadmissible, the
                        step is halved max_halve times and fails with the entry state
                        restored (src2d/timestepper2d.jl:74-104, 148-188)
"""
import numpy as np

import fivo2d as F
from fivo2d import ic

TAU0 = 0.4


def full_model():
    return F.build_model_2d(eos=F.LatticeHRGEOS(), enable_shear=True, eta_over_s=0.08, tauShear_coeff=0.2,
                            enable_bulk=True, zeta_over_s=0.1, tauPi_coeff=15.0, enable_diff=True,
                            kappa_coeff=0.1163, pi_clip_factor=1.0, consistent_m2=True)


def grid():
    return F.make_grid2d(48, 48, xmax=8.0, ymax=8.0)


def event(lib, steps=0):
    """A small lumpy full-sector event on `lib`, optionally evolved `steps` steps."""
    g, m = grid(), full_model()
    s = F.Solver(g, m, lib=lib)
    s.init_from_prims(TAU0, **ic.hotspot_event(g, seed=7, nspots=40, sigma=0.6, Tpeak=0.45))
    tau = TAU0
    if steps:
        r = s.run_sim_2d(tau0=TAU0, taufinal=1e9, CFL=0.2, CFLtau=0.05, integrator="ssprk3", max_steps=steps)
        assert r["ok"]
        tau = r["tau"]
    else:
        s.update_primitives_2d(TAU0)
    return s, tau


def interior_cells(g):
    sx, sy = g.interior()
    ix, iy = np.meshgrid(np.arange(g.Nxtot)[sx], np.arange(g.Nytot)[sy], indexing="ij")
    return (ix * g.Nytot + iy).ravel()


def pathological_states(U, g, layout, n=2000, seed=3):
    """A copy of U (Ntot, nv) with n interior cells made pathological, one of ten
    kinds each; returns (U', kinds) with kinds[cell] = kind or -1."""
    rng = np.random.default_rng(seed)
    U = U.copy()
    cells = rng.choice(interior_cells(g), size=n, replace=False)
    kinds = np.full(U.shape[0], -1)
    names = layout.names
    iNux = names.index("nux") if "nux" in names else None
    iPi = names.index("Pi") if "Pi" in names else None
    iPixx = names.index("pixx") if "pixx" in names else None
    for c in cells:
        k = int(rng.integers(10))
        kinds[c] = k
        E = abs(U[c, 3]) + 1e-3
        D = abs(U[c, 0])
        if k == 0:
            U[c, rng.integers(U.shape[1])] = np.nan
        elif k == 1:
            U[c, rng.integers(U.shape[1])] = np.inf
        elif k == 2:
            U[c, rng.integers(U.shape[1])] = -np.inf
        elif k == 3:
            U[c, 3] = -1e300
        elif k == 4:
            U[c, 0] = -(D + 1e-3)
        elif k == 5:
            U[c, 1], U[c, 2] = 3.0 * E, -0.5 * E
        elif k == 6 and iNux is not None:
            U[c, iNux], U[c, iNux + 1] = 10 * D / TAU0 + 1e-3, -10 * D / TAU0
        elif k == 7 and iPi is not None:
            U[c, iPi] = -10 * E
        elif k == 8 and iPixx is not None:
            U[c, iPixx], U[c, iPixx + 1] = 10 * E, -7 * E
        else:
            U[c, 3] = 1e-30
    return U, kinds


def stale_warm_starts(s, frac=0.5, seed=5):
    """Warm starts at log(T_MIN) (zero flow, phi = 0) on a share of the interior cells
    whose last recovery converged with T > 0.1 GeV. Returns the cells."""
    w = s.get_work(("yT", "ok", "x0_yT", "x0_phi", "x0_ux", "x0_uy"))
    cells = interior_cells(s.g)
    hot = cells[(w["ok"][cells] == 1) & (np.exp(w["yT"][cells]) > 0.1)]
    rng = np.random.default_rng(seed)
    pick = rng.choice(hot, size=max(1, int(frac * hot.size)), replace=False)
    x0 = {k: w[k].copy() for k in ("x0_yT", "x0_phi", "x0_ux", "x0_uy")}
    x0["x0_yT"][pick] = np.log(F.T_MIN)
    x0["x0_phi"][pick] = 0.0
    x0["x0_ux"][pick] = 0.0
    x0["x0_uy"][pick] = 0.0
    s.set_work(x0)
    return pick


def cold_starts(s, cells):
    """The same cells without a warm start (NaN): the reference's fresh seed."""
    w = s.get_work(("x0_yT", "x0_phi", "x0_ux", "x0_uy"))
    for k in w:
        w[k][cells] = np.nan
    s.set_work(w)


def charge_row_unsolvable(s, tau, k=40, seed=9):
    """On k interior cells with flow (|u| > 0.02) and charge, set nu along u so that
    nu^tau = u.nu / u^tau = 3 J^tau. Returns the cells."""
    U = s.get_state()
    w = s.get_work(("ux", "uy", "ok"))
    names = s.model.layout.names
    iNux = names.index("nux")
    cells = interior_cells(s.g)
    ux, uy = w["ux"][cells], w["uy"][cells]
    un = np.hypot(ux, uy)
    J = U[cells, 0] / tau
    cand = cells[(un > 0.02) & (J > 1e-10) & (w["ok"][cells] == 1)]
    rng = np.random.default_rng(seed)
    pick = rng.choice(cand, size=min(k, cand.size), replace=False)
    for c in pick:
        uxc, uyc = w["ux"][c], w["uy"][c]
        u2 = uxc * uxc + uyc * uyc
        ut = np.sqrt(1 + u2)
        a = 3.0 * (U[c, 0] / tau) * ut / u2          # nu = a u: nu^tau = a u^2 / u^tau
        U[c, iNux], U[c, iNux + 1] = a * uxc, a * uyc
    s.set_state(U)
    return pick


def superluminal_cell(s):
    """One interior cell near the centre with |S| = 1.3 E (not invertible)."""
    g = s.g
    U = s.get_state()
    c = (g.nghost + g.Nx // 2) * g.Nytot + g.nghost + g.Ny // 2
    U[c, 1], U[c, 2] = 1.3 * U[c, 3], 0.0
    s.set_state(U)
    return c
